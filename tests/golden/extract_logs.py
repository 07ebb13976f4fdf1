"""Extract the reference's recorded `sim 512 10000 10 1001 0.25 {cpu,gpu} PAO`
logs (tools/sim_512_1001_*.txt) into log_sim_512_1001.json: init sums, seed-out,
CFLM0/DT0 and the per-cadence ITERATION records.

Usage: python tests/golden/extract_logs.py /root/reference
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def parse(path):
    out = {"iterations": []}
    for line in open(path, encoding="utf-8"):
        m = re.match(r"\s*(Energy|WiWi|Time scale|Viscosity|PAO seed)\s*=\s*([-+0-9.eE]+)", line)
        if m:
            out[m.group(1)] = m.group(2)
        m = re.search(r"\[NEXTDT INIT\] CFLM=\s*([0-9.]+) DT=\s*([0-9.]+)", line)
        if m:
            out["CFLM0"], out["DT0"] = m.group(1), m.group(2)
        m = re.search(r"ITERATION\s+(\d+) T=([0-9.]+) DT=([0-9.]+) CN=([0-9.]+) CFLM=([0-9.]+) TAU_L=([0-9.]+)", line)
        if m:
            out["iterations"].append([int(m.group(1))] + [m.group(i) for i in range(2, 7)])
    return out


def main(ref):
    res = {}
    for tag in ("cpu_linux", "cpu_macOS", "gpu_linux"):
        res[tag] = parse(os.path.join(ref, "tools", f"sim_512_1001_{tag}.txt"))
    res["case"] = {"N": 512, "Re": 10000.0, "K0": 10.0, "STEPS": 1001, "CFL": 0.25,
                   "UPDATE": 100, "spectrum": "PAO", "method": "CNAB2", "seed": 1}
    with open(os.path.join(HERE, "log_sim_512_1001.json"), "w") as f:
        json.dump(res, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
