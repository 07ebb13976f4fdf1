"""C-ABI boundary and host logic, no GPU: libpalx.so loads and exports every
entry point include/palx.h declares; the host-side seeded init equals the
oracle's; argument errors surface as the reference's exception types."""
import os
import re

import numpy as np
import pytest

from oracle import dns_oracle as O
from palinstrophy_amd import _palx as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "palx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*int\s+(palx_\w+)\s*\(", text, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = P.load()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(P.SIGNATURES), set(names) ^ set(P.SIGNATURES)


def test_binding_constants_match_the_header():
    """Every enum constant of include/palx.h that the ctypes binding also
    defines (under its PALX_ name or without the prefix) has the same value."""
    text = open(os.path.join(ROOT, "include", "palx.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    pairs = re.findall(r"\b(PALX_[A-Z0-9_]+)\s*=\s*(-?(?:0x[0-9a-fA-F]+|\d+))\s*[,}\n]", text)
    assert len(pairs) >= 50
    matched = 0
    for name, val in pairs:
        for attr in (name, name[len("PALX_"):]):
            if hasattr(P, attr):
                assert getattr(P, attr) == int(val, 0), (name, attr, getattr(P, attr), val)
                matched += 1
                break
    assert matched >= 30, matched


def test_library_exports_no_other_c_function():
    """The only C-linkage functions libpalx.so defines are the header's entry
    points (helpers stay internal; C++ symbols are mangled, _Z...)."""
    import shutil
    import subprocess
    nm = shutil.which("nm")
    if nm is None:
        pytest.skip("nm not available")
    out = subprocess.run([nm, "-D", "--defined-only", P.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    funcs = {f[2] for f in (l.split() for l in out.splitlines()) if len(f) == 3 and f[1] in "TtWw"}
    c_funcs = {f for f in funcs if not f.startswith("_Z")}
    assert c_funcs == set(header_functions()), c_funcs ^ set(header_functions())


@pytest.mark.parametrize("N,spec,seed", [(16, "PAO", 1), (64, "KM3", 4895), (256, "PAO", 4895),
                                         (512, "PAO", 1), (256, "KM3", 3),
                                         (96, "KM3", 4895), (192, "PAO", 7), (384, "KM3", 1)])
def test_host_init_equals_oracle(N, spec, seed):
    S = O.create_state(N, 1e4, 10, 0.5, seed, spec, dtype=np.float32)
    uc0, st = P.pao_generate(N, 10.0, 1e4, seed, spec == "KM3")
    assert np.array_equal(uc0, S.uc0)           # value-exact (signed zeros aside)
    assert int(st[0]) == S.pao.seed_out
    assert st[1] == S.pao.Q2 and st[2] == S.pao.W2
    assert np.float32(st[4]) == S.pao.visc
    assert np.array_equal(P.calcom_f32(uc0), S.om2.astype(np.complex64))


def test_host_init_km3_matches_reference_golden():
    g = dict(np.load(os.path.join(ROOT, "tests", "golden", "km3_rk3_N32_s4895.npz")))
    uc0, _ = P.pao_generate(32, float(g["K0"]), float(g["Re"]), 4895, True)
    assert np.allclose(uc0, g["uc0"], rtol=2e-6, atol=1e-12)


def test_seed_kat_from_host_init():
    _, st = P.pao_generate(512, 10.0, 1e4, 1, False)
    assert int(st[0]) == 149


def test_invalid_arguments_raise_value_error():
    with pytest.raises(ValueError):
        P.pao_generate(7, 10.0, 1e4, 1, False)


def test_variant_query_is_a_host_call():
    """palx_variant_ok needs no GPU: every variant of N=4096 exists in both
    precisions; N=64 has the three base-plan variants; at N=16384 (FP32
    only) a complex float 24576-point transform needs 196 KB of LDS, more
    than a block may use, so only the split exchange (6, 7) is compiled, and
    likewise for N=8192 in FP64; unsupported sizes raise."""
    assert all(P.variant_ok(4096, p, v) for p in (4, 8) for v in range(P.MAX_VARIANTS))
    assert not P.variant_ok(4096, 4, P.MAX_VARIANTS) and not P.variant_ok(4096, 4, -1)
    assert [P.variant_ok(64, 8, v) for v in range(4)] == [True, True, True, False]
    assert [P.variant_ok(16384, 4, v) for v in range(P.MAX_VARIANTS)] == [False] * 6 + [True] * 2
    # N=8192: a 12288-point complex double transform needs 192 KB, its split form 99 KB
    assert [P.variant_ok(8192, 8, v) for v in range(P.MAX_VARIANTS)] == [False] * 6 + [True] * 2
    assert all(P.variant_ok(8192, 4, v) for v in range(P.MAX_VARIANTS))
    # N=12288 (9*2^k padded length, base plan only): variants 0-2 and the split 6
    assert [P.variant_ok(12288, 4, v) for v in range(P.MAX_VARIANTS)] == [True] * 3 + [False] * 3 + [True, False]
    assert [P.variant_ok(12288, 8, v) for v in range(P.MAX_VARIANTS)] == [False] * 6 + [True, False]
    with pytest.raises(ValueError):
        P.variant_ok(16384, 8, 6)


def test_seed_env_parsing(monkeypatch):
    from palinstrophy_amd import turbo_simulator as ts
    monkeypatch.delenv("SCIPYTURBO_SEED", raising=False)
    assert ts.pao_seed_from_env(5) == 5
    monkeypatch.setenv("SCIPYTURBO_SEED", "4895")
    assert ts.pao_seed_from_env(1) == 4895
    for bad in ("0", "5011", "x"):
        monkeypatch.setenv("SCIPYTURBO_SEED", bad)
        with pytest.raises(ValueError):
            ts.pao_seed_from_env(1)


def test_cpu_backend_is_refused():
    from palinstrophy_amd import turbo_simulator as ts
    with pytest.raises(RuntimeError):
        ts.get_xp("cpu")


def test_spectrum_average_trapezoid_and_csv(tmp_path):
    from palinstrophy_amd import turbo_simulator as ts

    class FakeS:
        Nbase, NX_full, NZ_full, NK_full = 8, 12, 12, 7
    avg = ts.SpectrumAverageState()
    e1 = np.array([1.0, 2.0, 0.0, 4.0])
    e2 = np.array([3.0, 2.0, 0.0, 0.0])
    c = np.array([1, 2, 3, 4], dtype=np.int32)
    s1 = ts.SpectrumSample(4, np.sqrt(2.0), 4.0, e1, c)
    s2 = ts.SpectrumSample(4, np.sqrt(2.0), 4.0, e2, c)
    assert ts.update_spectrum_average(avg, FakeS, s1, 0.0, 0)
    assert ts.spectrum_average_kind(avg) == "single snapshot"
    assert ts.update_spectrum_average(avg, FakeS, s2, 0.5, 1)
    assert avg.total_dt == 0.5 and avg.sample_count == 2
    assert np.allclose(ts.spectrum_average_mean_energy(avg), 0.5 * (e1 + e2))
    f = tmp_path / "spec.csv"
    ts.save_spectrum_average_csv(avg, str(f), ["N"], [8])
    lines = f.read_text().splitlines()
    assert lines[0].startswith("# N,AVERAGE_KIND") and lines[2] == "normalized_radius,shell_sum_energy,count"
    assert len(lines) == 3 + 3   # bins with energy > 0: 0, 1, 3


@pytest.mark.parametrize("N,km3", [(256, True), (384, False), (1024, True)])
def test_host_init_does_not_depend_on_the_thread_count(N, km3, monkeypatch):
    """The seeded init draws in loop order and keeps every sum in loop order;
    only per-mode work runs on threads, so 1, 3 and the default thread
    count give the same bits."""
    out = []
    for nt in ("1", "3", None):
        if nt is None:
            monkeypatch.delenv("PALX_INIT_THREADS", raising=False)
        else:
            monkeypatch.setenv("PALX_INIT_THREADS", nt)
        out.append(P.pao_generate(N, 10.0, 1e4, 4895, km3))
    for uc0, st in out[1:]:
        assert np.array_equal(uc0.view(np.uint32), out[0][0].view(np.uint32))
        assert np.array_equal(st, out[0][1])
