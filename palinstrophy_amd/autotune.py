"""Autotuner for the FFT pass kernels (radix sequence, threads per transform,
LDS padding) with a JSON cache.

Each padded length M = 3N/2 has ``palx_n_variants(N)`` compiled variants of
the row pass and of the CNAB2/RK3 column pass.  ``tune_ensemble`` times every
variant of every size class of an ensemble in its batched form (the members
of a class in one launch) and keeps the fastest; ``tune_handle`` does the same
for a single field through a one-member ensemble of that size.  Results are
cached per (device, library build, N, precision, batch size) in
``$PALX_TUNE_CACHE`` or ``~/.cache/palinstrophy_amd/tune.json``.

There is no counterpart in the reference, which delegates plans to cuFFT /
pocketfft (turbo_simulator.py:970-979, 1255-1261).
"""
from __future__ import annotations

import json
import os
import tempfile

from . import _palx as P


def cache_path() -> str:
    p = os.environ.get("PALX_TUNE_CACHE")
    if p:
        return p
    return os.path.join(os.path.expanduser("~"), ".cache", "palinstrophy_amd", "tune.json")


def library_id() -> str:
    """Identifies the kernel build: size and mtime of libpalx.so."""
    st = os.stat(P.LIB_PATH)
    return f"{st.st_size}-{int(st.st_mtime)}"


def device_id() -> str:
    import torch

    props = torch.cuda.get_device_properties(torch.cuda.current_device())
    return f"{props.name}|{getattr(props, 'gcnArchName', '')}"


def entry_key(device: str, lib: str, N: int, precision: int, count: int) -> str:
    return f"{device}|{lib}|N={int(N)}|prec={int(precision)}|batch={int(count)}"


def load_cache(path: str | None = None) -> dict:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def save_cache(data: dict, path: str | None = None) -> None:
    """Write atomically (temporary file + rename) so concurrent runs never
    read a torn file."""
    path = path or cache_path()
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    fd, tmp = tempfile.mkstemp(dir=os.path.dirname(path) or ".", prefix=".tune-", suffix=".json")
    with os.fdopen(fd, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune_ensemble(ens: "P.Ensemble", reps: int = 10, use_cache: bool = True, path: str | None = None) -> dict:
    """Pick the fastest variants for every class of `ens` (before its fields
    are loaded: timing the column pass overwrites member state).  Returns
    {N: {"vrow", "vcol", "ms_row", "ms_col", "cached"}}."""
    classes = ens.classes()
    dev, lib = device_id(), library_id()
    cache = load_cache(path) if use_cache else {}
    keys = {N: entry_key(dev, lib, N, ens.precision, cnt) for N, cnt, _, _ in classes}
    out = {}
    if use_cache and all(k in cache for k in keys.values()):
        for N, _, _, _ in classes:
            e = cache[keys[N]]
            ens.set_variants(N, e["vrow"], e["vcol"])
            out[N] = dict(e, cached=True)
        return out
    ms = ens.tune(reps)
    for c, (N, cnt, vr, vc) in enumerate(ens.classes()):
        e = {"vrow": vr, "vcol": vc,
             "ms_row": [float(x) for x in ms[c, 0]], "ms_col": [float(x) for x in ms[c, 1]]}
        cache[keys[N]] = e
        out[N] = dict(e, cached=False)
    if use_cache:
        save_cache(cache, path)
    return out


def tune_handle(h: "P.Handle", reps: int = 10, use_cache: bool = True, path: str | None = None) -> dict:
    """Tune the step passes of one field of size h.N (timed on a scratch
    one-member ensemble, so h's state is untouched) and apply them to h."""
    dev, lib = device_id(), library_id()
    key = entry_key(dev, lib, h.N, h.precision, 1)
    cache = load_cache(path) if use_cache else {}
    if use_cache and key in cache:
        e = cache[key]
        h.set_variants(e["vrow"], e["vcol"])
        return dict(e, cached=True)
    ens = P.Ensemble([h.N], h.precision)
    try:
        res = tune_ensemble(ens, reps=reps, use_cache=use_cache, path=path)[h.N]
    finally:
        ens.close()
    h.set_variants(res["vrow"], res["vcol"])
    return res
