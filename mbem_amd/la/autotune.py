"""Tile autotuner for the FP64 matrix-core GEMMs of the batched solves.

The library compiles a fixed set of tile configurations (``mbem_mgemm_config``:
block tile BM x BN, depth slice BK, waves along M and N).  For a given GEMM
shape the tuner times every configuration that fits, on the stream the solve
runs on, with HIP events, and keeps the fastest.  Results are cached per
(kernel, shape) in the process; ``MBEM_AUTOTUNE_CACHE`` names an optional JSON
file that persists them between processes.
"""
from __future__ import annotations

import ctypes
import json
import os

import torch

from .. import _lib

_CACHE: dict = {}
_LOADED_FILE = False


def configs() -> list:
    """[(index, {"bm", "bn", "bk", "wm", "wn"})] of the compiled tiles."""
    lib = _lib.load()
    out = []
    buf = (ctypes.c_int32 * 5)()
    for i in range(lib.mbem_mgemm_configs()):
        _lib.call("mbem_mgemm_config", i, ctypes.cast(buf, ctypes.c_void_p))
        out.append((i, dict(zip(("bm", "bn", "bk", "wm", "wn"), list(buf)))))
    return out


def _file():
    return os.environ.get("MBEM_AUTOTUNE_CACHE")


def _load_file():
    global _LOADED_FILE
    if _LOADED_FILE:
        return
    _LOADED_FILE = True
    path = _file()
    if path and os.path.exists(path):
        with open(path) as fh:
            for k, v in json.load(fh).items():
                _CACHE.setdefault(k, v)


def _save_file():
    path = _file()
    if not path:
        return
    tmp = path + ".tmp"
    with open(tmp, "w") as fh:
        json.dump(_CACHE, fh, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(key: str, run, candidates=None, reps: int = 3, warmup: int = 1) -> int:
    """Index of the fastest configuration for ``key``; ``run(cfg)`` launches
    the GEMM once with configuration ``cfg`` on the current stream (it must
    not allocate or synchronise).  ``candidates`` defaults to the compiled
    tiles; a caller may add its own ids (the library GEMM).  Timings are kept
    in ``report(key)``."""
    _load_file()
    if key in _CACHE:
        return int(_CACHE[key]["best"])
    cands = [i for i, _ in configs()] if candidates is None else list(candidates)
    if not cands:
        raise ValueError("no tile configuration to tune over")
    st = torch.cuda.current_stream()
    times = {}
    for cfg in cands:
        for _ in range(warmup):
            run(cfg)
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(reps):
            run(cfg)
        e1.record(st)
        e1.synchronize()
        times[str(cfg)] = e0.elapsed_time(e1) / reps
    best = min(times, key=times.get)
    _CACHE[key] = {"best": int(best), "ms": times}
    _save_file()
    return int(best)


def report(key: str):
    """{"best": cfg, "ms": {cfg: milliseconds}} for a tuned key, or None."""
    return _CACHE.get(key)


def clear():
    _CACHE.clear()
