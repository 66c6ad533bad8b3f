"""Launch-schedule autotuner for the slab kernels.

Three knobs shape a matvec's schedule on the device:

- the unroll depth of ``k_slab`` (loads in flight per lane: 4, 8 or 16),
  switchable per handle at no cost (``bw_hip_set_unroll``);
- the schedule of the MFMA kernels of the batched matvec (group pairs per
  step, depth of the load pipeline; ``bw_hip_set_mfma_schedule``, per
  precision), also free;
- the work-item size (bytes of factors one workgroup streams), fixed when the
  layout is built (``bw_hip_set_tuning`` before ``HipHmat``).

``tune_unroll`` times each unroll depth on the handle's own data with the
engine's HIP-event timers and keeps the fastest one for each pass; ``tune_mfma`` does the
same for the MFMA schedules.  ``tune_layout`` rebuilds
the handle for each candidate work-item size (a full re-layout and upload, so
only worth it for matrices that are applied many times) and returns the best.
Results are cached per matrix signature in ``~/.cache/bigwham_dev_amd/tune.json``
(``BIGWHAM_TUNE_CACHE`` overrides the path) so a later run of the same problem
skips the search.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

UNROLLS = (4, 8, 16)
MFMA_SCHEDULES = (0, 1, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15)   # 2 spills registers on gfx950
MFMA64_SCHEDULES = (0, 4, 5, 6, 7, 8, 9, 10, 11)       # 1-3 are 0 for FP64
WI_BYTES = (64 << 10, 96 << 10, 192 << 10, 384 << 10)


@dataclass
class TuneResult:
    unroll: dict = field(default_factory=dict)      # precision -> [pass-1, pass-2] unroll
    ms: dict = field(default_factory=dict)          # precision -> {unroll: ms}
    wi_bytes: int = 0


def _cache_path() -> str:
    return os.environ.get("BIGWHAM_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "bigwham_dev_amd",
                                       "tune.json"))


def signature(dev) -> str:
    """Key of a matrix for the cache: size, entries and device bytes."""
    return f"n{dev.n_rows}x{dev.n_cols}-e{dev.nb_entries}-b{dev.device_bytes}"


def _load_cache() -> dict:
    try:
        with open(_cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store_cache(key: str, value: dict) -> None:
    path = _cache_path()
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load_cache()
        data[key] = value
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
    except OSError:
        pass


def time_matvec(dev, x, y, precision: str, reps: int, stats: bool = False):
    """Average device time (ms) of one matvec over ``reps`` calls (with
    ``stats``, the dict of per-kernel averages)."""
    import torch
    dev.matvec(x, original=True, precision=precision, out=y)
    dev.set_timing(True)
    for _ in range(reps):
        dev.matvec(x, original=True, precision=precision, out=y)
    torch.cuda.synchronize()
    st = dev.timing_stats()
    dev.set_timing(False)
    return st if stats else float(st["total"])


def tune_unroll(dev, precisions=None, reps: int = 5, use_cache: bool = True) -> TuneResult:
    """Pick the fastest unroll depth of each pass, per precision, from one
    timing run per depth (the engine's per-kernel timers), and apply them to
    ``dev``.  ``unroll[p]`` is [pass 1, pass 2]; ``ms[p][u]`` the whole
    matvec at depth u in both passes."""
    import torch
    precisions = tuple(precisions or dev.precisions)
    key = "unroll2:" + signature(dev)
    res = TuneResult()
    cached = _load_cache().get(key) if use_cache else None
    if cached and all(p in cached for p in precisions):
        res.unroll = {p: [int(u) for u in cached[p]] for p in precisions}
    else:
        x = torch.ones(dev.n_cols, dtype=torch.float64, device=dev.torch_device)
        y = torch.empty(dev.n_rows, dtype=torch.float64, device=dev.torch_device)
        for p in precisions:
            res.ms[p] = {}
            per_pass = ({}, {})
            for u in UNROLLS:
                dev.set_unroll(**{p: u})
                st = time_matvec(dev, x, y, p, reps, stats=True)
                res.ms[p][u] = float(st["total"])
                per_pass[0][u], per_pass[1][u] = float(st["p1_slab"]), float(st["p2_slab"])
            res.unroll[p] = [min(d, key=d.get) for d in per_pass]
        if use_cache:
            _store_cache(key, dict(res.unroll))
    for p, (u1, u2) in res.unroll.items():
        dev.set_unroll_pass(p, 1, u1)
        dev.set_unroll_pass(p, 2, u2)
    return res


def tune_mfma(dev, nrhs: int = 32, reps: int = 5, precision: str = "fp32") -> dict:
    """Time the schedules of the batched MFMA kernel of one precision (group
    pairs per step, depth of the load pipeline) on `nrhs` right-hand sides
    and keep the fastest.  Returns {schedule: ms per batched call, "best":
    schedule}."""
    import time

    import torch
    X = torch.ones((dev.n_cols, nrhs), dtype=torch.float64, device=dev.torch_device)
    Y = torch.empty((dev.n_rows, nrhs), dtype=torch.float64, device=dev.torch_device)
    ms = {}
    for sch in (MFMA_SCHEDULES if precision == "fp32" else MFMA64_SCHEDULES):
        dev.set_mfma_schedule(sch, precision)
        dev.matvec_batched(X, precision=precision, out=Y)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            dev.matvec_batched(X, precision=precision, out=Y)
        torch.cuda.synchronize()
        ms[sch] = (time.perf_counter() - t0) / reps * 1e3
    best = min(ms, key=ms.get)
    dev.set_mfma_schedule(best, precision)
    return {**ms, "best": best}


def tune_layout(make_dev, precisions=("fp64", "fp32"), candidates=WI_BYTES, reps: int = 5):
    """Build the handle once per work-item size (``make_dev()`` after
    ``set_tuning``), keep the fastest; returns (dev, TuneResult)."""
    import torch
    from .hmat import set_tuning
    best = None
    for wi in candidates:
        set_tuning(wi, wi)
        dev = make_dev()
        r = tune_unroll(dev, precisions, reps, use_cache=False)
        x = torch.ones(dev.n_cols, dtype=torch.float64, device=dev.torch_device)
        y = torch.empty(dev.n_rows, dtype=torch.float64, device=dev.torch_device)
        score = sum(time_matvec(dev, x, y, p, reps) for p in precisions)
        if best is None or score < best[0]:
            if best is not None:
                best[1].close()
            r.wi_bytes = wi
            best = (score, dev, r)
        else:
            dev.close()
    set_tuning(0, 0)
    return best[1], best[2]
