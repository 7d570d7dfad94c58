"""Autotuner for the DP kernel's launch shape, measured on the device.

The DP kernel runs a pair on a segment of L lanes (L = 2, 4, 8 or 16), each
lane holding W = ceil(ncol / L) columns of the shorter trajectory in
registers (at most 24, 12 with two lanes), so choosing L per column count
chooses the tile width.  For each column-count bucket (BUCKET columns) the
tuner times DTW with every L that holds the bucket's columns, on synthetic
random-walk pairs of that shape (rows from the bucket up to 200), keeps the
fastest, and records by how much the built-in cost model's choice there
(model_lanes, the same rule as the library's) wins or loses.  It does so for
the FP32-coordinate kernels and for the FP64 ones (precision="fp64"), which
keep separate tables.  Then it times the persistent kernel at several
resident-wave counts per CU, and finally keeps a measured lanes table only
if it beats the cost model on whole workloads: the 100,000-pair batch with
lengths uniform on 20..200 and one with lognormal lengths (median 60).
Last, it times batches of 2,048 to 12,288 pairs (a caller's batch loop,
planner included) as direct launches (one wavefront per pair, no planner)
and as planned ones, and keeps the largest size at which direct wins.  The
DP kernel uses no LDS, so there is no LDS footprint to tune.  Tables go to
tsm_tune_set_lanes / tsm_tune_set_waves_per_cu / tsm_tune_set_direct_pairs
and can be saved as JSON; the package applies tuning_gfx950.json at import
when it is present.
Reference counterpart: the fixed 64 threads per pair of config.py:45 /
method_helper.py:69-70.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib

PKG = os.path.dirname(os.path.abspath(__file__))
SHIPPED = os.path.join(PKG, "tuning_gfx950.json")
WMAX = {32: 24, 64: 24}  # register columns per lane of the FP32 / FP64-coordinate kernels
BUCKET = 8
LANES = (2, 4, 8, 16)
NCOL_MAX = 256  # column counts the lanes table covers


def wmax(wm: int, lanes: int) -> int:
    """Largest W of an L-lane segment."""
    return wm // 2 if lanes == 2 else wm


def fits(wm: int, lanes: int, ncol: int) -> bool:
    return lanes in LANES and -(-ncol // lanes) <= wmax(wm, lanes)


def model_lanes(wm: int, ncol: int) -> int:
    """The library's cost model (tsm_launch.hpp model_lanes): per pair
    L * (rows + L - 1) * (1 + W) issue slots with rows = max(ncol, 32), the
    cheapest L that holds the columns, ties to fewer lanes."""
    rows = max(ncol, 32)
    best, cost = 16, 0.0
    for L in (16, 8, 4, 2):
        if not fits(wm, L, ncol):
            continue
        c = L * (rows + L - 1) * (1 + -(-ncol // L))
        if L == 16 or c <= cost:
            best, cost = L, c
    return best


def apply(table: dict) -> None:
    """Set the lanes tables and the waves-per-CU count from a tuning dict."""
    L = _lib.lib()
    for bits in (32, 64):
        lanes = table.get(f"lanes_fp{bits}")
        if lanes:
            arr = (ctypes.c_uint8 * 257)()
            for nc in range(1, 257):
                arr[nc] = int(lanes[min(nc, len(lanes) - 1)])
            _lib.check(L.tsm_tune_set_lanes(bits, arr, 257))
    if table.get("waves_per_cu"):
        _lib.check(L.tsm_tune_set_waves_per_cu(int(table["waves_per_cu"])))
    if table.get("direct_pairs") is not None:
        _lib.check(L.tsm_tune_set_direct_pairs(int(table["direct_pairs"])))


def load_shipped() -> dict | None:
    if not os.path.exists(SHIPPED):
        return None
    with open(SHIPPED) as f:
        table = json.load(f)
    apply(table)
    return table


def _pairs_with_ncol(rng, lo, hi, npairs, rows_hi=200):
    from .data import random_walks
    ncol = rng.integers(lo, hi + 1, npairs)
    nrow = np.maximum(ncol, rng.integers(lo, max(rows_hi, hi) + 1, npairs))
    lens = np.empty(2 * npairs, dtype=np.int64)
    lens[0::2] = nrow
    lens[1::2] = ncol
    ts = random_walks(rng, lens)
    pa = np.arange(0, 2 * npairs, 2, dtype=np.int32)
    return ts, pa, pa + 1


def _kernel_ms(measure, dset, pa, pb, reps=3, **kw):
    import torch

    from . import engine
    out = torch.empty(pa.numel(), dtype=torch.float64, device=pa.device)
    engine.pairs(measure, dset, pa, pb, validate=False, out=out, **kw)
    torch.cuda.synchronize()
    with engine.KernelTimer() as kt:
        for _ in range(reps):
            engine.pairs(measure, dset, pa, pb, validate=False, out=out, **kw)
        torch.cuda.synchronize()
    return kt.ms / max(kt.launches, 1)


def _workloads(rng, npairs):
    """The whole-workload checks: (name, set, pa, pb) on the device."""
    import torch

    from . import engine
    from .data import lengths_lognormal, lengths_uniform, random_walks
    out = []
    for name, lens in (("uniform 20..200", lengths_uniform(rng, 2 * npairs, 20, 200)),
                       ("lognormal median 60", lengths_lognormal(rng, 2 * npairs))):
        d = engine.DeviceTrajSet(random_walks(rng, lens))
        pa = torch.arange(0, 2 * npairs, 2, dtype=torch.int32, device="cuda")
        out.append((name, d, pa, pa + 1))
    return out


def tune(npairs_per_bucket: int = 65536, seed: int = 0, precisions=("fp32", "fp64"), waves=(4, 8, 12, 16),
         ncol_max: int = NCOL_MAX, workload_pairs: int = 100_000, verbose: bool = True,
         direct_sizes=(2048, 4096, 6144, 8192, 12288)) -> dict:
    """Measure and apply a tuning table; returns it (with the per-bucket
    timings under "buckets", the whole-workload comparison under
    "workloads", and the batch sizes up to which direct launches win under
    "direct_pairs" / "direct")."""
    import torch

    rng = np.random.default_rng(seed)
    L = _lib.lib()
    table = {"gpu": torch.cuda.get_device_name(0), "bucket": BUCKET, "measure": "dtw", "buckets": {},
             "workloads": {}}
    measured = {}
    # every timing below is of planned launches (a small call would
    # otherwise be a direct one, whatever its lanes table)
    _lib.check(L.tsm_tune_set_direct_pairs(0))
    try:
        _tune_planned(table, measured, rng, L, precisions, waves, npairs_per_bucket, ncol_max, workload_pairs,
                      verbose)
        table["direct_pairs"], table["direct"] = _direct_crossover(rng, direct_sizes, verbose)
    finally:
        _lib.check(L.tsm_tune_set_direct_pairs(int(table.get("direct_pairs", -1))))
    return table


def _call_ms(measure, dset, pa, pb, nb, reps=3):
    """Mean time of one call over nb consecutive batches of pa.numel() // nb
    pairs (planner and DP kernels, as a caller's batch loop runs them), best
    of reps, from events on the current stream."""
    import torch

    from . import engine
    B = pa.numel() // nb
    out = torch.empty(nb * B, dtype=torch.float64, device=pa.device)

    def run():
        for b in range(nb):
            sl = slice(b * B, (b + 1) * B)
            engine.pairs(measure, dset, pa[sl], pb[sl], validate=False, out=out[sl])

    run()
    best = None
    for _ in range(reps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run()
        e1.record()
        e1.synchronize()
        t = e0.elapsed_time(e1) / nb
        best = t if best is None else min(best, t)
    return best


def _direct_crossover(rng, sizes, verbose):
    """Largest batch size at which a direct launch beats the planned one on
    20..200-point DTW pairs (0 if none does), with the timings per size."""
    import torch

    from . import engine
    from .data import lengths_uniform, random_walks
    L = _lib.lib()
    n = max(sizes) * 8
    d = engine.DeviceTrajSet(random_walks(rng, lengths_uniform(rng, 2 * n, 20, 200)))
    pa = torch.arange(0, 2 * n, 2, dtype=torch.int32, device="cuda")
    best, rows = 0, {}
    for N in sizes:
        nb = n // N
        _lib.check(L.tsm_tune_set_direct_pairs(N))
        t_d = _call_ms("dtw", d, pa[:nb * N], pa[:nb * N] + 1, nb)
        _lib.check(L.tsm_tune_set_direct_pairs(0))
        t_p = _call_ms("dtw", d, pa[:nb * N], pa[:nb * N] + 1, nb)
        rows[str(N)] = {"direct_ms": round(t_d, 5), "planned_ms": round(t_p, 5)}
        if t_d < t_p:
            best = N
        if verbose:
            print(f"[tune] batch of {N} pairs: direct {t_d:.4f} ms, planned {t_p:.4f} ms", flush=True)
    return best, rows


def _tune_planned(table, measured, rng, L, precisions, waves, npairs_per_bucket, ncol_max, workload_pairs, verbose):
    import torch

    from . import engine
    for prec in precisions:
        bits = 32 if prec == "fp32" else 64
        wm = WMAX[bits]
        _lib.check(L.tsm_tune_set_lanes(bits, None, 0))
        lanes = [model_lanes(wm, max(nc, 1)) for nc in range(257)]
        rows = []
        for lo in range(1, ncol_max + 1, BUCKET):
            hi = min(lo + BUCKET - 1, ncol_max)
            cands = [c for c in LANES if fits(wm, c, hi)]
            ts, pa, pb = _pairs_with_ncol(rng, lo, hi, npairs_per_bucket)
            d = engine.DeviceTrajSet(ts)
            pa_d = torch.as_tensor(pa, device="cuda")
            pb_d = torch.as_tensor(pb, device="cuda")
            ms = {c: _kernel_ms("dtw", d, pa_d, pb_d, lanes=c, precision=prec) for c in cands}
            best = min(ms, key=ms.get)
            model = model_lanes(wm, hi)
            for nc in range(lo, hi + 1):
                lanes[nc] = best
            rows.append({"ncol": [lo, hi], "ms": {str(c): round(t, 5) for c, t in ms.items()}, "best": best,
                         "model": model, "model_over_best": round(ms[model] / ms[best] - 1.0, 4)})
            if verbose:
                print(f"[tune] {prec} ncol {lo:3d}-{hi:3d}: " + ", ".join(f"L={c} {t:.4f} ms" for c, t in ms.items())
                      + f" -> {best} (model {model}, +{100 * (ms[model] / ms[best] - 1):.1f} %)", flush=True)
        table["buckets"][prec] = rows
        measured[bits] = lanes
    # resident waves per CU on the 100k-pair uniform batch (cost-model tables)
    from .data import lengths_uniform, random_walks
    d = engine.DeviceTrajSet(random_walks(rng, lengths_uniform(rng, 2 * 100_000, 20, 200)))
    pa = torch.arange(0, 200_000, 2, dtype=torch.int32, device="cuda")
    best_w, best_t = None, None
    for w in waves:
        _lib.check(L.tsm_tune_set_waves_per_cu(int(w)))
        t = _kernel_ms("dtw", d, pa, pa + 1, reps=5)
        if verbose:
            print(f"[tune] waves per CU {w}: {t:.4f} ms", flush=True)
        if best_t is None or t < best_t:
            best_w, best_t = w, t
    table["waves_per_cu"] = int(best_w)
    _lib.check(L.tsm_tune_set_waves_per_cu(int(best_w)))
    # keep a measured lanes table only where it beats the cost model on the workloads
    loads = _workloads(rng, workload_pairs)
    for prec in precisions:
        bits = 32 if prec == "fp32" else 64
        arr = (ctypes.c_uint8 * 257)(*[int(x) for x in measured[bits]])
        res = {}
        for name, dd, pa_w, pb_w in loads:
            _lib.check(L.tsm_tune_set_lanes(bits, arr, 257))
            t_tuned = _kernel_ms("dtw", dd, pa_w, pb_w, reps=5, precision=prec)
            _lib.check(L.tsm_tune_set_lanes(bits, None, 0))
            t_model = _kernel_ms("dtw", dd, pa_w, pb_w, reps=5, precision=prec)
            res[name] = {"tuned_ms": round(t_tuned, 5), "model_ms": round(t_model, 5)}
        keep = sum(r["tuned_ms"] for r in res.values()) < sum(r["model_ms"] for r in res.values())
        table["workloads"][prec] = {**res, "kept": "measured table" if keep else "cost model"}
        if keep:
            table[f"lanes_fp{bits}"] = measured[bits]
            _lib.check(L.tsm_tune_set_lanes(bits, arr, 257))
        if verbose:
            print(f"[tune] {prec} workloads: {res} -> {table['workloads'][prec]['kept']}", flush=True)


def save(table: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="write the measured table here (JSON)")
    ap.add_argument("--pairs", type=int, default=65536)
    args = ap.parse_args()
    t = tune(npairs_per_bucket=args.pairs)
    if args.out:
        save(t, args.out)
    print(json.dumps({k: v for k, v in t.items() if k in ("gpu", "waves_per_cu", "workloads")}))
