"""Autotuner for the DP kernel's launch parameters, measured on the device.

For each column-count bucket (the shorter trajectory's length, in steps of
8) it times the kernel with 4 and with 8 lanes per pair on synthetic
random-walk pairs of that shape and keeps the faster; each bucket's batch is
large enough to fill the GPU at either width (a batch too small for 16 pairs
per task makes 4 lanes look slow for lack of tasks, not per pair).  Then it
times the persistent kernel at several resident-wave counts per CU on the
mixed 100k-pair workload, and finally keeps the measured table only if it
beats the built-in cost model on that workload.  W (register columns per lane) follows from the lanes choice as
ceil(ncol / L); the DP kernel uses no LDS, so there is no LDS footprint to
tune.  Results are applied through tsm_tune_set_lanes /
tsm_tune_set_waves_per_cu and can be saved as JSON; the package ships the
table measured on MI355X (tuning_gfx950.json) and applies it at import.
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
    nrow = np.maximum(ncol, rng.integers(lo, rows_hi + 1, npairs))
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


def tune(npairs_per_bucket: int = 65536, seed: int = 0, measures=(("dtw", 32), ("edr", 64)),
         waves=(4, 8, 12, 16), verbose: bool = True) -> dict:
    """Measure and apply a tuning table; returns it."""
    import torch

    from . import engine
    rng = np.random.default_rng(seed)
    table = {"gpu": torch.cuda.get_device_name(0), "bucket": BUCKET}
    for measure, bits in measures:
        wm = WMAX[bits]
        lanes = [8] * 257
        limit4 = 4 * wm  # beyond this only L = 8 fits in registers
        for lo in range(1, limit4 + 1, BUCKET):
            hi = min(lo + BUCKET - 1, limit4)
            ts, pa, pb = _pairs_with_ncol(rng, lo, hi, npairs_per_bucket)
            d = engine.DeviceTrajSet(ts)
            pa_d = torch.as_tensor(pa, device="cuda")
            pb_d = torch.as_tensor(pb, device="cuda")
            t4 = _kernel_ms(measure, d, pa_d, pb_d, lanes=4, eps=0.0025)
            t8 = _kernel_ms(measure, d, pa_d, pb_d, lanes=8, eps=0.0025)
            best = 4 if t4 <= t8 else 8
            for nc in range(lo, hi + 1):
                lanes[nc] = best
            if verbose:
                print(f"[tune] {measure} ncol {lo:3d}-{hi:3d}: L=4 {t4:.4f} ms, L=8 {t8:.4f} ms -> {best}", flush=True)
        table[f"lanes_fp{bits}"] = lanes
    apply(table)
    # resident waves per CU on the mixed 100k-pair shape
    from .data import lengths_uniform, random_walks
    ts = random_walks(rng, lengths_uniform(rng, 2 * 100_000, 20, 200))
    d = engine.DeviceTrajSet(ts)
    pa = torch.arange(0, 200_000, 2, dtype=torch.int32, device="cuda")
    pb = pa + 1
    best_w, best_t = None, None
    for w in waves:
        _lib.check(_lib.lib().tsm_tune_set_waves_per_cu(int(w)))
        t = _kernel_ms("dtw", d, pa, pb, reps=5)
        if verbose:
            print(f"[tune] waves per CU {w}: {t:.4f} ms", flush=True)
        if best_t is None or t < best_t:
            best_w, best_t = w, t
    table["waves_per_cu"] = int(best_w)
    apply(table)
    # keep the measured lanes tables only where they beat the cost model
    L = _lib.lib()
    for measure, bits in measures:
        t_tuned = _kernel_ms(measure, d, pa, pb, reps=5, eps=0.0025)
        _lib.check(L.tsm_tune_set_lanes(bits, None, 0))
        t_model = _kernel_ms(measure, d, pa, pb, reps=5, eps=0.0025)
        if verbose:
            print(f"[tune] {measure}: measured table {t_tuned:.4f} ms, cost model {t_model:.4f} ms", flush=True)
        if t_model <= t_tuned:
            table.pop(f"lanes_fp{bits}")
        else:
            apply({f"lanes_fp{bits}": table[f"lanes_fp{bits}"]})
    return table


def save(table: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(table, f)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=None, help="write the measured table here (JSON)")
    ap.add_argument("--pairs", type=int, default=65536)
    args = ap.parse_args()
    t = tune(npairs_per_bucket=args.pairs)
    if args.out:
        save(t, args.out)
    print(json.dumps({k: v for k, v in t.items() if not k.startswith("lanes")}))
