"""Benchmark: batched DTW over 100,000 random-walk pairs on one MI355X.

Workload (the six-measure 100k-pair configuration): 100,000 pairs of
independent seeded random-walk trajectories (seed 0), lengths uniform on
20..200, FP64 coordinates resident in HBM (plus the FP32 copy).  One step =
one call of the engine for DTW over all 100,000 pairs (planner + persistent
DP kernel), inputs already on the device.  `value` is pairs/s over the timed
steps.  The other five measures are timed the same way and reported under
"measures".

roofline: the DP kernel is bound by VALU FP32 issue.  Algorithmic FLOP per DP
cell (FMA = 2): DTW 9, Frechet 9, LCSS 9, EDR 10, ERP 11; Hausdorff 24 per
(point, segment) evaluation; STEDR 13 and STLCSS 12 (the EDR / LCSS cell plus
the time test: subtract, compare, and).  The spatio-temporal measures use
timestamps 15 +- 5 s apart from a start uniform in [0, 3600) s, delta = 60 s;
CDDS (eps 0.0025) reports segment pairs per second and no FLOP roofline (it is
FP64 and nearly all segment pairs stop at the time-overlap test).  SAR
(distance_eps 0.0025, time_eps 10 s, target_length 10, threads_per_traj 64:
the reference's defaults but for distance_eps, config.py:55-57) reports
points per second and the HBM fraction of its input bytes (timestamp and
FP64 x, y: 24 B per point) over the kernel time.
achieved = FLOP per launch / mean kernel time
(HIP events on the launch stream); peak = 157.3 TFLOP/s (MI355X vector FP32).

cpu_baseline: the FP64 full-matrix restatement of the DP (oracle/), one pair
per task, OpenMP over all host cores, on the same pairs.

knn: the learned-filter leg.  1,000 queries over 1,000,000 database random
walks (lognormal lengths, median 60, clipped to [10, 1000], seed 2024), k =
10.  Exact kNN = DTW of every query against every database trajectory (the
DP kernel over the grid) plus a per-row top-k.  The filter is trained in the
run (1,000 random anchors x their 200 nearest by the untrained embedding =
200,000 DTW labels from the DP kernel, 15 epochs of Adam through the
encoder's HIP forward/backward kernels); filtered kNN = embed the queries,
BF16-MFMA distance GEMM against the embedded database, 200 candidates per
query, exact DTW re-rank, top 10.  Recall@10 is judged against the exact
kNN; q/s excludes the one-off database embedding (reported as index_build_s).

cluster: the clustering leg.  10,000 random walks (lengths 98 % uniform
10..500, 2 % uniform 1000..2000), DTW over the upper triangle (49,995,000
pairs), the symmetric matrix expanded on the device, FasterPAM k-medoids
with k = 10; the oracle's FasterPAM runs on one host core beside it and must
give the same result.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3
FLOP_PER_CELL = {"dtw": 9, "frechet": 9, "lcss": 9, "edr": 10, "erp": 11, "hausdorff": 24, "stedr": 13, "stlcss": 12}
EPS = 0.0025
DELTA = 60.0
SAR = dict(time_step=10, target_length=10, replicas=64)
PEAK_HBM_GBS = 8000.0


def cells_of(measure, la, lb):
    la = la.astype(np.float64)
    lb = lb.astype(np.float64)
    if measure == "hausdorff":  # point-segment evaluations, both directions
        return float(np.sum(la * np.maximum(lb - 1, 0) + lb * np.maximum(la - 1, 0)))
    if measure == "cdds":  # segment pairs
        return float(np.sum(np.maximum(la - 1, 0) * np.maximum(lb - 1, 0)))
    if measure == "sar":  # points (the work is linear in them)
        return float(np.sum(la + lb))
    return float(np.sum(la * lb))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=100_000)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--measures", default="dtw,edr,lcss,erp,frechet,hausdorff,stedr,stlcss,cdds,sar",
                    help="measures reported under 'measures' (dtw is always the headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work at most")
    ap.add_argument("--no-knn", action="store_true", help="skip the learned kNN filter leg")
    ap.add_argument("--knn-db", type=int, default=1_000_000)
    ap.add_argument("--knn-queries", type=int, default=1000)
    ap.add_argument("--no-cluster", action="store_true", help="skip the all-pairs + k-medoids leg")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the DTW rates from host buffers (padded drop-in and packed upload, PCIe included)")
    ap.add_argument("--cluster-n", type=int, default=10_000)
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run one plain 256 MiB device copy (calibrates HBM byte counters under rocprofv3)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU: pairs would shard across GPUs by contiguous "
                                   "ranges of the work list, which is not built, so --gpus must be 1"}))
        sys.exit(2)

    import torch

    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import lengths_uniform, random_walks, with_timestamps

    t0 = time.time()
    rng = np.random.default_rng(0)
    ts = random_walks(rng, lengths_uniform(rng, 2 * args.pairs, 20, 200))
    ts = with_timestamps(np.random.default_rng(1), ts)  # for STEDR / STLCSS / CDDS / SAR only
    pa = np.arange(0, 2 * args.pairs, 2)
    pb = pa + 1
    gpoint = ts.xy.mean(0)
    gen_s = time.time() - t0
    print(f"[bench] generated {args.pairs} pairs ({ts.xy.shape[0]} points) in {gen_s:.1f}s", file=sys.stderr,
          flush=True)

    dev = torch.device("cuda:0")
    if args.copy_probe:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    dset = engine.DeviceTrajSet(ts, device=dev)
    pa_d = torch.as_tensor(pa, dtype=torch.int32, device=dev)
    pb_d = torch.as_tensor(pb, dtype=torch.int32, device=dev)
    out = torch.empty(args.pairs, dtype=torch.float64, device=dev)
    la, lb = ts.lengths[pa], ts.lengths[pb]

    def step(measure):
        if measure == "sar":
            engine.pairs(measure, dset, pa_d, pb_d, eps=EPS, validate=False, out=out, **SAR)
            return
        engine.pairs(measure, dset, pa_d, pb_d, eps=EPS, gpoint=gpoint, lanes=args.lanes, validate=False, out=out,
                     delta=DELTA)

    def timed(measure, steps, warmup):
        for _ in range(warmup):
            step(measure)
        torch.cuda.synchronize()
        with engine.KernelTimer() as kt:
            t = time.perf_counter()
            for _ in range(steps):
                step(measure)
            torch.cuda.synchronize()
            el = time.perf_counter() - t
        return el, kt.ms / max(kt.launches, 1), kt.launches

    measures = [m for m in args.measures.split(",") if m]
    report = {}
    for m in measures:
        steps = args.steps if m == "dtw" else max(3, args.steps // 4)
        el, kms, nl = timed(m, steps, args.warmup)
        cells = cells_of(m, la, lb)
        report[m] = {
            "pairs_per_s": args.pairs * steps / el,
            "ms_per_step": 1e3 * el / steps,
            "kernel_ms": kms,
            "cells_per_s": cells / (kms * 1e-3),
        }
        if m in FLOP_PER_CELL:
            report[m]["roofline_frac"] = cells * FLOP_PER_CELL[m] / (kms * 1e-3) / (PEAK_FP32_TFLOPS * 1e12)
        elif m == "sar":  # linear in the points: HBM fraction of the input bytes
            report[m]["unit"] = "points/s"
            report[m]["hbm_frac"] = 24.0 * cells / (kms * 1e-3) / (PEAK_HBM_GBS * 1e9)
        else:  # CDDS: FP64, mostly a time-overlap test per segment pair; no FLOP roofline
            report[m]["unit"] = "segment pairs/s"
        print(f"[bench] {m}: {report[m]}", file=sys.stderr, flush=True)

    el, kms, nl = timed("dtw", args.steps, args.warmup)
    cells = cells_of("dtw", la, lb)
    achieved = cells * FLOP_PER_CELL["dtw"] / (kms * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        from oracle import oracle as O
        threads = O.max_threads()
        n = 2000
        while True:
            t = time.perf_counter()
            O.pairs("dtw", ts.xy, ts.offsets, pa[:n], pb[:n], threads=threads)
            dt = time.perf_counter() - t
            if n >= args.pairs or dt * 2 * (1 + (n < args.pairs)) > args.cpu_budget or dt > args.cpu_budget / 4:
                break
            n = min(args.pairs, n * 4)
        cpu = {"value": n / dt, "unit": "pairs/s", "cores": threads, "kind": "port",
               "sample": f"first {n} of the {args.pairs} pairs, DTW, FP64 full-matrix restatement of the "
                         f"traj_dist recurrence (oracle/tsm_oracle.c, gcc -O2), one pair per OpenMP task, "
                         f"{threads} threads on {cpu_model()}; {n / dt * cells_of('dtw', la[:n], lb[:n]) / n:.3e} cells/s",
               "seconds": dt}
        print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)

    traffic = None
    try:  # HBM bytes per DTW launch from the committed rocprofv3 counter runs
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            traffic = json.load(f)["dp_kernel_dtw_fp32_100k_pairs"]["traffic_bytes"]
    except (OSError, KeyError, ValueError):
        traffic = None
    if args.pairs != 100_000:
        traffic = None

    host = host_leg(ts, pa, pb, dev) if "dtw" in measures and not args.no_host else None
    knn = None if args.no_knn else knn_leg(args.knn_queries, args.knn_db)
    clu = None if args.no_cluster else cluster_leg(args.cluster_n, cpu_too=not args.no_cpu)

    line = {
        "metric": "DTW pairs/s & cells/s (% VALU roofline); kNN q/s at recall@10\u22650.95 vs exact",
        "value": args.pairs * args.steps / el,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * el / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic seeded random walks (seed 0), FP64 coordinates resident in HBM",
        "config": {"workload": "100,000 random-walk pairs, lengths uniform 20..200, DTW (six-measure config)",
                   "pairs": args.pairs, "cells": cells, "lanes": args.lanes or "auto"},
        "cells_per_s": cells / (kms * 1e-3),
        "kernel_ms": kms,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic,
                     "note": "9 algorithmic FP32 FLOP per DTW cell over mean dp_kernel duration (HIP events)"},
        "cpu_baseline": cpu,
        "measures": report,
        "host_buffers": host,
        "knn": knn,
        "cluster": clu,
    }
    print(json.dumps(line))


def host_leg(ts, pa, pb, dev, reps: int = 3) -> dict:
    """The same DTW pairs handed over in host memory, timed from the host
    buffers to the results back in host memory (PCIe included), best of
    `reps` after one warm-up:
    - padded: the reference's boundary, core.dtw(t1, len1, t2, len2) on
      zero-padded float64 [pairs, max length, 2] arrays (core/dtw.py:56-77);
    - packed: the set uploaded once as one point buffer
      (engine.DeviceTrajSet: FP64 copy + centred FP32 copy) and the pairs
      given as index arrays, as drivers.pairwise does."""
    import torch

    from trajsimimeasures_amd import core, engine

    la, lb = ts.lengths[pa], ts.lengths[pb]
    width = int(max(la.max(), lb.max()))

    def padded(idx, lens):
        arr = np.zeros((len(idx), width, 2), dtype=np.float64)
        rows = np.repeat(np.arange(len(idx)), lens)
        cols = np.arange(int(lens.sum())) - np.repeat(np.cumsum(lens) - lens, lens)
        arr[rows, cols] = ts.xy[ts.offsets[idx][rows] + cols]
        return arr

    t1, t2 = padded(pa, la), padded(pb, lb)

    def best(fn):
        fn()
        ts_ = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            fn()
            ts_.append(time.perf_counter() - t)
        return min(ts_)

    t_pad = best(lambda: core.dtw(t1, la, t2, lb))
    t_pack = best(lambda: engine.pairs("dtw", engine.DeviceTrajSet(ts, device=dev), pa, pb).cpu().numpy())
    n = len(pa)
    out = {"padded_pairs_per_s": n / t_pad, "padded_s": t_pad, "padded_bytes": t1.nbytes + t2.nbytes,
           "packed_pairs_per_s": n / t_pack, "packed_s": t_pack, "packed_bytes": ts.xy.nbytes + ts.offsets.nbytes}
    print(f"[bench] host buffers (PCIe included): {out}", file=sys.stderr, flush=True)
    return out


def cluster_leg(n: int, k: int = 10, cpu_too: bool = True) -> dict:
    """The clustering workload: DTW over the upper triangle of n random walks
    (lengths 98 % uniform 10..500, 2 % uniform 1000..2000, seed 0), the
    symmetric matrix expanded on the device, FasterPAM with k medoids from
    RandomState(0) initial medoids.  The oracle's FasterPAM (one core, same
    matrix and initial medoids) is timed beside it and must give the same
    loss, medoids and labels."""
    import torch

    from trajsimimeasures_amd import cluster, engine
    from trajsimimeasures_amd.data import lengths_uniform_with_tail, random_walks

    def log(msg):
        print(f"[bench] cluster: {msg}", file=sys.stderr, flush=True)

    rng = np.random.default_rng(0)
    ts = random_walks(rng, lengths_uniform_with_tail(rng, n))
    dset = engine.DeviceTrajSet(ts)
    torch.cuda.synchronize()
    t = time.perf_counter()
    tri = engine.triu("dtw", dset, n)
    D = cluster.square_from_triu(tri, n)
    del tri
    torch.cuda.synchronize()
    t_mat = time.perf_counter() - t
    log(f"all-pairs DTW of {n} trajectories + symmetric matrix: {t_mat:.3f}s")
    med0 = np.random.RandomState(0).choice(n, k, False)
    cluster.fasterpam(D, med0)  # warm-up (module loads, workspace)
    torch.cuda.synchronize()
    t = time.perf_counter()
    res = cluster.fasterpam(D, med0)
    torch.cuda.synchronize()
    t_gpu = time.perf_counter() - t
    log(f"FasterPAM k={k}: {t_gpu * 1e3:.2f} ms, {res.n_iter} passes, {res.n_swap} swaps, loss {res.loss:.6f}")
    out = {"workload": f"{n} random walks (lengths 98% U(10..500), 2% U(1000..2000), seed 0), DTW upper "
                       f"triangle ({n * (n - 1) // 2} pairs), symmetric matrix on the device, FasterPAM k={k} "
                       f"from RandomState(0) initial medoids",
           "matrix_s": t_mat, "fasterpam_s": t_gpu, "n_iter": res.n_iter, "n_swap": res.n_swap, "loss": res.loss}
    if cpu_too:
        from oracle import oracle as O
        Dh = D.cpu().numpy()
        del D
        t = time.perf_counter()
        ref = O.fasterpam(Dh, med0)
        t_cpu = time.perf_counter() - t
        same = ref[0] == res.loss and np.array_equal(ref[2], res.medoids) and np.array_equal(ref[1], res.labels)
        log(f"oracle FasterPAM (1 core): {t_cpu:.2f}s, same result: {same}")
        out.update({"cpu_fasterpam_s": t_cpu, "cpu_cores": 1, "same_as_oracle": bool(same),
                    "speedup": t_cpu / t_gpu})
    return out


def knn_leg(nq: int, nd: int) -> dict:
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.engine import DeviceTrajSet

    def log(msg):
        print(f"[bench] knn: {msg}", file=sys.stderr, flush=True)

    t = time.time()
    ts = F.knn_workload(nq=nq, nd=nd)
    dset = DeviceTrajSet(ts)
    torch.cuda.synchronize()
    log(f"generated and uploaded {nq} + {nd} trajectories in {time.time() - t:.1f}s")
    truth, t_exact = F.exact_knn(dset, 0, nq, nq, nd, k=10)
    log(f"exact kNN {t_exact:.3f}s ({nq / t_exact:.1f} q/s)")
    db = torch.arange(nq, nq + nd, dtype=torch.int32, device="cuda")
    base = F.FilterIndex(F.Encoder(), dset, db)
    found0, _ = F.filtered_knn(base, 0, nq, k=10, C=200)
    r0 = F.recall_at(found0, truth)
    del base
    enc, info = F.train_filter(dset, db)
    log(f"trained: {info}")
    torch.cuda.synchronize()
    t = time.perf_counter()
    index = F.FilterIndex(enc, dset, db)
    torch.cuda.synchronize()
    t_index = time.perf_counter() - t
    F.filtered_knn(index, 0, nq, k=10, C=200)  # warm-up
    runs = [F.filtered_knn(index, 0, nq, k=10, C=200) for _ in range(3)]
    t_f = min(r[1] for r in runs)
    r1 = F.recall_at(runs[-1][0], truth)
    out = {"workload": f"{nq} queries x {nd} database random walks, lognormal lengths (median 60, sigma 0.8, "
                       f"[10, 1000]), seed 2024; k=10, 200 candidates, exact DTW re-rank",
           "exact_qps": nq / t_exact, "exact_s": t_exact, "filtered_qps": nq / t_f, "filtered_s": t_f,
           "speedup": t_exact / t_f, "recall_at_10": r1, "recall_at_10_untrained": r0,
           "index_build_s": t_index, "train": info}
    log(str(out))
    return out


if __name__ == "__main__":
    main()
