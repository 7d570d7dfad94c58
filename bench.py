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
achieved = FLOP per launch / mean kernel time (HIP events stamped by the
kernel's dispatch on its stream, on every tenth launch of the timed steps:
a timed dispatch costs ~6 us); peak = 157.3 TFLOP/s (MI355X vector FP32).

cpu_baseline: the FP64 full-matrix restatement of the DP (oracle/), one pair
per task, OpenMP over all host cores, on the same pairs: DTW (the headline
baseline), each of the six measures on a prefix of the pairs, and the
1,000-pair CPU configuration (DTW and discrete Frechet) in full.

knn: the learned-filter leg.  1,000 queries over 1,000,000 database random
walks (lognormal lengths, median 60, clipped to [10, 1000], seed 2024), k =
10.  Exact kNN = DTW of every query against every database trajectory (the
DP kernel over the grid) plus a per-row top-k.  The filter is trained in the
run (2,000 random anchors x their 100 nearest by the untrained embedding =
200,000 DTW labels from the DP kernel, 15 epochs of Adam through the
encoder's HIP forward/backward kernels); filtered kNN = embed the queries,
BF16-MFMA distance GEMM against the embedded database, 100 candidates per
query, exact DTW re-rank, top 10.  Recall@10 is judged against the exact
kNN; q/s excludes the one-off database embedding (reported as index_build_s).
Recall and q/s are also reported for 25..200 candidates, untrained, trained,
and trained on 200,000 random-pair labels instead.

cluster: the clustering leg.  10,000 random walks (lengths 98 % uniform
10..500, 2 % uniform 1000..2000), DTW over the upper triangle (49,995,000
pairs), the symmetric matrix expanded on the device, FasterPAM k-medoids
with k = 10; the oracle's FasterPAM runs on one host core beside it and must
give the same result.  The pairs touching a trajectory over 1,000 points are
rerun in FP64 for DTW and ERP, and the largest relative FP32 drift reported.

measures.<m>_fp64: DTW / ERP / Frechet with precision="fp64" (FP64
coordinates and accumulators: the drift-check path), against the 78.6
TFLOP/s FP64 vector peak; STEDR / STLCSS, whose arithmetic is FP64 too, are
quoted against that peak as well.

graph_replay: the headline step captured once into a hipGraph (planner +
persistent kernel) and replayed.  graphs: eager calls against replays for
the 512-pair batches of the reference's drivers and for the 1,000 x 100,000
query x database grid (DTW and EDR), replays checked bit for bit.
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
PEAK_FP64_TFLOPS = 78.6
FLOP_PER_CELL = {"dtw": 9, "frechet": 9, "lcss": 9, "edr": 10, "erp": 11, "hausdorff": 24, "stedr": 13, "stlcss": 12}
FP64_MEASURES = ("stedr", "stlcss")  # their DP arithmetic is FP64: quoted against the FP64 peak
FP64_RERUN = ("dtw", "erp", "frechet")  # precision="fp64": the drift-check path
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


def timer_every(steps):
    """Kernel-timing stride inside a timed region: about ten of its identical
    launches carry HIP events (a timed dispatch costs ~6 us)."""
    return max(1, steps // 10)


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
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--pairs", type=int, default=100_000)
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--measures", default="dtw,edr,lcss,erp,frechet,hausdorff,stedr,stlcss,cdds,sar",
                    help="measures reported under 'measures' (dtw is always the headline)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-budget", type=float, default=20.0, help="seconds of CPU baseline work at most")
    ap.add_argument("--no-knn", action="store_true", help="skip the learned kNN filter leg")
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the graph leg (query x database grid and 512-pair batches, eager vs replayed)")
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

    def step(measure, precision="fp32"):
        if measure == "sar":
            engine.pairs(measure, dset, pa_d, pb_d, eps=EPS, validate=False, out=out, **SAR)
            return
        engine.pairs(measure, dset, pa_d, pb_d, eps=EPS, gpoint=gpoint, lanes=args.lanes, validate=False, out=out,
                     delta=DELTA, precision=precision)

    def timed(measure, steps, warmup):
        for _ in range(warmup):
            step(measure)
        torch.cuda.synchronize()
        with engine.KernelTimer(every=timer_every(steps)) as kt:
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
            peak = PEAK_FP64_TFLOPS if m in FP64_MEASURES else PEAK_FP32_TFLOPS
            report[m]["roofline_frac"] = cells * FLOP_PER_CELL[m] / (kms * 1e-3) / (peak * 1e12)
            report[m]["peak_tflops"] = peak
        elif m == "sar":  # linear in the points: HBM fraction of the input bytes
            report[m]["unit"] = "points/s"
            report[m]["hbm_frac"] = 24.0 * cells / (kms * 1e-3) / (PEAK_HBM_GBS * 1e9)
        else:  # CDDS: FP64, mostly a time-overlap test per segment pair; no FLOP roofline
            report[m]["unit"] = "segment pairs/s"
        print(f"[bench] {m}: {report[m]}", file=sys.stderr, flush=True)
    # the FP64 DP (precision="fp64", the drift-check rerun), against the FP64 peak
    for m in FP64_RERUN:
        if m not in measures:
            continue
        steps = max(3, args.steps // 4)
        for _ in range(args.warmup):
            step(m, "fp64")
        torch.cuda.synchronize()
        with engine.KernelTimer(every=timer_every(steps)) as kt:
            t = time.perf_counter()
            for _ in range(steps):
                step(m, "fp64")
            torch.cuda.synchronize()
            el = time.perf_counter() - t
        kms = kt.ms / max(kt.launches, 1)
        cells = cells_of(m, la, lb)
        report[m + "_fp64"] = {"pairs_per_s": args.pairs * steps / el, "ms_per_step": 1e3 * el / steps,
                               "kernel_ms": kms, "cells_per_s": cells / (kms * 1e-3),
                               "roofline_frac": cells * FLOP_PER_CELL[m] / (kms * 1e-3) / (PEAK_FP64_TFLOPS * 1e12),
                               "peak_tflops": PEAK_FP64_TFLOPS}
        print(f"[bench] {m} fp64: {report[m + '_fp64']}", file=sys.stderr, flush=True)

    el, kms, nl = timed("dtw", args.steps, args.warmup)
    cells = cells_of("dtw", la, lb)
    achieved = cells * FLOP_PER_CELL["dtw"] / (kms * 1e-3) / 1e12

    # the same step captured once into a hipGraph (planner + persistent
    # kernel) and replayed: the launch model for repeated batches
    g = engine.PairsGraph("dtw", dset, pa_d, pb_d, lanes=args.lanes)
    for _ in range(args.warmup):
        g.launch()
    g.stream.synchronize()
    t = time.perf_counter()
    for _ in range(args.steps):
        g.launch()
    g.stream.synchronize()
    el_g = time.perf_counter() - t
    replay_equal = bool(torch.equal(g.out, engine.pairs("dtw", dset, pa_d, pb_d, lanes=args.lanes, validate=False)))
    g.close()
    graph_step = {"pairs_per_s": args.pairs * args.steps / el_g, "ms_per_step": 1e3 * el_g / args.steps,
                  "equal_to_eager": replay_equal}
    print(f"[bench] dtw graph replay: {graph_step}", file=sys.stderr, flush=True)

    cpu = None if args.no_cpu else cpu_leg(ts, pa, pb, gpoint, args.pairs, args.cpu_budget)

    traffic, traffic_src = None, None
    try:  # HBM bytes per DTW launch from the committed rocprofv3 counter runs
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            rec = json.load(f)["dp_kernel_dtw_fp32_100k_pairs"]
        traffic = rec["traffic_bytes"]
        traffic_src = (f"profiles/traffic.json: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of this workload, "
                       f"scaled by a copy kernel's factors; library sha256 {rec.get('library_sha256', '?')[:16]} "
                       f"({'this build' if rec.get('library_sha256') == library_sha256() else 'an earlier build'})")
    except (OSError, KeyError, ValueError):
        traffic = None
    if args.pairs != 100_000:
        traffic, traffic_src = None, None

    host = host_leg(ts, pa, pb, dev) if "dtw" in measures and not args.no_host else None
    graphs = None if args.no_graph else graph_leg(pa_d, pb_d, dset, args.pairs)
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
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "note": "9 algorithmic FP32 FLOP per DTW cell over mean dp_kernel duration (HIP events, every tenth launch of the timed steps)"},
        "graph_replay": graph_step,
        "cpu_baseline": cpu,
        "measures": report,
        "host_buffers": host,
        "graphs": graphs,
        "knn": knn,
        "cluster": clu,
    }
    print(json.dumps(line))


def library_sha256() -> str:
    import hashlib

    from trajsimimeasures_amd import _lib
    with open(_lib.LIB_PATH, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def cpu_leg(ts, pa, pb, gpoint, npairs: int, budget: float) -> dict:
    """The FP64 oracle (the traj_dist restatement, oracle/tsm_oracle.c) on the
    box's host cores, one pair per OpenMP task: DTW on a growing prefix of
    the 100k pairs until about budget/4 seconds (the headline baseline); each
    of the six measures on a prefix sized from DTW's rate to take about
    budget/12 seconds; and the 1,000-pair CPU configuration (its own seeded
    pairs, DTW and discrete Frechet) in full."""
    from oracle import oracle as O
    from trajsimimeasures_amd.data import lengths_uniform, random_walks

    threads = O.max_threads()
    la, lb = ts.lengths[pa], ts.lengths[pb]

    def run(m, a, b, xy=ts.xy, offs=ts.offsets):
        t = time.perf_counter()
        O.pairs(m, xy, offs, a, b, eps=EPS, gpoint=gpoint, threads=threads)
        return time.perf_counter() - t

    n = 2000
    while True:
        dt = run("dtw", pa[:n], pb[:n])
        if n >= npairs or dt * 2 * (1 + (n < npairs)) > budget or dt > budget / 4:
            break
        n = min(npairs, n * 4)
    dtw_rate = n / dt
    cpu = {"value": dtw_rate, "unit": "pairs/s", "cores": threads, "kind": "port",
           "sample": f"first {n} of the {npairs} pairs, DTW, FP64 full-matrix restatement of the traj_dist "
                     f"recurrence (oracle/tsm_oracle.c, gcc -O2), one pair per OpenMP task, {threads} threads on "
                     f"{cpu_model()}; {dtw_rate * cells_of('dtw', la[:n], lb[:n]) / n:.3e} cells/s",
           "seconds": dt, "measures": {}}
    for m in ("dtw", "edr", "lcss", "erp", "frechet", "hausdorff"):
        k = int(min(npairs, max(1000, dtw_rate * budget / 12 / (3 if m == "hausdorff" else 1))))
        dt_m = run(m, pa[:k], pb[:k])
        cpu["measures"][m] = {"pairs": k, "seconds": dt_m, "pairs_per_s": k / dt_m,
                              "cells_per_s": cells_of(m, la[:k], lb[:k]) / dt_m}
    rng = np.random.default_rng(0)  # the 1,000-pair CPU configuration, seed 0
    c1 = random_walks(rng, lengths_uniform(rng, 2000, 20, 200))
    a1 = np.arange(0, 2000, 2)
    cpu["config1"] = {"workload": "1,000 random-walk pairs (seed 0), lengths uniform 20..200, FP64"}
    for m in ("dtw", "frechet"):
        dt_m = run(m, a1, a1 + 1, c1.xy, c1.offsets)
        cpu["config1"][m] = {"seconds": dt_m, "pairs_per_s": 1000 / dt_m}
    print(f"[bench] cpu baseline: {cpu}", file=sys.stderr, flush=True)
    return cpu


def graph_leg(pa_d, pb_d, dset, npairs: int, reps: int = 3) -> dict:
    """Eager calls against hipGraph replays of the same launches:
    - the query x database configuration, 1,000 queries x 100,000 database
      random walks (lognormal lengths, median 60, sigma 0.8, [10, 1000],
      seed 0; 1e8 pairs) through grid addressing, DTW and EDR, best of
      `reps`, replays checked bit for bit against the eager results;
    - the reference's 512-pair batches (effi_batch_size_gpu, config.py:43)
      over the 100k-pair DTW workload: one engine call per batch against
      one captured 512-pair graph whose pair tensors are refilled per batch
      on the graph's stream."""
    import torch

    from trajsimimeasures_amd import engine
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks

    def log(msg):
        print(f"[bench] graphs: {msg}", file=sys.stderr, flush=True)

    res = {}
    # 512-pair batches
    B = 512
    nb = npairs // B
    out = torch.empty(nb * B, dtype=torch.float64, device=pa_d.device)

    def eager():
        for b in range(nb):
            sl = slice(b * B, (b + 1) * B)
            engine.pairs("dtw", dset, pa_d[sl], pb_d[sl], validate=False, out=out[sl])

    g = engine.PairsGraph("dtw", dset, pa_d[:B], pb_d[:B])

    def replay():
        with torch.cuda.stream(g.stream):
            for b in range(nb):
                sl = slice(b * B, (b + 1) * B)
                g.pa.copy_(pa_d[sl])
                g.pb.copy_(pb_d[sl])
                g.launch()
                out[sl].copy_(g.out)

    def best(fn, sync):
        fn()
        sync()
        ts_ = []
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            sync()
            ts_.append(time.perf_counter() - t)
        return min(ts_)

    t_e = best(eager, torch.cuda.synchronize)
    ref = out.clone()
    t_r = best(replay, g.stream.synchronize)
    same = bool(torch.equal(out, ref))
    g.close()
    # the same batches through the planner (direct launches switched off)
    from trajsimimeasures_amd import _lib
    _lib.check(_lib.lib().tsm_tune_set_direct_pairs(0))
    try:
        t_p = best(eager, torch.cuda.synchronize)
    finally:
        _lib.check(_lib.lib().tsm_tune_set_direct_pairs(-1))
    same = same and bool(torch.equal(out, ref))
    res["batches_512"] = {"workload": f"{nb} batches of {B} pairs of the 100k-pair DTW workload",
                          "eager_pairs_per_s": nb * B / t_e, "replay_pairs_per_s": nb * B / t_r,
                          "eager_planned_pairs_per_s": nb * B / t_p,
                          "note": "eager and replayed calls of this size are direct launches (one wavefront per "
                                  "pair, no planner); eager_planned: the same calls through the planner",
                          "equal": same}
    log(str(res["batches_512"]))
    # query x database, 1e8 pairs
    rng = np.random.default_rng(0)
    nq, nd = 1000, 100_000
    ts = random_walks(rng, lengths_lognormal(rng, nq + nd))
    qset = engine.DeviceTrajSet(ts)
    lens = ts.lengths.astype(np.float64)
    cells = float(lens[:nq].sum() * lens[nq:].sum())
    res["query_db"] = {"workload": f"{nq} queries x {nd} database random walks (lognormal lengths, median 60, "
                                   f"sigma 0.8, [10, 1000], seed 0): {nq * nd} pairs, {cells:.4e} cells"}
    for m in ("dtw", "edr"):
        holder = {}

        def run_eager():
            holder["out"] = engine.grid(m, qset, 0, nq, nq, nd, eps=EPS)

        t_e = best(run_eager, torch.cuda.synchronize)
        gg = engine.GridGraph(m, qset, 0, nq, nq, nd, eps=EPS)
        t_r = best(gg.launch, gg.stream.synchronize)
        same = bool(torch.equal(gg.out, holder["out"].reshape(-1)))
        gg.close()
        del gg, holder
        res["query_db"][m] = {"eager_pairs_per_s": nq * nd / t_e, "replay_pairs_per_s": nq * nd / t_r,
                              "eager_cells_per_s": cells / t_e, "replay_cells_per_s": cells / t_r, "equal": same}
        log(f"{m}: {res['query_db'][m]}")
    return res


def host_leg(ts, pa, pb, dev, reps: int = 3) -> dict:
    """The same DTW pairs handed over in host memory, timed from the host
    buffers to the results back in host memory (PCIe included), best of
    `reps` after one warm-up:
    - padded: the reference's boundary, core.dtw(t1, len1, t2, len2) on
      zero-padded float64 [pairs, max length, 2] arrays (core/dtw.py:56-77);
    - padded in the reference's batches: the same arrays in 512-pair
      slices (effi_batch_size_gpu, config.py:43), one core.dtw call each, as
      the reference's pairwise driver loops (test_trajsimi_time.py:222-276);
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
    B = 512

    def batches():
        for o in range(0, len(pa), B):
            core.dtw(t1[o:o + B], la[o:o + B], t2[o:o + B], lb[o:o + B])

    t_bat = best(batches)
    t_pack = best(lambda: engine.pairs("dtw", engine.DeviceTrajSet(ts, device=dev), pa, pb).cpu().numpy())
    n = len(pa)
    out = {"padded_pairs_per_s": n / t_pad, "padded_s": t_pad, "padded_bytes": t1.nbytes + t2.nbytes,
           "padded_512_batches_pairs_per_s": n / t_bat, "padded_512_batches_s": t_bat,
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
    # FP64 rerun of the pairs touching a trajectory over 1,000 points: the
    # FP32 drift of the cumulative-sum measures (DTW, and ERP with the set's
    # centroid as gap point)
    from trajsimimeasures_amd import drivers
    iu0, iu1 = np.triu_indices(n, 1)
    long = ts.lengths > 1000
    sel = long[iu0] | long[iu1]
    lpa, lpb = iu0[sel].astype(np.int32), iu1[sel].astype(np.int32)
    del iu0, iu1, sel
    drift = {"pairs": int(lpa.size)}
    for m in ("dtw", "erp"):
        t = time.perf_counter()
        _, _, drift[m] = drivers.fp64_drift(m, dset, lpa, lpb, gpoint=ts.xy.mean(0) if m == "erp" else None)
        log(f"FP64 rerun of {lpa.size} long-tail {m} pairs: max relative FP32 drift {drift[m]:.3e} "
            f"({time.perf_counter() - t:.2f}s)")
    out = {"workload": f"{n} random walks (lengths 98% U(10..500), 2% U(1000..2000), seed 0), DTW upper "
                       f"triangle ({n * (n - 1) // 2} pairs), symmetric matrix on the device, FasterPAM k={k} "
                       f"from RandomState(0) initial medoids",
           "matrix_s": t_mat, "fasterpam_s": t_gpu, "n_iter": res.n_iter, "n_swap": res.n_swap, "loss": res.loss,
           "fp64_drift_long_tail": drift}
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
    """Exact DTW kNN against the learned filter: recall@10 and q/s at
    C in {25, 50, 100, 150, 200} candidates for the untrained encoder (the
    resampled-L2 embedding), the trained one (the default training run: each
    of 2,000 anchors with its 100 nearest by the untrained embedding, ranking
    loss) and one trained on 200,000 random-pair DTW labels (2,000 anchors x
    100 random database trajectories, log-space distance regression as the
    reference fits random pairs' distances).  The headline pair is the
    default C."""
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
    Cs = (25, 50, 100, 150, 200)

    def sweep(index):
        row = {}
        for C in Cs:
            F.filtered_knn(index, 0, nq, k=10, C=C)  # warm-up
            runs = [F.filtered_knn(index, 0, nq, k=10, C=C) for _ in range(3)]
            row[C] = {"recall_at_10": F.recall_at(runs[-1][0], truth), "qps": nq / min(r[1] for r in runs)}
        return row

    untrained = sweep(F.FilterIndex(F.Encoder(), dset, db))
    log(f"untrained: {untrained}")
    enc, info = F.train_filter(dset, db)
    log(f"trained: {info}")
    torch.cuda.synchronize()
    t = time.perf_counter()
    index = F.FilterIndex(enc, dset, db)
    torch.cuda.synchronize()
    t_index = time.perf_counter() - t
    trained = sweep(index)
    log(f"trained: {trained}")
    del index
    enc_r, info_r = F.train_filter(dset, db, sampling="random")
    rnd = sweep(F.FilterIndex(enc_r, dset, db))
    log(f"random-pair labels: {info_r} {rnd}")
    C = F.DEFAULT_C
    out = {"workload": f"{nq} queries x {nd} database random walks, lognormal lengths (median 60, sigma 0.8, "
                       f"[10, 1000]), seed 2024; k=10, {C} candidates, exact DTW re-rank",
           "candidates": C, "exact_qps": nq / t_exact, "exact_s": t_exact,
           "filtered_qps": trained[C]["qps"], "filtered_s": nq / trained[C]["qps"],
           "speedup": trained[C]["qps"] * t_exact / nq, "recall_at_10": trained[C]["recall_at_10"],
           "recall_at_10_untrained": untrained[C]["recall_at_10"], "index_build_s": t_index, "train": info,
           "by_candidates": {"untrained": untrained, "trained": trained, "trained_on_random_pairs": rnd},
           "train_random_pairs": info_r}
    log(str({k: v for k, v in out.items() if k != "by_candidates"}))
    return out


if __name__ == "__main__":
    main()
