"""Callers of the similarity path: the reference's dispatch and collation
helpers and its three GPU drivers, on the engine.

  heuristic_fn_wrapper   utilities/method_helper.py:32-73
  heuristic_colllate_fn  utilities/method_helper.py:142-153
  pairwise               test_trajsimi_time.py:222-276   (pairs of trajectories)
  query_db               test_trajsimi_time_DQ.py:136-181 (|Q| x |D| grid)
  all_pairs              test_clustering_time.py:208-269  (upper triangle, symmetrised)
  clustering_kmedoids    test_clustering_time.py:353-367  (FasterPAM on the matrix, timed)
  cluster_kmedoids       all_pairs + clustering_kmedoids with the matrix kept on the device

The reference drivers pad every 512-pair batch on the host, gather it with
fancy indexing and launch one kernel per batch.  Here a trajectory list is
packed and uploaded once, pair addressing (explicit pairs, grid, triangle)
happens on the device, and one engine call covers the whole workload.  Each
driver returns the values and the reference's timing split
[running, collate, compute] in seconds.
"""
from __future__ import annotations

import time
from functools import partial

import numpy as np

from . import _lib, core
from .data import TrajSet, pack_list

__all__ = ["heuristic_fn_wrapper", "heuristic_colllate_fn", "is_heuristic", "pairwise", "query_db", "all_pairs",
           "symmetric_from_triu", "fp64_drift", "clustering_kmedoids", "cluster_kmedoids", "DEFAULT_EPS",
           "WORKLOAD_EPS", "DEFAULT_DELTA", "DEFAULT_SAR", "DEFAULT_CENTROIDS"]

# EDR/LCSS/CDDS/ST epsilon the reference binds (config.py:52
# trajsimi_edr_lcss_eps = 200, metres on its Mercator datasets)
DEFAULT_EPS = 200.0
# epsilon of the random-walk workloads in the unit square (pass it explicitly)
WORKLOAD_EPS = 0.0025
DEFAULT_DELTA = 60.0  # STEDR/STLCSS time threshold, seconds (config.py:53)
# SAR (distance_eps, time_eps, target_length, threads_per_traj): config.py:55-57
# and core/sar.py:248 (the reference's wrapper leaves threads_per_traj at 64)
DEFAULT_SAR = (200.0, 10, 10, 64)
DEFAULT_CENTROIDS = 10  # config.py:72 clustering_num_centroids

_GPU_FNS = {"dtw": core.dtw, "edr": core.edr, "erp": core.erp, "frechet": core.dfrechet,
            "hausdorff": core.hausdorff, "lcss": core.lcss, "stedr": core.stedr, "stlcss": core.stlcss,
            "cdds": core.cdds, "sar": core.sar}
_MEASURE = {"dtw": "dtw", "edr": "edr", "erp": "erp", "frechet": "frechet", "hausdorff": "hausdorff",
            "lcss": "lcss", "stedr": "stedr", "stlcss": "stlcss", "cdds": "cdds", "sar": "sar"}
_ST = ("stedr", "stlcss")


def is_heuristic(fn_name: str) -> bool:
    """utilities/method_helper.py:27-29."""
    heurs = ['dtw', 'erp', 'frechet', 'hausdorff', 'edr', 'lcss', 'sar', 'cdds']
    return any(h in fn_name for h in heurs)


def heuristic_fn_wrapper(fn_name: str, eps: float = DEFAULT_EPS, threads_per_traj: int = 64,
                         delta: float = DEFAULT_DELTA, sar=DEFAULT_SAR):
    """Name -> batch function with the reference's bound arguments
    (method_helper.py:32-73): eps for edr/lcss, eps and delta for
    stedr/stlcss, eps for cdds, threads_per_traj for all but sar, and
    distance_eps / time_eps / target_length for sar (from ``sar``).  The
    ``*_cpu`` names (traj_dist, core_cpu) are not part of this engine."""
    fn = _GPU_FNS.get(fn_name)
    if fn is None:
        raise NotImplementedError(f"{fn_name!r} is not served by this engine (GPU measures: {sorted(_GPU_FNS)})")
    if fn_name == "sar":
        return partial(fn, distance_eps=sar[0], time_eps=sar[1], target_length=sar[2])
    if fn_name in ("edr", "lcss", "cdds"):
        fn = partial(fn, eps=eps)
    elif fn_name in _ST:
        fn = partial(fn, eps=eps, delta=delta)
    return partial(fn, threads_per_traj=threads_per_traj)


def heuristic_colllate_fn(batch, is_pad: bool = True, quantize_fp32: bool = False):
    """method_helper.py:142-153: a list of (traj, traj2) -> padded
    (t1, len1, t2, len2).  The reference pads through float32 tensors
    (tool_funcs.py:230-234), rounding coordinates to FP32; quantize_fp32=True
    reproduces that, the default keeps FP64."""
    lst_trajs, lst_trajs2 = zip(*batch)
    len1 = [len(t) for t in lst_trajs]
    len2 = [len(t) for t in lst_trajs2]
    if not is_pad:
        return lst_trajs, len1, lst_trajs2, len2
    padded = core.pad_lists_to_array(list(lst_trajs) + list(lst_trajs2))
    if quantize_fp32:
        padded = padded.astype(np.float32).astype(np.float64)
    return np.copy(padded[:len(len1)]), len1, np.copy(padded[len(len1):]), len2


def _opts(method: str, eps, gpoint, lanes, precision, delta=DEFAULT_DELTA, sar=DEFAULT_SAR):
    if method not in _MEASURE:
        raise NotImplementedError(f"{method!r} is not served by this engine")
    if method == "sar":
        return dict(eps=float(sar[0]), time_step=sar[1], target_length=sar[2], replicas=sar[3])
    return dict(eps=eps if method in ("edr", "lcss", "cdds") + _ST else 0.0, gpoint=gpoint if method == "erp" else None,
                lanes=lanes, precision=precision, delta=delta if method in _ST else 0.0)


def _device_set(ts: TrajSet):
    import torch

    from .engine import DeviceTrajSet
    if not torch.cuda.is_available():
        raise RuntimeError("the engine needs a HIP device")
    return DeviceTrajSet(ts)


def pairwise(method: str, pairs, *, eps: float = DEFAULT_EPS, gpoint=None, lanes: int = 0, precision="fp32",
              delta: float = DEFAULT_DELTA, sar=DEFAULT_SAR):
    """Values of ``method`` for a list of (traj, traj2) pairs
    (test_trajsimi_time.py:222-276), or for a dataset with a ``packed()``
    form (``datasets.DatasetST_h5``: its pairs go to the device without
    building per-pair lists).  Returns (list of floats, [running, collate,
    compute] seconds)."""
    import torch

    from . import engine
    t0 = time.time()
    if hasattr(pairs, "packed"):
        ts, pa, pb = pairs.packed()
        n = int(pa.size)
    else:
        pairs = list(pairs)
        n = len(pairs)
        if n:
            a, b = zip(*pairs)
            ts = pack_list(list(a) + list(b))
            pa = np.arange(n, dtype=np.int32)
            pb = pa + n
    if not n:
        return [], [0.0, 0.0, 0.0]
    dset = _device_set(ts)
    collate = time.time() - t0
    t1 = time.time()
    out = engine.pairs(method, dset, pa, pb, validate=False, **_opts(method, eps, gpoint, lanes, precision, delta, sar))
    torch.cuda.synchronize()
    dists = out.cpu().numpy().tolist()
    comp = time.time() - t1
    return dists, [time.time() - t0, collate, comp]


def query_db(method: str, Q, D, *, eps: float = DEFAULT_EPS, gpoint=None, lanes: int = 0, precision="fp32",
              delta: float = DEFAULT_DELTA, sar=DEFAULT_SAR):
    """[|Q|, |D|] matrix of ``method`` between every query and every database
    trajectory (test_trajsimi_time_DQ.py:136-181, whose flattened order is
    query-major).  Returns (float64 ndarray, [running, collate, compute])."""
    import torch

    from . import engine
    t0 = time.time()
    Q, D = list(Q), list(D)
    ts = pack_list(Q + D)
    dset = _device_set(ts)
    collate = time.time() - t0
    t1 = time.time()
    out = engine.grid(method, dset, 0, len(Q), len(Q), len(D), **_opts(method, eps, gpoint, lanes, precision, delta, sar))
    torch.cuda.synchronize()
    mat = out.cpu().numpy()
    comp = time.time() - t1
    return mat, [time.time() - t0, collate, comp]


def symmetric_from_triu(vals: np.ndarray, n: int) -> np.ndarray:
    """Row-major strict upper triangle (itertools.combinations order) ->
    symmetric [n, n] matrix with a zero diagonal (test_clustering_time.py:256-269)."""
    vals = np.asarray(vals, dtype=np.float64)
    if vals.size != n * (n - 1) // 2:
        raise ValueError("triangle size does not match n")
    m = np.zeros((n, n), dtype=np.float64)
    iu = np.triu_indices(n, 1)
    m[iu] = vals
    return m + m.T


def all_pairs(method: str, trajs, *, eps: float = DEFAULT_EPS, gpoint=None, lanes: int = 0, precision="fp32",
              delta: float = DEFAULT_DELTA, sar=DEFAULT_SAR):
    """Symmetric all-pairs matrix over ``trajs`` computed on the upper
    triangle only (test_clustering_time.py:208-269).  Returns (float64
    [n, n], [running, collate, compute])."""
    import torch

    from . import engine
    t0 = time.time()
    ts = trajs if isinstance(trajs, TrajSet) else pack_list(list(trajs))
    n = ts.n
    dset = _device_set(ts)
    collate = time.time() - t0
    t1 = time.time()
    tri = engine.triu(method, dset, n, **_opts(method, eps, gpoint, lanes, precision, delta, sar))
    torch.cuda.synchronize()
    vals = tri.cpu().numpy()
    comp = time.time() - t1
    return symmetric_from_triu(vals, n), [time.time() - t0, collate, comp]


def clustering_kmedoids(dists, k: int = DEFAULT_CENTROIDS, random_state=None):
    """test_clustering_time.py:353-367: FasterPAM k-medoids of the distance
    matrix, timed; returns [clustering_time] in seconds, or [-1] when the
    clustering fails (the reference logs the error)."""
    import torch

    from . import cluster
    t0 = time.time()
    try:
        cluster.fasterpam(dists, k, random_state=random_state)
        torch.cuda.synchronize()
    except (ValueError, RuntimeError):
        return [-1]
    return [time.time() - t0]


def cluster_kmedoids(method: str, trajs, k: int = DEFAULT_CENTROIDS, *, eps: float = DEFAULT_EPS, gpoint=None,
                     lanes: int = 0, precision="fp32", delta: float = DEFAULT_DELTA, sar=DEFAULT_SAR,
                     max_iter: int = 100, init: str = "random", random_state=None):
    """The clustering workload end to end on the device: ``method`` over the
    upper triangle of ``trajs``, the symmetric matrix expanded on the device,
    FasterPAM with ``k`` medoids.  Returns (KMedoidsResult, [running,
    collate, matrix, clustering] seconds)."""
    import torch

    from . import cluster, engine
    t0 = time.time()
    ts = trajs if isinstance(trajs, TrajSet) else pack_list(list(trajs))
    n = ts.n
    dset = _device_set(ts)
    collate = time.time() - t0
    t1 = time.time()
    tri = engine.triu(method, dset, n, **_opts(method, eps, gpoint, lanes, precision, delta, sar))
    D = cluster.square_from_triu(tri, n)
    del tri
    torch.cuda.synchronize()
    t_mat = time.time() - t1
    t2 = time.time()
    res = cluster.fasterpam(D, k, max_iter=max_iter, init=init, random_state=random_state)
    t_clu = time.time() - t2
    return res, [time.time() - t0, collate, t_mat, t_clu]


def method_names():
    return sorted(_GPU_FNS)


def measure_id(name: str) -> int:
    return _lib.measure_id(_MEASURE.get(name, name))


def fp64_drift(method: str, dset, pa, pb, *, gpoint=None, lanes: int = 0):
    """Reruns pairs (pa[k], pb[k]) in FP64 on the GPU and returns
    (fp32 values, fp64 values, max relative drift).  The drift check of the
    cumulative-sum measures (DTW, ERP) on long trajectories."""
    import torch

    from . import engine
    if method not in ("dtw", "erp", "frechet", "hausdorff"):
        raise ValueError("FP64 reruns apply to dtw, erp, frechet and hausdorff")
    g = gpoint if method == "erp" else None
    v32 = engine.pairs(method, dset, pa, pb, gpoint=g, lanes=lanes, validate=False)
    v64 = engine.pairs(method, dset, pa, pb, gpoint=g, lanes=lanes, precision="fp64", validate=False)
    torch.cuda.synchronize()
    a, b = v32.cpu().numpy(), v64.cpu().numpy()
    den = np.maximum(np.abs(b), 1e-300)
    drift = float(np.max(np.abs(a - b) / den)) if a.size else 0.0
    return a, b, drift
