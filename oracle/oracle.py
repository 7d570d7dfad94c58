"""ctypes front of the FP64 CPU oracle (oracle/tsm_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker for tests/, __graft_entry__.smoke() and
bench.py's CPU-baseline leg.  The product (trajsimimeasures_amd) never
imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libtsm_oracle.so")

DTW, EDR, LCSS, ERP, FRECHET, HAUSDORFF, STEDR, STLCSS = 0, 1, 2, 3, 4, 5, 6, 7
NAMES = {"dtw": DTW, "edr": EDR, "lcss": LCSS, "erp": ERP, "frechet": FRECHET, "hausdorff": HAUSDORFF,
         "stedr": STEDR, "stlcss": STLCSS}

_lib = None
_P = ctypes.c_void_p


def build():
    """Compile the oracle with its Makefile (gcc -O2 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        d = ctypes.c_double
        i64 = ctypes.c_int64
        L.tsmo_eps_threshold.argtypes = [d, ctypes.c_int]
        L.tsmo_eps_threshold.restype = d
        for name in ("tsmo_dtw", "tsmo_frechet", "tsmo_hausdorff"):
            f = getattr(L, name)
            f.argtypes = [_P, i64, _P, i64]
            f.restype = d
        for name in ("tsmo_edr", "tsmo_lcss"):
            f = getattr(L, name)
            f.argtypes = [_P, i64, _P, i64, d]
            f.restype = d
        for name in ("tsmo_edr_count", "tsmo_lcss_count"):
            f = getattr(L, name)
            f.argtypes = [_P, i64, _P, i64, d]
            f.restype = i64
        L.tsmo_erp.argtypes = [_P, i64, _P, i64, _P]
        L.tsmo_erp.restype = d
        L.tsmo_pairs.argtypes = [ctypes.c_int, _P, _P, _P, _P, i64, d, _P, _P, _P, ctypes.c_int]
        L.tsmo_pairs.restype = ctypes.c_int
        for name in ("tsmo_stedr_count", "tsmo_stlcss_count"):
            f = getattr(L, name)
            f.argtypes = [_P, _P, i64, _P, _P, i64, d, d]
            f.restype = i64
        L.tsmo_pairs_st.argtypes = [ctypes.c_int, _P, _P, _P, _P, _P, i64, d, d, _P, _P, ctypes.c_int]
        L.tsmo_pairs_st.restype = ctypes.c_int
        L.tsmo_cdd.argtypes = [_P, _P, _P, _P, d]
        L.tsmo_cdd.restype = d
        L.tsmo_cdds.argtypes = [_P, _P, i64, _P, _P, i64, d]
        L.tsmo_cdds.restype = d
        L.tsmo_pairs_cdds.argtypes = [_P, _P, _P, _P, _P, i64, d, _P, ctypes.c_int]
        L.tsmo_pairs_cdds.restype = ctypes.c_int
        L.tsmo_sar.argtypes = [_P, _P, i64, _P, _P, i64, d, i64, i64, i64]
        L.tsmo_sar.restype = d
        L.tsmo_pairs_sar.argtypes = [_P, _P, _P, _P, _P, i64, d, i64, i64, i64, _P, ctypes.c_int]
        L.tsmo_pairs_sar.restype = ctypes.c_int
        L.tsmo_fasterpam.argtypes = [_P, i64, ctypes.c_int, _P, ctypes.c_int, _P, _P, _P, _P]
        L.tsmo_fasterpam.restype = ctypes.c_int
        L.tsmo_fasterpam_delta.argtypes = [_P, i64, ctypes.c_int, _P, i64, _P]
        L.tsmo_fasterpam_delta.restype = d
        L.tsmo_max_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def eps_threshold(eps: float, mode: str = "sqrt") -> float:
    """Squared-distance threshold equivalent to the EDR/LCSS match test."""
    return lib().tsmo_eps_threshold(float(eps), 0 if mode == "sqrt" else 1)


def _arr(t):
    a = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1, 2))
    return a


def single(measure: str, a, b, eps: float = 0.0, eps_mode: str = "sqrt", gpoint=(0.0, 0.0)) -> float:
    """One pair, reference orientation (a = t1)."""
    a, b = _arr(a), _arr(b)
    L = lib()
    m = NAMES[measure]
    if m == DTW:
        return L.tsmo_dtw(a.ctypes.data, len(a), b.ctypes.data, len(b))
    if m == FRECHET:
        return L.tsmo_frechet(a.ctypes.data, len(a), b.ctypes.data, len(b))
    if m == HAUSDORFF:
        return L.tsmo_hausdorff(a.ctypes.data, len(a), b.ctypes.data, len(b))
    if m == ERP:
        g = np.ascontiguousarray(np.asarray(gpoint, dtype=np.float64).reshape(2))
        return L.tsmo_erp(a.ctypes.data, len(a), b.ctypes.data, len(b), g.ctypes.data)
    thr = eps_threshold(eps, eps_mode)
    if m == EDR:
        return L.tsmo_edr(a.ctypes.data, len(a), b.ctypes.data, len(b), thr)
    return L.tsmo_lcss(a.ctypes.data, len(a), b.ctypes.data, len(b), thr)


def count(measure: str, a, b, eps: float, eps_mode: str = "sqrt") -> int:
    a, b = _arr(a), _arr(b)
    thr = eps_threshold(eps, eps_mode)
    f = lib().tsmo_edr_count if measure == "edr" else lib().tsmo_lcss_count
    return int(f(a.ctypes.data, len(a), b.ctypes.data, len(b), thr))


def pairs(measure: str, xy, offsets, pa, pb, eps: float = 0.0, eps_mode: str = "sqrt", gpoint=(0.0, 0.0),
          threads: int = 0):
    """Batched oracle over a packed set; returns (values float64, counts int64)."""
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pa = np.ascontiguousarray(pa, dtype=np.int64)
    pb = np.ascontiguousarray(pb, dtype=np.int64)
    n = pa.size
    out = np.empty(n, dtype=np.float64)
    cnt = np.empty(n, dtype=np.int64)
    g = np.ascontiguousarray(np.asarray(gpoint, dtype=np.float64).reshape(2))
    thr = eps_threshold(eps, eps_mode)
    rc = lib().tsmo_pairs(NAMES[measure], xy.ctypes.data, offsets.ctypes.data, pa.ctypes.data, pb.ctypes.data, n,
                          thr, g.ctypes.data, out.ctypes.data, cnt.ctypes.data, int(threads))
    if rc != 0:
        raise ValueError(f"oracle rejected measure {measure!r}")
    return out, cnt


def max_threads() -> int:
    return lib().tsmo_max_threads()


def st_count(measure: str, a, ta, b, tb, eps: float, delta: float, eps_mode: str = "sqrt") -> int:
    """STEDR / STLCSS match count of one pair: a, b [n, 2], ta, tb [n] times."""
    a, b = _arr(a), _arr(b)
    ta = np.ascontiguousarray(np.asarray(ta, dtype=np.float64).reshape(-1))
    tb = np.ascontiguousarray(np.asarray(tb, dtype=np.float64).reshape(-1))
    f = lib().tsmo_stedr_count if measure == "stedr" else lib().tsmo_stlcss_count
    return f(a.ctypes.data, ta.ctypes.data, len(a), b.ctypes.data, tb.ctypes.data, len(b),
             eps_threshold(eps, eps_mode), float(delta))


def pairs_st(measure: str, xy, t, offsets, pa, pb, eps: float, delta: float, eps_mode: str = "sqrt",
             threads: int = 0):
    """Batched STEDR / STLCSS: (values float64, counts int64)."""
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1))
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pa = np.ascontiguousarray(pa, dtype=np.int64)
    pb = np.ascontiguousarray(pb, dtype=np.int64)
    out = np.empty(pa.size, dtype=np.float64)
    cnt = np.empty(pa.size, dtype=np.int64)
    rc = lib().tsmo_pairs_st(NAMES[measure], xy.ctypes.data, t.ctypes.data, offsets.ctypes.data, pa.ctypes.data,
                             pb.ctypes.data, pa.size, eps_threshold(eps, eps_mode), float(delta), out.ctypes.data,
                             cnt.ctypes.data, int(threads))
    if rc != 0:
        raise ValueError(f"bad measure {measure!r}")
    return out, cnt


def cdd(e1p1, e1p2, e2p1, e2p2, eps: float) -> float:
    """Time two points moving along segments (x, y, t) -> (x, y, t) spend within eps."""
    v = [np.ascontiguousarray(np.asarray(e, dtype=np.float64).reshape(3)) for e in (e1p1, e1p2, e2p1, e2p2)]
    return lib().tsmo_cdd(v[0].ctypes.data, v[1].ctypes.data, v[2].ctypes.data, v[3].ctypes.data, float(eps))


def cdds_single(a, ta, b, tb, eps: float) -> float:
    a, b = _arr(a), _arr(b)
    ta = np.ascontiguousarray(np.asarray(ta, dtype=np.float64).reshape(-1))
    tb = np.ascontiguousarray(np.asarray(tb, dtype=np.float64).reshape(-1))
    return lib().tsmo_cdds(a.ctypes.data, ta.ctypes.data, len(a), b.ctypes.data, tb.ctypes.data, len(b), float(eps))


def pairs_cdds(xy, t, offsets, pa, pb, eps: float, threads: int = 0):
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1))
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pa = np.ascontiguousarray(pa, dtype=np.int64)
    pb = np.ascontiguousarray(pb, dtype=np.int64)
    out = np.empty(pa.size, dtype=np.float64)
    lib().tsmo_pairs_cdds(xy.ctypes.data, t.ctypes.data, offsets.ctypes.data, pa.ctypes.data, pb.ctypes.data,
                          pa.size, float(eps), out.ctypes.data, int(threads))
    return out


def sar_single(a, ta, b, tb, distance_eps: float, time_eps: int, target_length: int,
               threads_per_traj: int = 64) -> float:
    """SAR of one pair, the reference GPU form (core/sar.py:62-212)."""
    a, b = _arr(a), _arr(b)
    ta = np.ascontiguousarray(np.asarray(ta, dtype=np.float64).reshape(-1))
    tb = np.ascontiguousarray(np.asarray(tb, dtype=np.float64).reshape(-1))
    return lib().tsmo_sar(a.ctypes.data, ta.ctypes.data, len(a), b.ctypes.data, tb.ctypes.data, len(b),
                          float(distance_eps), int(time_eps), int(target_length), int(threads_per_traj))


def pairs_sar(xy, t, offsets, pa, pb, distance_eps: float, time_eps: int, target_length: int,
              threads_per_traj: int = 64, threads: int = 0):
    xy = np.ascontiguousarray(np.asarray(xy, dtype=np.float64).reshape(-1, 2))
    t = np.ascontiguousarray(np.asarray(t, dtype=np.float64).reshape(-1))
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    pa = np.ascontiguousarray(pa, dtype=np.int64)
    pb = np.ascontiguousarray(pb, dtype=np.int64)
    out = np.empty(pa.size, dtype=np.float64)
    lib().tsmo_pairs_sar(xy.ctypes.data, t.ctypes.data, offsets.ctypes.data, pa.ctypes.data, pb.ctypes.data,
                         pa.size, float(distance_eps), int(time_eps), int(target_length), int(threads_per_traj),
                         out.ctypes.data, int(threads))
    return out


def fasterpam(diss, medoids, max_iter: int = 100):
    """FasterPAM from the given initial medoids (see tsm_oracle.c); returns
    (loss, labels, medoids, n_iter, n_swap) like kmedoids.fasterpam."""
    D = np.ascontiguousarray(np.asarray(diss, dtype=np.float64))
    n = D.shape[0]
    med = np.ascontiguousarray(np.asarray(medoids, dtype=np.int64).copy())
    labels = np.empty(n, dtype=np.int32)
    loss = ctypes.c_double()
    it = ctypes.c_int()
    sw = ctypes.c_int()
    rc = lib().tsmo_fasterpam(D.ctypes.data, n, med.size, med.ctypes.data, int(max_iter), labels.ctypes.data,
                              ctypes.byref(loss), ctypes.byref(it), ctypes.byref(sw))
    if rc != 0:
        raise ValueError("bad k-medoids arguments")
    return loss.value, labels, med, it.value, sw.value


def fasterpam_delta(diss, medoids, j: int):
    """(loss change, medoid position swapped out) of the best swap bringing
    point j in, from the assignment to the given medoids."""
    D = np.ascontiguousarray(np.asarray(diss, dtype=np.float64))
    med = np.ascontiguousarray(np.asarray(medoids, dtype=np.int64))
    b = ctypes.c_int()
    v = lib().tsmo_fasterpam_delta(D.ctypes.data, D.shape[0], med.size, med.ctypes.data, int(j), ctypes.byref(b))
    return v, b.value
