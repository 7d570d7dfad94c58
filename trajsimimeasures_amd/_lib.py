"""ctypes binding of libtsm.so (include/tsm.h).

The library is the only compute path: if it is missing or fails to load,
``lib()`` raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("TSM_LIB", os.path.join(PKG, "libtsm.so"))

# Measure ids (tsm.h)
DTW, EDR, LCSS, ERP, FRECHET, HAUSDORFF, STEDR, STLCSS, CDDS, SAR = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9
MEASURES = {"dtw": DTW, "edr": EDR, "lcss": LCSS, "erp": ERP, "frechet": FRECHET, "dfrechet": FRECHET,
            "discret_frechet": FRECHET, "hausdorff": HAUSDORFF, "stedr": STEDR, "stlcss": STLCSS, "cdds": CDDS,
            "sar": SAR}
MEASURE_NAMES = {DTW: "dtw", EDR: "edr", LCSS: "lcss", ERP: "erp", FRECHET: "frechet", HAUSDORFF: "hausdorff",
                 STEDR: "stedr", STLCSS: "stlcss", CDDS: "cdds", SAR: "sar"}
COUNTED = (EDR, LCSS, STEDR, STLCSS)  # measures with integer match counts
TIMED = (STEDR, STLCSS, CDDS, SAR)  # measures that need per-point timestamps

EPS_SQRT, EPS_SQUARED = 0, 1
PREC_FP32, PREC_FP64 = 0, 1

ERRORS = {-1: "TSM_E_ARG", -2: "TSM_E_HIP", -3: "TSM_E_NOMEM", -4: "TSM_E_UNSUP"}

# Every function include/tsm.h declares; the load check verifies all exist.
EXPORTS = ("tsm_init", "tsm_last_error", "tsm_abi_version", "tsm_params_default", "tsm_xy_to_fp32",
           "tsm_xy_to_fp32_centred",
           "tsm_pairs", "tsm_pairs_noalloc", "tsm_pairs_grid", "tsm_pairs_triu", "tsm_pairs_host", "tsm_pairs_padded_host",
           "tsm_tune_set_lanes",
           "tsm_tune_set_waves_per_cu", "tsm_tune_set_wide_lanes", "tsm_tune_set_direct_pairs", "tsm_graph_capture_pairs", "tsm_graph_capture_grid",
           "tsm_graph_capture_triu", "tsm_graph_launch", "tsm_graph_destroy",
           "tsm_profile_enable", "tsm_profile_read")
# include/tsm_nn.h (argument types are declared by filter.py)
NN_EXPORTS = ("tsm_nn_gemm_nt", "tsm_nn_cast_bf16", "tsm_nn_resample", "tsm_nn_conv_fwd", "tsm_nn_conv_bwd",
              "tsm_nn_gru_h0", "tsm_nn_gru_h0_bwd", "tsm_nn_gru_step_fwd", "tsm_nn_gru_step_bwd", "tsm_nn_colsum",
              "tsm_nn_rank_loss", "tsm_nn_dist_loss", "tsm_nn_adam", "tsm_nn_rownorm_bf16", "tsm_nn_topk_f32", "tsm_nn_topk_f64",
              "tsm_nn_linear_f32", "tsm_nn_split_bf16", "tsm_nn_gemm_nt_filter", "tsm_nn_select_candidates",
              "tsm_nn_select_candidates_pairs", "tsm_nn_gemm_nt_acc", "tsm_nn_gru_forward", "tsm_nn_kth_smallest_f32")
# include/tsm_cluster.h
CLUSTER_EXPORTS = ("tsm_square_from_triu", "tsm_kmedoids_fasterpam")


class TsmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class TrajSetC(ctypes.Structure):
    _fields_ = [("xy32", ctypes.c_void_p), ("xy64", ctypes.c_void_p), ("offsets", ctypes.c_void_p),
                ("n", ctypes.c_int64), ("npts", ctypes.c_int64), ("max_len", ctypes.c_int64),
                ("coord_max", ctypes.c_double), ("t64", ctypes.c_void_p)]


class ParamsC(ctypes.Structure):
    _fields_ = [("eps", ctypes.c_double), ("eps_mode", ctypes.c_int), ("gx", ctypes.c_double),
                ("gy", ctypes.c_double), ("lanes", ctypes.c_int), ("precision", ctypes.c_int),
                ("delta", ctypes.c_double), ("time_step", ctypes.c_int64), ("target_length", ctypes.c_int32),
                ("replicas", ctypes.c_int32)]


_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


def _declare(L):
    L.tsm_init.argtypes = [ctypes.c_int]
    L.tsm_last_error.restype = ctypes.c_char_p
    L.tsm_abi_version.restype = ctypes.c_int
    L.tsm_params_default.argtypes = [ctypes.POINTER(ParamsC)]
    L.tsm_params_default.restype = None
    L.tsm_xy_to_fp32.argtypes = [_P, _P, _I64, _P]
    L.tsm_xy_to_fp32_centred.argtypes = [_P, _P, _I64, ctypes.c_double, ctypes.c_double, _P]
    L.tsm_pairs.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _P, _P, _I64, ctypes.POINTER(ParamsC), _P, _P, _P]
    L.tsm_pairs_noalloc.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _P, _P, _I64, ctypes.POINTER(ParamsC), _P,
                                    _P, _P]
    L.tsm_pairs_grid.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _I64, _I64, _I64, _I64,
                                 ctypes.POINTER(ParamsC), _P, _P, _P]
    L.tsm_pairs_triu.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _I64, ctypes.POINTER(ParamsC), _P, _P, _P]
    L.tsm_pairs_host.argtypes = [ctypes.c_int, _P, _P, _P, _I64, _I64, _P, _P, _I64, ctypes.POINTER(ParamsC), _P, _P]
    L.tsm_pairs_padded_host.argtypes = [ctypes.c_int, _P, _P, _I64, _P, _P, _I64, _I64, ctypes.c_int,
                                        ctypes.POINTER(ParamsC), _P, _P]
    L.tsm_tune_set_lanes.argtypes = [ctypes.c_int, _P, ctypes.c_int]
    L.tsm_square_from_triu.argtypes = [_P, _I64, _P, _P]
    L.tsm_kmedoids_fasterpam.argtypes = [_P, _I64, ctypes.c_int32, _P, ctypes.c_int32, _P, _P, _P, _P, _P]
    L.tsm_tune_set_waves_per_cu.argtypes = [ctypes.c_int]
    L.tsm_tune_set_wide_lanes.argtypes = [ctypes.c_int]
    L.tsm_tune_set_direct_pairs.argtypes = [ctypes.c_int64]
    L.tsm_graph_capture_pairs.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _P, _P, _I64,
                                          ctypes.POINTER(ParamsC), _P, _P, _P, ctypes.POINTER(ctypes.c_void_p)]
    L.tsm_graph_capture_grid.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _I64, _I64, _I64, _I64,
                                         ctypes.POINTER(ParamsC), _P, _P, _P, ctypes.POINTER(ctypes.c_void_p)]
    L.tsm_graph_capture_triu.argtypes = [ctypes.c_int, ctypes.POINTER(TrajSetC), _I64, ctypes.POINTER(ParamsC), _P,
                                         _P, _P, ctypes.POINTER(ctypes.c_void_p)]
    L.tsm_graph_launch.argtypes = [_P, _P]
    L.tsm_graph_destroy.argtypes = [_P]
    L.tsm_profile_enable.argtypes = [ctypes.c_int]
    L.tsm_profile_read.argtypes = [ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64)]
    for name in EXPORTS:
        if name not in ("tsm_last_error", "tsm_params_default"):
            getattr(L, name).restype = ctypes.c_int


def lib():
    """Load libtsm.so once (raises if it is missing or incomplete)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise ImportError(f"libtsm.so not built at {LIB_PATH}; run python -m trajsimimeasures_amd.build")
            L = ctypes.CDLL(LIB_PATH)
            missing = [n for n in EXPORTS + NN_EXPORTS + CLUSTER_EXPORTS if not hasattr(L, n)]
            if missing:
                raise ImportError(f"{LIB_PATH} lacks {missing}")
            _declare(L)
            _lib = L
            # the launch table measured on MI355X (autotune.py), if shipped
            from . import autotune
            autotune.load_shipped()
    return _lib


def check(rc: int):
    if rc != 0:
        msg = lib().tsm_last_error()
        raise TsmError(rc, msg.decode() if msg else "")


def params(eps: float = 0.0, eps_mode="sqrt", gpoint=None, lanes: int = 0, precision="fp32",
           delta: float = 0.0, time_step=None, target_length=None, replicas=None) -> ParamsC:
    p = ParamsC()
    lib().tsm_params_default(ctypes.byref(p))
    p.eps = float(eps)
    p.delta = float(delta)
    # SAR (core/sar.py:247-248): time_eps must be a whole number of seconds
    # (the reference steps a Python range by it)
    if time_step is not None:
        if float(time_step) != int(time_step) or int(time_step) < 1:
            raise ValueError("time_step (time_eps) must be a positive integer")
        p.time_step = int(time_step)
    if target_length is not None:
        if int(target_length) != target_length or not 1 <= int(target_length) < 2 ** 31:
            raise ValueError("target_length must be an integer in [1, 2^31)")
        p.target_length = int(target_length)
    if replicas is not None:
        if int(replicas) != replicas or not 1 <= int(replicas) < 2 ** 31:
            raise ValueError("threads_per_traj must be an integer in [1, 2^31)")
        p.replicas = int(replicas)
    if isinstance(eps_mode, str):
        if eps_mode not in ("sqrt", "squared"):
            raise ValueError("eps_mode must be 'sqrt' or 'squared'")
        eps_mode = EPS_SQRT if eps_mode == "sqrt" else EPS_SQUARED
    p.eps_mode = int(eps_mode)
    if gpoint is not None:
        g = [float(v) for v in list(__import__("numpy").asarray(gpoint, dtype="float64").reshape(-1))]
        if len(g) != 2:
            raise ValueError("gpoint must hold two coordinates")
        p.gx, p.gy = g
    p.lanes = int(lanes)
    if isinstance(precision, str):
        if precision not in ("fp32", "fp64"):
            raise ValueError("precision must be 'fp32' or 'fp64'")
        precision = PREC_FP32 if precision == "fp32" else PREC_FP64
    p.precision = int(precision)
    return p


def measure_id(m) -> int:
    if isinstance(m, str):
        if m not in MEASURES:
            raise ValueError(f"unknown measure {m!r}; expected one of {sorted(MEASURES)}")
        return MEASURES[m]
    m = int(m)
    if m not in MEASURE_NAMES:
        raise ValueError(f"unknown measure id {m}")
    return m
