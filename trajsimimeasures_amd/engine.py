"""Device-resident engine API over libtsm.so.

``DeviceTrajSet`` uploads a packed ``TrajSet`` once (FP64 coordinates, the
FP32 copy, offsets) into HBM through torch's allocator; ``pairs``, ``grid``
and ``triu`` run a measure over explicit pairs, a query x database grid or
the upper triangle of the all-pairs matrix, on torch's current stream, and
return device tensors.  torch supplies device memory and the stream only; the
compute is the HIP kernels of libtsm.so.
"""
from __future__ import annotations

import ctypes
import math

import numpy as np
import torch

from . import _lib
from .data import TrajSet

__all__ = ["DeviceTrajSet", "pairs", "grid", "triu", "KernelTimer", "PairsGraph", "GridGraph", "TriuGraph",
           "fp32_origin"]


def fp32_origin(xy: np.ndarray):
    """Origin of a set's FP32 copy and the bound on |coordinate - origin|:
    per axis the middle of the bounding box, rounded to a multiple of the
    FP32 spacing of the axis' largest |value| (FP32-valued coordinates of
    one binade per axis then stay exact after the subtraction).  The same
    rule as the library's own host entry point uses."""
    xy = np.asarray(xy, np.float64).reshape(-1, 2)
    if xy.size == 0:
        return (0.0, 0.0), 0.0
    # per column (numpy's axis-0 reduction of an [N, 2] array is ~5x slower)
    lo = np.array([xy[:, 0].min(), xy[:, 1].min()])
    hi = np.array([xy[:, 0].max(), xy[:, 1].max()])
    o = [0.0, 0.0]
    cmax = 0.0
    for k in range(2):
        big = max(abs(float(lo[k])), abs(float(hi[k])))
        if big > 0.0:
            q = math.ldexp(1.0, math.frexp(big)[1] - 1 - 23)
            o[k] = float(np.rint((0.5 * float(lo[k]) + 0.5 * float(hi[k])) / q) * q)
        cmax = max(cmax, abs(float(lo[k]) - o[k]), abs(float(hi[k]) - o[k]))
    return (o[0], o[1]), cmax


def _stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class DeviceTrajSet:
    """A TrajSet resident on the current HIP device."""

    def __init__(self, ts: TrajSet, device="cuda"):
        if ts.n < 1:
            raise ValueError("empty trajectory set")
        lens = ts.lengths
        if lens.min() < 1:
            raise ValueError("every trajectory needs at least one point")
        if not np.all(np.isfinite(ts.xy)):
            raise ValueError("coordinates must be finite")
        self.host = ts
        self.n = ts.n
        self.npts = int(ts.offsets[-1])
        self.max_len = int(lens.max())
        self.lengths = torch.as_tensor(lens, dtype=torch.int64, device=device)
        self.xy64 = torch.as_tensor(ts.xy, dtype=torch.float64, device=device).contiguous()
        self.offsets = torch.as_tensor(ts.offsets, dtype=torch.int64, device=device).contiguous()
        # the FP32 copy is taken relative to the set's centre (Mercator
        # coordinates are ~1e7 m); coord_max bounds its |coordinates| for the
        # FP32 match test of EDR/LCSS
        self.origin, self.coord_max = fp32_origin(ts.xy)
        self.xy32 = torch.empty((self.npts, 2), dtype=torch.float32, device=device)
        _lib.check(_lib.lib().tsm_xy_to_fp32_centred(self.xy64.data_ptr(), self.xy32.data_ptr(), self.npts,
                                                     self.origin[0], self.origin[1], _stream_ptr()))
        # per-point timestamps (STEDR / STLCSS), if the set has them
        self.t64 = None
        if ts.t is not None:
            if not np.all(np.isfinite(ts.t)):
                raise ValueError("timestamps must be finite")
            self.t64 = torch.as_tensor(ts.t, dtype=torch.float64, device=device).contiguous()

    def refill(self, xy, t=None):
        """Overwrite the coordinates (and timestamps) in place, keeping the
        lengths, the FP32 origin and the coordinate bound: the FP32 copy is
        regenerated relative to the same origin, and points farther than
        ``coord_max`` from it are refused, since the EDR/LCSS guard band
        (derived from that bound, also inside any graph captured on this
        set) would no longer cover them.  Stream-ordered on torch's current
        stream."""
        xy = np.ascontiguousarray(xy, dtype=np.float64).reshape(-1, 2)
        if xy.shape[0] != self.npts:
            raise ValueError(f"refill needs {self.npts} points, got {xy.shape[0]}")
        if not np.all(np.isfinite(xy)):
            raise ValueError("coordinates must be finite")
        dev = max(float(np.max(np.abs(xy[:, 0] - self.origin[0]))), float(np.max(np.abs(xy[:, 1] - self.origin[1]))))
        if dev > self.coord_max:
            raise ValueError(f"refilled points lie {dev:.6g} from the set's FP32 origin, beyond its bound "
                             f"{self.coord_max:.6g}; build a new DeviceTrajSet (and recapture its graphs)")
        if t is not None:
            if self.t64 is None:
                raise ValueError("this set has no timestamps")
            t = np.ascontiguousarray(t, dtype=np.float64).reshape(-1)
            if t.shape[0] != self.npts or not np.all(np.isfinite(t)):
                raise ValueError(f"refill needs {self.npts} finite timestamps")
            self.t64.copy_(torch.as_tensor(t), non_blocking=False)
        self.xy64.copy_(torch.as_tensor(xy), non_blocking=False)
        _lib.check(_lib.lib().tsm_xy_to_fp32_centred(self.xy64.data_ptr(), self.xy32.data_ptr(), self.npts,
                                                     self.origin[0], self.origin[1], _stream_ptr()))
        self.host = TrajSet(xy, self.host.offsets, t if t is not None else self.host.t)

    def cstruct(self) -> _lib.TrajSetC:
        return _lib.TrajSetC(self.xy32.data_ptr(), self.xy64.data_ptr(), self.offsets.data_ptr(), self.n,
                             self.npts, self.max_len, self.coord_max,
                             self.t64.data_ptr() if self.t64 is not None else None)


def _outs(n: int, measure: int, device, want_counts: bool):
    out = torch.empty(n, dtype=torch.float64, device=device)
    cnt = None
    if want_counts and measure in _lib.COUNTED:
        cnt = torch.empty(n, dtype=torch.int32, device=device)
    return out, cnt


def pairs(measure, dset: DeviceTrajSet, pa, pb, *, eps=0.0, eps_mode="sqrt", gpoint=None, lanes=0,
          precision="fp32", return_counts=False, validate=True, out=None, delta=0.0, time_step=None,
          target_length=None, replicas=None):
    """Measure over explicit pairs (pa[k], pb[k]) of ``dset`` (pa as t1).

    validate=False skips the (synchronising) index range check; ``out`` may
    supply a float64 device tensor of len(pa) to write into.  SAR reads
    ``eps`` (distance_eps), ``time_step`` (time_eps), ``target_length`` and
    ``replicas`` (the reference's threads_per_traj)."""
    m = _lib.measure_id(measure)
    dev = dset.xy64.device
    pa = torch.as_tensor(pa, dtype=torch.int32, device=dev).contiguous()
    pb = torch.as_tensor(pb, dtype=torch.int32, device=dev).contiguous()
    if pa.shape != pb.shape or pa.dim() != 1:
        raise ValueError("pa and pb must be 1-D and of equal length")
    n = pa.numel()
    if validate and n and (int(pa.min()) < 0 or int(pb.min()) < 0 or int(pa.max()) >= dset.n
                           or int(pb.max()) >= dset.n):
        raise IndexError("pair index outside the trajectory set")
    if out is None:
        out, cnt = _outs(n, m, dev, return_counts)
    else:
        if out.dtype != torch.float64 or out.numel() != n or not out.is_contiguous() or out.device != dev:
            raise ValueError("out must be a contiguous float64 tensor of len(pa) on the set's device")
        cnt = None
        if return_counts and m in _lib.COUNTED:
            cnt = torch.empty(n, dtype=torch.int32, device=dev)
    prm = _lib.params(eps, eps_mode, gpoint, lanes, precision, delta, time_step, target_length, replicas)
    s = dset.cstruct()
    _lib.check(_lib.lib().tsm_pairs(m, ctypes.byref(s), pa.data_ptr(), pb.data_ptr(), n, ctypes.byref(prm),
                                    out.data_ptr(), cnt.data_ptr() if cnt is not None else None, _stream_ptr()))
    return (out, cnt) if return_counts else out


def grid(measure, dset: DeviceTrajSet, q0: int, nq: int, d0: int, nd: int, *, eps=0.0, eps_mode="sqrt",
         gpoint=None, lanes=0, precision="fp32", return_counts=False, delta=0.0, time_step=None,
         target_length=None, replicas=None):
    """Measure over the grid (q0 + q, d0 + d); returns [nq, nd] (row-major)."""
    m = _lib.measure_id(measure)
    dev = dset.xy64.device
    out, cnt = _outs(nq * nd, m, dev, return_counts)
    prm = _lib.params(eps, eps_mode, gpoint, lanes, precision, delta, time_step, target_length, replicas)
    s = dset.cstruct()
    _lib.check(_lib.lib().tsm_pairs_grid(m, ctypes.byref(s), q0, nq, d0, nd, ctypes.byref(prm), out.data_ptr(),
                                         cnt.data_ptr() if cnt is not None else None, _stream_ptr()))
    out = out.view(nq, nd)
    if return_counts:
        return out, (cnt.view(nq, nd) if cnt is not None else None)
    return out


def triu(measure, dset: DeviceTrajSet, n: int | None = None, *, eps=0.0, eps_mode="sqrt", gpoint=None,
         lanes=0, precision="fp32", return_counts=False, delta=0.0, time_step=None, target_length=None,
         replicas=None):
    """Upper triangle over trajectories [0, n): n*(n-1)/2 values in
    itertools.combinations order."""
    m = _lib.measure_id(measure)
    n = dset.n if n is None else int(n)
    dev = dset.xy64.device
    npairs = n * (n - 1) // 2
    out, cnt = _outs(npairs, m, dev, return_counts)
    prm = _lib.params(eps, eps_mode, gpoint, lanes, precision, delta, time_step, target_length, replicas)
    s = dset.cstruct()
    _lib.check(_lib.lib().tsm_pairs_triu(m, ctypes.byref(s), n, ctypes.byref(prm), out.data_ptr(),
                                         cnt.data_ptr() if cnt is not None else None, _stream_ptr()))
    return (out, cnt) if return_counts else out


class KernelTimer:
    """HIP-event timing of the DP / Hausdorff / CDDS / SAR kernels inside
    libtsm.so, on the stream they were launched on: every ``every``-th launch
    is timed (a timed dispatch costs a few microseconds).  After the block,
    ``ms`` is the timed launches' total and ``launches`` their number."""

    def __init__(self, every: int = 1):
        if every < 1:
            raise ValueError("every must be >= 1")
        self.every = int(every)

    def __enter__(self):
        _lib.check(_lib.lib().tsm_profile_enable(self.every))
        return self

    def __exit__(self, *exc):
        ms = ctypes.c_double(0.0)
        n = ctypes.c_int64(0)
        _lib.check(_lib.lib().tsm_profile_read(ctypes.byref(ms), ctypes.byref(n)))
        _lib.check(_lib.lib().tsm_profile_enable(0))
        self.ms = ms.value
        self.launches = n.value
        return False


class _Graph:
    """A batch captured once into a hipGraph (planner + persistent kernel
    launches, no memsets) and replayed with ``launch()`` on its own stream
    (default priority) unless one is given.  Results land in ``self.out``
    (``self.counts`` for the counting measures).  Between launches the set's
    coordinates may be refilled in place (``DeviceTrajSet.refill``), and for
    explicit pairs the pair tensors; sizes may not change."""

    def _setup(self, measure, dset, npairs, stream, prm):
        m = _lib.measure_id(measure)
        dev = dset.xy64.device
        self.dset = dset
        self.measure = m
        self.out = torch.empty(npairs, dtype=torch.float64, device=dev)
        self.counts = torch.empty(npairs, dtype=torch.int32, device=dev) if m in _lib.COUNTED else None
        self.stream = stream if stream is not None else torch.cuda.Stream(device=dev)
        self._prm = prm
        self._set = dset.cstruct()
        self._h = ctypes.c_void_p()
        torch.cuda.current_stream().synchronize()

    def _args_tail(self):
        return (ctypes.byref(self._prm), self.out.data_ptr(),
                self.counts.data_ptr() if self.counts is not None else None,
                ctypes.c_void_p(self.stream.cuda_stream), ctypes.byref(self._h))

    def launch(self):
        """Replay the batch on the graph's stream; results land in self.out."""
        if not self._h:
            raise RuntimeError("graph already closed")
        _lib.check(_lib.lib().tsm_graph_launch(self._h, ctypes.c_void_p(self.stream.cuda_stream)))
        return self.out

    def close(self):
        if self._h:
            self.stream.synchronize()
            _lib.check(_lib.lib().tsm_graph_destroy(self._h))
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PairsGraph(_Graph):
    """Explicit pairs (pa[k], pb[k]) of ``dset`` captured into a hipGraph;
    ``pa``/``pb`` (device int32) may be refilled in place between launches."""

    def __init__(self, measure, dset: DeviceTrajSet, pa, pb, *, eps=0.0, eps_mode="sqrt", gpoint=None, lanes=0,
                 precision="fp32", stream=None, delta=0.0, time_step=None, target_length=None, replicas=None):
        dev = dset.xy64.device
        # the graph's own copies: refilling them must not write into the caller's tensors
        self.pa = torch.as_tensor(pa, dtype=torch.int32, device=dev).clone(memory_format=torch.contiguous_format)
        self.pb = torch.as_tensor(pb, dtype=torch.int32, device=dev).clone(memory_format=torch.contiguous_format)
        n = self.pa.numel()
        if self.pb.numel() != n or n < 1:
            raise ValueError("pa and pb must be non-empty and of equal length")
        self._setup(measure, dset, n, stream, _lib.params(eps, eps_mode, gpoint, lanes, precision, delta, time_step,
                                                          target_length, replicas))
        _lib.check(_lib.lib().tsm_graph_capture_pairs(self.measure, ctypes.byref(self._set), self.pa.data_ptr(),
                                                      self.pb.data_ptr(), n, *self._args_tail()))


class GridGraph(_Graph):
    """The query x database grid (q0 + q, d0 + d) of ``dset`` captured into a
    hipGraph; ``out`` is [nq * nd] row-major (``result()`` gives [nq, nd])."""

    def __init__(self, measure, dset: DeviceTrajSet, q0: int, nq: int, d0: int, nd: int, *, eps=0.0,
                 eps_mode="sqrt", gpoint=None, lanes=0, precision="fp32", stream=None, delta=0.0, time_step=None,
                 target_length=None, replicas=None):
        if nq < 1 or nd < 1:
            raise ValueError("the grid needs at least one query and one database trajectory")
        self.shape = (int(nq), int(nd))
        self._setup(measure, dset, nq * nd, stream, _lib.params(eps, eps_mode, gpoint, lanes, precision, delta,
                                                                time_step, target_length, replicas))
        _lib.check(_lib.lib().tsm_graph_capture_grid(self.measure, ctypes.byref(self._set), q0, nq, d0, nd,
                                                     *self._args_tail()))

    def result(self):
        return self.out.view(*self.shape)


class TriuGraph(_Graph):
    """The upper triangle over trajectories [0, n) of ``dset`` (pairs in
    itertools.combinations order) captured into a hipGraph."""

    def __init__(self, measure, dset: DeviceTrajSet, n: int | None = None, *, eps=0.0, eps_mode="sqrt",
                 gpoint=None, lanes=0, precision="fp32", stream=None, delta=0.0, time_step=None, target_length=None,
                 replicas=None):
        n = dset.n if n is None else int(n)
        if n < 2:
            raise ValueError("the triangle needs at least two trajectories")
        self.n = n
        self._setup(measure, dset, n * (n - 1) // 2, stream, _lib.params(eps, eps_mode, gpoint, lanes, precision,
                                                                         delta, time_step, target_length, replicas))
        _lib.check(_lib.lib().tsm_graph_capture_triu(self.measure, ctypes.byref(self._set), n, *self._args_tail()))
