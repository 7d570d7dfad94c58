"""Learned kNN filter for DTW: a trajectory encoder trained on distances the
exact DP kernel produces, used to shortlist candidates that the exact kernel
then re-ranks.

Encoder (all kernels in csrc/tsm_nn.hip, reached through include/tsm_nn.h):
  x  = trajectory resampled to T = 16 points along its index, centred to [-1, 1]
  h1 = ReLU(Conv1d(2 -> 64, kernel 3, padding 1)(x))
  h0 = tanh(u * (log(len) - 4));  hT = GRU(64 -> 128)(h1, h0)
  e  = W_o hT + b_o + S x_flat          (128-d; S a fixed orthonormal 32 -> 128 map)
With W_o = 0 the embedding is S x_flat, whose L2 distance is the resampled
trajectories' L2 distance: the non-learned method.  Training learns the
correction W_o hT + b_o.

Weights are FP32 masters with BF16 copies for the GEMMs (FP32 accumulation
on the matrix cores).  Training minimises a pairwise ranking loss on anchor
lists: each anchor's candidates (its nearest by the current embedding) carry
exact DTW labels from the DP kernel, and every true top-10 candidate is
pushed closer than every candidate ranked 15th or worse.

Search: distances |e_d|^2 - 2 e_q . e_d through the MFMA GEMM with the
embeddings split into BF16 hi + lo parts (the query's hi part against the
database's hi and lo parts: two products in one K = 256 GEMM), the C smallest per query by a radix select, then
the exact DTW of those C pairs by the DP kernel and the k smallest of them.

The reference's learned kNN path (test_knn_time.py:123-231) embeds with a
PyTorch model, searches a faiss IVF index and has no exact re-rank; its
training (nn/MLP_trajsimi_train.py:38-144) regresses embedding distances on
ground-truth matrices computed by traj_dist (preprocessing/trajsimi.py:123-149).
"""
from __future__ import annotations

import ctypes
import time

import numpy as np
import torch

from . import _lib
from .data import TrajSet, lengths_lognormal, random_walks
from .engine import DeviceTrajSet, pairs as dp_pairs

__all__ = ["Encoder", "FilterIndex", "exact_knn", "filtered_knn", "recall_at", "train_encoder", "hard_pairs",
           "random_pairs", "train_filter", "knn_workload", "DEFAULT_C"]

T = 16
C_CONV = 64
H = 128
D = 128
G3 = 3 * H
# candidates per query of the filtered kNN: the smallest count at which the
# trained encoder keeps recall@10 >= 0.95 on the 1,000 x 1,000,000 workload
# and the untrained one does not (0.956 against 0.938)
DEFAULT_C = 100


def _nn():
    L = _lib.lib()
    if not getattr(L, "_nn_declared", False):
        P, I64 = ctypes.c_void_p, ctypes.c_int64
        i, f = ctypes.c_int, ctypes.c_float
        L.tsm_nn_gemm_nt.argtypes = [i, i, i, P, i, P, i, P, i, f, f, P, P]
        L.tsm_nn_cast_bf16.argtypes = [P, i, i, i, P, i, i, P]
        L.tsm_nn_resample.argtypes = [P, P, P, i, i, P, P, P]
        L.tsm_nn_conv_fwd.argtypes = [P, i, i, i, P, P, P, P, P]
        L.tsm_nn_conv_bwd.argtypes = [P, P, P, i, i, i, P, P, P]
        L.tsm_nn_gru_h0.argtypes = [P, P, i, i, P, P, P]
        L.tsm_nn_gru_h0_bwd.argtypes = [P, P, P, i, i, P, P]
        L.tsm_nn_gru_step_fwd.argtypes = [P, I64, P, P, i, i, P, P, P, P, P, P, P]
        L.tsm_nn_gru_step_bwd.argtypes = [P, P, P, P, P, P, i, i, P, I64, P, P, P, P]
        L.tsm_nn_colsum.argtypes = [P, I64, i, I64, P, f, P]
        L.tsm_nn_rank_loss.argtypes = [P, i, i, i, i, i, f, P, P, P]
        L.tsm_nn_dist_loss.argtypes = [P, P, i, i, i, P, P, P]
        L.tsm_nn_adam.argtypes = [P, P, P, P, I64, f, f, f, f, I64, P]
        L.tsm_nn_rownorm_bf16.argtypes = [P, I64, i, P, P]
        L.tsm_nn_topk_f32.argtypes = [P, I64, I64, I64, i, P, P, P]
        L.tsm_nn_topk_f64.argtypes = [P, I64, I64, I64, i, P, P, P]
        L.tsm_nn_linear_f32.argtypes = [P, I64, i, P, i, P, f, P]
        L.tsm_nn_split_bf16.argtypes = [P, I64, i, i, P, P, P]
        L.tsm_nn_gemm_nt_filter.argtypes = [i, i, i, P, i, P, i, f, P, P, i, P, P, P, P]
        L.tsm_nn_select_candidates.argtypes = [P, P, P, i, i, i, P, P, P]
        L.tsm_nn_select_candidates_pairs.argtypes = [P, P, P, i, i, i, P, P, P, P, I64, P, P, P]
        L.tsm_nn_gemm_nt_acc.argtypes = [i, i, i, P, i, P, i, P, i, f, P]
        L.tsm_nn_gru_forward.argtypes = [P, i, i, i, P, P, P, P, P, P, P, P, P, i, P]
        L.tsm_nn_kth_smallest_f32.argtypes = [P, I64, I64, I64, i, P, P]
        for n in ("gemm_nt", "cast_bf16", "resample", "conv_fwd", "conv_bwd", "gru_h0", "gru_h0_bwd",
                  "gru_step_fwd", "gru_step_bwd", "colsum", "rank_loss", "dist_loss", "adam", "rownorm_bf16", "topk_f32",
                  "topk_f64", "linear_f32", "split_bf16", "gemm_nt_filter", "select_candidates",
                  "select_candidates_pairs", "gemm_nt_acc", "gru_forward", "kth_smallest_f32"):
            getattr(L, "tsm_nn_" + n).restype = ctypes.c_int
        L._nn_declared = True
    return L


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _bf16(t: torch.Tensor, transpose: bool = False) -> torch.Tensor:
    """FP32 [r, c] -> BF16 bits (uint16 in an int16 tensor), optionally [c, r]."""
    r, c = t.shape
    out = torch.empty((c, r) if transpose else (r, c), dtype=torch.int16, device=t.device)
    _lib.check(_nn().tsm_nn_cast_bf16(_p(t), r, c, t.stride(0), _p(out), out.stride(0), int(transpose), _s()))
    return out


def gemm_nt(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, alpha=1.0, beta=0.0, bias=None):
    """C = alpha * A . B^T + beta * C (+ bias); A [M, K], B [N, K] bf16 bits, C fp32 [M, N]."""
    M, K = A.shape
    N = B.shape[0]
    _lib.check(_nn().tsm_nn_gemm_nt(M, N, K, _p(A), A.stride(0), _p(B), B.stride(0), _p(C), C.stride(0),
                                    float(alpha), float(beta), _p(bias), _s()))
    return C


def gemm_acc(A: torch.Tensor, B: torch.Tensor, C: torch.Tensor, alpha=1.0):
    """C += alpha * A . B^T with K split across blocks (weight gradients)."""
    M, K = A.shape
    N = B.shape[0]
    _lib.check(_nn().tsm_nn_gemm_nt_acc(M, N, K, _p(A), A.stride(0), _p(B), B.stride(0), _p(C), C.stride(0),
                                        float(alpha), _s()))
    return C


def _skip_matrix(seed: int = 7) -> np.ndarray:
    """Fixed 128 x 32 map with orthonormal columns scaled by sqrt(8 / T):
    |S x - S y| = sqrt(8 / T) |x - y| for resampled trajectories x, y."""
    g = np.random.default_rng(seed).standard_normal((D, 2 * T))
    q, _ = np.linalg.qr(g)
    return (q * np.sqrt(8.0 / T)).astype(np.float32)


# Flat parameter layout: name -> shape.
LAYOUT = (("Wc", (C_CONV, 2, 3)), ("bc", (C_CONV,)), ("W_ih", (G3, C_CONV)), ("b_ih", (G3,)),
          ("W_hh", (G3, H)), ("b_hh", (G3,)), ("u", (H,)), ("W_o", (D, H)), ("b_o", (D,)))


def _views(flat: torch.Tensor) -> dict:
    out, o = {}, 0
    for name, shape in LAYOUT:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].view(shape)
        o += n
    return out


NPARAM = sum(int(np.prod(s)) for _, s in LAYOUT)


def init_params(seed: int = 0) -> np.ndarray:
    """PyTorch-style uniform fan-in initialisation; the head W_o, b_o starts
    at zero so the untrained encoder is the resampled-L2 embedding."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    v = _views(torch.from_numpy(p))
    kc = 1.0 / np.sqrt(6.0)
    v["Wc"][:] = torch.from_numpy(rng.uniform(-kc, kc, (C_CONV, 2, 3)).astype(np.float32))
    v["bc"][:] = torch.from_numpy(rng.uniform(-kc, kc, C_CONV).astype(np.float32))
    kg = 1.0 / np.sqrt(H)
    for name in ("W_ih", "b_ih", "W_hh", "b_hh"):
        v[name][:] = torch.from_numpy(rng.uniform(-kg, kg, tuple(v[name].shape)).astype(np.float32))
    v["u"][:] = torch.from_numpy((rng.standard_normal(H) * 0.1).astype(np.float32))
    return p


class Encoder:
    """The encoder's parameters on the device, with forward and backward
    passes made of libtsm.so kernels."""

    def __init__(self, params: np.ndarray | None = None, device="cuda"):
        self.device = torch.device(device)
        p = init_params() if params is None else np.asarray(params, dtype=np.float32)
        if p.size != NPARAM:
            raise ValueError(f"expected {NPARAM} parameters, got {p.size}")
        self.flat = torch.tensor(p, device=self.device)
        self.w = _views(self.flat)
        self.S = torch.tensor(_skip_matrix(), device=self.device)
        self.refresh()

    def refresh(self):
        """Rebuild the BF16 weight copies after an update."""
        w = self.w
        self.W_ih_b = _bf16(w["W_ih"])               # [384, 64]  (B operand of x W_ih^T)
        self.W_ihT_b = _bf16(w["W_ih"], True)        # [64, 384]  (B operand of dGx W_ih)
        self.W_hh_b = _bf16(w["W_hh"])               # [384, 128]
        self.W_hhT_b = _bf16(w["W_hh"], True)        # [128, 384]
        self.W_o_b = _bf16(w["W_o"])                 # [128, 128]
        self.W_oT_b = _bf16(w["W_o"], True)

    def params(self) -> np.ndarray:
        return self.flat.cpu().numpy().copy()

    # -------------------------------------------------------------- forward
    def resample(self, dset: DeviceTrajSet, ids: torch.Tensor | None, n: int):
        X = torch.empty((n, T, 2), dtype=torch.float32, device=self.device)
        lenf = torch.empty(n, dtype=torch.float32, device=self.device)
        _lib.check(_nn().tsm_nn_resample(_p(dset.xy64), _p(dset.offsets), _p(ids), n, T, _p(X), _p(lenf), _s()))
        return X, lenf

    def forward(self, X: torch.Tensor, lenf: torch.Tensor, keep: bool = False):
        """Embeddings [n, 128] (FP32) of resampled trajectories X [n, T, 2].
        keep=True stores the activations for backward()."""
        L, w, n, dev = _nn(), self.w, X.shape[0], self.device
        s = _s()
        H1 = torch.empty((n * T, C_CONV), dtype=torch.float32, device=dev)
        H1b = torch.empty((n * T, C_CONV), dtype=torch.int16, device=dev)
        _lib.check(L.tsm_nn_conv_fwd(_p(X), n, T, C_CONV, _p(w["Wc"]), _p(w["bc"]), _p(H1), _p(H1b), s))
        Gx = torch.empty((n * T, G3), dtype=torch.float32, device=dev)
        gemm_nt(H1b, self.W_ih_b, Gx, bias=w["b_ih"])
        Hs = torch.empty((T + 1, n, H), dtype=torch.float32, device=dev)
        Hb = torch.empty((n, H), dtype=torch.int16, device=dev)
        _lib.check(L.tsm_nn_gru_h0(_p(w["u"]), _p(lenf), n, H, _p(Hs[0]), _p(Hb), s))
        Gh = torch.empty((n, G3), dtype=torch.float32, device=dev)
        if keep:
            R, Z, NN, GHN = (torch.empty((T, n, H), dtype=torch.float32, device=dev) for _ in range(4))
        else:
            R, Z, NN, GHN = (torch.empty((1, n, H), dtype=torch.float32, device=dev) for _ in range(4))
        # T steps of recurrent GEMM + gate kernel, launched from C
        _lib.check(L.tsm_nn_gru_forward(_p(Gx), T, n, H, _p(self.W_hh_b), _p(w["b_hh"]), _p(Hs), _p(Hb), _p(Gh),
                                        _p(R), _p(Z), _p(NN), _p(GHN), int(keep), s))
        E = torch.empty((n, D), dtype=torch.float32, device=dev)
        gemm_nt(Hb, self.W_o_b, E, bias=w["b_o"])
        _lib.check(L.tsm_nn_linear_f32(_p(X), n, 2 * T, _p(self.S), D, _p(E), 1.0, s))  # fixed skip map, FP32
        if keep:
            self._cache = dict(X=X, lenf=lenf, H1=H1, H1b=H1b, Hs=Hs, R=R, Z=Z, NN=NN, GHN=GHN, n=n)
        return E

    # -------------------------------------------------------------- backward
    def backward(self, dE: torch.Tensor) -> torch.Tensor:
        """Gradient (flat, FP32) of the loss w.r.t. the parameters, given dE
        [n, 128] for the last forward(keep=True)."""
        c, L, w, dev = self._cache, _nn(), self.w, self.device
        n = c["n"]
        s = _s()
        grad = torch.zeros(NPARAM, dtype=torch.float32, device=dev)
        g = _views(grad)
        Hs = c["Hs"]
        # head: dW_o = dE^T hT, db_o = sum dE, dh = dE W_o
        gemm_acc(_bf16(dE, True), _bf16(Hs[T], True), g["W_o"])
        _lib.check(L.tsm_nn_colsum(_p(dE), n, D, D, _p(g["b_o"]), 0.0, s))
        dh = torch.empty((n, H), dtype=torch.float32, device=dev)
        gemm_nt(_bf16(dE), self.W_oT_b, dh)
        dGx = torch.empty((n * T, G3), dtype=torch.float32, device=dev)
        dGh = torch.empty((T, n, G3), dtype=torch.float32, device=dev)  # every step's, for one wgrad GEMM
        dGhb = torch.empty((n, G3), dtype=torch.int16, device=dev)
        for t in range(T - 1, -1, -1):
            dgx_t = dGx.view(n, T, G3)[:, t, :]
            _lib.check(L.tsm_nn_gru_step_bwd(_p(dh), _p(Hs[t]), _p(c["R"][t]), _p(c["Z"][t]), _p(c["NN"][t]),
                                             _p(c["GHN"][t]), n, H, _p(dgx_t), T * G3, _p(dGh[t]), _p(dGhb), _p(dh),
                                             s))
            # dh = dh * z + dGh W_hh  (the step wrote dh * z)
            gemm_nt(dGhb, self.W_hhT_b, dh, beta=1.0)
        _lib.check(L.tsm_nn_gru_h0_bwd(_p(dh), _p(Hs[0]), _p(c["lenf"]), n, H, _p(g["u"]), s))
        # recurrent weights over all steps at once: dW_hh = sum_t dGh_t^T h_t, db_hh = sum dGh
        dGh_all = dGh.view(T * n, G3)
        gemm_acc(_bf16(dGh_all, True), _bf16(Hs[:T].reshape(T * n, H), True), g["W_hh"])
        _lib.check(L.tsm_nn_colsum(_p(dGh_all), T * n, G3, G3, _p(g["b_hh"]), 0.0, s))
        # input projection: dW_ih = dGx^T h1, db_ih = sum dGx, dH1 = dGx W_ih
        gemm_acc(_bf16(dGx, True), _bf16(c["H1"], True), g["W_ih"])
        _lib.check(L.tsm_nn_colsum(_p(dGx), n * T, G3, G3, _p(g["b_ih"]), 0.0, s))
        dH1 = torch.empty((n * T, C_CONV), dtype=torch.float32, device=dev)
        gemm_nt(_bf16(dGx), self.W_ihT_b, dH1)
        _lib.check(L.tsm_nn_conv_bwd(_p(c["X"]), _p(c["H1"]), _p(dH1), n, T, C_CONV, _p(g["Wc"]), _p(g["bc"]), s))
        return grad

    # -------------------------------------------------------------- embedding
    def embed(self, dset: DeviceTrajSet, ids=None, batch: int = 65536) -> torch.Tensor:
        """FP32 embeddings [n, 128] of trajectories ids (default: all)."""
        if ids is None:
            ids = torch.arange(dset.n, dtype=torch.int32, device=self.device)
        ids = torch.as_tensor(ids, dtype=torch.int32, device=self.device)
        if 0 < ids.numel() <= batch:  # one batch: the forward pass's own output, no copy
            X, lenf = self.resample(dset, ids, ids.numel())
            return self.forward(X, lenf)
        out = torch.empty((ids.numel(), D), dtype=torch.float32, device=self.device)
        for o in range(0, ids.numel(), batch):
            sl = ids[o:o + batch]
            X, lenf = self.resample(dset, sl, sl.numel())
            out[o:o + sl.numel()] = self.forward(X, lenf)
        return out


class Adam:
    def __init__(self, n: int, device, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        self.m = torch.zeros(n, dtype=torch.float32, device=device)
        self.v = torch.zeros(n, dtype=torch.float32, device=device)
        self.lr, self.b1, self.b2, self.eps, self.step = lr, b1, b2, eps, 0

    def update(self, p: torch.Tensor, g: torch.Tensor):
        self.step += 1
        _lib.check(_nn().tsm_nn_adam(_p(p), _p(g), _p(self.m), _p(self.v), p.numel(), self.lr, self.b1, self.b2,
                                     self.eps, self.step, _s()))


def rank_loss(E: torch.Tensor, B: int, K: int, P: int = 10, Q0: int = 20, tau: float = 0.05):
    dE = torch.empty_like(E)
    loss = torch.empty(B, dtype=torch.float32, device=E.device)
    _lib.check(_nn().tsm_nn_rank_loss(_p(E), B, K, D, P, Q0, float(tau), _p(dE), _p(loss), _s()))
    return loss, dE


def dist_loss(E: torch.Tensor, Y: torch.Tensor, B: int, K: int):
    """Log-space regression of anchor-candidate embedding distances on the
    exact ones (Y [B, K] FP32); returns (per-list loss, dE)."""
    dE = torch.empty_like(E)
    loss = torch.empty(B, dtype=torch.float32, device=E.device)
    _lib.check(_nn().tsm_nn_dist_loss(_p(E), _p(Y), B, K, D, _p(dE), _p(loss), _s()))
    return loss, dE


def topk_smallest(V: torch.Tensor, k: int):
    """Indices (int32) and values of the k smallest entries of each row."""
    rows, cols = V.shape
    idx = torch.empty((rows, k), dtype=torch.int32, device=V.device)
    val = torch.empty((rows, k), dtype=V.dtype, device=V.device)
    f = _nn().tsm_nn_topk_f64 if V.dtype == torch.float64 else _nn().tsm_nn_topk_f32
    _lib.check(f(_p(V), rows, cols, V.stride(0), k, _p(idx), _p(val), _s()))
    return idx, val


def kth_smallest(V: torch.Tensor, k: int) -> torch.Tensor:
    """The k-th smallest value of each row (FP32), the largest of the k
    smallest: an LDS radix select for rows of up to 16,384 values, else the
    top-k's maximum."""
    rows, cols = V.shape
    if cols <= 16384:
        out = torch.empty(rows, dtype=torch.float32, device=V.device)
        _lib.check(_nn().tsm_nn_kth_smallest_f32(_p(V), rows, cols, V.stride(0), int(k), _p(out), _s()))
        return out
    _, vals = topk_smallest(V, k)
    return vals.max(dim=1).values.contiguous()


class FilterIndex:
    """A database embedded once (the index build), searched per query batch:
    BF16 hi/lo split embeddings so one K = 256 MFMA GEMM gives
    |e_d|^2 - 2 e_q . e_d to ~16 mantissa bits."""

    def __init__(self, enc: Encoder, dset: DeviceTrajSet, db_ids):
        self.enc, self.dset = enc, dset
        self.last_redo = 0
        self.db_ids = torch.as_tensor(db_ids, dtype=torch.int32, device=enc.device).contiguous()
        E = enc.embed(dset, self.db_ids)
        self.norm = torch.empty(E.shape[0], dtype=torch.float32, device=E.device)
        self.B = self._split(E, query_side=False, norm=self.norm)
        self._sample_norm = {}  # (stride, sample) -> norms of the threshold sample rows

    @staticmethod
    def _split(E: torch.Tensor, query_side: bool, norm=None) -> torch.Tensor:
        # database rows [hi, lo], query rows [hi, hi]: one K = 2D product gives
        # q_hi . (d_hi + d_lo), the database side to ~16 mantissa bits
        out = torch.empty((E.shape[0], 2 * D), dtype=torch.int16, device=E.device)
        _lib.check(_nn().tsm_nn_split_bf16(_p(E), E.shape[0], D, int(query_side), _p(out), _p(norm), _s()))
        return out

    def search(self, q_ids, C: int = DEFAULT_C, *, cap: int = 4096, sample: int = 16384, ks: int | None = None,
               defer_redo: bool = False, pairs: bool = False):
        """Candidate database positions [nq, C] by embedding distance (the C
        smallest |e_d|^2 - 2 e_q . e_d of each query, ascending).

        Large databases skip the [nq, nd] distance matrix: a strided sample of
        `sample` rows gives each query a threshold (its ks-th smallest sample
        distance, about ks * nd / sample database rows below it; by default
        ks aims at 4 C rows), one fused
        GEMM keeps only the distances below the threshold, and a per-row sort
        of those picks the C smallest.  A query whose list comes out shorter
        than C or longer than `cap` (rare by construction) is redone over the
        full distance row.

        defer_redo=True returns (candidates, flags, query operand) without
        that redo and without any host synchronisation: flagged rows hold
        column 0, and the caller redoes them (``redo_rows``) once it
        synchronises anyway.  pairs=True (with defer_redo) also returns the
        exact re-rank's pairs: (q_ids repeated C times, the candidates'
        database ids), both int32 [nq * C], written by the selection kernel
        on the filtered path."""
        if pairs and not defer_redo:
            raise ValueError("pairs=True needs defer_redo=True")
        Eq = self.enc.embed(self.dset, q_ids)
        A = self._split(Eq, query_side=True)
        nq, nd = A.shape[0], self.B.shape[0]
        if nd < 4 * sample or C > cap // 2:
            cand = self._search_full(A, C)
            if defer_redo:
                flag = torch.zeros(nq, dtype=torch.int32, device=A.device)
                if pairs:
                    return (cand, flag, A) + self._rerank_pairs(q_ids, cand)
                return cand, flag, A
            return cand
        stride = nd // sample
        if ks is None:
            # about 4 C rows below the threshold, and never fewer than 12
            # sample rows under it: the count below the ks-th sample order
            # statistic spreads like a Gamma(ks) variable, so a small ks
            # (ks = 3 at C = 50) leaves ~5 % of the queries short of C and
            # sends them to the full-row redo
            ks = max(int(round(4 * C * sample / nd)), 12)
        ks = max(1, min(ks, sample))
        Bs = self.B[::stride][:sample]
        ns = self._sample_norm.get((stride, sample))
        if ns is None:  # the sample rows' norms: index data, gathered once
            ns = self.norm[::stride][:sample].contiguous()
            self._sample_norm[(stride, sample)] = ns
        dist_s = torch.empty((nq, sample), dtype=torch.float32, device=A.device)
        gemm_nt(A, Bs, dist_s, alpha=-2.0, bias=ns)
        tau = kth_smallest(dist_s, ks)
        del dist_s
        cnt = torch.zeros(nq, dtype=torch.int32, device=A.device)
        key = torch.empty((nq, cap), dtype=torch.float32, device=A.device)
        idx = torch.empty((nq, cap), dtype=torch.int32, device=A.device)
        L = _nn()
        _lib.check(L.tsm_nn_gemm_nt_filter(nq, nd, 2 * D, _p(A), A.stride(0), _p(self.B), self.B.stride(0), -2.0,
                                           _p(self.norm), _p(tau), cap, _p(cnt), _p(key), _p(idx), _s()))
        cand = torch.empty((nq, C), dtype=torch.int32, device=A.device)
        flag = torch.empty(nq, dtype=torch.int32, device=A.device)
        if pairs:
            qi = torch.as_tensor(q_ids, dtype=torch.int32, device=A.device).contiguous()
            pa = torch.empty(nq * C, dtype=torch.int32, device=A.device)
            pb = torch.empty(nq * C, dtype=torch.int32, device=A.device)
            _lib.check(L.tsm_nn_select_candidates_pairs(_p(cnt), _p(key), _p(idx), nq, cap, C, _p(cand), _p(flag),
                                                        _p(qi), _p(self.db_ids), self.db_ids.numel(), _p(pa),
                                                        _p(pb), _s()))
            return cand, flag, A, pa, pb
        _lib.check(L.tsm_nn_select_candidates(_p(cnt), _p(key), _p(idx), nq, cap, C, _p(cand), _p(flag), _s()))
        if defer_redo:
            return cand, flag, A
        redo = torch.nonzero(flag).reshape(-1)
        self.last_redo = int(redo.numel())
        if self.last_redo:
            cand[redo] = self._search_full(A[redo].contiguous(), C)
        return cand

    def _rerank_pairs(self, q_ids, cand: torch.Tensor):
        """The exact re-rank's pairs of a candidate list [nq, C]: each query id
        C times, and the candidates' database ids."""
        qi = torch.as_tensor(q_ids, dtype=torch.int32, device=cand.device)
        return qi.repeat_interleave(cand.shape[1]), self.db_ids[cand.reshape(-1).long()]

    def _search_full(self, A: torch.Tensor, C: int, chunk: int = 1 << 20) -> torch.Tensor:
        """The C smallest of every full distance row (dense GEMM + radix select)."""
        nq, nd = A.shape[0], self.B.shape[0]
        cand = torch.empty((nq, C), dtype=torch.int32, device=A.device)
        rows = max(1, min(nq, chunk // max(nd // 1024, 1)))
        dist = torch.empty((rows, nd), dtype=torch.float32, device=A.device)
        for o in range(0, nq, rows):
            r = min(rows, nq - o)
            gemm_nt(A[o:o + r], self.B, dist[:r], alpha=-2.0, bias=self.norm)
            idx, _ = topk_smallest(dist[:r], C)
            cand[o:o + r] = idx
        return cand


def exact_knn(dset: DeviceTrajSet, q0: int, nq: int, d0: int, nd: int, k: int = 10, chunk_q: int | None = None):
    """Exact DTW k-nearest database positions [nq, k] (grid of the DP
    kernel, then per-row selection), and the elapsed seconds.  Queries go in
    chunks of about 2 GB of FP64 distances (at least 256 rows, so the
    per-row selection fills the GPU)."""
    from .engine import grid
    if chunk_q is None:
        chunk_q = max(256, (2 << 30) // (8 * max(nd, 1)))
    out = torch.empty((nq, k), dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for o in range(0, nq, chunk_q):
        r = min(chunk_q, nq - o)
        dist = grid("dtw", dset, q0 + o, r, d0, nd)
        idx, _ = topk_smallest(dist, k)
        out[o:o + r] = idx
        del dist
    torch.cuda.synchronize()
    return out, time.perf_counter() - t0


def _rerank(index: FilterIndex, qids: torch.Tensor, cand: torch.Tensor, k: int, pa=None, pb=None) -> torch.Tensor:
    """Exact DTW of each query against its candidates, the k smallest.  pa,
    pb: the pairs when the search already wrote them."""
    nq, C = cand.shape
    if pa is None:
        pa, pb = index._rerank_pairs(qids, cand)
    d = dp_pairs("dtw", index.dset, pa, pb, validate=False).view(nq, C)
    idx, _ = topk_smallest(d, k)
    return torch.gather(cand, 1, idx.long())


def filtered_knn(index: FilterIndex, q0: int, nq: int, k: int = 10, C: int = DEFAULT_C):
    """Embedding shortlist of C, exact DTW re-rank, k smallest: database
    positions [nq, k] and the elapsed seconds (index build excluded).  The
    whole path is queued without a host synchronisation; the rare queries
    whose shortlist needs the full distance row are redone after it."""
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    qids = torch.arange(q0, q0 + nq, dtype=torch.int32, device="cuda")
    cand, flag, A, pa, pb = index.search(qids, C, defer_redo=True, pairs=True)
    best = _rerank(index, qids, cand, k, pa, pb)
    redo = torch.nonzero(flag).reshape(-1)  # synchronises, after everything is queued
    index.last_redo = int(redo.numel())
    if index.last_redo:
        cr = index._search_full(A[redo].contiguous(), C)
        best[redo] = _rerank(index, qids[redo], cr, k)
    torch.cuda.synchronize()
    return best, time.perf_counter() - t0


def recall_at(found: torch.Tensor, truth: torch.Tensor) -> float:
    f = found.cpu().numpy()
    t = truth.cpu().numpy()
    return float(np.mean([len(set(f[i]) & set(t[i])) / t.shape[1] for i in range(t.shape[0])]))


def hard_pairs(enc: Encoder, dset: DeviceTrajSet, anchors, pool_ids, K: int = 200):
    """For each anchor: its K nearest pool trajectories by the current
    embedding (itself excluded), with exact DTW labels from the DP kernel,
    sorted by label.  Returns (anchor ids, candidate ids [NA, K], labels)."""
    anchors = torch.as_tensor(anchors, dtype=torch.int32, device="cuda")
    pool = torch.as_tensor(pool_ids, dtype=torch.int32, device="cuda")
    index = FilterIndex(enc, dset, pool)
    cand = index.search(anchors, K + 1)
    cid = pool[cand.long()]
    # drop the anchor itself when it is in the pool, keep K
    keep = torch.where(cid == anchors[:, None], torch.full_like(cid, -1), cid)
    cids = torch.empty((anchors.numel(), K), dtype=torch.int32, device="cuda")
    kk = keep.cpu().numpy()
    out = np.empty((anchors.numel(), K), dtype=np.int32)
    for i in range(kk.shape[0]):
        row = kk[i][kk[i] >= 0][:K]
        out[i] = row
    cids.copy_(torch.from_numpy(out))
    pa = anchors.repeat_interleave(K)
    y = dp_pairs("dtw", dset, pa, cids.reshape(-1), validate=False).view(anchors.numel(), K)
    order = torch.argsort(y, dim=1)
    return anchors, torch.gather(cids, 1, order), torch.gather(y, 1, order)


def train_encoder(enc: Encoder, dset: DeviceTrajSet, anchors, cands, labels, epochs: int = 20, batch: int = 32,
                  lr: float = 1e-3, tau: float = 0.05, seed: int = 0, log=None, callback=None, P: int = 10,
                  Q0: int = 20, loss: str = "rank"):
    """Adam over anchor lists (cands sorted by label) on the ranking loss
    (``loss="rank"``) or the log-space distance regression (``"dist"``).
    ``callback(epoch)`` runs after every epoch; returns the mean loss per epoch."""
    if loss not in ("rank", "dist"):
        raise ValueError("loss must be 'rank' or 'dist'")
    NA, K = cands.shape
    y32 = labels.to(torch.float32).contiguous() if loss == "dist" else None
    opt = Adam(NPARAM, enc.device, lr=lr)
    g = torch.Generator(device="cpu").manual_seed(seed)
    hist = []
    for ep in range(epochs):
        perm = torch.randperm(NA, generator=g)
        tot = 0.0
        nb = 0
        for o in range(0, NA - batch + 1, batch):
            sel = perm[o:o + batch].to("cuda")
            ids = torch.cat([anchors[sel][:, None], cands[sel]], dim=1).reshape(-1).contiguous()
            X, lenf = enc.resample(dset, ids, ids.numel())
            E = enc.forward(X, lenf, keep=True)
            if loss == "dist":
                lb, dE = dist_loss(E, y32[sel].contiguous(), batch, K)
            else:
                lb, dE = rank_loss(E, batch, K, P=P, Q0=Q0, tau=tau)
            grad = enc.backward(dE)
            opt.update(enc.flat, grad)
            enc.refresh()
            tot += float(lb.mean())
            nb += 1
        hist.append(tot / max(nb, 1))
        if log:
            log(f"epoch {ep}: loss {hist[-1]:.5f}")
        if callback:
            callback(ep)
    return hist


def knn_workload(seed: int = 2024, nq: int = 1000, nd: int = 1_000_000) -> TrajSet:
    """The kNN workload: nq query trajectories (ids [0, nq)) followed by nd
    database trajectories (ids [nq, nq + nd)), random walks with lognormal
    lengths (median 60, sigma 0.8, clipped to [10, 1000]), from one seed."""
    rng = np.random.default_rng(seed)
    return random_walks(rng, lengths_lognormal(rng, nq + nd))


def random_pairs(dset: DeviceTrajSet, anchors, pool_ids, K: int, seed: int = 0):
    """For each anchor, K pool trajectories drawn uniformly at random (with
    replacement), labelled with exact DTW by the DP kernel and sorted by
    label: the random-pair labels of the reference's training data
    (preprocessing/trajsimi.py:123-149 samples its trajectories at random)."""
    anchors = torch.as_tensor(anchors, dtype=torch.int32, device="cuda")
    pool = torch.as_tensor(pool_ids, dtype=torch.int32, device="cuda")
    rng = np.random.default_rng(seed + 1)
    pick = torch.as_tensor(rng.integers(0, pool.numel(), (anchors.numel(), K)), device="cuda")
    cids = pool[pick].contiguous()
    y = dp_pairs("dtw", dset, anchors.repeat_interleave(K), cids.reshape(-1), validate=False).view(anchors.numel(), K)
    order = torch.argsort(y, dim=1)
    return anchors, torch.gather(cids, 1, order), torch.gather(y, 1, order)


# Training defaults, chosen by recall@10 over 1,000 held-out queries of the
# 1,000 x 1,000,000 kNN workload (200,000 labels in every configuration).
# 2,000 anchors x 100 neighbours, lr 3e-4, tau 0.1, negatives from rank 15:
# recall@10 at C = 100 candidates 0.956 against 0.938 untrained (C = 50:
# 0.907 / 0.883).  1,000 x 200 at lr 1e-3, tau 0.05, negatives from rank 20
# gave 0.944; lr 3e-3 or 8,000 x 25 anchors lost recall.  Random pairs (the
# reference's training data) give lists whose best entries are nowhere near
# a trajectory's 10 nearest of 1,000,000: trained on them, by the ranking
# loss or by the log-space distance regression (lr 3e-4 to 3e-3), recall@10
# at C = 100 falls to 0.06-0.10, and the regression on nearest lists to
# 0.80: only a loss on the near-neighbour order keeps and improves it.
TRAIN_ANCHORS, TRAIN_K, TRAIN_LR, TRAIN_TAU, TRAIN_P, TRAIN_Q0, TRAIN_EPOCHS = 2000, 100, 3e-4, 0.1, 10, 15, 15


def train_filter(dset: DeviceTrajSet, db_ids, n_anchors: int = TRAIN_ANCHORS, K: int = TRAIN_K,
                 epochs: int = TRAIN_EPOCHS, lr: float = TRAIN_LR, tau: float = TRAIN_TAU, batch: int = 32,
                 seed: int = 0, log=None, P: int = TRAIN_P, Q0: int = TRAIN_Q0, sampling: str = "nearest",
                 loss: str | None = None):
    """The filter's training run, all on the device: n_anchors database
    trajectories drawn at random, each with K database trajectories --
    its K nearest by the untrained (resampled-L2) embedding
    (``sampling="nearest"``) or K drawn at random (``"random"``) -- labelled
    with exact DTW by the DP kernel (n_anchors * K distances), then
    ``epochs`` of Adam on the ranking loss (the P best of each list pushed
    ahead of every candidate ranked Q0 or worse; ``loss="rank"``, the
    default with nearest lists) or on the log-space regression of embedding
    distances on the labels (``"dist"``, the default with random lists, as
    the reference fits random pairs' distances).  Returns (encoder, info)."""
    db_ids = torch.as_tensor(db_ids, dtype=torch.int32, device="cuda")
    if db_ids.numel() <= K:
        raise ValueError("the database must hold more than K trajectories")
    if sampling not in ("nearest", "random"):
        raise ValueError("sampling must be 'nearest' or 'random'")
    if loss is None:
        loss = "rank" if sampling == "nearest" else "dist"
    rng = np.random.default_rng(seed)
    pick = torch.as_tensor(rng.choice(db_ids.numel(), n_anchors, replace=False), device="cuda")
    anchors = db_ids[pick]
    enc = Encoder()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if sampling == "nearest":
        A, Cn, Y = hard_pairs(enc, dset, anchors, db_ids, K=K)
    else:
        A, Cn, Y = random_pairs(dset, anchors, db_ids, K=K, seed=seed)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    hist = train_encoder(enc, dset, A, Cn, Y, epochs=epochs, batch=batch, lr=lr, tau=tau, seed=seed, log=log, P=P,
                         Q0=Q0, loss=loss)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    return enc, dict(labels=int(Cn.numel()), anchors=n_anchors, K=K, epochs=epochs, sampling=sampling, loss=loss,
                     label_s=t1 - t0, train_s=t2 - t1, loss_first=hist[0], loss_last=hist[-1])
