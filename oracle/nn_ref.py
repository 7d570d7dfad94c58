"""TEST INFRASTRUCTURE ONLY -- FP64 numpy restatement of the learned kNN
filter's encoder and its two losses (ranking; log-space distance regression,
after the reference's MSE fit of embedding distances in
nn/MLP_trajsimi_train.py:78-82), forward and backward: the checker for the
HIP kernels of trajsimimeasures_amd/filter.py (csrc/tsm_nn.hip).  Only tests
import it.

Written from the encoder's equations (Conv1d(2->64, k=3, pad 1) + ReLU,
GRU(64->128) with gate order r, z, n and h0 = tanh(u * lenf), head W_o h_T +
b_o plus a fixed skip S x); the GRU convention is the common one
(r = sig(.), z = sig(.), n = tanh(Gx_n + r * Gh_n), h' = (1-z) n + z h).
The filter has no counterpart in the reference's kernels (its learned
measures are PyTorch modules, e.g. nn/MLP.py), so nothing here follows a
reference file.  Pinned by central finite differences of the loss through
forward() against backward() (tests/test_nn_ref.py), and by the closed form
of a zero head (embedding distance = resampled L2 distance)."""
from __future__ import annotations

import numpy as np

T, C, H, D = 16, 64, 128, 128


def unpack(flat: np.ndarray) -> dict:
    shapes = (("Wc", (C, 2, 3)), ("bc", (C,)), ("W_ih", (3 * H, C)), ("b_ih", (3 * H,)), ("W_hh", (3 * H, H)),
              ("b_hh", (3 * H,)), ("u", (H,)), ("W_o", (D, H)), ("b_o", (D,)))
    out, o = {}, 0
    for name, shape in shapes:
        n = int(np.prod(shape))
        out[name] = flat[o:o + n].reshape(shape).astype(np.float64)
        o += n
    return out


def pack(g: dict) -> np.ndarray:
    order = ("Wc", "bc", "W_ih", "b_ih", "W_hh", "b_hh", "u", "W_o", "b_o")
    return np.concatenate([g[k].reshape(-1) for k in order])


def sig(x):
    return 1.0 / (1.0 + np.exp(-x))


def bf16(a):
    """Round to the nearest BF16 value (ties to even), kept as float64."""
    u = np.ascontiguousarray(a, dtype=np.float64).astype(np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _id(a):
    return a


def forward(p: dict, X: np.ndarray, lenf: np.ndarray, S: np.ndarray, round_bf16: bool = False):
    """X [n, T, 2], lenf [n] -> E [n, D] and the activations.  round_bf16
    rounds every matrix-product operand to BF16 first (the GEMM inputs of the
    HIP kernels); products and everything else stay FP64."""
    q = bf16 if round_bf16 else _id
    n = X.shape[0]
    pre = np.zeros((n, T, C))
    for k in range(3):
        sh = k - 1
        xs = np.zeros_like(X)
        if sh < 0:
            xs[:, 1:] = X[:, :-1]
        elif sh > 0:
            xs[:, :-1] = X[:, 1:]
        else:
            xs = X
        pre += np.einsum("nti,ci->ntc", xs, p["Wc"][:, :, k])
    pre += p["bc"]
    H1 = np.maximum(pre, 0.0)
    Gx = q(H1) @ q(p["W_ih"]).T + p["b_ih"]
    h = np.tanh(lenf[:, None] * p["u"][None])
    hs, cache = [h], []
    for t in range(T):
        Gh = q(h) @ q(p["W_hh"]).T + p["b_hh"]
        gx = Gx[:, t]
        r = sig(gx[:, :H] + Gh[:, :H])
        z = sig(gx[:, H:2 * H] + Gh[:, H:2 * H])
        nn = np.tanh(gx[:, 2 * H:] + r * Gh[:, 2 * H:])
        cache.append((r, z, nn, Gh[:, 2 * H:]))
        h = (1 - z) * nn + z * h
        hs.append(h)
    E = q(h) @ q(p["W_o"]).T + p["b_o"] + X.reshape(n, -1) @ S.T
    return E, dict(X=X, lenf=lenf, pre=pre, H1=H1, hs=hs, cache=cache, q=q)


def backward(p: dict, act: dict, dE: np.ndarray) -> dict:
    X, hs, cache, H1, pre, q = act["X"], act["hs"], act["cache"], act["H1"], act["pre"], act["q"]
    n = X.shape[0]
    g = {k: np.zeros_like(v) for k, v in p.items()}
    hT = hs[T]
    g["W_o"] = q(dE).T @ q(hT)
    g["b_o"] = dE.sum(0)
    dh = q(dE) @ q(p["W_o"])
    dGx = np.zeros((n, T, 3 * H))
    for t in range(T - 1, -1, -1):
        r, z, nn, ghn = cache[t]
        h = hs[t]
        dnn = dh * (1 - z)
        dz = dh * (h - nn)
        dan = dnn * (1 - nn ** 2)
        dar = dan * ghn * r * (1 - r)
        daz = dz * z * (1 - z)
        dGx[:, t] = np.concatenate([dar, daz, dan], 1)
        dGh = np.concatenate([dar, daz, dan * r], 1)
        g["W_hh"] += q(dGh).T @ q(h)
        g["b_hh"] += dGh.sum(0)
        dh = dh * z + q(dGh) @ q(p["W_hh"])
    h0 = hs[0]
    g["u"] = (dh * (1 - h0 ** 2) * act["lenf"][:, None]).sum(0)
    flat_dGx = dGx.reshape(n * T, -1)
    g["W_ih"] = q(flat_dGx).T @ q(H1.reshape(n * T, -1))
    g["b_ih"] = flat_dGx.sum(0)
    dH1 = (q(flat_dGx) @ q(p["W_ih"])).reshape(n, T, C)
    dpre = dH1 * (pre > 0)
    g["bc"] = dpre.sum((0, 1))
    for k in range(3):
        sh = k - 1
        xs = np.zeros_like(X)
        if sh < 0:
            xs[:, 1:] = X[:, :-1]
        elif sh > 0:
            xs[:, :-1] = X[:, 1:]
        else:
            xs = X
        g["Wc"][:, :, k] = np.einsum("ntc,nti->ci", dpre, xs)
    return g


def rank_loss(E: np.ndarray, B: int, K: int, P: int, Q0: int, tau: float, scales=None):
    """Mean over anchors of mean softplus((d2_p - d2_q) / (tau * s_b)); s_b =
    mean_k d2_k held constant (pass `scales` to freeze them).  Returns
    (loss, dE, scales)."""
    Dm = E.shape[1]
    Eb = E.reshape(B, 1 + K, Dm)
    diff = Eb[:, 1:] - Eb[:, :1]
    d2 = (diff ** 2).sum(-1)
    if scales is None:
        scales = d2.mean(1)
    loss = 0.0
    g = np.zeros_like(d2)
    for b in range(B):
        sc = scales[b] * tau
        x = (d2[b, :P, None] - d2[b, None, Q0:]) / sc
        sp = np.logaddexp(0.0, x)
        loss += sp.mean() / B
        s = sig(x) / (P * (K - Q0) * B * sc)
        g[b, :P] += s.sum(1)
        g[b, Q0:] -= s.sum(0)
    dEb = np.zeros_like(Eb)
    dEb[:, 1:] = 2 * diff * g[:, :, None]
    dEb[:, 0] = -dEb[:, 1:].sum(1)
    return loss, dEb.reshape(E.shape), scales


def dist_loss(E: np.ndarray, Y: np.ndarray, B: int, K: int, consts=None):
    """Mean over anchors of mean_k (log(d_k + c) - log(s y_k + c))^2 with
    d_k = |e_k - e_a|, s = sum d / sum y and c = 1e-3 mean d per anchor,
    both held constant (pass `consts` = (s, c) arrays to freeze them).
    Returns (loss, dE, consts)."""
    Dm = E.shape[1]
    Eb = E.reshape(B, 1 + K, Dm)
    Y = np.asarray(Y, dtype=np.float64).reshape(B, K)
    diff = Eb[:, 1:] - Eb[:, :1]
    d = np.sqrt((diff ** 2).sum(-1))
    if consts is None:
        consts = (d.sum(1) / Y.sum(1), np.maximum(1e-3 * d.mean(1), 1e-12))
    s, c = consts
    r = np.log(d + c[:, None]) - np.log(s[:, None] * Y + c[:, None])
    loss = float((r ** 2).mean(1).mean())
    g = 2 * r / ((d + c[:, None]) * K * B * d)
    dEb = np.zeros_like(Eb)
    dEb[:, 1:] = diff * g[:, :, None]
    dEb[:, 0] = -dEb[:, 1:].sum(1)
    return loss, dEb.reshape(E.shape), consts
