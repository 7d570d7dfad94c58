"""Pins the FP64 reference of the encoder and its loss (oracle/nn_ref.py) by
central finite differences: a dropped term, a wrong gate or a transposed
weight in either the forward or the backward restatement breaks agreement."""
import numpy as np

from oracle import nn_ref as R


def _setup(seed=0, B=2, K=5):
    rng = np.random.default_rng(seed)
    n = B * (1 + K)
    X = rng.uniform(-1, 1, (n, R.T, 2))
    lenf = rng.uniform(-1.5, 2.0, n)
    S = np.linalg.qr(rng.standard_normal((R.D, 2 * R.T)))[0]
    sizes = [R.C * 6, R.C, 3 * R.H * R.C, 3 * R.H, 3 * R.H * R.H, 3 * R.H, R.H, R.D * R.H, R.D]
    flat = np.concatenate([rng.uniform(-0.3, 0.3, s) for s in sizes])
    return X, lenf, S, flat, B, K


def _loss(flat, X, lenf, S, B, K, scales):
    p = R.unpack(flat)
    E, _ = R.forward(p, X, lenf, S)
    return R.rank_loss(E, B, K, 1, 3, 0.5, scales)[0]


def test_rank_loss_and_backward_finite_differences():
    X, lenf, S, flat, B, K = _setup()
    p = R.unpack(flat)
    E, act = R.forward(p, X, lenf, S)
    loss, dE, scales = R.rank_loss(E, B, K, 1, 3, 0.5)
    g = R.pack(R.backward(p, act, dE))
    rng = np.random.default_rng(1)
    # offsets of each tensor in the flat vector; probe 6 entries of each
    sizes = [R.C * 6, R.C, 3 * R.H * R.C, 3 * R.H, 3 * R.H * R.H, 3 * R.H, R.H, R.D * R.H, R.D]
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    h = 1e-6
    for s0, sz in zip(starts, sizes):
        for i in s0 + rng.choice(sz, 6, replace=False):
            e = np.zeros_like(flat)
            e[i] = h
            fd = (_loss(flat + e, X, lenf, S, B, K, scales) - _loss(flat - e, X, lenf, S, B, K, scales)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, (i, fd, g[i])


def test_forward_zero_head_is_resampled_l2():
    X, lenf, S, flat, B, K = _setup(2)
    p = R.unpack(flat)
    p["W_o"][:] = 0
    p["b_o"][:] = 0
    E, _ = R.forward(p, X, lenf, S)
    # orthonormal columns: embedding distances equal the resampled L2 distances
    d_e = np.linalg.norm(E[0] - E[1])
    d_x = np.linalg.norm(X[0].reshape(-1) - X[1].reshape(-1))
    assert abs(d_e - d_x) < 1e-12


def test_dist_loss_finite_differences():
    """d loss / d E of the distance regression by central differences in E
    (s and c frozen, as the kernel holds them), and the closed form of a
    perfect fit: embedding distances proportional to the labels give 0."""
    rng = np.random.default_rng(4)
    B, K = 3, 6
    E = rng.standard_normal((B * (1 + K), R.D))
    Y = rng.uniform(0.5, 3.0, (B, K))
    loss, dE, consts = R.dist_loss(E, Y, B, K)
    h = 1e-6
    for _ in range(40):
        i, j = rng.integers(0, E.shape[0]), rng.integers(0, R.D)
        e = np.zeros_like(E)
        e[i, j] = h
        fd = (R.dist_loss(E + e, Y, B, K, consts)[0] - R.dist_loss(E - e, Y, B, K, consts)[0]) / (2 * h)
        assert abs(fd - dE[i, j]) <= 1e-6 * max(1.0, abs(fd)) + 1e-10, (i, j, fd, dE[i, j])
    # anchors at 0, candidates at distance 2 y_k along axis 0: s = 2, residual 0
    E0 = np.zeros((B * (1 + K), R.D))
    for b in range(B):
        E0[b * (1 + K) + 1: (b + 1) * (1 + K), 0] = 2.0 * Y[b]
    l0, dE0, _ = R.dist_loss(E0, Y, B, K)
    assert l0 < 1e-24 and np.abs(dE0).max() < 1e-12
