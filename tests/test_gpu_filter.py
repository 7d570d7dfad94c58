"""GPU tests of the learned kNN filter kernels (include/tsm_nn.h): the BF16
MFMA GEMM against an FP64 product of the same BF16 values, the casts, the
top-k selection against a sort, the resampler, and the encoder's forward and
backward passes against the FP64 numpy restatement in oracle/nn_ref.py."""
import numpy as np
import pytest

from oracle import nn_ref as R

pytestmark = pytest.mark.gpu


def _bf16_round(a: np.ndarray) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), as float64."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16 << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


@pytest.mark.parametrize("M,N,K", [(64, 64, 32), (37, 90, 40), (130, 17, 136), (1, 300, 384), (200, 129, 8),
                                   (128, 128, 64), (300, 257, 136), (1000, 1100, 384), (129, 4000, 72)])
def test_gemm_nt_matches_fp64_product(gpu, M, N, K):
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(M * 1000 + N + K)
    a = rng.standard_normal((M, K)).astype(np.float32)
    b = rng.standard_normal((N, K)).astype(np.float32)
    c0 = rng.standard_normal((M, N)).astype(np.float32)
    bias = rng.standard_normal(N).astype(np.float32)
    A = F._bf16(torch.tensor(a, device="cuda"))
    B = F._bf16(torch.tensor(b, device="cuda"))
    for alpha, beta, use_bias in ((1.0, 0.0, False), (-2.0, 0.5, True)):
        C = torch.tensor(c0, device="cuda")
        F.gemm_nt(A, B, C, alpha=alpha, beta=beta, bias=torch.tensor(bias, device="cuda") if use_bias else None)
        ab, bb = _bf16_round(a), _bf16_round(b)
        ref = alpha * ab @ bb.T + beta * c0.astype(np.float64) + (bias if use_bias else 0.0)
        # FP32 accumulation of exact BF16 products: error bounded by K ulps of the absolute sum
        bound = 2e-6 * K * (np.abs(ab) @ np.abs(bb).T) * abs(alpha) + 1e-5
        err = np.abs(C.cpu().numpy() - ref)
        assert np.all(err <= bound), float(err.max())


def test_cast_bf16_rounds_to_nearest_even(gpu):
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(3)
    a = (rng.standard_normal((33, 70)) * 10.0 ** rng.integers(-3, 3, (33, 70))).astype(np.float32)
    # exact ties between two BF16 values round to the even one
    a[0, :4] = np.array([0x3F808000, 0x3F818000, 0xBF808000, 0x3F80C000], dtype=np.uint32).view(np.float32)
    t = torch.tensor(a, device="cuda")
    want = t.to(torch.bfloat16).view(torch.int16).cpu().numpy()
    got = F._bf16(t).cpu().numpy()
    assert np.array_equal(got, want)
    got_t = F._bf16(t, transpose=True).cpu().numpy()
    assert np.array_equal(got_t, want.T)


@pytest.mark.parametrize("dtype", ["float32", "float64"])
@pytest.mark.parametrize("rows,cols,k", [(3, 1000, 10), (5, 70000, 200), (2, 257, 257), (4, 5000, 1),
                                         (1000, 200, 10), (7, 512, 32), (5, 4, 1), (9, 63, 5), (6, 300, 31)])
def test_topk_smallest_matches_sort(gpu, dtype, rows, cols, k):
    """The k smallest of each row (as a set, with their values): the radix
    top-k and, for rows of at most 512 columns and k <= 32, one wave per row
    popping wave minima."""
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(cols + k)
    v = rng.standard_normal((rows, cols)).astype(dtype)
    v[0, ::7] = v[0, 3]           # many ties
    v[1 % rows, :50] = -np.abs(v[1 % rows, :50])
    idx, val = F.topk_smallest(torch.tensor(v, device="cuda"), k)
    idx, val = idx.cpu().numpy(), val.cpu().numpy()
    for r in range(rows):
        assert len(set(idx[r].tolist())) == k
        assert np.array_equal(v[r, idx[r]], val[r])
        assert np.array_equal(np.sort(val[r]), np.sort(v[r])[:k])


@pytest.mark.parametrize("rows,cols,k", [(7, 16384, 13), (300, 16384, 1), (5, 1000, 1000), (64, 4097, 200),
                                         (1000, 16384, 32), (20, 16384, 33), (9, 100, 32), (3, 1, 1),
                                         (5, 16383, 17)])
def test_kth_smallest_matches_kthvalue(gpu, rows, cols, k):
    """Both selects (k <= 32: per-lane sorts and wave-minimum merges; else
    the LDS radix select) give torch.kthvalue's value, with negative values,
    exact ties and duplicated rows in the data, and rows narrower than one
    key per lane."""
    import torch

    from trajsimimeasures_amd import filter as F
    g = torch.Generator(device="cuda").manual_seed(rows * cols + k)
    V = torch.randn((rows, cols), device="cuda", generator=g)
    V[:, ::7] = torch.round(V[:, ::7] * 4) / 4      # many exact ties
    V[0] = V[1]                                     # duplicate rows
    V[-1, :] = 2.5                                  # a constant row
    got = F.kth_smallest(V, k)
    want = torch.kthvalue(V, k, dim=1).values
    assert torch.equal(got, want)


def test_resample_matches_numpy(gpu):
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(11)
    lens = lengths_lognormal(rng, 300)
    lens[:3] = [1, 2, 16]
    ts = random_walks(rng, lens)
    dset = DeviceTrajSet(ts)
    enc = F.Encoder()
    ids = torch.tensor(rng.permutation(300)[:120].astype(np.int32), device="cuda")
    X, lenf = enc.resample(dset, ids, 120)
    X, lenf = X.cpu().numpy(), lenf.cpu().numpy()
    for r, i in enumerate(ids.cpu().numpy()):
        tr = ts.xy[ts.offsets[i]:ts.offsets[i + 1]]
        n = len(tr)
        s = np.linspace(0.0, n - 1, F.T)
        want = np.stack([np.interp(s, np.arange(n), tr[:, 0]), np.interp(s, np.arange(n), tr[:, 1])], 1)
        assert np.allclose(X[r], (want - 0.5) * 2.0, rtol=0, atol=2e-7)
        assert abs(lenf[r] - (np.log(n) - 4.0)) < 1e-6


def _cos(a, b):
    return float(np.dot(a.ravel(), b.ravel()) / (np.linalg.norm(a) * np.linalg.norm(b) + 1e-300))


def _rel(a, b):
    return float(np.linalg.norm(a - b) / (np.linalg.norm(b) + 1e-300))


def test_encoder_forward_and_gradients_match_fp64_reference(gpu):
    """Forward embedding, ranking-loss gradient and every parameter gradient
    of the HIP encoder against the FP64 reference on the same inputs: tightly
    against the reference whose matrix-product operands are rounded to BF16
    as the kernels' are, loosely (about 1%) against the unrounded one.  A
    wrong gate, index or transposed weight is off by order one in both."""
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(21)
    B, K, P, Q0 = 2, 15, 3, 6
    n = B * (1 + K)
    ts = random_walks(rng, lengths_lognormal(rng, n))
    dset = DeviceTrajSet(ts)
    p = F.init_params(4)
    v = F._views(torch.from_numpy(p))
    v["W_o"][:] = torch.from_numpy(rng.uniform(-0.1, 0.1, (F.D, F.H)).astype(np.float32))
    v["b_o"][:] = torch.from_numpy(rng.uniform(-0.1, 0.1, F.D).astype(np.float32))
    enc = F.Encoder(p)
    X, lenf = enc.resample(dset, None, n)
    E = enc.forward(X, lenf, keep=True)
    loss, dE = F.rank_loss(E, B, K, P=P, Q0=Q0, tau=0.5)
    grad = enc.backward(dE).cpu().numpy().astype(np.float64)
    torch.cuda.synchronize()

    pr = R.unpack(p.astype(np.float64))
    S = enc.S.cpu().numpy().astype(np.float64)
    Xr = X.cpu().numpy().astype(np.float64)
    lf = lenf.cpu().numpy().astype(np.float64)
    Eg = E.cpu().numpy().astype(np.float64)
    skip = Xr.reshape(n, -1) @ S.T
    dEg = dE.cpu().numpy().astype(np.float64)

    # loss and dE: FP32 kernel vs FP64 on the GPU's own embeddings
    lr_, dEr, _ = R.rank_loss(Eg, B, K, P, Q0, 0.5)
    assert abs(float(loss.mean()) - lr_) <= 1e-4 * abs(lr_) + 1e-7
    assert _rel(dEg, dEr) < 1e-4

    for round_bf16, tol_e, tol_g, min_cos in ((True, 1e-4, 2e-3, 0.99999), (False, 1e-2, 6e-2, 0.995)):
        Er, act = R.forward(pr, Xr, lf, S, round_bf16=round_bf16)
        assert _rel(Eg - skip, Er - skip) < tol_e, (round_bf16, _rel(Eg - skip, Er - skip))
        gr = R.pack(R.backward(pr, act, dEg))
        o = 0
        for name, shape in F.LAYOUT:
            m = int(np.prod(shape))
            a, b = grad[o:o + m], gr[o:o + m]
            o += m
            if name == "b_o":
                # the loss depends on embedding differences only, so its
                # gradient w.r.t. a shared offset is zero
                bound = 1e-5 * np.abs(dEg).sum(0).max()
                assert np.abs(a).max() <= bound and np.abs(b).max() <= bound
                continue
            assert _cos(a, b) > min_cos, (round_bf16, name, _cos(a, b))
            assert _rel(a, b) < tol_g, (round_bf16, name, _rel(a, b))


def test_adam_step_matches_formula(gpu):
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(2)
    p0 = rng.standard_normal(1000).astype(np.float32)
    g = [rng.standard_normal(1000).astype(np.float32) for _ in range(3)]
    pt = torch.tensor(p0, device="cuda")
    opt = F.Adam(1000, "cuda", lr=0.01)
    for gi in g:
        opt.update(pt, torch.tensor(gi, device="cuda"))
    m = np.zeros(1000)
    v = np.zeros(1000)
    p = p0.astype(np.float64)
    for t, gi in enumerate(g, 1):
        m = 0.9 * m + 0.1 * gi
        v = 0.999 * v + 0.001 * gi.astype(np.float64) ** 2
        p -= 0.01 * (m / (1 - 0.9 ** t)) / (np.sqrt(v / (1 - 0.999 ** t)) + 1e-8)
    assert np.allclose(pt.cpu().numpy(), p, rtol=1e-5, atol=1e-6)


def test_split_bf16_search_distances(gpu):
    """|e_d|^2 - 2 e_q . e_d through the hi/lo split GEMM (the query rounded
    to BF16, the database to ~16 mantissa bits as hi + lo) agrees with the
    FP64 value of |e_d|^2 - 2 bf16(e_q) . e_d to about 2^-16 of the scale;
    the rounding of e_q shifts every distance of a row by the same
    -2 (e_q - bf16(e_q)) . e_d term, small between near candidates."""
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(5)
    Eq = rng.standard_normal((40, F.D)).astype(np.float32)
    Ed = rng.standard_normal((300, F.D)).astype(np.float32)
    norm = torch.empty(300, dtype=torch.float32, device="cuda")
    Bm = F.FilterIndex._split(torch.tensor(Ed, device="cuda"), False, norm)
    Am = F.FilterIndex._split(torch.tensor(Eq, device="cuda"), True)
    dist = torch.empty((40, 300), dtype=torch.float32, device="cuda")
    F.gemm_nt(Am, Bm, dist, alpha=-2.0, bias=norm)
    eq = torch.tensor(Eq).to(torch.bfloat16).to(torch.float64).numpy()
    ed = Ed.astype(np.float64)
    want = (ed ** 2).sum(1)[None] - 2.0 * eq @ ed.T
    scale = (eq ** 2).sum(1)[:, None] + (ed ** 2).sum(1)[None]
    assert np.all(np.abs(dist.cpu().numpy() - want) <= 5e-5 * scale)


@pytest.mark.parametrize("K", [256, 384])
@pytest.mark.parametrize("M,N", [(1, 1000), (100, 3000), (128, 70001), (300, 20000), (1000, 70001), (1300, 9000)])
def test_search_filter_gemm_matches_dense(gpu, M, N, K):
    """The fused filter GEMM (K = 256 or 384: the persistent search kernel)
    appends exactly the (column, value) pairs of the dense GEMM that lie
    below each row's threshold, with bit-identical values (same MFMA
    accumulation order)."""
    import torch

    from trajsimimeasures_amd import _lib
    from trajsimimeasures_amd import filter as F
    print(f"[case] search filter gemm M={M} N={N} K={K}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(M + N)
    A = torch.randn((M, K), device="cuda", generator=g).to(torch.bfloat16)
    B = torch.randn((N, K), device="cuda", generator=g).to(torch.bfloat16)
    bias = torch.randn(N, device="cuda", generator=g)
    dense = torch.empty((M, N), dtype=torch.float32, device="cuda")
    F.gemm_nt(A.view(torch.int16), B.view(torch.int16), dense, alpha=-2.0, bias=bias)
    k = max(1, min(N // 50, 300))
    tau = torch.kthvalue(dense, k, dim=1).values.contiguous()  # about k values per row below tau
    cap = 4096
    cnt = torch.zeros(M, dtype=torch.int32, device="cuda")
    key = torch.empty((M, cap), dtype=torch.float32, device="cuda")
    idx = torch.empty((M, cap), dtype=torch.int32, device="cuda")
    L = F._nn()
    P = lambda t: t.data_ptr()
    _lib.check(L.tsm_nn_gemm_nt_filter(M, N, K, P(A), K, P(B), K, -2.0, P(bias), P(tau), cap, P(cnt), P(key),
                                       P(idx), F._s()))
    torch.cuda.synchronize()
    d, t, c = dense.cpu().numpy(), tau.cpu().numpy(), cnt.cpu().numpy()
    kk, ii = key.cpu().numpy(), idx.cpu().numpy()
    for r in range(M):
        want = np.nonzero(d[r] < t[r])[0]
        assert c[r] == want.size, (r, c[r], want.size)
        order = np.argsort(ii[r, :c[r]])
        assert np.array_equal(ii[r, :c[r]][order], want)
        assert np.array_equal(kk[r, :c[r]][order], d[r, want])


def test_threshold_search_matches_full_search(gpu):
    """The fused threshold search (sampled threshold, filtered GEMM, per-row
    sort) returns the same C nearest database rows as the full distance rows
    with a radix select, including queries forced onto the fallback (lists
    too short with ks = 1, overflowing with ks = 4000) and long lists
    (ks = 470).  By default the
    threshold aims at 4 C candidates per query, so fallbacks are rare."""
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(12)
    nq, nd = 300, 70000
    dset = DeviceTrajSet(random_walks(rng, lengths_lognormal(rng, nq + nd)))
    enc = F.Encoder()
    index = F.FilterIndex(enc, dset, torch.arange(nq, nq + nd, dtype=torch.int32, device="cuda"))
    q = torch.arange(nq, dtype=torch.int32, device="cuda")
    A = index._split(enc.embed(dset, q), query_side=True)
    full = index._search_full(A, 200).cpu().numpy()
    # ks = 470: ~2,000 survivors per query
    for ks, expect_redo in ((None, False), (1, True), (4000, True), (470, False)):
        got = index.search(q, 200, ks=ks).cpu().numpy()
        same = np.mean([len(set(got[i]) & set(full[i])) == 200 for i in range(nq)])
        assert same == 1.0, (ks, same)
        assert (index.last_redo > 0) == expect_redo or (not expect_redo and index.last_redo < 3)


@pytest.mark.parametrize("n", [1, 15, 16, 17, 1000, 4099])
def test_fused_gru_forward_matches_stepwise(gpu, n):
    """The inference GRU (one fused kernel for all T steps, keep=False) gives
    bit-identical embeddings to the stepwise recurrent GEMM + gate kernels
    (keep=True), for row counts around its 16-row blocks."""
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(n)
    dset = DeviceTrajSet(random_walks(rng, lengths_lognormal(rng, n)))
    p = F.init_params(seed=3)
    p += np.random.default_rng(7).normal(0, 0.05, p.size).astype(np.float32)  # every weight nonzero
    enc = F.Encoder(p)
    ids = torch.arange(n, dtype=torch.int32, device="cuda")
    X, lenf = enc.resample(dset, ids, n)
    e_fused = enc.forward(X, lenf, keep=False).clone()
    e_step = enc.forward(X, lenf, keep=True).clone()
    assert torch.equal(e_fused, e_step)
    assert torch.isfinite(e_fused).all() and float(e_fused.abs().max()) > 0


def test_deferred_redo_matches_synchronous_search(gpu):
    """search(defer_redo=True) queues the path without a host sync: rows
    whose survivor list is short or overflowed come back flagged and hold
    column 0 (a valid index), the others equal the synchronous search; the
    flagged rows redone over their full distance rows give the synchronous
    search's redo (as sets: that select has no fixed order), and
    filtered_knn (which defers) finds the same 10 as a re-rank of the
    synchronous candidates."""
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(13)
    nq, nd = 200, 70000
    dset = DeviceTrajSet(random_walks(rng, lengths_lognormal(rng, nq + nd)))
    index = F.FilterIndex(F.Encoder(), dset, torch.arange(nq, nq + nd, dtype=torch.int32, device="cuda"))
    q = torch.arange(nq, dtype=torch.int32, device="cuda")
    for ks in (1, 4000, None):
        sync = index.search(q, 200, ks=ks).cpu().numpy()
        cand, flag, A = index.search(q, 200, ks=ks, defer_redo=True)
        cand, fl = cand.cpu().numpy(), flag.cpu().numpy()
        if ks is not None:
            assert fl.sum() > 0, ks
        assert np.all(cand[fl == 1] == 0)
        assert np.array_equal(cand[fl == 0], sync[fl == 0])
        redo = np.nonzero(fl)[0]
        if redo.size:  # the full-row select returns the C nearest in no fixed order
            full = index._search_full(A[torch.as_tensor(redo, device="cuda")].contiguous(), 200).cpu().numpy()
            assert all(set(full[i]) == set(sync[r]) for i, r in enumerate(redo))
    best, _ = F.filtered_knn(index, 0, nq, k=10, C=200)
    ref = F._rerank(index, q, torch.as_tensor(sync, device="cuda"), 10)
    b, r = best.cpu().numpy(), ref.cpu().numpy()
    assert all(set(b[i]) == set(r[i]) for i in range(nq))


def test_filtered_knn_small_recall(gpu):
    """End to end on a small set: the untrained filter (resampled-L2
    embedding) shortlists candidates whose exact DTW re-rank recovers most of
    the exact 10-NN; with C = database size the recall is exactly 1."""
    import torch

    from trajsimimeasures_amd import filter as F
    from trajsimimeasures_amd.data import lengths_lognormal, random_walks
    from trajsimimeasures_amd.engine import DeviceTrajSet
    rng = np.random.default_rng(8)
    nq, nd = 20, 4000
    ts = random_walks(rng, lengths_lognormal(rng, nq + nd))
    dset = DeviceTrajSet(ts)
    truth, _ = F.exact_knn(dset, 0, nq, nq, nd, k=10)
    enc = F.Encoder()
    index = F.FilterIndex(enc, dset, torch.arange(nq, nq + nd, dtype=torch.int32, device="cuda"))
    found, _ = F.filtered_knn(index, 0, nq, k=10, C=nd)
    assert F.recall_at(found, truth) == 1.0
    found, _ = F.filtered_knn(index, 0, nq, k=10, C=200)
    assert F.recall_at(found, truth) > 0.8


def test_dist_loss_kernel_matches_reference(gpu):
    """The distance-regression loss kernel against its FP64 restatement
    (itself pinned by finite differences): loss and dE within 1e-4."""
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(5)
    for B, K in ((4, 100), (3, 7), (2, 1024)):
        E = rng.standard_normal((B * (1 + K), F.D)).astype(np.float32)
        Y = rng.uniform(0.01, 5.0, (B, K)).astype(np.float32)
        loss, dE = F.dist_loss(torch.as_tensor(E, device="cuda"), torch.as_tensor(Y, device="cuda"), B, K)
        lr_, dEr, _ = R.dist_loss(E.astype(np.float64), Y.astype(np.float64), B, K)
        assert abs(float(loss.mean()) - lr_) <= 1e-4 * abs(lr_) + 1e-7, (B, K)
        assert _rel(dE.cpu().numpy().astype(np.float64), dEr) < 1e-4, (B, K)


@pytest.mark.parametrize("C", [10, 100, 256, 300])
def test_select_candidates_matches_sort(gpu, C):
    """tsm_nn_select_candidates against a lexicographic sort of (value,
    column): rows of C - 1 and cap + 1 keys flagged, rows of C .. cap keys
    (the radix select when C <= 256, the bitonic sort at C = 300) giving
    their C smallest ascending; values repeat (ties broken by column) and
    include negatives."""
    import torch

    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(23 + C)
    cap = 4096
    counts = [C - 1, C, C + 1, 511, 512, 513, 700, 1000, 2048, 3000, cap, cap + 1]
    counts = [c for c in counts if c >= C - 1]
    nq = len(counts)
    key = np.zeros((nq, cap), dtype=np.float32)
    idx = np.zeros((nq, cap), dtype=np.int32)
    for r, c in enumerate(counts):
        m = min(c, cap)
        v = rng.normal(0.0, 1.0, m).astype(np.float32)
        v[rng.random(m) < 0.3] = np.float32(0.25)  # a large group of equal values
        v[rng.random(m) < 0.1] = np.float32(-0.5)
        key[r, :m] = v
        idx[r, :m] = rng.permutation(1_000_000)[:m].astype(np.int32)
    dev = "cuda"
    cnt_t = torch.as_tensor(np.array(counts, dtype=np.uint32).view(np.int32), device=dev)
    key_t = torch.as_tensor(key, device=dev)
    idx_t = torch.as_tensor(idx, device=dev)
    out = torch.full((nq, C), -7, dtype=torch.int32, device=dev)
    flag = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    from trajsimimeasures_amd import _lib
    _lib.check(F._nn().tsm_nn_select_candidates(F._p(cnt_t), F._p(key_t), F._p(idx_t), nq, cap, C, F._p(out),
                                                F._p(flag), F._s()))
    out, flag = out.cpu().numpy(), flag.cpu().numpy()
    for r, c in enumerate(counts):
        if c < C or c > cap:
            assert flag[r] == 1 and np.all(out[r] == 0), (r, c)
            continue
        assert flag[r] == 0, (r, c)
        order = np.lexsort((idx[r, :c], key[r, :c]))[:C]
        assert np.array_equal(out[r], idx[r, order]), (C, c)


@pytest.mark.parametrize("C", [100, 300])
def test_select_candidates_pairs(gpu, C):
    """tsm_nn_select_candidates_pairs: out and flag as the plain entry point
    gives them, and the re-rank pairs pa = row id repeated C times, pb = the
    column's id (flagged rows: column 0's id; a column past the id table:
    -1), against the torch construction the search used to do."""
    import torch

    from trajsimimeasures_amd import _lib
    from trajsimimeasures_amd import filter as F
    rng = np.random.default_rng(91 + C)
    cap, ncols = 4096, 5000
    counts = [C - 1, C, 700, 3000, cap + 1, 900]
    nq = len(counts)
    key = np.zeros((nq, cap), dtype=np.float32)
    idx = np.zeros((nq, cap), dtype=np.int32)
    for r, c in enumerate(counts):
        m = min(c, cap)
        v = rng.normal(0.0, 1.0, m).astype(np.float32)
        v[rng.random(m) < 0.2] = np.float32(0.5)
        key[r, :m] = v
        idx[r, :m] = rng.permutation(ncols)[:m].astype(np.int32)
    idx[5, 0], key[5, 0] = ncols + 3, np.float32(-9.0)  # a column outside the id table, ranked first
    dev = "cuda"
    cnt_t = torch.as_tensor(np.array(counts, dtype=np.uint32).view(np.int32), device=dev)
    key_t, idx_t = torch.as_tensor(key, device=dev), torch.as_tensor(idx, device=dev)
    row_ids = torch.as_tensor(rng.permutation(10_000)[:nq].astype(np.int32), device=dev)
    col_ids = torch.as_tensor(rng.permutation(1_000_000)[:ncols].astype(np.int32), device=dev)
    L = F._nn()
    out0 = torch.full((nq, C), -7, dtype=torch.int32, device=dev)
    flag0 = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    _lib.check(L.tsm_nn_select_candidates(F._p(cnt_t), F._p(key_t), F._p(idx_t), nq, cap, C, F._p(out0),
                                          F._p(flag0), F._s()))
    out = torch.full((nq, C), -7, dtype=torch.int32, device=dev)
    flag = torch.full((nq,), -7, dtype=torch.int32, device=dev)
    pa = torch.full((nq * C,), -7, dtype=torch.int32, device=dev)
    pb = torch.full((nq * C,), -7, dtype=torch.int32, device=dev)
    _lib.check(L.tsm_nn_select_candidates_pairs(F._p(cnt_t), F._p(key_t), F._p(idx_t), nq, cap, C, F._p(out),
                                                F._p(flag), F._p(row_ids), F._p(col_ids), ncols, F._p(pa),
                                                F._p(pb), F._s()))
    assert torch.equal(out, out0) and torch.equal(flag, flag0)
    o = out.cpu().numpy()
    assert list(flag.cpu().numpy()) == [1, 0, 0, 0, 1, 0]
    assert o[5, 0] == ncols + 3
    ids = col_ids.cpu().numpy()
    want_pb = np.where(o < ncols, ids[np.minimum(o, ncols - 1)], -1).reshape(-1)
    assert np.array_equal(pb.cpu().numpy(), want_pb)
    assert np.array_equal(pa.cpu().numpy(), np.repeat(row_ids.cpu().numpy(), C))
    # argument checks: missing id tables are refused before any launch
    with pytest.raises(_lib.TsmError):
        _lib.check(L.tsm_nn_select_candidates_pairs(F._p(cnt_t), F._p(key_t), F._p(idx_t), nq, cap, C, F._p(out),
                                                    F._p(flag), None, F._p(col_ids), ncols, F._p(pa), F._p(pb),
                                                    F._s()))
