"""FasterPAM k-medoids on the GPU (the clustering step of
test_clustering_time.py:353-367) against the oracle's FasterPAM on the same
matrix and initial medoids: equal loss (bit for bit: both sum in the same
fixed order), labels, medoids, passes and swaps.  Covers random and tied
(integer) matrices, k from 1 to 64, an early stop by max_iter, and the
reference's workload shape: the all-pairs DTW matrix of a trajectory set,
expanded from the triangle on the device."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _matrix(rng, n, integer=False):
    X = rng.random((n, 2)) * (50 if integer else 1)
    D = np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1))
    return np.round(D) if integer else D


def _same(got, want):
    assert got.loss == want[0], (got.loss, want[0])
    assert np.array_equal(got.labels, want[1])
    assert np.array_equal(got.medoids, want[2])
    assert (got.n_iter, got.n_swap) == (want[3], want[4])


@pytest.mark.parametrize("n,k,integer", [(100, 2, False), (700, 5, False), (3000, 10, False), (500, 64, False),
                                         (400, 7, True), (1500, 10, True)])
def test_fasterpam_matches_oracle(gpu, n, k, integer):
    from trajsimimeasures_amd import cluster
    print(f"[case] fasterpam n={n} k={k} integer={integer}", flush=True)
    rng = np.random.default_rng(n + k)
    D = _matrix(rng, n, integer)
    med0 = rng.choice(n, k, replace=False)
    got = cluster.fasterpam(D, med0)
    _same(got, O.fasterpam(D, med0))
    assert got.n_swap > 0


def test_fasterpam_k1_max_iter_and_inputs(gpu):
    import torch

    from trajsimimeasures_amd import cluster
    rng = np.random.default_rng(3)
    D = _matrix(rng, 600)
    _same(cluster.fasterpam(D, [17]), O.fasterpam(D, [17]))
    med0 = rng.choice(600, 8, replace=False)
    for it in (0, 1, 2):
        _same(cluster.fasterpam(D, med0, max_iter=it), O.fasterpam(D, med0, max_iter=it))
    # an integer k with a seed draws the kmedoids package's initial medoids;
    # a CUDA tensor input stays on the device and gives the same result
    a = cluster.fasterpam(D, 6, random_state=11)
    b = cluster.fasterpam(torch.as_tensor(D, device="cuda"), np.random.RandomState(11).choice(600, 6, False))
    _same(a, O.fasterpam(D, np.random.RandomState(11).choice(600, 6, False)))
    assert a.loss == b.loss and np.array_equal(a.labels, b.labels)


def test_cluster_all_pairs_dtw(gpu):
    """The clustering benchmark's shape: DTW over the upper triangle of a
    trajectory set, the symmetric matrix built on the device, FasterPAM
    with k = 10 (config.py:72)."""
    import torch

    from trajsimimeasures_amd import cluster, drivers, engine
    from trajsimimeasures_amd.data import lengths_uniform, random_walks
    print("[case] all-pairs dtw + fasterpam", flush=True)
    rng = np.random.default_rng(21)
    ts = random_walks(rng, lengths_uniform(rng, 2000, 10, 200))
    d = engine.DeviceTrajSet(ts)
    tri = engine.triu("dtw", d, ts.n)
    D = cluster.square_from_triu(tri, ts.n)
    torch.cuda.synchronize()
    Dh = D.cpu().numpy()
    assert np.array_equal(Dh, drivers.symmetric_from_triu(tri.cpu().numpy(), ts.n))
    med0 = np.random.RandomState(0).choice(ts.n, 10, False)
    got = cluster.fasterpam(D, med0)
    _same(got, O.fasterpam(Dh, med0))
    res, times = drivers.cluster_kmedoids("dtw", ts, 10, random_state=0)
    _same(res, O.fasterpam(Dh, med0))
    assert len(times) == 4 and all(t >= 0 for t in times)
