"""Pins the FasterPAM oracle (oracle/tsm_oracle.c, tsmo_fasterpam), the
k-medoids step of test_clustering_time.py:353-367 (kmedoids.fasterpam,
kmedoids==0.4.3, absent here):

- a plain sequential restatement (Python loops, running sums) agrees exactly
  -- medoids, labels, loss, passes, swaps -- on integer distance matrices,
  where every order of summation gives the same sums; this pins the control
  flow: eager swaps, candidate order, the stop at the last swap, tie rules;
- the loss change the oracle computes for a candidate equals the brute-force
  change of the best single swap (removal loss and shared-gain bookkeeping);
- the result is swap-optimal (no single swap lowers the loss), its loss is
  the sum of nearest-medoid distances and its labels the nearest medoids;
- planted, well separated clusters are recovered; k = 1 is the least row sum.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O


def fasterpam_seq(D, med, max_iter=100):
    """FasterPAM (Schubert & Rousseeuw 2021) with sums in plain sequence."""
    D = np.asarray(D)
    n, k = D.shape[0], len(med)
    med = list(med)
    near, seco = [], []
    for i in range(n):
        a, b = (0, D[i, med[0]]), (None, 0.0)
        for m in range(1, k):
            d = D[i, med[m]]
            if d < a[1] or i == med[m]:
                a, b = (m, d), a
            elif b[0] is None or d < b[1]:
                b = (m, d)
        near.append(a)
        seco.append(b)

    def removal():
        rl = [0.0] * k
        for o in range(n):
            rl[near[o][0]] += seco[o][1] - near[o][1]
        return rl

    rl = removal()
    loss = sum(near[o][1] for o in range(n))
    last, swaps, it = n, 0, 0
    while it < max_iter:
        it += 1
        before = swaps
        for j in range(n):
            if j == last:
                break
            if j == med[near[j][0]]:
                continue
            p, acc = list(rl), 0.0
            for o in range(n):
                d = D[j, o]
                if d < near[o][1]:
                    acc += d - near[o][1]
                    p[near[o][0]] += near[o][1] - seco[o][1]
                elif d < seco[o][1]:
                    p[near[o][0]] += d - seco[o][1]
            b = min(range(k), key=lambda m: (p[m], m))
            if p[b] + acc >= 0:
                continue
            swaps += 1
            last = j
            med[b] = j
            for o in range(n):
                if o == j:
                    if near[o][0] != b:
                        seco[o] = near[o]
                    near[o] = (b, 0.0)
                    continue
                d = D[j, o]
                if near[o][0] == b:
                    if d < seco[o][1]:
                        near[o] = (b, d)
                    else:
                        near[o] = seco[o]
                        seco[o] = _second(D, med, near[o][0], b, o, d)
                elif d < near[o][1]:
                    seco[o], near[o] = near[o], (b, d)
                elif d < seco[o][1]:
                    seco[o] = (b, d)
                elif seco[o][0] == b:
                    seco[o] = _second(D, med, near[o][0], b, o, d)
            rl = removal()
            loss = sum(near[o][1] for o in range(n))
        if swaps == before:
            break
    return loss, np.array([a[0] for a in near]), np.array(med), it, swaps


def _second(D, med, nn, b, o, d):
    s = (b, d)
    for i, mi in enumerate(med):
        if i in (nn, b):
            continue
        if D[o, mi] < s[1]:
            s = (i, D[o, mi])
    return s


def _points_matrix(rng, n, integer=False):
    X = rng.random((n, 2)) * (100 if integer else 1)
    D = np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1))
    return np.round(D) if integer else D


def _loss(D, med):
    return D[:, med].min(1).sum()


@pytest.mark.parametrize("seed", range(6))
def test_matches_sequential_restatement_on_integer_distances(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(20, 120))
    k = int(rng.integers(2, 9))
    D = _points_matrix(rng, n, integer=True)  # many ties, exact sums in any order
    med0 = rng.choice(n, k, replace=False)
    got = O.fasterpam(D, med0)
    want = fasterpam_seq(D, med0)
    assert got[0] == want[0]
    assert np.array_equal(got[1], want[1]) and np.array_equal(got[2], want[2])
    assert (got[3], got[4]) == (want[3], want[4])


def test_swap_delta_is_the_best_single_swap():
    rng = np.random.default_rng(7)
    for _ in range(20):
        n, k = 60, int(rng.integers(2, 7))
        D = _points_matrix(rng, n)
        med = rng.choice(n, k, replace=False)
        base = _loss(D, med)
        for j in rng.choice(np.setdiff1d(np.arange(n), med), 10, replace=False):
            delta, b = O.fasterpam_delta(D, med, j)
            brute = []
            for m in range(k):
                alt = med.copy()
                alt[m] = j
                brute.append(_loss(D, alt) - base)
            assert abs(delta - min(brute)) <= 1e-9 * max(1.0, base)
            assert abs(brute[b] - min(brute)) <= 1e-9 * max(1.0, base)


def test_result_is_swap_optimal_and_consistent():
    rng = np.random.default_rng(8)
    for n, k in ((80, 3), (150, 6), (40, 10)):
        D = _points_matrix(rng, n)
        loss, labels, med, it, sw = O.fasterpam(D, rng.choice(n, k, replace=False))
        assert len(set(med.tolist())) == k
        assert np.isclose(loss, _loss(D, med), rtol=1e-12)
        assert np.array_equal(D[np.arange(n), med[labels]], D[:, med].min(1))
        assert all(labels[m] == i for i, m in enumerate(med))
        for m, x in itertools.product(range(k), range(n)):
            if x in med:
                continue
            alt = med.copy()
            alt[m] = x
            assert _loss(D, alt) >= loss - 1e-9 * loss


def test_planted_clusters_and_brute_force():
    rng = np.random.default_rng(9)
    centers = np.array([[0, 0], [10, 0], [0, 10], [10, 10]])
    X = np.concatenate([c + rng.normal(0, 0.5, (25, 2)) for c in centers])
    D = np.sqrt(((X[:, None] - X[None]) ** 2).sum(-1))
    loss, labels, med, _, _ = O.fasterpam(D, [0, 1, 2, 3])  # all four start in one cluster
    assert sorted(m // 25 for m in med) == [0, 1, 2, 3]
    for c in range(4):
        assert len(set(labels[25 * c:25 * (c + 1)])) == 1
    # tiny case: the global optimum over every 3-subset
    Y = np.concatenate([c + rng.normal(0, 0.3, (3, 2)) for c in centers[:3]])
    E = np.sqrt(((Y[:, None] - Y[None]) ** 2).sum(-1))
    best = min(itertools.combinations(range(9), 3), key=lambda s: _loss(E, list(s)))
    got = O.fasterpam(E, [0, 1, 2])
    assert sorted(got[2].tolist()) == sorted(best) and np.isclose(got[0], _loss(E, list(best)))


def test_single_medoid_is_least_row_sum():
    rng = np.random.default_rng(10)
    D = _points_matrix(rng, 50)
    loss, labels, med, it, sw = O.fasterpam(D, [7])
    j = int(np.argmin(D.sum(1)))
    assert med[0] == j and np.isclose(loss, D[j].sum()) and not labels.any()
    assert (it, sw) == (1, int(j != 7))
