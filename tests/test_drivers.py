"""Driver layer: dispatch and collation mirrors (CPU) and the three GPU
drivers against the oracle."""
import itertools

import numpy as np
import pytest

from trajsimimeasures_amd import drivers


def test_dispatch_and_collate():
    f = drivers.heuristic_fn_wrapper("edr", eps=0.1, threads_per_traj=16)
    assert f.keywords == {"eps": 0.1, "threads_per_traj": 16} and f.func is drivers.core.edr
    with pytest.raises(NotImplementedError):
        drivers.heuristic_fn_wrapper("dtw_cpu")
    with pytest.raises(NotImplementedError):
        drivers.heuristic_fn_wrapper("sar_cpu")
    # SAR binds config.py:55-57 and leaves threads_per_traj at its default
    # (method_helper.py:64-71)
    h = drivers.heuristic_fn_wrapper("sar", threads_per_traj=16)
    assert h.keywords == {"distance_eps": 200.0, "time_eps": 10, "target_length": 10} and h.func is drivers.core.sar
    g = drivers.heuristic_fn_wrapper("stlcss", eps=0.2)
    assert g.keywords == {"eps": 0.2, "delta": drivers.DEFAULT_DELTA, "threads_per_traj": 64}
    assert g.func is drivers.core.stlcss
    assert drivers.is_heuristic("frechet_cpu") and not drivers.is_heuristic("TrajCL")
    # the default keywords are the reference's (method_helper.py:58-71 with
    # config.py:45,52-53: eps 200, delta 60, 64 threads per trajectory)
    for name in ("edr", "lcss", "cdds"):
        assert drivers.heuristic_fn_wrapper(name).keywords == {"eps": 200.0, "threads_per_traj": 64}
    for name in ("stedr", "stlcss"):
        assert drivers.heuristic_fn_wrapper(name).keywords == {"eps": 200.0, "delta": 60.0, "threads_per_traj": 64}
    for name in ("dtw", "erp", "frechet", "hausdorff"):
        assert drivers.heuristic_fn_wrapper(name).keywords == {"threads_per_traj": 64}
    batch = [([[0.1, 0.2]], [[0.3, 0.4], [0.5, 0.6]]), ([[1 / 3, 0.0], [0.7, 0.8]], [[0.9, 1.0]])]
    t1, l1, t2, l2 = drivers.heuristic_colllate_fn(batch)
    assert t1.shape == (2, 2, 2) and t2.shape == (2, 2, 2) and l1 == [1, 2] and l2 == [2, 1]
    assert t1[1, 0, 0] == 1 / 3 and t1[0, 1, 0] == 0.0
    q1, *_ = drivers.heuristic_colllate_fn(batch, quantize_fp32=True)
    assert q1[1, 0, 0] == float(np.float32(1 / 3))
    raw = drivers.heuristic_colllate_fn(batch, is_pad=False)
    assert raw[1] == [1, 2]


def test_symmetric_from_triu():
    n = 6
    vals = np.arange(n * (n - 1) // 2, dtype=np.float64) + 1
    m = drivers.symmetric_from_triu(vals, n)
    for k, (i, j) in enumerate(itertools.combinations(range(n), 2)):
        assert m[i, j] == m[j, i] == vals[k]
    assert np.all(np.diag(m) == 0)
    with pytest.raises(ValueError):
        drivers.symmetric_from_triu(vals[:-1], n)


@pytest.mark.gpu
def test_drivers_against_oracle(gpu):
    from oracle import oracle as O
    from tests.workloads import lognormal_set
    ts = lognormal_set(70, seed=12, hi=300)
    trajs = [ts.traj(i) for i in range(ts.n)]
    pairs = [(trajs[i], trajs[i + 35]) for i in range(35)]
    for m in ("dtw", "edr", "hausdorff"):
        eps = drivers.WORKLOAD_EPS
        d, tm = drivers.pairwise(m, pairs, eps=eps)
        ref, _ = O.pairs(m, ts.xy, ts.offsets, np.arange(35), np.arange(35) + 35, eps=eps)
        assert len(d) == 35 and len(tm) == 3
        np.testing.assert_allclose(d, ref, rtol=1e-5, atol=0)
        mat, _ = drivers.query_db(m, trajs[:5], trajs[5:], eps=eps)
        qa = np.repeat(np.arange(5), 65)
        qb = np.tile(np.arange(5, 70), 5)
        ref, _ = O.pairs(m, ts.xy, ts.offsets, qa, qb, eps=eps)
        np.testing.assert_allclose(mat.reshape(-1), ref, rtol=1e-5, atol=0)
        full, _ = drivers.all_pairs(m, trajs[:30], eps=eps)
        iu = np.triu_indices(30, 1)
        ref, _ = O.pairs(m, ts.xy, ts.offsets, iu[0], iu[1], eps=eps)
        np.testing.assert_allclose(full[iu], ref, rtol=1e-5, atol=0)
        assert np.array_equal(full, full.T)
