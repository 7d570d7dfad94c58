"""hipGraph capture/replay of a batch and the autotuner, on the device."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.workloads import EPS, uniform_pairs

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("measure", ["dtw", "edr", "erp", "hausdorff"])
def test_graph_replays_match_eager(gpu, measure):
    import torch

    from trajsimimeasures_amd import engine
    print(f"[case] graph replay {measure}", flush=True)
    ts, pa, pb = uniform_pairs(4000, seed=31)
    d = engine.DeviceTrajSet(ts)
    pa_dev = torch.as_tensor(pa, dtype=torch.int32, device="cuda")
    pb_dev = torch.as_tensor(pb, dtype=torch.int32, device="cuda")
    g = engine.PairsGraph(measure, d, pa_dev, pb_dev, eps=EPS, gpoint=(0.5, 0.5))
    eager = engine.pairs(measure, d, pa, pb, eps=EPS, gpoint=(0.5, 0.5)).cpu().numpy()
    for rep in range(3):
        g.out.fill_(-1.0)
        g.stream.synchronize()
        torch.cuda.synchronize()
        g.launch()
        g.stream.synchronize()
        got = g.out.cpu().numpy()
        assert np.array_equal(got, eager), (measure, rep)
        x = torch.randn(1 << 20, device="cuda").sum()  # eager work between replays
        torch.cuda.synchronize()
        del x
    # refill the pair list in place: replays follow the new contents
    perm = np.random.default_rng(3).permutation(pa.size)
    g.pa.copy_(torch.as_tensor(pb[perm], dtype=torch.int32))
    g.pb.copy_(torch.as_tensor(pa[perm], dtype=torch.int32))
    torch.cuda.synchronize()
    # the graph owns its pair tensors: the caller's are untouched
    assert np.array_equal(pa_dev.cpu().numpy(), pa) and np.array_equal(pb_dev.cpu().numpy(), pb)
    g.launch()
    g.stream.synchronize()
    ref, _ = O.pairs(measure, ts.xy, ts.offsets, pb[perm], pa[perm], eps=EPS, gpoint=(0.5, 0.5))
    got = g.out.cpu().numpy()
    if measure == "edr":
        assert np.array_equal(got, ref)
    else:
        assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5
    g.close()


def test_autotune_small(gpu):
    from trajsimimeasures_amd import autotune, engine
    t = autotune.tune(npairs_per_bucket=1024, waves=(4, 8), ncol_max=64, workload_pairs=4096, verbose=False,
                      direct_sizes=(256, 1024))
    # every bucket timed every lane count that holds its columns, and reports the model's margin
    for prec in ("fp32", "fp64"):
        rows = t["buckets"][prec]
        assert [r["ncol"][0] for r in rows] == list(range(1, 65, autotune.BUCKET))
        for r in rows:
            assert {int(c) for c in r["ms"]} == {c for c in autotune.LANES if autotune.fits(24, c, r["ncol"][1])}
            assert r["best"] in (2, 4, 8, 16) and r["model_over_best"] >= 0.0
        assert t["workloads"][prec]["kept"] in ("measured table", "cost model")
    # a measured lanes table is kept only where it beat the built-in cost model
    for key in ("lanes_fp32", "lanes_fp64"):
        if key in t:
            assert all(autotune.fits(24, t[key][nc], nc) for nc in range(1, 257))
    assert t["waves_per_cu"] in (4, 8)
    # the direct-launch crossover: the largest timed batch size at which direct won (0: none)
    assert set(t["direct"]) == {"256", "1024"} and t["direct_pairs"] in (0, 256, 1024)
    if t["direct_pairs"] > 0:
        r = t["direct"][str(t["direct_pairs"])]
        assert r["direct_ms"] < r["planned_ms"]
    ts, pa, pb = uniform_pairs(1500, seed=2)
    d = engine.DeviceTrajSet(ts)
    for m in ("dtw", "lcss"):
        got = engine.pairs(m, d, pa, pb, eps=EPS).cpu().numpy()
        ref, _ = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS)
        assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5
    # back to the shipped (or built-in) table
    import ctypes

    from trajsimimeasures_amd import _lib
    _lib.check(_lib.lib().tsm_tune_set_lanes(32, None, 0))
    _lib.check(_lib.lib().tsm_tune_set_lanes(64, None, 0))
    _lib.check(_lib.lib().tsm_tune_set_waves_per_cu(0))
    _lib.check(_lib.lib().tsm_tune_set_direct_pairs(-1))
    autotune.load_shipped()


def test_calls_on_two_streams_share_the_workspace_safely(gpu):
    """Back-to-back calls on two streams (no host synchronisation between
    them) reuse one workspace; the library orders them, so both results are
    right."""
    import torch

    from trajsimimeasures_amd import engine
    ts, pa, pb = uniform_pairs(4000, seed=9)
    d = engine.DeviceTrajSet(ts)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    outs = []
    for k in range(6):
        with torch.cuda.stream(s1 if k % 2 == 0 else s2):
            m = ("dtw", "edr")[k % 2]
            outs.append((m, engine.pairs(m, d, pa, pb, eps=EPS, validate=False)))
    torch.cuda.synchronize()
    for m, o in outs:
        ref, _ = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS)
        got = o.cpu().numpy()
        if m == "edr":
            assert np.array_equal(got, ref)
        else:
            assert np.max(np.abs(got - ref) / np.maximum(ref, 1e-300)) <= 1e-5
