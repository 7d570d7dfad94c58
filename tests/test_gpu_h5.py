"""Dataset-file pairs at Mercator scale on the GPU against the FP64 oracle.

The pairs come from DatasetST_h5 over tests/golden/h5_st.h5, a file the
HDF5 library wrote in the layout of preprocessing/pkl2h5py.py:16-27, with
coordinates around the Beijing Mercator origin (x ~ 1.296e7 m, y ~ 4.85e6 m,
points ~60 m apart, trajectories within +-20 km).  A plain FP32 cast of such
coordinates is off by up to 0.5 m; the engine's FP32 copy is taken relative
to the set's centre instead.

Bars (as in the other parity tests): DTW, ERP, Frechet, Hausdorff max
relative error <= 1e-5 against the FP64 oracle on the FP64 coordinates;
EDR, LCSS, STEDR, STLCSS counts bit-exact.
"""
import os

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

FP32_RTOL = 1e-5
EPS_M = 100.0  # EDR / LCSS threshold in metres
GOLD = os.path.join(os.path.dirname(__file__), "golden")
ALL = ["dtw", "edr", "lcss", "erp", "frechet", "hausdorff"]


def _dataset(temporal, n=400):
    from trajsimimeasures_amd.datasets import DatasetST_h5
    return DatasetST_h5(os.path.join(GOLD, "h5_st.h5"), temporal, n, 2, 30, seed=5)


def _rel(got, ref):
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-300)
    rel[got == ref] = 0.0
    return rel


@pytest.mark.parametrize("measure", ALL)
def test_mercator_pairs_from_dataset_file(gpu, measure):
    from trajsimimeasures_amd import engine
    print(f"[case] h5 mercator pairs measure={measure}", flush=True)
    ts, pa, pb = _dataset(False).packed()
    d = engine.DeviceTrajSet(ts)
    assert abs(d.origin[0] - 1.2958e7) < 5e4 and d.coord_max < 5e4
    g = ts.xy.mean(0)
    out = engine.pairs(measure, d, pa, pb, eps=EPS_M, gpoint=g, return_counts=measure in ("edr", "lcss"))
    ref, rcnt = O.pairs(measure, ts.xy, ts.offsets, pa, pb, eps=EPS_M, gpoint=g)
    if measure in ("edr", "lcss"):
        got, cnt = out[0].cpu().numpy(), out[1].cpu().numpy()
        assert np.array_equal(cnt, rcnt), f"{measure}: {int((cnt != rcnt).sum())} count mismatches"
        assert np.array_equal(got, ref)
        assert rcnt.sum() > 0  # pairs drawn twice from one trajectory match everywhere
    else:
        rel = _rel(out.cpu().numpy(), ref)
        assert rel.max() <= FP32_RTOL, f"{measure}: max rel err {rel.max():.3e}"


def test_centring_is_what_keeps_the_bar(gpu):
    """The same DTW pairs from an FP32 copy made without an origin miss the
    bar (the cast alone costs ~1e-4 relative on these pairs); the centred
    copy holds it (its cast costs ~1e-7)."""
    import torch

    from trajsimimeasures_amd import _lib, engine
    ts, pa, pb = _dataset(False).packed()
    d = engine.DeviceTrajSet(ts)
    ref, _ = O.pairs("dtw", ts.xy, ts.offsets, pa, pb)
    centred = _rel(engine.pairs("dtw", d, pa, pb).cpu().numpy(), ref).max()
    _lib.check(_lib.lib().tsm_xy_to_fp32(d.xy64.data_ptr(), d.xy32.data_ptr(), d.npts, engine._stream_ptr()))
    d.coord_max = float(np.abs(ts.xy).max())
    plain = _rel(engine.pairs("dtw", d, pa, pb).cpu().numpy(), ref).max()
    torch.cuda.synchronize()
    print(f"dtw max rel err: centred {centred:.2e}, plain FP32 cast {plain:.2e}", flush=True)
    assert centred <= FP32_RTOL < plain


def test_dropin_core_on_mercator_batch(gpu):
    """core.* (padded batches through tsm_pairs_host, which centres the
    FP32 copy in the library) on the dataset file's pairs."""
    from trajsimimeasures_amd import core
    ds = _dataset(False, n=128)
    t1l, t2l = zip(*(ds[k] for k in range(len(ds))))
    t1 = core.pad_lists_to_array(list(t1l))
    t2 = core.pad_lists_to_array(list(t2l))
    len1 = [len(t) for t in t1l]
    len2 = [len(t) for t in t2l]
    ts, pa, pb = ds.packed()
    for m, got in (("dtw", core.dtw(t1, len1, t2, len2)), ("edr", core.edr(t1, len1, t2, len2, EPS_M)),
                   ("hausdorff", core.hausdorff(t1, len1, t2, len2))):
        ref, _ = O.pairs(m, ts.xy, ts.offsets, pa, pb, eps=EPS_M)
        if m == "edr":
            assert np.array_equal(got, ref)
        else:
            assert _rel(got, ref).max() <= FP32_RTOL, m


@pytest.mark.parametrize("measure", ["stedr", "stlcss"])
def test_st_pairs_from_dataset_file(gpu, measure):
    from trajsimimeasures_amd import engine
    ts, pa, pb = _dataset(True).packed()
    assert ts.t is not None and ts.t.min() > 1.2e9
    d = engine.DeviceTrajSet(ts)
    got, cnt = engine.pairs(measure, d, pa, pb, eps=EPS_M, delta=600.0, return_counts=True)
    ref, rcnt = O.pairs_st(measure, ts.xy, ts.t, ts.offsets, pa, pb, EPS_M, 600.0)
    assert np.array_equal(cnt.cpu().numpy(), rcnt) and np.array_equal(got.cpu().numpy(), ref)


def test_pairwise_driver_takes_the_dataset(gpu):
    """drivers.pairwise over the dataset (its packed form) and over its
    items as lists (the reference's DataLoader path) both hold the bar."""
    from trajsimimeasures_amd import drivers
    ds = _dataset(False, n=200)
    ts, pa, pb = ds.packed()
    ref, _ = O.pairs("dtw", ts.xy, ts.offsets, pa, pb)
    packed, _ = drivers.pairwise("dtw", ds)
    lists, _ = drivers.pairwise("dtw", [ds[k] for k in range(len(ds))])
    assert len(packed) == len(lists) == 200
    assert _rel(np.array(packed), ref).max() <= FP32_RTOL
    assert _rel(np.array(lists), ref).max() <= FP32_RTOL
