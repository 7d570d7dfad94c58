"""The package's HDF5 reader and the reference's pair datasets over it.

The fixtures under tests/golden/ were written by the HDF5 library itself
(h5py 3.3.0 over HDF5 1.10.6, tests/golden/make_h5_fixtures.py) in the
layout preprocessing/pkl2h5py.py:16-27 writes, each with an .npz of what it
must read as: the reader is checked against the library's own files, not
against files it wrote.  The sampling of DatasetST_h5
(utilities/data_loader.py:67-73, 105-111) is checked by its defining
properties; the exact draw order follows the reference's sequence of
Python ``random`` calls and is parity unpinned (the reference loader needs
torch and h5py in one interpreter, which this image does not have).
"""
import os

import numpy as np
import pytest

from trajsimimeasures_amd.datasets import DatasetST_h5, DatasetSynthetic, read_h5_set, sample_h5_indices
from trajsimimeasures_amd.h5 import H5File, H5FormatError

GOLD = os.path.join(os.path.dirname(__file__), "golden")
READABLE = ["h5_spatial", "h5_st", "h5_many", "h5_latest", "h5_userblock", "h5_chunked"]


def _gold(name):
    return os.path.join(GOLD, name + ".h5"), np.load(os.path.join(GOLD, name + ".npz"))


@pytest.mark.parametrize("name", READABLE)
def test_reader_matches_library_written_file(name):
    path, exp = _gold(name)
    with H5File(path) as f:
        assert int(f.attrs["n"]) == int(exp["n"])
        assert np.array_equal(f.attrs["merc_range"], exp["merc_range"])
        assert ("ts_range" in f.attrs) == ("ts" in exp)
        lens = f["/trajs_len"][:]
        assert lens.dtype == np.int64 and np.array_equal(lens, exp["lengths"])
        assert sorted(f.keys("/trajs_merc"), key=int) == [str(i) for i in range(len(lens))]
        off = exp["offsets"]
        for i in range(len(lens)):
            xy = f["/trajs_merc/%d" % i][:]
            assert xy.dtype == np.float64 and xy.shape == (lens[i], 2)
            assert np.array_equal(xy, exp["xy"][off[i]:off[i + 1]])
            if "ts" in exp:
                assert np.array_equal(f["/trajs_ts/%d" % i][:], exp["ts"][off[i]:off[i + 1]])
        if "ts" in exp:
            assert np.array_equal(f.attrs["ts_range"], exp["ts_range"])


def test_reader_prefix_reads():
    path, exp = _gold("h5_chunked")
    with H5File(path) as f:
        full = f["/trajs_merc/3"][:]
        assert full.shape == (30, 2)
        for k in (0, 1, 7, 8, 9, 30, 45):
            part = f["/trajs_merc/3"][:k]
            assert np.array_equal(part, full[:k])
    path, exp = _gold("h5_spatial")
    with H5File(path) as f:
        assert np.array_equal(f["/trajs_merc/2"][:5], exp["xy"][3:8])


def test_reader_refuses_dense_link_storage():
    path, _ = _gold("h5_dense")
    with H5File(path) as f:
        assert f["/trajs_len"][:].size == 12  # the root group is compact
        with pytest.raises(H5FormatError, match="dense link storage"):
            f.keys("/trajs_merc")


def test_reader_refuses_non_hdf5(tmp_path):
    p = tmp_path / "junk.h5"
    p.write_bytes(bytes(range(256)) * 8)
    with pytest.raises(H5FormatError, match="no HDF5 signature"):
        H5File(str(p))
    e = tmp_path / "empty.h5"
    e.write_bytes(b"")
    with pytest.raises(H5FormatError):
        H5File(str(e))


def test_reader_missing_paths():
    path, _ = _gold("h5_spatial")
    with H5File(path) as f:
        assert "/trajs_len" in f and "/trajs_merc/11" in f
        assert "/trajs_merc/12" not in f and "/trajs_ts/0" not in f
        with pytest.raises(KeyError):
            f["/trajs_merc/12"]
        with pytest.raises(H5FormatError, match="not a dataset"):
            f["/trajs_merc"]


@pytest.mark.parametrize("name,temporal", [("h5_spatial", False), ("h5_st", True), ("h5_chunked", True)])
def test_read_h5_set_packs_the_file(name, temporal):
    path, exp = _gold(name)
    ts, merc, tsr = read_h5_set(path, use_temporal=temporal)
    assert np.array_equal(ts.offsets, exp["offsets"]) and np.array_equal(ts.xy, exp["xy"])
    assert np.array_equal(merc, exp["merc_range"])
    if temporal:
        assert np.array_equal(ts.t, exp["ts"].astype(np.float64)) and np.array_equal(tsr, exp["ts_range"])
    sub, _, _ = read_h5_set(path, indices=[3, 0, 3])
    assert np.array_equal(sub.lengths, exp["lengths"][[3, 0, 3]])
    assert np.array_equal(sub.traj(0), exp["xy"][exp["offsets"][3]:exp["offsets"][4]])


def test_dataset_st_h5_sampling_properties():
    path, exp = _gold("h5_st")
    lo, hi, n = 5, 12, 300
    ds = DatasetST_h5(path, True, n, lo, hi, seed=2000)
    assert len(ds) == n and np.array_equal(ds.merc_range, exp["merc_range"])
    assert np.array_equal(ds.ts_range, exp["ts_range"])
    lens, off = exp["lengths"], exp["offsets"]
    trunc = {}
    for pair in ds.sampled_indices:
        for i, m in pair:
            assert lens[i] >= lo  # filtered (data_loader.py:67)
            if lens[i] > hi:
                assert lo <= m <= hi  # truncated to a drawn length (data_loader.py:105-111)
            else:
                assert m == lens[i]
            assert trunc.setdefault(i, m) == m  # one length per trajectory for the whole dataset
    # both members are drawn from every kept trajectory, not only a few
    kept = {i for i in range(len(lens)) if lens[i] >= lo}
    assert len({p[0][0] for p in ds.sampled_indices} | {p[1][0] for p in ds.sampled_indices}) > 0.8 * len(kept)
    for k in (0, 1, n - 1):
        (i1, m1), (i2, m2) = ds.sampled_indices[k]
        t1, t2 = ds[k]
        for (i, m), t in (((i1, m1), t1), ((i2, m2), t2)):
            a = np.asarray(t)
            assert a.shape == (m, 3)
            assert np.array_equal(a[:, :2], exp["xy"][off[i]:off[i] + m])
            assert np.array_equal(a[:, 2], exp["ts"][off[i]:off[i] + m])
    s, pa, pb = ds.packed()
    for k in range(n):
        (i1, m1), (i2, m2) = ds.sampled_indices[k]
        assert np.array_equal(s.traj(int(pa[k])), exp["xy"][off[i1]:off[i1] + m1])
        assert np.array_equal(s.traj(int(pb[k])), exp["xy"][off[i2]:off[i2] + m2])
    assert len(set(zip(s.offsets[:-1].tolist(), s.lengths.tolist()))) == s.n  # each item stored once


def test_dataset_st_h5_is_deterministic_per_seed():
    path, _ = _gold("h5_many")
    a = DatasetST_h5(path, False, 500, 1, 2, seed=1)
    b = DatasetST_h5(path, False, 500, 1, 2, seed=1)
    c = DatasetST_h5(path, False, 500, 1, 2, seed=2)
    assert a.sampled_indices == b.sampled_indices != c.sampled_indices
    assert a[7] == b[7] and isinstance(a[7][0], list) and len(a[7][0][0]) == 2


def test_dataset_st_h5_needs_enough_trajectories():
    lens = [3, 50, 4, 2]
    with pytest.raises(AssertionError):
        sample_h5_indices(lens, 5, 10, 20, seed=0)  # one kept trajectory, sqrt(5) needed
    out = sample_h5_indices(lens, 4, 3, 20, seed=0)  # three kept, sqrt(4) = 2 needed
    assert len(out) == 4 and all(i in (0, 1, 2) for p in out for i, _ in p)


def test_dataset_synthetic_items():
    np.random.seed(0)
    ds = DatasetSynthetic(True, 10, 3, 9, seed=0)
    assert len(ds) == 10 and ds.merc_range == [0.0, 1.0, 0.0, 1.0]
    for k in range(10):
        t1, t2 = ds[k]
        for t in (t1, t2):
            a = np.asarray(t)
            assert 3 <= a.shape[0] <= 9 and a.shape[1] == 3
            assert np.all((a[:, :2] >= 0) & (a[:, :2] < 1))
            steps = np.diff(np.concatenate([[0], a[:, 2]]))
            assert np.all((steps >= 1) & (steps < 60))
    t1, t2 = DatasetSynthetic(False, 2, 4, 4, seed=0)[0]
    assert np.asarray(t1).shape == (4, 2) and np.asarray(t2).shape == (4, 2)


def test_fp32_origin_keeps_fp32_inputs_exact():
    """The origin of the engine's FP32 copy: for FP32-valued Mercator
    coordinates (what the reference's collate produces) x - o is exact in
    FP32; for FP64 inputs the cast error drops from ~0.5 m to FP32's
    spacing of the set's extent; cmax bounds |x - o|."""
    from trajsimimeasures_amd.engine import fp32_origin
    _, exp = _gold("h5_st")
    for xy in (exp["xy"].astype(np.float32).astype(np.float64), exp["xy"]):
        o, cmax = fp32_origin(xy)
        c = xy - np.asarray(o)
        assert np.all(np.abs(c) <= cmax) and cmax < 3e4
        lo, hi = xy.min(0), xy.max(0)
        assert np.all(np.abs(np.asarray(o) - 0.5 * (lo + hi)) <= 1.0)
        err = np.abs(c.astype(np.float32).astype(np.float64) - c)
        if xy.dtype == np.float64 and np.array_equal(xy, xy.astype(np.float32)):
            assert np.all(err == 0.0)
        assert err.max() <= np.spacing(np.float32(cmax)) / 2
        plain = np.abs(xy.astype(np.float32).astype(np.float64) - xy).max()
        if not np.array_equal(xy, xy.astype(np.float32)):
            assert plain > 100 * err.max()
    assert fp32_origin(np.zeros((0, 2))) == ((0.0, 0.0), 0.0)
    assert fp32_origin(np.array([[0.0, 0.0], [0.0, 0.0]])) == ((0.0, 0.0), 0.0)
    o, cmax = fp32_origin(np.array([[0.25, -3.0], [0.75, 5.0]]))
    assert o == (0.5, 1.0) and cmax == 4.0


def test_reader_refuses_truncated_and_corrupted_files(tmp_path):
    """A file cut short, or with its root group's B-tree signature damaged,
    raises H5FormatError (never an IndexError or a silent misread)."""
    path, _ = _gold("h5_spatial")
    raw = open(path, "rb").read()
    cut = tmp_path / "cut.h5"
    cut.write_bytes(raw[:2000])
    with pytest.raises(H5FormatError):
        with H5File(str(cut)) as f:
            for i in range(12):
                f["/trajs_merc/%d" % i][:]
    bad = bytearray(raw)
    i = raw.find(b"TREE")
    bad[i:i + 4] = b"XXXX"
    p = tmp_path / "bad.h5"
    p.write_bytes(bytes(bad))
    with pytest.raises(H5FormatError, match="B-tree"):
        with H5File(str(p)) as f:
            f.keys("/")
