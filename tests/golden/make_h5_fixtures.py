"""Write the HDF5 fixtures of the on-disk dataset format with the real HDF5
library (h5py), and the expected contents beside each as a plain .npz.

The files follow the layout preprocessing/pkl2h5py.py:16-27 writes (read by
utilities/data_loader.py:52-99 and query/cpp/utils/dataset.hpp:48-113):
file attributes ``n``, ``merc_range`` and, for spatio-temporal sets,
``ts_range``; an int64 dataset ``/trajs_len``; one float64 [len, 2] dataset
``/trajs_merc/<i>`` per trajectory and, for spatio-temporal sets, one
``/trajs_ts/<i>``.  The writes are made with h5py's defaults, as
pkl2h5py.py does, so the files carry the format that script produces.

The system interpreter has no h5py; run this with an interpreter that has
it (this image: /opt/conda/bin/python3.9):

    /opt/conda/bin/python3.9 tests/golden/make_h5_fixtures.py

Files written (each NAME.h5 with NAME.npz holding what it should read as):
  h5_spatial   12 trajectories, Mercator-scale coordinates (around the
               Beijing Mercator origin), lengths 1..40, no timestamps
  h5_st        80 trajectories with int64 timestamps and ts_range; the
               /trajs_merc group needs several symbol-table nodes
  h5_many      1,100 short trajectories: a two-level group B-tree

and, to cover the rest of what the reader accepts or refuses (these are not
pkl2h5py.py's defaults):
  h5_latest    libver="latest" (version 3 superblock, version 2 object
               headers, links in link messages), 6 trajectories with timestamps
  h5_userblock a 512-byte user block before the superblock
  h5_chunked   chunked datasets (chunks of up to 8 points) through the shuffle,
               deflate and Fletcher-32 filters
  h5_dense     libver="latest" with 12 trajectories: /trajs_merc keeps its
               links in dense storage, which the reader refuses by name
"""
import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def _walks(rng, lengths, x0, y0, step):
    out = []
    for n in lengths:
        th = rng.uniform(0, 2 * np.pi) + np.cumsum(rng.normal(0, 0.3, n))
        s = rng.lognormal(np.log(step), 0.5, n)
        xy = np.empty((n, 2))
        xy[:, 0] = x0 + rng.uniform(-2e4, 2e4) + np.cumsum(s * np.cos(th))
        xy[:, 1] = y0 + rng.uniform(-2e4, 2e4) + np.cumsum(s * np.sin(th))
        out.append(xy)
    return out


def write(name, lengths, with_ts, seed, file_kw=None, ds_kw=None):
    file_kw = file_kw or {}
    ds_kw = ds_kw or {}
    rng = np.random.default_rng(seed)
    trajs = _walks(rng, lengths, 12958000.0, 4852000.0, 60.0)
    allxy = np.concatenate(trajs)
    merc_range = [float(allxy[:, 0].min()), float(allxy[:, 0].max()),
                  float(allxy[:, 1].min()), float(allxy[:, 1].max())]
    ts = None
    if with_ts:
        ts = [1201000000 + int(rng.integers(0, 10**6)) + np.cumsum(rng.integers(1, 60, n)).astype(np.int64)
              for n in lengths]
    path = os.path.join(HERE, name + ".h5")
    with h5py.File(path, "w", **file_kw) as f:
        # the writes of preprocessing/pkl2h5py.py:16-27, with Python lists as there
        f.attrs["n"] = len(lengths)
        f.attrs["merc_range"] = merc_range
        if with_ts:
            allts = np.concatenate(ts)
            f.attrs["ts_range"] = [int(allts.min()), int(allts.max())]
        f.create_dataset("/trajs_len", data=[int(n) for n in lengths], **ds_kw)
        for i, xy in enumerate(trajs):
            kw = dict(ds_kw, chunks=(min(8, len(xy)), 2)) if ds_kw else {}
            f.create_dataset("/trajs_merc/%s" % i, data=xy.tolist(), **kw)
            if with_ts:
                f.create_dataset("/trajs_ts/%s" % i, data=ts[i].tolist(), **ds_kw)
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    exp = dict(n=np.int64(len(lengths)), merc_range=np.array(merc_range), lengths=np.asarray(lengths, np.int64),
               offsets=offsets, xy=allxy)
    if with_ts:
        exp["ts"] = np.concatenate(ts)
        exp["ts_range"] = np.array([int(exp["ts"].min()), int(exp["ts"].max())], np.int64)
    np.savez(os.path.join(HERE, name + ".npz"), **exp)
    print(name, os.path.getsize(path), "bytes,", len(lengths), "trajectories")


if __name__ == "__main__":
    r = np.random.default_rng(7)
    write("h5_spatial", [1, 2, 40, 17, 5, 33, 9, 28, 3, 12, 39, 21], False, 1)
    write("h5_st", list(r.integers(2, 30, 80)), True, 2)
    write("h5_many", list(r.integers(1, 4, 1100)), False, 3)
    write("h5_latest", [3, 1, 25, 8, 14, 30], True, 4, file_kw=dict(libver="latest"))
    write("h5_userblock", [4, 9, 2, 16, 7], False, 5, file_kw=dict(userblock_size=512))
    write("h5_chunked", [1, 8, 9, 30, 17, 2, 24, 16], True, 6,
          ds_kw=dict(compression="gzip", shuffle=True, fletcher32=True))
    write("h5_dense", list(r.integers(2, 10, 12)), False, 7, file_kw=dict(libver="latest"))
