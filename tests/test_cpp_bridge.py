"""The C++ adapter include/tsm_traj.hpp (the body of the reference's
Traj::TrajSimiGPU, query/cpp/spatial/traj.cpp:242-360, over the C ABI):
compiled with g++ against libtsm.so; on the CPU its failure is a clean error
code, on the GPU its values match the oracle."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "trajsimimeasures_amd")


def _build(tmp_path):
    exe = str(tmp_path / "traj_bridge")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "cpp", "traj_bridge.cpp"), "-L", PKG, "-ltsm", f"-Wl,-rpath,{PKG}",
           "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def _write(tmp_path, trajs, n):
    p = tmp_path / "pairs.txt"
    with open(p, "w") as f:
        f.write(f"{n}\n")
        for t in trajs:
            f.write(f"{len(t)} " + " ".join(f"{v:.9g}" for v in t.reshape(-1)) + "\n")
    return str(p)


def _trajs(n, seed):
    from trajsimimeasures_amd.data import lengths_uniform, random_walks
    rng = np.random.default_rng(seed)
    ts = random_walks(rng, lengths_uniform(rng, 2 * n, 5, 60))
    # the reference's Point holds floats
    return [ts.traj(i).astype(np.float32) for i in range(2 * n)]


def test_bridge_builds_and_fails_cleanly_without_device(tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("device present: the GPU test covers the call")
    exe = _build(tmp_path)
    f = _write(tmp_path, _trajs(3, 0), 3)
    r = subprocess.run([exe, f, "DTW"], capture_output=True, text=True)
    assert r.returncode == 2 and "failed" in r.stderr


@pytest.mark.gpu
def test_bridge_matches_oracle(gpu, tmp_path):
    from oracle import oracle as O
    n = 200
    trajs = _trajs(n, 1)
    exe = _build(tmp_path)
    f = _write(tmp_path, trajs, n)
    measures = ["DTW", "ERP", "Hausdorff", "Frechet", "EDR", "LCSS"]
    r = subprocess.run([exe, f] + measures, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got = {}
    for line in r.stdout.splitlines():
        m, i, v = line.split()
        got.setdefault(m, np.zeros(n))[int(i)] = float(v)
    from trajsimimeasures_amd.data import pack_list
    ts = pack_list([t.astype(np.float64) for t in trajs])
    pa, pb = np.arange(n), np.arange(n, 2 * n)
    names = {"DTW": "dtw", "ERP": "erp", "Hausdorff": "hausdorff", "Frechet": "frechet", "EDR": "edr", "LCSS": "lcss"}
    for m in measures:
        ref, _ = O.pairs(names[m], ts.xy, ts.offsets, pa, pb, eps=0.0025, gpoint=(0.0, 0.0))
        # values travel back as float (the reference's std::vector<float>)
        assert np.allclose(got[m], ref.astype(np.float32), rtol=1e-5, atol=1e-12), m
