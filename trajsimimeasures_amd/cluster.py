"""k-medoids clustering of the all-pairs matrix on the GPU: the clustering
step of the reference's benchmark (test_clustering_time.py:353-367,
``kmedoids.fasterpam(dists, Config.clustering_num_centroids)``; kmedoids==0.4.3)
behind the same call shape.

``fasterpam(diss, medoids, max_iter=100, init="random", random_state=None)``
takes the dissimilarity matrix as a float64 numpy array or a CUDA tensor
(kept on the device), and ``medoids`` as the number of clusters or the
initial medoid indices, and returns ``KMedoidsResult(loss, labels, medoids,
n_iter, n_swap)`` like the kmedoids package.  Initial medoids for an integer
``medoids``: ``init="random"`` draws ``RandomState(random_state).choice(n, k,
replace=False)`` (numpy's global generator when random_state is None),
``"first"`` takes 0..k-1.
"""
from __future__ import annotations

import ctypes
from collections import namedtuple

import numpy as np

from . import _lib

__all__ = ["KMedoidsResult", "fasterpam", "square_from_triu", "initial_medoids"]

KMedoidsResult = namedtuple("KMedoidsResult", ["loss", "labels", "medoids", "n_iter", "n_swap"])


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def initial_medoids(n: int, medoids, init: str = "random", random_state=None) -> np.ndarray:
    """Initial medoid indices: an explicit array is used as given; an integer
    k is drawn as described in the module doc."""
    if isinstance(medoids, (int, np.integer)):
        k = int(medoids)
        if not 1 <= k <= n:
            raise ValueError(f"need 1 <= k <= n, got k={k}, n={n}")
        if init == "random":
            if random_state is None or random_state is np.random:
                rs = np.random.mtrand._rand
            elif isinstance(random_state, np.random.RandomState):
                rs = random_state
            else:
                rs = np.random.RandomState(int(random_state))
            return np.asarray(rs.choice(n, k, False), dtype=np.int64)
        if init == "first":
            return np.arange(k, dtype=np.int64)
        raise ValueError(f"unknown initialization {init!r} (random or first)")
    med = np.asarray(medoids, dtype=np.int64).reshape(-1).copy()
    if med.size < 1 or med.min() < 0 or med.max() >= n or np.unique(med).size != med.size:
        raise ValueError("initial medoids must be distinct indices in [0, n)")
    return med


def square_from_triu(tri, n: int):
    """Symmetric [n, n] float64 CUDA tensor (zero diagonal) from the strict
    upper triangle in itertools.combinations order (a CUDA tensor)."""
    import torch
    tri = tri.to(torch.float64).contiguous()
    if tri.numel() != n * (n - 1) // 2:
        raise ValueError("triangle size does not match n")
    out = torch.empty((n, n), dtype=torch.float64, device=tri.device)
    _lib.check(_lib.lib().tsm_square_from_triu(tri.data_ptr() if tri.numel() else None, n, out.data_ptr(),
                                               _stream()))
    return out


def fasterpam(diss, medoids, max_iter: int = 100, init: str = "random", random_state=None, n_cpu: int = -1):
    """FasterPAM k-medoids of the dissimilarity matrix ``diss`` on the GPU
    (the kmedoids.fasterpam call shape; ``n_cpu`` is accepted and unused).

    Every swap that lowers the loss is taken as soon as a candidate is found
    (candidates in index order), a pass ends at the candidate of the last
    swap, and passes repeat until one makes no swap or ``max_iter`` passes
    ran.  Dissimilarities must not be NaN; at most 64 medoids."""
    import torch
    if isinstance(diss, torch.Tensor):
        D = diss.to(torch.float64)
        if not D.is_cuda:
            D = D.cuda()
    else:
        D = torch.as_tensor(np.ascontiguousarray(np.asarray(diss, dtype=np.float64)), device="cuda")
    D = D.contiguous()
    if D.dim() != 2 or D.shape[0] != D.shape[1]:
        raise ValueError("diss must be a square matrix")
    n = D.shape[0]
    med = np.ascontiguousarray(initial_medoids(n, medoids, init, random_state))
    labels = torch.empty(n, dtype=torch.int32, device=D.device)
    loss = ctypes.c_double()
    it = ctypes.c_int32()
    sw = ctypes.c_int32()
    _lib.check(_lib.lib().tsm_kmedoids_fasterpam(D.data_ptr(), n, int(med.size), med.ctypes.data, int(max_iter),
                                                 labels.data_ptr(), ctypes.byref(loss), ctypes.byref(it),
                                                 ctypes.byref(sw), _stream()))
    return KMedoidsResult(loss.value, labels.cpu().numpy().astype(np.int64), med, it.value, sw.value)
