"""Drop-in for the reference's core/cdds.py:97-124 (``cdds``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['cdds']


def cdds(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, eps, threads_per_traj=64):
    """Total time the two moving points stay within eps of each other,
    summed over every pair of segments (core/cdds.py:21-94); float64 [batch].

    t1, t2 are [batch, seq, 3] arrays of (x, y, timestamp).  A segment with no
    duration contributes 0 (core_cpu/cdds.pyx:26), and padding is never read
    (the reference sums an uninitialised buffer past each trajectory's end,
    core/cdds.py:115,124)."""
    return run_padded(_lib.CDDS, t1, len1, t2, len2, eps=eps, threads_per_traj=threads_per_traj)
