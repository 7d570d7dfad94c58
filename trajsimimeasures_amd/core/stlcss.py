"""Drop-in for the reference's core/stlcss.py:57-80 (``stlcss``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['stlcss']


def stlcss(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, eps, delta, threads_per_traj=64, *,
        eps_mode="sqrt", return_counts=False):
    """1 - STLCSS count / min(n1, n2) of each pair; float64 [batch].

    t1, t2 are [batch, seq, 3] arrays of (x, y, timestamp).  Two points
    match when they are within eps in space (eps_mode as for LCSS) and
    |t1 - t2| < delta in time.  Counts are exact integers; return_counts=True
    also returns them."""
    return run_padded(_lib.STLCSS, t1, len1, t2, len2, eps=eps, eps_mode=eps_mode, delta=delta,
                      threads_per_traj=threads_per_traj, return_counts=return_counts)
