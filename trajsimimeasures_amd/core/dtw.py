"""Drop-in for the reference's core/dtw.py:56-77 (``dtw``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['dtw']


def dtw(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, threads_per_traj=64, *,
        precision="fp32"):
    """DTW of pairs (t1[k][:len1[k]], t2[k][:len2[k]]): t1, t2 zero-padded
    [batch, seq, 2], returns float64 [batch].  The reference requires equal
    shapes (core/dtw.py:60)."""
    assert np.shape(t1) == np.shape(t2)
    return run_padded(_lib.DTW, t1, len1, t2, len2, threads_per_traj=threads_per_traj, precision=precision)
