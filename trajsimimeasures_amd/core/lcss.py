"""Drop-in for the reference's core/lcss.py:56-77 (``lcss``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['lcss']


def lcss(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, eps, threads_per_traj=64, *,
         eps_mode="sqrt", return_counts=False):
    """1 - LCSS count / min(n1, n2) of each pair (core/lcss.py:31-53);
    float64 [batch].  eps_mode and return_counts as for ``edr``."""
    return run_padded(_lib.LCSS, t1, len1, t2, len2, eps=eps, eps_mode=eps_mode, threads_per_traj=threads_per_traj,
                      return_counts=return_counts)
