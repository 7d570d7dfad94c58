"""Drop-in for the reference's core/discret_frechet.py:53-74 (``dfrechet``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['dfrechet']


def dfrechet(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, threads_per_traj=64, *,
             precision="fp32"):
    """Discrete Frechet distance of each pair; float64 [batch]."""
    return run_padded(_lib.FRECHET, t1, len1, t2, len2, threads_per_traj=threads_per_traj, precision=precision)
