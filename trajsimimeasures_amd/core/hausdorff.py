"""Drop-in for the reference's core/hausdorff.py:97-123 (``hausdorff``)."""
from typing import Union

import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['hausdorff']


def hausdorff(t1: Union[np.ndarray, list], len1: Union[np.ndarray, list], t2: Union[np.ndarray, list],
              len2: Union[np.ndarray, list], threads_per_traj=64, *, precision="fp32"):
    """Symmetric point-to-segment Hausdorff distance of each pair
    (core/hausdorff.py:12-69,121); float64 [batch].  Padding never enters
    the maximum."""
    return run_padded(_lib.HAUSDORFF, t1, len1, t2, len2, threads_per_traj=threads_per_traj, precision=precision)
