"""Drop-in for the reference's core/erp.py:94-120 (``erp``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['erp']


def erp(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, gpoint=None, threads_per_traj=64, *,
        precision="fp32"):
    """ERP of each pair with gap point ``gpoint`` (default (0, 0),
    core/erp.py:103) and the constant-sum boundary of core/erp.py:62-78;
    float64 [batch].  Unlike the reference (whose ``gpoint == None`` test
    raises for an ndarray), any 2-element sequence is accepted."""
    return run_padded(_lib.ERP, t1, len1, t2, len2, gpoint=gpoint, threads_per_traj=threads_per_traj,
                      precision=precision)
