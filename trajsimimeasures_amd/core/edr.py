"""Drop-in for the reference's core/edr.py:55-76 (``edr``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['edr']


def edr(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, eps, threads_per_traj=64, *,
        eps_mode="sqrt", return_counts=False):
    """EDR count / max(n1, n2) of each pair (zero-boundary variant,
    core/edr.py:29-52); float64 [batch].

    eps_mode "sqrt" matches points with sqrt(dx^2 + dy^2) < eps (the CPU
    convention, core_cpu/stedr.pyx:11-12,36); "squared" with
    dx^2 + dy^2 < eps*eps (the reference kernel, core/edr.py:22,42).  Counts
    are exact integers either way; return_counts=True also returns them."""
    return run_padded(_lib.EDR, t1, len1, t2, len2, eps=eps, eps_mode=eps_mode, threads_per_traj=threads_per_traj,
                      return_counts=return_counts)
