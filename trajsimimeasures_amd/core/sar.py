"""Drop-in for the reference's core/sar.py:247-286 (``sar``)."""
import numpy as np

from .. import _lib
from ._batch import run_padded

__all__ = ['sar']


def sar(t1: np.ndarray, len1: np.ndarray, t2: np.ndarray, len2: np.ndarray, distance_eps, time_eps, target_length,
        threads_per_traj=64):
    """Time the two trajectories stay within distance_eps of each other on a
    common time grid of time_eps seconds, compared in target_length bins
    over the stretch where both exist (core/sar.py:62-212); float64 [batch].

    t1, t2 are [batch, seq, 3] arrays of (x, y, timestamp); time_eps is a
    whole number of seconds.  As in the reference, when the common stretch
    has no more than target_length samples, each of its threads_per_traj
    threads contributes the same per-sample count, so that result is
    threads_per_traj * count * time_eps (threads_per_traj = 1 gives the count
    itself); binned results do not depend on threads_per_traj.  The last bin
    ends with the stretch (the reference reads one sample past it), and a
    pair without temporal overlap gives 0."""
    return run_padded(_lib.SAR, t1, len1, t2, len2, eps=float(distance_eps),
                      sar=(time_eps, target_length, threads_per_traj))
