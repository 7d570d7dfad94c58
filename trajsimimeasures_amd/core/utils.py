"""Drop-in for the reference's core/utils.py:21-27 (``pad_lists_to_array``)."""
import numpy as np

__all__ = ['pad_lists_to_array']


def pad_lists_to_array(lst):
    """Zero-pad a list of [n_i, d] point lists into float64 [len(lst), max n_i, d]."""
    dim2 = len(lst[0][0])
    dim1 = max(map(len, lst))
    arr = np.zeros([len(lst), dim1, dim2], dtype=np.float64)
    for i, j in enumerate(lst):
        arr[i][0:len(j)] = j
    return arr
