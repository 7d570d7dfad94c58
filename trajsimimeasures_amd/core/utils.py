"""Drop-in for the reference's core/utils.py:21-27 (``pad_lists_to_array``)."""
import numpy as np

__all__ = ['pad_lists_to_array']


def pad_lists_to_array(lst):
    """Zero-padded float64 [len(lst), max n_i, d] from a list of [n_i, d]
    point lists, d being the width of the first list's first point (as the
    reference reads it).  Built like the packed layout: the points of every
    list one after another, then one scatter into the padded block through
    their (list, position) indices."""
    width = len(lst[0][0])
    rows = [np.asarray(t, dtype=np.float64).reshape(-1, width) for t in lst]
    lens = np.fromiter((r.shape[0] for r in rows), dtype=np.int64, count=len(rows))
    out = np.zeros((len(rows), int(lens.max(initial=0)), width), dtype=np.float64)
    if lens.sum():
        which = np.repeat(np.arange(len(rows)), lens)
        starts = np.cumsum(lens) - lens
        pos = np.arange(int(lens.sum())) - np.repeat(starts, lens)
        out[which, pos] = np.concatenate(rows, axis=0)
    return out
