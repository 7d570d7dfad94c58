"""Independent brute-force definitions of the six measures for tiny inputs.

These do not reuse the DP recurrences: each measure is stated over explicit
alignments (monotone lattice paths, chains of matched index pairs, or points
against segments) and minimised/maximised by enumeration.  Used to pin the
oracle (tests/test_oracle.py).
"""
from __future__ import annotations

import itertools
import math

import numpy as np


def _d(p, q):
    return math.sqrt((p[0] - q[0]) ** 2 + (p[1] - q[1]) ** 2)


def _paths(n1, n2):
    """All monotone cell paths (1,1) -> (n1,n2) with steps (1,0), (0,1), (1,1)."""
    out = []

    def rec(i, j, acc):
        acc = acc + [(i, j)]
        if (i, j) == (n1, n2):
            out.append(acc)
            return
        if i < n1:
            rec(i + 1, j, acc)
        if j < n2:
            rec(i, j + 1, acc)
        if i < n1 and j < n2:
            rec(i + 1, j + 1, acc)

    rec(1, 1, [])
    return out


def dtw(a, b):
    """min over warping paths of the summed point distances."""
    return min(sum(_d(a[i - 1], b[j - 1]) for i, j in p) for p in _paths(len(a), len(b)))


def frechet(a, b):
    """min over warping paths of the largest point distance."""
    return min(max(_d(a[i - 1], b[j - 1]) for i, j in p) for p in _paths(len(a), len(b)))


def _grid_paths(starts, n1, n2):
    """Node paths in the (n1+1) x (n2+1) grid from a start node to (n1, n2)
    whose every move lands on an interior node (i, j >= 1); moves are
    'd' (i-1,j-1)->(i,j), 'u' (i-1,j)->(i,j), 'l' (i,j-1)->(i,j)."""
    res = []

    def rec(i, j, moves):
        if (i, j) == (n1, n2):
            res.append(moves)
            return
        for mv, (di, dj) in (("d", (1, 1)), ("u", (1, 0)), ("l", (0, 1))):
            ni, nj = i + di, j + dj
            if ni <= n1 and nj <= n2 and ni >= 1 and nj >= 1:
                rec(ni, nj, moves + [(mv, ni, nj)])

    for s in starts:
        rec(s[0], s[1], [(None, s[0], s[1])])
    return res


def edr_count(a, b, match):
    """Zero-boundary EDR: any boundary node is a free start (prefixes of
    either side are skipped at no cost); diagonal moves cost 0 on a match,
    else 1; other moves cost 1."""
    n1, n2 = len(a), len(b)
    starts = [(i, 0) for i in range(n1 + 1)] + [(0, j) for j in range(1, n2 + 1)]
    best = None
    for p in _grid_paths(starts, n1, n2):
        c = 0
        for mv, i, j in p[1:]:
            c += (0 if match(a[i - 1], b[j - 1]) else 1) if mv == "d" else 1
        best = c if best is None else min(best, c)
    return best


def lcss_count(a, b, match):
    """Largest chain of index pairs (i_k, j_k), strictly increasing in both,
    that all match."""
    n1, n2 = len(a), len(b)
    best = 0
    for k in range(1, min(n1, n2) + 1):
        for I in itertools.combinations(range(n1), k):
            for J in itertools.combinations(range(n2), k):
                if all(match(a[i], b[j]) for i, j in zip(I, J)):
                    best = k
                    break
            if best == k:
                break
    return best


def erp(a, b, g):
    """ERP with the constant-sum boundary: a path may start at (0,0) for 0,
    at (i,0), i >= 1, for Sa, or at (0,j), j >= 1, for Sb; a diagonal move
    into (i,j) costs d(a_i, b_j), a move down costs |a_i - g|, a move right
    |b_j - g|."""
    n1, n2 = len(a), len(b)
    ga = [_d(p, g) for p in a]
    gb = [_d(p, g) for p in b]
    sa, sb = sum(ga), sum(gb)
    starts = [(0, 0)] + [(i, 0) for i in range(1, n1 + 1)] + [(0, j) for j in range(1, n2 + 1)]
    best = math.inf
    for p in _grid_paths(starts, n1, n2):
        _, i0, j0 = p[0]
        c = 0.0 if (i0, j0) == (0, 0) else (sa if j0 == 0 else sb)
        for mv, i, j in p[1:]:
            c += _d(a[i - 1], b[j - 1]) if mv == "d" else (ga[i - 1] if mv == "u" else gb[j - 1])
        best = min(best, c)
    return best


def point_segment(p, s1, s2):
    """Textbook point-to-segment distance (projection clamped to [0, 1])."""
    p, s1, s2 = (np.asarray(v, dtype=np.float64) for v in (p, s1, s2))
    d = s2 - s1
    L2 = float(d @ d)
    if L2 == 0.0:
        return float(np.linalg.norm(p - s1))
    t = min(1.0, max(0.0, float((p - s1) @ d) / L2))
    return float(np.linalg.norm(p - (s1 + t * d)))


def hausdorff(a, b):
    """max over both directions of max over points of min over the other
    side's segments (0 when the other side has a single point)."""
    def directed(x, y):
        if len(y) < 2:
            return 0.0
        return max(min(point_segment(p, y[k], y[k + 1]) for k in range(len(y) - 1)) for p in x)

    return max(directed(a, b), directed(b, a))
