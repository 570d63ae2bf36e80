"""Octant ordering of a 3-D point set (drop-in for flux.octree).

Same semantics as octree.py:5-18 of the reference (bounding-box midpoint,
<= goes low, octants in itertools.product([<=, >], repeat=3) order).
"""
import itertools as it

import numpy as np


def get_octant_order(X, bbox=None):
    X = np.asarray(X)
    if bbox is not None:
        lo_b = np.array([b[0] for b in bbox])
        hi_b = np.array([b[1] for b in bbox])
    else:
        lo_b = np.min(X, axis=0)
        hi_b = np.max(X, axis=0)
    c = (lo_b + hi_b) / 2
    low = [X[:, d] <= c[d] for d in range(3)]
    out = []
    for s in it.product((True, False), repeat=3):
        m = np.ones(X.shape[0], dtype=bool)
        for d in range(3):
            m &= low[d] if s[d] else ~low[d]
        out.append(np.flatnonzero(m))
    return out
