"""Quadrant ordering of a 2-D point set (drop-in for flux.quadtree).

Same semantics as quadtree.py:5-18 of the reference: split at the midpoint of
the set's bounding box (or of `bbox`), points on the midpoint go to the
lower side, quadrants in itertools.product([<=, >], repeat=2) order, indices
ascending within each quadrant.
"""
import itertools as it

import numpy as np


def get_quadrant_order(X, bbox=None):
    X = np.asarray(X)
    if bbox is not None:
        (xmin, xmax), (ymin, ymax) = bbox
    else:
        xmin, ymin = np.min(X, axis=0)
        xmax, ymax = np.max(X, axis=0)
    xc, yc = (xmin + xmax) / 2, (ymin + ymax) / 2
    x, y = X[:, 0], X[:, 1]
    lo = (x <= xc, y <= yc)
    out = []
    for sx, sy in it.product((True, False), repeat=2):
        mx = lo[0] if sx else ~lo[0]
        my = lo[1] if sy else ~lo[1]
        out.append(np.flatnonzero(mx & my))
    return out
