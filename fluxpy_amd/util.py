"""Timing and byte-count helpers (drop-in for flux.util, util.py:1-24)."""
import time
from functools import singledispatch

import scipy.sparse

t0 = None


def tic():
    global t0
    t0 = time.perf_counter()


def toc():
    return time.perf_counter() - t0


@singledispatch
def nbytes(arg):
    return arg.nbytes


@nbytes.register
def _(arg: scipy.sparse.csr_matrix):
    return arg.data.nbytes + arg.indices.nbytes + arg.indptr.nbytes
