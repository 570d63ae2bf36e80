"""Timing and byte-count helpers (drop-in for flux.util, util.py:1-24).

``tic()`` / ``toc()`` share one module-level stopwatch, as the reference's
do; ``nbytes`` counts what an array holds, and for a SciPy CSR matrix the
three arrays it is stored in (the byte rule the compressed tree build
compares blocks with).
"""
import time

import scipy.sparse


class _Stopwatch:
    """The module's stopwatch: the perf_counter value of the last tic()."""
    started = None


_watch = _Stopwatch()


def tic():
    """Start, or restart, the stopwatch."""
    _watch.started = time.perf_counter()


def toc():
    """Seconds since the last tic()."""
    if _watch.started is None:
        raise RuntimeError("toc() called before tic()")
    return time.perf_counter() - _watch.started


def nbytes(arg):
    """Bytes of an array; a CSR matrix counts its values, column indices and
    row pointers."""
    if isinstance(arg, scipy.sparse.csr_matrix):
        return arg.data.nbytes + arg.indices.nbytes + arg.indptr.nbytes
    return arg.nbytes
