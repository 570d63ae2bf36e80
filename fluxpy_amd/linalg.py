"""Rank estimation of sparse blocks on the GPU (drop-in for flux.linalg,
linalg.py:8-62).

``estimate_rank`` keeps the reference's contract: k = 40 doubling, capped at
min(shape) - 1; the rank is the first index r with s_r / s_0 <= tol; None
once the k-term CSR factors need at least ``max_nbytes``; 0 / 1 for empty /
1 x 1 inputs; returns (U[:, :r], S[:r], Vt[:r], s_r / s_0) with U, Vt as
SciPy CSR.  The singular values and vectors come from the GPU sketch used by
the compressed tree build (Rademacher range sketch, FP64 Gram matrices,
batched Jacobi eigensolver) instead of ARPACK.
"""
import numpy as np
import scipy.sparse

from .util import nbytes


def _device_pair(spmat):
    from . import _kernels
    A = scipy.sparse.csr_matrix(spmat)
    if A.dtype not in (np.float32, np.float64):
        A = A.astype(np.float64)
    A.sort_indices()
    F = _kernels.DeviceCSR.from_scipy(A)
    return A, F, _kernels.csr_transpose(F)


def _svd_top(A, F, FT, k, seed=0x5eed):
    """(U (m x k), S (k,), Vt (k x n)) of the leading singular triplets."""
    from ._hbuild import L_MAX, _SvdBatch
    m, n = A.shape
    mm = min(m, n)
    l = min(2 * k + 8, mm, L_MAX)
    sb = _SvdBatch(F, FT, [(0, m, 0, n)], [l], seed)
    kk = min(k, int(sb.kept[0]))
    if kk == 0:
        return np.zeros((m, 0)), np.zeros(0), np.zeros((0, n)), sb.sigma[0]
    u, s, vt = sb.factors([(0, kk)], _torch_dtype(A.dtype))[0]
    return u.cpu().numpy(), s.cpu().numpy(), vt.cpu().numpy(), sb.sigma[0]


def _torch_dtype(dt):
    from . import _lib
    return _lib.torch_dtype(dt)


def sparse_svd(spmat, k):
    """Leading k singular triplets, descending, U and Vt as CSR (linalg.py:8-17)."""
    A, F, FT = _device_pair(spmat)
    U, S, Vt, _ = _svd_top(A, F, FT, k)
    pad = k - S.size
    if pad > 0:
        U = np.hstack([U, np.zeros((U.shape[0], pad), U.dtype)])
        Vt = np.vstack([Vt, np.zeros((pad, Vt.shape[1]), Vt.dtype)])
        S = np.concatenate([S, np.zeros(pad, S.dtype)])
    return scipy.sparse.csr_matrix(U), S, scipy.sparse.csr_matrix(Vt)


def estimate_rank(spmat, tol, max_nbytes=None, k0=40):
    assert tol < 1
    if spmat.shape[0] == 0 or spmat.shape[1] == 0:
        return 0
    if spmat.shape == (1, 1):
        return 1
    A, F, FT = _device_pair(spmat)
    m, k = min(spmat.shape), k0
    while True:
        k = min(k, m - 1)
        if not k >= 1:
            raise RuntimeError('bad value of k')
        U, S, Vt, sig = _svd_top(A, F, FT, k)
        pad = k - S.size
        if pad > 0:
            U = np.hstack([U, np.zeros((U.shape[0], pad), U.dtype)])
            Vt = np.vstack([Vt, np.zeros((pad, Vt.shape[1]), Vt.dtype)])
            S = np.concatenate([S, np.zeros(pad, S.dtype)])
        Uc, Vtc = scipy.sparse.csr_matrix(U), scipy.sparse.csr_matrix(Vt)
        svd_nbytes = nbytes(Uc) + nbytes(S) + nbytes(Vtc)
        if max_nbytes is not None and svd_nbytes >= max_nbytes:
            return None
        thresh = S / S[0]
        below = np.where(thresh <= tol)[0]
        if below.size > 0:
            r = below[0]
            return Uc[:, :r], S[:r], Vtc[:r, :], thresh[r]
        if k == m - 1:
            return None
        k *= 2


def nnz(mat, tol=None):
    if tol is None:
        return (mat != 0).sum()
    return (abs(mat) < tol).sum()


def sparsity(mat, tol=None):
    size = mat.size
    return 0 if size == 0 else nnz(mat, tol) / size
