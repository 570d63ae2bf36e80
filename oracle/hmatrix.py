"""CPU oracle of the hierarchical (compressed) form-factor matrix --
TEST INFRASTRUCTURE ONLY (see oracle/oracle.py for the usage rule).

Restates, with NumPy / SciPy / LAPACK:

* partition            quadtree.py:5-18, octree.py:5-18 (bbox midpoint of the
                       subset, <= / > predicates, product([le, gt]) order)
* tree build           compressed_form_factors.py:285-386 (make_block),
                       :512-568 (2d-tree block), :631-658 (quad/oct order)
* rank rule            linalg.py:20-50 (k = 40 doubling, capped at
                       min(shape) - 1; r = first i with s_i/s_0 <= tol; give
                       up when the k-term factors take as many bytes as the
                       sparse block) with EXACT singular values from a
                       dense LAPACK SVD in FP64
* byte rules           util.py:18-24 (CSR = data + indices + indptr, int32
                       indices), FormFactorCsrBlock.nbytes = data bytes only
                       (:219-221), tree nbytes = int64 index arrays +
                       children (:416-420)
* block matvec         compressed_form_factors.py:407-414

Factor sparsity model (the only part not restated from source): the svds of
the SciPy the reference pinned (scipy>=1.5.4, requirements.txt:34) takes
ARPACK eigenpairs of the Gram matrix of the shorter side and forms the other
factor as A v / s (tall) or A^T u / s (wide).  ARPACK's Ritz vectors are
exactly zero on the null lines of that Gram matrix, and the product with A
is exactly zero on the zero lines of A, so a zero row of the block gives a
zero row of U and a zero column gives a zero column of Vt on either shape:
nnz(U) = k * (nonzero rows), nnz(Vt) = k * (nonzero columns).  Pinned by
tests/test_host.py against real ARPACK factors of blocks with zero rows and
columns (tall, wide, square); the SciPy installed here (1.15) keeps the
same Vt pattern and has U dense up to LAPACK rounding on the zero rows.

``force_max_depth=True`` with max_depth > 1 hits ``assert False`` in the
reference (compressed_form_factors.py:300-301); the oracle implements the
recursion that branch was written to perform (always descend, then collapse
all-dense / all-sparse children), which is what the reference's
test_max_depth_2_for_stretched_sphere expects.
"""
import itertools as it

import numpy as np
import scipy.sparse


def quadrant_order(X):
    """quadtree.py:5-18."""
    xmin, ymin = np.min(X, axis=0)
    xmax, ymax = np.max(X, axis=0)
    xc, yc = (xmin + xmax) / 2, (ymin + ymax) / 2
    x, y = X.T
    out = []
    for xop, yop in it.product([np.less_equal, np.greater], repeat=2):
        out.append(np.where(xop(x, xc) & yop(y, yc))[0])
    return out


def octant_order(X):
    """octree.py:5-18."""
    lo = np.min(X, axis=0)
    hi = np.max(X, axis=0)
    c = (lo + hi) / 2
    out = []
    for ops in it.product([np.less_equal, np.greater], repeat=3):
        m = ops[0](X[:, 0], c[0]) & ops[1](X[:, 1], c[1]) & ops[2](X[:, 2], c[2])
        out.append(np.where(m)[0])
    return out


def csr_nbytes(m, nnz, itemsize):
    """util.py:18-24 for a CSR with int32 indices / indptr."""
    return nnz * itemsize + nnz * 4 + (m + 1) * 4


def factor_nnz(A, k):
    """(nnz(U), nnz(Vt)) of a k-term svds factorisation of block A: k per
    nonzero row of A, k per nonzero column of A."""
    nzr = int((np.diff(A.indptr) > 0).sum())
    nzc = int(np.unique(A.indices).size)
    return k * nzr, k * nzc


def svd_nbytes(A, k, itemsize):
    m, n = A.shape
    nu, nv = factor_nnz(A, k)
    return csr_nbytes(m, nu, itemsize) + k * itemsize + csr_nbytes(k, nv, itemsize)


class Leaf:
    def __init__(self, kind, shape, mat=None, u=None, s=None, vt=None, nbytes=0):
        self.kind = kind
        self.shape = shape
        self.mat = mat
        self.u, self.s, self.vt = u, s, vt
        self.nbytes = nbytes
        self.is_leaf = True

    @property
    def depth(self):
        return 0

    def matmat(self, x):
        if self.kind == "zero":
            return np.zeros((self.shape[0], x.shape[1]))
        if self.kind in ("dense", "csr"):
            return self.mat @ x
        return self.u @ (self.s[:, None] * (self.vt @ x))


class Tree:
    is_leaf = False

    def __init__(self, shape):
        self.shape = shape

    @property
    def depth(self):
        return 1 + max(b.depth for b in self.blocks.ravel())

    def matmat(self, x):
        y = np.zeros((self.shape[0], x.shape[1]))
        for i, ri in enumerate(self.row_inds):
            for j, cj in enumerate(self.col_inds):
                b = self.blocks[i, j]
                if b.is_leaf and b.kind == "zero":
                    continue
                y[ri] += b.matmat(x[cj])
        return y

    @property
    def nbytes(self):
        return (sum(8 * r.size for r in self.row_inds) + sum(8 * c.size for c in self.col_inds)
                + sum(b.nbytes for b in self.blocks.ravel()))


def _all_kind(tree, kinds):
    return all(b.is_leaf and b.kind in kinds for b in tree.blocks.ravel())


class HMatrix:
    """Oracle compressed matrix over the full sparse F (original order)."""

    def __init__(self, F, P, tol=1e-5, min_size=16384, max_depth=None, force_max_depth=False,
                 tree="quad"):
        if tol < 0:
            raise RuntimeError('tol should be a nonnegative float')
        self.F = scipy.sparse.csr_matrix(F)
        self.P = P
        self.itemsize = self.F.dtype.itemsize
        self.tol = tol
        self.min_size = min_size
        self.order = quadrant_order if tree == "quad" else octant_order
        self.dims = 2 if tree == "quad" else 3
        self.svd_calls = 0
        n = F.shape[0]
        I = np.arange(n)
        self.root = self._tree(I, I, None, max_depth, force_max_depth)

    # compressed_form_factors.py:514-568
    def _tree(self, I, J, spmat_par, max_depth, force):
        t = Tree((len(I), len(J)))
        t.row_inds = self.order(self.P[I, :self.dims])
        t.col_inds = self.order(self.P[J, :self.dims])
        t.I, t.J = I, J
        B = np.empty((len(t.row_inds), len(t.col_inds)), dtype=object)
        for i, ri in enumerate(t.row_inds):
            for j, cj in enumerate(t.col_inds):
                if spmat_par is None:
                    sp = self.F[I[ri]][:, J[cj]]
                else:
                    sp = spmat_par[ri][:, cj]
                B[i, j] = self._block(I[ri], J[cj], scipy.sparse.csr_matrix(sp), max_depth, force)
        t.blocks = B
        return t

    def _dense(self, sp):
        return Leaf("dense", sp.shape, mat=sp.toarray(), nbytes=sp.shape[0] * sp.shape[1]
                    * self.itemsize)

    def _csr(self, sp):
        return Leaf("csr", sp.shape, mat=sp, nbytes=sp.nnz * self.itemsize)

    # compressed_form_factors.py:285-386
    def _block(self, I, J, sp, max_depth, force):
        if force and max_depth is not None and max_depth > 1:
            child = self._tree(I, J, sp, max_depth - 1, force)
            if _all_kind(child, ("dense",)):
                return self._dense(sp)
            if _all_kind(child, ("csr", "zero")):
                return self._csr(sp)
            return child
        m, n = sp.shape
        if sp.nnz == 0:
            return Leaf("zero", sp.shape)
        nb_sparse = csr_nbytes(m, sp.nnz, self.itemsize)
        if m * n < self.min_size:
            if m * n * self.itemsize < nb_sparse:
                return self._dense(sp)
            return self._csr(sp)
        svd = self._svd(sp, nb_sparse)
        nb_svd = np.inf if svd is None else svd.nbytes
        child = None
        if max_depth is None or max_depth > 1:
            child = self._tree(I, J, sp, None if max_depth is None else max_depth - 1, False)
        nb_child = np.inf if child is None else child.nbytes
        nb_min = min(nb_sparse, nb_svd, nb_child)
        if nb_sparse == nb_min:
            return self._csr(sp)
        if nb_svd == nb_min:
            return svd
        if _all_kind(child, ("dense",)):
            return self._dense(sp)
        if _all_kind(child, ("csr", "zero")):
            return self._csr(sp)
        return child

    # linalg.py:20-50 with exact singular values
    def _svd(self, sp, max_nbytes):
        m, n = sp.shape
        mm = min(m, n)
        k = 40
        s = None
        while True:
            k = min(k, mm - 1)
            if k < 1:
                raise RuntimeError('bad value of k')
            if svd_nbytes(sp, k, self.itemsize) >= max_nbytes:
                return None
            if s is None:
                self.svd_calls += 1
                U, s, Vt = np.linalg.svd(sp.toarray().astype(np.float64), full_matrices=False)
            thresh = s[:k] / s[0]
            below = np.where(thresh <= self.tol)[0]
            if below.size > 0:
                r = int(below[0])
                dt = sp.dtype
                u = U[:, :r].astype(dt)
                vt = Vt[:r].astype(dt)
                ss = s[:r].astype(dt)
                nb = svd_nbytes(sp, r, self.itemsize)
                return Leaf("svd", sp.shape, u=u, s=ss, vt=vt, nbytes=nb)
            if k == mm - 1:
                return None
            k *= 2

    def matmat(self, x):
        x2 = x.reshape(x.shape[0], -1)
        y = self.root.matmat(x2)
        return y.reshape(x.shape[0], *x.shape[1:]) if x.ndim > 1 else y.ravel()

    @property
    def nbytes(self):
        return self.root.nbytes

    def leaves(self):
        """Yield (kind, rows (original ids), cols, leaf) for every leaf."""
        def rec(t):
            for i, ri in enumerate(t.row_inds):
                for j, cj in enumerate(t.col_inds):
                    b = t.blocks[i, j]
                    if b.is_leaf:
                        yield b.kind, t.I[ri], t.J[cj], b
                    else:
                        yield from rec(b)
        yield from rec(self.root)

    def stats(self):
        out = {}
        for kind, r, c, b in self.leaves():
            d = out.setdefault(kind, {"count": 0, "elements": 0, "nbytes": 0})
            d["count"] += 1
            d["elements"] += len(r) * len(c)
            d["nbytes"] += b.nbytes
        return out
