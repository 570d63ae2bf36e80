"""Form-factor matrices on the GPU (drop-in for flux.form_factors).

``get_form_factor_matrix`` keeps the reference's signature and return type
(a SciPy CSR, form_factors.py:11-72) but assembles on the GPU in one fused
pass: prefilter + ray test into a bitmask, per-row counts, scan, CSR fill.

``FormFactorMatrix`` is the LinearOperator the reference sketches
(form_factors.py:74-236) made usable: ``from_shape_model`` assembles the same
entries as ``get_form_factor_matrix`` into a device CSR and ``@`` runs on the
GPU.  NumPy inputs return NumPy; device tensors stay on the device.
"""
import pickle

import numpy as np
import scipy.sparse
import scipy.sparse.linalg

from . import _kernels
from . import _lib
from . import config


def get_form_factor_matrix(shape_model, I=None, J=None, eps=None):
    """Sparse midpoint-rule form factors F[I, J] (form_factors.py:11-72).

    F_ij = max(0, N_i.(P_j-P_i)) max(0, N_j.(P_i-P_j)) A_j / (pi |P_i-P_j|^4)
    for pairs whose unnormalised product exceeds ``eps`` and that the
    backend's ray test declares visible; the diagonal is zero.  Columns are
    positions in J.  Returns ``scipy.sparse.csr_matrix`` in the model dtype.
    """
    return _kernels.ff_csr(shape_model, I, J, eps).to_scipy()


class FormFactorMatrix(scipy.sparse.linalg.LinearOperator):
    """Form-factor LinearOperator backed by a device CSR."""

    def __init__(self, shape_model, I=None, J=None, eps=None, _csr=None):
        self.shape_model = shape_model
        self.I = np.arange(self.num_faces) if I is None else np.asarray(I)
        self.J = np.arange(self.num_faces) if J is None else np.asarray(J)
        self.eps = config.DEFAULT_EPS if eps is None else eps
        self._csr = _csr
        self._csr_t = None

    @classmethod
    def from_shape_model(cls, shape_model, I=None, J=None, eps=None):
        """Assemble F[I, J] on the GPU now and keep it resident."""
        ff = cls(shape_model, I, J, eps)
        ff.assemble()
        return ff

    @classmethod
    def from_csr(cls, shape_model, csr, I=None, J=None, eps=None):
        """Wrap an existing (SciPy or device) CSR of F[I, J]."""
        if not isinstance(csr, _kernels.DeviceCSR):
            csr = _kernels.DeviceCSR.from_scipy(csr)
        return cls(shape_model, I, J, eps, _csr=csr)

    def assemble(self):
        if self._csr is None:
            self._csr = _kernels.ff_csr(self.shape_model, self.I, self.J, self.eps)
        return self._csr

    @staticmethod
    def from_file(path):
        """Load a matrix saved by ``save`` (whitelisting loader of
        ``fluxpy_amd.persist``)."""
        from . import persist
        return persist.load(path)

    def save(self, path):
        with open(path, 'wb') as f:
            pickle.dump(self, f)

    def __reduce__(self):
        sp = self.tocsr()
        return (_rebuild_ffm, (self.shape_model, self.I, self.J, self.eps,
                               sp.data, sp.indices, sp.indptr, sp.shape))

    # -- reference properties (form_factors.py:95-137) ---------------------

    @property
    def num_faces(self):
        return self.shape_model.num_faces

    @property
    def dtype(self):
        return self.P.dtype

    @property
    def shape(self):
        return (len(self.I), len(self.J))

    @property
    def N(self):
        return self.shape_model.N

    @property
    def P(self):
        return self.shape_model.P

    @property
    def A(self):
        return self.shape_model.A

    @property
    def NI(self):
        return self.N[self.I]

    @property
    def PI(self):
        return self.P[self.I]

    @property
    def NI_PI(self):
        return np.sum(self.NI * self.PI, axis=1)

    @property
    def NJ(self):
        return self.N[self.J]

    @property
    def PJ(self):
        return self.P[self.J]

    @property
    def NJ_PJ(self):
        return np.sum(self.NJ * self.PJ, axis=1)

    # -- device access -----------------------------------------------------

    @property
    def device_csr(self):
        return self.assemble()

    @property
    def nbytes(self):
        return self.assemble().nbytes

    @property
    def nnz(self):
        return self.assemble().nnz

    def tocsr(self):
        return self.assemble().to_scipy()

    def toarray(self):
        return self.tocsr().toarray()

    def transpose_csr(self):
        """Device CSR of F^T (built once, for the adjoint products)."""
        if self._csr_t is None:
            self._csr_t = _kernels.csr_transpose(self.assemble())
        return self._csr_t

    # -- products ----------------------------------------------------------

    def apply(self, X, out=None):
        """Y = F X for a device tensor X (n or n x k)."""
        return self.assemble().matmat(X, out=out)

    def apply_t(self, X, out=None):
        """Y = F^T X for a device tensor X."""
        return self.transpose_csr().matmat(X, out=out)

    def _host_product(self, x, transpose=False):
        x = np.asarray(x)
        if x.dtype not in (np.float32, np.float64):
            x = x.astype(np.float64)
        xd = _lib.to_dev(x)
        y = self.apply_t(xd) if transpose else self.apply(xd)
        return y.cpu().numpy()

    def _matvec(self, x):
        return self._host_product(np.ravel(x)).reshape(-1)

    def _matmat(self, X):
        return self._host_product(X)

    def _rmatvec(self, x):
        return self._host_product(np.ravel(x), transpose=True).reshape(-1)

    def _rmatmat(self, X):
        return self._host_product(X, transpose=True)

    def __matmul__(self, x):
        t = _lib.torch()
        if isinstance(x, t.Tensor):
            return self.apply(x)
        return super().__matmul__(x)

    def dot(self, x):
        t = _lib.torch()
        if isinstance(x, t.Tensor):
            return self.apply(x)
        return super().dot(x)


def _rebuild_ffm(shape_model, I, J, eps, data, indices, indptr, shape):
    sp = scipy.sparse.csr_matrix((data, indices, indptr), shape=shape)
    return FormFactorMatrix.from_csr(shape_model, sp, I, J, eps)
