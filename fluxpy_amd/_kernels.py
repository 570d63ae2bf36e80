"""Device-side building blocks composed from libfluxhip entry points.

Holds the device CSR container and the form-factor assembly, insolation,
Jacobi and vector helpers that the drop-in modules (form_factors, solve,
model, compressed_form_factors) are built on.  All arithmetic runs in the
HIP library; torch only allocates device memory and supplies the stream.
"""
import numpy as np
import scipy.sparse

from . import _lib
from . import config

# Row chunk of the form-factor mask so that the bitmask stays below ~1 GiB.
_MASK_BYTES = 1 << 30


class DeviceCSR:
    """CSR matrix resident on the GPU: int64 indptr, int32 indices, values in
    float32 or float64."""

    def __init__(self, indptr, indices, data, shape):
        self.indptr = indptr
        self.indices = indices
        self.data = data
        self.shape = tuple(int(s) for s in shape)

    @property
    def dtype(self):
        return np.dtype(str(self.data.dtype).replace("torch.", ""))

    @property
    def nnz(self):
        return int(self.indices.numel())

    @property
    def nbytes(self):
        return (self.indptr.numel() * 8 + self.indices.numel() * 4
                + self.data.numel() * self.data.element_size())

    @classmethod
    def from_scipy(cls, A, dtype=None):
        A = scipy.sparse.csr_matrix(A)
        A.sum_duplicates()
        if dtype is not None:
            A = A.astype(dtype)
        if A.dtype not in (np.float32, np.float64):
            A = A.astype(np.float64)
        return cls(_lib.to_dev(A.indptr.astype(np.int64)),
                   _lib.to_dev(A.indices.astype(np.int32)),
                   _lib.to_dev(A.data), A.shape)

    def to_scipy(self):
        return scipy.sparse.csr_matrix(
            (self.data.cpu().numpy(), self.indices.cpu().numpy(),
             self.indptr.cpu().numpy()), shape=self.shape)

    def astype(self, torch_dtype):
        """Same pattern, values converted on the GPU (round to nearest even)."""
        t = _lib.torch()
        src = self.data
        if src.dtype == torch_dtype:
            return DeviceCSR(self.indptr, self.indices, src.clone(), self.shape)
        if src.dtype == t.float64 and torch_dtype in (t.bfloat16, t.float16):
            src = convert(src, t.float32)
        return DeviceCSR(self.indptr, self.indices, convert(src, torch_dtype), self.shape)

    def matmat(self, X, out=None):
        """Y = A X for a device block X (n or n x k), in X's dtype."""
        t = _lib.torch()
        vec = X.dim() == 1
        X2 = X.reshape(X.shape[0], -1).contiguous()
        if X2.shape[0] != self.shape[1]:
            raise ValueError(f"dimension mismatch: {self.shape} @ {tuple(X.shape)}")
        k = X2.shape[1]
        Y = out if out is not None else t.empty((self.shape[0], k), dtype=X2.dtype,
                                                 device=X2.device)
        _lib.call("flux_csr_spmm", self.shape[0], self.shape[1], self.indptr.data_ptr(),
                  self.indices.data_ptr(), self.data.data_ptr(),
                  _lib.dtype_code(self.data.dtype), X2.data_ptr(), k, Y.data_ptr(),
                  _lib.dtype_code(X2.dtype), _lib.stream())
        return Y.reshape(-1) if vec else Y


def csr_transpose(A):
    """Device CSR of A^T (rows sorted), built on the GPU."""
    t = _lib.require_gpu()
    m, n = A.shape
    dev = A.data.device
    t_indptr = t.empty(n + 1, dtype=t.int64, device=dev)
    t_indices = t.empty(max(A.nnz, 1), dtype=t.int32, device=dev)
    t_data = t.empty(max(A.nnz, 1), dtype=A.data.dtype, device=dev)
    if n == 0:
        t_indptr.zero_()
    _lib.call("flux_csr_transpose", m, n, A.indptr.data_ptr(), A.indices.data_ptr(),
              A.data.data_ptr(), _lib.dtype_code(A.data.dtype), t_indptr.data_ptr(),
              t_indices.data_ptr(), t_data.data_ptr(), _lib.stream())
    return DeviceCSR(t_indptr, t_indices[:A.nnz], t_data[:A.nnz], (n, m))


# --------------------------------------------------------------------------
# Form factors
# --------------------------------------------------------------------------

def ff_csr(shape_model, I=None, J=None, eps=None, dtype=None):
    """Device CSR of get_form_factor_matrix(shape_model, I, J, eps)
    (form_factors.py:11-72): rows I, column positions into J."""
    t = _lib.require_gpu()
    if eps is None:
        eps = config.DEFAULT_EPS
    nf = shape_model.num_faces
    I = np.arange(nf, dtype=np.int64) if I is None else np.asarray(I).astype(np.int64).ravel()
    J = np.arange(nf, dtype=np.int64) if J is None else np.asarray(J).astype(np.int64).ravel()
    for X in (I, J):
        if X.size and (X.min() < 0 or X.max() >= nf):
            raise IndexError("facet index out of range")
    mdt = np.dtype(shape_model.dtype)
    if mdt not in (np.float32, np.float64):
        raise RuntimeError(f"unsupported dtype {mdt}")
    dev = _lib.device()
    tdt = _lib.torch_dtype(mdt)
    P = _lib.to_dev(shape_model.P)
    N = _lib.to_dev(shape_model.N)
    A = _lib.to_dev(shape_model.A)
    m, n = I.size, J.size
    W = (n + 63) // 64
    Jd = _lib.to_dev(J)
    st = _lib.stream()
    code = _lib.dtype_code(mdt)
    rows_per_chunk = max(1, min(m, _MASK_BYTES // max(1, W * 8)))
    parts = []
    for r0 in range(0, m, rows_per_chunk):
        Ic = I[r0:r0 + rows_per_chunk]
        mc = Ic.size
        Id = _lib.to_dev(Ic)
        mask = t.empty((mc, max(W, 1)), dtype=t.int64, device=dev)
        row_nnz = t.empty(mc, dtype=t.int64, device=dev)
        _lib.call("flux_ff_mask", shape_model._scene.handle, P.data_ptr(), N.data_ptr(), code,
                  Id.data_ptr(), mc, Jd.data_ptr(), n, float(eps), shape_model._mode,
                  mask.data_ptr(), row_nnz.data_ptr(), st)
        indptr = t.empty(mc + 1, dtype=t.int64, device=dev)
        total = _lib.ctypes.c_int64()
        _lib.call("flux_scan_i64", row_nnz.data_ptr(), mc, indptr.data_ptr(),
                  _lib.ctypes.byref(total), st)
        nnz = total.value
        indices = t.empty(max(nnz, 1), dtype=t.int32, device=dev)
        data = t.empty(max(nnz, 1), dtype=tdt, device=dev)
        _lib.call("flux_ff_fill", P.data_ptr(), N.data_ptr(), A.data_ptr(), code, Id.data_ptr(),
                  mc, Jd.data_ptr(), n, mask.data_ptr(), indptr.data_ptr(),
                  indices.data_ptr(), data.data_ptr(), st)
        parts.append((indptr, indices[:nnz], data[:nnz]))
        del mask
    if not parts:  # m == 0
        return DeviceCSR(t.zeros(1, dtype=t.int64, device=dev),
                         t.empty(0, dtype=t.int32, device=dev),
                         t.empty(0, dtype=tdt, device=dev), (m, n))
    if len(parts) == 1:
        indptr, indices, data = parts[0]
    else:
        offs = 0
        ip = [t.zeros(1, dtype=t.int64, device=dev)]
        for p_, _, _ in parts:
            ip.append(p_[1:] + offs)
            offs += int(p_[-1])
        indptr = t.cat(ip)
        indices = t.cat([q for _, q, _ in parts])
        data = t.cat([d for _, _, d in parts])
    csr = DeviceCSR(indptr.contiguous(), indices.contiguous(), data.contiguous(), (m, n))
    if dtype is not None:
        csr = csr.astype(dtype)
    return csr


# --------------------------------------------------------------------------
# Insolation
# --------------------------------------------------------------------------

def direct_irradiance(shape_model, F0, Dsuns, device=False):
    """N x M block: column k = 1-D-sun insolation for direction Dsuns[k]."""
    t = _lib.require_gpu()
    D = np.ascontiguousarray(np.atleast_2d(np.asarray(Dsuns, dtype=np.float64)))
    if D.shape[1] != 3:
        raise ValueError("Dsuns must have shape (M, 3)")
    M = D.shape[0]
    dt = _lib.torch_dtype(shape_model.dtype)
    E = t.empty((shape_model.num_faces, M), dtype=dt, device=_lib.device())
    Nd = _lib.to_dev(shape_model.N)
    Pd = _lib.to_dev(shape_model.P)
    Dd = _lib.to_dev(D)
    _lib.call("flux_direct_irradiance", shape_model._scene.handle, Pd.data_ptr(),
              Nd.data_ptr(), _lib.dtype_code(shape_model.dtype), Dd.data_ptr(), M, float(F0),
              shape_model._mode, E.data_ptr(), _lib.stream())
    return E if device else E.cpu().numpy()


# --------------------------------------------------------------------------
# Vector helpers (per-column, deterministic)
# --------------------------------------------------------------------------

class VecOps:
    """Reusable work buffers for the per-column reductions of one block size."""

    def __init__(self, nrhs):
        t = _lib.require_gpu()
        self.nrhs = nrhs
        self.work = t.empty(max(64 * nrhs, 1), dtype=t.float64, device=_lib.device())
        self.out = t.empty(max(nrhs, 1), dtype=t.float64, device=_lib.device())

    def _reduce(self, x, y, mode):
        n = x.shape[0]
        k = x.shape[1] if x.dim() == 2 else 1
        _lib.call("flux_vec_reduce", n, k, x.data_ptr(), _lib.ptr(y), _lib.dtype_code(x.dtype),
                  mode, self.out.data_ptr(), self.work.data_ptr(), _lib.stream())
        return self.out[:k]

    def dot(self, x, y):
        return self._reduce(x, y, 0)

    def maxdiff(self, x, y):
        return self._reduce(x, y, 1)

    def maxabs(self, x):
        return self._reduce(x, None, 2)


def axpby(alpha, x, beta, y):
    """y[:, c] = alpha[c] x[:, c] + beta[c] y[:, c] (alpha/beta: FP64 device
    vectors or None for 1)."""
    n = x.shape[0]
    k = x.shape[1] if x.dim() == 2 else 1
    _lib.call("flux_vec_axpby", n, k, _lib.ptr(alpha), x.data_ptr(), _lib.ptr(beta),
              y.data_ptr(), _lib.dtype_code(x.dtype), _lib.stream())
    return y


def rowscale(rho, x, out=None, one_minus=False, add=0.0):
    """out[i, :] = s_i x[i, :] + add with s_i = rho_i (or 1 - rho_i)."""
    t = _lib.torch()
    out = t.empty_like(x) if out is None else out
    n = x.shape[0]
    k = x.shape[1] if x.dim() == 2 else 1
    rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
    _lib.call("flux_vec_rowscale", n, k, _lib.ptr(rho_t), rho_s, int(one_minus), float(add),
              x.data_ptr(), out.data_ptr(), _lib.dtype_code(x.dtype), _lib.stream())
    return out


def clamp0(x, out=None):
    t = _lib.torch()
    out = t.empty_like(x) if out is None else out
    _lib.call("flux_vec_clamp0", x.numel(), x.data_ptr(), out.data_ptr(),
              _lib.dtype_code(x.dtype), _lib.stream())
    return out


def convert(x, torch_dtype):
    t = _lib.torch()
    if x.dtype == torch_dtype:
        return x.clone()
    out = t.empty(x.shape, dtype=torch_dtype, device=x.device)
    _lib.call("flux_vec_convert", x.numel(), x.data_ptr(), _lib.dtype_code(x.dtype),
              out.data_ptr(), _lib.dtype_code(torch_dtype), _lib.stream())
    return out


def temperature(B, Q, rho, emiss, fsurf=0.0, clamp=True, sigma=None):
    t = _lib.torch()
    from scipy.constants import Stefan_Boltzmann
    sigma = Stefan_Boltzmann if sigma is None else sigma
    out = t.empty_like(B)
    n = B.shape[0]
    k = B.shape[1] if B.dim() == 2 else 1
    rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
    em_t, em_s = (emiss, 0.0) if isinstance(emiss, t.Tensor) else (None, float(emiss))
    _lib.call("flux_temperature", n, k, B.data_ptr(), Q.data_ptr(), _lib.ptr(rho_t), rho_s,
              _lib.ptr(em_t), em_s, float(fsurf), int(clamp), float(sigma), out.data_ptr(),
              _lib.dtype_code(B.dtype), _lib.stream())
    return out
