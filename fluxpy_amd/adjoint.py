"""Gradients of a temperature loss by the adjoint radiosity solve.

The reference has no derivatives; this module differentiates its
steady-state map (model.py:8-42 in the single-solve form
T = ((1 - rho)(E + F B) / (emiss sigma))^(1/4)) with respect to the
per-facet albedo rho, emissivity emiss and insolation E, for k suns at once:

    (I - diag(rho) F) B = E            forward radiosity (solve.py:25-45,
                                       'left' placement; B kept per sun)
    Phi = E + F B,  T = ((1 - rho) Phi / (emiss sigma))^(1/4)

Backward, for gT = dL/dT (flux_temperature_adjoint_seed / _finish):

    gPhi = gT (1 - rho) / (4 T^3 emiss sigma)
    (I - F^T diag(rho)) lam = F^T gPhi  adjoint solve on the F^T product
    dL/dE = gPhi + lam
    dL/drho = sum_s [lam (F B) - gT T / (4 (1 - rho))]
    dL/demiss = -sum_s gT T / (4 emiss)

F may be a CompressedFormFactorMatrix (the F^T product is the transposed
near field plus the far-field factors read transposed), a FormFactorMatrix
or a SciPy sparse matrix.  Every product, solve and element-wise step runs
in libfluxhip; torch holds the tensors and records the graph.
"""
import numpy as np
import scipy.sparse
from scipy.constants import Stefan_Boltzmann

from . import _kernels
from . import _lib
from .solve import _CsrOperator, apply_original, device_operator, solve_radiosity


class _Transposed:
    """F^T as an operator solve_radiosity accepts."""

    def __init__(self, FF):
        self.FF = FF
        self.shape = FF.shape[::-1]

    def device_operator(self):
        FF = self.FF
        if hasattr(FF, "device_operator"):
            return FF.device_operator(transpose=True)
        if hasattr(FF, "transpose_csr"):
            return _CsrOperator(FF.transpose_csr())
        if scipy.sparse.issparse(FF):
            return _CsrOperator(_kernels.DeviceCSR.from_scipy(scipy.sparse.csr_matrix(FF.T)))
        raise TypeError(f"no adjoint product for {type(FF).__name__}")


def transposed(FF):
    """An operator whose products are F^T x (for solve_radiosity)."""
    return _Transposed(FF)


def _forward(FF, rho, emiss, E, sigma, tol):
    """B, F B and T for device tensors rho (n), emiss (n), E (n x k)."""
    B, niter = solve_radiosity(FF, E, rho, 'left', 'jacobi', tol=tol)
    FB = apply_original(device_operator(FF), B)
    Phi = FB.clone()
    _kernels.axpby(None, E, None, Phi)             # Phi = E + F B
    zero = _lib.torch().zeros_like(Phi)
    T = _kernels.temperature(Phi, zero, rho, emiss, 0.0, True, sigma)
    return B, FB, T, niter


def _backward(FF, rho, emiss, FB, T, gT, sigma, tol):
    t = _lib.torch()
    n, k = T.shape
    dev = T.device
    gPhi = t.empty_like(T)
    grho = t.empty(n, dtype=t.float64, device=dev)
    gemiss = t.empty_like(grho)
    st = _lib.stream()
    gTc = gT.contiguous().to(t.float64)
    _lib.call("flux_temperature_adjoint_seed", n, k, T.data_ptr(), gTc.data_ptr(),
              rho.data_ptr(), emiss.data_ptr(), float(sigma), gPhi.data_ptr(),
              grho.data_ptr(), gemiss.data_ptr(), st)
    FT = transposed(FF)
    gB = apply_original(FT.device_operator(), gPhi)
    # (I - F^T diag(rho)) lam = gB  is the 'right'-placement system of F^T
    lam, niter = solve_radiosity(FT, gB, rho, 'right', 'jacobi', tol=tol)
    lam = lam.reshape(n, k).contiguous()
    gE = t.empty_like(T)
    _lib.call("flux_temperature_adjoint_finish", n, k, gPhi.data_ptr(), lam.data_ptr(),
              FB.data_ptr(), gE.data_ptr(), grho.data_ptr(), st)
    return grho, gemiss, gE, niter


def _as_dev64(x, n=None):
    t = _lib.require_gpu()
    if isinstance(x, t.Tensor):
        out = x.to(device=_lib.device(), dtype=t.float64)
    else:
        out = _lib.to_dev(np.asarray(x, dtype=np.float64))
    if n is not None and out.numel() == 1:
        out = out.reshape(1).expand(n)
    return out.contiguous()


def _function():
    t = _lib.torch()

    class RadiosityTemperature(t.autograd.Function):
        """T(rho, emiss, E) with the adjoint backward described above.
        forward(rho (n), emiss (n), E (n x k), FF, sigma, tol) -> T (n x k)."""

        @staticmethod
        def forward(ctx, rho, emiss, E, FF, sigma, tol):
            B, FB, T, niter = _forward(FF, rho, emiss, E, sigma, tol)
            ctx.FF, ctx.sigma, ctx.tol = FF, sigma, tol
            ctx.niter_forward = niter
            ctx.save_for_backward(rho, emiss, FB, T)
            return T

        @staticmethod
        def backward(ctx, gT):
            rho, emiss, FB, T = ctx.saved_tensors
            grho, gemiss, gE, _ = _backward(ctx.FF, rho, emiss, FB, T, gT, ctx.sigma, ctx.tol)
            return grho, gemiss, gE, None, None, None

    return RadiosityTemperature


_FN = None


def steady_state_temperature(FF, rho, emiss, E, sigma=Stefan_Boltzmann, tol=None):
    """Differentiable T = ((1 - rho)(E + F B) / (emiss sigma))^(1/4) with
    (I - diag(rho) F) B = E, for device (or host) rho (n), emiss (n) and E
    (n or n x k).  Returns a device tensor of E's shape; rho, emiss and E
    that require grad receive gradients through the adjoint solve.
    ``tol``: Jacobi tolerance of both solves (default: the reference's
    finfo(float64).resolution * max|rhs|)."""
    global _FN
    t = _lib.require_gpu()
    if _FN is None:
        _FN = _function()
    n = FF.shape[0]
    leaf = [v if isinstance(v, t.Tensor) else None for v in (rho, emiss, E)]
    r = _as_dev64(rho, n) if leaf[0] is None or leaf[0].device.type != "cuda" else rho
    e = _as_dev64(emiss, n) if leaf[1] is None or leaf[1].device.type != "cuda" else emiss
    x = _as_dev64(E) if leaf[2] is None or leaf[2].device.type != "cuda" else E
    for v, name in ((r, "rho"), (e, "emiss")):
        if v.dtype != t.float64 or v.dim() != 1 or v.shape[0] != n:
            raise ValueError(f"{name} must be {n} FP64 values")
    if x.dtype != t.float64 or x.shape[0] != n:
        raise ValueError(f"E must have {n} FP64 rows")
    vec = x.dim() == 1
    X = x.reshape(n, -1)
    T = _FN.apply(r.contiguous(), e.contiguous(), X.contiguous(), FF, float(sigma), tol)
    return T.reshape(n) if vec else T


def temperature_loss_gradient(FF, rho, emiss, E, T_obs, sigma=Stefan_Boltzmann, tol=None):
    """L = sum (T - T_obs)^2 and (dL/drho, dL/demiss, dL/dE) as device
    tensors, by one forward and one adjoint solve."""
    t = _lib.require_gpu()
    r = _as_dev64(rho, FF.shape[0]).clone().requires_grad_(True)
    e = _as_dev64(emiss, FF.shape[0]).clone().requires_grad_(True)
    x = _as_dev64(E).clone().requires_grad_(True)
    T = steady_state_temperature(FF, r, e, x, sigma, tol)
    To = _as_dev64(T_obs).reshape(T.shape)
    diff = T - To
    L = (diff * diff).sum()
    gr, ge, gx = t.autograd.grad(L, (r, e, x))
    return float(L.detach().cpu()), gr, ge, gx
