"""Radiosity solvers on the GPU (drop-in for flux.solve, solve.py:4-109).

``solve_radiosity(FF, E, rho, albedo_placement, method, tol)`` keeps the
reference's signature, defaults and return value ``(B, niter)``:

* ``jacobi`` -- B <- E + F (rho B) ('right') or E + rho (F B) ('left'),
  starting at B = E, stopping when max |B1 - B| <= tol with the default
  tol = finfo(E.dtype).resolution * max|E| (solve.py:6-8, 25-45).  Like the
  reference it returns the iterate *before* the last update.
* ``cg``     -- the reference's CG on (I - rho F) (solve.py:47-106),
  column by column for block right-hand sides.
* ``gmres``  -- restarted GMRES(m) with modified Gram-Schmidt in the
  working precision; ``tol`` is the relative residual |r| / |E| per column
  (default 1e-12).  Not in the reference; added for the 1e-12 solves.

Per-facet albedo is ``diag(rho)``: 'right' is F diag(rho), 'left' is
diag(rho) F.  E may be a NumPy array (result returned as NumPy) or a device
tensor (result stays on the device).  Unlike the reference, iterations are
capped by ``maxiter`` (a warning is emitted when it is reached): the
reference loops forever when rho * rowsum(F) >= 1.
"""
import warnings

import numpy as np
import scipy.sparse

from . import _kernels
from . import _lib


class _CsrOperator:
    """Device CSR with the fused Jacobi step."""

    def __init__(self, csr):
        self.csr = csr
        self.shape = csr.shape
        self.fused_jacobi = True

    def apply(self, X, out=None):
        return self.csr.matmat(X, out=out)

    def jacobi_step(self, E, B, rho, left, B1, colmax):
        t = _lib.torch()
        n = B.shape[0]
        k = B.shape[1]
        rho_t, rho_s = (rho, 0.0) if isinstance(rho, t.Tensor) else (None, float(rho))
        c = self.csr
        _lib.call("flux_csr_jacobi_step", n, c.indptr.data_ptr(), c.indices.data_ptr(),
                  c.data.data_ptr(), _lib.dtype_code(c.data.dtype), E.data_ptr(), B.data_ptr(),
                  k, _lib.ptr(rho_t), rho_s, int(left), B1.data_ptr(),
                  _lib.dtype_code(B.dtype), colmax.data_ptr(), _lib.stream())


def _to_internal(op, X):
    return op.to_internal(X) if hasattr(op, "to_internal") else X


def _from_internal(op, X):
    return op.from_internal(X) if hasattr(op, "from_internal") else X


def _rho_internal(op, rho):
    t = _lib.torch()
    if isinstance(rho, t.Tensor) and hasattr(op, "to_internal"):
        return op.to_internal(rho.reshape(-1, 1)).reshape(-1)
    return rho


def apply_original(op, X):
    """F X with X and the result in facet order, whatever order `op` uses."""
    return _from_internal(op, op.apply(_to_internal(op, X)))


def device_operator(FF):
    """The device operator behind FF (ours, a SciPy sparse matrix or a dense
    array).  Anything else is rejected: there is no host execution path."""
    if hasattr(FF, "device_operator"):
        return FF.device_operator()
    if hasattr(FF, "apply") and hasattr(FF, "jacobi_step"):
        return FF  # already a device operator
    if hasattr(FF, "device_csr"):
        return _CsrOperator(FF.device_csr)
    if isinstance(FF, _kernels.DeviceCSR):
        return _CsrOperator(FF)
    if scipy.sparse.issparse(FF):
        return _CsrOperator(_kernels.DeviceCSR.from_scipy(FF))
    if isinstance(FF, np.ndarray) and FF.ndim == 2:
        return _CsrOperator(_kernels.DeviceCSR.from_scipy(scipy.sparse.csr_matrix(FF)))
    raise TypeError(
        f"solve_radiosity: unsupported operator {type(FF).__name__}; pass a fluxpy_amd "
        "form-factor matrix, a SciPy sparse matrix or a dense array")


def _prep_rho(rho, n, tdt):
    """Scalar rho stays a float; per-facet rho becomes a device vector."""
    t = _lib.torch()
    if isinstance(rho, t.Tensor):
        r = rho.to(device=_lib.device(), dtype=tdt).reshape(-1)
        return r.contiguous() if r.numel() > 1 else float(r.item())
    if np.ndim(rho) == 0:
        return float(rho)
    r = np.asarray(rho).reshape(-1)
    if r.size == 1:
        return float(r[0])
    if r.size != n:
        raise ValueError(f"rho must be a scalar or have {n} entries (got {r.size})")
    return _lib.to_dev(r).to(tdt).contiguous()


def _apply_K(op, X, rho, left, out=None):
    """out = (I - F diag(rho)) X  (right)  or  (I - diag(rho) F) X  (left)."""
    if left:
        Y = op.apply(X)
        _kernels.rowscale(rho, Y, out=Y)
    else:
        Y = op.apply(_kernels.rowscale(rho, X))
    # out = X - Y
    out = X.clone() if out is None else out.copy_(X)
    t = _lib.torch()
    k = X.shape[1]
    minus = t.full((k,), -1.0, dtype=t.float64, device=X.device)
    _kernels.axpby(minus, Y, None, out)
    return out


def solve_radiosity(FF, E, rho=1, albedo_placement='right', method='jacobi', tol=None,
                    maxiter=None, restart=30, x0=None):
    if albedo_placement not in {'left', 'right'}:
        raise Exception('albedo_placement must be "left" or "right"')
    if method not in {'jacobi', 'cg', 'gmres'}:
        raise Exception('method must be one of: jacobi, cg, gmres')
    t = _lib.require_gpu()
    host = not isinstance(E, t.Tensor)
    if host:
        E = np.asarray(E)
        if E.dtype not in (np.float32, np.float64):
            E = E.astype(np.float64)
    Ed = _lib.to_dev(E)
    vec = Ed.dim() == 1
    E2 = Ed.reshape(Ed.shape[0], -1).contiguous()
    op = device_operator(FF)
    if op.shape[0] != E2.shape[0] or op.shape[1] != E2.shape[0]:
        raise ValueError(f"shape mismatch: F {op.shape}, E {tuple(Ed.shape)}")
    rho_d = _rho_internal(op, _prep_rho(rho, E2.shape[0], E2.dtype))
    E2 = _to_internal(op, E2)
    left = albedo_placement == 'left'
    if method == 'jacobi':
        if tol is None:
            emax = float(abs(E).max()) if host else float(Ed.abs().max())
            tol = np.finfo(np.dtype(str(E2.dtype).replace("torch.", ""))).resolution * emax
        B, niter = _jacobi(op, E2, rho_d, left, tol, maxiter or 100000)
    elif method == 'cg':
        if tol is None:
            emax = float(abs(E).max()) if host else float(Ed.abs().max())
            tol = np.finfo(np.dtype(str(E2.dtype).replace("torch.", ""))).resolution * emax
        B, niter = _cg(op, E2, rho_d, left, tol, maxiter or 100000)
    else:
        if x0 is not None:
            x0 = _to_internal(op, _lib.to_dev(x0).reshape(E2.shape).to(E2.dtype))
        B, niter = _gmres(op, E2, rho_d, left, 1e-12 if tol is None else tol,
                          maxiter or 10000, restart, x0)
    B = _from_internal(op, B)
    B = B.reshape(-1) if vec else B
    return (B.cpu().numpy() if host else B), niter


def _jacobi(op, E, rho, left, tol, maxiter):
    """solve.py:25-45, fused into one kernel per iteration for CSR operators."""
    t = _lib.torch()
    k = E.shape[1]
    tol_c = np.broadcast_to(np.asarray(tol, dtype=np.float64), (k,))
    B = E.clone()
    B1 = t.empty_like(E)
    colmax = t.empty(k, dtype=t.float64, device=E.device)
    niter = 0
    while True:
        niter += 1
        op.jacobi_step(E, B, rho, left, B1, colmax)
        d = colmax.cpu().numpy()
        if np.all(d <= tol_c):
            break
        if niter >= maxiter:
            warnings.warn(f"Jacobi did not converge in {maxiter} iterations "
                          f"(max |B1 - B| = {d.max():.3e} > tol)", RuntimeWarning)
            break
        B, B1 = B1, B
    return B, niter


def _cg(op, E, rho, left, tol, maxiter):
    """solve.py:47-106 (CG applied to I - rho F), column by column."""
    t = _lib.torch()
    k = E.shape[1]
    vo = _kernels.VecOps(k)
    X = t.zeros_like(E)
    R = E.clone()
    if float(vo.maxabs(R).max()) <= tol:
        return X, 0
    P = R.clone()
    niter = 0
    while True:
        FP = op.apply(P)
        # KK_P = P - rho*(FF@P)  (solve.py:75)
        KP = P.clone()
        _kernels.rowscale(rho, FP, out=FP)
        minus = t.full((k,), -1.0, dtype=t.float64, device=E.device)
        _kernels.axpby(minus, FP, None, KP)
        RR = vo.dot(R, R).clone()
        alpha = RR / vo.dot(KP, P)
        _kernels.axpby(alpha, P, None, X)
        _kernels.axpby(-alpha, KP, None, R)
        if float(vo.maxabs(R).max()) <= tol:
            return X, niter + 1
        beta = vo.dot(R, R) / RR
        _kernels.axpby(None, R, beta, P)
        niter += 1
        if niter >= maxiter:
            warnings.warn(f"CG did not converge in {maxiter} iterations", RuntimeWarning)
            return X, niter


GMRES_CHECK = 4     # Arnoldi steps between the host's convergence checks


def _gmres(op, E, rho, left, rtol, maxiter, restart, x0):
    """Restarted GMRES(m), one Krylov space per column, resident on the GPU:
    the basis V (m + 1 blocks of n x k), the Hessenberg matrix H and every
    Arnoldi step (the product, classical Gram-Schmidt with one
    re-orthogonalisation, the normalisation) stay on the device
    (``flux_gs_*``).  The host reads the new columns of H every GMRES_CHECK
    steps, applies the Givens rotations and tests |g_{j+1}| <= rtol |E_c|
    per column, and reads |r| once per restart: a few synchronisations per
    cycle instead of one per inner product.  Up to GMRES_CHECK - 1 steps
    may run past convergence within a check window.  FP64 throughout; the
    result has E's dtype."""
    t = _lib.torch()
    n, k = E.shape
    dev = E.device
    out_dtype = E.dtype
    E = E.to(t.float64).contiguous()
    f64 = t.float64
    vo = _kernels.VecOps(k)
    X = (t.zeros_like(E) if x0 is None
         else _lib.to_dev(x0).reshape(n, k).to(f64).contiguous().clone())
    bnorm = np.sqrt(np.maximum(vo.dot(E, E).cpu().numpy(), 0.0))
    bnorm[bnorm == 0] = 1.0
    m = max(1, int(restart))
    V = t.empty((m + 1, n, k), dtype=f64, device=dev)
    H = t.zeros((m + 1, m, k), dtype=f64, device=dev)
    work = t.empty(int(_lib.load().flux_gs_work_elems(n, k, min(m, 32))), dtype=f64, device=dev)
    st = _lib.stream()
    nk = n * k

    def vptr(i):
        return V.data_ptr() + i * nk * 8

    def hptr(i, j):
        return H.data_ptr() + (i * m + j) * k * 8

    niter = 0
    while True:
        R = E.clone()
        if x0 is not None or niter > 0:
            KX = _apply_K(op, X, rho, left)
            minus = t.full((k,), -1.0, dtype=f64, device=dev)
            _kernels.axpby(minus, KX, None, R)
        nrm2 = vo.dot(R, R)
        beta = np.sqrt(np.maximum(nrm2.cpu().numpy(), 0.0))
        if np.all(beta <= rtol * bnorm) or niter >= maxiter:
            if niter >= maxiter and not np.all(beta <= rtol * bnorm):
                warnings.warn(f"GMRES did not converge in {maxiter} iterations", RuntimeWarning)
            return X.to(out_dtype), niter
        _lib.call("flux_gs_normalize", n, k, R.data_ptr(), nrm2.data_ptr(), None, vptr(0), st)
        Hh = np.zeros((m + 1, m, k))
        cs = np.zeros((m, k))
        sn = np.zeros((m, k))
        g = np.zeros((m + 1, k))
        g[0] = beta
        done = beta <= rtol * bnorm
        j = j0 = 0
        while j < m:
            w = V[j + 1]
            _apply_K(op, V[j], rho, left, out=w)
            for accumulate in (0, 1):
                for s0 in range(0, j + 1, 32):
                    nv = min(32, j + 1 - s0)
                    _lib.call("flux_gs_project", n, k, nv, vptr(s0), nk, w.data_ptr(),
                              hptr(s0, j), m * k, accumulate, work.data_ptr(), st)
            nrm2 = vo.dot(w, w)
            _lib.call("flux_gs_normalize", n, k, w.data_ptr(), nrm2.data_ptr(),
                      hptr(j + 1, j), w.data_ptr(), st)
            j += 1
            niter += 1
            if j % GMRES_CHECK and j < m and niter < maxiter:
                continue
            # new Hessenberg columns j0 .. j-1 to the host; Givens rotations
            Hh[:, j0:j] = H[:, j0:j].cpu().numpy()
            for jj in range(j0, j):
                for i in range(jj):
                    a, b_ = Hh[i, jj].copy(), Hh[i + 1, jj].copy()
                    Hh[i, jj] = cs[i] * a + sn[i] * b_
                    Hh[i + 1, jj] = -sn[i] * a + cs[i] * b_
                a, b_ = Hh[jj, jj].copy(), Hh[jj + 1, jj].copy()
                r = np.hypot(a, b_)
                safe = r > 0
                cs[jj] = np.where(safe, a / np.where(safe, r, 1.0), 1.0)
                sn[jj] = np.where(safe, b_ / np.where(safe, r, 1.0), 0.0)
                Hh[jj, jj] = np.where(safe, r, a)
                Hh[jj + 1, jj] = 0.0
                g[jj + 1] = -sn[jj] * g[jj]
                g[jj] = cs[jj] * g[jj]
                done = done | (np.abs(g[jj + 1]) <= rtol * bnorm)
            j0 = j
            if np.all(done) or niter >= maxiter:
                break
        # back substitution per column, then X += V y on the device
        y = np.zeros((j, k))
        for i in range(j - 1, -1, -1):
            s_ = g[i] - (Hh[i, i + 1:j] * y[i + 1:j]).sum(0)
            d = Hh[i, i]
            y[i] = np.where(d != 0, s_ / np.where(d != 0, d, 1.0), 0.0)
        yd = _lib.to_dev(np.ascontiguousarray(y))
        _lib.call("flux_gs_combine", n, k, j, vptr(0), nk, yd.data_ptr(), k, X.data_ptr(), st)


def _panel_ready(FF):
    """True for a compressed matrix: its batched solves run on the row-block
    panels of its near field, plus its low-rank far field if it has one; and
    for a row-chunked matrix whose panels were built (build_panels)."""
    from .compressed_form_factors import CompressedFormFactorMatrix
    from .streamed import StreamedFormFactorMatrix
    if isinstance(FF, StreamedFormFactorMatrix):
        return FF.panels is not None
    return (isinstance(FF, CompressedFormFactorMatrix) and not getattr(FF, "_derived", False)
            and getattr(FF, "_near", None) is not None)


def _panel_engine(FF, residual):
    """(BF16 panels, residual panels, far field for the sweeps, far field
    for the residual) of FF, built once and kept on FF; the far-field parts
    are None without low-rank leaves."""
    from . import _hbuild
    from ._panel import PanelMatrix, SlicedPanelMatrix
    from .streamed import StreamedFormFactorMatrix
    if isinstance(FF, StreamedFormFactorMatrix):
        if residual != "int8":
            raise ValueError("a row-chunked matrix's panels take residual='int8'")
        lo, hi = FF.panels
        return lo, hi, None, None
    cache = FF.__dict__.setdefault("_panel_cache", {})
    if "bf16" not in cache:
        cache["bf16"] = PanelMatrix(FF._near, "bf16")
    if residual not in cache:
        lo = cache["bf16"]
        cache[residual] = (SlicedPanelMatrix(FF._near, lo) if residual == "int8"
                           else lo.sibling(FF._near, "fp32"))
    far_lo = far_hi = None
    if FF._far_used:
        far_hi = _hbuild.far_master(FF)
        far_lo = _hbuild.far16(FF, "bf16")
    return cache["bf16"], cache[residual], far_lo, far_hi


def _solve_panels(FF, E2, rho, left, rtol, inner_iters, max_refine, residual):
    from ._panel import PanelRefinedSolver
    t = _lib.torch()
    n, k = E2.shape
    op = device_operator(FF)
    # the albedo in panel order is kept while the caller's tensor is unchanged
    # (same storage, same version counter)
    cache = FF.__dict__.setdefault("_panel_rho", {})
    rkey = ((rho.data_ptr(), rho._version, tuple(rho.shape), rho.dtype)
            if isinstance(rho, t.Tensor) else None)
    if rkey is not None and cache.get("key") == rkey:
        rho64 = cache["rho64"]
    else:
        rho64 = _rho_internal(op, _prep_rho(rho, n, t.float64))
        cache.clear()
        if rkey is not None:
            cache.update(key=rkey, rho64=rho64)
    Ei = _to_internal(op, E2)
    P_lo, P_hi, far_lo, far_hi = _panel_engine(FF, residual)
    solvers = FF.__dict__.setdefault("_panel_solvers", {})
    key = (residual, k, bool(left))
    S = solvers.get(key)
    if S is None:
        S = solvers[key] = PanelRefinedSolver(P_lo, P_hi, n, k, rho64, left=left,
                                              far_lo=far_lo, far_hi=far_hi)
    else:
        S.set_rho(rho64)
    _, info = S.solve(Ei, rtol=rtol, inner_iters=inner_iters, max_rounds=max_refine,
                      finish=lambda B: _from_internal(op, B))
    return info.pop("finished"), info


def solve_radiosity_refined(FF_lo, FF_hi, E, rho=1, albedo_placement='right', rtol=1e-7,
                            inner_iters=4, max_refine=30, return_info=False, residual="int8"):
    """Mixed-precision radiosity solve by iterative refinement.

    Inner solves use ``FF_lo`` (low-precision values, e.g. BF16 with FP32
    vectors and accumulation): ``inner_iters`` Jacobi sweeps on the
    correction equation (I - F rho) d = r, with no host synchronisation
    (an int, or one count per round with the last repeated).
    The residual r = E - (I - F rho) B is formed in FP64 with ``FF_hi`` (the
    FP32 master values, FP64 vectors and accumulation).  Stops when every
    column satisfies max|r_c| <= rtol * max|E_c|; the error of B against the
    exact solution of the master system is then below about
    rtol / (1 - max rho * rowsum F) relative to max|E_c|.

    When ``FF_lo is FF_hi`` is a compressed matrix with no low-rank leaves,
    the solve runs on its row-block panels (built on first use and kept):
    BF16 panels on MFMA for the inner sweeps, and for the residual either
    int8-sliced panels on int8 MFMA (``residual='int8'``, FP64-accurate:
    about 1e-9 of max|F_r.| max|x| per column) or the FP32 master panels on
    FP64 MFMA (``residual='fp64'``).

    E: N x k (NumPy or device).  Returns B (FP64) or (B, info).
    """
    t = _lib.require_gpu()
    if residual not in ("int8", "fp64"):
        raise ValueError("residual must be 'int8' or 'fp64'")
    host = not isinstance(E, t.Tensor)
    Ed = _lib.to_dev(np.asarray(E, dtype=np.float64) if host else E.to(t.float64))
    vec = Ed.dim() == 1
    E2 = Ed.reshape(Ed.shape[0], -1).contiguous()
    n, k = E2.shape
    left = albedo_placement == 'left'
    if FF_lo is FF_hi and _panel_ready(FF_lo):
        B, info = _solve_panels(FF_lo, E2, rho, left, rtol, inner_iters, max_refine, residual)
        out = B.reshape(-1) if vec else B
        out = out.cpu().numpy() if host else out
        return (out, info) if return_info else out
    op_lo = device_operator(FF_lo)
    op_hi = device_operator(FF_hi)
    # both operators must share one ordering (they are views of one matrix)
    rho64 = _rho_internal(op_hi, _prep_rho(rho, n, t.float64))
    E2 = _to_internal(op_hi, E2)
    rho32 = _kernels.convert(rho64, t.float32) if isinstance(rho64, t.Tensor) else rho64
    vo = _kernels.VecOps(k)
    emax = vo.maxabs(E2).cpu().numpy()
    thresh = rtol * np.where(emax > 0, emax, 1.0)
    B = t.zeros_like(E2)
    B1 = t.empty_like(E2)
    colmax = t.empty(k, dtype=t.float64, device=E2.device)
    R = E2.clone()                 # residual of B = 0
    R32 = t.empty((n, k), dtype=t.float32, device=E2.device)
    D = t.empty_like(R32)
    D1 = t.empty_like(R32)
    cm32 = t.empty(k, dtype=t.float64, device=E2.device)
    sched = [int(inner_iters)] if np.ndim(inner_iters) == 0 else [int(v) for v in inner_iters]
    if not sched or min(sched) < 1:
        raise ValueError("inner_iters: at least one sweep per round")
    rounds = 0
    inner_total = 0
    hist = []
    while True:
        rounds += 1
        # low-precision correction: D <- R + F(rho D) repeated, D0 = R
        R32.copy_(_kernels.convert(R, t.float32))
        D.copy_(R32)
        sweeps = sched[min(rounds - 1, len(sched) - 1)]
        for _ in range(sweeps):
            op_lo.jacobi_step(R32, D, rho32, left, D1, cm32)
            D, D1 = D1, D
        inner_total += sweeps
        # B += D (FP64 += FP32)
        _kernels.axpby(None, _kernels.convert(D, t.float64), None, B)
        # FP64 residual: B1 = E + F(rho B); R = B1 - B; colmax = max|R|
        op_hi.jacobi_step(E2, B, rho64, left, B1, colmax)
        R.copy_(B1)
        minus = t.full((k,), -1.0, dtype=t.float64, device=E2.device)
        _kernels.axpby(minus, B, None, R)
        rmax = colmax.cpu().numpy()
        hist.append(float((rmax / np.where(emax > 0, emax, 1.0)).max()))
        if np.all(rmax <= thresh):
            break
        if rounds >= max_refine:
            warnings.warn(f"refinement stopped after {rounds} rounds (residual "
                          f"{hist[-1]:.2e} > {rtol:.1e})", RuntimeWarning)
            break
    B = _from_internal(op_hi, B)
    out = B.reshape(-1) if vec else B
    out = out.cpu().numpy() if host else out
    if return_info:
        return out, {"rounds": rounds, "inner_iters": inner_total, "residual_history": hist}
    return out
