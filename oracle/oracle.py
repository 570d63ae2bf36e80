"""CPU oracle for the fluxpy hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this module, and only as the checker.  The product
(package ``fluxpy_amd``) never imports it.

Contents and what each part restates (paths relative to src/flux/):

* ``geometry``            shape.py:16-45 (NumPy, same expressions)
* ``OracleScene``         cgal/aabb_wrapper.cpp:64-129, cgal/aabb.pyx:58-87,
                          shape.py:349-421 -- C code in flux_oracle.c
* ``form_factor_matrix``  form_factors.py:11-72 -- C code in flux_oracle.c
* ``direct_irradiance``   shape.py:190-244, 1-D sun branch
* ``solve_radiosity``     solve.py:4-45 (Jacobi left/right, reference
                          stopping rule) and a dense FP64 LU solve
* ``steady_state_temp``   model.py:8-42
* ``conduction_q``, ``conduction_t``
                          pcc/conductionQ.c:10-186, pcc/conductionT.c:4-80,
                          pcc/tridag.c:5-34 -- C code in pcc_oracle.c
* ``thermal_model_run``   model.py:108-234 (ThermalModel iteration) over
                          a SciPy F and the conduction above

What pins each part (beyond restating it) is listed in tests/test_oracle.py:
closed forms, the sphere invariants of tests/test_shape.py /
tests/test_form_factors.py of the reference, brute force vs BVH, and
LAPACK.  Functions with no such check say "parity unpinned" in their
docstring.
"""
import ctypes
import os

import numpy as np
import scipy.sparse

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            from . import build as _b  # noqa
            _b.build()
        L = ctypes.CDLL(LIB)
        vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        L.oracle_scene_create.restype = vp
        L.oracle_scene_create.argtypes = [vp, i64, vp, i64, i32]
        L.oracle_scene_destroy.argtypes = [vp]
        L.oracle_scene_centroids.restype = ctypes.POINTER(ctypes.c_double)
        L.oracle_scene_centroids.argtypes = [vp]
        L.oracle_vis_mn.argtypes = [vp, vp, i64, vp, i64, i32, vp, i32, vp]
        L.oracle_embree_pair_rays.argtypes = [vp, vp, i64, i32, vp, vp, vp]
        L.oracle_embree_sun_rays.argtypes = [vp, vp, i64, vp, i32, vp, vp]
        L.oracle_occluded.argtypes = [vp, vp, i64, vp, i64, i32, i32, vp, vp, i32, vp]
        L.oracle_intersect1.argtypes = [vp, vp, vp, i64, vp, vp]
        L.oracle_ff_csr.argtypes = [vp, vp, vp, vp, i32, vp, i64, vp, i64, ctypes.c_double,
                                    i32, vp, vp, vp, vp]
        dbl = ctypes.c_double
        L.oracle_conduction_q.restype = i32
        L.oracle_conduction_q.argtypes = [i64, i32, vp, dbl, vp, vp, vp, vp, vp, vp, vp, vp, vp]
        L.oracle_conduction_t.restype = i32
        L.oracle_conduction_t.argtypes = [i64, i32, vp, dbl, vp, vp, vp, vp, vp, vp, vp]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


# --------------------------------------------------------------------------
# Geometry: shape.py:16-45, the reference's own NumPy expressions.
# --------------------------------------------------------------------------

def geometry(V, F):
    """(P, N, A) exactly as shape.py:16-45 computes them."""
    VF = V[F]
    P = VF.mean(axis=1)
    C = np.cross(VF[:, 1, :] - VF[:, 0, :], VF[:, 2, :] - VF[:, 0, :])
    Cn = np.sqrt(np.sum(C ** 2, axis=1))
    N = C / Cn.reshape(C.shape[0], 1)
    A = Cn / 2
    return P, N, A


# --------------------------------------------------------------------------
# Ray tracing: the C oracle.
# --------------------------------------------------------------------------

MODE_CGAL, MODE_EMBREE = 0, 1


class OracleScene:
    """Brute-force (use_bvh=False) or CPU-BVH scene over an FP64 copy of V."""

    def __init__(self, V, F, use_bvh=True):
        self.V = np.ascontiguousarray(V, dtype=np.float64)
        self.F = np.ascontiguousarray(F, dtype=np.int64)
        self.nf = self.F.shape[0]
        self.h = lib().oracle_scene_create(_p(self.V), self.V.shape[0], _p(self.F), self.nf,
                                           int(bool(use_bvh)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().oracle_scene_destroy(self.h)
            self.h = None

    def centroids(self):
        p = lib().oracle_scene_centroids(self.h)
        return np.ctypeslib.as_array(p, shape=(self.nf, 3)).copy()

    def vis_mn(self, I, J, mode=MODE_CGAL, P=None):
        I = np.ascontiguousarray(I, dtype=np.int64)
        J = np.ascontiguousarray(J, dtype=np.int64)
        out = np.zeros((I.size, J.size), dtype=np.uint8)
        Pm = None if P is None else np.ascontiguousarray(P, dtype=np.float64)
        f32 = int(P is not None and np.asarray(P).dtype == np.float32)
        lib().oracle_vis_mn(self.h, _p(I), I.size, _p(J), J.size, mode, _p(Pm), f32, _p(out))
        return out.astype(bool)

    def occluded(self, I, D, per_face=False, mode=MODE_CGAL, P=None, N=None):
        I = np.ascontiguousarray(I, dtype=np.int64)
        D = np.ascontiguousarray(np.atleast_2d(D), dtype=np.float64)
        cols = 1 if per_face else D.shape[0]
        out = np.zeros((I.size, cols), dtype=np.uint8)
        Pm = None if P is None else np.ascontiguousarray(P, dtype=np.float64)
        Nm = None if N is None else np.ascontiguousarray(N, dtype=np.float64)
        f32 = int(P is not None and np.asarray(P).dtype == np.float32)
        lib().oracle_occluded(self.h, _p(I), I.size, _p(D), D.shape[0], int(per_face), mode,
                              _p(Pm), _p(Nm), f32, _p(out))
        return out.astype(bool)

    def intersect1(self, x, d):
        x = np.ascontiguousarray(np.atleast_2d(x), dtype=np.float64)
        d = np.ascontiguousarray(np.atleast_2d(d), dtype=np.float64)
        hit = np.zeros(x.shape[0], dtype=np.int64)
        xt = np.zeros_like(x)
        lib().oracle_intersect1(self.h, _p(x), _p(d), x.shape[0], _p(hit), _p(xt))
        return hit, xt

    def form_factor_matrix(self, P, N, A, I=None, J=None, eps=1e-5, mode=MODE_CGAL):
        """form_factors.py:11-72: CSR with int32 column positions."""
        P = np.ascontiguousarray(P)
        dt = P.dtype
        N = np.ascontiguousarray(N, dtype=dt)
        A = np.ascontiguousarray(A, dtype=dt)
        nf = P.shape[0]
        I = np.arange(nf, dtype=np.int64) if I is None else np.ascontiguousarray(I, np.int64)
        J = np.arange(nf, dtype=np.int64) if J is None else np.ascontiguousarray(J, np.int64)
        code = 1 if dt == np.float64 else 0
        nnz = np.zeros(I.size, dtype=np.int64)
        lib().oracle_ff_csr(self.h, _p(P), _p(N), _p(A), code, _p(I), I.size, _p(J), J.size,
                            float(eps), mode, _p(nnz), None, None, None)
        indptr = np.zeros(I.size + 1, dtype=np.int64)
        np.cumsum(nnz, out=indptr[1:])
        cols = np.zeros(int(indptr[-1]), dtype=np.int32)
        vals = np.zeros(int(indptr[-1]), dtype=dt)
        lib().oracle_ff_csr(self.h, _p(P), _p(N), _p(A), code, _p(I), I.size, _p(J), J.size,
                            float(eps), mode, None, _p(indptr), _p(cols), _p(vals))
        return scipy.sparse.csr_matrix((vals, cols, indptr), shape=(I.size, J.size))


# --------------------------------------------------------------------------
# Insolation, solvers, temperature (NumPy restatements).
# --------------------------------------------------------------------------

def embree_pair_rays(Pi, Pj):
    """Embree's ray inputs for the pairs (Pi[p], Pj[p]) as shape.py:355-385
    forms them, in the dtype of Pi (flux_oracle.c oemb_pair): (origins,
    directions, mask of the pairs that get a ray)."""
    Pi = np.asarray(Pi)
    f32 = int(Pi.dtype == np.float32)
    a = np.ascontiguousarray(Pi, dtype=np.float64)
    b = np.ascontiguousarray(Pj, dtype=np.float64)
    m = a.shape[0]
    o = np.zeros((m, 3))
    d = np.zeros((m, 3))
    ok = np.zeros(m, dtype=np.uint8)
    lib().oracle_embree_pair_rays(_p(a), _p(b), m, f32, _p(o), _p(d), _p(ok))
    return o, d, ok.astype(bool)


def embree_sun_rays(P, N, D):
    """Embree's occlusion-ray inputs (shape.py:410-411) for the facets of P,
    N toward the direction D (flux_oracle.c oemb_sun)."""
    P = np.asarray(P)
    f32 = int(P.dtype == np.float32)
    a = np.ascontiguousarray(P, dtype=np.float64)
    b = np.ascontiguousarray(N, dtype=np.float64)
    Dd = np.ascontiguousarray(D, dtype=np.float64).reshape(3)
    m = a.shape[0]
    o = np.zeros((m, 3))
    d = np.zeros((m, 3))
    lib().oracle_embree_sun_rays(_p(a), _p(b), m, _p(Dd), f32, _p(o), _p(d))
    return o, d


def direct_irradiance(N, occluded, F0, d):
    """1-D sun branch of shape.py:224-226 with the dot product in FP64.

    E = F0 max(0, N . d) where not occluded, else 0, stored in N's dtype.
    ``N . d`` uses ((n0 d0 + n1 d1) + n2 d2).
    """
    N64 = N.astype(np.float64)
    cosi = (N64[:, 0] * d[0] + N64[:, 1] * d[1]) + N64[:, 2] * d[2]
    E = np.zeros(N.shape[0], dtype=N.dtype)
    vis = ~occluded
    E[vis] = F0 * np.maximum(0, cosi[vis])
    return E


def solve_radiosity_jacobi(F, E, rho=1.0, albedo_placement="right", tol=None,
                           maxiter=100000):
    """solve.py:4-45 with per-facet rho read as diag(rho) on multi-RHS blocks."""
    if tol is None:
        tol = np.finfo(E.dtype).resolution * abs(E).max()
    rho = np.asarray(rho)
    if rho.ndim == 1 and E.ndim == 2:
        rho = rho[:, None]
    B = E.copy()
    niter = 0
    while True:
        niter += 1
        if albedo_placement == "left":
            B1 = E + rho * (F @ B)
        else:
            B1 = E + F @ (rho * B)
        if abs(B1 - B).max() <= tol or niter >= maxiter:
            break
        B = B1
    # like solve.py:28-34, the iterate before the last update is returned
    return B, niter


def solve_radiosity_dense(F, E, rho=1.0, albedo_placement="right"):
    """Dense FP64 LU solve of (I - F diag(rho)) B = E (right) or
    (I - diag(rho) F) B = E (left)."""
    Fd = F.toarray() if scipy.sparse.issparse(F) else np.asarray(F)
    Fd = Fd.astype(np.float64)
    n = Fd.shape[0]
    r = np.broadcast_to(np.asarray(rho, dtype=np.float64), (n,))
    K = np.eye(n) - (Fd * r[None, :] if albedo_placement == "right" else r[:, None] * Fd)
    return np.linalg.solve(K, E.astype(np.float64))


from scipy.constants import Stefan_Boltzmann as SIGMA  # as model.py:3


def steady_state_temp(F, E, rho, emiss, Fsurf=0.0, clamp=True, solver=None):
    """model.py:8-42.  ``solver(F, E, rho)`` defaults to the Jacobi above."""
    solver = solver or (lambda F_, E_, r_: solve_radiosity_jacobi(F_, E_, r_, "right")[0])
    if E.ndim == 1:
        B = solver(F, E, rho)
        if clamp:
            B = np.maximum(0, B)
        IR = F @ ((1 - rho) * B + Fsurf)
        Q = solver(F, IR, 1)
        if clamp:
            Q = np.maximum(0, Q)
        tot = (1 - rho) * B + emiss * Q + Fsurf
        if clamp:
            tot = np.maximum(0, tot)
        return (tot / (emiss * SIGMA)) ** 0.25
    B_arr = solver(F, E, np.asarray(rho)[:, None] if np.ndim(rho) == 1 else rho)
    T = []
    for B in B_arr.T:
        if clamp:
            B = np.maximum(0, B)
        IR = F @ ((1 - rho) * B)
        Q = solver(F, IR, 1)
        if clamp:
            Q = np.maximum(0, Q)
        T.append((((1 - rho) * B + emiss * Q) / (emiss * SIGMA)) ** 0.25)
    return np.array(T)


# --------------------------------------------------------------------------
# Time-dependent thermal model: 1-D conduction per facet and the iteration.
# --------------------------------------------------------------------------

def _f64(a, n=None):
    """Contiguous FP64 array; a scalar becomes n copies when n is given."""
    a = np.asarray(a, dtype=np.float64)
    if n is not None and a.ndim == 0:
        a = np.full(n, float(a))
    a = np.ascontiguousarray(a)
    if n is not None and a.size != n:
        raise ValueError(f"expected {n} values, got {a.size}")
    return a


def conduction_q(z, dt, Qn, Qnp1, T, ti, rhoc, emiss, Fgeotherm):
    """pcc/conductionQ.c:10-186 for every row of T (nf x (nz + 1), updated in
    place).  Returns (Fsurf, iters): the surface heat flux and the number of
    predictor-corrector repeats per facet.  Raises on a zero pivot (the
    reference exits the process, tridag.c:16-28)."""
    z = _f64(z)
    nz = z.size - 1
    assert T.dtype == np.float64 and T.flags.c_contiguous and T.shape[1] == nz + 1
    nf = T.shape[0]
    Qn, Qnp1 = _f64(Qn, nf), _f64(Qnp1, nf)
    emiss, Fg = _f64(emiss, nf), _f64(Fgeotherm, nf)
    ti, rhoc = _f64(ti, nz + 1), _f64(rhoc, nz + 1)
    Fsurf = np.empty(nf)
    iters = np.zeros(nf, dtype=np.int32)
    rc = lib().oracle_conduction_q(nf, nz, _p(z), float(dt), _p(Qn), _p(Qnp1), _p(T), _p(ti),
                                   _p(rhoc), _p(emiss), _p(Fg), _p(Fsurf), _p(iters))
    if rc:
        raise RuntimeError("tridag: rewrite equations" if rc == 1 else "tridag failed")
    return Fsurf, iters


def conduction_t(z, dt, T, Tsurf, Tsurfp1, ti, rhoc, Fgeotherm):
    """pcc/conductionT.c:4-80 for every row of T (updated in place; T[:, 0]
    is left as it is, as in the reference).  Returns Fsurf."""
    z = _f64(z)
    nz = z.size - 1
    assert T.dtype == np.float64 and T.flags.c_contiguous and T.shape[1] == nz + 1
    nf = T.shape[0]
    Ts, Ts1, Fg = _f64(Tsurf, nf), _f64(Tsurfp1, nf), _f64(Fgeotherm, nf)
    ti, rhoc = _f64(ti, nz + 1), _f64(rhoc, nz + 1)
    Fsurf = np.empty(nf)
    rc = lib().oracle_conduction_t(nf, nz, _p(z), float(dt), _p(T), _p(Ts), _p(Ts1), _p(ti),
                                   _p(rhoc), _p(Fg), _p(Fsurf))
    if rc:
        raise RuntimeError("tridag: rewrite equations" if rc == 1 else "tridag failed")
    return Fsurf


def _absorbed(E, Qrefl, QIR, rho, emiss):
    """model.py:224: the flux absorbed by the surface."""
    return (1 - rho) * (E + Qrefl) + emiss * QIR


def thermal_model_run(F, E_steps, t, rho, emiss, z, T0, ti, rhoc, Fgeotherm, E0,
                      out_dtype=np.float64, return_flux=False, absorbed=_absorbed):
    """model.py:108-234 (one sun per time): the ThermalModel iteration.

    F: SciPy CSR (N x N) in the model dtype; E_steps[i] the direct
    irradiance at time t[i] (what get_direct_irradiance returns, i = 1..);
    E0 the irradiance of the initial state.  ``F @ x`` is rounded to
    ``out_dtype`` as the reference's compressed operator does
    (compressed_form_factors.py:408).  Yields T (N x nz+1) after each step,
    or (T, E, Qrefl, QIR) with return_flux.  ``absorbed`` is the formula of
    model.py:224 (a test replaces it to plant a mistake)."""
    nf = F.shape[0]
    rho = np.asarray(rho, dtype=np.float64)
    emiss = _f64(emiss, nf)
    Fg = _f64(Fgeotherm, nf)
    Q0 = (1 - rho) * E0 + Fg
    Tsurf = (Q0 / (SIGMA * emiss)) ** 0.25
    T = np.ascontiguousarray(np.broadcast_to(np.asarray(T0, dtype=np.float64),
                                             (nf, np.size(z))).copy())
    Qprev = Q0
    Qrefl = QIR = 0
    apply = lambda x: np.asarray(F @ x).astype(out_dtype)
    for i in range(1, len(t)):
        E = E_steps[i]
        dt = t[i] - t[i - 1]
        Qrefl = rho * apply(E + Qrefl)          # model.py:60-63 (clamp)
        Qrefl = np.maximum(0, Qrefl)
        QIR = apply(emiss * SIGMA * Tsurf ** 4 + (1 - emiss) * QIR)
        QIR = np.maximum(0, QIR)
        Q = absorbed(E, Qrefl, QIR, rho, emiss)
        conduction_q(z, dt, Qprev, Q, T, ti, rhoc, emiss, Fg)
        Qprev = Q
        Tsurf = T[:, 0].copy()
        yield (T.copy(), E, Qrefl, QIR) if return_flux else T.copy()
