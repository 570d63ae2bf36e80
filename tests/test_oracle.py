"""Pins of the CPU oracle (oracle/) -- CPU only.

Each oracle function is checked against something that does not restate it:
closed forms, exact rational arithmetic, the invariants of the reference's
own tests (tests/test_shape.py:16-112, tests/test_form_factors.py:18-57 of
fluxpy), brute force against the BVH, and LAPACK.
"""
import itertools as it
from fractions import Fraction

import numpy as np
import pytest

from fluxpy_amd import meshes


# --------------------------------------------------------------------------
# geometry (shape.py:16-45)
# --------------------------------------------------------------------------

def test_geometry_closed_form_right_triangles(oracle_mod):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 2.0]])
    F = np.array([[0, 1, 2], [0, 2, 1], [0, 1, 3]])
    P, N, A = oracle_mod.geometry(V, F)
    assert np.allclose(A, [0.5, 0.5, 1.0])
    assert np.allclose(N[0], [0, 0, 1]) and np.allclose(N[1], [0, 0, -1])
    assert np.allclose(N[2], [0, -1, 0])
    assert np.allclose(P[0], [1 / 3, 1 / 3, 0])


def test_geometry_icosphere_area_converges_to_4pi(oracle_mod):
    errs = []
    for n in (2, 4, 8, 16):
        V, F = meshes.icosphere(n)
        _, N, A = oracle_mod.geometry(V, F)
        errs.append(abs(A.sum() - 4 * np.pi))
        assert np.allclose(np.linalg.norm(N, axis=1), 1)
    # second-order convergence of the polyhedral area: halving h quarters it
    assert all(3.5 < a / b < 4.5 for a, b in zip(errs[1:], errs[2:]))
    assert errs[-1] < 2e-2


def test_host_geometry_helpers_match_reference_expressions(oracle_mod, rng):
    """The package's host helpers give bit-identical P, N, A to the NumPy
    expressions of shape.py:16-45, in both dtypes."""
    from fluxpy_amd import shape
    for dt in (np.float64, np.float32):
        V, F = meshes.spherical_crater(npts=12, seed=3)
        V = (V * rng.uniform(0.5, 2.0, 3)).astype(dt)
        P, N, A = oracle_mod.geometry(V, F)
        assert np.array_equal(shape.get_centroids(V, F), P)
        N2, A2 = shape.get_surface_normals_and_face_areas(V, F)
        assert np.array_equal(N2, N) and np.array_equal(A2, A)
        assert np.array_equal(shape.get_face_areas(V, F), A)
        assert np.array_equal(shape.get_surface_normals(V, F), N)
        assert shape.get_centroids(V, F).dtype == dt


# --------------------------------------------------------------------------
# ray/triangle test against exact rational arithmetic
# --------------------------------------------------------------------------

def _exact_hit(o, d, tri):
    """Exact (Fraction) ray/triangle intersection: returns t or None."""
    o = [Fraction(x) for x in o]
    d = [Fraction(x) for x in d]
    A, B, C = ([Fraction(x) for x in tri[k]] for k in range(3))

    def sub(a, b):
        return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    e1, e2 = sub(B, A), sub(C, A)
    n = cross(e1, e2)
    den = dot(n, d)
    if den == 0:
        return None
    t = dot(n, sub(A, o)) / den
    if t < 0:
        return None
    x = [o[k] + t * d[k] for k in range(3)]
    # barycentric signs
    s1 = dot(cross(sub(B, A), sub(x, A)), n)
    s2 = dot(cross(sub(C, B), sub(x, B)), n)
    s3 = dot(cross(sub(A, C), sub(x, C)), n)
    if (s1 >= 0 and s2 >= 0 and s3 >= 0) or (s1 <= 0 and s2 <= 0 and s3 <= 0):
        return t
    return None


def test_triangle_test_matches_exact_arithmetic(oracle_mod, rng):
    """Random rays against one triangle: hit/miss and t agree with exact
    rational arithmetic away from the edges."""
    tri = np.array([[0.1, 0.2, 0.3], [1.3, 0.1, 0.5], [0.4, 1.2, 0.9]])
    V = tri.copy()
    F = np.array([[0, 1, 2]])
    S = oracle_mod.OracleScene(V, F, use_bvh=False)
    nchk = 0
    for _ in range(400):
        o = rng.uniform(-1, 2, 3)
        target = rng.uniform(-0.3, 1.3, 3)
        d = target - o
        tex = _exact_hit(o, d, tri)
        hit, xt = S.intersect1(o, d)
        # skip rays within 1e-9 of an edge (decided by rounding)
        bary_margin = _edge_margin(o, d, tri)
        if bary_margin < 1e-9:
            continue
        nchk += 1
        if tex is None:
            assert hit[0] == -1
        else:
            assert hit[0] == 0
            xe = o + float(tex) * d
            assert np.allclose(xt[0], xe, rtol=1e-12, atol=1e-12)
    assert nchk > 300


def _edge_margin(o, d, tri):
    A, B, C = tri
    n = np.cross(B - A, C - A)
    den = n @ d
    if abs(den) < 1e-14:
        return 0.0
    t = n @ (A - o) / den
    x = o + t * d
    ws = [np.cross(B - A, x - A) @ n, np.cross(C - B, x - B) @ n, np.cross(A - C, x - C) @ n]
    return min(abs(w) for w in ws) / (n @ n)


def test_shared_edge_hits_both_triangles(oracle_mod):
    """Watertight: a ray through a common edge is reported on both facets
    (closest hit = smaller index)."""
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.0]])
    F = np.array([[0, 1, 2], [1, 3, 2]])
    S = oracle_mod.OracleScene(V, F, use_bvh=False)
    o = np.array([0.5, 0.5, 1.0])  # right above the diagonal edge
    hit, xt = S.intersect1(o, np.array([0, 0, -1.0]))
    assert hit[0] == 0 and np.allclose(xt[0], [0.5, 0.5, 0])
    occ = S.occluded(np.array([0]), np.array([[0, 0, 1.0]]))
    assert not occ[0, 0]


# --------------------------------------------------------------------------
# visibility and occlusion: reference sphere invariants + brute force
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 5])
def test_sphere_visibility_invariants(oracle_mod, n):
    """tests/test_shape.py:16-60 of the reference, CGAL semantics."""
    V, F = meshes.icosphere(n)
    P, N, A = oracle_mod.geometry(V, F)
    nf = F.shape[0]
    I = np.arange(nf)
    S = oracle_mod.OracleScene(V, F)
    vis = S.vis_mn(I, I)
    assert (vis == ~np.eye(nf, dtype=bool)).all()
    # Embree semantics: diagonal visible by default (|d| = 0 <= 1e-3,
    # shape.py:360-362,392), every other pair visible on a convex sphere
    visE = S.vis_mn(I, I, mode=oracle_mod.MODE_EMBREE, P=P)
    assert visE.all()


@pytest.mark.parametrize("n", [4, 5])
def test_sphere_occlusion_invariant(oracle_mod, n, rng):
    """tests/test_shape.py:62-84: occluded <=> N.D < 0 for outward normals.
    Embree semantics lift the origin 1e-3 along N, so rays within that
    grazing band (-0.02 < N.D < 0 on the unit sphere) may escape; those are
    the only mismatches allowed and all of them are listed as grazing."""
    V, F = meshes.icosphere(n)
    P, N, A = oracle_mod.geometry(V, F)
    I = np.arange(F.shape[0])
    S = oracle_mod.OracleScene(V, F)
    grazing = 0
    for _ in range(20):
        D = rng.standard_normal(3)
        D /= np.linalg.norm(D)
        nd = N @ D
        occ = S.occluded(I, D)[:, 0]
        assert (occ == (nd < 0)).all()
        occE = S.occluded(I, D, mode=oracle_mod.MODE_EMBREE, P=P, N=N)[:, 0]
        bad = occE != (nd < 0)
        assert ((nd[bad] > -0.02) & (nd[bad] < 0)).all()
        grazing += bad.sum()
    assert grazing < 0.05 * 20 * len(I)


def test_bvh_equals_brute_force(oracle_mod, rng):
    V, F = meshes.spherical_crater(npts=14, seed=11)
    P, N, A = oracle_mod.geometry(V, F)
    nf = F.shape[0]
    I = rng.choice(nf, 40, replace=False)
    J = np.arange(nf)
    Sb = oracle_mod.OracleScene(V, F, use_bvh=True)
    Sf = oracle_mod.OracleScene(V, F, use_bvh=False)
    for mode in (0, 1):
        assert (Sb.vis_mn(I, J, mode, P) == Sf.vis_mn(I, J, mode, P)).all()
    D = np.array([[np.cos(0.2), 0, np.sin(0.2)], [0.1, 0.3, 0.2]])
    assert (Sb.occluded(J, D) == Sf.occluded(J, D)).all()
    x = rng.uniform(-1, 1, (50, 3)) + [0, 0, 1]
    d = rng.standard_normal((50, 3))
    hb, xb = Sb.intersect1(x, d)
    hf, xf = Sf.intersect1(x, d)
    assert (hb == hf).all() and np.array_equal(xb, xf)


def test_single_occluder_mask(oracle_mod):
    """Two facing horizontal facets with a blocker between them."""
    tri_lo = [[0, 0, 0], [1, 0, 0], [0, 1, 0]]
    tri_hi = [[0, 0, 2], [0, 1, 2], [1, 0, 2]]       # faces down
    blk = [[-1, -1, 1], [3, -1, 1], [-1, 3, 1]]       # big blocker at z = 1
    V = np.array(tri_lo + tri_hi + blk, dtype=float)
    F = np.array([[0, 1, 2], [3, 4, 5], [6, 7, 8]])
    S = oracle_mod.OracleScene(V, F)
    vis = S.vis_mn(np.arange(3), np.arange(3))
    assert not vis[0, 1] and not vis[1, 0]
    assert vis[0, 2] and vis[1, 2] and vis[2, 0] and vis[2, 1]
    assert not vis.diagonal().any()
    S2 = oracle_mod.OracleScene(V[:6], F[:2])
    vis2 = S2.vis_mn(np.arange(2), np.arange(2))
    assert vis2[0, 1] and vis2[1, 0]


def test_visibility_symmetry_on_crater(oracle_mod):
    """vis(i, j) == vis(j, i) on a crater except for a few rays whose
    decision is grazing or on a shared edge."""
    V, F = meshes.spherical_crater(npts=16, seed=5)
    nf = F.shape[0]
    I = np.arange(nf)
    S = oracle_mod.OracleScene(V, F)
    vis = S.vis_mn(I, I)
    asym = (vis != vis.T).sum() // 2
    assert asym <= 0.002 * vis.sum()


# --------------------------------------------------------------------------
# form factors (form_factors.py:11-72)
# --------------------------------------------------------------------------

def test_form_factor_closed_form_parallel_facets(oracle_mod):
    """Facing parallel facets with aligned centroids: F_ij = A_j/(pi d^2)."""
    d = 3.0
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, d], [0, 2, d], [2, 0, d]], float)
    F = np.array([[0, 1, 2], [3, 4, 5]])
    P, N, A = oracle_mod.geometry(V, F)
    # shift the upper facet so centroids align vertically
    V[3:, :2] += P[0, :2] - P[1, :2]
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    FF = S.form_factor_matrix(P, N, A).toarray()
    assert np.isclose(FF[0, 1], A[1] / (np.pi * d ** 2), rtol=1e-14)
    assert np.isclose(FF[1, 0], A[0] / (np.pi * d ** 2), rtol=1e-14)
    assert FF[0, 0] == 0 and FF[1, 1] == 0


def _mutant_ff_value(c, Pi, Pj, Aj):
    """A planted mistake (r^2 instead of r^4 in the denominator), used only to
    show the closed-form check above would catch it."""
    s2 = ((Pi - Pj) ** 2).sum()
    return c * Aj / (np.pi * s2)


def test_planted_mistake_is_caught_by_closed_form():
    d = 3.0
    Pi, Pj = np.zeros(3), np.array([0, 0, d])
    c = d * d
    assert not np.isclose(_mutant_ff_value(c, Pi, Pj, 2.0), 2.0 / (np.pi * d ** 2))


@pytest.mark.parametrize("n", [4, 5])
def test_sphere_form_factor_invariants(oracle_mod, n):
    """tests/test_form_factors.py:18-57: inward normals -> every off-diagonal
    entry nonzero, diagonal zero; outward normals -> F == 0.  Also the
    sphere's known value A_j / (4 pi R^2) and row sums near 1."""
    V, F = meshes.icosphere(n)
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    FF = S.form_factor_matrix(P, -N, A).toarray()
    off = ~np.eye(F.shape[0], dtype=bool)
    assert (FF[off] != 0).all() and (FF.diagonal() == 0).all()
    Rc = np.linalg.norm(P, axis=1).mean()
    ratio = FF[off] / (A[None, :] / (4 * np.pi * Rc ** 2)).repeat(F.shape[0], 0)[off]
    assert abs(np.median(ratio) - 1) < 0.05
    assert np.allclose(FF.sum(1), 1, atol=0.02)
    assert S.form_factor_matrix(P, N, A).nnz == 0


def test_form_factor_reciprocity(oracle_mod):
    V, F = meshes.spherical_crater(npts=14, seed=2)
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    FF = S.form_factor_matrix(P, N, A).toarray()
    L = A[:, None] * FF
    sym = (FF != 0) & (FF.T != 0)
    assert np.allclose(L[sym], L.T[sym], rtol=1e-12)


def test_form_factor_flat_plane_is_zero(oracle_mod):
    x = np.linspace(-1, 1, 9)
    V, F = meshes.grid_mesh(x, x, np.zeros((9, 9)))
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    assert S.form_factor_matrix(P, N, A).nnz == 0


def test_form_factor_rows_subset_and_f32(oracle_mod, rng):
    """F[I, J] equals the slice of the full matrix; FP32 models give FP32."""
    V, F = meshes.spherical_crater(npts=14, seed=4)
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    full = S.form_factor_matrix(P, N, A).toarray()
    I = rng.choice(F.shape[0], 30, replace=False)
    J = rng.choice(F.shape[0], 50, replace=False)
    sub = S.form_factor_matrix(P, N, A, I, J).toarray()
    assert np.array_equal(sub, full[I][:, J])
    P32, N32, A32 = oracle_mod.geometry(V.astype(np.float32), F)
    f32 = S.form_factor_matrix(P32, N32, A32)
    assert f32.dtype == np.float32
    assert np.allclose(f32.toarray(), full, rtol=2e-4, atol=1e-9)


# --------------------------------------------------------------------------
# solvers and temperature (solve.py, model.py)
# --------------------------------------------------------------------------

def test_jacobi_matches_dense_lu(oracle_mod):
    V, F = meshes.spherical_crater(npts=16, seed=1)
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    FF = S.form_factor_matrix(P, N, A)
    d = np.array([np.cos(np.deg2rad(10)), 0, np.sin(np.deg2rad(10))])
    E = oracle_mod.direct_irradiance(N, S.occluded(np.arange(len(A)), d)[:, 0], 1365.0, d)
    for place in ("right", "left"):
        rho = np.linspace(0.1, 0.2, len(A))
        B, it = oracle_mod.solve_radiosity_jacobi(FF, E, rho, place)
        Bd = oracle_mod.solve_radiosity_dense(FF, E, rho, place)
        assert abs(B - Bd).max() <= 1e-12 * abs(Bd).max()
        assert it > 3


def test_steady_state_temp_flat_plane_closed_form(oracle_mod):
    """With F = 0 the IR solve is trivial: T = ((1-rho) E / (eps sigma))^(1/4)."""
    import scipy.sparse
    E = np.linspace(100, 1000, 20)
    F0 = scipy.sparse.csr_matrix((20, 20))
    T = oracle_mod.steady_state_temp(F0, E, 0.12, 0.95)
    assert np.allclose(T, ((1 - 0.12) * E / (0.95 * oracle_mod.SIGMA)) ** 0.25, rtol=1e-14)


def test_insolation_is_cosine_law(oracle_mod):
    x = np.linspace(-1, 1, 5)
    V, F = meshes.grid_mesh(x, x, np.zeros((5, 5)))
    P, N, A = oracle_mod.geometry(V, F)
    S = oracle_mod.OracleScene(V, F)
    e = np.deg2rad(30)
    d = np.array([np.cos(e), 0, np.sin(e)])
    E = oracle_mod.direct_irradiance(N, S.occluded(np.arange(len(A)), d)[:, 0], 1000.0, d)
    assert np.allclose(E, 1000 * np.sin(e))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_embree_ray_inputs_follow_the_reference_expressions(oracle_mod, dt):
    """The oracle's Embree ray inputs equal those of fluxpy's own NumPy
    expressions (shape.py:351-385 for pairs, 405-411 for occlusion) on the
    same arrays, stored as FP32 (Embree's ray arrays).  For FP32 models
    NumPy 1.x and 2.x agree on them (eps = 1e3 * float32 resolution rounds
    once to FP32 either way); for FP64 models eps is the FP64 product, as in
    the NumPy 1.x the reference requires.  Pairs closer than eps get no
    ray."""
    rng = np.random.default_rng(17)
    m = 400
    P = rng.uniform(-1, 1, (m, 3)).astype(dt)
    Pj = (P + rng.normal(0, 1, (m, 3)) * np.where(rng.random((m, 1)) < 0.1, 1e-4, 1.0)).astype(dt)
    eps = 1e3 * np.finfo(np.float32).resolution if dt == np.float32 \
        else 1e3 * float(np.finfo(np.float32).resolution)
    # shape.py:357-364, 377-378 for one row at a time
    D = (Pj - P).astype(dt)
    D_norm = np.sqrt(np.sum(D ** 2, axis=1))
    mask = D_norm > eps
    Dn = D[mask] / D_norm[mask].reshape(-1, 1)
    org = P[mask] + eps * Dn
    o, d, ok = oracle_mod.embree_pair_rays(P, Pj)
    assert np.array_equal(ok, mask) and (~mask).sum() > 10
    assert np.array_equal(o[ok], org.astype(np.float32).astype(np.float64))
    assert np.array_equal(d[ok], Dn.astype(np.float32).astype(np.float64))
    N = rng.normal(0, 1, (m, 3))
    N = (N / np.linalg.norm(N, axis=1, keepdims=True)).astype(dt)
    Dsun = np.array([0.3, -0.2, 0.93])
    Dsun /= np.linalg.norm(Dsun)
    o, d = oracle_mod.embree_sun_rays(P, N, Dsun)
    assert np.array_equal(o, (P + eps * N).astype(np.float32).astype(np.float64))
    assert np.array_equal(d, np.tile(Dsun.astype(np.float32).astype(np.float64), (m, 1)))
