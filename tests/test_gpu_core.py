"""GPU parity tests of the visibility / form-factor / solve path against the
CPU oracle.  Integer and mask results must be bit-exact; FP64 solver outputs
agree to 1e-12 relative, FP32 to the stated tolerances."""
import pickle

import numpy as np
import pytest

from fluxpy_amd import meshes

pytestmark = pytest.mark.gpu


def _lib_device():
    from fluxpy_amd import _lib
    return _lib.device()


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def crater():
    return meshes.spherical_crater(npts=20, seed=21)


def _model(V, F, cls="cgal", dtype=np.float64):
    from fluxpy_amd import shape
    C = shape.CgalTrimeshShapeModel if cls == "cgal" else shape.EmbreeTrimeshShapeModel
    return C(V.astype(dtype), F)


# --------------------------------------------------------------------------
# geometry and scene
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_geometry_kernel_bit_exact(torch_dev, oracle_mod, dt):
    from fluxpy_amd.shape import get_geometry_gpu
    V, F = meshes.fractal_terrain(grid=61, levels=6, seed=9)
    V = V.astype(dt)
    P, N, A = get_geometry_gpu(V, F)
    Po, No, Ao = oracle_mod.geometry(V, F)
    assert np.array_equal(P.cpu().numpy(), Po)
    assert np.array_equal(A.cpu().numpy(), Ao)
    assert np.array_equal(N.cpu().numpy(), No)


def test_scene_structure(torch_dev, crater):
    from fluxpy_amd import _lib
    sm = _model(*crater)
    nn = _lib.ctypes.c_int64()
    _lib.call("flux_scene_num_nodes", sm._scene.handle, _lib.ctypes.byref(nn))
    nn = nn.value
    assert nn == 2 * sm.num_faces - 1
    buf = torch_dev.empty(8 * nn * 8, dtype=torch_dev.float32, device=_lib_device())
    _lib.call("flux_scene_copy_nodes", sm._scene.handle, buf.data_ptr(), _lib.stream())
    layouts = buf.cpu().numpy().reshape(8, nn, 8)
    V = sm.V[sm.F]
    boxes0 = None
    for oct in range(8):
        nodes = layouts[oct]
        esc = nodes[:, 3].copy().view(np.int32)
        tri = nodes[:, 7].copy().view(np.int32)
        leaves = tri[tri >= 0]
        assert sorted(leaves.tolist()) == list(range(sm.num_faces))
        assert (esc[tri >= 0] == np.arange(nn)[tri >= 0] + 1).all()
        assert (esc > np.arange(nn)).all() and (esc <= nn).all() and esc[0] == nn
        lo, hi = nodes[:, :3], nodes[:, 4:7]
        for k in np.where(tri >= 0)[0]:
            t = tri[k]
            assert (lo[k] <= V[t].min(0)).all() and (hi[k] >= V[t].max(0)).all()
        # internal boxes contain their two children (preorder: first child =
        # k+1, second = its escape), and the first child's box centre is not
        # later along this octant's diagonal
        sg = np.array([-1.0 if (oct >> c) & 1 else 1.0 for c in range(3)], dtype=np.float32)
        for k in np.where(tri < 0)[0]:
            l, r = k + 1, esc[k + 1]
            for c in (l, r):
                assert (lo[k] <= lo[c]).all() and (hi[k] >= hi[c]).all()
            kl = np.float32(0)
            kr = np.float32(0)
            for c in range(3):
                kl = np.float32(kl + sg[c] * (lo[l, c] + hi[l, c]))
                kr = np.float32(kr + sg[c] * (lo[r, c] + hi[r, c]))
            assert kl <= kr
        # every layout holds the same set of boxes
        key = sorted(zip(map(tuple, lo.tolist()), map(tuple, hi.tolist()), tri.tolist()))
        boxes0 = key if boxes0 is None else boxes0
        assert key == boxes0


def test_scene_single_facet_and_empty_queries(torch_dev):
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.0]])
    F = np.array([[0, 1, 2]])
    sm = _model(V, F)
    assert sm.get_visibility_matrix().tolist() == [[False]]
    assert sm.intersect1(np.array([0.2, 0.2, 1.0]), np.array([0, 0, -1.0]))[0] == 0
    assert sm.intersect1(np.array([2.0, 2, 1]), np.array([0, 0, -1.0])) is None
    assert sm.get_visibility([], [0]).shape == (0, 1)


def test_bad_face_index_raises(torch_dev):
    from fluxpy_amd import shape
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.0]])
    # V[F] raises IndexError in the geometry helpers, as in the reference
    with pytest.raises((IndexError, ValueError)):
        shape.CgalTrimeshShapeModel(V, np.array([[0, 1, 3]]))
    with pytest.raises(ValueError):
        shape._Scene(V, np.array([[0, 1, 3]]))


# --------------------------------------------------------------------------
# reference tests ported (tests/test_shape.py, tests/test_form_factors.py)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [4, 5])
@pytest.mark.parametrize("cls", ["cgal", "embree"])
def test_visibility_matrix_for_sphere(torch_dev, n, cls):
    V, F = meshes.icosphere(n)
    sm = _model(V, F, cls)
    sm.N[(sm.N * sm.P).sum(1) < 0] *= -1
    nf = sm.num_faces
    offdiag = ~np.eye(nf, dtype=bool)
    vis = sm.get_visibility_matrix(oriented=False)
    if cls == "cgal":
        assert (vis == offdiag).all()
    else:  # Embree: the diagonal is visible by default (shape.py:360-362,392)
        assert vis[offdiag].all() and vis.diagonal().all()
    vis = sm.get_visibility_matrix(oriented=True)
    assert not vis.any()
    sm.N *= -1
    vis = sm.get_visibility_matrix(oriented=True)
    assert vis[offdiag].all()


@pytest.mark.parametrize("n", [4, 5])
@pytest.mark.parametrize("cls", ["cgal", "embree"])
def test_is_occluded_for_sphere(torch_dev, n, cls, rng):
    V, F = meshes.icosphere(n)
    sm = _model(V, F, cls)
    sm.N[(sm.N * sm.P).sum(1) < 0] *= -1
    for _ in range(5):
        D = rng.standard_normal(3)
        D /= np.linalg.norm(D)
        occ = sm.is_occluded(np.arange(sm.num_faces), D)
        nd = sm.N @ D
        bad = occ != (nd < 0)
        if cls == "cgal":
            assert not bad.any()
        else:  # grazing band of the 1e-3 normal offset
            assert ((nd[bad] > -0.02) & (nd[bad] < 0)).all()


@pytest.mark.parametrize("cls", ["cgal", "embree"])
def test_visibility_and_1_to_N_equivalent(torch_dev, cls):
    V, F = meshes.icosphere(4)
    sm = _model(V, F, cls)
    sm.N[(sm.N * sm.P).sum(1) > 0] *= -1
    I = np.arange(sm.num_faces)
    vis = sm.get_visibility(I, I)
    for i in I[::7]:
        assert (vis[i] == sm.get_visibility_1_to_N(i, I)).all()


@pytest.mark.parametrize("n", [4, 5])
@pytest.mark.parametrize("cls", ["cgal", "embree"])
def test_form_factor_matrix_for_sphere(torch_dev, n, cls):
    from fluxpy_amd.form_factors import get_form_factor_matrix
    V, F = meshes.icosphere(n)
    sm = _model(V, F, cls)
    outward = (sm.P * sm.N).sum(1) > 0
    sm.N[outward] *= -1
    FF = get_form_factor_matrix(sm).toarray()
    off = ~np.eye(sm.num_faces, dtype=bool)
    assert (FF.diagonal() == 0).all() and (FF[off] != 0).all()
    sm.N *= -1
    assert get_form_factor_matrix(sm).nnz == 0


# --------------------------------------------------------------------------
# bit-exact parity with the oracle
# --------------------------------------------------------------------------

@pytest.mark.parametrize("mode", ["cgal", "embree"])
def test_visibility_bit_exact_crater(torch_dev, oracle_mod, crater, mode):
    V, F = crater
    sm = _model(V, F, mode)
    I = np.arange(sm.num_faces)
    S = oracle_mod.OracleScene(V, F)
    m = 0 if mode == "cgal" else 1
    vis_o = S.vis_mn(I, I, m, sm.P)
    vis = sm.get_visibility(I, I)
    bad = np.argwhere(vis != vis_o)
    assert bad.size == 0, f"{len(bad)} mismatched pairs, first: {bad[:10].tolist()}"


def test_oriented_visibility_subsets(torch_dev, oracle_mod, crater):
    """get_visibility(I, J, oriented=True) with I != J: the oracle's mask
    with every pair whose facet-id direction P[J[q]] - P[I[p]] is not in
    front of N[I[p]] cleared; rows equal get_visibility_1_to_N; on I = J =
    arange it equals the reference's positional expression."""
    V, F = crater
    sm = _model(V, F)
    n = sm.num_faces
    S = oracle_mod.OracleScene(V, F)
    rng = np.random.default_rng(17)
    I = np.sort(rng.choice(n, 120, replace=False))
    J = rng.permutation(n)[:300]
    vis = sm.get_visibility(I, J, oriented=True)
    ref = S.vis_mn(I, J).astype(bool)
    front = ((sm.P[J][None, :, :] - sm.P[I][:, None, :]) * sm.N[I][:, None, :]).sum(-1) > 0
    assert np.array_equal(vis, ref & front)
    assert (ref & ~front).any()          # the orientation test removed pairs
    for p in (0, 17, 119):
        assert np.array_equal(sm.get_visibility_1_to_N(I[p], J, oriented=True), vis[p])
    # full index sets: identical to the reference's positional form
    A = np.arange(n)
    full = sm.get_visibility_matrix(oriented=True)
    base = S.vis_mn(A, A).astype(bool)
    I_, J_ = np.where(base)
    drop = ((sm.P[J_] - sm.P[I_]) * sm.N[I_]).sum(1) <= 0
    base[I_[drop], J_[drop]] = False
    assert np.array_equal(full, base)


@pytest.mark.parametrize("dt", [np.float64, np.float32])
@pytest.mark.parametrize("mode", ["cgal", "embree"])
def test_form_factor_bit_exact(torch_dev, oracle_mod, crater, dt, mode):
    from fluxpy_amd.form_factors import get_form_factor_matrix
    V, F = crater
    sm = _model(V, F, mode, dt)
    S = oracle_mod.OracleScene(V, F)
    m = 0 if mode == "cgal" else 1
    FF = get_form_factor_matrix(sm)
    FFo = S.form_factor_matrix(sm.P, sm.N, sm.A, mode=m)
    assert FF.dtype == dt
    assert FF.shape == FFo.shape
    assert np.array_equal(FF.indptr, FFo.indptr)
    assert np.array_equal(FF.indices, FFo.indices)
    assert np.array_equal(FF.data, FFo.data)


def test_form_factor_subsets_and_eps(torch_dev, oracle_mod, crater, rng):
    from fluxpy_amd.form_factors import get_form_factor_matrix
    V, F = crater
    sm = _model(V, F)
    S = oracle_mod.OracleScene(V, F)
    I = rng.choice(sm.num_faces, 77, replace=False)
    J = rng.choice(sm.num_faces, 300, replace=False)
    for eps in (1e-5, 1e-3, 0.0):
        FF = get_form_factor_matrix(sm, I, J, eps=eps)
        FFo = S.form_factor_matrix(sm.P, sm.N, sm.A, I, J, eps=eps)
        assert (FF != FFo).nnz == 0 and np.array_equal(FF.indices, FFo.indices)
    assert get_form_factor_matrix(sm, [], J).shape == (0, 300)


def test_occlusion_intersect_bit_exact(torch_dev, oracle_mod, crater, rng):
    V, F = crater
    S = oracle_mod.OracleScene(V, F)
    I = np.arange(F.shape[0])
    D = meshes.sun_directions(7, seed=3)
    for cls, m in (("cgal", 0), ("embree", 1)):
        sm = _model(V, F, cls)
        for d in D:
            occ = sm.is_occluded(I, d)
            occ_o = S.occluded(I, d, mode=m, P=sm.P, N=sm.N)[:, 0]
            assert np.array_equal(occ, occ_o)
    sm = _model(V, F)
    occ2 = sm.is_occluded(I, D)
    assert occ2.shape == (len(I), len(D))
    assert np.array_equal(occ2, S.occluded(I, D))
    for _ in range(20):
        x = np.r_[rng.uniform(-1, 1, 2), 1.0]
        d = np.r_[rng.uniform(-0.3, 0.3, 2), -1.0]
        got = sm.intersect1(x, d)
        h, xt = S.intersect1(x, d)
        if h[0] < 0:
            assert got is None
        else:
            assert got[0] == h[0] and np.array_equal(got[1], xt[0])


def test_direct_irradiance_bit_exact(torch_dev, oracle_mod, crater):
    V, F = crater
    sm = _model(V, F)
    S = oracle_mod.OracleScene(V, F)
    I = np.arange(sm.num_faces)
    D = meshes.sun_directions(5, seed=4)
    E = sm.get_direct_irradiance_multi(1365.0, D)
    for k, d in enumerate(D):
        Eo = oracle_mod.direct_irradiance(sm.N, S.occluded(I, d)[:, 0], 1365.0, d)
        assert np.array_equal(E[:, k], Eo)
        assert np.array_equal(sm.get_direct_irradiance(1365.0, d), Eo)


@pytest.mark.slow
def test_form_factor_rows_at_40k(torch_dev, oracle_mod, rng):
    """Full 41,472-facet terrain: a random sample of rows equals the oracle."""
    from fluxpy_amd.form_factors import get_form_factor_matrix
    V, F = meshes.fractal_terrain()
    sm = _model(V, F, dtype=np.float32)
    S = oracle_mod.OracleScene(V, F)
    I = np.sort(rng.choice(sm.num_faces, 48, replace=False))
    FF = get_form_factor_matrix(sm, I)
    FFo = S.form_factor_matrix(sm.P, sm.N, sm.A, I)
    assert np.array_equal(FF.indptr, FFo.indptr)
    assert np.array_equal(FF.indices, FFo.indices)
    assert np.array_equal(FF.data, FFo.data)


# --------------------------------------------------------------------------
# device CSR, operator protocol, solvers, temperature
# --------------------------------------------------------------------------

def test_csr_transpose_and_products(torch_dev, crater, rng):
    from fluxpy_amd.form_factors import FormFactorMatrix
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    sp = Fm.tocsr()
    spT = Fm.transpose_csr().to_scipy()
    ref = sp.T.tocsr()
    ref.sort_indices()
    assert np.array_equal(spT.indptr, ref.indptr) and np.array_equal(spT.indices, ref.indices)
    assert np.array_equal(spT.data, ref.data)
    x = rng.standard_normal(sm.num_faces)
    X = rng.standard_normal((sm.num_faces, 5))
    assert np.allclose(Fm @ x, sp @ x, rtol=1e-13, atol=1e-15)
    assert np.allclose(Fm @ X, sp @ X, rtol=1e-13, atol=1e-15)
    assert np.allclose(Fm.T @ x, sp.T @ x, rtol=1e-13, atol=1e-15)
    Xd = torch_dev.tensor(X, device=_lib_device())
    assert np.allclose((Fm @ Xd).cpu().numpy(), sp @ X, rtol=1e-13, atol=1e-15)
    Fm2 = pickle.loads(pickle.dumps(Fm))
    assert (Fm2.tocsr() != sp).nnz == 0


@pytest.mark.parametrize("place", ["right", "left"])
def test_jacobi_parity(torch_dev, oracle_mod, crater, place):
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.solve import solve_radiosity
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    sp = Fm.tocsr()
    D = meshes.sun_directions(3, seed=8)
    E = sm.get_direct_irradiance_multi(1365.0, D)
    for rho in (0.12, np.linspace(0.08, 0.16, sm.num_faces)):
        B, it = solve_radiosity(Fm, E[:, 0], rho, place)
        Bo, ito = oracle_mod.solve_radiosity_jacobi(sp, E[:, 0], rho, place)
        assert abs(it - ito) <= 1
        assert abs(B - Bo).max() <= 1e-12 * abs(Bo).max()
        Bb, itb = solve_radiosity(Fm, E, rho, place)
        Bbo = oracle_mod.solve_radiosity_dense(sp, E, rho, place)
        assert abs(Bb - Bbo).max() <= 1e-12 * abs(Bbo).max()


def test_cg_gmres_against_dense(torch_dev, oracle_mod, crater):
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.solve import solve_radiosity
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    sp = Fm.tocsr()
    E = sm.get_direct_irradiance_multi(1365.0, meshes.sun_directions(2, seed=2))
    Bd = oracle_mod.solve_radiosity_dense(sp, E, 0.3, "left")
    B, it = solve_radiosity(Fm, E[:, 0].copy(), 0.3, "left", method="cg")
    assert abs(B - Bd[:, 0]).max() <= 1e-10 * abs(Bd).max()
    for place in ("right", "left"):
        Bd = oracle_mod.solve_radiosity_dense(sp, E, 0.3, place)
        B, it = solve_radiosity(Fm, E, 0.3, place, method="gmres", tol=1e-12)
        assert abs(B - Bd).max() <= 1e-11 * abs(Bd).max()
        assert it < 60


@pytest.mark.parametrize("restart", [3, 7, 40])
def test_gmres_restarts_against_dense(torch_dev, oracle_mod, crater, restart):
    """The device-resident GMRES with restarts shorter than convergence
    (several cycles), longer than one projection call (40 > 32 vectors),
    three columns converging at different rates, a starting guess and FP32
    input, against the oracle's dense LU."""
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.solve import solve_radiosity
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    sp = Fm.tocsr()
    E = sm.get_direct_irradiance_multi(1365.0, meshes.sun_directions(3, seed=8))
    E[:, 2] *= 1e-3
    Bd = oracle_mod.solve_radiosity_dense(sp, E, 0.6, "right")
    B, it = solve_radiosity(Fm, E, 0.6, "right", method="gmres", tol=1e-12, restart=restart)
    assert abs(B - Bd).max(0).max() <= 1e-10 * abs(Bd).max()
    assert (abs(B - Bd).max(0) <= 1e-10 * abs(Bd).max(0)).all()
    x0 = Bd * (1 + 1e-3)
    B2, it2 = solve_radiosity(Fm, E, 0.6, "right", method="gmres", tol=1e-12, restart=restart,
                              x0=x0)
    assert abs(B2 - Bd).max() <= 1e-10 * abs(Bd).max() and it2 <= it
    B3, _ = solve_radiosity(Fm, E.astype(np.float32), 0.6, "right", method="gmres", tol=1e-6,
                            restart=restart)
    assert B3.dtype == np.float32 and abs(B3 - Bd).max() <= 1e-5 * abs(Bd).max()


def test_steady_state_temperature(torch_dev, oracle_mod, crater):
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.model import compute_steady_state_temp
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    sp = Fm.tocsr()
    D = meshes.sun_directions(3, seed=5)
    E = sm.get_direct_irradiance_multi(1365.0, D)
    T = compute_steady_state_temp(Fm, E[:, 0].copy(), 0.12, 0.95)
    To = oracle_mod.steady_state_temp(sp, E[:, 0].copy(), 0.12, 0.95)
    assert abs(T - To).max() <= 1e-10 * To.max()
    T2 = compute_steady_state_temp(Fm, E, 0.12, 0.95)
    To2 = oracle_mod.steady_state_temp(sp, E, 0.12, 0.95)
    assert T2.shape == To2.shape == (3, sm.num_faces)
    assert abs(T2 - To2).max() <= 1e-10 * To2.max()


def test_solver_errors(torch_dev, crater):
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.solve import solve_radiosity
    V, F = crater
    sm = _model(V, F)
    Fm = FormFactorMatrix.from_shape_model(sm)
    E = np.ones(sm.num_faces)
    with pytest.raises(Exception):
        solve_radiosity(Fm, E, 0.1, "middle")
    with pytest.raises(Exception):
        solve_radiosity(Fm, E, 0.1, method="lu")
    with pytest.raises(ValueError):
        solve_radiosity(Fm, np.ones(3), 0.1)
    with pytest.warns(RuntimeWarning):
        solve_radiosity(Fm, E, 1.0, maxiter=3)
