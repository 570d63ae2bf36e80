"""GPU tests of the low-rank path: the drop-in rank estimation
(linalg.py:8-50 of fluxpy) against LAPACK singular values, the sketch's
range check, the compressed operator in every stored precision, the
(I - rho F) kernel operator, and the plain configuration (spherical crater,
1,922 facets, FP64) at its full size."""
import numpy as np
import pytest

from fluxpy_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    torch.cuda.set_device(0)
    return torch


@pytest.fixture(scope="module")
def crater32(oracle_mod):
    """Config-1 mesh: spherical crater, Delaunay of 32 x 32 jittered points."""
    from fluxpy_amd.form_factors import get_form_factor_matrix
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    V, F = meshes.spherical_crater(npts=32, seed=1001)
    sm = CgalTrimeshShapeModel(V, F)
    return sm, get_form_factor_matrix(sm)


def _far_block(sm, FF):
    """A well-separated block of F: facets of the bowl (r < 0.45) with
    x < -0.1 against those with x > 0.1 (113 x 111, every pair visible),
    numerically low-rank (s_i / s_0 below 1e-3 from i = 9)."""
    P = sm.P
    inside = np.hypot(P[:, 0], P[:, 1]) < 0.45
    r = np.where(inside & (P[:, 0] < -0.1))[0]
    c = np.where(inside & (P[:, 0] > 0.1))[0]
    return FF.tocsr()[r][:, c]


# --------------------------------------------------------------------------
# linalg drop-ins
# --------------------------------------------------------------------------

@pytest.mark.parametrize("dt", [np.float64, np.float32])
def test_sparse_svd_matches_lapack(torch_dev, crater32, dt):
    from fluxpy_amd import linalg
    sm, FF = crater32
    A = _far_block(sm, FF).astype(dt)
    assert A.nnz > 1000
    s_ref = np.linalg.svd(A.toarray().astype(np.float64), compute_uv=False)
    k = 6
    U, S, Vt = linalg.sparse_svd(A, k)
    assert U.shape == (A.shape[0], k) and Vt.shape == (k, A.shape[1]) and S.shape == (k,)
    assert np.all(np.diff(S) <= 0)
    # leading singular values: a sketch of 2k + 8 = 20 columns resolves
    # s_i to about (s_20 / s_i)^2 relative (5e-5 for s_3 of this block;
    # FP32 storage adds 2^-24 relative to s_0)
    S32 = np.linalg.svd(A.toarray().astype(np.float64), compute_uv=False)
    assert np.allclose(S[:4], S32[:4], rtol=2e-4, atol=1e-7 * s_ref[0])
    rel = 1e-12 if dt == np.float64 else 1e-6
    # orthonormal singular vectors
    Ud, Vd = U.toarray(), Vt.toarray()
    # (the factors come back in the block's dtype; Vt's rows are X^T Z^T / s,
    # whose rounding grows like eps s_0^2 / s_i^2)
    orth = 1e-10 if dt == np.float64 else 1e-6
    assert abs(Ud.T @ Ud - np.eye(k)).max() < orth
    assert abs(Vd @ Vd.T - np.eye(k)).max() < max(orth, 1e-8)
    # and a k-term approximation no worse than the optimum by more than the
    # sketch's tail
    err = np.linalg.norm(A.toarray() - (Ud * S) @ Vd, 2)
    assert err <= 1.5 * s_ref[k] + 10 * rel * s_ref[0]


def test_estimate_rank_rule(torch_dev, crater32):
    """Rank = first r with s_r / s_0 <= tol (linalg.py:42-46), compared with
    LAPACK's singular values away from the threshold; None when the factors
    would take max_nbytes; 0 / 1 for empty / 1 x 1 input."""
    import scipy.sparse
    from fluxpy_amd import linalg
    sm, FF = crater32
    A = _far_block(sm, FF)
    s_ref = np.linalg.svd(A.toarray(), compute_uv=False)
    ratio = s_ref / s_ref[0]
    for tol in (1e-2, 1e-3, 1e-4):
        out = linalg.estimate_rank(A, tol)
        assert out is not None
        U, S, Vt, thr = out
        r = S.size
        r_ref = int(np.where(ratio <= tol)[0][0])
        near = np.any(abs(np.log(ratio[max(0, r_ref - 1):r_ref + 2] / tol)) < np.log(1.2))
        assert r == r_ref or near, (tol, r, r_ref)
        assert thr <= tol
        assert U.shape == (A.shape[0], r) and Vt.shape == (r, A.shape[1])
        assert scipy.sparse.issparse(U) and scipy.sparse.issparse(Vt)
        approx = (U.toarray() * S) @ Vt.toarray()
        assert np.linalg.norm(A.toarray() - approx, 2) <= 3 * tol * s_ref[0]
    # byte cap: the rank-r factors of a 160 x 160 block cannot fit in 100 bytes
    assert linalg.estimate_rank(A, 1e-4, max_nbytes=100) is None
    assert linalg.estimate_rank(scipy.sparse.csr_matrix((0, 5)), 1e-3) == 0
    assert linalg.estimate_rank(scipy.sparse.csr_matrix(np.ones((1, 1))), 1e-3) == 1


def test_sketch_reports_unsorted_rows(torch_dev):
    """A CSR whose row is not sorted by column holds ids outside the block
    between the two binary-search bounds; the sketch kernel skips them and
    the build raises instead of reading outside the block."""
    from fluxpy_amd import _hbuild, _kernels, _lib
    t = torch_dev
    n = 64
    rng = np.random.default_rng(0)
    indptr = np.arange(0, n * n + 1, n, dtype=np.int64)
    indices = np.tile(np.arange(n, dtype=np.int32), n)
    # row 5 interleaves the two halves: 32, 0, 33, 1, ...
    row5 = np.empty(n, dtype=np.int32)
    row5[0::2], row5[1::2] = np.arange(32, 64), np.arange(0, 32)
    indices[5 * n:6 * n] = row5

    def lower_bound(a, lo, hi, key):
        while lo < hi:
            mid = (lo + hi) // 2
            lo, hi = (mid + 1, hi) if a[mid] < key else (lo, mid)
        return lo

    # the kernel's two binary searches over that row for columns [32, 64)
    k0 = lower_bound(row5, 0, n, 32)
    k1 = lower_bound(row5, k0, n, 64)
    assert ((row5[k0:k1] < 32) | (row5[k0:k1] >= 64)).any()
    data = rng.random(n * n)
    dev = _lib.device()
    F = _kernels.DeviceCSR(t.tensor(indptr, device=dev), t.tensor(indices, device=dev),
                           t.tensor(data, device=dev), (n, n))
    FT = _kernels.csr_transpose(F)
    with pytest.raises(_lib.FluxError, match="sorted"):
        _hbuild._SvdBatch(F, FT, [(0, 32, 32, 64)], [16], 1)
    # the same block of a sorted matrix passes
    indices[5 * n:6 * n] = np.arange(n, dtype=np.int32)
    F = _kernels.DeviceCSR(t.tensor(indptr, device=dev), t.tensor(indices, device=dev),
                           t.tensor(data, device=dev), (n, n))
    sb = _hbuild._SvdBatch(F, _kernels.csr_transpose(F), [(0, 32, 32, 64)], [16], 1)
    assert sb.sigma[0].size == 16


# --------------------------------------------------------------------------
# the compressed operator in each stored precision, and (I - rho F)
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def compressed64(crater32):
    from fluxpy_amd.compressed_form_factors import CompressedFormFactorMatrix
    sm, FF = crater32
    H = CompressedFormFactorMatrix(sm, tol=1e-2, min_size=1024)
    assert H.build_stats["svd_leaves"] > 0
    return H


@pytest.mark.parametrize("values,tol", [("fp64", 1e-13), ("fp32", 2e-6), ("fp16", 2e-3),
                                        ("bf16", 1.6e-2)])
def test_device_operator_precisions(torch_dev, crater32, compressed64, values, tol):
    """Near-field values and far-field factors stored in `values` (products
    accumulated in FP32, FP64 for 'fp64') against the same compressed
    operator in FP64; the stated bound is a few units of the storage
    format's rounding relative to max |y|."""
    from fluxpy_amd import _lib
    sm, FF = crater32
    H = compressed64
    rng = np.random.default_rng(5)
    X = np.abs(rng.standard_normal((sm.num_faces, 3)))
    op64 = H.device_operator("fp64")
    opv = H.device_operator(values)
    assert opv.nfar == op64.nfar > 0
    Xd = _lib.to_dev(X)
    Xi = op64.to_internal(Xd)
    Y64 = op64.apply(Xi).cpu().numpy()
    Xv = Xi if values == "fp64" else Xi.to(torch_dev.float32)
    Yv = opv.apply(Xv).double().cpu().numpy()
    err = abs(Yv - Y64).max() / abs(Y64).max()
    assert err <= tol, (values, err)
    # and the FP64 operator against the dense F within the compression bound
    Yd = FF @ X
    Y = H.apply(Xd).cpu().numpy()
    assert abs(Y - Yd).max() / abs(Yd).max() <= 50 * 1e-2


@pytest.mark.parametrize("place", ["left", "right"])
def test_compressed_kernel_matrix(torch_dev, crater32, compressed64, place):
    """K x = x - rho F x / x - F (rho x) (compressed_form_factors.py:875-913)
    against the dense (I - rho F) x with the compressed F's error bound, and
    exactly against the same expression through the compressed product."""
    from fluxpy_amd.compressed_form_factors import CompressedKernelMatrix
    sm, FF = crater32
    H = compressed64
    K = CompressedKernelMatrix(H, 0.3, place)
    x = np.random.default_rng(6).standard_normal(sm.num_faces)
    y = K @ x
    Hx = H @ (0.3 * x) if place == "right" else 0.3 * (H @ x)
    assert np.allclose(y, x - Hx, rtol=0, atol=1e-14 * abs(x).max())
    yd = x - 0.3 * (FF @ x)
    assert abs(y - yd).max() <= 0.3 * 50 * 1e-2 * abs(FF @ x).max()


# --------------------------------------------------------------------------
# plain configuration at its stated size
# --------------------------------------------------------------------------

def test_plain_configuration_full_size(torch_dev, oracle_mod, crater32):
    """Spherical crater (32 x 32 jittered points, 1,922 facets, FP64), one
    sun at 10 degrees elevation, rho = 0.12, eps = 0.95: F bit-exact, GMRES
    to 1e-12 against the oracle's dense LU, both temperature forms against
    the oracle."""
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.model import compute_steady_state_temp, steady_state_temp_closed
    from fluxpy_amd.solve import solve_radiosity
    sm, FF = crater32
    assert sm.num_faces == 1922
    S = oracle_mod.OracleScene(sm.V, sm.F)
    FFo = S.form_factor_matrix(sm.P, sm.N, sm.A)
    assert np.array_equal(FF.indptr, FFo.indptr) and np.array_equal(FF.indices, FFo.indices)
    assert np.array_equal(FF.data, FFo.data)
    d = np.array([np.cos(np.deg2rad(10)), 0, np.sin(np.deg2rad(10))])
    E = sm.get_direct_irradiance(1365.0, d)
    I = np.arange(sm.num_faces)
    Eo = oracle_mod.direct_irradiance(sm.N, S.occluded(I, d)[:, 0], 1365.0, d)
    assert np.array_equal(E, Eo)
    Fm = FormFactorMatrix.from_shape_model(sm)
    B, it = solve_radiosity(Fm, E, 0.12, method="gmres", tol=1e-12)
    Bo = oracle_mod.solve_radiosity_dense(FFo, Eo, 0.12)
    assert abs(B - Bo).max() <= 1e-11 * abs(Bo).max()
    # residual of the returned B in the oracle's F
    r = Eo - (B - 0.12 * (FFo @ B))
    assert np.linalg.norm(r) <= 1e-11 * np.linalg.norm(Eo)
    T = compute_steady_state_temp(Fm, E, 0.12, 0.95)
    To = oracle_mod.steady_state_temp(FFo, Eo, 0.12, 0.95)
    assert abs(T - To).max() <= 1e-10 * To.max()
    from scipy.constants import Stefan_Boltzmann
    Tc = steady_state_temp_closed(Fm, E, B, 0.12, 0.95)
    Tco = ((1 - 0.12) * (Eo + FFo @ Bo) / (0.95 * Stefan_Boltzmann)) ** 0.25
    assert abs(Tc - Tco).max() <= 1e-10 * Tco.max()
    # a planted change (emissivity 0.9) moves the closed form well outside
    Tbad = ((1 - 0.12) * (Eo + FFo @ Bo) / (0.9 * Stefan_Boltzmann)) ** 0.25
    assert abs(Tc - Tbad).max() > 1e-3 * Tco.max()


# --------------------------------------------------------------------------
# the far field on MFMA (16-bit factors) and the refined solve with it
# --------------------------------------------------------------------------

@pytest.mark.parametrize("values,tol", [("bf16", 1e-2), ("fp16", 1.5e-3)])
@pytest.mark.parametrize("left", [False, True])
def test_far_field_mfma(torch_dev, compressed64, values, tol, left):
    """FarField16.mm (v_mfma_f32_16x16x16 bf16 / f16) against the FP64
    far-field product of the same 16-bit X with the master factors; the
    bound is three roundings of the storage format (factors, S folded into
    Vt, the projected T) relative to max |y|."""
    from fluxpy_amd import _hbuild, _lib
    t = torch_dev
    H = compressed64
    master = _hbuild.far_master(H)
    far = _hbuild.far16(H, values)
    n = H.shape[0]
    gen = t.Generator(device="cuda").manual_seed(3)
    for k in (64, 16, 3):
        X16 = t.rand((n, k), generator=gen, device="cuda").to(far.dtype)
        rs32 = t.rand(n, generator=gen, device="cuda") if left else None
        Y = t.zeros((n, k), dtype=t.float32, device="cuda")
        far.mm(X16, Y, rscale=rs32)
        Yref = t.zeros((n, k), dtype=t.float64, device="cuda")
        master.far_apply(X16.to(t.float64), Yref,
                         rscale=None if rs32 is None else rs32.to(t.float64))
        err = float((Y.double() - Yref).abs().max() / Yref.abs().max())
        assert err <= tol, (values, k, left, err)
        # repeated calls give identical bits (fixed summation order)
        Y2 = t.zeros_like(Y)
        far.mm(X16, Y2, rscale=rs32)
        assert t.equal(Y, Y2)


@pytest.mark.parametrize("left", [False, True])
def test_refined_solve_with_far_field(torch_dev, oracle_mod, crater32, compressed64, left):
    """The batched BF16 + refinement solve on a compressed matrix with
    low-rank leaves (near-field panels + MFMA far field in the sweeps, int8
    residual + FP64 far field) against GMRES on the same compressed operator
    (FP64, 1e-13): within 1e-6 relative per sun."""
    from fluxpy_amd.solve import solve_radiosity, solve_radiosity_refined
    sm, FF = crater32
    H = compressed64
    rho = np.random.default_rng(12).uniform(0.1, 0.3, sm.num_faces)
    E = sm.get_direct_irradiance_multi(1365.0, meshes.sun_directions(8, seed=13))
    place = "left" if left else "right"
    B, info = solve_radiosity_refined(H, H, E, rho, albedo_placement=place, rtol=1e-7,
                                      inner_iters=2, return_info=True)
    Bg, _ = solve_radiosity(H, E, rho, place, method="gmres", tol=1e-13)
    err = abs(B - Bg).max(0) / abs(Bg).max(0)
    assert err.max() <= 1e-6, (err.max(), info)
    assert info["residual_history"][-1] <= 1e-7
    # and within the compression error of the exact (dense) solution
    Bd = oracle_mod.solve_radiosity_dense(FF, E, rho, place)
    assert abs(B - Bd).max() <= 0.05 * abs(Bd).max()


def _far_problem(rng, n, ranks, fdt):
    """Far blocks on an n x n matrix with nested and overlapping ranges:
    (desc rows r0, r1, c0, c1, rank, offU, offVt, offS; U, S, Vt flat)."""
    desc, Us, Ss, Vts = [], [], [], []
    oU = oV = oS = 0
    for r in ranks:
        r0 = int(rng.integers(0, n - 40))
        r1 = int(min(n, r0 + rng.integers(20, 400)))
        c0 = int(rng.integers(0, n - 40))
        c1 = int(min(n, c0 + rng.integers(20, 400)))
        m, nc = r1 - r0, c1 - c0
        desc.append([r0, r1, c0, c1, r, oU, oV, oS])
        Us.append(rng.standard_normal((m, r)).astype(fdt))
        Ss.append(rng.uniform(0.1, 1.0, r).astype(fdt))
        Vts.append(rng.standard_normal((r, nc)).astype(fdt))
        oU += m * r
        oV += r * nc
        oS += r
    return np.array(desc, dtype=np.int64), Us, Ss, Vts


@pytest.mark.parametrize("fdt,xdt", [(np.float64, np.float64), (np.float32, np.float64),
                                     (np.float32, np.float32)])
@pytest.mark.parametrize("trans", [0, 1])
@pytest.mark.parametrize("k", [1, 5, 64, 70])
def test_far_matmat_against_dense(torch_dev, fdt, xdt, trans, k):
    """flux_far_matmat (FP64 vectors on FP64 MFMA, FP32 vectors per element)
    against the dense sum of U_b diag(S_b) Vt_b (or its transpose) in FP64:
    ranks 1 to 300 (above 256 the projection runs in rank chunks), ranges
    that nest and overlap, row scaling, Y accumulated into."""
    t = torch_dev
    from fluxpy_amd import _hbuild, _lib
    rng = np.random.default_rng(11 + trans + 3 * k)
    n = 700
    ranks = [1, 3, 16, 17, 40, 64, 100, 300, 7, 33]
    desc, Us, Ss, Vts = _far_problem(rng, n, ranks, fdt)
    Fd = np.zeros((n, n))
    for (r0, r1, c0, c1, *_), u, s, vt in zip(desc, Us, Ss, Vts):
        Fd[r0:r1, c0:c1] += (u.astype(np.float64) * s.astype(np.float64)) @ vt.astype(np.float64)
    if trans:
        Fd = Fd.T
    lo, hi = (desc[:, 2], desc[:, 3]) if trans else (desc[:, 0], desc[:, 1])
    tptr, tblk = _hbuild.tile_lists(lo, hi, n)
    items, red, nslots = _hbuild.far_plan(tptr, tblk, ranks, max_steps=5)
    assert len(red) > 0           # some tiles are split into partial sums
    tdt = t.float64 if fdt == np.float64 else t.float32
    xt = t.float64 if xdt == np.float64 else t.float32
    dev = "cuda"
    U = t.tensor(np.concatenate([u.ravel() for u in Us]), dtype=tdt, device=dev)
    S = t.tensor(np.concatenate(Ss), dtype=tdt, device=dev)
    Vt = t.tensor(np.concatenate([v.ravel() for v in Vts]), dtype=tdt, device=dev)
    X = rng.standard_normal((n, k))
    Y0 = rng.standard_normal((n, k))
    rs = rng.uniform(0.5, 1.5, n)
    ts = np.concatenate([[0], np.cumsum(np.array(ranks) * k)])
    d_desc = t.tensor(desc, device=dev)
    d_ts = t.tensor(ts[:-1], device=dev)
    d_tptr = t.tensor(tptr, device=dev)
    d_tblk = t.tensor(tblk, device=dev)
    d_items = t.tensor(items, device=dev)
    d_red = t.tensor(red, device=dev)
    d_rs = t.tensor(rs, device=dev)
    for scaled in (False, True):
        Xd = t.tensor(X, dtype=xt, device=dev)
        Yd = t.tensor(Y0, dtype=xt, device=dev)
        T = t.empty(int(_lib.load().flux_far_work_elems(int(ts[-1]), k, nslots,
                                                         _lib.dtype_code(xt))),
                    dtype=xt, device=dev)
        _lib.call("flux_far_matmat", n, k, len(ranks), d_desc.data_ptr(), d_ts.data_ptr(),
                  int(ts[-1]), U.data_ptr(), S.data_ptr(), Vt.data_ptr(),
                  _lib.dtype_code(tdt), max(ranks), d_tptr.data_ptr(), d_tblk.data_ptr(),
                  len(items), d_items.data_ptr(), len(red), d_red.data_ptr(), nslots,
                  Xd.data_ptr(), Yd.data_ptr(), _lib.dtype_code(xt), T.data_ptr(), trans,
                  d_rs.data_ptr() if scaled else None, _lib.stream())
        got = Yd.cpu().numpy().astype(np.float64)
        Xr = X.astype(np.float32).astype(np.float64) if xdt == np.float32 else X
        want = Y0 + (rs[:, None] if scaled else 1.0) * (Fd @ Xr)
        scale = np.abs(Fd).sum(1).max() * np.abs(X).max()
        tol = 1e-13 if xdt == np.float64 else 3e-6
        assert np.abs(got - want).max() <= tol * scale, (scaled, np.abs(got - want).max() / scale)
