"""GPU tests of the int8-sliced panels (the FP64-accurate refinement residual
on int8 MFMA): the slices decode to the CSR values within the slicing bound,
products and residuals stay within the a-priori error bound against FP64,
and refined solves with this residual meet 1e-6 against the dense FP64
solution (through the solver and through the public solve API)."""
import numpy as np
import pytest

from fluxpy_amd import meshes

pytestmark = pytest.mark.gpu

SLICES = 5


def _dev():
    from fluxpy_amd import _lib
    return _lib.device()


def _exp(m):
    """frexp exponent (m < 2^e), 0 where m == 0."""
    e = np.frexp(m)[1]
    return np.where(m > 0, e, 0)


@pytest.fixture(scope="module")
def setup():
    import torch
    torch.cuda.set_device(0)
    from fluxpy_amd.compressed_form_factors import _ClusterTree
    from fluxpy_amd import _kernels
    from fluxpy_amd._panel import PanelMatrix, SlicedPanelMatrix
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    V, F = meshes.fractal_terrain(grid=49, levels=6, seed=11)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    ct = _ClusterTree(np.asarray(sm.P, dtype=np.float64), 2)
    csr = _kernels.ff_csr(sm, ct.perm, ct.perm)
    Pb = PanelMatrix(csr, "bf16")
    P8 = SlicedPanelMatrix(csr, Pb, keep_dense=True)
    Pb._default_plan = Pb.plan
    return torch, sm, csr, Pb, P8, ct


def _bound(sp, X):
    """A-priori error bound of the sliced product per (row, rhs):
    6.2 * 2^(e_r + e_k - 35) * nnz_r plus FP64 rounding."""
    absp = abs(sp)
    er = _exp(np.asarray(absp.max(axis=1).todense()).ravel())
    ek = _exp(abs(X).max(axis=0))
    nnz = np.diff(sp.indptr)
    b = 6.2 * np.ldexp(1.0, er[:, None] + ek[None, :] - 35) * nnz[:, None]
    return b + 4e-16 * (absp @ abs(X))


def test_panel8_layout_and_slices(setup):
    torch, sm, csr, Pb, P8, ct = setup
    sp = csr.to_scipy().tocsr()
    sp.sort_indices()
    n = sp.shape[0]
    colptr = P8.colptr.cpu().numpy()
    cols = P8.cols.cpu().numpy()
    qptr = P8.qptr.cpu().numpy()
    nq = (np.diff(colptr) // 16 + 3) // 4
    assert qptr[0] == 0 and np.array_equal(np.diff(qptr), nq)
    assert P8.quads == qptr[-1]
    rowmax = np.asarray(abs(sp).max(axis=1).todense()).ravel()
    er = P8.rowexp.cpu().numpy()[:n]
    assert np.array_equal(er, _exp(rowmax))
    panel = P8.dense.cpu().numpy()[:P8.quads * 20480].reshape(P8.quads, SLICES, 64, 64)
    assert abs(panel.astype(np.int32)).max() <= 64
    w = np.ldexp(1.0, -6 - 7 * np.arange(SLICES))
    for b in range(len(colptr) - 1):
        r0, r1 = 64 * b, min(n, 64 * (b + 1))
        q = panel[qptr[b]:qptr[b + 1]].astype(np.float64)       # (nq, 5, 64, 64)
        dec = np.einsum("qirc,i->rqc", q, w).reshape(64, -1)     # rows x positions
        dec = dec[:r1 - r0] * np.ldexp(1.0, er[r0:r1])[:, None]
        ncl = colptr[b + 1] - colptr[b]
        nd = np.unique(sp[r0:r1].indices).size
        want = np.zeros_like(dec)
        want[:, :nd] = sp[r0:r1][:, cols[colptr[b]:colptr[b] + nd]].toarray()
        # |v - decoded| <= 2^(e_r - 35); exact zeros elsewhere (padding included)
        tol = np.ldexp(1.0, er[r0:r1] - 35)[:, None]
        assert (abs(dec - want) <= tol).all(), b
        assert (dec[:, ncl:] == 0).all() and (dec[want == 0] == 0).all()
    # packed form: lane chunk l of subtile s = 4 q + g is rows 16 g + (l & 15),
    # columns 16 (l >> 4) .. + 15 of quad q; lmask bit l says it holds a
    # nonzero, and the nonzero chunks are stored slice-major from 80 sbase[s]
    lm = P8.lmask.cpu().numpy()[:4 * P8.quads].view(np.uint64)
    sb = P8.sbase.cpu().numpy()
    packed = P8.panel.cpu().numpy()
    # (quad, slice, group g, row li, chunk lg, 16 bytes) -> (s, slice, lane, 16)
    ch = panel.reshape(P8.quads, SLICES, 4, 16, 4, 16).transpose(0, 2, 1, 4, 3, 5)
    ch = ch.reshape(4 * P8.quads, SLICES, 64, 16)
    nzc = (ch != 0).any(axis=(1, 3))                          # (s, lane)
    bits = (lm[:, None] >> np.arange(64, dtype=np.uint64)) & np.uint64(1)
    assert np.array_equal(bits.astype(bool), nzc)
    cnt = nzc.sum(axis=1)
    assert sb[0] == 0 and np.array_equal(np.diff(sb), cnt) and sb[-1] == P8.chunks
    assert packed.size >= 80 * P8.chunks
    for s_ in range(4 * P8.quads):
        c = int(cnt[s_])
        if c == 0:
            continue
        got = packed[80 * sb[s_]:80 * (sb[s_] + c)].reshape(SLICES, c, 16)
        assert np.array_equal(got, ch[s_][:, nzc[s_]]), s_
    # some subtiles are zero, some lane chunks inside nonzero subtiles too
    assert 0 < (lm != 0).sum() < lm.size
    assert nzc[lm != 0].mean() < 1.0


def test_panel8_packed_equals_dense(setup):
    """The packed panels (zero lane chunks read from the zero quad, zero
    subtiles' MFMAs skipped) give the dense panels' product bit for bit, for
    every slice count."""
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    for k in (64, 8):
      X = torch.rand((n, k), dtype=torch.float64, device=_dev())
      for ns in (5, 4, 3):
        Y1, Y0 = torch.empty_like(X), torch.empty_like(X)
        P8.mm_f64(X, Y1, nslices=ns)
        P8.packed = False
        try:
            P8.mm_f64(X, Y0, nslices=ns)
        finally:
            P8.packed = True
        assert torch.equal(Y1, Y0), ns


@pytest.mark.parametrize("k", [64, 16, 5, 70, 33])
def test_panel8_product_within_bound(setup, k):
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    rng = np.random.default_rng(100 + k)
    X = rng.standard_normal((n, k))
    X[:, 0] = 0.0                                    # an all-zero right-hand side
    if k > 2:
        X[:, 1] *= np.exp(rng.uniform(-20, 0, n))    # wide dynamic range
    Xd = torch.tensor(X, device=_dev())
    Y = torch.empty_like(Xd)
    P8.mm_f64(Xd, Y)
    ref = sp @ X
    err = abs(Y.cpu().numpy() - ref)
    bound = _bound(sp, X)
    assert (err <= bound).all(), (err / np.maximum(bound, 1e-300)).max()
    assert (Y.cpu().numpy()[:, 0] == 0).all()
    # the bound is absolute (relative to max|p_r.| max|x_.k|); typical
    # entries of the ordinary columns are far below FP32 accuracy relative to
    # |P| |X| (column 1 spans e^-20 and is left out)
    ordinary = np.arange(k) >= 2
    scale = (abs(sp) @ abs(X))[:, ordinary]
    rel = err[:, ordinary][scale > 0] / scale[scale > 0]
    assert rel.size == 0 or np.median(rel) < 1e-9, np.median(rel)


@pytest.mark.parametrize("k", [64, 8])
@pytest.mark.parametrize("left", [False, True])
def test_panel8_residual(setup, left, k):
    """k = 8 runs the groups of 16 right-hand sides, k = 64 those of 64."""
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    rng = np.random.default_rng(7 + int(left))
    E = rng.uniform(0, 1000, (n, k))
    B = E * rng.uniform(1.0, 1.2, (n, k))
    rho = rng.uniform(0.1, 0.14, n)
    t = lambda a: torch.tensor(a, device=_dev())
    R = torch.empty((n, k), dtype=torch.float64, device=_dev())
    cm = torch.empty(k, dtype=torch.float64, device=_dev())
    if left:
        P8.mm_f64(t(B), R, E=t(E), B=t(B), rho=t(rho), left=True, colmax=cm)
        Rr = E + rho[:, None] * (sp @ B) - B
        bound = rho[:, None] * _bound(sp, B)
    else:
        P8.mm_f64(t(B), R, E=t(E), B=t(B), colmax=cm, xscale=t(rho))
        X = rho[:, None] * B
        Rr = E + sp @ X - B
        bound = _bound(sp, X)
    err = abs(R.cpu().numpy() - Rr)
    assert (err <= bound + 1e-13 * abs(E).max()).all()
    assert np.allclose(cm.cpu().numpy(), abs(R.cpu().numpy()).max(0), rtol=0, atol=0)
    assert np.allclose(cm.cpu().numpy(), abs(Rr).max(0), rtol=1e-9)
    # a scalar albedo on the right
    P8.mm_f64(t(B), R, E=t(E), B=t(B), xscale=0.12)
    Rs = E + sp @ (0.12 * B) - B
    assert (abs(R.cpu().numpy() - Rs) <= _bound(sp, 0.12 * B) + 1e-13 * abs(E).max()).all()


@pytest.mark.parametrize("left", [False, True])
def test_refined_solve_int8_residual(setup, oracle_mod, left):
    from fluxpy_amd._panel import PanelRefinedSolver
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    D = meshes.sun_directions(64, seed=3)
    E = sm.get_direct_irradiance_multi(1365.0, D).astype(np.float64)
    Ep = E[ct.perm]
    rho = np.random.default_rng(0).uniform(0.1, 0.14, n)
    S = PanelRefinedSolver(Pb, P8, n, 64, torch.tensor(rho, device=_dev()), left=left)
    B, info = S.solve(torch.tensor(Ep, device=_dev()), rtol=1e-7, inner_iters=2)
    Bd = oracle_mod.solve_radiosity_dense(sp, Ep, rho, "left" if left else "right")
    err = abs(B.cpu().numpy() - Bd).max(0) / abs(Bd).max(0)
    assert err.max() <= 1e-6, err.max()
    assert info["rounds"] <= 6
    # the same solver object reused with a scalar albedo
    S.set_rho(0.12)
    B2, _ = S.solve(torch.tensor(Ep, device=_dev()), rtol=1e-7, inner_iters=2)
    Bd2 = oracle_mod.solve_radiosity_dense(sp, Ep, np.full(n, 0.12), "left" if left else "right")
    assert (abs(B2.cpu().numpy() - Bd2).max(0) / abs(Bd2).max(0)).max() <= 1e-6


@pytest.mark.parametrize("nslices", [3, 4])
def test_panel8_fewer_slices_within_bound(setup, nslices):
    """The leading-slice products: error <= (S + 1.2) 2^(e_r + e_k - 7 S)
    nnz_r (the five-slice bound with 35 replaced by 7 S)."""
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    rng = np.random.default_rng(7 + nslices)
    X = rng.standard_normal((n, 64))
    Xd = torch.tensor(X, device=_dev())
    Y = torch.empty_like(Xd)
    P8.mm_f64(Xd, Y, nslices=nslices)
    absp = abs(sp)
    er = _exp(np.asarray(absp.max(axis=1).todense()).ravel())
    ek = _exp(abs(X).max(axis=0))
    nnz = np.diff(sp.indptr)
    bound = ((nslices + 1.2) * np.ldexp(1.0, er[:, None] + ek[None, :] - 7 * nslices)
             * nnz[:, None] + 4e-16 * (absp @ abs(X)))
    err = abs(Y.cpu().numpy() - sp @ X)
    assert (err <= bound).all(), (err / bound).max()
    # and it is not the five-slice product
    Y5 = torch.empty_like(Xd)
    P8.mm_f64(Xd, Y5)
    assert not torch.equal(Y, Y5)


@pytest.mark.parametrize("force_full", [False, True])
def test_refined_solve_update_residual(setup, oracle_mod, monkeypatch, force_full):
    """Rounds after the first form the residual as an update on four slices
    (no full recomputation needed here); with the magnitude check forced to
    fail, every such round recomputes its residual in full.  Both meet 1e-6
    against the dense FP64 solution."""
    from fluxpy_amd._panel import PanelRefinedSolver
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    D = meshes.sun_directions(64, seed=4)
    Ep = sm.get_direct_irradiance_multi(1365.0, D).astype(np.float64)[ct.perm]
    rho = np.random.default_rng(1).uniform(0.1, 0.14, n)
    if force_full:
        monkeypatch.setattr(PanelRefinedSolver, "_update_ok", lambda self, b, d: False)
    S = PanelRefinedSolver(Pb, P8, n, 64, torch.tensor(rho, device=_dev()))
    B, info = S.solve(torch.tensor(Ep, device=_dev()), rtol=1e-7, inner_iters=1,
                      max_rounds=4)
    Bd = oracle_mod.solve_radiosity_dense(sp, Ep, rho, "right")
    err = abs(B.cpu().numpy() - Bd).max(0) / abs(Bd).max(0)
    assert err.max() <= 1e-6, err.max()
    assert info["residual_history"][-1] <= 1e-7
    if force_full:
        assert info["full_residual_fallbacks"] == info["rounds"] - 1 >= 1
    else:
        assert info["full_residual_fallbacks"] == 0
    # the update residual of the last round equals a full recomputation
    # within the four-slice bound: compare against the FP64 residual
    R = S.Rs[1 if info["rounds"] % 2 == 0 else 0].cpu().numpy()
    Bn = B.cpu().numpy()
    R_exact = Ep + sp @ (rho[:, None] * Bn) - Bn
    assert abs(R - R_exact).max() <= 1e-9 * abs(Ep).max()


@pytest.mark.parametrize("residual", ["int8", "fp64"])
def test_public_refined_solve_on_panels(oracle_mod, residual):
    """solve_radiosity_refined(H, H, ...) on a compressed matrix with no
    low-rank leaves runs on the panels and meets 1e-6 in facet order."""
    import torch
    from fluxpy_amd import _hbuild
    from fluxpy_amd.compressed_form_factors import CompressedFormFactorMatrix
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    from fluxpy_amd.solve import _panel_ready, solve_radiosity_refined
    V, F = meshes.fractal_terrain(grid=33, levels=5, seed=5)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    old = _hbuild.LOWRANK
    _hbuild.LOWRANK = False
    try:
        H = CompressedFormFactorMatrix(sm, tol=1e-6, min_size=1024)
    finally:
        _hbuild.LOWRANK = old
    assert _panel_ready(H)
    n = sm.num_faces
    D = meshes.sun_directions(8, seed=9)
    E = sm.get_direct_irradiance_multi(1365.0, D).astype(np.float64)
    rho = np.random.default_rng(2).uniform(0.1, 0.14, n)
    B, info = solve_radiosity_refined(H, H, E, rho, rtol=1e-7, inner_iters=2,
                                      return_info=True, residual=residual)
    assert isinstance(B, np.ndarray) and B.shape == E.shape
    Fd = H.toarray().astype(np.float64)
    import scipy.sparse
    Bd = oracle_mod.solve_radiosity_dense(scipy.sparse.csr_matrix(Fd), E, rho, "right")
    err = abs(B - Bd).max(0) / abs(Bd).max(0)
    assert err.max() <= 1e-6, err.max()
    # device input stays on the device; a second call reuses the cached panels
    Bt = solve_radiosity_refined(H, H, torch.tensor(E, device=_dev()), rho, rtol=1e-7,
                                 inner_iters=2, residual=residual)
    assert isinstance(Bt, torch.Tensor)
    assert abs(Bt.cpu().numpy() - Bd).max() / abs(Bd).max() <= 1e-6


def test_plan_covers_blocks(setup):
    """The balanced plan tiles every block's quads exactly once, longest
    pieces first, with consistent split counts and slot ranges."""
    from fluxpy_amd._panel import PanelPlan
    torch, sm, csr, Pb, P8, ct = setup
    colptr = Pb.colptr.cpu().numpy()
    nblk = len(colptr) - 1
    nq = (np.diff(colptr) // 16 + 3) // 4
    for cap in (1, 3, 48, 10 ** 6):
        pl = PanelPlan(Pb.colptr, nblk, cap=cap)
        it = pl.items.cpu().numpy()[:4 * pl.n_items].reshape(-1, 4)
        ns = pl.nsplit.cpu().numpy()[:nblk]
        s0 = pl.slot0.cpu().numpy()[:nblk]
        size = it[:, 2] - it[:, 1]
        assert size.max() <= max(cap, 1)
        # each eighth (one XCD's run) is sorted longest first and holds a
        # contiguous range of blocks
        q, r = divmod(pl.n_items, 8)
        lo = 0
        for x in range(8):
            hi = lo + q + (1 if x < r else 0)
            assert (np.diff(size[lo:hi]) <= 0).all()
            if hi > lo:
                blk = np.unique(it[lo:hi, 0])
                assert np.array_equal(blk, np.arange(blk.min(), blk.max() + 1))
            lo = hi
        assert np.array_equal(ns, np.maximum(1, -(-nq // cap)))
        assert pl.n_items == ns.sum() and pl.n_slots == ns[ns > 1].sum()
        assert pl.n_split == (ns > 1).sum()
        if pl.n_split:
            assert np.array_equal(pl.split_blocks.cpu().numpy()[:pl.n_split],
                                  np.flatnonzero(ns > 1))
        for b in range(nblk):
            mine = it[it[:, 0] == b]
            mine = mine[np.argsort(mine[:, 3])]
            assert np.array_equal(mine[:, 3], np.arange(ns[b]))
            assert mine[0, 1] == 0 and mine[-1, 2] == nq[b]
            assert (mine[1:, 1] == mine[:-1, 2]).all()
            assert (s0[b] == -1) == (ns[b] == 1)
        used = np.concatenate([s0[b] + np.arange(ns[b]) for b in range(nblk) if ns[b] > 1] or
                              [np.zeros(0, int)])
        assert np.array_equal(np.sort(used), np.arange(pl.n_slots))


@pytest.mark.parametrize("k", [64, 70, 8])
def test_planned_products_match_unplanned(setup, k):
    """Split blocks (cap of 2 quads, so nearly every block is cut) give the
    same products, residuals and residual column maxima as one workgroup per
    block, and the same bits on a second call."""
    from fluxpy_amd._panel import PanelPlan
    torch, sm, csr, Pb, P8, ct = setup
    n = csr.shape[0]
    rng = np.random.default_rng(50 + k)
    X = torch.tensor(rng.standard_normal((n, k)), device=_dev())
    E = torch.tensor(rng.standard_normal((n, k)), device=_dev())
    Xb = X.to(torch.float32).to(torch.bfloat16)
    nblk = (n + 63) // 64
    tiny = PanelPlan(Pb.colptr, nblk, cap=2)
    assert tiny.n_slots > 0
    base = torch.tensor(rng.uniform(0, 1, (n, k)), dtype=torch.float32, device=_dev())
    rho = torch.tensor(rng.uniform(0.1, 0.5, n), dtype=torch.float32, device=_dev())
    outs, sweeps = {}, {}
    for name, plan in (("none", False), ("tiny", tiny), ("tiny2", tiny)):
        Pb.plan = P8.plan = plan
        y32 = torch.empty((n, k), dtype=torch.float32, device=_dev())
        Pb.mm_bf16(Xb, y32)
        # a Jacobi sweep: base, albedo on the right, next iterate, change maxima
        s32 = torch.empty((n, k), dtype=torch.float32, device=_dev())
        xn = torch.empty((n, k), dtype=torch.bfloat16, device=_dev())
        scm = torch.empty(k, dtype=torch.float64, device=_dev())
        Pb.mm_bf16(Xb, s32, base=base, rho=rho, xnext=xn, prev=base, colmax=scm)
        y64 = torch.empty((n, k), dtype=torch.float64, device=_dev())
        P8.mm_f64(X, y64)
        # residual with column maxima (the split blocks' epilogue)
        r64 = torch.empty((n, k), dtype=torch.float64, device=_dev())
        cm = torch.empty(k, dtype=torch.float64, device=_dev())
        P8.mm_f64(X, r64, E=E, B=X, colmax=cm, xscale=0.12)
        torch.cuda.synchronize()
        outs[name] = (y32.cpu().numpy(), y64.cpu().numpy(), r64.cpu().numpy(),
                      cm.cpu().numpy())
        sweeps[name] = (s32.cpu().numpy(), xn.float().cpu().numpy(), scm.cpu().numpy())
    Pb.plan = P8.plan = Pb.__dict__.get("_default_plan", tiny)
    a32, a64, ar, acm = outs["none"]
    b32, b64, br, bcm = outs["tiny"]
    c32, c64, cr, ccm = outs["tiny2"]
    assert np.array_equal(b32, c32) and np.array_equal(b64, c64)      # deterministic
    assert np.array_equal(br, cr) and np.array_equal(bcm, ccm)
    assert np.allclose(b32, a32, rtol=1e-5, atol=1e-5 * abs(a32).max())
    assert np.allclose(b64, a64, rtol=1e-14, atol=1e-14 * abs(a64).max())
    assert np.allclose(br, ar, rtol=0, atol=1e-13 * abs(ar).max())
    assert np.array_equal(bcm, abs(br).max(axis=0))
    assert np.allclose(bcm, acm, rtol=1e-12)
    (sa, xa, ca), (sb, xb, cb), (sc, xc, cc) = sweeps["none"], sweeps["tiny"], sweeps["tiny2"]
    assert np.array_equal(sb, sc) and np.array_equal(xb, xc) and np.array_equal(cb, cc)
    assert np.allclose(sb, sa, rtol=1e-5, atol=1e-5 * abs(sa).max())
    assert np.isfinite(xb).all()
    assert np.allclose(xb, rho.cpu().numpy()[:, None] * sb, rtol=1e-2, atol=1e-6)
    basen = base.cpu().numpy()
    assert np.array_equal(cb, abs(sb.astype(np.float64) - basen).max(axis=0))
    assert np.allclose(cb, ca, rtol=1e-4)
