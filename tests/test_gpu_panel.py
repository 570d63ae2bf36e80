"""GPU tests of the near-field panel format and the batched refined solve:
panel products equal the CSR products (FP64 exactly up to summation order,
BF16 within BF16 rounding), and the refined solve meets 1e-6 against the
FP64 solution."""
import numpy as np
import pytest

from fluxpy_amd import meshes

pytestmark = pytest.mark.gpu


def _dev():
    from fluxpy_amd import _lib
    return _lib.device()


@pytest.fixture(scope="module")
def setup():
    import torch
    torch.cuda.set_device(0)
    from fluxpy_amd.compressed_form_factors import _ClusterTree
    from fluxpy_amd import _kernels
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    V, F = meshes.fractal_terrain(grid=49, levels=6, seed=11)
    sm = CgalTrimeshShapeModel(V.astype(np.float32), F)
    ct = _ClusterTree(np.asarray(sm.P, dtype=np.float64), 2)
    csr = _kernels.ff_csr(sm, ct.perm, ct.perm)
    return torch, sm, csr


def test_panel_structure(setup):
    from fluxpy_amd._panel import PanelMatrix
    torch, sm, csr = setup
    P = PanelMatrix(csr, "fp32")
    sp = csr.to_scipy()
    colptr = P.colptr.cpu().numpy()
    cols = P.cols.cpu().numpy()
    assert (np.diff(colptr) % 16 == 0).all()
    order = P.order.cpu().numpy()
    for b in range(len(colptr) - 1):
        blk = sp[64 * b:64 * (b + 1)].tocsr()
        want = np.unique(blk.indices)
        got = cols[colptr[b]:colptr[b + 1]]
        # columns ordered by (16-row occupancy mask, index), padding with 0
        mask = np.zeros(want.size, dtype=np.int64)
        for r in range(blk.shape[0]):
            mask[np.searchsorted(want, blk.indices[blk.indptr[r]:blk.indptr[r + 1]])] |= 1 << (r // 16)
        key = np.lexsort((want, mask))
        assert np.array_equal(got[:want.size], want[key])
        assert (got[want.size:] == 0).all()
        # order maps the r-th smallest column to its position
        assert np.array_equal(got[order[colptr[b]:colptr[b] + want.size]], want)
    assert 0.0 < P.fill <= 1.0


@pytest.mark.parametrize("k", [64, 16, 5, 70])
def test_panel_products_match_csr(setup, k):
    from fluxpy_amd._panel import PanelMatrix
    torch, sm, csr = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    rng = np.random.default_rng(k)
    X = rng.standard_normal((n, k))
    P32 = PanelMatrix(csr, "fp32")
    Xd = torch.tensor(X, device=_dev())
    Y = torch.empty_like(Xd)
    P32.mm_f64(Xd, Y)
    ref = sp @ X
    assert np.allclose(Y.cpu().numpy(), ref, rtol=1e-12, atol=1e-15 * abs(ref).max())
    # residual epilogue
    E = rng.standard_normal((n, k))
    B = rng.standard_normal((n, k))
    rho = rng.uniform(0.1, 0.14, n)
    cm = torch.empty(k, dtype=torch.float64, device=_dev())
    R = torch.empty_like(Xd)
    P32.mm_f64(torch.tensor(rho[:, None] * B, device=_dev()), R, E=torch.tensor(E, device=_dev()),
               B=torch.tensor(B, device=_dev()), rho=torch.tensor(rho, device=_dev()),
               colmax=cm)
    Rr = E + sp @ (rho[:, None] * B) - B
    assert np.allclose(R.cpu().numpy(), Rr, rtol=1e-12, atol=1e-13)
    assert np.allclose(cm.cpu().numpy(), abs(Rr).max(0), rtol=1e-12)
    # BF16 product against the CSR of the BF16-rounded values
    Pb = P32.sibling(csr, "bf16")
    Xb = torch.tensor(X, dtype=torch.float32, device=_dev()).to(torch.bfloat16)
    out = torch.empty((n, k), dtype=torch.float32, device=_dev())
    Pb.mm_bf16(Xb, out)
    spb = sp.copy()
    spb.data = torch.tensor(spb.data).to(torch.bfloat16).to(torch.float64).numpy()
    refb = spb @ Xb.to(torch.float64).cpu().numpy()
    assert np.allclose(out.cpu().numpy(), refb, rtol=2e-5, atol=2e-6 * abs(refb).max())


def test_bf16_packed_panels(setup):
    """Packed BF16 panels: every nonzero 8-byte lane chunk of the dense
    panels sits at its packed position and every other chunk is zero; the
    products read from the packed panels equal the dense panels' bit for
    bit, with and without a split-block plan, at both stage widths, for
    right-hand-side counts on the vector and the scalar paths."""
    from fluxpy_amd._panel import PanelMatrix, PanelPlan
    torch, sm, csr = setup
    n = csr.shape[0]
    P = PanelMatrix(csr, "bf16", keep_dense=True)
    assert P.packed
    dense = P.dense.view(torch.int64).cpu().numpy()           # 8-byte units
    packed = P.panel.view(torch.int64).cpu().numpy()
    lmask = P.lmask.cpu().numpy().view(np.uint64)
    sbase = P.sbase.cpu().numpy()
    nsub = P.total_cols // 4
    s = np.arange(nsub)[:, None]
    lane = np.arange(64)[None, :]
    unit = (s >> 2) * 256 + (s & 3) * 64 + (lane & 15) * 4 + (lane >> 4)
    assert np.array_equal(np.sort(unit.ravel()), np.arange(dense.size))
    chunks = dense[unit]                                        # (subtile, lane)
    bits = ((lmask[:nsub, None] >> lane.astype(np.uint64)) & np.uint64(1)).astype(bool)
    assert np.array_equal(bits, chunks != 0)
    below = np.cumsum(bits, axis=1) - bits
    pos = sbase[:nsub, None] + below
    assert P.lanes == bits.sum() and np.array_equal(np.sort(pos[bits]), np.arange(P.lanes))
    assert np.array_equal(packed[pos[bits]], chunks[bits])
    assert 0 < P.nonzero_subtiles < nsub and P.lanes < dense.size
    rng = np.random.default_rng(3)
    tiny = PanelPlan(P.colptr, (n + 63) // 64, cap=2)
    assert tiny.n_slots > 0
    for k in (64, 70, 16, 5):
        Xb = torch.tensor(rng.standard_normal((n, k)), dtype=torch.float32,
                          device=_dev()).to(torch.bfloat16)
        base = torch.tensor(rng.uniform(0, 1, (n, k)), dtype=torch.float32, device=_dev())
        rho = torch.tensor(rng.uniform(0.1, 0.5, n), dtype=torch.float32, device=_dev())
        for plan in (None, tiny):
            for stage in (4, 8):
                P.plan, P.stage = plan, stage
                res = {}
                for pk in (True, False):
                    P.packed = pk
                    y = torch.empty((n, k), dtype=torch.float32, device=_dev())
                    P.mm_bf16(Xb, y)
                    s32 = torch.empty_like(y)
                    xn = torch.empty_like(Xb)
                    cm = torch.empty(k, dtype=torch.float64, device=_dev())
                    P.mm_bf16(Xb, s32, base=base, rho=rho, xnext=xn, prev=base, colmax=cm)
                    res[pk] = (y, s32, xn, cm)
                for a, b in zip(res[True], res[False]):
                    assert torch.equal(a, b), (k, plan is None, stage)
    P.packed = True


@pytest.mark.parametrize("left", [False, True])
def test_refined_panel_solve(setup, oracle_mod, left):
    from fluxpy_amd._panel import PanelMatrix, PanelRefinedSolver
    torch, sm, csr = setup
    n = csr.shape[0]
    sp = csr.to_scipy().astype(np.float64)
    D = meshes.sun_directions(64, seed=3)
    E = sm.get_direct_irradiance_multi(1365.0, D).astype(np.float64)
    # rows of E in cluster order: the CSR is already permuted, so permute E
    from fluxpy_amd.compressed_form_factors import _ClusterTree
    ct = _ClusterTree(np.asarray(sm.P, dtype=np.float64), 2)
    Ep = E[ct.perm]
    rho = np.random.default_rng(0).uniform(0.1, 0.14, n)
    P32 = PanelMatrix(csr, "fp32")
    Pb = P32.sibling(csr, "bf16")
    S = PanelRefinedSolver(Pb, P32, n, 64, torch.tensor(rho, device=_dev()), left=left)
    B, info = S.solve(torch.tensor(Ep, device=_dev()), rtol=1e-7)
    Bd = oracle_mod.solve_radiosity_dense(sp, Ep, rho, "left" if left else "right")
    err = abs(B.cpu().numpy() - Bd).max(0) / abs(Bd).max(0)
    assert err.max() <= 1e-6, err.max()
    assert info["rounds"] <= 6


@pytest.mark.parametrize("n,k,scale", [(1000, 64, 1), (5003, 70, 1), (37, 5, 0), (0, 8, 1)])
def test_refine_prepare_with_column_maxima(n, k, scale):
    """flux_refine_prepare: r32 = fp32(r), xb = bf16(s r32) and, with colmax,
    max|r| per column; the same r32 and xb as the pass without maxima."""
    import torch
    from fluxpy_amd import _lib
    print(f"case: refine_prepare n={n} k={k} scale={scale}", flush=True)
    rng = np.random.default_rng(n + k)
    r = torch.tensor(rng.standard_normal((n, k)) * rng.uniform(0.1, 100.0, k),
                     device=_dev())
    rho = torch.tensor(rng.uniform(0.1, 0.5, n), dtype=torch.float32, device=_dev())
    st = _lib.stream()
    outs = []
    for with_max in (False, True):
        r32 = torch.full((n, k), -1.0, dtype=torch.float32, device=_dev())
        xb = torch.zeros((n, k), dtype=torch.bfloat16, device=_dev())
        cm = torch.full((k,), -7.0, dtype=torch.float64, device=_dev())
        work = torch.empty(max(int(_lib.load().flux_refine_prepare_work_bytes(k)), 16),
                           dtype=torch.uint8, device=_dev())
        _lib.call("flux_refine_prepare", n, k, r.data_ptr(), rho.data_ptr(), 1.0, scale,
                  r32.data_ptr(), xb.data_ptr(), cm.data_ptr() if with_max else None,
                  work.data_ptr() if with_max else None, st)
        torch.cuda.synchronize()
        outs.append((r32.cpu(), xb.float().cpu(), cm.cpu()))
    (a32, axb, _), (b32, bxb, bcm) = outs
    ref32 = r.float().cpu()
    s = rho.cpu()[:, None] if scale else 1.0
    assert torch.equal(a32, ref32) and torch.equal(b32, ref32)
    assert torch.equal(axb, (s * ref32).to(torch.bfloat16).float())
    assert torch.equal(bxb, axb)
    want = r.abs().amax(dim=0).cpu() if n else torch.zeros(k, dtype=torch.float64)
    assert torch.equal(bcm, want)


def test_timing_events(setup):
    """The launch timer's events (no cache writeback on record) time a BF16
    product: a positive duration no longer than the host-side wall time of
    the synchronised launch, and one (kind, start, end) record per launch."""
    import time
    from fluxpy_amd._panel import PanelMatrix
    torch, sm, csr = setup
    n = csr.shape[0]
    P = PanelMatrix(csr, "bf16")
    Xb = torch.ones((n, 64), dtype=torch.bfloat16, device=_dev())
    out = torch.empty((n, 64), dtype=torch.float32, device=_dev())
    P.mm_bf16(Xb, out)
    torch.cuda.synchronize()
    P.timing = []
    t0 = time.perf_counter()
    P.mm_bf16(Xb, out)
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t0) * 1e3
    P.timing, rec = None, P.timing
    assert len(rec) == 1 and rec[0][0] == "bf16"
    ms = rec[0][1].elapsed_time(rec[0][2])
    assert 0.0 < ms <= wall_ms
