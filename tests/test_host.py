"""Host-side logic of the package (CPU only): partition, cluster tree,
byte rules, mesh generators, API argument checks."""
import numpy as np
import pytest

from fluxpy_amd import meshes


def test_quadrant_octant_order_match_oracle(rng):
    from fluxpy_amd.quadtree import get_quadrant_order
    from fluxpy_amd.octree import get_octant_order
    from oracle.hmatrix import quadrant_order, octant_order
    for dt in (np.float32, np.float64):
        X = rng.random((500, 3)).astype(dt)
        X[:7] = X[7]
        for a, b in zip(get_quadrant_order(X[:, :2]), quadrant_order(X[:, :2])):
            assert np.array_equal(a, b)
        for a, b in zip(get_octant_order(X), octant_order(X)):
            assert np.array_equal(a, b)
    # bbox argument and degenerate inputs
    X = np.array([[0.0, 0.0], [1.0, 1.0], [0.5, 0.5]])
    parts = get_quadrant_order(X, bbox=((0, 1), (0, 1)))
    assert [p.tolist() for p in parts] == [[0, 2], [], [], [1]]
    parts = get_quadrant_order(np.array([[0.3, 0.3]]))
    assert [p.tolist() for p in parts] == [[0], [], [], []]


@pytest.mark.parametrize("dims", [2, 3])
def test_cluster_tree_is_consistent(dims):
    from fluxpy_amd.compressed_form_factors import _ClusterTree
    from oracle.hmatrix import quadrant_order, octant_order
    V, F = meshes.spherical_crater(npts=18, seed=3)
    P = V[F].mean(axis=1)
    ct = _ClusterTree(P, dims)
    n = P.shape[0]
    assert np.array_equal(np.sort(ct.perm), np.arange(n))
    order = quadrant_order if dims == 2 else octant_order

    def check(c):
        ids = ct.perm[c.start:c.end]
        assert np.array_equal(np.sort(ids), c.idx)
        assert np.all(np.diff(c.idx) > 0)
        if c.children is not None:
            ref = order(P[c.idx, :dims])
            for ch, part, rp in zip(c.children, c.parts, ref):
                assert np.array_equal(part, rp)
                assert np.array_equal(ch.idx, c.idx[rp])
            assert c.children[0].start == c.start and c.children[-1].end == c.end
            for a, b in zip(c.children, c.children[1:]):
                assert a.end == b.start
            for ch in c.children:
                check(ch)
    check(ct.root)


def test_byte_rules_match_reference_formulas():
    from fluxpy_amd.compressed_form_factors import _csr_nbytes, _svd_nbytes
    import scipy.sparse
    A = scipy.sparse.random(50, 70, density=0.2, format="csr", dtype=np.float32,
                            random_state=1)
    assert _csr_nbytes(50, A.nnz, 4) == A.data.nbytes + A.indices.nbytes + A.indptr.nbytes
    # 90 nonzero rows, 33 nonzero columns: U holds 5 x 90 values, Vt 5 x 33
    assert _svd_nbytes(100, 40, 90, 33, 5, 8) == (5 * 90 * 12 + 101 * 4) + 5 * 8 + \
        (5 * 33 * 12 + 6 * 4)


def test_fractal_mesh_counts():
    V, F = meshes.fractal_terrain(grid=145)
    assert F.shape == (41472, 3)
    C = np.cross(V[F[:, 1]] - V[F[:, 0]], V[F[:, 2]] - V[F[:, 0]])
    assert (C[:, 2] > 0).all()
    assert np.isclose(np.ptp(V[:, 2]), 0.2)
    V2, F2 = meshes.fractal_terrain(grid=145)
    assert np.array_equal(V, V2)


def test_direct_instantiation_and_argument_errors():
    from fluxpy_amd.shape import TrimeshShapeModel
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0.0]])
    with pytest.raises(RuntimeError):
        TrimeshShapeModel(V, np.array([[0, 1, 2]]))
    from fluxpy_amd.compressed_form_factors import CompressedKernelMatrix
    with pytest.raises(ValueError):
        CompressedKernelMatrix(None, 1.5)
    with pytest.raises(ValueError):
        CompressedKernelMatrix(None, 0.5, "middle")


# --------------------------------------------------------------------------
# device pointers handed to a launch
# --------------------------------------------------------------------------

def _product_sources():
    import glob
    import os
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return sorted(glob.glob(os.path.join(root, "fluxpy_amd", "*.py"))) + \
        [os.path.join(root, "bench.py")]


def test_argument_temporary_is_freed_before_the_call():
    """A tensor created inside an argument list and reduced to data_ptr()
    is released before the callee runs: its memory can then be handed to the
    next allocation while the pointer is still in use."""
    import weakref
    import torch
    freed = []

    def make():
        t = torch.empty(64)
        weakref.finalize(t, freed.append, True)
        return t

    def callee(p0, p1):
        return len(freed)

    # (outside the assert statements: pytest's assertion rewriting keeps
    # intermediate values alive)
    n_inline = callee(make().data_ptr(), make().data_ptr())
    keep = make()
    freed.clear()
    n_named = callee(keep.data_ptr(), keep.data_ptr())
    assert n_inline == 2
    assert n_named == 0


def test_launch_arguments_hold_no_device_temporaries():
    """No product source takes data_ptr() (or _lib.ptr()) of a call result:
    every device array passed to a launch is bound to a name first."""
    import ast
    bad = []
    for path in _product_sources():
        tree = ast.parse(open(path).read(), path)
        for node in ast.walk(tree):
            if not isinstance(node, ast.Call) or not isinstance(node.func, ast.Attribute):
                continue
            if node.func.attr == "data_ptr" and isinstance(node.func.value, ast.Call):
                bad.append(f"{path}:{node.lineno}")
            if node.func.attr == "ptr" and node.args and isinstance(node.args[0], ast.Call):
                bad.append(f"{path}:{node.lineno}")
    assert not bad, f"device temporaries passed by pointer: {bad}"


def test_sketch_descriptor_checks():
    """The host-side checks of the batched sketch launches accept in-range
    descriptors and reject every out-of-range one before a launch."""
    from fluxpy_amd import _hbuild, _lib
    l, r, m = np.array([8, 5]), np.array([3, 2]), np.array([10, 7])
    offC = np.array([0, 24])
    offY = np.array([0, 80])
    z = np.zeros(2, dtype=np.int64)
    ok = np.stack([m, r, l, offY, l, z, offC, r, z, [0, 30], r, z], 1)
    _hbuild._check_gemm(ok, 80 + 35, 24 + 10, 30 + 14)
    with pytest.raises(_lib.FluxError):
        _hbuild._check_gemm(ok, 80 + 34, 24 + 10, 30 + 14)     # Y one short
    with pytest.raises(_lib.FluxError):
        _hbuild._check_gemm(ok, 80 + 35, 24 + 10, 30 + 13)     # C one short
    bad = ok.copy()
    bad[1, 7] = 1                                              # ldb below N
    with pytest.raises(_lib.FluxError):
        _hbuild._check_gemm(bad, 10 ** 6, 10 ** 6, 10 ** 6)
    # C^T Y^T with both operands stored transposed (the Vt product's shape)
    tr = np.stack([r, m, l, offC, r, z + 1, offY, l, z + 1, [0, 30], m, z], 1)
    _hbuild._check_gemm(tr, 24 + 10, 80 + 35, 30 + 14)
    with pytest.raises(_lib.FluxError):
        _hbuild._check_gemm(tr, 24 + 9, 80 + 35, 30 + 14)
    _hbuild._check_extent("x", [0, 5], [5, 3], 8)
    with pytest.raises(_lib.FluxError):
        _hbuild._check_extent("x", [0, 5], [5, 4], 8)
    with pytest.raises(_lib.FluxError):
        _hbuild._check_extent("x", [-1], [1], 8)


# --------------------------------------------------------------------------
# the factor byte model of the rank rule
# --------------------------------------------------------------------------

def _svds_gram_side(A, k):
    """The svds algorithm of the SciPy the reference pinned (scipy>=1.5.4,
    requirements.txt:34; scipy/sparse/linalg/eigen/arpack/arpack.py, svds):
    ARPACK eigenpairs of the Gram matrix of the shorter side, then the other
    side's vectors as A v / s (tall) or A^T u / s (wide)."""
    import scipy.sparse.linalg as sla
    m, n = A.shape
    if m >= n:
        G = sla.LinearOperator((n, n), matvec=lambda x: A.T @ (A @ x), dtype=np.float64)
        w, v = sla.eigsh(G, k=k, v0=np.ones(n))
        s = np.sqrt(w)
        return (A @ v) / s, s, v.T
    G = sla.LinearOperator((m, m), matvec=lambda x: A @ (A.T @ x), dtype=np.float64)
    w, u = sla.eigsh(G, k=k, v0=np.ones(m))
    s = np.sqrt(w)
    return u, s, ((A.T @ u) / s).T


@pytest.mark.parametrize("shape", [(120, 60), (60, 120), (90, 90), (200, 41)])
def test_factor_nnz_model_matches_svds(shape):
    """oracle/hmatrix.py factor_nnz (and the product's _factor_nnz) against
    the zero pattern of real ARPACK factors of blocks with zero rows and
    columns; a swapped tall / wide rule fails."""
    import scipy.sparse
    from oracle import hmatrix as OH
    from fluxpy_amd.compressed_form_factors import _factor_nnz
    rng = np.random.default_rng(sum(shape))
    m, n = shape
    A = rng.standard_normal(shape) * (rng.random(shape) < 0.4)
    A[rng.choice(m, m // 5, replace=False)] = 0
    A[:, rng.choice(n, n // 4, replace=False)] = 0
    A = scipy.sparse.csr_matrix(A)
    nzr = int((np.diff(A.indptr) > 0).sum())
    nzc = int(np.unique(A.indices).size)
    for k in (4, 12):
        U, s, Vt = _svds_gram_side(A, k)
        got = (scipy.sparse.csr_matrix(U).nnz, scipy.sparse.csr_matrix(Vt).nnz)
        assert OH.factor_nnz(A, k) == got
        assert _factor_nnz(m, n, nzr, nzc, k) == got
        # planted mistakes: zero lines counted on the longer side only, or
        # on neither side
        longer_only = (k * nzr, k * n) if m >= n else (k * m, k * nzc)
        assert longer_only != got and (k * m, k * n) != got


def test_current_svds_zero_pattern_is_documented():
    """The svds of the SciPy installed here (1.15) orthogonalises the Gram
    eigenvectors and takes a dense SVD of A V: the factor of the shorter
    side keeps the exact zero lines, the other one is dense up to LAPACK
    rounding there.  The byte model follows the pinned SciPy, which is a
    lower bound here; this records the difference."""
    import scipy.sparse
    import scipy.sparse.linalg as sla
    rng = np.random.default_rng(7)
    for m, n in ((120, 60), (60, 120)):
        A = rng.standard_normal((m, n)) * (rng.random((m, n)) < 0.4)
        A[rng.choice(m, m // 5, replace=False)] = 0
        A[:, rng.choice(n, n // 4, replace=False)] = 0
        A = scipy.sparse.csr_matrix(A)
        nzr = int((np.diff(A.indptr) > 0).sum())
        nzc = int(np.unique(A.indices).size)
        U, s, Vt = sla.svds(A, 6, random_state=1)
        nu, nv = scipy.sparse.csr_matrix(U).nnz, scipy.sparse.csr_matrix(Vt).nnz
        # the shorter side's factor keeps the exact zero lines; the longer
        # side's holds the model's count plus LAPACK rounding on zero lines
        if m >= n:
            assert nv == 6 * nzc and 6 * nzr <= nu <= 6 * m
        else:
            assert nu == 6 * nzr and 6 * nzc <= nv <= 6 * n


def test_far_tile_lists_match_brute_force():
    """The far-field expansion visits, per output tile of rows, every block
    whose output range meets the tile, in block order (the summation order
    of each output element); ranges outside [0, n) are refused."""
    from fluxpy_amd import _hbuild, _lib
    rng = np.random.default_rng(5)
    for n, tl in ((1, 16), (15, 16), (17, 16), (1000, 16), (127, 128), (129, 128),
                  (3000, 128)):
        nb = 40
        lo = rng.integers(0, n, nb)
        hi = np.minimum(n, lo + rng.integers(1, 300, nb))
        tptr, tblk = _hbuild.tile_lists(lo, hi, n, tl)
        assert len(tptr) == (n + tl - 1) // tl + 1 and tptr[-1] == len(tblk)
        for t in range(len(tptr) - 1):
            want = [b for b in range(nb) if lo[b] < tl * t + tl and hi[b] > tl * t]
            assert tblk[tptr[t]:tptr[t + 1]].tolist() == want
    with pytest.raises(_lib.FluxError):
        _hbuild.tile_lists([0], [17], 16)
    with pytest.raises(_lib.FluxError):
        _hbuild.tile_lists([3], [3], 16)


def test_far_plan_covers_every_list_once():
    """The FP64 far-field work plan: per tile, the runs tile its block list
    in order, none exceeds the step budget unless it is a single block, a
    tile with one run adds into Y and the slots of split tiles are
    consecutive and listed for the ordered reduction."""
    from fluxpy_amd import _hbuild
    rng = np.random.default_rng(8)
    n = 3000
    lo = rng.integers(0, n - 1, 200)
    hi = np.minimum(n, lo + rng.integers(1, 700, 200))
    ranks = rng.integers(1, 200, 200)
    tptr, tblk = _hbuild.tile_lists(lo, hi, n)
    for ms in (1, 4, 8, 1000):
        items, red, nslots = _hbuild.far_plan(tptr, tblk, ranks, max_steps=ms)
        steps = (ranks + 15) // 16
        redmap = {int(r[0]): (int(r[1]), int(r[2])) for r in red}
        slots = []
        for t in range(len(tptr) - 1):
            mine = items[items[:, 0] == t]
            if tptr[t] == tptr[t + 1]:
                assert len(mine) == 0
                continue
            assert mine[0, 1] == tptr[t] and mine[-1, 2] == tptr[t + 1]
            assert (mine[1:, 1] == mine[:-1, 2]).all()
            for _, q0, q1, _ in mine:
                assert q1 > q0 and (q1 - q0 == 1 or steps[tblk[q0:q1]].sum() <= ms)
            if len(mine) == 1:
                assert mine[0, 3] == -1 and t not in redmap
            else:
                s0, ns = redmap[t]
                assert ns == len(mine) and mine[:, 3].tolist() == list(range(s0, s0 + ns))
                slots += mine[:, 3].tolist()
        assert sorted(slots) == list(range(nslots))


def test_morton_order_is_z_order():
    """morton_order sorts points by the bit interleaving of their quantised
    coordinates (x in the even bits): checked against a bit-by-bit key on
    random points, on a 4 x 4 grid (the Z pattern) and with ties."""
    from fluxpy_amd.streamed import morton_order
    rng = np.random.default_rng(3)
    P = rng.uniform(-2, 5, (500, 2))
    perm = morton_order(P)
    assert sorted(perm.tolist()) == list(range(500))
    lo, span = P.min(0), P.max(0) - P.min(0)
    q = np.minimum(((P - lo) / span * (1 << 21)).astype(np.int64), (1 << 21) - 1)

    def key(x, y):
        k = 0
        for b in range(21):
            k |= ((int(x) >> b) & 1) << (2 * b) | ((int(y) >> b) & 1) << (2 * b + 1)
        return k

    keys = np.array([key(x, y) for x, y in q])
    assert (np.diff(keys[perm]) >= 0).all()
    g = np.array([[x, y] for y in range(4) for x in range(4)], dtype=float)
    order = morton_order(g)
    assert [tuple(g[i].astype(int)) for i in order[:4]] == [(0, 0), (1, 0), (0, 1), (1, 1)]
    assert morton_order(np.zeros((5, 2))).tolist() == [0, 1, 2, 3, 4]
