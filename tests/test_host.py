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
    # tall block with zero rows: U rows of zero block rows are zero
    assert _svd_nbytes(100, 40, 90, 40, 5, 8) == (5 * 90 * 12 + 101 * 4) + 5 * 8 + \
        (5 * 40 * 12 + 6 * 4)


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
