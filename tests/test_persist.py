"""Loading saved matrices: fluxpy's FF.bin layout (compressed_form_factors.py
:35-803, shape.py:55, 114-115 of fluxpy) through the whitelisting loader.

fluxpy itself cannot be imported here, so its files are written by stand-in
classes registered under fluxpy's module and class names, with the
attributes fluxpy's constructors set (read from its source): the pickle
bytes then name exactly what a file of fluxpy's names.  The stand-ins are
removed before loading, so the loader cannot reach them."""
import contextlib
import pickle
import sys
import types

import numpy as np
import pytest
import scipy.sparse

from fluxpy_amd import persist


@contextlib.contextmanager
def fluxpy_names():
    """Stand-in modules flux, flux.shape, flux.compressed_form_factors."""
    flux = types.ModuleType("flux")
    shape = types.ModuleType("flux.shape")
    cff = types.ModuleType("flux.compressed_form_factors")

    class _Shape:
        def __init__(self, V, F, N=None, P=None, A=None):
            self.V, self.F, self.N, self.P, self.A = V, F, N, P, A

        def __reduce__(self):  # shape.py:114-115
            return (self.__class__, (self.V, self.F, self.N, self.P, self.A))

    for name in ("CgalTrimeshShapeModel", "EmbreeTrimeshShapeModel"):
        cls = type(name, (_Shape,), {"__module__": "flux.shape", "__qualname__": name})
        setattr(shape, name, cls)
    for name in sorted(persist._REF_BLOCKS):
        cls = type(name, (), {"__module__": "flux.compressed_form_factors", "__qualname__": name})
        setattr(cff, name, cls)
    saved = {k: sys.modules.get(k) for k in ("flux", "flux.shape", "flux.compressed_form_factors")}
    sys.modules.update({"flux": flux, "flux.shape": shape, "flux.compressed_form_factors": cff})
    try:
        yield shape, cff
    finally:
        for k, v in saved.items():
            if v is None:
                sys.modules.pop(k, None)
            else:
                sys.modules[k] = v


def _obj(cls, **attrs):
    o = cls.__new__(cls)
    o.__dict__.update(attrs)
    return o


def fluxpy_file(shape_cls_name, V, F, N, P, A, kinds, tol=1e-5, min_size=16384, octree=False,
                protocol=4):
    """Bytes of a fluxpy CompressedFormFactorMatrix whose root block has the
    given 2-D array of leaf kinds ('dense', 'csr', 'svd', 'zero', 'null',
    or a nested array for a child tree block)."""
    with fluxpy_names() as (shape, cff):
        sm = getattr(shape, shape_cls_name)(V, F, N, P, A)
        mat = _obj(cff.CompressedFormFactorMatrix, shape_model=sm, _tol=tol,
                   _min_size=min_size, _max_depth=None, _force_max_depth=False)

        def block(kind, root):
            if isinstance(kind, np.ndarray):
                tree = cff.FormFactorOctreeBlock if octree else cff.FormFactorQuadtreeBlock
                kids = np.empty(kind.shape, dtype=object)
                for idx, k in np.ndenumerate(kind):
                    kids[idx] = block(k, root)
                return _obj(tree, _root=root, shape=(4, 4), _row_block_inds=[np.arange(2)] * 2,
                            _col_block_inds=[np.arange(2)] * 2, _blocks=kids)
            if kind == "dense":
                return _obj(cff.FormFactorDenseBlock, _root=root, shape=(2, 2), _mat=np.eye(2))
            if kind == "csr":
                return _obj(cff.FormFactorCsrBlock, _root=root, shape=(2, 2),
                            _spmat=scipy.sparse.csr_matrix(np.eye(2)))
            if kind == "svd":
                return _obj(cff.FormFactorSvdBlock, _root=root, shape=(2, 2), _k=1,
                            _s=np.ones(1), _u=np.ones((2, 1)),
                            _vt=scipy.sparse.csr_matrix(np.ones((1, 2))))
            cls = {"zero": cff.FormFactorZeroBlock, "null": cff.FormFactorNullBlock}[kind]
            return _obj(cls, _root=root, shape=(2, 2))

        mat._root = block(np.asarray(kinds, dtype=object), mat)
        return pickle.dumps(mat, protocol=protocol)


def _tiny():
    V = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [1, 1, 0.1]], dtype=np.float64)
    F = np.array([[0, 1, 2], [1, 3, 2]])
    N = np.tile([0.0, 0.0, 1.0], (2, 1))
    P = V[F].mean(1)
    A = np.full(2, 0.5)
    return V, F, N, P, A


@pytest.mark.parametrize("protocol", [0, 2, 4, 5])
def test_fluxpy_file_loads_as_records(protocol):
    """A fluxpy file loads into records holding its arrays, with no fluxpy
    module reachable, under every pickle protocol."""
    V, F, N, P, A = _tiny()
    kinds = np.empty((2, 2), dtype=object)
    kinds[:] = [["dense", "svd"], ["zero", "csr"]]
    data = fluxpy_file("CgalTrimeshShapeModel", V, F, N, P, A, kinds, protocol=protocol)
    assert b"flux.compressed_form_factors" in data and "flux" not in sys.modules
    rec = persist._Loader(__import__("io").BytesIO(data)).load()
    assert rec.kind == "CompressedFormFactorMatrix"
    assert rec._tol == 1e-5 and rec._min_size == 16384
    sa = rec.shape_model
    assert sa.backend == "cgal" and np.array_equal(sa.V, V) and np.array_equal(sa.F, F)
    assert np.array_equal(sa.N, N) and np.array_equal(sa.P, P) and np.array_equal(sa.A, A)
    blocks = rec._root._blocks
    assert [persist._kind(b) for b in blocks.ravel()] == ["dense", "svd", "zero", "csr"]
    assert np.array_equal(blocks[0, 0]._mat, np.eye(2))
    assert scipy.sparse.issparse(blocks[1, 1]._spmat)
    assert blocks[0, 1]._k == 1 and np.array_equal(blocks[0, 1]._u, np.ones((2, 1)))


def test_tree_comparison_counts_leaves_and_mismatches():
    V, F, N, P, A = _tiny()
    inner = np.empty((2, 2), dtype=object)
    inner[:] = [["dense", "csr"], ["csr", "dense"]]
    kinds = np.empty((2, 2), dtype=object)
    kinds[0, 0], kinds[0, 1], kinds[1, 0], kinds[1, 1] = inner, "svd", "zero", "csr"
    rec = persist._Loader(__import__("io").BytesIO(
        fluxpy_file("EmbreeTrimeshShapeModel", V, F, N, P, A, kinds))).load()
    assert rec.shape_model.backend == "embree"
    leaves, mism, counts = persist.compare_trees(rec._root, rec._root)
    assert (leaves, mism) == (7, 0)
    assert counts == {"dense": 2, "csr": 3, "svd": 1, "zero": 1}
    other = np.empty((2, 2), dtype=object)
    other[:] = [["dense", "svd"], ["zero", "csr"]]
    rec2 = persist._Loader(__import__("io").BytesIO(
        fluxpy_file("EmbreeTrimeshShapeModel", V, F, N, P, A, other))).load()
    leaves, mism, _ = persist.compare_trees(rec._root, rec2._root)
    assert mism == 1  # the tree block against a dense leaf


class _Evil:
    def __reduce__(self):
        import os
        return (os.system, ("echo loaded > /dev/null",))


class _Eval:
    def __reduce__(self):
        return (eval, ("1 + 1",))


@pytest.mark.parametrize("obj", [_Evil(), _Eval(), {"x": _Eval()}, [1, _Evil()]])
def test_loader_refuses_other_names(obj):
    """Names outside the list are refused before anything is called."""
    with pytest.raises(pickle.UnpicklingError, match="refusing to load"):
        persist.loads(pickle.dumps(obj))


def test_loader_refuses_other_fluxpy_objects():
    with fluxpy_names() as (shape, _):
        data = pickle.dumps(shape.CgalTrimeshShapeModel(*_tiny()))
    with pytest.raises(pickle.UnpicklingError, match="only a fluxpy CompressedFormFactorMatrix"):
        persist.loads(data)


@pytest.mark.gpu
def test_fluxpy_file_rebuilds_the_same_tree(tmp_path):
    """A fluxpy file written from a tree with this build's leaf kinds loads
    as a CompressedFormFactorMatrix rebuilt on the GPU, with no leaf kind
    differing and the same products as the matrix built directly; this
    build's own files load back through the same loader."""
    import torch
    torch.cuda.set_device(0)
    from fluxpy_amd import meshes
    from fluxpy_amd.compressed_form_factors import CompressedFormFactorMatrix
    from fluxpy_amd.form_factors import FormFactorMatrix
    from fluxpy_amd.shape import CgalTrimeshShapeModel
    V, F = meshes.spherical_crater(npts=24, seed=7)
    sm = CgalTrimeshShapeModel(V, F)
    H = CompressedFormFactorMatrix(sm, tol=1e-3, min_size=256)

    def kinds_of(block):
        kids = getattr(block, "_blocks", None)
        if kids is None:
            return persist._kind(block)
        out = np.empty(kids.shape, dtype=object)
        for idx, b in np.ndenumerate(kids):
            out[idx] = kinds_of(b)
        return out

    data = fluxpy_file("CgalTrimeshShapeModel", sm.V, sm.F, sm.N, sm.P, sm.A,
                       kinds_of(H._root), tol=1e-3, min_size=256)
    path = tmp_path / "FF.bin"
    path.write_bytes(data)
    H2 = CompressedFormFactorMatrix.from_file(str(path))
    chk = H2.reference_tree_check
    assert chk["mismatches"] == 0 and chk["leaves"] > 4
    x = np.random.default_rng(3).standard_normal(sm.num_faces)
    assert np.array_equal(H2 @ x, H @ x)
    H.save(str(tmp_path / "own.bin"))
    H3 = CompressedFormFactorMatrix.from_file(str(tmp_path / "own.bin"))
    assert np.array_equal(H3 @ x, H @ x)
    FF = FormFactorMatrix.from_shape_model(sm)
    FF.save(str(tmp_path / "ff.bin"))
    FF2 = FormFactorMatrix.from_file(str(tmp_path / "ff.bin"))
    assert np.array_equal(FF2 @ x, FF @ x)
