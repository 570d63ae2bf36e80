"""Loading saved form-factor matrices: this package's own pickles and the
``FF.bin`` files fluxpy writes (compressed_form_factors.py:748-751, 801-803).

A pickle names the classes and functions its loader calls.  The loader here
(``load``) resolves only an explicit list of names:

* NumPy's array reconstruction and SciPy's sparse matrix classes;
* this package's own classes and restore functions (its pickles call the
  device rebuilds of ``__reduce__``);
* fluxpy's classes (``flux.shape.*ShapeModel``, ``flux.compressed_form_factors
  .*Block`` and ``.CompressedFormFactorMatrix``), which are mapped onto inert
  records: nothing of fluxpy is imported or run, and the records only hold
  the arrays and numbers the file contains.

Every other name raises ``pickle.UnpicklingError`` before anything is
called.  A fluxpy file becomes this package's ``CompressedFormFactorMatrix``
rebuilt on the GPU from the pickled shape model (V, F, N, P, A) and the
pickled build parameters (tol, min_size, max_depth, force_max_depth, the
root block class).  The leaves of the stored tree are then compared with the
rebuilt one, position by position, and the comparison is kept on the result
as ``reference_tree_check``: the leaf kinds are the reference's decisions,
and the rebuilt operator reproduces them where the singular values do not
sit at the threshold.  The factors themselves are recomputed, not copied,
because the rebuilt operator keeps its own cluster-ordered device layout.
"""
import importlib
import io
import pickle
import warnings

import numpy as np

# NumPy / SciPy names a pickle of arrays and sparse matrices uses, by the
# module paths of the releases that write them
_ARRAY_NAMES = {
    ("numpy.core.multiarray", "_reconstruct"), ("numpy._core.multiarray", "_reconstruct"),
    ("numpy.core.multiarray", "scalar"), ("numpy._core.multiarray", "scalar"),
    ("numpy.core.numeric", "_frombuffer"), ("numpy._core.numeric", "_frombuffer"),
    ("numpy", "ndarray"), ("numpy", "dtype"),
    ("copyreg", "_reconstructor"), ("builtins", "object"),
    ("_codecs", "encode"),            # bytes in protocols 0 and 1
}
# protocols 0 and 1 write Python 2 module names
_PY2 = {("copy_reg", "_reconstructor"): ("copyreg", "_reconstructor"),
        ("__builtin__", "object"): ("builtins", "object")}
_SPARSE = {"csr_matrix", "csc_matrix", "coo_matrix"}
_SPARSE_MODULES = {"scipy.sparse", "scipy.sparse.csr", "scipy.sparse._csr", "scipy.sparse.csc",
                   "scipy.sparse._csc", "scipy.sparse.coo", "scipy.sparse._coo"}
# this package's names its own pickles use
_OWN = {
    "fluxpy_amd._hbuild": {"restore"},
    "fluxpy_amd.form_factors": {"_rebuild_ffm", "FormFactorMatrix"},
    "fluxpy_amd.shape": {"CgalTrimeshShapeModel", "EmbreeTrimeshShapeModel"},
    "fluxpy_amd.quadtree": {"get_quadrant_order"},
    "fluxpy_amd.octree": {"get_octant_order"},
}
_OWN_CFF = "fluxpy_amd.compressed_form_factors"
# fluxpy's block classes (compressed_form_factors.py:35-700)
_REF_BLOCKS = {
    "CompressedFormFactorMatrix", "FormFactorQuadtreeBlock", "FormFactorOctreeBlock",
    "FormFactor2dTreeBlock", "FormFactorBlockMatrix", "FormFactorDenseBlock",
    "FormFactorSparseBlock", "FormFactorCsrBlock", "FormFactorSvdBlock",
    "FormFactorZeroBlock", "FormFactorNullBlock", "FormFactorPartitionBlock",
}
_REF_SHAPES = {"CgalTrimeshShapeModel", "EmbreeTrimeshShapeModel"}


class _Record:
    """State of a fluxpy object: its class name and attribute dict."""

    kind = None

    def __setstate__(self, state):
        if isinstance(state, tuple) and len(state) == 2:  # (dict, slots)
            state = {**(state[0] or {}), **(state[1] or {})}
        if not isinstance(state, dict):
            raise pickle.UnpicklingError(f"unexpected state for {self.kind}")
        self.__dict__.update(state)


def _record_class(name):
    return type(name, (_Record,), {"kind": name})


class _ShapeArgs:
    """A fluxpy shape model as its __reduce__ stores it: the constructor
    arguments (V, F, N, P, A) of shape.py:55, 114-115."""

    def __init__(self, backend, *args):
        if not 2 <= len(args) <= 5:
            raise pickle.UnpicklingError("shape model: expected (V, F, N, P, A)")
        self.backend = backend
        V, F, N, P, A = (list(args) + [None] * 5)[:5]
        self.V, self.F, self.N, self.P, self.A = V, F, N, P, A


class _Loader(pickle.Unpickler):
    def find_class(self, module, name):
        module, name = _PY2.get((module, name), (module, name))
        if (module, name) in _ARRAY_NAMES:
            return getattr(importlib.import_module(module), name)
        if module in _SPARSE_MODULES and name in _SPARSE:
            import scipy.sparse
            return getattr(scipy.sparse, name)
        if module in _OWN and name in _OWN[module]:
            return getattr(importlib.import_module(module), name)
        if module == _OWN_CFF:
            mod = importlib.import_module(module)
            obj = getattr(mod, name, None)
            if isinstance(obj, type) and obj.__module__ == _OWN_CFF:
                return obj
        if module == "flux.compressed_form_factors" and name in _REF_BLOCKS:
            return _record_class(name)
        if module == "flux.shape" and name in _REF_SHAPES:
            backend = "cgal" if name.startswith("Cgal") else "embree"
            return lambda *args: _ShapeArgs(backend, *args)
        raise pickle.UnpicklingError(f"refusing to load {module}.{name}")


def loads(data):
    """``load`` from bytes."""
    return _convert(_Loader(io.BytesIO(data)).load())


def load(path):
    """Load a saved form-factor matrix (this package's pickles, or fluxpy's
    FF.bin files, converted as described in the module docstring)."""
    with open(path, "rb") as f:
        return _convert(_Loader(f).load())


# --------------------------------------------------------------------------
# fluxpy files -> this package's objects
# --------------------------------------------------------------------------

def _convert(obj):
    if isinstance(obj, _Record) and obj.kind == "CompressedFormFactorMatrix":
        return _from_reference(obj)
    if isinstance(obj, (_Record, _ShapeArgs)):
        raise pickle.UnpicklingError(
            "only a fluxpy CompressedFormFactorMatrix can be loaded from a fluxpy file")
    return obj


def _shape_model(sa):
    from .shape import CgalTrimeshShapeModel, EmbreeTrimeshShapeModel
    cls = CgalTrimeshShapeModel if sa.backend == "cgal" else EmbreeTrimeshShapeModel
    V = np.asarray(sa.V)
    F = np.asarray(sa.F)
    N = None if sa.N is None else np.asarray(sa.N)
    A = None if sa.A is None else np.asarray(sa.A)
    return cls(V, F, N=N, A=A)


def _kind(block):
    """Leaf kind of a block (a fluxpy record or one of this package's)."""
    name = block.kind if isinstance(block, _Record) else type(block).__name__
    return {"FormFactorDenseBlock": "dense", "FormFactorCsrBlock": "csr",
            "FormFactorSparseBlock": "csr", "FormFactorSvdBlock": "svd",
            "FormFactorZeroBlock": "zero", "FormFactorNullBlock": "null"}.get(name, "tree")


def _children(block):
    blocks = getattr(block, "_blocks", None)
    return None if blocks is None else np.asarray(blocks, dtype=object)


def compare_trees(ref_root, root):
    """Walk both block trees in step: (leaves compared, positions whose
    kinds differ, counts of the reference's leaf kinds)."""
    leaves = mismatches = 0
    counts = {}
    stack = [(ref_root, root)]
    while stack:
        a, b = stack.pop()
        ka, kb = _kind(a), _kind(b)
        if ka == "tree":
            ca, cb = _children(a), _children(b) if kb == "tree" else None
            if cb is None or ca.shape != cb.shape:
                mismatches += 1
                continue
            stack.extend(zip(ca.ravel(), cb.ravel()))
            continue
        leaves += 1
        counts[ka] = counts.get(ka, 0) + 1
        mismatches += ka != kb
    return leaves, mismatches, counts


def _from_reference(rec):
    from .compressed_form_factors import (CompressedFormFactorMatrix,
                                          FormFactorOctreeBlock, FormFactorQuadtreeBlock)
    st = rec.__dict__
    sa = st.get("shape_model")
    if not isinstance(sa, _ShapeArgs):
        raise pickle.UnpicklingError("fluxpy file: no shape model")
    root = st.get("_root")
    root_cls = FormFactorOctreeBlock if (isinstance(root, _Record) and
                                         root.kind == "FormFactorOctreeBlock") \
        else FormFactorQuadtreeBlock
    sm = _shape_model(sa)
    H = CompressedFormFactorMatrix(sm, tol=float(st.get("_tol", 1e-5)),
                                   min_size=int(st.get("_min_size", 16384)),
                                   max_depth=st.get("_max_depth"),
                                   force_max_depth=bool(st.get("_force_max_depth", False)),
                                   RootBlock=root_cls)
    leaves, mism, counts = compare_trees(root, H._root)
    H.reference_tree_check = {"leaves": leaves, "mismatches": mism, "reference_leaves": counts}
    if mism:
        warnings.warn(f"fluxpy file: {mism} of {leaves} leaves differ in kind from the "
                      "rebuilt tree (singular values at the threshold)", RuntimeWarning)
    return H
