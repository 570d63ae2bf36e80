"""ctypes binding of libfluxhip.so (the C ABI declared in include/fluxhip.h).

Device memory and streams come from PyTorch-ROCm: arguments are
``tensor.data_ptr()`` values and the current HIP stream.  There is no CPU
fallback: if the library is missing or no GPU is present, every compute entry
raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfluxhip.so")

F32, F64, BF16, F16 = 0, 1, 2, 3
MODE_CGAL, MODE_EMBREE = 0, 1

_c_i64 = ctypes.c_int64
_c_int = ctypes.c_int
_c_dbl = ctypes.c_double
_vp = ctypes.c_void_p

# name -> (restype, argtypes).  Kept in the order of include/fluxhip.h.
SIGNATURES = {
    "flux_last_error": (ctypes.c_char_p, []),
    "flux_version": (ctypes.c_char_p, []),
    "flux_geometry": (_c_int, [_vp, _c_i64, _vp, _c_i64, _c_int, _vp, _vp, _vp, _vp]),
    "flux_scene_create": (_c_int, [_vp, _c_i64, _vp, _c_i64, _c_int, _vp,
                                   ctypes.POINTER(_vp)]),
    "flux_scene_destroy": (_c_int, [_vp]),
    "flux_scene_num_faces": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "flux_scene_num_nodes": (_c_int, [_vp, ctypes.POINTER(_c_i64)]),
    "flux_scene_centroids": (_c_int, [_vp, ctypes.POINTER(_vp)]),
    "flux_scene_copy_nodes": (_c_int, [_vp, _vp, _vp]),
    "flux_vis_mn": (_c_int, [_vp, _vp, _c_i64, _vp, _c_i64, _c_int, _vp, _c_int, _vp, _vp]),
    "flux_occluded": (_c_int, [_vp, _vp, _c_i64, _vp, _c_i64, _c_int, _c_int, _vp, _vp,
                               _c_int, _vp, _vp]),
    "flux_intersect1": (_c_int, [_vp, _vp, _vp, _c_i64, _vp, _vp, _vp]),
    "flux_ff_mask": (_c_int, [_vp, _vp, _vp, _c_int, _vp, _c_i64, _vp, _c_i64, _c_dbl,
                              _c_int, _vp, _vp, _vp]),
    "flux_ff_candidates": (_c_int, [_vp, _vp, _vp, _c_int, _vp, _c_i64, _vp, _c_i64, _c_dbl,
                                    _vp, _vp, _vp]),
    "flux_ff_fill":(_c_int, [_vp, _vp, _vp, _c_int, _vp, _c_i64, _vp, _c_i64, _vp, _vp,
                              _vp, _vp, _vp]),
    "flux_scan_i64": (_c_int, [_vp, _c_i64, _vp, ctypes.POINTER(_c_i64), _vp]),
    "flux_csr_spmm": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_int, _vp, _c_i64, _vp,
                               _c_int, _vp]),
    "flux_csr_rows_chunk": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_int, _vp, _c_i64, _vp,
                                     _c_int, _vp, _vp, _c_dbl, _c_int, _vp, _vp]),
    "flux_csr_jacobi_step": (_c_int, [_c_i64, _vp, _vp, _vp, _c_int, _vp, _vp, _c_i64, _vp,
                                      _c_dbl, _c_int, _vp, _c_int, _vp, _vp]),
    "flux_csr_transpose": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_int, _vp, _vp, _vp,
                                    _vp]),
    "flux_direct_irradiance":(_c_int, [_vp, _vp, _vp, _c_int, _vp, _c_i64, _c_dbl, _c_int,
                                        _vp, _vp]),
    "flux_block_stats": (_c_int, [_vp, _vp, _c_i64, _vp, _vp, _vp, _vp]),
    "flux_block_spmm": (_c_int, [_vp, _vp, _vp, _c_int, _c_i64, _vp, _vp, _vp, _c_i64, _vp,
                                 _vp, _vp, _vp, ctypes.c_uint64, _vp, _vp]),
    "flux_batched_gram": (_c_int, [_c_i64, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_batched_gemm": (_c_int, [_c_i64, _vp, _c_i64, _vp, _vp, _vp, _vp]),
    "flux_batched_syevj": (_c_int, [_c_i64, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _c_int,
                                    _vp]),
    "flux_batched_ns_factor": (_c_int, [_c_i64, _vp, _vp, _vp, _vp]),
    "flux_scale_basis": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_dbl, _vp, _vp, _vp]),
    "flux_sigma_from_lambda": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "flux_scale_cols_inv_sqrt": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "flux_far_matmat": (_c_int, [_c_i64, _c_i64, _c_i64, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_int,
                                 _c_i64, _vp, _vp, _c_i64, _vp, _c_i64, _vp, _c_i64, _vp, _vp,
                                 _c_int, _vp, _c_int, _vp, _vp]),
    "flux_far_work_elems": (_c_i64, [_c_i64, _c_i64, _c_i64, _c_int]),
    "flux_far_pack16": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _vp, _vp,
                                 _vp]),
    "flux_far_mm16": (_c_int, [_c_i64, _vp, _vp, _vp, _c_int, _vp, _c_i64, _c_i64, _vp, _vp,
                               _c_i64, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_far_work16_elems": (_c_i64, [_c_i64, _c_i64]),
    "flux_permute_rows": (_c_int, [_c_i64, _c_i64, _vp, _c_int, _vp, _vp, _c_int, _vp]),
    "flux_jacobi_update": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _vp,
                                    _c_int, _vp, _vp]),
    "flux_panel_count": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp]),
    "flux_panel_order": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp]),
    "flux_panel_fill": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_int, _vp, _c_i64, _vp, _vp,
                                 _c_int, _vp, _vp]),
    "flux_panel_plan": (_c_int, [_c_i64, _vp, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "flux_panel_mm_bf16": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp,
                                    ctypes.c_float, _c_int, _vp, _vp, _vp, _vp, _vp, _c_i64,
                                    _vp, _vp, _vp, _c_i64, _vp, _c_int, _vp, _vp, _vp]),
    "flux_panel_lanemask": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_timer_create": (_c_int, [_vp]),
    "flux_timer_record": (_c_int, [_vp, _vp]),
    "flux_timer_elapsed": (_c_int, [_vp, _vp, _vp]),
    "flux_timer_destroy": (_c_int, [_vp]),
    "flux_panel_pack": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_panel_mm_f64": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_dbl,
                                   _c_int, _vp, _vp, _vp]),
    "flux_refine_prepare_work_bytes": (_c_i64, [_c_i64]),
    "flux_refine_prepare": (_c_int, [_c_i64, _c_i64, _vp, _vp, ctypes.c_float, _c_int, _vp,
                                     _vp, _vp, _vp, _vp]),
    "flux_refine_accumulate": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_scale_to_bf16": (_c_int, [_c_i64, _c_i64, _vp, _vp, ctypes.c_float, _c_int, _vp,
                                    _vp]),
    "flux_panel8_layout": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_panel8_fill": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_int, _vp, _c_i64, _vp, _vp,
                                  _vp, _vp, _vp]),
    "flux_panel8_lanemask": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_panel8_pack": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_slice_x_bytes": (_c_i64, [_c_i64, _c_i64]),
    "flux_slice_x": (_c_int, [_c_i64, _c_i64, _vp, _vp, _c_dbl, _c_int, _vp, _vp]),
    "flux_panel8_mm": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _vp,
                                _c_dbl, _c_int, _vp, _vp, _vp, _c_i64, _vp, _vp, _vp, _c_i64,
                                _vp, _c_int, _vp, _vp, _vp]),
    "flux_vec_reduce":(_c_int, [_c_i64, _c_i64, _vp, _vp, _c_int, _c_int, _vp, _vp, _vp]),
    "flux_vec_reduce_noalloc": (_c_int, [_c_i64, _c_i64, _vp, _vp, _c_int, _c_int, _vp, _vp]),
    "flux_vec_axpby": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _c_int, _vp]),
    "flux_vec_rowscale": (_c_int, [_c_i64, _c_i64, _vp, _c_dbl, _c_int, _c_dbl, _vp, _vp,
                                   _c_int, _vp]),
    "flux_vec_clamp0": (_c_int, [_c_i64, _vp, _vp, _c_int, _vp]),
    "flux_gs_work_elems": (_c_i64, [_c_i64, _c_i64, _c_int]),
    "flux_gs_project": (_c_int, [_c_i64, _c_i64, _c_int, _vp, _c_i64, _vp, _vp, _c_i64, _c_int,
                                 _vp, _vp]),
    "flux_gs_combine": (_c_int, [_c_i64, _c_i64, _c_int, _vp, _c_i64, _vp, _c_i64, _vp, _vp]),
    "flux_gs_normalize": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp]),
    "flux_vec_convert": (_c_int, [_c_i64, _vp, _c_int, _vp, _c_int, _vp]),
    "flux_temperature": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _c_dbl, _vp, _c_dbl, _c_dbl,
                                  _c_int, _c_dbl, _vp, _c_int, _vp]),
    "flux_temperature_adjoint_seed": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _c_dbl, _vp,
                                               _vp, _vp, _vp]),
    "flux_temperature_adjoint_finish": (_c_int, [_c_i64, _c_i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "flux_conduction_work_bytes": (_c_i64, [_c_i64, _c_int]),
    "flux_conduction_q": (_c_int, [_c_i64, _c_int, _vp, _c_dbl, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp, _vp]),
    "flux_conduction_t": (_c_int, [_c_i64, _c_int, _vp, _c_dbl, _vp, _vp, _vp, _vp, _vp, _vp,
                                   _vp, _vp, _vp, _vp, _vp]),
    "flux_thermal_sources": (_c_int, [_c_i64, _vp, _vp, _vp, _vp, _vp, _c_dbl, _c_int, _vp,
                                      _vp]),
    "flux_thermal_update": (_c_int, [_c_i64, _vp, _c_int, _vp, _vp, _vp, _c_int, _vp, _vp, _vp,
                                     _vp]),
}


class FluxError(RuntimeError):
    """An error reported by libfluxhip (status < 0)."""


_lib = None


def load(path=None):
    """Load libfluxhip.so and bind every entry point.  Raises if absent."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise FluxError(
            f"{p} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " or `python fluxpy_amd/_build.py`")
    lib = ctypes.CDLL(p)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error():
    return load().flux_last_error().decode(errors="replace")


def check(rc, what=""):
    if rc != 0:
        raise FluxError(f"{what}: {last_error()} (status {rc})")


def call(name, *args):
    """Call entry point `name` and raise FluxError on a negative status."""
    rc = getattr(load(), name)(*args)
    check(rc, name)
    return rc


# --- torch plumbing ------------------------------------------------------

_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        _torch = t
    return _torch


_gpu_ok = False


def require_gpu():
    """torch, after checking (once per process) that a HIP device is visible
    and the library loads."""
    global _gpu_ok
    t = torch()
    if not _gpu_ok:
        if not t.cuda.is_available():
            raise FluxError("no HIP device is visible: fluxpy_amd has no CPU fallback")
        load()
        _gpu_ok = True
    return t


def device():
    return require_gpu().device("cuda", torch().cuda.current_device())


def stream():
    """Current torch (HIP) stream as a void* value."""
    t = require_gpu()
    return t._C._cuda_getCurrentRawStream(t.cuda.current_device())


def ptr(t):
    return None if t is None else t.data_ptr()


def dtype_code(dt):
    t = torch()
    if dt in (np.float64, np.dtype(np.float64), t.float64):
        return F64
    if dt in (np.float32, np.dtype(np.float32), t.float32):
        return F32
    if dt in (t.bfloat16,):
        return BF16
    if dt in (np.float16, np.dtype(np.float16), t.float16):
        return F16
    raise TypeError(f"unsupported dtype {dt}")


def to_dev(a, dtype=None):
    """numpy array / torch tensor -> contiguous device tensor."""
    t = require_gpu()
    if isinstance(a, t.Tensor):
        out = a.to(device=device(), dtype=dtype or a.dtype)
        return out.contiguous()
    arr = np.ascontiguousarray(a)
    out = t.from_numpy(arr).to(device())
    if dtype is not None:
        out = out.to(dtype)
    return out.contiguous()


def torch_dtype(np_dtype):
    t = torch()
    return {np.dtype(np.float64): t.float64, np.dtype(np.float32): t.float32}[np.dtype(np_dtype)]
