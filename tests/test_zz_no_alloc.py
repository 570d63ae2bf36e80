"""The allocation-free column reduction (flux_vec_reduce_noalloc) against
the workspace entry point (flux_vec_reduce), through the C ABI: identical
results on two calls in a row (the second smaller), every mode and dtype,
and a clear error above the built-in maximum."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

MAX_COLS = 4096


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    torch.cuda.set_device(0)
    return torch


def _reduce_both(t, lib, st, x, y, mode):
    from fluxpy_amd import _lib
    n, k = x.shape
    code = _lib.dtype_code(x.dtype)
    work = t.empty(64 * k, dtype=t.float64, device=x.device)
    ref = t.empty(k, dtype=t.float64, device=x.device)
    got = t.full((k,), np.nan, dtype=t.float64, device=x.device)
    assert lib.flux_vec_reduce(n, k, x.data_ptr(), y.data_ptr(), code, mode, ref.data_ptr(),
                               work.data_ptr(), st) == 0
    assert lib.flux_vec_reduce_noalloc(n, k, x.data_ptr(), y.data_ptr(), code, mode,
                                       got.data_ptr(), st) == 0
    t.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("dtype", ["float64", "float32"])
def test_noalloc_reduce_matches_workspace_entry(torch_dev, dtype):
    from fluxpy_amd import _lib
    t = torch_dev
    lib = _lib.load()
    st = _lib.stream()
    gen = t.Generator(device="cuda").manual_seed(11)
    # two calls in a row, the second smaller, so the shared scratch holds
    # the larger call's partials when the smaller one runs
    for n, k in ((300_000, 1000), (7_001, 3)):
        x = t.randn((n, k), generator=gen, device="cuda", dtype=getattr(t, dtype))
        y = t.randn((n, k), generator=gen, device="cuda", dtype=getattr(t, dtype))
        for mode in (0, 1, 2):
            ref, got = _reduce_both(t, lib, st, x, y, mode)
            assert np.array_equal(ref, got), (n, k, mode)
    # the largest accepted width
    x = t.randn((512, MAX_COLS), generator=gen, device="cuda", dtype=t.float64)
    ref, got = _reduce_both(t, lib, st, x, x, 0)
    assert np.array_equal(ref, got)


def test_noalloc_reduce_rejects_too_many_columns(torch_dev):
    from fluxpy_amd import _lib
    t = torch_dev
    lib = _lib.load()
    x = t.zeros((8, MAX_COLS + 1), dtype=t.float64, device="cuda")
    out = t.zeros(MAX_COLS + 1, dtype=t.float64, device="cuda")
    rc = lib.flux_vec_reduce_noalloc(8, MAX_COLS + 1, x.data_ptr(), x.data_ptr(), _lib.F64, 0,
                                     out.data_ptr(), _lib.stream())
    assert rc != 0
    assert "4096" in _lib.last_error()
