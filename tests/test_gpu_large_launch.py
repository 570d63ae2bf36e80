"""GPU regression test: element-wise kernels over more than 2^32 elements.
A dispatch counts work-items in 32 bits, so a launch of one thread per
element wraps past 2^32 and leaves the tail untouched; the panel fill zeroes
its panels with a grid-stride launch.  Row-block panels of the 819,200-facet
terrain reach 4.7e9 BF16 entries per column piece."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_panel_fill_zeroes_past_2_32_entries():
    import torch
    torch.cuda.set_device(0)
    from fluxpy_amd import _lib
    dev = _lib.device()
    total_cols = (1 << 26) + 4096                    # 64 * total_cols > 2^32 entries
    nel = total_cols * 64
    assert nel > 2 ** 32
    panel = torch.ones(nel, dtype=torch.bfloat16, device=dev)      # stale contents
    cols = torch.empty(total_cols, dtype=torch.int32, device=dev)
    # one row block of 64 rows with a single nonzero, F[0, 0] = 0.5
    indptr = torch.tensor([0] + [1] * 64, dtype=torch.int64, device=dev)
    indices = torch.zeros(1, dtype=torch.int32, device=dev)
    data = torch.tensor([0.5], dtype=torch.float32, device=dev)
    colptr = torch.tensor([0, total_cols], dtype=torch.int64, device=dev)
    _lib.call("flux_panel_fill", 64, 1, indptr.data_ptr(), indices.data_ptr(), data.data_ptr(),
              _lib.dtype_code(data.dtype), colptr.data_ptr(), total_cols, cols.data_ptr(),
              panel.data_ptr(), _lib.dtype_code(torch.bfloat16), None, _lib.stream())
    torch.cuda.synchronize()
    assert int(torch.count_nonzero(panel).item()) == 1
    assert float(panel[0].item()) == 0.5
    assert int(cols[0].item()) == 0 and int(torch.count_nonzero(cols).item()) == 0
    del panel, cols
    torch.cuda.empty_cache()
