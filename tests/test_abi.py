"""The C ABI library loads and exports every symbol include/*.h declares
(no compute calls: runs without a GPU)."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = open(h).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        names.update(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(flux_\w+)\s*\(", txt, re.M))
    return names


def test_header_declares_entry_points():
    names = _declared()
    assert "flux_scene_create" in names and "flux_ff_mask" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    from fluxpy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfluxhip.so not built")
    lib = _lib.load()
    missing = [n for n in sorted(_declared()) if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_ctypes_table_matches_header():
    from fluxpy_amd import _lib
    assert set(_lib.SIGNATURES) == _declared()


def test_error_reporting_without_gpu():
    """Argument checks fail before any device work and set last_error."""
    from fluxpy_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfluxhip.so not built")
    lib = _lib.load()
    rc = lib.flux_scene_create(None, 0, None, -1, 1, None, _lib.ctypes.byref(_lib.ctypes.c_void_p()))
    assert rc < 0
    assert "negative" in _lib.last_error()
    rc = lib.flux_vis_mn(None, None, 0, None, 0, 0, None, 1, None, None)
    assert rc < 0 and "NULL scene" in _lib.last_error()
    assert lib.flux_version().decode().startswith("fluxhip")


def test_slice_workspace_layout_without_gpu():
    """flux_slice_x_bytes (host arithmetic, no device work) and the Python
    side's group width agree on the X-slice workspace: right-hand sides in
    groups of G (16 for at most 16, else 64), np = groups x G; np int32
    exponents, np doubles of column maxima (the view SlicedPanelMatrix.xmax
    reads at bytes 4 np .. 12 np), 5 G slice bytes per facet and group, the
    block column maxima (ceil(n / 64) x np doubles) and 512 x np doubles of
    slicer partials.  The boundary at 16 / 17 right-hand sides is where the
    two widths meet."""
    from fluxpy_amd import _lib
    from fluxpy_amd._panel import x8_width
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libfluxhip.so not built")
    lib = _lib.load()
    assert (x8_width(1), x8_width(16), x8_width(17), x8_width(64)) == (16, 16, 64, 64)
    for n, k in ((41472, 64), (41472, 8), (1000, 16), (1000, 17), (131072, 70), (5, 1)):
        g = x8_width(k)
        nh = (k + g - 1) // g
        np_ = nh * g
        want = np_ * 12 + n * nh * 5 * g + ((n + 63) // 64) * np_ * 8 + 512 * np_ * 8
        assert lib.flux_slice_x_bytes(n, k) == want, (n, k)
    assert lib.flux_slice_x_bytes(-1, 4) == -1
