"""CPU checks of the C-ABI library: it loads and exports every entry point
declared in include/ngt_transport.h (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ngt_transport.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ngt_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for name in ("ngt_sample_grid_lut_fwd", "ngt_sample_grid_lut_bwd", "ngt_sample_grid_transform_fwd",
                 "ngt_sample_grid_transform_bwd", "ngt_reduce_grid_blend_fwd", "ngt_reduce_grid_blend_bwd",
                 "ngt_raymarch_grid_transform_fwd", "ngt_raymarch_grid_transform_bwd", "ngt_advect_fwd",
                 "ngt_advect_bwd", "ngt_last_error"):
        assert name in syms


def test_library_exports_all_declared_symbols():
    from ngt import _lib
    path = _lib.lib_path()
    if not os.path.exists(path):
        pytest.skip("library not built")
    lib = ctypes.CDLL(path)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    # the python binding covers exactly the declared ABI
    assert sorted(_lib.SYMBOLS) == declared_symbols()


def test_error_path_without_gpu():
    """Argument validation happens on the host before any HIP call."""
    from ngt import _lib
    if not os.path.exists(_lib.lib_path()):
        pytest.skip("library not built")
    lib = _lib.load()
    shape = (ctypes.c_int64 * 5)(1, 4, 4, 4, 3)
    lut = (ctypes.c_int64 * 5)(1, 4, 4, 4, 4)
    rc = lib.ngt_sample_grid_lut_fwd(None, shape, None, lut, 1, 0, 0, 0, 0.0, 0.0, 1, 1, 0, None, None, 0, None)
    assert rc == 2
    assert lib.ngt_last_error().decode() == "Only 1,2 and 4 Channel supported."
    with pytest.raises(_lib.NGTUnimplemented):
        _lib.check(rc)
    bad = (ctypes.c_int64 * 5)(1, 0, 4, 4, 1)
    rc = lib.ngt_reduce_grid_blend_fwd(None, bad, 0, 0, None, None)
    assert rc == 1 and "Empty input" in lib.ngt_last_error().decode()
    shape1 = (ctypes.c_int64 * 5)(1, 4, 4, 4, 1)
    rc = lib.ngt_reduce_grid_blend_fwd(None, shape1, 7, 0, None, None)
    assert rc == 1 and lib.ngt_last_error().decode() == "invalid blending mode"


def test_mip_argument_checks_without_gpu():
    """Mip arguments are validated on the host (resample_grid.cc:338-347 plus
    the pyramid limits) before any HIP call."""
    from ngt import _lib
    if not os.path.exists(_lib.lib_path()):
        pytest.skip("library not built")
    lib = _lib.load()
    shape = (ctypes.c_int64 * 5)(1, 8, 8, 8, 1)
    lut = (ctypes.c_int64 * 5)(1, 4, 4, 4, 4)
    rc = lib.ngt_sample_grid_lut_fwd(None, shape, None, lut, 1, 0, 2, 0, 0.0, 0.0, 1, 1, 0, None, None, 0, None)
    assert rc == 1 and lib.ngt_last_error().decode() == "when using mipmaps num_mipmaps must be larger than 0."
    rc = lib.ngt_sample_grid_lut_fwd(None, shape, None, lut, 1, 0, 2, 4, 0.0, 0.0, 1, 1, 0, None, None, 0, None)
    assert rc == 1 and "num_mipmaps too large" in lib.ngt_last_error().decode()
    need = lib.ngt_sample_grid_lut_fwd_workspace_size(shape, 2, 3)
    assert need == 256 * ((4 * (64 + 8 + 1) + 255) // 256)
    rc = lib.ngt_sample_grid_lut_fwd(None, shape, None, lut, 1, 0, 2, 3, 0.0, 0.0, 1, 1, 0, None, None, need - 1,
                                     None)
    assert rc == 3
    assert lib.ngt_sample_grid_lut_fwd_workspace_size(shape, 0, 3) == 0
    # the reference gradient op accepts mipmapping NONE only (resample_grid.cc:79, 135)
    rc = lib.ngt_sample_grid_lut_bwd(None, shape, None, None, lut, 1, 0, 2, 2, 0.0, 0.0, 1, 1, 0, 1, None, None, None,
                                     1 << 20, None)
    assert rc == 2 and "reference gradient supports mipmapping 'NONE' only" in lib.ngt_last_error().decode()
    m = (ctypes.c_float * 16)(*[1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])
    f = (ctypes.c_float * 6)(1, 2, -1, 1, 1, -1)
    od = (ctypes.c_int32 * 3)(4, 4, 4)
    rc = lib.ngt_sample_grid_transform_bwd(None, shape, None, m, m, m, f, 1, od, 1, 0, 1, 2, 0.0, 0, 0.5, 1, 1, None,
                                           None, 1 << 20, None)
    assert rc == 2
    assert lib.ngt_sample_grid_transform_bwd_workspace_size(shape, 1, 2, 0) == 256 * ((4 * (64 + 8) + 255) // 256)
    rc = lib.ngt_lod_transform(None, m, m, f, 1, od, 0, 1, 0, None, None)
    assert rc == 1 and lib.ngt_last_error().decode() == "Invalid input_shape"
