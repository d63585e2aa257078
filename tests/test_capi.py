"""The C-ABI library loads on a machine without a GPU and exports every entry point include/lbm.h
declares; calls that touch no device (version, defaults, error reporting) behave as documented."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "lbm.h")
LIB = os.path.join(ROOT, "cells_amd", "liblbm_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lbm_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("lbm_create_2d", "lbm_create_3d", "lbm_enqueue_steps", "lbm_read_forces", "lbm_set_sdf",
                 "lbm_adjoint_forward", "lbm_adjoint_backward", "lbm_batch_create_2d"):
        assert must in names
    assert len(names) >= 40


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_function():
    lib = C.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_host_only_calls():
    from cells_amd import lbm

    lib = lbm.load_library()  # binds every signature; raises if a symbol is missing
    assert lib.lbm_abi_version() == 2
    c2 = lbm.Config2D()
    lib.lbm_config2d_defaults(C.byref(c2))
    assert abs(c2.lambda_ - 3 / 16) < 1e-15 and c2.force_every == 1
    c3 = lbm.Config3D()
    lib.lbm_config3d_defaults(C.byref(c3))
    assert c3.lambda_ != c3.lambda_  # NaN: follow tau
    # argument validation fails before any device work
    h = C.c_void_p()
    assert lib.lbm_create_2d(None, 0, C.byref(h)) != 0
    assert lib.lbm_last_error()
    assert lib.lbm_adjoint_forward(None, None, None) != 0
    assert lib.lbm_copy_roofline(0, -1, 19, 4, 0, 1, C.byref(C.c_double())) != 0
    assert lib.lbm_copy_roofline(0, 1024, 19, 4, 96, 1, C.byref(C.c_double())) != 0
    assert lib.lbm_pull_copy_roofline(0, 2048, 8, 8, 4, 1, 1, C.byref(C.c_double()), C.byref(C.c_int64())) != 0
    assert lib.lbm_pull_copy_roofline(0, 64, 2, 8, 4, 1, 1, C.byref(C.c_double()), C.byref(C.c_int64())) != 0
    assert lib.lbm_pull_copy_roofline(0, 64, 8, 8, 4, 2, 1, C.byref(C.c_double()), C.byref(C.c_int64())) != 0
    assert lib.lbm_copy_roofline_filled(0, 1024, 19, 4, 256, 2, 1, C.byref(C.c_double())) != 0
    assert lib.lbm_enqueue_steps_noalloc(None, 1) != 0
    assert lib.lbm_set_launch(None, 64, 0) != 0
    assert lib.lbm_population_stride(None, None) != 0
    assert lib.lbm_set_population_stride(None, 0) != 0
    assert lib.lbm_population_layout(None, None, None) != 0
    assert lib.lbm_set_population_layout(None, 1, 0) != 0
    assert lib.lbm_rebuild_flags(None) != 0
    assert lib.lbm_set_open_faces(None, 1) != 0
    assert lib.lbm_brush(None, 1.0, 1.0, 1.0, 0, 1) != 0
    assert lib.lbm_brush_stroke(None, None, 0, 1.0, 0) != 0
    assert lib.lbm_device_name(0, None, 0) != 0
    c3.precision = 4  # beyond LBM_FP16_SR
    assert lib.lbm_create_3d(C.byref(c3), 0, C.byref(h)) != 0 and b"precision" in lib.lbm_last_error()
