"""The C-ABI library loads on a CPU-only host and exports every function that
include/sgdct.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os

import pytest

from sgd_amd import _lib


def test_header_declares_reference_entry_points():
    names = set(_lib.header_functions())
    for pre in ("b0", "b1", "SF"):
        assert "%s_forward_projection" % pre in names
        assert "%s_backward_projection" % pre in names
    assert {"b0_init", "b1_init", "SF_init", "sgdct_destroy", "sgdct_forward_dev",
            "sgdct_backward_dev", "sgdct_last_error"} <= names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.fail("libsgdct.so not built (run __graft_entry__.build())")
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _lib.header_functions() if not hasattr(L, n)]
    assert not missing, missing


def test_binding_sets_signatures():
    L = _lib.lib()
    assert L.b1_init.restype is ctypes.c_void_p
    assert L.sgdct_last_error.restype is ctypes.c_char_p


def test_split_config_rejects_bad_arguments_without_a_device():
    """Argument checks run before any device work: a null handle or a split
    outside 0..32 is SGDCT_EINVAL with a message."""
    L = _lib.lib()
    assert L.sgdct_split_config(None, 0, 0) == -1
    assert "split" in _lib.last_error()
    assert L.sgdct_split_config(None, 33, 0) == -1


def test_tv_descent_rejects_bad_arguments_without_a_device():
    """Null pointers, empty shapes, a non-positive eps or a negative step
    count are SGDCT_EINVAL before any device work; zero steps do nothing."""
    L = _lib.lib()
    p = ctypes.c_void_p(16)  # never dereferenced: every call below fails its checks or returns first
    assert L.sgdct_tv_descent(None, p, p, p, 8, 8, 1, 1e-8, 1, None) == -1
    assert "tv_descent" in _lib.last_error()
    assert L.sgdct_tv_descent(p, p, p, p, 0, 8, 1, 1e-8, 1, None) == -1
    assert L.sgdct_tv_descent(p, p, p, p, 8, 8, 1, 0.0, 1, None) == -1
    assert L.sgdct_tv_descent(p, p, p, p, 8, 8, 1, 1e-8, -1, None) == -1
    assert L.sgdct_tv_descent(p, p, p, p, 8, 8, 1, 1e-8, 0, None) == 0


def test_init_and_calls_reject_bad_arguments_without_a_device():
    """Invalid geometry, sizes past the staged maxima (512 voxels per column,
    512 detector rows) and null handles are errors -- a NULL handle or a
    negative code, with a message -- before any device work."""
    L = _lib.lib()
    ang = (ctypes.c_double * 4)(0.0, 0.5, 1.0, 1.5)
    assert L.b1_init(0, 8, 4, 16, 1.0, 1.0, 1.0, 1000.0, 500.0, ang, None, None) is None
    assert "geometry" in _lib.last_error()
    assert L.b0_init(8, 8, 4, 16, 0.0, 1.0, 1.0, 1000.0, 500.0, ang, None, None) is None
    assert L.b1_init(8, 8, 4, 16, 1.0, 1.0, 1.0, 1000.0, 500.0, None, None, None) is None
    assert L.SF_init(8, 8, 4, 0, 1.0, 1.0, 1.0, 1000.0, 500.0, ang) is None
    assert L.h0_init(9, 8, 4, 16, 0.0, 1.0, 1000.0, 500.0, ang, None, None) is None
    assert "dt" in _lib.last_error()
    cone = (1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 500.0)
    assert L.b1_3d_init(8, 8, 0, 4, 16, 16, *cone, ang) is None
    assert L.b1_3d_init(8, 8, 513, 4, 16, 16, *cone, ang) is None
    assert L.sf3d_init(8, 8, 16, 4, 16, 513, *cone, ang, 1) is None
    assert L.b1_forward_projection(None, None, None) < 0
    assert "null" in _lib.last_error()
    assert L.h0_backward_projection(None, None, None) < 0
    assert L.sgdct_forward_dev(None, None, None, 1, None) < 0
    assert L.sgdct_backward_dev(None, None, None, 1, None) < 0
