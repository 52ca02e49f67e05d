"""ctypes binding of sgd_amd/libsgdct.so (declared in include/sgdct.h).

There is no CPU fallback: every projector call goes to the gfx950 kernels,
and a missing library or a failed call raises.
"""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsgdct.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "sgdct.h")

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.c_void_p
_h = ctypes.c_void_p
_lib = None


class SgdctError(RuntimeError):
    pass


def header_functions(path=HEADER):
    """Names of the functions include/sgdct.h declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text)
    skip = {"if", "defined"}
    return sorted({n for n in names if n not in skip and not n.isupper()})


def lib():
    """Load the library (raises SgdctError if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise SgdctError("%s is missing: build it with `python -m sgd_amd.build` "
                         "(or __graft_entry__.build())" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    init12 = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [_dp, _dp, _dp]
    for name in ("b0_init", "b1_init"):
        f = getattr(L, name)
        f.argtypes = init12
        f.restype = _h
    L.SF_init.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [_dp]
    L.SF_init.restype = _h
    L.sgdct_sf_init_offsets.argtypes = init12
    L.sgdct_sf_init_offsets.restype = _h
    for name in ("h0_init", "hl_init"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [_dp, _dp, _dp]
        f.restype = _h
    for pre in ("b0", "b1", "SF", "h0", "hl"):
        for kind in ("forward", "backward"):
            f = getattr(L, "%s_%s_projection" % (pre, kind))
            f.argtypes = [_dp, _dp, _h]
            f.restype = ctypes.c_int
    L.sgdct_destroy.argtypes = [_h]
    L.sgdct_destroy.restype = None
    for name in ("sgdct_forward_dev", "sgdct_backward_dev"):
        f = getattr(L, name)
        f.argtypes = [_h, _fp, _fp, ctypes.c_int, _fp]
        f.restype = ctypes.c_int
    L.sgdct_project_dev_noalloc.argtypes = [_h, ctypes.c_int, _fp, _fp, ctypes.c_int, _fp]
    L.sgdct_project_dev_noalloc.restype = ctypes.c_int
    L.sgdct_split_config.argtypes = [_h, ctypes.c_int, ctypes.c_int]
    L.sgdct_split_config.restype = ctypes.c_int
    L.sgdct_shape.argtypes = [_h, ctypes.POINTER(ctypes.c_int)]
    L.sgdct_shape.restype = ctypes.c_int
    L.sgdct_shape3.argtypes = [_h, ctypes.POINTER(ctypes.c_int)]
    L.sgdct_shape3.restype = ctypes.c_int
    L.b1_3d_init.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double] * 7 + [_dp]
    L.b1_3d_init.restype = _h
    L.b1_3d_forward_projection.argtypes = [_dp, _dp, _h]
    L.b1_3d_forward_projection.restype = ctypes.c_ulonglong
    L.b1_3d_backward_projection.argtypes = [_dp, _dp, _h]
    L.b1_3d_backward_projection.restype = ctypes.c_int
    L.sf3d_init.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double] * 7 + [_dp, ctypes.c_int]
    L.sf3d_init.restype = _h
    for name in ("sf3d_forward_projection", "sf3d_backward_projection"):
        f = getattr(L, name)
        f.argtypes = [_dp, _dp, _h]
        f.restype = ctypes.c_int
    _v, _i, _ll, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong, ctypes.c_float
    sigs = {
        "sgdct_nn_conv3x3": [_v, _v, _v, _v] + [_i] * 7 + [_v],
        "sgdct_nn_wgrad_workspace": [_i] * 5,
        "sgdct_nn_wgrad": [_v, _v, _v, _v, _ll] + [_i] * 5 + [_v],
        "sgdct_nn_conv3x3_ex": [_v, _v, _i, _v, _v, _v, _v] + [_i] * 7 + [_v],
        "sgdct_nn_wgrad_ex": [_v, _v, _i, _v, _v, _v, _v, _ll] + [_i] * 5 + [_v],
        "sgdct_nn_colsum_workspace": [_ll, _i],
        "sgdct_nn_colsum": [_v, _v, _v, _ll, _ll, _i, _v],
        "sgdct_nn_pad_input": [_v, _v, _ll, _v],
        "sgdct_nn_pool2": [_v, _v] + [_i] * 4 + [_v],
        "sgdct_nn_pool2_back": [_v, _v, _v, _v] + [_i] * 4 + [_v],
        "sgdct_nn_upcat": [_v, _v, _v] + [_i] * 5 + [_v],
        "sgdct_nn_upcat_back": [_v, _v, _v, _v, _v] + [_i] * 6 + [_v],
        "sgdct_nn_head": [_v] * 5 + [_ll, _i, _v],
        "sgdct_nn_head_workspace": [_ll, _i],
        "sgdct_nn_head_back": [_v] * 10 + [_ll, _ll, _i, _v],
        "sgdct_nn_relu_mask": [_v, _v, _v, _ll, _v],
        "sgdct_nn_flip_weights": [_v, _v, _i, _i, _v],
        "sgdct_nn_adam": [_v] * 5 + [_ll, _f, _f, _f, _f, _i, _v],
        "sgdct_nn_fill": [_v, _ll, _f, _v],
        "sgdct_nn_to_bf16": [_v, _v, _ll, _v],
        "sgdct_axpy_batch": [_v, _v, _v, _ll, _ll, _v],
        "sgdct_tv_descent": [_v, _v, _v, _v, _i, _i, _i, _f, _i, _v],
    }
    for name, args in sigs.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = _ll if name.endswith("_workspace") else _i
    for name in ("pixel_matrix_generation", "bilinear_matrix_generation"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [_dp, ctypes.c_int]
        f.restype = ctypes.c_int
    for name in ("h0_matrix_generation", "hl_matrix_generation", "hex_matrix_generation"):
        f = getattr(L, name)
        f.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 4 + [_dp, ctypes.c_int]
        f.restype = ctypes.c_int
    L.sgdct_batch_create.argtypes = [ctypes.POINTER(_h), ctypes.c_int]
    L.sgdct_batch_create.restype = _h
    L.sgdct_batch_destroy.argtypes = [_h]
    L.sgdct_batch_destroy.restype = None
    L.sgdct_batch_layout.argtypes = [_h, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]
    L.sgdct_batch_layout.restype = ctypes.c_int
    L.sgdct_batch_config.argtypes = [_h, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.sgdct_batch_config.restype = ctypes.c_int
    for name in ("sgdct_batch_forward", "sgdct_batch_backward"):
        f = getattr(L, name)
        f.argtypes = [_h, _fp, _fp, _fp]
        f.restype = ctypes.c_int
    L.sgdct_batch_graph_launch.argtypes = [_h, ctypes.c_int, _fp, _fp, _fp]
    L.sgdct_batch_graph_launch.restype = ctypes.c_int
    L.sgdct_eval_count.argtypes = [_h]
    L.sgdct_eval_count.restype = ctypes.c_longlong
    L.sgdct_tv_gradient.argtypes = [_fp, _fp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _fp]
    L.sgdct_sirt_residual.argtypes = [_fp, _fp, _fp, _fp, ctypes.c_longlong, ctypes.c_longlong, _fp]
    L.sgdct_sirt_update.argtypes = [_fp, _fp, _fp, ctypes.c_float, ctypes.c_int, ctypes.c_longlong,
                                    ctypes.c_longlong, _fp]
    L.sgdct_axpy.argtypes = [_fp, _fp, ctypes.c_float, ctypes.c_longlong, _fp]
    for name in ("sgdct_tv_gradient", "sgdct_sirt_residual", "sgdct_sirt_update", "sgdct_axpy"):
        getattr(L, name).restype = ctypes.c_int
    L.sgdct_last_error.argtypes = []
    L.sgdct_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def last_error():
    return lib().sgdct_last_error().decode(errors="replace")


def check(rc, what):
    if rc < 0:
        raise SgdctError("%s failed (%d): %s" % (what, rc, last_error()))
    return rc


def dptr(a):
    return a.ctypes.data_as(_dp)
