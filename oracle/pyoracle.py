"""ctypes front end of the CPU oracle (oracle/sgdct_oracle.c).

Test infrastructure only: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it to check results; the product (sgd_amd) never does.

Two builds of the same source are exposed:
  Oracle("ref")    FP32 footprint arithmetic as in layers/A_*.cu, FP64 sums
  Oracle("exact")  the same formulas evaluated in FP64
Images are (ny, nx) arrays indexed [iy, ix]; sinograms are (np, nu).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
BASES = {"B0": 0, "B1": 1, "SF": 2, "H0": 3, "Hl": 4}

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build():
    """Compile both oracle libraries (gcc; no GPU toolchain needed)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _ptr(a):
    return a.ctypes.data_as(_dp)


class Oracle:
    _libs = {}

    def __init__(self, mode="exact"):
        if mode not in ("ref", "exact"):
            raise ValueError(mode)
        path = os.path.join(_HERE, "liboracle_%s.so" % mode)
        if not os.path.exists(path):
            build()
        if path not in Oracle._libs:
            lib = ctypes.CDLL(path)
            lib.or_geo_create.restype = ctypes.c_void_p
            lib.or_geo_create.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double] * 5 + [_dp, _dp]
            lib.or_geo_destroy.argtypes = [ctypes.c_void_p]
            lib.or_geo_u.restype = ctypes.c_double
            lib.or_geo_u.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double]
            lib.or_geo_view.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp]
            for name in ("or_forward", "or_backward", "or_forward_exact"):
                f = getattr(lib, name)
                f.argtypes = [ctypes.c_int, ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            lib.or_footprint.restype = ctypes.c_int
            lib.or_footprint.argtypes = [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 4 + [_ip, _dp, ctypes.c_int]
            lib.or_eval_count.restype = ctypes.c_longlong
            lib.or_eval_count.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
            lib.or_strip_exact.restype = ctypes.c_double
            lib.or_strip_exact.argtypes = [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 4
            lib.or_strip_simpson.restype = ctypes.c_double
            lib.or_strip_simpson.argtypes = [ctypes.c_int, ctypes.c_void_p] + [ctypes.c_int] * 4 + [
                ctypes.c_double, ctypes.c_int, _dp]
            for name in ("or_b0_conv", "or_b1_conv", "or_box3"):
                f = getattr(lib, name)
                f.restype = ctypes.c_double
                f.argtypes = [ctypes.c_double] * 4
            lib.or_s2tri.restype = ctypes.c_double
            lib.or_s2tri.argtypes = [ctypes.c_double] * 3
            lib.or_precision_bits.restype = ctypes.c_int
            for name in ("or_hex_forward", "or_hex_backward"):
                f = getattr(lib, name)
                f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double, _dp, _dp, ctypes.c_int, ctypes.c_int,
                              ctypes.c_int]
            lib.or_hex_footprint.restype = ctypes.c_int
            lib.or_hex_footprint.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4 + [
                _ip, _dp, ctypes.c_int]
            lib.or_hex_strip_exact.restype = ctypes.c_double
            lib.or_hex_strip_exact.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4
            lib.or_hl_conv.restype = ctypes.c_double
            lib.or_hl_conv.argtypes = [ctypes.c_double] * 5
            lib.or_geo3_create.restype = ctypes.c_void_p
            lib.or_geo3_create.argtypes = [ctypes.c_int] * 6 + [ctypes.c_double] * 7 + [_dp]
            lib.or_geo3_destroy.argtypes = [ctypes.c_void_p]
            for name in ("or_b1_3d_forward", "or_b1_3d_backward"):
                f = getattr(lib, name)
                f.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, ctypes.c_int]
            lib.or_b1_3d_cell_exact.restype = ctypes.c_double
            lib.or_b1_3d_cell_exact.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7
            for name in ("or_sf3d_forward", "or_sf3d_backward"):
                f = getattr(lib, name)
                f.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int]
            lib.or_sf3d_cell_exact.restype = ctypes.c_double
            lib.or_sf3d_cell_exact.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7
            Oracle._libs[path] = lib
        self.lib = Oracle._libs[path]
        self.mode = mode

    # -- primitives ---------------------------------------------------------
    def b0_conv(self, y, v1, v2, v3):
        return self.lib.or_b0_conv(y, v1, v2, v3)

    def b1_conv(self, x1, x2, a, b):
        return self.lib.or_b1_conv(x1, x2, a, b)

    def s2tri(self, a, b, x):
        return self.lib.or_s2tri(a, b, x)

    def box3(self, y, v1, v2, v3):
        return self.lib.or_box3(y, v1, v2, v3)

    def hl_conv(self, y, h1, h2, h3, h4):
        return self.lib.or_hl_conv(y, h1, h2, h3, h4)


class Geometry:
    """2-D fan-beam geometry (utils/geo.cu:232-269) held by one oracle build."""

    def __init__(self, oracle, nx, ny, angles, nu, dx, dy, du, lsd, lso, drho=None):
        self.o = oracle
        self.nx, self.ny, self.nu = int(nx), int(ny), int(nu)
        self.angles = np.ascontiguousarray(angles, dtype=np.float64)
        self.np = len(self.angles)
        self.dx, self.dy, self.du, self.lsd, self.lso = map(float, (dx, dy, du, lsd, lso))
        self._drho = None if drho is None else np.ascontiguousarray(drho, dtype=np.float64)
        self.h = oracle.lib.or_geo_create(self.nx, self.ny, self.np, self.nu, self.dx, self.dy, self.du,
                                          self.lsd, self.lso, _ptr(self.angles),
                                          None if self._drho is None else _ptr(self._drho))

    def __del__(self):
        try:
            self.o.lib.or_geo_destroy(self.h)
        except Exception:
            pass

    def u_of(self, p, x, y):
        return self.o.lib.or_geo_u(self.h, p, x, y)

    def view(self, p):
        s, d, u = (np.zeros(2) for _ in range(3))
        self.o.lib.or_geo_view(self.h, p, _ptr(s), _ptr(d), _ptr(u))
        return s, d, u

    def forward(self, basis, img, views=None, pad=0):
        img = np.ascontiguousarray(img, dtype=np.float64).reshape(self.ny, self.nx)
        p0, p1 = views if views is not None else (0, self.np)
        sino = np.zeros((self.np, self.nu))
        self.o.lib.or_forward(BASES[basis], self.h, _ptr(img), _ptr(sino), p0, p1, pad)
        return sino

    def backward(self, basis, sino, views=None, pad=0):
        sino = np.ascontiguousarray(sino, dtype=np.float64).reshape(self.np, self.nu)
        p0, p1 = views if views is not None else (0, self.np)
        img = np.zeros((self.ny, self.nx))
        self.o.lib.or_backward(BASES[basis], self.h, _ptr(sino), _ptr(img), p0, p1, pad)
        return img

    def forward_exact(self, basis, img, views=None, pad=1):
        img = np.ascontiguousarray(img, dtype=np.float64).reshape(self.ny, self.nx)
        p0, p1 = views if views is not None else (0, self.np)
        sino = np.zeros((self.np, self.nu))
        self.o.lib.or_forward_exact(BASES[basis], self.h, _ptr(img), _ptr(sino), p0, p1, pad)
        return sino

    def footprint(self, basis, p, ix, iy, pad=0, cap=64):
        us = np.zeros(cap, dtype=np.int32)
        ws = np.zeros(cap)
        n = self.o.lib.or_footprint(BASES[basis], self.h, p, ix, iy, pad,
                                    us.ctypes.data_as(_ip), _ptr(ws), cap)
        n = min(n, cap)
        return us[:n].copy(), ws[:n].copy()

    def eval_count(self, basis, views=None):
        p0, p1 = views if views is not None else (0, self.np)
        return self.o.lib.or_eval_count(BASES[basis], self.h, p0, p1)

    def strip_exact(self, basis, p, u, ix, iy):
        return self.o.lib.or_strip_exact(BASES[basis], self.h, p, u, ix, iy)

    def strip_simpson(self, basis, p, u, ix, iy, eps=1e-12, depth=30):
        err = ctypes.c_double(0.0)
        v = self.o.lib.or_strip_simpson(BASES[basis], self.h, p, u, ix, iy, eps, depth, ctypes.byref(err))
        return v, err.value


class HexGeometry(Geometry):
    """Hexagonal lattice of nl rows x nc columns (layers/A_H0.cu, A_Hl.cu):
    odd rows are shifted by half a column and hold nc-1 points.  Images are
    (nl, nc) arrays indexed [il, ic]; the last column of odd rows is unused."""

    def __init__(self, oracle, nl, nc, angles, nu, dt, du, lsd, lso, drho=None):
        super().__init__(oracle, nc, nl, angles, nu, dt * 3 ** 0.5, dt * 0.5, du, lsd, lso, drho=drho)
        self.nl, self.nc, self.dt = int(nl), int(nc), float(dt)

    def mask(self):
        m = np.ones((self.nl, self.nc), dtype=bool)
        m[1::2, -1] = False
        return m

    def forward(self, basis, img, views=None, pad=0):
        img = np.ascontiguousarray(img, dtype=np.float64).reshape(self.nl, self.nc)
        p0, p1 = views if views is not None else (0, self.np)
        sino = np.zeros((self.np, self.nu))
        self.o.lib.or_hex_forward(BASES[basis], self.h, self.dt, _ptr(img), _ptr(sino), p0, p1, pad)
        return sino

    def backward(self, basis, sino, views=None, pad=0):
        sino = np.ascontiguousarray(sino, dtype=np.float64).reshape(self.np, self.nu)
        p0, p1 = views if views is not None else (0, self.np)
        img = np.zeros((self.nl, self.nc))
        self.o.lib.or_hex_backward(BASES[basis], self.h, self.dt, _ptr(sino), _ptr(img), p0, p1, pad)
        return img

    def footprint(self, basis, p, ic, il, pad=0, cap=64):
        us = np.zeros(cap, dtype=np.int32)
        ws = np.zeros(cap)
        n = self.o.lib.or_hex_footprint(BASES[basis], self.h, self.dt, p, ic, il, pad,
                                        us.ctypes.data_as(_ip), _ptr(ws), cap)
        n = min(n, cap)
        return us[:n].copy(), ws[:n].copy()

    def strip_exact(self, basis, p, u, ic, il):
        return self.o.lib.or_hex_strip_exact(BASES[basis], self.h, self.dt, p, u, ic, il)

    def centre(self, ic, il, basis):
        dy = 1.5 * self.dt if basis == "H0" else 0.5 * self.dt
        dx = 3 ** 0.5 * self.dt
        cy = (il - 0.5 * (self.nl - 1)) * dy
        cx = (ic - 0.5 * (self.nc - (2 if il & 1 else 1))) * dx
        return cx, cy


class Geometry3:
    """Cone-beam geometry of layers/A_B1_3D.cu: nx x ny x nz voxels, an nu x nv
    flat detector.  Volumes are (nz, ny, nx) arrays [iz, iy, ix]; projections
    (np, nv, nu) [p, j, u]."""

    def __init__(self, oracle, nx, ny, nz, angles, nu, nv, dx, dy, dz, du, dv, lsd, lso):
        self.o = oracle
        self.lib = oracle.lib
        self.nx, self.ny, self.nz, self.nu, self.nv = nx, ny, nz, nu, nv
        self.angles = np.ascontiguousarray(angles, dtype=np.float64)
        self.np = self.angles.size
        self.d = (dx, dy, dz, du, dv, lsd, lso)
        self.h = self.lib.or_geo3_create(nx, ny, nz, self.np, nu, nv, dx, dy, dz, du, dv, lsd, lso,
                                         _ptr(self.angles))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.or_geo3_destroy(self.h)
            self.h = None

    def _views(self, views):
        return (0, self.np) if views is None else views

    def forward(self, vol, views=None):
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(self.nz, self.ny, self.nx)
        out = np.zeros((self.np, self.nv, self.nu))
        p0, p1 = self._views(views)
        self.lib.or_b1_3d_forward(self.h, _ptr(vol), _ptr(out), p0, p1)
        return out

    def backward(self, proj, views=None):
        proj = np.ascontiguousarray(proj, dtype=np.float64).reshape(self.np, self.nv, self.nu)
        out = np.zeros((self.nz, self.ny, self.nx))
        p0, p1 = self._views(views)
        self.lib.or_b1_3d_backward(self.h, _ptr(proj), _ptr(out), p0, p1)
        return out

    def cell_exact(self, p, u, j, ix, iy, iz, panels=8):
        return self.lib.or_b1_3d_cell_exact(self.h, p, u, j, ix, iy, iz, panels)

    # separable footprints (layers/SF_projector.cu), trap_v: SF-TT, else SF-TR
    def sf_forward(self, vol, trap_v=False, views=None):
        vol = np.ascontiguousarray(vol, dtype=np.float64).reshape(self.nz, self.ny, self.nx)
        out = np.zeros((self.np, self.nv, self.nu))
        p0, p1 = self._views(views)
        self.lib.or_sf3d_forward(self.h, int(bool(trap_v)), _ptr(vol), _ptr(out), p0, p1)
        return out

    def sf_backward(self, proj, trap_v=False, views=None):
        proj = np.ascontiguousarray(proj, dtype=np.float64).reshape(self.np, self.nv, self.nu)
        out = np.zeros((self.nz, self.ny, self.nx))
        p0, p1 = self._views(views)
        self.lib.or_sf3d_backward(self.h, int(bool(trap_v)), _ptr(proj), _ptr(out), p0, p1)
        return out

    def sf_cell_exact(self, p, u, j, ix, iy, iz, panels=8):
        return self.lib.or_sf3d_cell_exact(self.h, p, u, j, ix, iy, iz, panels)


def matrix_generation(oracle, basis, np_, nu, nx, ny, dx, dy, du, lsd, lso, buffer_size=None, fixed=True):
    """COO triplets (rows, cols, vals) of pixel_matrix_generation (B0) /
    bilinear_matrix_generation (B1); `fixed` selects the B1 projector's
    interval instead of the reference kernel's one-bin-off one; raises if
    buffer_size is too small."""
    lib = oracle.lib
    f = lib.or_matrix_generation
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int] * 5 + [ctypes.c_double] * 5 + [_dp, ctypes.c_longlong, ctypes.c_int]
    cap = buffer_size or np_ * nu * nx * ny
    buf = np.zeros((cap, 3))
    n = f(BASES[basis], np_, nu, nx, ny, dx, dy, du, lsd, lso, _ptr(buf), cap, int(bool(fixed)))
    if n < 0:
        raise ValueError("buffer too small")
    return buf[:n, 0].astype(np.int64), buf[:n, 1].astype(np.int64), buf[:n, 2].copy()
