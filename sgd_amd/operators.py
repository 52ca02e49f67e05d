"""pylops-style LinearOperators over the gfx950 projectors.

Mirrors python/operators/operators.py of the reference: same class names,
constructor arguments, shapes, dtype float64, and the same flattened host
layouts (image x[ix*ny + iy], sinogram y[p*nu + u]; cone beam: volume raw
[iz][iy][ix], projections [p][v][u]).  `_matvec` is A and `_rmatvec` is A^T,
each one call of the C ABI.  The *_operator_matrix classes
(operators.py:11-89) multiply by an explicit CSR system matrix: the
reference's ./matrixes/A_<basis>_sparse_<sizes>.npz when that file exists,
otherwise the matrix built by the C ABI's exporters (pixel_/
bilinear_/h0_/hl_matrix_generation).  When pylops is importable the classes
derive from pylops.LinearOperator; otherwise from a small base that provides
shape, dtype, matvec, rmatvec, H, dot and @.
"""
import os

import numpy as np

from . import _lib

try:  # pragma: no cover - pylops is optional
    from pylops import LinearOperator as _Base

    _HAVE_PYLOPS = True
except Exception:  # pylops absent: minimal LinearOperator surface
    _HAVE_PYLOPS = False

    class _Base:
        def __init__(self, dtype=np.float64, shape=(0, 0)):
            self.dtype = np.dtype(dtype)
            self.shape = tuple(int(s) for s in shape)

        def matvec(self, x):
            x = np.asarray(x, dtype=self.dtype)
            if x.size != self.shape[1]:
                raise ValueError("matvec expects %d elements, got %d" % (self.shape[1], x.size))
            return self._matvec(x.ravel())

        def rmatvec(self, y):
            y = np.asarray(y, dtype=self.dtype)
            if y.size != self.shape[0]:
                raise ValueError("rmatvec expects %d elements, got %d" % (self.shape[0], y.size))
            return self._rmatvec(y.ravel())

        def dot(self, x):
            return self.matvec(x)

        def __matmul__(self, x):
            return self.matvec(x)

        def __mul__(self, x):
            return self.matvec(x)

        @property
        def H(self):
            return _Adjoint(self)

        T = H


class _Adjoint:
    def __init__(self, op):
        self.op = op
        self.shape = (op.shape[1], op.shape[0])
        self.dtype = op.dtype

    def matvec(self, y):
        return self.op.rmatvec(y)

    def rmatvec(self, x):
        return self.op.matvec(x)

    def __matmul__(self, y):
        return self.matvec(y)

    dot = matvec

    @property
    def H(self):
        return self.op


class _ProjectorOperator(_Base):
    _prefix = None

    def _setup(self, handle):
        if not handle:
            raise _lib.SgdctError("%s_init failed: %s" % (self._prefix, _lib.last_error()))
        self.cpointer = handle
        L = _lib.lib()
        self._fwd = getattr(L, "%s_forward_projection" % self._prefix)
        self._bwd = getattr(L, "%s_backward_projection" % self._prefix)
        self.conv_count = None

    def _matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros((self.shape[0],), dtype=np.float64)
        self.conv_count = _lib.check(self._fwd(_lib.dptr(y), _lib.dptr(x), self.cpointer), "forward projection")
        return y

    def _rmatvec(self, y):
        y = np.ascontiguousarray(y, dtype=np.float64).ravel()
        x = np.zeros((self.shape[1],), dtype=np.float64)
        _lib.check(self._bwd(_lib.dptr(y), _lib.dptr(x), self.cpointer), "back projection")
        return x

    def close(self):
        if getattr(self, "cpointer", None):
            _lib.lib().sgdct_destroy(self.cpointer)
            self.cpointer = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _opt(a):
    if a is None:
        return None, None
    arr = np.ascontiguousarray(a, dtype=np.float64)
    return arr, _lib.dptr(arr)


class _FanOperator(_ProjectorOperator):
    def __init__(self, nx, ny, nnp, nu, dx, dy, du, lsd, lso, angles, dz=None, drho=None,
                 dtype=np.float64):
        super().__init__(dtype=dtype, shape=(nnp * nu, nx * ny))
        self.vol_shape = (nx, ny)
        self.proj_shape = (nnp, nu)
        self.lsd = lsd
        self.lso = lso
        self.dimension = (dx, dy)
        self.blur = du
        self.angles = angles
        if dz is not None and drho is None:  # operators.py:106-107
            raise AssertionError("Error: drho should be provided if dz is provided.")
        ang, ang_p = _opt(angles)
        if ang.size != nnp:
            raise ValueError("angles must have nnp=%d entries" % nnp)
        self._keep = [ang, _opt(dz)[0], _opt(drho)[0]]
        init = getattr(_lib.lib(), "%s_init" % self._prefix)
        self._setup(init(nx, ny, nnp, nu, dx, dy, du, lsd, lso, ang_p, _opt(dz)[1], _opt(drho)[1]))


class A_B0_operator(_FanOperator):
    """Pixel basis (operators.py:91-154 -> layers/A_B0.cu)."""
    _prefix = "b0"


class A_B1_operator(_FanOperator):
    """Bilinear basis (operators.py:156-219 -> layers/A_B1.cu)."""
    _prefix = "b1"


class A_SF_operator(_ProjectorOperator):
    """Separable footprint (operators.py:351-399 -> layers/A_SF.cu)."""
    _prefix = "SF"

    def __init__(self, nx, ny, nnp, nu, dx, dy, du, lsd, lso, angles, dtype=np.float64):
        super().__init__(dtype=dtype, shape=(nnp * nu, nx * ny))
        self.vol_shape = (nx, ny)
        self.proj_shape = (nnp, nu)
        self.lsd = lsd
        self.lso = lso
        self.dimension = (dx, dy)
        self.blur = du
        self.angles = angles
        ang, ang_p = _opt(angles)
        if ang.size != nnp:
            raise ValueError("angles must have nnp=%d entries" % nnp)
        self._keep = [ang]
        self._setup(_lib.lib().SF_init(nx, ny, nnp, nu, dx, dy, du, lsd, lso, ang_p))


class _HexOperator(_ProjectorOperator):
    def __init__(self, nx, ny, nnp, nu, dt, du, lsd, lso, angles, dz=None, drho=None, dtype=np.float64):
        # operators.py:221-284: (nx, ny) are passed to h0_init/hl_init as (nl, nc)
        super().__init__(dtype=dtype, shape=(nnp * nu, nx * ny))
        self.vol_shape = (nx, ny)
        self.proj_shape = (nnp, nu)
        self.lsd = lsd
        self.lso = lso
        self.dimension = dt
        self.blur = du
        self.angles = angles
        if dz is not None and drho is None:
            raise AssertionError("Error: drho should be provided if dz is provided.")
        ang, ang_p = _opt(angles)
        if ang.size != nnp:
            raise ValueError("angles must have nnp=%d entries" % nnp)
        self._keep = [ang, _opt(dz)[0], _opt(drho)[0]]
        init = getattr(_lib.lib(), "%s_init" % self._prefix)
        self._setup(init(nx, ny, nnp, nu, dt, du, lsd, lso, ang_p, _opt(dz)[1], _opt(drho)[1]))


class A_H0_operator(_HexOperator):
    """Hexagon-indicator basis (operators.py:221-284 -> layers/A_H0.cu)."""
    _prefix = "h0"


class A_Hl_operator(_HexOperator):
    """Hexagonal three-direction box-spline basis (operators.py:286-349 -> layers/A_Hl.cu)."""
    _prefix = "hl"


class A_B1_3D_operator(_ProjectorOperator):
    """Cone-beam bilinear footprint (operators.py:401-449 -> layers/A_B1_3D.cu).

    Shape (nnp*nu*nv, nx*ny*nz); the volume is flattened raw as [iz][iy][ix]
    (vol_shape (nz, ny, nx)) and the projections as [p][v][u]; proj_shape
    keeps the reference's (nnp, nu, nv).  A^T is the transpose of A (the
    reference's back-projector returns zeros, A_B1_3D.cu:408, 436)."""
    _prefix = "b1_3d"

    def __init__(self, nx, ny, nz, nnp, nu, nv, dx, dy, dz, du, dv, lsd, lso, angles, dtype=np.float64):
        super().__init__(dtype=dtype, shape=(nnp * nu * nv, nx * ny * nz))
        self.vol_shape = (nz, ny, nx)
        self.proj_shape = (nnp, nu, nv)
        self.lsd = lsd
        self.lso = lso
        self.dimension = (dx, dy, dz)
        self.blur = (du, dv)
        self.angles = angles
        ang, ang_p = _opt(angles)
        if ang.size != nnp:
            raise ValueError("angles must have nnp=%d entries" % nnp)
        self._keep = [ang]
        self._setup(_lib.lib().b1_3d_init(nx, ny, nz, nnp, nu, nv, dx, dy, dz, du, dv, lsd, lso, ang_p))

    def _matvec(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64).ravel()
        y = np.zeros((self.shape[0],), dtype=np.float64)
        rc = self._fwd(_lib.dptr(y), _lib.dptr(x), self.cpointer)
        if rc >= 1 << 63:  # a negative status in the unsigned count (include/sgdct.h)
            _lib.check(rc - (1 << 64), "cone-beam forward projection")
        self.conv_count = rc
        return y


class A_SF_3D_operator(_ProjectorOperator):
    """Cone-beam separable footprints (layers/SF_projector.cu, which the
    reference exposes only as a C++ layer): trap_v selects SF-TT (trapezoid
    in v), else SF-TR (rectangle); same shapes and layouts as
    A_B1_3D_operator, A^T the transpose of A."""
    _prefix = "sf3d"

    def __init__(self, nx, ny, nz, nnp, nu, nv, dx, dy, dz, du, dv, lsd, lso, angles, trap_v=True,
                 dtype=np.float64):
        super().__init__(dtype=dtype, shape=(nnp * nu * nv, nx * ny * nz))
        self.vol_shape = (nz, ny, nx)
        self.proj_shape = (nnp, nu, nv)
        self.lsd = lsd
        self.lso = lso
        self.dimension = (dx, dy, dz)
        self.blur = (du, dv)
        self.angles = angles
        self.trap_v = bool(trap_v)
        ang, ang_p = _opt(angles)
        if ang.size != nnp:
            raise ValueError("angles must have nnp=%d entries" % nnp)
        self._keep = [ang]
        self._setup(_lib.lib().sf3d_init(nx, ny, nz, nnp, nu, nv, dx, dy, dz, du, dv, lsd, lso, ang_p,
                                         int(self.trap_v)))


def exported_matrix(exporter, args, shape):
    """CSR matrix from one of the C ABI's COO exporters (pixel_/bilinear_/
    h0_/hl_matrix_generation, called with `args` before the buffer): a
    counting call sizes the buffer, a second call fills it."""
    import scipy.sparse as sp
    f = getattr(_lib.lib(), exporter)
    n = _lib.check(f(*args, None, 0), exporter)
    buf = np.zeros((max(n, 1), 3))
    n2 = _lib.check(f(*args, _lib.dptr(buf), n), exporter)
    rows, cols = buf[:n2, 0].astype(np.int64), buf[:n2, 1].astype(np.int64)
    return sp.csr_matrix((buf[:n2, 2], (rows, cols)), shape=shape)


class _MatrixOperator(_Base):
    """Explicit-matrix operator (operators.py:11-89): y = A x, x = A^T y with
    a CSR matrix, read from matrix_dir/A_<name>_sparse_<d0>_<d1>_<nnp>_<nu>.npz
    when present (scipy.sparse.load_npz, which loads no pickled objects), else
    exported by the GPU."""

    def _load(self, name, dims, exporter, args, matrix_dir):
        import scipy.sparse as sp
        path = os.path.join(matrix_dir, "A_%s_sparse_%s.npz" % (name, "_".join(str(int(d)) for d in dims)))
        if os.path.exists(path):
            self.A = sp.csr_matrix(sp.load_npz(path))
            self.source = path
        else:
            self.A = exported_matrix(exporter, args, self.shape)
            self.source = exporter
        if self.A.shape != self.shape:
            raise ValueError("%s holds a %s matrix, expected %s" % (self.source, self.A.shape, self.shape))

    def _matvec(self, x):
        return self.A @ np.asarray(x, dtype=np.float64).ravel()

    def _rmatvec(self, y):
        return self.A.T @ np.asarray(y, dtype=np.float64).ravel()


class A_B0_operator_matrix(_MatrixOperator):
    """operators.py:11-27: pixel basis, np views over [0, 2 pi (np-1)/np]."""

    def __init__(self, nx, ny, nnp, nu, dx, dy, du, lsd, lso, dtype=np.float64, matrix_dir="./matrixes"):
        super().__init__(dtype=dtype, shape=(nnp * nu, nx * ny))
        self.vol_shape = (nx, ny)
        self.proj_shape = (nnp, nu)
        self.lsd, self.lso, self.dimension, self.blur = lsd, lso, (dx, dy), du
        self._load("B0", (nx, ny, nnp, nu), "pixel_matrix_generation",
                   (nnp, nu, nx, ny, dx, dy, du, lsd, lso), matrix_dir)


class A_B1_operator_matrix(_MatrixOperator):
    """operators.py:29-45: bilinear basis (rows are the B1 projector's
    footprints; the reference's matrix kernel is one bin off)."""

    def __init__(self, nx, ny, nnp, nu, dx, dy, du, lsd, lso, dtype=np.float64, matrix_dir="./matrixes"):
        super().__init__(dtype=dtype, shape=(nnp * nu, nx * ny))
        self.vol_shape = (nx, ny)
        self.proj_shape = (nnp, nu)
        self.lsd, self.lso, self.dimension, self.blur = lsd, lso, (dx, dy), du
        self._load("B1", (nx, ny, nnp, nu), "bilinear_matrix_generation",
                   (nnp, nu, nx, ny, dx, dy, du, lsd, lso), matrix_dir)


class A_H0_operator_matrix(_MatrixOperator):
    """operators.py:47-63: hexagon-indicator basis on nl x nc lattice points,
    columns il*nc + ic."""

    def __init__(self, nl, nc, nnp, nu, dt, du, lsd, lso, dtype=np.float64, matrix_dir="./matrixes"):
        super().__init__(dtype=dtype, shape=(nnp * nu, nl * nc))
        self.vol_shape = (nl, nc)
        self.proj_shape = (nnp, nu)
        self.lsd, self.lso, self.dimension, self.blur = lsd, lso, dt, du
        self._load("H0", (nl, nc, nnp, nu), "h0_matrix_generation", (nnp, nu, nl, nc, du, dt, lsd, lso), matrix_dir)


class A_Hl_operator_matrix(_MatrixOperator):
    """operators.py:65-81: three-direction box-spline basis on nl x nc
    lattice points, columns il*nc + ic."""

    def __init__(self, nl, nc, nnp, nu, dt, du, lsd, lso, dtype=np.float64, matrix_dir="./matrixes"):
        super().__init__(dtype=dtype, shape=(nnp * nu, nl * nc))
        self.vol_shape = (nl, nc)
        self.proj_shape = (nnp, nu)
        self.lsd, self.lso, self.dimension, self.blur = lsd, lso, dt, du
        self._load("Hl", (nl, nc, nnp, nu), "hl_matrix_generation", (nnp, nu, nl, nc, du, dt, lsd, lso), matrix_dir)
