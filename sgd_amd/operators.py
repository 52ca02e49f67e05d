"""pylops-style LinearOperators over the gfx950 projectors.

Mirrors python/operators/operators.py:91-399 of the reference: same class
names, constructor arguments, shapes (nnp*nu, nx*ny), dtype float64, and the
same flattened host layouts (image x[ix*ny + iy], sinogram y[p*nu + u]).
`_matvec` is A and `_rmatvec` is A^T, each one call of the C ABI.  When pylops
is importable the classes derive from pylops.LinearOperator; otherwise from a
small base that provides shape, dtype, matvec, rmatvec, H, dot and @.
"""
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
