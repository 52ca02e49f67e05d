"""Device-resident FP32 projectors on torch tensors, and A as an autograd op.

`Projector` owns one C-ABI handle and runs A / A^T on CUDA(HIP) tensors on
torch's current stream (sgdct_forward_dev / sgdct_backward_dev): images are
(..., ny, nx) and sinograms (..., np, nu), FP32, contiguous.  `project` is A
wrapped as a torch.autograd.Function whose backward is the hand-written A^T
kernel (no autograd tracing through the footprint arithmetic).
"""
import ctypes

import numpy as np
import torch

from . import _lib

BASES = {"B0": "b0", "B1": "b1", "SF": "SF", "H0": "h0", "Hl": "hl"}


class Projector:
    def __init__(self, basis, nx, ny, angles, nu, dx, dy, du, lsd, lso, drho=None, dz=None):
        """Square lattices (B0, B1, SF): nx x ny pixels of dx x dy.  For the
        hexagonal bases use Projector.hexagonal.  `drho` / `dz`: per-view
        radial / axial source offsets (layers/A_B1.hh:24; dz does not change a
        2-D fan projection); SF takes them through sgdct_sf_init_offsets."""
        if basis not in BASES:
            raise ValueError("unknown basis %r" % (basis,))
        if not torch.cuda.is_available():
            raise _lib.SgdctError("sgd_amd projectors need a ROCm GPU (torch.cuda unavailable)")
        self.basis = basis
        self.nx, self.ny, self.nu = int(nx), int(ny), int(nu)
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        self.np = int(ang.size)
        self.angles = ang
        self.geom = dict(dx=float(dx), dy=float(dy), du=float(du), lsd=float(lsd), lso=float(lso))
        L = _lib.lib()
        torch.cuda.init()
        pre = BASES[basis]

        def per_view(a, name):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64)
            if a.size != self.np:
                raise ValueError("%s needs one value per view (%d), got %d" % (name, self.np, a.size))
            return a

        dr, dzv = per_view(drho, "drho"), per_view(dz, "dz")
        self._drho, self._dz = dr, dzv
        p_dr = None if dr is None else _lib.dptr(dr)
        p_dz = None if dzv is None else _lib.dptr(dzv)
        if basis in ("H0", "Hl"):
            # hexagonal: nx = nc columns, ny = nl rows, dx = dt (lattice constant)
            h = getattr(L, pre + "_init")(self.ny, self.nx, self.np, self.nu, dx, du, lsd, lso, _lib.dptr(ang),
                                          p_dz, p_dr)
        elif pre == "SF":
            if dr is None and dzv is None:
                h = L.SF_init(self.nx, self.ny, self.np, self.nu, dx, dy, du, lsd, lso, _lib.dptr(ang))
            else:
                h = L.sgdct_sf_init_offsets(self.nx, self.ny, self.np, self.nu, dx, dy, du, lsd, lso,
                                            _lib.dptr(ang), p_dz, p_dr)
        else:
            h = getattr(L, pre + "_init")(self.nx, self.ny, self.np, self.nu, dx, dy, du, lsd, lso,
                                          _lib.dptr(ang), p_dz, p_dr)
        if not h:
            raise _lib.SgdctError("%s_init failed: %s" % (pre, _lib.last_error()))
        self.handle = h
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._img_shape = (self.ny, self.nx)
        self._sino_shape = (self.np, self.nu)

    @classmethod
    def hexagonal(cls, basis, nl, nc, angles, nu, dt, du, lsd, lso, drho=None, dz=None):
        """H0 / Hl on nl rows x nc columns (images (nl, nc), [il][ic]); odd
        rows hold nc-1 points, their last column is ignored by A and zeroed
        by A^T."""
        if basis not in ("H0", "Hl"):
            raise ValueError("hexagonal bases are H0 and Hl")
        return cls(basis, nc, nl, angles, nu, dt, dt, du, lsd, lso, drho=drho, dz=dz)

    @classmethod
    def cone_beam(cls, nx, ny, nz, angles, nu, nv, dx, dy, dz, du, dv, lsd, lso, model="B1"):
        """Cone-beam projector, model "B1" (b1_3d_init), "SF-TR" or "SF-TT"
        (sf3d_init): volumes (nz, ny, nx), projections (np, nv, nu); A^T is
        the transpose of A."""
        if model not in ("B1", "SF-TR", "SF-TT"):
            raise ValueError("cone-beam models are B1, SF-TR and SF-TT")
        if not torch.cuda.is_available():
            raise _lib.SgdctError("sgd_amd projectors need a ROCm GPU (torch.cuda unavailable)")
        self = cls.__new__(cls)
        self.basis = "B1_3D" if model == "B1" else model
        self.nx, self.ny, self.nz, self.nu, self.nv = int(nx), int(ny), int(nz), int(nu), int(nv)
        ang = np.ascontiguousarray(angles, dtype=np.float64)
        self.np = int(ang.size)
        self.angles = ang
        self.geom = dict(dx=float(dx), dy=float(dy), dz=float(dz), du=float(du), dv=float(dv), lsd=float(lsd),
                         lso=float(lso))
        torch.cuda.init()
        L = _lib.lib()
        if model == "B1":
            h = L.b1_3d_init(self.nx, self.ny, self.nz, self.np, self.nu, self.nv, dx, dy, dz, du, dv, lsd, lso,
                             _lib.dptr(ang))
        else:
            h = L.sf3d_init(self.nx, self.ny, self.nz, self.np, self.nu, self.nv, dx, dy, dz, du, dv, lsd, lso,
                            _lib.dptr(ang), int(model == "SF-TT"))
        if not h:
            raise _lib.SgdctError("cone-beam init failed: %s" % _lib.last_error())
        self.handle = h
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._img_shape = (self.nz, self.ny, self.nx)
        self._sino_shape = (self.np, self.nv, self.nu)
        return self

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().sgdct_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def img_shape(self):
        return self._img_shape

    @property
    def sino_shape(self):
        return self._sino_shape

    def eval_count(self):
        return _lib.check(_lib.lib().sgdct_eval_count(self.handle), "eval_count")

    def set_split(self, fwd=0, bwd=0):
        """Fan-beam launch shape: A line segments (hexagonal handles: segments
        of lattice rows or, for steep strips, of vertical lattice lines), A^T
        view groups (0 = automatic)."""
        _lib.check(_lib.lib().sgdct_split_config(self.handle, int(fwd), int(bwd)), "split_config")

    def _check_in(self, t, tail):
        if not (t.is_cuda and t.dtype == torch.float32):
            raise TypeError("expected a float32 CUDA tensor")
        k = len(tail)
        if tuple(t.shape[-k:]) != tail:
            raise ValueError("trailing shape %s, expected %s" % (tuple(t.shape[-k:]), tail))
        return t.contiguous()

    def forward(self, img, out=None):
        img = self._check_in(img, self.img_shape)
        lead = img.shape[:-len(self.img_shape)]
        batch = int(np.prod(lead)) if len(lead) else 1
        if out is None:
            out = torch.empty(lead + self.sino_shape, device=img.device, dtype=torch.float32)
        stream = torch.cuda.current_stream(img.device).cuda_stream
        _lib.check(_lib.lib().sgdct_forward_dev(self.handle, ctypes.c_void_p(img.data_ptr()),
                                                ctypes.c_void_p(out.data_ptr()), batch,
                                                ctypes.c_void_p(stream)), "forward")
        return out

    def backward(self, sino, out=None):
        sino = self._check_in(sino, self.sino_shape)
        lead = sino.shape[:-len(self.sino_shape)]
        batch = int(np.prod(lead)) if len(lead) else 1
        if out is None:
            out = torch.empty(lead + self.img_shape, device=sino.device, dtype=torch.float32)
        stream = torch.cuda.current_stream(sino.device).cuda_stream
        _lib.check(_lib.lib().sgdct_backward_dev(self.handle, ctypes.c_void_p(sino.data_ptr()),
                                                 ctypes.c_void_p(out.data_ptr()), batch,
                                                 ctypes.c_void_p(stream)), "backward")
        return out

    __call__ = forward


class _ProjectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, img, proj):
        ctx.proj = proj
        return proj.forward(img)

    @staticmethod
    def backward(ctx, grad):
        return ctx.proj.backward(grad.contiguous()), None


class _BackProjectFn(torch.autograd.Function):
    @staticmethod
    def forward(ctx, sino, proj):
        ctx.proj = proj
        return proj.backward(sino)

    @staticmethod
    def backward(ctx, grad):
        return ctx.proj.forward(grad.contiguous()), None


def project(img, proj):
    """A as a differentiable op; d/d img is the A^T kernel."""
    return _ProjectFn.apply(img, proj)


def backproject(sino, proj):
    """A^T as a differentiable op; d/d sino is the A kernel."""
    return _BackProjectFn.apply(sino, proj)
