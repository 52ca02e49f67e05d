"""Many fan-beam slices in one persistent launch (sgdct_batch_*), hipGraph
replay, and the autotuner that picks the launch shape.

The reference projects one slice per call with one kernel launch per view
(layers/A_B1.cu:512-541, A_B0.cu:417-446, A_SF.cu:362-392).  `SliceBatch`
packs any mix of B0 / B1 / SF `Projector`s -- uneven image sizes, view counts
and bin counts -- into one work queue per direction that a single persistent
kernel drains (see sgd_amd/csrc/batch.hip).  Images (and sinograms) of all
slices live back to back in one device buffer at `img_off` (`sino_off`).

`tune()` times the candidate launch shapes -- persistent blocks per CU for A
and A^T, or one launch per slice through each projector's own split kernels
(blocks per CU 0), and the slice order of the persistent queue (decreasing
per-item cost, or as given) -- with
HIP events on the current stream and caches the winner per device and
workload signature in a JSON file ($SGDCT_TUNE_CACHE, default
~/.cache/sgd_amd/autotune.json).
"""
import ctypes
import hashlib
import json
import os

import numpy as np
import torch

from . import _lib
from .torch_ops import Projector

# candidate persistent blocks per CU; 0 = one launch per slice
FWD_BLOCKS = (0, 2, 4, 8, 16)
BWD_BLOCKS = (0, 1, 2, 4, 8)
ORDERS = (0, 1)


def default_cache_path():
    return os.environ.get("SGDCT_TUNE_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "sgd_amd", "autotune.json"))


def uneven_slices(nslice=256, seed=0, sizes=(128, 256, 384, 512, 768), views=(180, 1440),
                  bases=("B0", "B1", "SF"), fov=256.0, lsd=1000.0, lso=500.0, helical=False):
    """Specs of fan slices of uneven size: image edge uniform over `sizes`
    (field of view `fov` mm), view count uniform over [views[0], views[1]]
    with a per-slice start angle, 1.5x as many bins as the image edge,
    detector pitch covering the field of view, basis uniform over `bases`.
    With `helical`, every view also carries the source offsets of a slice
    rebinned from a helical scan (sgd_amd.workloads.helical_offsets: axial dz
    over one pitch, radial drho of a flying focal spot)."""
    from .workloads import helical_offsets
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(nslice):
        n = int(rng.choice(sizes))
        nv = int(rng.integers(views[0], views[1] + 1))
        nu = int(round(1.5 * n))
        r = fov / np.sqrt(2.0)
        du = 1.02 * 2.0 * r * lsd / (lso - r) / nu
        start = float(rng.uniform(0, 2 * np.pi))
        specs.append(dict(basis=str(rng.choice(bases)), nx=n, ny=n, np=nv, nu=nu, dx=fov / n, dy=fov / n, du=du,
                          lsd=lsd, lso=lso, angles=start + 2 * np.pi * np.arange(nv) / nv))
        if helical:
            specs[-1]["dz"], specs[-1]["drho"] = helical_offsets(nv, seed=1000003 * seed + i)
    return specs


def make_projectors(specs):
    return [Projector(s["basis"], s["nx"], s["ny"], s["angles"], s["nu"], s["dx"], s["dy"], s["du"], s["lsd"],
                      s["lso"], drho=s.get("drho"), dz=s.get("dz")) for s in specs]


class SliceBatch:
    def __init__(self, projectors):
        self.projectors = list(projectors)
        for p in self.projectors:
            if p.basis not in ("B0", "B1", "SF"):
                raise ValueError("batched slices must be B0, B1 or SF projectors")
        n = len(self.projectors)
        L = _lib.lib()
        arr = (ctypes.c_void_p * n)(*[p.handle for p in self.projectors])
        h = L.sgdct_batch_create(arr, n)
        if not h:
            raise _lib.SgdctError("sgdct_batch_create failed: %s" % _lib.last_error())
        self.handle = h
        io = (ctypes.c_longlong * (n + 1))()
        so = (ctypes.c_longlong * (n + 1))()
        _lib.check(L.sgdct_batch_layout(h, io, so), "batch layout")
        self.img_off, self.sino_off = list(io), list(so)
        self.device = self.projectors[0].device
        self.config = (8, 8, 0)

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().sgdct_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- packing -------------------------------------------------------------
    def empty_images(self):
        return torch.empty(self.img_off[-1], device=self.device)

    def empty_sinos(self):
        return torch.empty(self.sino_off[-1], device=self.device)

    def pack_images(self, imgs):
        return torch.cat([t.reshape(-1).float() for t in imgs]).contiguous()

    def pack_sinos(self, sinos):
        return torch.cat([t.reshape(-1).float() for t in sinos]).contiguous()

    def image(self, packed, i):
        p = self.projectors[i]
        return packed[self.img_off[i]:self.img_off[i + 1]].view(p.ny, p.nx)

    def sino(self, packed, i):
        p = self.projectors[i]
        return packed[self.sino_off[i]:self.sino_off[i + 1]].view(p.np, p.nu)

    # -- launches ------------------------------------------------------------
    def set_config(self, fwd_blocks_per_cu, bwd_blocks_per_cu, order):
        _lib.check(_lib.lib().sgdct_batch_config(self.handle, int(fwd_blocks_per_cu), int(bwd_blocks_per_cu),
                                                 int(order)), "batch config")
        self.config = (int(fwd_blocks_per_cu), int(bwd_blocks_per_cu), int(order))

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def forward(self, img, out=None, graph=False):
        assert img.is_cuda and img.dtype == torch.float32 and img.numel() == self.img_off[-1]
        img = img.contiguous()
        out = self.empty_sinos() if out is None else out
        L = _lib.lib()
        if graph:
            rc = L.sgdct_batch_graph_launch(self.handle, 0, ctypes.c_void_p(img.data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()), self._stream())
        else:
            rc = L.sgdct_batch_forward(self.handle, ctypes.c_void_p(img.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                       self._stream())
        _lib.check(rc, "batch forward")
        return out

    def backward(self, sino, out=None, graph=False):
        assert sino.is_cuda and sino.dtype == torch.float32 and sino.numel() == self.sino_off[-1]
        sino = sino.contiguous()
        out = self.empty_images() if out is None else out
        L = _lib.lib()
        if graph:
            rc = L.sgdct_batch_graph_launch(self.handle, 1, ctypes.c_void_p(sino.data_ptr()),
                                            ctypes.c_void_p(out.data_ptr()), self._stream())
        else:
            rc = L.sgdct_batch_backward(self.handle, ctypes.c_void_p(sino.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                        self._stream())
        _lib.check(rc, "batch backward")
        return out

    def forward_per_slice(self, img, out=None):
        """The same A with one launch per slice (the per-projector kernels)."""
        out = self.empty_sinos() if out is None else out
        for i, p in enumerate(self.projectors):
            p.forward(self.image(img, i), out=self.sino(out, i))
        return out

    def backward_per_slice(self, sino, out=None):
        out = self.empty_images() if out is None else out
        for i, p in enumerate(self.projectors):
            p.backward(self.sino(sino, i), out=self.image(out, i))
        return out

    # -- autotuner -----------------------------------------------------------
    def signature(self):
        desc = [(p.basis, p.nx, p.ny, p.np, p.nu) for p in self.projectors]
        name = torch.cuda.get_device_name(self.device)
        return name + ":" + hashlib.sha1(json.dumps(desc).encode()).hexdigest()[:16]

    def _time(self, fn, reps):
        fn()
        times = []
        for _ in range(reps):
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record()
            fn()
            b.record()
            b.synchronize()
            times.append(a.elapsed_time(b))
        return float(np.median(times))

    def tune(self, img=None, sino=None, cache_path=None, reps=3, force=False):
        """Pick (A blocks/CU, A^T blocks/CU, slice order) by timing every
        candidate; returns the chosen record (also stored in the cache)."""
        path = cache_path or default_cache_path()
        key = self.signature()
        cache = {}
        if os.path.exists(path):
            with open(path) as f:
                cache = json.load(f)
        if key in cache and not force:
            rec = cache[key]
            self.set_config(rec["fwd_blocks_per_cu"], rec["bwd_blocks_per_cu"], rec["order"])
            return dict(rec, cached=True)
        img = self.empty_images().uniform_() if img is None else img
        sino = self.empty_sinos().uniform_() if sino is None else sino
        out_s, out_i = self.empty_sinos(), self.empty_images()
        best = None
        trials = []
        for order in ORDERS:
            f_ms = {}
            for fb in FWD_BLOCKS:
                self.set_config(fb, BWD_BLOCKS[0], order)
                f_ms[fb] = self._time(lambda: self.forward(img, out=out_s), reps)
            b_ms = {}
            fb_best = min(f_ms, key=f_ms.get)
            for bb in BWD_BLOCKS:
                self.set_config(fb_best, bb, order)
                b_ms[bb] = self._time(lambda: self.backward(sino, out=out_i), reps)
            bb_best = min(b_ms, key=b_ms.get)
            total = f_ms[fb_best] + b_ms[bb_best]
            trials.append(dict(order=order, fwd_ms=f_ms, bwd_ms=b_ms))
            if best is None or total < best[0]:
                best = (total, fb_best, bb_best, order, f_ms[fb_best], b_ms[bb_best])
        _, fb, bb, order, fms, bms = best
        self.set_config(fb, bb, order)
        rec = dict(fwd_blocks_per_cu=fb, bwd_blocks_per_cu=bb, order=order, fwd_ms=fms, bwd_ms=bms,
                   trials=[{k: (v if not isinstance(v, dict) else {str(a): b for a, b in v.items()})
                            for k, v in t.items()} for t in trials])
        cache[key] = rec
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(cache, f, indent=1)
        os.replace(tmp, path)
        return dict(rec, cached=False)
