"""sgdct_project_dev_noalloc (partial sums in the library's static device
array) against sgdct_forward_dev / sgdct_backward_dev (partial sums in the
handle's workspace), both called through the C ABI on device buffers.

Calls in a row share the one array: a large call (512 x 512 pixels, 720
views, 768 bins, two images: one image's A partial sums at the automatic
16-segment split take 67.5 MiB, so the call lowers the split to what fits in
32 MiB) is followed by smaller ones on other handles, and every output is
checked against the existing entry point's.  Same split: bit for bit; lowered
split: FP64 rounding of the partial sums, <= 1e-6 relative L2 (the parity
bar for FP32 results is 1e-5).
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _run(L, P, direction, inp, batch, noalloc):
    import torch
    out = torch.full((batch,) + (P.sino_shape if direction == 0 else P.img_shape), float("nan"),
                     dtype=torch.float32, device=inp.device)
    s = ctypes.c_void_p(torch.cuda.current_stream(inp.device).cuda_stream)
    if noalloc:
        rc = L.sgdct_project_dev_noalloc(P.handle, direction, _vp(inp), _vp(out), batch, s)
    elif direction == 0:
        rc = L.sgdct_forward_dev(P.handle, _vp(inp), _vp(out), batch, s)
    else:
        rc = L.sgdct_backward_dev(P.handle, _vp(inp), _vp(out), batch, s)
    torch.cuda.synchronize()
    assert rc == 0, (rc, L.sgdct_last_error().decode())
    return out


def _compare(L, P, batch, seed, name):
    import torch
    g = torch.Generator(device=P.device).manual_seed(seed)
    for direction in (0, 1):
        shape = (batch,) + (P.img_shape if direction == 0 else P.sino_shape)
        inp = torch.randn(shape, generator=g, device=P.device, dtype=torch.float32)
        ref = _run(L, P, direction, inp, batch, noalloc=False)
        got = _run(L, P, direction, inp, batch, noalloc=True)
        assert torch.isfinite(got).all()
        rel = float((got.double() - ref.double()).norm() / ref.double().norm())
        same = bool(torch.equal(got, ref))
        print("%s %s batch %d: rel %.2e, bitwise %s" % (name, "A" if direction == 0 else "A^T", batch, rel, same),
              flush=True)
        assert rel <= 1e-6, (name, direction, rel)


def test_noalloc_matches_workspace_entry_points(cuda):
    from sgd_amd import _lib
    from sgd_amd.torch_ops import Projector
    L = _lib.lib()
    ang = 2 * np.pi * np.arange(720) / 720
    print("no-alloc: B1 512x512 / 720 views / 768 bins, batch 2", flush=True)
    big = Projector("B1", 512, 512, ang, 768, 0.6, 0.6, 0.8, 1000.0, 500.0)
    _compare(L, big, 2, 1, "B1 512^2")
    print("no-alloc: B1 96x80 / 120 views / 160 bins, batch 1", flush=True)
    small = Projector("B1", 96, 80, ang[::6].copy(), 160, 2.0, 2.0, 2.4, 1000.0, 500.3)
    _compare(L, small, 1, 2, "B1 96x80")
    print("no-alloc: Hl 37 x 32 / 90 views / 96 bins, batch 3", flush=True)
    hexp = Projector.hexagonal("Hl", 37, 32, ang[::8].copy(), 96, 2.0, 2.4, 1000.0, 500.3)
    _compare(L, hexp, 3, 3, "Hl 37x32")
    print("no-alloc: B1 512x512 again after the smaller calls", flush=True)
    _compare(L, big, 1, 4, "B1 512^2")


def test_noalloc_rejects_what_it_cannot_serve(cuda):
    """A split fixed by sgdct_split_config whose partial sums of one image
    exceed the 32 MiB array, and cone-beam handles, are errors with a message;
    nothing is launched."""
    import torch
    from sgd_amd import _lib
    from sgd_amd.torch_ops import Projector
    L = _lib.lib()
    ang = 2 * np.pi * np.arange(720) / 720
    P = Projector("B1", 512, 512, ang, 768, 0.6, 0.6, 0.8, 1000.0, 500.0)
    P.set_split(16, 16)
    x = torch.zeros(P.img_shape, device=cuda)
    y = torch.zeros(P.sino_shape, device=cuda)
    s = ctypes.c_void_p(torch.cuda.current_stream(cuda).cuda_stream)
    assert L.sgdct_project_dev_noalloc(P.handle, 0, _vp(x), _vp(y), 1, s) == -1
    assert "scratch" in _lib.last_error()
    assert L.sgdct_project_dev_noalloc(P.handle, 1, _vp(y), _vp(x), 1, s) == 0  # 16 groups of 512^2: 32 MiB fits
    assert L.sgdct_project_dev_noalloc(P.handle, 2, _vp(y), _vp(x), 1, s) == -1
    C = Projector.cone_beam(8, 8, 4, ang[:4].copy(), 16, 8, 1.0, 1.0, 1.0, 2.0, 2.0, 1000.0, 500.0)
    v = torch.zeros(C.img_shape, device=cuda)
    p = torch.zeros(C.sino_shape, device=cuda)
    assert L.sgdct_project_dev_noalloc(C.handle, 0, _vp(v), _vp(p), 1, s) == -4
    assert "cone-beam" in _lib.last_error()
    torch.cuda.synchronize()
