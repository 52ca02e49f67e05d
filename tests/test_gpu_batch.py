"""Batched persistent launch of uneven fan slices (sgdct_batch_*), its hipGraph
replay and the autotuner.  Bars:
  * batched A / A^T equal the per-projector kernels to the last FP32 bit
    (every work item runs the same device function; the per-projector
    launches may split a bin's walk or the views into FP64 partial sums, whose
    different association can move a result by one FP32 rounding), for every
    launch shape the autotuner may pick;
  * graph replays (native capture, and torch.cuda.CUDAGraph capture of the
    same entry point) equal eager results on every replay, with eager work
    and synchronisations between replays and a changed input;
  * the autotuner writes, reuses and applies its cache.
The per-projector kernels themselves are pinned against the oracle in
test_gpu_parity.py.
"""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(nslice=24, seed=3):
    from sgd_amd.batch import SliceBatch, make_projectors, uneven_slices
    specs = uneven_slices(nslice, seed=seed, sizes=(32, 48, 64, 96, 128), views=(40, 160), fov=96.0)
    return SliceBatch(make_projectors(specs))


def _inputs(B, cuda, seed=0):
    import torch
    g = torch.Generator(device=cuda).manual_seed(seed)
    img = torch.rand(B.img_off[-1], device=cuda, generator=g)
    sino = torch.rand(B.sino_off[-1], device=cuda, generator=g)
    return img, sino


def _same(a, b):
    """Equal up to one FP32 rounding per element."""
    import torch
    return bool(torch.allclose(a, b, rtol=2.4e-7, atol=1e-30))


def test_batched_equals_per_slice(cuda):
    import torch
    print("batch vs per-slice", flush=True)
    B = _batch()
    img, sino = _inputs(B, cuda)
    ref_s = B.forward_per_slice(img)
    ref_i = B.backward_per_slice(sino)
    for fb, bb, order in [(8, 4, 0), (2, 1, 1), (16, 8, 0)]:
        B.set_config(fb, bb, order)
        s = B.forward(img)
        i = B.backward(sino)
        torch.cuda.synchronize()
        assert _same(s, ref_s), (fb, bb, order)
        assert _same(i, ref_i), (fb, bb, order)
    # blocks per CU 0: one launch per slice inside the batch call, bit for bit
    # the per-projector launches; mixed with a persistent direction
    for fb, bb in [(0, 0), (0, 8), (8, 0)]:
        B.set_config(fb, bb, 0)
        s = B.forward(img)
        i = B.backward(sino)
        torch.cuda.synchronize()
        assert torch.equal(s, ref_s) if fb == 0 else _same(s, ref_s), (fb, bb)
        assert torch.equal(i, ref_i) if bb == 0 else _same(i, ref_i), (fb, bb)
    kinds = {p.basis for p in B.projectors}
    assert kinds == {"B0", "B1", "SF"}


def test_native_graph_replay(cuda):
    import torch
    print("batch native graph", flush=True)
    B = _batch(seed=5)
    img, sino = _inputs(B, cuda, seed=1)
    out_s, out_i = B.empty_sinos(), B.empty_images()
    stream = torch.cuda.Stream()
    B.set_config(0, 8, 0)  # A captured as one launch per slice, A^T persistent
    with torch.cuda.stream(stream):
        for rep in range(3):
            B.forward(img, out=out_s, graph=True)
            B.backward(sino, out=out_i, graph=True)
            stream.synchronize()
            eager_s = B.forward(img)
            eager_i = B.backward(sino)
            stream.synchronize()
            assert torch.equal(out_s, eager_s), rep
            assert torch.equal(out_i, eager_i), rep
            img.mul_(1.25).add_(0.01)  # eager work between replays, new input values
            sino.mul_(0.8)
            stream.synchronize()


def test_torch_graph_capture(cuda):
    import torch
    print("batch torch graph", flush=True)
    B = _batch(seed=7)
    img, sino = _inputs(B, cuda, seed=2)
    out_s, out_i = B.empty_sinos(), B.empty_images()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        B.forward(img, out=out_s)  # warm up outside the capture
        B.backward(sino, out=out_i)
    torch.cuda.current_stream().wait_stream(s)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        B.forward(img, out=out_s)
        B.backward(sino, out=out_i)
    for rep in range(3):
        img.mul_(0.9)
        sino.add_(0.5)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert _same(out_s, B.forward_per_slice(img)), rep
        assert _same(out_i, B.backward_per_slice(sino)), rep


def test_autotuner_cache(cuda, tmp_path):
    import torch
    print("batch autotuner", flush=True)
    B = _batch(nslice=16, seed=9)
    path = str(tmp_path / "tune.json")
    rec = B.tune(cache_path=path, reps=1)
    assert not rec["cached"]
    data = json.load(open(path))
    assert B.signature() in data
    B.set_config(2, 1, 1)
    rec2 = B.tune(cache_path=path)
    assert rec2["cached"] and B.config == (rec["fwd_blocks_per_cu"], rec["bwd_blocks_per_cu"], rec["order"])
    img, sino = _inputs(B, cuda, seed=4)
    assert _same(B.forward(img), B.forward_per_slice(img))
    assert _same(B.backward(sino), B.backward_per_slice(sino))
    print("tuned", {k: rec[k] for k in ("fwd_blocks_per_cu", "bwd_blocks_per_cu", "order", "fwd_ms", "bwd_ms")})


def test_degenerate_slices_in_a_batch(cuda):
    """Slices of one pixel, view or bin, and a few cells wide, mixed with
    ordinary ones: the persistent launch equals the per-projector kernels."""
    import torch
    from sgd_amd.batch import SliceBatch, make_projectors, uneven_slices
    print("batch with degenerate slices", flush=True)
    specs = uneven_slices(6, seed=5, sizes=(32, 48), views=(20, 40), fov=96.0)
    for basis, nx, ny, npv, nu in [("B1", 1, 1, 1, 1), ("SF", 3, 2, 2, 3), ("B0", 2, 5, 3, 1), ("B1", 5, 4, 7, 2)]:
        specs.append(dict(basis=basis, nx=nx, ny=ny, np=npv, nu=nu, dx=2.0, dy=2.0, du=3.0, lsd=1000.0, lso=500.3,
                          angles=2 * np.pi * np.arange(npv) / npv + 0.3))
    B = SliceBatch(make_projectors(specs))
    img, sino = _inputs(B, cuda, seed=4)
    ref_s, ref_i = B.forward_per_slice(img), B.backward_per_slice(sino)
    for fb, bb, order in [(8, 4, 0), (2, 1, 1)]:
        B.set_config(fb, bb, order)
        s, i = B.forward(img), B.backward(sino)
        torch.cuda.synchronize()
        assert _same(s, ref_s), (fb, bb, order)
        assert _same(i, ref_i), (fb, bb, order)
