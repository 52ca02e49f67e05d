"""The bench workload end to end: B1, SF and B0 at 512 x 512 pixels (0.6 mm),
720 views, 768 bins (0.8 mm), lso 500 mm, lsd 1000 mm -- exactly what
bench.py times (sgd_amd.workloads.fan512_inputs: the seed-0 phantom and its
Poisson-noisy log sinogram at 1e5 photons), at the automatic launch split
(24 line segments for A, 24 view groups for A^T).

  * A(phantom) and A^T(noisy sinogram) against the FP64 oracle over all 720
    views: <= 1e-5 relative L2, and no farther from the FP32 oracle (the
    reference's own arithmetic) than that is from the FP64 one;
  * adjoint with Gaussian x, y at this size: <= 1e-6;
  * the split 24 / 24 (the automatic one) equals the unsplit 1 / 1 to one
    FP32 rounding;
  * the reference-layout FP64 host entry point (b1_forward_projection,
    layers/A_B1.cu:605-678) returns the device path's values and the
    reference's evaluation count.
"""
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_EXACT = 1e-5
TOL_ADJ = 1e-6


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.fixture(scope="module")
def fan512(cuda):
    from sgd_amd import workloads as wl
    P = wl.fan512_projector("B1")
    img, sino = wl.fan512_inputs(P)
    return img, sino


def _geoms(basis):
    from pyoracle import Geometry, Oracle
    from sgd_amd import workloads as wl
    g = wl.FAN512
    ang = wl.angles(g["views"])
    args = (g["nx"], g["ny"], ang, g["nu"], g["dx"], g["dy"], g["du"], g["lsd"], g["lso"])
    return Geometry(Oracle("exact"), *args), Geometry(Oracle("ref"), *args)


def _run(P, arr, fwd):
    import torch
    t = torch.tensor(arr, dtype=torch.float32, device=P.device)
    out = P.forward(t) if fwd else P.backward(t)
    torch.cuda.synchronize()
    return out.double().cpu().numpy()


@pytest.mark.parametrize("basis", ["B1", "SF", "B0"])
def test_fan512_against_oracle(basis, fan512):
    from sgd_amd import workloads as wl
    img, sino = fan512
    print("fan512 %s: 512x512 / 720 views / 768 bins, automatic split" % basis, flush=True)
    P = wl.fan512_projector(basis)
    ge, gr = _geoms(basis)
    ya = _run(P, img, True)
    t0 = time.perf_counter()
    ye = ge.forward(basis, img)
    t1 = time.perf_counter()
    yr = gr.forward(basis, img)
    print("  A   gpu-exact %.2e gpu-ref %.2e ref-exact %.2e (oracle %.1f s)"
          % (rel(ya, ye), rel(ya, yr), rel(yr, ye), t1 - t0), flush=True)
    assert rel(ya, ye) <= TOL_EXACT
    assert rel(ya, yr) <= max(TOL_EXACT, 1.5 * rel(yr, ye))
    xa = _run(P, sino, False)
    t0 = time.perf_counter()
    xe = ge.backward(basis, sino)
    t1 = time.perf_counter()
    xr = gr.backward(basis, sino)
    print("  A^T gpu-exact %.2e gpu-ref %.2e ref-exact %.2e (oracle %.1f s)"
          % (rel(xa, xe), rel(xa, xr), rel(xr, xe), t1 - t0), flush=True)
    assert rel(xa, xe) <= TOL_EXACT
    assert rel(xa, xr) <= max(TOL_EXACT, 1.5 * rel(xr, xe))


@pytest.mark.parametrize("basis", ["B1", "SF", "B0"])
def test_fan512_adjoint_and_split(basis, cuda):
    import torch
    from sgd_amd import workloads as wl
    print("fan512 %s: adjoint, split 24/24 vs 1/1" % basis, flush=True)
    P = wl.fan512_projector(basis)
    g = torch.Generator(device=cuda).manual_seed(21)
    x = torch.randn(P.img_shape, device=cuda, generator=g)
    y = torch.randn(P.sino_shape, device=cuda, generator=g)
    Ax, Aty = P.forward(x).double(), P.backward(y).double()
    err = float(abs((Ax * y.double()).sum() - (x.double() * Aty).sum()) / (Ax.norm() * y.double().norm()))
    print("  adjoint %.2e" % err, flush=True)
    assert err <= TOL_ADJ
    P.set_split(24, 24)
    y24, x24 = P.forward(x), P.backward(y)
    P.set_split(1, 1)
    y1, x1 = P.forward(x), P.backward(y)
    torch.cuda.synchronize()
    print("  split 24/24 vs 1/1: A %s, A^T %s" % ("bitwise" if torch.equal(y24, y1) else "rounding",
                                                 "bitwise" if torch.equal(x24, x1) else "rounding"))
    assert torch.allclose(y24, y1, rtol=2.4e-7, atol=1e-30)
    assert torch.allclose(x24, x1, rtol=2.4e-7, atol=1e-30)
    # the automatic split at this size is 24 segments / 24 view groups
    P.set_split(0, 0)
    assert torch.equal(P.forward(x), y24) and torch.equal(P.backward(y), x24)


def test_fan512_host_abi(fan512):
    """b1_forward_projection / b1_backward_projection on the reference's FP64
    host layouts x[ix*ny + iy], b[p*nu + u] at the bench size."""
    from sgd_amd import operators as ops
    from sgd_amd import workloads as wl
    img, sino = fan512
    g = wl.FAN512
    ang = wl.angles(g["views"])
    print("fan512 B1 host ABI", flush=True)
    op = ops.A_B1_operator(g["nx"], g["ny"], g["views"], g["nu"], g["dx"], g["dy"], g["du"], g["lsd"], g["lso"],
                           ang)
    P = wl.fan512_projector("B1")
    b = op.matvec(img.T.ravel()).reshape(g["views"], g["nu"])
    assert np.array_equal(b, _run(P, img, True))
    ge, _ = _geoms("B1")
    ce = ge.eval_count("B1")
    print("  evaluation count %d (oracle %d)" % (op.conv_count, ce))
    assert abs(op.conv_count - ce) <= 1e-6 * ce
    xt = op.rmatvec(sino.ravel()).reshape(g["nx"], g["ny"]).T
    assert np.array_equal(xt, _run(P, sino, False))
