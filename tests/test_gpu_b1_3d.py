"""Cone-beam B1 projector (layers/A_B1_3D.cu) on the GPU against the FP64
oracle (oracle/sgdct_oracle.c, or_b1_3d_*):

  * A and its transpose A^T to <= 1e-5 relative L2 on random volumes and
    projections (the reference's own back-projector returns zeros, so A^T is
    pinned by the oracle's transpose and by the adjoint test);
  * adjointness <= 1e-6 at a mid size and at the 256^3 / 360-view /
    512 x 384-detector workload;
  * the reference-named host ABI (FP64 arrays, raw [iz][iy][ix] volume,
    [p][v][u] projections) against the device path, and its 64-bit count.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LSD, LSO = 200.0, 100.0
CASES = {
    "small": dict(nx=24, ny=20, nz=18, np_=12, nu=48, nv=40, d=1.0, du=1.2, dv=1.1),
    "mid": dict(nx=64, ny=60, nz=48, np_=16, nu=128, nv=96, d=0.8, du=1.0, dv=1.0),
    # the largest column and detector height the kernels stage at once (512
    # voxels, 512 rows)
    "tall": dict(nx=8, ny=6, nz=512, np_=4, nu=24, nv=512, d=0.2, du=0.5, dv=0.4),
    # past them: columns of 600 voxels (two voxel chunks) on 640 detector rows
    # (two row windows of the axial transpose, two row groups in-plane)
    "taller": dict(nx=6, ny=5, nz=600, np_=3, nu=16, nv=640, d=0.2, du=0.5, dv=0.4),
    # a detector three times wider than the volume: the strips of the outer
    # bins miss it, so their walks have no line
    "wide": dict(nx=16, ny=14, nz=10, np_=8, nu=96, nv=24, d=1.0, du=1.0, dv=1.0),
}


def _angles(np_):
    return 2 * np.pi * np.arange(np_) / np_ + 0.05


def _setup(c, oracle):
    from pyoracle import Geometry3
    from sgd_amd.torch_ops import Projector
    ang = _angles(c["np_"])
    d = c["d"]
    P = Projector.cone_beam(c["nx"], c["ny"], c["nz"], ang, c["nu"], c["nv"], d, d, d, c["du"], c["dv"], LSD, LSO)
    G = Geometry3(oracle, c["nx"], c["ny"], c["nz"], ang, c["nu"], c["nv"], d, d, d, c["du"], c["dv"], LSD, LSO)
    return P, G


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("case", sorted(CASES))
def test_parity_with_oracle(cuda, oracle_exact, case):
    import torch
    print("b1_3d parity", case, flush=True)
    P, G = _setup(CASES[case], oracle_exact)
    rng = np.random.default_rng(7)
    x = rng.random(P.img_shape).astype(np.float32)
    y = rng.random(P.sino_shape).astype(np.float32)
    ax = P.forward(torch.tensor(x, device=cuda)).cpu().numpy()
    aty = P.backward(torch.tensor(y, device=cuda)).cpu().numpy()
    ex, ey = G.forward(x.astype(np.float64)), G.backward(y.astype(np.float64))
    e1, e2 = _rel(ax, ex), _rel(aty, ey)
    print("A rel %.2e  A^T rel %.2e" % (e1, e2))
    assert e1 <= 1e-5 and e2 <= 1e-5


def _adjoint(P, cuda, seed):
    import torch
    g = torch.Generator(device=cuda).manual_seed(seed)
    x = torch.rand(P.img_shape, device=cuda, generator=g)
    y = torch.rand(P.sino_shape, device=cuda, generator=g)
    ax, aty = P.forward(x), P.backward(y)
    lhs = float((ax.double() * y.double()).sum())
    rhs = float((x.double() * aty.double()).sum())
    return abs(lhs - rhs) / float(ax.double().norm() * y.double().norm())


def test_adjoint_mid_and_full(cuda):
    from sgd_amd.torch_ops import Projector
    c = CASES["mid"]
    P = Projector.cone_beam(c["nx"], c["ny"], c["nz"], _angles(c["np_"]), c["nu"], c["nv"], c["d"], c["d"], c["d"],
                            c["du"], c["dv"], LSD, LSO)
    err = _adjoint(P, cuda, 1)
    print("adjoint mid %.2e" % err, flush=True)
    assert err <= 1e-6
    print("adjoint 256^3 x 360 views, 512 x 384 detector", flush=True)
    P = Projector.cone_beam(256, 256, 256, _angles(360), 512, 384, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 500.0)
    err = _adjoint(P, cuda, 2)
    print("adjoint full %.2e" % err)
    assert err <= 1e-6


def test_host_abi(cuda, oracle_exact):
    import torch
    from sgd_amd import _lib
    print("b1_3d host abi", flush=True)
    c = CASES["small"]
    P, G = _setup(c, oracle_exact)
    L = _lib.lib()
    rng = np.random.default_rng(11)
    x = rng.random(P.img_shape)
    b = np.zeros(P.sino_shape)
    cnt = L.b1_3d_forward_projection(_lib.dptr(b), _lib.dptr(x), P.handle)
    assert 0 < cnt < 2 ** 63, _lib.last_error()
    assert cnt == L.sgdct_eval_count(P.handle)
    dev = P.forward(torch.tensor(x, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    assert _rel(b, dev) <= 1e-6
    assert _rel(b, G.forward(x)) <= 1e-5
    xb = np.full(P.img_shape, 7.0)
    assert L.b1_3d_backward_projection(_lib.dptr(b), _lib.dptr(xb), P.handle) == 0
    assert _rel(xb, G.backward(b)) <= 1e-5
    shape = (ctypes.c_int * 6)()
    assert L.sgdct_shape3(P.handle, shape) == 5
    assert list(shape) == [c["nx"], c["ny"], c["nz"], c["np_"], c["nu"], c["nv"]]
    # a fan-beam entry point refuses a cone-beam handle
    assert L.b1_forward_projection(_lib.dptr(b), _lib.dptr(x), P.handle) < 0
