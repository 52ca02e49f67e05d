"""Cone-beam separable footprints (SF-TR / SF-TT, layers/SF_projector.cu with
the fixes listed in include/sgdct.h) on the GPU against the FP64 oracle
(or_sf3d_*):
  * A and its transpose to <= 1e-5 relative L2 on random volumes and
    projections, both models;
  * adjointness <= 1e-6 at a mid size and at a 256^3 / 360-view /
    512 x 384-detector workload;
  * the host ABI (sf3d_*) against the device path.
"""
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


def _setup(c, oracle, model):
    from pyoracle import Geometry3
    from sgd_amd.torch_ops import Projector
    ang = _angles(c["np_"])
    d = c["d"]
    P = Projector.cone_beam(c["nx"], c["ny"], c["nz"], ang, c["nu"], c["nv"], d, d, d, c["du"], c["dv"], LSD, LSO,
                            model=model)
    G = Geometry3(oracle, c["nx"], c["ny"], c["nz"], ang, c["nu"], c["nv"], d, d, d, c["du"], c["dv"], LSD, LSO)
    return P, G


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("model", ["SF-TR", "SF-TT"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_parity_with_oracle(cuda, oracle_exact, case, model):
    import torch
    print("sf3d parity", case, model, flush=True)
    P, G = _setup(CASES[case], oracle_exact, model)
    tt = model == "SF-TT"
    rng = np.random.default_rng(8)
    x = rng.random(P.img_shape).astype(np.float32)
    y = rng.random(P.sino_shape).astype(np.float32)
    ax = P.forward(torch.tensor(x, device=cuda)).cpu().numpy()
    aty = P.backward(torch.tensor(y, device=cuda)).cpu().numpy()
    e1, e2 = _rel(ax, G.sf_forward(x.astype(np.float64), tt)), _rel(aty, G.sf_backward(y.astype(np.float64), tt))
    print("A rel %.2e  A^T rel %.2e" % (e1, e2))
    assert e1 <= 1e-5 and e2 <= 1e-5


@pytest.mark.parametrize("model", ["SF-TR", "SF-TT"])
def test_adjoint(cuda, model):
    import torch
    from sgd_amd.torch_ops import Projector
    print("sf3d adjoint", model, flush=True)
    for shape in (dict(n=(64, 60, 48), np_=16, nu=128, nv=96, d=0.8, lsd=LSD, lso=LSO),
                  dict(n=(256, 256, 256), np_=360, nu=512, nv=384, d=1.0, lsd=1000.0, lso=500.0)):
        nx, ny, nz = shape["n"]
        P = Projector.cone_beam(nx, ny, nz, _angles(shape["np_"]), shape["nu"], shape["nv"], shape["d"], shape["d"],
                                shape["d"], 1.0, 1.0, shape["lsd"], shape["lso"], model=model)
        g = torch.Generator(device=cuda).manual_seed(3)
        x = torch.rand(P.img_shape, device=cuda, generator=g)
        y = torch.rand(P.sino_shape, device=cuda, generator=g)
        ax, aty = P.forward(x), P.backward(y)
        err = abs(float((ax.double() * y.double()).sum()) - float((x.double() * aty.double()).sum())) / float(
            ax.double().norm() * y.double().norm())
        print("adjoint %s %.2e" % (shape["n"], err))
        assert err <= 1e-6


def test_host_abi(cuda, oracle_exact):
    import torch
    from sgd_amd import _lib
    print("sf3d host abi", flush=True)
    P, G = _setup(CASES["small"], oracle_exact, "SF-TT")
    L = _lib.lib()
    x = np.random.default_rng(12).random(P.img_shape)
    b = np.zeros(P.sino_shape)
    assert L.sf3d_forward_projection(_lib.dptr(b), _lib.dptr(x), P.handle) == 0
    dev = P.forward(torch.tensor(x, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    assert _rel(b, dev) <= 1e-6 and _rel(b, G.sf_forward(x, True)) <= 1e-5
    xb = np.full(P.img_shape, 3.0)
    assert L.sf3d_backward_projection(_lib.dptr(b), _lib.dptr(xb), P.handle) == 0
    assert _rel(xb, G.sf_backward(b, True)) <= 1e-5
    assert L.b1_3d_forward_projection(_lib.dptr(b), _lib.dptr(x), P.handle) >= 2 ** 63  # wrong handle kind
