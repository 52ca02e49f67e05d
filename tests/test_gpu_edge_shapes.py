"""Degenerate shapes through every projector family on the GPU, against the
FP64 oracle (oracle/sgdct_oracle.c): a single pixel, view and bin; images,
lattices and volumes a few cells wide; one detector bin or row.  Each case
checks A and A^T to <= 1e-5 relative L2 and prints its shape before touching
the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


FAN = [(1, 1, 1, 1), (3, 2, 2, 3), (2, 5, 3, 1), (5, 4, 7, 2)]  # nx, ny, views, bins


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
@pytest.mark.parametrize("shape", FAN, ids=["x".join(map(str, s)) for s in FAN])
def test_fan_degenerate(basis, shape, cuda):
    import torch
    from pyoracle import Geometry, Oracle
    from sgd_amd.torch_ops import Projector
    nx, ny, npv, nu = shape
    print("fan %s nx=%d ny=%d np=%d nu=%d" % (basis, nx, ny, npv, nu), flush=True)
    ang = 2 * np.pi * np.arange(npv) / npv + 0.3
    P = Projector(basis, nx, ny, ang, nu, 2.0, 2.0, 3.0, 1000.0, 500.3)
    G = Geometry(Oracle("exact"), nx, ny, ang, nu, 2.0, 2.0, 3.0, 1000.0, 500.3)
    rng = np.random.default_rng(1)
    x = rng.random((ny, nx))
    y = rng.random((npv, nu))
    ya = P.forward(torch.tensor(x, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    xa = P.backward(torch.tensor(y, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    e1, e2 = rel(ya, G.forward(basis, x)), rel(xa, G.backward(basis, y))
    print("A %.2e  A^T %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL


HEX = [(1, 1, 1, 1), (3, 2, 2, 3), (2, 3, 3, 2)]  # rows, columns, views, bins


@pytest.mark.parametrize("basis", ["H0", "Hl"])
@pytest.mark.parametrize("shape", HEX, ids=["x".join(map(str, s)) for s in HEX])
def test_hex_degenerate(basis, shape, cuda):
    import torch
    from pyoracle import HexGeometry, Oracle
    from sgd_amd.torch_ops import Projector
    nl, nc, npv, nu = shape
    print("hex %s nl=%d nc=%d np=%d nu=%d" % (basis, nl, nc, npv, nu), flush=True)
    ang = 2 * np.pi * np.arange(npv) / npv + 0.3
    P = Projector.hexagonal(basis, nl, nc, ang, nu, 2.0, 3.0, 1000.0, 500.3)
    G = HexGeometry(Oracle("exact"), nl, nc, ang, nu, 2.0, 3.0, 1000.0, 500.3)
    rng = np.random.default_rng(2)
    img = rng.random((nl, nc))
    img[1::2, -1] = 0.0  # odd rows hold nc - 1 points
    y = rng.random((npv, nu))
    ya = P.forward(torch.tensor(img, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    xa = P.backward(torch.tensor(y, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    xe = G.backward(basis, y)
    e1, e2 = rel(ya, G.forward(basis, img)), rel(xa, xe)
    print("A %.2e  A^T %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL


CONE = [(1, 1, 1, 1, 1, 1), (2, 3, 2, 2, 3, 2), (3, 2, 5, 3, 1, 4)]  # nx, ny, nz, views, nu, nv


@pytest.mark.parametrize("model", ["B1", "SF-TR", "SF-TT"])
@pytest.mark.parametrize("shape", CONE, ids=["x".join(map(str, s)) for s in CONE])
def test_cone_degenerate(model, shape, cuda):
    import torch
    from pyoracle import Geometry3, Oracle
    from sgd_amd.torch_ops import Projector
    nx, ny, nz, npv, nu, nv = shape
    print("cone %s nx=%d ny=%d nz=%d np=%d nu=%d nv=%d" % (model, nx, ny, nz, npv, nu, nv), flush=True)
    ang = 2 * np.pi * np.arange(npv) / npv + 0.3
    P = Projector.cone_beam(nx, ny, nz, ang, nu, nv, 1.0, 1.0, 1.0, 1.5, 1.5, 200.0, 100.0, model=model)
    G = Geometry3(Oracle("exact"), nx, ny, nz, ang, nu, nv, 1.0, 1.0, 1.0, 1.5, 1.5, 200.0, 100.0)
    rng = np.random.default_rng(3)
    x = rng.random(P.img_shape)
    y = rng.random(P.sino_shape)
    ya = P.forward(torch.tensor(x, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    xa = P.backward(torch.tensor(y, dtype=torch.float32, device=cuda)).double().cpu().numpy()
    if model == "B1":
        ye, xe = G.forward(x), G.backward(y)
    else:
        tt = model == "SF-TT"
        ye, xe = G.sf_forward(x, tt), G.sf_backward(y, tt)
    e1, e2 = rel(ya, ye), rel(xa, xe)
    print("A %.2e  A^T %.2e" % (e1, e2))
    assert e1 <= TOL and e2 <= TOL
