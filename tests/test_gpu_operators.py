"""The pylops operator surface beyond the fan-beam projectors:

  * A_B1_3D_operator (operators.py:401-449) and A_SF_3D_operator through the
    C ABI's FP64 host entry points equal the FP32 device path
    (Projector.cone_beam) on the same inputs, with the reference's raw volume
    and [p][v][u] projection layouts; adjoint through the ABI <= 1e-6;
  * A_{B0,B1,H0,Hl}_operator_matrix (operators.py:11-89) with no saved
    matrix: the GPU exporter's CSR matrix times dx dy equals the projector
    (matvec and rmatvec, <= 1e-5 relative: B0 / B1 / Hl agree to ~4e-8, H0 at
    views where a hexagon edge is parallel to the detector to ~4e-6, as the H0
    projector does with the oracle there), and for the hexagonal exporters
    (h0_/hl_matrix_generation, the dead hl_matrix_generate of
    layers/A_Hl.cu:370-451) the FP64 oracle's projector (<= 1e-5); the
    counting call returns the filled count; the triplets come in the
    reference's order.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


CONE = dict(nx=24, ny=20, nz=12, nnp=18, nu=40, nv=16, dx=1.0, dy=1.1, dz=0.9, du=1.6, dv=1.5, lsd=400.0, lso=200.0)


@pytest.mark.parametrize("model", ["B1", "SF-TT", "SF-TR"])
def test_cone_operators_match_device_path(model, cuda):
    import torch
    from sgd_amd import operators as ops
    from sgd_amd.torch_ops import Projector
    c = CONE
    ang = 2 * np.pi * np.arange(c["nnp"]) / c["nnp"] + 0.1
    print("cone operator %s" % model, flush=True)
    geo = (c["nx"], c["ny"], c["nz"], c["nnp"], c["nu"], c["nv"], c["dx"], c["dy"], c["dz"], c["du"], c["dv"],
           c["lsd"], c["lso"], ang)
    op = ops.A_B1_3D_operator(*geo) if model == "B1" else ops.A_SF_3D_operator(*geo, trap_v=(model == "SF-TT"))
    assert op.shape == (c["nnp"] * c["nu"] * c["nv"], c["nx"] * c["ny"] * c["nz"])
    assert op.vol_shape == (c["nz"], c["ny"], c["nx"]) and op.proj_shape == (c["nnp"], c["nu"], c["nv"])
    P = Projector.cone_beam(c["nx"], c["ny"], c["nz"], ang, c["nu"], c["nv"], c["dx"], c["dy"], c["dz"], c["du"],
                            c["dv"], c["lsd"], c["lso"], model=model)
    rng = np.random.default_rng(3)
    x = rng.random(op.shape[1]).astype(np.float32).astype(np.float64)
    y = rng.standard_normal(op.shape[0]).astype(np.float32).astype(np.float64)
    b = op.matvec(x)
    bd = P.forward(torch.tensor(x.reshape(P.img_shape), dtype=torch.float32, device=cuda))
    assert rel(b, bd.double().cpu().numpy().ravel()) <= 1e-6
    if model == "B1":
        assert op.conv_count > 0
    xt = op.rmatvec(y)
    xd = P.backward(torch.tensor(y.reshape(P.sino_shape), dtype=torch.float32, device=cuda))
    assert rel(xt, xd.double().cpu().numpy().ravel()) <= 1e-6
    err = abs(np.vdot(b, y) - np.vdot(x, xt)) / (np.linalg.norm(b) * np.linalg.norm(y))
    print("  adjoint through the ABI %.2e" % err)
    assert err <= 1e-6


FAN = dict(nx=21, ny=17, nnp=24, nu=40, dx=2.0, dy=1.8, du=2.4, lsd=400.0, lso=200.3)
HEX = dict(nl=19, nc=15, nnp=24, nu=40, dt=1.4, du=2.2, lsd=400.0, lso=200.3)


@pytest.mark.parametrize("basis", ["B0", "B1", "H0", "Hl"])
def test_matrix_operators_equal_projectors(basis, cuda, tmp_path):
    import torch
    from sgd_amd import _lib
    from sgd_amd import operators as ops
    from sgd_amd.torch_ops import Projector
    print("matrix operator %s" % basis, flush=True)
    rng = np.random.default_rng(7)
    if basis in ("B0", "B1"):
        f = FAN
        ang = 2 * np.pi * np.arange(f["nnp"]) / f["nnp"]
        cls = ops.A_B0_operator_matrix if basis == "B0" else ops.A_B1_operator_matrix
        op = cls(f["nx"], f["ny"], f["nnp"], f["nu"], f["dx"], f["dy"], f["du"], f["lsd"], f["lso"],
                 matrix_dir=str(tmp_path))
        P = Projector(basis, f["nx"], f["ny"], ang, f["nu"], f["dx"], f["dy"], f["du"], f["lsd"], f["lso"])
        area = f["dx"] * f["dy"]
        img = rng.random((f["ny"], f["nx"]))
        xm = img.T.ravel()  # columns ix * ny + iy
        exporter, args = ("pixel_matrix_generation" if basis == "B0" else "bilinear_matrix_generation",
                          (f["nnp"], f["nu"], f["nx"], f["ny"], f["dx"], f["dy"], f["du"], f["lsd"], f["lso"]))
    else:
        h = HEX
        ang = 2 * np.pi * np.arange(h["nnp"]) / h["nnp"]
        cls = ops.A_H0_operator_matrix if basis == "H0" else ops.A_Hl_operator_matrix
        op = cls(h["nl"], h["nc"], h["nnp"], h["nu"], h["dt"], h["du"], h["lsd"], h["lso"], matrix_dir=str(tmp_path))
        P = Projector.hexagonal(basis, h["nl"], h["nc"], ang, h["nu"], h["dt"], h["du"], h["lsd"], h["lso"])
        area = 3 ** 0.5 * h["dt"] * (1.5 if basis == "H0" else 0.5) * h["dt"]
        img = rng.random((h["nl"], h["nc"]))
        img[1::2, -1] = 0.0
        xm = img.ravel()  # columns il * nc + ic
        exporter, args = ("h0_matrix_generation" if basis == "H0" else "hl_matrix_generation",
                          (h["nnp"], h["nu"], h["nl"], h["nc"], h["du"], h["dt"], h["lsd"], h["lso"]))
    assert op.source == exporter
    y_dev = P.forward(torch.tensor(img, dtype=torch.float32, device=cuda)).double().cpu().numpy().ravel()
    y_mat = op.matvec(xm) * area
    print("  A: matrix vs projector %.2e (nnz %d)" % (rel(y_mat, y_dev), op.A.nnz))
    if basis in ("H0", "Hl"):
        from pyoracle import HexGeometry, Oracle
        ge = HexGeometry(Oracle("exact"), h["nl"], h["nc"], ang, h["nu"], h["dt"], h["du"], h["lsd"], h["lso"])
        ye = ge.forward(basis, img).ravel()
        print("  A: matrix vs FP64 oracle %.2e" % rel(y_mat, ye))
        assert rel(y_mat, ye) <= 1e-5
    assert rel(y_mat, y_dev) <= 1e-5
    sino = rng.standard_normal(op.shape[0])
    x_dev = P.backward(torch.tensor(sino.reshape(P.sino_shape), dtype=torch.float32, device=cuda))
    x_dev = x_dev.double().cpu().numpy()
    x_mat = op.rmatvec(sino) * area
    x_mat = x_mat.reshape(P.nx, P.ny).T if basis in ("B0", "B1") else x_mat.reshape(P.ny, P.nx)
    assert rel(x_mat, x_dev) <= 1e-5
    # counting call == filled count; triplets in the order (row, column)
    L = _lib.lib()
    n = getattr(L, exporter)(*args, None, 0)
    buf = np.zeros((n, 3))
    assert getattr(L, exporter)(*args, _lib.dptr(buf), n) == n == op.A.nnz
    key = buf[:, 0].astype(np.int64) * op.shape[1] + buf[:, 1].astype(np.int64)
    assert np.all(np.diff(key) > 0) and np.all(buf[:, 2] > 0)
