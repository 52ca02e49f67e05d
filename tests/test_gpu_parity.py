"""GPU parity of the B0 / B1 / SF projectors against the CPU oracle.

The oracle (oracle/sgdct_oracle.c) restates the reference's footprint formulas;
its FP64 build ("exact") evaluates them without rounding noise, its FP32 build
("ref") reproduces the reference's own FP32 arithmetic.  Bars:
  * GPU vs FP64 oracle: relative L2 <= 1e-5 (A and A^T)
  * GPU vs FP32 oracle: no farther than the FP32 oracle is from the FP64 one
    (times 1.5), i.e. within the reference's own rounding noise
  * adjoint |<Ax,y> - <x,A^T y>| / (|Ax||y|) <= 1e-6
Each case prints its name before touching the GPU.

Hand-made geometries that contain axis-aligned views use an odd pixel count
and a source distance like 500.3 mm.  With an even count the pixel corners on
the central axis project exactly onto the central bin edge at 0/90/180/270
degrees (and with "round" magnifications other corners land exactly on other
edges); there the reference's B0 corner rule floor(u) (layers/A_B0.cu:194-195)
is discontinuous, its footprint reaches past the corner range, and any two
implementations -- the reference's own FP32 build included -- may keep or drop
a bin with a nonzero value.  The real workloads (fan256 here) are tested
at their exact numbers, where such ties weigh ~4e-7.
"""
import numpy as np
import pytest

from sgd_amd.phantom import phantom

pytestmark = pytest.mark.gpu

TOL_EXACT = 1e-5
TOL_ADJ = 1e-6


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


CASES = {
    # name: (nx, ny, angles, nu, dx, dy, du, lsd, lso, drho)
    "tiny": (64, 64, 2 * np.pi * np.arange(90) / 90, 96, 2.4, 2.4, 3.2, 1000.0, 500.0, None),
    "diagonals": (47, 47, np.array([0, np.pi / 4, np.pi / 2, 3 * np.pi / 4, np.pi, 5 * np.pi / 4,
                                     3 * np.pi / 2, 7 * np.pi / 4, 0.3, 2.0, -0.7, 4.0 + 1e-9]),
                  80, 3.0, 3.0, 3.5, 1000.0, 500.3, None),
    "nonsquare": (96, 64, 2 * np.pi * np.arange(60) / 60 + 0.01, 128, 1.5, 2.0, 2.2, 800.0, 350.0, None),
    "offdetector": (63, 63, 2 * np.pi * np.arange(40) / 40, 24, 2.4, 2.4, 3.0, 1000.0, 500.3, None),
    "drho": (63, 63, 2 * np.pi * np.arange(36) / 36, 96, 2.4, 2.4, 3.2, 1000.0, 500.0,
             5.0 * np.sin(np.arange(36))),
    "tight": (64, 64, 2 * np.pi * np.arange(45) / 45, 160, 0.5, 0.5, 0.4, 60.0, 30.07, None),
    # 4 mm pixels on 0.25 mm bins: a 16-pixel tile spans ~512 bins, past the
    # 128-bin LDS window, so A^T reads its bin records from global memory
    "widewindow": (48, 48, 2 * np.pi * np.arange(30) / 30 + 0.05, 1024, 4.0, 4.0, 0.25, 1000.0, 500.3, None),
    "fan256": (256, 256, 2 * np.pi * np.arange(360) / 360, 512, 1.0, 1.0, 1.0, 1000.0, 500.0, None),
}


def make(basis, case, cuda):
    from pyoracle import Geometry, Oracle
    from sgd_amd.torch_ops import Projector

    nx, ny, ang, nu, dx, dy, du, lsd, lso, drho = CASES[case]
    print("case %s basis %s nx=%d ny=%d np=%d nu=%d" % (case, basis, nx, ny, len(ang), nu), flush=True)
    P = Projector(basis, nx, ny, ang, nu, dx, dy, du, lsd, lso, drho=drho)
    ge = Geometry(Oracle("exact"), nx, ny, ang, nu, dx, dy, du, lsd, lso, drho=drho)
    gr = Geometry(Oracle("ref"), nx, ny, ang, nu, dx, dy, du, lsd, lso, drho=drho)
    return P, ge, gr


def gpu_fwd(P, img):
    import torch
    x = torch.tensor(img, dtype=torch.float32, device="cuda")
    y = P.forward(x)
    torch.cuda.synchronize()
    return y.double().cpu().numpy()


def gpu_bwd(P, sino):
    import torch
    y = torch.tensor(sino, dtype=torch.float32, device="cuda")
    x = P.backward(y)
    torch.cuda.synchronize()
    return x.double().cpu().numpy()


def test_sanity_library_loads_on_gpu(cuda):
    from sgd_amd import lib
    assert lib() is not None


@pytest.mark.parametrize("basis", ["B1", "B0", "SF"])
@pytest.mark.parametrize("case", ["tiny", "diagonals", "nonsquare", "offdetector", "drho", "tight", "widewindow"])
def test_forward_backward_parity(basis, case, cuda):
    P, ge, gr = make(basis, case, cuda)
    img = phantom(7, P.nx, P.ny, ge.dx, ge.dy)
    ya = gpu_fwd(P, img)
    ye, yr = ge.forward(basis, img), gr.forward(basis, img)
    e_exact, e_ref, e_rr = rel(ya, ye), rel(ya, yr), rel(yr, ye)
    print("A   gpu-exact %.2e gpu-ref %.2e ref-exact %.2e" % (e_exact, e_ref, e_rr))
    assert e_exact <= TOL_EXACT
    assert e_ref <= max(TOL_EXACT, 1.5 * e_rr)

    sino = ye + 0.1 * np.random.default_rng(3).standard_normal(ye.shape)
    xa = gpu_bwd(P, sino)
    xe, xr = ge.backward(basis, sino), gr.backward(basis, sino)
    e_exact, e_ref, e_rr = rel(xa, xe), rel(xa, xr), rel(xr, xe)
    print("A^T gpu-exact %.2e gpu-ref %.2e ref-exact %.2e" % (e_exact, e_ref, e_rr))
    assert e_exact <= TOL_EXACT
    assert e_ref <= max(TOL_EXACT, 1.5 * e_rr)


@pytest.mark.parametrize("basis", ["B1", "B0", "SF"])
def test_adjoint(basis, cuda):
    import torch
    P, ge, gr = make(basis, "tiny", cuda)
    rng = np.random.default_rng(11)
    x = rng.standard_normal((P.ny, P.nx))
    y = rng.standard_normal((P.np, P.nu))
    Ax = gpu_fwd(P, x)
    Aty = gpu_bwd(P, y)
    err = abs(np.vdot(Ax, y) - np.vdot(x, Aty)) / (np.linalg.norm(Ax) * np.linalg.norm(y))
    print("adjoint %s %.2e" % (basis, err))
    assert err <= TOL_ADJ


@pytest.mark.parametrize("basis", ["B0", "B1"])
def test_fan256_full_size(basis, cuda):
    P, ge, gr = make(basis, "fan256", cuda)
    img = phantom(0, P.nx, P.ny, ge.dx, ge.dy)
    ya = gpu_fwd(P, img)
    ye = ge.forward(basis, img)
    print("fan256 %s A gpu-exact %.2e" % (basis, rel(ya, ye)))
    assert rel(ya, ye) <= TOL_EXACT
    xa = gpu_bwd(P, ye)
    xe = ge.backward(basis, ye)
    print("fan256 %s A^T gpu-exact %.2e" % (basis, rel(xa, xe)))
    assert rel(xa, xe) <= TOL_EXACT


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
def test_split_launches_agree(basis, cuda):
    """A with 1..32 line segments and A^T with 1..32 view groups (FP64
    partial sums added in a fixed order) give the unsplit result to one FP32
    rounding, and each split is deterministic."""
    import torch
    P, ge, gr = make(basis, "fan256", cuda)
    g = torch.Generator(device="cuda").manual_seed(3)
    x = torch.rand(P.img_shape, device="cuda", generator=g)
    y = torch.rand(P.sino_shape, device="cuda", generator=g)
    P.set_split(1, 1)
    y1, x1 = P.forward(x), P.backward(y)
    for fs, bs in ((2, 3), (7, 5), (16, 16), (24, 32), (0, 0)):
        P.set_split(fs, bs)
        ys, xs = P.forward(x), P.backward(y)
        torch.cuda.synchronize()
        assert torch.allclose(ys, y1, rtol=2.4e-7, atol=1e-30), (fs, bs)
        assert torch.allclose(xs, x1, rtol=2.4e-7, atol=1e-30), (fs, bs)
        assert torch.equal(P.forward(x), ys) and torch.equal(P.backward(y), xs), (fs, bs)


def test_batch_and_determinism(cuda):
    import torch
    P, ge, gr = make("B1", "tiny", cuda)
    imgs = np.stack([phantom(s, P.nx, P.ny, ge.dx, ge.dy) for s in range(3)])
    x = torch.tensor(imgs, dtype=torch.float32, device="cuda")
    yb = P.forward(x)
    y1 = torch.stack([P.forward(x[i]) for i in range(3)])
    torch.cuda.synchronize()
    assert torch.equal(yb, y1)
    assert torch.equal(P.forward(x), yb)
    xb = P.backward(yb)
    x1 = torch.stack([P.backward(yb[i]) for i in range(3)])
    assert torch.equal(xb, x1)
    assert torch.equal(P.backward(yb), xb)


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
def test_host_abi_matches_device_path(basis, cuda):
    """The reference-layout FP64 entry points (x[ix*ny+iy], b[p*nu+u])."""
    from sgd_amd import operators as ops
    nx, ny, ang, nu, dx, dy, du, lsd, lso, _ = CASES["nonsquare"]
    print("host abi %s" % basis, flush=True)
    cls = {"B0": ops.A_B0_operator, "B1": ops.A_B1_operator, "SF": ops.A_SF_operator}[basis]
    op = cls(nx, ny, len(ang), nu, dx, dy, du, lsd, lso, ang)
    img = phantom(5, nx, ny, dx, dy)            # [iy, ix]
    b = op.matvec(img.T.ravel())                # host layout x[ix*ny + iy]
    P, ge, gr = make(basis, "nonsquare", cuda)
    ye = ge.forward(basis, img)
    assert rel(b.reshape(len(ang), nu), ye) <= TOL_EXACT
    assert op.conv_count == ge.eval_count(basis) or abs(op.conv_count - ge.eval_count(basis)) <= 1e-6 * op.conv_count
    xt = op.rmatvec(ye.ravel()).reshape(nx, ny).T
    assert rel(xt, ge.backward(basis, ye)) <= TOL_EXACT
