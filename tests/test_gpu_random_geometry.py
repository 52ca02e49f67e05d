"""GPU parity of the fan-beam projectors (B0, B1, SF) on seeded random
geometries, against the FP64 oracle: A and A^T (white-noise inputs) within
1e-5 relative L2, the adjoint within 1e-6.

Each seed draws the image size and aspect (5..80 pixels a side), pixel
sizes (0.3..2.5 mm, dx != dy), bin width, bin count, view count and random
view angles, the source distance and magnification (lsd / lso in 1.2..3),
and for every other seed per-view source offsets drho (+-10 mm).  The image
stays inside 0.8 lso of the rotation centre, so every pixel is in front of
the source.  Continuous random angles and sizes put no pixel corner exactly
on a bin edge, where the reference's corner rule is discontinuous (the
hand-made cases of test_gpu_parity cover those ties)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEEDS = range(16)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def draw(seed):
    rng = np.random.default_rng(1000 + seed)
    nx, ny = (int(v) for v in rng.integers(5, 81, 2))
    dx, dy = (float(v) for v in rng.uniform(0.3, 2.5, 2))
    lso = float(rng.uniform(150.0, 800.0))
    # keep the image's half-diagonal below 0.8 lso
    s = min(1.0, 0.8 * lso / (0.5 * np.hypot(nx * dx, ny * dy)))
    dx, dy = dx * s, dy * s
    lsd = lso * float(rng.uniform(1.2, 3.0))
    du = float(rng.uniform(0.3, 2.5))
    nv = int(rng.integers(3, 41))
    ang = np.sort(rng.uniform(0.0, 2 * np.pi, nv))
    # enough bins to cover most of the image, sometimes fewer (clamped edges)
    span = np.hypot(nx * dx, ny * dy) * lsd / lso / du
    nu = int(max(8, min(200, span * rng.uniform(0.6, 1.3))))
    drho = rng.uniform(-10.0, 10.0, nv) if seed % 2 else None
    return nx, ny, ang, nu, dx, dy, du, lsd, lso, drho


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
@pytest.mark.parametrize("seed", SEEDS)
def test_random_geometry(cuda, oracle_exact, basis, seed):
    import torch
    from pyoracle import Geometry
    from sgd_amd.torch_ops import Projector
    nx, ny, ang, nu, dx, dy, du, lsd, lso, drho = draw(seed)
    print("seed %d %s: %dx%d px %.3fx%.3f mm, %d views, %d bins of %.3f mm, lso %.1f lsd %.1f, drho %s"
          % (seed, basis, nx, ny, dx, dy, len(ang), nu, du, lso, lsd, drho is not None), flush=True)
    P = Projector(basis, nx, ny, ang, nu, dx, dy, du, lsd, lso, drho=drho)
    G = Geometry(oracle_exact, nx, ny, ang, nu, dx, dy, du, lsd, lso, drho=drho)
    rng = np.random.default_rng(seed)
    x = rng.random((ny, nx)).astype(np.float32)
    y = rng.random((len(ang), nu)).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    ax, aty = P.forward(xt), P.backward(yt)
    torch.cuda.synchronize()
    ax_o = G.forward(basis, x.astype(np.float64))
    aty_o = G.backward(basis, y.astype(np.float64))
    e1, e2 = rel(ax.double().cpu().numpy(), ax_o), rel(aty.double().cpu().numpy(), aty_o)
    lhs = float((ax.double() * yt.double()).sum())
    rhs = float((xt.double() * aty.double()).sum())
    ea = abs(lhs - rhs) / (float(ax.double().norm()) * float(yt.double().norm()))
    print("  A %.2e  A^T %.2e  adjoint %.2e" % (e1, e2, ea), flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5, (seed, basis, e1, e2)
    assert ea <= 1e-6, (seed, basis, ea)


def draw_hex(seed):
    rng = np.random.default_rng(2000 + seed)
    nl, nc = (int(v) for v in rng.integers(5, 61, 2))
    dt = float(rng.uniform(0.3, 2.5))
    lso = float(rng.uniform(150.0, 800.0))
    dt = min(dt, 0.8 * lso / (0.5 * np.hypot(nc * dt, nl * dt)) * dt)
    lsd = lso * float(rng.uniform(1.2, 3.0))
    du = float(rng.uniform(0.3, 2.5))
    nv = int(rng.integers(3, 41))
    ang = np.sort(rng.uniform(0.0, 2 * np.pi, nv))
    span = np.hypot(nc * dt, nl * dt) * lsd / lso / du
    nu = int(max(8, min(200, span * rng.uniform(0.6, 1.3))))
    drho = rng.uniform(-10.0, 10.0, nv) if seed % 2 else None
    return nl, nc, ang, nu, dt, du, lsd, lso, drho


@pytest.mark.parametrize("basis", ["H0", "Hl"])
@pytest.mark.parametrize("seed", range(8))
def test_random_hex_geometry(cuda, oracle_exact, basis, seed):
    """The hexagonal lattices the same way (Hl rows dt/2 apart: 2 nl - 1)."""
    import torch
    from pyoracle import HexGeometry
    from sgd_amd.torch_ops import Projector
    nl, nc, ang, nu, dt, du, lsd, lso, drho = draw_hex(seed)
    if basis == "Hl":
        nl = 2 * nl - 1
    print("seed %d %s: %d x %d points, dt %.3f mm, %d views, %d bins of %.3f mm, lso %.1f lsd %.1f, drho %s"
          % (seed, basis, nl, nc, dt, len(ang), nu, du, lso, lsd, drho is not None), flush=True)
    P = Projector.hexagonal(basis, nl, nc, ang, nu, dt, du, lsd, lso, drho=drho)
    G = HexGeometry(oracle_exact, nl, nc, ang, nu, dt, du, lsd, lso, drho=drho)
    rng = np.random.default_rng(seed)
    x = rng.random((nl, nc)).astype(np.float32)
    x[1::2, -1] = 0.0
    y = rng.random((len(ang), nu)).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    ax, aty = P.forward(xt), P.backward(yt)
    torch.cuda.synchronize()
    e1 = rel(ax.double().cpu().numpy(), G.forward(basis, x.astype(np.float64)))
    e2 = rel(aty.double().cpu().numpy(), G.backward(basis, y.astype(np.float64)))
    print("  A %.2e  A^T %.2e" % (e1, e2), flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5, (seed, basis, e1, e2)


def draw_cone(seed):
    rng = np.random.default_rng(3000 + seed)
    nx, ny = (int(v) for v in rng.integers(4, 25, 2))
    nz = int(rng.integers(2, 41))
    d = float(rng.uniform(0.3, 2.0))
    lso = float(rng.uniform(100.0, 400.0))
    d = min(d, 0.6 * lso / np.hypot(nx, ny))
    lsd = lso * float(rng.uniform(1.2, 3.0))
    du, dv = (float(v) for v in rng.uniform(0.3, 2.0, 2))
    nv_ = int(rng.integers(2, 13))
    ang = np.sort(rng.uniform(0.0, 2 * np.pi, nv_))
    mag = lsd / lso
    nu = int(max(8, min(96, np.hypot(nx, ny) * d * mag / du * rng.uniform(0.7, 1.3))))
    nv = int(max(4, min(96, nz * d * mag / dv * rng.uniform(0.7, 1.3))))
    return nx, ny, nz, ang, nu, nv, d, du, dv, lsd, lso


@pytest.mark.parametrize("model", ["B1", "SF-TR", "SF-TT"])
@pytest.mark.parametrize("seed", range(6))
def test_random_cone_geometry(cuda, oracle_exact, model, seed):
    """Cone beam the same way: random volumes (4..24 columns a side, 2..40
    voxels tall), detectors and magnifications."""
    import torch
    from pyoracle import Geometry3
    from sgd_amd.torch_ops import Projector
    nx, ny, nz, ang, nu, nv, d, du, dv, lsd, lso = draw_cone(seed)
    print("seed %d %s: %dx%dx%d voxels of %.3f mm, %d views, %d x %d cells of %.3f x %.3f mm, lso %.1f lsd %.1f"
          % (seed, model, nx, ny, nz, d, len(ang), nu, nv, du, dv, lso, lsd), flush=True)
    P = Projector.cone_beam(nx, ny, nz, ang, nu, nv, d, d, d, du, dv, lsd, lso, model=model)
    G = Geometry3(oracle_exact, nx, ny, nz, ang, nu, nv, d, d, d, du, dv, lsd, lso)
    rng = np.random.default_rng(seed)
    x = rng.random((nz, ny, nx)).astype(np.float32)
    y = rng.random((len(ang), nv, nu)).astype(np.float32)
    xt, yt = torch.from_numpy(x).to(cuda), torch.from_numpy(y).to(cuda)
    ax, aty = P.forward(xt), P.backward(yt)
    torch.cuda.synchronize()
    if model == "B1":
        ax_o, aty_o = G.forward(x.astype(np.float64)), G.backward(y.astype(np.float64))
    else:
        tv = model == "SF-TT"
        ax_o = G.sf_forward(x.astype(np.float64), trap_v=tv)
        aty_o = G.sf_backward(y.astype(np.float64), trap_v=tv)
    e1, e2 = rel(ax.double().cpu().numpy(), ax_o), rel(aty.double().cpu().numpy(), aty_o)
    print("  A %.2e  A^T %.2e" % (e1, e2), flush=True)
    assert e1 <= 1e-5 and e2 <= 1e-5, (seed, model, e1, e2)
