"""Pins the oracle's whole projectors (restating layers/A_B0.cu, A_B1.cu,
A_SF.cu and utils/geo.cu) against mathematics that does not reuse them:

  * geometry: or_geo_u equals the detector coordinate found by intersecting
    the source->point line with the detector line (numpy solve); 45-degree
    snapping (utils/geo.cu:254-257);
  * footprint model vs FP64 exact strip integrals (average over the bin of the
    line integral of the pixel indicator / bilinear hat along the true fan
    rays): agreement to the local-parallel model error (<= 5e-4 at these
    geometries, <= 1.5e-3 at lso = 30 mm) -- a wrong index, sign, scale or
    bin convention costs O(1);
  * the exact strip integral itself vs an independent fine trapezoid rule of
    numpy chord lengths, and vs the adaptive-Simpson oracle within its
    Richardson error estimate;
  * A / A^T adjointness, 90-degree rotation <-> view shift, analytic
    ellipse sinograms (a flipped detector axis gives 0.65 there);
  * the FP32 reference-mode build stays within 1e-4 of the FP64 build.
"""
import numpy as np
import pytest

from sgd_amd.phantom import analytic_sinogram, ellipse_params, phantom, rasterize


def geo(o, nx=48, np_=32, nu=72, dx=3.0, du=3.6, lsd=1000.0, lso=500.0, extra=(), drho=None, ny=None):
    from pyoracle import Geometry
    ang = np.concatenate([2 * np.pi * np.arange(np_) / np_ + 0.05, np.asarray(extra, dtype=float)])
    return Geometry(o, nx, ny or nx, ang, nu, dx, dx, du, lsd, lso, drho=drho)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_projection_matches_line_intersection(oracle_exact):
    g = geo(oracle_exact, extra=(np.pi / 4, 3 * np.pi / 4, 7 * np.pi / 4), drho=None)
    rng = np.random.default_rng(0)
    for p in range(g.np):
        src, dtv, pu = g.view(p)
        for _ in range(5):
            X = rng.uniform(-60, 60, 2)
            # src + s (X - src) = dtv + t pu
            M = np.array([X - src, -pu]).T
            s, t = np.linalg.solve(M, dtv - src)
            assert g.u_of(p, *X) == pytest.approx(t + g.nu / 2, abs=1e-9)


def test_view_placement_and_snapping(oracle_exact):
    g = geo(oracle_exact, np_=4, extra=(np.pi / 4 + 5e-8, 3 * np.pi / 4 - 5e-8, 1.0))
    src, dtv, pu = g.view(0)  # angle 0.05
    c, s = np.cos(0.05), np.sin(0.05)
    assert np.allclose(src, [-500 * c, -500 * s]) and np.allclose(dtv, [500 * c, 500 * s])
    assert np.allclose(pu, [-3.6 * s, 3.6 * c])
    for p, snapped in ((4, np.pi / 4), (5, 3 * np.pi / 4)):
        src, _, _ = g.view(p)
        assert np.allclose(src, [-500 * np.cos(snapped), -500 * np.sin(snapped)], atol=1e-12, rtol=0)
    src, _, _ = g.view(6)
    assert np.allclose(src, [-500 * np.cos(1.0), -500 * np.sin(1.0)], atol=1e-12, rtol=0)


def chord_np(P, e, x0, x1, y0, y1):
    t = []
    with np.errstate(divide="ignore", invalid="ignore"):
        tx = np.sort([(x0 - P[0]) / e[0], (x1 - P[0]) / e[0]])
        ty = np.sort([(y0 - P[1]) / e[1], (y1 - P[1]) / e[1]])
    lo, hi = max(tx[0], ty[0]), min(tx[1], ty[1])
    return max(hi - lo, 0.0)


def test_exact_strip_vs_independent_quadrature(oracle_exact):
    g = geo(oracle_exact)
    for p, ix, iy in ((3, 24, 24), (17, 5, 40), (29, 40, 3)):
        us, _ = g.footprint("B0", p, ix, iy, pad=1)
        src, dtv, pu = g.view(p)
        cx, cy = (ix - 23.5) * 3.0, (iy - 23.5) * 3.0
        for u in us:
            taus = np.linspace(0, 1, 20001)
            vals = []
            for t in taus:
                q = dtv + pu * (u + t - g.nu / 2)
                e = (q - src) / np.linalg.norm(q - src)
                vals.append(chord_np(src, e, cx - 1.5, cx + 1.5, cy - 1.5, cy + 1.5))
            want = np.trapezoid(vals, taus)
            assert g.strip_exact("B0", p, u, ix, iy) == pytest.approx(want, abs=2e-7)
            v, err = g.strip_simpson("B0", p, u, ix, iy, eps=1e-12, depth=40)
            assert abs(v - g.strip_exact("B0", p, u, ix, iy)) <= err + 1e-9


@pytest.mark.parametrize("basis,tol", [("B0", 1e-3), ("B1", 5e-4), ("SF", 1.5e-3)])
@pytest.mark.parametrize("lso,tolscale", [(500.0, 1.0), (30.0, 8.0)])
def test_model_vs_exact_strip(oracle_exact, basis, tol, lso, tolscale):
    if lso == 30.0:
        g = geo(oracle_exact, nx=64, dx=0.5, nu=160, du=0.4, lsd=60.0, lso=30.0,
                extra=(3 * np.pi / 4, 7 * np.pi / 4))
    else:
        g = geo(oracle_exact, extra=(3 * np.pi / 4, 7 * np.pi / 4))
    strip_basis = "B0" if basis == "SF" else basis  # SF approximates the pixel-indicator strip
    n = g.nx
    checked = 0
    for ix, iy in ((n // 2, n // 2), (n // 4, 3 * n // 4), (1, n - 2), (n - 1, 0)):
        for p in range(g.np):
            us, ws = g.footprint(basis, p, ix, iy, pad=1)
            ex = np.array([g.strip_exact(strip_basis, p, u, ix, iy) for u in us])
            if np.linalg.norm(ex) == 0:
                continue
            assert rel(ws, ex) <= tol * tolscale, (basis, p, ix, iy, rel(ws, ex))
            checked += 1
    assert checked >= 60


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
def test_adjoint_exact(oracle_exact, basis):
    g = geo(oracle_exact, nx=40, ny=36, extra=(np.pi / 4, 3 * np.pi / 4))
    rng = np.random.default_rng(1)
    x = rng.standard_normal((g.ny, g.nx))
    y = rng.standard_normal((g.np, g.nu))
    Ax, Aty = g.forward(basis, x), g.backward(basis, y)
    assert abs(np.vdot(Ax, y) - np.vdot(x, Aty)) / (np.linalg.norm(Ax) * np.linalg.norm(y)) <= 1e-13


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
def test_rotation_by_quarter_turn_shifts_views(oracle_exact, basis):
    g = geo(oracle_exact, np_=32)
    img = phantom(4, g.nx, g.ny, g.dx, g.dy)
    n = g.nx
    # object rotated by +90 degrees about the centre: new(x, y) = old(y, -x)
    iy, ix = np.mgrid[0:n, 0:n]
    rot = img[n - 1 - ix, iy]
    a = g.forward(basis, img)[:32]
    b = g.forward(basis, rot)[:32]
    # the scanner at angle t sees the rotated object as the original at t - 90 deg
    assert rel(b, np.roll(a, 8, axis=0)) <= 1e-12


@pytest.mark.parametrize("basis,tol", [("B0", 2e-2), ("B1", 2e-2), ("SF", 2e-2)])
def test_analytic_ellipse_sinogram(oracle_exact, basis, tol):
    from pyoracle import Geometry
    nx, dx, nu, du = 128, 2.4, 192, 3.2
    ang = 2 * np.pi * np.arange(64) / 64
    g = Geometry(oracle_exact, nx, nx, ang, nu, dx, dx, du, 1000.0, 500.0)
    prm = ellipse_params(3, 10, fov_half=nx * dx / 2, densities=(0.1, 1.0))
    img = rasterize(prm, nx, nx, dx, dx, supersample=8, clip=False)
    an = analytic_sinogram(prm, ang, nu, du, 1000.0, 500.0)
    s = g.forward(basis, img)
    assert rel(s, an) <= tol
    assert rel(s[:, ::-1], an) > 0.3  # a flipped detector axis is far off


@pytest.mark.parametrize("basis", ["B0", "B1", "SF"])
def test_reference_mode_within_fp32_noise(oracle_ref, oracle_exact, basis):
    gr, ge = geo(oracle_ref, nx=64, dx=2.4, nu=96, du=3.2), geo(oracle_exact, nx=64, dx=2.4, nu=96, du=3.2)
    img = phantom(2, 64, 64, 2.4, 2.4)
    assert rel(gr.forward(basis, img), ge.forward(basis, img)) <= 1e-4
    y = ge.forward(basis, img)
    assert rel(gr.backward(basis, y), ge.backward(basis, y)) <= 1e-4


def test_per_view_mass_is_constant(oracle_exact):
    """sum_u A(1)[p, u] * du * (lso/lsd)-free form: for a centred disk phantom the
    bin-integrated sinogram mass (sum_u y_u * width of bin u at the centre)
    equals the disk area for every view."""
    g = geo(oracle_exact, nx=64, dx=2.4, nu=128, du=3.2, np_=24)
    prm = np.array([[0.0, 0.0, 40.0, 40.0, 0.0, 1.0]])
    img = rasterize(prm, 64, 64, 2.4, 2.4, supersample=4, clip=False)
    for basis in ("B0", "B1"):
        s = g.forward(basis, img)
        mass = s.sum(axis=1) * 3.2 * 500.0 / 1000.0
        assert np.ptp(mass) / mass.mean() <= 2e-3
        assert mass.mean() == pytest.approx(img.sum() * 2.4 * 2.4, rel=5e-3)
