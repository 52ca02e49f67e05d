"""Pins the oracle's hexagonal projectors (restating layers/A_H0.cu and
layers/A_Hl.cu) against mathematics:

  * H0 vs the FP64 exact strip integral of the hexagon indicator (chord
    through a regular hexagon of circumradius dt, averaged over the bin):
    agreement to the local-parallel model error.  This fixes the H0 scale:
    without the 1/3 (A_H0.cu:245 sums three unit-integral rhombi) the error
    would be 2 (200 %);
  * Hl vs the exact strip integral of the Courant hat max(0, 1 - |x|_hex)
    (the three-direction box spline on the triangular lattice);
  * the FP64 truncated-power box spline (A_Hl.cu:21-46) vs quadrature of
    box_p * box_q * box_r averaged over a window, including a degenerate
    direction (|h| <= 1e-8 dropped);
  * adjointness and the culling fix: every lattice row keeps its footprint.
"""
import numpy as np
import pytest
from scipy import integrate


def hexgeo(o, basis, nl=21, nc=19, np_=24, nu=64, dt=2.0, du=3.2, lso=500.3, extra=()):
    from pyoracle import HexGeometry
    ang = np.concatenate([2 * np.pi * np.arange(np_) / np_ + 0.05, np.asarray(extra, float)])
    if basis == "Hl":
        nl = 2 * nl - 1
    return HexGeometry(o, nl, nc, ang, nu, dt, du, 1000.0, lso)


@pytest.mark.parametrize("basis,tol", [("H0", 6e-4), ("Hl", 4e-4)])
def test_hex_model_vs_exact_strip(oracle_exact, basis, tol):
    g = hexgeo(oracle_exact, basis, extra=(0.0, np.pi / 3, np.pi / 2))
    checked = 0
    for ic, il in ((g.nc // 2, g.nl // 2), (2, 3), (g.nc - 3, g.nl - 4), (7, g.nl // 2 + 1)):
        for p in range(g.np):
            us, ws = g.footprint(basis, p, ic, il, pad=1)
            ex = np.array([g.strip_exact(basis, p, u, ic, il) for u in us])
            if np.linalg.norm(ex) == 0:
                continue
            assert np.linalg.norm(ws - ex) / np.linalg.norm(ex) <= tol, (basis, p, ic, il)
            checked += 1
    assert checked >= 80


def test_hex_exact_strip_mass(oracle_exact):
    """sum_u strip * (bin width at the cell) = area (H0) / volume (Hl) of the basis."""
    for basis, area in (("H0", 1.5 * 3 ** 0.5 * 4.0), ("Hl", 0.5 * 3 ** 0.5 * 4.0)):
        g = hexgeo(oracle_exact, basis, nu=128, du=1.0)
        ic, il = g.nc // 2, g.nl // 2
        cx, cy = g.centre(ic, il, basis)
        for p in range(0, g.np, 5):
            src, dtv, pu = g.view(p)
            tot = 0.0
            for u in range(g.nu):
                v = g.strip_exact(basis, p, u, ic, il)
                if v == 0:
                    continue
                s = dtv + pu * (u + 0.5 - g.nu / 2)
                d = s - src
                depth = ((np.array([cx, cy]) - src) @ d) / np.linalg.norm(d)
                tot += v * np.linalg.norm(pu) * depth / np.linalg.norm(d) * (np.linalg.norm(d) / (d @ (dtv - src) / np.linalg.norm(dtv - src)))
            assert tot == pytest.approx(area, rel=2e-3)


def test_hl_truncated_powers_vs_quadrature(oracle_exact):
    def box3(t, p, q, r):
        f = lambda s: (abs(s) <= p / 2) / p
        g2 = lambda tt: integrate.quad(lambda s: f(s) * ((abs(tt - s) <= q / 2) / q), -p / 2, p / 2,
                                       points=[tt - q / 2, tt + q / 2], limit=100)[0]
        return integrate.quad(lambda s: g2(t - s) * ((abs(s) <= r / 2) / r), -r / 2, r / 2, limit=100)[0]
    p, q = 1.2, 0.7
    r = p - q
    w = 0.4
    for y in (-0.9, -0.3, 0.0, 0.25, 0.8):
        # hl_conv(y, -w, va, vb, vc) is the average of the centred kernel over [y, y + w]
        want = integrate.quad(lambda t: box3(t, p, q, r), y, y + w, limit=50, epsabs=1e-10)[0] / w
        got = oracle_exact.hl_conv(y, -w, p, -q, -r)
        assert got == pytest.approx(want, abs=2e-7)
    # degenerate direction dropped (|h| <= 1e-8): triangle of half-width q averaged over [y, y + w]
    tri = lambda t: max(0.0, 1 - abs(t) / q) / q
    for y in (-0.5, 0.1):
        want = integrate.quad(tri, y, y + w, points=[0.0])[0] / w
        assert oracle_exact.hl_conv(y, -w, q, -q, 1e-9) == pytest.approx(want, abs=1e-9)


@pytest.mark.parametrize("basis", ["H0", "Hl"])
def test_hex_adjoint_and_every_row_projects(oracle_exact, basis):
    g = hexgeo(oracle_exact, basis)
    rng = np.random.default_rng(0)
    x = rng.standard_normal((g.nl, g.nc)) * g.mask()
    y = rng.standard_normal((g.np, g.nu))
    Ax, Aty = g.forward(basis, x), g.backward(basis, y)
    assert abs(np.vdot(Ax, y) - np.vdot(x, Aty)) / (np.linalg.norm(Ax) * np.linalg.norm(y)) <= 1e-13
    assert np.all(Aty[1::2, -1] == 0)
    # per-cell mass is the same for every row (the reference's H0 culling
    # dropped off-centre rows): sum_u A(e_cell) equal across rows within 2 %
    masses = []
    for il in range(0, g.nl, 4):
        us, ws = g.footprint(basis, 0, g.nc // 2, il)
        masses.append(ws.sum())
    assert np.ptp(masses) / np.mean(masses) < 0.1
