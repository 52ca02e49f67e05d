"""FORBILD phantom, line integrals, noise model and image metrics on the CPU
(sgd_amd.forbild, restating python/FORBILD.py, which needs numba, absent
here; its Python is not run).  Pins:
  * the reference notebook's own printed output: the 3000 x 3000 raster of
    [-50, 50)^2 has mean 0.6484057988888887 and maximum 5.85
    (python/main copy 2.ipynb cell 24);
  * exact line integrals against closed forms: an ellipse's chord
    2ab sqrt(a^2 c^2 + b^2 s^2 - p^2) / (a^2 c^2 + b^2 s^2), a circle cut by a
    clip line; and against fine quadrature of the raster along random lines
    of the full phantom;
  * the fan geometry: view 0's central ray is the x axis, and a bin's rays
    sample its detector interval;
  * the noise model: with the counts replaced by their means the noisy
    sinogram becomes -100 ln(mean exp(-L/100)), and over seeds its mean and
    variance follow the Poisson delta-method closed forms;
  * SSIM against a brute-force window-by-window evaluation of its formula,
    PSNR's closed form, the samplers at pixel centres.
"""
import numpy as np
import pytest

from sgd_amd import forbild as fb


def test_tables_and_notebook_raster():
    E, C = fb.phantom_tables()
    assert E.shape == (80, 11) and C.shape == (4, 50)
    assert int(E[:, 6].sum()) == C.shape[1]  # every clip line used once
    xs, ys = fb.eval_grid(3000)
    base = fb.raster_grid(xs[0], ys[:, 0])
    assert abs(base.mean() - 0.6484057988888887) <= 1e-12 and base.max() == 5.85
    x2, y2 = fb.eval_grid(256)
    assert np.array_equal(fb.raster(x2, y2), fb.raster_grid(x2[0], y2[:, 0]))
    # resolution pattern and ear switches change the tables as the reference's
    Er, Cr = fb.phantom_tables(resolution=True, ear=True)
    assert Er.shape[0] == 160 and Cr.shape[1] == 50
    En, Cn = fb.phantom_tables(resolution=False, ear=False)
    assert En.shape[0] == 17 and Cn.shape[1] == 12 and int(En[:, 6].sum()) == 12


def _one(x0, y0, a, b, ang, f=1.0, clips=()):
    phi = np.radians(ang)
    row = [x0, y0, a, b, ang, f, len(clips), np.cos(phi) / a, np.sin(phi) / a, -np.sin(phi) / b, np.cos(phi) / b]
    C = np.zeros((4, max(len(clips), 1)))
    for k, (d, al) in enumerate(clips):
        C[:, k] = (d, al, np.cos(np.radians(al)), np.sin(np.radians(al)))
    return np.array([row]), C[:, :len(clips)] if clips else np.zeros((4, 0))


def test_ellipse_chord_closed_form():
    rng = np.random.default_rng(0)
    x0, y0, a, b, ang = 3.0, -2.0, 7.0, 4.0, 35.0
    tab = _one(x0, y0, a, b, ang, f=1.3)
    th = rng.uniform(0, 2 * np.pi, 200)
    s = rng.uniform(-8, 8, 200) + x0 * np.cos(th) + y0 * np.sin(th)
    L = fb.line_integrals(th, s, tab)
    # in the ellipse's frame: normal angle th - ang, offset p
    t = th - np.radians(ang)
    p = s - (x0 * np.cos(th) + y0 * np.sin(th))
    r2 = a * a * np.cos(t) ** 2 + b * b * np.sin(t) ** 2
    chord = np.where(r2 > p * p, 2 * a * b * np.sqrt(np.maximum(r2 - p * p, 0)) / r2, 0.0)
    assert np.allclose(L, 1.3 * chord, rtol=1e-12, atol=1e-12)
    assert (chord > 0).sum() > 100


def test_clipped_circle():
    # circle of radius 5 at the origin, points with x < 2 kept (clip angle 0)
    tab = _one(0.0, 0.0, 5.0, 5.0, 0.0, clips=[(2.0, 0.0)])
    # horizontal lines y = s (normal angle 90 deg): chord from -sqrt(25-s^2) to min(2, sqrt(25-s^2))
    s = np.linspace(-4.9, 4.9, 50)
    L = fb.line_integrals(np.full_like(s, np.pi / 2), s, tab)
    h = np.sqrt(25 - s * s)
    assert np.allclose(L, np.minimum(2.0, h) + h, rtol=1e-9, atol=1e-9)
    # vertical lines x = s (normal angle 0, parallel to the clip line: turned by 1e-10)
    s = np.array([-3.0, 0.0, 1.9, 2.1, 4.0])
    L = fb.line_integrals(np.zeros_like(s), s, tab)
    full = 2 * np.sqrt(25 - s * s)
    assert np.allclose(L, np.where(s < 2.0, full, 0.0), atol=1e-6)


def test_line_integrals_match_raster_quadrature():
    """Full phantom: midpoint quadrature of the raster along 24 random lines
    (step 2.5 um) within the step times the number of edge crossings."""
    rng = np.random.default_rng(2)
    th = rng.uniform(0, np.pi, 24)
    s = rng.uniform(-40, 40, 24)
    L = fb.line_integrals(th, s)
    h = 0.0025
    t = np.arange(-60, 60, h) + h / 2
    for k in range(24):
        x = s[k] * np.cos(th[k]) - t * np.sin(th[k])
        y = s[k] * np.sin(th[k]) + t * np.cos(th[k])
        f = fb.raster(x, y)
        q = f.sum() * h
        jumps = np.abs(np.diff(f)).sum()  # total variation bounds the quadrature error: step x jumps
        assert abs(q - L[k]) <= h * jumps + 1e-9, (k, q, L[k])
    assert np.max(L) > 50


def test_fan_geometry_and_noise_model():
    E, C = fb.phantom_tables()
    # view 0, the two bins around the detector centre: their inner rays meet
    # at the x axis, which the horizontal line y = 0 integrates
    sino = fb.fan_sinogram(1, 4, 0.5, 312.5, 156.25, samples=1, mode="mean")
    assert sino.shape == (1, 4)
    Lc = fb.line_integrals(np.array([np.pi / 2]), np.array([0.0]))[0]
    # a single sample per bin sits at the bin centre u = +-0.25, +-0.75
    for iu, u in enumerate([-0.75, -0.25, 0.25, 0.75]):
        th = np.pi / 2 + np.arctan2(u, 312.5)
        sv = u * 156.25 / np.hypot(312.5, u)
        assert abs(sino[0, iu] - fb.line_integrals(np.array([th]), np.array([sv]))[0]) < 1e-12
    assert abs(fb.line_integrals(np.array([np.pi / 2 + 1e-12]), np.array([0.0]))[0] - Lc) < 1e-6
    # expected counts: -100 ln(mean exp(-L/100))
    args = (3, 6, 2.0, 312.5, 156.25)
    Lbin = fb.fan_sinogram(*args, samples=200, mode="mean")
    ex = fb.fan_sinogram(*args, samples=200, mode="expected")
    lme = fb.fan_sinogram(*args, samples=200, mode="log_mean_exp")
    assert np.all(ex <= Lbin + 1e-12) and np.all(lme >= Lbin - 1e-12)  # Jensen, both directions
    # Poisson: over seeds, mean and variance of b by the delta method
    draws = np.stack([fb.fan_sinogram(*args, samples=200, mode="poisson", seed=k) for k in range(300)])
    lam_total = 200 * 100.0 * np.exp(-ex / 100.0)  # sum of the per-ray means
    var = 100.0 ** 2 / lam_total
    z = (draws.mean(0) - ex) / np.sqrt(var / draws.shape[0])
    assert np.all(np.abs(z) < 5.0)
    ratio = draws.var(0, ddof=1) / var
    assert np.all((ratio > 0.7) & (ratio < 1.4))


def test_ssim_psnr_and_samplers():
    from numpy.lib.stride_tricks import sliding_window_view
    rng = np.random.default_rng(4)
    X = rng.random((14, 17))
    Y = X + 0.2 * rng.standard_normal(X.shape)
    R = 2.0
    # brute force: symmetric extension, 7x7 windows, sample (co)variances
    Xp, Yp = np.pad(X, 3, mode="symmetric"), np.pad(Y, 3, mode="symmetric")
    wx, wy = sliding_window_view(Xp, (7, 7)), sliding_window_view(Yp, (7, 7))
    mx, my = wx.mean((-1, -2)), wy.mean((-1, -2))
    vx = wx.var((-1, -2), ddof=1)
    vy = wy.var((-1, -2), ddof=1)
    vxy = ((wx - mx[..., None, None]) * (wy - my[..., None, None])).sum((-1, -2)) / 48.0
    c1, c2 = (0.01 * R) ** 2, (0.03 * R) ** 2
    S = ((2 * mx * my + c1) * (2 * vxy + c2)) / ((mx ** 2 + my ** 2 + c1) * (vx + vy + c2))
    assert abs(fb.ssim(Y, X, R) - S[3:-3, 3:-3].mean()) < 1e-12
    assert fb.ssim(X, X, R) == pytest.approx(1.0, abs=1e-12)
    assert fb.psnr(X + 0.1, X, 2.0) == pytest.approx(10 * np.log10(4.0 / 0.01), rel=1e-12)
    img = rng.random((6, 5))  # [ix, iy]
    nx, ny, dx, dy = 6, 5, 0.5, 0.7
    xc = (np.arange(nx) - (nx - 1) / 2) * dx
    yc = (np.arange(ny) - (ny - 1) / 2) * dy
    X2, Y2 = np.meshgrid(xc, yc, indexing="ij")
    assert np.allclose(fb.sample_nearest(img, nx, ny, dx, dy, X2, Y2), img)
    assert np.allclose(fb.sample_bilinear(img, nx, ny, dx, dy, X2, Y2), img)
    mid = fb.sample_bilinear(img, nx, ny, dx, dy, (xc[1] + xc[2]) / 2, yc[3])
    assert mid == pytest.approx(0.5 * (img[1, 3] + img[2, 3]))


def test_notebook_curve_fixture():
    """The committed fixture holds the notebook's 32 printed iterations per
    basis (2..64), PSNR rising from ~15 dB."""
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "forbild_cgls_curves.json")))
    for b in ("B0", "B1"):
        assert [r["iters"] for r in g[b]] == list(range(2, 65, 2))
        assert 15.0 < g[b][0]["psnr"] < 15.5 and max(r["psnr"] for r in g[b]) > 26.0
