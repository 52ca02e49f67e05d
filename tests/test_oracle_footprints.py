"""Pins the oracle's 1-D footprints (restating layers/spline_conv.hh) against
independent mathematics, never against a retyped copy of the same formula:

  * B1: S(x) = int_0^x tri_a * tri_b, checked against (i) numerical
    quadrature of the convolution of two unit-integral hats and (ii) the
    box-spline truncated-power form
        S(x) = 1/2 - [(a+b-x)_+^4 - 2(a-x)_+^4 - 2(b-x)_+^4 + (a-b-x)_+^4] / (24 a^2 b^2)
    evaluated in 50-digit arithmetic (mpmath);
  * B1 bin average linear_convolution_1d(x1*sgn, x2, a, b) == mean of the
    kernel over [ta, tb] for every sign pattern, and the 1tri1box branch;
  * B0: the 3-direction box spline == quadrature of the average of the
    trapezoid box_p * box_q over a window of width w;
  * unit integrals and symmetry.
A dropped term, a wrong coefficient or a swapped branch in the restatement
moves these by far more than the tolerances used.
"""
import mpmath
import numpy as np
import pytest
from scipy import integrate


def tri(t, h):
    return np.maximum(0.0, 1.0 - np.abs(t) / h) / h


def k_b1(t, a, b):
    """(tri_a * tri_b)(t) by quadrature."""
    f = lambda s: tri(s, a) * tri(t - s, b)
    pts = sorted({-a, 0.0, a, t - b, t, t + b})
    return integrate.quad(f, -a - b - 1, a + b + 1, points=pts, limit=200, epsabs=1e-14)[0]


def s_b1_quad(x, a, b):
    pts = [p for p in (b, a - b, a, a + b, abs(a - b)) if 0 < p < x]
    return integrate.quad(lambda t: k_b1(t, a, b), 0, x, points=pts or None, limit=200, epsabs=1e-13)[0]


def s_b1_tp(x, a, b):
    mpmath.mp.dps = 50
    x, a, b = mpmath.mpf(x), mpmath.mpf(a), mpmath.mpf(b)
    p = lambda z: z ** 4 if z > 0 else mpmath.mpf(0)
    v = mpmath.mpf(1) / 2 - (p(a + b - x) - 2 * p(a - x) - 2 * p(b - x) + p(a - b - x)) / (24 * a * a * b * b)
    return float(v)


AB = [(1.0, 0.3), (1.0, 0.5), (1.0, 0.62), (0.8, 0.8), (0.6, 0.05), (0.6, 0.0123), (0.9, 0.45)]


@pytest.mark.parametrize("a,b", AB)
def test_b1_antiderivative_vs_truncated_powers(oracle_exact, oracle_ref, a, b):
    xs = np.linspace(0.0, a + b + 0.2, 97)
    for x in xs:
        want = s_b1_tp(x, a, b)
        assert abs(oracle_exact.s2tri(a, b, x) - want) <= 1e-11
        assert abs(oracle_ref.s2tri(a, b, x) - want) <= 2e-5


@pytest.mark.parametrize("a,b", [(1.0, 0.3), (0.8, 0.55), (0.6, 0.05)])
def test_b1_antiderivative_vs_quadrature(oracle_exact, a, b):
    for x in (0.1 * b, 0.5 * b, b, 0.5 * (a + b) * 0.9, a - 0.5 * b, a, a + 0.5 * b, a + b):
        assert abs(oracle_exact.s2tri(a, b, x) - s_b1_quad(x, a, b)) <= 2e-9


@pytest.mark.parametrize("a,b", AB)
def test_b1_unit_integral(oracle_exact, a, b):
    assert oracle_exact.s2tri(a, b, a + b) == pytest.approx(0.5, abs=1e-13)
    assert oracle_exact.s2tri(a, b, 10 * (a + b)) == pytest.approx(0.5, abs=1e-13)
    assert oracle_exact.s2tri(a, b, 0.0) == pytest.approx(0.0, abs=1e-15)


def b1_ref_call(o, ta, tb, a, b):
    """The reference's call pattern (layers/A_B1.cu:192-197)."""
    x1, x2 = abs(ta), abs(tb)
    sgn = 1.0 if ta * tb > 0 else -1.0
    return o.b1_conv(x1 * sgn, x2, a, b)


@pytest.mark.parametrize("a,b", [(1.0, 0.3), (0.7, 0.6), (0.6, 0.0)])
def test_b1_bin_average(oracle_exact, a, b):
    rng = np.random.default_rng(0)
    for _ in range(40):
        ta = rng.uniform(-2, 2)
        tb = ta + rng.uniform(0.05, 0.8)
        if b > 0:
            quad = integrate.quad(lambda t: k_b1(t, a, b), ta, tb, limit=200, epsabs=1e-12)[0] / (tb - ta)
            sodd = lambda t: np.sign(t) * s_b1_tp(abs(t), a, b)
            want = (sodd(tb) - sodd(ta)) / (tb - ta)
            assert abs(quad - want) <= 2e-7
        else:
            want = integrate.quad(lambda t: tri(t, a), ta, tb, points=[0.0], epsabs=1e-13)[0] / (tb - ta)
        got = b1_ref_call(oracle_exact, ta, tb, a, b)
        assert abs(got - want) <= 1e-10 * max(1.0, abs(want))


def trap(t, p, q):
    return np.clip(np.minimum((p + q) / 2 - np.abs(t), q), 0, None) / (p * q)


@pytest.mark.parametrize("p,q,w", [(1.0, 0.4, 0.5), (0.7, 0.7, 0.3), (0.6, 0.02, 0.9), (0.5, 0.3, 1.4)])
def test_b0_three_direction_spline(oracle_exact, oracle_ref, p, q, w):
    for y in np.linspace(0, (p + q + w) / 2 + 0.1, 41):
        want = integrate.quad(lambda t: trap(t, p, q), y - w / 2, y + w / 2,
                              points=[-(p + q) / 2, -(p - q) / 2, (p - q) / 2, (p + q) / 2], limit=100,
                              epsabs=1e-14)[0] / w
        # linear_convolution_1d re-centres by (v1+v2+v3)/2 (layers/A_B0.cu:24)
        shift = (p + q + w) / 2
        assert abs(oracle_exact.b0_conv(y + shift, p, q, w) - want) <= 1e-10
        assert abs(oracle_ref.b0_conv(y + shift, p, q, w) - want) <= 2e-5


def test_b0_two_direction_limit(oracle_exact):
    # smallest direction <= 1e-7 -> univariate3dirboxspline_fast2 (box_p * box_w)
    p, w = 0.9, 0.4
    for y in np.linspace(0, 0.8, 17):
        want = integrate.quad(lambda t: (np.abs(t) <= p / 2) / p, y - w / 2, y + w / 2, points=[-p / 2, p / 2])[0] / w
        assert abs(oracle_exact.box3(y, p, 1e-9, w) - want) <= 1e-12


@pytest.mark.parametrize("p,q,w", [(1.0, 0.4, 0.5), (0.6, 0.02, 0.9)])
def test_b0_unit_integral_and_symmetry(oracle_exact, p, q, w):
    ys = np.linspace(-(p + q + w), p + q + w, 4001)
    vals = np.array([oracle_exact.box3(y, p, q, w) for y in ys])
    assert np.trapezoid(vals, ys) == pytest.approx(1.0, abs=1e-5)
    assert np.allclose(vals, vals[::-1], atol=1e-14)
