"""Pins the FP64 oracle's SF footprints (oracle/sgdct_oracle.c sf_pixel;
layers/A_SF.cu:87-225) on a large grid whose pixel size float cannot hold:
2048 x 2048 pixels of 0.15 mm, 2880 views, 3072 bins of 0.2 mm.  Each weight
is rebuilt from the view geometry alone, without the oracle's projection
matrices or its overlap formula:
  * the trapezoid's corners are where the rays from the source through the
    pixel's four corners meet the detector line E(k) = dtv + pu (k - nu/2);
  * the overlap with bin u is F(u + 1) - F(u), F the trapezoid's
    antiderivative;
  * the amplitude is dx / max(|cos|, |sin|) of the ray through the centre.
They must agree to 1e-10.  The second test shows the resolution: the same
rebuild with the image centre rounded the way the reference's projection
matrix rounds it (float (nx - 1) * float(dx) / 2, utils/geo.cu:141-152)
misses the oracle by more than 1e-6, so a float term in the FP64 geometry
fails the first test."""
import numpy as np
import pytest

NX = NY = 2048
NV, NU = 2880, 3072
D, DU, LSD, LSO = 0.15, 0.2, 1000.0, 500.0
VIEWS = (0, 137, 721, 1440, 2177, 2879)


def _pixels():
    rng = np.random.default_rng(3)
    pts = [(0, 0), (NX - 1, NY - 1), (0, NY - 1), (NX // 2, NY // 2)]
    pts += [tuple(int(v) for v in rng.integers(0, NX, 2)) for _ in range(12)]
    return pts


def _antiderivative(q, s):
    q0, q1, q2, q3 = q
    r, f = q1 - q0, q3 - q2
    if s <= q0:
        return 0.0
    if s <= q1:
        return (s - q0) ** 2 / (2 * r)
    if s <= q2:
        return r / 2 + (s - q1)
    if s <= q3:
        return r / 2 + (q2 - q1) + f / 2 - (q3 - s) ** 2 / (2 * f)
    return r / 2 + (q2 - q1) + f / 2


def _footprint(g, p, ix, iy, cx, cy):
    """{u: weight} from geometry; cx, cy the image-centre offsets in mm."""
    s, d, pu = g.view(p)
    ks = []
    for ox in (-0.5, 0.5):
        for oy in (-0.5, 0.5):
            ax, ay = (ix + ox) * D - cx - s[0], (iy + oy) * D - cy - s[1]
            rx, ry = d[0] - s[0], d[1] - s[1]
            t = (ax * ry - ay * rx) / (ay * pu[0] - ax * pu[1])
            ks.append(t + 0.5 * NU)
    q = sorted(ks)
    wx, wy = ix * D - cx - s[0], iy * D - cy - s[1]
    amp = D * np.hypot(wx, wy) / max(abs(wx), abs(wy))
    out = {}
    for u in range(max(int(np.floor(q[0])), 0), min(int(np.floor(q[3])), NU - 1) + 1):
        out[u] = (_antiderivative(q, u + 1.0) - _antiderivative(q, float(u)), amp)
    return out


def _max_error(g, cx, cy):
    worst, checked = 0.0, 0
    for p in VIEWS:
        for ix, iy in _pixels():
            ref = _footprint(g, p, ix, iy, cx, cy)
            # (an image corner can lie outside the fan: empty list)
            us, ws = g.footprint("SF", p, ix, iy)
            for u, w in zip(us.tolist(), ws.tolist()):
                gm, amp = ref[u]
                worst = max(worst, abs(w - gm * amp) / amp)
                checked += 1
            # bins left out must hold overlaps at or below the 1e-7 cut
            # (A_SF.cu); more than that counts as error
            for u, (gm, _) in ref.items():
                if u not in set(us.tolist()):
                    worst = max(worst, gm - 1e-7)
    assert checked >= 200, checked
    return worst


def _geometry(oracle_exact):
    from pyoracle import Geometry
    ang = 2 * np.pi * np.arange(NV) / NV
    return Geometry(oracle_exact, NX, NY, ang, NU, D, D, DU, LSD, LSO)


def test_sf_footprints_from_geometry(oracle_exact):
    g = _geometry(oracle_exact)
    c = 0.5 * (NX - 1) * D
    assert _max_error(g, c, c) <= 1e-10


def test_float_centre_is_resolved(oracle_exact):
    g = _geometry(oracle_exact)
    cf = float(np.float32(np.float32(NX - 1) * np.float32(D)) / np.float32(2))
    assert abs(cf - 0.5 * (NX - 1) * D) > 1e-6
    assert _max_error(g, cf, cf) > 1e-6
