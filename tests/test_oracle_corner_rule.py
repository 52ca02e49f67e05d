"""Pins the oracle's corner rule (oracle/sgdct_oracle.c corner_range, the
reference's layers/A_B0.cu:158-195 and A_B1.cu:100-151): a pixel's bins are
[clamp(floor(min k)), clamp(floor(max k))] over the detector coordinates k of
its four corners at index offsets +-h (h = 1/2 for B0 and SF, 1 for B1),
clamped into [0, nu-1].  Here k is recomputed from the view geometry alone
(the ray from the source through the corner meets the detector line
E(k) = dtv + pu (k - nu/2)), without the oracle's projective matrices, and the
oracle's per-view evaluation counts must equal the sums of these ranges
exactly.  A wrong half-width (B0 and B1 swapped), a floor / ceil or an
off-by-one in the clamp changes the counts: B0's footprint past its corner
range is only ~6.5e-5 of the matrix norm, below what the footprint pins
resolve."""
import numpy as np
import pytest

NX, NY, NU = 24, 20, 60
DX, DY, DU, LSD, LSO = 1.0, 1.1, 1.2, 200.0, 100.0
# no view within a degree of a multiple of 45 degrees (the reference treats
# the diagonals separately)
ANGLES = np.array([0.1, 0.7, 1.3, 2.0, 2.9, 3.7, 4.4, 5.5, 6.05])


def _ranges(g, p, h):
    s, d, pu = g.view(p)
    ix, iy = np.meshgrid(np.arange(NX), np.arange(NY), indexing="ij")
    ks = []
    for ox in (-h, h):
        for oy in (-h, h):
            X = (ix + ox - 0.5 * (NX - 1)) * DX
            Y = (iy + oy - 0.5 * (NY - 1)) * DY
            # s + lam (P - s) = d + t pu  ->  solve for t (Cramer)
            ax, ay = X - s[0], Y - s[1]
            rx, ry = d[0] - s[0], d[1] - s[1]
            det = ax * (-pu[1]) - ay * (-pu[0])
            t = (ax * ry - ay * rx) / det
            ks.append(t + 0.5 * NU)
    ks = np.array(ks)
    lo = np.clip(np.floor(ks.min(axis=0)), 0, NU - 1).astype(int)
    hi = np.clip(np.floor(ks.max(axis=0)), 0, NU - 1).astype(int)
    return lo, hi


@pytest.mark.parametrize("basis,h", [("B0", 0.5), ("SF", 0.5), ("B1", 1.0)])
def test_corner_rule_counts_from_geometry(oracle_exact, basis, h):
    from pyoracle import Geometry
    g = Geometry(oracle_exact, NX, NY, ANGLES, NU, DX, DY, DU, LSD, LSO)
    for p in range(len(ANGLES)):
        lo, hi = _ranges(g, p, h)
        assert (hi >= lo).all()
        assert g.eval_count(basis, views=(p, p + 1)) == int((hi - lo + 1).sum()), (basis, p)


def test_corner_rule_footprint_bins(oracle_exact):
    """The bins the oracle's B0 / B1 footprints return lie in those ranges
    (the footprints drop values <= 1e-7, as layers/A_B1.cu:203 and
    A_B0.cu:225 do, so the lists can be shorter), and the ranges' interior
    bins are all there."""
    from pyoracle import Geometry
    g = Geometry(oracle_exact, NX, NY, ANGLES, NU, DX, DY, DU, LSD, LSO)
    for basis, h in (("B0", 0.5), ("B1", 1.0)):
        for p in (0, 4, 8):
            lo, hi = _ranges(g, p, h)
            for ix, iy in ((0, 0), (5, 7), (23, 19), (12, 3)):
                us, _ = g.footprint(basis, p, ix, iy, pad=0)
                rng = set(range(lo[ix, iy], hi[ix, iy] + 1))
                assert set(us.tolist()) <= rng, (basis, p, ix, iy)
                assert set(range(lo[ix, iy] + 1, hi[ix, iy])) <= set(us.tolist()), (basis, p, ix, iy)


def test_half_width_is_visible_in_counts(oracle_exact):
    """The check has teeth: the B1 half-width gives B0's count a different
    value on this geometry."""
    from pyoracle import Geometry
    g = Geometry(oracle_exact, NX, NY, ANGLES, NU, DX, DY, DU, LSD, LSO)
    wrong = sum(int((hi - lo + 1).sum()) for hi, lo in
                (tuple(reversed(_ranges(g, p, 1.0))) for p in range(len(ANGLES))))
    assert wrong != g.eval_count("B0")
