"""Pins the oracle's cone-beam separable-footprint projector (restating
layers/SF_projector.cu:81-234 with its fixes, see or_sf3d_*) against
mathematics that does not reuse it:

  * exact cell integrals: the FP64 mean over each detector cell of the chord
    length of the true cone-beam rays through one voxel.  The model's mass per
    voxel and view (lsd^2 dx dy dz / (du dv r^2)) equals the exact mass to the
    model's cone-angle error (within 6 % here; a dropped dx or 1/r^2 costs
    30 % or more), and its shape to <= 6 %, for SF-TR and SF-TT, at unit and
    non-unit voxels;
  * forward / transpose adjointness, 90-degree rotation <-> view shift of the
    row-summed projections, FP32 build within 2e-5 of the FP64 build.
"""
import numpy as np
import pytest

LSD, LSO = 200.0, 100.0


def geo3(o, d=1.0, nx=9, ny=9, nz=41, angles=(0.1, 0.9, 2.0), nu=64, nv=100):
    from pyoracle import Geometry3
    return Geometry3(o, nx, ny, nz, np.asarray(angles, float), nu, nv, d, d, d, 1.0, 1.0, LSD, LSO)


@pytest.mark.parametrize("d", [1.0, 0.7])
@pytest.mark.parametrize("trap_v", [False, True])
def test_single_voxels_match_exact_cells(oracle_exact, d, trap_v):
    g = geo3(oracle_exact, d)
    for ix, iy, iz in [(4, 4, 20), (2, 6, 35), (7, 1, 3)]:
        vol = np.zeros((g.nz, g.ny, g.nx))
        vol[iz, iy, ix] = 1.0
        a = g.sf_forward(vol, trap_v=trap_v)
        for p in range(g.np):
            js, us = np.nonzero(a[p])
            j0, j1, u0, u1 = js.min() - 1, js.max() + 1, us.min() - 1, us.max() + 1
            E = np.array([[g.sf_cell_exact(p, u, j, ix, iy, iz, panels=12) for u in range(u0, u1 + 1)]
                          for j in range(j0, j1 + 1)])
            m = a[p, j0:j1 + 1, u0:u1 + 1]
            ratio = m.sum() / E.sum()
            shape = np.linalg.norm(m / m.sum() - E / E.sum()) / np.linalg.norm(E / E.sum())
            assert 0.94 <= ratio <= 1.015, (d, trap_v, ix, iy, iz, p, ratio)
            assert shape <= 0.06, (d, trap_v, ix, iy, iz, p, shape)


def test_adjoint_rotation_and_fp32(oracle_exact, oracle_ref):
    rng = np.random.default_rng(4)
    np_ = 8
    ang = 2 * np.pi * np.arange(np_) / np_ + 0.05
    kw = dict(nx=12, ny=12, nz=10, angles=ang, nu=40, nv=30)
    g = geo3(oracle_exact, 1.0, **kw)
    x = rng.random((g.nz, g.ny, g.nx))
    y = rng.random((g.np, g.nv, g.nu))
    for trap_v in (False, True):
        ax = g.sf_forward(x, trap_v)
        aty = g.sf_backward(y, trap_v)
        assert abs((ax * y).sum() - (x * aty).sum()) / (np.linalg.norm(ax) * np.linalg.norm(y)) < 1e-12
        xr = np.rot90(x, k=1, axes=(1, 2)).copy()
        s, sr = ax.sum(axis=1), g.sf_forward(xr, trap_v).sum(axis=1)
        assert np.linalg.norm(sr - np.roll(s, -(np_ // 4), axis=0)) / np.linalg.norm(s) < 1e-9
        gr = geo3(oracle_ref, 1.0, **kw)
        assert np.linalg.norm(gr.sf_forward(x, trap_v) - ax) / np.linalg.norm(ax) < 2e-5
