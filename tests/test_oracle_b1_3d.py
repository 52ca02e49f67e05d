"""Pins the oracle's cone-beam B1 projector (restating layers/A_B1_3D.cu:82-312)
against mathematics that does not reuse it:

  * exact cell integrals: the FP64 mean over each detector cell of the line
    integral of one voxel's basis (bilinear hat in x, y times a box in z)
    along the true cone-beam rays.  The reference's model equals that mean
    times the voxel's back-projected row height dv / M (M = lsd / depth of
    the voxel centre): it scales each voxel's axial trapezoid to unit area
    and multiplies by dz (A_B1_3D.cu:209-210).  Per voxel and view the model
    mass matches the exact mass to 1e-3 (a dropped path-length weight costs
    2 % at the off-plane voxels used here), and the cell values agree to the
    model's own error (<= 8 %: the reference builds the axial trapezoid from
    x +- 1 at the column's iy only);
  * the v axis: a voxel and its mirror image in z project to row-mirrored
    images, and the row centroid sits where the voxel centre projects;
  * forward / transpose adjointness, and 90-degree rotation <-> view shift;
  * the FP32 reference-mode build stays within 2e-5 of the FP64 build.
"""
import numpy as np

LSD, LSO = 200.0, 100.0


def geo3(o, nx=9, ny=9, nz=41, angles=(0.1, 0.9, 2.0, 3.0, 4.4), nu=64, nv=128, d=1.0, du=1.0, dv=1.0):
    from pyoracle import Geometry3
    return Geometry3(o, nx, ny, nz, np.asarray(angles, dtype=float), nu, nv, d, d, d, du, dv, LSD, LSO)


def _voxel(g, ix, iy, iz):
    vol = np.zeros((g.nz, g.ny, g.nx))
    vol[iz, iy, ix] = 1.0
    return vol


def _depth(g, p, ix, iy):
    dx = g.d[0]
    a = g.angles[p]
    src = np.array([-LSO * np.cos(a), -LSO * np.sin(a)])
    c = np.array([(ix - 0.5 * (g.nx - 1)) * dx, (iy - 0.5 * (g.ny - 1)) * dx])
    return (c - src) @ (-src / LSO)


def test_single_voxels_match_exact_cell_integrals(oracle_exact):
    g = geo3(oracle_exact)
    for ix, iy, iz in [(4, 4, 20), (2, 6, 35), (7, 1, 3)]:
        a = g.forward(_voxel(g, ix, iy, iz))
        for p in range(g.np):
            js, us = np.nonzero(a[p])
            j0, j1, u0, u1 = js.min() - 1, js.max() + 1, us.min() - 1, us.max() + 1
            E = np.array([[g.cell_exact(p, u, j, ix, iy, iz, panels=3) for u in range(u0, u1 + 1)]
                          for j in range(j0, j1 + 1)])
            scaled = E * g.d[4] / (LSD / _depth(g, p, ix, iy))
            model = a[p, j0:j1 + 1, u0:u1 + 1]
            assert abs(model.sum() / scaled.sum() - 1) < 1e-3, (ix, iy, iz, p)
            assert np.linalg.norm(model - scaled) / np.linalg.norm(scaled) < 0.08, (ix, iy, iz, p)


def test_v_axis_mirror_and_centroid(oracle_exact):
    g = geo3(oracle_exact)
    for ix, iy, iz in [(3, 5, 30), (6, 2, 8)]:
        a = g.forward(_voxel(g, ix, iy, iz))
        b = g.forward(_voxel(g, ix, iy, g.nz - 1 - iz))
        assert np.allclose(a, b[:, ::-1, :], rtol=1e-9, atol=1e-12)
        for p in range(g.np):
            rows = a[p].sum(axis=1)
            centroid = (rows * (np.arange(g.nv) + 0.5)).sum() / rows.sum()
            z = (iz - 0.5 * (g.nz - 1)) * g.d[2]
            v_centre = g.nv / 2 + (LSD / g.d[4]) * z / _depth(g, p, ix, iy)
            assert abs(centroid - v_centre) < 0.05, (ix, iy, iz, p, centroid, v_centre)


def test_adjoint_and_rotation(oracle_exact, oracle_ref):
    rng = np.random.default_rng(3)
    np_ = 8
    ang = 2 * np.pi * np.arange(np_) / np_ + 0.05
    g = geo3(oracle_exact, nx=12, ny=12, nz=10, angles=ang, nu=40, nv=24, du=1.5, dv=1.5)
    x = rng.random((g.nz, g.ny, g.nx))
    y = rng.random((g.np, g.nv, g.nu))
    ax = g.forward(x)
    aty = g.backward(y)
    assert abs((ax * y).sum() - (x * aty).sum()) / (np.linalg.norm(ax) * np.linalg.norm(y)) < 1e-12
    # rotating the volume by 90 degrees about z shifts the views by np/4 in
    # the row-summed projections (the in-plane part; the axial trapezoid uses
    # x +- 1 only, so single rows are not rotation-equivariant)
    xr = np.rot90(x, k=1, axes=(1, 2)).copy()
    s, sr = ax.sum(axis=1), g.forward(xr).sum(axis=1)
    assert np.linalg.norm(sr - np.roll(s, -(np_ // 4), axis=0)) / np.linalg.norm(s) < 1e-9
    gr = geo3(oracle_ref, nx=12, ny=12, nz=10, angles=ang, nu=40, nv=24, du=1.5, dv=1.5)
    assert np.linalg.norm(gr.forward(x) - ax) / np.linalg.norm(ax) < 2e-5
