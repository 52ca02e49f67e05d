"""Seeded ellipse phantoms, their analytic fan-beam sinograms, Poisson noise.

The workloads are stated in terms of a random-ellipse phantom: ten ellipses
with centres 0.8*sqrt(U)*(cos 2piU', sin 2piU') (units of the field-of-view
half-width), semi-axes U[0.05, 0.4], orientation U[0, pi), density
U[-0.5, 1.0], summed and clipped at zero.  Draws come from numpy's MT19937
bit generator seeded with `seed`.

Line integrals of ellipses are closed-form, so `analytic_sinogram` gives an
independent check of a projector's geometry (orientation of views, direction
of the detector axis, pixel-centre convention).  The noise model is
counts ~ Poisson(N0 exp(-p)), b = -ln(max(counts, 1) / N0).
"""
import numpy as np

__all__ = ["ellipse_params", "rasterize", "phantom", "analytic_sinogram", "poisson_noise"]


def ellipse_params(seed, n=10, fov_half=1.0, densities=(-0.5, 1.0)):
    """(n, 6) array of (cx, cy, a, b, angle, density) in physical units."""
    rng = np.random.Generator(np.random.MT19937(seed))
    out = np.zeros((n, 6))
    for k in range(n):
        r = 0.8 * np.sqrt(rng.uniform())
        phi = 2 * np.pi * rng.uniform()
        a = rng.uniform(0.05, 0.4)
        b = rng.uniform(0.05, 0.4)
        ang = rng.uniform(0.0, np.pi)
        dens = rng.uniform(*densities)
        out[k] = (r * np.cos(phi) * fov_half, r * np.sin(phi) * fov_half, a * fov_half, b * fov_half, ang, dens)
    return out


def rasterize(params, nx, ny, dx, dy, supersample=1, clip=True):
    """Image indexed [iy, ix]; pixel centres at ((ix-(nx-1)/2)dx, (iy-(ny-1)/2)dy)."""
    s = int(supersample)
    offs = (np.arange(s) + 0.5) / s - 0.5
    xs = (np.arange(nx) - 0.5 * (nx - 1)) * dx
    ys = (np.arange(ny) - 0.5 * (ny - 1)) * dy
    img = np.zeros((ny, nx))
    for oy in offs:
        for ox in offs:
            X, Y = np.meshgrid(xs + ox * dx, ys + oy * dy)
            acc = np.zeros((ny, nx))
            for cx, cy, a, b, ang, d in params:
                c, sn = np.cos(ang), np.sin(ang)
                xr = (X - cx) * c + (Y - cy) * sn
                yr = -(X - cx) * sn + (Y - cy) * c
                acc += d * ((xr / a) ** 2 + (yr / b) ** 2 <= 1.0)
            img += acc
    img /= s * s
    if clip:
        np.maximum(img, 0.0, out=img)
    return img


def phantom(seed, nx, ny, dx, dy, n=10, supersample=1):
    """Random-ellipse phantom on an nx x ny grid covering the square FOV."""
    fov_half = 0.5 * min(nx * dx, ny * dy)
    return rasterize(ellipse_params(seed, n, fov_half), nx, ny, dx, dy, supersample)


def _chords(params, px, py, ex, ey):
    """Sum over ellipses of density * chord length of rays P + t E (unit E)."""
    total = np.zeros(np.broadcast(px, ex).shape)
    for cx, cy, a, b, ang, d in params:
        c, s = np.cos(ang), np.sin(ang)
        qx = (px - cx) * c + (py - cy) * s
        qy = -(px - cx) * s + (py - cy) * c
        fx = ex * c + ey * s
        fy = -ex * s + ey * c
        A = (fx / a) ** 2 + (fy / b) ** 2
        B = 2 * (qx * fx / a ** 2 + qy * fy / b ** 2)
        C = (qx / a) ** 2 + (qy / b) ** 2 - 1.0
        disc = B * B - 4 * A * C
        total += d * np.where(disc > 0, np.sqrt(np.maximum(disc, 0.0)) / A, 0.0)
    return total


def analytic_sinogram(params, angles, nu, du, lsd, lso, nquad=8):
    """Bin-averaged analytic fan-beam sinogram (np, nu) of the unclipped sum.

    View geometry as utils/geo.cu:232-262: source R(b)(-lso, 0), detector
    centre R(b)(lsd-lso, 0), detector step R(b)(0, du); bin u spans detector
    coordinates [u - nu/2, u + 1 - nu/2) * du.
    """
    angles = np.asarray(angles, dtype=np.float64)
    xg, wg = np.polynomial.legendre.leggauss(nquad)
    tau = 0.5 * (xg + 1.0)
    w = 0.5 * wg
    out = np.zeros((len(angles), nu))
    us = np.arange(nu)[:, None] + tau[None, :] - 0.5 * nu
    for k, b in enumerate(angles):
        c, s = np.cos(b), np.sin(b)
        px, py = -lso * c, -lso * s
        qx = (lsd - lso) * c - du * s * us
        qy = (lsd - lso) * s + du * c * us
        ex, ey = qx - px, qy - py
        n = np.sqrt(ex * ex + ey * ey)
        out[k] = (_chords(params, px, py, ex / n, ey / n) * w[None, :]).sum(axis=1)
    return out


def poisson_noise(sino, n0, seed):
    """Log-transformed Poisson counts: -ln(max(Poisson(n0 e^-p), 1) / n0)."""
    rng = np.random.Generator(np.random.MT19937(seed))
    counts = rng.poisson(n0 * np.exp(-np.asarray(sino, dtype=np.float64)))
    return -np.log(np.maximum(counts, 1) / n0)
