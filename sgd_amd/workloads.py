"""Named fan-beam workloads, built the same way by bench.py, the tests and
sgd_amd.measure.

fan512      B0 / B1 / SF at 512 x 512 pixels of 0.6 mm, 720 views over 2 pi,
            768 bins of 0.8 mm, source-isocentre 500 mm, source-detector
            1000 mm.  The image is the seed-0 random-ellipse phantom; A^T is
            applied to its Poisson-noisy log sinogram (1e5 incident photons
            per ray, attenuation 0.1 / mm).
hex296      H0 / Hl on 296 rows x 256 columns of a hexagonal lattice (lattice
            constant 0.5 mm), 1024 views over 2 pi, 1024 bins of 0.6 mm, same
            source distances.
helical     fan slices rebinned from a helical scan: each view carries its
            source's radial offset drho (flying focal spot) and axial offset
            dz from the slice plane (layers/A_B1.hh:24, utils/geo.cu:243-251).
"""
import numpy as np

from .phantom import phantom, poisson_noise

FAN512 = dict(nx=512, ny=512, dx=0.6, dy=0.6, views=720, nu=768, du=0.8, lsd=1000.0, lso=500.0)
FAN512_N0 = 1e5
FAN512_MU = 0.1  # per mm: sinogram values of the phantom stay well inside the photon budget

HEX296 = dict(nl=296, nc=256, dt=0.5, views=1024, nu=1024, du=0.6, lsd=1000.0, lso=500.0)


def angles(n, start=0.0):
    return start + 2 * np.pi * np.arange(n) / n


def fan512_projector(basis="B1"):
    from .torch_ops import Projector
    g = FAN512
    return Projector(basis, g["nx"], g["ny"], angles(g["views"]), g["nu"], g["dx"], g["dy"], g["du"], g["lsd"],
                     g["lso"])


def fan512_inputs(P):
    """(image, noisy sinogram) as FP64 numpy arrays: the seed-0 phantom and
    -ln(max(Poisson(N0 exp(-mu A x)), 1) / N0) / mu with A the projector P
    (its own FP32 result, widened)."""
    import torch
    g = FAN512
    img = phantom(0, g["nx"], g["ny"], g["dx"], g["dy"])
    x = torch.tensor(img, dtype=torch.float32, device=P.device)
    y0 = P.forward(x).double().cpu().numpy()
    sino = poisson_noise(y0 * FAN512_MU, FAN512_N0, seed=1) / FAN512_MU
    return img, sino


def hex296_projector(basis):
    from .torch_ops import Projector
    g = HEX296
    return Projector.hexagonal(basis, g["nl"], g["nc"], angles(g["views"]), g["nu"], g["dt"], g["du"], g["lsd"],
                               g["lso"])


def helical_offsets(nviews, seed, pitch=20.0, focal=0.6, jitter=0.05):
    """Per-view (dz, drho) of a fan slice rebinned from a helical scan: the
    source rises by `pitch` mm over the slice's 2 pi of views, centred on the
    slice plane; a flying focal spot alternates the source radius by
    +-`focal` mm, plus Gaussian jitter of `jitter` mm."""
    rng = np.random.default_rng(seed)
    k = np.arange(nviews)
    dz = pitch * (k / nviews - 0.5)
    drho = focal * np.where(k % 2 == 0, 1.0, -1.0) + jitter * rng.standard_normal(nviews)
    return dz, drho
