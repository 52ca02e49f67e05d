"""FORBILD head phantom: analytic description, rasterisation, exact line
integrals, fan-beam sinograms with the reference's noise model, image
metrics, and the CGLS accuracy experiment of the reference's notebook.

Follows python/FORBILD.py of the reference (which needs numba, absent here):
  * `phantom_tables` -- _analytical_forbild_phantom (FORBILD.py:19-190): the
    FORBILD head (Z. Yu et al., Phys. Med. Biol. 57 (2012) 3897, App. 1) as
    ellipses (x0, y0, a, b, angle, density, number of clip lines) and clip
    lines (distance, angle), with the reference's additions (a right-ear
    structure with air cavities, four-clipped high-contrast bars and an
    optional left-ear resolution pattern), all lengths scaled by 4 (the head
    spans about 77 x 96 mm).
  * `raster` -- discrete_phantom (FORBILD.py:192-218): density at points.
  * `line_integrals` -- batch_line_integral (FORBILD.py:220-279): exact
    integral along the line {x : x . (cos t, sin t) = s}: per ellipse the two
    chord ends from the quadratic, each end moved onto a clip line when it
    lies on the clipped side.
  * `fan_sinogram` -- forbild_sinogram (FORBILD.py:398-412, bin average of the
    line integral, or log of the bin average of exp(line integral), the
    `beers_law` form whose 1800-view, 1710-bin version the notebook combines
    into its 900 x 855 sinogram) and forbild_sinogram_noise
    (FORBILD.py:355-379: 200 rays per bin, counts ~ Poisson(100
    exp(-L / 100)) each, b = 100 ln(20000 / sum counts)).  The fan geometry is
    the projectors' (utils/geo.cu:232-262): the ray through detector
    coordinate u of view b has normal angle b + pi/2 + atan(u / lsd) and
    offset s = lso u / sqrt(lsd^2 + u^2).
  * `psnr`, `ssim`, `sample_nearest`, `sample_bilinear` -- the notebook's
    evaluation (main copy 2.ipynb cells 24-28: skimage's PSNR and SSIM on a
    3000 x 3000 raster of the phantom, data range 2; B0 images sampled at the
    containing pixel, B1 bilinearly between pixel centres).

`python -m sgd_amd.forbild` runs the notebook's experiment on the device:
CGLS (the Krylov method pylops' regularized_inversion runs) for B0 and B1 at
512 x 512 pixels of 100/512 mm, 900 views, 855 bins of 0.390625 mm, lsd
312.5 mm, lso 156.25 mm, with PSNR and SSIM every second iteration.
"""
import argparse
import json
import time

import numpy as np

__all__ = ["phantom_tables", "raster", "raster_grid", "line_integrals", "fan_sinogram", "psnr", "ssim", "sample_nearest",
           "sample_bilinear", "NOTEBOOK"]

# The notebook's reconstruction workload (main copy 2.ipynb cell 7).
NOTEBOOK = dict(nx=512, ny=512, dx=100.0 / 512, dy=100.0 / 512, views=900, nu=855, du=25.0 / 128 * 2,
                lsd=312.5, lso=156.25)


def phantom_tables(resolution=False, ear=True):
    """(E, C): E (n, 11) rows (x0, y0, a, b, angle_deg, density, nclip,
    cos/a, sin/a, -sin/b, cos/b); C (4, m) rows (distance, angle_deg, cos,
    sin) of the clip lines, used in order by the ellipses with nclip > 0.
    Lengths in mm (the published table in cm, times 4)."""
    sha = 0.2 * np.sqrt(3)
    y016b = -14.294530834372887
    a16b = 0.443194085308632
    b16b = 3.892760834372886
    # x0, y0, a, b, angle, density, clip lines (FORBILD.py:34-61)
    base = [[-4.7, 4.3, 1.79989, 1.79989, 0, 0.010, 0],
            [4.7, 4.3, 1.79989, 1.79989, 0, 0.010, 0],
            [-1.08, -9, 0.4, 0.4, 0, 0.0025, 0],
            [1.08, -9, 0.4, 0.4, 0, -0.0025, 0],
            [0, 0, 9.6, 12, 0, 1.800, 0],
            [0, 8.4, 1.8, 3.0, 0, -1.050, 0],
            [1.9, 5.4, 0.41633, 1.17425, -31.07698, 0.750, 0],
            [-1.9, 5.4, 0.41633, 1.17425, 31.07698, 0.750, 0],
            [-4.3, 6.8, 1.8, 0.24, -30, 0.750, 0],
            [4.3, 6.8, 1.8, 0.24, 30, 0.750, 0],
            [0, -3.6, 1.8, 3.6, 0, -0.005, 0],
            [6.39395, -6.39395, 1.2, 0.42, 58.1, 0.005, 0],
            [0, 3.6, 2, 2, 0, 0.750, 4],
            [0, 9.6, 1.8, 3.0, 0, 1.800, 4],
            [0, 0, 9.0, 11.4, 0, 0.750, 3],
            [0, y016b, a16b, b16b, 0, 0.750, 1],
            [0, 0, 9.0, 11.4, 0, -0.750, 1 if ear else 0]]
    extra = [[9.1, 0, 4.2, 1.8, 0, 0.750, 1],  # right ear
             [-3, 0, 3, 3, 0, 0.8, 4], [-4, 0, 3, 3, 0, 0.8, 4], [-5, 0, 3, 3, 0, 0.8, 4]]
    extra += [[-4, -6, 3, 3, 0, 0.8, 4]] * 6
    E = np.array(base + extra, dtype=np.float64)
    # air cavities of the right ear (FORBILD.py:66-79): 53 small circles
    cav_x = np.arange(8.8, 5.6, -0.4)[:, None]
    cav_y = np.zeros([9, 1])
    for j in range(1, 4):
        kj = 8 - 2 * (j // 3)
        dj = 0.2 * (j % 2)
        cav_x = np.vstack((cav_x, cav_x[0:kj] - dj, cav_x[0:kj] - dj))
        cav_y = np.vstack((cav_y, j * sha * np.ones([kj, 1]), -j * sha * np.ones([kj, 1])))
    E_cav = np.hstack((cav_x, cav_y, np.ones([cav_x.shape[0], 1]) * [0.15, 0.15, 0, -1.800, 0]))
    # left-ear resolution pattern (FORBILD.py:81-108): 80 small circles
    r, x0, y0 = 3.0, -7.0, -1.0
    d0 = 0.04 * r
    dxy = [0.0357 * r, 0.0312 * r, 0.0278 * r, 0.0250 * r]
    ab = (0.5 * np.ones([5, 1]) * dxy).T.ravel()[:, None] * np.ones([1, 4])
    abr = ab.T.ravel()[:, None]
    xs = np.zeros([0, 1])
    ys = np.zeros([0, 1])
    for i in range(1, 5):
        ys = np.vstack((ys, (y0 + np.arange(0, 5) * 2 * dxy[i - 1])[:, None]))
        xs = np.vstack((xs, (x0 + 2 * (i - 1) * d0) * np.ones([5, 1])))
    xs = (xs * np.ones([1, 4])).T.ravel()[:, None]
    ys = np.vstack([ys, ys + 12 * d0, ys + 24 * d0, ys + 36 * d0])
    E_res = np.hstack([xs, ys, abr, abr, np.ones([80, 1]) * [0, 0.75, 0]])
    # clip lines (FORBILD.py:110-133): distance, angle in degrees
    dist = [1.2, 1.2, 0.27884, 0.27884, 0.60687, 0.60687, 0.2, 0.2, -2.605, -2.605, -10.71177, y016b + 10.71177,
            8.88740, -0.21260]
    ang = [0, 180, 90, 270, 90, 270, 0, 180, 15, 165, 90, 270, 0, 0]
    bars = [(0.1, 0.1, 2, 2)] * 4 + [(0.1 / 2 * np.sqrt(3), 0.1 / 2 * np.sqrt(3), 2, 2), (0.1 / 2, 0.1 / 2, 2, 2),
                                     (0.1, 0.1, 2, 2), (0.1 / 2 * np.sqrt(3), 0.1 / 2 * np.sqrt(3), 2, 2),
                                     (0.1 / 2, 0.1 / 2, 2, 2)]
    for k, d in enumerate(bars):
        dist += list(d)
        ang += [a + 30 * max(k - 3, 0) for a in (0, 180, 90, 270)]
    C = np.array([dist, ang], dtype=np.float64)
    if not resolution and not ear:
        E, C = E[:17], C[:, :12]
    elif not resolution and ear:
        E = np.vstack([E, E_cav])
    elif resolution and not ear:
        E, C = np.vstack([E_res, E[:17]]), C[:, :12]
    else:
        E = np.vstack([E_res, E, E_cav])
    E = E.copy()
    E[:, :4] *= 4.0
    phi = np.radians(E[:, 4])
    DQ = np.stack([np.cos(phi) / E[:, 2], np.sin(phi) / E[:, 2], -np.sin(phi) / E[:, 3], np.cos(phi) / E[:, 3]], 1)
    E = np.hstack([E, DQ])
    Cr = np.radians(C[1])
    C = np.vstack([C[0] * 4.0, C[1], np.cos(Cr), np.sin(Cr)])
    return E, C


def raster(x, y, tables=None):
    """Phantom density at points (x, y) (arrays of one shape, mm)."""
    E, C = tables if tables is not None else phantom_tables()
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    img = np.zeros(x.shape)
    nc = 0
    for row in E:
        vx, vy = x - row[0], y - row[1]
        q1 = row[7] * vx + row[8] * vy
        q2 = row[9] * vx + row[10] * vy
        inside = q1 * q1 + q2 * q2 <= 1.0
        for _ in range(int(row[6])):
            inside &= C[2, nc] * vx + C[3, nc] * vy < C[0, nc]
            nc += 1
        img += np.where(inside, row[5], 0.0)
    return img


def raster_grid(xs, ys, tables=None):
    """raster() on the grid (ys[i], xs[j]) (1-D coordinate arrays), each
    ellipse evaluated only on the rows and columns of its bounding box: the
    same values, point by point, as raster() of the meshgrid."""
    E, C = tables if tables is not None else phantom_tables()
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    img = np.zeros((ys.size, xs.size))
    nc = 0
    for row in E:
        nclip = int(row[6])
        r = max(row[2], row[3])
        ci = np.nonzero(np.abs(xs - row[0]) <= r)[0]
        ri = np.nonzero(np.abs(ys - row[1]) <= r)[0]
        if ci.size and ri.size:
            c0, c1, r0, r1 = ci.min(), ci.max() + 1, ri.min(), ri.max() + 1
            vx = xs[None, c0:c1] - row[0]
            vy = ys[r0:r1, None] - row[1]
            q1 = row[7] * vx + row[8] * vy
            q2 = row[9] * vx + row[10] * vy
            inside = q1 * q1 + q2 * q2 <= 1.0
            for k in range(nclip):
                inside &= C[2, nc + k] * vx + C[3, nc + k] * vy < C[0, nc + k]
            img[r0:r1, c0:c1] += np.where(inside, row[5], 0.0)
        nc += nclip
    return img


def line_integrals(theta, s, tables=None, xp=np):
    """Integral of the phantom along the lines x . (cos theta, sin theta) = s
    (direction (-sin theta, cos theta)); theta, s arrays of one shape.  `xp`
    is numpy, or torch for FP64 tensors on a device."""
    E, C = tables if tables is not None else phantom_tables()
    # lines parallel to a clip line are turned by 1e-10 rad (FORBILD.py:226-233)
    par = None
    for k in range(C.shape[1]):
        pk = xp.abs(-xp.sin(theta) * C[2, k] + xp.cos(theta) * C[3, k]) < 1e-10
        par = pk if par is None else (par | pk)
    theta = xp.where(par, theta + 1e-10, theta + 0.0)
    st, ct = xp.sin(theta), xp.cos(theta)
    sx, sy = s * ct, s * st
    out = xp.zeros_like(s)
    nc = 0
    for row in E:
        x0, y0, f, nclip = row[0], row[1], row[5], int(row[6])
        s0x, s0y = sx - x0, sy - y0
        d1 = row[7] * (-st) + row[8] * ct
        d2 = row[9] * (-st) + row[10] * ct
        e1 = row[7] * s0x + row[8] * s0y
        e2 = row[9] * s0x + row[10] * s0y
        A = d1 * d1 + d2 * d2
        B = 2.0 * (d1 * e1 + d2 * e2)
        Cq = e1 * e1 + e2 * e2 - 1.0
        disc = B * B - 4.0 * A * Cq
        hit = disc > 0
        root = xp.sqrt(xp.where(hit, disc, xp.zeros_like(disc)))
        tp = 0.5 * (-B + root) / A
        tq = 0.5 * (-B - root) / A
        for _ in range(nclip):
            d, cx, cy = C[0, nc], C[2, nc], C[3, nc]
            tz = (d - cx * s0x - cy * s0y) / (-st * cx + ct * cy)
            ep = (s0x - tp * st) * cx + (s0y + tp * ct) * cy
            eq = (s0x - tq * st) * cx + (s0y + tq * ct) * cy
            tp = xp.where(ep >= d, tz, tp)
            tq = xp.where(eq >= d, tz, tq)
            nc += 1
        out = out + xp.where(hit, f * xp.abs(tp - tq), xp.zeros_like(tp))
    return out


def fan_sinogram(nnp, nu, du, lsd, lso, samples=200, mode="mean", n0=100.0, scale=100.0, seed=0, device=None,
                 views_per_chunk=8, tables=None):
    """Fan-beam sinogram (nnp, nu) of the phantom, views b_k = 2 pi k / nnp,
    bin iu covering detector coordinates [iu - nu/2, iu + 1 - nu/2) du,
    `samples` rays at the centres of equal sub-intervals of each bin:
      mode "mean":          bin average of the line integral L
      mode "log_mean_exp":  log of the bin average of exp(L)
      mode "poisson":       counts ~ Poisson(n0 exp(-L / scale)) per ray,
                            scale * ln(n0 * samples / sum counts)
      mode "expected":      "poisson" with the counts replaced by their means.
    On a torch device when `device` is given (FP64), else numpy."""
    tables = tables if tables is not None else phantom_tables()
    step = du / samples
    u = (np.arange(nu)[:, None] - nu / 2) * du + (np.arange(samples)[None, :] + 0.5) * step  # (nu, samples)
    gamma = np.arctan2(u, lsd)
    soff = u * lso / np.sqrt(lsd * lsd + u * u)
    betas = 2 * np.pi * np.arange(nnp) / nnp
    out = np.zeros((nnp, nu))
    if device is not None:
        import torch
        g = torch.Generator(device=device).manual_seed(int(seed))
        gt = torch.tensor(gamma, device=device, dtype=torch.float64)
        st = torch.tensor(soff, device=device, dtype=torch.float64)
    else:
        rng = np.random.Generator(np.random.MT19937(seed))
    for p0 in range(0, nnp, views_per_chunk):
        bs = betas[p0:p0 + views_per_chunk]
        if device is not None:
            b = torch.tensor(bs, device=device, dtype=torch.float64)[:, None, None]
            L = line_integrals(b + np.pi / 2 + gt[None], st[None].expand(len(bs), -1, -1).contiguous(), tables,
                               xp=_TorchNS)
            if mode == "mean":
                v = L.mean(-1)
            elif mode == "log_mean_exp":
                v = torch.logsumexp(L, -1) - np.log(samples)
            elif mode in ("poisson", "expected"):
                lam = n0 * torch.exp(-L / scale)
                c = torch.poisson(lam, generator=g) if mode == "poisson" else lam
                v = scale * torch.log(n0 * samples / c.sum(-1))
            else:
                raise ValueError(mode)
            out[p0:p0 + len(bs)] = v.cpu().numpy()
        else:
            b = bs[:, None, None]
            L = line_integrals(b + np.pi / 2 + gamma[None], np.broadcast_to(soff, (len(bs),) + soff.shape).copy(),
                               tables)
            if mode == "mean":
                v = L.mean(-1)
            elif mode == "log_mean_exp":
                m = L.max(-1, keepdims=True)
                v = (m[..., 0] + np.log(np.exp(L - m).mean(-1)))
            elif mode in ("poisson", "expected"):
                lam = n0 * np.exp(-L / scale)
                c = rng.poisson(lam) if mode == "poisson" else lam
                v = scale * np.log(n0 * samples / c.sum(-1))
            else:
                raise ValueError(mode)
            out[p0:p0 + len(bs)] = v
    return out


class _TorchNS:
    """The few numpy functions line_integrals uses, on torch tensors."""

    @staticmethod
    def sin(x):
        import torch
        return torch.sin(x)

    @staticmethod
    def cos(x):
        import torch
        return torch.cos(x)

    @staticmethod
    def abs(x):
        import torch
        return torch.abs(x)

    @staticmethod
    def sqrt(x):
        import torch
        return torch.sqrt(x)

    @staticmethod
    def where(c, a, b):
        import torch
        return torch.where(c, a, b)

    @staticmethod
    def zeros_like(x):
        import torch
        return torch.zeros_like(x)


# ---- evaluation (main copy 2.ipynb cells 24-28) -----------------------------

def eval_grid(points=3000, half=50.0):
    """Coordinates of the notebook's reference raster: points x points cell
    centres of [-half, half)^2, x along columns, y decreasing down the rows."""
    x = np.linspace(-half, half, points, endpoint=False)
    x += (x[1] - x[0]) / 2
    return np.meshgrid(x, x[::-1])


def sample_nearest(img_xy, nx, ny, dx, dy, xs, ys):
    """Value of the pixel containing (x, y); img_xy indexed [ix, iy] (the
    reference's host layout reshaped), pixel ix spanning [ix - nx/2, ix + 1 -
    nx/2) dx."""
    ix = np.clip(np.floor((xs + nx / 2 * dx) / dx).astype(np.int64), 0, nx - 1)
    iy = np.clip(np.floor((ys + ny / 2 * dy) / dy).astype(np.int64), 0, ny - 1)
    return img_xy[ix, iy]


def sample_bilinear(img_xy, nx, ny, dx, dy, xs, ys):
    """Bilinear interpolation between pixel centres, edge values continued
    past the outer centres (the notebook pads by one pixel, mode 'edge')."""
    p = np.pad(img_xy, 1, mode="edge")
    fx = (xs + nx / 2 * dx) / dx + 0.5
    fy = (ys + ny / 2 * dy) / dy + 0.5
    x1 = np.floor(fx).astype(np.int64)
    y1 = np.floor(fy).astype(np.int64)
    wx, wy = fx - x1, fy - y1
    return (p[x1, y1] * (1 - wx) * (1 - wy) + p[x1, y1 + 1] * (1 - wx) * wy + p[x1 + 1, y1] * wx * (1 - wy) +
            p[x1 + 1, y1 + 1] * wx * wy)


def psnr(img, ref, data_range):
    mse = float(np.mean((np.asarray(img, dtype=np.float64) - ref) ** 2))
    return 10.0 * np.log10(data_range * data_range / mse)


def ssim(img, ref, data_range, win=7, k1=0.01, k2=0.03):
    """Mean structural similarity (Z. Wang et al., IEEE TIP 13 (2004) 600)
    with the defaults of skimage.metrics.structural_similarity, which the
    notebook calls: a win x win uniform window (symmetric extension at the
    borders), sample (co)variances, C1 = (k1 R)^2, C2 = (k2 R)^2, and the mean
    over the image minus a (win - 1) / 2 border."""
    from scipy.ndimage import uniform_filter
    X = np.asarray(img, dtype=np.float64)
    Y = np.asarray(ref, dtype=np.float64)
    n = win * win
    cov = n / (n - 1.0)
    ux, uy = uniform_filter(X, win), uniform_filter(Y, win)
    vx = cov * (uniform_filter(X * X, win) - ux * ux)
    vy = cov * (uniform_filter(Y * Y, win) - uy * uy)
    vxy = cov * (uniform_filter(X * Y, win) - ux * uy)
    c1, c2 = (k1 * data_range) ** 2, (k2 * data_range) ** 2
    S = ((2 * ux * uy + c1) * (2 * vxy + c2)) / ((ux * ux + uy * uy + c1) * (vx + vy + c2))
    pad = (win - 1) // 2
    return float(S[pad:-pad, pad:-pad].mean())


def experiment(niter=64, every=2, samples=200, points=3000, log=print, at=None):
    """The notebook's accuracy experiment on the device (see the module
    docstring); returns per-iteration PSNR / SSIM for B0 and B1, evaluated
    every `every` iterations (or at the iteration counts in `at`)."""
    import torch

    from .solvers import cgls
    from .torch_ops import Projector
    g = NOTEBOOK
    dev = torch.device("cuda", torch.cuda.current_device())
    t0 = time.perf_counter()
    sino = fan_sinogram(g["views"], g["nu"], g["du"], g["lsd"], g["lso"], samples=samples, mode="log_mean_exp",
                        device=dev)
    t_sino = time.perf_counter() - t0
    xs, ys = eval_grid(points)
    base = raster_grid(xs[0], ys[:, 0])
    log("sinogram %.1f s, reference raster mean %.4f max %.3f" % (t_sino, base.mean(), base.max()))
    ang = 2 * np.pi * np.arange(g["views"]) / g["views"]
    b = torch.tensor(sino, dtype=torch.float32, device=dev)
    out = dict(workload=dict(g, samples=samples, sinogram="log_mean_exp", eval_points=points, data_range=2.0),
               seconds=dict(sinogram=t_sino), bases={})
    for basis in ("B0", "B1"):
        P = Projector(basis, g["nx"], g["ny"], ang, g["nu"], g["dx"], g["dy"], g["du"], g["lsd"], g["lso"])
        curve = []
        t1 = time.perf_counter()

        def on_iter(k, x):
            if (at is not None and k not in at) or (at is None and k % every):
                return
            img_xy = x.double().cpu().numpy().T  # [ix, iy]
            sampler = sample_nearest if basis == "B0" else sample_bilinear
            ev = sampler(img_xy, g["nx"], g["ny"], g["dx"], g["dy"], xs, ys)
            rec = dict(iters=k, psnr=psnr(ev, base, 2.0), ssim=ssim(ev, base, 2.0))
            curve.append(rec)
            log("%s iter %2d  PSNR %.2f dB  SSIM %.4f" % (basis, k, rec["psnr"], rec["ssim"]))

        res = cgls(P, b, niter, callback=on_iter)
        best = max(curve, key=lambda r: r["psnr"])
        out["bases"][basis] = dict(curve=curve, best_psnr=best, best_ssim=max(curve, key=lambda r: r["ssim"]),
                                   residuals=res.residuals, seconds=time.perf_counter() - t1)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="FORBILD CGLS experiment (B0, B1; PSNR / SSIM vs iterations)")
    ap.add_argument("--niter", type=int, default=64)
    ap.add_argument("--samples", type=int, default=200)
    ap.add_argument("--points", type=int, default=3000)
    ap.add_argument("--out", default="forbild_report.json")
    a = ap.parse_args(argv)
    out = experiment(a.niter, samples=a.samples, points=a.points, log=lambda m: print(m, flush=True))
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    summary = {k: dict(best_psnr=v["best_psnr"], best_ssim=v["best_ssim"]) for k, v in out["bases"].items()}
    print(json.dumps(summary), flush=True)
    return out


if __name__ == "__main__":
    main()
