"""SIRT and ASD-POCS on the device, built on the gfx950 A / A^T kernels.

The reference's optimizer/asd_pocs.cu is an empty file and its notebooks
reconstruct with pylops CGLS (python/main copy 2.ipynb cell 14); these solvers
are built from the published algorithms:

* SIRT: x <- max(x + lam * V * A^T(W * (b - A x)), 0), W = 1/(A 1), V = 1/(A^T 1).
* ASD-POCS (E. Y. Sidky and X. Pan, "Image reconstruction in circular
  cone-beam computed tomography by constrained, total-variation minimization",
  Phys. Med. Biol. 53 (2008) 4777, Fig. 2): a data-consistency step (here the
  SIRT step with relaxation beta), positivity, then ng steepest-descent steps
  on the total variation whose length dtvg adapts so the TV change stays below
  r_max times the data-step change; beta and dtvg decay geometrically.

All per-pixel work runs in hand-written kernels (sgdct_sirt_residual,
sgdct_sirt_update, sgdct_tv_gradient, sgdct_axpy); norms are torch reductions.
Iterations stop when the relative data residual ||A x - b|| / ||b|| reaches
`tol` (or after `niter`).
"""
import ctypes
import time

import torch

from . import _lib


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream(t):
    return ctypes.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


class SirtWeights:
    """W = 1 / (A 1) on the sinogram, V = 1 / (A^T 1) on the image (zero where
    the sums vanish)."""

    def __init__(self, P, eps=1e-8):
        dev = torch.device("cuda", torch.cuda.current_device())
        one_img = torch.ones(P.img_shape, device=dev)
        one_sino = torch.ones(P.sino_shape, device=dev)
        rs = P.forward(one_img)
        cs = P.backward(one_sino)
        self.W = torch.where(rs > eps, 1.0 / rs.clamp_min(eps), torch.zeros_like(rs)).contiguous()
        self.V = torch.where(cs > eps, 1.0 / cs.clamp_min(eps), torch.zeros_like(cs)).contiguous()


def sirt_step(P, w, x, b, lam=1.0, positivity=True, ax=None):
    """One SIRT update of x (in place); returns A x before the update."""
    L = _lib.lib()
    if ax is None:
        ax = P.forward(x)
    r = torch.empty_like(ax)
    # W and V hold one sinogram / image; a batch repeats them every `per` values
    n, per = ax.numel(), w.W.numel()
    if n % per or b.numel() != n:
        raise ValueError("sirt_step: sinogram of %d values does not hold whole copies of the %d-value weights"
                         % (n, per))
    _lib.check(L.sgdct_sirt_residual(_vp(b), _vp(ax), _vp(w.W), _vp(r), n, per, _stream(x)), "sirt_residual")
    c = P.backward(r)
    n, per = x.numel(), w.V.numel()
    if n % per:
        raise ValueError("sirt_step: image of %d values does not hold whole copies of the %d-value weights" % (n, per))
    _lib.check(L.sgdct_sirt_update(_vp(x), _vp(c), _vp(w.V), float(lam), int(bool(positivity)), n, per, _stream(x)),
               "sirt_update")
    return ax


def rel_residual(P, x, b):
    ax = P.forward(x)
    return float(torch.linalg.vector_norm((ax - b).double()) / torch.linalg.vector_norm(b.double()))


def sirt(P, b, niter, x0=None, lam=1.0, positivity=True, weights=None):
    """niter SIRT iterations; returns (x, list of relative residuals before each step)."""
    b = b.contiguous().float()
    w = weights or SirtWeights(P)
    x = torch.zeros(P.img_shape, device=b.device) if x0 is None else x0.clone().contiguous().float()
    hist = []
    bn = float(torch.linalg.vector_norm(b.double()))
    for _ in range(niter):
        ax = sirt_step(P, w, x, b, lam, positivity)
        hist.append(float(torch.linalg.vector_norm((ax - b).double())) / bn)
    return x, hist


def tv(x, eps=1e-8):
    """Smoothed isotropic TV (the functional whose gradient tv_gradient computes)."""
    gx = torch.zeros_like(x)
    gy = torch.zeros_like(x)
    gx[..., :, 1:] = x[..., :, 1:] - x[..., :, :-1]
    gy[..., 1:, :] = x[..., 1:, :] - x[..., :-1, :]
    return torch.sqrt(eps + gx * gx + gy * gy).sum()


def tv_gradient(x, eps=1e-8, out=None):
    x = x.contiguous()
    g = torch.empty_like(x) if out is None else out
    ny, nx = x.shape[-2:]
    batch = x.numel() // (nx * ny)
    _lib.check(_lib.lib().sgdct_tv_gradient(_vp(x), _vp(g), nx, ny, batch, float(eps), _stream(x)), "tv_gradient")
    return g


def tv_descent(x, step, nsteps, eps=1e-8, g=None, work=None):
    """In place on x (ny, nx) or (B, ny, nx): `nsteps` steps
    x <- x - step * grad TV(x) / ||grad TV(x)|| per image, run on the device
    without host synchronisation (`step`: one FP64 value per image, a device
    tensor; 0 leaves the image unchanged, as does a zero gradient)."""
    ny, nx = x.shape[-2:]
    batch = x.numel() // (nx * ny)
    g = torch.empty_like(x) if g is None else g
    work = torch.empty(64 * batch, dtype=torch.float64, device=x.device) if work is None else work
    step = step.to(torch.float64).contiguous()
    assert x.is_contiguous() and g.shape == x.shape and step.numel() == batch and work.numel() >= 64 * batch
    _lib.check(_lib.lib().sgdct_tv_descent(_vp(x), _vp(g), _vp(work), _vp(step), nx, ny, batch, float(eps),
                                           int(nsteps), _stream(x)), "tv_descent")
    return x


class AsdPocsResult:
    """`iters`: iterations run before the stopping check that met `tol` (the
    data steps that produced `x`, counting an abandoned warm-start iteration),
    or niter; `residuals`: the relative residual entering each iteration, NaN
    for the iteration that switched to the fallback image."""

    def __init__(self, x, iters, residuals, seconds, fell_back):
        self.x, self.iters, self.residuals, self.seconds, self.fell_back = x, iters, residuals, seconds, fell_back


def asd_pocs(P, b, niter=200, tol=None, x0=None, beta=1.0, beta_red=0.995, ng=20, alpha=0.2, r_max=0.95,
             alpha_red=0.95, eps_data=0.0, tv_eps=1e-8, weights=None, fallback=None):
    """ASD-POCS (Sidky & Pan 2008).  alpha is the TV step ratio (dtvg =
    alpha * |data step| at the first iteration).  Stops after niter iterations
    or when ||A x - b|| / ||b|| <= tol.  If `fallback` is an image and the
    relative residual rises above the previous iteration's during the first
    iterations from x0, the run restarts from `fallback` (the warm start is
    abandoned) -- the plain method then decides when to stop.  That iteration
    takes no TV step, counts as an iteration and decays beta like any other
    (asd_pocs_batch follows the same rules)."""
    t0 = time.perf_counter()
    b = b.contiguous().float()
    w = weights or SirtWeights(P)
    x = torch.zeros(P.img_shape, device=b.device) if x0 is None else x0.clone().contiguous().float().clamp_min(0)
    bn = float(torch.linalg.vector_norm(b.double()))
    dtvg = None
    hist = []
    fell_back = False
    g = torch.empty_like(x)
    step = torch.zeros(1, dtype=torch.float64, device=b.device)
    work = torch.empty(64, dtype=torch.float64, device=b.device)
    k = 0
    while k < niter:
        x_prev = x.clone()
        ax = sirt_step(P, w, x, b, beta, positivity=True)
        res = float(torch.linalg.vector_norm((ax - b).double())) / bn
        hist.append(res)
        if tol is not None and res <= tol:
            x.copy_(x_prev)  # the image whose residual met the tolerance
            break
        if fallback is not None and not fell_back and len(hist) >= 2 and hist[-1] > hist[-2]:
            x = fallback.clone().contiguous().float()
            fell_back = True
            dtvg = None
            hist.append(float("nan"))
            beta *= beta_red
            k += 1
            continue
        dp = float(torch.linalg.vector_norm((x - x_prev).double()))
        if dtvg is None:
            dtvg = alpha * dp
        x_tv0 = x.clone()
        step.fill_(dtvg)
        tv_descent(x, step, ng, tv_eps, g=g, work=work)
        dg = float(torch.linalg.vector_norm((x - x_tv0).double()))
        if dg > r_max * dp and res > eps_data:
            dtvg *= alpha_red
        beta *= beta_red
        k += 1
    torch.cuda.synchronize()
    return AsdPocsResult(x, k, hist, time.perf_counter() - t0, fell_back)


def _norms(t):
    return torch.linalg.vector_norm(t.double(), dim=tuple(range(1, t.dim())))


class AsdPocsBatchResult:
    def __init__(self, x, iters, residuals, fell_back, seconds, history=None):
        self.x, self.iters, self.residuals, self.fell_back, self.seconds = x, iters, residuals, fell_back, seconds
        self.history = history


def asd_pocs_batch(P, b, niter=200, tol=None, x0=None, fallback=None, beta=1.0, beta_red=0.995, ng=20, alpha=0.2,
                   r_max=0.95, alpha_red=0.95, eps_data=0.0, tv_eps=1e-8, weights=None, check_every=10,
                   history=False):
    """ASD-POCS (Sidky & Pan 2008) on a batch of sinograms b (B, np, nu) with
    per-image TV step control and stopping.  An image stops at the first
    iteration whose entering residual ||A x - b|| / ||b|| is <= tol (a number,
    or one per image as a (B,) tensor; its iteration count is the number of
    data steps taken before that).  With
    `fallback` (B, ny, nx), an image whose residual rises between two
    iterations of its warm-started run restarts once from its fallback image.
    The data step's relaxation beta follows one schedule for the batch.
    With history=True (and no tol) the result's `history` holds every
    iteration's residual, (niter, B): entry k is the residual of the iterate
    after k iterations (entry 0 that of the starting image).

    Stopped images leave the working batch at the next check (every
    `check_every` iterations), so the work per iteration follows the number
    of images still running.  Every image's iterates are the ones it would
    have alone: the projector gives each image of a batch the result it gets
    by itself, and all other steps are per image."""
    t0 = time.perf_counter()
    b = b.contiguous().float()
    nb = b.shape[0]
    w = weights or SirtWeights(P)
    dev = b.device
    x_out = torch.zeros((nb,) + tuple(P.img_shape), device=dev) if x0 is None else \
        x0.clone().contiguous().float().clamp_min(0)
    iters = torch.full((nb,), niter, dtype=torch.int64, device=dev)
    last_res = torch.full((nb,), float("nan"), dtype=torch.float64, device=dev)
    fell = torch.zeros(nb, dtype=torch.bool, device=dev)
    # working set: the images still running, idx = their positions in the batch
    idx = torch.arange(nb, device=dev)
    x = x_out.clone()
    bw = b
    bn = _norms(b)
    fb = fallback.float().contiguous() if fallback is not None else None
    fellw = fell.clone()
    active = torch.ones(nb, dtype=torch.bool, device=dev)
    prev_res = None
    dtvg = torch.full((nb,), float("nan"), dtype=torch.float64, device=dev)
    tolw = None if tol is None else torch.as_tensor(tol, dtype=torch.float64, device=dev).expand(nb).clone()
    g = torch.empty_like(x)
    if history and tol is not None:
        raise ValueError("asd_pocs_batch: history needs tol=None (every image runs all iterations)")
    hist = [] if history else None
    for k in range(niter):
        x_prev = x.clone()
        ax = sirt_step(P, w, x, bw, beta, positivity=True)
        res = _norms(ax - bw) / bn
        if hist is not None:
            hist.append(res)
        last_res[idx] = torch.where(active, res, last_res[idx])
        if tolw is not None:
            done = active & (res <= tolw)
            iters[idx] = torch.where(done, torch.full_like(idx, k), iters[idx])
            active = active & ~done
        rise = torch.zeros_like(active)
        if fb is not None and prev_res is not None:
            # an image that switches to its fallback takes no TV step in this
            # iteration and sets its TV step length at its next data step
            rise = active & ~fellw & (res > prev_res)
            x = torch.where(rise[:, None, None], fb, x)
            fellw = fellw | rise
            dtvg = torch.where(rise, torch.full_like(dtvg, float("nan")), dtvg)
        prev_res = res
        x = torch.where(active[:, None, None], x, x_prev).contiguous()
        dp = _norms(x - x_prev)
        dtvg = torch.where(torch.isnan(dtvg) & ~rise, alpha * dp, dtvg)
        x_tv0 = x.clone()
        tv_descent(x, torch.where(active & ~rise, dtvg, torch.zeros_like(dtvg)), ng, tv_eps, g=g)
        dg = _norms(x - x_tv0)
        dtvg = torch.where((dg > r_max * dp) & (res > eps_data), dtvg * alpha_red, dtvg)
        beta *= beta_red
        if tol is not None and (k + 1) % check_every == 0:
            n_act = int(active.sum())
            if n_act == 0:
                break
            if n_act < idx.numel():
                x_out[idx] = x
                fell[idx] = fellw
                keep = active.nonzero().squeeze(1)
                idx, x, bw, bn = idx[keep], x[keep].contiguous(), bw[keep].contiguous(), bn[keep]
                dtvg, prev_res, fellw, active = dtvg[keep], prev_res[keep], fellw[keep], active[keep]
                tolw = tolw[keep]
                fb = fb[keep].contiguous() if fb is not None else None
                g = torch.empty_like(x)
    x_out[idx] = x
    fell[idx] = fellw
    torch.cuda.synchronize()
    return AsdPocsBatchResult(x_out, iters.cpu().numpy(), last_res.cpu().numpy(), fell.cpu().numpy(),
                              time.perf_counter() - t0,
                              torch.stack(hist).cpu().numpy() if hist else None)


def sirt_batch(P, b, niter, weights=None, x0=None, lam=1.0):
    """niter SIRT iterations on a batch of sinograms (B, np, nu)."""
    b = b.contiguous().float()
    w = weights or SirtWeights(P)
    x = torch.zeros((b.shape[0],) + tuple(P.img_shape), device=b.device) if x0 is None else x0.clone().float()
    for _ in range(niter):
        sirt_step(P, w, x, b, lam, positivity=True)
    return x


class CglsResult:
    def __init__(self, x, residuals, seconds):
        self.x, self.residuals, self.seconds = x, residuals, seconds


def cgls(P, b, niter, x0=None, callback=None):
    """CGLS (Hestenes-Stiefel conjugate gradients on A^T A x = A^T b, in the
    least-squares form of Paige & Saunders): the Krylov iterates that LSQR --
    what pylops' regularized_inversion runs for the reference's notebook
    reconstructions (python/main copy 2.ipynb cells 14, 27-28) -- produces in
    exact arithmetic.  Vectors FP32 on the device, inner products in FP64.
    callback(k, x) after iteration k; returns the iterate and the relative
    residuals ||b - A x_k|| / ||b||, k = 0..niter."""
    t0 = time.perf_counter()
    b = b.contiguous().float()
    x = torch.zeros(P.img_shape, device=b.device) if x0 is None else x0.clone().contiguous().float()
    r = b - P.forward(x) if x0 is not None else b.clone()
    s = P.backward(r)
    p = s.clone()
    gamma = float(torch.dot(s.double().ravel(), s.double().ravel()))
    bn = float(torch.linalg.vector_norm(b.double()))
    hist = [float(torch.linalg.vector_norm(r.double())) / bn]
    for k in range(1, niter + 1):
        q = P.forward(p)
        qq = float(torch.dot(q.double().ravel(), q.double().ravel()))
        if qq == 0.0 or gamma == 0.0:
            break
        alpha = gamma / qq
        x.add_(p, alpha=alpha)
        r.sub_(q, alpha=alpha)
        s = P.backward(r)
        gamma_new = float(torch.dot(s.double().ravel(), s.double().ravel()))
        p.mul_(gamma_new / gamma).add_(s)
        gamma = gamma_new
        hist.append(float(torch.linalg.vector_norm(r.double())) / bn)
        if callback is not None:
            callback(k, x)
    torch.cuda.synchronize()
    return CglsResult(x, hist, time.perf_counter() - t0)
