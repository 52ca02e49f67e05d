"""Learned warm start for ASD-POCS: data generation, training, evaluation.

Workload (the sparse-view reconstruction of the learned-warm-start
configuration): 512 x 512 images of 0.6 mm pixels, 90 views over 2 pi, 768
bins of 0.8 mm, B1 basis, source-isocentre 500 mm, source-detector 1000 mm;
random-ellipse phantoms (sgd_amd.phantom's generator) scaled by
mu = 0.02 / mm, Poisson noise at 5e4 incident photons per ray,
b = -ln(max(counts, 1) / N0) / mu.

  pairs     input  = SIRT(5) of the noisy sinogram,
            target = ASD-POCS(200) of the same sinogram (sgd_amd.solvers);
  network   sgd_amd.unet.UNet (4 levels, 32 base channels, bf16 / FP32-acc.);
  use       x0 = UNet(SIRT(5)) seeds ASD-POCS, which runs until the relative
            data residual ||A x - b|| / ||b|| reaches the tolerance and
            restarts from the SIRT(5) image if its residual rises (the plain
            method still decides when to stop: the stopping rule is the same).

Tolerance.  The Poisson noise of this workload is 0.7 % of ||b|| (relative
||b - b_clean|| / ||b|| of 0.0068..0.009 on seeded phantoms), and with 90 x
768 rays against 512^2 unknowns no iterate of these methods comes near 1e-3:
unregularised SIRT without the positivity constraint is at 2.3e-3 after 3000
iterations, with positivity at 3.6e-3, and ASD-POCS with the controls below
at 5e-3..8e-3 (floor_study, `--phase floor`).  The evaluation therefore stops
each image at a multiple tau of its own noise level (the discrepancy
principle; the level is estimated from the sinogram by the Poisson model,
noise_level), tau = 2 and 1.5, and also reports the fixed tolerance 1e-2,
which a few images cannot reach within 3000 iterations; --tol sets others.

`python -m sgd_amd.warmstart --help` runs the experiment (or one phase of it)
and writes a JSON report.  No reference counterpart (the reference's
optimizer/asd_pocs.cu is empty).
"""
import argparse
import json
import time

import numpy as np
import torch

from .phantom import ellipse_params
from .solvers import SirtWeights, asd_pocs_batch, sirt_batch
from .torch_ops import Projector
from .unet import UNet

GEOMETRY = dict(nx=512, ny=512, dx=0.6, dy=0.6, nu=768, du=0.8, lsd=1000.0, lso=500.0, views=90)
MU = 0.02
N0 = 5e4
# ASD-POCS controls: TV step fraction alpha, its reduction alpha_red when the
# TV change exceeds r_max times the data change, ng TV steps per iteration,
# data-step relaxation beta reduced by beta_red per iteration.  With these the
# plain method reaches 1e-2 in 170..410 iterations on this workload.
ASD = dict(alpha=0.02, r_max=0.5, ng=20, beta=1.0, beta_red=0.999, alpha_red=0.95)


def projector(geom=GEOMETRY):
    ang = 2 * np.pi * np.arange(geom["views"]) / geom["views"]
    return Projector("B1", geom["nx"], geom["ny"], ang, geom["nu"], geom["dx"], geom["dy"], geom["du"], geom["lsd"],
                     geom["lso"])


def phantoms(seeds, nx, ny, dx, dy, device):
    """Random-ellipse phantoms (the generator of sgd_amd.phantom), rasterised
    on the device: (len(seeds), ny, nx) FP32."""
    xs = (torch.arange(nx, device=device, dtype=torch.float64) - 0.5 * (nx - 1)) * dx
    ys = (torch.arange(ny, device=device, dtype=torch.float64) - 0.5 * (ny - 1)) * dy
    Y, X = torch.meshgrid(ys, xs, indexing="ij")
    fov_half = 0.5 * min(nx * dx, ny * dy)
    out = torch.zeros((len(seeds), ny, nx), device=device, dtype=torch.float64)
    for i, s in enumerate(seeds):
        for cx, cy, a, b, ang, d in ellipse_params(int(s), 10, fov_half):
            c, sn = np.cos(ang), np.sin(ang)
            xr = (X - cx) * c + (Y - cy) * sn
            yr = -(X - cx) * sn + (Y - cy) * c
            out[i] += d * (((xr / a) ** 2 + (yr / b) ** 2) <= 1.0)
    return out.clamp_min_(0.0).float()


def noisy_sinograms(P, x, seed, n0=N0, mu=MU):
    """b = -ln(max(Poisson(n0 exp(-mu A x)), 1) / n0) / mu."""
    y = P.forward(x).double()
    g = torch.Generator(device=x.device).manual_seed(int(seed))
    counts = torch.poisson(n0 * torch.exp(-mu * y), generator=g).clamp_min_(1.0)
    return (-torch.log(counts / n0) / mu).float()


def noise_level(b, n0=N0, mu=MU):
    """Relative noise of each sinogram of b (B, np, nu) under the Poisson
    model of noisy_sinograms: Var(b) = exp(mu b) / (mu^2 n0) per ray, so
    sigma = sqrt(sum exp(mu b) / (mu^2 n0)) / ||b||, estimated from the data."""
    bd = b.double()
    var = torch.exp(mu * bd).sum(dim=(1, 2)) / (mu * mu * n0)
    return torch.sqrt(var) / torch.linalg.vector_norm(bd, dim=(1, 2))


def parse_tol(spec):
    """'2s' -> ('2s', 2.0, True): 2 x each image's noise level; '1e-2' ->
    ('1e-2', 0.01, False): a fixed relative residual."""
    spec = spec.strip()
    if spec.endswith("s"):
        return spec, float(spec[:-1]), True
    return spec, float(spec), False


def make_pairs(P, seeds, batch=64, sirt_iters=5, asd_iters=200, log=print):
    """(inputs, targets, sinograms) for the phantoms of `seeds`."""
    w = SirtWeights(P)
    dev = P.device
    ins, outs, sinos = [], [], []
    t0 = time.perf_counter()
    for k in range(0, len(seeds), batch):
        sb = seeds[k:k + batch]
        x = phantoms(sb, P.nx, P.ny, P.geom["dx"], P.geom["dy"], dev)
        b = noisy_sinograms(P, x, seed=int(sb[0]) + 7919)
        ins.append(sirt_batch(P, b, sirt_iters, weights=w))
        outs.append(asd_pocs_batch(P, b, niter=asd_iters, weights=w, **ASD).x)
        sinos.append(b)
        log("pairs %d/%d  %.1fs" % (k + len(sb), len(seeds), time.perf_counter() - t0))
    return torch.cat(ins), torch.cat(outs), torch.cat(sinos)


def train(net, inputs, targets, epochs=8, batch=8, lr=1e-3, seed=0, log=print, P=None, sinos=None,
          data_weight=0.0, weights=None):
    """Adam on mean((y - target)^2) with a cosine learning rate.  With
    data_weight = lam > 0 (and the projector P and the pairs' sinograms) the
    output gradient becomes that of
        mean((y - t)^2) + lam * (the SIRT-metric data misfit of y),
    i.e. (y - t) + lam * g with g = V A^T W (A y - b), the negative of one
    unconstrained SIRT step from y: the network is steered towards outputs
    whose data residual is small, not only towards the target pixels."""
    from .solvers import sirt_step
    n = inputs.shape[0]
    if data_weight > 0 and (P is None or sinos is None):
        raise ValueError("train: data_weight > 0 needs the projector P and the sinograms")
    w = (weights or SirtWeights(P)) if data_weight > 0 else None
    rng = np.random.default_rng(seed)
    losses = []
    steps = epochs * (n // batch)
    t0 = time.perf_counter()
    s = 0
    for ep in range(epochs):
        order = rng.permutation(n)
        run = torch.zeros((), dtype=torch.float64, device=inputs.device)
        nstep = 0
        for k in range(0, n - batch + 1, batch):
            idx = torch.as_tensor(order[k:k + batch], device=inputs.device)
            lr_s = lr * 0.5 * (1 + np.cos(np.pi * s / steps))
            t = targets[idx]
            if data_weight > 0:
                y = net.forward(inputs[idx])
                z = y.clone()
                sirt_step(P, w, z, sinos[idx].contiguous(), 1.0, positivity=False)  # z = y - g
                net.backward(t + data_weight * (z - y))
                net.step(lr_s)
            else:
                y = net.train_step(inputs[idx], t, lr=lr_s)
            loss = ((y - t).double() ** 2).mean()
            run += loss
            nstep += 1
            if s % 50 == 0:
                losses.append((s, float(loss)))
            s += 1
        log("epoch %d  mean loss %.4e  %.1fs" % (ep, float(run) / max(nstep, 1), time.perf_counter() - t0))
    return losses


def evaluate(P, net, b, tols=("2s", "1.5s", "1e-2"), max_iters=3000, batch=64, sirt_iters=5, log=print,
             plain_cache=None):
    """Plain ASD-POCS (from zero) vs warm-started (from UNet(SIRT(5)), with
    the SIRT(5) image as fallback) to each relative-residual tolerance in
    `tols` (see parse_tol: a number, or a multiple of each image's own noise
    level, the discrepancy principle): per-image iterations and wall times per
    method and tolerance.  The warm time includes SIRT(5) and the network.
    A dict passed as plain_cache keeps the plain runs for the next call on
    the same sinograms (the plain method does not depend on the network)."""
    w = SirtWeights(P)
    specs = [parse_tol(t) for t in tols]
    res = {name: dict(plain_iters=[], warm_iters=[], plain_s=0.0, warm_s=0.0, fell_back=0, tol=[])
           for name, _, _ in specs}
    for k in range(0, b.shape[0], batch):
        bb = b[k:k + batch]
        sigma = noise_level(bb)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        xs = sirt_batch(P, bb, sirt_iters, weights=w)
        x0 = net.forward(xs).clone()
        torch.cuda.synchronize()
        t_seed = time.perf_counter() - t0
        for name, val, rel in specs:
            tol = val * sigma if rel else val
            key = (k, name, max_iters)
            if plain_cache is not None and key in plain_cache:
                r = plain_cache[key]
            else:
                r = asd_pocs_batch(P, bb, niter=max_iters, tol=tol, weights=w, **ASD)
                r.x = None
                if plain_cache is not None:
                    plain_cache[key] = r
            r2 = asd_pocs_batch(P, bb, niter=max_iters, tol=tol, x0=x0, fallback=xs, weights=w, **ASD)
            e = res[name]
            e["plain_s"] += r.seconds
            e["warm_s"] += t_seed + r2.seconds
            e["plain_iters"] += r.iters.tolist()
            e["warm_iters"] += r2.iters.tolist()
            e["fell_back"] += int(r2.fell_back.sum())
            e["tol"] += (tol.tolist() if rel else [tol] * bb.shape[0])
            log("eval %d/%d tol %s  plain %.0f its %.1fs  warm %.0f its %.1fs" % (
                k + bb.shape[0], b.shape[0], name, np.mean(r.iters), r.seconds, np.mean(r2.iters),
                t_seed + r2.seconds))
    return res


def summarize(ev, max_iters):
    out = {}
    for name, e in ev.items():
        pi, wi = np.array(e["plain_iters"]), np.array(e["warm_iters"])
        both = (pi < max_iters) & (wi < max_iters)
        out[name] = dict(
            tol_min=float(np.min(e["tol"])), tol_max=float(np.max(e["tol"])),
            plain_mean=float(pi.mean()), warm_mean=float(wi.mean()),
            plain_median=float(np.median(pi)), warm_median=float(np.median(wi)),
            plain_reached=int((pi < max_iters).sum()), warm_reached=int((wi < max_iters).sum()),
            both_reached_plain_mean=float(pi[both].mean()) if both.any() else None,
            both_reached_warm_mean=float(wi[both].mean()) if both.any() else None,
            plain_s=e["plain_s"], warm_s=e["warm_s"],
            speedup_iterations=float(pi.mean() / max(wi.mean(), 1e-9)),
            speedup_time=float(e["plain_s"] / max(e["warm_s"], 1e-9)), fell_back=e["fell_back"])
    return out


def curve_stats(hp, hw, sig, tol=1e-2, marks=(0, 1, 2, 5, 10, 20, 50, 100, 150, 200, 300, 500, 750, 1000, 1499)):
    """Statistics of residual curves hp, hw (iterations, images) of the plain
    and the warm-started run and the images' noise levels sig (images,):
    see residual_curves."""
    iters, n = hp.shape
    pmin = np.minimum.accumulate(hp, axis=0)  # non-increasing envelope of each plain curve
    marks = [m for m in marks if m < iters]
    out = dict(images=n, iters=iters, tol=tol, noise_level_min=float(sig.min()), noise_level_max=float(sig.max()),
               marks=marks, plain_median=[], warm_median=[], plain_median_over_noise=[],
               warm_median_over_noise=[], offset_median=[], offset_censored=[])
    for m in marks:
        out["plain_median"].append(float(np.median(hp[m])))
        out["warm_median"].append(float(np.median(hw[m])))
        out["plain_median_over_noise"].append(float(np.median(hp[m] / sig)))
        out["warm_median_over_noise"].append(float(np.median(hw[m] / sig)))
        d = []
        for i in range(n):
            j = np.nonzero(pmin[:, i] <= hw[m, i])[0]
            if j.size:
                d.append(int(j[0]) - m)
        # median over the images whose plain run gets as low within iters
        out["offset_median"].append(float(np.median(d)) if d else None)
        out["offset_censored"].append(n - len(d))

    def first(h):
        r = np.full(n, iters, dtype=np.int64)
        for i in range(n):
            j = np.nonzero(h[:, i] <= tol)[0]
            if j.size:
                r[i] = j[0]
        return r
    fp, fw = first(hp), first(hw)
    both = (fp < iters) & (fw < iters)
    out.update(tol_plain_reached=int((fp < iters).sum()), tol_warm_reached=int((fw < iters).sum()),
               tol_plain_mean=float(fp.mean()), tol_warm_mean=float(fw.mean()),
               tol_plain_median=float(np.median(fp)), tol_warm_median=float(np.median(fw)),
               tol_both_plain_mean=float(fp[both].mean()) if both.any() else None,
               tol_both_warm_mean=float(fw[both].mean()) if both.any() else None,
               tol_saved_median=float(np.median((fp - fw)[both])) if both.any() else None,
               final_plain_over_noise=[float(v) for v in np.quantile(hp[-1] / sig, (0.0, 0.5, 1.0))])
    return out


def residual_curves(P, net, b, iters=1500, batch=64, sirt_iters=5, tol=1e-2,
                    marks=(0, 1, 2, 5, 10, 20, 50, 100, 150, 200, 300, 500, 750, 1000, 1499), log=print,
                    plain_cache=None):
    """Residual curves of plain and warm-started ASD-POCS run for `iters`
    iterations without stopping (same controls and seeds as evaluate), and
    what they say about a fixed tolerance:

      median residual and residual / own noise level at the `marks`;
      offset d(k) = E(k) - k, where E(k) is the first plain iteration whose
        residual is at or below the warm residual after k iterations (the
        plain iterations the warm start has saved by iteration k); censored
        when the plain run never gets that low within `iters` (the median is
        over the other images, None when all are censored);
      iterations to `tol` read off the curves (first k with residual <= tol),
        for both methods.
    plain_cache: as in evaluate."""
    w = SirtWeights(P)
    hp, hw, sig = [], [], []
    for k in range(0, b.shape[0], batch):
        bb = b[k:k + batch]
        sig.append(noise_level(bb).cpu().numpy())
        xs = sirt_batch(P, bb, sirt_iters, weights=w)
        x0 = net.forward(xs).clone()
        key = ("curve", k, iters)
        if plain_cache is not None and key in plain_cache:
            hp.append(plain_cache[key])
        else:
            hp.append(asd_pocs_batch(P, bb, niter=iters, weights=w, history=True, **ASD).history)
            if plain_cache is not None:
                plain_cache[key] = hp[-1]
        hw.append(asd_pocs_batch(P, bb, niter=iters, x0=x0, weights=w, history=True, **ASD).history)
        log("curves %d/%d" % (k + bb.shape[0], b.shape[0]))
    hp, hw, sig = np.concatenate(hp, axis=1), np.concatenate(hw, axis=1), np.concatenate(sig)
    out = curve_stats(hp, hw, sig, tol=tol, marks=marks)
    log(json.dumps(out))
    return out


def floor_study(P, seeds=range(900, 908), iters=3000, log=print):
    """How far the data residual of these methods goes on this workload:
    the drawn noise level, SIRT with and without the positivity constraint and
    ASD-POCS (ASD controls) after `iters` iterations, per phantom."""
    from .solvers import _norms, sirt_step
    seeds = list(seeds)
    x = phantoms(seeds, P.nx, P.ny, P.geom["dx"], P.geom["dy"], P.device)
    b = noisy_sinograms(P, x, seed=3)
    clean = P.forward(x)
    w = SirtWeights(P)
    bn = _norms(b)
    out = dict(seeds=seeds, iters=iters, noise=(_norms(b - clean) / bn).tolist(),
               noise_estimate=noise_level(b).tolist())
    for pos in (True, False):
        xx = torch.zeros_like(x)
        for _ in range(iters):
            ax = sirt_step(P, w, xx, b, 1.0, positivity=pos)
        out["sirt_positivity" if pos else "sirt_free"] = (_norms(ax - b) / bn).tolist()
    r = asd_pocs_batch(P, b, niter=iters, weights=w, **ASD)
    out["asd_pocs"] = r.residuals.tolist()
    log(json.dumps(out))
    return out


def save_weights(net, path):
    p = net.parameters_numpy()
    np.savez(path, **{k + "_W": v[0] for k, v in p.items()}, **{k + "_b": v[1] for k, v in p.items()})


def load_weights(path, levels=4, c0=32):
    from .unet import layer_specs
    with np.load(path) as z:
        params = {name: (z[name + "_W"], z[name + "_b"]) for name, _, _, _ in layer_specs(levels, c0)}
    return UNet(levels, c0, params=params)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--phase", choices=["all", "train", "eval", "floor", "curves"], default="all",
                    help="train: generate pairs, train, save --weights; eval: load --weights, evaluate; "
                         "curves: load --weights, residual curves (--curves images); "
                         "floor: residual floor of SIRT / ASD-POCS on 8 phantoms")
    ap.add_argument("--weights", default=None, help="weights file (.npz) written by train / read by eval")
    ap.add_argument("--pairs", type=int, default=4096)
    ap.add_argument("--test", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--tol", default="2s,1.5s,1e-2",
                    help="comma-separated tolerances: relative residuals, or multiples of each image's noise "
                         "level with an 's' suffix")
    ap.add_argument("--max-iters", type=int, default=3000)
    ap.add_argument("--asd-iters", type=int, default=200)
    ap.add_argument("--gen-batch", type=int, default=64)
    ap.add_argument("--curves", type=int, default=0,
                    help="residual curves on this many test images (after eval with --phase all)")
    ap.add_argument("--curve-iters", type=int, default=1500)
    ap.add_argument("--data-weight", default="0",
                    help="comma-separated data-consistency weights of the training loss (see train); one "
                         "network is trained per weight on the same pairs and each is evaluated")
    ap.add_argument("--out", default="warmstart_report.json")
    a = ap.parse_args(argv)
    if a.phase in ("eval", "curves") and not a.weights:
        ap.error("--phase %s needs --weights" % a.phase)
    if a.phase == "curves" and a.curves <= 0:
        a.curves = 64

    def log(msg):
        print(msg, flush=True)

    P = projector()
    if a.phase == "floor":
        out = floor_study(P, iters=a.max_iters, log=log)
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
        return out
    tols = a.tol.split(",")
    report = dict(workload=dict(GEOMETRY, mu=MU, n0=N0, asd=ASD, pairs=a.pairs, test=a.test, tols=tols,
                                max_iters=a.max_iters, asd_iters_target=a.asd_iters, epochs=a.epochs,
                                batch=a.batch, lr=a.lr, data_weight=a.data_weight),
                  phase=a.phase, seconds={})
    dws = [float(v) for v in a.data_weight.split(",")]
    nets = []
    report["variants"] = {}
    if a.phase in ("all", "train"):
        t0 = time.perf_counter()
        xin, xt, sinos = make_pairs(P, list(range(a.pairs)), batch=a.gen_batch, asd_iters=a.asd_iters, log=log)
        report["seconds"]["generate"] = time.perf_counter() - t0
        for dw in dws:
            label = "data_weight=%g" % dw
            net = UNet(4, 32, seed=0)
            t0 = time.perf_counter()
            log("training, " + label)
            losses = train(net, xin, xt, epochs=a.epochs, batch=a.batch, lr=a.lr, log=log, P=P, sinos=sinos,
                           data_weight=dw)
            report["variants"][label] = dict(train_seconds=time.perf_counter() - t0, train_loss=losses[-5:])
            if a.weights:
                save_weights(net, a.weights if len(dws) == 1 else a.weights.replace(".npz", "_dw%g.npz" % dw))
            nets.append((label, net))
        del xin, xt, sinos
    else:
        nets.append(("weights=" + a.weights, load_weights(a.weights)))
        report["variants"][nets[0][0]] = {}
    if a.phase in ("all", "eval", "curves"):
        ntest = max(a.test, a.curves) if a.phase != "curves" else a.curves
        seeds = list(range(10 ** 6, 10 ** 6 + ntest))
        x = torch.cat([phantoms(seeds[k:k + 64], P.nx, P.ny, P.geom["dx"], P.geom["dy"], P.device)
                       for k in range(0, len(seeds), 64)])
        b = torch.cat([noisy_sinograms(P, x[k:k + 64], seed=seeds[k] + 7919) for k in range(0, len(seeds), 64)])
        del x
        plain_cache = {}
        for label, net in nets:
            v = report["variants"][label]
            if a.phase != "curves" and a.test > 0:
                t0 = time.perf_counter()
                ev = evaluate(P, net, b[:a.test], tols=tols, max_iters=a.max_iters, log=log, plain_cache=plain_cache)
                v["evaluate_seconds"] = time.perf_counter() - t0
                v["tolerances"] = summarize(ev, a.max_iters)
            if a.curves > 0:
                t0 = time.perf_counter()
                v["curves"] = residual_curves(P, net, b[:a.curves], iters=a.curve_iters, log=log,
                                              plain_cache=plain_cache)
                v["curves_seconds"] = time.perf_counter() - t0
    with open(a.out, "w") as f:
        json.dump(report, f, indent=1)
    log(json.dumps({k: report[k] for k in ("phase", "seconds", "variants") if k in report}))
    return report


if __name__ == "__main__":
    main()
