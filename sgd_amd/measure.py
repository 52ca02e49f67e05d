"""Timings of the workloads beside the headline pair (one JSON line each).

    python -m sgd_amd.measure fan     # every 2-D basis at its workload
    python -m sgd_amd.measure asd     # ASD-POCS, 50 iterations at the 512^2 / 720-view workload
    python -m sgd_amd.measure batch   # 256 uneven fan slices: per-slice launches,
                                      # persistent launch, autotuned, graph replay
    python -m sgd_amd.measure cone    # cone-beam B1 / SF-TR / SF-TT, 256^3 volume,
                                      # 360 views, 512 x 384 detector
    python -m sgd_amd.measure unet    # U-Net (4 levels, 32 channels) inference and
                                      # training step, batch 8 at 512 x 512
    python -m sgd_amd.measure conv    # the U-Net's 3x3 convolutions alone (32 -> 32 and
                                      # 64 -> 64 channels, batch 8 at 512 x 512)
    python -m sgd_amd.measure hexangle  # hexagonal A by view angle (1024 views within
                                        # 0.06 degrees of each angle, batch 8)
    python -m sgd_amd.measure hostabi   # the B1 pair through the reference-named host
                                        # entry points (FP64 host arrays, PCIe included)

Times are HIP-event averages on the launch stream with inputs resident in
device memory, except hostabi's (wall clock per call, host arrays in and out).
"""
import argparse
import json
import sys
import time

import numpy as np
import torch


def _time(fn, reps):
    # warm up for at least 20 calls and 50 ms: after the host-side set-up of a
    # projector the GPU has been idle and its clocks take tens of
    # milliseconds to come back (the first 30 B1 A calls on a new handle
    # average ~10 % longer than the same calls a moment later)
    t0 = time.perf_counter()
    n = 0
    while n < 20 or time.perf_counter() - t0 < 0.05:
        fn()
        torch.cuda.synchronize()
        n += 1
    s = torch.cuda.current_stream()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record(s)
    for _ in range(reps):
        fn()
    b.record(s)
    b.synchronize()
    return a.elapsed_time(b) / reps


def _maxrel(*pairs):
    """Largest elementwise |a - b| / max|b| over (a, b) pairs."""
    it = iter(pairs)
    return max(float((a - b).abs().max() / b.abs().max().clamp_min(1e-30)) for a, b in zip(it, it))


def batch(reps=3, tune_cache=None):
    from .batch import SliceBatch, make_projectors, uneven_slices
    t0 = time.perf_counter()
    B = SliceBatch(make_projectors(uneven_slices(256, seed=0)))
    build_s = time.perf_counter() - t0
    img = B.empty_images().uniform_()
    sino = B.empty_sinos().uniform_()
    out_s, out_i = B.empty_sinos(), B.empty_images()
    rec = dict(workload="256 uneven fan slices (B0/B1/SF, 128^2..768^2 pixels, 180..1440 views, 1.5 bins per "
                        "pixel edge), uniform random inputs",
               pixels=B.img_off[-1], bins=B.sino_off[-1], build_s=build_s)
    rec["per_slice_ms"] = dict(A=_time(lambda: B.forward_per_slice(img, out_s), reps),
                               AT=_time(lambda: B.backward_per_slice(sino, out_i), reps))
    ref_s, ref_i = out_s.clone(), out_i.clone()
    B.set_config(8, 8, 0)
    rec["persistent_default_ms"] = dict(A=_time(lambda: B.forward(img, out_s), reps),
                                        AT=_time(lambda: B.backward(sino, out_i), reps))
    rec["persistent_vs_per_slice_max_rel"] = _maxrel(out_s, ref_s, out_i, ref_i)
    t = B.tune(img, sino, cache_path=tune_cache, reps=2, force=True)
    rec["tuned"] = {k: t[k] for k in ("fwd_blocks_per_cu", "bwd_blocks_per_cu", "order", "fwd_ms", "bwd_ms")}
    rec["tune_trials"] = t["trials"]
    s = torch.cuda.Stream()  # graphs are captured on a created stream, not the null stream
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        rec["graph_ms"] = dict(A=_time(lambda: B.forward(img, out_s, graph=True), reps),
                               AT=_time(lambda: B.backward(sino, out_i, graph=True), reps))
    s.synchronize()
    rec["graph_vs_per_slice_max_rel"] = _maxrel(out_s, ref_s, out_i, ref_i)
    return rec


def fan(reps=10):
    """Every 2-D basis at its workload: B0 at 256^2 (1 mm) with 360 views and
    512 bins of 1 mm, B0 / B1 / SF at 512^2 (0.6 mm) with 720 views and 768
    bins of 0.8 mm, H0 / Hl on a 296 x 256 hexagonal lattice (dt 0.5 mm,
    1024 views, 1024 bins of 0.6 mm) for one image and a batch of 64; lso 500,
    lsd 1000 mm; uniform random inputs."""
    from .torch_ops import Projector
    recs = []
    cases = [("B0", 256, 256, 360, 512, 1.0, 1.0, 1), ("B0", 512, 512, 720, 768, 0.6, 0.8, 1),
             ("B1", 512, 512, 720, 768, 0.6, 0.8, 1), ("SF", 512, 512, 720, 768, 0.6, 0.8, 1),
             ("H0", 296, 256, 1024, 1024, 0.5, 0.6, 1), ("Hl", 296, 256, 1024, 1024, 0.5, 0.6, 1),
             ("H0", 296, 256, 1024, 1024, 0.5, 0.6, 64), ("Hl", 296, 256, 1024, 1024, 0.5, 0.6, 64)]
    for basis, n1, n2, nv, nu, d, du, batch in cases:
        ang = 2 * np.pi * np.arange(nv) / nv
        if basis in ("H0", "Hl"):
            P = Projector.hexagonal(basis, n1, n2, ang, nu, d, du, 1000.0, 500.0)
            shape = "%dx%d hexagonal lattice (dt %.1f mm)" % (n1, n2, d)
        else:
            P = Projector(basis, n2, n1, ang, nu, d, d, du, 1000.0, 500.0)
            shape = "%dx%d (%.1f mm)" % (n2, n1, d)
        x = torch.rand((batch,) + tuple(P.img_shape), device=P.device)
        y = torch.rand((batch,) + tuple(P.sino_shape), device=P.device)
        yo, xo = torch.empty_like(y), torch.empty_like(x)
        a_ms = _time(lambda: P.forward(x, out=yo), reps)
        at_ms = _time(lambda: P.backward(y, out=xo), reps)
        evals = P.eval_count()
        recs.append(dict(workload="%s fan beam, %s, %d views, %d bins (%.1f mm), batch %d" % (basis, shape, nv, nu, du,
                                                                                           batch),
                         basis=basis, batch=batch, A_ms=a_ms, AT_ms=at_ms, evals_per_image=evals,
                         A_evals_per_s=evals * batch / (a_ms * 1e-3)))
        del P
    return recs


def hostabi(reps=20):
    """The B1 pair at 512^2 (0.6 mm), 720 views, 768 bins (0.8 mm) through the
    reference-named host entry points (b1_forward_projection /
    b1_backward_projection via the pylops-compatible A_B1_operator): FP64 host
    arrays in and out, so each call includes the host-side rounding to FP32,
    the upload through pinned buffers, the kernels, the download and the
    widening back to FP64.  Wall clock per call after 5 untimed calls."""
    from .operators import A_B1_operator
    nx = ny = 512
    nv, nu = 720, 768
    op = A_B1_operator(nx, ny, nv, nu, 0.6, 0.6, 0.8, 1000.0, 500.0, 2 * np.pi * np.arange(nv) / nv)
    rng = np.random.default_rng(0)
    x, y = rng.random(nx * ny), rng.random(nv * nu)

    def wall(fn):
        for _ in range(5):
            fn()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        return (time.perf_counter() - t0) / reps * 1e3

    a_ms, at_ms = wall(lambda: op._matvec(x)), wall(lambda: op._rmatvec(y))
    op.close()
    return dict(workload="B1 fan beam, 512x512 (0.6 mm), 720 views, 768 bins (0.8 mm), FP64 host arrays through "
                         "b1_forward_projection / b1_backward_projection",
                A_ms=a_ms, AT_ms=at_ms, pairs_per_s=1e3 / (a_ms + at_ms),
                host_bytes_per_pair=8 * 2 * (nx * ny + nv * nu), pcie_bytes_per_pair=4 * 2 * (nx * ny + nv * nu))


def hexangle(reps=5, angles=(0, 5, 10, 15, 20, 25, 30, 45, 60, 90, 120, 150, 165)):
    """Hexagonal A (H0, Hl) by view angle: 1024 views within +-0.03 degrees
    of each angle, 296 x 256 lattice (dt 0.5 mm), 1024 bins of 0.6 mm, lso
    500 / lsd 1000 mm, batch 8 of uniform random images; ms per image."""
    from .torch_ops import Projector
    recs = []
    x = torch.rand((8, 296, 256), device=torch.device("cuda", torch.cuda.current_device()))
    for basis in ("Hl", "H0"):
        ms = {}
        for adeg in angles:
            ang = np.radians(adeg + np.linspace(-0.03, 0.03, 1024))
            P = Projector.hexagonal(basis, 296, 256, ang, 1024, 0.5, 0.6, 1000.0, 500.0)
            yo = torch.empty((8,) + tuple(P.sino_shape), device=x.device)
            ms[str(adeg)] = _time(lambda: P.forward(x, out=yo), reps) / 8
            del P
        v = np.array(list(ms.values()))
        recs.append(dict(workload="%s A by view angle, 296x256 hexagonal lattice (dt 0.5 mm), 1024 views within "
                                  "0.06 deg of each angle, 1024 bins (0.6 mm), batch 8" % basis,
                         basis=basis, A_ms_per_image=ms, max_over_median=float(v.max() / np.median(v))))
    return recs


def asd(niter=50):
    """Plain ASD-POCS on B1 data: 512^2 phantom (0.6 mm), 720 views,
    768 bins (0.8 mm), Poisson noise at 1e5 photons on p = 0.1 A x, 50
    iterations with TV step ratio alpha = 0.2 (Sidky & Pan's controls); SIRT
    beside it at the same iteration count."""
    from .phantom import phantom, poisson_noise
    from .solvers import SirtWeights, asd_pocs, sirt
    from .torch_ops import Projector
    ang = 2 * np.pi * np.arange(720) / 720
    P = Projector("B1", 512, 512, ang, 768, 0.6, 0.6, 0.8, 1000.0, 500.0)
    xt = torch.tensor(phantom(0, 512, 512, 0.6, 0.6), dtype=torch.float32, device=P.device)
    y = P.forward(xt).double().cpu().numpy()
    b = torch.tensor(poisson_noise(0.1 * y, 1e5, seed=1) / 0.1, dtype=torch.float32, device=P.device)
    w = SirtWeights(P)
    asd_pocs(P, b, niter=2, weights=w)  # warm-up
    r = asd_pocs(P, b, niter=niter, alpha=0.2, weights=w)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    xs, hist = sirt(P, b, niter, weights=w)
    torch.cuda.synchronize()
    ts = time.perf_counter() - t0
    err = lambda x: float(torch.linalg.vector_norm((x - xt).double()) / torch.linalg.vector_norm(xt.double()))
    return dict(workload="ASD-POCS, B1 fan beam, 512^2 (0.6 mm), 720 views, 768 bins (0.8 mm), Poisson 1e5, "
                         "%d iterations, TV step ratio 0.2, 20 TV steps per iteration" % niter,
                asd_seconds=r.seconds, asd_ms_per_iteration=1e3 * r.seconds / niter,
                asd_final_residual=r.residuals[-1], asd_image_error=err(r.x),
                sirt_seconds=ts, sirt_final_residual=hist[-1], sirt_image_error=err(xs))


def cone(reps=3):
    from .torch_ops import Projector
    ang = 2 * np.pi * np.arange(360) / 360
    recs = []
    for model in ("B1", "SF-TR", "SF-TT"):
        P = Projector.cone_beam(256, 256, 256, ang, 512, 384, 1.0, 1.0, 1.0, 1.0, 1.0, 1000.0, 500.0, model=model)
        x = torch.rand(P.img_shape, device=P.device)
        y = torch.rand(P.sino_shape, device=P.device)
        recs.append(dict(workload="cone beam %s, 256^3 voxels (1 mm), 360 views, 512 x 384 detector (1 mm), "
                                  "lsd 1000 / lso 500 mm" % model, model=model,
                         A_ms=_time(lambda: P.forward(x), reps), AT_ms=_time(lambda: P.backward(y), reps)))
    return recs


def conv(reps=10, batch_size=8, size=512):
    """sgdct_nn_conv3x3_ex, bias + ReLU, bf16 output."""
    import ctypes
    from . import _lib
    L = _lib.lib()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())
    recs = []
    for ci, co in ((32, 32), (64, 64)):
        x = torch.rand(batch_size, size, size, ci, device="cuda").to(torch.bfloat16)
        w = (torch.rand(co, 9 * ci, device="cuda") * 0.1).to(torch.bfloat16)
        b = torch.zeros(co, device="cuda")
        y = torch.empty(batch_size, size, size, co, device="cuda", dtype=torch.bfloat16)

        def run():
            _lib.check(L.sgdct_nn_conv3x3_ex(vp(x), None, 0, vp(w), vp(b), vp(y), None, batch_size, size, size, ci,
                                             co, 1, 0, None), "conv3x3")
        ms = _time(run, reps)
        flop = 2.0 * 9 * ci * co * batch_size * size * size
        recs.append(dict(workload="3x3 convolution %d -> %d channels, batch %d at %dx%d, bf16" % (ci, co, batch_size,
                                                                                                  size, size),
                         us=ms * 1e3, tflops=flop / ms / 1e9))
    return recs


def unet(reps=5, batch_size=8, size=512):
    from .unet import UNet, layer_specs
    net = UNet(4, 32, seed=0)
    x = torch.rand(batch_size, size, size, device="cuda")
    t = x * 0.9
    for _ in range(2):
        net.train_step(x, t, lr=1e-4)
    flop = 0.0
    for name, ci, co, k in layer_specs(4, 32):
        level = int(name[3]) if name != "out" else 0
        flop += 2.0 * k * k * ci * co * batch_size * (size >> level) ** 2
    fwd = _time(lambda: net.forward(x), reps)
    train = _time(lambda: net.train_step(x, t, lr=1e-4), reps)
    return dict(workload="U-Net 4 levels / 32 channels, bf16 operands, FP32 accumulation, batch %d at %dx%d"
                % (batch_size, size, size), conv_gflop_forward=flop / 1e9,
                forward_ms=fwd, train_step_ms=train,
                forward_tflops=flop / (fwd * 1e-3) / 1e12, train_tflops=3 * flop / (train * 1e-3) / 1e12)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("what", choices=["fan", "asd", "batch", "cone", "unet", "conv", "hexangle", "hostabi"])
    ap.add_argument("--tune-cache", default=None)
    a = ap.parse_args(argv)
    if a.what == "batch":
        recs = [batch(tune_cache=a.tune_cache)]
    elif a.what == "cone":
        recs = cone()
    elif a.what == "conv":
        recs = conv()
    elif a.what == "fan":
        recs = fan()
    elif a.what == "hexangle":
        recs = hexangle()
    elif a.what == "hostabi":
        recs = [hostabi()]
    elif a.what == "asd":
        recs = [asd()]
    else:
        recs = [unet()]
    for r in recs:
        print(json.dumps(r), flush=True)


if __name__ == "__main__":
    sys.exit(main())
