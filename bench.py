"""Throughput of the B1 fan-beam projector pair on one MI355X.

One step = one A (image -> sinogram) plus one A^T (sinogram -> image) of the
bilinear (B1) box-spline projector at 512 x 512 pixels (0.6 mm), 720 views
over 2 pi, 768 bins (0.8 mm), source-isocentre 500 mm, source-detector
1000 mm, FP32, inputs resident in HBM.  The image is a seeded random-ellipse
phantom; A^T is applied to its Poisson-noisy (1e5 photons) log sinogram.

Prints one JSON line (rank 0).  `roofline` prices the dominant kernel with the
view-visit byte model 4 np (nx ny + nu) bytes per launch (every view reads or
updates every pixel once and its detector row once), timed per launch with
HIP events on the launch stream; `valu_roofline` prices the same call by its
VALU instructions against the FP32 issue peak (the resource that binds these
kernels).  `cpu_baseline` times the CPU restatement of
the reference formulas (FP32 footprint, FP64 sums, OpenMP) on a bounded
sample of views and scales to pairs/s.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from sgd_amd.workloads import FAN512  # noqa: E402

NX, NY = FAN512["nx"], FAN512["ny"]
NP_VIEWS = FAN512["views"]
NU = FAN512["nu"]
DX = FAN512["dx"]
DU = FAN512["du"]
LSD = FAN512["lsd"]
LSO = FAN512["lso"]
HBM_PEAK_GBS = 8000.0
# HBM bytes per B1 call at this workload, summed over the call's kernels
# (the projector kernel and the partial-sum kernel), from rocprofv3 --pmc
# FETCH_SIZE then WRITE_SIZE (KiB per dispatch; FETCH_SIZE scaled by 2.00 and
# WRITE_SIZE by 1.00, the factors a 256 MiB device copy gives, --copy-calib):
#   A:   fwd_bin 16744 x 2 + 103680, sum_parts 51851 x 2 + 2160 KiB
#   A^T: bwd_pix 73486 x 2 + 49152, sum_parts 24587 x 2 + 1024 KiB
# (24 line segments / 24 view groups, each writing FP64 partial sums;
# profiles/b1_fanbeam_512x720x768.md).
TRAFFIC_BYTES_B1 = {"A": 243030 * 1024.0, "A^T": 246322 * 1024.0}
# VALU wave-instructions per B1 launch at this workload (the projector kernel;
# the partial-sum kernel adds 5e5 / 3e5), rocprofv3 --pmc SQ_INSTS_VALU
# (profiles/b1_fanbeam_512x720x768.md).  Peak issue rate of non-packed FP32
# VALU instructions: 256 CUs x 4 SIMDs x 2.4 GHz / 2 cycles per wave64
# instruction = 1.2288e12 / s, the FP32 vector peak of 157.3 TFLOP/s divided by
# the 128 FLOP of one wave64 FMA (packed FP32 is not used).
VALU_INSTS_B1 = {"A": 7.8705e8, "A^T": 8.1794e8}
VALU_PEAK_PER_S = 256 * 4 * 2.4e9 / 2
METRIC = "A+Aᵀ pairs/s at 512²/720v (%HBM peak)"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--basis", default="B1", choices=["B1", "B0", "SF"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-views", type=int, default=720)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--copy-calib", type=int, default=0,
                    help="also run one device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def cpu_baseline(img, sino, ang, basis, nviews, min_seconds=10.0):
    """Oracle in reference mode: A and A^T over views [0, nviews), repeated
    until at least `min_seconds` of CPU work, scaled to one full pair."""
    import numpy as np

    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from pyoracle import Geometry, Oracle

    g = Geometry(Oracle("ref"), NX, NY, ang, NU, DX, DX, DU, LSD, LSO)
    nviews = min(nviews, NP_VIEWS)
    reps = 0
    t0 = time.perf_counter()
    while True:
        g.forward(basis, img, views=(0, nviews))
        g.backward(basis, sino, views=(0, nviews))
        reps += 1
        dt = time.perf_counter() - t0
        if dt >= min_seconds:
            break
    per_pair = dt * NP_VIEWS / (nviews * reps)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    model = "unknown CPU"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / per_pair,
        "unit": "pairs/s",
        "cores": threads,
        "kind": "port",
        "sample": "%s A + A^T over %d of %d views at 512x512/768 bins, %d times (%.1f s), scaled to one full "
                  "pair; oracle/sgdct_oracle.c FP32-footprint build, OpenMP %d threads on %s (%d logical CPUs "
                  "visible)" % (basis, nviews, NP_VIEWS, reps, dt, threads, model, os.cpu_count() or 0),
    }


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: there is no multi-GPU path "
              "(views would shard across GPUs with one all-reduce of A^T)", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    from sgd_amd import workloads

    dev = torch.device("cuda", 0)
    ang = workloads.angles(NP_VIEWS)
    P = workloads.fan512_projector(args.basis)
    img_np, sino_np = workloads.fan512_inputs(P)
    x = torch.tensor(img_np, dtype=torch.float32, device=dev)
    y = torch.tensor(sino_np, dtype=torch.float32, device=dev)
    yo = torch.empty_like(y)
    xo = torch.empty_like(x)
    stream = torch.cuda.current_stream(dev)

    if args.copy_calib:
        src = torch.ones(args.copy_calib * (1 << 18), dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    for _ in range(args.warmup):
        P.forward(x, out=yo)
        P.backward(y, out=xo)
    torch.cuda.synchronize()

    ev = [[torch.cuda.Event(enable_timing=True) for _ in range(4)] for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        e = ev[k]
        e[0].record(stream)
        P.forward(x, out=yo)
        e[1].record(stream)
        e[2].record(stream)
        P.backward(y, out=xo)
        e[3].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0

    fwd_ms = sum(e[0].elapsed_time(e[1]) for e in ev) / args.steps
    bwd_ms = sum(e[2].elapsed_time(e[3]) for e in ev) / args.steps
    ms_per_step = wall * 1e3 / args.steps
    value = args.steps / wall
    bytes_per_launch = 4.0 * NP_VIEWS * (NX * NY + NU)
    dom_ms, dom = (fwd_ms, "A") if fwd_ms >= bwd_ms else (bwd_ms, "A^T")
    achieved = bytes_per_launch / (dom_ms * 1e-3) / 1e9
    evals = P.eval_count()
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "%s fan-beam A + A^T pair, 512x512 image (0.6 mm), 720 views over 2pi, 768 bins "
                        "(0.8 mm), lso 500 mm, lsd 1000 mm" % args.basis,
            "basis": args.basis, "nx": NX, "ny": NY, "views": NP_VIEWS, "bins": NU,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": TRAFFIC_BYTES_B1[dom] if args.basis == "B1" else None,
            "kernel": dom,
            "bytes_per_launch": bytes_per_launch,
        },
        "valu_roofline": ({
            "bound": "valu",
            "achieved": VALU_INSTS_B1[dom] / (dom_ms * 1e-3),
            "peak": VALU_PEAK_PER_S,
            "unit": "VALU wave-instructions/s",
            "frac": VALU_INSTS_B1[dom] / (dom_ms * 1e-3) / VALU_PEAK_PER_S,
            "insts_per_launch": VALU_INSTS_B1[dom],
            "kernel": dom,
        } if args.basis == "B1" else None),
        "kernel_ms": {"A": fwd_ms, "A^T": bwd_ms},
        "footprint_evals_per_A": evals,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(img_np, sino_np, ang, args.basis, args.cpu_views, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
