"""The hexagonal workload's autograd checks at full size: H0 and Hl on 296
rows x 256 columns (lattice constant 0.5 mm), 1024 views, 1024 bins of 0.6 mm,
a batch of 64 lattice images, A wrapped as a torch.autograd.Function whose
backward is the hand-written A^T kernel.

  * central finite differences, h = 1e-3, on 256 random lattice points, loss
    L = sum(w * A x) summed in FP64.  Bar, per sample: the difference of the
    two FP32 projections is exact except for their rounding, which only
    touches the bins the perturbed point projects to (every other bin is
    computed from identical inputs in the same order and comes out bit for
    bit the same).  With u_i the FP32 spacing at entry i of A(x +- h e),
    sigma = sqrt(sum_i w_i^2 (u+_i^2 + u-_i^2)) / (2h) bounds the finite
    difference's rounding error at about 2.5 standard deviations (each entry
    carries two roundings of at most u/2, standard deviation u / sqrt(12)
    each); |FD - grad| <= 4 sigma + 1e-5 |grad|;
  * the adjoint dot-product test on the batch with Gaussian x, y: <= 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("basis", ["H0", "Hl"])
def test_hex296_autograd_batch64(basis, cuda):
    import torch
    from sgd_amd import workloads as wl
    from sgd_amd.torch_ops import project
    P = wl.hex296_projector(basis)
    nl, nc = P.ny, P.nx
    print("hex296 %s: %d x %d lattice, %d views, %d bins, batch 64" % (basis, nl, nc, P.np, P.nu), flush=True)
    gen = torch.Generator(device=cuda).manual_seed(0)
    x = torch.rand((64, nl, nc), device=cuda, generator=gen)
    x[:, 1::2, -1] = 0
    w = torch.randn((64, P.np, P.nu), device=cuda, generator=gen)
    xr = x.clone().requires_grad_(True)
    (project(xr, P) * w).sum().backward()
    g = xr.grad
    rng = np.random.default_rng(0)
    h = 1e-3
    idx = []
    while len(idx) < 256:
        b_, i_, j_ = (int(v) for v in (rng.integers(64), rng.integers(nl), rng.integers(nc)))
        if not (i_ & 1 and j_ == nc - 1):
            idx.append((b_, i_, j_))
    ratios, rels = [], []
    inf = torch.tensor(float("inf"), device=cuda)
    with torch.no_grad():
        for (b_, i_, j_) in idx:
            xp = x[b_].clone()
            xp[i_, j_] += h
            xm = x[b_].clone()
            xm[i_, j_] -= h
            ap, am = P.forward(xp), P.forward(xm)
            wd = w[b_].double()
            fd = float(((ap.double() - am.double()) * wd).sum() / (2 * h))
            m = ap != am
            up = (torch.nextafter(ap, inf) - ap).double()
            um = (torch.nextafter(am, inf) - am).double()
            sig = float(torch.sqrt((wd[m] ** 2 * (up[m] ** 2 + um[m] ** 2)).sum()) / (2 * h))
            gv = float(g[b_, i_, j_])
            err = abs(fd - gv)
            ratios.append(err / (4 * sig + 1e-5 * abs(gv)))
            rels.append(err / max(abs(gv), 1e-30))
    print("  FD: max |FD - grad| / bar %.3f, median relative error %.2e, max %.2e"
          % (max(ratios), float(np.median(rels)), max(rels)), flush=True)
    assert max(ratios) <= 1.0
    xg = torch.randn((64, nl, nc), device=cuda, generator=gen)
    xg[:, 1::2, -1] = 0
    yg = torch.randn((64, P.np, P.nu), device=cuda, generator=gen)
    Ax = P.forward(xg).double()
    lhs = (Ax * yg.double()).sum()
    rhs = (xg.double() * P.backward(yg).double()).sum()
    err = float(abs(lhs - rhs) / (Ax.norm() * yg.double().norm()))
    print("  batch-64 adjoint %.2e" % err, flush=True)
    assert err <= 1e-6
