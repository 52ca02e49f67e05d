"""Learned warm start (sgd_amd/warmstart.py) end to end at a reduced size:
  * the batched ASD-POCS equals the single-image one image by image (same
    iterate to FP32 noise, same iteration count at a tolerance);
  * a U-Net trained for a few epochs on pairs the method generates seeds
    ASD-POCS so that it reaches the tolerance in fewer iterations than from
    zero, on phantoms it was not trained on, with the fallback available.
The full-size experiment (512^2, 90 views, 4096 pairs, 256 test phantoms,
tolerance 1e-2, the noise level) is `python -m sgd_amd.warmstart`.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SMALL = dict(nx=128, ny=128, dx=2.4, dy=2.4, nu=192, du=3.2, lsd=1000.0, lso=500.0, views=45)


def test_batched_asd_pocs_matches_single(cuda):
    import torch
    from sgd_amd import warmstart as ws
    from sgd_amd.solvers import SirtWeights, asd_pocs, asd_pocs_batch
    print("asd-pocs batch vs single", flush=True)
    P = ws.projector(SMALL)
    x = ws.phantoms([11, 12, 13], 128, 128, 2.4, 2.4, cuda)
    b = ws.noisy_sinograms(P, x, seed=5)
    w = SirtWeights(P)
    rb = asd_pocs_batch(P, b, niter=25, weights=w, **ws.ASD)
    for i in range(3):
        r = asd_pocs(P, b[i], niter=25, weights=w, **ws.ASD)
        e = float((rb.x[i] - r.x).norm() / r.x.norm())
        assert e <= 1e-4, (i, e)
    tol = 0.05
    # images leave the working batch as they stop (check_every=1: at once)
    rb = asd_pocs_batch(P, b, niter=400, tol=tol, weights=w, check_every=1, **ws.ASD)
    for i in range(3):
        r = asd_pocs(P, b[i], niter=400, tol=tol, weights=w, **ws.ASD)
        assert abs(int(rb.iters[i]) - (len(r.residuals) - 1)) <= 1, (i, rb.iters[i], len(r.residuals))
        if int(rb.iters[i]) == len(r.residuals) - 1:
            e = float((rb.x[i] - r.x).norm() / r.x.norm())
            assert e <= 1e-4, (i, e)


def test_trained_network_needs_fewer_iterations(cuda):
    import torch
    from sgd_amd import warmstart as ws
    from sgd_amd.unet import UNet
    print("warm start mini experiment", flush=True)
    P = ws.projector(SMALL)
    xin, xt, _ = ws.make_pairs(P, list(range(128)), batch=64, asd_iters=100)
    net = UNet(4, 16, seed=0)
    ws.train(net, xin, xt, epochs=4, batch=16, lr=2e-3)
    seeds = list(range(5000, 5016))
    x = ws.phantoms(seeds, 128, 128, 2.4, 2.4, cuda)
    b = ws.noisy_sinograms(P, x, seed=17)
    ev = ws.evaluate(P, net, b, tols=("0.03",), max_iters=600, batch=16)["0.03"]
    pi, wi = np.mean(ev["plain_iters"]), np.mean(ev["warm_iters"])
    print("plain %.1f its (%.2fs), warm %.1f its (%.2fs), fell back %d" % (pi, ev["plain_s"], wi, ev["warm_s"],
                                                                          ev["fell_back"]))
    assert wi < pi
