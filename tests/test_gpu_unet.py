"""The warm-start U-Net's gfx950 kernels (sgd_amd/csrc/unet.hip) against the
CPU oracle (oracle/unet_oracle.py, pinned by finite differences):
  * forward output and every parameter gradient equal the oracle's, which
    rounds to bfloat16 at the same points (relative L2 <= 1e-2 for the output,
    <= 4e-2 per layer and <= 1.5e-2 over all gradients), and stay within
    bfloat16 precision of the oracle's unrounded FP64 network;
  * a few Adam steps on a fixed batch lower the loss.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-30))


@pytest.mark.parametrize("levels,c0,hw", [(3, 8, 16), (4, 32, 32)])
def test_forward_and_gradients_match_oracle(cuda, levels, c0, hw):
    import torch
    from unet_oracle import UNetOracle
    from sgd_amd.unet import UNet, init_params
    print("unet parity levels %d c0 %d %dx%d" % (levels, c0, hw, hw), flush=True)
    rng = np.random.default_rng(1)
    params = init_params(levels, c0, seed=2)
    params = {k: (W, rng.normal(0, 0.05, b.shape).astype(np.float32)) for k, (W, b) in params.items()}
    x = rng.random((2, hw, hw)).astype(np.float32)
    target = (x + 0.2 * rng.standard_normal(x.shape)).astype(np.float32)
    net = UNet(levels, c0, params=params)
    y = net.forward(torch.tensor(x, device=cuda))
    net.backward(torch.tensor(target, device=cuda))
    torch.cuda.synchronize()
    yg = y.cpu().numpy()
    gg = net.grads_numpy()
    # per-layer bars allow for bfloat16 rounding flips (one ulp, 2^-8, of an
    # activation or gradient element rounded the other way), which weigh most
    # in the deepest layers' few pixels; the whole gradient vector is tighter
    for use_bf16, tol_y, tol_g, tol_all in ((True, 1e-2, 4e-2, 1.5e-2), (False, 3e-2, 8e-2, 4e-2)):
        o = UNetOracle(params, levels, c0, use_bf16=use_bf16)
        yo, tape = o.forward(x)
        go = o.backward(tape, yo, target)
        ey = _rel(yg - x, yo - x)
        print("bf16 oracle" if use_bf16 else "fp64 oracle", "output rel %.2e" % ey)
        assert ey <= tol_y
        worst = 0.0
        for name in go:
            for k in (0, 1):
                e = _rel(gg[name][k], go[name][k])
                worst = max(worst, e)
                assert e <= tol_g, (name, k, e)
        flat = lambda g: np.concatenate([np.ravel(g[n][k]) for n in sorted(g) for k in (0, 1)])
        e_all = _rel(flat(gg), flat(go))
        print("  worst layer gradient rel %.2e, all gradients rel %.2e" % (worst, e_all))
        assert e_all <= tol_all


def test_adam_lowers_the_loss(cuda):
    import torch
    from sgd_amd.unet import UNet
    print("unet training sanity", flush=True)
    rng = np.random.default_rng(3)
    x = torch.tensor(rng.random((4, 32, 32)).astype(np.float32), device=cuda)
    target = x * 0.5 + 0.25
    net = UNet(4, 16, seed=4)
    losses = []
    for _ in range(30):
        y = net.train_step(x, target, lr=2e-3)
        losses.append(float(((y - target) ** 2).mean()))
    print("loss %.4e -> %.4e" % (losses[0], losses[-1]))
    assert losses[-1] < 0.3 * losses[0]
