"""Pins the U-Net oracle (oracle/unet_oracle.py): its hand-written backward
equals central finite differences of its own forward (FP64, no bfloat16
rounding) for weights and biases of every layer, and the bfloat16 rounding
helper rounds to nearest even."""
import numpy as np


def test_bf16_rounding():
    from unet_oracle import bf16
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, -2.5, 3.0e-3, 65504.0])
    r = bf16(x)
    assert r[0] == 1.0 and r[1] == 1.0  # tie to even
    assert r[2] == 1.0 + 2 ** -7  # 1 + 1.5 ulp rounds up
    assert r[3] == -2.5
    assert abs(r[4] - 3.0e-3) / 3.0e-3 < 2 ** -8
    assert r[5] == 65536.0


def test_backward_matches_finite_differences():
    from unet_oracle import UNetOracle, init_params, layer_specs
    levels, c0 = 3, 4
    rng = np.random.default_rng(0)
    params = init_params(levels, c0, seed=1)
    params = {k: (W.astype(np.float64), rng.normal(0, 0.1, b.shape)) for k, (W, b) in params.items()}
    x = rng.random((2, 8, 8))
    target = x + 0.3 * rng.standard_normal(x.shape)
    net = UNetOracle(params, levels, c0, use_bf16=False)
    y, tape = net.forward(x)
    grads = net.backward(tape, y, target)
    h = 1e-6
    for name, _, _, _ in layer_specs(levels, c0):
        W, b = params[name]
        for which in ("W", "b"):
            arr = W if which == "W" else b
            for _ in range(3):
                idx = tuple(rng.integers(0, s) for s in arr.shape)
                old = arr[idx]
                arr[idx] = old + h
                lp = net.loss(UNetOracle(params, levels, c0, use_bf16=False).forward(x)[0], target)
                arr[idx] = old - h
                lm = net.loss(UNetOracle(params, levels, c0, use_bf16=False).forward(x)[0], target)
                arr[idx] = old
                fd = (lp - lm) / (2 * h)
                an = grads[name][0 if which == "W" else 1][idx]
                assert abs(fd - an) <= 1e-6 + 1e-5 * abs(fd), (name, which, idx, fd, an)
