"""CPU oracle of the warm-start U-Net (numpy, FP64), forward and backward.

TEST INFRASTRUCTURE ONLY: tests/ use it to check the gfx950 training and
inference kernels of sgd_amd/unet.py; the product never imports it.

The network (no reference counterpart; the closest precedent is the
reference's Layer::forward/back_prop pattern, layers/layer.hh:15-16):
  levels l = 0..L-1 with c_l = c0 * 2^l channels at resolution H / 2^l,
  encoder  e_0 = R(conv(R(conv(x8, 8 -> c0)), c0 -> c0))
           e_l = R(conv(R(conv(pool(e_{l-1}), c_{l-1} -> c_l)), c_l -> c_l))
  decoder  d_{L-1} = e_{L-1}
           d_l = R(conv(R(conv(cat(up(d_{l+1}), e_l), c_{l+1} + c_l -> c_l)), c_l -> c_l))
  output   y = x + conv1x1(d_0, c0 -> 1) + b
with R = ReLU, conv a 3x3 zero-padded convolution with bias, pool a 2x2
average, up a 2x2 nearest-neighbour copy, cat a channel concatenation
(upsampled part first), x8 the 1-channel input padded with 7 zero channels.
Tensors are NHWC.  Loss: mean squared error against the target.

With bf16=True the oracle rounds to bfloat16 (round to nearest even) at the
points where the GPU kernels store bfloat16: the padded input, every
post-ReLU activation, pooled operands, the weights, and every gradient that
feeds a matrix-core product (the conv outputs' gradients).  Sums are FP64.
"""
import numpy as np


def bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.ascontiguousarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layer_specs(levels=4, c0=32, cin=8):
    """(name, cin, cout, ksize) of every convolution, in execution order."""
    specs = []
    c = [c0 << l for l in range(levels)]
    for l in range(levels):
        specs.append(("enc%d_a" % l, cin if l == 0 else c[l - 1], c[l], 3))
        specs.append(("enc%d_b" % l, c[l], c[l], 3))
    for l in range(levels - 2, -1, -1):
        specs.append(("dec%d_a" % l, c[l + 1] + c[l], c[l], 3))
        specs.append(("dec%d_b" % l, c[l], c[l], 3))
    specs.append(("out", c0, 1, 1))
    return specs


def init_params(levels=4, c0=32, seed=0):
    """He-normal weights W[cout][kh][kw][cin], zero biases (FP32 values)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co, k in layer_specs(levels, c0):
        fan_in = ci * k * k
        std = np.sqrt(2.0 / fan_in) if name != "out" else np.sqrt(1.0 / fan_in) * 0.1
        params[name] = (rng.normal(0.0, std, (co, k, k, ci)).astype(np.float32), np.zeros(co, np.float32))
    return params


def _pad_in(x):
    b, h, w = x.shape
    x8 = np.zeros((b, h, w, 8))
    x8[..., 0] = x
    return x8


def _patches(x):
    b, h, w, c = x.shape
    xp = np.zeros((b, h + 2, w + 2, c))
    xp[:, 1:-1, 1:-1] = x
    return np.stack([xp[:, dy:dy + h, dx:dx + w] for dy in range(3) for dx in range(3)], axis=3)


def conv3(x, W, b):
    co = W.shape[0]
    return np.einsum("byxtc,otc->byxo", _patches(x), W.reshape(co, 9, -1), optimize=True) + b


def conv3_back(x, W, dy):
    co, _, _, ci = W.shape
    P = _patches(x)
    dW = np.einsum("byxo,byxtc->otc", dy, P, optimize=True).reshape(W.shape)
    dP = np.einsum("byxo,otc->byxtc", dy, W.reshape(co, 9, ci), optimize=True)
    bb, h, w, _ = x.shape
    dxp = np.zeros((bb, h + 2, w + 2, ci))
    t = 0
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + h, kx:kx + w] += dP[:, :, :, t]
            t += 1
    return dxp[:, 1:-1, 1:-1], dW, dy.sum(axis=(0, 1, 2))


def pool(x):
    b, h, w, c = x.shape
    return x.reshape(b, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool_back(d):
    return np.repeat(np.repeat(d, 2, axis=1), 2, axis=2) / 4.0


def up(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up_back(d):
    b, h, w, c = d.shape
    return d.reshape(b, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


class UNetOracle:
    def __init__(self, params, levels=4, c0=32, use_bf16=True):
        self.levels, self.c0, self.use_bf16 = levels, c0, use_bf16
        r = bf16 if use_bf16 else (lambda a: np.asarray(a, dtype=np.float64))
        self.r = r
        self.p = {k: (r(W), np.asarray(b, np.float64)) for k, (W, b) in params.items()}

    def forward(self, x):
        """x: (B, H, W) input images.  Returns the output and a tape."""
        r, L = self.r, self.levels
        tape = {"x": np.asarray(x, np.float64)}
        h = r(_pad_in(x))
        tape["in8"] = h
        enc = []
        for l in range(L):
            src = h if l == 0 else r(pool(enc[-1]))
            tape["enc%d_in" % l] = src
            a = r(np.maximum(conv3(src, *self.p["enc%d_a" % l]), 0))
            tape["enc%d_a" % l] = a
            e = r(np.maximum(conv3(a, *self.p["enc%d_b" % l]), 0))
            tape["enc%d_b" % l] = e
            enc.append(e)
        d = enc[-1]
        for l in range(L - 2, -1, -1):
            src = np.concatenate([up(d), enc[l]], axis=3)
            tape["dec%d_in" % l] = src
            a = r(np.maximum(conv3(src, *self.p["dec%d_a" % l]), 0))
            tape["dec%d_a" % l] = a
            d = r(np.maximum(conv3(a, *self.p["dec%d_b" % l]), 0))
            tape["dec%d_b" % l] = d
        W, b = self.p["out"]
        y = tape["x"] + np.einsum("byxc,oc->byxo", d, W.reshape(1, -1))[..., 0] + b[0]
        return y, tape

    def loss(self, y, target):
        return float(np.mean((y - target) ** 2))

    def backward(self, tape, y, target):
        """Gradients of the MSE loss: {name: (dW, db)}."""
        r, L = self.r, self.levels
        grads = {}
        dy = 2.0 * (y - np.asarray(target, np.float64)) / y.size
        top = "dec0_b" if L > 1 else "enc0_b"
        d0 = tape[top]
        W, _ = self.p["out"]
        grads["out"] = (np.einsum("byx,byxc->c", dy, d0).reshape(W.shape), np.array([dy.sum()]))
        gdec = {0: dy[..., None] * W.reshape(1, 1, 1, -1)}  # wrt post-ReLU d_l
        genc = [np.zeros_like(tape["enc%d_b" % l]) for l in range(L)]
        for l in range(0, L - 1):
            g = gdec[l]
            gy = r(g * (tape["dec%d_b" % l] > 0))
            ga, dW, db = conv3_back(tape["dec%d_a" % l], self.p["dec%d_b" % l][0], gy)
            grads["dec%d_b" % l] = (dW, db)
            gy = r(ga * (tape["dec%d_a" % l] > 0))
            gin, dW, db = conv3_back(tape["dec%d_in" % l], self.p["dec%d_a" % l][0], gy)
            grads["dec%d_a" % l] = (dW, db)
            cu = gin.shape[3] - tape["enc%d_b" % l].shape[3]
            genc[l] += gin[..., cu:]
            gup = up_back(gin[..., :cu])
            if l + 1 == L - 1:
                genc[L - 1] += gup
            else:
                gdec[l + 1] = gup
        for l in range(L - 1, -1, -1):
            gy = r(genc[l] * (tape["enc%d_b" % l] > 0))
            ga, dW, db = conv3_back(tape["enc%d_a" % l], self.p["enc%d_b" % l][0], gy)
            grads["enc%d_b" % l] = (dW, db)
            gy = r(ga * (tape["enc%d_a" % l] > 0))
            gin, dW, db = conv3_back(tape["enc%d_in" % l], self.p["enc%d_a" % l][0], gy)
            grads["enc%d_a" % l] = (dW, db)
            if l > 0:
                genc[l - 1] += pool_back(gin)
        return grads
