"""Warm-start U-Net trained and run on the gfx950 matrix cores.

No reference counterpart (the reference's learned part is a cuBLAS MLP,
network.cu / layers/linear.cu; its Layer::forward/back_prop pattern,
layers/layer.hh:15-16, is the closest precedent).  The network maps a
5-iteration SIRT image to an estimate of the 200-iteration ASD-POCS image,
which then seeds ASD-POCS (sgd_amd/warmstart.py):

  levels l = 0..L-1, c_l = c0 2^l channels at H / 2^l,
  e_0 = R(conv(R(conv(x8))))                     x8: x padded to 8 channels
  e_l = R(conv(R(conv(pool(e_{l-1})))))          pool: 2x2 average
  d_{L-1} = e_{L-1}
  d_l = R(conv(R(conv([up(d_{l+1}), e_l]))))     up: 2x2 nearest copy
  y   = x + conv1x1(d_0) + b

R is ReLU and conv a 3x3 zero-padded convolution with bias.  Activations are
NHWC bfloat16, every product runs on the matrix cores with FP32 accumulation
(sgd_amd/csrc/unet.hip), weights keep an FP32 master copy updated by Adam.
torch only allocates the device buffers; every arithmetic step is a kernel of
libsgdct.so called on torch's current stream.
"""
import ctypes

import numpy as np
import torch

from . import _lib


def layer_specs(levels=4, c0=32, cin=8):
    """(name, cin, cout, ksize) of every convolution, in execution order."""
    c = [c0 << l for l in range(levels)]
    specs = []
    for l in range(levels):
        specs.append(("enc%d_a" % l, cin if l == 0 else c[l - 1], c[l], 3))
        specs.append(("enc%d_b" % l, c[l], c[l], 3))
    for l in range(levels - 2, -1, -1):
        specs.append(("dec%d_a" % l, c[l + 1] + c[l], c[l], 3))
        specs.append(("dec%d_b" % l, c[l], c[l], 3))
    specs.append(("out", c0, 1, 1))
    return specs


def init_params(levels=4, c0=32, seed=0):
    """He-normal conv weights W[cout][3][3][cin], a small output head, zero
    biases (numpy FP32)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, ci, co, k in layer_specs(levels, c0):
        fan_in = ci * k * k
        std = np.sqrt(2.0 / fan_in) if name != "out" else np.sqrt(1.0 / fan_in) * 0.1
        params[name] = (rng.normal(0.0, std, (co, k, k, ci)).astype(np.float32), np.zeros(co, np.float32))
    return params


def _vp(t):
    return ctypes.c_void_p(t.data_ptr())


class _Conv:
    def __init__(self, name, cin, cout, W, b, device):
        self.name, self.cin, self.cout = name, cin, cout
        self.w32 = torch.tensor(np.ascontiguousarray(W, np.float32).reshape(cout, -1), device=device)
        self.b32 = torch.tensor(np.asarray(b, np.float32), device=device)
        self.w16 = torch.empty(self.w32.shape, dtype=torch.bfloat16, device=device)
        self.wt16 = torch.empty((cin, 9 * cout), dtype=torch.bfloat16, device=device)
        self.gw = torch.zeros_like(self.w32)
        self.gb = torch.zeros_like(self.b32)
        self.state = [torch.zeros_like(self.w32), torch.zeros_like(self.w32), torch.zeros_like(self.b32),
                      torch.zeros_like(self.b32)]


class UNet:
    def __init__(self, levels=4, c0=32, params=None, seed=0, device=None):
        if not torch.cuda.is_available():
            raise _lib.SgdctError("the U-Net kernels need a ROCm GPU")
        self.levels, self.c0 = levels, c0
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        params = params if params is not None else init_params(levels, c0, seed)
        self.convs = {}
        for name, ci, co, k in layer_specs(levels, c0):
            W, b = params[name]
            if name == "out":
                self.head_w = torch.tensor(np.asarray(W, np.float32).reshape(-1), device=self.device)
                self.head_b = torch.tensor(np.asarray(b, np.float32).reshape(1), device=self.device)
                self.head_g = torch.zeros(c0 + 1, device=self.device)
                self.head_state = [torch.zeros(c0 + 1, device=self.device), torch.zeros(c0 + 1, device=self.device)]
                self.head_wb = torch.zeros(c0 + 1, device=self.device)
            else:
                self.convs[name] = _Conv(name, ci, co, W, b, self.device)
        self._sync_bf16()
        self.steps = 0
        self._buf = {}

    # -- plumbing ------------------------------------------------------------
    def _s(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _chk(self, rc, what):
        _lib.check(rc, what)

    def _sync_bf16(self):
        L = _lib.lib()
        for c in self.convs.values():
            self._chk(L.sgdct_nn_to_bf16(_vp(c.w32), _vp(c.w16), c.w32.numel(), self._s()), "to_bf16")

    def _get(self, key, shape, dtype):
        t = self._buf.get(key)
        if t is None or tuple(t.shape) != tuple(shape) or t.dtype != dtype:
            t = torch.empty(shape, dtype=dtype, device=self.device)
            self._buf[key] = t
        return t

    def parameters_numpy(self):
        out = {}
        for name, c in self.convs.items():
            out[name] = (c.w32.cpu().numpy().reshape(c.cout, 3, 3, c.cin), c.b32.cpu().numpy())
        out["out"] = (self.head_w.cpu().numpy().reshape(1, 1, 1, self.c0), self.head_b.cpu().numpy())
        return out

    def _conv(self, c, inp, out, B, H, W, relu=True, w=None, bias=True, out_f32=False, cin=None, cout=None,
              skip=None, mask=None):
        """3x3 convolution.  inp may be (low, skip_tensor): the decoder's
        concatenation [upsample2(low), skip] read in place; mask (an
        activation) zeroes the output where it is not positive."""
        L = _lib.lib()
        w = c.w16 if w is None else w
        b = _vp(c.b32) if bias else None
        low, in2, c0 = (inp[0], inp[1], inp[0].shape[3]) if isinstance(inp, tuple) else (inp, None, 0)
        self._chk(L.sgdct_nn_conv3x3_ex(_vp(low), _vp(in2) if in2 is not None else None, c0, _vp(w), b, _vp(out),
                                        _vp(mask) if mask is not None else None, B, H, W, cin or c.cin,
                                        cout or c.cout, int(relu), int(out_f32), self._s()), "conv3x3 " + c.name)

    # -- forward -------------------------------------------------------------
    def forward(self, x):
        """x: (B, H, W) FP32 device tensor, H and W divisible by 2^(levels-1).
        Returns y (B, H, W) FP32; activations are kept for backward."""
        L = _lib.lib()
        x = x.contiguous().float()
        B, H, W = x.shape
        if H % (1 << (self.levels - 1)) or W % (1 << (self.levels - 1)):
            raise ValueError("image size must be divisible by 2^(levels-1)")
        bf = torch.bfloat16
        self.shape = (B, H, W)
        self.x = x
        t = {}
        t["in8"] = self._get("in8", (B, H, W, 8), bf)
        self._chk(L.sgdct_nn_pad_input(_vp(x), _vp(t["in8"]), B * H * W, self._s()), "pad_input")
        enc = []
        for l in range(self.levels):
            h, w = H >> l, W >> l
            ca, cb = self.convs["enc%d_a" % l], self.convs["enc%d_b" % l]
            if l == 0:
                src = t["in8"]
            else:
                src = self._get("enc%d_in" % l, (B, h, w, ca.cin), bf)
                self._chk(L.sgdct_nn_pool2(_vp(enc[-1]), _vp(src), B, h, w, ca.cin, self._s()), "pool2")
            t["enc%d_in" % l] = src
            a = self._get("enc%d_a" % l, (B, h, w, ca.cout), bf)
            self._conv(ca, src, a, B, h, w)
            e = self._get("enc%d_b" % l, (B, h, w, cb.cout), bf)
            self._conv(cb, a, e, B, h, w)
            t["enc%d_a" % l], t["enc%d_b" % l] = a, e
            enc.append(e)
        d = enc[-1]
        for l in range(self.levels - 2, -1, -1):
            h, w = H >> l, W >> l
            ca, cb = self.convs["dec%d_a" % l], self.convs["dec%d_b" % l]
            # the concatenation [upsample2(d), enc[l]] is read in place by the
            # convolution and by its weight gradient
            src = (d, enc[l])
            a = self._get("dec%d_a" % l, (B, h, w, ca.cout), bf)
            self._conv(ca, src, a, B, h, w)
            d = self._get("dec%d_b" % l, (B, h, w, cb.cout), bf)
            self._conv(cb, a, d, B, h, w)
            t["dec%d_in" % l], t["dec%d_a" % l], t["dec%d_b" % l] = src, a, d
        self.tape = t
        self.d0 = d
        y = self._get("y", (B, H, W), torch.float32)
        self._chk(L.sgdct_nn_head(_vp(d), _vp(self.head_w), _vp(self.head_b), _vp(x), _vp(y), B * H * W, self.c0,
                                  self._s()), "head")
        return y

    __call__ = forward

    # -- backward ------------------------------------------------------------
    def _workspace(self, floats):
        ws = self._buf.get("ws")
        if ws is None or ws.numel() < floats:
            ws = torch.empty(int(floats), device=self.device)
            self._buf["ws"] = ws
        return ws

    def _conv_grads(self, c, x_in, gy, B, h, w):
        """dW, db of conv c from its input activation (or (low, skip) for a
        decoder's concatenated input) and bf16 output gradient."""
        L = _lib.lib()
        M = B * h * w
        ws = self._workspace(L.sgdct_nn_wgrad_workspace(B, h, w, c.cin, c.cout))
        low, x2, c0 = (x_in[0], x_in[1], x_in[0].shape[3]) if isinstance(x_in, tuple) else (x_in, None, 0)
        self._chk(L.sgdct_nn_wgrad_ex(_vp(low), _vp(x2) if x2 is not None else None, c0, _vp(gy), _vp(c.gw),
                                      _vp(c.gb), _vp(ws), ws.numel(), B, h, w, c.cin, c.cout, self._s()),
                  "wgrad " + c.name)

    def _input_grad(self, c, gy, out, B, h, w, mask=None):
        """Gradient of conv c's input: FP32, or with mask (the input's own
        ReLU activation) the bf16 output gradient of the layer before."""
        L = _lib.lib()
        self._chk(L.sgdct_nn_flip_weights(_vp(c.w16), _vp(c.wt16), c.cin, c.cout, self._s()), "flip")
        self._conv(c, gy, out, B, h, w, relu=False, w=c.wt16, bias=False, out_f32=mask is None, cin=c.cout,
                   cout=c.cin, mask=mask)

    def backward(self, target):
        """Gradients of mean((y - target)^2) into every parameter's .gw/.gb
        and the head's head_g (w then b)."""
        L = _lib.lib()
        B, H, W = self.shape
        t = self.tape
        f32, bf = torch.float32, torch.bfloat16
        target = target.contiguous().float()
        y = self._buf["y"]
        M = B * H * W
        L_ = self.levels
        dy = self._get("dy", (B, H, W), f32)
        # Every gradient that ends in a ReLU is written masked, in bf16, by
        # its last contribution: the head (decoder level 0), the 2x2 sum of
        # the upsampling's adjoint (the level below in the decoder, and the
        # bottom encoder level), the pooling's adjoint (encoder levels above
        # the bottom; their skip part arrives first, from the decoder).
        gyb = {l: self._get("gy%d_b" % l, tuple(t["dec%d_b" % l].shape), bf) for l in range(L_ - 1)}
        gyeb = {l: self._get("gye%d_b" % l, tuple(t["enc%d_b" % l].shape), bf) for l in range(L_)}
        ws = self._workspace(L.sgdct_nn_head_workspace(M, self.c0))
        self._chk(L.sgdct_nn_head_back(_vp(y), _vp(target), _vp(self.head_w), _vp(self.d0), _vp(dy), None,
                                       _vp(gyb[0]), _vp(t["dec0_b"]), _vp(self.head_g), _vp(ws), ws.numel(), M,
                                       self.c0, self._s()), "head_back")
        genc = [self._get("genc%d" % l, tuple(t["enc%d_b" % l].shape), f32) for l in range(L_ - 1)]
        for l in range(0, L_ - 1):
            h, w = H >> l, W >> l
            ca, cb = self.convs["dec%d_a" % l], self.convs["dec%d_b" % l]
            gy = gyb[l]
            self._conv_grads(cb, t["dec%d_a" % l], gy, B, h, w)
            gy2 = self._get("gy%d_a" % l, (B, h, w, ca.cout), bf)
            self._input_grad(cb, gy, gy2, B, h, w, mask=t["dec%d_a" % l])
            self._conv_grads(ca, t["dec%d_in" % l], gy2, B, h, w)
            gin = self._get("gin%d" % l, (B, h, w, ca.cin), f32)
            self._input_grad(ca, gy2, gin, B, h, w)
            cs = genc[l].shape[3]
            cu = ca.cin - cs
            if l + 1 < L_ - 1:
                low_out, low_mask = gyb[l + 1], t["dec%d_b" % (l + 1)]
            else:
                low_out, low_mask = gyeb[L_ - 1], t["enc%d_b" % (L_ - 1)]
            self._chk(L.sgdct_nn_upcat_back(_vp(gin), None, _vp(low_out), _vp(low_mask), _vp(genc[l]), 0, B, h, w,
                                            cu, cs, self._s()), "upcat_back")
        for l in range(L_ - 1, -1, -1):
            h, w = H >> l, W >> l
            ca, cb = self.convs["enc%d_a" % l], self.convs["enc%d_b" % l]
            gy = gyeb[l]
            self._conv_grads(cb, t["enc%d_a" % l], gy, B, h, w)
            gy2 = self._get("gye%d_a" % l, (B, h, w, ca.cout), bf)
            self._input_grad(cb, gy, gy2, B, h, w, mask=t["enc%d_a" % l])
            self._conv_grads(ca, t["enc%d_in" % l], gy2, B, h, w)
            if l > 0:
                gin = self._get("gine%d" % l, (B, h, w, ca.cin), f32)
                self._input_grad(ca, gy2, gin, B, h, w)
                self._chk(L.sgdct_nn_pool2_back(_vp(gin), _vp(genc[l - 1]), _vp(gyeb[l - 1]),
                                                _vp(t["enc%d_b" % (l - 1)]), B, h, w, ca.cin, self._s()),
                          "pool2_back")

    def grads_numpy(self):
        out = {}
        for name, c in self.convs.items():
            out[name] = (c.gw.cpu().numpy().reshape(c.cout, 3, 3, c.cin), c.gb.cpu().numpy())
        g = self.head_g.cpu().numpy()
        out["out"] = (g[:self.c0].reshape(1, 1, 1, self.c0), g[self.c0:])
        return out

    # -- optimiser -----------------------------------------------------------
    def step(self, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        L = _lib.lib()
        self.steps += 1
        s = self._s()
        for c in self.convs.values():
            self._chk(L.sgdct_nn_adam(_vp(c.w32), _vp(c.w16), _vp(c.gw), _vp(c.state[0]), _vp(c.state[1]),
                                      c.w32.numel(), lr, b1, b2, eps, self.steps, s), "adam")
            self._chk(L.sgdct_nn_adam(_vp(c.b32), None, _vp(c.gb), _vp(c.state[2]), _vp(c.state[3]),
                                      c.b32.numel(), lr, b1, b2, eps, self.steps, s), "adam")
        # head: w and b live in head_wb for the update, then copied back
        self.head_wb[:self.c0].copy_(self.head_w)
        self.head_wb[self.c0:].copy_(self.head_b)
        self._chk(L.sgdct_nn_adam(_vp(self.head_wb), None, _vp(self.head_g), _vp(self.head_state[0]),
                                  _vp(self.head_state[1]), self.c0 + 1, lr, b1, b2, eps, self.steps, s), "adam")
        self.head_w.copy_(self.head_wb[:self.c0])
        self.head_b.copy_(self.head_wb[self.c0:])

    def train_step(self, x, target, lr=1e-3):
        y = self.forward(x)
        self.backward(target)
        self.step(lr)
        return y

    def state_dict(self):
        return {k: (v[0].copy(), v[1].copy()) for k, v in self.parameters_numpy().items()}
