"""The warm-start U-Net's 3x3 convolution (sgdct_nn_conv3x3_ex, unet.hip)
against torch's FP64 conv2d on the CPU, on bf16-valued inputs and weights:

  * FP32 output (the input-gradient form): <= 1e-5 relative L2 (the kernel
    accumulates the bf16 products in FP32; only the summation order differs);
  * bf16 output with bias, ReLU and an activation mask: elementwise within
    one bf16 rounding (2^-8 relative) of the reference;
  * image sizes that are not multiples of the 16 x 16 / 8 x 16 tiles, input
    channels 8, 32, 64 and the decoder's concatenation [upsample2(low), skip]
    (64 + 32 channels, read in place), output channels 32, 64, 128.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _vp(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _ref(x_nhwc, wk, bias, cin, cout):
    """FP64 conv2d with zero padding; wk[o][(3 ky + kx) cin + c]."""
    import torch
    x = torch.tensor(x_nhwc, dtype=torch.float64).permute(0, 3, 1, 2)
    w = torch.tensor(wk, dtype=torch.float64).reshape(cout, 3, 3, cin).permute(0, 3, 1, 2)
    y = torch.nn.functional.conv2d(x, w, padding=1)
    if bias is not None:
        y = y + torch.tensor(bias, dtype=torch.float64)[None, :, None, None]
    return y.permute(0, 2, 3, 1).numpy()


def _bf16(a):
    import torch
    return torch.tensor(a, dtype=torch.float32).to(torch.bfloat16)


CASES = [
    # B, H, W, cin, cout, concat (low channels)
    (2, 40, 56, 32, 32, 0),
    (2, 24, 40, 8, 32, 0),
    (2, 20, 36, 64, 64, 0),
    (1, 17, 30, 32, 128, 0),
    (2, 24, 40, 96, 32, 64),
]


@pytest.mark.parametrize("B,H,W,cin,cout,c0", CASES)
def test_conv3x3_against_fp64(cuda, B, H, W, cin, cout, c0):
    import torch
    from sgd_amd import _lib
    L = _lib.lib()
    print("conv3x3 B %d %dx%d cin %d cout %d concat %d" % (B, H, W, cin, cout, c0), flush=True)
    rng = np.random.default_rng(cin * 7 + cout)
    w = rng.normal(0, 0.1, (cout, 9 * cin)).astype(np.float32)
    bias = rng.normal(0, 0.1, cout).astype(np.float32)
    wb = _bf16(w).to(cuda)
    if c0:
        low = rng.random((B, H // 2, W // 2, c0)).astype(np.float32)
        skip = rng.random((B, H, W, cin - c0)).astype(np.float32)
        lowb, skipb = _bf16(low).to(cuda), _bf16(skip).to(cuda)
        up = np.repeat(np.repeat(lowb.float().cpu().numpy(), 2, axis=1), 2, axis=2)
        xin = np.concatenate([up, skipb.float().cpu().numpy()], axis=3)
        src, src2 = lowb, skipb
    else:
        x = rng.random((B, H, W, cin)).astype(np.float32) - 0.3
        src, src2 = _bf16(x).to(cuda), None
        xin = src.float().cpu().numpy()
    wr = wb.float().cpu().numpy()
    # FP32 output, no bias or activation
    out = torch.empty((B, H, W, cout), dtype=torch.float32, device=cuda)
    assert L.sgdct_nn_conv3x3_ex(_vp(src), _vp(src2), c0, _vp(wb), None, _vp(out), None, B, H, W, cin, cout, 0, 1,
                                 None) == 0, _lib.last_error()
    torch.cuda.synchronize()
    ref = _ref(xin, wr, None, cin, cout)
    e = float(np.linalg.norm(out.double().cpu().numpy() - ref) / np.linalg.norm(ref))
    print("  FP32 output rel %.2e" % e)
    assert e <= 1e-5
    # bf16 output with bias, ReLU and a mask
    mask = _bf16(rng.random((B, H, W, cout)) - 0.3).to(cuda)
    b32 = torch.tensor(bias, device=cuda)
    outb = torch.empty((B, H, W, cout), dtype=torch.bfloat16, device=cuda)
    assert L.sgdct_nn_conv3x3_ex(_vp(src), _vp(src2), c0, _vp(wb), _vp(b32), _vp(outb), _vp(mask), B, H, W, cin,
                                 cout, 1, 0, None) == 0, _lib.last_error()
    torch.cuda.synchronize()
    refb = np.maximum(_ref(xin, wr, bias, cin, cout), 0.0) * (mask.float().cpu().numpy() > 0)
    got = outb.float().cpu().numpy()
    assert np.all(np.abs(got - refb) <= 2.0 ** -8 * np.abs(refb) + 1e-6)
    assert (got == 0).mean() > 0.2  # the mask and ReLU zero a part
