"""Cone-beam projectors over several passes (capi.cpp cone_forward /
cone_backward): with 256 x 256 columns and 384 detector rows a row-profile
pass holds 16 views, so 40 views run as passes of 16, 16 and 8, A's
alternating between two row-profile buffers and two streams.  Each view's
projection must equal, bit for bit, what single-pass handles over the views
[0, 16), [16, 32) and [32, 40) give, on the default stream and on a created
stream; A^T, which sums the passes, must equal the sum of the three handles'
A^T to FP32 rounding.  B1 and SF-TR."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NX = NY = 256
NZ, NU, NV = 4, 48, 384
D, DU, DV, LSD, LSO = 0.8, 1.6, 1.6, 1000.0, 500.0


@pytest.mark.parametrize("model", ["B1", "SF-TR"])
def test_passes_equal_single_pass_handles(cuda, model):
    import torch
    from sgd_amd.torch_ops import Projector
    ang = 2 * np.pi * np.arange(40) / 40 + 0.1
    print("cone %s: 40 views in passes vs 3 single-pass handles" % model, flush=True)
    P = Projector.cone_beam(NX, NY, NZ, ang, NU, NV, D, D, D, DU, DV, LSD, LSO, model=model)
    parts = [Projector.cone_beam(NX, NY, NZ, ang[a:b], NU, NV, D, D, D, DU, DV, LSD, LSO, model=model)
             for a, b in ((0, 16), (16, 32), (32, 40))]
    g = torch.Generator(device=cuda).manual_seed(3)
    x = torch.rand((NZ, NY, NX), device=cuda, generator=g)
    y = torch.rand((40, NV, NU), device=cuda, generator=g)
    ref = torch.cat([Q.forward(x) for Q in parts])
    torch.cuda.synchronize()
    for stream in (None, torch.cuda.Stream()):
        if stream is None:
            a = P.forward(x)
        else:
            with torch.cuda.stream(stream):
                a = P.forward(x)
            stream.synchronize()
        torch.cuda.synchronize()
        assert torch.equal(a, ref), float((a - ref).abs().max())
    at = P.backward(y)
    at_ref = parts[0].backward(y[:16]) + parts[1].backward(y[16:32]) + parts[2].backward(y[32:])
    torch.cuda.synchronize()
    e = float((at - at_ref).norm() / at_ref.norm())
    print("  A bitwise equal; A^T rel %.2e" % e)
    assert e <= 1e-6
