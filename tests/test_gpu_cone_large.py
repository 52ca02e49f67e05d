"""Cone-beam projectors at a large size: a 512^3 volume of 0.5 mm voxels,
360 views, a 1024 x 768 detector of 0.8 mm cells (lso 500, lsd 1000 mm;
8x the voxels and 4x the detector cells of the 256^3 workload, 805 MB of row
profiles per view, so a pass holds two views).  The detector is shorter than
the volume's cone near the source, so rows are also cut at its edges.  For
B1 (A_B1_3D.cu), SF-TR and SF-TT (SF_projector.cu, fixed):
  * A on two views and A^T of a projection set holding only those views,
    against the FP64 oracle: <= 1e-5 relative L2 (A^T of white-noise rows is
    the sensitive one: as FP32 detector rows up to ~800 the axial
    trapezoids' corners would be ~3e-5 rows from exact, which A^T's
    per-voxel anchoring avoids; the reference's own FP32 formulas, the
    oracle's FP32 build, are printed beside it, 1.1e-5 .. 1.7e-5);
  * the adjoint <A x, y> = <x, A^T y> (FP64 sums) on full-size random data,
    <= 1e-6."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N, NV_, NU, NVR = 512, 360, 1024, 768
D, DU, LSD, LSO = 0.5, 0.8, 1000.0, 500.0
VIEWS = (37, 181)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


@pytest.mark.parametrize("model", ["B1", "SF-TR", "SF-TT"])
def test_large_cone(cuda, oracle_exact, oracle_ref, model):
    import torch
    from pyoracle import Geometry3
    from sgd_amd.torch_ops import Projector
    ang = 2 * np.pi * np.arange(NV_) / NV_
    print("large cone %s: %d^3, %d views, %d x %d detector" % (model, N, NV_, NU, NVR), flush=True)
    P = Projector.cone_beam(N, N, N, ang, NU, NVR, D, D, D, DU, DU, LSD, LSO, model=model)
    G = Geometry3(oracle_exact, N, N, N, ang, NU, NVR, D, D, D, DU, DU, LSD, LSO)
    G32 = Geometry3(oracle_ref, N, N, N, ang, NU, NVR, D, D, D, DU, DU, LSD, LSO)
    tv = model == "SF-TT"
    fwd = G.forward if model == "B1" else (lambda v, views: G.sf_forward(v, trap_v=tv, views=views))
    bwd = G.backward if model == "B1" else (lambda p, views: G.sf_backward(p, trap_v=tv, views=views))
    bwd32 = G32.backward if model == "B1" else (lambda p, views: G32.sf_backward(p, trap_v=tv, views=views))
    rng = np.random.default_rng(21)
    x = rng.random((N, N, N), dtype=np.float32)
    xt = torch.from_numpy(x).to(cuda)
    y = P.forward(xt)
    torch.cuda.synchronize()
    x64 = x.astype(np.float64)
    for p in VIEWS:
        yo = fwd(x64, views=(p, p + 1))[p]
        e = rel(y[p].double().cpu().numpy(), yo)
        print("  A view %d rel %.2e" % (p, e), flush=True)
        assert e <= 1e-5, (model, p, e)
    del yo
    s = torch.zeros((NV_, NVR, NU), dtype=torch.float32, device=cuda)
    z_ref = np.zeros((N, N, N))
    s_host = np.zeros((NV_, NVR, NU))
    for p in VIEWS:
        s_host[p] = rng.random((NVR, NU))
        s[p] = torch.from_numpy(s_host[p]).float().to(cuda)
        s_host[p] = s[p].double().cpu().numpy()  # the values the GPU sees
    z32 = np.zeros((N, N, N))
    for p in VIEWS:
        z_ref += bwd(s_host, views=(p, p + 1))
        z32 += bwd32(s_host, views=(p, p + 1))
    z = P.backward(s).double().cpu().numpy()
    e, e32 = rel(z, z_ref), rel(z32, z_ref)
    print("  A^T on %d views rel %.2e (FP32 formulas %.2e)" % (len(VIEWS), e, e32), flush=True)
    assert e <= 1e-5, (model, e, e32)
    yr = torch.rand((NV_, NVR, NU), device=cuda, dtype=torch.float32)
    aty = P.backward(yr)
    lhs = float((y.double() * yr.double()).sum())
    rhs = float((xt.double() * aty.double()).sum())
    e = abs(lhs - rhs) / abs(rhs)
    print("  adjoint rel %.2e" % e, flush=True)
    assert e <= 1e-6
