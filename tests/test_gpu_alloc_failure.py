"""Allocation failures leave handles and the runtime usable.  A cone-beam
handle over 8192 x 8192 columns of one voxel and 8192 detector rows gets its
256 MiB column-major volume buffer, but one view's row profiles (2 TiB)
cannot be allocated.  Each projection call must then fail with the
allocation error before launching anything, the second call too, since a
half-built workspace is released rather than kept.  A small projection on
another handle afterwards must succeed: the failed allocation's error is not
left for the next launch check to report."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_cone_workspace_failure(cuda):
    import torch
    from sgd_amd._lib import SgdctError
    from sgd_amd.torch_ops import Projector
    n, nv, nu = 8192, 8192, 16
    print("cone handle %d x %d x 1 voxels, %d x %d detector" % (n, n, nu, nv), flush=True)
    P = Projector.cone_beam(n, n, 1, np.array([0.3]), nu, nv, 0.05, 0.05, 0.05, 1.0, 1.0, 1000.0, 500.0)
    x = torch.zeros((1, n, n), dtype=torch.float32, device=cuda)
    for _ in range(2):
        with pytest.raises(SgdctError, match="cone-beam workspace"):
            P.forward(x)
    with pytest.raises(SgdctError, match="cone-beam workspace"):
        P.backward(torch.zeros((1, nv, nu), dtype=torch.float32, device=cuda))
    del P, x
    torch.cuda.synchronize()
    Q = Projector("B1", 32, 32, 2 * np.pi * np.arange(8) / 8, 48, 1.0, 1.0, 1.0, 200.0, 100.0)
    y = Q.forward(torch.ones((32, 32), dtype=torch.float32, device=cuda))
    torch.cuda.synchronize()
    assert float(y.sum()) > 0
