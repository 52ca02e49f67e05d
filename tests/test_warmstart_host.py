"""Host-side pieces of the warm-start experiment (sgd_amd/warmstart.py):
the Poisson noise-level estimate that sets the discrepancy-principle
tolerances, and the tolerance parser."""
import numpy as np
import torch


def test_noise_level_matches_drawn_noise():
    from sgd_amd import warmstart as ws
    g = torch.Generator().manual_seed(0)
    # line integrals of 0..200 mm of water-like attenuation, 3 sinograms
    y = torch.rand((3, 90, 768), generator=g, dtype=torch.float64) * 200.0
    counts = torch.poisson(ws.N0 * torch.exp(-ws.MU * y), generator=g).clamp_min(1.0)
    b = -torch.log(counts / ws.N0) / ws.MU
    est = ws.noise_level(b).numpy()
    true = (torch.linalg.vector_norm(b - y, dim=(1, 2)) / torch.linalg.vector_norm(b, dim=(1, 2))).numpy()
    assert np.all(np.abs(est / true - 1.0) < 0.03), (est, true)


def test_parse_tol():
    from sgd_amd import warmstart as ws
    assert ws.parse_tol("2s") == ("2s", 2.0, True)
    assert ws.parse_tol(" 1e-2") == ("1e-2", 0.01, False)
