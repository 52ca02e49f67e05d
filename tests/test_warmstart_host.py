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


def test_curve_stats_recovers_a_shift():
    """A warm curve that is the plain curve shifted by d iterations: the
    offset is d wherever the shifted point lies inside the plain run, and
    censored beyond; iterations to the tolerance differ by d."""
    from sgd_amd import warmstart as ws
    iters, d = 1000, 150
    k = np.arange(iters + d, dtype=np.float64)[:, None]
    floors = np.array([0.005, 0.006, 0.0095])
    f = floors + 0.05 * np.exp(-k / 200.0)  # (iters + d, 3)
    hp, hw = f[:iters], f[d:d + iters]
    sig = floors / 0.75
    out = ws.curve_stats(hp, hw, sig, tol=1e-2, marks=(0, 10, 500, 849, 900))
    assert out["offset_median"] == [d] * 4 + [None]
    assert out["offset_censored"] == [0, 0, 0, 0, 3]
    # the third image's floor is 0.0095: it reaches 1e-2 at k = 200 ln(0.05 / 0.0005)
    kp = np.ceil(200.0 * np.log(0.05 / np.array([0.005, 0.004, 0.0005]))).astype(int)
    assert out["tol_plain_reached"] == 3 and out["tol_warm_reached"] == 3
    assert out["tol_both_plain_mean"] == kp.mean()
    assert out["tol_both_warm_mean"] == (kp - d).mean()
    assert out["tol_saved_median"] == d
    assert out["plain_median"][0] == f[0, 1]
