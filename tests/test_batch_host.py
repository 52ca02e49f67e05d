"""Host-side logic of the batched-slice workload (no GPU): the uneven slice
generator follows its stated distribution and geometry rules."""
import numpy as np


def test_uneven_slices_distribution():
    from sgd_amd.batch import uneven_slices
    specs = uneven_slices(256, seed=0)
    sizes = {s["nx"] for s in specs}
    assert sizes == {128, 256, 384, 512, 768}
    assert all(s["nx"] == s["ny"] and s["nu"] == round(1.5 * s["nx"]) for s in specs)
    views = np.array([s["np"] for s in specs])
    assert views.min() >= 180 and views.max() <= 1440 and len(set(views)) > 100
    assert {s["basis"] for s in specs} == {"B0", "B1", "SF"}
    for s in specs:
        # the detector covers the field of view's circumscribed circle
        r = 256.0 / np.sqrt(2.0)
        half_fan = np.arcsin(r / s["lso"])
        assert s["nu"] * s["du"] / 2 >= s["lsd"] * np.tan(half_fan)
        assert len(s["angles"]) == s["np"] and abs(s["dx"] * s["nx"] - 256.0) < 1e-9
    # same seed, same workload
    again = uneven_slices(256, seed=0)
    assert all(a["nx"] == b["nx"] and a["np"] == b["np"] and a["basis"] == b["basis"] for a, b in zip(specs, again))
