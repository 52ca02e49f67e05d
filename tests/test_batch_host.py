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


def test_helical_slices_carry_per_view_offsets():
    """helical=True adds per-view source offsets and changes nothing else: dz
    rises over one pitch centred on the slice plane, drho alternates with the
    flying focal spot (+-0.6 mm) plus small jitter."""
    from sgd_amd.batch import uneven_slices
    plain = uneven_slices(32, seed=4)
    hel = uneven_slices(32, seed=4, helical=True)
    for a, b in zip(plain, hel):
        assert all(np.array_equal(a[k], b[k]) if k == "angles" else a[k] == b[k] for k in a)
        dz, drho = b["dz"], b["drho"]
        assert dz.shape == drho.shape == (b["np"],)
        assert np.all(np.diff(dz) > 0) and abs(dz[0] + 10.0) < 1e-12 and dz[-1] < 10.0
        assert np.all(np.abs(drho[0::2] - 0.6) < 0.3) and np.all(np.abs(drho[1::2] + 0.6) < 0.3)
    assert not np.array_equal(hel[0]["drho"], hel[1]["drho"][:hel[0]["np"]])
