"""The batched-slice workload at its stated size: 256 fan slices rebinned from
a helical scan (per-view source offsets dz and drho), image edges uniform over
{128, 256, 384, 512, 768} (256 mm field of view), 180..1440 views, 1.5x as
many bins as the image edge, bases mixed B0 / B1 / SF, one persistent launch
per direction with the launch shape the autotuner picks.

  * the persistent launch equals the per-slice kernels to one FP32 rounding
    (A and A^T), at the autotuned shape;
  * a captured hipGraph replays to the eager result, three times, with eager
    work and new input values between the replays;
  * eight slices -- the largest image with the most views among them, the
    slice with the most views overall, and slices of every basis -- against
    the FP64 oracle (with the same per-view offsets) on 16 sampled views each:
    A's rows and A^T of a sinogram that is nonzero on those views only,
    <= 1e-5 relative L2.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_EXACT = 1e-5


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _same(a, b):
    import torch
    return bool(torch.allclose(a, b, rtol=2.4e-7, atol=1e-30))


@pytest.fixture(scope="module")
def helical256(cuda):
    from sgd_amd.batch import SliceBatch, make_projectors, uneven_slices
    print("helical batch: 256 slices, creating projectors", flush=True)
    specs = uneven_slices(256, seed=0, helical=True)
    B = SliceBatch(make_projectors(specs))
    return specs, B


def _checked_slices(specs):
    order = sorted(range(len(specs)), key=lambda i: (specs[i]["nx"], specs[i]["np"]))
    picks = [order[-1], max(range(len(specs)), key=lambda i: specs[i]["np"])]
    rng = np.random.default_rng(1)
    for basis in ("B0", "B1", "SF"):
        cand = [i for i in range(len(specs)) if specs[i]["basis"] == basis and i not in picks]
        picks += [int(c) for c in rng.choice(cand, 2, replace=False)]
    return sorted(set(picks))[:8]


def test_helical256_persistent_equals_per_slice(helical256, cuda, tmp_path):
    import torch
    specs, B = helical256
    sizes = sorted({s["nx"] for s in specs})
    views = [s["np"] for s in specs]
    print("helical batch: sizes %s, views %d..%d, bases %s, %d image / %d sinogram values"
          % (sizes, min(views), max(views), sorted({s["basis"] for s in specs}), B.img_off[-1], B.sino_off[-1]),
          flush=True)
    assert sizes == [128, 256, 384, 512, 768] and all("drho" in s and "dz" in s for s in specs)
    rec = B.tune(cache_path=str(tmp_path / "tune.json"), reps=1)
    print("helical batch: tuned A %d blocks/CU, A^T %d, order %d (A %.1f ms, A^T %.1f ms)"
          % (rec["fwd_blocks_per_cu"], rec["bwd_blocks_per_cu"], rec["order"], rec["fwd_ms"], rec["bwd_ms"]),
          flush=True)
    g = torch.Generator(device=cuda).manual_seed(0)
    img = torch.rand(B.img_off[-1], device=cuda, generator=g)
    sino = torch.rand(B.sino_off[-1], device=cuda, generator=g)
    s, i = B.forward(img), B.backward(sino)
    ref_s, ref_i = B.forward_per_slice(img), B.backward_per_slice(sino)
    torch.cuda.synchronize()
    assert _same(s, ref_s)
    assert _same(i, ref_i)
    out_s, out_i = B.empty_sinos(), B.empty_images()
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        for rep in range(3):
            B.forward(img, out=out_s, graph=True)
            B.backward(sino, out=out_i, graph=True)
            stream.synchronize()
            assert torch.equal(out_s, B.forward(img)), rep
            assert torch.equal(out_i, B.backward(sino)), rep
            img.mul_(0.75).add_(0.125)
            sino.mul_(1.5)
            stream.synchronize()
    print("helical batch: persistent == per-slice, graph replays == eager", flush=True)


def test_helical256_sampled_views_against_oracle(helical256, cuda):
    import torch
    from pyoracle import Geometry, Oracle
    specs, B = helical256
    picks = _checked_slices(specs)
    assert len(picks) == 8 and {specs[i]["basis"] for i in picks} == {"B0", "B1", "SF"}
    assert max(specs[i]["nx"] for i in picks) == 768
    g = torch.Generator(device=cuda).manual_seed(5)
    img = torch.rand(B.img_off[-1], device=cuda, generator=g)
    sino = torch.zeros(B.sino_off[-1], device=cuda)
    sampled = {}
    for k in picks:
        npv = specs[k]["np"]
        sampled[k] = np.unique(np.linspace(0, npv - 1, 16).round().astype(int))
        view = B.sino(sino, k)
        view[torch.as_tensor(sampled[k], device=cuda)] = torch.rand((len(sampled[k]), specs[k]["nu"]), device=cuda,
                                                                    generator=g)
    ys, xs = B.forward(img), B.backward(sino)
    torch.cuda.synchronize()
    ox = Oracle("exact")
    for k in picks:
        s = specs[k]
        print("slice %d: %s %dx%d, %d views, %d bins" % (k, s["basis"], s["nx"], s["ny"], s["np"], s["nu"]),
              flush=True)
        ge = Geometry(ox, s["nx"], s["ny"], s["angles"], s["nu"], s["dx"], s["dy"], s["du"], s["lsd"], s["lso"],
                      drho=s["drho"])
        im = B.image(img, k).double().cpu().numpy()
        ya = B.sino(ys, k).double().cpu().numpy()
        sn = B.sino(sino, k).double().cpu().numpy()
        ye = np.zeros_like(ya)
        xe = np.zeros((s["ny"], s["nx"]))
        for p in sampled[k]:
            ye[p] = ge.forward(s["basis"], im, views=(int(p), int(p) + 1))[p]
            xe += ge.backward(s["basis"], sn, views=(int(p), int(p) + 1))
        ea = rel(ya[sampled[k]], ye[sampled[k]])
        eb = rel(B.image(xs, k).double().cpu().numpy(), xe)
        print("  A rows %.2e, A^T %.2e" % (ea, eb), flush=True)
        assert ea <= TOL_EXACT and eb <= TOL_EXACT
