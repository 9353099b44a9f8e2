"""Launch-schedule autotuner of the binned kernel (chip_vset_schedule) on the GPU.

  * the first major cycle times the 2 x 4 candidates (threads per workgroup x
    groups per work item) and keeps the fastest of them;
  * every schedule gives the same results: degridded visibilities are
    bit-identical (a group's sum does not depend on how groups are split into
    work items or spread over wavefronts), PSF grids agree to 1e-6 of their
    peak, dirty grids of random-sign visibilities to 1e-5 and their residual
    images to 3e-5 of peak (float atomics add in another order; the data are
    pure noise, so the residual image is a small difference of large grid
    sums, and the grid correction amplifies the rounding at the image
    corners);
  * schedules outside the kernel's limits are rejected.
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def setup(gpu, n_chan=3, nx=128):
    enu = synth.y_array(3, arm_m=3000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 1.0, 30.0)
    n_rows = uvw.shape[0]
    scale = np.linspace(1.0e9, 1.5e9, n_chan) / synth.C
    rng = np.random.default_rng(41)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, (n_rows, n_chan)).astype(np.float32)
    flag = rng.uniform(size=(n_rows, n_chan)) < 0.1
    cell = 1.0 / (3.0 * np.max(np.abs(uvw[:, :2])) * scale[-1])
    maxw = np.max(np.abs(uvw[:, 2])) * scale[-1]
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(maxw, planes=8)
    batch = gpu.RowChannelBatch(t(uvw), t(scale), t(vis), t(wt), t(flag.astype(np.uint8)))
    return g, batch, n_chan, nx


def test_first_major_cycle_keeps_the_fastest_candidate(gpu):
    g, batch, n_chan, nx = setup(gpu)
    vs = g.visibility_set(batch)
    assert vs.schedule() == {"threads": 512, "groups_per_item": 1024, "timed": []}
    psf = g.new_grids(n_chan)
    vs.grid(gpu.MODE_PSF, psf=psf)
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])
    _, _, peaks = g.finish_products(psf, None, nsum)
    model = torch.zeros((n_chan, nx, nx))
    model[:, 40, 50] = 1.0
    vs.major_cycle(model, nsum, peaks)
    sch = vs.schedule()
    timed = sch["timed"]
    assert sorted((a, b) for a, b, _ in timed) == sorted((a, b) for a in (256, 512) for b in (256, 512, 1024, 2048))
    assert all(ms > 0 for _, _, ms in timed)
    best = min(timed, key=lambda x: x[2])
    assert (sch["threads"], sch["groups_per_item"]) == best[:2]
    # a second cycle keeps the schedule without timing again
    vs.major_cycle(model, nsum, peaks)
    assert vs.schedule() == sch


@pytest.mark.parametrize("threads,gpi", [(64, 1), (128, 100), (256, 256), (512, 4096)])
def test_every_schedule_gives_the_same_results(gpu, threads, gpi):
    g, batch, n_chan, nx = setup(gpu)
    ref = g.visibility_set(batch)
    ref.set_schedule(512, 1024)
    vs = g.visibility_set(batch)
    vs.set_schedule(threads, gpi)
    assert vs.schedule()["threads"] == threads and vs.schedule()["groups_per_item"] == gpi
    assert vs.stats()["n_items"] >= ref.stats()["n_items"] if gpi <= 1024 else True

    pr, dr = g.new_grids(n_chan), g.new_grids(n_chan)
    ps, ds = g.new_grids(n_chan), g.new_grids(n_chan)
    ref.grid(gpu.MODE_DIRTY, psf=pr, dirty=dr)
    vs.grid(gpu.MODE_DIRTY, psf=ps, dirty=ds)
    assert rel(ps.cpu().numpy(), pr.cpu().numpy()) < 1e-6
    assert rel(ds.cpu().numpy(), dr.cpu().numpy()) < 1e-5

    rng = np.random.default_rng(9)
    img = np.zeros((n_chan, nx, nx), np.float32)
    for c in range(n_chan):
        img[c, rng.integers(20, nx - 20, 10), rng.integers(20, nx - 20, 10)] = rng.uniform(0.1, 1.0, 10)
    mg = g.model_to_grid(t(img))
    a = ref.degrid(mg).cpu().numpy()
    b = vs.degrid(mg).cpu().numpy()
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))

    nsum = np.array([s["normalization_sumwt"] for s in ref.sums])
    _, _, peaks = g.finish_products(pr, None, nsum)
    r_ref = ref.major_cycle(t(img), nsum, peaks).cpu().numpy()
    r_vs = vs.major_cycle(t(img), nsum, peaks).cpu().numpy()
    print(f"schedule ({threads}, {gpi}): residual {rel(r_vs, r_ref):.2e} of peak")
    assert rel(r_vs, r_ref) < 3e-5
    # fixed schedules are not autotuned
    assert vs.schedule()["timed"] == [] and vs.schedule()["threads"] == threads


def test_schedule_limits(gpu):
    g, batch, n_chan, nx = setup(gpu, n_chan=1, nx=64)
    vs = g.visibility_set(batch)
    for threads, gpi in ((100, 1024), (1024, 1024), (512, 0), (0, 7)):
        with pytest.raises(gpu.InvalidRequest):
            vs.set_schedule(threads, gpi)
    vs.set_schedule(0, 0)  # back to autotuning
    assert vs.schedule()["threads"] == 512
