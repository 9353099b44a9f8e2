"""W tables recomputed on the fly (chip_gridder_wkernel_budget) against the
resident tables, on the GPU.

Bars:
  * kernels, supports and the per-(plane, offset) normalisation table of a
    gridder whose tables are recomputed are bit-identical to the resident ones
    (same screen, one FFT per plane, same crop and normalisation);
  * degridded visibilities agree to 1e-6 of their peak: the same taps, but
    the register-slot layout that decides which lanes sum a footprint
    follows the work item's largest support, and work items are cut
    differently once the stream is split by plane chunk;
  * FP32 PSF grids agree to 2e-6 of their peak, dirty grids and the adjoint
    grid (random-sign visibilities and values) to 1e-5 of their peak (their
    cells are sums of terms of both signs, so the order of the float atomics,
    which changes once the work is split by plane chunk, shows more),
    residual images to 1e-5 of
    peak (the grid correction amplifies grid rounding up to ~25x per axis at
    the image corners); FP64 plain grids to 1e-12;
  * a visibility set binned before the residency changed is rejected.
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu

NX, ARCSEC, PLANES, SAMPLING, MAXW = 128, 600.0, 16, 4, 300.0


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def make(gpu, budget=None):
    cell = np.deg2rad(ARCSEC / 3600.0)
    g = gpu.StandardGridder(NX, NX, cell, cell)
    chunk = None
    if budget is not None:
        chunk = g.set_wkernel_budget(budget)
        assert chunk == 0  # nothing attached yet
    g.attach_wproject(MAXW, planes=PLANES, sampling=SAMPLING)
    return g


def budget_for(g, planes_per_chunk):
    """A budget that holds the tables of `planes_per_chunk` planes plus one FFT plane."""
    i = g.info
    axis = 2 * i.norm_radius + 1
    per_plane = (i.width * i.width + axis * axis * (2 * i.max_support + 1) ** 2) * 8
    return planes_per_chunk * per_plane + i.gx * i.gy * 8


def samples(gpu, g, seed=31, n=30000):
    umax = 0.45 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(seed, n, umax, wmax=MAXW)
    u[: n // 3] *= 0.1
    v[: n // 3] *= 0.1
    b = gpu.VisibilityBatch(t(u), t(v), t(w), t(wt), t(vis), t(sf), t(gr.astype(np.uint8)))
    return b


def test_recomputed_tables_are_bit_identical(gpu):
    res = make(gpu)
    otf = make(gpu, budget=0)
    chunk = otf.set_wkernel_budget(budget_for(res, 3))
    assert chunk == 3
    s0, k0, n0 = res.wproject_tables()
    s1, k1, n1 = otf.wproject_tables()
    assert np.array_equal(s0, s1)
    assert np.array_equal(k0.view(np.uint32), k1.view(np.uint32))
    assert np.array_equal(n0.view(np.uint32), n1.view(np.uint32))
    # the budget may also be set before attaching; a budget above the tables keeps them resident
    otf2 = make(gpu, budget=budget_for(res, 5))
    assert otf2.info.planes == PLANES
    assert otf2.set_wkernel_budget(budget_for(res, 5)) == 5
    assert otf2.set_wkernel_budget(budget_for(res, PLANES + 1)) == 0
    assert otf2.set_wkernel_budget(0) == 0


@pytest.mark.parametrize("per_chunk", [1, 4])
def test_on_the_fly_matches_resident(gpu, per_chunk):
    res = make(gpu)
    otf = make(gpu)
    assert otf.set_wkernel_budget(budget_for(res, per_chunk)) == per_chunk
    b = samples(gpu, res)
    vr, vo = res.visibility_set(b), otf.visibility_set(b)
    assert vr.stats()["n_valid"] == vo.stats()["n_valid"]
    for a, c in zip(vr.sums, vo.sums):
        assert a["gridded"] == c["gridded"]
        assert abs(a["normalization_sumwt"] - c["normalization_sumwt"]) <= 1e-12 * a["normalization_sumwt"]

    # dirty + PSF grids
    pr, dr = res.new_grids(1), res.new_grids(1)
    po, do = otf.new_grids(1), otf.new_grids(1)
    vr.grid(gpu.MODE_DIRTY, psf=pr, dirty=dr)
    vo.grid(gpu.MODE_DIRTY, psf=po, dirty=do)
    assert rel(po.cpu().numpy(), pr.cpu().numpy()) < 2e-6
    assert rel(do.cpu().numpy(), dr.cpu().numpy()) < 1e-5

    # forward operator
    rng = np.random.default_rng(3)
    img = np.zeros((NX, NX), np.float32)
    img[rng.integers(20, NX - 20, 12), rng.integers(20, NX - 20, 12)] = rng.uniform(0.1, 1.0, 12)
    model = res.model_to_grid(t(img))
    d_r = vr.degrid(model).cpu().numpy()
    d_o = vo.degrid(model).cpu().numpy()
    assert rel(d_o, d_r) < 1e-6

    # adjoint
    n = b.u_lambda.numel()
    y = t((rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)).cuda()
    ar, ao = res.new_grids(1), otf.new_grids(1)
    vr.grid_values(y, ar)
    vo.grid_values(y, ao)
    assert rel(ao.cpu().numpy(), ar.cpu().numpy()) < 1e-5

    # residual refresh (major cycle) of the same model image
    nsum = vr.sums[0]["normalization_sumwt"]
    _, _, peak = res.finish_products(pr.clone(), dr.clone(), nsum)
    rr = vr.major_cycle(t(img[None]), nsum, peak).cpu().numpy()
    ro = vo.major_cycle(t(img[None]), nsum, peak).cpu().numpy()
    print(f"residual image, recomputed vs resident tables: {rel(ro, rr):.2e} of peak")
    assert rel(ro, rr) < 1e-5

    # plain FP64 path, one pass per chunk
    p64r, d64r = res.new_grids(1, gpu.F64), res.new_grids(1, gpu.F64)
    p64o, d64o = otf.new_grids(1, gpu.F64), otf.new_grids(1, gpu.F64)
    sr = res.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=p64r, dirty=d64r)[0]
    so = otf.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=p64o, dirty=d64o)[0]
    assert sr["gridded"] == so["gridded"]
    assert abs(sr["normalization_sumwt"] - so["normalization_sumwt"]) <= 1e-12 * sr["normalization_sumwt"]
    assert rel(d64o.cpu().numpy(), d64r.cpu().numpy()) < 1e-12
    assert rel(p64o.cpu().numpy(), p64r.cpu().numpy()) < 1e-12
    m64 = model.to(torch.complex128)
    q_r = res.degrid_plain(b, m64, gpu.F64).cpu().numpy()
    q_o = otf.degrid_plain(b, m64, gpu.F64).cpu().numpy()
    assert np.array_equal(q_r.view(np.uint64), q_o.view(np.uint64))


def test_set_binned_before_a_residency_change_is_rejected(gpu):
    g = make(gpu)
    b = samples(gpu, g, n=2000)
    vs = g.visibility_set(b)
    g.set_wkernel_budget(budget_for(g, 2))
    with pytest.raises(RuntimeError, match="residency"):
        vs.grid(gpu.MODE_DIRTY, psf=g.new_grids(1), dirty=g.new_grids(1))
    vs2 = g.visibility_set(b)
    vs2.grid(gpu.MODE_DIRTY, psf=g.new_grids(1), dirty=g.new_grids(1))
