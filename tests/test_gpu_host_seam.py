"""The Metal W-projection seam replaced one to one over host arrays.

accumulate_w_project_grid_metal_samples(projector, samples, model_grid,
psf_grid, residual_grid, mode) (lib.rs:71863-71871) takes the reference's
own projector tables and its packed, pre-planned 20-byte samples
(WProjectMetalSample, lib.rs:71629-71754).  Here the oracle's WProjector
stands in for the reference's: its tables go to the device through
chip_gridder_attach_wproject_tables, its plans are packed as the reference
packs them, and numpy grids go through chip_buffer_* and
chip_grid_wproject_samples -- the calls a crate that links only the library
makes.  Bars: the tables read back bit-identical, the device-computed
normalisation table bit-identical to the oracle's (same kernels, same
summation order), grids within 1e-5 of the oracle's Complex32 grids (the
Metal arithmetic rounds weight * (V * cwt), the CPU path (V * w) * cwt).
"""
import numpy as np
import pytest

import synth

pytestmark = pytest.mark.gpu


def _setup(gpu, oracle_mod):
    nx, cell = 128, np.deg2rad(600.0 / 3600.0)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    wp = oracle_mod.WProjector(o, 300.0, explicit_planes=16, sampling=4)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject_tables(wp.kernels, wp.supports, wp.w_scale, wp.sampling)
    return o, wp, g


def _packed(gpu, wp, seed, n):
    u, v, w, wt, sf, vis, gr = synth.random_samples(seed, n, 0.45 * wp.gridder.gx * wp.gridder.du, wmax=300.0,
                                                    edge_cases=False)
    plan, _ = wp.plan(u, v, w)
    ok = plan[:, 0] >= 0
    u, v, w, wt, vis, plan = u[ok], v[ok], w[ok], wt[ok], vis[ok], plan[ok]
    samples = gpu.pack_wproject_samples(plan[:, 0], plan[:, 1], plan[:, 2], plan[:, 3], plan[:, 4], plan[:, 5],
                                        wp.supports[plan[:, 4]], wt, vis)
    return u, v, w, wt, vis, samples


def test_tables_attach_bit_identical(gpu, oracle_mod):
    o, wp, g = _setup(gpu, oracle_mod)
    assert g.info.planes == wp.planes and g.info.sampling == wp.sampling and g.info.w_scale == wp.w_scale
    sup, ker, nt = g.wproject_tables()
    assert np.array_equal(sup, wp.supports)
    assert np.array_equal(ker.view(np.uint32), wp.kernels.view(np.uint32))
    assert np.array_equal(nt.view(np.uint32), wp.norm_table.view(np.uint32))


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_packed_samples_against_oracle(gpu, oracle_mod, mode):
    o, wp, g = _setup(gpu, oracle_mod)
    u, v, w, wt, vis, samples = _packed(gpu, wp, 41, 20000)
    assert len(samples) > 10000
    shape = (g.info.gx, g.info.gy)
    model = None
    if mode == 2:
        img = np.zeros((o.nx, o.ny), np.float32)
        rng = np.random.default_rng(3)
        img[rng.integers(20, 108, 12), rng.integers(20, 108, 12)] = rng.uniform(0.2, 1.0, 12)
        model = wp.model_grid(img)
    psf = np.zeros(shape, np.complex64) if mode != 2 else None
    res = np.zeros(shape, np.complex64) if mode != 0 else None
    gpu.grid_wproject_samples_host(g, samples, mode, model=model, psf=psf, residual=res)
    ones = np.ones(len(u), np.float32)
    opsf, ores, sums = wp.grid(u, v, w, wt, ones, vis, mode=mode, model=model)
    assert sums.gridded == len(u)
    if psf is not None:
        assert np.max(np.abs(psf - opsf)) <= 1e-5 * np.max(np.abs(opsf))
    if res is not None:
        assert np.max(np.abs(res - ores)) <= 1e-5 * np.max(np.abs(ores))


def test_packed_samples_accumulate_and_reject(gpu, oracle_mod):
    o, wp, g = _setup(gpu, oracle_mod)
    _, _, _, _, _, samples = _packed(gpu, wp, 42, 4000)
    shape = (g.info.gx, g.info.gy)
    once = np.zeros(shape, np.complex64)
    gpu.grid_wproject_samples_host(g, samples, 0, psf=once)
    twice = once.copy()
    gpu.grid_wproject_samples_host(g, samples, 0, psf=twice)  # accumulates into the caller's grid
    assert np.max(np.abs(twice - 2 * once)) <= 1e-5 * np.max(np.abs(twice))
    bad = samples[:3].copy()
    bad["loc"][1] = 1 | (1 << 16)  # anchor (1, 1): the footprint leaves the grid
    with pytest.raises(gpu.InvalidRequest, match="outside the grid"):
        gpu.grid_wproject_samples_host(g, bad, 0, psf=np.zeros(shape, np.complex64))
