"""GPU density weighting (uniform / Briggs) against the oracle.

Bars: density cells (which samples land where, which are rejected) are
integer work and exact; densities are exact when the weights are sums of
halves (every partial sum is representable), otherwise within 1e-6 relative
(the GPU sums in FP64 and rounds once, the reference adds Floats in order);
f2 within 1e-6 relative (FP64 sums in another order); reweighted values
within 1e-6 relative, zeros exactly where the reference gives zero.
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def _cube(seed, n_chan=5, weights="halves"):
    enu = synth.y_array(seed, arm_m=3000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 0.5, 30.0)
    n_rows = uvw.shape[0]
    freq = np.linspace(1.0e9, 1.5e9, n_chan)
    scale = freq / synth.C
    rng = np.random.default_rng(seed)
    if weights == "halves":
        wt = rng.choice([0.5, 1.0, 1.5, 2.0], (n_rows, n_chan)).astype(np.float32)
    else:
        wt = rng.uniform(0.1, 3.0, (n_rows, n_chan)).astype(np.float32)
    wt[rng.uniform(size=wt.shape) < 0.01] = 0.0
    wt[rng.uniform(size=wt.shape) < 0.005] = np.nan
    flag = rng.uniform(size=(n_rows, n_chan)) < rng.uniform(0.0, 0.6, n_chan)[None, :]
    vis = np.ones((n_rows, n_chan), np.complex64)
    return uvw, scale, wt, flag, vis


@pytest.mark.parametrize("weights", ["halves", "uniform-random"])
@pytest.mark.parametrize("mode", ["uniform", "briggs"])
def test_per_channel_density_weights(gpu, oracle_mod, mode, weights):
    uvw, scale, wt, flag, vis = _cube(31, weights=weights)
    n_rows, n_chan = wt.shape
    nx = 300
    cell = 1.0 / (3.0 * float(np.max(np.hypot(uvw[:, 0], uvw[:, 1])) * scale[-1]))
    g = gpu.StandardGridder(nx, nx, cell, cell)
    b = gpu.RowChannelBatch(t(uvw), t(scale), t(vis), t(wt), t(flag.astype(np.uint8)))
    m = gpu.WEIGHT_UNIFORM if mode == "uniform" else gpu.WEIGHT_BRIGGS
    got, f2, dens = g.density_weights(b, m, robust=0.5, per_plane=True)
    got, dens = got.cpu().numpy(), dens.cpu().numpy()

    # oracle: sample (row, chan) at index row * n_chan + chan, wavelengths uvw * f / c
    u = (uvw[:, 0:1] * scale[None, :]).ravel()
    v = (uvw[:, 1:2] * scale[None, :]).ravel()
    plane = np.tile(np.arange(n_chan, dtype=np.uint32), n_rows)
    usable = ~flag.ravel()
    om = oracle_mod.WEIGHT_UNIFORM if mode == "uniform" else oracle_mod.WEIGHT_BRIGGS
    ref, rdens, rf2 = oracle_mod.density_weights(nx, nx, cell, cell, u, v, wt.ravel(), om, robust=0.5,
                                                 plane=plane, n_planes=n_chan, usable=usable)
    assert np.array_equal(dens > 0, rdens > 0)
    if weights == "halves":
        assert np.array_equal(dens, rdens)
    else:
        assert np.max(np.abs(dens - rdens) / np.maximum(rdens, 1e-30)) < 1e-6
    assert np.max(np.abs(f2 - rf2) / rf2) < 1e-6
    assert np.array_equal(got == 0, ref == 0)
    nz = (ref != 0) & np.isfinite(ref)
    assert np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])) < 1e-6
    # flagged samples keep their weight
    assert np.array_equal(got[~usable], wt.ravel()[~usable], equal_nan=True)


def test_combined_density_weights(gpu, oracle_mod):
    """One density over every sample, VisImagingWeight convention (Float
    coordinates truncated), with the edge cases the filters reject."""
    nx = 200
    cell = np.deg2rad(20.0 / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    umax = 0.5 / cell
    u, v, w, wt, sf, vis, gr = synth.random_samples(77, 40000, umax)
    wt = np.where(np.isfinite(wt), np.round(wt * 2) / 2, wt).astype(np.float32)  # halves
    b = gpu.VisibilityBatch(t(u), t(v), t(w), t(wt), t(vis), t(sf), t(gr.astype(np.uint8)))
    for m, om in ((gpu.WEIGHT_UNIFORM, oracle_mod.WEIGHT_UNIFORM), (gpu.WEIGHT_BRIGGS, oracle_mod.WEIGHT_BRIGGS)):
        got, f2, dens = g.density_weights(b, m, robust=-0.5, per_plane=False)
        got, dens = got.cpu().numpy(), dens.cpu().numpy()
        ref, rdens, rf2 = oracle_mod.density_weights(nx, nx, cell, cell, u, v, wt, om, robust=-0.5, usable=gr)
        assert np.array_equal(dens, rdens)
        assert abs(f2[0] - rf2[0]) <= 1e-6 * abs(rf2[0])
        assert np.array_equal(got == 0, ref == 0)
        nz = (ref != 0) & np.isfinite(ref)
        assert np.max(np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])) < 1e-6
