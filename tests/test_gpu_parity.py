"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bars (stated per test):
  * integer work (plans, cell and kernel indices, accept/skip counts) is
    bit-exact;
  * tables built on the host (tap weights, corrections) are bit-exact;
  * FP64 plain-path grids match the oracle's FP64 grids to 1e-12 relative
    (only the summation order differs);
  * FP32 grids, degridded visibilities and images match to the tolerance
    written in each test (FP32 MACs and FP32 FFT against FP64);
  * W kernels match the oracle's FP64-FFT kernels to FFT tolerance, with
    supports equal wherever the deciding value is not within 1e-5 of the
    1e-3 threshold.
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def t(x, dtype=None):
    x = torch.from_numpy(np.ascontiguousarray(x))
    return x if dtype is None else x.to(dtype)


def batch_of(pkg, u, v, w, wt, sf, vis, gridable):
    return pkg.VisibilityBatch(t(u), t(v), t(w), t(wt), t(vis), t(sf), t(gridable.astype(np.uint8)))


# --------------------------------------------------------------------------
# standard (spheroidal) path
# --------------------------------------------------------------------------

STD_CASES = [
    # nx, cell (arcsec), oversampling, padding
    (64, 20.0, 100, 0),
    (100, 7.0, 128, 0),
    (96, 3.0, 128, 1),
]


@pytest.mark.parametrize("nx,arcsec,os_,pad", STD_CASES)
def test_standard_tables_bit_exact(gpu, oracle_mod, nx, arcsec, os_, pad):
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell, padding_mode=pad, oversampling=os_)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell, padding_mode=pad, oversampling=os_)
    assert g.grid_shape == o.grid_shape and g.image_blc == o.blc
    assert g.info.du == o.du and g.info.dv == o.dv
    taps, cx, cy = g.tables()
    assert np.array_equal(taps, o.taps)
    assert np.array_equal(cx, o.correction_x) and np.array_equal(cy, o.correction_y)


@pytest.mark.parametrize("nx,arcsec,os_,pad", STD_CASES)
def test_standard_plan_bit_exact(gpu, oracle_mod, nx, arcsec, os_, pad):
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell, padding_mode=pad, oversampling=os_)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell, padding_mode=pad, oversampling=os_)
    umax = 0.55 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(1, 20000, umax, edge_cases=False)
    hu, hv = synth.half_cell_samples(o.du, o.dv, 30, os_)
    u, v = np.concatenate([u, hu]), np.concatenate([v, hv])
    n = len(u)
    ones = np.ones(n, np.float32)
    b = batch_of(gpu, u, v, np.zeros(n), ones, ones, np.ones(n, np.complex64), np.ones(n, bool))
    plan, _ = g.plan(b)
    plan = plan.cpu().numpy()
    ref = o.plan(u, v)
    assert np.array_equal(plan[:, :4], ref.astype(np.int32))
    assert 0 < np.sum(ref[:, 0] >= 0) < n  # both accepted and rejected samples present


def with_sources(seed, u, v, w, vis, nx, cell, n_src=8, noise=0.05):
    """Replace the finite visibilities by point sources plus noise, so that an
    image error 'of peak' is measured against a real dirty-image peak."""
    rng = np.random.default_rng(seed)
    src = [(rng.uniform(0.2, 1.0), rng.integers(-nx // 3, nx // 3) * cell,
            rng.integers(-nx // 3, nx // 3) * cell) for _ in range(n_src)]
    ok = np.isfinite(u) & np.isfinite(v) & np.isfinite(w)
    sky = np.zeros_like(vis)
    sky[ok] = synth.point_source_vis((u[ok], v[ok], w[ok]), src)
    noisy = (sky + noise * (rng.normal(size=len(u)) + 1j * rng.normal(size=len(u)))).astype(np.complex64)
    return np.where(np.isfinite(vis), noisy, vis).astype(np.complex64)


def _std_setup(gpu, oracle_mod, seed=5, n=30000, nx=100, arcsec=7.0, os_=128, sources=False):
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell, oversampling=os_)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell, oversampling=os_)
    umax = 0.5 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(seed, n, umax)
    # cluster half the samples near the uv centre (dense tiles, atomics contention)
    u[: n // 2] *= 0.1
    v[: n // 2] *= 0.1
    if sources:
        vis = with_sources(seed, u, v, w, vis, nx, cell)
    return g, o, (u, v, w, wt, sf, vis, gr)


def test_standard_dirty_grids(gpu, oracle_mod):
    g, o, (u, v, w, wt, sf, vis, gr) = _std_setup(gpu, oracle_mod)
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    opsf, odirty, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=1)

    # FP64 plain path: identical arithmetic, summation order differs
    psf64, dirty64 = g.new_grids(1, gpu.F64), g.new_grids(1, gpu.F64)
    sums = g.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=psf64, dirty=dirty64)[0]
    assert (sums["gridded"], sums["skipped"]) == (osums.gridded, osums.skipped)
    assert abs(sums["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
    assert rel(psf64[0].cpu().numpy(), opsf) < 1e-12
    assert rel(dirty64[0].cpu().numpy(), odirty) < 1e-12

    # FP32 binned path
    vs = g.visibility_set(b)
    assert vs.sums[0]["gridded"] == osums.gridded
    assert abs(vs.sums[0]["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
    psf32, dirty32 = g.new_grids(1), g.new_grids(1)
    vs.grid(gpu.MODE_DIRTY, psf=psf32, dirty=dirty32)
    assert rel(psf32[0].cpu().numpy(), opsf) < 2e-6
    assert rel(dirty32[0].cpu().numpy(), odirty) < 2e-6
    # PSF-only mode equals the PSF of the dirty mode
    psf_only = g.new_grids(1)
    vs.grid(gpu.MODE_PSF, psf=psf_only)
    assert rel(psf_only[0].cpu().numpy(), opsf) < 2e-6


def test_standard_residual_and_degrid(gpu, oracle_mod):
    g, o, (u, v, w, wt, sf, vis, gr) = _std_setup(gpu, oracle_mod, seed=9)
    rng = np.random.default_rng(4)
    model_img = np.zeros((g.nx, g.ny), np.float32)
    model_img[rng.integers(10, 90, 30), rng.integers(10, 90, 30)] = rng.uniform(0.1, 1.0, 30)
    omodel = o.model_grid(model_img)
    gmodel = g.model_to_grid(torch.from_numpy(model_img))
    assert rel(gmodel[0].cpu().numpy(), omodel) < 2e-6

    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    _, ores, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=2, model=omodel)
    vs = g.visibility_set(b)
    res32 = g.new_grids(1)
    vs.grid(gpu.MODE_RESIDUAL, model=gmodel, dirty=res32)
    assert rel(res32[0].cpu().numpy(), ores) < 1e-5

    m64 = gmodel.to(torch.complex128)
    res64 = g.new_grids(1, gpu.F64)
    g.grid_plain(b, gpu.MODE_RESIDUAL, gpu.F64, model=m64, dirty=res64)
    assert rel(res64[0].cpu().numpy(), ores) < 1e-5

    # degrid (forward operator)
    odeg = o.degrid(u, v, gmodel[0].cpu().numpy())
    gdeg = vs.degrid(gmodel).cpu().numpy()
    plan, _ = g.plan(b)
    acc = plan[:, 0].cpu().numpy() >= 0
    assert np.all(gdeg[~acc] == 0)
    assert rel(gdeg[acc], odeg[acc]) < 1e-5
    pdeg = g.degrid_plain(b, gmodel, gpu.F32).cpu().numpy()
    assert rel(pdeg[acc], odeg[acc]) < 1e-5


def test_standard_dirty_products(gpu, oracle_mod):
    g, o, (u, v, w, wt, sf, vis, gr) = _std_setup(gpu, oracle_mod, seed=13, sources=True)
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    opsf, odirty, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=1)
    ref_psf, ref_res, ref_peak = o.dirty_products(opsf, odirty, osums)

    vs = g.visibility_set(b)
    psf32, dirty32 = g.new_grids(1), g.new_grids(1)
    vs.grid(gpu.MODE_DIRTY, psf=psf32, dirty=dirty32)
    psf, res, peak = g.finish_products(psf32, dirty32, vs.sums[0]["normalization_sumwt"])
    # FP32 grids + FP32 FFT against FP64 grids + FP64 FFT: <= 1e-5 of peak
    assert abs(peak[0] - ref_peak) <= 1e-5 * ref_peak
    assert rel(psf[0].cpu().numpy(), ref_psf) < 1e-5
    assert rel(res[0].cpu().numpy(), ref_res) < 1e-5

    psf64, dirty64 = g.new_grids(1, gpu.F64), g.new_grids(1, gpu.F64)
    s64 = g.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=psf64, dirty=dirty64)[0]
    p64, r64, pk64 = g.finish_products(psf64, dirty64, s64["normalization_sumwt"], precision=gpu.F64)
    assert rel(p64[0].cpu().numpy(), ref_psf) < 1e-6
    assert rel(r64[0].cpu().numpy(), ref_res) < 1e-6


# --------------------------------------------------------------------------
# W-projection
# --------------------------------------------------------------------------

W_CASES = [
    # nx, cell arcsec, planes, sampling, max|w|
    (128, 600.0, 16, 4, 300.0),
    (128, 300.0, 16, 8, 300.0),
    (96, 30.0, 4, None, 2000.0),
]


def _w_setup(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw):
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(maxw, planes=planes, sampling=sampling)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    wp = oracle_mod.WProjector(o, maxw, explicit_planes=planes, sampling=sampling)
    return g, o, wp


@pytest.mark.parametrize("nx,arcsec,planes,sampling,maxw", W_CASES)
def test_w_kernels_match_oracle(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw):
    g, o, wp = _w_setup(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw)
    info = g.info
    assert info.planes == wp.planes and info.sampling == wp.sampling
    assert info.w_scale == wp.w_scale
    sup, ker, nt = g.wproject_tables()
    # a support may differ only where the deciding value lies within 1e-5 of
    # the 1e-3 threshold (FP32 against FP64 FFT rounding)
    margin_ok = np.asarray(wp.support_margin) > 1e-5
    assert np.array_equal(sup[margin_ok], wp.supports[margin_ok])
    # plane 0 fixes the common scale (its integral normalises every plane)
    assert sup[0] == wp.supports[0]
    same = np.nonzero(sup == wp.supports)[0]
    assert len(same) >= len(sup) - 1
    s = g.info.sampling
    R = (s + 1) // 2
    quarter = max(g.info.gx, g.info.gy) // 2 - 1
    ax = 2 * R + 1
    peak = np.max(np.abs(wp.kernels))
    nt_g = nt.reshape(len(sup), ax * ax)
    nt_o = wp.norm_table.reshape(len(sup), ax * ax)
    for k in same:
        # FP32 FFT (GPU) against FP64 FFT (oracle), relative to the kernel
        # peak, over the plane's cropped footprint and its zero padding
        cw = max(min(int(sup[k]) * s + R + 1, quarter), 1)
        w_common = min(ker.shape[1], wp.kernels.shape[1])
        assert np.max(np.abs(ker[k, :w_common, :w_common] - wp.kernels[k, :w_common, :w_common])) < 2e-5 * peak
        assert not np.any(ker[k, cw:, :]) and not np.any(ker[k, :, cw:])
        assert np.max(np.abs(nt_g[k] - nt_o[k])) < 2e-5 * np.max(np.abs(nt_o))


@pytest.mark.parametrize("nx,arcsec,planes,sampling,maxw", W_CASES)
def test_w_plan_bit_exact(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw):
    g, o, wp = _w_setup(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw)
    sup, ker, nt = g.wproject_tables()
    wpg = wp.with_kernels(ker, sup, nt)  # same tables: isolates the planner
    umax = 0.55 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(21, 20000, umax, wmax=1.2 * maxw, edge_cases=False)
    hu, hv = synth.half_cell_samples(o.du, o.dv, 30, g.info.sampling)
    u, v = np.concatenate([u, hu]), np.concatenate([v, hv])
    n = len(u)
    w = np.concatenate([w, np.linspace(-maxw, maxw, len(hu))])
    ones = np.ones(n, np.float32)
    b = batch_of(gpu, u, v, w, ones, ones, np.ones(n, np.complex64), np.ones(n, bool))
    plan, norm = g.plan(b)
    ref, rnorm = wpg.plan(u, v, w)
    assert np.array_equal(plan.cpu().numpy(), ref.astype(np.int32))
    acc = ref[:, 0] >= 0
    assert np.array_equal(norm.cpu().numpy()[acc], rnorm[acc])
    assert len(np.unique(ref[acc, 4])) > 1


@pytest.mark.parametrize("nx,arcsec,planes,sampling,maxw", W_CASES[:2])
def test_w_dirty_residual_degrid(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw):
    g, o, wp = _w_setup(gpu, oracle_mod, nx, arcsec, planes, sampling, maxw)
    sup, ker, nt = g.wproject_tables()
    wpg = wp.with_kernels(ker, sup, nt)
    umax = 0.5 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(31, 30000, umax, wmax=maxw)
    u[:10000] *= 0.1
    v[:10000] *= 0.1
    vis = with_sources(31, u, v, w, vis, nx, np.deg2rad(arcsec / 3600.0))
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    opsf, odirty, osums = wpg.grid(u, v, w, wt, sf, vis, gridable=gr, mode=1)

    vs = g.visibility_set(b)
    s = vs.sums[0]
    assert (s["gridded"], s["skipped"]) == (osums.gridded, osums.skipped)
    assert abs(s["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
    assert abs(s["reported_sumwt"] - osums.reported_sumwt) <= 1e-12 * osums.reported_sumwt
    psf32, dirty32 = g.new_grids(1), g.new_grids(1)
    vs.grid(gpu.MODE_DIRTY, psf=psf32, dirty=dirty32)
    assert rel(psf32[0].cpu().numpy(), opsf) < 5e-6
    assert rel(dirty32[0].cpu().numpy(), odirty) < 5e-6

    psf64, dirty64 = g.new_grids(1, gpu.F64), g.new_grids(1, gpu.F64)
    s64 = g.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=psf64, dirty=dirty64)[0]
    assert s64["gridded"] == osums.gridded
    assert rel(psf64[0].cpu().numpy(), opsf) < 5e-6
    assert rel(dirty64[0].cpu().numpy(), odirty) < 5e-6

    # images
    ref_psf, ref_res, ref_peak = wpg.dirty_products(opsf, odirty, osums)
    psf, res, peak = g.finish_products(psf32, dirty32, s["normalization_sumwt"])
    assert abs(peak[0] - ref_peak) <= 1e-5 * ref_peak
    assert rel(psf[0].cpu().numpy(), ref_psf) < 1e-5
    assert rel(res[0].cpu().numpy(), ref_res) < 1e-5

    # residual refresh against a model
    rng = np.random.default_rng(8)
    model_img = np.zeros((nx, nx), np.float32)
    model_img[rng.integers(20, nx - 20, 20), rng.integers(20, nx - 20, 20)] = rng.uniform(0.1, 1.0, 20)
    omodel = wpg.model_grid(model_img)
    gmodel = g.model_to_grid(torch.from_numpy(model_img))
    assert rel(gmodel[0].cpu().numpy(), omodel) < 2e-6
    _, ores, _ = wpg.grid(u, v, w, wt, sf, vis, gridable=gr, mode=2, model=omodel)
    res32 = g.new_grids(1)
    vs.grid(gpu.MODE_RESIDUAL, model=gmodel, dirty=res32)
    assert rel(res32[0].cpu().numpy(), ores) < 1e-5

    odeg = wpg.degrid(u, v, w, gmodel[0].cpu().numpy())
    gdeg = vs.degrid(gmodel).cpu().numpy()
    plan, _ = g.plan(b)
    acc = plan[:, 0].cpu().numpy() >= 0
    assert rel(gdeg[acc], odeg[acc]) < 1e-5
    pdeg = g.degrid_plain(b, gmodel.to(torch.complex128), gpu.F64).cpu().numpy()
    assert rel(pdeg[acc], odeg[acc]) < 1e-5


# --------------------------------------------------------------------------
# adjoint identities (size-independent properties)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("wproj", [False, True])
def test_degrid_grid_adjoint(gpu, oracle_mod, wproj):
    nx = 128
    cell = np.deg2rad((600.0 if wproj else 20.0) / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    if wproj:
        g.attach_wproject(300.0, planes=16)
    umax = 0.45 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(41, 20000, umax, wmax=300.0 if wproj else 0.0,
                                                    edge_cases=False)
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    vs = g.visibility_set(b)
    rng = np.random.default_rng(2)
    G = (rng.normal(size=(1, g.info.gx, g.info.gy)) + 1j * rng.normal(size=(1, g.info.gx, g.info.gy)))
    G = torch.from_numpy(G.astype(np.complex64)).cuda()
    y = (rng.normal(size=len(u)) + 1j * rng.normal(size=len(u))).astype(np.complex64)
    Ax = vs.degrid(G).cpu().numpy().astype(np.complex128)
    AHy = g.new_grids(1)
    vs.grid_values(torch.from_numpy(y).cuda(), AHy)
    lhs = np.vdot(Ax, y.astype(np.complex128))
    rhs = np.vdot(G.cpu().numpy().astype(np.complex128).ravel(), AHy.cpu().numpy().astype(np.complex128).ravel())
    assert abs(lhs - rhs) <= 1e-5 * abs(lhs)

    # model transform and its adjoint: Re<F x, H> = <x, Re(F^H H)>
    x = rng.normal(size=(nx, nx)).astype(np.float32)
    Fx = g.model_to_grid(torch.from_numpy(x)).cpu().numpy().astype(np.complex128)
    H = G.clone()
    FhH = g.grid_to_model_adjoint(H)[0].cpu().numpy().astype(np.float64)
    lhs2 = np.vdot(Fx.ravel(), G.cpu().numpy().astype(np.complex128).ravel()).real
    rhs2 = float(np.sum(x.astype(np.float64) * FhH))
    assert abs(lhs2 - rhs2) <= 1e-5 * abs(lhs2)


# --------------------------------------------------------------------------
# rows x channels (batched cube)
# --------------------------------------------------------------------------

def test_cube_rows_by_channels(gpu, oracle_mod):
    enu = synth.y_array(3, arm_m=2000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 0.5, 60.0)
    n_rows, n_chan = uvw.shape[0], 6
    freq = np.linspace(1.0e9, 2.0e9, n_chan)
    scale = freq / synth.C
    rng = np.random.default_rng(12)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, (n_rows, n_chan)).astype(np.float32)
    flag = rng.uniform(size=(n_rows, n_chan)) < rng.uniform(0.0, 0.6, n_chan)[None, :]
    maxuv = np.max(np.abs(uvw[:, :2])) * scale[-1]
    nx = 96
    cell = 1.0 / (3.0 * maxuv)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(np.max(np.abs(uvw[:, 2])) * scale[-1], planes=8)
    batch = gpu.RowChannelBatch(t(uvw), t(scale), t(vis), t(wt), t(flag.astype(np.uint8)))
    vs = g.visibility_set(batch)
    psf, dirty = g.new_grids(n_chan), g.new_grids(n_chan)
    vs.grid(gpu.MODE_DIRTY, psf=psf, dirty=dirty)
    psf64, dirty64 = g.new_grids(n_chan, gpu.F64), g.new_grids(n_chan, gpu.F64)
    sums64 = g.grid_plain(batch, gpu.MODE_DIRTY, gpu.F64, psf=psf64, dirty=dirty64)

    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    sup, ker, nt = g.wproject_tables()
    wp = oracle_mod.WProjector(o, np.max(np.abs(uvw[:, 2])) * scale[-1], explicit_planes=8).with_kernels(ker, sup, nt)
    for c in range(n_chan):
        u, v, w = (uvw[:, k] * scale[c] for k in range(3))
        opsf, odirty, osums = wp.grid(u, v, w, wt[:, c], np.ones(n_rows, np.float32), vis[:, c],
                                      gridable=~flag[:, c], mode=1)
        assert vs.sums[c]["gridded"] == osums.gridded == sums64[c]["gridded"]
        assert abs(vs.sums[c]["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
        assert rel(dirty[c].cpu().numpy(), odirty) < 5e-6
        assert rel(psf[c].cpu().numpy(), opsf) < 5e-6
        assert rel(dirty64[c].cpu().numpy(), odirty) < 5e-6


# --------------------------------------------------------------------------
# fused major cycle (model -> residual image in one call)
# --------------------------------------------------------------------------

def _point_models(seed, n_chan, nx, n=20):
    rng = np.random.default_rng(seed)
    m = np.zeros((n_chan, nx, nx), np.float32)
    for c in range(n_chan):
        m[c, rng.integers(15, nx - 15, n), rng.integers(15, nx - 15, n)] = rng.uniform(0.1, 1.0, n)
    return m


@pytest.mark.parametrize("strategy", [0, 1])
def test_major_cycle_standard(gpu, oracle_mod, strategy):
    """Standard path: compute_residual_standard_tiled (lib.rs:68862-68932)
    against the oracle's FP64 residual grid and FP64 FFT.

    Bar: 2e-5 of the residual peak.  The path stores FP32 grids and runs FP32
    transforms; the largest errors sit at the image corners, where the grid
    correction 1/psi multiplies the transform's rounding by up to ~25 per
    axis (padding 1.2), and FP32 storage of the residual grid alone accounts
    for ~4e-6 there.  The reference validates its own f32 FFT backends at
    3e-4 (fft_backend.rs:727-731); the 1e-5 dirty-image bar of the FP32
    batched path is measured on the full-size cube by the benchmark."""
    g, o, (u, v, w, wt, sf, vis, gr) = _std_setup(gpu, oracle_mod, seed=21, sources=True)
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    opsf, _, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=0)
    vs = g.visibility_set(b)
    vs.set_cycle_strategy(strategy)
    psf32 = g.new_grids(1)
    vs.grid(gpu.MODE_PSF, psf=psf32)
    _, _, peak = g.finish_products(psf32, None, vs.sums[0]["normalization_sumwt"])
    nsum = vs.sums[0]["normalization_sumwt"]
    for seed in (3, 4):  # the second call reuses the workspace
        model = _point_models(seed, 1, g.nx)
        _, ores, _ = o.grid(u, v, wt, sf, vis, gridable=gr, mode=2, model=o.model_grid(model[0]))
        _, ref_res, ref_peak = o.dirty_products(opsf, ores, osums)
        res = vs.major_cycle(torch.from_numpy(model), nsum, peak).cpu().numpy()
        print(f"strategy {strategy} model {seed}: {rel(res[0], ref_res):.2e} of the residual peak")
        assert abs(peak[0] - ref_peak) <= 1e-5 * ref_peak
        assert rel(res[0], ref_res) < 2e-5
    assert vs.cycle_info()["strategy"] == strategy


@pytest.mark.parametrize("strategy", [-1, 0, 1])
def test_major_cycle_w_cube(gpu, oracle_mod, strategy):
    """W-projection cube, rows x channels: compute_residual_w_project_replay
    (lib.rs:70895-70963) per channel against the oracle (FP32 grids as the
    reference, FP64 FFT); 2e-5 of the residual peak, for the reasons given in
    test_major_cycle_standard (here the correction also divides by the sinc
    of the W kernel sampling).  Also equal to the unfused path model_to_grid
    -> grid(RESIDUAL) -> finish_products to 2e-5 of the peak."""
    enu = synth.y_array(3, arm_m=2000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 0.5, 60.0)
    n_rows, n_chan = uvw.shape[0], 5
    freq = np.linspace(1.0e9, 1.6e9, n_chan)
    scale = freq / synth.C
    rng = np.random.default_rng(17)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, (n_rows, n_chan)).astype(np.float32)
    flag = rng.uniform(size=(n_rows, n_chan)) < 0.2
    maxuv = np.max(np.abs(uvw[:, :2])) * scale[-1]
    nx = 96
    cell = 1.0 / (3.0 * maxuv)
    maxw = np.max(np.abs(uvw[:, 2])) * scale[-1]
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(maxw, planes=8)
    batch = gpu.RowChannelBatch(t(uvw), t(scale), t(vis), t(wt), t(flag.astype(np.uint8)))
    vs = g.visibility_set(batch)
    vs.set_cycle_strategy(strategy)
    psf = g.new_grids(n_chan)
    vs.grid(gpu.MODE_PSF, psf=psf)
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])
    _, _, peaks = g.finish_products(psf.clone(), None, nsum)

    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    sup, ker, nt = g.wproject_tables()
    wp = oracle_mod.WProjector(o, maxw, explicit_planes=8).with_kernels(ker, sup, nt)
    for seed in (5, 6):
        model = _point_models(seed, n_chan, nx)
        res = vs.major_cycle(torch.from_numpy(model), nsum, peaks).cpu().numpy()
        gm = g.model_to_grid(torch.from_numpy(model))
        r32 = g.new_grids(n_chan)
        vs.grid(gpu.MODE_RESIDUAL, model=gm, dirty=r32)
        _, unfused, _ = g.finish_products(None, r32, nsum, psf_peak=peaks)
        unfused = unfused.cpu().numpy()
        for c in range(n_chan):
            u, v, w = (uvw[:, k] * scale[c] for k in range(3))
            opsf, ores, osums = wp.grid(u, v, w, wt[:, c], np.ones(n_rows, np.float32), vis[:, c],
                                        gridable=~flag[:, c], mode=2, model=wp.model_grid(model[c]))
            opsf, _, _ = wp.grid(u, v, w, wt[:, c], np.ones(n_rows, np.float32), vis[:, c],
                                 gridable=~flag[:, c], mode=0)
            _, ref_res, ref_peak = wp.dirty_products(opsf, ores, osums)
            print(f"channel {c}: major cycle {rel(res[c], ref_res):.2e}, unfused {rel(unfused[c], ref_res):.2e} of peak")
            assert abs(peaks[c] - ref_peak) <= 1e-5 * ref_peak
            assert rel(res[c], ref_res) < 2e-5
        assert rel(res, unfused) < 2e-5
    info = vs.cycle_info()
    assert info["strategy"] == (gpu.CYCLE_FUSED if strategy < 0 else strategy)
    assert (info["columns"] > 0) == (info["strategy"] == gpu.CYCLE_FUSED)


def test_major_cycle_rejects_bad_normalization(gpu, oracle_mod):
    g, o, (u, v, w, wt, sf, vis, gr) = _std_setup(gpu, oracle_mod, seed=22)
    vs = g.visibility_set(batch_of(gpu, u, v, w, wt, sf, vis, gr))
    model = torch.zeros((1, g.nx, g.ny))
    with pytest.raises(gpu.NoUsableSamples):
        vs.major_cycle(model, 0.0, 1.0)
    with pytest.raises(gpu.Normalization):
        vs.major_cycle(model, 1.0, float("nan"))
    with pytest.raises(gpu.InvalidRequest):
        vs.major_cycle(torch.zeros((2, g.nx, g.ny)), 1.0, 1.0)


@pytest.mark.parametrize("nx,ny", [(64, 64), (96, 64), (200, 150), (520, 512), (1400, 1400), (3413, 1707),
                                   (2400, 3413)])
def test_major_cycle_fused_shapes(gpu, oracle_mod, nx, ny):
    """Fused transforms on square and rectangular grids whose lengths use
    every radix (2, 3, 4, 5): the residual image against the oracle (FP64
    grids and FFT) to 1e-5 of peak, and against the hipFFT strategy.  The
    three large shapes have axes above 1536 points (1728 x 1728, 4096 x 2048
    and 2880 x 4096 grids): 128 threads per transform and half-block column
    workgroups."""
    cell = np.deg2rad(5.0 / 3600.0)
    g = gpu.StandardGridder(nx, ny, cell, cell)
    o = oracle_mod.StandardGridder(nx, ny, cell, cell)
    umax = 0.5 * min(g.info.gx * g.info.du, g.info.gy * g.info.dv)
    u, v, w, wt, sf, vis, gr = synth.random_samples(40 + nx, 20000, umax)
    u[:8000] *= 0.15
    v[:8000] *= 0.15
    vis = with_sources(41, u, v, w, vis, min(nx, ny), cell)
    b = batch_of(gpu, u, v, w, wt, sf, vis, gr)
    vs = g.visibility_set(b)
    nsum = vs.sums[0]["normalization_sumwt"]
    opsf, _, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=0)
    rng = np.random.default_rng(nx)
    model = np.zeros((1, nx, ny), np.float32)
    model[0, rng.integers(5, nx - 5, 12), rng.integers(5, ny - 5, 12)] = rng.uniform(0.1, 1.0, 12)
    _, ores, _ = o.grid(u, v, wt, sf, vis, gridable=gr, mode=2, model=o.model_grid(model[0]))
    _, ref_res, ref_peak = o.dirty_products(opsf, ores, osums)
    out = {}
    for strategy in (gpu.CYCLE_FUSED, gpu.CYCLE_HIPFFT):
        vs.set_cycle_strategy(strategy)
        out[strategy] = vs.major_cycle(torch.from_numpy(model), nsum, ref_peak).cpu().numpy()
        assert vs.cycle_info()["strategy"] == strategy
        print(f"strategy {strategy}: {rel(out[strategy][0], ref_res):.2e} of peak")
        assert rel(out[strategy][0], ref_res) < 1e-5
    assert rel(out[gpu.CYCLE_FUSED], out[gpu.CYCLE_HIPFFT]) < 1e-5
