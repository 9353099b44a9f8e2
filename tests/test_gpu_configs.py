"""GPU parity at the sizes of the benchmark configurations.

* Plain spheroidal gridding at full size (27 antennas in a 1 km disk, 2 h at
  10 s = 252,720 rows, 1.4 GHz, 512^2 image, cell 1/(3 max|uv|), 7x7
  support): plans bit-exact, sums of weights and counts exact, dirty image and
  PSF against the oracle (FP64 grids, FP64 FFT), at the configuration's
  oversampling 128 and the reference's hard-coded 100 (gridder.rs:407).
* W-projection with large supports (up to 32 and 38 cells, oversampling 8):
  the binned kernel switches to smaller uv tiles and the full LDS; grids,
  degridded visibilities and the major cycle against the oracle.
* Differentiable imaging at size (1024^2 sky of Gaussians and a disk, 2M
  visibilities, 32 W planes): dot-product adjoint, central finite differences
  (eps 1e-3) on 64 seeded pixels of chi^2 + 1e-3 TV, the gradient against the
  direct DFT on a 10k-visibility subset, and 200 Adam steps, with the W
  kernels recomputed on the fly (no table kept between calls).
"""
import importlib

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


def batch(pkg, u, v, w, wt, sf, vis, gridable):
    return pkg.VisibilityBatch(t(u), t(v), t(w), t(wt), t(vis), t(sf), t(gridable.astype(np.uint8)))


# --------------------------------------------------------------------------
# plain spheroidal gridding, full size
# --------------------------------------------------------------------------

def config1_data(seed=101):
    enu = synth.disk_array(seed, 27, 1000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 2.0, 10.0)
    scale = 1.4e9 / synth.C
    u, v, w = (np.ascontiguousarray(uvw[:, k] * scale) for k in range(3))
    cell = 1.0 / (3.0 * float(np.max(np.hypot(u, v))))
    nx = 512
    rng = np.random.default_rng(seed)
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), 20))
    half = 0.5 * nx * cell
    sources = [(f, rng.uniform(-half, half), rng.uniform(-half, half)) for f in flux]
    vis = synth.point_source_vis((u, v, w), sources)
    noise = 0.05 / np.sqrt(2.0) * (rng.normal(size=len(u)) + 1j * rng.normal(size=len(u)))
    vis = (vis + noise).astype(np.complex64)
    return u, v, w, np.ones(len(u), np.float32), vis, nx, cell


@pytest.mark.parametrize("os_", [128, 100])
def test_config1_standard_full_size(gpu, oracle_mod, os_):
    u, v, w, wt, vis, nx, cell = config1_data()
    n = len(u)
    assert n == 252_720
    sf, gr = np.ones(n, np.float32), np.ones(n, bool)
    g = gpu.StandardGridder(nx, nx, cell, cell, oversampling=os_)
    o = oracle_mod.StandardGridder(nx, nx, cell, cell, oversampling=os_)
    assert g.grid_shape == o.grid_shape == (640, 640)  # 512 -> 640 (gridder.rs:5577-5584)
    b = batch(gpu, u, v, w, wt, sf, vis, gr)

    plan, _ = g.plan(b)
    assert np.array_equal(plan.cpu().numpy()[:, :4], o.plan(u, v).astype(np.int32))

    opsf, odirty, osums = o.grid(u, v, wt, sf, vis, gridable=gr, mode=1)
    ref_psf, ref_dirty, ref_peak = o.dirty_products(opsf, odirty, osums)

    vs = g.visibility_set(b)
    s = vs.sums[0]
    assert (s["gridded"], s["skipped"]) == (osums.gridded, osums.skipped)
    assert abs(s["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
    psf32, dirty32 = g.new_grids(1), g.new_grids(1)
    vs.grid(gpu.MODE_DIRTY, psf=psf32, dirty=dirty32)
    assert rel(psf32[0].cpu().numpy(), opsf) < 2e-6
    assert rel(dirty32[0].cpu().numpy(), odirty) < 2e-6
    psf, dirty, peak = g.finish_products(psf32, dirty32, s["normalization_sumwt"])
    # batched FP32 against FP64: the metric's 1e-5 of peak
    assert abs(peak[0] - ref_peak) <= 1e-5 * ref_peak
    assert rel(psf[0].cpu().numpy(), ref_psf) < 1e-5
    assert rel(dirty[0].cpu().numpy(), ref_dirty) < 1e-5

    psf64, dirty64 = g.new_grids(1, gpu.F64), g.new_grids(1, gpu.F64)
    s64 = g.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=psf64, dirty=dirty64)[0]
    assert abs(s64["normalization_sumwt"] - osums.normalization_sumwt) <= 1e-12 * osums.normalization_sumwt
    assert rel(dirty64[0].cpu().numpy(), odirty) < 1e-12
    p64, d64, _ = g.finish_products(psf64, dirty64, s64["normalization_sumwt"], precision=gpu.F64)
    assert rel(p64[0].cpu().numpy(), ref_psf) < 1e-6
    assert rel(d64[0].cpu().numpy(), ref_dirty) < 1e-6


# --------------------------------------------------------------------------
# W-projection with large supports
# --------------------------------------------------------------------------

LARGE_W = [
    # nx, cell arcsec, planes, sampling -> max support (from the oracle's kernels)
    (448, 240.0, 64, 8),   # supports 4..32: 32x32 tiles, (32 + 64)^2 LDS tile above 64 KB
    (512, 180.0, 64, 8),   # supports 4..38: 16x16 tiles
]


@pytest.mark.parametrize("nx,arcsec,planes,sampling", LARGE_W)
def test_w_large_supports(gpu, oracle_mod, nx, arcsec, planes, sampling):
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(1.0, planes=planes, sampling=sampling)
    sup, ker, nt = g.wproject_tables()
    assert sup.max() >= 30
    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    wp = oracle_mod.WProjector(o, 1.0, explicit_planes=planes, sampling=sampling).with_kernels(ker, sup, nt)
    max_w = (planes - 1) ** 2 / g.info.w_scale  # every plane is used
    umax = 0.5 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(nx, 30000, umax, wmax=1.2 * max_w)
    u[:10000] *= 0.2
    v[:10000] *= 0.2
    b = batch(gpu, u, v, w, wt, sf, vis, gr)

    plan, _ = g.plan(b)
    oplan, _ = wp.plan(u, v, w)
    # chip_plan also applies the validity filters (raw_w_project_sample_is_valid,
    # lib.rs:73666-73676); the oracle's planner does not
    valid = (gr & np.isfinite(wt) & (wt > 0) & np.isfinite(sf) & (sf > 0) & np.isfinite(vis.real)
             & np.isfinite(vis.imag) & np.isfinite(u) & np.isfinite(v) & np.isfinite(w))
    expect = np.where(valid[:, None], oplan, -1).astype(np.int32)
    assert np.array_equal(plan.cpu().numpy(), expect)
    assert len(np.unique(oplan[(oplan[:, 0] >= 0) & valid, 4])) >= planes // 2

    opsf, odirty, osums = wp.grid(u, v, w, wt, sf, vis, gridable=gr, mode=1)
    vs = g.visibility_set(b)
    s = vs.sums[0]
    assert (s["gridded"], s["skipped"]) == (osums.gridded, osums.skipped)
    psf32, dirty32 = g.new_grids(1), g.new_grids(1)
    vs.grid(gpu.MODE_DIRTY, psf=psf32, dirty=dirty32)
    assert rel(psf32[0].cpu().numpy(), opsf) < 5e-6
    assert rel(dirty32[0].cpu().numpy(), odirty) < 5e-6

    rng = np.random.default_rng(nx)
    model = np.zeros((nx, nx), np.float32)
    model[rng.integers(40, nx - 40, 25), rng.integers(40, nx - 40, 25)] = rng.uniform(0.1, 1.0, 25)
    gm = g.model_to_grid(t(model))
    acc = plan[:, 0].cpu().numpy() >= 0
    odeg = wp.degrid(u, v, w, gm[0].cpu().numpy())
    assert rel(vs.degrid(gm).cpu().numpy()[acc], odeg[acc]) < 1e-5
    _, ores, _ = wp.grid(u, v, w, wt, sf, vis, gridable=gr, mode=2, model=wp.model_grid(model))
    r32 = g.new_grids(1)
    vs.grid(gpu.MODE_RESIDUAL, model=gm, dirty=r32)
    assert rel(r32[0].cpu().numpy(), ores) < 1e-5

    # major cycle with the large supports (fused transforms)
    _, _, peak = g.finish_products(psf32.clone(), None, s["normalization_sumwt"])
    _, ref_res, _ = wp.dirty_products(opsf, ores, osums)
    res = vs.major_cycle(t(model[None]), s["normalization_sumwt"], peak).cpu().numpy()
    assert rel(res[0], ref_res) < 2e-5


def test_two_phase_large_supports(gpu, oracle_mod):
    """Residual refreshes of half-supports 16 - 28 (supports 4..28: both
    large classes within the two-phase kernel's range) against the oracle,
    and the two-phase kernel against the chunked LDS-atomic kernel on the
    same set (grids and major cycle within FP32 accumulation order)."""
    nx, arcsec, planes, sampling = 384, 240.0, 64, 8
    cell = np.deg2rad(arcsec / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(1.0, planes=planes, sampling=sampling)
    sup, ker, nt = g.wproject_tables()
    assert 23 <= sup.max() <= 31 and (sup >= 16).sum() >= 8
    o = oracle_mod.StandardGridder(nx, nx, cell, cell)
    wp = oracle_mod.WProjector(o, 1.0, explicit_planes=planes, sampling=sampling).with_kernels(ker, sup, nt)
    max_w = (planes - 1) ** 2 / g.info.w_scale
    umax = 0.5 * g.info.gx * g.info.du
    u, v, w, wt, sf, vis, gr = synth.random_samples(nx + 1, 30000, umax, wmax=1.2 * max_w)
    u[:10000] *= 0.2
    v[:10000] *= 0.2
    b = batch(gpu, u, v, w, wt, sf, vis, gr)
    plan, _ = g.plan(b)
    used = plan[:, 4].cpu().numpy()[plan[:, 0].cpu().numpy() >= 0]
    assert (sup[used] >= 23).sum() > 1000 and ((sup[used] >= 16) & (sup[used] < 23)).sum() > 1000

    rng = np.random.default_rng(nx)
    model = np.zeros((nx, nx), np.float32)
    model[rng.integers(40, nx - 40, 25), rng.integers(40, nx - 40, 25)] = rng.uniform(0.1, 1.0, 25)
    gm = g.model_to_grid(t(model))
    opsf, _, osums = wp.grid(u, v, w, wt, sf, vis, gridable=gr, mode=1)
    _, ores, _ = wp.grid(u, v, w, wt, sf, vis, gridable=gr, mode=2, model=wp.model_grid(model))
    vs = g.visibility_set(b)
    s = vs.sums[0]
    res = {}
    for two_phase in (True, False):
        vs.set_two_phase(two_phase)
        r32 = g.new_grids(1)
        vs.grid(gpu.MODE_RESIDUAL, model=gm, dirty=r32)
        res[two_phase] = r32[0].cpu().numpy()
        assert rel(res[two_phase], ores) < 1e-5, two_phase
    assert rel(res[True], res[False]) < 1e-5

    psf32 = g.new_grids(1)
    vs.grid(gpu.MODE_PSF, psf=psf32)
    _, _, peak = g.finish_products(psf32, None, s["normalization_sumwt"])
    _, ref_res, _ = wp.dirty_products(opsf, ores, osums)
    cyc = {}
    for two_phase in (True, False):
        vs.set_two_phase(two_phase)
        cyc[two_phase] = vs.major_cycle(t(model[None]), s["normalization_sumwt"], peak).cpu().numpy()[0]
        assert rel(cyc[two_phase], ref_res) < 2e-5, two_phase
    assert rel(cyc[True], cyc[False]) < 1e-5


# --------------------------------------------------------------------------
# differentiable imaging at size
# --------------------------------------------------------------------------

def config5_sky(nx, rng):
    yy, xx = np.mgrid[0:nx, 0:nx].astype(np.float64)
    sky = np.zeros((nx, nx))
    for _ in range(50):
        x0, y0 = rng.uniform(100, nx - 100, 2)
        fwhm = rng.uniform(1.0, 10.0)
        sig = fwhm / 2.3548
        sky += np.exp(rng.uniform(np.log(0.01), np.log(1.0))) * np.exp(
            -((xx - x0) ** 2 + (yy - y0) ** 2) / (2 * sig * sig))
    cx, cy = rng.uniform(300, nx - 300, 2)
    sky[(xx - cx) ** 2 + (yy - cy) ** 2 <= 50.0 ** 2] += 0.01  # 100 px uniform disk
    return sky.astype(np.float32)


def tv(img, delta=1e-6):
    dx = img[1:, :-1] - img[:-1, :-1]
    dy = img[:-1, 1:] - img[:-1, :-1]
    return torch.sum(torch.sqrt(dx * dx + dy * dy + delta))


def test_config5_differentiable_at_size(gpu, oracle_mod):
    ag = importlib.import_module("casa_rs_amd.autograd")
    rng = np.random.default_rng(505)
    nx = 1024
    enu = synth.y_array(5, arm_m=6000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 8.0, 5.0)[:2_000_000]
    scale = 1.4e9 / synth.C
    u, v, w = (np.ascontiguousarray(uvw[:, k] * scale) for k in range(3))
    n = len(u)
    assert n == 2_000_000
    cell = 1.0 / (3.0 * float(np.max(np.hypot(u, v))))
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_wproject(float(np.max(np.abs(w))), planes=32)
    # W kernels recomputed on the fly: no table is kept between calls, every
    # degrid / grid recomputes them one plane at a time
    assert g.set_wkernel_budget(1) == 1
    ones = torch.ones(n, dtype=torch.float32)
    b = gpu.VisibilityBatch(t(u), t(v), t(w), ones, torch.zeros(n, dtype=torch.complex64), ones)
    vs = g.visibility_set(b)
    op = ag.DegridOperator(g, vs)
    dev = torch.device("cuda")

    sky = torch.from_numpy(config5_sky(nx, rng)).to(dev)
    data = op.forward_linear(sky)
    data = data + (0.01 / np.sqrt(2.0)) * torch.randn_like(data)  # sigma = 10 mJy
    wts = torch.full((n,), 1.0 / 0.01 ** 2, dtype=torch.float32, device=dev)

    # dot-product adjoint test.  The inner products sum 2M terms of random sign
    # (they are far smaller than |A x| |y|), so the FP32 rounding of each
    # model visibility (~1e-6 relative) adds up to ~1e-6 |A x| |y| / sqrt(n);
    # the bar is ten times that
    x = torch.randn(nx, nx, device=dev)
    y = torch.randn(n, dtype=torch.complex64, device=dev)
    ax = op.forward_linear(x).to(torch.complex128)
    lhs = torch.sum(ax.conj() * y.to(torch.complex128)).real.item()
    rhs = torch.sum(x.double() * op.adjoint_linear(y).double()).item()
    scale = torch.linalg.vector_norm(ax).item() * torch.linalg.vector_norm(y.to(torch.complex128)).item()
    assert abs(lhs - rhs) <= 1e-5 * scale / np.sqrt(n)

    def loss(img):
        r = op(img) - data
        return torch.sum(wts * (r.real ** 2 + r.imag ** 2)) + 1e-3 * tv(img)

    def loss_fp64(img):
        # the same loss through the plain FP64 path (FP64 model transform and
        # degridding MACs), so that the finite differences of a loss of ~1e10
        # are not swamped by the FP32 rounding of 2M model visibilities
        pred = g.degrid_plain(b, g.model_to_grid(img, gpu.F64), gpu.F64)  # complex128
        r = pred - data.to(torch.complex128)
        return (torch.sum(wts.double() * (r.real ** 2 + r.imag ** 2)) + 1e-3 * tv(img.double())).item()

    # central finite differences on 64 seeded pixels against the FP32 autograd gradient
    x0 = (sky + 1e-3 * torch.randn_like(sky)).contiguous()
    xg = x0.clone().requires_grad_(True)
    loss(xg).backward()
    grad = xg.grad.double().cpu().numpy()
    gmax = np.max(np.abs(grad))
    eps = 1e-3
    pix = list(zip(rng.integers(200, nx - 200, 64), rng.integers(200, nx - 200, 64)))
    bad = 0
    for (a, c) in pix:
        xp, xm = x0.clone(), x0.clone()
        xp[a, c] += eps
        xm[a, c] -= eps
        fd = (loss_fp64(xp) - loss_fp64(xm)) / (2 * eps)
        if abs(fd - grad[a, c]) > 2e-2 * max(abs(grad[a, c]), 1e-3 * gmax):
            bad += 1
            print(f"pixel ({a}, {c}): finite difference {fd:.6e}, gradient {grad[a, c]:.6e}, max |grad| {gmax:.3e}")
    assert bad == 0

    # gradient of the chi^2 term against the direct DFT on a 10k subset
    sub = rng.choice(n, 10_000, replace=False)
    us, vsub, ws = u[sub], v[sub], w[sub]
    bs = gpu.VisibilityBatch(t(us), t(vsub), t(ws), torch.ones(len(sub)), torch.zeros(len(sub), dtype=torch.complex64),
                             torch.ones(len(sub)))
    op_s = ag.DegridOperator(g, g.visibility_set(bs))
    r = (rng.normal(size=len(sub)) + 1j * rng.normal(size=len(sub))).astype(np.complex64)
    got = 2.0 * op_s.adjoint_linear(t(r).to(dev)).double().cpu().numpy()
    x0w, y0w, win = nx // 2 - 32, nx // 2 - 32, 64
    ref = oracle_mod.direct_adjoint_window(nx, nx, cell, cell, x0w, y0w, win, win, us, vsub, ws, r)
    err = np.max(np.abs(got[x0w:x0w + win, y0w:y0w + win] - ref)) / np.max(np.abs(ref))
    assert err < 3e-2  # the reference's own direct-vs-gridded bound (lib.rs:89880-89887)

    # 200 Adam steps from an empty image reduce the loss
    img = torch.zeros(nx, nx, device=dev, requires_grad=True)
    opt = torch.optim.Adam([img], lr=1e-2)
    first = None
    for step in range(200):
        opt.zero_grad()
        L = loss(img)
        L.backward()
        opt.step()
        if first is None:
            first = L.item()
    assert np.isfinite(L.item()) and L.item() < 0.5 * first
