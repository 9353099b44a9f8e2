"""AWProject CF-atlas gridding on the GPU against the oracle (oracle/casa_oracle.c,
co_aw_*), and the Airy A-term CF builder against a numpy restatement.

Bars:
  plans (anchors, offsets, conjugation, rejection codes) and both
  normalisations are bit-exact: integer work and sequential f32 / f64 sums in
  the reference order;
  phased taps and the plain FP32 degrid (Complex32 sum in tap order, num-complex
  division) are bit-exact: the same f32 / f64 operations in the same order;
  plain FP64 grids equal the oracle's to 1e-12 relative (Complex32
  contributions added in another order in f64);
  binned FP32 grids and degrids are within 1e-5 of the peak (FP32 accumulation
  in another order);
  Airy CF cells within 2e-5 of the plane peak against an FP64 numpy FFT of the
  same screens (the device FFT is FP32), with identical supports.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def t(x):
    return torch.from_numpy(np.ascontiguousarray(x))


def _lit_cell(shape, fr, fi):
    x, y = np.meshgrid(np.arange(shape[0], dtype=np.float32), np.arange(shape[1], dtype=np.float32), indexing="ij")
    return (fr(x) + 1j * fi(y)).astype(np.complex64)


def _batch(gpu, u, v, w, wt, vis, cell=None, pointing=None, gridable=None):
    return gpu.VisibilityBatch(t(np.asarray(u, np.float64)), t(np.asarray(v, np.float64)), t(np.asarray(w, np.float64)),
                               t(np.asarray(wt, np.float32)), t(np.asarray(vis, np.complex64)), None,
                               None if gridable is None else t(gridable.astype(np.uint8)),
                               None if cell is None else t(np.asarray(cell, np.int32)),
                               None if pointing is None else t(np.asarray(pointing, np.int32)))


def test_reference_literal_plan_and_tap(gpu, oracle_mod):
    """gridder.rs:4343-4534 through the C ABI."""
    o = oracle_mod
    k = _lit_cell((12, 12), lambda x: np.float32(0.01) * x + np.float32(0.2), lambda y: np.float32(-0.02) * y + np.float32(0.1))
    grad = (0.07, -0.11)
    g = gpu.StandardGridder(32, 32, 1.0e-4, 1.0e-4, padding_mode=gpu.abi.PAD_NONE)
    g.attach_aw([(k, 2, 2)], sampling=2, phase_gradients=[grad])
    assert g.info.aw == 1 and g.info.wproject == 1 and g.info.planes == 2
    du = g.info.du
    b = _batch(gpu, [0.3 * du, 0.3 * du, 0.3 * du], [-0.3 * du] * 3, [12.0, -12.0, 0.0], [1.0] * 3, [1.25 - 0.5j] * 3)
    plan, n32, n64 = g.aw_plan(b)
    plan = plan.cpu().numpy()
    assert plan[0].tolist() == [16, 16, -1, -1, 1, 0]
    assert plan[1].tolist() == [16, 16, -1, -1, 0, 0] and plan[2].tolist() == [16, 16, -1, -1, 0, 0]
    atlas = o.AwAtlas([(k, 2, 2)], 2)
    for i, w in enumerate([12.0, -12.0, 0.0]):
        rc, geo, r32, r64 = o.aw_plan(32, 32, du, du, atlas, 0, 0.3 * du, -0.3 * du, w)
        assert rc == 0
        assert n32[i].item() == complex(r32) and n64[i].item() == r64
    # the single gridded sample equals the oracle's f64 grid exactly
    psf = g.new_grids(1, gpu.F64)
    dirty = g.new_grids(1, gpu.F64)
    one = _batch(gpu, [0.3 * du], [-0.3 * du], [12.0], [1.0], [1.25 - 0.5j])
    sums = g.grid_plain(one, gpu.MODE_DIRTY, gpu.F64, psf=psf, dirty=dirty)
    ref, st = o.aw_grid_f64(32, 32, du, du, atlas, [grad], [0.3 * du], [-0.3 * du], [12.0], [0],
                            np.array([1.25 - 0.5j], np.complex64))
    assert st[0] == 0
    assert np.array_equal(dirty[0].cpu().numpy(), ref)
    refp, _ = o.aw_grid_f64(32, 32, du, du, atlas, [grad], [0.3 * du], [-0.3 * du], [12.0], [0],
                            np.array([1.0], np.complex64))
    assert np.array_equal(psf[0].cpu().numpy(), refp)
    assert sums[0]["gridded"] == 1
    assert sums[0]["normalization_sumwt"] == abs(n64[0].item())


def test_attach_validation(gpu):
    g = gpu.StandardGridder(32, 32, 1.0e-4, 1.0e-4, padding_mode=gpu.abi.PAD_NONE)
    big = np.ones((12, 12), np.complex64)
    with pytest.raises(gpu.InvalidRequest):
        g.attach_aw([(big, 3, 2)], sampling=2)  # origin 6 < 3 * 2 fails the upper side: 6 + 6 >= 12
    g.attach_aw([(big, 2, 2)], sampling=2)
    with pytest.raises(gpu.InvalidRequest):
        g.attach_aw([(big, 2, 2)], sampling=2)
    with pytest.raises(gpu.InvalidRequest):
        g.wproject_tables()


def _random_atlas(seed, sampling):
    rng = np.random.default_rng(seed)
    cells = []
    for xs, ys in [(3, 3), (5, 4), (2, 6), (7, 7)]:
        sx = 2 * (xs * sampling + sampling) + 1 + rng.integers(0, 2)  # odd and even shapes
        sy = 2 * (ys * sampling + sampling) + rng.integers(0, 2)
        x, y = np.meshgrid(np.arange(sx) - sx / 2, np.arange(sy) - sy / 2, indexing="ij")
        env = np.exp(-(x * x + y * y) / (2.0 * (1.5 * sampling * max(xs, ys)) ** 2))
        pix = (env * (1.0 + 0.2 * rng.normal(size=env.shape)) + 0.1j * env * rng.normal(size=env.shape))
        cells.append((pix.astype(np.complex64), xs, ys))
    return cells


def _stream(seed, n, umax, cells_n, pointings_n):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, n)
    v = rng.uniform(-umax, umax, n)
    w = rng.uniform(-300.0, 300.0, n)
    w[::17] = 0.0
    u[::97] = np.nan
    u[5::101] = 10 * umax
    vis = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    wt = rng.uniform(0.2, 2.0, n).astype(np.float32)
    cell = rng.integers(0, cells_n, n)
    pointing = rng.integers(0, pointings_n, n)
    return u, v, w, vis, wt, cell, pointing


@pytest.mark.parametrize("sampling", [4, 5])
def test_random_atlas_parity(gpu, oracle_mod, sampling):
    o = oracle_mod
    cells = _random_atlas(3, sampling)
    grads = [(0.0, 0.0), (0.031, -0.017), (-0.052, 0.044)]
    nx = 96
    cell_rad = 2.0e-4
    g = gpu.StandardGridder(nx, nx, cell_rad, cell_rad)
    g.attach_aw(cells, sampling=sampling, phase_gradients=grads)
    gx, gy, du, dv = g.info.gx, g.info.gy, g.info.du, g.info.dv
    atlas = o.AwAtlas(cells, sampling)
    n = 6000
    u, v, w, vis, wt, cidx, pidx = _stream(11, n, 0.45 * gx * du, len(cells), len(grads))
    b = _batch(gpu, u, v, w, wt, vis, cidx, pidx)

    # plans, bit-exact
    plan, n32, n64 = g.aw_plan(b)
    plan = plan.cpu().numpy()
    n32 = n32.cpu().numpy()
    n64 = n64.cpu().numpy()
    for i in range(0, n, 7):
        rc, geo, r32, r64 = o.aw_plan(gx, gy, du, dv, atlas, int(cidx[i]), u[i], v[i], w[i])
        assert plan[i, 5] == rc, i
        if rc in (0, 4):
            assert tuple(plan[i, :5]) == geo
        if rc == 0:
            assert n32[i] == r32 and n64[i] == r64
    ok = plan[:, 5] == 0
    assert 0.5 * n < ok.sum() < n

    # plain FP64 grids against the oracle's sequential f64 grids
    psf = g.new_grids(1, gpu.F64)
    dirty = g.new_grids(1, gpu.F64)
    sums = g.grid_plain(b, gpu.MODE_DIRTY, gpu.F64, psf=psf, dirty=dirty)
    ref_d, st = o.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cidx, (vis * wt).astype(np.complex64), pidx)
    ref_p, _ = o.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cidx, wt.astype(np.complex64), pidx)
    assert np.array_equal(st == 0, ok)
    for got, ref in ((dirty, ref_d), (psf, ref_p)):
        got = got[0].cpu().numpy()
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
    nsum = float(np.sum(wt[ok].astype(np.float64) * np.abs(n64[ok])))
    assert sums[0]["gridded"] == ok.sum()
    assert abs(sums[0]["normalization_sumwt"] - nsum) <= 1e-12 * nsum

    # binned FP32 grids
    vs = g.visibility_set(b)
    p32 = g.new_grids(1, gpu.F32)
    d32 = g.new_grids(1, gpu.F32)
    vs.grid(gpu.MODE_DIRTY, psf=p32, dirty=d32)
    for got, ref in ((d32, ref_d), (p32, ref_p)):
        got = got[0].cpu().numpy().astype(np.complex128)
        assert np.max(np.abs(got - ref)) <= 1e-5 * np.max(np.abs(ref))
    vsum = vs.sums[0]
    assert abs(vsum["normalization_sumwt"] - nsum) <= 1e-12 * nsum

    # degrid: plain FP32 bit-exact, binned within the FP32 bar
    rng = np.random.default_rng(5)
    model = (rng.normal(size=(gx, gy)) + 1j * rng.normal(size=(gx, gy))).astype(np.complex64)
    ref_v, _ = o.aw_degrid(gx, gy, du, dv, atlas, grads, u, v, w, cidx, model, pidx)
    mt = t(model[None]).cuda()
    got = g.degrid_plain(b, mt, gpu.F32).cpu().numpy()
    assert np.array_equal(got[ok], ref_v[ok])
    outb = vs.degrid(mt)
    gotb = outb.cpu().numpy()
    scale = np.max(np.abs(ref_v[ok]))
    assert np.max(np.abs(gotb[ok] - ref_v[ok])) <= 1e-5 * scale

    # residual: (V - pred) * w gridded with the same taps
    resid_vals = ((vis - ref_v) * wt).astype(np.complex64)
    ref_r, _ = o.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cidx, resid_vals, pidx)
    r32 = g.new_grids(1, gpu.F32)
    vs.grid(gpu.MODE_RESIDUAL, model=mt, dirty=r32)
    got = r32[0].cpu().numpy().astype(np.complex128)
    assert np.max(np.abs(got - ref_r)) <= 1e-5 * np.max(np.abs(ref_r))


def test_w_plane_cell_selection(gpu, oracle_mod):
    """Without cf_cell the cell is round(sqrt(WIncr |w|)) clamped (awprojection.rs:1239-1246)."""
    cells = _random_atlas(8, 4)
    g = gpu.StandardGridder(64, 64, 3.0e-4, 3.0e-4)
    w_incr = 0.02
    g.attach_aw(cells, sampling=4, w_increment=w_incr)
    w = np.array([0.0, 10.0, -10.0, 60.0, 130.0, -500.0, 1e5])
    b = _batch(gpu, np.zeros_like(w), np.zeros_like(w), w, np.ones_like(w), np.ones_like(w))
    expect = np.clip(np.round(np.sqrt(w_incr * np.abs(w))), 0, len(cells) - 1).astype(int)
    plan_w, _, _ = g.aw_plan(b)
    b2 = _batch(gpu, np.zeros_like(w), np.zeros_like(w), w, np.ones_like(w), np.ones_like(w), cell=expect)
    plan_c, _, _ = g.aw_plan(b2)
    assert torch.equal(plan_w, plan_c)
    # tap tables: plane index encodes the cell (chip_plan reports it)
    raw, _ = g.plan(b)
    raw = raw.cpu().numpy()
    assert np.array_equal(raw[:, 4] // 2, expect)


def _airy_cells_numpy(o, gx, gy, cell_x, cell_y, s, n_w, max_abs_w, dish, blockage, freq, pa):
    conv = max(gx, gy)
    inner = max(conv // s, 2)
    half = inner // 2
    corr = o.gridder_correction_axis(inner)
    s0 = abs(cell_x) * s * gx / conv
    s1 = abs(cell_y) * s * gy / conv
    w_scale = (n_w - 1) ** 2 / (1.05 * max_abs_w) if n_w > 1 else 1.0
    idx = np.arange(-half, half)
    l = s0 * idx.astype(np.float64)
    m = s1 * idx.astype(np.float64)
    L, M = np.meshgrid(l, m, indexing="ij")
    rsq = L * L + M * M
    inside = rsq < 1.0
    lr = L * math.cos(pa) - M * math.sin(pa)
    mr = L * math.sin(pa) + M * math.cos(pa)
    vp = o.airy_offsets(lr.ravel(), mr.ravel(), freq, dish, blockage).reshape(L.shape)
    c = corr[idx + half]
    amp = ((c[:, None] * c[None, :]).astype(np.float32) * (vp * vp).astype(np.float32)).astype(np.float32)
    R = (s + 1) // 2
    out, peak, sups = [], None, []
    for k in range(n_w):
        wk = (k * k) / w_scale if n_w > 1 else 0.0
        with np.errstate(invalid="ignore"):
            phase = 2.0 * np.pi * wk * (np.sqrt(1.0 - rsq) - 1.0)
        re = (amp * np.cos(phase).astype(np.float32)).astype(np.float32)
        im = (amp * np.sin(phase).astype(np.float32)).astype(np.float32)
        scr = np.zeros((conv, conv), np.complex128)
        xi = np.where(idx >= 0, idx, idx + conv)
        vals = np.where(inside, re + 1j * im, 0)
        scr[np.ix_(xi, xi)] = vals
        T = np.fft.fft2(scr)
        if k == 0:
            peak = np.float32(abs(T[0, 0]))
        sup = o.find_w_support(T, peak, s)
        sups.append(sup)
        sh = 2 * (sup * s + R + 1)
        d = np.arange(sh) - sh // 2
        di = np.where(d < 0, d + conv, d)
        out.append(T[np.ix_(di, di)] / float(peak))
    o0 = out[0].shape[0] // 2
    S0 = sups[0]
    pb = sum(out[0][o0 + ix * s, o0 + iy * s].real for iy in range(-S0, S0 + 1) for ix in range(-S0, S0 + 1))
    return [c / pb for c in out], sups


@pytest.mark.parametrize("pa", [0.0, 0.6])
def test_airy_cells_match_numpy(gpu, oracle_mod, pa):
    o = oracle_mod
    nx, cell = 160, math.radians(30.0 / 3600.0)
    g = gpu.StandardGridder(nx, nx, cell, cell)
    s, n_w, max_w = 4, 3, 400.0
    g.attach_aw_airy(25.0, 2.0, 1.4e9, n_w, max_w, sampling=s, pa_rad=pa)
    got = g.aw_cells()
    ref, sups = _airy_cells_numpy(o, g.info.gx, g.info.gy, cell, cell, s, n_w, max_w, 25.0, 2.0, 1.4e9, pa)
    assert [c[1] for c in got] == sups and [c[2] for c in got] == sups
    for (gc, _, _), rc in zip(got, ref):
        gc = gc.cpu().numpy()
        assert gc.shape == rc.shape
        assert np.max(np.abs(gc - rc)) <= 2e-5 * np.max(np.abs(rc))
    # an Airy pattern is circularly symmetric: plane zero is real and even
    c0 = got[0][0].cpu().numpy()
    o0 = c0.shape[0] // 2
    core = c0[1:, 1:]
    assert np.max(np.abs(core - core[::-1, ::-1])) <= 1e-4 * np.max(np.abs(core))
    assert np.max(np.abs(c0.imag)) <= 1e-4 * np.max(np.abs(c0))


def test_airy_frequency_atlas_and_cell_selection(gpu):
    # an atlas at several CF frequencies holds, cell for cell, the atlases
    # built at each frequency alone (cell f * n_w + k), and every sample of a
    # rows x channels batch selects the nearest CF frequency and its W plane
    # round(sqrt(WIncr |w|)) (select_key_for_sample, awprojection.rs:1127-1180)
    nx, cell, s, n_w, max_w = 96, np.deg2rad(60.0 / 3600.0), 4, 5, 800.0
    freqs = np.array([1.05e9, 1.2e9, 1.35e9])
    multi = gpu.StandardGridder(nx, nx, cell, cell)
    multi.attach_aw_airy(25.0, 2.0, freqs, n_w, max_w, sampling=s)
    cells = multi.aw_cells()
    assert len(cells) == len(freqs) * n_w
    for f, fr in enumerate(freqs):
        one = gpu.StandardGridder(nx, nx, cell, cell)
        one.attach_aw_airy(25.0, 2.0, float(fr), n_w, max_w, sampling=s)
        for k, (c, xs, ys) in enumerate(one.aw_cells()):
            m, mx, my = cells[f * n_w + k]
            assert (mx, my) == (xs, ys)
            assert torch.equal(m, c)
    rng = np.random.default_rng(3)
    n_rows = 500
    chan = np.linspace(1.0e9, 1.4e9, 7)
    ls = chan / 299_792_458.0
    uvw = np.stack([rng.uniform(-300, 300, n_rows), rng.uniform(-300, 300, n_rows),
                    rng.uniform(-250, 250, n_rows)], axis=1)
    batch = gpu.RowChannelBatch(torch.from_numpy(uvw), torch.from_numpy(ls),
                                torch.zeros((n_rows, len(chan)), dtype=torch.complex64),
                                torch.ones((n_rows, len(chan)), dtype=torch.float32))
    got = multi.aw_select_cells(batch).cpu().numpy()
    nearest = np.argmin(np.abs(freqs[None, :] - (ls * 299_792_458.0)[:, None]), axis=1)
    w = uvw[:, 2:3] * ls[None, :]
    plane = np.clip(np.round(np.sqrt(multi.info.w_scale * np.abs(w))), 0, n_w - 1).astype(np.int64)
    assert np.array_equal(got, nearest[None, :] * n_w + plane)
    assert len(np.unique(got)) > n_w
