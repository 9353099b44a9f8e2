"""Pins of the CPU oracle (oracle/) against the reference's own test literals,
closed forms and brute force.  These run on the CPU only.

Reference literals used (bglenden/casa-rs crates/casa-imaging):
  gridder.rs:5577-5584  512x512 @ 0.08 arcsec -> 640x640 grid
  gridder.rs:5644-5650  257x257, composite padding 1.3 -> 360x360
  gridder.rs:5078-5107  8 explicit planes at 0.25 arcsec: w_scale = 49 / (0.25/cell);
                        cached normalisation equals the direct sum
  gridder.rs:5110-5121  automatic planes = floor(1.05*100*sin(1.0)) for 1000x500 @ 1e-3/4e-3
  gridder.rs:5124-5140  explicit max uvw 644.5775195221761, auto 105.0 at 80 arcsec
  gridder.rs:4764-4771  spheroidal kernel is zero beyond the support
  gridder.rs:4867-4891  degridding is the adjoint of gridding for one sample (1e-4)
  gridder.rs:5178-5188  a constant grid degrids to the constant (1e-4)
  execution.rs:18825-18851  plan accepts 2 of 4 samples with sumwt 5
"""
import numpy as np
import pytest
from scipy import integrate


def test_padding_literals(oracle_mod):
    o = oracle_mod
    g = o.StandardGridder(512, 512, 0.08 / 206_264.806_247, 0.08 / 206_264.806_247)
    assert g.grid_shape == (640, 640)
    assert g.blc == (64, 64)
    cell = np.deg2rad(8.638_889_530_690e-7)
    g2 = o.StandardGridder(257, 257, cell, cell, padding_mode=o.StandardGridder.PAD_COMPOSITE_1_3)
    assert g2.grid_shape == (360, 360)
    # 1024 -> 1250, 3413 -> 4096 (composite 2*3*5 rule); brute-force the rule
    for n, p in [(1024, 1.2), (3413, 1.2), (100, 1.2), (257, 1.3), (999, 1.2)]:
        got = o.composite_padded_len(n, p)
        cand = max(int(np.floor(p * n - 0.5)), n)
        cand += cand % 2
        while True:
            m = cand
            for f in (2, 3, 5):
                while m > 1 and m % f == 0:
                    m //= f
            if m == 1:
                break
            cand += 2
        assert got == cand
    assert o.composite_padded_len(1024, 1.2) == 1250
    assert o.composite_padded_len(3413, 1.2) == 4096


def _grdsf_eigen_residual(grdsf_fn):
    c = 3.0 * np.pi  # support m = 6 cells: c = pi m / 2
    nus = np.linspace(0.0, 1.0, 41)

    def ft(nu):
        return integrate.quad(lambda e: (1 - e * e) * grdsf_fn(abs(e)) * np.cos(c * e * nu), -1, 1,
                              limit=200, epsabs=1e-13, epsrel=1e-12)[0]

    f = np.array([ft(n) for n in nus])
    g = np.array([grdsf_fn(n) for n in nus])
    return np.max(np.abs(f / f[0] - g / g[0]))


def test_grdsf_is_the_alpha1_spheroidal_eigenfunction(oracle_mod):
    """grdsf approximates psi_{1,0}(c = 3 pi, nu), defined by the eigen-equation
    lambda psi(nu) = integral_{-1}^{1} (1 - eta^2) psi(eta) exp(i c eta nu) d eta:
    the gridding kernel (1 - eta^2) psi transforms into the correction psi.  The
    rational approximation satisfies it to < 1e-6; a wrong coefficient, sign or
    branch point breaks it by orders of magnitude (checked with a planted error)."""
    o = oracle_mod
    assert _grdsf_eigen_residual(o.grdsf) < 1e-6
    assert o.grdsf(-0.1) == 0.0 and o.grdsf(1.1) == 0.0

    def planted(nu):  # P0[2] off by 1e-3 relative
        v = o.grdsf(nu)
        if nu < 0.75:
            d = nu * nu - 0.5625
            v += 6.278660e-1 * 1e-3 * d * d / (1 + 8.212018e-1 * d + 2.078043e-1 * d * d)
        return v

    assert _grdsf_eigen_residual(planted) > 1e-5


def test_spheroidal_kernel_compact(oracle_mod):
    o = oracle_mod
    assert o.spheroidal_kernel(3.5, 3.0) == 0.0
    assert o.spheroidal_kernel(3.0, 3.0) == 0.0
    assert o.spheroidal_kernel(0.0, 3.0) == pytest.approx(o.grdsf(0.0), rel=1e-7)


def test_tap_weights_normalised_symmetric(oracle_mod):
    o = oracle_mod
    for os_ in (100, 128):
        g = o.StandardGridder(64, 64, 1e-5, 1e-5, oversampling=os_)
        taps = g.taps
        assert taps.shape == (os_ + 1, 7)
        assert np.allclose(taps.sum(axis=1), 1.0, atol=2e-7)
        # offset +k and -k are mirror images (up to the order of the f32 norm sum)
        assert np.max(np.abs(taps[::-1, ::-1] - taps)) <= 1.2e-7
        # row for offset 0: taps at |delta| = 3 vanish (1 - nu^2 = 0)
        assert taps[os_ // 2, 0] == 0.0 and taps[os_ // 2, 6] == 0.0


def test_correction_axis_is_inverse_grdsf(oracle_mod):
    o = oracle_mod
    n = 640
    c = o.correction_axis(n)
    nu = np.abs(np.arange(n) - n / 2) / (n / 2)
    psi = np.array([o.grdsf(x) for x in nu])
    expect = np.where(psi > 1e-6, 1.0 / psi, 0.0).astype(np.float32)
    assert np.array_equal(c, expect)
    assert c[n // 2] == np.float32(1.0 / o.grdsf(0.0))


def test_sinc_axis(oracle_mod):
    o = oracle_mod
    s = o.sinc_axis(640, 4)
    assert s[320] == 1.0
    x = np.pi * (np.arange(640) - 320) / (640 * 4)
    ref = np.where(np.arange(640) == 320, 1.0, np.sin(x) / np.where(x == 0, 1, x))
    assert np.max(np.abs(s - ref)) < 1e-6


def test_w_scale_and_plane_literals(oracle_mod):
    o = oracle_mod
    cell = np.deg2rad(0.25 / 3600.0)
    g = o.StandardGridder(64, 64, cell, cell, padding_mode=1)
    wp = o.WProjector(g, 20_000.0, explicit_planes=8)
    assert abs(wp.w_scale - 49.0 / (0.25 / cell)) <= np.finfo(float).eps
    plan, norm = wp.plan([12_000.25], [-9_000.75], [-4_000.0])
    assert plan[0, 0] >= 0
    # cached normalisation equals the direct sum (gridder.rs:5078-5107)
    import ctypes

    ok = ctypes.c_int()
    k = np.ascontiguousarray(wp.kernels[plan[0, 4]])
    direct = o.lib().co_w_sample_normalization(
        k.ctypes.data_as(ctypes.c_void_p), wp.width, int(wp.supports[plan[0, 4]]), wp.sampling,
        int(plan[0, 2]), int(plan[0, 3]), ctypes.byref(ok))
    assert ok.value == 1 and abs(direct - norm[0]) <= np.finfo(np.float32).eps

    assert o.suggested_w_plane_count(1000, 500, 1e-3, 4e-3, 100.0) == int(1.05 * 100.0 * np.sin(1.0))
    assert o.suggested_w_plane_count(1000, 500, 1e-3, 4e-3, 0.0) == 1
    cell80 = np.deg2rad(80.0 / 3600.0)
    assert abs(o.w_max_uvw(cell80, 100.0, 16) - 644.577_519_522_176_1) < 1e-12
    assert o.w_max_uvw(cell80, 100.0, None) == 105.0


def test_plan_counts_literal(oracle_mod):
    o = oracle_mod
    g = o.StandardGridder(32, 32, 1e-4, 1e-4, padding_mode=o.StandardGridder.PAD_NONE)
    psf, dirty, s = g.grid(np.array([0.0, 4.0, 8.0, 12.0]), np.array([0.0, 1.0, 2.0, 3.0]),
                           np.array([1, 2, 0, 3], np.float32), np.array([1, 2, 1, np.nan], np.float32),
                           np.ones(4, np.complex64))
    assert (s.gridded, s.skipped) == (2, 2)
    assert abs(s.normalization_sumwt - 5.0) < 1e-6 and abs(s.reported_sumwt - 5.0) < 1e-6


def _brute_plan(u, du, n, os_):
    """Brute-force restatement of one axis: nearest cell (ties away from 0),
    offset in units of 1/os rounded the same way."""
    pos = u / du + n / 2.0
    anchor = np.floor(np.abs(pos) + 0.5) * np.sign(pos)
    off = (anchor - pos) * os_
    off = np.floor(np.abs(off) + 0.5) * np.sign(off)
    return anchor, off


def test_standard_plan_brute_force(oracle_mod):
    o = oracle_mod
    rng = np.random.default_rng(3)
    cell = 1e-5
    g = o.StandardGridder(100, 100, cell, cell, oversampling=128)
    du = g.du
    # random points plus exact half-cell and half-offset ties
    u = np.concatenate([rng.uniform(-70 * du, 70 * du, 3000), (np.arange(-20, 20) + 0.5) * du,
                        (np.arange(-20, 20) + 0.5 / 128) * du])
    v = np.concatenate([rng.uniform(-70 * du, 70 * du, 3000), (np.arange(-20, 20) + 0.5) * du,
                        (np.arange(-20, 20) - 0.5 / 128) * du])
    plan = g.plan(u, v)
    ax, ox = _brute_plan(u, du, g.gx, 128)
    ay, oy = _brute_plan(-v, g.dv, g.gy, 128)
    ok = (ax - 3 >= 0) & (ax + 3 < g.gx) & (ay - 3 >= 0) & (ay + 3 < g.gy)
    assert np.array_equal(plan[:, 0] >= 0, ok)
    assert np.array_equal(plan[ok, 0], (ax[ok] - 3).astype(np.int64))
    assert np.array_equal(plan[ok, 1], (ay[ok] - 3).astype(np.int64))
    assert np.array_equal(plan[ok, 2], (ox[ok] + 64).astype(np.int64))
    assert np.array_equal(plan[ok, 3], (oy[ok] + 64).astype(np.int64))


def test_adjoint_single_sample(oracle_mod):
    """gridder.rs:4867-4891 restated: <G s, b> = <s, G^H b>."""
    o = oracle_mod
    g = o.StandardGridder(32, 32, 1e-4, 1e-4, padding_mode=o.StandardGridder.PAD_NONE)
    sample = np.complex64(0.5 - 0.25j)
    _, a, _ = g.grid(np.array([120.0]), np.array([-80.0]), np.ones(1, np.float32),
                     np.ones(1, np.float32), np.array([sample]))
    x, y = np.meshgrid(np.arange(32), np.arange(32), indexing="ij")
    b = ((x - 10.0) * 0.01 + 1j * (y - 6.0) * -0.02).astype(np.complex64)
    lhs = np.sum(np.conj(a) * b)
    rhs = np.conj(sample) * g.degrid(np.array([120.0]), np.array([-80.0]), b)[0]
    assert abs(lhs - rhs) < 1e-4


def test_constant_grid_degrids_to_constant(oracle_mod):
    o = oracle_mod
    g = o.StandardGridder(32, 32, 1e-4, 1e-4, padding_mode=o.StandardGridder.PAD_NONE)
    grid = np.full((32, 32), 2.5 - 1.0j, np.complex64)
    val = g.degrid(np.array([55.0]), np.array([-40.0]), grid)[0]
    assert abs(val.real - 2.5) < 1e-4 and abs(val.imag + 1.0) < 1e-4


def test_w_kernel_plane_zero_is_spheroid(oracle_mod):
    """With w = 0 the screen is psi(l) psi(m) only; its transform is real up to
    the unpaired edge sample at -inner/2 (psi(1) ~ 4e-3 of psi(0)), and the
    normalised plane sums to 1 at integer taps."""
    o = oracle_mod
    g = o.StandardGridder(64, 64, 1e-4, 1e-4)
    wp = o.WProjector(g, 100.0, explicit_planes=4)
    k0 = wp.kernels[0]
    assert np.max(np.abs(k0.imag)) < 1e-3 * np.max(np.abs(k0))
    S, s = wp.supports[0], wp.sampling
    idx = np.abs(np.arange(-S, S + 1) * s)
    assert abs(np.sum(k0[np.ix_(idx, idx)].real, dtype=np.float64) - 1.0) < 1e-5
    # larger w needs at least as much support
    assert np.all(np.diff(wp.supports) >= 0)


def test_direct_dft_matches_gridded_dirty_image(oracle_mod):
    """FFT pipeline vs the direct DFT oracle (lib.rs:89810-89888, rms < 3e-2)."""
    o = oracle_mod
    rng = np.random.default_rng(11)
    nx, cell = 64, np.deg2rad(20.0 / 3600.0)
    g = o.StandardGridder(nx, nx, cell, cell)
    n = 400
    umax = 0.4 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = np.zeros(n)
    l0, m0 = 5 * cell, -3 * cell
    vis = np.exp(-2j * np.pi * (u * l0 + v * m0)).astype(np.complex64)
    wt = np.ones(n, np.float32)
    psf_g, d_g, sums = g.grid(u, v, wt, wt, vis)
    psf, res, peak = g.dirty_products(psf_g, d_g, sums)
    dpsf, dres, dpeak, _ = o.direct_psf_and_residual(nx, nx, cell, cell, u, v, w, wt, wt, vis)
    rms = np.sqrt(np.mean((res - dres) ** 2)) / np.max(np.abs(dres))
    assert rms < 3e-2
    px = np.unravel_index(np.argmax(res), res.shape)
    assert px == (nx // 2 + 5, nx // 2 + 3)


def test_direct_adjoint_window_matches_full_image(oracle_mod):
    """The windowed direct adjoint is the same sum restricted to a window."""
    o = oracle_mod
    rng = np.random.default_rng(13)
    n, nx, ny = 500, 24, 20
    cell = np.deg2rad(30.0 / 3600.0)
    u, v, w = rng.uniform(-800, 800, n), rng.uniform(-800, 800, n), rng.uniform(-50, 50, n)
    vals = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    full = o.direct_adjoint(nx, ny, cell, cell, u, v, w, vals)
    win = o.direct_adjoint_window(nx, ny, cell, cell, 5, 3, 11, 9, u, v, w, vals)
    assert np.array_equal(win, full[5:16, 3:12])


# ---------------------------------------------------------------------------
# density weighting: cell literals and the f2 formula from the reference tests
# ---------------------------------------------------------------------------

def _f32_bits(b):
    return float(np.array([b], np.uint32).view(np.float32)[0])


def test_density_cells_match_reference_literals(oracle_mod):
    o = oracle_mod
    vis, cube = o.VIS_IMAGING_WEIGHT, o.CUBE_BRIGGS
    du = dv = 1.0 / (32.0 * 1.0e-4)
    # density_cells_truncate_to_image_weight_grid (gridder.rs:5190-5214)
    cell = lambda conv, u, v: o.density_cell(32, 32, 1.0e-4, 1.0e-4, conv, u, v)
    assert cell(vis, 0.0, 0.0) == (16, 16)
    assert cell(vis, 0.99 * du, -0.99 * dv) == (15, 15)
    assert cell(vis, 1.01 * du, -1.01 * dv) == (14, 14)
    assert cell(vis, -1.01 * du, 1.01 * dv) == (17, 17)
    # cube_density_cells_round_to_nearest_weight_grid_location (:5293-5317)
    assert cell(cube, 0.49 * du, -0.49 * dv) == (16, 16)
    assert cell(cube, 0.51 * du, -0.51 * dv) == (17, 17)
    # visibility_weight_density_cell_uses_casa_separate_float_operations (:5217-5242):
    # a fused multiply-add would give cell 2
    c = 0.08 / 206_264.806_247
    assert o.density_cell(1280, 1280, c, c, vis, _f32_bits(0x499CA142), 0.0) == (3, 640)
    # ..._uses_casa_float_scale_and_origin (:5244-5262) and
    # ..._narrows_uv_before_float_transform (:5264-5290)
    c = np.deg2rad(0.6 / 3600.0)
    assert o.density_cell(4096, 4096, c, c, vis, 0.0, _f32_bits(0x4767CA0C)) == (2048, 2755)
    assert o.density_cell(4096, 4096, c, c, vis, 1_091.081_732_177_734_5, 0.0) == (2035, 2048)
    # cells on the edge row/column 0 are rejected (anchor must be > 0)
    assert cell(cube, -16.0 * du, 0.0) is None


def test_briggs_f2_matches_reference_literal(oracle_mod):
    # cube_briggs_f2_and_denominator_follow_casa_formula (weighting.rs:3454-3458)
    d = np.array([[0.0, 2.0], [3.0, 0.0]], np.float32)
    assert abs(oracle_mod.briggs_f2(d, 0.0) - 9.615385) < 1e-5
    # robust shifts f2 by 100^-robust; the total doubles without conjugates
    assert abs(oracle_mod.briggs_f2(d, 1.0) / oracle_mod.briggs_f2(d, 0.0) - 0.01) < 1e-6
    assert abs(oracle_mod.briggs_f2(d, 0.0, includes_conjugates=False) / 9.615385 - 2.0) < 1e-6


def test_density_weights_brute_force(oracle_mod):
    """Uniform and Briggs reweighting against a plain numpy restatement."""
    o = oracle_mod
    rng = np.random.default_rng(21)
    nx, cell, n = 40, 1.0e-3, 3000
    umax = 0.5 / cell
    u, v = rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)
    w = rng.choice([0.5, 1.0, 2.0], n).astype(np.float32)
    w[::97] = 0.0
    plane = rng.integers(0, 3, n).astype(np.uint32)
    for mode in (o.WEIGHT_UNIFORM, o.WEIGHT_BRIGGS):
        got, dens, f2 = o.density_weights(nx, nx, cell, cell, u, v, w, mode, robust=0.5, plane=plane, n_planes=3)
        ref_d = np.zeros((3, nx, nx), np.float64)
        cells = [o.density_cell(nx, nx, cell, cell, o.CUBE_BRIGGS, a, b) for a, b in zip(u, v)]
        mir = [o.density_cell(nx, nx, cell, cell, o.CUBE_BRIGGS, -a, -b) for a, b in zip(u, v)]
        for i in range(n):
            if w[i] > 0:
                for cc in (cells[i], mir[i]):
                    if cc:
                        ref_d[plane[i]][cc] += w[i]
        assert np.array_equal(dens, ref_d.astype(np.float32))  # sums of 0.5, 1, 2 are exact
        for p in range(3):
            pos = ref_d[p][ref_d[p] > 0]
            f2_ref = (5.0 * 10 ** -0.5) ** 2 / (np.sum(pos ** 2) / np.sum(ref_d[p])) if mode == o.WEIGHT_BRIGGS else 1.0
            assert abs(f2[p] - f2_ref) <= 1e-6 * f2_ref
        for i in range(n):
            c = cells[i]
            if c is None or w[i] <= 0:
                assert got[i] == 0.0
                continue
            d = np.float32(dens[plane[i]][c])
            ref = w[i] / d if mode == o.WEIGHT_UNIFORM else w[i] / (d * f2[plane[i]] + np.float32(1.0))
            assert got[i] == np.float32(ref)


# ---------------------------------------------------------------------------
# phase-centre rotation
# ---------------------------------------------------------------------------

def _uvw_toward(b, ra, dec):
    """uvw of baseline b (equatorial XYZ) toward (ra, dec), MS sign convention
    (u, v opposite to the textbook u = b . (-sin ra, cos ra, 0))."""
    u = np.array([-np.sin(ra), np.cos(ra), 0.0])
    v = np.array([-np.sin(dec) * np.cos(ra), -np.sin(dec) * np.sin(ra), np.cos(dec)])
    s = np.array([np.cos(dec) * np.cos(ra), np.cos(dec) * np.sin(ra), np.sin(dec)])
    return np.array([-(b @ u), -(b @ v), b @ s])


def test_phase_rotation_reprojects_uvw(oracle_mod):
    """gridft_uvw_rotation_matrix(a, b) maps uvw toward a onto uvw toward b and
    the phase vector gives uvw' . phrot = w_b - w_a (the geometric delay)."""
    o = oracle_mod
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = (rng.uniform(0, 2 * np.pi), rng.uniform(-1.2, 1.2))
        bdir = (a[0] + rng.uniform(-0.05, 0.05), a[1] + rng.uniform(-0.05, 0.05))
        R, ph = o.phase_rotation(a, bdir)
        assert np.max(np.abs(R @ R.T - np.eye(3))) < 1e-15
        base = rng.normal(size=(50, 3)) * 3000.0
        uvw_a = np.array([_uvw_toward(x, *a) for x in base])
        uvw_b = np.array([_uvw_toward(x, *bdir) for x in base])
        out, shift, _ = o.rotate_rows(R, ph, uvw_a, [1.0e9], np.ones((50, 1), np.complex64))
        assert np.max(np.abs(out - uvw_b)) < 1e-9
        assert np.max(np.abs(shift - (uvw_b[:, 2] - uvw_a[:, 2]))) < 1e-9
    R, ph = o.phase_rotation((0.3, 0.5), (0.3, 0.5))
    assert np.max(np.abs(R - np.eye(3))) < 1e-15 and np.all(ph == 0.0)


def test_phase_rotate_visibility_literal(oracle_mod):
    """phase_rotate_visibility_applies_geometric_phasor (casa-imaging lib.rs:78080-78090)."""
    o = oracle_mod
    R, ph = np.eye(3), np.array([0.0, 0.0, 1.0])  # shift = w
    f = 1.0e9
    quarter = 299_792_458.0 / (4.0 * f)
    uvw = np.array([[0.0, 0.0, quarter], [0.0, 0.0, 0.0]])
    _, shift, vis = o.rotate_rows(R, ph, uvw, [f], np.ones((2, 1), np.complex64))
    assert abs(vis[0, 0].real) < 1e-6 and abs(vis[0, 0].imag + 1.0) < 1e-6
    assert vis[1, 0] == 1.0  # zero shift leaves the sample unchanged
    _, _, vis = o.rotate_rows(R, ph, uvw, [0.0], np.ones((2, 1), np.complex64))
    assert np.all(vis == 1.0)  # zero frequency too
