"""Pins of the oracle's FP64 gridding modes, the references that the FP32
device results are measured against at full size (CPU only).

* WProjector.grid64 agrees with the Complex32 restatement WProjector.grid
  (pinned elsewhere against the reference's literals and the direct
  transform) to FP32 rounding, with identical sums and counts; its residual
  mode satisfies the adjoint identity Re <M, G(-p)> = -sum |p|^2, with p the
  restatement's own degrid.
* aw_grid64 agrees with the reference-precision AW restatement (aw_grid_f64,
  aw_degrid) to FP32 rounding and satisfies the same identity.
* w_dirty_products64 agrees with WProjector.dirty_products (f32 narrowing)
  to FP32 rounding.
A planted mistake (the conjugation of the degrid dropped) fails the
identity.
"""
import numpy as np

import synth


def _w_setup(O):
    cell = np.deg2rad(600.0 / 3600.0)
    g = O.StandardGridder(128, 128, cell, cell)
    wp = O.WProjector(g, 300.0, explicit_planes=16, sampling=4)
    u, v, w, wt, sf, vis, gr = synth.random_samples(31, 20000, 0.5 * g.gx * g.du, wmax=300.0)
    return g, wp, (u, v, w, wt, sf, vis, gr)


def test_w_grid64_matches_complex32_restatement(oracle_mod):
    O = oracle_mod
    g, wp, (u, v, w, wt, sf, vis, gr) = _w_setup(O)
    p32, d32, s32 = wp.grid(u, v, w, wt, sf, vis, gridable=gr, mode=1)
    p64, d64, s64 = wp.grid64(u, v, w, wt, sf, vis, gridable=gr, mode=1)
    assert (s32.gridded, s32.skipped) == (s64.gridded, s64.skipped)
    assert abs(s32.normalization_sumwt - s64.normalization_sumwt) <= 1e-12 * s64.normalization_sumwt
    for a, b in ((p32, p64), (d32, d64)):
        assert np.max(np.abs(a - b)) <= 1e-6 * np.max(np.abs(b))
    a = wp.dirty_products(p32, d32, s32)
    b = wp.dirty_products64(p64, d64, s64)
    assert abs(a[2] - b[2]) <= 1e-6 * b[2]
    assert np.max(np.abs(a[0] - b[0])) <= 2e-6
    assert np.max(np.abs(a[1] - b[1])) <= 2e-5 * np.max(np.abs(b[1]))


def test_w_residual64_adjoint_identity(oracle_mod):
    O = oracle_mod
    g, wp, (u, v, w, wt, sf, vis, gr) = _w_setup(O)
    n = len(u)
    model = np.zeros((g.nx, g.ny), np.float32)
    rng = np.random.default_rng(4)
    model[rng.integers(20, 108, 15), rng.integers(20, 108, 15)] = rng.uniform(0.2, 1.0, 15)
    m64 = wp.model_grid64(model)
    ok = np.isfinite(u) & np.isfinite(w)
    u, v, w = u[ok], v[ok], w[ok]
    ones = np.ones(len(u), np.float32)
    zero = np.zeros(len(u), np.complex64)
    _, r64, s = wp.grid64(u, v, w, ones, ones, zero, mode=2, model=m64)
    p = wp.degrid(u, v, w, m64.astype(np.complex64))
    acc = np.isfinite(p)
    assert s.gridded == int(acc.sum()) > 0.5 * len(u)
    lhs = float(np.sum(np.conj(m64) * r64).real)
    rhs = -float(np.sum(np.abs(p[acc].astype(np.complex128)) ** 2))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
    # planted mistake: grid the prediction's conjugate (a degrid without conj)
    _, bad, _ = wp.grid64(u, v, w, ones, ones, zero, mode=2, model=np.conj(m64))
    assert abs(float(np.sum(np.conj(m64) * bad).real) - rhs) > 0.1 * abs(rhs)


def _aw_setup(O):
    rng = np.random.default_rng(12)
    cells = []
    for k in range(3):
        n = 45 + 2 * k
        x, y = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
        r2 = ((x - n // 2) ** 2 + (y - n // 2) ** 2).astype(np.float32)
        pix = (np.exp(-r2 / np.float32(60.0 + 10 * k)) * np.exp(1j * np.float32(0.002 * (k + 1)) * r2)).astype(np.complex64)
        cells.append((pix, 2, 2))
    atlas = O.AwAtlas(cells, 8)
    gx = gy = 96
    du = dv = 1.0 / (gx * 2.0e-3)
    grads = np.array([[0.0, 0.0], [0.013, -0.021]])
    n = 4000
    u = rng.uniform(-0.4, 0.4, n) * gx * du
    v = rng.uniform(-0.4, 0.4, n) * gy * dv
    w = rng.uniform(-50.0, 50.0, n)
    cell_of = rng.integers(0, 3, n).astype(np.uint32)
    pointing = rng.integers(0, 2, n).astype(np.uint32)
    vis = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    return atlas, gx, gy, du, dv, grads, u, v, w, cell_of, pointing, vis, wt


def test_aw_grid64_matches_reference_precision(oracle_mod):
    O = oracle_mod
    atlas, gx, gy, du, dv, grads, u, v, w, cell_of, pointing, vis, wt = _aw_setup(O)
    vals = (vis * wt).astype(np.complex64)
    ref, st = O.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, vals, pointing)
    psf, dirty, nsum, count = O.aw_grid64(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, vis, wt, mode=1,
                                          pointing_of=pointing)
    assert count == int(np.sum(st == 0)) > 0.8 * len(u)
    assert np.max(np.abs(dirty - ref)) <= 1e-6 * np.max(np.abs(ref))
    wref, _ = O.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, wt.astype(np.complex64), pointing)
    assert np.max(np.abs(psf - wref)) <= 1e-6 * np.max(np.abs(wref))


def test_aw_residual64_adjoint_identity(oracle_mod):
    O = oracle_mod
    atlas, gx, gy, du, dv, grads, u, v, w, cell_of, pointing, vis, wt = _aw_setup(O)
    rng = np.random.default_rng(5)
    model = (rng.normal(size=(gx, gy)) + 1j * rng.normal(size=(gx, gy))).astype(np.complex128)
    ones = np.ones(len(u), np.float32)
    zero = np.zeros(len(u), np.complex64)
    _, res, _, count = O.aw_grid64(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, zero, ones, mode=2,
                                   model=model, pointing_of=pointing)
    p, st = O.aw_degrid(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, model.astype(np.complex64), pointing)
    acc = st == 0
    assert count == int(acc.sum())
    # p_i = sum_k conj(t_ik) M_k / conj(n_i), so the residual grid of zero
    # data, R = -sum_i p_i t_i, has <M, R> = -sum_i p_i conj(p_i conj(n_i))
    # = -sum_i |p_i|^2 n_i
    plan_n = np.array([complex(O.aw_plan(gx, gy, du, dv, atlas, int(cell_of[i]), u[i], v[i], w[i])[3])
                       if acc[i] else 0j for i in range(len(u))])
    lhs = complex(np.sum(np.conj(model) * res))
    rhs = -complex(np.sum(np.abs(p[acc].astype(np.complex128)) ** 2 * plan_n[acc]))
    assert abs(lhs - rhs) <= 1e-5 * abs(rhs)
