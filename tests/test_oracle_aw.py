"""Pins of the oracle's AWProject CF-cell projector and Airy A-term (CPU only).

Reference literals and invariants used (bglenden/casa-rs crates/casa-imaging):
  gridder.rs:4343-4534  32x32 unpadded grid, 12x12 cell, sampling 2, support 2:
                        sample (0.3 du, -0.3 dv, 12) plans to loc (16, 16),
                        off (-1, -1), conjugated; the f64 normalisation is the
                        exact sum of the conjugated taps at origin 6 + 2i - 1;
                        a gridded tap equals value * raw * exp(i phase) (1e-5);
                        w <= 0 is not conjugated and its norm is the conjugate
  gridder.rs:4537-4582  f64 grids receive the Complex32 product value * tap,
                        not the product of promoted operands
  gridder.rs:4585-4613  degridding times conj(norm) is the adjoint of gridding (2e-5)
  gridder.rs:4616-4633  a 68x68 cell with support 16 cannot fit a 32x32 grid
  lib.rs:86437-86486    ALMA / ACA Airy support radii at 100 GHz
  lib.rs:86489-86503    ALMA Airy trace sample 0.999848 (5e-5)
The Airy table is also checked against scipy.special.j1, an independent
implementation of the Bessel function.
"""
import math

import numpy as np
import pytest
from scipy import special

DU = 1.0 / (32 * 1.0e-4)  # unpadded 32x32 grid at 1e-4 rad: du = 1 / (n cell)


def _cell(shape, fn, supports=2):
    x, y = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    re, im = fn(x.astype(np.float32), y.astype(np.float32))
    return (np.asarray(re, np.float32) + 1j * np.asarray(im, np.float32)).astype(np.complex64), supports, supports


def test_aw_plan_offsets_conjugation_and_phase(oracle_mod):
    o = oracle_mod
    cell = _cell((12, 12), lambda x, y: (np.float32(0.01) * x + np.float32(0.2), np.float32(-0.02) * y + np.float32(0.1)))
    atlas = o.AwAtlas([cell], sampling=2)
    grad = np.array([0.07, -0.11])
    assert o.aw_kernel_origin(12) == 6 and o.aw_kernel_origin(13) == 7
    rc, geo, n32, n64 = o.aw_plan(32, 32, DU, DU, atlas, 0, 0.3 * DU, -0.3 * DU, 12.0)
    assert rc == o.AW_OK
    assert geo == (16, 16, -1, -1, 1)
    k = cell[0]
    taps = [np.conj(k[6 + 2 * ix - 1, 6 + 2 * iy - 1]) for iy in range(-2, 3) for ix in range(-2, 3)]
    exp64 = sum(complex(np.float64(t.real), np.float64(t.imag)) for t in taps)
    assert n64 == exp64
    assert abs(complex(n32) - exp64) < 1e-5
    # one tap of the gridded sample against value * raw * exp(i phase)
    value = np.complex64(1.25 - 0.5j)
    grid, st = o.aw_grid_f64(32, 32, DU, DU, atlas, grad[None, :], [0.3 * DU], [-0.3 * DU], [12.0], [0],
                             np.array([value]))
    assert st[0] == 0
    ix, iy = -1, 2
    raw = np.conj(k[6 + 2 * ix - 1, 6 + 2 * iy - 1])
    phase = (2 * ix - 1) * 0.07 + (2 * iy - 1) * -0.11
    expected = complex(value) * complex(raw) * complex(math.cos(phase), math.sin(phase))
    assert abs(grid[16 + ix, 16 + iy] - expected) < 1e-5
    # the phased tap itself: pointing phasor narrowed per axis, product through f64
    t = o.aw_taps(atlas, 0, geo, grad)
    px = np.complex64(complex(np.float32(math.cos((2 * ix - 1) * 0.07)), np.float32(math.sin((2 * ix - 1) * 0.07))))
    py = np.complex64(complex(np.float32(math.cos((2 * iy - 1) * -0.11)), np.float32(math.sin((2 * iy - 1) * -0.11))))
    ph = np.complex64(complex(px) * complex(py))
    assert t[iy + 2, ix + 2] == np.complex64(raw * ph)
    # w <= 0: not conjugated, norm conjugated
    for w in (-12.0, 0.0):
        rc, geo2, m32, _ = o.aw_plan(32, 32, DU, DU, atlas, 0, 0.3 * DU, -0.3 * DU, w)
        assert rc == 0 and geo2[4] == 0
        assert abs(complex(m32) - exp64.conjugate()) < 1e-5


def test_aw_f64_grid_keeps_complex32_contribution(oracle_mod):
    o = oracle_mod
    cell = _cell((12, 12), lambda x, y: (np.float32(0.013) * x + np.float32(0.21), np.float32(-0.017) * y + np.float32(0.09)))
    atlas = o.AwAtlas([cell], sampling=2)
    value = np.complex64(12345.678 - 9876.543j)
    grid, st = o.aw_grid_f64(32, 32, DU, DU, atlas, [[0.07, -0.11]], [0.0], [0.0], [12.0], [0], np.array([value]))
    assert st[0] == 0
    ix, iy = -1, 2
    tap = np.conj(cell[0][6 + 2 * ix, 6 + 2 * iy])
    phase = (2 * ix) * 0.07 + (2 * iy) * -0.11
    tap = np.complex64(tap * np.complex64(complex(np.float32(math.cos(phase)), np.float32(math.sin(phase)))))
    casa = np.complex64(value * tap)
    got = grid[16 + ix, 16 + iy]
    assert got == complex(casa)
    promoted = complex(value) * complex(tap)
    assert got != promoted


def test_aw_degrid_is_normalized_adjoint(oracle_mod):
    o = oracle_mod
    cell = _cell((13, 13), lambda x, y: (np.float32(0.03) * x + np.float32(0.1), np.float32(0.01) * y - np.float32(0.08)))
    atlas = o.AwAtlas([cell], sampling=2)
    du = 1.0 / (24 * 1.0e-4)
    grad = [[0.04, -0.02]]
    x = np.complex64(0.75 - 0.4j)
    gx, _ = o.aw_grid_f64(24, 24, du, du, atlas, grad, [0.0], [0.0], [5.0], [0], np.array([x]))
    a, b = np.meshgrid(np.arange(24), np.arange(24), indexing="ij")
    y = (np.float32(0.001) * a + 1j * (np.float32(-0.002) * b)).astype(np.complex64)
    lhs = np.sum(np.conj(gx) * y)
    pred, st = o.aw_degrid(24, 24, du, du, atlas, grad, [0.0], [0.0], [5.0], [0], y)
    rc, _, n32, _ = o.aw_plan(24, 24, du, du, atlas, 0, 0.0, 0.0, 5.0)
    rhs = np.conj(x) * (complex(pred[0]) * np.conj(complex(n32)))
    assert st[0] == 0 and rc == 0
    assert abs(lhs - rhs) < 2e-5
    # the division is num-complex's: (a c + b d, b c - a d) / (c^2 + d^2) in f32
    assert abs(complex(pred[0]) * np.conj(complex(n32)) - complex(np.sum(np.conj(o.aw_taps(atlas, 0, (12, 12, 0, 0, 1), grad[0])).ravel() * y[10:15, 10:15].T.ravel()))) < 1e-5


def test_aw_rejections(oracle_mod):
    o = oracle_mod
    ones = (np.ones((68, 68), np.complex64), 16, 16)
    atlas = o.AwAtlas([ones], sampling=2)
    rc, *_ = o.aw_plan(32, 32, DU, DU, atlas, 0, 0.0, 0.0, 0.0)
    assert rc == o.AW_OUTSIDE_GRID
    small = o.AwAtlas([(np.ones((12, 12), np.complex64), 2, 2)], sampling=2)
    assert o.aw_plan(32, 32, DU, DU, small, 0, np.nan, 0.0, 1.0)[0] == o.AW_NON_FINITE_COORDINATE
    assert o.aw_plan(32, 32, DU, DU, small, 0, 0.0, 0.0, np.inf)[0] == o.AW_NON_FINITE_COORDINATE
    # a cell whose stored extent is shorter than support*sampling + |off| on one side
    short = o.AwAtlas([(np.ones((10, 12), np.complex64), 2, 2)], sampling=2)
    assert o.aw_plan(32, 32, DU, DU, short, 0, 0.3 * DU, 0.0, 1.0)[0] == o.AW_OK  # off -1 fits
    assert o.aw_plan(32, 32, DU, DU, short, 0, -0.3 * DU, 0.0, 1.0)[0] == o.AW_KERNEL_INDEX_OUTSIDE_CELL
    zero = o.AwAtlas([(np.zeros((12, 12), np.complex64), 2, 2)], sampling=2)
    assert o.aw_plan(32, 32, DU, DU, zero, 0, 0.0, 0.0, 1.0)[0] == o.AW_INVALID_NORMALIZATION


def test_airy_reference_literals(oracle_mod):
    o = oracle_mod
    f = 2.2972554e11
    r = math.radians(3.42261 / (60.0 * (f / 1.0e9)))
    assert abs(o.airy_radius(r, f, 10.7, 0.75) - 0.999848) < 5.0e-5
    rad = lambda arcsec: math.radians(arcsec / 3600.0)
    assert o.airy_radius(rad(120.0), 100e9, 10.7, 0.75) != 0.0
    assert o.airy_radius(rad(250.0), 100e9, 6.25, 0.75) != 0.0
    assert o.airy_radius(rad(160.0), 100e9, 10.7, 0.75) == 0.0
    assert o.airy_radius(rad(320.0), 100e9, 6.25, 0.75) == 0.0
    assert o.airy_max_radius(10.7, 0.75) == 250.0 and o.airy_max_radius(6.25, 0.75) == 500.0
    assert o.airy_max_radius(25.0, 2.0) == 1.784 * 60.0


@pytest.mark.parametrize("dish,blockage", [(25.0, 2.0), (25.0, 0.0), (10.7, 0.75)])
def test_airy_table_matches_scipy_bessel(oracle_mod, dish, blockage):
    o = oracle_mod
    t = o.airy_table(dish, blockage)
    mr = o.airy_max_radius(dish, blockage)
    idx = np.arange(10000, dtype=np.float64)
    x = idx * (mr * 7.016 / (1.566 * 60.0) * dish / 24.5) / 9999.0
    with np.errstate(divide="ignore", invalid="ignore"):
        if blockage <= 0.0:
            ref = 2.0 * special.j1(x) / x
        else:
            a = (dish / blockage) ** 2
            lr = dish / blockage
            ref = (a * 2.0 * special.j1(x) / x - 2.0 * special.j1(x * lr) / (x * lr)) / (a - 1.0)
    ref[0] = 1.0
    ref = ref.astype(np.float32)
    assert t[0] == 1.0
    ulp = np.abs(t.view(np.int32).astype(np.int64) - ref.view(np.int32).astype(np.int64))
    # the two Bessel implementations agree to a few f64 ulp; after narrowing
    # to f32 at most one f32 step apart
    assert ulp.max() <= 1
    # unblocked aperture: first null of 2 J1(x)/x at x = 3.8317
    if blockage == 0.0:
        k = int(np.argmax(t < 0))
        assert abs(x[k] - 3.8317) < 2 * (x[1] - x[0])


def test_airy_offsets_lookup(oracle_mod):
    o = oracle_mod
    t = o.airy_table(25.0, 2.0)
    l = np.linspace(-0.02, 0.02, 101)
    m = np.full_like(l, 0.003)
    got = o.airy_offsets(l, m, 1.4e9, 25.0, 2.0, t)
    # same truncated index as the single-radius form when the radius is the hypotenuse of f32 degrees
    for li, mi, g in zip(l, m, got):
        ld, md = np.float32(li * 57.29577951308232), np.float32(mi * 57.29577951308232)
        rd = np.sqrt(np.float32(ld * ld + md * md), dtype=np.float32)
        r = float(np.float32(float(rd) * 60.0 * 1.4))
        mr = o.airy_max_radius(25.0, 2.0)
        ref = 0.0 if r > mr else t[int(min(max(math.floor(r * 9999.0 / mr), 0), 9999))]
        assert g == ref
    assert got[50] == t[int(math.floor(float(np.float32(float(np.float32(0.003 * 57.29577951308232)) * 60 * 1.4)) * 9999 / (1.784 * 60)))]
    # symmetric in l
    assert np.array_equal(got, got[::-1])
