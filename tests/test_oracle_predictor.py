"""The oracle's standard MFS model predictor (StandardMfsModelPredictor,
lib.rs:3375-3404; degrid_sample_product_planned_sectdgrid,
gridder.rs:1227-1266) pinned by the padding literal, closed forms and the
direct transform, with planted mistakes that must fail."""
import numpy as np

# standard_mfs_model_predictor_matches_casacore_for_structured_model,
# lib.rs:89744-89768: geometry, structured model and sample coordinates
NX = 257
CELL = np.deg2rad(8.638_889_530_690e-7)
REFERENCE_UV = [
    (4_806.297_926_382_51, 41_290.840_313_424_32),
    (-38_890.191_177_123_3, -12_300.584_882_047_77),
    (24_915.177_739_689_71, -34_020.365_105_376_14),
    (-9_024.365_419_946_97, 7_115.436_092_750_48),
]


def structured_model():
    """0.0025 ring - 0.0007 shoulder in f32 (lib.rs:89749-89759)."""
    x = np.arange(NX, dtype=np.float32)[:, None]
    y = np.arange(NX, dtype=np.float32)[None, :]
    dx = (x - np.float32(129.25)) / np.float32(27.0)
    dy = (y - np.float32(126.5)) / np.float32(19.0)
    ring = np.exp(-(dx * dx + dy * dy))
    shoulder = np.exp(-(((x - np.float32(87.0)) / np.float32(13.0)) ** 2
                        + ((y - np.float32(169.0)) / np.float32(21.0)) ** 2))
    return (np.float32(0.0025) * ring - np.float32(0.0007) * shoulder).astype(np.float32)


def random_uv(pred, n, seed, frac=0.45):
    rng = np.random.default_rng(seed)
    umax = frac * pred.gridder.gx * pred.gridder.du
    return rng.uniform(-umax, umax, n), rng.uniform(-umax, umax, n)


def test_predictor_padding_literal(oracle_mod):
    # new_with_casa_composite_padding: 257 -> 360 (gridder.rs:5644-5650)
    p = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, structured_model())
    assert p.gridder.grid_shape == (360, 360)


def test_constant_grid_predicts_the_constant(oracle_mod):
    p = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, structured_model())
    c = np.complex64(2.5 - 1.25j)
    grid = np.full(p.gridder.grid_shape, c, np.complex64)
    u, v = random_uv(p, 500, 1, frac=0.49)
    got = p.predict(u, v, model_grid=grid)
    # sum(c w) / sum(w) = c up to the rounding of 49 products and sums
    assert np.max(np.abs(got - c)) <= 64 * np.finfo(np.float32).eps * abs(c)
    # off the grid or not finite: no taps, predicted as 0 (lib.rs:3398-3400)
    du = p.gridder.du
    bad_u = np.array([0.6 * 360 * du, np.nan, np.inf, 0.0])
    bad_v = np.array([0.0, 0.0, 0.0, -(180 - 2.6) * du])
    assert np.all(p.predict(bad_u, bad_v, model_grid=grid) == 0)


def test_zero_model_predicts_zero(oracle_mod):
    p = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, np.zeros((NX, NX), np.float32))
    assert p.model_grid is None
    assert np.all(p.predict(np.array([100.0]), np.array([50.0])) == 0)


def test_tap_sum_normalisation_equals_normalised_taps(oracle_mod):
    # sum g wx wy / sum wx wy == sum g (wx / sum wx)(wy / sum wy): the
    # predictor's degrid is the normalised-tap degrid up to rounding
    p = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, structured_model())
    u, v = random_uv(p, 2000, 2)
    a = p.predict(u, v)
    b = p.gridder.degrid(u, v, p.model_grid)
    assert np.max(np.abs(a - b)) <= 1e-6 * np.max(np.abs(a))


def test_predictor_against_direct_transform(oracle_mod):
    O = oracle_mod
    model = structured_model()
    p = O.StandardMfsModelPredictor(NX, NX, CELL, CELL, model)
    u = np.array([a for a, _ in REFERENCE_UV])
    v = np.array([b for _, b in REFERENCE_UV])
    direct = O.direct_predict(NX, NX, CELL, CELL, model, u, v, np.zeros(len(u)))
    # the reference test's own samples: |V| ~ 3.43, gridding error ~ 0.015
    assert np.max(np.abs(p.predict(u, v) - direct)) < 0.02
    ur, vr = random_uv(p, 400, 3)
    dr = O.direct_predict(NX, NX, CELL, CELL, model, ur, vr, np.zeros(len(ur)))
    peak = np.max(np.abs(dr))
    assert np.max(np.abs(p.predict(ur, vr) - dr)) < 0.03 * peak
    # planted mistakes: the model grid without the grid correction, and v
    # with the other sign
    g = p.gridder
    raw = np.zeros(g.grid_shape, np.complex64)
    bx, by = g.blc
    raw[bx:bx + NX, by:by + NX] = model
    no_apod = O.centered_fft2(raw).astype(np.complex64)
    assert np.max(np.abs(p.predict(ur, vr, model_grid=no_apod) - dr)) > 0.05 * peak
    assert np.max(np.abs(p.predict(ur, -vr) - dr)) > 0.2 * peak
