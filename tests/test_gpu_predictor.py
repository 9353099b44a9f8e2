"""The model predictor on the GPU (chip_predict_model) against the oracle's
restatement of StandardMfsModelPredictor (lib.rs:3375-3404).

Bars: the degrid itself is bit-exact for the same model grid (FP64 anchors
and offsets, products and sums in the reference's order); the model grid
(FP32 FFT on the device against the oracle's FP64 FFT) within 2e-6 of its
peak; end to end within 0.03 of the peak of the direct transform, as the
oracle's own bound."""
import numpy as np
import pytest
import torch

from test_oracle_predictor import CELL, NX, REFERENCE_UV, random_uv, structured_model

pytestmark = pytest.mark.gpu


def test_predictor_bit_exact_for_the_same_grid(gpu, oracle_mod):
    model = structured_model()
    o = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, model)
    p = gpu.StandardMfsModelPredictor(NX, NX, CELL, CELL, torch.from_numpy(model))
    assert p.gridder.grid_shape == o.gridder.grid_shape == (360, 360)
    gm = p.model_grid[0].cpu().numpy()
    assert np.max(np.abs(gm - o.model_grid)) < 2e-6 * np.max(np.abs(o.model_grid))
    u, v = random_uv(o, 20000, 4, frac=0.5)
    du = o.gridder.du
    # rounding ties of the anchor and of the 1/100 offset, off-grid and
    # non-finite coordinates
    k = np.arange(-40, 40)
    tie_u = np.concatenate([(k + 0.5) * du, (k + 0.005) * du, (k - 0.005) * du, [1e9, np.nan, 0.0]])
    tie_v = np.concatenate([(k - 0.5) * du, (k + 0.005) * du, (k + 0.5) * du, [0.0, 0.0, np.inf]])
    u = np.concatenate([u, tie_u])
    v = np.concatenate([v, tie_v])
    got = p.predict(u, v).cpu().numpy()
    ref = o.predict(u, v, model_grid=gm)
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))
    assert np.count_nonzero(got == 0) >= 3


def test_predictor_end_to_end_against_direct_transform(gpu, oracle_mod):
    model = structured_model()
    p = gpu.StandardMfsModelPredictor(NX, NX, CELL, CELL, torch.from_numpy(model))
    u = np.array([a for a, _ in REFERENCE_UV])
    v = np.array([b for _, b in REFERENCE_UV])
    direct = oracle_mod.direct_predict(NX, NX, CELL, CELL, model, u, v, np.zeros(len(u)))
    assert np.max(np.abs(p.predict(u, v).cpu().numpy() - direct)) < 0.02
    o = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, model)
    ur, vr = random_uv(o, 400, 3)
    dr = oracle_mod.direct_predict(NX, NX, CELL, CELL, model, ur, vr, np.zeros(len(ur)))
    assert np.max(np.abs(p.predict(ur, vr).cpu().numpy() - dr)) < 0.03 * np.max(np.abs(dr))


def test_predictor_rows_by_channels(gpu, oracle_mod):
    # a cube: one model per channel, uvw in metres scaled per channel
    rng = np.random.default_rng(9)
    n_rows, n_chan = 3000, 4
    freq = np.linspace(1.0e9, 1.2e9, n_chan)
    ls = freq / 299_792_458.0
    base = structured_model()
    models = np.stack([base * np.float32(1.0 + 0.1 * c) for c in range(n_chan)])
    o = oracle_mod.StandardMfsModelPredictor(NX, NX, CELL, CELL, base)
    umax = 0.45 * o.gridder.gx * o.gridder.du / ls[-1]
    uvw = np.stack([rng.uniform(-umax, umax, n_rows), rng.uniform(-umax, umax, n_rows),
                    rng.uniform(-50, 50, n_rows)], axis=1)
    p = gpu.StandardMfsModelPredictor(NX, NX, CELL, CELL, torch.from_numpy(models))
    assert p.n_grids == n_chan
    got = p.predict(uvw_m=uvw, lambda_scale=ls).cpu().numpy().reshape(n_rows, n_chan)
    for c in range(n_chan):
        gm = p.model_grid[c].cpu().numpy()
        ref = o.predict(uvw[:, 0] * ls[c], uvw[:, 1] * ls[c], model_grid=gm)
        assert np.array_equal(np.ascontiguousarray(got[:, c]).view(np.uint32), ref.view(np.uint32))


def test_predictor_edge_cases(gpu):
    z = gpu.StandardMfsModelPredictor(NX, NX, CELL, CELL, torch.zeros((NX, NX)))
    assert z.model_grid is None
    assert torch.all(z.predict(np.array([10.0, 20.0]), np.array([5.0, 1.0])) == 0)
    assert z.predict(np.zeros(0), np.zeros(0)).numel() == 0
    g = gpu.StandardGridder(64, 64, CELL, CELL)
    g.attach_wproject(100.0, planes=4)
    s = gpu.StandardMfsModelPredictor(NX, NX, CELL, CELL, torch.from_numpy(structured_model()))
    s.gridder = g  # a W gridder is refused: the predictor degrids with spheroid taps
    with pytest.raises(gpu.InvalidRequest):
        s.predict(np.array([10.0]), np.array([5.0]), model_grid=g.new_grids(1))
