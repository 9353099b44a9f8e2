"""Hogbom minor cycle on the GPU (chip_hogbom) against the oracle's CPU loop.

Bar: bit-exact.  The peak search is a comparison, the component one f32
product and each residual update one f32 product then one subtraction, so
residual and model planes, component counts and stop codes must equal the
oracle's (run_hogbom_minor_cycle_cpu_with_progress, lib.rs:59748-59885) bit
for bit: on the reference's own literal case, on cubes with masks, plateaus
of equal |value|, PSFs smaller and larger than the image, per-plane cycle and
nsigma thresholds, a NaN at the search seed, no cleanable pixel, on 1024^2
planes with a beam-like PSF, and on one large plane, which the GPU splits
into many row slices (peak candidates reduced across slices, slices the PSF
window misses kept from the previous component).
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

CASE_PSF = np.array([0, 0, 0, .05, 0, 0, 0, 0, .1, .2, .1, 0, 0, .1, .2, .4, .2, .05,
                     .05, .2, .4, 1, .4, .1, 0, .1, .2, .4, .2, .05, 0, 0, .05, .1, .05, 0],
                    np.float32).reshape(6, 6)
CASE_RES = np.array([0, 0, 0, 0, 0, 0, 0, 0, .05, .1, .05, 0, 0, .05, .15, .4, .15, 0,
                     0, .1, .4, 1.2, .4, .05, 0, .05, .15, .4, .15, 0, 0, 0, 0, .05, 0, 0],
                    np.float32).reshape(6, 6)


def run_both(gpu, oracle_mod, res, psf, gain, niter, threshold=0.0, cyc=0.0, nsig=0.0, mask=None, model=None):
    res = np.asarray(res, np.float32)
    if res.ndim == 2:
        res = res[None]
    model = np.zeros_like(res) if model is None else np.asarray(model, np.float32)
    r_dev = torch.from_numpy(res.copy()).cuda()
    m_dev = torch.from_numpy(model.copy()).cuda()
    upd, rep, stop = gpu.run_hogbom_minor_cycle(r_dev, torch.from_numpy(np.ascontiguousarray(psf, np.float32)), m_dev, gain,
                                                niter, threshold, cyc, nsig, mask=None if mask is None else
                                                torch.from_numpy(np.ascontiguousarray(mask)))
    o_r, o_m, o_upd, o_stop = oracle_mod.hogbom(res, psf, model, gain, niter + 1, threshold, cyc, nsig,
                                                mask=mask)
    return (r_dev.cpu().numpy(), m_dev.cpu().numpy(), upd, rep, stop), (o_r, o_m, o_upd, o_stop)


def assert_same(g, o):
    r, m, upd, _, stop = g
    o_r, o_m, o_upd, o_stop = o
    assert np.array_equal(upd.astype(np.uint64), o_upd)
    assert np.array_equal(stop, o_stop)
    assert np.array_equal(r.view(np.uint32), o_r.view(np.uint32))
    assert np.array_equal(m.view(np.uint32), o_m.view(np.uint32))


def test_reference_literal(gpu, oracle_mod):
    # hogbom_minor_cycle_matches_hclean_inclusive_iteration_budget (lib.rs:95317-95403)
    g, o = run_both(gpu, oracle_mod, CASE_RES, CASE_PSF, 0.5, 1)
    assert_same(g, o)
    r, m, upd, rep, stop = g
    assert int(upd[0]) == 2 and int(rep[0]) == 1 and int(stop[0]) == gpu.CLEAN_BUDGET
    assert abs(float(m[0][3, 3]) - 0.9) < 1e-6
    # with a cycle threshold of 0.15 and niter 4 (lib.rs:95522-95600's inputs)
    g, o = run_both(gpu, oracle_mod, CASE_RES, CASE_PSF, 0.5, 4, cyc=0.15)
    assert_same(g, o)


@pytest.mark.parametrize("kx,ky,masked,ties", [(31, 27, False, False), (9, 9, True, False), (48, 40, True, True),
                                               (200, 180, False, True)])
def test_random_cubes(gpu, oracle_mod, kx, ky, masked, ties):
    rng = np.random.default_rng(kx * 7 + ky)
    planes, nx, ny = 6, 96, 80
    res = rng.normal(size=(planes, nx, ny)).astype(np.float32)
    if ties:
        res = np.round(res * 2.0) / 2.0
    psf = rng.uniform(-0.15, 0.15, size=(planes, kx, ky)).astype(np.float32)
    psf[:, kx // 2, ky // 2] = 1.0
    mask = (rng.uniform(size=(planes, nx, ny)) < 0.6) if masked else None
    cyc = rng.uniform(0.2, 1.5, planes).astype(np.float32)
    nsig = rng.uniform(0.0, 1.5, planes).astype(np.float32)
    g, o = run_both(gpu, oracle_mod, res, psf, 0.1, 150, threshold=0.05, cyc=cyc, nsig=nsig, mask=mask)
    assert_same(g, o)
    print("components", g[2], "stops", g[4])


def test_nan_seed_and_no_pixels(gpu, oracle_mod):
    rng = np.random.default_rng(3)
    res = rng.normal(size=(3, 16, 12)).astype(np.float32)
    res[0, 0, 0] = np.nan          # the seed of the reference's search: it stays the peak
    res[1, 5, 7] = np.nan          # elsewhere a NaN never wins
    mask = np.ones((3, 16, 12), bool)
    mask[2] = False                # nothing to clean
    psf = np.zeros((5, 5), np.float32)
    psf[2, 2] = 1.0
    psf[1, 2] = psf[2, 1] = 0.3
    g, o = run_both(gpu, oracle_mod, res, np.broadcast_to(psf, (3, 5, 5)), 0.2, 10, mask=mask)
    r, m, upd, _, stop = g
    o_r, o_m, o_upd, o_stop = o
    assert np.array_equal(upd.astype(np.uint64), o_upd) and np.array_equal(stop, o_stop)
    assert int(stop[2]) == gpu.CLEAN_NO_PIXELS and int(upd[2]) == 0
    # NaN propagates the same way on both sides
    for a, b in ((r, o_r), (m, o_m)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(np.nan_to_num(a).view(np.uint32), np.nan_to_num(b).view(np.uint32))


def test_full_size_planes(gpu, oracle_mod):
    """1024^2 planes, a beam-like PSF of the image's size, 60 components."""
    rng = np.random.default_rng(8)
    planes, n = 3, 1024
    yy, xx = np.mgrid[0:n, 0:n].astype(np.float32)
    c = n // 2
    psf = (np.exp(-((xx - c) ** 2 + (yy - c) ** 2) / (2 * 3.0 ** 2))
           * np.cos(0.05 * (xx - c)) * np.cos(0.04 * (yy - c))).astype(np.float32)
    res = (0.01 * rng.normal(size=(planes, n, n))).astype(np.float32)
    for p in range(planes):
        for _ in range(5):
            x0, y0 = rng.integers(100, n - 100, 2)
            res[p] += rng.uniform(0.2, 1.0) * np.roll(np.roll(psf, x0 - c, 0), y0 - c, 1)
    g, o = run_both(gpu, oracle_mod, res, np.broadcast_to(psf, (planes, n, n)), 0.1, 59, cyc=0.0, nsig=0.0)
    assert_same(g, o)
    assert all(int(u) == 60 for u in g[2])


@pytest.mark.parametrize("nan_seed", [False, True])
def test_single_large_plane(gpu, oracle_mod, nan_seed):
    """One 2048 x 1536 plane (1024 row slices of two rows), a 65 x 49 PSF patch,
    a mask whose first unmasked pixel lies beyond 10^4 pixels, and plateaus
    of equal |value| spread over many slices."""
    rng = np.random.default_rng(21 + nan_seed)
    nx, ny = 2048, 1536
    res = np.round(rng.normal(size=(1, nx, ny)) * 4.0).astype(np.float32) / 4.0
    mask = rng.uniform(size=(1, nx, ny)) < 0.8
    mask.reshape(-1)[:10007] = False
    seed = int(np.flatnonzero(mask.reshape(-1))[0])
    if nan_seed:
        res.reshape(-1)[seed] = np.nan
    psf = rng.uniform(-0.1, 0.1, size=(65, 49)).astype(np.float32)
    psf[32, 24] = 1.0
    g, o = run_both(gpu, oracle_mod, res, psf[None], 0.1, 199, threshold=0.01, mask=mask)
    r, m, upd, _, stop = g
    o_r, o_m, o_upd, o_stop = o
    assert np.array_equal(upd.astype(np.uint64), o_upd) and np.array_equal(stop, o_stop)
    for a, b in ((r, o_r), (m, o_m)):
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(np.nan_to_num(a).view(np.uint32), np.nan_to_num(b).view(np.uint32))
    if not nan_seed:
        assert int(upd[0]) == 200


def test_rejects_bad_arguments(gpu):
    dev = torch.device("cuda")
    r = torch.zeros((2, 8, 8), device=dev)
    with pytest.raises(gpu.InvalidRequest):
        gpu.run_hogbom_minor_cycle(r.double(), torch.ones(3, 3), torch.zeros_like(r), 0.1, 5)
    with pytest.raises(gpu.InvalidRequest):
        gpu.run_hogbom_minor_cycle(r, torch.ones(3, 3), torch.zeros((2, 8, 7), device=dev), 0.1, 5)
