"""The oracle's Hogbom minor cycle (run_hogbom_minor_cycle_cpu_with_progress,
lib.rs:59748-59885) pinned without the reference:

* the reference's own literal (hogbom_minor_cycle_matches_hclean_inclusive_
  iteration_budget, lib.rs:95317-95403): niter 1 in CASA's inclusive mode
  commits two components and leaves model[3, 3] = 0.9;
* a closed form: with a delta PSF a point of flux a keeps its position and
  after k components the model holds a (1 - (1 - g)^k);
* a vectorised restatement in numpy (argmax of |r| with the y-major tie
  rule, float32 product then subtraction over the overlap window) that must
  agree bit for bit, including ties, masks and a PSF smaller than the image;
* the stop rule of minor_cycle_stop_reason (lib.rs:74653-74677) on its
  branches.
"""
import numpy as np
import pytest

CASE_PSF = np.array([0, 0, 0, .05, 0, 0, 0, 0, .1, .2, .1, 0, 0, .1, .2, .4, .2, .05,
                     .05, .2, .4, 1, .4, .1, 0, .1, .2, .4, .2, .05, 0, 0, .05, .1, .05, 0],
                    np.float32).reshape(6, 6)
CASE_RES = np.array([0, 0, 0, 0, 0, 0, 0, 0, .05, .1, .05, 0, 0, .05, .15, .4, .15, 0,
                     0, .1, .4, 1.2, .4, .05, 0, .05, .15, .4, .15, 0, 0, 0, 0, .05, 0, 0],
                    np.float32).reshape(6, 6)


def test_reference_inclusive_budget_literal(oracle_mod):
    # niter 1 in CasaInclusive mode -> budget 2 (hogbom_component_budget, lib.rs:64541-64551)
    r, m, upd, stop = oracle_mod.hogbom(CASE_RES, CASE_PSF, np.zeros((6, 6), np.float32), 0.5, budget=2)
    assert int(upd[0]) == 2 and int(stop[0]) == oracle_mod.STOP_NONE
    assert abs(float(m[0][3, 3]) - 0.9) < 1e-6
    assert np.count_nonzero(m[0]) == 1


def test_delta_psf_closed_form(oracle_mod):
    n, a, g = 16, np.float32(2.5), np.float32(0.3)
    psf = np.zeros((n, n), np.float32)
    psf[n // 2, n // 2] = 1.0
    img = np.zeros((n, n), np.float32)
    img[5, 11] = a
    for k in (1, 4, 9):
        r, m, upd, _ = oracle_mod.hogbom(img, psf, np.zeros_like(img), g, budget=k)
        assert int(upd[0]) == k
        assert np.count_nonzero(m[0]) == 1 and m[0][5, 11] > 0
        assert abs(float(m[0][5, 11]) - a * (1.0 - (1.0 - g) ** k)) < 1e-5
        assert abs(float(r[0][5, 11]) - a * (1.0 - g) ** k) < 1e-5


def numpy_hogbom(residual, psf, model, gain, budget, threshold, cyc, nsig, mask, stop_fn):
    """Vectorised restatement for tiny planes."""
    r, m = residual.astype(np.float32).copy(), model.astype(np.float32).copy()
    nx, ny = r.shape
    kx, ky = psf.shape
    xs, ys = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    key = ys * nx + xs
    allowed = np.ones((nx, ny), bool) if mask is None else mask.astype(bool)
    done, why = 0, 0
    while done < budget:
        if not allowed.any():
            why = 6
            break
        a = np.where(allowed, np.abs(r), -1.0)
        best = a.max()
        cand = allowed & (a == best)
        k = key[cand].min()
        px, py = int(k % nx), int(k // nx)
        pv = r[px, py]
        s = stop_fn(abs(pv), threshold, cyc, nsig)
        if s:
            why = s
            break
        comp = np.float32(gain) * pv
        m[px, py] = m[px, py] + comp
        xmin, ymin = px - kx // 2, py - ky // 2
        x0, y0, x1, y1 = max(xmin, 0), max(ymin, 0), min(xmin + kx - 1, nx - 1), min(ymin + ky - 1, ny - 1)
        if x0 <= x1 and y0 <= y1:
            p = comp * psf[x0 - xmin:x1 - xmin + 1, y0 - ymin:y1 - ymin + 1]
            r[x0:x1 + 1, y0:y1 + 1] = r[x0:x1 + 1, y0:y1 + 1] - p
        done += 1
    return r, m, done, why


@pytest.mark.parametrize("seed,kx,with_mask,ties", [(1, 12, False, False), (2, 7, True, False), (3, 12, False, True),
                                                     (4, 20, True, True)])
def test_against_numpy_restatement(oracle_mod, seed, kx, with_mask, ties):
    rng = np.random.default_rng(seed)
    nx, ny = 12, 10
    img = rng.normal(size=(nx, ny)).astype(np.float32)
    if ties:  # plateaus of equal |value| (signs differ): the y-major key decides
        img = np.round(img * 2) / 2
    psf = rng.uniform(-0.2, 0.2, size=(kx, kx)).astype(np.float32)
    psf[kx // 2, kx // 2] = 1.0
    mask = (rng.uniform(size=(nx, ny)) < 0.7) if with_mask else None
    ref = numpy_hogbom(img, psf, np.zeros_like(img), 0.1, 40, 0.0, 0.05, 0.0, mask, oracle_mod.minor_cycle_stop)
    r, m, upd, stop = oracle_mod.hogbom(img, psf, np.zeros_like(img), 0.1, 40, cycle_threshold=0.05, mask=mask)
    assert int(upd[0]) == ref[2] and int(stop[0]) == ref[3]
    assert np.array_equal(r[0].view(np.uint32), ref[0].view(np.uint32))
    assert np.array_equal(m[0].view(np.uint32), ref[1].view(np.uint32))


def test_stop_rule_branches(oracle_mod):
    f = oracle_mod.minor_cycle_stop
    assert f(0.5, 1.0, 0.0, 0.0) == oracle_mod.STOP_GLOBAL          # below the global threshold
    assert f(0.5, 0.1, 0.6, 0.2) == oracle_mod.STOP_CYCLE           # cycle threshold dominates
    assert f(0.7, 0.1, 0.6, 0.2) == oracle_mod.STOP_NONE
    assert f(0.5, 0.1, 0.2, 0.6) == oracle_mod.STOP_NSIGMA          # nsigma above cycle and global
    assert f(0.7, 0.1, 0.2, 0.6) == oracle_mod.STOP_NONE
    # tiny global thresholds stop within max(1 %, 2e-8) of the threshold
    # (threshold_reached_with_tolerance, lib.rs:74628-74635)
    t = np.float32(1.0e-6)
    assert f(t + np.float32(1.5e-8), t, 0.0, 0.0) == oracle_mod.STOP_GLOBAL
    assert f(t + np.float32(5e-8), t, 0.0, 0.0) == oracle_mod.STOP_NONE


def test_no_cleanable_pixels(oracle_mod):
    img = np.ones((4, 4), np.float32)
    r, m, upd, stop = oracle_mod.hogbom(img, np.ones((3, 3), np.float32), np.zeros_like(img), 0.5, 5,
                                        mask=np.zeros((4, 4), bool))
    assert int(upd[0]) == 0 and int(stop[0]) == oracle_mod.STOP_NO_PIXELS
