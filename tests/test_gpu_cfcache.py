"""A CASA CF cache on the GPU: the cache's imaging cells as the AW atlas, the
per-sample CASA cell selection on the device, and AWProject gridding with the
selected cells, against the oracle.

Bars: atlas pixels and supports bit-exact to the cells the reader loads;
per-sample cell indices identical to oracle.cf_select_key (FP64 steps in the
same order); plain FP64 grids within 1e-12 and binned FP32 grids within 1e-5
of the peak against the oracle's sequential FP64 gridding with those cells.
"""
import numpy as np
import pytest
import torch

from cf_cache_writer import write_test_cell

pytestmark = pytest.mark.gpu

C = 299_792_458.0
SAMPLING = 4
W_INCR = 0.01


def _cell_pixels(rng, support, shape):
    x, y = np.meshgrid(np.arange(shape) - shape / 2, np.arange(shape) - shape / 2, indexing="ij")
    env = np.exp(-(x * x + y * y) / (2.0 * (1.2 * SAMPLING * support) ** 2))
    return (env * (1.0 + 0.2 * rng.normal(size=env.shape)) + 0.1j * env * rng.normal(size=env.shape)).astype(
        np.complex64)


def _write_cache(root, seed=4):
    """2 frequencies x 3 W planes (w_k = k^2 / WIncr, supports 2, 3, 4) x Mueller
    0 and 15 x PAs -30 and 40: 24 paired cells with distinct pixels."""
    rng = np.random.default_rng(seed)
    freqs, pas, muellers = [1.1e9, 1.3e9], [-30.0, 40.0], [0, 15]
    for fi, f in enumerate(freqs):
        for k, sup in enumerate([2, 3, 4]):
            shape = 2 * (sup * SAMPLING + SAMPLING // 2 + 1)
            for mi, m in enumerate(muellers):
                for pi, pa in enumerate(pas):
                    stem = f"0_{mi}_CF_{fi}_{k}_{pi}.im"
                    kw = dict(mueller_element=m, pa=pa, w_incr=W_INCR, sampling=float(SAMPLING), support=sup,
                              crpix=(shape / 2, shape / 2), cdelt=(-1.0, 1.0), big_endian=bool((fi + k) % 2))
                    write_test_cell(root, f"CFS_{stem}", f, k * k / W_INCR, False,
                                    pixels=_cell_pixels(rng, sup, shape), **kw)
                    write_test_cell(root, f"WTCFS_{stem}", f, k * k / W_INCR, True,
                                    pixels=_cell_pixels(rng, sup, shape), **kw)


def _key_index(keys, k):
    return -1 if k is None else keys.index(k)


def test_cache_atlas_selection_and_gridding(gpu, oracle_mod, tmp_path):
    o = oracle_mod
    _write_cache(str(tmp_path))
    cache = gpu.AwConvolutionFunctionCache.open(str(tmp_path))
    keys = [(k.frequency_hz, k.w_value_lambda, k.mueller_element, k.parallactic_angle_deg) for k in cache.keys()]
    assert len(keys) == 24

    nx, cell = 96, 2.0e-4
    grads = [(0.0, 0.0), (0.021, -0.034)]
    g = gpu.StandardGridder(nx, nx, cell, cell)
    g.attach_aw_cache(cache, phase_gradients=grads)
    cells = g.aw_cells()
    host_cells = []
    for k in range(len(keys)):
        im, _ = cache.load(k)
        meta = cache.metadata(k)[0]
        pix, xs, ys = cells[k]
        assert (xs, ys) == (meta["x_support"], meta["y_support"])
        assert np.array_equal(pix.cpu().numpy().view(np.uint32), im.view(np.uint32))
        host_cells.append((im, xs, ys))

    rng = np.random.default_rng(12)
    n_rows = 900
    chan = np.array([1.02e9, 1.18e9, 1.2e9, 1.37e9])
    ls = chan / C
    gx, du = g.info.gx, g.info.du
    umax = 0.4 * gx * du / ls.max()
    uvw = np.stack([rng.uniform(-umax, umax, n_rows), rng.uniform(-umax, umax, n_rows),
                    rng.uniform(-400.0, 400.0, n_rows) / ls.max()], axis=1)
    uvw[::13, 2] = 0.0
    pa_row = rng.uniform(-200.0, 200.0, n_rows)
    pa_row[::7] = 5.0  # equidistant from -30 and 40: ties to the smaller bit pattern
    vis = (rng.normal(size=(n_rows, len(chan))) + 1j * rng.normal(size=(n_rows, len(chan)))).astype(np.complex64)
    wt = rng.uniform(0.3, 2.0, (n_rows, len(chan))).astype(np.float32)
    pointing = torch.from_numpy(rng.integers(0, len(grads), n_rows).astype(np.int32))
    batch = gpu.RowChannelBatch(torch.from_numpy(uvw), torch.from_numpy(ls), torch.from_numpy(vis),
                                torch.from_numpy(wt), pointing=pointing)
    w_lambda = uvw[:, 2:3] * ls[None, :]

    f_ref = 1.2e9
    for mueller, conj, pred in ((0, False, False), (15, True, False), (0, False, True), (5, False, False)):
        got = g.cf_select_cells(batch, mueller, chan_freq_hz=chan, reference_frequency_hz=f_ref,
                                parallactic_angle_deg=pa_row, conjugate_beams=conj, prediction=pred).cpu().numpy()
        want = np.array([[_key_index(keys, o.cf_select_key(keys, W_INCR, chan[c], f_ref, w_lambda[r, c], mueller,
                                                           pa_row[r], conj, pred))
                          for c in range(len(chan))] for r in range(n_rows)])
        assert np.array_equal(got, want), (mueller, conj, pred)
        if mueller == 5:
            assert np.all(got == -1)
        else:
            assert len(np.unique(got)) >= 8
    # a scalar PA and lambda_scale * c frequencies
    got = g.cf_select_cells(batch, 0, reference_frequency_hz=f_ref, parallactic_angle_deg=-100.0).cpu().numpy()
    want = np.array([[_key_index(keys, o.cf_select_key(keys, W_INCR, ls[c] * C, f_ref, w_lambda[r, c], 0, -100.0))
                      for c in range(len(chan))] for r in range(n_rows)])
    assert np.array_equal(got, want)

    # gridding with the selected cells, channel by channel, against the oracle
    sel = g.cf_select_cells(batch, 0, chan_freq_hz=chan, reference_frequency_hz=f_ref, parallactic_angle_deg=pa_row)
    assert int((sel < 0).sum()) == 0
    batch.cf_cell = sel
    atlas = o.AwAtlas(host_cells, SAMPLING)
    psf = g.new_grids(len(chan), gpu.F64)
    dirty = g.new_grids(len(chan), gpu.F64)
    g.grid_plain(batch, gpu.MODE_DIRTY, gpu.F64, psf=psf, dirty=dirty)
    vs = g.visibility_set(batch)
    p32 = g.new_grids(len(chan), gpu.F32)
    d32 = g.new_grids(len(chan), gpu.F32)
    vs.grid(gpu.MODE_DIRTY, psf=p32, dirty=d32)
    s = sel.cpu().numpy()
    pidx = pointing.numpy()
    for c in range(len(chan)):
        u = uvw[:, 0] * ls[c]
        v = uvw[:, 1] * ls[c]
        w = uvw[:, 2] * ls[c]
        ref_d, st = o.aw_grid_f64(gx, gx, du, g.info.dv, atlas, grads, u, v, w, s[:, c],
                                  (vis[:, c] * wt[:, c]).astype(np.complex64), pidx)
        ref_p, _ = o.aw_grid_f64(gx, gx, du, g.info.dv, atlas, grads, u, v, w, s[:, c],
                                 wt[:, c].astype(np.complex64), pidx)
        assert (st == 0).sum() > 0.8 * n_rows
        for got, ref, bar in ((dirty[c], ref_d, 1e-12), (psf[c], ref_p, 1e-12), (d32[c], ref_d, 1e-5),
                              (p32[c], ref_p, 1e-5)):
            got = got.cpu().numpy().astype(np.complex128)
            assert np.max(np.abs(got - ref)) <= bar * np.max(np.abs(ref)), (c, bar)
