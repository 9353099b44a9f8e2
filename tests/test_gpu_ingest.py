"""Row-stream ingestion on the GPU (chip_ingest_*) against the oracle's
co_ingest_rows, and the ingested stream gridded like the same samples
handed over as a RowChannelBatch.

Bar: bit-exact for uvw, weights, sumwt factors and exclusion flags, and for
visibilities without a rotation.  With a phase-centre rotation the phasor's
FP64 cos/sin may differ from the host library in the last bit before
narrowing to Float (as in test_gpu_rotate.py), so rotated visibilities agree
within two Float ulps and almost all bit for bit.
"""
import time

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rows(seed, n, n_corr=4, n_chan=8, spectrum=False):
    rng = np.random.default_rng(seed)
    uvw = rng.uniform(-2500.0, 2500.0, (n, 3))
    data = (rng.normal(size=(n, n_corr, n_chan)) + 1j * rng.normal(size=(n, n_corr, n_chan))).astype(np.complex64)
    flag = rng.uniform(size=(n, n_corr, n_chan)) < 0.08
    wshape = (n, n_corr, n_chan) if spectrum else (n, n_corr)
    weight = rng.uniform(0.5, 2.0, wshape).astype(np.float32)
    # edge cases: non-finite visibilities and weights, zero weights, flagged rows
    data.reshape(-1)[rng.integers(0, data.size, 7)] = np.nan
    weight.reshape(-1)[rng.integers(0, weight.size, 5)] = 0.0
    weight.reshape(-1)[rng.integers(0, weight.size, 3)] = np.inf
    row_flag = rng.uniform(size=n) < 0.05
    return uvw, data, flag, weight, row_flag


def _push_blocks(ing, arrays, sizes):
    uvw, data, flag, weight, row_flag = arrays
    r = 0
    for s in sizes:
        ing.push(uvw[r:r + s], data[r:r + s], flag[r:r + s], weight[r:r + s], row_flag[r:r + s])
        r += s
    assert r == uvw.shape[0]


def _assert_same(got, want, rotated):
    assert np.array_equal(got["uvw_m"], want["uvw_m"])
    assert np.array_equal(got["flag"], want["flag"])
    assert np.array_equal(got["weight"].view(np.uint32), want["weight"].view(np.uint32))
    assert np.array_equal(got["sumwt_factor"], want["sumwt_factor"])
    g, w = got["visibility"], want["visibility"]
    fin = np.isfinite(w)
    assert np.array_equal(np.isfinite(g), fin)
    if not rotated:
        assert np.array_equal(g[fin].view(np.float32), w[fin].view(np.float32))
    else:
        ulp = np.spacing(np.abs(w[fin]).astype(np.float32))
        assert np.all(np.abs(g[fin] - w[fin]) <= 2 * ulp)
        assert np.mean(g[fin] == w[fin]) > 0.999


@pytest.mark.parametrize("pair,spectrum,rotate", [(None, False, False), ("half_sum", True, False),
                                                   ("half_difference", False, True),
                                                   ("positive_half_imag_difference", True, True),
                                                   ("negative_half_imag_difference", False, False)])
def test_ingest_matches_oracle(gpu, oracle_mod, pair, spectrum, rotate):
    n, n_corr, n_chan = 1500, 4, 8
    arrays = _rows(7 + n_chan, n, n_corr, n_chan, spectrum)
    freq = np.linspace(1.0e9, 1.5e9, n_chan)
    rot = gpu.phase_rotation((1.2, 0.52), (1.2 + 3e-3, 0.52 - 2e-3)) if rotate else None
    corr = 1 if pair is None else (0, 3)
    ing = gpu.RowIngest(n_corr, freq, corr=corr, pair=pair, sumwt_factor=2.0, weight_spectrum=spectrum,
                        rotation=rot, max_rows=n, block_rows=400)
    # five blocks through two staging buffers: ragged sizes, a one-row block
    _push_blocks(ing, arrays, [400, 1, 399, 300, 400])
    uvw, data, flag, weight, row_flag = arrays
    want = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=corr, pair=pair, sumwt_factor=2.0,
                                  row_flag=row_flag, rotation=rot)
    _assert_same(ing.download(), want, rotate)


def test_ingest_capacity_and_block_limits(gpu):
    n_corr, n_chan = 2, 3
    uvw, data, flag, weight, row_flag = _rows(3, 12, n_corr, n_chan)
    ing = gpu.RowIngest(n_corr, np.linspace(1e9, 1.1e9, n_chan), corr=1, max_rows=10, block_rows=6)
    ing.push(uvw[:6], data[:6], flag[:6], weight[:6])
    with pytest.raises(gpu.InvalidRequest):
        ing.push(uvw[:7], data[:7], flag[:7], weight[:7])   # larger than a block
    with pytest.raises(gpu.InvalidRequest):
        ing.push(uvw[:5], data[:5], flag[:5], weight[:5])   # beyond max_rows: nothing appended
    ing.push(uvw[6:10], data[6:10], flag[6:10], weight[6:10])
    got = ing.download()
    assert got["uvw_m"].shape == (10, 3) and np.array_equal(got["uvw_m"], uvw[:10])
    with pytest.raises(gpu.InvalidRequest):
        gpu.RowIngest(n_corr, [1e9], corr=2, max_rows=4, block_rows=2)   # corr outside the rows
    with pytest.raises(gpu.InvalidRequest):
        gpu.RowIngest(n_corr, [1e9], corr=(0, 1), pair="sum", max_rows=4, block_rows=2)


def test_ingested_stream_grids_like_a_batch(gpu, oracle_mod):
    """The ingested stream as a batch: dirty and PSF grids and sums equal to
    those of the oracle's converted arrays handed over as a RowChannelBatch."""
    import synth
    enu = synth.y_array(3, arm_m=2000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 0.5, 60.0)
    n, n_corr, n_chan = uvw.shape[0], 2, 5
    _, data, flag, weight, row_flag = _rows(11, n, n_corr, n_chan)
    freq = np.linspace(1.0e9, 1.6e9, n_chan)
    scale = freq / synth.C
    maxuv = np.max(np.abs(uvw[:, :2])) * scale[-1]
    cell = 1.0 / (3.0 * maxuv)
    g = gpu.StandardGridder(96, 96, cell, cell)
    ing = gpu.RowIngest(n_corr, freq, corr=(0, 1), pair="half_sum", max_rows=n, block_rows=(n + 2) // 3)
    b = (n + 2) // 3
    _push_blocks(ing, (uvw, data, flag, weight, row_flag), [b, b, n - 2 * b])
    want = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 1), pair="half_sum", row_flag=row_flag)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a))
    batch = gpu.RowChannelBatch(t(want["uvw_m"]), t(scale), t(want["visibility"]), t(want["weight"]),
                                t(want["flag"].astype(np.uint8)), sumwt_factor=t(want["sumwt_factor"]))
    out = []
    for src in (ing, batch):
        vs = g.visibility_set(src)
        psf, dirty = g.new_grids(n_chan), g.new_grids(n_chan)
        vs.grid(gpu.MODE_DIRTY, psf=psf, dirty=dirty)
        out.append((psf.cpu().numpy(), dirty.cpu().numpy(), vs.sums))
    (p0, d0, s0), (p1, d1, s1) = out
    assert s0 == s1
    assert np.max(np.abs(p0 - p1)) <= 1e-6 * np.max(np.abs(p1))
    assert np.max(np.abs(d0 - d1)) <= 1e-6 * np.max(np.abs(d1))


def test_ingest_overlaps_blocks(gpu):
    """Throughput of pushes through the two staging buffers (PCIe included),
    printed; every block lands in order."""
    n_corr, n_chan, blk, k = 4, 64, 8192, 12
    uvw, data, flag, weight, row_flag = _rows(5, blk, n_corr, n_chan)
    ing = gpu.RowIngest(n_corr, np.linspace(1e9, 2e9, n_chan), corr=(0, 3), pair="half_sum",
                        max_rows=blk * k, block_rows=blk)
    t0 = time.perf_counter()
    for i in range(k):
        u = uvw.copy()
        u[:, 0] += i  # tag the block
        ing.push(u, data, flag, weight, row_flag)
    ing.to_samples()  # waits for every block
    dt = time.perf_counter() - t0
    mb = k * (uvw.nbytes + data.nbytes + flag.nbytes + weight.nbytes + row_flag.nbytes) / 1e6
    print(f"ingest: {k} blocks of {blk} rows x {n_corr} corr x {n_chan} chan, {mb:.0f} MB in {dt * 1e3:.1f} ms "
          f"({mb / dt / 1e3:.2f} GB/s host -> resident samples)")
    got = ing.download()
    for i in range(k):
        assert np.array_equal(got["uvw_m"][i * blk:(i + 1) * blk, 0], uvw[:, 0] + i)
