"""chip_grid_plain_noalloc (no allocation; sums in the library's fixed device
scratch) against chip_grid_plain on the same inputs, through the C ABI.

Bar (as for the plain path): gridded / skipped counts exact, FP64 grids and
sums within 1e-12 relative (only the order of the FP64 atomic adds differs),
FP32 grids within 1e-6 relative.  Two calls in a row, the second with fewer
rows and channels, so that a scratch left uncleared by the first call would
show in the second's sums."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rows(pkg, seed, n_rows, n_chan, umax):
    rng = np.random.default_rng(seed)
    uvw = np.stack([rng.uniform(-umax, umax, n_rows), rng.uniform(-umax, umax, n_rows),
                    rng.uniform(-200.0, 200.0, n_rows)], axis=1)
    freq = np.linspace(1.0e9, 1.3e9, n_chan)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    wt = rng.uniform(0.2, 2.0, (n_rows, n_chan)).astype(np.float32)
    flag = rng.uniform(size=(n_rows, n_chan)) < 0.1
    return pkg.RowChannelBatch(torch.from_numpy(uvw), torch.from_numpy(pkg.RowChannelBatch.lambda_scales(freq)),
                               torch.from_numpy(vis), torch.from_numpy(wt), torch.from_numpy(flag))


def _grid(pkg, g, batch, precision, name, mode):
    s, keep = batch.to_samples()
    psf = g.new_grids(batch.n_chan, precision)
    dirty = g.new_grids(batch.n_chan, precision)
    sums = (pkg.abi.Sums * batch.n_chan)()
    pkg.abi.call(name, g.handle, ctypes.byref(s), mode, precision, None,
                 ctypes.c_void_p(psf.data_ptr()), ctypes.c_void_p(dirty.data_ptr()), sums)
    torch.cuda.synchronize()
    return psf.cpu().numpy(), dirty.cpu().numpy(), [x.as_dict() for x in sums]


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


@pytest.mark.parametrize("wproject", [False, True])
def test_noalloc_matches_plain_twice(gpu, wproject):
    cell = np.deg2rad(20.0 / 3600.0)
    g = gpu.StandardGridder(96, 96, cell, cell)
    if wproject:
        g.attach_wproject(200.0 * 1.3e9 / 299_792_458.0, planes=8)
    umax = 0.4 * g.info.gx * g.info.du * 299_792_458.0 / 1.3e9
    for seed, n_rows, n_chan in ((1, 4000, 6), (2, 700, 3)):
        batch = _rows(gpu, seed, n_rows, n_chan, umax)
        for precision, bar in ((gpu.F64, 1e-12), (gpu.F32, 1e-6)):
            a = _grid(gpu, g, batch, precision, "chip_grid_plain", gpu.MODE_DIRTY)
            b = _grid(gpu, g, batch, precision, "chip_grid_plain_noalloc", gpu.MODE_DIRTY)
            assert _rel(b[0], a[0]) <= bar and _rel(b[1], a[1]) <= bar
            for sa, sb in zip(a[2], b[2]):
                assert (sa["gridded"], sa["skipped"]) == (sb["gridded"], sb["skipped"])
                assert sa["gridded"] > 0
                for k in ("normalization_sumwt", "reported_sumwt"):
                    assert abs(sa[k] - sb[k]) <= 1e-12 * abs(sa[k])


def test_noalloc_rejects_too_many_channels(gpu):
    cell = np.deg2rad(20.0 / 3600.0)
    g = gpu.StandardGridder(64, 64, cell, cell)
    n_chan = 4097  # CHIP_NOALLOC_MAX_CHAN + 1
    batch = _rows(gpu, 3, 2, n_chan, 100.0)
    with pytest.raises(gpu.InvalidRequest, match="at most 4096 channels"):
        _grid(gpu, g, batch, gpu.F64, "chip_grid_plain_noalloc", gpu.MODE_PSF)
