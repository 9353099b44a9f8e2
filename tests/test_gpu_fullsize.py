"""The benchmark workloads at full size against the CPU FP64 oracle.

Each test builds the workload exactly as bench.py does (bench.setup_workload:
same seed, array, sky, flags, weights, W kernels or AW atlas), runs the
batched FP32 path over every channel -- the dirty products and the fused
major cycle that bench.py times -- and checks seeded channels (planes of the
mosaic) against the FP64 oracle on the same inputs (complex128 grids, FP64
products, FP64 model grids and transforms; bench.oracle_fp64_check).

Bar: the metric's 1e-5 of the FP64 image peak for the dirty image, the PSF
and the major-cycle residual.

* cube256 (config 3): 631,800 rows x 256 channels, Briggs weights, 1250^2
  grids, 32 W planes.
* w4096 (config 2): 5.05 M rows x 16 channels, 4096^2 grid, 128 W planes
  (supports 4..7).
* w4096wide: the same data over a 2.4x wider field, the automatic plane
  count (1456 planes) and supports growing from 4 to 31 (63 px): the
  large-support regime of the binned kernel at full size.
* mosaic7 (config 4): 7 pointings x 16 channels x 4 correlations, Airy x W
  CF cells, 2048^2 grid.
* A nearly converged model on cube16: the data are the FP64 prediction of
  the model plus 1 % of the sky, so that the residual refresh subtracts
  nearly equal quantities (the regime of the late major cycles).  The FP32
  prediction has the rounding of the full data (~1e-7 of |V|), so relative to
  a residual 100x smaller than the data the floor is ~1e-5; bar 2e-5 of the
  residual's own peak (2e-7 of the dirty-image peak).
"""
import numpy as np
import pytest
import torch

import bench

pytestmark = pytest.mark.gpu

SEED = 2026


def _run(pkg, name, log=print):
    ws = bench.setup_workload(pkg, torch, name, SEED, torch.device("cuda", 0), log=log)
    g, data = ws["g"], ws["data"]
    vs = g.visibility_set(ws["batch"])
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])
    psf, dirty = g.new_grids(data["n_chan"]), g.new_grids(data["n_chan"])
    vs.grid(pkg.MODE_DIRTY, psf=psf, dirty=dirty)
    p32, d32, peak32 = g.finish_products(psf, dirty, nsum)
    del psf, dirty
    r32 = vs.major_cycle(data["model"], nsum, peak32)
    torch.cuda.synchronize()
    return ws, vs, nsum, p32, d32, peak32, r32


def _pick(n_chan, k):
    return np.sort(np.random.default_rng(SEED + 1).choice(n_chan, k, replace=False))


@pytest.mark.parametrize("name,k", [("cube256", 2), ("mosaic7", 2), ("w4096", 1), ("w4096wide", 1)])
def test_fullsize_against_fp64_oracle(gpu, name, k):
    print(f"workload {name}", flush=True)
    ws, vs, nsum, p32, d32, peak32, r32 = _run(gpu, name)
    g, data = ws["g"], ws["data"]
    st = vs.stats()
    assert st["n_valid"] > 0.85 * ws["n_unflagged"]
    sup = ws["sup"]
    if name == "w4096":
        assert g.info.planes == 128 and sup.min() == 4 and sup.max() >= 7, (sup.min(), sup.max())
    if name == "w4096wide":
        assert g.info.planes > 1000 and sup.min() == 4 and 30 <= sup.max() <= 32, (sup.min(), sup.max())
    out = bench.oracle_fp64_check(g, data, ws["mosaic"], _pick(data["n_chan"], k), p32, d32, r32)
    for c, n in zip(out["channels"], out["gridded"]):
        assert n == vs.sums[c]["gridded"]
    assert out["dirty_err"] <= 1e-5, out
    assert out["psf_err"] <= 1e-5, out
    assert out["residual_err"] <= 1e-5, out


def test_nearly_converged_model_against_fp64_oracle(gpu):
    ws = bench.setup_workload(gpu, torch, "cube16", SEED, torch.device("cuda", 0))
    g, data = ws["g"], ws["data"]
    # data = the FP64 prediction of the model (plain FP64 degrid of the FP64
    # model grid) + 1 % of the synthetic sky
    mgrid = g.model_to_grid(data["model"], gpu.F64)
    pred = g.degrid_plain(ws["batch"], mgrid, gpu.F64).view(data["n_rows"], data["n_chan"])
    del mgrid
    vis = (pred + 0.01 * data["vis"].to(torch.complex128)).to(torch.complex64)
    data = dict(data, vis=vis)
    batch = gpu.RowChannelBatch(data["uvw"], data["scale"], vis, data["weight"], data["flag"])
    vs = g.visibility_set(batch)
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])
    psf, dirty = g.new_grids(data["n_chan"]), g.new_grids(data["n_chan"])
    vs.grid(gpu.MODE_DIRTY, psf=psf, dirty=dirty)
    p32, d32, peak32 = g.finish_products(psf, dirty, nsum)
    r32 = vs.major_cycle(data["model"], nsum, peak32)
    torch.cuda.synchronize()
    # the residual is about 1 % of the dirty image
    ratio = float(torch.max(torch.abs(r32)) / torch.max(torch.abs(d32)))
    assert ratio < 0.05, ratio
    out = bench.oracle_fp64_check(g, data, False, _pick(data["n_chan"], 2), p32, d32, r32)
    assert out["residual_err"] <= 2e-5, out
    assert out["residual_err"] * ratio <= 1e-6, out
    assert out["dirty_err"] <= 1e-5, out
