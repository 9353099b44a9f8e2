#!/usr/bin/env python3
"""Throughput of the W-projection grid+degrid major cycle on one MI355X.

Workload (default ``cube256``, the batched spectral cube): 27 antennas in a Y
with 21 km arms (seeded jitter), latitude 34 deg, phase centre dec +30 deg,
1 h at 2 s (1800 x 351 = 631,800 rows) x 256 channels 1.0-2.0 GHz, per-channel
flag fraction U(0, 0.6), Briggs robust 0 weights from a density grid, 200
Gaussian sources (FWHM 1-10 px, flux log-uniform 10 mJy-1 Jy), 1024^2 image
per channel (1250^2 grid), 32 W planes (explicit, W sampling 4), cell =
1/(3 max|uv|).

One step = one residual refresh of every channel: model cube -> apodise +
forward FFT -> degrid and grid (V - model)*w -> inverse FFT, crop, correct,
normalise -> residual cube.  The batched FP32 path runs every channel in one
binned launch over a resident, pre-sorted visibility stream; the baseline is
the plain per-channel FP64 path (one launch per channel, FP64 MACs and
atomics, FP64 FFTs) on the same GPU.  ``value`` is the batched FP32 Mvis/s;
the FP64 rate, the ratio and the dirty/residual image errors between the two
(relative to the FP64 image peak) are reported beside it.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

C = 299_792_458.0
HBM_PEAK_GBS = 8000.0
PASS_NAMES = ("k_fwd_rows", "k_fwd_cols", "k_zero_touched", "k_binned", "k_inv_cols", "k_inv_rows")
FP32_PEAK_TFLOPS = 157.3

WORKLOADS = {
    "cube256": dict(hours=1.0, dt=2.0, n_chan=256, f0=1.0e9, f1=2.0e9, nx=1024, planes=32,
                    sampling=None, n_src=200, flag_max=0.6, arm_m=21_000.0),
    # reduced cube for quick checks (same geometry, fewer rows and channels)
    "cube16": dict(hours=0.25, dt=2.0, n_chan=16, f0=1.0e9, f1=2.0e9, nx=1024, planes=32,
                   sampling=None, n_src=200, flag_max=0.6, arm_m=21_000.0),
    # W-projection drop-in at size: same array, 8 h at 2 s (5.05 M rows) x 16
    # channels 1.0-2.0 GHz (80.9 M visibilities), 3413^2 image on a 4096^2
    # grid at cell 1/(3 max|uv|), 128 explicit W planes (spanning w up to
    # 0.25/cell, CASA's rule for an explicit plane count, which here is the
    # data's max|w|), W oversampling 8, natural weights, no flags; supports 4..7
    "w4096": dict(hours=8.0, dt=2.0, n_chan=16, f0=1.0e9, f1=2.0e9, nx=3413, planes=128,
                  sampling=8, n_src=200, flag_max=0.0, arm_m=21_000.0, weighting="natural"),
    # the same data imaged over a 2.4x wider field (cell 1/(0.93 max|uv|)),
    # where the W term is strong: the reference's automatic plane count
    # (floor(1.05 max|w| sin(half field)) = 1456 planes over the data's w
    # range) and supports growing from 4 to 31 (9 to 63 px); the ~10 % of
    # samples whose footprint would leave the grid are skipped by the planner
    "w4096wide": dict(hours=8.0, dt=2.0, n_chan=16, f0=1.0e9, f1=2.0e9, nx=3413, planes=None,
                      sampling=8, n_src=200, flag_max=0.0, arm_m=21_000.0, weighting="natural", cell_k=0.93),
    # AW-projection mosaic: 7 pointings on a hexagon at half-power spacing,
    # RR/RL/LR/LL, Airy A-term (25 m dish, 2 m blockage), 6 h at 10 s,
    # 16 channels 1.0-1.5 GHz, 2048^2 grid (1707^2 image), 64 W planes,
    # oversampling 8, 100 Gaussian sources, per-pointing flags U(0, 0.4); CF
    # cells at 8 frequencies (one per channel pair), nearest per sample
    "mosaic7": dict(kind="mosaic", hours=6.0, dt=10.0, n_chan=16, n_corr=4, f0=1.0e9, f1=1.5e9, nx=1707,
                    planes=64, sampling=8, n_src=100, flag_max=0.4, arm_m=2_500.0, pointings=7,
                    dish=25.0, blockage=2.0, cf_freqs=8),
    # reduced mosaic for quick checks (same geometry, fewer rows)
    "mosaic7s": dict(kind="mosaic", hours=1.0, dt=10.0, n_chan=4, n_corr=4, f0=1.0e9, f1=1.5e9, nx=1707,
                     planes=64, sampling=8, n_src=100, flag_max=0.4, arm_m=2_500.0, pointings=7,
                     dish=25.0, blockage=2.0, cf_freqs=2),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def y_array(rng, n_ant=27, arm_m=21_000.0, jitter_m=50.0):
    per_arm = n_ant // 3
    frac = (np.arange(1, per_arm + 1) / per_arm) ** 1.716
    pos = []
    for arm in range(3):
        az = np.deg2rad(5.0 + 120.0 * arm)
        for f in frac:
            r = arm_m * f
            pos.append([r * np.sin(az) + rng.normal(0, jitter_m), r * np.cos(az) + rng.normal(0, jitter_m),
                        rng.normal(0, 2.0)])
    return np.asarray(pos)


def uvw_tracks(enu, lat_deg, dec_deg, hours, dt_s):
    lat, dec = np.deg2rad(lat_deg), np.deg2rad(dec_deg)
    e, nn, up = enu[:, 0], enu[:, 1], enu[:, 2]
    ant = np.stack([-np.sin(lat) * nn + np.cos(lat) * up, e, np.cos(lat) * nn + np.sin(lat) * up], 1)
    i, j = np.triu_indices(len(enu), 1)
    b = ant[j] - ant[i]
    n_t = int(round(hours * 3600.0 / dt_s))
    ha = np.deg2rad(15.0 * (np.arange(n_t) * dt_s / 3600.0 - hours / 2.0))
    sh, ch = np.sin(ha)[:, None], np.cos(ha)[:, None]
    sd, cd = np.sin(dec), np.cos(dec)
    bx, by, bz = b[None, :, 0], b[None, :, 1], b[None, :, 2]
    u = sh * bx + ch * by
    v = -sd * ch * bx + sd * sh * by + cd * bz
    w = cd * ch * bx - cd * sh * by + sd * bz
    return np.stack([u, v, w], -1).reshape(-1, 3)


def make_cube(torch, wl, seed, dev):
    """Synthetic cube on the GPU (data generation is not timed)."""
    rng = np.random.default_rng(seed)
    enu = y_array(rng, arm_m=wl["arm_m"])
    uvw = uvw_tracks(enu, 34.0, 30.0, wl["hours"], wl["dt"])
    n_rows, n_chan = uvw.shape[0], wl["n_chan"]
    freq = np.linspace(wl["f0"], wl["f1"], n_chan)
    scale = freq / C
    max_uv = float(np.max(np.hypot(uvw[:, 0], uvw[:, 1])) * scale[-1])
    cell = 1.0 / (wl.get("cell_k", 3.0) * max_uv)
    nx = wl["nx"]
    # sources: Gaussians inside the inner 80% of the field
    ns = wl["n_src"]
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), ns))
    lpix = rng.uniform(-0.4 * nx, 0.4 * nx, ns)
    mpix = rng.uniform(-0.4 * nx, 0.4 * nx, ns)
    fwhm = rng.uniform(1.0, 10.0, ns)
    sig = fwhm / (2.0 * np.sqrt(2.0 * np.log(2.0))) * cell
    l, m = lpix * cell, mpix * cell
    n1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    U = torch.tensor(uvw, dtype=torch.float64, device=dev)
    vis = torch.empty((n_rows, n_chan), dtype=torch.complex64, device=dev)
    L_ = torch.tensor(l, dtype=torch.float64, device=dev)
    M_ = torch.tensor(m, dtype=torch.float64, device=dev)
    N_ = torch.tensor(n1, dtype=torch.float64, device=dev)
    S_ = torch.tensor(flux, dtype=torch.float64, device=dev)
    G_ = torch.tensor(2.0 * np.pi ** 2 * sig ** 2, dtype=torch.float64, device=dev)
    for c in range(n_chan):
        u = U[:, 0:1] * scale[c]
        v = U[:, 1:2] * scale[c]
        w = U[:, 2:3] * scale[c]
        # w enters with the sign the W projector images (WProjector and
        # AwProjector apply e^{-2 pi i w (n-1)}, the opposite of the
        # reference's direct mode), so that the sky is in focus
        ph = -2.0 * np.pi * (u * L_ + v * M_ - w * N_)
        amp = S_ * torch.exp(-G_ * (u * u + v * v))
        vis[:, c] = torch.complex((amp * torch.cos(ph)).sum(1), (amp * torch.sin(ph)).sum(1)).to(torch.complex64)
    frac = rng.uniform(0.0, wl["flag_max"], n_chan)
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    flag = torch.rand((n_rows, n_chan), generator=gen, device=dev) < torch.tensor(frac, device=dev)[None, :]
    weight = torch.ones((n_rows, n_chan), dtype=torch.float32, device=dev)  # natural; Briggs below
    model = torch.zeros((n_chan, nx, nx), dtype=torch.float32, device=dev)
    # CLEAN-like point components at the source positions (x = l/cell + nx/2, y = ny/2 - m/cell)
    xs = np.clip(np.round(lpix + nx / 2).astype(int), 0, nx - 1)
    ys = np.clip(np.round(nx / 2 - mpix).astype(int), 0, nx - 1)
    model[:, xs, ys] = torch.tensor(0.5 * flux, dtype=torch.float32, device=dev)[None, :]
    return dict(uvw=U, scale=torch.tensor(scale, dtype=torch.float64, device=dev), vis=vis, flag=flag,
                weight=weight, model=model, cell=cell, nx=nx, n_rows=n_rows, n_chan=n_chan,
                freq=freq, max_uv=max_uv, max_w=float(np.max(np.abs(uvw[:, 2])) * scale[-1]))


def airy_power(torch, r, freq, dish, blockage):
    """Airy power pattern for data synthesis, in the form of the voltage table
    the A-term uses: V = (a^2 jinc(x) - jinc(a x)) / (a^2 - 1), a = dish /
    blockage, x = pi dish sin(r) / lambda, jinc(z) = 2 J1(z) / z."""
    a = dish / blockage
    x = torch.clamp(torch.pi * dish * torch.sin(r) * freq / C, min=1e-9)

    def jinc(z):
        return 2.0 * torch.special.bessel_j1(z) / z

    v = (a * a * jinc(x) - jinc(x * a)) / (a * a - 1.0)
    return v * v


def make_mosaic(torch, wl, seed, dev):
    """7-pointing AW mosaic on the GPU: rows cycle through the pointings in
    time; every row carries 16 channels x 4 correlations as 64 planes
    (plane = chan * 4 + corr).  Visibilities of a pointing are the sky
    attenuated by that pointing's Airy power pattern, phase-referenced to the
    common phase centre."""
    rng = np.random.default_rng(seed)
    enu = y_array(rng, arm_m=wl["arm_m"])
    uvw = uvw_tracks(enu, 34.0, 30.0, wl["hours"], wl["dt"])
    n_base = len(enu) * (len(enu) - 1) // 2
    n_rows = uvw.shape[0]
    n_t = n_rows // n_base
    nc, ncorr = wl["n_chan"], wl["n_corr"]
    freq = np.linspace(wl["f0"], wl["f1"], nc)
    plane_freq = np.repeat(freq, ncorr)
    scale = plane_freq / C
    max_uv = float(np.max(np.hypot(uvw[:, 0], uvw[:, 1])) * freq[-1] / C)
    cell = 1.0 / (3.0 * max_uv)
    nx = wl["nx"]
    # pointings: centre + hexagon at half the power-beam FWHM of the band centre
    fwhm = 1.02 * (C / (0.5 * (wl["f0"] + wl["f1"]))) / wl["dish"]
    sep = 0.5 * fwhm
    pl = [(0.0, 0.0)] + [(sep * np.cos(np.pi / 3 * k), sep * np.sin(np.pi / 3 * k)) for k in range(6)]
    pl = np.asarray(pl[: wl["pointings"]])
    pointing_of_t = np.arange(n_t) % len(pl)
    pointing = np.repeat(pointing_of_t, n_base).astype(np.int32)
    ns = wl["n_src"]
    flux = np.exp(rng.uniform(np.log(0.01), np.log(1.0), ns))
    half = 0.4 * nx * cell
    l = rng.uniform(-half, half, ns)
    m = rng.uniform(-half, half, ns)
    sfw = rng.uniform(1.0, 10.0, ns) / (2.0 * np.sqrt(2.0 * np.log(2.0))) * cell
    n1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    U = torch.tensor(uvw, dtype=torch.float64, device=dev)
    P = torch.tensor(pointing, dtype=torch.int64, device=dev)
    L_ = torch.tensor(l, dtype=torch.float64, device=dev)
    M_ = torch.tensor(m, dtype=torch.float64, device=dev)
    N_ = torch.tensor(n1, dtype=torch.float64, device=dev)
    S_ = torch.tensor(flux, dtype=torch.float64, device=dev)
    G_ = torch.tensor(2.0 * np.pi ** 2 * sfw ** 2, dtype=torch.float64, device=dev)
    PL = torch.tensor(pl, dtype=torch.float64, device=dev)
    # distance of every source from every pointing centre
    dr = torch.sqrt((L_[None, :] - PL[:, 0:1]) ** 2 + (M_[None, :] - PL[:, 1:2]) ** 2)
    pol = torch.tensor(0.05 * np.exp(1j * rng.uniform(0, 2 * np.pi, ns)), dtype=torch.complex128, device=dev)
    vis = torch.empty((n_rows, nc * ncorr), dtype=torch.complex64, device=dev)
    for c in range(nc):
        f = freq[c]
        pb = airy_power(torch, dr, f, wl["dish"], wl["blockage"])  # (pointings, sources)
        u = U[:, 0:1] * (f / C)
        v = U[:, 1:2] * (f / C)
        w = U[:, 2:3] * (f / C)
        ph = -2.0 * np.pi * (u * L_ + v * M_ - w * N_)  # the W projector's w sign, as make_cube
        amp = S_ * torch.exp(-G_ * (u * u + v * v)) * pb[P]
        e = torch.complex(torch.cos(ph), torch.sin(ph))
        par = (amp * e).sum(1)
        crs = (amp * e * pol).sum(1)
        vis[:, 4 * c + 0] = par.to(torch.complex64)
        vis[:, 4 * c + 1] = crs.to(torch.complex64)
        vis[:, 4 * c + 2] = crs.conj().to(torch.complex64)
        vis[:, 4 * c + 3] = par.to(torch.complex64)
    frac = rng.uniform(0.0, wl["flag_max"], len(pl))
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    flag = torch.rand((n_rows, nc * ncorr), generator=gen, device=dev) < \
        torch.tensor(frac, device=dev)[P][:, None]
    weight = torch.ones((n_rows, nc * ncorr), dtype=torch.float32, device=dev)
    model = torch.zeros((nc * ncorr, nx, nx), dtype=torch.float32, device=dev)
    xs = np.clip(np.round(l / cell + nx / 2).astype(int), 0, nx - 1)
    ys = np.clip(np.round(nx / 2 - m / cell).astype(int), 0, nx - 1)
    for k in range(nc * ncorr):
        if k % 4 in (0, 3):
            model[k, xs, ys] = torch.tensor(0.5 * flux, dtype=torch.float32, device=dev)
    return dict(uvw=U, scale=torch.tensor(scale, dtype=torch.float64, device=dev), vis=vis, flag=flag,
                weight=weight, model=model, cell=cell, nx=nx, n_rows=n_rows, n_chan=nc * ncorr,
                freq=plane_freq, max_uv=max_uv, max_w=float(np.max(np.abs(uvw[:, 2])) * freq[-1] / C),
                pointing=torch.tensor(pointing, dtype=torch.int32, device=dev), pointings=pl)


def native_oracle(log):
    """Route the oracle through a -march=native build for this host (the
    portable build when the compiler is not there); returns (build, cpu)."""
    import tempfile

    import oracle

    try:
        oracle.use_native_build(tempfile.mkdtemp(prefix="oracle_native_"))
        build = "gcc -O3 -march=native -ffp-contract=off"
    except Exception as e:  # the baseline still runs, on the portable build
        log(f"native oracle build failed ({e!r}); timing the portable build")
        build = "gcc -O3 -march=x86-64-v2 -ffp-contract=off"
    return build, oracle.cpu_model()


def cpu_baseline_aw(pkg, g, data, cpu_seconds, log):
    """Oracle AW residual refresh (degrid from a complex64 model grid, then
    grid (V - p) w into a Complex64 grid, sequential as the reference's
    AwProjector, one core) on a bounded sample of rows of one plane."""
    import oracle

    build, cpu_name = native_oracle(log)
    cells = [(c.cpu().numpy(), xs, ys) for c, xs, ys in g.aw_cells()]
    atlas = oracle.AwAtlas(cells, g.info.sampling)
    grads = data["gradients"]
    uvw = data["uvw"].cpu().numpy()
    scale = data["scale"].cpu().numpy()
    pointing = data["pointing"].cpu().numpy()
    gx, gy, du, dv = g.info.gx, g.info.gy, g.info.du, g.info.dv
    w_incr = g.info.w_scale
    c = 0
    model = g.model_to_grid(data["model"][c])[0].cpu().numpy()
    keep = ~data["flag"][:, c].cpu().numpy()
    rows = np.nonzero(keep)[0]
    rng = np.random.default_rng(1)
    rng.shuffle(rows)
    nvis, spent, chunk = 0, 0.0, 2000
    grid = None
    for k in range(0, len(rows), chunk):
        r = np.sort(rows[k:k + chunk])
        u, v, w = (uvw[r, j] * scale[c] for j in range(3))
        cell_of = data["cf_cell"][r, c].cpu().numpy().astype(np.uint32)
        vis = data["vis"][r, c].cpu().numpy()
        wt = data["weight"][r, c].cpu().numpy()
        t0 = time.perf_counter()
        pred, _ = oracle.aw_degrid(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, model, pointing[r])
        vals = ((vis - pred) * wt).astype(np.complex64)
        gk, _ = oracle.aw_grid_f64(gx, gy, du, dv, atlas, grads, u, v, w, cell_of, vals, pointing[r])
        spent += time.perf_counter() - t0
        grid = gk if grid is None else grid
        nvis += len(r)
        if spent >= cpu_seconds:
            break
    log(f"cpu baseline (AW): {nvis} vis in {spent:.2f} s on 1 thread ({cpu_name}, {build})")
    return {"value": nvis / spent / 1e6, "unit": "Mvis/s", "cores": 1, "kind": "port", "cpu_model": cpu_name,
            "build": build,
            "sample": f"oracle AWProject residual refresh (degrid + grid into Complex64 grids, sequential as "
                      f"the reference's AwProjector), {nvis} unflagged visibilities of plane 0, FFTs excluded"}


def cpu_baseline(pkg, g, data, cpu_seconds, log):
    """Oracle residual refresh (degrid + grid, FP32 like the reference W path)
    on host cores for a bounded sample of channels."""
    import torch
    import oracle

    build, cpu_name = native_oracle(log)
    # every host thread of the run, up to 16: the host share of one GPU on
    # the machines this runs on
    threads = min(16, len(os.sched_getaffinity(0)))
    nx, cell = data["nx"], data["cell"]
    og = oracle.StandardGridder(nx, nx, cell, cell)
    sup, ker, nt = g.wproject_tables()
    wp = oracle.WProjector.from_tables(og, g.info.planes, g.info.sampling, g.info.w_scale, ker, sup, nt)
    uvw = data["uvw"].cpu().numpy()
    scale = data["scale"].cpu().numpy()
    order = np.random.default_rng(1).permutation(data["n_chan"])
    chunk = 2_097_152  # rows per call: whole channels of the cube, a few blocks per channel at 5 M rows
    nvis, spent, chans = 0, 0.0, 0
    for c in order:
        u, v, w = (uvw[:, k] * scale[c] for k in range(3))
        vis = data["vis"][:, c].cpu().numpy()
        wt = data["weight"][:, c].cpu().numpy()
        gr = ~data["flag"][:, c].cpu().numpy()
        model = g.model_to_grid(data["model"][c])[0].cpu().numpy()
        chans += 1
        for r0 in range(0, len(u), chunk):
            r = slice(r0, r0 + chunk)
            t0 = time.perf_counter()
            wp.grid(u[r], v[r], w[r], wt[r], np.ones_like(wt[r]), vis[r], gridable=gr[r], mode=2, model=model,
                    threads=threads)
            spent += time.perf_counter() - t0
            nvis += int(gr[r].sum())
            if spent >= cpu_seconds:
                break
        if spent >= cpu_seconds:
            break
    log(f"cpu baseline: {nvis} vis of {chans} channels in {spent:.2f} s on {threads} threads ({cpu_name}, {build})")
    return {"value": nvis / spent / 1e6, "unit": "Mvis/s", "cores": threads, "kind": "port", "cpu_model": cpu_name,
            "build": build, "host_threads": len(os.sched_getaffinity(0)),
            "sample": f"oracle W-projection residual refresh (degrid+grid, FP32 grids as the "
                      f"reference) of {nvis} unflagged visibilities from {chans} of {data['n_chan']} channels "
                      f"(in blocks of {chunk} rows), FFTs excluded"}


def oracle_fp64_check(g, data, mosaic, channels, psf32, dirty32, res32, model=None, log=log, threads=None):
    """The batched FP32 images of the sampled ``channels`` (planes of a mosaic)
    against the FP64 oracle on the same inputs: complex128 grids with FP64
    products (W: WProjector.grid64 over the device's W tables; AW: aw_grid64
    over the device's CF cells and phase gradients), FP64 model grids and FP64
    transforms.  psf32 / dirty32: the FP32 dirty products (n_chan, nx, nx);
    res32: the major-cycle residual of ``model`` (default data["model"]).
    Errors are max |FP32 - FP64| over the image, relative to the FP64 image's
    peak.  Returns a dict of per-channel and maximum errors."""
    import oracle

    t0 = time.perf_counter()
    nx, cell = data["nx"], data["cell"]
    model = data["model"] if model is None else model
    uvw = data["uvw"].cpu().numpy()
    scale = data["scale"].cpu().numpy()
    og = oracle.StandardGridder(nx, nx, cell, cell)
    s = g.info.sampling
    threads = threads or min(16, len(os.sched_getaffinity(0)))
    # private FP64 grids per thread: keep them within a few GB
    threads = max(1, min(threads, int(4e9 // (2 * 16 * g.info.gx * g.info.gy))))
    if mosaic:
        cells = [(c.cpu().numpy(), xs, ys) for c, xs, ys in g.aw_cells()]
        atlas = oracle.AwAtlas(cells, s)
        pointing = data["pointing"].cpu().numpy().astype(np.uint32)
    else:
        sup, ker, nt = g.wproject_tables()
        wp = oracle.WProjector.from_tables(og, g.info.planes, s, g.info.w_scale, ker, sup, nt)
    out = {"channels": [int(c) for c in channels], "dirty": [], "psf": [], "residual": [], "gridded": []}
    for c in channels:
        u, v, w = (uvw[:, k] * scale[c] for k in range(3))
        vis = data["vis"][:, c].cpu().numpy()
        wt = data["weight"][:, c].cpu().numpy()
        usable = ~data["flag"][:, c].cpu().numpy()
        m = model[c].cpu().numpy()
        m64 = oracle.w_model_grid64(og, s, m)
        if mosaic:
            cell_of = data["cf_cell"][:, c].cpu().numpy().astype(np.uint32)
            gd = (g.info.gx, g.info.gy, g.info.du, g.info.dv, atlas, data["gradients"], u, v, w, cell_of)
            p64, d64, nsum, count = oracle.aw_grid64(*gd, vis, wt, mode=1, pointing_of=pointing, usable=usable)
            _, r64, _, _ = oracle.aw_grid64(*gd, vis, wt, mode=2, model=m64, pointing_of=pointing, usable=usable)
        else:
            ones = np.ones_like(wt)
            p64, d64, sm = wp.grid64(u, v, w, wt, ones, vis, gridable=usable, mode=1, threads=threads)
            _, r64, _ = wp.grid64(u, v, w, wt, ones, vis, gridable=usable, mode=2, model=m64, threads=threads)
            nsum, count = sm.normalization_sumwt, sm.gridded
        psf, dirty, peak = oracle.w_dirty_products64(og, s, p64, d64, nsum)
        _, res, _ = oracle.w_dirty_products64(og, s, None, r64, nsum, psf_peak=peak)
        del p64, d64, r64

        def err(a, b):
            return float(np.max(np.abs(np.asarray(a, np.float64) - b)) / np.max(np.abs(b)))

        out["dirty"].append(err(dirty32[c].cpu().numpy(), dirty))
        out["psf"].append(float(np.max(np.abs(psf32[c].cpu().numpy() - psf))))
        out["residual"].append(err(res32[c].cpu().numpy(), res))
        out["gridded"].append(int(count))
    out.update(dirty_err=max(out["dirty"]), psf_err=max(out["psf"]), residual_err=max(out["residual"]),
               seconds=time.perf_counter() - t0, threads=threads)
    log(f"FP64 oracle on {'planes' if mosaic else 'channels'} {out['channels']} ({out['seconds']:.1f} s, "
        f"{threads} threads): dirty {out['dirty_err']:.2e}, residual {out['residual_err']:.2e} of peak, "
        f"PSF {out['psf_err']:.2e}")
    return out


def measured_traffic(workload):
    """HBM bytes per dispatch of each kernel of the step, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes, each counter scaled by its reading for a
    device copy of known size (profiles/step_traffic_<workload>.json, written
    by --reduce-traffic): {kernel: bytes}, or None when no measurement is
    committed."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", f"step_traffic_{workload}.json")
    try:
        with open(path) as f:
            return {k: float(v["bytes_per_dispatch"]) for k, v in json.load(f)["kernels"].items()}
    except (OSError, KeyError, ValueError, TypeError):
        return None


def _profile_csv(directory, suffix):
    import glob
    found = glob.glob(os.path.join(directory, "**", "*" + suffix), recursive=True)
    if not found:
        raise SystemExit(f"no *{suffix} under {directory}")
    import csv
    with open(found[0]) as f:
        return list(csv.DictReader(f))


def reduce_trace(directory, out_txt, command, top=40):
    """rocprofv3 --kernel-trace --stats output -> short per-kernel table
    (calls, mean, median and total device time, share)."""
    rows = _profile_csv(directory, "kernel_stats.csv")
    rows.sort(key=lambda r: -float(r["TotalDurationNs"]))
    med = {}
    try:
        for r in _profile_csv(directory, "kernel_trace.csv"):
            med.setdefault(r["Kernel_Name"], []).append(int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
    except SystemExit:
        pass
    lines = [f"command: {command}", "",
             f"{'calls':>7} {'avg_us':>10} {'median_us':>10} {'total_ms':>10} {'pct':>6}  kernel"]
    for r in rows[:top]:
        d = med.get(r["Name"])
        m = f"{float(np.median(d)) / 1e3:10.2f}" if d else f"{'-':>10}"
        lines.append(f"{int(r['Calls']):7d} {float(r['AverageNs']) / 1e3:10.2f} {m} "
                     f"{float(r['TotalDurationNs']) / 1e6:10.2f} {float(r['Percentage']):6.2f}  {r['Name'][:140]}")
    with open(out_txt, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def reduce_traffic(fetch_dir, write_dir, out_json, out_txt, commands):
    """Two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE, KiB per dispatch) ->
    HBM bytes per dispatch of every kernel of the step, each counter scaled by
    its reading for the 1 GiB device copies that --calibrate-copy runs."""
    res, lines = {}, [f"command: {c}" for c in commands] + [""]
    scales = {}
    per_counter = {}
    for name, d, counter in (("fetch", fetch_dir, "FETCH_SIZE"), ("write", write_dir, "WRITE_SIZE")):
        per = {}
        for r in _profile_csv(d, "counter_collection.csv"):
            if r["Counter_Name"] == counter:
                per.setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
        copies = sorted(v for k, vals in per.items() if "copy" in k.lower() for v in vals)[-3:]
        if not copies:
            raise SystemExit("no copy kernel dispatches: run the profiled bench with --calibrate-copy")
        cal = sum(copies) / len(copies)
        scales[name] = 1024.0 * 1024.0 / cal
        per_counter[name] = per
        lines.append(f"{counter}: a 1 GiB copy reads as {cal:.1f} KiB -> scale {scales[name]:.4f}")
    kernels = {}
    for k in PASS_NAMES:
        entry = {}
        for name, counter in (("fetch", "FETCH_SIZE"), ("write", "WRITE_SIZE")):
            matches = [(kn, v) for kn, v in per_counter[name].items() if k + "<" in kn or k + "(" in kn]
            if k == "k_binned":
                # the residual pass of a step is one launch per support class:
                # every residual-mode instantiation and the two-phase kernel of
                # the largest supports, summed
                matches = [(kn, v) for kn, v in per_counter[name].items()
                           if "k_binned<2," in kn or "k_binned_wide(" in kn]
            if not matches:
                break
            if k == "k_binned":
                # per step: every dispatch of every symbol over the number of
                # steps profiled (the dispatch count of a symbol launched once
                # per step; a symbol that serves two support classes is
                # launched twice per step)
                kname = " + ".join(sorted(kn[:60] for kn, _ in matches))
                n_disp = min(len(v) for _, v in matches)
                raw = sum(sum(v) for _, v in matches) / n_disp
            else:
                kname, vals = max(matches, key=lambda kv: len(kv[1]))
                raw = sum(vals) / len(vals)
                n_disp = len(vals)
            entry[name] = dict(counter=counter, kernel=kname[:400], dispatches=n_disp, kib_per_dispatch=raw,
                               bytes_per_dispatch=raw * scales[name] * 1024.0)
        if len(entry) == 2:
            entry["bytes_per_dispatch"] = entry["fetch"]["bytes_per_dispatch"] + entry["write"]["bytes_per_dispatch"]
            kernels[k] = entry
            lines.append(f"{k}: {entry['fetch']['dispatches']} dispatches, fetch "
                         f"{entry['fetch']['bytes_per_dispatch'] / 1e9:.4f} GB + write "
                         f"{entry['write']['bytes_per_dispatch'] / 1e9:.4f} GB = "
                         f"{entry['bytes_per_dispatch'] / 1e9:.4f} GB per dispatch")
    with open(out_json, "w") as f:
        json.dump(dict(kernels=kernels, scales=scales, commands=list(commands)), f, indent=1)
    with open(out_txt, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def reduce_counters(directory, kernel, out_txt, command, kernel_ms=None):
    """rocprofv3 --pmc run -> per-counter mean over the dispatches of `kernel`
    (substrings separated by '|': any of them), with the FP32 operation rate
    when the FP32 instruction counters are there."""
    per = {}  # counter -> kernel -> values
    keys = kernel.split("|")
    for r in _profile_csv(directory, "counter_collection.csv"):
        if any(k in r["Kernel_Name"] for k in keys):
            per.setdefault(r["Counter_Name"], {}).setdefault(r["Kernel_Name"], []).append(float(r["Counter_Value"]))
    if not per:
        raise SystemExit(f"no dispatches of {kernel}")
    names = sorted({kn for d in per.values() for kn in d})
    lines = [f"command: {command}", "", f"kernel: {kernel} (per step: the sum over {len(names)} instantiation(s), "
                                        "one launch per support class)"]
    lines += [f"  {kn[:110]}" for kn in names]
    # per step: every dispatch of every instantiation over the number of steps
    # profiled (the dispatch count of an instantiation launched once per step)
    mean = {c: sum(sum(v) for v in d.values()) / min(len(v) for v in d.values()) for c, d in per.items()}
    for k in sorted(mean):
        n = min(len(v) for v in per[k].values())
        lines.append(f"  {k:28s} {mean[k]:16.4f}   ({n} dispatches each)")
    f32 = ("SQ_INSTS_VALU_FMA_F32", "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_TRANS_F32")
    if all(k in mean for k in f32):
        ops = (2 * mean[f32[0]] + mean[f32[1]] + mean[f32[2]] + mean[f32[3]]) * 64
        lines.append(f"  FP32 operations executed per dispatch: {ops:.4e}")
        if kernel_ms:
            tf = ops / (float(kernel_ms) * 1e-3) / 1e12
            lines.append(f"  at {float(kernel_ms):.3f} ms per dispatch: {tf:.2f} TFLOP/s executed = "
                         f"{100 * tf / FP32_PEAK_TFLOPS:.1f} % of the FP32 peak ({FP32_PEAK_TFLOPS} TFLOP/s)")
    with open(out_txt, "w") as f:
        f.write("\n".join(lines) + "\n")
    print("\n".join(lines))


def setup_workload(pkg, torch, name, seed, dev, log=log):
    """The workload ``name`` on the device: synthetic data (``make_cube`` or
    ``make_mosaic``), a gridder with its W kernels or AW Airy atlas attached,
    Briggs robust 0 weights for the cubes that ask for them, and the
    RowChannelBatch the visibility set is built from."""
    wl = WORKLOADS[name]
    mosaic = wl.get("kind") == "mosaic"
    t0 = time.perf_counter()
    data = (make_mosaic if mosaic else make_cube)(torch, wl, seed, dev)
    torch.cuda.synchronize()
    log(f"data: {data['n_rows']} rows x {data['n_chan']} chan, cell {np.rad2deg(data['cell']) * 3600:.4f} arcsec, "
        f"max|uv| {data['max_uv']:.0f}, max|w| {data['max_w']:.0f} ({time.perf_counter() - t0:.1f} s)")

    nx, cell, n_chan = data["nx"], data["cell"], data["n_chan"]
    g = pkg.StandardGridder(nx, nx, cell, cell)
    t0 = time.perf_counter()
    if mosaic:
        # pointing phase gradients (lib.rs:56657-56671) from each pointing's
        # pixel offset; CF cells from the Airy power pattern at cf_freqs CF
        # frequencies across the band (the centres of equal sub-bands), each
        # sample using the nearest (awprojection.rs:1127-1180)
        grads = [pkg.mosaic_phase_gradient(g.info.gx, g.info.gy, (lp / cell, -mp / cell), wl["sampling"])
                 for lp, mp in data["pointings"]]
        data["gradients"] = np.asarray(grads)
        chan_freq = np.linspace(wl["f0"], wl["f1"], wl["n_chan"])
        nf = wl.get("cf_freqs", 1)
        cf_freqs = chan_freq.reshape(nf, -1).mean(axis=1)
        g.attach_aw_airy(wl["dish"], wl["blockage"], cf_freqs, wl["planes"], data["max_w"],
                         sampling=wl["sampling"], phase_gradients=grads)
        kernel_build_s = time.perf_counter() - t0
        sup = np.array([c[1] for c in g.aw_cells()])
        data["cf_freqs"] = cf_freqs
        log(f"AW CF atlas: {len(cf_freqs)} frequencies x {wl['planes']} W planes = {len(sup)} Airy x W cells x "
            f"{len(grads)} pointings = {g.info.planes} tap tables, sampling {g.info.sampling}, supports "
            f"{sup.min()}..{sup.max()} ({kernel_build_s * 1e3:.1f} ms)")
    else:
        g.attach_wproject(data["max_w"], planes=wl["planes"], sampling=wl["sampling"])
        kernel_build_s = time.perf_counter() - t0
        sup, _, _ = g.wproject_tables()
        log(f"W kernels: {g.info.planes} planes, sampling {g.info.sampling}, width {g.info.width}, "
            f"supports {sup.min()}..{sup.max()} ({kernel_build_s * 1e3:.1f} ms)")
    if not mosaic and wl.get("weighting", "briggs") == "briggs":
        # Briggs robust 0 from one density grid per channel (library, cube convention)
        natural = pkg.RowChannelBatch(data["uvw"], data["scale"], data["vis"], data["weight"], data["flag"])
        t0 = time.perf_counter()
        wts, f2, _ = g.density_weights(natural, pkg.WEIGHT_BRIGGS, 0.0, per_plane=True, want_density=False)
        torch.cuda.synchronize()
        data["weight"] = wts.view(data["n_rows"], data["n_chan"])
        del natural, wts
        log(f"Briggs robust 0 weights from {n_chan} density grids ({time.perf_counter() - t0:.2f} s), "
            f"f2 {f2.min():.3g}..{f2.max():.3g}")
    batch = pkg.RowChannelBatch(data["uvw"], data["scale"], data["vis"], data["weight"], data["flag"],
                                pointing=data.get("pointing"))
    if mosaic:
        data["cf_cell"] = g.aw_select_cells(batch)
        batch.cf_cell = data["cf_cell"]
    n_unflagged = int((~data["flag"]).sum())
    return dict(wl=wl, mosaic=mosaic, data=data, g=g, batch=batch, sup=sup, kernel_build_s=kernel_build_s,
                n_unflagged=n_unflagged)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="cube256", choices=sorted(WORKLOADS))
    ap.add_argument("--fp64-steps", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-minor", action="store_true", help="skip the Hogbom minor-cycle timing")
    ap.add_argument("--oracle-channels", type=int, default=None,
                    help="channels (planes) checked against the CPU FP64 oracle (default 2, w4096 1; 0: none)")
    ap.add_argument("--seed", type=int, default=2026)
    ap.add_argument("--wave-groups", type=int, default=0, choices=(0, 1, 2),
                    help="groups per wavefront of the binned kernel (0: two where the halo allows)")
    ap.add_argument("--schedule", default="", metavar="TILE,ORDER,THREADS,GROUPS",
                    help="fix the binned kernel's schedule instead of autotuning it (profiling runs)")
    ap.add_argument("--compare-hipfft", type=int, default=1,
                    help="also time the step with hipFFT 2-D plans (reported beside value)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run a plain 1 GiB device copy (reference for HBM byte counters)")
    ap.add_argument("--reduce-trace", nargs=3, metavar=("DIR", "OUT_TXT", "COMMAND"),
                    help="summarise a rocprofv3 --kernel-trace --stats run (no GPU work)")
    ap.add_argument("--reduce-traffic", nargs=6,
                    metavar=("FETCH_DIR", "WRITE_DIR", "OUT_JSON", "OUT_TXT", "FETCH_CMD", "WRITE_CMD"),
                    help="reduce rocprofv3 FETCH_SIZE / WRITE_SIZE runs to bytes per dispatch of every step kernel "
                         "(no GPU work)")
    ap.add_argument("--reduce-counters", nargs="+", metavar="ARG",
                    help="DIR KERNEL OUT_TXT COMMAND [KERNEL_MS]: summarise a rocprofv3 --pmc run (no GPU work)")
    args = ap.parse_args()
    if args.reduce_counters:
        reduce_counters(*args.reduce_counters)
        return
    if args.reduce_trace:
        reduce_trace(*args.reduce_trace)
        return
    if args.reduce_traffic:
        a = args.reduce_traffic
        reduce_traffic(a[0], a[1], a[2], a[3], a[4:6])
        return
    if args.oracle_channels is None:
        args.oracle_channels = 1 if args.workload.startswith("w4096") else 2
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every GPU box here has one, and the sharded path "
              f"is described, not built (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    import torch
    import __graft_entry__

    pkg = __graft_entry__.load_package()
    dev = torch.device("cuda", 0)
    wl = WORKLOADS[args.workload]
    if args.calibrate_copy:
        # 1 GiB read + 1 GiB written by torch's copy kernel: the scale for
        # FETCH_SIZE / WRITE_SIZE readings of the gridding kernels.
        src = torch.ones(1 << 28, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy: 3 x 1 GiB device copies done")
    ws = setup_workload(pkg, torch, args.workload, args.seed, dev)
    data, g, batch, sup = ws["data"], ws["g"], ws["batch"], ws["sup"]
    mosaic, kernel_build_s, n_unflagged = ws["mosaic"], ws["kernel_build_s"], ws["n_unflagged"]
    nx, cell, n_chan = data["nx"], data["cell"], data["n_chan"]
    t0 = time.perf_counter()
    vs = g.visibility_set(batch)
    torch.cuda.synchronize()
    plan_s = time.perf_counter() - t0
    st = vs.stats()
    log(f"plan+bin: {st['n_valid']} accepted of {n_unflagged} unflagged, {st['n_groups']} groups of "
        f"identical plans, {st['n_items']} work items, {st['taps']} taps ({plan_s * 1e3:.1f} ms)")
    nsum = np.array([s["normalization_sumwt"] for s in vs.sums])

    # initial dirty images: batched FP32 vs plain FP64
    psf32, dirty32 = g.new_grids(n_chan), g.new_grids(n_chan)
    vs.grid(pkg.MODE_DIRTY, psf=psf32, dirty=dirty32)
    p32, d32, peak32 = g.finish_products(psf32, dirty32, nsum)
    del psf32, dirty32
    result_extra = {}
    if not args.no_fp64:
        psf64, dirty64 = g.new_grids(n_chan, pkg.F64), g.new_grids(n_chan, pkg.F64)
        sums64 = g.grid_plain(batch, pkg.MODE_DIRTY, pkg.F64, psf=psf64, dirty=dirty64)
        nsum64 = np.array([s["normalization_sumwt"] for s in sums64])
        p64, d64, peak64 = g.finish_products(psf64, dirty64, nsum64, precision=pkg.F64)
        del psf64, dirty64
        ref_peak = torch.amax(torch.abs(d64), dim=(1, 2))
        per_chan = torch.amax(torch.abs(d32 - d64), dim=(1, 2)) / ref_peak
        bad = (~torch.isfinite(per_chan)).nonzero().flatten().tolist()
        if bad:
            c = bad[0]
            log(f"non-finite dirty error in {len(bad)} channels, e.g. {c}: ref peak {ref_peak[c].item()}, "
                f"FP32 finite {bool(torch.isfinite(d32[c]).all())}, FP64 finite {bool(torch.isfinite(d64[c]).all())}, "
                f"nsum32 {nsum[c]}, nsum64 {nsum64[c]}, peak32 {peak32[c]}, peak64 {peak64[c]}")
        dirty_err = float(torch.max(per_chan))
        psf_err = float(torch.max(torch.abs(p32 - p64)))
        result_extra.update(dirty_err_vs_fp64=dirty_err, psf_err_vs_fp64=psf_err)
        log(f"dirty image error FP32 batched vs FP64 plain: {dirty_err:.2e} of peak; PSF {psf_err:.2e}")
    torch.cuda.empty_cache()

    model = data["model"]
    n_vis = st["n_valid"]  # visibilities the planner accepted: the throughput numerator
    kernel_ms, step_s = [], []
    r32 = torch.empty((n_chan, nx, nx), dtype=torch.float32, device=dev)

    def step32(record=True):
        # model -> apodise + FFT -> degrid + grid (V - p) w -> inverse FFT ->
        # residual image; the call returns once its work completed
        t = time.perf_counter()
        vs.major_cycle(model, nsum, peak32, out=r32)
        if record:
            step_s.append(time.perf_counter() - t)
            kernel_ms.append(vs.stats()["last_kernel_ms"])
        return r32

    wave_groups = vs.set_wave_groups(args.wave_groups)
    if args.schedule:
        # a fixed binned-kernel schedule (profiling runs: no autotuner launches)
        tile, order, threads, gpi = (int(x) for x in args.schedule.split(","))
        vs.set_schedule(threads, gpi, tile=tile, order=order)
    # the first major cycle sets up the workspace, times the transform shapes
    # and (unless fixed) autotunes the binned schedule
    first_s = None
    for k in range(args.warmup):
        t0 = time.perf_counter()
        step32(record=False)
        if k == 0:
            first_s = time.perf_counter() - t0
    info = vs.cycle_info()
    log(f"major-cycle transforms: {'fused row/column passes' if info['strategy'] == pkg.CYCLE_FUSED else 'hipFFT 2-D plans'}"
        f", column block {info['columns']} x {info['column_threads']} threads, {info['rows_per_group']} rows per "
        f"rows group, timed forward+inverse {info['transform_ms']:.2f} ms")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        step32()
    torch.cuda.synchronize()
    t32 = (time.perf_counter() - t0) / args.steps
    rate32 = n_vis / t32 / 1e6
    if first_s is None:  # no warm-up: the first timed step was the first cycle
        first_s = step_s[0]
    log(f"batched FP32 major cycle: {t32 * 1e3:.2f} ms/step, {rate32:.1f} Mvis/s; steps "
        f"{min(step_s) * 1e3:.2f}..{max(step_s) * 1e3:.2f} ms; binned kernel {np.mean(kernel_ms):.3f} ms")
    sched = vs.schedule()
    wave_groups = vs.set_wave_groups(args.wave_groups)  # the value in use after the autotuner
    log(f"binned schedule: {sched['tile']}^2 uv tiles, groups by {('anchor', 'plan')[sched['order']]}, "
        f"{sched['threads']} threads x {sched['groups_per_item']} groups per item ("
        + ("fixed by --schedule" if args.schedule else
           f"autotuned over {len(sched['timed'])} candidates tile/order/threads x groups: "
           + ", ".join(f"{tl}/{od}/{a}x{b} {ms:.3f} ms" for tl, od, a, b, ms in sched["timed"])) + ")")
    cycles = 10  # a typical number of major cycles per imaging run
    amortised = (plan_s + first_s + (cycles - 1) * t32) / cycles
    result_extra.update(
        step_ms={"min": min(step_s) * 1e3, "median": float(np.median(step_s)) * 1e3, "max": max(step_s) * 1e3},
        plan_inclusive={"plan_bin_ms": plan_s * 1e3, "first_cycle_ms": first_s * 1e3,
                        "first_cycle_mvis_s": n_vis / (plan_s + first_s) / 1e6,
                        "cycles": cycles, "amortised_ms_per_cycle": amortised * 1e3,
                        "amortised_mvis_s": n_vis / amortised / 1e6},
        groups={"n_groups": st["n_groups"], "vis_per_group": n_vis / max(st["n_groups"], 1),
                "n_work_items": st["n_items"]},
        transforms="fused" if info["strategy"] == pkg.CYCLE_FUSED else "hipfft",
        transform_shape=[info["columns"], info["column_threads"], info["rows_per_group"]],
        transform_fwd_inv_ms=info["transform_ms"],
        binned_schedule={"tile": sched["tile"], "order": ("anchor", "plan")[sched["order"]],
                         "groups_per_wavefront": wave_groups,
                         "threads": sched["threads"], "groups_per_item": sched["groups_per_item"],
                         "candidates_timed": len(sched["timed"])})
    pass_ms = None
    if info["strategy"] == pkg.CYCLE_FUSED:
        # device time of every pass of the step, from events between the
        # passes, over a few more cycles (not part of the timed steps)
        vs.cycle_passes(1)
        acc = []
        for _ in range(max(3, min(args.steps, 5))):
            step32(record=False)
            acc.append(vs.cycle_passes(-1))
        vs.cycle_passes(0)
        pass_ms = np.mean(np.asarray(acc), axis=0)
        log("passes (ms): " + ", ".join(f"{n} {m:.3f}" for n, m in zip(PASS_NAMES, pass_ms)))

        # the same step replayed from a captured HIP graph (one launch per
        # cycle), after as many warm-up replays as the eager steps had
        vs.set_cycle_graph(True)
        r_g = torch.empty_like(r32)
        for _ in range(max(1, args.warmup)):
            vs.major_cycle(model, nsum, peak32, out=r_g)
        torch.cuda.synchronize()
        graph_s = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            t1 = time.perf_counter()
            vs.major_cycle(model, nsum, peak32, out=r_g)
            graph_s.append(time.perf_counter() - t1)
        torch.cuda.synchronize()
        tg = (time.perf_counter() - t0) / args.steps
        diff = float(torch.max(torch.abs(r_g - r32)) / torch.max(torch.abs(r32)))
        result_extra.update(graph_ms_per_step=tg * 1e3, graph_mvis_s=n_vis / tg / 1e6,
                            graph_step_ms={"min": min(graph_s) * 1e3, "median": float(np.median(graph_s)) * 1e3,
                                           "max": max(graph_s) * 1e3},
                            graph_vs_eager_max_diff=diff, graph_replays=vs.cycle_replays())
        log(f"same step replayed from a HIP graph: {tg * 1e3:.2f} ms/step ({n_vis / tg / 1e6:.1f} Mvis/s; steps "
            f"{min(graph_s) * 1e3:.2f}..{max(graph_s) * 1e3:.2f} ms); graph vs eager residual diff {diff:.2e} of peak")
        vs.set_cycle_graph(False)
        del r_g
    if info["strategy"] == pkg.CYCLE_FUSED and args.compare_hipfft:
        # the same step with hipFFT 2-D plans and separate placement / crop kernels
        r_h = torch.empty_like(r32)
        vs.set_cycle_strategy(pkg.CYCLE_HIPFFT)
        vs.major_cycle(model, nsum, peak32, out=r_h)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(3):
            vs.major_cycle(model, nsum, peak32, out=r_h)
        torch.cuda.synchronize()
        th = (time.perf_counter() - t0) / 3
        diff = float(torch.max(torch.abs(r_h - r32)) / torch.max(torch.abs(r_h)))
        result_extra.update(hipfft_ms_per_step=th * 1e3, fused_vs_hipfft_max_diff=diff)
        log(f"same step with hipFFT 2-D plans: {th * 1e3:.2f} ms/step; fused vs hipFFT residual diff {diff:.2e} of peak")
        vs.set_cycle_strategy(pkg.CYCLE_AUTO)
        del r_h
        torch.cuda.empty_cache()

    if args.oracle_channels > 0:
        # the batched FP32 dirty and residual images of sampled channels
        # against the CPU FP64 oracle on the same inputs (outside the timed
        # region)
        k = min(args.oracle_channels, n_chan)
        pick = np.sort(np.random.default_rng(args.seed + 1).choice(n_chan, k, replace=False))
        result_extra["oracle_fp64"] = oracle_fp64_check(g, data, mosaic, pick, p32, d32, r32)

    if not args.no_fp64:
        torch.cuda.empty_cache()
        mgrid64 = g.new_grids(n_chan, pkg.F64)
        res64 = g.new_grids(n_chan, pkg.F64)

        def step64():
            g.model_to_grid(model, pkg.F64, out=mgrid64)
            res64.zero_()
            g.grid_plain(batch, pkg.MODE_RESIDUAL, pkg.F64, model=mgrid64, dirty=res64)
            return g.finish_products(None, res64, nsum64, psf_peak=peak64, precision=pkg.F64)[1]

        step64()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.fp64_steps):
            r64 = step64()
        torch.cuda.synchronize()
        t64 = (time.perf_counter() - t0) / args.fp64_steps
        rate64 = n_vis / t64 / 1e6
        rpk = torch.amax(torch.abs(r64), dim=(1, 2))
        res_err = float(torch.max(torch.amax(torch.abs(r32 - r64), dim=(1, 2)) / torch.clamp(rpk, min=1e-30)))
        result_extra.update(fp64_plain_mvis_s=rate64, fp64_plain_ms_per_step=t64 * 1e3,
                            speedup_vs_fp64_plain=rate32 / rate64, residual_err_vs_fp64=res_err)
        log(f"plain FP64 per-channel major cycle: {t64 * 1e3:.2f} ms/step, {rate64:.1f} Mvis/s "
            f"(batched FP32 is {rate32 / rate64:.2f}x); residual error {res_err:.2e} of peak")
        del mgrid64, res64
        torch.cuda.empty_cache()

    if not args.no_minor:
        # Hogbom minor cycle on every plane's residual image with its own PSF
        # (run_hogbom_minor_cycle, lib.rs:59640-59885): 100 components per plane
        # (niter 99, CASA's inclusive budget), gain 0.1, no thresholds
        res_c = r32.clone()
        model_c = torch.zeros_like(res_c)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        upd, _, stop = pkg.run_hogbom_minor_cycle(res_c, p32, model_c, 0.1, 99)
        torch.cuda.synchronize()
        tm = time.perf_counter() - t0
        mc = {"planes": n_chan, "components_per_plane": int(upd.max()), "ms": tm * 1e3,
              "us_per_component_all_planes": tm * 1e6 / max(int(upd.max()), 1)}
        if not args.no_cpu:
            import oracle
            t0 = time.perf_counter()
            o_r, o_m, o_upd, _ = oracle.hogbom(r32[0].cpu().numpy(), p32[0].cpu().numpy(),
                                               np.zeros((nx, nx), np.float32), 0.1, 100)
            tc = time.perf_counter() - t0
            same = bool(np.array_equal(o_r[0].view(np.uint32), res_c[0].cpu().numpy().view(np.uint32))
                        and np.array_equal(o_m[0].view(np.uint32), model_c[0].cpu().numpy().view(np.uint32)))
            mc.update(cpu_oracle_s_one_plane=tc, cpu_one_thread=True, plane0_bit_identical_to_oracle=same)
        result_extra["minor_cycle"] = mc
        log(f"Hogbom minor cycle: {int(upd.max())} components on each of {n_chan} planes in {tm * 1e3:.1f} ms"
            + (f"; oracle one plane {mc['cpu_oracle_s_one_plane']:.2f} s, plane 0 bit-identical: "
               f"{mc['plane0_bit_identical_to_oracle']}" if "cpu_oracle_s_one_plane" in mc else ""))
        del res_c, model_c
        torch.cuda.empty_cache()

    # Roofline of the dominant kernel, k_binned<RESIDUAL, W> (degrid + grid of
    # one major cycle).  Algorithmic bytes per launch: 13 B per gridded
    # visibility (vis 8 + weight 4 + flag 1) + 24 B per row (uvw f64) + the
    # model grid read once and the residual grid written once (8 B per cell
    # each).  Algorithmic flops: 16 per tap (complex degrid MAC 8 + grid MAC 8).
    kms = float(np.mean(kernel_ms))
    gx, gy = g.info.gx, g.info.gy
    hy = gy // 2 + 1  # Hermitian half-spectrum columns
    cells = n_chan * gx * gy
    bytes_alg = n_vis * (8 + 4 + 1) + data["n_rows"] * 24 + cells * 8 * 2
    flops_alg = st["taps"] * 16.0
    gbs = bytes_alg / (kms * 1e-3) / 1e9
    tfl = flops_alg / (kms * 1e-3) / 1e12
    hbm_time = bytes_alg / (HBM_PEAK_GBS * 1e9)
    flop_time = flops_alg / (FP32_PEAK_TFLOPS * 1e12)
    if hbm_time >= flop_time:
        roof = {"bound": "hbm", "achieved": gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": gbs / HBM_PEAK_GBS}
    else:
        # compute-bound: the FP32 vector ALUs (MFMA is not used), priced at
        # the FP32 vector peak of 157.3 TFLOP/s
        roof = {"bound": "valu_fp32", "achieved": tfl, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": tfl / FP32_PEAK_TFLOPS}
    traffic = measured_traffic(args.workload)
    roof["traffic"] = traffic.get("k_binned") if traffic else None
    roof.update(kernel="k_binned<RESIDUAL, W>", kernel_ms=kms, kernel_ms_source="HIP events around each launch on "
                "its stream, mean over the timed steps", algorithmic_bytes=bytes_alg,
                algorithmic_flops=flops_alg, fp32_tflops=tfl, hbm_gbs=gbs,
                hbm_frac=gbs / HBM_PEAK_GBS, fp32_frac=tfl / FP32_PEAK_TFLOPS,
                measured_hbm_gbs=roof["traffic"] / (kms * 1e-3) / 1e9 if roof["traffic"] else None)
    if pass_ms is not None:
        # every pass of the fused step: algorithmic bytes (each input read
        # once, each output written once, the whole half spectrum of every
        # channel, no uv-coverage pruning) / event-timed device ms
        img = n_chan * nx * nx * 4
        half_img = n_chan * nx * hy * 8
        half_grid = n_chan * gx * hy * 8
        alg = {"k_fwd_rows": img + half_img, "k_fwd_cols": half_img + half_grid, "k_zero_touched": half_grid,
               "k_binned": bytes_alg, "k_inv_cols": half_grid + half_img, "k_inv_rows": half_img + img}
        passes = []
        for name, ms in zip(PASS_NAMES, pass_ms):
            a = alg[name] / (ms * 1e-3) / 1e9
            t = traffic.get(name) if traffic else None
            passes.append({"kernel": name, "ms": float(ms), "algorithmic_bytes": int(alg[name]), "achieved_gbs": a,
                           "frac": a / HBM_PEAK_GBS, "traffic": t,
                           "measured_hbm_gbs": t / (ms * 1e-3) / 1e9 if t else None})
        roof["passes"] = passes
        step_alg = sum(alg.values())
        roof["step"] = {"ms": t32 * 1e3, "passes_ms": float(np.sum(pass_ms)), "algorithmic_bytes": int(step_alg),
                        "achieved_gbs": step_alg / t32 / 1e9, "frac": step_alg / t32 / 1e9 / HBM_PEAK_GBS}

    cpu = None
    if not args.no_cpu:
        try:
            cpu = (cpu_baseline_aw if mosaic else cpu_baseline)(pkg, g, data, args.cpu_seconds, log)
        except Exception as e:  # the baseline is reported, never required
            log(f"cpu baseline failed: {e!r}")

    out = {
        "metric": "Mvis/s grid+degrid, batched FP32 vs plain FP64, same GPU, err <=1e-5 peak",
        "value": rate32,
        "unit": "Mvis/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t32 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": (f"{args.workload}: AW-projection mosaic residual refresh (degrid+grid+FFTs), "
                         f"{len(data['pointings'])} pointings, {data['n_rows']} rows x {n_chan} planes "
                         f"(channels x correlations), {n_unflagged} unflagged vis ({n_vis} gridded), {nx}^2 image / "
                         f"{g.info.gx}^2 grid, {len(sup)} W planes of Airy CFs, sampling {g.info.sampling}" if mosaic else
                         f"{args.workload}: W-projection residual refresh (degrid+grid+FFTs), "
                         f"{data['n_rows']} rows x {n_chan} channels, {n_unflagged} unflagged vis ({n_vis} gridded), "
                         f"{nx}^2 image / {g.info.gx}^2 grid, cell 1/({WORKLOADS[args.workload].get('cell_k', 3.0)} "
                         f"max|uv|), {g.info.planes} W planes with supports {int(sup.min())}..{int(sup.max())}, "
                         f"sampling {g.info.sampling}"),
            "n_rows": data["n_rows"], "n_chan": n_chan, "n_vis_unflagged": n_unflagged,
            "n_vis_gridded": n_vis, "grid": g.info.gx, "image": nx,
            "w_planes": len(sup) if mosaic else g.info.planes,
            "w_sampling": g.info.sampling, "w_support_range": [int(sup.min()), int(sup.max())],
            "grid_point_updates_per_pass": st["taps"], "plan_bin_ms": plan_s * 1e3,
            "w_kernel_build_ms": kernel_build_s * 1e3, "parallelism": "single GPU",
        },
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    out["value_counts"] = "visibilities the planner accepted (gridded), per second of the timed steps"
    out.update(result_extra)
    # taps of every accepted sample; a group of identical plans costs one
    # footprint on the device, so the kernel executes fewer
    out["grid_point_updates_per_s"] = st["taps"] / t32
    print(json.dumps(out))


if __name__ == "__main__":
    main()
