"""Seeded synthetic visibility sets for the parity tests (numpy, CPU)."""
from __future__ import annotations

import numpy as np

C = 299_792_458.0


def random_samples(seed, n, umax, wmax=0.0, edge_cases=True):
    """Per-sample wavelengths with the edge cases the reference filters:
    non-gridable, zero / NaN / infinite weights, NaN sumwt factor, NaN
    visibilities, non-finite coordinates, samples off the grid, and exact
    half-cell coordinates."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, n)
    v = rng.uniform(-umax, umax, n)
    w = rng.uniform(-wmax, wmax, n) if wmax > 0 else np.zeros(n)
    vis = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    wt = rng.uniform(0.2, 3.0, n).astype(np.float32)
    sf = rng.uniform(0.5, 1.5, n).astype(np.float32)
    gridable = np.ones(n, bool)
    if edge_cases and n >= 64:
        k = np.arange(n)
        gridable[k % 97 == 5] = False
        wt[k % 89 == 7] = 0.0
        wt[k % 113 == 9] = np.nan
        wt[k % 127 == 11] = np.inf
        sf[k % 131 == 13] = np.nan
        vis[k % 83 == 3] = np.complex64(np.nan + 1j)
        u[k % 139 == 17] = np.nan
        u[k % 149 == 19] = 50.0 * umax  # off the grid
        w[k % 151 == 23] = np.inf
    return u, v, w, wt, sf, vis, gridable


def half_cell_samples(du, dv, n_cells=20, os_=128):
    """Coordinates exactly on rounding ties of the anchor and the offset."""
    k = np.arange(-n_cells, n_cells)
    u = np.concatenate([(k + 0.5) * du, (k + 0.5 / os_) * du, (k - 0.5 / os_) * du, k * du])
    v = np.concatenate([(k + 0.5) * dv, (k - 0.5 / os_) * dv, (k + 0.5) * dv, -(k + 0.5) * dv])
    return u, v


def y_array(seed, n_ant=27, arm_m=21_000.0, jitter_m=50.0):
    """Antennas on a Y with three arms at 120 degrees (power-law spacing along
    each arm, seeded jitter).  Returns ENU positions (n_ant, 3) in metres."""
    rng = np.random.default_rng(seed)
    per_arm = n_ant // 3
    frac = (np.arange(1, per_arm + 1) / per_arm) ** 1.716
    pos = []
    for arm in range(3):
        az = np.deg2rad(5.0 + 120.0 * arm)
        for f in frac:
            r = arm_m * f
            e, nn = r * np.sin(az), r * np.cos(az)
            pos.append([e + rng.normal(0, jitter_m), nn + rng.normal(0, jitter_m), rng.normal(0, 2.0)])
    return np.asarray(pos)


def disk_array(seed, n_ant=27, radius_m=1000.0):
    rng = np.random.default_rng(seed)
    r = radius_m * np.sqrt(rng.uniform(0, 1, n_ant))
    t = rng.uniform(0, 2 * np.pi, n_ant)
    return np.stack([r * np.cos(t), r * np.sin(t), np.zeros(n_ant)], axis=1)


def uvw_tracks(enu, lat_deg, dec_deg, hours, dt_s):
    """uvw (rows, 3) in metres for every baseline and time step (earth-rotation
    synthesis with a phase centre at hour angle h, declination dec)."""
    lat, dec = np.deg2rad(lat_deg), np.deg2rad(dec_deg)
    e, nn, up = enu[:, 0], enu[:, 1], enu[:, 2]
    # ENU -> equatorial XYZ (local meridian)
    X = -np.sin(lat) * nn + np.cos(lat) * up
    Y = e
    Z = np.cos(lat) * nn + np.sin(lat) * up
    ant = np.stack([X, Y, Z], axis=1)
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
    return np.stack([u, v, w], axis=-1).reshape(-1, 3)


def point_source_vis(uvw_l, sources):
    """Visibilities exp(-2 pi i (u l + v m + w (n - 1))) summed over (flux, l, m)."""
    u, v, w = uvw_l
    out = np.zeros(u.shape, np.complex128)
    for flux, l, m in sources:
        n1 = np.sqrt(1.0 - l * l - m * m) - 1.0
        out += flux * np.exp(-2j * np.pi * (u * l + v * m + w * n1))
    return out.astype(np.complex64)
