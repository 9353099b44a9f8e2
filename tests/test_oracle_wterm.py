"""The oracle's W term, pinned by physics rather than by its own formula.

Three checks, each with a planted-mistake control that must fail:

1. The reference's own invariant (lib.rs:95026-95099,
   ``direct_w_term_mode_recovers_off_axis_sources_better_than_2d_mode``),
   restated literally: five samples with |w| up to 300 wavelengths, a unit
   point source at pixel (42, 20) of a 64^2 image at 4e-3 rad cells; the
   direct transform and W projection must both put more flux on the source
   pixel than 2-D gridding.
2. The direct transform against the reference's point-source synthesiser
   (``point_source_visibilities_with_mode``, lib.rs:86969-87009): every phase
   cancels at the source pixel, so the residual there equals the flux (up to
   float rounding) whatever the w-phase; a flipped w sign does not.
3. W projection against the direct transform at w-phases of 3.7-9.8 rad.
   The reference's WProjector (screen phase +2 pi w (n - 1), gridder.rs:3453;
   kernel conjugated for w > 0, :3621; value * kernel, :3696) applies the
   phase e^{-2 pi i w (n - 1)} to the image, the opposite sign of its direct
   mode (lib.rs:73909-73910), which applies e^{+2 pi i w (n - 1)}: CASA's
   gridders negate u and v but not w before gridding, and the reference keeps
   that for W projection.  The oracle keeps the reference's sign (bit-level
   parity is the contract), so the physics pin is
   W-projection(u, v, w) == direct(u, v, -w) to within the W-plane and
   kernel-oversampling quantisation (<= 0.1 of the peak here), while 2-D
   gridding misses by more than the peak and a screen of the other sign by
   ~0.9 of it.
"""
import numpy as np

# lib.rs:95028-95034
REFERENCE_SAMPLES = [
    (-80.0, -45.0, 300.0),
    (-50.0, 35.0, -225.0),
    (37.5, -52.5, 262.5),
    (65.0, 42.5, -275.0),
    (20.0, 15.0, 175.0),
]
NX = 64
CELL = 4.0e-3

SOURCES = [((42.0, 20.0), 1.0), ((20.0, 45.0), 0.7), ((50.0, 50.0), 0.5), ((14.0, 16.0), 0.4),
           ((32.0, 32.0), 0.3)]


def point_source_visibilities(u, v, w, cell, shape, offset, flux):
    """point_source_visibilities_with_mode with include_w_term = true
    (lib.rs:86969-87009): flux * e^{-2 pi i (u l + v m + w (n - 1))}, cos and
    sin evaluated in f64 and narrowed to f32 before the product."""
    cx, cy = shape[0] / 2.0, shape[1] / 2.0
    l = (offset[0] - cx) * cell
    m = (cy - offset[1]) * cell
    nm1 = np.sqrt(1.0 - l * l - m * m) - 1.0
    ph = -2.0 * np.pi * (np.asarray(u) * l + np.asarray(v) * m + np.asarray(w) * nm1)
    f = np.float32(flux)
    return (f * np.cos(ph).astype(np.float32) + 1j * (f * np.sin(ph).astype(np.float32))).astype(np.complex64)


def w_phase(offset, w):
    l = (offset[0] - NX / 2) * CELL
    m = (NX / 2 - offset[1]) * CELL
    return 2.0 * np.pi * abs(w) * (1.0 - np.sqrt(1.0 - l * l - m * m))


def sky_samples(seed=5, n=3000):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-90.0, 90.0, n)
    v = rng.uniform(-90.0, 90.0, n)
    w = rng.uniform(-300.0, 300.0, n)
    return u, v, w


def sky_visibilities(u, v, w):
    vis = np.zeros(len(u), np.complex64)
    for off, flux in SOURCES:
        vis = (vis + point_source_visibilities(u, v, w, CELL, (NX, NX), off, flux)).astype(np.complex64)
    return vis


def residual_2d(O, u, v, vis):
    g = O.StandardGridder(NX, NX, CELL, CELL)
    ones = np.ones(len(u), np.float32)
    psf, dirty, sums = g.grid(u, v, ones, ones, vis, mode=1)
    return g.dirty_products(psf, dirty, sums)[1]


def residual_wproject(O, u, v, w, vis, screen_sign=1.0):
    g = O.StandardGridder(NX, NX, CELL, CELL)
    # automatic plane count (w_project_planes: None), as the reference test
    wp = O.WProjector(g, float(np.max(np.abs(w))), screen_sign=screen_sign)
    ones = np.ones(len(u), np.float32)
    psf, dirty, sums = wp.grid(u, v, w, ones, ones, vis, mode=1)
    return wp, wp.dirty_products(psf, dirty, sums)[1]


def residual_direct(O, u, v, w, vis, w_sign=1.0):
    ones = np.ones(len(u), np.float32)
    return O.direct_psf_and_residual(NX, NX, CELL, CELL, u, v, w, ones, ones, vis, w_sign=w_sign)


def test_reference_off_axis_invariant(oracle_mod):
    O = oracle_mod
    u, v, w = (np.array(a, np.float64) for a in zip(*REFERENCE_SAMPLES))
    vis = point_source_visibilities(u, v, w, CELL, (NX, NX), (42.0, 20.0), 1.0)
    two_d = residual_2d(O, u, v, vis)
    wp, wproj = residual_wproject(O, u, v, w, vis)
    psf_d, direct, _, _ = residual_direct(O, u, v, w, vis)
    # automatic plane count of the reference (gridder.rs:3822-3830):
    # floor(1.05 * 300 * sin(64 * 4e-3 / 2))
    assert wp.planes == int(1.05 * 300.0 * np.sin(NX * CELL / 2.0)) == 40
    # lib.rs:95093-95096
    assert direct[42, 20] > two_d[42, 20]
    assert wproj[42, 20] > two_d[42, 20]
    assert O.peak_abs_value(psf_d) >= 1.0
    # every w-phase at the source is above 1 rad, so 2-D gridding is far off
    assert min(w_phase((42.0, 20.0), x) for x in w) > 1.0
    assert two_d[42, 20] < 0.0 and abs(direct[42, 20] - 1.0) < 1e-5


def test_direct_transform_recovers_point_source_exactly(oracle_mod):
    O = oracle_mod
    u, v, w = sky_samples(11, 2000)
    for off in [(42.0, 20.0), (50.0, 50.0), (14.0, 16.0)]:
        assert w_phase(off, 300.0) > 3.0
        vis = point_source_visibilities(u, v, w, CELL, (NX, NX), off, 1.0)
        _, res, _, _ = residual_direct(O, u, v, w, vis)
        x, y = int(off[0]), int(off[1])
        # the phases cancel at the source: the residual there is the flux,
        # and it is the image peak
        assert abs(res[x, y] - 1.0) < 1e-4
        assert np.unravel_index(np.argmax(res), res.shape) == (x, y)
        # planted mistake: the transform evaluated with the other w sign
        _, bad, _, _ = residual_direct(O, u, v, w, vis, w_sign=-1.0)
        assert bad[x, y] < 0.5


def test_wproject_matches_direct_transform_of_its_w_convention(oracle_mod):
    O = oracle_mod
    u, v, w = sky_samples()
    assert max(w_phase(off, 300.0) for off, _ in SOURCES) > 9.0
    # the sky as the W projector's convention sees it: w enters with the
    # opposite sign to u and v
    vis = sky_visibilities(u, v, -w)
    _, ref, _, _ = residual_direct(O, u, v, -w, vis)
    peak = float(np.max(np.abs(ref)))

    def err(img):
        return float(np.max(np.abs(img - ref))) / peak

    _, wproj = residual_wproject(O, u, v, w, vis)
    e_wp = err(wproj)
    e_2d = err(residual_2d(O, u, v, vis))
    _, flipped = residual_wproject(O, u, v, w, vis, screen_sign=-1.0)
    e_flip = err(flipped)
    assert e_wp < 0.1, e_wp
    assert e_2d > 0.5, e_2d
    # planted mistake: the screen phase with the other sign
    assert e_flip > 0.5, e_flip
    # source pixels within 5 % of their flux
    for off, flux in SOURCES:
        x, y = int(off[0]), int(off[1])
        assert abs(wproj[x, y] - ref[x, y]) < 0.05 * peak

    # and the reference's direct mode (w with the sign of u and v) is the
    # transform the W projector does not reproduce: the convention gap
    vis_d = sky_visibilities(u, v, w)
    _, ref_d, _, _ = residual_direct(O, u, v, w, vis_d)
    _, wproj_d = residual_wproject(O, u, v, w, vis_d)
    assert float(np.max(np.abs(wproj_d - ref_d))) / float(np.max(np.abs(ref_d))) > 0.5
