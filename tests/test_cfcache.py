"""CASA CF-cache reader (chip_cfcache_*) and the oracle's key selection, on the
CPU (the reader is host code; no device is needed).

Caches are written by tests/cf_cache_writer.py in casacore's PagedImage layout.
The reference ships no CASA-written cache, so the on-disk format is pinned to
the reference's own reader and writer of that layout (casa-tables
table_control.rs, tiled_stman.rs; casa-aipsio aipsio.rs) and the cache
semantics to the literals of the reference's tests (awprojection.rs:2160-2560):
  * `test_reference_paired_cache_literals`  <- paired_cache_indexes_and_loads_pixels_on_demand
  * `test_reference_mueller_prediction`     <- the conjugate-Mueller prediction test
  * `test_reference_conjugate_frequency_map`<- sample_selection_uses_casa_forward_conjugate_frequency_map
  * `test_reference_rejections`             <- cache_rejects_an_unpaired_cell / _pair_metadata_mismatch /
                                               _pair_uv_coordinate_mismatch, cache_identity_binds_...
The oracle's selection (oracle.cf_select_key) is pinned by those literals, by
its two restated paths (scanned and direct) agreeing on complete caches as the
reference asserts, and by a brute-force minimum over every key; a planted
error in it fails the brute force.  Bars: pixels bit-exact, keys and indices
exact.
"""
import os

import numpy as np
import pytest

from cf_cache_writer import D, I, S, write_test_cell


def _cache(pkg, path):
    return pkg.AwConvolutionFunctionCache.open(path)


def _key_tuple(k):
    return (k.frequency_hz, k.w_value_lambda, k.mueller_element, k.parallactic_angle_deg)


def _paired(root, frequencies, w_values, muellers=(0,), pas=(30.0,), **kw):
    for fi, f in enumerate(frequencies):
        for wi, w in enumerate(w_values):
            for mi, m in enumerate(muellers):
                for pi, p in enumerate(pas):
                    stem = f"0_{mi}_CF_{fi}_{wi}_{pi}.im"
                    write_test_cell(root, f"CFS_{stem}", f, w, False, mueller_element=m, pa=p, **kw)
                    write_test_cell(root, f"WTCFS_{stem}", f, w, True, mueller_element=m, pa=p, **kw)


def test_reference_paired_cache_literals(pkg, tmp_path):
    _paired(str(tmp_path), [1.1e9, 1.2e9], [0.0, 2.0])
    c = _cache(pkg, tmp_path)
    inv = c.inventory()
    assert inv.paired_cells == 4
    assert inv.frequencies_hz == [1.1e9, 1.2e9]
    assert inv.w_values_lambda == [0.0, 2.0]
    assert inv.mueller_elements == [0]
    assert inv.parallactic_angles_deg == [30.0]
    im, wt = c.metadata(0)
    assert im["shape"] == (16, 16) and wt["shape"] == (32, 32)
    assert im["sampling"] == 2 and (im["x_support"], im["y_support"]) == (2, 2)
    assert (wt["x_support"], wt["y_support"]) == (4, 4)
    incs = {c.metadata(k)[j]["uv_increment"] for k in range(4) for j in range(2)}
    assert incs == {(-2.0, 2.0), (-1.0, 1.0)}
    assert im["telescope_name"] == "EVLA" and im["band_name"] == "EVLA_L" and im["polarization"] == 5
    assert c.metadata_fingerprint == _cache(pkg, tmp_path).metadata_fingerprint

    key = c.nearest_key(1.19e9, 1.8, 0, 31.0)
    assert (key.frequency_hz, key.w_value_lambda) == (1.2e9, 2.0)
    a, b = c.load(key)
    assert a.shape == (16, 16) and b.shape == (32, 32)
    assert a[0, 0] == np.complex64(3.0 - 1.0j) and b[0, 0] == np.complex64(7.0 + 2.0j)
    assert np.all(a == np.complex64(3.0 - 1.0j)) and np.all(b == np.complex64(7.0 + 2.0j))

    d = c.select_key_for_sample(1.19e9, 1.0e9, 1.9, 0, 31.0, False)
    assert (d.frequency_hz, d.w_value_lambda) == (1.2e9, 2.0)
    cj = c.select_key_for_sample(0.81e9, 1.0e9, 2.1, 0, 31.0, True)
    assert (cj.frequency_hz, cj.w_value_lambda) == (1.2e9, 2.0)


def test_reference_mueller_prediction(pkg, tmp_path):
    root = str(tmp_path)
    for m in (0, 15):
        for weight in (False, True):
            fam = "WTCFS" if weight else "CFS"
            write_test_cell(root, f"{fam}_0_{m}_CF_0_0_0.im", 1.1e9, 0.0, weight, mueller_element=m)
    c = _cache(pkg, root)
    assert c.select_key_for_prediction_sample(1.1e9, 1.0, 0, 30.0).mueller_element == 15
    assert c.select_key_for_prediction_sample(1.1e9, 0.0, 0, 30.0).mueller_element == 0
    assert c.select_key_for_prediction_sample(1.1e9, -1.0, 15, 30.0).mueller_element == 15
    # forward gridding maps the other way round
    assert c.select_key_for_sample(1.1e9, 1.0e9, 1.0, 0, 30.0, False).mueller_element == 0
    assert c.select_key_for_sample(1.1e9, 1.0e9, -1.0, 0, 30.0, False).mueller_element == 15
    assert c.select_key_for_sample(1.1e9, 1.0e9, 1.0, 3, 30.0, False) is None


def test_reference_conjugate_frequency_map(pkg, tmp_path):
    root = str(tmp_path)
    for fi, (f, cf) in enumerate([(1.0e9, 1.0e9), (2.0e9, 4.0e9)]):
        for weight in (False, True):
            fam = "WTCFS" if weight else "CFS"
            write_test_cell(root, f"{fam}_0_0_CF_{fi}_0_0.im", f, 0.0, weight, conjugate_frequency_hz=cf)
    c = _cache(pkg, root)
    # sqrt(2 * (2 GHz)^2 - (1 GHz)^2) = sqrt(7) GHz selects the 2 GHz CF
    assert c.select_key_for_sample(1.0e9, 2.0e9, 1.0, 0, 30.0, True).frequency_hz == 2.0e9
    assert c.select_key_for_prediction_sample(1.0e9, 0.0, 0, 30.0).frequency_hz == 1.0e9
    # a non-positive radicand selects nothing
    assert c.select_key_for_sample(3.0e9, 1.0e9, 1.0, 0, 30.0, True) is None


def test_reference_rejections(pkg, tmp_path):
    err = pkg.abi.ConvolutionFunctionCache
    r1 = str(tmp_path / "unpaired")
    write_test_cell(r1, "CFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, False)
    with pytest.raises(err, match="missing weight WTCFS"):
        _cache(pkg, r1)
    r2 = str(tmp_path / "sampling")
    write_test_cell(r2, "CFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, False)
    write_test_cell(r2, "WTCFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, True, sampling=3.0)
    with pytest.raises(err, match="sampling does not match"):
        _cache(pkg, r2)
    r3 = str(tmp_path / "uv")
    write_test_cell(r3, "CFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, False)
    write_test_cell(r3, "WTCFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, True, uv_reference_value=(1.0, 0.0))
    with pytest.raises(err, match="does not cover the same UV world window"):
        _cache(pkg, r3)
    # the identity binds the UV affine coordinate
    fps = []
    for name, ref in (("a", (0.0, 0.0)), ("b", (1.0, 0.0))):
        r = str(tmp_path / name)
        write_test_cell(r, "CFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, False, uv_reference_value=ref)
        write_test_cell(r, "WTCFS_0_0_CF_0_0_0.im", 1.1e9, 0.0, True, uv_reference_value=ref)
        c = _cache(pkg, r)
        fps.append((c.metadata_fingerprint, c.metadata(0)[0]["uv_reference_value"]))
    assert fps[0][0] != fps[1][0] and fps[0][1] == (0.0, 0.0) and fps[1][1] == (1.0, 0.0)
    with pytest.raises(err, match="no CFS"):
        os.makedirs(tmp_path / "empty")
        _cache(pkg, str(tmp_path / "empty"))


def test_cache_axis_rejections(pkg, tmp_path):
    err = pkg.abi.ConvolutionFunctionCache
    # incomplete frequency x W product
    r = str(tmp_path / "incomplete")
    _paired(r, [1.1e9, 1.2e9], [0.0])
    write_test_cell(r, "CFS_0_0_CF_0_1_0.im", 1.1e9, 2.0, False)
    write_test_cell(r, "WTCFS_0_0_CF_0_1_0.im", 1.1e9, 2.0, True)
    with pytest.raises(err, match="complete frequency x W x Mueller x PA product"):
        _cache(pkg, r)
    # W values off the quadratic k^2 / WIncr ladder
    r = str(tmp_path / "ladder")
    _paired(r, [1.1e9], [0.0, 2.5])
    with pytest.raises(err, match="CASA quadratic value"):
        _cache(pkg, r)
    # WIncr differing between pairs
    r = str(tmp_path / "wincr")
    write_test_cell(r, "CFS_a.im", 1.1e9, 0.0, False)
    write_test_cell(r, "WTCFS_a.im", 1.1e9, 0.0, True)
    write_test_cell(r, "CFS_b.im", 1.2e9, 0.0, False, w_incr=0.25)
    write_test_cell(r, "WTCFS_b.im", 1.2e9, 0.0, True, w_incr=0.25)
    with pytest.raises(err, match="WIncr must be identical"):
        _cache(pkg, r)
    # missing and mistyped miscinfo fields, bad supports
    for name, override, msg in (("nosamp", {"Sampling": None}, "missing miscinfo field Sampling"),
                                ("typed", {"WValue": I(0)}, "WValue must be floating point"),
                                ("frac", {"Sampling": D(2.5)}, "positive integral"),
                                ("zero", {"Xsupport": I(0)}, "must be positive"),
                                ("band", {"BandName": S("  ")}, "must be non-empty"),
                                ("big", {"Xsupport": I(9)}, "cannot hold support")):
        r = str(tmp_path / name)
        write_test_cell(r, "CFS_x.im", 1.1e9, 0.0, False, override=override)
        write_test_cell(r, "WTCFS_x.im", 1.1e9, 0.0, True)
        with pytest.raises(err, match=msg):
            _cache(pkg, r)
    # non-finite pixels fail at load, not at open
    r = str(tmp_path / "nan")
    px = np.full((16, 16), 1.0 + 0.0j, np.complex64)
    px[3, 5] = np.nan
    write_test_cell(r, "CFS_x.im", 1.1e9, 0.0, False, pixels=px)
    write_test_cell(r, "WTCFS_x.im", 1.1e9, 0.0, True)
    c = _cache(pkg, r)
    with pytest.raises(err, match="non-finite"):
        c.load(0)
    # a duplicate cell of one key
    r = str(tmp_path / "dup")
    write_test_cell(r, "CFS_x.im", 1.1e9, 0.0, False)
    write_test_cell(r, "CFS_y.im", 1.1e9, 0.0, False)
    write_test_cell(r, "WTCFS_x.im", 1.1e9, 0.0, True)
    with pytest.raises(err, match="duplicate imaging cell"):
        _cache(pkg, r)


@pytest.mark.parametrize("layout", [
    dict(big_endian=True),
    dict(big_endian=False),
    dict(big_endian=False, tsm_version=3),
    dict(big_endian=True, tile=[5, 7, 1, 1]),
    dict(big_endian=False, tile=[6, 4, 1, 1], manager="TiledShapeStMan"),
    dict(big_endian=True, manager="TiledShapeStMan", tsm_version=3, stokes_as_codes=True),
])
def test_pixel_layouts_round_trip(pkg, tmp_path, layout):
    """Tiles in either endianness, TiledStMan versions 1-3, partial edge tiles,
    both tiled managers: the pixels come back bit for bit as (x, y)."""
    rng = np.random.default_rng(11)
    a = (rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16))).astype(np.complex64)
    b = (rng.normal(size=(32, 32)) + 1j * rng.normal(size=(32, 32))).astype(np.complex64)
    a[1, 2] = np.complex64(-0.0 + 3.0e-38j)  # sign of zero and a subnormal-adjacent value survive
    root = str(tmp_path)
    write_test_cell(root, "CFS_0.im", 1.3e9, 0.0, False, pixels=a, **dict(layout))
    write_test_cell(root, "WTCFS_0.im", 1.3e9, 0.0, True, pixels=b, **dict(layout))
    c = _cache(pkg, root)
    ga, gb = c.load(0)
    assert ga.view(np.uint32).tobytes() == a.view(np.uint32).tobytes()
    assert gb.view(np.uint32).tobytes() == b.view(np.uint32).tobytes()
    assert c.metadata(0)[0]["polarization"] == 5


def _random_cache(rng, nf, nw, muellers, pas, w_incr):
    f = sorted(set(np.round(rng.uniform(0.9e9, 1.6e9, nf), -3)))
    w = [k * k / w_incr for k in range(nw)]
    return [(fv, wv, m, p) for fv in f for wv in w for m in muellers for p in pas]


def test_oracle_selection_pinned(oracle_mod):
    """The oracle's scanned and direct restatements agree on complete caches
    (as awprojection.rs:2231-2262 asserts of the reference's two paths) and
    equal a brute-force lexicographic minimum over every key; a planted
    tie-break error (larger bits first) breaks the agreement."""
    o = oracle_mod
    rng = np.random.default_rng(5)
    for trial in range(30):
        keys = _random_cache(rng, int(rng.integers(1, 4)), int(rng.integers(1, 5)), [0, 5, 10, 15],
                             [-170.0, -30.0, 0.0, 30.0, 150.0][: int(rng.integers(1, 6))], 0.5)
        for _ in range(40):
            f = float(rng.uniform(0.8e9, 1.7e9))
            if rng.random() < 0.2:  # exact ties between two CF frequencies
                fs = sorted({k[0] for k in keys})
                if len(fs) > 1:
                    f = (fs[0] + fs[1]) / 2
            w = float(rng.uniform(-30, 30))
            m = int(rng.choice([0, 5, 10, 15]))
            pa = float(rng.choice([-180.0, 180.0, 90.0, rng.uniform(-400, 400)]))
            conj = bool(rng.random() < 0.3)
            pred = bool(rng.random() < 0.3)
            a = o.cf_select_key(keys, 0.5, f, 1.2e9, w, m, pa, conj, pred)
            b = o.cf_select_key(keys, 0.5, f, 1.2e9, w, m, pa, conj, pred, direct=True)
            assert a == b
            # brute force: the definition of select_key_for_sample_scanned
            mm = (m if w > 0 else 15 - m) if not pred else (15 - m if w > 0 else m)
            req = f if (pred or not conj) else (np.sqrt(2 * 1.2e9 ** 2 - f * f) if 2 * 1.2e9 ** 2 > f * f else None)
            pool = [k for k in keys if k[2] == mm]
            if req is None or not pool:
                assert a is None
                continue
            ws = sorted({k[1] for k in pool})
            wi = int(min(max(np.floor(np.sqrt(0.5 * abs(w)) + 0.5), 0), len(ws) - 1))
            cand = [k for k in pool if k[1] == ws[wi]]
            best = min(cand, key=lambda k: (abs(k[0] - req), abs(o.cf_circular_degrees(k[3] - pa)),
                                            o.cf_stable_key(k)))
            assert a == best
    # planted: ties to the larger frequency would differ from the brute force
    keys = [(1.0e9, 0.0, 0, 30.0), (2.0e9, 0.0, 0, 30.0)]
    assert o.cf_select_key(keys, 0.5, 1.5e9, 1.0e9, 1.0, 0, 30.0)[0] == 1.0e9
    planted = max([k for k in keys], key=lambda k: (-abs(k[0] - 1.5e9), o.cf_stable_key(k)))
    assert planted[0] != o.cf_select_key(keys, 0.5, 1.5e9, 1.0e9, 1.0, 0, 30.0)[0]


def test_library_selection_matches_oracle(pkg, oracle_mod, tmp_path):
    """Host selection of the library (nearest_key, select_key_for_sample and
    select_key_for_prediction_sample) equals the oracle on a cache with two
    frequencies, three W planes, Mueller 0 and 15 and three PAs."""
    o = oracle_mod
    root = str(tmp_path)
    _paired(root, [1.1e9, 1.25e9], [0.0, 2.0, 8.0], muellers=(0, 15), pas=(-170.0, 30.0, 150.0))
    c = _cache(pkg, root)
    keys = [_key_tuple(k) for k in c.keys()]
    assert keys == sorted(keys, key=o.cf_stable_key) and len(keys) == 36
    rng = np.random.default_rng(9)
    for _ in range(600):
        f = float(rng.uniform(0.9e9, 1.5e9)) if rng.random() > 0.1 else 1.175e9
        w = float(rng.uniform(-40, 40))
        m = int(rng.choice([0, 15, 3]))
        pa = float(rng.uniform(-500, 500)) if rng.random() > 0.1 else -10.0
        conj = bool(rng.random() < 0.4)
        got = c.select_key_for_sample(f, 1.2e9, w, m, pa, conj)
        want = o.cf_select_key(keys, 0.5, f, 1.2e9, w, m, pa, conj, False)
        assert (None if got is None else _key_tuple(got)) == want
        got = c.select_key_for_prediction_sample(f, w, m, pa)
        want = o.cf_select_key(keys, 0.5, f, 0.0, w, m, pa, False, True)
        assert (None if got is None else _key_tuple(got)) == want
        got = c.nearest_key(f, w, m, pa)
        want = o.cf_nearest_key(keys, f, w, m, pa)
        assert (None if got is None else _key_tuple(got)) == want
    assert c.select_key_for_sample(float("nan"), 1.2e9, 0.0, 0, 0.0, False) is None
    assert c.select_key_for_sample(1.1e9, 1.2e9, float("inf"), 0, 0.0, False) is None
