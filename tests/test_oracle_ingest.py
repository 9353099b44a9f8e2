"""The oracle's row-stream ingestion (co_ingest_rows: execution.rs:2168-2315,
types.rs:1593-1631) pinned without the reference:

* the pair collapses by what the algebra fixes: a half sum of a hand with
  itself is that hand, a half difference with itself is zero, the
  imaginary variants are +-i times the half difference;
* the exclusion rules of the routed-sample builder on their branches
  (flags, FLAG_ROW, non-finite visibilities with and without PSF-only use,
  non-positive and non-finite pair weights, 0.5 (w0 + w1));
* with a phase-centre rotation, an explicit correlation equals the oracle's
  row rotation (co_rotate_rows, pinned by the reference's phasor literal in
  test_oracle.py) of that correlation, and a pair equals the collapse of its
  two rotated hands (casars-imager lib.rs:19207-19211 rotates both hands).
"""
import numpy as np
import pytest


def _rows(seed, n=40, n_corr=4, n_chan=6):
    rng = np.random.default_rng(seed)
    uvw = rng.uniform(-3000.0, 3000.0, (n, 3))
    data = (rng.normal(size=(n, n_corr, n_chan)) + 1j * rng.normal(size=(n, n_corr, n_chan))).astype(np.complex64)
    flag = rng.uniform(size=(n, n_corr, n_chan)) < 0.1
    weight = rng.uniform(0.5, 2.0, (n, n_corr)).astype(np.float32)
    freq = np.linspace(1.0e9, 1.4e9, n_chan)
    return uvw, data, flag, weight, freq


def test_collapse_closed_forms(oracle_mod):
    uvw, data, flag, weight, freq = _rows(1)
    flag[:] = False
    same = data.copy()
    same[:, 1] = same[:, 0]
    hs = oracle_mod.ingest_rows(uvw, same, flag, weight, freq, corr=(0, 1), pair="half_sum")
    assert np.array_equal(hs["visibility"], same[:, 0])
    hd = oracle_mod.ingest_rows(uvw, same, flag, weight, freq, corr=(0, 1), pair="half_difference")
    assert np.all(hd["visibility"] == 0)
    d = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 3), pair="half_difference")["visibility"]
    half = (data[:, 0].astype(np.complex128) - data[:, 3]) / 2
    assert np.max(np.abs(d - half)) <= 1e-6 * np.max(np.abs(half))
    pos = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 3),
                                 pair="positive_half_imag_difference")["visibility"]
    neg = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 3),
                                 pair="negative_half_imag_difference")["visibility"]
    assert np.array_equal(pos, (1j * d).astype(np.complex64))
    assert np.array_equal(neg, -pos)


def test_exclusion_rules(oracle_mod):
    uvw, data, flag, weight, freq = _rows(2, n=6, n_corr=2, n_chan=3)
    flag[:] = False
    flag[0, 0, 0] = True               # explicit corr 0 flagged
    flag[1, 1, 1] = True               # second hand flagged: pair excluded, explicit kept
    data[2, 0, 2] = np.nan             # non-finite visibility
    weight[3, 1] = 0.0                 # a non-positive hand weight
    weight[4, 0] = np.inf              # a non-finite hand weight
    row_flag = np.zeros(6, np.uint8)
    row_flag[5] = 1
    e = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=0, row_flag=row_flag)
    assert e["flag"][0, 0] and not e["flag"][1, 1] and np.all(e["flag"][5])
    assert not e["flag"][2, 2] and np.isnan(e["visibility"][2, 2].real)       # PSF only
    assert not e["flag"][3, 0] and not e["flag"][4, 0]                     # explicit: no weight filter here
    e2 = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=0, allow_psf_only=False)
    assert e2["flag"][2, 2]
    p = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 1), pair="half_sum", sumwt_factor=2.0,
                               row_flag=row_flag)
    assert p["flag"][0, 0] and p["flag"][1, 1] and p["flag"][2, 2]
    assert np.all(p["flag"][3]) and np.all(p["flag"][4]) and np.all(p["flag"][5])
    ok = ~p["flag"]
    expect_w = (0.5 * (weight[:, 0] + weight[:, 1]).astype(np.float32)).astype(np.float32)
    assert np.array_equal(p["weight"][ok], np.broadcast_to(expect_w[:, None], (6, 3))[ok])
    assert np.all(p["sumwt_factor"] == 2.0)


@pytest.mark.parametrize("spectrum", [False, True])
def test_rotation_matches_row_rotation(oracle_mod, spectrum):
    uvw, data, flag, weight, freq = _rows(3)
    if spectrum:
        weight = np.random.default_rng(4).uniform(0.5, 2.0, data.shape).astype(np.float32)
    rot = oracle_mod.phase_rotation((1.2, 0.52), (1.2 + 2e-3, 0.52 - 1e-3))
    e = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=2, rotation=rot)
    ouvw, _, ovis = oracle_mod.rotate_rows(rot[0], rot[1], uvw, freq, data[:, 2])
    assert np.array_equal(e["uvw_m"], ouvw)
    assert np.array_equal(e["visibility"], ovis)
    p = oracle_mod.ingest_rows(uvw, data, flag, weight, freq, corr=(0, 3), pair="half_sum", rotation=rot)
    a = oracle_mod.rotate_rows(rot[0], rot[1], uvw, freq, data[:, 0])[2]
    b = oracle_mod.rotate_rows(rot[0], rot[1], uvw, freq, data[:, 3])[2]
    want = np.empty_like(a)
    want.real = (a.real + b.real) * np.float32(0.5)
    want.imag = (a.imag + b.imag) * np.float32(0.5)
    assert np.array_equal(p["visibility"], want)
    if not spectrum:
        assert np.array_equal(e["weight"], np.broadcast_to(weight[:, 2][:, None], e["weight"].shape))
    else:
        assert np.array_equal(e["weight"], weight[:, 2, :])
