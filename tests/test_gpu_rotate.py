"""Phase-centre rotation of a row stream on the GPU against the oracle.

uvw' and the phase shift are FP64 arithmetic in the reference's order and
bit-exact; the phasor's FP64 cos/sin may differ from the host library in the
last bit before narrowing to Float, so rotated visibilities agree within one
Float ulp of their magnitude (the reference's own phasor test uses 1e-6).
"""
import numpy as np
import pytest
import torch

import synth

pytestmark = pytest.mark.gpu


def test_rotate_rows_matches_oracle(gpu, oracle_mod):
    enu = synth.y_array(4, arm_m=10_000.0)
    uvw = synth.uvw_tracks(enu, 34.0, 30.0, 1.0, 20.0)
    n_rows, n_chan = uvw.shape[0], 24
    freq = np.linspace(1.0e9, 2.0e9, n_chan)
    rng = np.random.default_rng(9)
    vis = (rng.normal(size=(n_rows, n_chan)) + 1j * rng.normal(size=(n_rows, n_chan))).astype(np.complex64)
    image_dir, row_dir = (1.2, 0.52), (1.2 + 3e-3, 0.52 - 2e-3)
    uvrot, phrot = gpu.phase_rotation(image_dir, row_dir)
    ouvw, oshift, ovis = oracle_mod.rotate_rows(uvrot, phrot, uvw, freq, vis)

    v = torch.from_numpy(vis).cuda()
    guvw, gshift = gpu.rotate_rows(torch.from_numpy(uvw), freq, v, uvrot, phrot)
    assert np.array_equal(guvw.cpu().numpy(), ouvw)
    assert np.array_equal(gshift.cpu().numpy(), oshift)
    got = v.cpu().numpy()
    ulp = np.spacing(np.abs(ovis).astype(np.float32))
    assert np.all(np.abs(got - ovis) <= 2 * ulp)
    assert np.mean(got == ovis) > 0.999  # almost every sample bit-exact
    # the rotation is unitary: magnitudes are kept
    assert np.max(np.abs(np.abs(got) - np.abs(vis)) / np.abs(vis)) < 1e-6
