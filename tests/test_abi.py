"""CPU checks of the C ABI: the library loads, exports every symbol the header
declares, and fails loudly (no CPU fallback) when there is no device."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "casa_imaging_hip.h")


def header_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(chip_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree(pkg):
    assert sorted(pkg.abi.EXPORTS) == header_functions()


def test_header_constants_match_binding(pkg):
    text = open(HEADER).read()
    defs = dict(re.findall(r"#define (CHIP_[A-Z0-9_]+) \(?(-?\d+)\)?", text))
    for name in ("MODE_PSF", "MODE_DIRTY", "MODE_RESIDUAL", "F32", "F64", "PAD_COMPOSITE_1_2",
                 "PAD_COMPOSITE_1_3", "PAD_NONE", "CYCLE_AUTO", "CYCLE_HIPFFT", "CYCLE_FUSED",
                 "WEIGHT_UNIFORM", "WEIGHT_BRIGGS", "AW_OK", "AW_NON_FINITE_COORDINATE", "AW_OUTSIDE_GRID",
                 "AW_KERNEL_INDEX_OUTSIDE_CELL", "AW_INVALID_NORMALIZATION", "AW_NO_CELL"):
        assert int(defs["CHIP_" + name]) == getattr(pkg.abi, name), name


def test_library_exports_every_header_symbol(pkg):
    L = pkg.abi.load()
    for name in header_functions():
        assert hasattr(L, name), name
    assert L.chip_version().decode().startswith("casa-imaging-hip")


def test_error_path_without_device(pkg):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(pkg.Unsupported):
        pkg.Context(0)


def test_context_binds_the_callers_stream_once(pkg, monkeypatch):
    """Library work is launched on torch's current stream (bound lazily, once
    per stream change), so kernels are ordered after torch's work on the same
    tensors."""
    import importlib

    im = importlib.import_module("casa_rs_amd.imaging")
    calls = []
    streams = iter([11, 11, 22])

    class S:
        def __init__(self, v):
            self.cuda_stream = v

    monkeypatch.setattr(im.torch.cuda, "current_stream", lambda d=None: S(next(streams)))
    monkeypatch.setattr(im.abi, "call", lambda name, *a: calls.append((name, a[1], a[2].value)))
    c = im.Context.__new__(im.Context)
    c.device, c._h, c._bound = 0, "H", None
    assert [c.handle, c.handle, c.handle] == ["H", "H", "H"]
    assert calls == [("chip_ctx_set_stream", 1, 11), ("chip_ctx_set_stream", 1, 22)]
    c._h = None


def test_invalid_geometry_is_rejected_before_device_work(pkg):
    """A NULL context is an InvalidRequest, not a crash."""
    L = pkg.abi.load()
    geo = pkg.abi.Geometry(4, 4, 1e-5, 1e-5, 0, 100)
    out = ctypes.c_void_p()
    err = ctypes.c_void_p()
    rc = L.chip_gridder_create(None, ctypes.byref(geo), ctypes.byref(out), ctypes.byref(err))
    assert rc == pkg.abi.CHIP_ERR_INVALID
    assert b"NULL" in ctypes.cast(err, ctypes.c_char_p).value
    L.chip_error_free(err)


def test_phase_rotation_host_matches_oracle(pkg, oracle_mod):
    """chip_phase_rotation runs on the host (no device needed): FP64 with the
    reference's operation order, bit-exact with the oracle."""
    import numpy as np
    rng = np.random.default_rng(8)
    for _ in range(50):
        a = (rng.uniform(0, 2 * np.pi), rng.uniform(-1.5, 1.5))
        b = (rng.uniform(0, 2 * np.pi), rng.uniform(-1.5, 1.5))
        R, ph = pkg.phase_rotation(a, b)
        oR, oph = oracle_mod.phase_rotation(a, b)
        assert np.array_equal(R, oR) and np.array_equal(ph, oph)
