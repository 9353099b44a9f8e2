"""CPU oracle for the casa-imaging gridding / degridding path.

TEST INFRASTRUCTURE ONLY.  This package restates the reference CPU algorithm of
bglenden/casa-rs ``crates/casa-imaging`` so that the HIP library can be checked
against it.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU
baseline leg may import it, and only as the checker.  The product library
(``casa-rs_amd``) never imports, links or calls anything here.

The per-sample arithmetic lives in ``casa_oracle.c`` (built into
``oracle/_build/libcasa_oracle.so`` by ``oracle/Makefile``); this module adds
the array-level pieces with numpy:

* centered 2-D FFTs in float64 (pocketfft) in place of rustfft 6.4.1 -- the
  reference's FFT *bits* are not reproducible without rustfft, so every FFT
  result here is pinned to tolerance, not bit-exactly (``fft.rs:27-100``,
  ``:630-700``);
* ``WProjector`` kernel construction (``gridder.rs:3403-3586``): screen with
  float64 phase, forward FFT (float64 here, float32 in the reference), support
  from the 1e-3 threshold, crop, plane-zero normalisation;
* dirty-product finishing (``lib.rs:48987-49131`` standard,
  ``lib.rs:71009-71039`` / ``:70883-70885`` W-projection).

Parity status (which checks pin each function) is listed in the tests that
exercise it; functions whose results depend on FFT rounding are pinned to
tolerance only.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libcasa_oracle.so")
_lib = None

SUPPORT = 3
TAPS = 7


def build() -> str:
    """Compile the C oracle with its Makefile (gcc)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def use_native_build(directory) -> str:
    """Compile the oracle for this host (-march=native, same floating-point
    flags) into ``directory`` and route every later call through it; used to
    time the CPU baseline on the host that runs it.  Returns the path."""
    import subprocess

    global _lib
    subprocess.run(["make", "-s", "-C", _HERE, "native", f"NATIVE_DIR={directory}"], check=True)
    path = os.path.join(directory, "libcasa_oracle_native.so")
    L = ctypes.CDLL(path)
    _declare(L)
    _lib = L
    return path


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


class _Std(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64),
        ("cell_x", ctypes.c_double), ("cell_y", ctypes.c_double),
        ("gx", ctypes.c_uint64), ("gy", ctypes.c_uint64),
        ("blc_x", ctypes.c_uint64), ("blc_y", ctypes.c_uint64),
        ("os", ctypes.c_uint64),
        ("du", ctypes.c_double), ("dv", ctypes.c_double),
    ]


class _Sums(ctypes.Structure):
    _fields_ = [
        ("normalization_sumwt", ctypes.c_double),
        ("reported_sumwt", ctypes.c_double),
        ("gridded", ctypes.c_uint64),
        ("skipped", ctypes.c_uint64),
    ]


class _WProj(ctypes.Structure):
    _fields_ = [
        ("gx", ctypes.c_uint64), ("gy", ctypes.c_uint64),
        ("du", ctypes.c_double), ("dv", ctypes.c_double),
        ("sampling", ctypes.c_uint64),
        ("w_scale", ctypes.c_double),
        ("planes", ctypes.c_uint64),
        ("width", ctypes.c_uint64),
        ("support", ctypes.c_void_p),
        ("kernels", ctypes.c_void_p),
        ("norm_radius", ctypes.c_int64),
        ("norm_table", ctypes.c_void_p),
    ]


_P = ctypes.c_void_p
_U64 = ctypes.c_uint64
_D = ctypes.c_double


def _declare(L):
    L.co_grdsf.restype = ctypes.c_double
    L.co_grdsf.argtypes = [_D]
    L.co_spheroidal_kernel.restype = ctypes.c_float
    L.co_spheroidal_kernel.argtypes = [_D, _D]
    L.co_padded_len.restype = _U64
    L.co_padded_len.argtypes = [_U64, _D]
    L.co_is_composite.restype = ctypes.c_int
    L.co_is_composite.argtypes = [_U64]
    L.co_composite_padded_len.restype = _U64
    L.co_composite_padded_len.argtypes = [_U64, _D]
    for name in ("co_correction_axis", "co_gridder_correction_axis"):
        getattr(L, name).restype = None
        getattr(L, name).argtypes = [_U64, _P]
    L.co_sinc_axis.restype = None
    L.co_sinc_axis.argtypes = [_U64, _U64, _P]
    L.co_std_init.restype = ctypes.c_int
    L.co_std_init.argtypes = [ctypes.POINTER(_Std), _U64, _U64, _D, _D, ctypes.c_int, _U64]
    L.co_std_kernel_table.restype = None
    L.co_std_kernel_table.argtypes = [_U64, _P]
    L.co_std_tap_weights.restype = None
    L.co_std_tap_weights.argtypes = [_U64, _P]
    L.co_std_plan_many.restype = None
    L.co_std_plan_many.argtypes = [ctypes.POINTER(_Std), _U64, _P, _P, _P]
    L.co_std_grid.restype = None
    L.co_std_grid.argtypes = [ctypes.POINTER(_Std), _P, _U64, _P, _P, _P, _P, _P, _P, _P,
                              ctypes.c_int, ctypes.c_int, _P, _P, ctypes.POINTER(_Sums)]
    L.co_std_degrid.restype = None
    L.co_std_degrid.argtypes = [ctypes.POINTER(_Std), _P, _U64, _P, _P, _P, _P]
    L.co_std_predict_many.restype = None
    L.co_std_predict_many.argtypes = [ctypes.POINTER(_Std), _U64, _P, _P, _P, _P]
    L.co_w_plane_sum.restype = ctypes.c_float
    L.co_w_plane_sum.argtypes = [_P, _U64, ctypes.c_int64, _U64]
    L.co_w_sample_normalization.restype = ctypes.c_float
    L.co_w_sample_normalization.argtypes = [_P, _U64, ctypes.c_int64, _U64, ctypes.c_int64,
                                            ctypes.c_int64, ctypes.POINTER(ctypes.c_int)]
    L.co_w_plan_many.restype = None
    L.co_w_plan_many.argtypes = [ctypes.POINTER(_WProj), _U64, _P, _P, _P, _P, _P]
    L.co_w_grid.restype = None
    L.co_w_grid.argtypes = [ctypes.POINTER(_WProj), _U64, _P, _P, _P, _P, _P, _P, _P, _P,
                            ctypes.c_int, ctypes.c_int, _P, _P, ctypes.POINTER(_Sums)]
    L.co_w_grid64.restype = None
    L.co_w_grid64.argtypes = [ctypes.POINTER(_WProj), _U64, _P, _P, _P, _P, _P, _P, _P, _P,
                              ctypes.c_int, ctypes.c_int, _P, _P, ctypes.POINTER(_Sums)]
    L.co_aw_grid64.restype = None
    L.co_aw_grid64.argtypes = [_U64, _U64, _D, _D, _U64, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P,
                               _P, ctypes.c_int, _P, _P, _P]
    L.co_w_degrid.restype = None
    L.co_w_degrid.argtypes = [ctypes.POINTER(_WProj), _U64, _P, _P, _P, _P, _P]
    L.co_direct_adjoint.restype = None
    L.co_direct_adjoint.argtypes = [_U64, _U64, _D, _D, _U64, _P, _P, _P, _P, ctypes.c_int, _P]
    L.co_density_cell.restype = ctypes.c_int
    L.co_density_cell.argtypes = [_U64, _U64, _D, _D, ctypes.c_int, _D, _D, _P, _P]
    L.co_density_grid.restype = None
    L.co_density_grid.argtypes = [_U64, _U64, _D, _D, ctypes.c_int, ctypes.c_int, _U64, _P, _P, _P, _P, _P, _P]
    L.co_briggs_f2.restype = ctypes.c_float
    L.co_briggs_f2.argtypes = [_U64, _P, ctypes.c_float, ctypes.c_int]
    L.co_density_reweight.restype = None
    L.co_density_reweight.argtypes = [_U64, _U64, _D, _D, ctypes.c_int, ctypes.c_int, _U64, _P, _P, _P, _P, _P,
                                      _P, _P, _P]
    L.co_uvw_rotation_matrix.restype = None
    L.co_uvw_rotation_matrix.argtypes = [_P, _P, _P]
    L.co_phase_rotation_vector.restype = None
    L.co_phase_rotation_vector.argtypes = [_P, _P, _P]
    L.co_rotate_rows.restype = None
    L.co_rotate_rows.argtypes = [_U64, _U64, _P, _P, _P, _P, _P, _P, _P]
    L.co_direct_adjoint_window.restype = None
    L.co_direct_adjoint_window.argtypes = [_U64, _U64, _D, _D, _U64, _U64, _U64, _U64, _U64, _P, _P, _P, _P,
                                           ctypes.c_int, _P]
    L.co_direct_predict.restype = None
    L.co_direct_predict.argtypes = [_U64, _U64, _D, _D, _P, _U64, _P, _P, _P, _P]
    L.co_aw_kernel_origin.restype = _U64
    L.co_aw_kernel_origin.argtypes = [_U64]
    L.co_aw_plan.restype = ctypes.c_int
    L.co_aw_plan.argtypes = [_U64, _U64, _D, _D, _U64, _P, _P, _D, _D, _D, _P, _P, _P]
    L.co_aw_taps.restype = None
    L.co_aw_taps.argtypes = [_U64, _P, _P, _P, _P, _P]
    L.co_aw_grid_f64.restype = None
    L.co_aw_grid_f64.argtypes = [_U64, _U64, _D, _D, _U64, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P]
    L.co_aw_degrid.restype = None
    L.co_aw_degrid.argtypes = [_U64, _U64, _D, _D, _U64, _P, _P, _P, _U64, _P, _P, _P, _P, _P, _P, _P, _P]
    L.co_airy_max_radius.restype = _D
    L.co_airy_max_radius.argtypes = [_D, _D]
    L.co_airy_from_index.restype = ctypes.c_float
    L.co_airy_from_index.argtypes = [_D, _D, _D, _D]
    L.co_airy_table.restype = None
    L.co_airy_table.argtypes = [_D, _D, _P]
    L.co_airy_offsets.restype = ctypes.c_float
    L.co_airy_offsets.argtypes = [_D, _D, _D, _D, _D, _P]
    L.co_airy_radius.restype = ctypes.c_float
    L.co_airy_radius.argtypes = [_D, _D, _D, _D]
    L.co_airy_offsets_many.restype = None
    L.co_minor_cycle_stop.argtypes = [ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float]
    L.co_minor_cycle_stop.restype = ctypes.c_int
    L.co_hogbom.argtypes = [_U64, _U64, _U64, _P, _U64, _U64, _P, _P, _P, ctypes.c_float, ctypes.c_float, _P, _P,
                            _U64, _P, _P]
    L.co_hogbom.restype = None
    L.co_airy_offsets_many.argtypes = [_U64, _P, _P, _D, _D, _D, _P, _P]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64_as_f32(a):
    """complex64 array viewed as interleaved float32 (re, im)."""
    a = np.ascontiguousarray(a, dtype=np.complex64)
    return a.view(np.float32)


def default_threads() -> int:
    return max(1, len(os.sched_getaffinity(0)))


# --------------------------------------------------------------------------
# Closed forms (gridder.rs:4152-4268)
# --------------------------------------------------------------------------

def grdsf(nu: float) -> float:
    return lib().co_grdsf(float(nu))


def spheroidal_kernel(distance: float, support: float) -> float:
    return lib().co_spheroidal_kernel(float(distance), float(support))


def padded_len(n: int, padding: float) -> int:
    return int(lib().co_padded_len(n, padding))


def is_composite(n: int) -> bool:
    return bool(lib().co_is_composite(n))


def composite_padded_len(n: int, padding: float) -> int:
    return int(lib().co_composite_padded_len(n, padding))


def correction_axis(n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().co_correction_axis(n, _ptr(out))
    return out


def gridder_correction_axis(n: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().co_gridder_correction_axis(n, _ptr(out))
    return out


def sinc_axis(n: int, sampling: int) -> np.ndarray:
    out = np.empty(n, np.float32)
    lib().co_sinc_axis(n, sampling, _ptr(out))
    return out


# --------------------------------------------------------------------------
# Centered FFTs (fft.rs:27-100, :630-700, :820-839)
# --------------------------------------------------------------------------

def centered_ifft2(grid: np.ndarray) -> np.ndarray:
    """fftshift(ifft2(ifftshift(grid))), scaled 1/(nx*ny), in float64."""
    g = np.asarray(grid, dtype=np.complex128)
    return np.fft.fftshift(np.fft.ifft2(np.fft.ifftshift(g)))


def centered_fft2(grid: np.ndarray) -> np.ndarray:
    """Unnormalised forward centered 2-D FFT in float64 (fft.rs:34)."""
    g = np.asarray(grid, dtype=np.complex128)
    return np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(g)))


def peak_abs_value(image: np.ndarray) -> np.float32:
    """lib.rs:75036-75040 (max |v| folded from 0)."""
    a = np.abs(np.asarray(image, dtype=np.float32))
    return np.float32(max(np.float32(0.0), np.nanmax(a) if a.size else np.float32(0.0)))


# --------------------------------------------------------------------------
# StandardGridder (gridder.rs:322-2070)
# --------------------------------------------------------------------------

@dataclass
class Sums:
    normalization_sumwt: float = 0.0
    reported_sumwt: float = 0.0
    gridded: int = 0
    skipped: int = 0


class StandardGridder:
    """Restatement of ``StandardGridder`` (gridder.rs:322-443)."""

    PAD_COMPOSITE_1_2 = 0
    PAD_COMPOSITE_1_3 = 1
    PAD_NONE = 2

    def __init__(self, nx, ny, cell_x, cell_y, padding_mode=0, oversampling=100):
        self._s = _Std()
        rc = lib().co_std_init(ctypes.byref(self._s), nx, ny, cell_x, cell_y, padding_mode,
                               oversampling)
        if rc != 0:
            raise ValueError("invalid image geometry")
        self.os = oversampling
        self.taps = np.empty((oversampling // 2 * 2 + 1, TAPS), np.float32)
        lib().co_std_tap_weights(oversampling, _ptr(self.taps))
        self.kernel_table = np.empty(oversampling * (SUPPORT + 1), np.float32)
        lib().co_std_kernel_table(oversampling, _ptr(self.kernel_table))
        self.correction_x = correction_axis(self.gx)
        self.correction_y = correction_axis(self.gy)

    nx = property(lambda s: s._s.nx)
    ny = property(lambda s: s._s.ny)
    gx = property(lambda s: s._s.gx)
    gy = property(lambda s: s._s.gy)
    blc = property(lambda s: (s._s.blc_x, s._s.blc_y))
    du = property(lambda s: s._s.du)
    dv = property(lambda s: s._s.dv)
    cell = property(lambda s: (s._s.cell_x, s._s.cell_y))

    @property
    def grid_shape(self):
        return (self.gx, self.gy)

    def plan(self, u, v) -> np.ndarray:
        """plan_positive_taps per sample: (x_start, y_start, wix, wiy), -1 if rejected."""
        u, v = _f64(u), _f64(v)
        out = np.empty((len(u), 4), np.int64)
        lib().co_std_plan_many(ctypes.byref(self._s), len(u), _ptr(u), _ptr(v), _ptr(out))
        return out

    def grid(self, u, v, weight, sumwt_factor, vis, gridable=None, mode=1, model=None,
             threads=None, psf=None, dirty=None):
        """Accumulate into f64 complex grids; returns (psf, dirty, Sums).

        mode 0: PSF, 1: PSF + dirty (execution.rs:18255-18338),
        2: degrid ``model`` (complex64 grid) and grid the residual
        (execution.rs:8890-8940, gridder.rs:1315-1396)."""
        n = len(u)
        u, v = _f64(u), _f64(v)
        weight, sumwt_factor = _f32(weight), _f32(sumwt_factor)
        vis_f = _c64_as_f32(vis)
        g = None if gridable is None else np.ascontiguousarray(gridable, dtype=np.uint8)
        shape = (self.gx, self.gy)
        if psf is None:
            psf = np.zeros(shape, np.complex128)
        if dirty is None:
            dirty = np.zeros(shape, np.complex128)
        m = None if model is None else _c64_as_f32(model)
        sums = _Sums()
        lib().co_std_grid(ctypes.byref(self._s), _ptr(self.taps), n, _ptr(u), _ptr(v),
                          _ptr(weight), _ptr(sumwt_factor), _ptr(g), _ptr(vis_f), _ptr(m), mode,
                          threads or default_threads(), _ptr(psf), _ptr(dirty), ctypes.byref(sums))
        return psf, dirty, Sums(sums.normalization_sumwt, sums.reported_sumwt, sums.gridded,
                                sums.skipped)

    def degrid(self, u, v, model) -> np.ndarray:
        """degrid_sample_taps_planned_normalized (gridder.rs:1278-1312); NaN if rejected."""
        u, v = _f64(u), _f64(v)
        m = _c64_as_f32(model)
        out = np.empty(len(u), np.complex64)
        lib().co_std_degrid(ctypes.byref(self._s), _ptr(self.taps), len(u), _ptr(u), _ptr(v),
                            _ptr(m), _ptr(out.view(np.float32)))
        return out

    def predict_sectdgrid(self, u, v, model_grid) -> np.ndarray:
        """degrid_sample_product_planned_sectdgrid (gridder.rs:1227-1266) per
        sample, 0 where the sample has no taps (StandardMfsModelPredictor::
        predict, lib.rs:3387-3403)."""
        u, v = _f64(u), _f64(v)
        m = _c64_as_f32(model_grid)
        out = np.empty(len(u), np.complex64)
        lib().co_std_predict_many(ctypes.byref(self._s), len(u), _ptr(u), _ptr(v), _ptr(m),
                                  _ptr(out.view(np.float32)))
        return out

    def apodize_model(self, model: np.ndarray) -> np.ndarray:
        """gridder.rs:1554-1570: model x correction placed at the image origin."""
        out = np.zeros((self.gx, self.gy), np.complex64)
        bx, by = self.blc
        corr = np.outer(self.correction_x[bx:bx + self.nx],
                        self.correction_y[by:by + self.ny]).astype(np.float32)
        out[bx:bx + self.nx, by:by + self.ny] = (np.asarray(model, np.float32) * corr)
        return out

    def apodize_w_project_model(self, model, sampling):
        """gridder.rs:1572-1592 (correction x sinc per axis)."""
        sx = sinc_axis(self.gx, sampling)
        sy = sinc_axis(self.gy, sampling)
        bx, by = self.blc
        cx = self.correction_x[bx:bx + self.nx]
        cy = self.correction_y[by:by + self.ny]
        # factor = cx*cy*sx*sy evaluated left to right in f32
        f = (cx[:, None] * cy[None, :]).astype(np.float32)
        f = (f * sx[bx:bx + self.nx][:, None]).astype(np.float32)
        f = (f * sy[by:by + self.ny][None, :]).astype(np.float32)
        out = np.zeros((self.gx, self.gy), np.complex64)
        out[bx:bx + self.nx, by:by + self.ny] = np.asarray(model, np.float32) * f
        return out

    def model_grid(self, model: np.ndarray) -> np.ndarray:
        """centered_fft2(apodize_model(model)) as complex64 (lib.rs:68878-68885)."""
        return centered_fft2(self.apodize_model(model)).astype(np.complex64)

    def corrected_image_from_grid_f64(self, raw: np.ndarray) -> np.ndarray:
        """gridder.rs:1646-1657: Re(raw) narrowed to f32 times (cx*cy)."""
        bx, by = self.blc
        re = raw[bx:bx + self.nx, by:by + self.ny].real.astype(np.float32)
        corr = (self.correction_x[bx:bx + self.nx][:, None]
                * self.correction_y[by:by + self.ny][None, :]).astype(np.float32)
        return (re * corr).astype(np.float32)

    def corrected_w_project_image_from_grid(self, raw: np.ndarray, sampling: int) -> np.ndarray:
        """gridder.rs:1685-1715: Re(raw * (cx*cy)/(sinc_x*sinc_y)), 0 where |sinc| <= 1e-6."""
        sinc = sinc_axis(max(self.gx, self.gy), sampling)
        bx, by = self.blc
        raw32 = np.asarray(raw, np.complex64)[bx:bx + self.nx, by:by + self.ny]
        factor = (self.correction_x[bx:bx + self.nx][:, None]
                  * self.correction_y[by:by + self.ny][None, :]).astype(np.float32)
        sf = (sinc[bx:bx + self.nx][:, None] * sinc[by:by + self.ny][None, :]).astype(np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            ratio = (factor / sf).astype(np.float32)
        img = (raw32.real * ratio).astype(np.float32)
        return np.where(np.abs(sf) > np.float32(1e-6), img, np.float32(0.0)).astype(np.float32)

    def dirty_products(self, psf_grid, dirty_grid, sums: Sums):
        """dirty_grids_to_psf_and_residual_owned (lib.rs:48987-49131), f64 FFT."""
        if sums.normalization_sumwt <= 0 or sums.reported_sumwt <= 0:
            raise ValueError("no usable samples")
        nsum = np.float32(sums.normalization_sumwt)
        psf = self.corrected_image_from_grid_f64(centered_ifft2(psf_grid))
        psf = (psf / nsum).astype(np.float32)
        peak = peak_abs_value(psf)
        psf = (psf / peak).astype(np.float32)
        res = self.corrected_image_from_grid_f64(centered_ifft2(dirty_grid))
        res = ((res / nsum) / peak).astype(np.float32)
        return psf, res, peak


class StandardMfsModelPredictor:
    """StandardMfsModelPredictor (lib.rs:3375-3404): a gridder with CASA's
    composite padding 1.3 (new_with_casa_composite_padding, gridder.rs:379-383)
    and the reference's oversampling 100, the model grid
    centered_fft2(apodize_model(model)) when the model has a non-zero pixel,
    and predict(u, v) by the tap-sum-normalised degrid."""

    def __init__(self, nx, ny, cell_x, cell_y, model, oversampling=100):
        self.gridder = StandardGridder(nx, ny, cell_x, cell_y, padding_mode=StandardGridder.PAD_COMPOSITE_1_3,
                                       oversampling=oversampling)
        model = np.asarray(model, np.float32)
        self.model_grid = self.gridder.model_grid(model) if np.any(np.abs(model) > 0) else None

    def predict(self, u, v, model_grid=None):
        """Predicted visibilities (complex64); ``model_grid`` overrides the
        predictor's own (e.g. a grid made elsewhere, to compare the degrid
        alone)."""
        grid = self.model_grid if model_grid is None else model_grid
        if grid is None:
            return np.zeros(len(u), np.complex64)
        return self.gridder.predict_sectdgrid(u, v, grid)


# --------------------------------------------------------------------------
# WProjector (gridder.rs:3391-3830)
# --------------------------------------------------------------------------

W_MAX_SAFETY = 1.05


def suggested_w_plane_count(nx, ny, cell_x, cell_y, max_abs_w) -> int:
    """gridder.rs:3822-3830."""
    if not (np.isfinite(max_abs_w) and max_abs_w > 0.0):
        return 1
    half = max(nx * abs(cell_x), ny * abs(cell_y)) / 2.0
    return int(W_MAX_SAFETY * max_abs_w * abs(np.sin(half)))


def w_max_uvw(cell_x, max_abs_w, explicit_planes) -> float:
    """gridder.rs:3808-3820."""
    if explicit_planes is not None:
        return 0.25 / abs(cell_x)
    return W_MAX_SAFETY * max_abs_w


def find_w_support(transformed: np.ndarray, peak: np.float32, sampling: int) -> int:
    """find_w_project_support_from_transformed, gridder.rs:3832-3855."""
    conv = transformed.shape[0]
    quarter = conv // 2 - 1
    t32 = np.asarray(transformed, np.complex64)
    col = np.abs((t32[1:quarter, 0] / peak).astype(np.complex64))
    row = np.abs((t32[0, 1:quarter] / peak).astype(np.complex64))
    hit = np.nonzero((col > np.float32(1e-3)) | (row > np.float32(1e-3)))[0]
    max_support = quarter // sampling - 1
    if hit.size == 0:
        return max(max_support, 1)
    trial = int(hit[-1]) + 1
    sup = int(np.floor(np.float32(np.float32(trial) / np.float32(sampling)) + np.float32(0.5))) + 1
    return min(sup, max(max_support, 1))


def w_screen(conv, inner, s0, s1, w_lambda, corr) -> np.ndarray:
    """Screen of WProjector::new (gridder.rs:3438-3471), complex64, origin-wrapped."""
    half = inner // 2
    idx = np.arange(-half, half)
    l = s0 * idx.astype(np.float64)
    m = s1 * idx.astype(np.float64)
    rsq = l[:, None] * l[:, None] + (m * m)[None, :]
    inside = rsq < 1.0
    with np.errstate(invalid="ignore"):
        phase = 2.0 * np.pi * w_lambda * (np.sqrt(1.0 - rsq) - 1.0)
    c = corr[idx + half]
    correction = (c[:, None] * c[None, :]).astype(np.float32)
    re = (correction * np.cos(phase).astype(np.float32)).astype(np.float32)
    im = (correction * np.sin(phase).astype(np.float32)).astype(np.float32)
    screen = np.zeros((conv, conv), np.complex64)
    xs = np.where(idx >= 0, idx, idx + conv)
    vals = (re + 1j * im).astype(np.complex64)
    vals[~inside] = 0
    screen[np.ix_(xs, xs)] = vals
    return screen


@dataclass
class WProjector:
    """Restatement of ``WProjector::new`` (gridder.rs:3403-3586).

    The FFT is float64 here and float32 (rustfft) in the reference, so kernel
    values agree to FFT tolerance; the support integers agree unless a value
    lies within that tolerance of the 1e-3 threshold (``support_margin`` records
    how far the deciding samples are from it)."""

    gridder: StandardGridder
    max_abs_w: float
    explicit_planes: int | None = None
    sampling: int | None = None
    planes: int = 0
    w_scale: float = 1.0
    width: int = 0
    supports: np.ndarray = field(default_factory=lambda: np.zeros(0, np.int32))
    kernels: np.ndarray = field(default_factory=lambda: np.zeros((0, 0, 0), np.complex64))
    norm_radius: int = 0
    norm_table: np.ndarray = field(default_factory=lambda: np.zeros(0, np.float32))
    support_margin: list = field(default_factory=list)
    # Sign of the screen phase.  +1 is the reference's (gridder.rs:3453); a
    # test sets -1 to plant a sign mistake and show that its checks catch it.
    screen_sign: float = 1.0

    def __post_init__(self):
        g = self.gridder
        nx, ny = g.nx, g.ny
        cx, cy = g.cell
        if self.explicit_planes is not None:
            planes = self.explicit_planes
        else:
            planes = max(suggested_w_plane_count(nx, ny, cx, cy, self.max_abs_w), 1)
        self.planes = planes
        if self.sampling is None:
            self.sampling = 4 if planes > 1 else 1  # gridder.rs:3413
        s = self.sampling
        gx, gy = g.gx, g.gy
        conv = max(gx, gy)
        inner = max(conv // s, 2)
        corr = gridder_correction_axis(inner)
        s0 = abs(cx) * s * gx / conv
        s1 = abs(cy) * s * gy / conv
        maxw = w_max_uvw(cx, self.max_abs_w, self.explicit_planes)
        if planes > 1 and np.isfinite(maxw) and maxw > 0:
            self.w_scale = float((planes - 1) * (planes - 1)) / maxw
        else:
            self.w_scale = 1.0
        R = (s + 1) // 2
        self.norm_radius = R
        quarter = conv // 2 - 1
        raw = []
        sups = []
        peak0 = None
        for k in range(planes):
            w = float(k * k) / self.w_scale if planes > 1 else 0.0
            screen = w_screen(conv, inner, s0, s1, self.screen_sign * w, corr)
            t = np.fft.fft2(screen.astype(np.complex128)).astype(np.complex64)
            pk = np.float32(np.abs(t[0, 0]))
            if peak0 is None:
                peak0 = pk
            if not peak0 > 0:
                raise ValueError("wproject plane-zero kernel peak is zero")
            sup = find_w_support(t, peak0, s)
            mag = np.abs(np.concatenate([t[1:quarter, 0], t[0, 1:quarter]]) / peak0)
            self.support_margin.append(float(np.min(np.abs(mag - 1e-3))) if mag.size else 0.0)
            cw = max(min(sup * s + R + 1, quarter), 1)
            raw.append((t[:cw, :cw] / peak0).astype(np.complex64))
            sups.append(sup)
        width = max(r.shape[0] for r in raw)
        kern = np.zeros((planes, width, width), np.complex64)
        for k, r in enumerate(raw):
            kern[k, :r.shape[0], :r.shape[1]] = r
        self.supports = np.asarray(sups, np.int32)
        self.width = width
        k0 = np.ascontiguousarray(kern[0])
        pb = lib().co_w_plane_sum(_ptr(k0.view(np.float32)), width, int(sups[0]), s)
        pb = np.float32(pb)
        if not (np.isfinite(pb) and pb > 0):
            raise ValueError("wproject plane-zero kernel integral is non-positive")
        kern = (kern / pb).astype(np.complex64)
        self.kernels = np.ascontiguousarray(kern)
        axis = 2 * R + 1
        table = np.zeros(planes * axis * axis, np.float32)
        ok = ctypes.c_int()
        for k in range(planes):
            kk = np.ascontiguousarray(self.kernels[k])
            base = k * axis * axis
            for ox in range(-R, R + 1):
                for oy in range(-R, R + 1):
                    val = lib().co_w_sample_normalization(
                        _ptr(kk.view(np.float32)), width, int(sups[k]), s, ox, oy, ctypes.byref(ok))
                    table[base + (ox + R) * axis + (oy + R)] = val if ok.value else 0.0
        self.norm_table = table
        self._make_struct()

    def _make_struct(self):
        g = self.gridder
        self._sup32 = np.ascontiguousarray(self.supports, np.int32)
        self._k32 = np.ascontiguousarray(self.kernels).view(np.float32)
        self._n32 = np.ascontiguousarray(self.norm_table, np.float32)
        self._s = _WProj(g.gx, g.gy, g.du, g.dv, self.sampling, self.w_scale, self.planes,
                         self.width, self._sup32.ctypes.data, self._k32.ctypes.data,
                         self.norm_radius, self._n32.ctypes.data)

    @classmethod
    def from_tables(cls, gridder, planes, sampling, w_scale, kernels, supports, norm_table):
        """A projector over given tables (no kernel construction)."""
        self = cls.__new__(cls)
        self.gridder, self.max_abs_w, self.explicit_planes = gridder, float("nan"), planes
        self.planes, self.sampling, self.w_scale = int(planes), int(sampling), float(w_scale)
        self.norm_radius = (int(sampling) + 1) // 2
        self.support_margin = []
        return self.with_kernels(kernels, supports, norm_table)

    def with_kernels(self, kernels, supports, norm_table):
        """Return a copy that uses externally built kernels (e.g. the GPU's), so
        that gridding can be compared independently of FFT rounding."""
        import copy

        other = copy.copy(self)
        other.kernels = np.ascontiguousarray(kernels, np.complex64)
        other.supports = np.ascontiguousarray(supports, np.int32)
        other.norm_table = np.ascontiguousarray(norm_table, np.float32)
        other.width = other.kernels.shape[1]
        other._make_struct()
        return other

    def plan(self, u, v, w):
        """WProjector::plan_sample: (loc_x, loc_y, off_x, off_y, plane, conj), norm."""
        u, v, w = _f64(u), _f64(v), _f64(w)
        out = np.empty((len(u), 6), np.int64)
        norm = np.empty(len(u), np.float32)
        lib().co_w_plan_many(ctypes.byref(self._s), len(u), _ptr(u), _ptr(v), _ptr(w), _ptr(out),
                             _ptr(norm))
        return out, norm

    def grid(self, u, v, w, weight, sumwt_factor, vis, gridable=None, mode=1, model=None,
             threads=None, psf=None, dirty=None):
        """Complex32 grids as the reference (lib.rs:71472-71540).  mode 0 PSF,
        1 PSF + dirty, 2 residual against ``model`` (complex64 grid or None)."""
        n = len(u)
        u, v, w = _f64(u), _f64(v), _f64(w)
        weight, sumwt_factor = _f32(weight), _f32(sumwt_factor)
        vis_f = _c64_as_f32(vis)
        g = None if gridable is None else np.ascontiguousarray(gridable, dtype=np.uint8)
        shape = (self.gridder.gx, self.gridder.gy)
        if psf is None:
            psf = np.zeros(shape, np.complex64)
        if dirty is None:
            dirty = np.zeros(shape, np.complex64)
        m = None if model is None else _c64_as_f32(model)
        sums = _Sums()
        lib().co_w_grid(ctypes.byref(self._s), n, _ptr(u), _ptr(v), _ptr(w), _ptr(weight),
                        _ptr(sumwt_factor), _ptr(g), _ptr(vis_f), _ptr(m), mode,
                        threads or default_threads(), _ptr(psf.view(np.float32)),
                        _ptr(dirty.view(np.float32)), ctypes.byref(sums))
        return psf, dirty, Sums(sums.normalization_sumwt, sums.reported_sumwt, sums.gridded,
                                sums.skipped)

    def grid64(self, u, v, w, weight, sumwt_factor, vis, gridable=None, mode=1, model=None,
               threads=None, psf=None, dirty=None):
        """``grid`` accumulated in FP64 (complex128 grids, f64 products; the
        model grid, for mode 2, complex128): the FP64 reference FP32 results
        are measured against, not the reference's Complex32 W grids."""
        n = len(u)
        u, v, w = _f64(u), _f64(v), _f64(w)
        weight, sumwt_factor = _f32(weight), _f32(sumwt_factor)
        vis_f = _c64_as_f32(vis)
        g = None if gridable is None else np.ascontiguousarray(gridable, dtype=np.uint8)
        shape = (self.gridder.gx, self.gridder.gy)
        if psf is None:
            psf = np.zeros(shape, np.complex128)
        if dirty is None:
            dirty = np.zeros(shape, np.complex128)
        m = None if model is None else np.ascontiguousarray(model, np.complex128).view(np.float64)
        sums = _Sums()
        lib().co_w_grid64(ctypes.byref(self._s), n, _ptr(u), _ptr(v), _ptr(w), _ptr(weight),
                          _ptr(sumwt_factor), _ptr(g), _ptr(vis_f), _ptr(m), mode,
                          threads or default_threads(), _ptr(psf.view(np.float64)),
                          _ptr(dirty.view(np.float64)), ctypes.byref(sums))
        return psf, dirty, Sums(sums.normalization_sumwt, sums.reported_sumwt, sums.gridded,
                                sums.skipped)

    def model_grid64(self, model):
        """model_grid with the transform kept in FP64 (complex128)."""
        return w_model_grid64(self.gridder, self.sampling, model)

    def dirty_products64(self, psf_grid, dirty_grid, sums: Sums, psf_peak=None):
        """dirty_products in FP64 end to end (see w_dirty_products64)."""
        return w_dirty_products64(self.gridder, self.sampling, psf_grid, dirty_grid,
                                  sums.normalization_sumwt, psf_peak)

    def degrid(self, u, v, w, model):
        u, v, w = _f64(u), _f64(v), _f64(w)
        m = _c64_as_f32(model)
        out = np.empty(len(u), np.complex64)
        lib().co_w_degrid(ctypes.byref(self._s), len(u), _ptr(u), _ptr(v), _ptr(w), _ptr(m),
                          _ptr(out.view(np.float32)))
        return out

    def model_grid(self, model):
        """centered_fft2(apodize_w_project_model(model)) (lib.rs:70909-70913)."""
        return centered_fft2(self.gridder.apodize_w_project_model(model, self.sampling)).astype(
            np.complex64)

    def dirty_products(self, psf_grid, dirty_grid, sums: Sums):
        """raw_psf_grid_to_w_project_state + residual (lib.rs:70870-70885, :71009-71039).

        The reference transforms the f32 grids in f64 and narrows the result
        to f32 (execute_dirty_product_host_f32_batch, lib.rs:2610-2641)."""
        if sums.normalization_sumwt <= 0 or sums.reported_sumwt <= 0:
            raise ValueError("no usable samples")
        g = self.gridder
        s = self.sampling
        nsum = np.float32(sums.normalization_sumwt)
        raw_psf = centered_ifft2(psf_grid).astype(np.complex64)
        psf = g.corrected_w_project_image_from_grid(raw_psf, s)
        psf = ((np.float32(2.0) * psf) / nsum).astype(np.float32)
        peak = peak_abs_value(psf)
        psf = (psf / peak).astype(np.float32)
        raw_res = centered_ifft2(dirty_grid).astype(np.complex64)
        res = g.corrected_w_project_image_from_grid(raw_res, s)
        res = (((np.float32(2.0) * res) / nsum) / peak).astype(np.float32)
        return psf, res, peak


def w_model_grid64(gridder, sampling, model):
    """centered_fft2(apodize_w_project_model(model)) (lib.rs:70909-70913) with
    the transform kept in FP64 (complex128)."""
    return centered_fft2(gridder.apodize_w_project_model(model, sampling))


def w_dirty_products64(gridder, sampling, psf_grid, dirty_grid, nsum, psf_peak=None):
    """The W-projection dirty products (lib.rs:71009-71039, gridder.rs:1685-1715)
    in FP64 end to end: f64 transform, correction cx*cy/sinc (the f32
    correction and sinc values of the reference, 0 where |sinc| <= 1e-6),
    2 v / nsum, PSF peak, residual / peak.  With ``psf_grid`` None,
    ``psf_peak`` normalises the residual.  Returns (psf, residual, peak)."""
    g = gridder
    sinc = sinc_axis(max(g.gx, g.gy), sampling).astype(np.float64)
    bx, by = g.blc
    nx, ny = g.nx, g.ny
    cx = g.correction_x[bx:bx + nx].astype(np.float64)
    cy = g.correction_y[by:by + ny].astype(np.float64)
    sf = sinc[bx:bx + nx][:, None] * sinc[by:by + ny][None, :]
    ok = np.abs(sf) > 1e-6
    factor = np.where(ok, (cx[:, None] * cy[None, :]) / np.where(ok, sf, 1.0), 0.0)

    def image(grid):
        return centered_ifft2(np.asarray(grid, np.complex128))[bx:bx + nx, by:by + ny].real * factor

    psf = None
    if psf_grid is not None:
        psf = 2.0 * image(psf_grid) / float(nsum)
        peak = float(np.max(np.abs(psf)))
        psf = psf / peak
    else:
        peak = float(psf_peak)
    res = 2.0 * image(dirty_grid) / float(nsum) / peak
    return psf, res, peak


# --------------------------------------------------------------------------
# Direct DFT (lib.rs:73709-73910)
# --------------------------------------------------------------------------

def direct_adjoint(nx, ny, cell_x, cell_y, u, v, w, values, threads=None, image=None, w_sign=1.0):
    """accumulate_direct_adjoint over samples: image += 2 Re(value e^{i phase}).

    ``w_sign`` multiplies w (+1 is the reference's phase, lib.rs:73909-73910;
    -1 evaluates the transform in the opposite w convention)."""
    u, v, w = _f64(u), _f64(v), _f64(np.asarray(w, np.float64) * w_sign)
    vals = _c64_as_f32(values)
    if image is None:
        image = np.zeros((nx, ny), np.float32)
    lib().co_direct_adjoint(nx, ny, cell_x, cell_y, len(u), _ptr(u), _ptr(v), _ptr(w), _ptr(vals),
                            threads or default_threads(), _ptr(image))
    return image


def direct_adjoint_window(nx, ny, cell_x, cell_y, x0, y0, wx, wy, u, v, w, values, threads=None):
    """direct_adjoint restricted to the pixels [x0, x0+wx) x [y0, y0+wy)."""
    u, v, w = _f64(u), _f64(v), _f64(w)
    vals = _c64_as_f32(values)
    image = np.zeros((wx, wy), np.float32)
    lib().co_direct_adjoint_window(nx, ny, cell_x, cell_y, x0, y0, wx, wy, len(u), _ptr(u), _ptr(v),
                                   _ptr(w), _ptr(vals), threads or default_threads(), _ptr(image))
    return image


def direct_predict(nx, ny, cell_x, cell_y, model, u, v, w):
    """direct_predict_visibility for every sample (lib.rs:73870-73887)."""
    u, v, w = _f64(u), _f64(v), _f64(w)
    m = _f32(model)
    out = np.empty(len(u), np.complex64)
    lib().co_direct_predict(nx, ny, cell_x, cell_y, _ptr(m), len(u), _ptr(u), _ptr(v), _ptr(w),
                            _ptr(out.view(np.float32)))
    return out


def direct_psf_and_residual(nx, ny, cell_x, cell_y, u, v, w, weight, sumwt_factor, vis,
                            model=None, gridable=None, threads=None, w_sign=1.0):
    """compute_psf_direct + compute_residual_direct (lib.rs:73709-73823).

    ``w_sign`` multiplies w before the transform (+1 is the reference's)."""
    n = len(u)
    w = np.asarray(w, np.float64) * w_sign
    weight = _f32(weight)
    sumwt_factor = _f32(sumwt_factor)
    vis = np.asarray(vis, np.complex64)
    g = np.ones(n, bool) if gridable is None else np.asarray(gridable, bool)
    ok_psf = g & np.isfinite(weight) & (weight > 0) & np.isfinite(sumwt_factor) & (sumwt_factor > 0)
    nsum = float(np.sum(2.0 * weight[ok_psf].astype(np.float64)))
    psf = direct_adjoint(nx, ny, cell_x, cell_y, u[ok_psf], v[ok_psf], w[ok_psf],
                         weight[ok_psf].astype(np.complex64), threads)
    psf = (psf / np.float32(nsum)).astype(np.float32)
    peak = peak_abs_value(psf)
    psf = (psf / peak).astype(np.float32)
    ok = g & np.isfinite(weight) & (weight > 0) & np.isfinite(vis.real) & np.isfinite(vis.imag)
    if model is not None and np.any(np.abs(model) > 0):
        pred = direct_predict(nx, ny, cell_x, cell_y, model, u[ok], v[ok], w[ok])
    else:
        pred = np.zeros(int(ok.sum()), np.complex64)
    resid = ((vis[ok] - pred) * weight[ok]).astype(np.complex64)
    img = direct_adjoint(nx, ny, cell_x, cell_y, u[ok], v[ok], w[ok], resid, threads)
    img = ((img / np.float32(nsum)) / peak).astype(np.float32)
    return psf, img, peak, nsum


# ---------------------------------------------------------------------------
# density weighting (weighting.rs)
# ---------------------------------------------------------------------------

VIS_IMAGING_WEIGHT = 0   # DensityCellConvention::VisImagingWeight (Float ops, truncation)
CUBE_BRIGGS = 1          # CubeBriggsWeightorDensity / Lookup (Double, round half away)
WEIGHT_UNIFORM = 1
WEIGHT_BRIGGS = 2


def density_cell(nx, ny, cell_x, cell_y, convention, u, v):
    """weight_density_cell_anchor (gridder.rs:1888-1934): (x, y) or None."""
    x, y = ctypes.c_int64(), ctypes.c_int64()
    ok = lib().co_density_cell(nx, ny, cell_x, cell_y, convention, float(u), float(v), ctypes.byref(x),
                               ctypes.byref(y))
    return (x.value, y.value) if ok else None


def density_weights(nx, ny, cell_x, cell_y, u, v, weight, mode, robust=0.0, plane=None, n_planes=1,
                    usable=None):
    """Uniform / Briggs reweighting of every sample (weighting.rs:1180-1232).

    plane None: one combined density (VisImagingWeight convention, conjugates
    included); else per-plane densities (cube convention: build with
    conjugates, look up without).  Returns (weights, density, f2)."""
    u, v = _f64(u), _f64(v)
    weight = _f32(weight)
    n = len(u)
    per_plane = plane is not None
    conv = CUBE_BRIGGS if per_plane else VIS_IMAGING_WEIGHT
    pl = None if plane is None else np.ascontiguousarray(plane, dtype=np.uint32)
    us = None if usable is None else np.ascontiguousarray(usable, dtype=np.uint8)
    planes = n_planes if per_plane else 1
    density = np.zeros((planes, nx, ny), np.float32)
    lib().co_density_grid(nx, ny, cell_x, cell_y, conv, 1, n, _ptr(u), _ptr(v), _ptr(weight),
                          None if pl is None else _ptr(pl), None if us is None else _ptr(us), _ptr(density))
    f2 = np.ones(planes, np.float32)
    if mode == WEIGHT_BRIGGS:
        for p in range(planes):
            f2[p] = lib().co_briggs_f2(nx * ny, _ptr(density[p]), ctypes.c_float(robust), 1)
    out = np.empty(n, np.float32)
    lib().co_density_reweight(nx, ny, cell_x, cell_y, conv, mode, n, _ptr(u), _ptr(v), _ptr(weight),
                              None if pl is None else _ptr(pl), None if us is None else _ptr(us),
                              _ptr(density), _ptr(f2), _ptr(out))
    return out, density, f2


def briggs_f2(density, robust, includes_conjugates=True):
    d = np.ascontiguousarray(density, np.float32)
    return float(lib().co_briggs_f2(d.size, _ptr(d), ctypes.c_float(robust), 1 if includes_conjugates else 0))


# ---------------------------------------------------------------------------
# phase-centre rotation (casars-imager/src/lib.rs:48030-48159, casa-imaging lib.rs:248-259)
# ---------------------------------------------------------------------------

def phase_rotation(source_dir, target_dir):
    """(uvrot 3x3, phrot 3): gridft_uvw_rotation_matrix(source, target) and
    uvw_phase_rotation_vector_from_angles(source, target); directions (ra, dec) rad."""
    src = np.ascontiguousarray(source_dir, np.float64)
    tgt = np.ascontiguousarray(target_dir, np.float64)
    rot = np.empty(9, np.float64)
    ph = np.empty(3, np.float64)
    lib().co_uvw_rotation_matrix(_ptr(src), _ptr(tgt), _ptr(rot))
    lib().co_phase_rotation_vector(_ptr(src), _ptr(tgt), _ptr(ph))
    return rot.reshape(3, 3), ph


def rotate_rows(uvrot, phrot, uvw_m, freq_hz, vis):
    """Rotated uvw (rows, 3), phase shift (rows,) and phase-rotated vis (rows, chan)."""
    uvw = np.ascontiguousarray(uvw_m, np.float64)
    f = np.ascontiguousarray(freq_hz, np.float64)
    v = np.array(vis, dtype=np.complex64, order="C", copy=True)
    n_rows, n_chan = v.shape
    out = np.empty_like(uvw)
    shift = np.empty(n_rows, np.float64)
    lib().co_rotate_rows(n_rows, n_chan, _ptr(np.ascontiguousarray(uvrot, np.float64)),
                         _ptr(np.ascontiguousarray(phrot, np.float64)), _ptr(uvw), _ptr(f),
                         _ptr(v.view(np.float32)), _ptr(out), _ptr(shift))
    return out, shift, v


PAIR_TRANSFORMS = {"half_sum": 0, "half_difference": 1, "positive_half_imag_difference": 2,
                   "negative_half_imag_difference": 3}


def ingest_rows(uvw_m, data, flag, weight, freq_hz, corr=(0,), pair=None, sumwt_factor=1.0,
                allow_psf_only=True, row_flag=None, rotation=None):
    """Rows x channels samples of one plane from row blocks (co_ingest_rows:
    execution.rs:2168-2315, types.rs:1593-1631).  data / flag (rows, corr,
    chan); weight (rows, corr) or (rows, corr, chan).  Returns uvw (rows, 3),
    visibility (rows, chan) complex64, weight, sumwt_factor, flag (bool)."""
    uvw = np.ascontiguousarray(uvw_m, np.float64)
    d = np.ascontiguousarray(data, np.complex64)
    n, n_corr, n_chan = d.shape
    fl = np.ascontiguousarray(flag, np.uint8)
    w = np.ascontiguousarray(weight, np.float32)
    spectrum = w.ndim == 3
    f = np.ascontiguousarray(freq_hz, np.float64)
    corr = tuple(int(c) for c in np.atleast_1d(corr))
    c1 = corr[1] if pair is not None else 0
    rf = None if row_flag is None else np.ascontiguousarray(row_flag, np.uint8)
    rot = np.zeros(9) if rotation is None else np.ascontiguousarray(rotation[0], np.float64).ravel()
    ph = np.zeros(3) if rotation is None else np.ascontiguousarray(rotation[1], np.float64).ravel()
    uo = np.empty((n, 3), np.float64)
    vo = np.empty((n, n_chan), np.complex64)
    wo = np.empty((n, n_chan), np.float32)
    so = np.empty((n, n_chan), np.float32)
    fo = np.empty((n, n_chan), np.uint8)
    L = lib()
    L.co_ingest_rows(ctypes.c_uint64(n), ctypes.c_uint32(n_corr), ctypes.c_uint32(n_chan),
                     ctypes.c_int(0 if pair is None else 1), ctypes.c_uint32(corr[0]), ctypes.c_uint32(c1),
                     ctypes.c_int(PAIR_TRANSFORMS[pair] if pair is not None else 0), ctypes.c_float(sumwt_factor),
                     ctypes.c_int(1 if allow_psf_only else 0), ctypes.c_int(1 if spectrum else 0),
                     ctypes.c_int(0 if rotation is None else 1), _ptr(rot), _ptr(ph), _ptr(f), _ptr(uvw),
                     _ptr(d.view(np.float32)), _ptr(fl), _ptr(w), None if rf is None else _ptr(rf), _ptr(uo),
                     _ptr(vo.view(np.float32)), _ptr(wo), _ptr(so), _ptr(fo))
    return {"uvw_m": uo, "visibility": vo, "weight": wo, "sumwt_factor": so, "flag": fo.astype(bool)}


# ---------------------------------------------------------------------------
# AWProject CF cells (gridder.rs:2157-2815) and the Airy A-term (lib.rs:57436-57605)
# ---------------------------------------------------------------------------

AW_OK = 0
AW_NON_FINITE_COORDINATE = 1
AW_OUTSIDE_GRID = 2
AW_KERNEL_INDEX_OUTSIDE_CELL = 3
AW_INVALID_NORMALIZATION = 4


class _CfCell(ctypes.Structure):
    _fields_ = [("shape_x", ctypes.c_uint64), ("shape_y", ctypes.c_uint64),
                ("x_support", ctypes.c_uint64), ("y_support", ctypes.c_uint64),
                ("offset", ctypes.c_uint64)]


def aw_kernel_origin(length: int) -> int:
    return int(lib().co_aw_kernel_origin(length))


class AwAtlas:
    """CF cells sharing one oversampling: each cell is (pixels complex64
    [shape_x, shape_y] indexed (x, y), x_support, y_support)."""

    def __init__(self, cells, sampling):
        self.sampling = int(sampling)
        self.cells = [(np.ascontiguousarray(p, np.complex64), int(xs), int(ys)) for p, xs, ys in cells]
        arr = (_CfCell * len(self.cells))()
        off = 0
        for k, (p, xs, ys) in enumerate(self.cells):
            arr[k] = _CfCell(p.shape[0], p.shape[1], xs, ys, off)
            off += p.size
        self.table = arr
        self.pixels = np.concatenate([p.ravel() for p, _, _ in self.cells]) if self.cells else np.zeros(0, np.complex64)
        self.offsets = np.array([arr[k].offset for k in range(len(self.cells))], np.uint64)

    def _cell_ptr(self, k):
        return ctypes.cast(ctypes.byref(self.table, k * ctypes.sizeof(_CfCell)), ctypes.c_void_p)


def aw_plan(gx, gy, du, dv, atlas, cell, u, v, w):
    """plan_sample_detailed for one sample: (status, (loc_x, loc_y, off_x,
    off_y, conj), norm complex64, grid_normalization complex128)."""
    geo = np.zeros(5, np.int64)
    n32 = np.zeros(2, np.float32)
    n64 = np.zeros(2, np.float64)
    rc = lib().co_aw_plan(gx, gy, du, dv, atlas.sampling, atlas._cell_ptr(cell), _ptr(atlas.pixels.view(np.float32)),
                          u, v, w, _ptr(geo), _ptr(n32), _ptr(n64))
    return int(rc), tuple(int(x) for x in geo), np.complex64(complex(n32[0], n32[1])), complex(n64[0], n64[1])


def aw_taps(atlas, cell, geo, gradient):
    """Phased taps of a plan, complex64 [2*ys+1][2*xs+1] (iy outer)."""
    _, xs, ys = atlas.cells[cell]
    out = np.zeros((2 * ys + 1, 2 * xs + 1), np.complex64)
    g = np.ascontiguousarray(gradient, np.float64)
    lib().co_aw_taps(atlas.sampling, atlas._cell_ptr(cell), _ptr(atlas.pixels.view(np.float32)),
                     _ptr(np.asarray(geo, np.int64)), _ptr(g), _ptr(out.view(np.float32)))
    return out


def _aw_stream(atlas, gradients, u, v, w, cell_of, pointing_of):
    n = len(u)
    return (n, _f64(u), _f64(v), _f64(w), np.ascontiguousarray(cell_of, np.uint32),
            np.ascontiguousarray(np.zeros(n) if pointing_of is None else pointing_of, np.uint32),
            np.ascontiguousarray(gradients, np.float64).reshape(-1))


def aw_grid_f64(gx, gy, du, dv, atlas, gradients, u, v, w, cell_of, values, pointing_of=None):
    """Sequential grid_sample_planned_f64 over the samples: complex128 grid
    [gx, gy] and the per-sample plan status."""
    n, u, v, w, c, p, g = _aw_stream(atlas, gradients, u, v, w, cell_of, pointing_of)
    vals = _c64_as_f32(values)
    grid = np.zeros((gx, gy), np.complex128)
    st = np.zeros(n, np.int32)
    lib().co_aw_grid_f64(gx, gy, du, dv, atlas.sampling, ctypes.cast(atlas.table, ctypes.c_void_p),
                         _ptr(atlas.pixels.view(np.float32)), _ptr(g), n, _ptr(u), _ptr(v), _ptr(w), _ptr(c),
                         _ptr(p), _ptr(vals), _ptr(grid.view(np.float64)), _ptr(st))
    return grid, st


def aw_degrid(gx, gy, du, dv, atlas, gradients, u, v, w, cell_of, model, pointing_of=None):
    """degrid_sample_planned per sample from a complex64 grid: complex64 values and plan status."""
    n, u, v, w, c, p, g = _aw_stream(atlas, gradients, u, v, w, cell_of, pointing_of)
    m = _c64_as_f32(model)
    out = np.zeros(n, np.complex64)
    st = np.zeros(n, np.int32)
    lib().co_aw_degrid(gx, gy, du, dv, atlas.sampling, ctypes.cast(atlas.table, ctypes.c_void_p),
                       _ptr(atlas.pixels.view(np.float32)), _ptr(g), n, _ptr(u), _ptr(v), _ptr(w), _ptr(c),
                       _ptr(p), _ptr(m), _ptr(out.view(np.float32)), _ptr(st))
    return out, st


def aw_grid64(gx, gy, du, dv, atlas, gradients, u, v, w, cell_of, vis, weight, mode=1, model=None,
              pointing_of=None, usable=None):
    """co_aw_grid64: the AW gridding in FP64 throughout (complex128 PSF and
    dirty / residual grids; ``model`` complex128 for mode 2).  Returns (psf,
    dirty, sum of w |normalisation|, gridded count)."""
    n, u, v, w, c, p, g = _aw_stream(atlas, gradients, u, v, w, cell_of, pointing_of)
    vis_f = _c64_as_f32(vis)
    wt = _f32(weight)
    us = None if usable is None else np.ascontiguousarray(usable, np.uint8)
    psf = np.zeros((gx, gy), np.complex128) if mode != 2 else None
    dirty = np.zeros((gx, gy), np.complex128) if mode != 0 else None
    m = None if model is None else np.ascontiguousarray(model, np.complex128).view(np.float64)
    sums = np.zeros(2, np.float64)
    lib().co_aw_grid64(gx, gy, du, dv, atlas.sampling, ctypes.cast(atlas.table, ctypes.c_void_p),
                       _ptr(atlas.pixels.view(np.float32)), _ptr(g), n, _ptr(u), _ptr(v), _ptr(w), _ptr(c),
                       _ptr(p), _ptr(vis_f), _ptr(wt), _ptr(us), _ptr(m), mode,
                       _ptr(None if psf is None else psf.view(np.float64)),
                       _ptr(None if dirty is None else dirty.view(np.float64)), _ptr(sums))
    return psf, dirty, float(sums[0]), int(sums[1])


def airy_max_radius(dish, blockage) -> float:
    return float(lib().co_airy_max_radius(dish, blockage))


def airy_from_index(index, max_radius, dish, blockage) -> np.float32:
    return np.float32(lib().co_airy_from_index(float(index), max_radius, dish, blockage))


def airy_table(dish, blockage) -> np.ndarray:
    out = np.empty(10000, np.float32)
    lib().co_airy_table(dish, blockage, _ptr(out))
    return out


def airy_radius(radius, freq, dish, blockage) -> np.float32:
    return np.float32(lib().co_airy_radius(radius, freq, dish, blockage))


def airy_offsets(l, m, freq, dish, blockage, table=None):
    """AiryVoltageLookup::evaluate_offsets over arrays of (l, m) radians."""
    if table is None:
        table = airy_table(dish, blockage)
    l = _f64(l)
    m = _f64(np.broadcast_to(m, l.shape))
    out = np.empty(l.shape, np.float32)
    lib().co_airy_offsets_many(l.size, _ptr(l), _ptr(np.ascontiguousarray(m)), freq, dish, blockage,
                               _ptr(np.ascontiguousarray(table, np.float32)), _ptr(out))
    return out


# --------------------------------------------------------------------------
# Hogbom minor cycle (lib.rs:59748-59885)
# --------------------------------------------------------------------------

STOP_NONE, STOP_GLOBAL, STOP_NSIGMA, STOP_CYCLE, STOP_NO_PIXELS = 0, 1, 2, 3, 6


def minor_cycle_stop(peak_abs, threshold, cycle_threshold, nsigma_threshold) -> int:
    """minor_cycle_stop_reason (lib.rs:74653-74677) as a stop code."""
    return lib().co_minor_cycle_stop(float(peak_abs), float(threshold), float(cycle_threshold),
                                     float(nsigma_threshold))


def hogbom(residual, psf, model, gain, budget, threshold=0.0, cycle_threshold=0.0, nsigma_threshold=0.0,
           mask=None):
    """One Hogbom minor cycle on every plane of ``residual`` (planes, nx, ny),
    in place on float32 copies.  Returns (residual, model, updates, stop codes)."""
    r = np.array(residual, dtype=np.float32, copy=True, order="C")
    m = np.array(model, dtype=np.float32, copy=True, order="C")
    if r.ndim == 2:
        r, m = r[None], m[None]
    k = _f32(psf)
    if k.ndim == 2:
        k = k[None]
    planes, nx, ny = r.shape
    mk = None if mask is None else np.ascontiguousarray(np.broadcast_to(mask, r.shape), dtype=np.uint8)
    ct = _f32(np.broadcast_to(np.asarray(cycle_threshold, np.float32), (planes,)))
    ns = _f32(np.broadcast_to(np.asarray(nsigma_threshold, np.float32), (planes,)))
    upd = np.zeros(planes, np.uint64)
    stop = np.zeros(planes, np.int32)
    lib().co_hogbom(planes, nx, ny, _ptr(r), k.shape[1], k.shape[2], _ptr(k), _ptr(m), _ptr(mk), float(gain),
                    float(threshold), _ptr(ct), _ptr(ns), int(budget), _ptr(upd), _ptr(stop))
    return r, m, upd, stop


# --------------------------------------------------------------------------
# CASA CF-cache key selection (awprojection.rs:1068-1340), restated on key
# tuples (frequency_hz, w_value_lambda, mueller_element, parallactic_angle_deg).
# "Stable key" order compares the IEEE-754 bit patterns (awprojection.rs:66-95).

def _bits(x: float) -> int:
    return int(np.array(float(x), np.float64).view(np.uint64))


def cf_stable_key(key):
    f, w, m, p = key
    return (_bits(f), _bits(w), int(m), _bits(p))


def cf_circular_degrees(v: float) -> float:
    """(v + 180).rem_euclid(360) - 180 (awprojection.rs:2046-2048)."""
    r = float(np.fmod(v + 180.0, 360.0))
    if r < 0.0:
        r += 360.0
    return r - 180.0


def cf_sorted_unique(values):
    """sorted_f64_bits (awprojection.rs:1974-1978): unique bit patterns in value order."""
    uniq = {_bits(v): float(v) for v in values}
    return sorted(uniq.values(), key=lambda v: (v, not np.signbit(v)))


def cf_nearest_key(keys, frequency_hz, w_value_lambda, mueller_element, parallactic_angle_deg):
    """nearest_key (awprojection.rs:1071-1104): Mueller exact, minimum of the
    relative frequency + relative W + |dPA|/360 distance, ties by stable key."""
    if not (np.isfinite(frequency_hz) and frequency_hz > 0 and np.isfinite(w_value_lambda)
            and np.isfinite(parallactic_angle_deg)):
        return None
    fs, ws = max(abs(frequency_hz), 1.0), max(abs(w_value_lambda), 1.0)
    best = None
    for k in keys:
        if k[2] != mueller_element:
            continue
        d = (abs((k[0] - frequency_hz) / fs) + abs((k[1] - w_value_lambda) / ws)
             + abs(cf_circular_degrees(k[3] - parallactic_angle_deg)) / 360.0)
        cand = (d, cf_stable_key(k))
        if best is None or cand < best[0]:
            best = (cand, k)
    return None if best is None else best[1]


def _cf_requested(sample_frequency_hz, image_reference_frequency_hz, conjugate_beams):
    if conjugate_beams:
        rad = 2.0 * image_reference_frequency_hz * image_reference_frequency_hz - sample_frequency_hz ** 2
        if not (np.isfinite(rad) and rad > 0.0):
            return None
        return float(np.sqrt(rad))
    return sample_frequency_hz


def _cf_w_index(w_increment, n_w, w_lambda):
    if n_w <= 1 or w_increment == 0.0:
        return 0
    r = float(np.sqrt(w_increment * abs(w_lambda)))
    f = float(np.floor(r))
    r = f + 1.0 if r - f >= 0.5 else f  # f64::round, half away from zero (r >= 0; r - f is exact)
    return int(min(max(r, 0.0), float(n_w - 1)))


def cf_select_scanned(keys, w_increment, requested_frequency_hz, w_lambda, mueller, pa):
    """select_key_for_sample_scanned (awprojection.rs:1212-1258): candidates of
    the Mueller element, the W plane round(sqrt(WIncr |w|)) among their sorted W
    values, then the lexicographic minimum of (|df|, |circular dPA|, stable key)."""
    cands = [k for k in keys if k[2] == mueller]
    if not cands:
        return None
    ws = cf_sorted_unique([k[1] for k in cands])
    sel_w = ws[_cf_w_index(w_increment, len(ws), w_lambda)]
    pool = [k for k in cands if _bits(k[1]) == _bits(sel_w)]
    return min(pool, key=lambda k: (abs(k[0] - requested_frequency_hz),
                                    abs(cf_circular_degrees(k[3] - pa)), cf_stable_key(k)))


def cf_select_direct(keys, w_increment, requested_frequency_hz, w_lambda, mueller, pa):
    """select_key_for_sample_direct (awprojection.rs:1260-1322): per-axis
    nearest values from the inventory, then an exact key lookup."""
    if mueller not in {k[2] for k in keys}:
        return None
    ws = cf_sorted_unique([k[1] for k in keys])
    fs = cf_sorted_unique([k[0] for k in keys])
    ps = cf_sorted_unique([k[3] for k in keys])
    w = ws[_cf_w_index(w_increment, len(ws), w_lambda)]
    f = min(fs, key=lambda v: (abs(v - requested_frequency_hz), _bits(v)))
    p = min(ps, key=lambda v: (abs(cf_circular_degrees(v - pa)), _bits(v)))
    want = cf_stable_key((f, w, mueller, p))
    for k in keys:
        if cf_stable_key(k) == want:
            return k
    return None


def cf_select_key(keys, w_increment, sample_frequency_hz, image_reference_frequency_hz, w_lambda, mueller_element,
                  parallactic_angle_deg, conjugate_beams=False, prediction=False, direct=False):
    """select_key_for_sample (prediction False, awprojection.rs:1106-1172) and
    select_key_for_prediction_sample (prediction True, :1174-1210)."""
    if not (np.isfinite(sample_frequency_hz) and sample_frequency_hz > 0 and np.isfinite(w_lambda)
            and np.isfinite(parallactic_angle_deg)):
        return None
    if prediction:
        req = sample_frequency_hz
        m = 15 - mueller_element if w_lambda > 0.0 else mueller_element
    else:
        if not (np.isfinite(image_reference_frequency_hz) and image_reference_frequency_hz > 0):
            return None
        req = _cf_requested(sample_frequency_hz, image_reference_frequency_hz, conjugate_beams)
        if req is None:
            return None
        m = mueller_element if w_lambda > 0.0 else 15 - mueller_element
    pick = cf_select_direct if direct else cf_select_scanned
    return pick(keys, w_increment, req, w_lambda, m, parallactic_angle_deg)
