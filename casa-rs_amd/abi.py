"""ctypes binding of libcasa_imaging_hip.so (declarations in include/casa_imaging_hip.h).

The library is loaded from ``casa-rs_amd/lib/`` (built in place by
``casa-rs_amd/csrc/Makefile``).  There is no fallback: if the library is
missing or no HIP device is present, loading or context creation raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libcasa_imaging_hip.so")

CHIP_OK = 0
CHIP_ERR_INVALID = 1
CHIP_ERR_NO_SAMPLES = 2
CHIP_ERR_NORMALIZATION = 3
CHIP_ERR_DEVICE = 4
CHIP_ERR_CF_CACHE = 5

MODE_PSF = 0
MODE_DIRTY = 1
MODE_RESIDUAL = 2

WEIGHT_UNIFORM = 1
WEIGHT_BRIGGS = 2

CYCLE_AUTO = -1
CYCLE_HIPFFT = 0
CYCLE_FUSED = 1

# chip_hogbom stop codes (CleanStopReason order + 1; 0 = budget used up)
CLEAN_BUDGET = 0
CLEAN_GLOBAL = 1
CLEAN_NSIGMA = 2
CLEAN_CYCLE = 3
CLEAN_NO_PIXELS = 6

F32 = 0
F64 = 1

AW_OK = 0
AW_NON_FINITE_COORDINATE = 1
AW_OUTSIDE_GRID = 2
AW_KERNEL_INDEX_OUTSIDE_CELL = 3
AW_INVALID_NORMALIZATION = 4
AW_NO_CELL = 5

PAD_COMPOSITE_1_2 = 0
PAD_COMPOSITE_1_3 = 1
PAD_NONE = 2

# Every symbol the header declares; tests check that the library exports them.
EXPORTS = (
    "chip_ctx_create", "chip_ctx_destroy", "chip_ctx_set_stream", "chip_error_free", "chip_version",
    "chip_gridder_create", "chip_gridder_destroy", "chip_gridder_info_get", "chip_gridder_tables",
    "chip_max_abs_w", "chip_gridder_attach_wproject", "chip_wproject_tables", "chip_gridder_wkernel_budget",
    "chip_plan", "chip_buffer_alloc", "chip_buffer_free", "chip_buffer_upload", "chip_buffer_download",
    "chip_buffer_zero", "chip_gridder_attach_wproject_tables", "chip_grid_wproject_samples",
    "chip_density_weights", "chip_phase_rotation", "chip_rotate_rows",
    "chip_vset_create", "chip_vset_destroy", "chip_vset_sums", "chip_vset_stats", "chip_vset_grid",
    "chip_vset_degrid", "chip_vset_grid_values", "chip_vset_major_cycle", "chip_vset_cycle_info",
    "chip_vset_cycle_strategy", "chip_vset_cycle_graph", "chip_vset_wave_groups", "chip_vset_two_phase", "chip_vset_cycle_passes", "chip_vset_item_order", "chip_vset_schedule", "chip_vset_schedule_info",
    "chip_grid_plain", "chip_grid_plain_noalloc", "chip_degrid_plain",
    "chip_finish_products", "chip_model_to_grid", "chip_grid_to_model_adjoint", "chip_predict_model",
    "chip_gridder_attach_aw", "chip_gridder_attach_aw_airy", "chip_gridder_attach_aw_airy_freqs",
    "chip_aw_select_cells", "chip_aw_cells", "chip_aw_plan",
    "chip_airy_table", "chip_hogbom",
    "chip_ingest_create", "chip_ingest_destroy", "chip_ingest_push", "chip_ingest_samples", "chip_ingest_download",
    "chip_cfcache_open", "chip_cfcache_close", "chip_cfcache_info_get", "chip_cfcache_keys", "chip_cfcache_metadata",
    "chip_cfcache_nearest_key", "chip_cfcache_select_key", "chip_cfcache_load", "chip_gridder_attach_aw_cache",
    "chip_cfcache_select_cells",
)


class ImagingError(RuntimeError):
    """Base error; subclasses mirror ImagingError variants (error.rs:8-24)."""

    code = CHIP_ERR_INVALID


class InvalidRequest(ImagingError):
    code = CHIP_ERR_INVALID


class NoUsableSamples(ImagingError):
    code = CHIP_ERR_NO_SAMPLES


class Normalization(ImagingError):
    code = CHIP_ERR_NORMALIZATION


class Unsupported(ImagingError):
    code = CHIP_ERR_DEVICE


class ConvolutionFunctionCache(ImagingError):
    code = CHIP_ERR_CF_CACHE


_ERRORS = {
    CHIP_ERR_INVALID: InvalidRequest,
    CHIP_ERR_NO_SAMPLES: NoUsableSamples,
    CHIP_ERR_NORMALIZATION: Normalization,
    CHIP_ERR_DEVICE: Unsupported,
    CHIP_ERR_CF_CACHE: ConvolutionFunctionCache,
}


class Geometry(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_uint64), ("ny", ctypes.c_uint64),
        ("cell_x", ctypes.c_double), ("cell_y", ctypes.c_double),
        ("padding_mode", ctypes.c_uint32), ("oversampling", ctypes.c_uint32),
    ]


class GridderInfo(ctypes.Structure):
    _fields_ = [
        ("gx", ctypes.c_uint64), ("gy", ctypes.c_uint64),
        ("blc_x", ctypes.c_uint64), ("blc_y", ctypes.c_uint64),
        ("du", ctypes.c_double), ("dv", ctypes.c_double),
        ("oversampling", ctypes.c_uint32), ("wproject", ctypes.c_uint32),
        ("planes", ctypes.c_uint32), ("sampling", ctypes.c_uint32),
        ("width", ctypes.c_uint32), ("max_support", ctypes.c_uint32),
        ("norm_radius", ctypes.c_uint32), ("aw", ctypes.c_uint32),
        ("w_scale", ctypes.c_double),
    ]


class Sums(ctypes.Structure):
    _fields_ = [
        ("normalization_sumwt", ctypes.c_double),
        ("reported_sumwt", ctypes.c_double),
        ("gridded", ctypes.c_uint64),
        ("skipped", ctypes.c_uint64),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Samples(ctypes.Structure):
    _fields_ = [
        ("n_rows", ctypes.c_uint64), ("n_chan", ctypes.c_uint64),
        ("u_dev", ctypes.c_void_p), ("v_dev", ctypes.c_void_p), ("w_dev", ctypes.c_void_p),
        ("uvw_m_dev", ctypes.c_void_p), ("lambda_scale_dev", ctypes.c_void_p),
        ("vis_dev", ctypes.c_void_p), ("weight_dev", ctypes.c_void_p),
        ("sumwt_factor_dev", ctypes.c_void_p), ("gridable_dev", ctypes.c_void_p),
        ("flag_dev", ctypes.c_void_p),
        ("cf_cell_dev", ctypes.c_void_p), ("pointing_dev", ctypes.c_void_p),
    ]


class CfCell(ctypes.Structure):
    _fields_ = [
        ("shape_x", ctypes.c_uint64), ("shape_y", ctypes.c_uint64),
        ("x_support", ctypes.c_uint32), ("y_support", ctypes.c_uint32),
        ("offset", ctypes.c_uint64),
    ]


# row-stream ingestion: polarization and pair collapse (include/casa_imaging_hip.h)
POL_EXPLICIT = 0
POL_PAIR = 1
PAIR_HALF_SUM = 0
PAIR_HALF_DIFFERENCE = 1
PAIR_POSITIVE_HALF_IMAG_DIFF = 2
PAIR_NEGATIVE_HALF_IMAG_DIFF = 3


class IngestParams(ctypes.Structure):
    _fields_ = [
        ("n_corr", ctypes.c_uint32), ("n_chan", ctypes.c_uint32), ("polarization", ctypes.c_int32),
        ("corr", ctypes.c_uint32 * 2), ("transform", ctypes.c_int32), ("sumwt_factor", ctypes.c_float),
        ("allow_psf_only", ctypes.c_int32), ("weight_spectrum", ctypes.c_int32), ("rotate", ctypes.c_int32),
        ("uvrot", ctypes.c_double * 9), ("phrot", ctypes.c_double * 3),
        ("max_rows", ctypes.c_uint64), ("block_rows", ctypes.c_uint64),
    ]


class HogbomParams(ctypes.Structure):
    _fields_ = [("gain", ctypes.c_float), ("threshold", ctypes.c_float), ("budget", ctypes.c_uint32)]


class AiryParams(ctypes.Structure):
    _fields_ = [
        ("dish_m", ctypes.c_double), ("blockage_m", ctypes.c_double),
        ("freq_hz", ctypes.c_double), ("pa_rad", ctypes.c_double),
        ("sampling", ctypes.c_uint32), ("n_w", ctypes.c_uint32),
        ("max_abs_w", ctypes.c_double),
    ]


_lib = None

_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_ERR = ctypes.POINTER(ctypes.c_char_p)
_U32 = ctypes.c_uint32
_U64 = ctypes.c_uint64


class CfKey(ctypes.Structure):
    _fields_ = [("frequency_hz", ctypes.c_double), ("w_value_lambda", ctypes.c_double),
                ("mueller_element", ctypes.c_int32), ("parallactic_angle_deg", ctypes.c_double)]


class CfKernelMeta(ctypes.Structure):
    _fields_ = [
        ("shape_x", ctypes.c_uint64), ("shape_y", ctypes.c_uint64),
        ("sampling", ctypes.c_uint64), ("x_support", ctypes.c_uint64), ("y_support", ctypes.c_uint64),
        ("uv_reference_value", ctypes.c_double * 2), ("uv_reference_pixel", ctypes.c_double * 2),
        ("uv_increment", ctypes.c_double * 2), ("uv_pc", ctypes.c_double * 4),
        ("diameter_m", ctypes.c_double), ("conjugate_frequency_hz", ctypes.c_double),
        ("w_increment", ctypes.c_double),
        ("conjugate_polarization", ctypes.c_int32), ("polarization", ctypes.c_int32),
        ("rotationally_symmetric", ctypes.c_int32),
        ("telescope_name", ctypes.c_char * 64), ("band_name", ctypes.c_char * 64),
    ]


class CfCacheInfo(ctypes.Structure):
    _fields_ = [("paired_cells", ctypes.c_uint64), ("n_freq", ctypes.c_uint32), ("n_w", ctypes.c_uint32),
                ("n_mueller", ctypes.c_uint32), ("n_pa", ctypes.c_uint32), ("w_increment", ctypes.c_double),
                ("metadata_fingerprint", ctypes.c_uint64)]


def load():
    """Load the library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise Unsupported(
            f"{LIB_PATH} is missing: build it with `make -C casa-rs_amd/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "chip_version": (ctypes.c_char_p, []),
        "chip_error_free": (None, [_P]),
        "chip_ctx_create": (ctypes.c_int, [ctypes.c_int, _PP, _PP]),
        "chip_ctx_destroy": (None, [_P]),
        "chip_ctx_set_stream": (ctypes.c_int, [_P, ctypes.c_int, _P, _PP]),
        "chip_gridder_create": (ctypes.c_int, [_P, ctypes.POINTER(Geometry), _PP, _PP]),
        "chip_gridder_destroy": (None, [_P]),
        "chip_gridder_info_get": (ctypes.c_int, [_P, ctypes.POINTER(GridderInfo), _PP]),
        "chip_gridder_tables": (ctypes.c_int, [_P, _P, _P, _P, _PP]),
        "chip_max_abs_w": (ctypes.c_int, [_P, ctypes.POINTER(Samples), ctypes.POINTER(ctypes.c_double), _PP]),
        "chip_gridder_attach_wproject": (ctypes.c_int, [_P, ctypes.c_double, _U32, _U32, _PP]),
        "chip_wproject_tables": (ctypes.c_int, [_P, _P, _P, _P, _PP]),
        "chip_gridder_wkernel_budget": (ctypes.c_int, [_P, _U64, _P, _PP]),
        "chip_plan": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _P, _P, _PP]),
        "chip_phase_rotation": (ctypes.c_int, [_P, _P, _P, _P, _PP]),
        "chip_rotate_rows": (ctypes.c_int, [_P, _U64, _U64, _P, _P, _P, _P, _P, _P, _P, _PP]),
        "chip_density_weights": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _U32, ctypes.c_float, _U32, _P, _P,
                                                _P, _PP]),
        "chip_vset_create": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _PP, _PP]),
        "chip_ingest_create": (ctypes.c_int, [_P, ctypes.POINTER(IngestParams), _P, _PP, _PP]),
        "chip_ingest_destroy": (None, [_P]),
        "chip_ingest_push": (ctypes.c_int, [_P, _U64, _P, _P, _P, _P, _P, _PP]),
        "chip_ingest_samples": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _PP]),
        "chip_ingest_download": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _PP]),
        "chip_vset_destroy": (None, [_P]),
        "chip_vset_sums": (ctypes.c_int, [_P, _P, _PP]),
        "chip_vset_stats": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _PP]),
        "chip_vset_grid": (ctypes.c_int, [_P, _U32, _P, _P, _P, _PP]),
        "chip_vset_degrid": (ctypes.c_int, [_P, _P, _P, _PP]),
        "chip_vset_grid_values": (ctypes.c_int, [_P, _P, _P, _PP]),
        "chip_vset_major_cycle": (ctypes.c_int, [_P, _P, _P, _P, _P, _PP]),
        "chip_vset_cycle_info": (ctypes.c_int, [_P, _P, _P, _P, _PP]),
        "chip_vset_cycle_strategy": (ctypes.c_int, [_P, ctypes.c_int, _PP]),
        "chip_vset_cycle_graph": (ctypes.c_int, [_P, ctypes.c_int, _P, _PP]),
        "chip_vset_cycle_passes": (ctypes.c_int, [_P, ctypes.c_int, _P, _PP]),
        "chip_vset_item_order": (ctypes.c_int, [_P, _U32, _P, _PP]),
        "chip_vset_schedule": (ctypes.c_int, [_P, _U32, _U32, _U32, _U32, _PP]),
        "chip_vset_wave_groups": (ctypes.c_int, [_P, _U32, _P, _PP]),
        "chip_vset_two_phase": (ctypes.c_int, [_P, ctypes.c_int, _PP]),
        "chip_hogbom": (ctypes.c_int, [_P, ctypes.POINTER(HogbomParams), _U64, _U64, _U64, _P, _U64, _U64, _P, _P,
                                       _P, _P, _P, _P, _P, _PP]),
        "chip_vset_schedule_info": (ctypes.c_int, [_P, _P, _P, _P, _P, _P, _P, _PP]),
        "chip_grid_plain": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _U32, _U32, _P, _P, _P, _P, _PP]),
        "chip_grid_plain_noalloc": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _U32, _U32, _P, _P, _P, _P,
                                                   _PP]),
        "chip_degrid_plain": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _U32, _P, _P, _PP]),
        "chip_finish_products": (ctypes.c_int, [_P, _U32, _U64, _P, _P, _P, _P, _P, _P, _PP]),
        "chip_model_to_grid": (ctypes.c_int, [_P, _U32, _U64, _P, _P, _PP]),
        "chip_grid_to_model_adjoint": (ctypes.c_int, [_P, _U64, _P, _P, _PP]),
        "chip_predict_model": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _P, _U64, _P, _PP]),
        "chip_buffer_alloc": (ctypes.c_int, [_P, _U64, _PP, _PP]),
        "chip_buffer_free": (None, [_P, _P]),
        "chip_buffer_upload": (ctypes.c_int, [_P, _P, _P, _U64, _PP]),
        "chip_buffer_download": (ctypes.c_int, [_P, _P, _P, _U64, _PP]),
        "chip_buffer_zero": (ctypes.c_int, [_P, _P, _U64, _PP]),
        "chip_gridder_attach_wproject_tables": (ctypes.c_int, [_P, _U32, _U32, ctypes.c_double, _U32, _P, _P, _P,
                                                               _PP]),
        "chip_grid_wproject_samples": (ctypes.c_int, [_P, _U64, _P, _P, _P, _P, _U32, _PP]),
        "chip_gridder_attach_aw": (ctypes.c_int, [_P, _U32, ctypes.POINTER(CfCell), _P, _U64, _U32, _U32, _P,
                                                  ctypes.c_double, _PP]),
        "chip_gridder_attach_aw_airy": (ctypes.c_int, [_P, ctypes.POINTER(AiryParams), _U32, _P, _PP]),
        "chip_gridder_attach_aw_airy_freqs": (ctypes.c_int, [_P, ctypes.POINTER(AiryParams), _U32, _P, _U32, _P,
                                                             _PP]),
        "chip_aw_select_cells": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _P, _PP]),
        "chip_aw_cells": (ctypes.c_int, [_P, _P, _P, _P, _P, _PP]),
        "chip_aw_plan": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _P, _P, _P, _PP]),
        "chip_airy_table": (ctypes.c_int, [ctypes.c_double, ctypes.c_double, _P, _P, _PP]),
        "chip_cfcache_open": (ctypes.c_int, [ctypes.c_char_p, _PP, _PP]),
        "chip_cfcache_close": (None, [_P]),
        "chip_cfcache_info_get": (ctypes.c_int, [_P, ctypes.POINTER(CfCacheInfo), _P, _P, _P, _P, _PP]),
        "chip_cfcache_keys": (ctypes.c_int, [_P, _P, _PP]),
        "chip_cfcache_metadata": (ctypes.c_int, [_P, _U64, ctypes.POINTER(CfKernelMeta),
                                                 ctypes.POINTER(CfKernelMeta), _PP]),
        "chip_cfcache_nearest_key": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                                    ctypes.c_double, ctypes.POINTER(ctypes.c_int64), _PP]),
        "chip_cfcache_select_key": (ctypes.c_int, [_P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_int32, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                                   ctypes.POINTER(ctypes.c_int64), _PP]),
        "chip_cfcache_load": (ctypes.c_int, [_P, _U64, _P, _P, _PP]),
        "chip_gridder_attach_aw_cache": (ctypes.c_int, [_P, _P, _U32, _P, _PP]),
        "chip_cfcache_select_cells": (ctypes.c_int, [_P, ctypes.POINTER(Samples), _P, ctypes.c_double,
                                                     ctypes.c_int32, _P, ctypes.c_double, ctypes.c_int,
                                                     ctypes.c_int, _P, _PP]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def call(name, *args):
    """Call an entry point that ends in ``char **err`` and raise on failure."""
    L = load()
    err = ctypes.c_void_p()
    rc = getattr(L, name)(*args, ctypes.byref(err))
    if rc != CHIP_OK:
        msg = ctypes.cast(err, ctypes.c_char_p).value.decode() if err.value else name
        if err.value:
            L.chip_error_free(err)
        raise _ERRORS.get(rc, ImagingError)(f"{name}: {msg}")
    return rc
