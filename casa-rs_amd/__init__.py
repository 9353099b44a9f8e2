"""casa-imaging gridding / degridding on MI355X (gfx950).

The compute path is libcasa_imaging_hip.so (hand-written HIP kernels behind a
C ABI, header ``include/casa_imaging_hip.h``); this package is the host-side
mirror of the reference interface over that ABI plus the autograd surface.
It has no CPU fallback: without the built library or a HIP device, every
entry point raises.
"""
from . import abi
from .abi import (CLEAN_BUDGET, CLEAN_CYCLE, CLEAN_GLOBAL, CLEAN_NO_PIXELS, CLEAN_NSIGMA, AW_INVALID_NORMALIZATION, AW_KERNEL_INDEX_OUTSIDE_CELL, AW_NO_CELL,
                  AW_NON_FINITE_COORDINATE, AW_OK, AW_OUTSIDE_GRID, CYCLE_AUTO, CYCLE_FUSED,
                  CYCLE_HIPFFT, F32, F64, MODE_DIRTY, MODE_PSF, MODE_RESIDUAL, WEIGHT_BRIGGS,
                  WEIGHT_UNIFORM, ImagingError, InvalidRequest, NoUsableSamples, Normalization,
                  Unsupported)
from .imaging import (PACKED_W_SAMPLE, AwConvolutionFunctionCache, AwConvolutionFunctionInventory,
                      AwConvolutionFunctionKey, Context, HostBuffer, RowChannelBatch, RowIngest, StandardGridder,
                      StandardMfsModelPredictor, VisibilityBatch, VisibilitySet, grid_wproject_samples_host,
                      pack_wproject_samples,
                      airy_table, default_context, mosaic_phase_gradient, phase_rotation,
                      rotate_rows, run_hogbom_minor_cycle)

__all__ = [
    "abi", "CYCLE_AUTO", "CYCLE_FUSED", "CYCLE_HIPFFT", "WEIGHT_BRIGGS", "WEIGHT_UNIFORM", "F32", "F64", "MODE_DIRTY", "MODE_PSF", "MODE_RESIDUAL", "ImagingError",
    "InvalidRequest", "NoUsableSamples", "Normalization", "Unsupported", "Context",
    "RowChannelBatch", "StandardGridder", "VisibilityBatch", "VisibilitySet", "default_context",
    "phase_rotation", "rotate_rows", "airy_table", "mosaic_phase_gradient",
    "AW_OK", "AW_NON_FINITE_COORDINATE", "AW_OUTSIDE_GRID", "AW_KERNEL_INDEX_OUTSIDE_CELL",
    "AW_INVALID_NORMALIZATION", "AW_NO_CELL", "run_hogbom_minor_cycle", "CLEAN_BUDGET", "CLEAN_GLOBAL",
    "CLEAN_NSIGMA", "CLEAN_CYCLE", "CLEAN_NO_PIXELS", "RowIngest", "StandardMfsModelPredictor",
    "PACKED_W_SAMPLE", "HostBuffer", "grid_wproject_samples_host", "pack_wproject_samples",
    "AwConvolutionFunctionCache", "AwConvolutionFunctionInventory", "AwConvolutionFunctionKey",
]
