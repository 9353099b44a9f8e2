/*
 * casa_imaging_hip.h -- C ABI of libcasa_imaging_hip.so, the MI355X (gfx950)
 * gridding / degridding path of casa-imaging.
 *
 * The entry points replace the device seams of bglenden/casa-rs
 * `crates/casa-imaging` that return ImagingError::Unsupported on non-macOS
 * hosts (lib.rs:71813-71838, :72454-72472, :68057-68070).  Conventions follow
 * the reference's own foreign-oracle shim (crates/casa-test-support/src/cpp/
 * casacore_cpp_gridder_shim.cpp:23-60, casacore_cpp_common.h:47-55):
 *   - every call returns an int status (CHIP_OK = 0) and, on failure, sets
 *     *err to a malloc'd message that the caller frees with chip_error_free;
 *   - the status maps onto ImagingError (crates/casa-imaging/src/error.rs:8-24):
 *       CHIP_ERR_INVALID       -> InvalidRequest
 *       CHIP_ERR_NO_SAMPLES    -> NoUsableSamples
 *       CHIP_ERR_NORMALIZATION -> Normalization
 *       CHIP_ERR_DEVICE        -> Unsupported (no usable HIP device / runtime error)
 *   - calls are synchronous on return (the Metal path commits and waits,
 *     lib.rs:72289-72290); one context per host thread.
 *
 * Memory: every array argument whose name ends in `_dev` is device memory on
 * the context's device (hipMalloc'd or a torch CUDA tensor).  Grids are the
 * reference's Array2 layout: logical (x, y), y fastest, flat index x*gy + y
 * (gridder.rs:61-66), complex values interleaved (re, im), one grid per
 * channel, channels outermost.  Grid precision CHIP_F32 means complex64
 * (float2) grids, CHIP_F64 means complex128 (double2) grids.
 */
#ifndef CASA_IMAGING_HIP_H
#define CASA_IMAGING_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHIP_OK 0
#define CHIP_ERR_INVALID 1
#define CHIP_ERR_NO_SAMPLES 2
#define CHIP_ERR_NORMALIZATION 3
#define CHIP_ERR_DEVICE 4
#define CHIP_ERR_CF_CACHE 5 /* ImagingError::ConvolutionFunctionCache (error.rs): unreadable or inconsistent CF cache */

/* Grid modes of accumulate_w_project_grid_metal_samples (lib.rs:71688-71690). */
#define CHIP_MODE_PSF 0      /* PSF grid only */
#define CHIP_MODE_DIRTY 1    /* PSF + dirty (V*w) grids */
#define CHIP_MODE_RESIDUAL 2 /* degrid model grid, grid (V - model)*w */

/* Density weighting modes of chip_density_weights (WeightingMode, types.rs:145-160). */
#define CHIP_WEIGHT_UNIFORM 1
#define CHIP_WEIGHT_BRIGGS 2

/* Transform strategies of chip_vset_major_cycle. */
#define CHIP_CYCLE_AUTO (-1)
#define CHIP_CYCLE_HIPFFT 0
#define CHIP_CYCLE_FUSED 1

/* Stop codes of chip_hogbom: CleanStopReason (types.rs:2215-2230) in
 * declaration order plus one; 0 = the component budget was used up. */
#define CHIP_CLEAN_BUDGET 0
#define CHIP_CLEAN_GLOBAL 1    /* GlobalThresholdReached */
#define CHIP_CLEAN_NSIGMA 2    /* NsigmaThresholdReached */
#define CHIP_CLEAN_CYCLE 3     /* CycleThresholdReached */
#define CHIP_CLEAN_NO_PIXELS 6 /* NoCleanablePixels */

#define CHIP_F32 0 /* binned, LDS-tiled, FP32 complex MACs, complex64 grids */
#define CHIP_F64 1 /* plain per-channel launch, FP64 MACs and atomics, complex128 grids */

/* Padding rule of StandardGridder (gridder.rs:371-390). */
#define CHIP_PAD_COMPOSITE_1_2 0 /* StandardGridder::new */
#define CHIP_PAD_COMPOSITE_1_3 1 /* new_with_casa_composite_padding (model predictor) */
#define CHIP_PAD_NONE 2          /* new_unpadded */

typedef struct chip_ctx chip_ctx;
typedef struct chip_gridder chip_gridder;
typedef struct chip_vset chip_vset;
typedef struct chip_ingest chip_ingest;
typedef struct chip_cfcache chip_cfcache;

/* ImageGeometry (types.rs:885-925) plus the gridder's padding rule and tap-table
 * oversampling (the reference fixes 100, gridder.rs:407). */
typedef struct {
    uint64_t nx, ny;
    double cell_x, cell_y; /* radians, finite and positive */
    uint32_t padding_mode;
    uint32_t oversampling; /* even, 2..254 */
} chip_geometry;

typedef struct {
    uint64_t gx, gy;       /* padded grid shape */
    uint64_t blc_x, blc_y; /* image origin in the grid, gridder.rs:403-406 */
    double du, dv;         /* grid spacing in wavelengths, gridder.rs:426-427 */
    uint32_t oversampling;
    uint32_t wproject;     /* 1 once chip_gridder_attach_wproject succeeded */
    uint32_t planes;       /* W planes */
    uint32_t sampling;     /* W kernel oversampling */
    uint32_t width;        /* W kernel quarter-plane width */
    uint32_t max_support;  /* largest W half-support */
    uint32_t norm_radius;  /* (sampling + 1) / 2 */
    uint32_t aw;           /* 1 once an AWProject CF atlas is attached (then wproject = 1 too,
                              planes = tap tables: cells x pointings x 2) */
    double w_scale;
} chip_gridder_info;

/* Per-channel accumulation summary (WProjectGridAccumulation, lib.rs:70978-70993;
 * StandardMfsDirtyAccumulation).  Sums are accumulated in FP64 on the device. */
typedef struct {
    double normalization_sumwt;
    double reported_sumwt;
    uint64_t gridded;
    uint64_t skipped;
} chip_sums;

/* Visibility input, device resident.  Two shapes share one descriptor:
 *  - per-sample wavelengths (VisibilityBatch, types.rs:1376-1405): set u,v,w
 *    (length n_rows), n_chan = 1, uvw_m = lambda_scale = NULL;
 *  - rows x channels (casars-imager's MS rows): set uvw_m (n_rows x 3 metres)
 *    and lambda_scale (n_chan values of freq/c, casars-imager/src/lib.rs:557),
 *    u = v = w = NULL.  Sample (row, chan) is index row*n_chan + chan and its
 *    wavelength coordinates are uvw_m[row][k] * lambda_scale[chan] in FP64.
 * weight is required; sumwt_factor (NULL = 1), gridable (NULL = all true) and
 * flag (NULL = none flagged, nonzero = flagged) are optional, one per sample. */
typedef struct {
    uint64_t n_rows;
    uint64_t n_chan;
    const double *u_dev, *v_dev, *w_dev;
    const double *uvw_m_dev;
    const double *lambda_scale_dev;
    const float *vis_dev; /* complex64 interleaved, one per sample */
    const float *weight_dev;
    const float *sumwt_factor_dev;
    const uint8_t *gridable_dev;
    const uint8_t *flag_dev;
    /* AWProject only (both optional): the CF cell of every sample, as chosen
     * by the caller's cache lookup (AwConvolutionFunctionCache::
     * select_key_for_sample, awprojection.rs:1117-1180); NULL selects the cell
     * as the W plane round(sqrt(w_increment * |w|)) clamped to the cells
     * (awprojection.rs:1239-1246).  pointing_dev: per row, index into the
     * atlas's phase gradients; NULL = pointing 0. */
    const uint32_t *cf_cell_dev;
    const uint32_t *pointing_dev;
} chip_samples;

/* One AWProject convolution-function cell (AwConvolutionFunctionKernelMetadata,
 * awprojection.rs:97-130, with its pixels AwConvolutionFunctionCell::imaging,
 * :141-149): stored pixel shape, half-supports in grid pixels, and the index
 * (in complex64 elements) of its first pixel in the atlas pixel array.  Pixels
 * are (x, y) with y fastest, origin at aw_kernel_origin (gridder.rs:2809-2815):
 * n/2 for even n, n/2 + 1 for odd n. */
typedef struct {
    uint64_t shape_x, shape_y;
    uint32_t x_support, y_support;
    uint64_t offset;
} chip_cf_cell;

/* Airy A-term CF construction (chip_gridder_attach_aw_airy). */
typedef struct {
    double dish_m;      /* PrimaryBeamModel::Airy dish_diameter_m */
    double blockage_m;  /* blockage_diameter_m (0 = unblocked) */
    double freq_hz;     /* frequency the voltage pattern is evaluated at */
    double pa_rad;      /* parallactic-angle rotation of the sky coordinates */
    uint32_t sampling;  /* CF oversampling (even, 2..126) */
    uint32_t n_w;       /* W planes: w_k = k^2 / w_scale, as chip_gridder_attach_wproject */
    double max_abs_w;   /* largest |w| in wavelengths: w_scale = (n_w - 1)^2 / (1.05 max|w|) */
} chip_airy_params;

/* AWProject plan rejections (AwProjectSamplePlanRejection, gridder.rs:2133-2138). */
#define CHIP_AW_OK 0
#define CHIP_AW_NON_FINITE_COORDINATE 1
#define CHIP_AW_OUTSIDE_GRID 2
#define CHIP_AW_KERNEL_INDEX_OUTSIDE_CELL 3
#define CHIP_AW_INVALID_NORMALIZATION 4
#define CHIP_AW_NO_CELL 5 /* cf_cell_dev / pointing_dev index outside the atlas (no reference counterpart) */

/* ---- context ------------------------------------------------------------ */
int chip_ctx_create(int device, chip_ctx **out, char **err);
void chip_ctx_destroy(chip_ctx *ctx);
/* external != 0: launch subsequent work on `stream` (a hipStream_t of the
 * context's device, NULL meaning the legacy default stream), e.g. the caller's
 * current stream, so that the library's kernels are ordered after the
 * caller's own work on the same buffers.  external == 0 restores the
 * context's own stream.  Every call still returns only after its work on that
 * stream has completed.  No reference counterpart (the Metal path owns its
 * command queue, lib.rs:72270-72290). */
int chip_ctx_set_stream(chip_ctx *ctx, int external, void *stream, char **err);
void chip_error_free(char *err);
const char *chip_version(void);

/* ---- gridder (StandardGridder::new_with_padding_factor, gridder.rs:392-443) */
int chip_gridder_create(chip_ctx *ctx, const chip_geometry *geometry, chip_gridder **out,
                        char **err);
void chip_gridder_destroy(chip_gridder *g);
int chip_gridder_info_get(const chip_gridder *g, chip_gridder_info *out, char **err);
/* Host copies of the tap table ((os+1) x 7, gridder.rs:2026-2051) and grid
 * correction axes (gridder.rs:4175-4188). Any pointer may be NULL. */
int chip_gridder_tables(const chip_gridder *g, float *taps, float *corr_x, float *corr_y,
                        char **err);

/* Largest |w| over usable samples (w_project_raw_sample_summary, lib.rs:73642-73664). */
int chip_max_abs_w(chip_gridder *g, const chip_samples *s, double *out, char **err);

/* WProjector::new (gridder.rs:3403-3586) built on the GPU: screen in FP64
 * phase, FP32 FFT per plane, 1e-3 support threshold, crop, plane-zero peak and
 * integral normalisation, per-(plane, offset) normalisation table.
 * explicit_planes = 0 selects the automatic plane count (gridder.rs:3801-3830);
 * sampling = 0 selects the reference rule (4 if planes > 1 else 1, :3413). */
int chip_gridder_attach_wproject(chip_gridder *g, double max_abs_w, uint32_t explicit_planes,
                                 uint32_t sampling, char **err);
/* Residency of the W tables (WProjector::new, gridder.rs:3403-3586: the
 * reference keeps them for the projector's lifetime).  max_table_bytes = 0
 * (default) keeps them resident on the device.  Otherwise, when the tables of
 * every plane (quarter planes plus the gathered per-offset taps) exceed
 * max_table_bytes, they are not stored: each call that grids or degrids
 * recomputes them (screen, FFT, crop, normalisation, gather) for chunks of
 * planes whose tables, plus one FFT plane of scratch, fit in max_table_bytes,
 * running each chunk's samples before the next chunk (memory is held only
 * during the call).  Results are bit-identical to the resident tables.  Takes
 * effect immediately (an attached W projector is rebuilt); visibility sets
 * created before must be created again.  chunk_planes (may be NULL) receives
 * the planes per chunk, 0 when resident.  The AW atlas is always resident.
 * No reference counterpart. */
int chip_gridder_wkernel_budget(chip_gridder *g, uint64_t max_table_bytes, uint32_t *chunk_planes,
                                char **err);
/* Host copies: supports[planes], kernels[planes][width][width] complex64,
 * norm_table[planes][2R+1][2R+1].  Any pointer may be NULL. */
int chip_wproject_tables(const chip_gridder *g, int32_t *supports, float *kernels,
                         float *norm_table, char **err);

/* ---- AWProject CF atlas -------------------------------------------------- */
/* Attach n_cells CF cells sharing one oversampling (AwProjector::new,
 * gridder.rs:2194-2247: every cell must hold origin +- support*sampling) and
 * n_pointings pointing phase gradients (host, 2 doubles per pointing: radians
 * per oversampled CF pixel along x and y, the projector's
 * phase_gradient_rad_per_sample; mosaic_projector_phase_gradient_rad_per_
 * sample_from_pixel_offset, lib.rs:56657-56671, derives them).  pixels_dev
 * holds n_pixels complex64 values and is copied.  The library precomputes,
 * per (cell, pointing, w-sign, sub-cell offset), the phased taps of
 * for_each_grid_tap (gridder.rs:2636-2666) and the f32 / f64 normalisations
 * of plan_sample_detailed (:2253-2311); afterwards the gridder grids, degrids
 * and plans through AWProject with every entry point that takes samples.
 * w_increment (the cache's WIncr) selects cells when chip_samples.cf_cell_dev
 * is NULL.  Grid weight and sums: a sample grids value*weight (PSF: weight)
 * with its phased taps; normalization_sumwt and reported_sumwt accumulate
 * weight * |grid_normalization| (lib.rs:26535, one hand per sample). */
int chip_gridder_attach_aw(chip_gridder *g, uint32_t n_cells, const chip_cf_cell *cells,
                           const float *pixels_dev, uint64_t n_pixels, uint32_t sampling,
                           uint32_t n_pointings, const double *phase_gradients, double w_increment,
                           char **err);
/* Build the CF cells of an Airy A-term on the device and attach them: per W
 * plane k, the FFT of the screen corr(l) corr(m) PB(l', m') exp(2 pi i w_k
 * (sqrt(1 - l^2 - m^2) - 1)) with PB = V^2 the Airy power pattern
 * (AiryVoltageLookup::evaluate_offsets, lib.rs:57505-57532, at freq_hz) of
 * the sky coordinates (l', m') rotated by pa_rad, on the W-projection screen
 * geometry (gridder.rs:3434-3470), normalised by the plane-zero peak and
 * integral (gridder.rs:3485-3550); each plane becomes a full-plane cell of
 * even shape 2*(S*sampling + R + 1), S its support from the 1e-3 threshold. */
int chip_gridder_attach_aw_airy(chip_gridder *g, const chip_airy_params *p, uint32_t n_pointings,
                                const double *phase_gradients, char **err);
/* The Airy CF cells of chip_gridder_attach_aw_airy at n_freq CF frequencies
 * (p->freq_hz is not read): cell f * n_w + k is frequency freq_hz[f], W plane
 * k, each frequency normalised by its own plane-zero peak and integral; an
 * AWProject cache with a frequency axis (AwConvolutionFunctionCache,
 * awprojection.rs:920-1180).  Samples select their cell through cf_cell_dev
 * (chip_aw_select_cells).  n_freq * n_w * pointings * 2 <= 8191 tap tables. */
int chip_gridder_attach_aw_airy_freqs(chip_gridder *g, const chip_airy_params *p, uint32_t n_freq,
                                      const double *freq_hz, uint32_t n_pointings, const double *phase_gradients,
                                      char **err);
/* CF cell of every sample of a rows x channels input for an atlas from
 * chip_gridder_attach_aw_airy(_freqs): the CF frequency nearest to the
 * sample's (lambda_scale * c) times n_w, plus the W plane round(sqrt(WIncr
 * |w|)) clamped to n_w - 1 (select_key_for_sample, awprojection.rs:1127-1180,
 * without a CASA cache's conjugate-beam and Mueller keys).  cf_cell_dev:
 * uint32 per sample, for chip_samples.cf_cell_dev. */
int chip_aw_select_cells(chip_gridder *g, const chip_samples *s, uint32_t *cf_cell_dev, char **err);
/* Read back the attached atlas: cells (n_cells entries) and, when
 * pixels_dev != NULL, the pixels (complex64, n_pixels).  Either may be NULL. */
int chip_aw_cells(const chip_gridder *g, uint32_t *n_cells, chip_cf_cell *cells, uint64_t *n_pixels,
                  float *pixels_dev, char **err);
/* plan_sample_detailed per sample (no weight / flag filters): plan_dev
 * 6 int32 per sample (loc_x, loc_y, off_x, off_y, conjugate, CHIP_AW_* code),
 * norm32_dev 2 floats (normalization, Complex32), norm64_dev 2 doubles
 * (grid_normalization, Complex64); either norm output may be NULL. */
int chip_aw_plan(chip_gridder *g, const chip_samples *s, int32_t *plan_dev, float *norm32_dev,
                 double *norm64_dev, char **err);
/* CASA PBMath1DAiry table of AiryVoltageLookup::new (lib.rs:57479-57503):
 * 10,000 voltage values and the table's maximum radius (arcmin * GHz,
 * casa_airy_max_radius_arcmin_ghz, lib.rs:57592-57605).  Host only. */
int chip_airy_table(double dish_m, double blockage_m, float *values, double *max_radius, char **err);

/* ---- CASA AWProject CF caches (CFS_*.im / WTCFS_*.im PagedImages) -------- */
/* Scientific key of one cache cell (AwConvolutionFunctionKey,
 * awprojection.rs:36-62): spectral reference frequency, miscinfo WValue,
 * MuellerElement and ParallacticAngle. */
typedef struct {
    double frequency_hz;
    double w_value_lambda;
    int32_t mueller_element;
    double parallactic_angle_deg;
} chip_cf_key;

/* AwConvolutionFunctionKernelMetadata (awprojection.rs:97-128).  The UU/VV
 * affine coordinate is world = crval + cdelt * PC * (pixel - crpix) with
 * uv_pc row-major; names longer than 63 bytes are truncated. */
typedef struct {
    uint64_t shape_x, shape_y;
    uint64_t sampling, x_support, y_support;
    double uv_reference_value[2], uv_reference_pixel[2], uv_increment[2], uv_pc[4];
    double diameter_m, conjugate_frequency_hz, w_increment;
    int32_t conjugate_polarization, polarization, rotationally_symmetric;
    char telescope_name[64], band_name[64];
} chip_cf_kernel_meta;

/* AwConvolutionFunctionInventory and the identity's fingerprint
 * (awprojection.rs:165-190, :1820-1972). */
typedef struct {
    uint64_t paired_cells;
    uint32_t n_freq, n_w, n_mueller, n_pa;
    double w_increment;
    uint64_t metadata_fingerprint;
} chip_cfcache_info;

/* AwConvolutionFunctionCache::open (awprojection.rs:920-1045): index and
 * validate every CFS_*.im / WTCFS_*.im casacore image in `path` (pairs,
 * metadata, a complete frequency x W x Mueller x PA product, W values
 * k^2 / WIncr).  Host only; no device needed.  Errors: CHIP_ERR_CF_CACHE. */
int chip_cfcache_open(const char *path, chip_cfcache **out, char **err);
void chip_cfcache_close(chip_cfcache *c);
/* The inventory: sorted unique frequencies (n_freq), W values (n_w), Mueller
 * elements (n_mueller) and PAs (n_pa); any array may be NULL. */
int chip_cfcache_info_get(const chip_cfcache *c, chip_cfcache_info *info, double *freqs, double *w_values,
                          int32_t *muellers, double *pas, char **err);
/* keys() (awprojection.rs:1058-1061): paired_cells keys in key order
 * (frequency, W, Mueller, PA bit patterns).  Cell index k below is keys[k]. */
int chip_cfcache_keys(const chip_cfcache *c, chip_cf_key *keys, char **err);
/* metadata(key) (:1063-1069): imaging and weight metadata of cell `index`. */
int chip_cfcache_metadata(const chip_cfcache *c, uint64_t index, chip_cf_kernel_meta *imaging,
                          chip_cf_kernel_meta *weight, char **err);
/* nearest_key (:1071-1104).  *index = -1 where the reference returns None. */
int chip_cfcache_nearest_key(const chip_cfcache *c, double frequency_hz, double w_value_lambda, int32_t mueller,
                             double parallactic_angle_deg, int64_t *index, char **err);
/* select_key_for_sample (prediction = 0, :1106-1172) or
 * select_key_for_prediction_sample (prediction != 0; reference_frequency_hz
 * and conjugate_beams unused, :1174-1210).  *index = -1 for None. */
int chip_cfcache_select_key(const chip_cfcache *c, double sample_frequency_hz, double reference_frequency_hz,
                            double w_lambda, int32_t mueller, double parallactic_angle_deg, int conjugate_beams,
                            int prediction, int64_t *index, char **err);
/* load(key) (:1380-1412): host pixels of cell `index`, complex64 (x, y) with
 * y fastest: imaging shape_x * shape_y values, weight likewise; either may be
 * NULL.  Non-finite pixels are a CHIP_ERR_CF_CACHE error. */
int chip_cfcache_load(const chip_cfcache *c, uint64_t index, float *imaging, float *weight, char **err);
/* Every imaging cell of the cache, in key order, as the gridder's AW atlas
 * (chip_gridder_attach_aw with the cache's sampling and WIncr): cell k of the
 * atlas is cache cell k.  All cells must share one sampling. */
int chip_gridder_attach_aw_cache(chip_gridder *g, const chip_cfcache *c, uint32_t n_pointings,
                                 const double *phase_gradients, char **err);
/* Per-sample cell of a rows x channels input for an atlas from
 * chip_gridder_attach_aw_cache, with chip_cfcache_select_key's semantics on
 * the device: sample frequency chan_freq_hz[chan] (host array, or NULL for
 * lambda_scale * c), w = uvw_m[2] * lambda_scale, PA pa_deg_dev[row] (device,
 * or NULL for pa_deg), one Mueller element for the call.  cf_cell_dev: uint32
 * per sample for chip_samples.cf_cell_dev, 0xffffffff where no cell is
 * selected (the planner rejects it). */
int chip_cfcache_select_cells(chip_gridder *g, const chip_samples *s, const double *chan_freq_hz,
                              double reference_frequency_hz, int32_t mueller, const double *pa_deg_dev,
                              double pa_deg, int conjugate_beams, int prediction, uint32_t *cf_cell_dev,
                              char **err);

/* ---- phase-centre rotation of a row stream ------------------------------- */
/* Host FP64: uvrot = gridft_uvw_rotation_matrix(source_dir, target_dir)
 * (casars-imager/src/lib.rs:48074-48090, row-major 3x3) and phrot =
 * uvw_phase_rotation_vector_from_angles(source_dir, target_dir) (:48038-48055);
 * directions are (ra, dec) in radians.  The imager passes (image phase
 * centre, row phase centre) (:47952-47956). */
int chip_phase_rotation(const double *source_dir, const double *target_dir, double *uvrot, double *phrot,
                        char **err);
/* Device FP64 (lib.rs:47064-47069; casa-imaging lib.rs:248-259): per row
 * uvw' = uvw * uvrot (row vector) into uvw_out_dev (may be NULL, may alias
 * uvw_m_dev) and shift = phrot . uvw' into phase_shift_m_dev (may be NULL);
 * per (row, chan), if vis_dev is not NULL, vis *= (cos phi, sin phi) in
 * Float with phi = -2 pi shift f / c (skipped when shift or f is 0).  vis_dev
 * is complex64 [row][chan], freq_hz_dev n_chan values. */
int chip_rotate_rows(chip_ctx *ctx, uint64_t n_rows, uint64_t n_chan, const double *uvrot, const double *phrot,
                     const double *uvw_m_dev, const double *freq_hz_dev, float *vis_dev, double *uvw_out_dev,
                     double *phase_shift_m_dev, char **err);

/* ---- row-stream ingestion ------------------------------------------------ */
/* Host row blocks (MeasurementSet rows as casars-imager reads them: uvw in
 * metres, DATA / FLAG as [corr][chan], WEIGHT per corr or WEIGHT_SPECTRUM per
 * (corr, chan), FLAG_ROW) turned on the GPU into the rows x channels samples
 * of one imaging plane, with double-buffered host-to-device copies: the
 * blocks the caller pushes go through two pinned staging buffers, so the
 * caller reads block k + 1 while block k is copied and converted (the
 * reference's two-live-block read-ahead, ARCHITECTURE.md:461-475).
 * Per (row, chan) of a block (execution.rs:2168-2315):
 *  - polarization CHIP_POL_EXPLICIT (StandardMfsVisibilityPolarization::
 *    Explicit, types.rs:1611-1619): sample = corr[0]; flagged -> excluded;
 *    a non-finite visibility adds to the PSF only when allow_psf_only, else
 *    is excluded; weight = WEIGHT_SPECTRUM or WEIGHT of the corr;
 *  - CHIP_POL_PAIR (CollapsedPair, types.rs:1620-1631): either corr flagged
 *    -> excluded; visibility = collapse(v0, v1) (StandardMfsPairCollapse
 *    Transform::collapse, types.rs:1593-1610, Float complex arithmetic);
 *    non-finite -> excluded; both weights finite and > 0 or excluded;
 *    weight = 0.5 (w0 + w1);
 *  - FLAG_ROW set -> every sample of the row excluded;
 *  - rotate != 0: uvw' = uvw * uvrot and each corr's visibility phase-rotated
 *    before the collapse (chip_rotate_rows' arithmetic; casars-imager
 *    lib.rs:19207-19211 rotates both hands of a pair);
 *  - sumwt_factor on every sample; lambda scale freq / c per channel
 *    (casars-imager/src/lib.rs:557).
 * Excluded samples keep flag = 1 in the output. */
#define CHIP_POL_EXPLICIT 0
#define CHIP_POL_PAIR 1
#define CHIP_PAIR_HALF_SUM 0                   /* (v0 + v1) * 0.5 */
#define CHIP_PAIR_HALF_DIFFERENCE 1            /* (v0 - v1) * 0.5 */
#define CHIP_PAIR_POSITIVE_HALF_IMAG_DIFF 2    /* (v0 - v1) * (0 + 0.5i) */
#define CHIP_PAIR_NEGATIVE_HALF_IMAG_DIFF 3    /* (v0 - v1) * (0 - 0.5i) */

typedef struct {
    uint32_t n_corr;          /* correlations per row in the blocks */
    uint32_t n_chan;          /* channels per row */
    int32_t polarization;     /* CHIP_POL_* */
    uint32_t corr[2];         /* explicit: corr[0]; pair: corr[0], corr[1] */
    int32_t transform;        /* CHIP_PAIR_* (pair only) */
    float sumwt_factor;       /* reported-sumwt multiplicity of every sample */
    int32_t allow_psf_only;   /* explicit: non-finite visibilities still add to the PSF */
    int32_t weight_spectrum;  /* weights per (corr, chan) instead of per corr */
    int32_t rotate;           /* apply uvrot / phrot (chip_phase_rotation) */
    double uvrot[9];
    double phrot[3];
    uint64_t max_rows;        /* capacity of the resident row stream */
    uint64_t block_rows;      /* largest block one push accepts */
} chip_ingest_params;

/* freq_hz: host, n_chan values.  Allocates the resident stream (max_rows x
 * n_chan samples) and two pinned staging buffers of block_rows rows. */
int chip_ingest_create(chip_ctx *ctx, const chip_ingest_params *p, const double *freq_hz, chip_ingest **out,
                       char **err);
void chip_ingest_destroy(chip_ingest *ing);
/* Host arrays of one block of n_rows <= block_rows rows: uvw_m [row][3],
 * data complex64 [row][corr][chan], flag [row][corr][chan] (nonzero =
 * flagged), weight [row][corr] or [row][corr][chan], row_flag [row] (NULL =
 * none).  Returns once the block is in a staging buffer and its copy and
 * conversion are queued (it waits only for the copy of block k - 2); the
 * rows are appended after those of earlier pushes.  Exceeding max_rows is an
 * error (CHIP_ERR_INVALID) and appends nothing.  No reference counterpart as
 * a function (the reference converts rows on the CPU). */
int chip_ingest_push(chip_ingest *ing, uint64_t n_rows, const double *uvw_m, const float *data,
                     const uint8_t *flag, const float *weight, const uint8_t *row_flag, char **err);
/* Waits for every pushed block and describes the resident stream as rows x
 * channels samples (chip_samples: uvw_m, lambda_scale, vis, weight,
 * sumwt_factor, flag); the pointers stay valid until the next push or
 * chip_ingest_destroy. */
int chip_ingest_samples(chip_ingest *ing, chip_samples *out, char **err);
/* Host copies of the resident stream (any pointer may be NULL): uvw_m
 * [rows][3], vis complex64 / weight / sumwt_factor / flag [rows][chan]. */
int chip_ingest_download(chip_ingest *ing, double *uvw_m, float *vis, float *weight, float *sumwt_factor,
                         uint8_t *flag, char **err);

/* ---- density weighting -------------------------------------------------- */
/* Uniform / Briggs reweighting from a density grid of image size
 * (apply_weighting_to_owned_batches_by_sample_groups, weighting.rs:1180-1232;
 * density build :1747-1856; reweight :2501-2592; cells gridder.rs:1888-1934).
 * per_plane != 0 (WeightDensityMode::PerPlane): one density grid per channel,
 * cube convention (Double coordinates rounded; built with the conjugate
 * samples, looked up without); per_plane == 0: one combined grid over every
 * sample, VisImagingWeight convention (Float coordinates truncated).
 * weight_out_dev: one float per sample in input order: w / d (uniform) or
 * w / (d * f2 + 1) (Briggs, f2 = (5 * 10^-robust)^2 / (sum d^2 / sum d));
 * samples off the density grid or with a non-positive weight get 0; flagged
 * or non-gridable samples keep their weight.  f2_host (may be NULL): one value
 * per density plane; density_dev (may be NULL): planes x nx x ny floats. */
int chip_density_weights(chip_gridder *g, const chip_samples *s, uint32_t mode, float robust,
                         uint32_t per_plane, float *weight_out_dev, float *f2_host, float *density_dev,
                         char **err);

/* ---- host arrays: device buffers ---------------------------------------- */
/* A crate that links only this library moves the caller's host arrays
 * (borrowed for the call, as newBufferWithBytesNoCopy does in the Metal seam,
 * lib.rs:72046-72064) through these: device memory of the context's device,
 * copies and zero fill on the context's stream, each call synchronous.  No
 * reference counterpart beyond that Metal buffer handling. */
int chip_buffer_alloc(chip_ctx *ctx, uint64_t bytes, void **dev_out, char **err);
void chip_buffer_free(chip_ctx *ctx, void *dev);
int chip_buffer_upload(chip_ctx *ctx, void *dev, const void *host, uint64_t bytes, char **err);
int chip_buffer_download(chip_ctx *ctx, void *host, const void *dev, uint64_t bytes, char **err);
int chip_buffer_zero(chip_ctx *ctx, void *dev, uint64_t bytes, char **err);

/* ---- the Metal W-projection seam, one to one ------------------------------ */
/* WProjector over the caller's own tables (the reference's projector:
 * flattened_kernel_weights planes x width x width complex64 quarter planes,
 * kernel_support per plane, w_scale, sampling; gridder.rs:3640-3678), host
 * arrays copied at the call.  norm_table (planes x (2R+1)^2 floats, R =
 * (sampling + 1) / 2, WProjector::sample_normalization, gridder.rs:3759-3779)
 * may be NULL: it is then computed on the device from the kernels.  Tables
 * stay resident (chip_gridder_wkernel_budget refuses them). */
int chip_gridder_attach_wproject_tables(chip_gridder *g, uint32_t planes, uint32_t sampling, double w_scale,
                                        uint32_t width, const int32_t *supports, const float *kernels,
                                        const float *norm_table, char **err);
/* WProjectMetalSample (lib.rs:71629-71635): 20 bytes, packed by
 * w_project_metal_sample_from_plan (lib.rs:71708-71754). */
typedef struct {
    uint32_t loc;                   /* loc_x | loc_y << 16 */
    uint32_t plane_support_offsets; /* plane (10 bits) | support << 10 (5) | conjugate << 15 |
                                       (off_x + 128) << 16 | (off_y + 128) << 24 */
    float weight;
    float visibility_re, visibility_im;
} chip_wproject_sample;
/* accumulate_w_project_grid_metal_samples(projector, samples, model_grid,
 * psf_grid, residual_grid, mode) (lib.rs:71863-71871, the non-mac stub
 * :71825-71838): n packed, pre-planned samples (device array of
 * chip_wproject_sample) gridded into psf_dev / residual_dev (complex64 x-major
 * grids, accumulated: the Metal path's chunk grids start at zero) with the
 * arithmetic of W_PROJECT_METAL_SHADER (lib.rs:72656-72724): mode PSF adds
 * weight * cwt to psf_dev, DIRTY adds that and weight * (V * cwt) to
 * residual_dev, RESIDUAL adds weight * ((V - p) * cwt) with p the degrid of
 * model_dev (NULL: zero model).  CHIP_ERR_INVALID when a sample's footprint
 * leaves the grid or the kernel (the other samples are gridded). */
int chip_grid_wproject_samples(chip_gridder *g, uint64_t n, const void *samples_dev, const float *model_dev,
                               float *psf_dev, float *residual_dev, uint32_t mode, char **err);

/* ---- FP64 planning ------------------------------------------------------ */
/* Per-sample plan, bit-exact with the reference CPU planner:
 *  standard (no W attached): plan_positive_taps (gridder.rs:557-566, :1958-1974)
 *     -> {x_start, y_start, wix, wiy, 0, 0}
 *  W-projection: WProjector::plan_sample (gridder.rs:3588-3624)
 *     -> {loc_x, loc_y, off_x, off_y, plane, conj}
 * Rejected samples (validity filters of execution.rs:18270-18300 /
 * lib.rs:73666-73676, out of grid) get all -1.  plan_dev: int32[n][6];
 * norm_dev (may be NULL): W normalisation per sample, NaN when rejected. */
int chip_plan(chip_gridder *g, const chip_samples *s, int32_t *plan_dev, float *norm_dev,
              char **err);

/* ---- resident visibility set: plan once, bin, grid/degrid many times ------ */
/* Plans every sample in FP64, sorts the accepted ones by (channel, uv tile)
 * with an integer radix sort, and stores a packed 20-byte stream plus work
 * items of bounded size.  The input arrays may be freed afterwards. */
int chip_vset_create(chip_gridder *g, const chip_samples *s, chip_vset **out, char **err);
void chip_vset_destroy(chip_vset *vs);
/* sums_host: n_chan entries (may be NULL). */
int chip_vset_sums(const chip_vset *vs, chip_sums *sums_host, char **err);
/* Stream statistics: accepted samples, groups of identical plans, work items,
 * total taps sum (2S+1)^2 over accepted samples (the exact grid-point updates
 * of one gridding pass), and the device duration in ms of the last binned
 * kernel launch (measured with events on the launch stream; -1 before the
 * first launch).  Any pointer may be NULL.  No reference counterpart: the
 * reference reports stage timings through ImagingStageTimings
 * (lib.rs:3496-3503). */
int chip_vset_stats(chip_vset *vs, uint64_t *n_valid, uint64_t *n_groups, uint64_t *n_items,
                    uint64_t *taps, double *last_kernel_ms, char **err);
/* One launch over every (channel, tile) work item.  Grids are CHIP_F32
 * complex64, n_chan grids each; they are accumulated into, not cleared.
 * mode PSF/DIRTY use psf_dev (and dirty_dev); RESIDUAL reads model_dev
 * (may be NULL = zero model) and writes dirty_dev. */
int chip_vset_grid(chip_vset *vs, uint32_t mode, const float *model_dev, float *psf_dev,
                   float *dirty_dev, char **err);
/* Degrid model_dev into vis_out_dev (complex64, one per input sample, in input
 * order); rejected samples are written as 0.  This is the forward operator
 * (StandardGridder::degrid_sample_taps_planned_normalized, gridder.rs:1278-1312;
 * WProjector::degrid_sample_planned, :3726-3747). */
int chip_vset_degrid(chip_vset *vs, const float *model_dev, float *vis_out_dev, char **err);
/* Exact adjoint of chip_vset_degrid: grids values_dev (complex64, one per input
 * sample, input order) with unit weight into grid_dev (accumulated), i.e.
 * grid += value * tap (gridder.rs:3680-3699 with value in place of V*w). */
int chip_vset_grid_values(chip_vset *vs, const float *values_dev, float *grid_dev, char **err);
/* Major-cycle residual refresh of every channel in one call
 * (compute_residual_w_project_replay, lib.rs:70895-70963; standard:
 * compute_residual_standard_tiled, lib.rs:68862-68932):
 *   apodise model (gridder.rs:1554-1592) -> centered FFT -> degrid + grid
 *   (V - p) w -> centered inverse FFT -> crop/correct -> / nsum / psf_peak
 *   (x 2 for W-projection).
 * model_dev: n_chan images float32 [chan][nx][ny]; residual_out_dev: same
 * shape; nsum, psf_peak: host arrays of n_chan (normalization sum of weights,
 * PSF peak of each channel, as returned with the PSF).  Grids and FFT plans
 * live in a workspace owned by the set, allocated on the first call (two
 * complex64 grid cubes); with the fused strategy the first call also times
 * the candidate launch shapes of the passes and keeps the fastest.
 * Synchronous on the context's stream. */
int chip_vset_major_cycle(chip_vset *vs, const float *model_dev, const double *nsum,
                          const float *psf_peak, float *residual_out_dev, char **err);
/* Transform strategy of later major cycles (frees the current workspace):
 *   CHIP_CYCLE_AUTO   fused when the grid allows it, else hipFFT (default);
 *   CHIP_CYCLE_HIPFFT hipFFT 2-D real-to-complex (model) and complex-to-real
 *                     (Hermitian part of the residual grid) plans over
 *                     [chan][x][y], separate placement and crop kernels;
 *   CHIP_CYCLE_FUSED  row / column FFT passes with placement, apodisation,
 *                     Hermitian half spectrum, crop and normalisation fused
 *                     in (grid lengths factoring into 2, 3, 5; CHIP_ERR_INVALID
 *                     at the next major cycle otherwise).
 * No reference counterpart (the reference's DirtyProductFftPolicy picks its
 * host FFT backend, lib.rs:1280-1310). */
int chip_vset_cycle_strategy(chip_vset *vs, int strategy, char **err);
/* Launch schedule of the binned kernel: the uv tile edge (a power of two
 * whose (tile + 2 halo)^2 LDS tile fits the device), the order of the groups
 * inside a (channel, tile) bin (0: by anchor cell, so consecutive footprints
 * overlap; 1: by plan, so consecutive groups share their tap table) -- both
 * applied by re-binning the stream on the device, group membership unchanged
 * -- threads per workgroup (a multiple of 64, at most 512) and groups of
 * identical plans per work item (the splitting of each bin).  Unless fixed
 * here, the first major cycle times every candidate of {the set's tile, 16,
 * 32, 64 where they fit} x {256, 512} threads x {256, 512, 1024} groups with
 * anchor order, then the same threads x groups with plan order at the
 * fastest tile, on its own grids, and keeps the fastest (re-binning sorts
 * the groups, not the samples); it starts no new candidate after 3 s of
 * timed launches, and a
 * candidate whose launch takes over 20 ms is timed once without a warm-up
 * (default before that: 32 or the largest fitting tile, anchor order, 512,
 * 1024).
 * tile = 0 keeps the current tile; tile = threads = groups_per_item = 0 asks
 * for autotuning again at the next major cycle.  Results do not depend on the
 * schedule beyond the order of float additions.  No reference counterpart
 * (the Metal path dispatches fixed-size chunks, lib.rs:72180-72290). */
/* Groups the binned kernel processes per wavefront: 2 runs two groups at
 * once, one per half-wavefront (32 lanes per footprint), which shares the
 * per-group work that does not scale with the footprint; it applies to
 * halos up to 4 and otherwise falls back to 1.  groups = 0 (the default)
 * leaves the choice to the schedule autotuner of the first major cycle
 * (one until then).  in_use (may be NULL) receives the value in use.
 * Results agree within FP32 accumulation order.  No reference counterpart
 * (a launch-shape knob of the batched kernel). */
int chip_vset_wave_groups(chip_vset *vs, uint32_t groups, uint32_t *in_use, char **err);
/* Residual refreshes of work items whose largest half-support is 16 - 31
 * (footprint sides 33 - 63) run a two-phase kernel by default: predictions of
 * the item's groups first, then gridding with the footprint's x residues
 * spread over wavefronts and register accumulation.  enable = 0 selects the
 * chunked kernel with per-tap LDS atomics instead (both agree within FP32
 * accumulation order).  Other modes and larger supports always take the
 * chunked kernel.  No reference counterpart (a kernel choice of the batched
 * path). */
int chip_vset_two_phase(chip_vset *vs, int enable, char **err);
int chip_vset_schedule(chip_vset *vs, uint32_t tile, uint32_t order, uint32_t threads, uint32_t groups_per_item,
                       char **err);
/* Schedule in use and, after autotuning, the timed candidates: n_timed
 * quintuples (tile, order, threads, groups per item, ms) in timed (order + 2:
 * timed with two groups per wavefront, chip_vset_wave_groups; at most 96
 * quintuples; may be NULL).  Any pointer may be NULL. */
int chip_vset_schedule_info(const chip_vset *vs, uint32_t *tile, uint32_t *order, uint32_t *threads,
                            uint32_t *groups_per_item, uint32_t *n_timed, double *timed, char **err);
/* Graph replay of later major cycles: enable = 1 captures the next major
 * cycle's kernels (forward passes, zero fill, binned residual kernel, inverse
 * passes) into a HIP graph once and replays it with one launch per call, on
 * the context's stream; a call with other model / residual pointers captures
 * again.  Needs the fused strategy (CHIP_ERR_INVALID at the major cycle
 * otherwise) and resident W tables (CHIP_ERR_INVALID here otherwise).  While
 * replaying, chip_vset_stats reports no binned-kernel time (-1).  enable = 0
 * returns to eager launches, -1 leaves the setting unchanged.  replays (may
 * be NULL) receives the number of major cycles run by replay so far.  No
 * reference counterpart (the Metal path encodes one command buffer per
 * dispatch, lib.rs:72270-72290). */
int chip_vset_cycle_graph(chip_vset *vs, int enable, uint64_t *replays, char **err);
/* Launch order of the binned kernel's work items.  order 0: by (plane,
 * tile); order 1: when the set's planes come in runs of equal lambda scale
 * (the correlations of one channel of a rows x channels input share every
 * plan and tap table), the items of one (run, tile, split) are placed so that
 * they run together on one XCD (workgroups are dealt to the 8 XCDs round
 * robin; empty padding items keep the alignment), and the tap tables one
 * reads again come from that XCD's L2 (the default when such runs exist;
 * mosaic7's binned kernel 16.5 -> 15.9 ms).  order UINT32_MAX leaves the order.
 * sibling (may be NULL) receives the run length found when the set was
 * created (1: none).  Results do not depend on the order beyond the order of
 * float additions.  No reference counterpart. */
int chip_vset_item_order(chip_vset *vs, uint32_t order, uint32_t *sibling, char **err);
/* Per-pass timing of later eager major cycles with the fused strategy:
 * enable = 1 records events before, between and after the six passes (a few
 * microseconds per cycle), 0 stops, -1 leaves the setting unchanged.
 * pass_ms (6 doubles, may be NULL) receives the last timed cycle's device ms
 * of: forward rows, forward columns, zero fill, binned residual kernel,
 * inverse columns, inverse rows (-1 before one was timed).  No reference
 * counterpart (the reference's ImagingStageTimings, lib.rs:3496-3503, time
 * host stages). */
int chip_vset_cycle_passes(chip_vset *vs, int enable, double *pass_ms, char **err);
/* Strategy in use (-1 before the first major cycle); shape[3] of the fused
 * passes (columns per column-pass workgroup, its threads, rows per rows-pass
 * workgroup; zeros for hipFFT); the timed forward + inverse transform ms of
 * the chosen shape (-1 if not timed). */
int chip_vset_cycle_info(const chip_vset *vs, int *strategy, uint32_t *shape, double *transform_ms,
                         char **err);

/* ---- plain per-channel path (CHIP_F64 or CHIP_F32 grids, no binning) ------ */
/* One launch per channel, one thread per sample, global atomics.  CHIP_F64
 * accumulates complex128 grids with FP64 MACs; CHIP_F32 complex64 grids with
 * FP32 MACs.  sums_host: n_chan entries (may be NULL). */
int chip_grid_plain(chip_gridder *g, const chip_samples *s, uint32_t mode, uint32_t precision,
                    const void *model_dev, void *psf_dev, void *dirty_dev, chip_sums *sums_host,
                    char **err);
/* chip_grid_plain without any allocation, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  No counterpart in the reference.  Same arguments,
 * grids and results; the per-channel FP64 sums live in a fixed device array
 * of the library (3 doubles per channel, CHIP_NOALLOC_MAX_CHAN channels at
 * most; more is CHIP_ERR_INVALID), cleared by a kernel at every call, so
 * calls that overlap on one GPU race on it.  W tables must be resident
 * (CHIP_ERR_INVALID under a chip_gridder_wkernel_budget that recomputes
 * them).  Synchronous, like chip_grid_plain. */
#define CHIP_NOALLOC_MAX_CHAN 4096
int chip_grid_plain_noalloc(chip_gridder *g, const chip_samples *s, uint32_t mode, uint32_t precision,
                            const void *model_dev, void *psf_dev, void *dirty_dev, chip_sums *sums_host,
                            char **err);
/* Plain degridding; vis_out_dev is complex128 for CHIP_F64 (FP64 MACs end to
 * end), complex64 for CHIP_F32; rejected samples get 0. */
int chip_degrid_plain(chip_gridder *g, const chip_samples *s, uint32_t precision,
                      const void *model_dev, void *vis_out_dev, char **err);

/* ---- grid <-> image ------------------------------------------------------ */
/* Dirty products (dirty_grids_to_psf_and_residual_owned, lib.rs:48987-49131;
 * W-projection: raw_psf_grid_to_w_project_state, lib.rs:71009-71039 and
 * :70883-70885): centered inverse FFT scaled 1/(gx*gy), crop, grid correction
 * (x 1/sinc for W), PSF / sumwt -> peak -> / peak, residual / sumwt / peak.
 * psf_grids_dev / dirty_grids_dev hold `batch` grids and are overwritten by
 * the transform.  psf_grids_dev may be NULL (residual only): then
 * psf_peak_host supplies the peaks.  Otherwise psf_peak_host receives them.
 * normalization_sumwt_host: batch values. */
int chip_finish_products(chip_gridder *g, uint32_t precision, uint64_t batch, void *psf_grids_dev,
                         void *dirty_grids_dev, const double *normalization_sumwt_host,
                         float *psf_out_dev, float *residual_out_dev, float *psf_peak_host,
                         char **err);
/* Model image(s) -> model grid(s): apodize (x grid correction, x sinc for W,
 * gridder.rs:1554-1592) into a zero grid, unnormalised centered forward FFT
 * (fft.rs:34). model_dev: batch x nx x ny float; grid_out_dev: batch grids. */
int chip_model_to_grid(chip_gridder *g, uint32_t precision, uint64_t batch, const float *model_dev,
                       void *grid_out_dev, char **err);
/* Adjoint of chip_model_to_grid for real images: Re(factor * crop(F^H grid))
 * with F^H the unnormalised centered inverse transform (the conjugate
 * transpose of the forward model transform).  grid_dev (complex64) is
 * overwritten.  image_out_dev: batch x nx x ny float. */
int chip_grid_to_model_adjoint(chip_gridder *g, uint64_t batch, float *grid_dev,
                               float *image_out_dev, char **err);

/* ---- MODEL_DATA prediction ---------------------------------------------- */
/* Model visibilities of the standard MFS model predictor
 * (StandardMfsModelPredictor::predict, lib.rs:3387-3403), the reference's
 * MODEL_DATA writer: per sample, the 7x7 spheroid taps before normalisation
 * (sample_taps_unnormalized, gridder.rs:1997-2023) read from the model grid
 * and divided by their sum (degrid_sample_product_planned_sectdgrid,
 * gridder.rs:1227-1266); 0 for a sample whose taps leave the grid or whose
 * coordinates are not finite (the predictor's fallback).  The gridder is a
 * standard one (CHIP_ERR_INVALID with W or AW kernels attached); the
 * reference's predictor is StandardGridder::new_with_casa_composite_padding
 * (CHIP_PAD_COMPOSITE_1_3, oversampling 100) and its model grid
 * centered_fft2(apodize_model(model)), i.e. chip_model_to_grid with CHIP_F32
 * (lib.rs:3376-3384).  Samples: u_dev / v_dev per sample (n_chan = 1) or
 * uvw_m_dev with lambda_scale_dev; vis, weight, flags and w are not read.
 * model_grids_dev: n_model_grids complex64 grids, 1 (shared) or n_chan (one
 * per channel).  vis_out_dev: complex64, one per sample in input order. */
int chip_predict_model(chip_gridder *g, const chip_samples *s, const float *model_grids_dev,
                       uint64_t n_model_grids, float *vis_out_dev, char **err);

/* ---- minor cycle --------------------------------------------------------- */
/* Hogbom minor-cycle controls (CleanConfig, types.rs): loop gain, global
 * threshold (Jy/beam) and component budget (hogbom_component_budget,
 * lib.rs:64541-64551: niter, plus one in CASA's inclusive mode). */
typedef struct {
    float gain;
    float threshold;
    uint32_t budget;
} chip_hogbom_params;

/* One Hogbom minor cycle on every plane of a cube, on the device
 * (run_hogbom_minor_cycle_cpu_with_progress, lib.rs:59748-59885): per plane,
 * until the budget or a stop rule (minor_cycle_stop_reason, lib.rs:74653-
 * 74677, with the plane's cycle and nsigma thresholds): peak = largest |value|
 * over unmasked pixels, ties to the smallest y*nx + x, the first unmasked
 * pixel in x-major order kept when it is NaN (lib.rs:74064-74115); component
 * = gain * peak; model[peak] += component; residual -= component * psf shifted
 * to the peak with centre (kx/2, ky/2) over the overlap (lib.rs:74805-74868).
 * Planes are bit-identical to the reference's CPU loop.  residual_dev,
 * model_dev: n_planes x nx x ny float (updated in place); psf_dev: n_planes x
 * kx x ky float; mask_dev: n_planes x nx x ny bytes (nonzero = cleanable) or
 * NULL; cycle_threshold, nsigma_threshold: host arrays of n_planes;
 * updates_host, stop_host (CHIP_CLEAN_*): host arrays of n_planes.  One
 * workgroup per plane; synchronous. */
int chip_hogbom(chip_ctx *ctx, const chip_hogbom_params *params, uint64_t n_planes, uint64_t nx, uint64_t ny,
                float *residual_dev, uint64_t kx, uint64_t ky, const float *psf_dev, float *model_dev,
                const uint8_t *mask_dev, const float *cycle_threshold, const float *nsigma_threshold,
                uint32_t *updates_host, int32_t *stop_host, char **err);

#ifdef __cplusplus
}
#endif

#endif /* CASA_IMAGING_HIP_H */
