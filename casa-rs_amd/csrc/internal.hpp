// Internal declarations shared by the HIP translation units of
// libcasa_imaging_hip.so.  Nothing here is part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>
#include <hipfft/hipfft.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/casa_imaging_hip.h"

namespace chip {

constexpr int kSupport = 3;  // STANDARD_GRIDDER_SUPPORT, gridder.rs:15
constexpr int kTaps = 7;     // STANDARD_GRIDDER_TAP_COUNT, gridder.rs:16

// Error carrying a C-ABI status code.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Error(code, msg); }

#define CHIP_HIP(expr)                                                                     \
    do {                                                                                   \
        hipError_t _e = (expr);                                                            \
        if (_e != hipSuccess)                                                              \
            ::chip::fail(CHIP_ERR_DEVICE, std::string(#expr " failed: ") +                 \
                                              hipGetErrorString(_e));                      \
    } while (0)

#define CHIP_FFT(expr)                                                                     \
    do {                                                                                   \
        hipfftResult _r = (expr);                                                          \
        if (_r != HIPFFT_SUCCESS)                                                          \
            ::chip::fail(CHIP_ERR_DEVICE,                                                  \
                         std::string(#expr " failed with hipfft status ") +                \
                             std::to_string(static_cast<int>(_r)));                        \
    } while (0)

// Owning device buffer.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    explicit DevBuf(size_t n) { alloc(n); }
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
        return *this;
    }
    ~DevBuf() { release(); }
    void alloc(size_t n) {
        release();
        if (n) CHIP_HIP(hipMalloc(&p, n));
        bytes = n;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

// hipFFT plans cached per (gx, gy, batch, type).
struct FftCache {
    std::map<std::tuple<uint64_t, uint64_t, uint64_t, int>, hipfftHandle> plans;
    hipfftHandle get(uint64_t gx, uint64_t gy, uint64_t batch, hipfftType type, hipStream_t s);
    void clear();
    ~FftCache() { clear(); }
};

}  // namespace chip

namespace chip {
// Everything the W tables of a WProjector are computed from (wkernel.hip).
struct WkState {
    uint32_t conv = 0, inner = 0, quarter = 0, planes = 0, sampling = 1;
    int32_t R = 0;
    double s0 = 0.0, s1 = 0.0, w_scale = 1.0;
    uint32_t width = 0, stride = 0;  // common quarter-plane edge, compact taps per (plane, offset)
    float peak0 = 0.0f, pb = 0.0f;   // plane-zero FFT peak and integral at integer taps
    std::vector<uint32_t> crop;      // per plane: cropped quarter-plane edge
    DevBuf d_corr;                   // screen correction axis (inner floats)
    // attach arguments, to rebuild when the residency budget changes
    double max_abs_w = 0.0;
    uint32_t explicit_planes = 0, sampling_arg = 0;
};
}  // namespace chip

namespace chip {

// One CASA CF-cache cell key (AwConvolutionFunctionKey, awprojection.rs:36-62).
struct CfKey {
    double frequency_hz = 0.0, w_value_lambda = 0.0;
    int32_t mueller_element = 0;
    double parallactic_angle_deg = 0.0;
};

// AwConvolutionFunctionKernelMetadata (awprojection.rs:97-128); uv holds
// crval, crpix, cdelt (two each) and the row-major PC matrix.
struct CfKernelMeta {
    std::string path, relative;
    uint64_t shape[2] = {0, 0};
    uint64_t sampling = 0, x_support = 0, y_support = 0;
    double uv[10] = {};
    std::string telescope, band;
    double diameter_m = 0.0, conj_freq_hz = 0.0;
    int32_t conj_poln = 0, polarization = 0;
    double w_increment = 0.0;
    bool rotationally_symmetric = false;
};

// Selection axes of a validated cache: sorted unique frequencies, W values,
// Mueller elements and PAs, and the cell of every index combination.
struct CfAxes {
    std::vector<double> freqs, w_values, pas;
    std::vector<int32_t> muellers;
    std::vector<uint32_t> cell_of;  // ((f * n_w + w) * n_mueller + m) * n_pa + p
    double w_increment = 0.0;
};

// A validated, indexed cache (AwConvolutionFunctionCache, awprojection.rs:920-1045);
// cells are in key order (frequency, W, Mueller, PA bit patterns).
struct CfCache : CfAxes {
    std::string root;
    std::vector<CfKey> keys;
    std::vector<CfKernelMeta> imaging, weight;
    uint64_t fingerprint = 0;
};

void cf_cache_open(CfCache &c, const std::string &root);
CfKernelMeta cf_read_kernel_metadata(const std::string &dir, const std::string &relative, CfKey &key);
std::vector<float> cf_read_kernel_pixels(const CfKernelMeta &m);
int64_t cf_nearest_key(const CfCache &c, double f, double w, int32_t mueller, double pa);
int64_t cf_select_key(const CfCache &c, double f, double f_ref, double w, int32_t mueller, double pa,
                      bool conjugate_beams, bool prediction);

}  // namespace chip

struct chip_cfcache {
    chip::CfCache c;
};

struct chip_ctx {
    int device = 0;
    hipStream_t own_stream = nullptr;  // created with the context
    hipStream_t stream = nullptr;      // where work is launched: own_stream or the caller's
    chip::FftCache fft;
};

// Host-side tables of a StandardGridder and, optionally, its WProjector.
struct chip_gridder {
    chip_ctx *ctx = nullptr;
    int device = 0;  // kept so that destruction never dereferences ctx
    chip_geometry geom{};
    uint64_t gx = 0, gy = 0, blc_x = 0, blc_y = 0;
    double du = 0, dv = 0;
    uint32_t os = 0;
    std::vector<float> taps;          // (os+1) x 7
    std::vector<float> corr_x, corr_y;
    chip::DevBuf d_taps, d_corr_x, d_corr_y;
    chip::DevBuf d_taps_raw;          // (os+1) x 7 before normalisation (model predictor)

    // W projection
    bool wproject = false;
    uint32_t planes = 1, sampling = 1, width = 0, max_support = 0;
    int32_t norm_radius = 0;
    double w_scale = 1.0;
    std::vector<int32_t> supports;
    std::vector<float> norm_table;  // planes x (2R+1)^2
    chip::DevBuf d_kernels;         // planes x width x width float2
    chip::DevBuf d_supports, d_norm_table;
    // Taps of every (plane, off_x, off_y) gathered contiguously so that the
    // lanes of a wavefront read one sample's taps with coalesced loads.
    chip::DevBuf d_compact;         // planes x (2R+1)^2 x tap_stride float2
    uint32_t tap_stride = 0;
    // Residency of the W tables (d_kernels, d_compact).  wk_budget = 0 keeps
    // them resident; otherwise, when they exceed wk_budget bytes, they are not
    // stored: every call that grids or degrids recomputes them in chunks of
    // wk_chunk planes (screen, FFT, crop, normalisation, gather) and runs the
    // chunk's samples before the next chunk overwrites them.  wk_chunk = 0:
    // resident.  wk holds what the recomputation needs.
    uint64_t wk_budget = 0;
    uint32_t wk_chunk = 0;
    bool wk_external = false;  // tables given by the caller (chip_gridder_attach_wproject_tables)
    chip::WkState wk;
    std::vector<float> sinc;        // build_sinc_axis(max(gx, gy), sampling)
    std::vector<float> sinc_x, sinc_y;  // per-axis sinc used by model apodisation
    chip::DevBuf d_sinc, d_sinc_x, d_sinc_y;

    // AWProject CF atlas (aw.hip).  When attached, wproject is set too and the
    // W tap-table fields describe the atlas's tap tables: "plane" t =
    // (cell * pointings + pointing) * 2 + conjugate, supports[t] the cell's
    // larger half-support (tables are square, zero outside the cell's own
    // x/y supports), d_compact the phased taps.
    bool aw = false;
    uint32_t aw_cells = 0, aw_pointings = 1;
    double aw_w_incr = 0.0;
    std::vector<chip_cf_cell> aw_meta;
    uint64_t aw_n_pixels = 0;
    chip::DevBuf d_aw_pixels;  // complex64 copy of the cells
    chip::DevBuf d_aw_cell;    // AwCell per cell
    chip::DevBuf d_aw_n32;     // cells x (2R+1)^2 float2: unconjugated Complex32 normalisation
    chip::DevBuf d_aw_n64;     // cells x (2R+1)^2 double2: unconjugated Complex64 normalisation
    chip::DevBuf d_aw_mag;     // cells x (2R+1)^2 double: |grid_normalization|, 0 where invalid
    chip::DevBuf d_aw_tnorm;   // tables x (2R+1)^2 float2: normalisation with the table's w sign
    // Airy atlas built on the device: W planes per CF frequency and the CF
    // frequencies (cell f * aw_n_w + k is frequency f, W plane k)
    uint32_t aw_n_w = 0;
    std::vector<double> aw_cf_freq;
    // Atlas attached from a CASA CF cache: its selection axes (cells in key order)
    bool aw_from_cache = false;
    chip::CfAxes aw_cache_axes;
};

namespace chip {

// ---- host math (tables.cpp) ----
double grdsf(double nu);
float spheroidal_kernel(double distance, double support);
uint64_t padded_len(uint64_t n, double padding);
bool is_composite(uint64_t n);
uint64_t composite_padded_len(uint64_t n, double padding);
std::vector<float> correction_axis(uint64_t n);
std::vector<float> gridder_correction_axis(uint64_t n);
std::vector<float> sinc_axis(uint64_t n, uint64_t sampling);
std::vector<float> tap_weights(uint32_t os);
std::vector<float> tap_weights_raw(uint32_t os);
uint32_t suggested_w_plane_count(const chip_geometry &g, double max_abs_w);
double w_max_uvw(const chip_geometry &g, double max_abs_w, bool explicit_planes);
uint32_t find_w_support(const std::vector<float> &axis_mag_over_peak, uint32_t conv,
                        uint32_t sampling);

// ---- device launch helpers ----
// Parameters of the FP64 planner, passed by value to kernels.
struct PlanParams {
    uint64_t gx, gy;
    double du, dv;
    double half_gx, half_gy;  // gx/2.0, gy/2.0
    uint32_t os;              // standard tap oversampling
    uint32_t wproject;
    uint32_t sampling;
    uint32_t planes;
    double w_scale;
    int32_t norm_radius;
    const int32_t *supports;  // device
    const float *norm_table;  // device
    // AWProject (aw = 1): per-cell geometry, per (cell, offset) |grid_normalization|
    uint32_t aw;
    uint32_t aw_cells, aw_pointings;
    double aw_w_incr;
    const struct AwCell *aw_cell;
    const double *aw_mag;
    const uint32_t *cf_cell;   // per sample, or NULL (W-plane selection)
    const uint32_t *pointing;  // per row, or NULL
};

// Device view of one CF cell.
struct AwCell {
    int64_t shape_x, shape_y;
    int32_t xs, ys;      // half-supports
    int32_t cx, cy;      // aw_kernel_origin of each axis
    uint64_t offset;     // first pixel (complex64 elements)
};

PlanParams plan_params(const chip_gridder &g);

// Source descriptor for device kernels (mirror of chip_samples).
struct Source {
    uint64_t n_rows, n_chan;
    const double *u, *v, *w, *uvw_m, *lambda_scale;
    const float2 *vis;
    const float *weight, *sumwt_factor;
    const uint8_t *gridable, *flag;
    const uint32_t *cf_cell, *pointing;
};
Source make_source(const chip_samples &s);
// plan_params with the source's per-sample AWProject selections
PlanParams plan_params(const chip_gridder &g, const Source &src);
void validate_source(const chip_samples &s);

void launch_plan_raw(const chip_gridder &g, const Source &src, int32_t *plan, float *norm,
                     hipStream_t st);
double launch_max_abs_w(const chip_gridder &g, const Source &src, hipStream_t st);
void build_wproject(chip_gridder &g, double max_abs_w, uint32_t explicit_planes,
                    uint32_t sampling);

void grid_plain(chip_gridder &g, const Source &src, uint32_t mode, uint32_t precision,
                const void *model, void *psf, void *dirty, chip_sums *sums);
void degrid_plain(chip_gridder &g, const Source &src, uint32_t precision, const void *model,
                  void *out);
void noalloc_scratch_init(int device);
void attach_wproject_tables(chip_gridder &g, uint32_t planes, uint32_t sampling, double w_scale, uint32_t width,
                            const int32_t *supports, const float *kernels_host, const float *norm_host);
void grid_wproject_samples(chip_gridder &g, uint64_t n, const void *samples, const float *model, float *psf,
                           float *residual, uint32_t mode);
void grid_plain_noalloc(chip_gridder &g, const Source &src, uint32_t mode, uint32_t precision,
                        const void *model, void *psf, void *dirty, chip_sums *sums);

void finish_products(chip_gridder &g, uint32_t precision, uint64_t batch, void *psf, void *dirty,
                     const double *nsumwt, float *psf_out, float *res_out, float *peaks);
void model_to_grid(chip_gridder &g, uint32_t precision, uint64_t batch, const float *model,
                   void *grid_out);
void grid_to_model_adjoint(chip_gridder &g, uint64_t batch, float *grid, float *image_out);
void predict_model(chip_gridder &g, const chip_samples &s, const float *grids, uint64_t n_grids, float *out);
void density_weights(chip_gridder &g, const chip_samples &cs, uint32_t mode, float robust, uint32_t per_plane,
                     float *weight_out, float *f2_host, float *density_out);
void phase_rotation(const double src[2], const double tgt[2], double uvrot[9], double phrot[3]);
void hogbom(chip_ctx &ctx, const chip_hogbom_params &hp, uint64_t n_planes, uint64_t nx, uint64_t ny,
            float *residual, uint64_t kx, uint64_t ky, const float *psf, float *model, const uint8_t *mask,
            const float *cycle_threshold, const float *nsigma_threshold, uint32_t *updates, int32_t *stop);
void attach_aw(chip_gridder &g, uint32_t n_cells, const chip_cf_cell *cells, const float *pixels_dev,
               uint64_t n_pixels, uint32_t sampling, uint32_t n_pointings, const double *gradients,
               double w_increment);
void attach_aw_airy(chip_gridder &g, const chip_airy_params &p, uint32_t n_freq, const double *freqs,
                    uint32_t n_pointings, const double *gradients);
void aw_select_cells(chip_gridder &g, const chip_samples &cs, uint32_t *out);
void attach_aw_cache(chip_gridder &g, const CfCache &c, uint32_t n_pointings, const double *gradients);
void cf_select_cells(chip_gridder &g, const chip_samples &cs, const double *chan_freq_hz, double f_ref,
                     int32_t mueller, const double *pa_deg_dev, double pa_deg, int conjugate_beams, int prediction,
                     uint32_t *out);
void aw_plan_raw(chip_gridder &g, const Source &src, int32_t *plan, float *n32, double *n64);
double airy_max_radius(double dish, double blockage);
std::vector<float> airy_table(double dish, double blockage);
void rotate_rows(chip_ctx &ctx, uint64_t n_rows, uint64_t n_chan, const double uvrot[9], const double phrot[3],
                 const double *uvw, const double *freq, float *vis, double *uvw_out, double *shift_out);

// Row-stream ingestion (chip_ingest_*): the resident rows x channels
// stream and two staging slots (pinned host + device) of block_rows rows.
struct IngestSlot {
    void *host = nullptr;        // pinned
    DevBuf dev;
    hipEvent_t copied = nullptr;    // the H2D copy from `host` completed
    hipEvent_t consumed = nullptr;  // the conversion from `dev` completed
    bool used = false;
};
void ingest_destroy_slots(IngestSlot *slots, int n);

// Largest dynamic LDS one workgroup may use on `device` (160 KB on gfx950).
inline size_t lds_limit(int device) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess || v <= 0)
        return 64 * 1024;
    return static_cast<size_t>(v);
}

// LDS bytes of the binned kernel's tile for the most demanding mode
// (four float arrays of (tile + 2 halo) x ((tile + 2 halo) | 1) cells, plus
// the standard tap table).
size_t binned_tile_lds(const chip_gridder &g, uint32_t tile, uint32_t halo);

// Major-cycle workspace (cycle.hip): grids and FFT plans kept with the set.
struct CycleWorkspace;

}  // namespace chip

// Resident, binned visibility stream.
struct chip_ingest {
    chip_ctx *ctx = nullptr;
    int device = 0;
    chip_ingest_params p{};
    uint64_t rows = 0;              // rows appended so far
    chip::DevBuf lscale;            // n_chan freq / c
    chip::DevBuf freq;              // n_chan Hz
    chip::DevBuf uvw, vis, wt, sf, flag;
    chip::IngestSlot slot[2];
    int next = 0;
    hipStream_t copy = nullptr;     // H2D copies (default priority)
    ~chip_ingest();
};

struct chip_vset {
    chip_gridder *g = nullptr;
    int device = 0;  // kept so that destruction never dereferences g
    uint64_t n_samples = 0;   // input samples (rows x channels)
    uint64_t n_chan = 0;
    uint64_t n_valid = 0;     // accepted samples in the stream
    uint32_t tile = 32;       // uv tile edge (cells)
    uint32_t halo = 3;        // max half-support
    uint32_t tiles_x = 0, tiles_y = 0;
    uint64_t n_items = 0;
    uint64_t items_version = 0;  // bumped whenever the work items are rewritten
    // W tables recomputed on the fly: the stream is binned by (plane chunk,
    // channel, tile); chunk_items[c] is the first work item of chunk c (plus
    // n_items at the end).  wk_chunk = 0: one chunk.
    uint32_t wk_chunk = 0;
    std::vector<uint32_t> chunk_items;
    // items are ordered by (plane chunk, support class of their largest
    // half-support, bin); range_first[chunk * n_classes + class] is the first
    // item of each range (plus n_items at the end).  The standard gridder has
    // one class.
    std::vector<uint32_t> range_first;
    uint32_t n_classes = 1;
    // residual launches of half-supports 16 - 31 take the two-phase kernel
    // (chip_vset_two_phase); false: the chunked LDS-atomic path
    bool wide = true;
    // bins (plane chunk, channel, tile): groups per bin and first group, kept
    // so that the schedule autotuner can split the work items again
    chip::DevBuf bin_groups, bin_first;
    uint32_t n_bins = 0, n_tiles_all = 0;
    // launch schedule of the binned kernel: threads per workgroup and groups
    // per work item (chosen by the autotuner at the first major cycle unless
    // fixed by chip_vset_schedule)
    uint32_t threads = 512, groups_per_item = 1024;
    // groups per wavefront of the binned kernel: 0 = the autotuner's choice
    // (tuned_wave_groups, one until tuned), 1 = one, 2 = two where the halo
    // allows it (up to 4)
    uint32_t wave_groups = 0, tuned_wave_groups = 1;
    uint32_t order = 0;  // groups of a bin by anchor cell (0) or by plan (1)
    // Planes whose samples share their plans (rows x channels input with equal
    // lambda scales in runs of `sibling`, e.g. the correlations of one
    // channel): with item_order 1 their work items of one (tile, split) run
    // on one XCD at the same time, so that the tap tables they read again
    // come from that XCD's L2 (padding items with no groups keep the XCD
    // round robin aligned).
    uint32_t sibling = 1;
    uint32_t item_order = 0;
    bool schedule_fixed = false, schedule_tuned = false;
    std::vector<double> tune_ms;  // timed candidates: (tile, order, threads, groups per item, ms)
    // stream (sorted by channel, tile)
    chip::DevBuf plan;    // uint2 per sample
    chip::DevBuf vis;     // float2
    chip::DevBuf wt;      // float
    chip::DevBuf index;   // uint32 original sample index
    chip::DevBuf items;   // WorkItem (chan, tile, first group, group count)
    chip::DevBuf grec;    // uint4 per group: first stream position, end, packed plan (x, y)
    chip::DevBuf gsum_d;  // float4 per group: sum w, sum w V (finite V)      (PSF, DIRTY)
    chip::DevBuf gsum_r;  // float4 per group: sum w, sum w V (usable samples) (RESIDUAL)
    uint64_t n_groups = 0;
    std::vector<chip_sums> sums;
    uint64_t taps = 0;    // sum over accepted samples of (2S+1)^2
    // events around the last binned kernel launch
    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    bool timed = false;
    std::shared_ptr<chip::CycleWorkspace> cycle;
    int cycle_strategy = CHIP_CYCLE_AUTO;
    bool cycle_graph = false;  // major cycles replay a captured HIP graph
    bool cycle_passes = false;  // eager fused cycles time each pass with events
    chip::DevBuf cycle_norm;  // per channel: sumwt (f32), then psf peak
    ~chip_vset() {
        if (ev_start) (void)hipEventDestroy(ev_start);
        if (ev_stop) (void)hipEventDestroy(ev_stop);
    }
};
