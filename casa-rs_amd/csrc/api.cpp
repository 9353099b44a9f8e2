// extern "C" entry points of libcasa_imaging_hip.so.  Each wrapper converts
// exceptions into the status / message convention of the header.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "internal.hpp"

namespace chip {
void vset_build(chip_vset &vs, const chip_samples &cs);
void vset_grid(chip_vset &vs, uint32_t mode, const float *model, float *psf, float *dirty);
void vset_degrid(chip_vset &vs, const float *model, float *vis_out);
void vset_grid_values(chip_vset &vs, const float *values, float *grid);
void vset_major_cycle(chip_vset &vs, const float *model, const double *nsum, const float *peaks,
                      float *out);
void vset_cycle_info(const chip_vset &vs, int *strategy, uint32_t shape[3], double *transform_ms);
void ingest_create(chip_ingest &ing, const chip_ingest_params &p, const double *freq_hz);
void ingest_push(chip_ingest &ing, uint64_t n_rows, const double *uvw, const float *data, const uint8_t *flag,
                 const float *weight, const uint8_t *row_flag);
void ingest_samples(chip_ingest &ing, chip_samples &out);
void ingest_download(chip_ingest &ing, double *uvw, float *vis, float *weight, float *sumwt_factor, uint8_t *flag);
void vset_cycle_strategy(chip_vset &vs, int strategy);
void vset_cycle_graph(chip_vset &vs, int enable);
void vset_schedule(chip_vset &vs, uint32_t tile, uint32_t order, uint32_t threads, uint32_t gpi);
int binned_wave_groups(const chip_vset &vs);
uint64_t vset_cycle_replays(const chip_vset &vs);
void vset_cycle_passes(chip_vset &vs, int enable, double *pass_ms);
void vset_item_order(chip_vset &vs, uint32_t order);
void wk_all_kernels(const chip_gridder &g, float2 *kernels_dev);
}  // namespace chip

using namespace chip;

namespace {

char *dup_message(const char *m) {
    const size_t n = std::strlen(m);
    char *r = static_cast<char *>(std::malloc(n + 1));
    if (r) std::memcpy(r, m, n + 1);
    return r;
}

template <class F>
int guarded(char **err, F &&f) {
    if (err) *err = nullptr;
    try {
        f();
        return CHIP_OK;
    } catch (const Error &e) {
        if (err) *err = dup_message(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        if (err) *err = dup_message("host allocation failed");
        return CHIP_ERR_DEVICE;
    } catch (const std::exception &e) {
        if (err) *err = dup_message(e.what());
        return CHIP_ERR_INVALID;
    }
}

void set_device(const chip_ctx *ctx) { CHIP_HIP(hipSetDevice(ctx->device)); }

}  // namespace

namespace {

const CfCache &cache_of(const chip_cfcache *c) {
    if (!c) fail(CHIP_ERR_INVALID, "NULL CF cache");
    return c->c;
}

void copy_meta(const CfKernelMeta &m, chip_cf_kernel_meta *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->shape_x = m.shape[0];
    o->shape_y = m.shape[1];
    o->sampling = m.sampling;
    o->x_support = m.x_support;
    o->y_support = m.y_support;
    for (int k = 0; k < 2; ++k) {
        o->uv_reference_value[k] = m.uv[k];
        o->uv_reference_pixel[k] = m.uv[2 + k];
        o->uv_increment[k] = m.uv[4 + k];
    }
    for (int k = 0; k < 4; ++k) o->uv_pc[k] = m.uv[6 + k];
    o->diameter_m = m.diameter_m;
    o->conjugate_frequency_hz = m.conj_freq_hz;
    o->w_increment = m.w_increment;
    o->conjugate_polarization = m.conj_poln;
    o->polarization = m.polarization;
    o->rotationally_symmetric = m.rotationally_symmetric ? 1 : 0;
    std::strncpy(o->telescope_name, m.telescope.c_str(), sizeof(o->telescope_name) - 1);
    std::strncpy(o->band_name, m.band.c_str(), sizeof(o->band_name) - 1);
}

}  // namespace

extern "C" {

const char *chip_version(void) { return "casa-imaging-hip 0.1.0 (gfx950)"; }

void chip_error_free(char *err) { std::free(err); }

int chip_ctx_create(int device, chip_ctx **out, char **err) {
    return guarded(err, [&] {
        if (!out) fail(CHIP_ERR_INVALID, "out is NULL");
        int count = 0;
        if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
            fail(CHIP_ERR_DEVICE, "no HIP device is available");
        if (device < 0 || device >= count) fail(CHIP_ERR_INVALID, "device index out of range");
        std::unique_ptr<chip_ctx> c(new chip_ctx);
        c->device = device;
        CHIP_HIP(hipSetDevice(device));
        noalloc_scratch_init(device);  // device address of the no-allocation scratch
        CHIP_HIP(hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking));
        c->stream = c->own_stream;
        *out = c.release();
    });
}

int chip_ctx_set_stream(chip_ctx *ctx, int external, void *stream, char **err) {
    return guarded(err, [&] {
        if (!ctx) fail(CHIP_ERR_INVALID, "NULL context");
        ctx->stream = external ? static_cast<hipStream_t>(stream) : ctx->own_stream;
    });
}

void chip_ctx_destroy(chip_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    ctx->fft.clear();
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->own_stream);
    delete ctx;
}

int chip_gridder_create(chip_ctx *ctx, const chip_geometry *geo, chip_gridder **out, char **err) {
    return guarded(err, [&] {
        if (!ctx || !geo || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        // ImageGeometry::validate (types.rs:910-925)
        if (geo->nx < 8 || geo->ny < 8) fail(CHIP_ERR_INVALID, "image shape must be at least 8x8");
        if (!(std::isfinite(geo->cell_x) && geo->cell_x > 0.0 && std::isfinite(geo->cell_y) &&
              geo->cell_y > 0.0))
            fail(CHIP_ERR_INVALID, "cell sizes must be finite positive radians");
        if (geo->oversampling < 2 || geo->oversampling > 254 || geo->oversampling % 2)
            fail(CHIP_ERR_INVALID, "oversampling must be even and in 2..254");
        set_device(ctx);
        std::unique_ptr<chip_gridder> g(new chip_gridder);
        g->ctx = ctx;
        g->device = ctx->device;
        g->geom = *geo;
        switch (geo->padding_mode) {
            case CHIP_PAD_COMPOSITE_1_2:
                g->gx = composite_padded_len(geo->nx, 1.2);
                g->gy = composite_padded_len(geo->ny, 1.2);
                break;
            case CHIP_PAD_COMPOSITE_1_3:
                g->gx = composite_padded_len(geo->nx, 1.3);
                g->gy = composite_padded_len(geo->ny, 1.3);
                break;
            case CHIP_PAD_NONE:
                g->gx = geo->nx;
                g->gy = geo->ny;
                break;
            default:
                fail(CHIP_ERR_INVALID, "unknown padding mode");
        }
        if (g->gx > 65535 || g->gy > 65535) fail(CHIP_ERR_INVALID, "grid edge above 65535");
        // gridder.rs:403-406, :426-427
        g->blc_x = (g->gx - geo->nx + (g->gx % 2 == 0 ? 1 : 0)) / 2;
        g->blc_y = (g->gy - geo->ny + (g->gy % 2 == 0 ? 1 : 0)) / 2;
        g->du = 1.0 / (static_cast<double>(g->gx) * geo->cell_x);
        g->dv = 1.0 / (static_cast<double>(g->gy) * geo->cell_y);
        g->os = geo->oversampling;
        g->taps = tap_weights(g->os);
        g->corr_x = correction_axis(g->gx);
        g->corr_y = correction_axis(g->gy);
        hipStream_t st = ctx->stream;
        g->d_taps.alloc(g->taps.size() * sizeof(float));
        g->d_corr_x.alloc(g->corr_x.size() * sizeof(float));
        g->d_corr_y.alloc(g->corr_y.size() * sizeof(float));
        CHIP_HIP(hipMemcpyAsync(g->d_taps.p, g->taps.data(), g->d_taps.bytes, hipMemcpyHostToDevice, st));
        CHIP_HIP(hipMemcpyAsync(g->d_corr_x.p, g->corr_x.data(), g->d_corr_x.bytes, hipMemcpyHostToDevice, st));
        CHIP_HIP(hipMemcpyAsync(g->d_corr_y.p, g->corr_y.data(), g->d_corr_y.bytes, hipMemcpyHostToDevice, st));
        const std::vector<float> raw = tap_weights_raw(g->os);
        g->d_taps_raw.alloc(raw.size() * sizeof(float));
        CHIP_HIP(hipMemcpyAsync(g->d_taps_raw.p, raw.data(), g->d_taps_raw.bytes, hipMemcpyHostToDevice, st));
        CHIP_HIP(hipStreamSynchronize(st));
        *out = g.release();
    });
}

// Destroy functions only touch the object itself (its device index is copied
// at creation), so children and parents may be destroyed in any order.
void chip_gridder_destroy(chip_gridder *g) {
    if (!g) return;
    (void)hipSetDevice(g->device);
    delete g;
}

int chip_gridder_info_get(const chip_gridder *g, chip_gridder_info *out, char **err) {
    return guarded(err, [&] {
        if (!g || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        chip_gridder_info r{};
        r.gx = g->gx; r.gy = g->gy; r.blc_x = g->blc_x; r.blc_y = g->blc_y;
        r.du = g->du; r.dv = g->dv; r.oversampling = g->os;
        r.wproject = g->wproject ? 1u : 0u;
        r.planes = g->planes; r.sampling = g->sampling; r.width = g->width;
        r.max_support = g->max_support;
        r.norm_radius = static_cast<uint32_t>(g->norm_radius);
        r.aw = g->aw ? 1u : 0u;
        r.w_scale = g->w_scale;
        *out = r;
    });
}

int chip_gridder_tables(const chip_gridder *g, float *taps, float *corr_x, float *corr_y, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        if (taps) std::memcpy(taps, g->taps.data(), g->taps.size() * sizeof(float));
        if (corr_x) std::memcpy(corr_x, g->corr_x.data(), g->corr_x.size() * sizeof(float));
        if (corr_y) std::memcpy(corr_y, g->corr_y.data(), g->corr_y.size() * sizeof(float));
    });
}

int chip_max_abs_w(chip_gridder *g, const chip_samples *s, double *out, char **err) {
    return guarded(err, [&] {
        if (!g || !s || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        *out = launch_max_abs_w(*g, make_source(*s), g->ctx->stream);
    });
}

int chip_gridder_attach_wproject(chip_gridder *g, double max_abs_w, uint32_t explicit_planes,
                                 uint32_t sampling, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        if (g->wproject) fail(CHIP_ERR_INVALID, "W projection already attached");
        set_device(g->ctx);
        build_wproject(*g, max_abs_w, explicit_planes, sampling);
    });
}

int chip_gridder_attach_aw(chip_gridder *g, uint32_t n_cells, const chip_cf_cell *cells,
                           const float *pixels_dev, uint64_t n_pixels, uint32_t sampling,
                           uint32_t n_pointings, const double *phase_gradients, double w_increment,
                           char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        set_device(g->ctx);
        attach_aw(*g, n_cells, cells, pixels_dev, n_pixels, sampling, n_pointings, phase_gradients, w_increment);
    });
}

int chip_gridder_attach_aw_airy(chip_gridder *g, const chip_airy_params *p, uint32_t n_pointings,
                                const double *phase_gradients, char **err) {
    return guarded(err, [&] {
        if (!g || !p) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        attach_aw_airy(*g, *p, 1, &p->freq_hz, n_pointings, phase_gradients);
    });
}

int chip_gridder_attach_aw_airy_freqs(chip_gridder *g, const chip_airy_params *p, uint32_t n_freq,
                                      const double *freq_hz, uint32_t n_pointings, const double *phase_gradients,
                                      char **err) {
    return guarded(err, [&] {
        if (!g || !p) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        attach_aw_airy(*g, *p, n_freq, freq_hz, n_pointings, phase_gradients);
    });
}

int chip_aw_select_cells(chip_gridder *g, const chip_samples *s, uint32_t *cf_cell_dev, char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        aw_select_cells(*g, *s, cf_cell_dev);
    });
}

int chip_aw_cells(const chip_gridder *g, uint32_t *n_cells, chip_cf_cell *cells, uint64_t *n_pixels,
                  float *pixels_dev, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        if (!g->aw) fail(CHIP_ERR_INVALID, "no AW atlas attached");
        set_device(g->ctx);
        if (n_cells) *n_cells = g->aw_cells;
        if (n_pixels) *n_pixels = g->aw_n_pixels;
        if (cells) std::memcpy(cells, g->aw_meta.data(), g->aw_meta.size() * sizeof(chip_cf_cell));
        if (pixels_dev) {
            CHIP_HIP(hipMemcpyAsync(pixels_dev, g->d_aw_pixels.p, g->d_aw_pixels.bytes, hipMemcpyDeviceToDevice,
                                    g->ctx->stream));
            CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
        }
    });
}

int chip_aw_plan(chip_gridder *g, const chip_samples *s, int32_t *plan_dev, float *norm32_dev,
                 double *norm64_dev, char **err) {
    return guarded(err, [&] {
        if (!g || !s || !plan_dev) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        aw_plan_raw(*g, make_source(*s), plan_dev, norm32_dev, norm64_dev);
        CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int chip_airy_table(double dish_m, double blockage_m, float *values, double *max_radius, char **err) {
    return guarded(err, [&] {
        if (!(std::isfinite(dish_m) && dish_m > 0.0 && std::isfinite(blockage_m) && blockage_m >= 0.0))
            fail(CHIP_ERR_INVALID, "dish and blockage diameters must be finite, dish positive");
        if (values) {
            const std::vector<float> t = airy_table(dish_m, blockage_m);
            std::memcpy(values, t.data(), t.size() * sizeof(float));
        }
        if (max_radius) *max_radius = airy_max_radius(dish_m, blockage_m);
    });
}

// ---- CASA CF caches (cfcache.cpp, aw.hip) ----

int chip_cfcache_open(const char *path, chip_cfcache **out, char **err) {
    return guarded(err, [&] {
        if (!path || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        *out = nullptr;
        auto c = std::make_unique<chip_cfcache>();
        cf_cache_open(c->c, path);
        *out = c.release();
    });
}

void chip_cfcache_close(chip_cfcache *c) { delete c; }

int chip_cfcache_info_get(const chip_cfcache *c, chip_cfcache_info *info, double *freqs, double *w_values,
                          int32_t *muellers, double *pas, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (info) {
            info->paired_cells = x.keys.size();
            info->n_freq = static_cast<uint32_t>(x.freqs.size());
            info->n_w = static_cast<uint32_t>(x.w_values.size());
            info->n_mueller = static_cast<uint32_t>(x.muellers.size());
            info->n_pa = static_cast<uint32_t>(x.pas.size());
            info->w_increment = x.w_increment;
            info->metadata_fingerprint = x.fingerprint;
        }
        if (freqs) std::copy(x.freqs.begin(), x.freqs.end(), freqs);
        if (w_values) std::copy(x.w_values.begin(), x.w_values.end(), w_values);
        if (muellers) std::copy(x.muellers.begin(), x.muellers.end(), muellers);
        if (pas) std::copy(x.pas.begin(), x.pas.end(), pas);
    });
}

int chip_cfcache_keys(const chip_cfcache *c, chip_cf_key *keys, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (!keys) fail(CHIP_ERR_INVALID, "NULL output");
        for (size_t k = 0; k < x.keys.size(); ++k)
            keys[k] = chip_cf_key{x.keys[k].frequency_hz, x.keys[k].w_value_lambda, x.keys[k].mueller_element,
                                  x.keys[k].parallactic_angle_deg};
    });
}

int chip_cfcache_metadata(const chip_cfcache *c, uint64_t index, chip_cf_kernel_meta *imaging,
                          chip_cf_kernel_meta *weight, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (index >= x.keys.size()) fail(CHIP_ERR_INVALID, "CF cell index out of range");
        copy_meta(x.imaging[index], imaging);
        copy_meta(x.weight[index], weight);
    });
}

int chip_cfcache_nearest_key(const chip_cfcache *c, double frequency_hz, double w_value_lambda, int32_t mueller,
                             double parallactic_angle_deg, int64_t *index, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (!index) fail(CHIP_ERR_INVALID, "NULL output");
        *index = cf_nearest_key(x, frequency_hz, w_value_lambda, mueller, parallactic_angle_deg);
    });
}

int chip_cfcache_select_key(const chip_cfcache *c, double sample_frequency_hz, double reference_frequency_hz,
                            double w_lambda, int32_t mueller, double parallactic_angle_deg, int conjugate_beams,
                            int prediction, int64_t *index, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (!index) fail(CHIP_ERR_INVALID, "NULL output");
        *index = cf_select_key(x, sample_frequency_hz, reference_frequency_hz, w_lambda, mueller,
                               parallactic_angle_deg, conjugate_beams != 0, prediction != 0);
    });
}

int chip_cfcache_load(const chip_cfcache *c, uint64_t index, float *imaging, float *weight, char **err) {
    return guarded(err, [&] {
        const CfCache &x = cache_of(c);
        if (index >= x.keys.size()) fail(CHIP_ERR_INVALID, "CF cell index out of range");
        // both cells are read (and checked) before either output is written
        const std::vector<float> im = cf_read_kernel_pixels(x.imaging[index]);
        const std::vector<float> wt = cf_read_kernel_pixels(x.weight[index]);
        if (imaging) std::memcpy(imaging, im.data(), im.size() * sizeof(float));
        if (weight) std::memcpy(weight, wt.data(), wt.size() * sizeof(float));
    });
}

int chip_gridder_attach_aw_cache(chip_gridder *g, const chip_cfcache *c, uint32_t n_pointings,
                                 const double *phase_gradients, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        const CfCache &x = cache_of(c);
        set_device(g->ctx);
        attach_aw_cache(*g, x, n_pointings, phase_gradients);
    });
}

int chip_cfcache_select_cells(chip_gridder *g, const chip_samples *s, const double *chan_freq_hz,
                              double reference_frequency_hz, int32_t mueller, const double *pa_deg_dev,
                              double pa_deg, int conjugate_beams, int prediction, uint32_t *cf_cell_dev,
                              char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        cf_select_cells(*g, *s, chan_freq_hz, reference_frequency_hz, mueller, pa_deg_dev, pa_deg,
                        conjugate_beams, prediction, cf_cell_dev);
    });
}

int chip_wproject_tables(const chip_gridder *g, int32_t *supports, float *kernels, float *norm_table,
                         char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        if (!g->wproject || g->aw) fail(CHIP_ERR_INVALID, "no W projection attached");
        set_device(g->ctx);
        if (supports) std::memcpy(supports, g->supports.data(), g->supports.size() * sizeof(int32_t));
        if (norm_table) std::memcpy(norm_table, g->norm_table.data(), g->norm_table.size() * sizeof(float));
        if (kernels) {
            const uint64_t bytes = static_cast<uint64_t>(g->width) * g->width * g->planes * sizeof(float2);
            DevBuf tmp(bytes);
            wk_all_kernels(*g, tmp.as<float2>());
            CHIP_HIP(hipMemcpy(kernels, tmp.p, bytes, hipMemcpyDeviceToHost));
        }
    });
}

int chip_gridder_attach_wproject_tables(chip_gridder *g, uint32_t planes, uint32_t sampling, double w_scale,
                                        uint32_t width, const int32_t *supports, const float *kernels,
                                        const float *norm_table, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        set_device(g->ctx);
        attach_wproject_tables(*g, planes, sampling, w_scale, width, supports, kernels, norm_table);
    });
}

int chip_grid_wproject_samples(chip_gridder *g, uint64_t n, const void *samples_dev, const float *model_dev,
                               float *psf_dev, float *residual_dev, uint32_t mode, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        set_device(g->ctx);
        grid_wproject_samples(*g, n, samples_dev, model_dev, psf_dev, residual_dev, mode);
    });
}

int chip_buffer_alloc(chip_ctx *ctx, uint64_t bytes, void **dev_out, char **err) {
    return guarded(err, [&] {
        if (!ctx || !dev_out) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        *dev_out = nullptr;
        if (bytes) CHIP_HIP(hipMalloc(dev_out, bytes));
    });
}

void chip_buffer_free(chip_ctx *ctx, void *dev) {
    if (!ctx || !dev) return;
    (void)hipSetDevice(ctx->device);
    (void)hipFree(dev);
}

int chip_buffer_upload(chip_ctx *ctx, void *dev, const void *host, uint64_t bytes, char **err) {
    return guarded(err, [&] {
        if (!ctx || (bytes && (!dev || !host))) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        if (bytes) CHIP_HIP(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
        CHIP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int chip_buffer_download(chip_ctx *ctx, void *host, const void *dev, uint64_t bytes, char **err) {
    return guarded(err, [&] {
        if (!ctx || (bytes && (!dev || !host))) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        if (bytes) CHIP_HIP(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CHIP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int chip_buffer_zero(chip_ctx *ctx, void *dev, uint64_t bytes, char **err) {
    return guarded(err, [&] {
        if (!ctx || (bytes && !dev)) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        if (bytes) CHIP_HIP(hipMemsetAsync(dev, 0, bytes, ctx->stream));
        CHIP_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int chip_gridder_wkernel_budget(chip_gridder *g, uint64_t max_table_bytes, uint32_t *chunk_planes, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        if (g->wk_external) fail(CHIP_ERR_INVALID, "W tables given by the caller stay resident");
        set_device(g->ctx);
        g->wk_budget = max_table_bytes;
        if (g->wproject && !g->aw)
            build_wproject(*g, g->wk.max_abs_w, g->wk.explicit_planes, g->wk.sampling_arg);
        if (chunk_planes) *chunk_planes = (g->wproject && !g->aw) ? g->wk_chunk : 0u;
    });
}

int chip_plan(chip_gridder *g, const chip_samples *s, int32_t *plan_dev, float *norm_dev, char **err) {
    return guarded(err, [&] {
        if (!g || !s || !plan_dev) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        launch_plan_raw(*g, make_source(*s), plan_dev, norm_dev, g->ctx->stream);
        CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int chip_phase_rotation(const double *source_dir, const double *target_dir, double *uvrot, double *phrot,
                        char **err) {
    return guarded(err, [&] {
        if (!source_dir || !target_dir || !uvrot || !phrot) fail(CHIP_ERR_INVALID, "NULL argument");
        for (int k = 0; k < 2; ++k)
            if (!std::isfinite(source_dir[k]) || !std::isfinite(target_dir[k]))
                fail(CHIP_ERR_INVALID, "directions must be finite");
        phase_rotation(source_dir, target_dir, uvrot, phrot);
    });
}

int chip_rotate_rows(chip_ctx *ctx, uint64_t n_rows, uint64_t n_chan, const double *uvrot, const double *phrot,
                     const double *uvw_m_dev, const double *freq_hz_dev, float *vis_dev, double *uvw_out_dev,
                     double *phase_shift_m_dev, char **err) {
    return guarded(err, [&] {
        if (!ctx) fail(CHIP_ERR_INVALID, "NULL context");
        set_device(ctx);
        rotate_rows(*ctx, n_rows, n_chan, uvrot, phrot, uvw_m_dev, freq_hz_dev, vis_dev, uvw_out_dev,
                    phase_shift_m_dev);
    });
}

int chip_ingest_create(chip_ctx *ctx, const chip_ingest_params *p, const double *freq_hz, chip_ingest **out,
                       char **err) {
    return guarded(err, [&] {
        if (!ctx || !p || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        std::unique_ptr<chip_ingest> ing(new chip_ingest);
        ing->ctx = ctx;
        ing->device = ctx->device;
        ingest_create(*ing, *p, freq_hz);
        *out = ing.release();
    });
}

void chip_ingest_destroy(chip_ingest *ing) {
    if (!ing) return;
    (void)hipSetDevice(ing->device);
    delete ing;
}

int chip_ingest_push(chip_ingest *ing, uint64_t n_rows, const double *uvw_m, const float *data,
                     const uint8_t *flag, const float *weight, const uint8_t *row_flag, char **err) {
    return guarded(err, [&] {
        if (!ing) fail(CHIP_ERR_INVALID, "NULL ingest");
        set_device(ing->ctx);
        ingest_push(*ing, n_rows, uvw_m, data, flag, weight, row_flag);
    });
}

int chip_ingest_samples(chip_ingest *ing, chip_samples *out, char **err) {
    return guarded(err, [&] {
        if (!ing || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ing->ctx);
        ingest_samples(*ing, *out);
    });
}

int chip_ingest_download(chip_ingest *ing, double *uvw_m, float *vis, float *weight, float *sumwt_factor,
                         uint8_t *flag, char **err) {
    return guarded(err, [&] {
        if (!ing) fail(CHIP_ERR_INVALID, "NULL ingest");
        set_device(ing->ctx);
        ingest_download(*ing, uvw_m, vis, weight, sumwt_factor, flag);
    });
}

int chip_density_weights(chip_gridder *g, const chip_samples *s, uint32_t mode, float robust,
                         uint32_t per_plane, float *weight_out_dev, float *f2_host, float *density_dev,
                         char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        density_weights(*g, *s, mode, robust, per_plane, weight_out_dev, f2_host, density_dev);
    });
}

int chip_vset_create(chip_gridder *g, const chip_samples *s, chip_vset **out, char **err) {
    return guarded(err, [&] {
        if (!g || !s || !out) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        std::unique_ptr<chip_vset> vs(new chip_vset);
        vs->g = g;
        vs->device = g->device;
        vset_build(*vs, *s);
        CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
        *out = vs.release();
    });
}

void chip_vset_destroy(chip_vset *vs) {
    if (!vs) return;
    (void)hipSetDevice(vs->device);
    delete vs;
}

int chip_vset_sums(const chip_vset *vs, chip_sums *sums, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        if (sums) std::memcpy(sums, vs->sums.data(), vs->sums.size() * sizeof(chip_sums));
    });
}

int chip_vset_stats(chip_vset *vs, uint64_t *n_valid, uint64_t *n_groups, uint64_t *n_items,
                    uint64_t *taps, double *last_kernel_ms, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        if (n_valid) *n_valid = vs->n_valid;
        if (n_groups) *n_groups = vs->n_groups;
        if (n_items) *n_items = vs->n_items;
        if (taps) *taps = vs->taps;
        if (last_kernel_ms) {
            *last_kernel_ms = -1.0;
            if (vs->timed) {
                set_device(vs->g->ctx);
                float ms = 0.0f;
                CHIP_HIP(hipEventSynchronize(vs->ev_stop));
                CHIP_HIP(hipEventElapsedTime(&ms, vs->ev_start, vs->ev_stop));
                *last_kernel_ms = ms;
            }
        }
    });
}

int chip_vset_grid(chip_vset *vs, uint32_t mode, const float *model_dev, float *psf_dev,
                   float *dirty_dev, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        set_device(vs->g->ctx);
        vset_grid(*vs, mode, model_dev, psf_dev, dirty_dev);
        CHIP_HIP(hipStreamSynchronize(vs->g->ctx->stream));
    });
}

int chip_vset_degrid(chip_vset *vs, const float *model_dev, float *vis_out_dev, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        set_device(vs->g->ctx);
        vset_degrid(*vs, model_dev, vis_out_dev);
        CHIP_HIP(hipStreamSynchronize(vs->g->ctx->stream));
    });
}

int chip_vset_grid_values(chip_vset *vs, const float *values_dev, float *grid_dev, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        set_device(vs->g->ctx);
        vset_grid_values(*vs, values_dev, grid_dev);
        CHIP_HIP(hipStreamSynchronize(vs->g->ctx->stream));
    });
}

int chip_vset_major_cycle(chip_vset *vs, const float *model_dev, const double *nsum,
                          const float *psf_peak, float *residual_out_dev, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        set_device(vs->g->ctx);
        vset_major_cycle(*vs, model_dev, nsum, psf_peak, residual_out_dev);
    });
}

int chip_vset_wave_groups(chip_vset *vs, uint32_t groups, uint32_t *in_use, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        if (groups > 2) fail(CHIP_ERR_INVALID, "groups per wavefront must be 0 (automatic), 1 or 2");
        vs->wave_groups = groups;
        if (in_use) *in_use = static_cast<uint32_t>(binned_wave_groups(*vs));
    });
}

int chip_vset_two_phase(chip_vset *vs, int enable, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        vs->wide = enable != 0;
    });
}

int chip_vset_schedule(chip_vset *vs, uint32_t tile, uint32_t order, uint32_t threads, uint32_t groups_per_item,
                       char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        set_device(vs->g->ctx);
        vset_schedule(*vs, tile, order, threads, groups_per_item);
    });
}

int chip_vset_schedule_info(const chip_vset *vs, uint32_t *tile, uint32_t *order, uint32_t *threads,
                            uint32_t *groups_per_item, uint32_t *n_timed, double *timed, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        if (tile) *tile = vs->tile;
        if (order) *order = vs->order;
        if (threads) *threads = vs->threads;
        if (groups_per_item) *groups_per_item = vs->groups_per_item;
        if (n_timed) *n_timed = static_cast<uint32_t>(vs->tune_ms.size() / 5);
        if (timed) std::memcpy(timed, vs->tune_ms.data(), vs->tune_ms.size() * sizeof(double));
    });
}

int chip_vset_item_order(chip_vset *vs, uint32_t order, uint32_t *sibling, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        set_device(vs->g->ctx);
        if (order != UINT32_MAX) vset_item_order(*vs, order);
        if (sibling) *sibling = vs->sibling;
    });
}

int chip_vset_cycle_passes(chip_vset *vs, int enable, double *pass_ms, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        if (enable < -1 || enable > 1) fail(CHIP_ERR_INVALID, "enable must be -1, 0 or 1");
        vset_cycle_passes(*vs, enable, pass_ms);
    });
}

int chip_vset_cycle_graph(chip_vset *vs, int enable, uint64_t *replays, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL visibility set");
        if (enable != -1 && enable != 0 && enable != 1) fail(CHIP_ERR_INVALID, "enable must be -1, 0 or 1");
        set_device(vs->g->ctx);
        if (replays) *replays = vset_cycle_replays(*vs);
        if (enable != -1) vset_cycle_graph(*vs, enable);
    });
}

int chip_vset_cycle_strategy(chip_vset *vs, int strategy, char **err) {
    return guarded(err, [&] {
        if (!vs) fail(CHIP_ERR_INVALID, "NULL vset");
        set_device(vs->g->ctx);
        vset_cycle_strategy(*vs, strategy);
    });
}

int chip_vset_cycle_info(const chip_vset *vs, int *strategy, uint32_t *shape, double *transform_ms,
                         char **err) {
    return guarded(err, [&] {
        if (!vs || !strategy || !shape || !transform_ms) fail(CHIP_ERR_INVALID, "NULL argument");
        vset_cycle_info(*vs, strategy, shape, transform_ms);
    });
}

int chip_grid_plain(chip_gridder *g, const chip_samples *s, uint32_t mode, uint32_t precision,
                    const void *model_dev, void *psf_dev, void *dirty_dev, chip_sums *sums_host,
                    char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        grid_plain(*g, make_source(*s), mode, precision, model_dev, psf_dev, dirty_dev, sums_host);
    });
}

int chip_grid_plain_noalloc(chip_gridder *g, const chip_samples *s, uint32_t mode, uint32_t precision,
                            const void *model_dev, void *psf_dev, void *dirty_dev, chip_sums *sums_host,
                            char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        grid_plain_noalloc(*g, make_source(*s), mode, precision, model_dev, psf_dev, dirty_dev, sums_host);
    });
}

int chip_degrid_plain(chip_gridder *g, const chip_samples *s, uint32_t precision,
                      const void *model_dev, void *vis_out_dev, char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        validate_source(*s);
        set_device(g->ctx);
        degrid_plain(*g, make_source(*s), precision, model_dev, vis_out_dev);
        CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int chip_predict_model(chip_gridder *g, const chip_samples *s, const float *model_grids_dev,
                       uint64_t n_model_grids, float *vis_out_dev, char **err) {
    return guarded(err, [&] {
        if (!g || !s) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        predict_model(*g, *s, model_grids_dev, n_model_grids, vis_out_dev);
        CHIP_HIP(hipStreamSynchronize(g->ctx->stream));
    });
}

int chip_finish_products(chip_gridder *g, uint32_t precision, uint64_t batch, void *psf_grids_dev,
                         void *dirty_grids_dev, const double *nsum, float *psf_out_dev,
                         float *residual_out_dev, float *psf_peak_host, char **err) {
    return guarded(err, [&] {
        if (!g || !nsum || !psf_peak_host) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        finish_products(*g, precision, batch, psf_grids_dev, dirty_grids_dev, nsum, psf_out_dev,
                        residual_out_dev, psf_peak_host);
    });
}

int chip_model_to_grid(chip_gridder *g, uint32_t precision, uint64_t batch, const float *model_dev,
                       void *grid_out_dev, char **err) {
    return guarded(err, [&] {
        if (!g) fail(CHIP_ERR_INVALID, "NULL gridder");
        set_device(g->ctx);
        model_to_grid(*g, precision, batch, model_dev, grid_out_dev);
    });
}

int chip_grid_to_model_adjoint(chip_gridder *g, uint64_t batch, float *grid_dev, float *image_out_dev,
                               char **err) {
    return guarded(err, [&] {
        if (!g || !grid_dev || !image_out_dev) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(g->ctx);
        grid_to_model_adjoint(*g, batch, grid_dev, image_out_dev);
    });
}

int chip_hogbom(chip_ctx *ctx, const chip_hogbom_params *params, uint64_t n_planes, uint64_t nx, uint64_t ny,
                float *residual_dev, uint64_t kx, uint64_t ky, const float *psf_dev, float *model_dev,
                const uint8_t *mask_dev, const float *cycle_threshold, const float *nsigma_threshold,
                uint32_t *updates_host, int32_t *stop_host, char **err) {
    return guarded(err, [&] {
        if (!ctx || !params || !updates_host || !stop_host) fail(CHIP_ERR_INVALID, "NULL argument");
        set_device(ctx);
        hogbom(*ctx, *params, n_planes, nx, ny, residual_dev, kx, ky, psf_dev, model_dev, mask_dev, cycle_threshold,
               nsigma_threshold, updates_host, stop_host);
    });
}

}  // extern "C"
