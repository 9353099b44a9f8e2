/*
 * casa_oracle.c -- CPU oracle for the casa-imaging gridding/degridding path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file restates, in plain C, the reference CPU
 * algorithm of bglenden/casa-rs `crates/casa-imaging` so that the HIP library can
 * be checked against it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * CPU-baseline leg may load it; the product library never links or calls it.
 *
 * Every function names the reference file:line it follows.  Arithmetic is kept
 * in the same precision and the same operation order as the reference
 * (f32 tap products widened to f64, f32 complex products for W-projection,
 * sequential f32 sums for kernel normalisation) and the file is compiled with
 * -ffp-contract=off so that no multiply-add is fused where Rust fuses none.
 *
 * Grid layout follows ndarray's Array2 default: logical (x, y), y fastest,
 * flat index x*ny + y  (gridder.rs:61-66, :711-715).  Complex grids are stored
 * interleaved (re, im).
 */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define CO_SUPPORT 3
#define CO_TAPS 7

/* ------------------------------------------------------------------------- */
/* Closed-form pieces                                                         */
/* ------------------------------------------------------------------------- */

/* Schwab rational approximation of the prolate spheroidal function.
 * gridder.rs:4230-4268.  `powi` is evaluated by repeated squaring, which is
 * how both LLVM's constant-exponent expansion and __powidf2 evaluate it. */
double co_grdsf(double nu) {
    static const double P0[5] = {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1};
    static const double P1[5] = {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2};
    static const double Q0[3] = {1.0, 8.212018e-1, 2.078043e-1};
    static const double Q1[3] = {1.0, 9.599102e-1, 2.918724e-1};
    if (!(nu >= 0.0 && nu <= 1.0)) return 0.0;
    const double *p, *q;
    double nu_end;
    if (nu < 0.75) { p = P0; q = Q0; nu_end = 0.75; }
    else { p = P1; q = Q1; nu_end = 1.0; }
    double d = nu * nu - nu_end * nu_end;
    double pw[5];
    pw[0] = 1.0;
    pw[1] = d;
    pw[2] = d * d;
    pw[3] = d * pw[2];
    pw[4] = pw[2] * pw[2];
    double num = 0.0, den = 0.0;
    for (int k = 0; k < 5; ++k) num = num + p[k] * pw[k];
    for (int k = 0; k < 3; ++k) den = den + q[k] * pw[k];
    if (den == 0.0) return 0.0;
    return num / den;
}

/* gridder.rs:4200-4209 */
float co_spheroidal_kernel(double distance, double support) {
    if (!(isfinite(distance) && distance <= support)) return 0.0f;
    double nu = distance / support;
    if (nu > 1.0) return 0.0f;
    return (float)((1.0 - nu * nu) * co_grdsf(nu));
}

/* gridder.rs:4152-4156 */
uint64_t co_padded_len(uint64_t n, double padding) {
    double f = floor(padding * (double)n - 0.5);
    uint64_t padded = f <= 0.0 ? 0 : (uint64_t)f;
    if (padded < n) padded = n;
    return (padded % 2 == 0) ? padded : padded + 1;
}

/* gridder.rs:4166-4173 */
int co_is_composite(uint64_t value) {
    static const uint64_t factors[3] = {2, 3, 5};
    for (int i = 0; i < 3; ++i)
        while (value > 1 && value % factors[i] == 0) value /= factors[i];
    return value == 1;
}

/* gridder.rs:4158-4164 */
uint64_t co_composite_padded_len(uint64_t n, double padding) {
    uint64_t padded = co_padded_len(n, padding);
    while (!co_is_composite(padded)) padded += 2;
    return padded;
}

/* Grid correction 1/psi, zero where psi <= 1e-6.  gridder.rs:4175-4188 */
void co_correction_axis(uint64_t size, float *out) {
    double center = (double)size / 2.0;
    for (uint64_t i = 0; i < size; ++i) {
        double nu = fabs((double)i - center) / center;
        if (nu < 0.0) nu = 0.0;
        if (nu > 1.0) nu = 1.0;
        double v = co_grdsf(nu);
        out[i] = v > 1.0e-6 ? (float)(1.0 / v) : 0.0f;
    }
}

/* psi itself, used to taper the W screen.  gridder.rs:4190-4198 */
void co_gridder_correction_axis(uint64_t size, float *out) {
    double center = (double)size / 2.0;
    for (uint64_t i = 0; i < size; ++i) {
        double nu = fabs((double)i - center) / center;
        if (nu < 0.0) nu = 0.0;
        if (nu > 1.0) nu = 1.0;
        out[i] = (float)co_grdsf(nu);
    }
}

/* Float sinc of the sampling-cell offset.  gridder.rs:4211-4228 */
void co_sinc_axis(uint64_t size, uint64_t conv_sampling, float *out) {
    int64_t half = (int64_t)(size / 2);
    for (uint64_t i = 0; i < size; ++i) {
        float offset = (float)((int64_t)i - half);
        float denominator = (float)size * (float)conv_sampling;
        float argument = (float)(3.14159265358979323846 * (double)offset / (double)denominator);
        out[i] = (i == size / 2) ? 1.0f : sinf(argument) / argument;
    }
}

/* ------------------------------------------------------------------------- */
/* StandardGridder (prolate spheroidal, 7x7)                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t nx, ny;       /* image shape */
    double cell_x, cell_y; /* radians */
    uint64_t gx, gy;       /* grid shape */
    uint64_t blc_x, blc_y; /* image origin inside the grid */
    uint64_t os;           /* oversampling of the tap table */
    double du, dv;         /* grid spacing in wavelengths */
} co_std;

/* StandardGridder::new_with_padding_factor, gridder.rs:392-443.
 * padding_mode 0: composite padding 1.2 (`new`, :371-373, :385-390)
 * padding_mode 1: composite padding 1.3 (`new_with_casa_composite_padding`, :379-383)
 * padding_mode 2: unpadded (`new_unpadded`, :375-377).
 * The reference fixes os = 100 (:407); `os` is a parameter here so that the
 * oversampling-128 configuration can be restated with the same formulas. */
int co_std_init(co_std *g, uint64_t nx, uint64_t ny, double cell_x, double cell_y,
                int padding_mode, uint64_t os) {
    if (nx < 8 || ny < 8) return 1; /* ImageGeometry::validate, types.rs:910-925 */
    if (!(isfinite(cell_x) && cell_x > 0.0 && isfinite(cell_y) && cell_y > 0.0)) return 1;
    if (os < 2) return 1;
    g->nx = nx; g->ny = ny; g->cell_x = cell_x; g->cell_y = cell_y; g->os = os;
    if (padding_mode == 0) { g->gx = co_composite_padded_len(nx, 1.2); g->gy = co_composite_padded_len(ny, 1.2); }
    else if (padding_mode == 1) { g->gx = co_composite_padded_len(nx, 1.3); g->gy = co_composite_padded_len(ny, 1.3); }
    else { g->gx = nx; g->gy = ny; }
    g->blc_x = (g->gx - nx + (g->gx % 2 == 0 ? 1 : 0)) / 2;
    g->blc_y = (g->gy - ny + (g->gy % 2 == 0 ? 1 : 0)) / 2;
    g->du = 1.0 / ((double)g->gx * cell_x);
    g->dv = 1.0 / ((double)g->gy * cell_y);
    return 0;
}

/* kernel_table[os*(S+1)], first os*S entries filled.  gridder.rs:407-416 */
void co_std_kernel_table(uint64_t os, float *table) {
    uint64_t len = os * (CO_SUPPORT + 1);
    for (uint64_t i = 0; i < len; ++i) table[i] = 0.0f;
    for (uint64_t i = 0; i < os * CO_SUPPORT; ++i) {
        double distance = (double)i / ((double)CO_SUPPORT * (double)os);
        table[i] = co_spheroidal_kernel(distance * (double)CO_SUPPORT, (double)CO_SUPPORT);
    }
}

/* Per-offset 7-tap weights normalised to sum 1.  gridder.rs:2026-2051.
 * Output rows: 2*(os/2)+1, row r <-> offset r - os/2. */
void co_std_tap_weights(uint64_t os, float *out) {
    uint64_t len = os * (CO_SUPPORT + 1);
    float *table = (float *)malloc(len * sizeof(float));
    co_std_kernel_table(os, table);
    int64_t half = (int64_t)os / 2;
    uint64_t row = 0;
    for (int64_t offset = -half; offset <= half; ++offset, ++row) {
        float w[CO_TAPS];
        float norm = 0.0f;
        for (int tap = 0; tap < CO_TAPS; ++tap) {
            int64_t delta = tap - CO_SUPPORT;
            int64_t lookup = delta * (int64_t)os + offset;
            if (lookup < 0) lookup = -lookup;
            float weight = (uint64_t)lookup < len ? table[lookup] : 0.0f;
            w[tap] = weight;
            norm += weight;
        }
        if (norm > 0.0f)
            for (int tap = 0; tap < CO_TAPS; ++tap) w[tap] /= norm;
        memcpy(out + row * CO_TAPS, w, sizeof(w));
    }
    free(table);
}

/* sample_tap_span, gridder.rs:1958-1974 (with normalized_tap_weight_index :1991-1995).
 * Returns 1 and writes (start, weight_index) when the 7 taps fit the axis. */
static int co_tap_span(double coordinate, uint64_t size, uint64_t os, int64_t *start,
                       int64_t *weight_index) {
    if (!isfinite(coordinate)) return 0;
    /* `as isize` saturates; anything this far out is rejected anyway. */
    if (coordinate > 4.0e18 || coordinate < -4.0e18) return 0;
    double anchor_f = round(coordinate);
    int64_t anchor = (int64_t)anchor_f;
    int64_t offset = (int64_t)round(((double)anchor - coordinate) * (double)os);
    int64_t s = anchor - CO_SUPPORT, e = anchor + CO_SUPPORT;
    if (s < 0 || e >= (int64_t)size) return 0;
    int64_t half = (int64_t)os / 2;
    int64_t index = offset + half;
    if (index < 0 || index >= 2 * half + 1) return 0;
    *start = s;
    *weight_index = index;
    return 1;
}

/* grid_coordinate_x/y, gridder.rs:1877-1886 */
static double co_std_coord_x(const co_std *g, double u) { return u / g->du + (double)g->gx / 2.0; }
static double co_std_coord_y(const co_std *g, double v) { return -v / g->dv + (double)g->gy / 2.0; }

/* plan_positive_taps, gridder.rs:557-566.  out = {x_start, y_start, wix, wiy}. */
int co_std_plan(const co_std *g, double u, double v, int64_t *out) {
    int64_t xs, xw, ys, yw;
    if (!co_tap_span(co_std_coord_x(g, u), g->gx, g->os, &xs, &xw)) return 0;
    if (!co_tap_span(co_std_coord_y(g, v), g->gy, g->os, &ys, &yw)) return 0;
    out[0] = xs; out[1] = ys; out[2] = xw; out[3] = yw;
    return 1;
}

void co_std_plan_many(const co_std *g, uint64_t n, const double *u, const double *v, int64_t *out) {
    for (uint64_t i = 0; i < n; ++i) {
        if (!co_std_plan(g, u[i], v[i], out + 4 * i)) {
            out[4 * i] = out[4 * i + 1] = out[4 * i + 2] = out[4 * i + 3] = -1;
        }
    }
}

/* sample_taps_unnormalized, gridder.rs:1997-2023: indices and raw spheroid
 * weights of the 7 taps of one axis, or 0 when the coordinate is not finite,
 * a tap leaves [0, size) or the weights sum to zero. */
static int co_taps_unnormalized(double coordinate, uint64_t size, uint64_t os, const float *table,
                                uint64_t len, int64_t *indices, float *weights) {
    if (!isfinite(coordinate)) return 0;
    if (coordinate > 4.0e18 || coordinate < -4.0e18) return 0;
    int64_t anchor = (int64_t)round(coordinate);
    int64_t offset = (int64_t)round(((double)anchor - coordinate) * (double)os);
    float norm = 0.0f;
    for (int tap = 0; tap < CO_TAPS; ++tap) {
        int64_t index = anchor - CO_SUPPORT + tap;
        if (index < 0 || index >= (int64_t)size) return 0;
        int64_t delta = index - anchor;
        int64_t lookup = delta * (int64_t)os + offset;
        if (lookup < 0) lookup = -lookup;
        float weight = (uint64_t)lookup < len ? table[lookup] : 0.0f;
        indices[tap] = index;
        weights[tap] = weight;
        norm += weight;
    }
    return norm > 0.0f;
}

/* StandardMfsModelPredictor::predict (lib.rs:3387-3403) over a model grid
 * (complex64, x-major): degrid_sample_product_planned_sectdgrid
 * (gridder.rs:1227-1266) -- sum of grid * (wx * wy) and of wx * wy, x taps
 * outer, y taps inner, then value / norm -- or, when that is None, the
 * normalised product taps of plan_sample, or 0 when the sample has no plan.
 * With the spheroid table every in-grid tap set has a positive sum, so the
 * second branch returns the first branch's rejections as 0 (plan_sample
 * rejects the same out-of-grid and non-finite coordinates). */
void co_std_predict_many(const co_std *g, uint64_t n, const double *u, const double *v,
                         const float *grid, float *out) {
    uint64_t len = g->os * (CO_SUPPORT + 1);
    float *table = (float *)malloc(len * sizeof(float));
    co_std_kernel_table(g->os, table);
    for (uint64_t i = 0; i < n; ++i) {
        int64_t xi[CO_TAPS], yi[CO_TAPS];
        float xw[CO_TAPS], yw[CO_TAPS];
        float re = 0.0f, im = 0.0f;
        if (co_taps_unnormalized(co_std_coord_x(g, u[i]), g->gx, g->os, table, len, xi, xw) &&
            co_taps_unnormalized(co_std_coord_y(g, v[i]), g->gy, g->os, table, len, yi, yw)) {
            float vr = 0.0f, vi = 0.0f, norm = 0.0f;
            for (int a = 0; a < CO_TAPS; ++a) {
                for (int b = 0; b < CO_TAPS; ++b) {
                    float weight = xw[a] * yw[b];
                    uint64_t idx = (uint64_t)xi[a] * g->gy + (uint64_t)yi[b];
                    vr += grid[2 * idx] * weight;
                    vi += grid[2 * idx + 1] * weight;
                    norm += weight;
                }
            }
            if (norm > 0.0f && isfinite(norm)) {
                re = vr / norm;
                im = vi / norm;
            }
        }
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
    free(table);
}

typedef struct {
    double normalization_sumwt;
    double reported_sumwt;
    uint64_t gridded;
    uint64_t skipped;
} co_sums;

/* ---- threaded private-grid accumulation with ordered merge ----
 * The reference splits the sample range into `threads` contiguous chunks, each
 * worker accumulates into private grids, and the partial grids are added to
 * the output in worker order (lib.rs:71204-71286, merge :71133-71139;
 * standard streaming lib.rs:68306-68428). */

typedef struct {
    const co_std *g;
    const float *taps;
    uint64_t begin, end;
    const double *u, *v;
    const float *weight, *sumwt_factor, *vis;
    const uint8_t *gridable;
    const float *model; /* complex f32 model grid (residual mode) */
    int mode;           /* 0 psf, 1 dirty (psf + residual), 2 residual (degrid+grid) */
    double *psf, *dirty;
    co_sums sums;
} co_std_job;

static inline int co_finite_vis(const float *vis) { return isfinite(vis[0]) && isfinite(vis[1]); }

/* accumulate_batches, execution.rs:18255-18338 (mode 0/1), and the residual
 * refresh per-sample path, execution.rs:8890-8940 with
 * degrid_model_and_grid_residual_taps_planned_f64, gridder.rs:1315-1396 (mode 2). */
static void *co_std_worker(void *arg) {
    co_std_job *j = (co_std_job *)arg;
    const co_std *g = j->g;
    const uint64_t stride = g->gy;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        if (j->gridable && !j->gridable[i]) { j->sums.skipped++; continue; }
        float weight = j->weight[i];
        float sf = j->sumwt_factor[i];
        const float *vis = j->vis + 2 * i;
        if (j->mode == 2) {
            if (!(isfinite(weight) && weight > 0.0f)) { j->sums.skipped++; continue; }
            if (!co_finite_vis(vis)) { j->sums.skipped++; continue; }
        } else if (!(isfinite(weight) && weight > 0.0f && isfinite(sf) && sf > 0.0f)) {
            j->sums.skipped++;
            continue;
        }
        int64_t plan[4];
        if (!co_std_plan(g, j->u[i], j->v[i], plan)) { j->sums.skipped++; continue; }
        float grid_weight = weight * sf;
        if (j->mode == 2 && !(isfinite(grid_weight) && grid_weight > 0.0f)) { j->sums.skipped++; continue; }
        double sumwt = (double)grid_weight;
        const float *wx = j->taps + plan[2] * CO_TAPS;
        const float *wy = j->taps + plan[3] * CO_TAPS;
        j->sums.gridded++;
        if (j->mode != 2) {
            j->sums.normalization_sumwt += sumwt;
            j->sums.reported_sumwt += sumwt;
            int vis_ok = co_finite_vis(vis);
            double rr = (double)vis[0] * sumwt, ri = (double)vis[1] * sumwt;
            for (int tx = 0; tx < CO_TAPS; ++tx) {
                uint64_t xi = (uint64_t)(plan[0] + tx) * stride;
                float xw = wx[tx];
                for (int ty = 0; ty < CO_TAPS; ++ty) {
                    double w = (double)(xw * wy[ty]);
                    uint64_t idx = xi + (uint64_t)plan[1] + ty;
                    j->psf[2 * idx] += sumwt * w;
                    if (j->mode == 1 && vis_ok) {
                        j->dirty[2 * idx] += rr * w;
                        j->dirty[2 * idx + 1] += ri * w;
                    }
                }
            }
        } else {
            float pr = 0.0f, pi = 0.0f;
            if (j->model) {
                for (int tx = 0; tx < CO_TAPS; ++tx) {
                    uint64_t xi = (uint64_t)(plan[0] + tx) * stride;
                    float xw = wx[tx];
                    for (int ty = 0; ty < CO_TAPS; ++ty) {
                        uint64_t idx = xi + (uint64_t)plan[1] + ty;
                        float w = xw * wy[ty];
                        pr += j->model[2 * idx] * w;
                        pi += j->model[2 * idx + 1] * w;
                    }
                }
            }
            float rvr = vis[0] - pr, rvi = vis[1] - pi;
            double rr = (double)rvr * sumwt, ri = (double)rvi * sumwt;
            for (int tx = 0; tx < CO_TAPS; ++tx) {
                uint64_t xi = (uint64_t)(plan[0] + tx) * stride;
                float xw = wx[tx];
                for (int ty = 0; ty < CO_TAPS; ++ty) {
                    double w = (double)(xw * wy[ty]);
                    uint64_t idx = xi + (uint64_t)plan[1] + ty;
                    j->dirty[2 * idx] += rr * w;
                    j->dirty[2 * idx + 1] += ri * w;
                }
            }
        }
    }
    return NULL;
}

static void co_add_f64(double *dst, const double *src, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) dst[i] += src[i];
}

/* Accumulates into psf/dirty (f64 complex interleaved, gx*gy) in place.
 * mode 0: PSF only, 1: PSF + dirty, 2: residual (model may be NULL). */
void co_std_grid(const co_std *g, const float *taps, uint64_t n, const double *u, const double *v,
                 const float *weight, const float *sumwt_factor, const uint8_t *gridable,
                 const float *vis, const float *model, int mode, int threads, double *psf,
                 double *dirty, co_sums *sums) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n) threads = n > 0 ? (int)n : 1;
    uint64_t cells = g->gx * g->gy * 2;
    uint64_t chunk = (n + threads - 1) / threads;
    co_std_job *jobs = (co_std_job *)calloc(threads, sizeof(co_std_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        co_std_job *j = &jobs[t];
        j->g = g; j->taps = taps; j->u = u; j->v = v; j->weight = weight;
        j->sumwt_factor = sumwt_factor; j->vis = vis; j->gridable = gridable; j->model = model;
        j->mode = mode;
        j->begin = (uint64_t)t * chunk;
        j->end = j->begin + chunk < n ? j->begin + chunk : n;
        if (j->begin > n) j->begin = n;
        if (threads == 1) { j->psf = psf; j->dirty = dirty; }
        else {
            j->psf = mode != 2 ? (double *)calloc(cells, sizeof(double)) : NULL;
            j->dirty = mode != 0 ? (double *)calloc(cells, sizeof(double)) : NULL;
        }
        pthread_create(&tids[t], NULL, co_std_worker, j);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    for (int t = 0; t < threads; ++t) {
        co_std_job *j = &jobs[t];
        sums->normalization_sumwt += j->sums.normalization_sumwt;
        sums->reported_sumwt += j->sums.reported_sumwt;
        sums->gridded += j->sums.gridded;
        sums->skipped += j->sums.skipped;
        if (threads > 1) {
            if (j->psf) { co_add_f64(psf, j->psf, cells); free(j->psf); }
            if (j->dirty) { co_add_f64(dirty, j->dirty, cells); free(j->dirty); }
        }
    }
    free(jobs);
    free(tids);
}

/* degrid_sample_taps_planned_normalized, gridder.rs:1278-1312.
 * Rejected samples produce NaN.  model: complex f32 grid. */
void co_std_degrid(const co_std *g, const float *taps, uint64_t n, const double *u, const double *v,
                   const float *model, float *out) {
    for (uint64_t i = 0; i < n; ++i) {
        int64_t plan[4];
        if (!co_std_plan(g, u[i], v[i], plan)) { out[2 * i] = out[2 * i + 1] = NAN; continue; }
        const float *wx = taps + plan[2] * CO_TAPS;
        const float *wy = taps + plan[3] * CO_TAPS;
        float re = 0.0f, im = 0.0f;
        for (int tx = 0; tx < CO_TAPS; ++tx) {
            uint64_t xi = (uint64_t)(plan[0] + tx) * g->gy;
            float xw = wx[tx];
            for (int ty = 0; ty < CO_TAPS; ++ty) {
                float w = xw * wy[ty];
                uint64_t idx = xi + (uint64_t)plan[1] + ty;
                /* Complex32 * f32 then += : (re*w, im*w) */
                float tr = model[2 * idx] * w, ti = model[2 * idx + 1] * w;
                re += tr;
                im += ti;
            }
        }
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

/* ------------------------------------------------------------------------- */
/* WProjector                                                                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t gx, gy;
    double du, dv;
    uint64_t sampling;
    double w_scale;
    uint64_t planes;
    uint64_t width;          /* common kernel width (quarter plane) */
    const int32_t *support;  /* per plane */
    const float *kernels;    /* planes * width * width complex f32, flat x*width + y */
    int64_t norm_radius;     /* (sampling + 1) / 2 */
    const float *norm_table; /* planes * (2R+1)^2 */
} co_wproj;

/* Only the pieces of WProjector::new that are sums; the screen FFT is done by
 * the Python side of the oracle in f64 (gridder.rs:3438-3502).
 * w_project_plane_sum, gridder.rs:3991-4002 (sequential f32 sum of Re at
 * integer taps, iy outer, ix inner). */
float co_w_plane_sum(const float *kernel, uint64_t width, int64_t support, uint64_t sampling) {
    float sum = 0.0f;
    for (int64_t iy = -support; iy <= support; ++iy) {
        int64_t ky = iy * (int64_t)sampling; if (ky < 0) ky = -ky;
        for (int64_t ix = -support; ix <= support; ++ix) {
            int64_t kx = ix * (int64_t)sampling; if (kx < 0) kx = -kx;
            sum += kernel[2 * ((uint64_t)kx * width + (uint64_t)ky)];
        }
    }
    return sum;
}

/* compute_w_project_sample_normalization, gridder.rs:3782-3799.
 * Returns 0 and sets *ok = 0 when a tap falls outside the stored quarter plane. */
float co_w_sample_normalization(const float *kernel, uint64_t width, int64_t support,
                                uint64_t sampling, int64_t off_x, int64_t off_y, int *ok) {
    float sum = 0.0f;
    *ok = 1;
    for (int64_t iy = -support; iy <= support; ++iy) {
        int64_t ky = iy * (int64_t)sampling + off_y; if (ky < 0) ky = -ky;
        for (int64_t ix = -support; ix <= support; ++ix) {
            int64_t kx = ix * (int64_t)sampling + off_x; if (kx < 0) kx = -kx;
            if ((uint64_t)kx >= width || (uint64_t)ky >= width) { *ok = 0; return 0.0f; }
            sum += kernel[2 * ((uint64_t)kx * width + (uint64_t)ky)];
        }
    }
    return sum;
}

/* w_plane_index, gridder.rs:3749-3757 */
static uint64_t co_w_plane_index(const co_wproj *p, double w) {
    if (p->planes <= 1) return 0;
    double c = p->w_scale * fabs(w);
    if (c < 0.0) c = 0.0;
    c = sqrt(c);
    double r = round(c);
    if (r < 0.0) r = 0.0;
    if (r > (double)(p->planes - 1)) r = (double)(p->planes - 1);
    return (uint64_t)r;
}

/* WProjector::plan_sample, gridder.rs:3588-3624 (sample_normalization :3759-3779).
 * out = {loc_x, loc_y, off_x, off_y, plane, conj}; *norm = normalization. */
int co_w_plan(const co_wproj *p, double u, double v, double w, int64_t *out, float *norm) {
    double pos_x = u / p->du + (double)p->gx / 2.0;
    double pos_y = -v / p->dv + (double)p->gy / 2.0;
    if (!(isfinite(pos_x) && isfinite(pos_y) && isfinite(w))) return 0;
    if (pos_x > 4.0e18 || pos_x < -4.0e18 || pos_y > 4.0e18 || pos_y < -4.0e18) return 0;
    int64_t loc_x = (int64_t)round(pos_x), loc_y = (int64_t)round(pos_y);
    int64_t off_x = (int64_t)round(((double)loc_x - pos_x) * (double)p->sampling);
    int64_t off_y = (int64_t)round(((double)loc_y - pos_y) * (double)p->sampling);
    uint64_t plane = co_w_plane_index(p, w);
    int64_t support = p->support[plane];
    if (support <= 0 || loc_x - support < 0 || loc_y - support < 0 ||
        loc_x + support >= (int64_t)p->gx || loc_y + support >= (int64_t)p->gy)
        return 0;
    int64_t R = p->norm_radius;
    float nrm;
    if (off_x <= R && off_x >= -R && off_y <= R && off_y >= -R) {
        uint64_t axis = (uint64_t)(2 * R + 1);
        nrm = p->norm_table[plane * axis * axis + (uint64_t)(off_x + R) * axis + (uint64_t)(off_y + R)];
    } else {
        int ok;
        nrm = co_w_sample_normalization(p->kernels + 2 * plane * p->width * p->width, p->width,
                                        support, p->sampling, off_x, off_y, &ok);
        if (!ok) return 0;
    }
    out[0] = loc_x; out[1] = loc_y; out[2] = off_x; out[3] = off_y;
    out[4] = (int64_t)plane; out[5] = w > 0.0 ? 1 : 0;
    *norm = nrm;
    return 1;
}

void co_w_plan_many(const co_wproj *p, uint64_t n, const double *u, const double *v,
                    const double *w, int64_t *out, float *norm) {
    for (uint64_t i = 0; i < n; ++i) {
        if (!co_w_plan(p, u[i], v[i], w[i], out + 6 * i, norm + i)) {
            for (int k = 0; k < 6; ++k) out[6 * i + k] = -1;
            norm[i] = NAN;
        }
    }
}

/* raw_w_project_sample_is_valid, lib.rs:73666-73676 */
static int co_w_sample_valid(float weight, float sf, const float *vis, double u, double v, double w) {
    return isfinite(weight) && weight > 0.0f && isfinite(sf) && sf > 0.0f && isfinite(vis[0]) &&
           isfinite(vis[1]) && isfinite(u) && isfinite(v) && isfinite(w);
}

typedef struct {
    const co_wproj *p;
    uint64_t begin, end;
    const double *u, *v, *w;
    const float *weight, *sumwt_factor, *vis;
    const uint8_t *gridable;
    const float *model;
    int mode; /* 0 psf, 1 dirty (psf+residual), 2 residual */
    float *psf, *dirty;
    co_sums sums;
} co_w_job;

/* WProjector::grid_sample_planned / _pair / degrid_sample_planned,
 * gridder.rs:3680-3747, driven per sample as in
 * accumulate_w_project_{psf,dirty,residual}_streaming_sample, lib.rs:71472-71540. */
static void *co_w_worker(void *arg) {
    co_w_job *j = (co_w_job *)arg;
    const co_wproj *p = j->p;
    const uint64_t W = p->width, stride = p->gy;
    const int64_t s = (int64_t)p->sampling;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        const float *vis = j->vis + 2 * i;
        if ((j->gridable && !j->gridable[i]) ||
            !co_w_sample_valid(j->weight[i], j->sumwt_factor[i], vis, j->u[i], j->v[i], j->w[i])) {
            j->sums.skipped++;
            continue;
        }
        int64_t plan[6];
        float nrm;
        if (!co_w_plan(p, j->u[i], j->v[i], j->w[i], plan, &nrm)) { j->sums.skipped++; continue; }
        float weight = j->weight[i];
        /* w_project_accumulation_for_sample, lib.rs:71460-71470 */
        j->sums.normalization_sumwt += 2.0 * (double)weight * (double)nrm;
        j->sums.reported_sumwt += (double)weight * (double)j->sumwt_factor[i];
        j->sums.gridded++;
        const int64_t S = p->support[plan[4]];
        const float *K = p->kernels + 2 * (uint64_t)plan[4] * W * W;
        const int conj = (int)plan[5];
        float val_r, val_i;
        if (j->mode == 2) {
            float pr = 0.0f, pi = 0.0f;
            if (j->model) {
                for (int64_t iy = -S; iy <= S; ++iy) {
                    int64_t ky = iy * s + plan[3]; if (ky < 0) ky = -ky;
                    for (int64_t ix = -S; ix <= S; ++ix) {
                        int64_t kx = ix * s + plan[2]; if (kx < 0) kx = -kx;
                        float cr = K[2 * (kx * W + ky)], ci = K[2 * (kx * W + ky) + 1];
                        if (conj) ci = -ci;
                        /* conj(cwt) * grid */
                        ci = -ci;
                        uint64_t idx = (uint64_t)(plan[0] + ix) * stride + (uint64_t)(plan[1] + iy);
                        float gr = j->model[2 * idx], gi = j->model[2 * idx + 1];
                        float tr = cr * gr - ci * gi;
                        float ti = cr * gi + ci * gr;
                        pr += tr;
                        pi += ti;
                    }
                }
            }
            /* (visibility - predicted) * weight, lib.rs:71536 */
            float dr = vis[0] - pr, di = vis[1] - pi;
            val_r = dr * weight;
            val_i = di * weight;
        } else {
            val_r = vis[0] * weight;
            val_i = vis[1] * weight;
        }
        for (int64_t iy = -S; iy <= S; ++iy) {
            int64_t ky = iy * s + plan[3]; if (ky < 0) ky = -ky;
            for (int64_t ix = -S; ix <= S; ++ix) {
                int64_t kx = ix * s + plan[2]; if (kx < 0) kx = -kx;
                float cr = K[2 * (kx * W + ky)], ci = K[2 * (kx * W + ky) + 1];
                if (conj) ci = -ci;
                uint64_t idx = (uint64_t)(plan[0] + ix) * stride + (uint64_t)(plan[1] + iy);
                if (j->mode != 2) {
                    /* PSF value = weight + 0i: (w*cr - 0*ci, w*ci + 0*cr) */
                    float pr_ = weight * cr - 0.0f * ci;
                    float pi_ = weight * ci + 0.0f * cr;
                    j->psf[2 * idx] += pr_;
                    j->psf[2 * idx + 1] += pi_;
                }
                if (j->mode != 0) {
                    float tr = val_r * cr - val_i * ci;
                    float ti = val_r * ci + val_i * cr;
                    j->dirty[2 * idx] += tr;
                    j->dirty[2 * idx + 1] += ti;
                }
            }
        }
    }
    return NULL;
}

static void co_add_f32(float *dst, const float *src, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i) dst[i] += src[i];
}

/* Threaded as accumulate_w_project_streaming_parallel, lib.rs:71204-71286,
 * with Complex32 private grids merged in worker order (add_complex32_grid,
 * lib.rs:69280). */
void co_w_grid(const co_wproj *p, uint64_t n, const double *u, const double *v, const double *w,
               const float *weight, const float *sumwt_factor, const uint8_t *gridable,
               const float *vis, const float *model, int mode, int threads, float *psf,
               float *dirty, co_sums *sums) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n) threads = n > 0 ? (int)n : 1;
    uint64_t cells = p->gx * p->gy * 2;
    uint64_t chunk = (n + threads - 1) / threads;
    co_w_job *jobs = (co_w_job *)calloc(threads, sizeof(co_w_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        co_w_job *j = &jobs[t];
        j->p = p; j->u = u; j->v = v; j->w = w; j->weight = weight; j->sumwt_factor = sumwt_factor;
        j->vis = vis; j->gridable = gridable; j->model = model; j->mode = mode;
        j->begin = (uint64_t)t * chunk;
        j->end = j->begin + chunk < n ? j->begin + chunk : n;
        if (j->begin > n) j->begin = n;
        if (threads == 1) { j->psf = psf; j->dirty = dirty; }
        else {
            j->psf = mode != 2 ? (float *)calloc(cells, sizeof(float)) : NULL;
            j->dirty = mode != 0 ? (float *)calloc(cells, sizeof(float)) : NULL;
        }
        pthread_create(&tids[t], NULL, co_w_worker, j);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    for (int t = 0; t < threads; ++t) {
        co_w_job *j = &jobs[t];
        sums->normalization_sumwt += j->sums.normalization_sumwt;
        sums->reported_sumwt += j->sums.reported_sumwt;
        sums->gridded += j->sums.gridded;
        sums->skipped += j->sums.skipped;
        if (threads > 1) {
            if (j->psf) { co_add_f32(psf, j->psf, cells); free(j->psf); }
            if (j->dirty) { co_add_f32(dirty, j->dirty, cells); free(j->dirty); }
        }
    }
    free(jobs);
    free(tids);
}

/* The W path of co_w_grid accumulated in FP64 throughout: Complex64 grids, the
 * (Complex32) kernel taps widened to f64, V * w, the degridded model and
 * (V - p) * w in f64, with an f64 model grid.  Not a restatement of the
 * reference's Complex32 W grids (lib.rs:71472-71540): this is the FP64
 * reference that FP32 results are measured against; plans, validity filters
 * and sums are co_w_grid's. */
typedef struct {
    const co_wproj *p;
    uint64_t begin, end;
    const double *u, *v, *w;
    const float *weight, *sumwt_factor, *vis;
    const uint8_t *gridable;
    const double *model;
    int mode;
    double *psf, *dirty;
    co_sums sums;
} co_w_job64;

static void *co_w_worker64(void *arg) {
    co_w_job64 *j = (co_w_job64 *)arg;
    const co_wproj *p = j->p;
    const uint64_t W = p->width, stride = p->gy;
    const int64_t s = (int64_t)p->sampling;
    for (uint64_t i = j->begin; i < j->end; ++i) {
        const float *vis = j->vis + 2 * i;
        if ((j->gridable && !j->gridable[i]) ||
            !co_w_sample_valid(j->weight[i], j->sumwt_factor[i], vis, j->u[i], j->v[i], j->w[i])) {
            j->sums.skipped++;
            continue;
        }
        int64_t plan[6];
        float nrm;
        if (!co_w_plan(p, j->u[i], j->v[i], j->w[i], plan, &nrm)) { j->sums.skipped++; continue; }
        const double weight = (double)j->weight[i];
        j->sums.normalization_sumwt += 2.0 * weight * (double)nrm;
        j->sums.reported_sumwt += weight * (double)j->sumwt_factor[i];
        j->sums.gridded++;
        const int64_t S = p->support[plan[4]];
        const float *K = p->kernels + 2 * (uint64_t)plan[4] * W * W;
        const double csign = plan[5] ? -1.0 : 1.0;
        double val_r, val_i;
        if (j->mode == 2) {
            double pr = 0.0, pi = 0.0;
            if (j->model) {
                for (int64_t iy = -S; iy <= S; ++iy) {
                    int64_t ky = iy * s + plan[3]; if (ky < 0) ky = -ky;
                    for (int64_t ix = -S; ix <= S; ++ix) {
                        int64_t kx = ix * s + plan[2]; if (kx < 0) kx = -kx;
                        double cr = (double)K[2 * (kx * W + ky)];
                        double ci = -csign * (double)K[2 * (kx * W + ky) + 1]; /* conj(cwt) */
                        uint64_t idx = (uint64_t)(plan[0] + ix) * stride + (uint64_t)(plan[1] + iy);
                        double gr = j->model[2 * idx], gi = j->model[2 * idx + 1];
                        pr += cr * gr - ci * gi;
                        pi += cr * gi + ci * gr;
                    }
                }
            }
            val_r = ((double)vis[0] - pr) * weight;
            val_i = ((double)vis[1] - pi) * weight;
        } else {
            val_r = (double)vis[0] * weight;
            val_i = (double)vis[1] * weight;
        }
        for (int64_t iy = -S; iy <= S; ++iy) {
            int64_t ky = iy * s + plan[3]; if (ky < 0) ky = -ky;
            for (int64_t ix = -S; ix <= S; ++ix) {
                int64_t kx = ix * s + plan[2]; if (kx < 0) kx = -kx;
                double cr = (double)K[2 * (kx * W + ky)];
                double ci = csign * (double)K[2 * (kx * W + ky) + 1];
                uint64_t idx = (uint64_t)(plan[0] + ix) * stride + (uint64_t)(plan[1] + iy);
                if (j->mode != 2) {
                    j->psf[2 * idx] += weight * cr;
                    j->psf[2 * idx + 1] += weight * ci;
                }
                if (j->mode != 0) {
                    j->dirty[2 * idx] += val_r * cr - val_i * ci;
                    j->dirty[2 * idx + 1] += val_r * ci + val_i * cr;
                }
            }
        }
    }
    return NULL;
}

void co_w_grid64(const co_wproj *p, uint64_t n, const double *u, const double *v, const double *w,
                 const float *weight, const float *sumwt_factor, const uint8_t *gridable,
                 const float *vis, const double *model, int mode, int threads, double *psf,
                 double *dirty, co_sums *sums) {
    if (threads < 1) threads = 1;
    if ((uint64_t)threads > n) threads = n > 0 ? (int)n : 1;
    uint64_t cells = p->gx * p->gy * 2;
    uint64_t chunk = (n + threads - 1) / threads;
    co_w_job64 *jobs = (co_w_job64 *)calloc(threads, sizeof(co_w_job64));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        co_w_job64 *j = &jobs[t];
        j->p = p; j->u = u; j->v = v; j->w = w; j->weight = weight; j->sumwt_factor = sumwt_factor;
        j->vis = vis; j->gridable = gridable; j->model = model; j->mode = mode;
        j->begin = (uint64_t)t * chunk;
        j->end = j->begin + chunk < n ? j->begin + chunk : n;
        if (j->begin > n) j->begin = n;
        if (threads == 1) { j->psf = psf; j->dirty = dirty; }
        else {
            j->psf = mode != 2 ? (double *)calloc(cells, sizeof(double)) : NULL;
            j->dirty = mode != 0 ? (double *)calloc(cells, sizeof(double)) : NULL;
        }
        pthread_create(&tids[t], NULL, co_w_worker64, j);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    for (int t = 0; t < threads; ++t) {
        co_w_job64 *j = &jobs[t];
        sums->normalization_sumwt += j->sums.normalization_sumwt;
        sums->reported_sumwt += j->sums.reported_sumwt;
        sums->gridded += j->sums.gridded;
        sums->skipped += j->sums.skipped;
        if (threads > 1) {
            if (j->psf) { co_add_f64(psf, j->psf, cells); free(j->psf); }
            if (j->dirty) { co_add_f64(dirty, j->dirty, cells); free(j->dirty); }
        }
    }
    free(jobs);
    free(tids);
}

/* WProjector::degrid_sample_planned, gridder.rs:3726-3747.  NaN when rejected. */
void co_w_degrid(const co_wproj *p, uint64_t n, const double *u, const double *v, const double *w,
                 const float *model, float *out) {
    const uint64_t W = p->width, stride = p->gy;
    const int64_t s = (int64_t)p->sampling;
    for (uint64_t i = 0; i < n; ++i) {
        int64_t plan[6];
        float nrm;
        if (!co_w_plan(p, u[i], v[i], w[i], plan, &nrm)) { out[2 * i] = out[2 * i + 1] = NAN; continue; }
        const int64_t S = p->support[plan[4]];
        const float *K = p->kernels + 2 * (uint64_t)plan[4] * W * W;
        float pr = 0.0f, pi = 0.0f;
        for (int64_t iy = -S; iy <= S; ++iy) {
            int64_t ky = iy * s + plan[3]; if (ky < 0) ky = -ky;
            for (int64_t ix = -S; ix <= S; ++ix) {
                int64_t kx = ix * s + plan[2]; if (kx < 0) kx = -kx;
                float cr = K[2 * (kx * W + ky)], ci = K[2 * (kx * W + ky) + 1];
                if (plan[5]) ci = -ci;
                ci = -ci;
                uint64_t idx = (uint64_t)(plan[0] + ix) * stride + (uint64_t)(plan[1] + iy);
                float gr = model[2 * idx], gi = model[2 * idx + 1];
                float tr = cr * gr - ci * gi;
                float ti = cr * gi + ci * gr;
                pr += tr;
                pi += ti;
            }
        }
        out[2 * i] = pr;
        out[2 * i + 1] = pi;
    }
}

/* ------------------------------------------------------------------------- */
/* Direct DFT (FP64 phase), lib.rs:73709-73910                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    uint64_t nx, ny;
    double cell_x, cell_y;
    uint64_t x0, y0, wy;  /* window origin and height: pixel p is (x0 + p / wy, y0 + p % wy) */
    uint64_t begin, end; /* pixel range of the window */
    uint64_t n;
    const double *u, *v, *w;
    const float *value; /* complex per sample */
    float *image;
} co_dft_job;

/* build_direct_pixel_coordinates, lib.rs:73823-73845 */
static void co_pixel_lmn(uint64_t nx, uint64_t ny, double cx, double cy, uint64_t x, uint64_t y,
                         double *l, double *m, double *nm1) {
    double center_x = (double)nx / 2.0, center_y = (double)ny / 2.0;
    *l = ((double)x - center_x) * cx;
    *m = (center_y - (double)y) * cy;
    double r2 = (*l) * (*l) + (*m) * (*m);
    *nm1 = r2 < 1.0 ? sqrt(1.0 - r2) - 1.0 : -1.0;
}

/* accumulate_direct_adjoint, lib.rs:73890-73910; pixels are split over
 * threads (each pixel still sums samples in order). */
static void *co_dft_adjoint_worker(void *arg) {
    co_dft_job *j = (co_dft_job *)arg;
    const double TAU = 6.28318530717958647692;
    for (uint64_t p = j->begin; p < j->end; ++p) {
        uint64_t x = j->x0 + p / j->wy, y = j->y0 + p % j->wy;
        double l, m, nm1;
        co_pixel_lmn(j->nx, j->ny, j->cell_x, j->cell_y, x, y, &l, &m, &nm1);
        float acc = j->image[p];
        for (uint64_t i = 0; i < j->n; ++i) {
            double phase = TAU * (j->u[i] * l + j->v[i] * m + j->w[i] * nm1);
            float c = (float)cos(phase), s = (float)sin(phase);
            float contribution = 2.0f * (j->value[2 * i] * c - j->value[2 * i + 1] * s);
            acc += contribution;
        }
        j->image[p] = acc;
    }
    return NULL;
}

/* The pixels of the window [x0, x0 + wx) x [y0, y0 + wy) of an nx x ny image;
 * image holds wx * wy values, y fastest. */
void co_direct_adjoint_window(uint64_t nx, uint64_t ny, double cell_x, double cell_y, uint64_t x0,
                              uint64_t y0, uint64_t wx, uint64_t wy, uint64_t n, const double *u,
                              const double *v, const double *w, const float *value, int threads,
                              float *image) {
    if (threads < 1) threads = 1;
    uint64_t pixels = wx * wy;
    uint64_t chunk = (pixels + threads - 1) / threads;
    co_dft_job *jobs = (co_dft_job *)calloc(threads, sizeof(co_dft_job));
    pthread_t *tids = (pthread_t *)calloc(threads, sizeof(pthread_t));
    for (int t = 0; t < threads; ++t) {
        co_dft_job *j = &jobs[t];
        j->nx = nx; j->ny = ny; j->cell_x = cell_x; j->cell_y = cell_y;
        j->x0 = x0; j->y0 = y0; j->wy = wy;
        j->begin = (uint64_t)t * chunk;
        j->end = j->begin + chunk < pixels ? j->begin + chunk : pixels;
        if (j->begin > pixels) j->begin = pixels;
        j->n = n; j->u = u; j->v = v; j->w = w; j->value = value; j->image = image;
        pthread_create(&tids[t], NULL, co_dft_adjoint_worker, j);
    }
    for (int t = 0; t < threads; ++t) pthread_join(tids[t], NULL);
    free(jobs);
    free(tids);
}

void co_direct_adjoint(uint64_t nx, uint64_t ny, double cell_x, double cell_y, uint64_t n,
                       const double *u, const double *v, const double *w, const float *value,
                       int threads, float *image) {
    co_direct_adjoint_window(nx, ny, cell_x, cell_y, 0, 0, nx, ny, n, u, v, w, value, threads, image);
}

/* direct_predict_visibility, lib.rs:73870-73887: components are the non-zero
 * model pixels in (x, y) order. */
void co_direct_predict(uint64_t nx, uint64_t ny, double cell_x, double cell_y, const float *model,
                       uint64_t n, const double *u, const double *v, const double *w, float *out) {
    const double TAU = 6.28318530717958647692;
    for (uint64_t i = 0; i < n; ++i) {
        float re = 0.0f, im = 0.0f;
        for (uint64_t x = 0; x < nx; ++x) {
            for (uint64_t y = 0; y < ny; ++y) {
                float value = model[x * ny + y];
                if (fabsf(value) <= 0.0f) continue;
                double l, m, nm1;
                co_pixel_lmn(nx, ny, cell_x, cell_y, x, y, &l, &m, &nm1);
                double phase = TAU * (u[i] * l + v[i] * m + w[i] * nm1);
                re += value * (float)cos(phase);
                im -= value * (float)sin(phase);
            }
        }
        out[2 * i] = re;
        out[2 * i + 1] = im;
    }
}

/* ------------------------------------------------------------------------- */
/* Density (uniform / Briggs) weighting                                      */
/* ------------------------------------------------------------------------- */

/* weight_density_cell_anchor, gridder.rs:1888-1934, with the density origin
 * and scale of StandardGridder::new_with_padding_factor (gridder.rs:421-441):
 * center = n / 2, scale = n * cell, limit = n.
 * convention 0 (VisImagingWeight): u, v, scale and origin as Float, the product
 *   and the sum as separate Float operations, x from -u, then truncation;
 * convention 1 (cube Briggs): x = u*scale + center, y = -v*scale + center in
 *   Double, rounded half away from zero.
 * Accepted cells satisfy 0 < x < nx and 0 < y < ny.  Returns 1 if accepted. */
int co_density_cell(uint64_t nx, uint64_t ny, double cell_x, double cell_y, int convention, double u,
                    double v, int64_t *ax, int64_t *ay) {
    const double cx = (double)nx / 2.0, cy = (double)ny / 2.0;
    const double su = (double)nx * cell_x, sv = (double)ny * cell_y;
    double x, y;
    if (convention == 0) {
        const float uf = (float)u, vf = (float)v;
        const float xs = (-uf) * (float)su;
        const float ys = vf * (float)sv;
        const float xf = xs + (float)cx;
        const float yf = ys + (float)cy;
        x = (double)xf;
        y = (double)yf;
    } else {
        x = u * su + cx;
        y = -v * sv + cy;
    }
    if (!(isfinite(x) && isfinite(y))) return 0;
    if (convention != 0) {
        x = round(x);
        y = round(y);
    }
    /* Rust's `as isize` truncates toward zero and saturates: anything this far
     * out is rejected by the bounds below either way */
    if (x >= 9.0e18 || x <= -9.0e18 || y >= 9.0e18 || y <= -9.0e18) return 0;
    const int64_t X = (int64_t)x, Y = (int64_t)y;
    if (X <= 0 || Y <= 0 || X >= (int64_t)nx || Y >= (int64_t)ny) return 0;
    *ax = X;
    *ay = Y;
    return 1;
}

/* accumulate_density_grid_serial / accumulate_density_sample_serial
 * (weighting.rs:1747-1856): samples in order, weight added to the sample's
 * cell and, with mirror, to the cell of (-u, -v), in Float.  plane[i] selects
 * the density plane of sample i (NULL: plane 0); usable[i] == 0 skips it. */
void co_density_grid(uint64_t nx, uint64_t ny, double cell_x, double cell_y, int convention, int mirror,
                     uint64_t n, const double *u, const double *v, const float *weight,
                     const uint32_t *plane, const uint8_t *usable, float *density) {
    for (uint64_t i = 0; i < n; ++i) {
        if (usable && !usable[i]) continue;
        const float w = weight[i];
        if (!(isfinite(w) && w > 0.0f)) continue;
        float *d = density + (plane ? (uint64_t)plane[i] * nx * ny : 0);
        int64_t x, y;
        if (co_density_cell(nx, ny, cell_x, cell_y, convention, u[i], v[i], &x, &y)) d[x * ny + y] += w;
        if (mirror && co_density_cell(nx, ny, cell_x, cell_y, convention, -u[i], -v[i], &x, &y))
            d[x * ny + y] += w;
    }
}

/* Briggs f2 of one density plane (weighting.rs:1211-1224, :2758-2777):
 * (5 * 10^-robust)^2 / (sum of squares of the positive cells / total), the
 * total being the plane's sum, or twice it when the density holds no
 * conjugate samples (robust_density_weight_sum_for_f2, :1493-1501). */
float co_briggs_f2(uint64_t cells, const float *density, float robust, int includes_conjugates) {
    double sum = 0.0, sumsq = 0.0;
    for (uint64_t i = 0; i < cells; ++i) {
        sum += (double)density[i];
        if (density[i] > 0.0f) sumsq += (double)density[i] * (double)density[i];
    }
    const double total = includes_conjugates ? sum : 2.0 * sum;
    if (!(sumsq > 0.0 && total > 0.0)) return 0.0f;
    const double s = 5.0 * pow(10.0, -(double)robust);
    return (float)(s * s / (sumsq / total));
}

/* reweight_owned_batch_uniform_in_place / _briggs_in_place
 * (weighting.rs:2501-2592) with casa_briggs_denominator (:2440-2443):
 * uniform w / d, Briggs w / (d * f2 + 1) as separate Float operations; a
 * sample off the density grid, or with a non-positive weight or density,
 * gets 0.  mode 1 uniform, 2 Briggs; f2 per plane. */
void co_density_reweight(uint64_t nx, uint64_t ny, double cell_x, double cell_y, int convention, int mode,
                         uint64_t n, const double *u, const double *v, const float *weight,
                         const uint32_t *plane, const uint8_t *usable, const float *density,
                         const float *f2, float *out) {
    for (uint64_t i = 0; i < n; ++i) {
        const float w = weight[i];
        if (usable && !usable[i]) {
            out[i] = w;
            continue;
        }
        const uint64_t p = plane ? plane[i] : 0;
        int64_t x, y;
        if (!co_density_cell(nx, ny, cell_x, cell_y, convention, u[i], v[i], &x, &y)) {
            out[i] = 0.0f;
            continue;
        }
        const float d = density[p * nx * ny + (uint64_t)x * ny + (uint64_t)y];
        if (!(isfinite(w) && w > 0.0f && isfinite(d) && d > 0.0f)) {
            out[i] = 0.0f;
            continue;
        }
        if (mode == 1) {
            out[i] = w / d;
        } else {
            const float product = d * f2[p];
            const float denominator = product + 1.0f;
            out[i] = w / denominator;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Phase-centre rotation (casars-imager/src/lib.rs:47064-47069,              */
/* :48030-48159; casa-imaging/src/lib.rs:248-259)                            */
/* ------------------------------------------------------------------------- */

static void co_mat3_mul(const double a[3][3], const double b[3][3], double out[3][3]) {
    double t[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
    memcpy(out, t, sizeof(t));
}

/* gridft_axis_rotation: axis 0 = X, 1 = Z */
static void co_axis_rotation(double angle, int axis, double m[3][3]) {
    const double s = sin(angle), c = cos(angle);
    if (axis == 0) {
        double x[3][3] = {{1.0, 0.0, 0.0}, {0.0, c, -s}, {0.0, s, c}};
        memcpy(m, x, sizeof(x));
    } else {
        double z[3][3] = {{c, -s, 0.0}, {s, c, 0.0}, {0.0, 0.0, 1.0}};
        memcpy(m, z, sizeof(z));
    }
}

/* gridft_euler_rotation of two operations, starting from the identity */
static void co_euler2(double a1, int ax1, double a2, int ax2, double out[3][3]) {
    double m[3][3] = {{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}, r[3][3];
    co_axis_rotation(a1, ax1, r);
    co_mat3_mul(m, r, m);
    co_axis_rotation(a2, ax2, r);
    co_mat3_mul(m, r, m);
    memcpy(out, m, sizeof(m));
}

static const double CO_HALF_PI = 1.57079632679489661923;

/* rot3 of the target direction (X then Z) */
static void co_target_rotation(const double tgt[2], double out[3][3]) {
    co_euler2(CO_HALF_PI - tgt[1], 0, -(tgt[0] - CO_HALF_PI), 1, out);
}

/* gridft_uvw_rotation_matrix (:48074-48090): transpose(rot3 * rot1), row-major out[9] */
void co_uvw_rotation_matrix(const double src[2], const double tgt[2], double *out) {
    double rot1[3][3], rot3[3][3], p[3][3];
    co_euler2(-(CO_HALF_PI - src[0]), 1, src[1] - CO_HALF_PI, 0, rot1);
    co_target_rotation(tgt, rot3);
    co_mat3_mul(rot3, rot1, p);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) out[3 * r + c] = p[c][r];
}

/* direction_cosines_from_angles / uvw_phase_rotation_vector_from_angles (:48030-48055) */
void co_phase_rotation_vector(const double src[2], const double tgt[2], double *out) {
    double rot3[3][3];
    co_target_rotation(tgt, rot3);
    const double tc[3] = {cos(tgt[1]) * cos(tgt[0]), cos(tgt[1]) * sin(tgt[0]), sin(tgt[1])};
    const double sc[3] = {cos(src[1]) * cos(src[0]), cos(src[1]) * sin(src[0]), sin(src[1])};
    const double d[3] = {tc[0] - sc[0], tc[1] - sc[1], tc[2] - sc[2]};
    for (int r = 0; r < 3; ++r) out[r] = rot3[r][0] * d[0] + rot3[r][1] * d[1] + rot3[r][2] * d[2];
}

/* Apply per row: uvw' = uvw * uvrot (row vector), shift = phrot . uvw'; per
 * (row, chan): phase_rotate_visibility with phase = -TAU * shift * f / c,
 * phasor cast to Float, complex Float multiply.  vis [row][chan] in place. */
void co_rotate_rows(uint64_t n_rows, uint64_t n_chan, const double *uvrot, const double *phrot,
                    const double *uvw, const double *freq, float *vis, double *uvw_out, double *shift_out) {
    const double TAU = 6.28318530717958647692, C = 299792458.0;
    for (uint64_t r = 0; r < n_rows; ++r) {
        const double *x = uvw + 3 * r;
        double o[3];
        for (int k = 0; k < 3; ++k) o[k] = x[0] * uvrot[k] + x[1] * uvrot[3 + k] + x[2] * uvrot[6 + k];
        const double shift = phrot[0] * o[0] + phrot[1] * o[1] + phrot[2] * o[2];
        if (uvw_out) memcpy(uvw_out + 3 * r, o, sizeof(o));
        if (shift_out) shift_out[r] = shift;
        for (uint64_t c = 0; c < n_chan; ++c) {
            if (shift == 0.0 || freq[c] == 0.0) continue;
            const double phase = -TAU * shift * freq[c] / C;
            const float pr = (float)cos(phase), pi = (float)sin(phase);
            float *v = vis + 2 * (r * n_chan + c);
            const float re = v[0] * pr - v[1] * pi;
            const float im = v[0] * pi + v[1] * pr;
            v[0] = re;
            v[1] = im;
        }
    }
}

/* Row-stream ingestion of one imaging plane: the routed-sample builder
 * (execution.rs:2168-2315) per (row, chan) of row blocks with DATA / FLAG
 * [corr][chan] and WEIGHT [corr] or WEIGHT_SPECTRUM [corr][chan]:
 * pol 0 (Explicit, types.rs:1611-1619): excluded when flagged, or when the
 * visibility is non-finite and allow_psf_only == 0; weight of the corr;
 * pol 1 (CollapsedPair, types.rs:1620-1631): excluded when either hand is
 * flagged, the collapse (types.rs:1593-1610) is non-finite or either weight
 * is not finite and positive; weight 0.5 (w0 + w1).  row_flag excludes the
 * row.  rotate: uvw' and each hand's phase rotation as co_rotate_rows,
 * before the collapse (casars-imager lib.rs:19207-19211).  Outputs per
 * sample (row, chan): vis complex64, weight, sumwt_factor, flag (1 =
 * excluded); uvw_out per row. */
static void co_collapse(int t, const float *a, const float *b, float *out) {
    if (t == 0) {
        out[0] = (a[0] + b[0]) * 0.5f;
        out[1] = (a[1] + b[1]) * 0.5f;
        return;
    }
    const float dr = a[0] - b[0], di = a[1] - b[1];
    if (t == 1) {
        out[0] = dr * 0.5f;
        out[1] = di * 0.5f;
        return;
    }
    const float im = t == 2 ? 0.5f : -0.5f; /* Complex32::new(0, +-0.5), full complex product */
    out[0] = dr * 0.0f - di * im;
    out[1] = dr * im + di * 0.0f;
}

static void co_phase_rotate(float *v, double shift, double f) {
    if (shift == 0.0 || f == 0.0) return;
    const double phase = -6.28318530717958647692 * shift * f / 299792458.0;
    const float pr = (float)cos(phase), pi = (float)sin(phase);
    const float re = v[0] * pr - v[1] * pi, im = v[0] * pi + v[1] * pr;
    v[0] = re;
    v[1] = im;
}

void co_ingest_rows(uint64_t n_rows, uint32_t n_corr, uint32_t n_chan, int pol, uint32_t c0, uint32_t c1,
                    int transform, float sumwt_factor, int allow_psf_only, int spectrum, int rotate,
                    const double *uvrot, const double *phrot, const double *freq, const double *uvw,
                    const float *data, const uint8_t *flag, const float *weight, const uint8_t *row_flag,
                    double *uvw_out, float *vis_out, float *wt_out, float *sf_out, uint8_t *flag_out) {
    const uint64_t per = (uint64_t)n_corr * n_chan;
    for (uint64_t r = 0; r < n_rows; ++r) {
        const double *x = uvw + 3 * r;
        double o[3] = {x[0], x[1], x[2]}, shift = 0.0;
        if (rotate) {
            for (int k = 0; k < 3; ++k) o[k] = x[0] * uvrot[k] + x[1] * uvrot[3 + k] + x[2] * uvrot[6 + k];
            shift = phrot[0] * o[0] + phrot[1] * o[1] + phrot[2] * o[2];
        }
        memcpy(uvw_out + 3 * r, o, sizeof(o));
        for (uint32_t c = 0; c < n_chan; ++c) {
            const uint64_t i0 = r * per + (uint64_t)c0 * n_chan + c, i1 = r * per + (uint64_t)c1 * n_chan + c;
            const float w0 = spectrum ? weight[i0] : weight[r * n_corr + c0];
            const float w1 = spectrum ? weight[i1] : weight[r * n_corr + c1];
            int excluded = row_flag && row_flag[r];
            float v[2];
            float w;
            if (pol == 0) {
                excluded = excluded || flag[i0];
                v[0] = data[2 * i0];
                v[1] = data[2 * i0 + 1];
                if (!(isfinite(v[0]) && isfinite(v[1])) && !allow_psf_only) excluded = 1;
                if (rotate) co_phase_rotate(v, shift, freq[c]);
                w = w0;
            } else {
                excluded = excluded || flag[i0] || flag[i1];
                float a[2] = {data[2 * i0], data[2 * i0 + 1]}, b[2] = {data[2 * i1], data[2 * i1 + 1]};
                if (rotate) {
                    co_phase_rotate(a, shift, freq[c]);
                    co_phase_rotate(b, shift, freq[c]);
                }
                co_collapse(transform, a, b, v);
                if (!(isfinite(v[0]) && isfinite(v[1])) ||
                    !(isfinite(w0) && w0 > 0.0f && isfinite(w1) && w1 > 0.0f))
                    excluded = 1;
                w = 0.5f * (w0 + w1);
            }
            const uint64_t o_i = r * n_chan + c;
            vis_out[2 * o_i] = v[0];
            vis_out[2 * o_i + 1] = v[1];
            wt_out[o_i] = w;
            sf_out[o_i] = sumwt_factor;
            flag_out[o_i] = (uint8_t)(excluded ? 1 : 0);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* AWProject convolution-function cells and the Airy A-term                   */
/* ------------------------------------------------------------------------- */

/* aw_kernel_origin, gridder.rs:2809-2815 */
uint64_t co_aw_kernel_origin(uint64_t length) { return length % 2 == 0 ? length / 2 : length / 2 + 1; }

/* One CF cell: stored shape, supports and the offset (in complex elements) of
 * its pixels [x][y] (y fastest) in a shared pixel array. */
typedef struct {
    uint64_t shape_x, shape_y;
    uint64_t x_support, y_support;
    uint64_t offset;
} co_cf_cell;

/* awproject_sample_geometry, gridder.rs:2699-2748.  out = loc_x, loc_y,
 * off_x, off_y, conjugate.  Returns 0, or the AwProjectSamplePlanRejection
 * (gridder.rs:2133-2138 order): 1 NonFiniteCoordinate, 2 OutsideGrid,
 * 3 KernelIndexOutsideCell. */
int co_aw_geometry(uint64_t gx, uint64_t gy, double du, double dv, uint64_t sampling,
                   const co_cf_cell *cell, double u, double v, double w, int64_t *out) {
    const double pos_x = u / du + (double)gx / 2.0;
    const double pos_y = -v / dv + (double)gy / 2.0;
    if (!(isfinite(pos_x) && isfinite(pos_y) && isfinite(w))) return 1;
    /* |pos| beyond 4e18 saturates the reference's cast and is off-grid either way */
    if (fabs(pos_x) > 4.0e18 || fabs(pos_y) > 4.0e18) return 2;
    const int64_t loc_x = (int64_t)round(pos_x), loc_y = (int64_t)round(pos_y);
    const double s = (double)sampling;
    const int64_t off_x = (int64_t)round(((double)loc_x - pos_x) * s);
    const int64_t off_y = (int64_t)round(((double)loc_y - pos_y) * s);
    const int64_t xs = (int64_t)cell->x_support, ys = (int64_t)cell->y_support;
    if (loc_x - xs < 0 || loc_y - ys < 0 || loc_x + xs >= (int64_t)gx || loc_y + ys >= (int64_t)gy) return 2;
    const int64_t cx = (int64_t)co_aw_kernel_origin(cell->shape_x);
    const int64_t cy = (int64_t)co_aw_kernel_origin(cell->shape_y);
    const int64_t si = (int64_t)sampling;
    if (cx - xs * si + off_x < 0 || cy - ys * si + off_y < 0 || cx + xs * si + off_x >= (int64_t)cell->shape_x ||
        cy + ys * si + off_y >= (int64_t)cell->shape_y)
        return 3;
    out[0] = loc_x;
    out[1] = loc_y;
    out[2] = off_x;
    out[3] = off_y;
    out[4] = w > 0.0 ? 1 : 0;
    return 0;
}

/* casa_aw_axis_phase / casa_aw_phase_gradient_from_axes, gridder.rs:2758-2766:
 * each axis phasor in f64 narrowed to f32, the product in f64 narrowed to f32. */
static void co_aw_phasor(double phase_x, double phase_y, float *out) {
    const float xr = (float)cos(phase_x), xi = (float)sin(phase_x);
    const float yr = (float)cos(phase_y), yi = (float)sin(phase_y);
    const double re = (double)xr * (double)yr - (double)xi * (double)yi;
    const double im = (double)xr * (double)yi + (double)xi * (double)yr;
    out[0] = (float)re;
    out[1] = (float)im;
}

/* plan_sample_detailed, gridder.rs:2253-2311: geometry, then the f32 and f64
 * sums of the (conjugated when w > 0) unphased taps, iy outer, ix inner.
 * Returns the rejection code, 4 = InvalidNormalization.  norm32 = 2 floats,
 * norm64 = 2 doubles. */
int co_aw_plan(uint64_t gx, uint64_t gy, double du, double dv, uint64_t sampling, const co_cf_cell *cell,
               const float *pixels, double u, double v, double w, int64_t *geo, float *norm32, double *norm64) {
    const int rc = co_aw_geometry(gx, gy, du, dv, sampling, cell, u, v, w, geo);
    if (rc) return rc;
    const int64_t xs = (int64_t)cell->x_support, ys = (int64_t)cell->y_support, si = (int64_t)sampling;
    const int64_t cx = (int64_t)co_aw_kernel_origin(cell->shape_x), cy = (int64_t)co_aw_kernel_origin(cell->shape_y);
    const float *px = pixels + 2 * cell->offset;
    float nr = 0.0f, ni = 0.0f;
    double dr = 0.0, di = 0.0;
    for (int64_t iy = -ys; iy <= ys; ++iy) {
        const int64_t ky = cy + iy * si + geo[3];
        for (int64_t ix = -xs; ix <= xs; ++ix) {
            const int64_t kx = cx + ix * si + geo[2];
            const float *t = px + 2 * ((uint64_t)kx * cell->shape_y + (uint64_t)ky);
            const float tr = t[0], ti = geo[4] ? -t[1] : t[1];
            nr += tr;
            ni += ti;
            dr += (double)tr;
            di += (double)ti;
        }
    }
    norm32[0] = nr;
    norm32[1] = ni;
    norm64[0] = dr;
    norm64[1] = di;
    if (!(isfinite(nr) && isfinite(ni) && (nr != 0.0f || ni != 0.0f) && isfinite(dr) && isfinite(di) &&
          (dr != 0.0 || di != 0.0)))
        return 4;
    return 0;
}

/* for_each_grid_tap, gridder.rs:2636-2666: the phased taps of a plan, in
 * (iy outer, ix inner) order: raw CF pixel, conjugated when w > 0, times the
 * pointing phasor of (ix*s + off_x, iy*s + off_y), Complex32 product. */
void co_aw_taps(uint64_t sampling, const co_cf_cell *cell, const float *pixels, const int64_t *geo,
                const double *gradient, float *out) {
    const int64_t xs = (int64_t)cell->x_support, ys = (int64_t)cell->y_support, si = (int64_t)sampling;
    const int64_t cx = (int64_t)co_aw_kernel_origin(cell->shape_x), cy = (int64_t)co_aw_kernel_origin(cell->shape_y);
    const float *px = pixels + 2 * cell->offset;
    uint64_t k = 0;
    for (int64_t iy = -ys; iy <= ys; ++iy) {
        const int64_t ycoord = iy * si + geo[3];
        for (int64_t ix = -xs; ix <= xs; ++ix) {
            const int64_t xcoord = ix * si + geo[2];
            const float *t = px + 2 * ((uint64_t)(cx + xcoord) * cell->shape_y + (uint64_t)(cy + ycoord));
            const float tr = t[0], ti = geo[4] ? -t[1] : t[1];
            float ph[2];
            co_aw_phasor((double)xcoord * gradient[0], (double)ycoord * gradient[1], ph);
            out[2 * k] = tr * ph[0] - ti * ph[1];
            out[2 * k + 1] = tr * ph[1] + ti * ph[0];
            ++k;
        }
    }
}

/* Sequential AW gridding over samples in order (grid_sample_planned_f64,
 * gridder.rs:2318-2343): contribution = value * tap as Complex32, added to a
 * Complex64 grid [gx][gy].  Per sample: cell index, pointing gradient
 * (2 doubles per pointing), value (complex64).  Rejected samples are skipped;
 * status[i] receives the plan code. */
void co_aw_grid_f64(uint64_t gx, uint64_t gy, double du, double dv, uint64_t sampling, const co_cf_cell *cells,
                    const float *pixels, const double *gradients, uint64_t n, const double *u, const double *v,
                    const double *w, const uint32_t *cell_of, const uint32_t *pointing_of, const float *values,
                    double *grid, int32_t *status) {
    float *taps = NULL;
    uint64_t cap = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const co_cf_cell *c = cells + cell_of[i];
        int64_t geo[5];
        float n32[2];
        double n64[2];
        const int rc = co_aw_plan(gx, gy, du, dv, sampling, c, pixels, u[i], v[i], w[i], geo, n32, n64);
        if (status) status[i] = rc;
        if (rc) continue;
        const uint64_t T = (2 * c->x_support + 1) * (2 * c->y_support + 1);
        if (T > cap) {
            free(taps);
            taps = (float *)malloc(2 * T * sizeof(float));
            cap = T;
        }
        co_aw_taps(sampling, c, pixels, geo, gradients + 2 * pointing_of[i], taps);
        const float vr = values[2 * i], vi = values[2 * i + 1];
        const int64_t xs = (int64_t)c->x_support, ys = (int64_t)c->y_support;
        uint64_t k = 0;
        for (int64_t iy = -ys; iy <= ys; ++iy)
            for (int64_t ix = -xs; ix <= xs; ++ix, ++k) {
                const float tr = taps[2 * k], ti = taps[2 * k + 1];
                const float cr = vr * tr - vi * ti, ci = vr * ti + vi * tr;
                double *g = grid + 2 * ((uint64_t)(geo[0] + ix) * gy + (uint64_t)(geo[1] + iy));
                g[0] += (double)cr;
                g[1] += (double)ci;
            }
    }
    free(taps);
}

/* degrid_sample_planned, gridder.rs:2359-2371: Complex32 sum of
 * conj(tap) * grid in tap order, divided by conj(normalization) with
 * num-complex 0.4's division ((a c + b d) / |z|^2, (b c - a d) / |z|^2).
 * model: complex64 [gx][gy].  Rejected samples give 0 and their status. */
void co_aw_degrid(uint64_t gx, uint64_t gy, double du, double dv, uint64_t sampling, const co_cf_cell *cells,
                  const float *pixels, const double *gradients, uint64_t n, const double *u, const double *v,
                  const double *w, const uint32_t *cell_of, const uint32_t *pointing_of, const float *model,
                  float *out, int32_t *status) {
    float *taps = NULL;
    uint64_t cap = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const co_cf_cell *c = cells + cell_of[i];
        int64_t geo[5];
        float n32[2];
        double n64[2];
        const int rc = co_aw_plan(gx, gy, du, dv, sampling, c, pixels, u[i], v[i], w[i], geo, n32, n64);
        if (status) status[i] = rc;
        out[2 * i] = out[2 * i + 1] = 0.0f;
        if (rc) continue;
        const uint64_t T = (2 * c->x_support + 1) * (2 * c->y_support + 1);
        if (T > cap) {
            free(taps);
            taps = (float *)malloc(2 * T * sizeof(float));
            cap = T;
        }
        co_aw_taps(sampling, c, pixels, geo, gradients + 2 * pointing_of[i], taps);
        const int64_t xs = (int64_t)c->x_support, ys = (int64_t)c->y_support;
        float ar = 0.0f, ai = 0.0f;
        uint64_t k = 0;
        for (int64_t iy = -ys; iy <= ys; ++iy)
            for (int64_t ix = -xs; ix <= xs; ++ix, ++k) {
                const float tr = taps[2 * k], ti = -taps[2 * k + 1];
                const float *g = model + 2 * ((uint64_t)(geo[0] + ix) * gy + (uint64_t)(geo[1] + iy));
                const float pr = tr * g[0] - ti * g[1];
                const float pi = tr * g[1] + ti * g[0];
                ar += pr;
                ai += pi;
            }
        const float c_ = n32[0], d_ = -n32[1];
        const float ns = c_ * c_ + d_ * d_;
        const float re = ar * c_ + ai * d_;
        const float im = ai * c_ - ar * d_;
        out[2 * i] = re / ns;
        out[2 * i + 1] = im / ns;
    }
    free(taps);
}

/* The AW gridding of co_aw_grid_f64 / co_aw_degrid in FP64 throughout: the
 * phased Complex32 taps widened to f64, prediction sum of conj(tap) * model
 * over an f64 model grid divided by conj of the f64 normalisation, value
 * (V - p) * w (mode 2), V * w (mode 1, with w * tap into the PSF grid) or w
 * (mode 0) times the tap, all in f64 into f64 grids.  The FP64 reference that
 * FP32 results are measured against; plans are co_aw_plan's.  Samples with
 * usable = 0, a non-positive / non-finite weight or a non-finite V are
 * skipped (the W-path acceptance rule).  sums[0] += w |n64| (the sum of
 * weights the device accumulates per accepted sample), sums[1] counts the
 * gridded samples. */
void co_aw_grid64(uint64_t gx, uint64_t gy, double du, double dv, uint64_t sampling, const co_cf_cell *cells,
                  const float *pixels, const double *gradients, uint64_t n, const double *u, const double *v,
                  const double *w, const uint32_t *cell_of, const uint32_t *pointing_of, const float *vis,
                  const float *weight, const uint8_t *usable, const double *model, int mode, double *psf,
                  double *dirty, double *sums) {
    float *taps = NULL;
    uint64_t cap = 0;
    for (uint64_t i = 0; i < n; ++i) {
        const float wt = weight[i];
        if ((usable && !usable[i]) || !(isfinite(wt) && wt > 0.0f)) continue;
        if (!(isfinite(vis[2 * i]) && isfinite(vis[2 * i + 1]))) continue;
        const co_cf_cell *c = cells + cell_of[i];
        int64_t geo[5];
        float n32[2];
        double n64[2];
        if (co_aw_plan(gx, gy, du, dv, sampling, c, pixels, u[i], v[i], w[i], geo, n32, n64)) continue;
        const uint64_t T = (2 * c->x_support + 1) * (2 * c->y_support + 1);
        if (T > cap) {
            free(taps);
            taps = (float *)malloc(2 * T * sizeof(float));
            cap = T;
        }
        co_aw_taps(sampling, c, pixels, geo, gradients + 2 * pointing_of[i], taps);
        const double dw = (double)wt;
        sums[0] += dw * hypot(n64[0], n64[1]);
        sums[1] += 1.0;
        const int64_t xs = (int64_t)c->x_support, ys = (int64_t)c->y_support;
        double val_r = 0.0, val_i = 0.0;
        if (mode == 2) {
            double ar = 0.0, ai = 0.0;
            if (model) {
                uint64_t k = 0;
                for (int64_t iy = -ys; iy <= ys; ++iy)
                    for (int64_t ix = -xs; ix <= xs; ++ix, ++k) {
                        const double tr = taps[2 * k], ti = -(double)taps[2 * k + 1];
                        const double *g = model + 2 * ((uint64_t)(geo[0] + ix) * gy + (uint64_t)(geo[1] + iy));
                        ar += tr * g[0] - ti * g[1];
                        ai += tr * g[1] + ti * g[0];
                    }
            }
            /* / conj(n64) */
            const double cr = n64[0], ci = -n64[1], ns = cr * cr + ci * ci;
            const double pr = (ar * cr + ai * ci) / ns, pi = (ai * cr - ar * ci) / ns;
            val_r = ((double)vis[2 * i] - pr) * dw;
            val_i = ((double)vis[2 * i + 1] - pi) * dw;
        } else if (mode == 1) {
            val_r = (double)vis[2 * i] * dw;
            val_i = (double)vis[2 * i + 1] * dw;
        }
        uint64_t k = 0;
        for (int64_t iy = -ys; iy <= ys; ++iy)
            for (int64_t ix = -xs; ix <= xs; ++ix, ++k) {
                const double tr = taps[2 * k], ti = taps[2 * k + 1];
                const uint64_t idx = 2 * ((uint64_t)(geo[0] + ix) * gy + (uint64_t)(geo[1] + iy));
                if (mode != 2 && psf) {
                    psf[idx] += dw * tr;
                    psf[idx + 1] += dw * ti;
                }
                if (mode != 0 && dirty) {
                    dirty[idx] += val_r * tr - val_i * ti;
                    dirty[idx + 1] += val_r * ti + val_i * tr;
                }
            }
    }
    free(taps);
}

/* casa_airy_max_radius_arcmin_ghz, lib.rs:57592-57605 */
double co_airy_max_radius(double dish, double blockage) {
    if (fabs(dish - 10.7) <= 1.0e-6 && fabs(blockage - 0.75) <= 1.0e-6) return 250.0;
    if (fabs(dish - 6.25) <= 1.0e-6 && fabs(blockage - 0.75) <= 1.0e-6) return 500.0;
    return 1.784 * 60.0;
}

/* airy_voltage_pattern_from_table_index, lib.rs:57557-57578 (j1: the C
 * library's Bessel function of the first kind, order one). */
float co_airy_from_index(double index, double max_radius, double dish, double blockage) {
    const double n1 = 9999.0;
    const double xmax = max_radius * 7.016 / (1.566 * 60.0) * dish / 24.5;
    const double x = index * xmax / n1;
    if (fabs(x) <= 2.220446049250313e-16) return 1.0f;
    if (blockage <= 0.0) return (float)(2.0 * j1(x) / x);
    const double area_ratio = (dish / blockage) * (dish / blockage);
    const double area_norm = area_ratio - 1.0;
    const double lr = dish / blockage;
    return (float)((area_ratio * 2.0 * j1(x) / x - 2.0 * j1(x * lr) / (x * lr)) / area_norm);
}

/* AiryVoltageLookup::new, lib.rs:57479-57503: 10,000 entries. */
void co_airy_table(double dish, double blockage, float *out) {
    const double mr = co_airy_max_radius(dish, blockage);
    for (int k = 0; k < 10000; ++k) out[k] = co_airy_from_index((double)k, mr, dish, blockage);
}

static int64_t co_airy_index(double radius_arcmin_ghz, double max_radius) {
    double idx = floor(radius_arcmin_ghz * 9999.0 / max_radius);
    if (idx < 0.0) idx = 0.0;
    if (idx > 9999.0) idx = 9999.0;
    return (int64_t)idx;
}

/* AiryVoltageLookup::evaluate_offsets, lib.rs:57505-57532: degrees narrowed
 * to f32, radius in f32, arcmin*GHz narrowed to f32, truncated table index. */
float co_airy_offsets(double l, double m, double freq, double dish, double blockage, const float *table) {
    const double DEG = 57.29577951308232; /* f64::to_degrees multiplier, 180/pi */
    if (!(isfinite(l) && isfinite(m) && isfinite(freq) && freq > 0.0)) return 0.0f;
    const float ld = (float)(l * DEG), md = (float)(m * DEG);
    const float rd = sqrtf(ld * ld + md * md);
    const double r = (double)(float)((double)rd * 60.0 * (freq / 1.0e9));
    const double mr = co_airy_max_radius(dish, blockage);
    if (r > mr) return 0.0f;
    return table[co_airy_index(r, mr)];
}

/* airy_voltage_pattern, lib.rs:57436-57453: one radius in radians. */
float co_airy_radius(double radius, double freq, double dish, double blockage) {
    const double DEG = 57.29577951308232;
    if (!(isfinite(radius) && radius >= 0.0 && isfinite(freq) && freq > 0.0)) return 0.0f;
    const double r = (double)(float)(radius * DEG * 60.0 * (freq / 1.0e9));
    const double mr = co_airy_max_radius(dish, blockage);
    if (r > mr) return 0.0f;
    return co_airy_from_index((double)co_airy_index(r, mr), mr, dish, blockage);
}

void co_airy_offsets_many(uint64_t n, const double *l, const double *m, double freq, double dish, double blockage,
                          const float *table, float *out) {
    for (uint64_t i = 0; i < n; ++i) out[i] = co_airy_offsets(l[i], m[i], freq, dish, blockage, table);
}

/* ------------------------------------------------------------------------- */
/* Hogbom minor cycle (run_hogbom_minor_cycle_cpu_with_progress,             */
/* lib.rs:59748-59885), one image plane at a time.                            */
/* ------------------------------------------------------------------------- */

/* Stop codes: CleanStopReason (types.rs:2215-2230) in declaration order,
 * plus 1; 0 = the component budget was used up. */
enum { CO_STOP_NONE = 0, CO_STOP_GLOBAL = 1, CO_STOP_NSIGMA = 2, CO_STOP_CYCLE = 3, CO_STOP_NO_PIXELS = 6 };

/* threshold_reached_with_tolerance, lib.rs:74628-74635 */
static int co_threshold_reached_with_tolerance(float peak_abs, float threshold) {
    if (threshold <= 0.0f) return peak_abs <= threshold;
    const float tol = fmaxf(0.01f * threshold, 2.0e-8f);
    return peak_abs <= threshold || fabsf(peak_abs - threshold) <= tol;
}

/* minor_cycle_stop_reason, lib.rs:74653-74677 */
int co_minor_cycle_stop(float peak_abs, float threshold, float cycle_threshold, float nsigma_threshold) {
    const int floor_active = threshold > 0.0f && (0.01f * threshold) < 2.0e-8f && cycle_threshold <= threshold;
    if (peak_abs < threshold || (floor_active && co_threshold_reached_with_tolerance(peak_abs, threshold)))
        return CO_STOP_GLOBAL;
    if (cycle_threshold >= nsigma_threshold) return peak_abs < cycle_threshold ? CO_STOP_CYCLE : CO_STOP_NONE;
    if (nsigma_threshold > threshold) return peak_abs < nsigma_threshold ? CO_STOP_NSIGMA : CO_STOP_NONE;
    return CO_STOP_NONE;
}

/* peak_location_masked_in_bounds_zero_tolerance over the whole plane,
 * lib.rs:74064-74115: visited x-major; the first visited pixel seeds the
 * best, a later one replaces it when |v| is larger, or equal with a smaller
 * y-major key y*nx + x.  Returns 0 when no pixel is unmasked. */
static int co_hclean_peak(const float *img, const uint8_t *mask, uint64_t nx, uint64_t ny, uint64_t *px,
                          uint64_t *py, float *value) {
    int found = 0;
    float best = 0.0f;
    uint64_t bkey = 0;
    for (uint64_t x = 0; x < nx; ++x) {
        for (uint64_t y = 0; y < ny; ++y) {
            if (mask && !mask[x * ny + y]) continue;
            const float v = img[x * ny + y];
            const uint64_t key = y * nx + x;
            if (!found || fabsf(v) > fabsf(best) || (fabsf(v) == fabsf(best) && key < bkey)) {
                found = 1;
                best = v;
                bkey = key;
                *px = x;
                *py = y;
            }
        }
    }
    *value = best;
    return found;
}

/* subtract_shifted_kernel_in_window over the full image, lib.rs:74805-74868:
 * image[x, y] -= scale * kernel[x - px + kx/2, y - py + ky/2] where both
 * overlap (one f32 product, then the subtraction). */
static void co_subtract_shifted_psf(float *img, uint64_t nx, uint64_t ny, const float *psf, uint64_t kx,
                                    uint64_t ky, uint64_t px, uint64_t py, float scale) {
    const int64_t xmin = (int64_t)px - (int64_t)(kx / 2), ymin = (int64_t)py - (int64_t)(ky / 2);
    const int64_t xmax = xmin + (int64_t)kx - 1, ymax = ymin + (int64_t)ky - 1;
    if (nx == 0 || ny == 0 || xmax < 0 || ymax < 0 || xmin > (int64_t)nx - 1 || ymin > (int64_t)ny - 1) return;
    const int64_t x0 = xmin > 0 ? xmin : 0, y0 = ymin > 0 ? ymin : 0;
    const int64_t x1 = xmax < (int64_t)nx - 1 ? xmax : (int64_t)nx - 1;
    const int64_t y1 = ymax < (int64_t)ny - 1 ? ymax : (int64_t)ny - 1;
    for (int64_t x = x0; x <= x1; ++x)
        for (int64_t y = y0; y <= y1; ++y) {
            const float p = scale * psf[(uint64_t)(x - xmin) * ky + (uint64_t)(y - ymin)];
            img[(uint64_t)x * ny + (uint64_t)y] = img[(uint64_t)x * ny + (uint64_t)y] - p;
        }
}

/* One Hogbom minor cycle on each of n_planes planes (residual, model: nx*ny
 * each; psf: kx*ky each; mask: nx*ny bytes per plane or NULL), lib.rs:59761-
 * 59810: up to `budget` components (hogbom_component_budget, lib.rs:64541-
 * 64551: niter, plus one in CASA's inclusive mode).  Writes the components
 * taken and the stop code per plane. */
void co_hogbom(uint64_t n_planes, uint64_t nx, uint64_t ny, float *residual, uint64_t kx, uint64_t ky,
               const float *psf, float *model, const uint8_t *mask, float gain, float threshold,
               const float *cycle_threshold, const float *nsigma_threshold, uint64_t budget, uint64_t *updates,
               int32_t *stop) {
    for (uint64_t c = 0; c < n_planes; ++c) {
        float *r = residual + c * nx * ny;
        float *m = model + c * nx * ny;
        const float *k = psf + c * kx * ky;
        const uint8_t *mk = mask ? mask + c * nx * ny : NULL;
        uint64_t done = 0;
        int32_t why = CO_STOP_NONE;
        while (done < budget) {
            uint64_t px = 0, py = 0;
            float pv = 0.0f;
            if (!co_hclean_peak(r, mk, nx, ny, &px, &py, &pv)) {
                why = CO_STOP_NO_PIXELS;
                break;
            }
            const int s = co_minor_cycle_stop(fabsf(pv), threshold, cycle_threshold[c], nsigma_threshold[c]);
            if (s) {
                why = s;
                break;
            }
            const float comp = gain * pv;
            m[px * ny + py] = m[px * ny + py] + comp;
            co_subtract_shifted_psf(r, nx, ny, k, kx, ky, px, py, comp);
            ++done;
        }
        updates[c] = done;
        stop[c] = why;
    }
}
