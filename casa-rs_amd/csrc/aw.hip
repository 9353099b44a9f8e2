// AWProject convolution-function atlas on the GPU (AwProjector,
// gridder.rs:2157-2815) and the Airy A-term CF builder (lib.rs:57436-57605).
//
// The reference evaluates, per sample, every tap of its CF cell as
//   tap = CF[origin + i*s + off] (conjugated when w > 0) * phasor(i*s + off),
// phasor = axis phasors cos/sin in f64 narrowed to f32, multiplied in f64 and
// narrowed to f32 (casa_aw_phase_gradient_from_axes), the tap product in f32.
// A tap therefore depends only on (cell, pointing, w sign, sub-cell offset,
// tap), so the atlas precomputes those values once into the same "compact"
// tap tables the W-projection path uses (table t = (cell * pointings +
// pointing) * 2 + conjugate, offsets (off_x, off_y), taps (ix, iy) with iy
// fastest).  The binned and plain gridders then read AW taps exactly as W
// taps, bit-identical to the per-sample evaluation.  Normalisations
// (plan_sample_detailed: sequential f32 and f64 sums of the conjugated,
// unphased taps) are computed once per (cell, offset) on the host in the
// reference order; the conjugated variant is the exact negation of the
// imaginary sum.
#include <cmath>
#include <cstring>

#include "device.hpp"

namespace chip {

namespace {

unsigned nb(uint64_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

// Phased taps of every (table, offset); zero outside the cell's own supports.
__global__ __launch_bounds__(256) void k_aw_compact(const float2 *pixels, const AwCell *cells, uint32_t pointings,
                                                    uint32_t n_tables, uint32_t sampling, int32_t R,
                                                    uint32_t stride, const float2 *phx, const float2 *phy,
                                                    uint32_t ph_len, int32_t ph_half, float2 *out) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t n_off = static_cast<uint32_t>(2 * R + 1);
    const uint64_t total = static_cast<uint64_t>(n_tables) * n_off * n_off * stride;
    if (i >= total) return;
    const uint32_t f = static_cast<uint32_t>(i % stride);
    const uint64_t po = i / stride;
    const uint32_t t = static_cast<uint32_t>(po / (n_off * n_off));
    const uint32_t off = static_cast<uint32_t>(po - static_cast<uint64_t>(t) * n_off * n_off);
    const int ox = static_cast<int>(off / n_off) - R, oy = static_cast<int>(off % n_off) - R;
    const uint32_t conj = t & 1u;
    const uint32_t pt = (t >> 1) % pointings;
    const uint32_t cell = (t >> 1) / pointings;
    const AwCell c = cells[cell];
    const int S = c.xs > c.ys ? c.xs : c.ys;
    const int side = 2 * S + 1;
    float2 v = make_float2(0.0f, 0.0f);
    if (static_cast<int>(f) < side * side) {
        const int ix = static_cast<int>(f) / side - S, iy = static_cast<int>(f) % side - S;
        const int s = static_cast<int>(sampling);
        const int xc = ix * s + ox, yc = iy * s + oy;
        const int64_t kx = c.cx + xc, ky = c.cy + yc;
        if (ix >= -c.xs && ix <= c.xs && iy >= -c.ys && iy <= c.ys && kx >= 0 && ky >= 0 && kx < c.shape_x &&
            ky < c.shape_y && xc >= -ph_half && xc <= ph_half && yc >= -ph_half && yc <= ph_half) {
            const float2 raw = pixels[c.offset + static_cast<uint64_t>(kx) * c.shape_y + ky];
            const float tr = raw.x, ti = conj ? -raw.y : raw.y;
            const float2 a = phx[static_cast<uint64_t>(pt) * ph_len + (xc + ph_half)];
            const float2 b = phy[static_cast<uint64_t>(pt) * ph_len + (yc + ph_half)];
            const double pre = static_cast<double>(a.x) * static_cast<double>(b.x) -
                               static_cast<double>(a.y) * static_cast<double>(b.y);
            const double pim = static_cast<double>(a.x) * static_cast<double>(b.y) +
                               static_cast<double>(a.y) * static_cast<double>(b.x);
            const float pr = static_cast<float>(pre), pi = static_cast<float>(pim);
            const float re = tr * pr - ti * pi;
            const float im = tr * pi + ti * pr;
            v = make_float2(re, im);
        }
    }
    out[i] = v;
}

// plan_sample_detailed per sample (no weight / flag filters).
__global__ __launch_bounds__(256) void k_aw_plan_raw(PlanParams p, Source s, const float2 *n32,
                                                     const double2 *n64, int32_t *plan, float *o32,
                                                     double *o64) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= s.n_rows * s.n_chan) return;
    const SampleIn smp = load_sample(s, i);
    int32_t *o = plan + 6 * i;
    for (int k = 0; k < 5; ++k) o[k] = 0;
    float2 m32 = make_float2(0.0f, 0.0f);
    double2 m64 = make_double2(0.0, 0.0);
    const uint32_t cell = aw_cell_of(p, smp);
    const uint32_t pt = p.pointing ? p.pointing[smp.row] : 0u;
    int code = CHIP_AW_NO_CELL;
    if (cell < p.aw_cells && pt < p.aw_pointings) {
        int64_t lx = 0, ly = 0, ox = 0, oy = 0;
        code = aw_geometry(p, p.aw_cell[cell], smp.u, smp.v, smp.w, lx, ly, ox, oy);
        const int64_t R = p.norm_radius, axis = 2 * R + 1;
        if (code == CHIP_AW_OK && (ox > R || ox < -R || oy > R || oy < -R)) code = CHIP_AW_KERNEL_INDEX_OUTSIDE_CELL;
        if (code == CHIP_AW_OK) {
            const uint64_t k = static_cast<uint64_t>((static_cast<int64_t>(cell) * axis + (ox + R)) * axis + (oy + R));
            const bool conj = smp.w > 0.0;
            m32 = n32[k];
            m64 = n64[k];
            if (conj) {
                m32.y = -m32.y;
                m64.y = -m64.y;
            }
            if (!(p.aw_mag[k] > 0.0)) code = CHIP_AW_INVALID_NORMALIZATION;
            o[0] = static_cast<int32_t>(lx);
            o[1] = static_cast<int32_t>(ly);
            o[2] = static_cast<int32_t>(ox);
            o[3] = static_cast<int32_t>(oy);
            o[4] = conj ? 1 : 0;
        }
    }
    o[5] = code;
    if (o32) {
        o32[2 * i] = m32.x;
        o32[2 * i + 1] = m32.y;
    }
    if (o64) {
        o64[2 * i] = m64.x;
        o64[2 * i + 1] = m64.y;
    }
}

// ---- Airy A-term CF builder -------------------------------------------------

struct AiryScreen {
    uint32_t conv, inner;
    double s0, s1;
    double w_scale;
    uint32_t planes, first;
    double cos_pa, sin_pa;
    double freq, max_radius;
};

// AiryVoltageLookup::evaluate_offsets (lib.rs:57505-57532): degrees narrowed
// to f32, f32 radius, arcmin*GHz narrowed to f32, truncated table index.
__device__ __forceinline__ float airy_lookup(const AiryScreen &a, const float *table, double l, double m) {
#pragma clang fp contract(off)
    const double deg = 57.29577951308232;  // f64::to_degrees multiplier (180 / pi)
    const float ld = static_cast<float>(l * deg), md = static_cast<float>(m * deg);
    const float rd = sqrtf(ld * ld + md * md);
    const double r = static_cast<double>(static_cast<float>(static_cast<double>(rd) * 60.0 * (a.freq / 1.0e9)));
    if (r > a.max_radius) return 0.0f;
    double idx = floor(r * 9999.0 / a.max_radius);
    idx = idx < 0.0 ? 0.0 : (idx > 9999.0 ? 9999.0 : idx);
    return table[static_cast<int>(idx)];
}

// Origin-wrapped screens of a batch of W planes (the W-projection screen of
// gridder.rs:3434-3470 times the Airy power pattern of the rotated l, m).
__global__ __launch_bounds__(256) void k_aw_screen(AiryScreen a, const float *corr, const float *table,
                                                   float2 *screens, uint32_t batch) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const int half = static_cast<int>(a.inner / 2);
    const uint64_t side = 2 * static_cast<uint64_t>(half);
    const uint64_t per = side * side;
    if (i >= per * batch) return;
    const uint32_t b = static_cast<uint32_t>(i / per);
    const uint64_t r = i - static_cast<uint64_t>(b) * per;
    const int iy = static_cast<int>(r / side) - half;
    const int ix = static_cast<int>(r % side) - half;
    const uint32_t plane = a.first + b;
    const double w = a.planes > 1 ? static_cast<double>(static_cast<uint64_t>(plane) * plane) / a.w_scale : 0.0;
    const double m = a.s1 * static_cast<double>(iy);
    const double l = a.s0 * static_cast<double>(ix);
    const double rsq = l * l + m * m;
    if (rsq >= 1.0) return;
    const double lr = l * a.cos_pa - m * a.sin_pa;
    const double mr = l * a.sin_pa + m * a.cos_pa;
    const float vp = airy_lookup(a, table, lr, mr);
    const float amp = (corr[ix + half] * corr[iy + half]) * (vp * vp);
    const double phase = 6.283185307179586476925 * w * (sqrt(1.0 - rsq) - 1.0);
    const uint32_t x = ix >= 0 ? static_cast<uint32_t>(ix) : static_cast<uint32_t>(ix + static_cast<int>(a.conv));
    const uint32_t y = iy >= 0 ? static_cast<uint32_t>(iy) : static_cast<uint32_t>(iy + static_cast<int>(a.conv));
    const uint64_t cells = static_cast<uint64_t>(a.conv) * a.conv;
    screens[b * cells + static_cast<uint64_t>(x) * a.conv + y] =
        make_float2(amp * static_cast<float>(cos(phase)), amp * static_cast<float>(sin(phase)));
}

__global__ void k_aw_peak(const float2 *t, float *peak) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *peak = hypotf(t[0].x, t[0].y);
}

// max(|T[c,0]|, |T[0,c]|) / peak for c < quarter, per plane in the batch.
__global__ void k_aw_axis_mag(const float2 *t, uint32_t conv, uint32_t quarter, uint32_t batch,
                              const float *peak, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= quarter * batch) return;
    const uint32_t b = i / quarter, c = i - b * quarter;
    const uint64_t base = static_cast<uint64_t>(b) * conv * conv;
    const float p = *peak;
    const float2 x = t[base + static_cast<uint64_t>(c) * conv];
    const float2 y = t[base + c];
    out[i] = fmaxf(hypotf(x.x / p, x.y / p), hypotf(y.x / p, y.y / p));
}

// Full-plane cell of one transformed plane: cell[o + d] = T[d mod conv] / peak
// for d in [-o, shape - o), o = shape / 2.
__global__ void k_aw_crop(const float2 *t, uint32_t conv, uint32_t shape, const float *peak, float2 *cell) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= shape * shape) return;
    const int o = static_cast<int>(shape / 2);
    const int x = static_cast<int>(i / shape) - o, y = static_cast<int>(i % shape) - o;
    const uint32_t sx = static_cast<uint32_t>(x < 0 ? x + static_cast<int>(conv) : x);
    const uint32_t sy = static_cast<uint32_t>(y < 0 ? y + static_cast<int>(conv) : y);
    const float p = *peak;
    const float2 v = t[static_cast<uint64_t>(sx) * conv + sy];
    cell[i] = make_float2(v.x / p, v.y / p);
}

__global__ void k_aw_scale(float2 *k, uint64_t n, float pb) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = k[i];
    k[i] = make_float2(v.x / pb, v.y / pb);
}

int64_t aw_origin(uint64_t n) { return static_cast<int64_t>(n % 2 == 0 ? n / 2 : n / 2 + 1); }

}  // namespace

// casa_airy_max_radius_arcmin_ghz (lib.rs:57592-57605)
double airy_max_radius(double dish, double blockage) {
    if (std::fabs(dish - 10.7) <= 1.0e-6 && std::fabs(blockage - 0.75) <= 1.0e-6) return 250.0;
    if (std::fabs(dish - 6.25) <= 1.0e-6 && std::fabs(blockage - 0.75) <= 1.0e-6) return 500.0;
    return 1.784 * 60.0;
}

// AiryVoltageLookup::new (lib.rs:57479-57503) with
// airy_voltage_pattern_from_table_index (lib.rs:57557-57578).
std::vector<float> airy_table(double dish, double blockage) {
    const double mr = airy_max_radius(dish, blockage);
    const double xmax = mr * 7.016 / (1.566 * 60.0) * dish / 24.5;
    std::vector<float> t(10000);
    for (int k = 0; k < 10000; ++k) {
        const double x = static_cast<double>(k) * xmax / 9999.0;
        if (std::fabs(x) <= 2.220446049250313e-16) {
            t[k] = 1.0f;
        } else if (blockage <= 0.0) {
            t[k] = static_cast<float>(2.0 * ::j1(x) / x);
        } else {
            const double ar = (dish / blockage) * (dish / blockage);
            const double lr = dish / blockage;
            t[k] = static_cast<float>((ar * 2.0 * ::j1(x) / x - 2.0 * ::j1(x * lr) / (x * lr)) / (ar - 1.0));
        }
    }
    return t;
}

void attach_aw(chip_gridder &g, uint32_t n_cells, const chip_cf_cell *cells, const float *pixels_dev,
               uint64_t n_pixels, uint32_t sampling, uint32_t n_pointings, const double *gradients,
               double w_increment) {
    if (g.wproject) fail(CHIP_ERR_INVALID, "a W projection or AW atlas is already attached");
    if (n_cells == 0 || !cells || !pixels_dev || n_pixels == 0)
        fail(CHIP_ERR_INVALID, "an AW atlas needs cells and pixels");
    if (sampling == 0 || sampling > 127) fail(CHIP_ERR_INVALID, "AW sampling must be 1..127");
    if (n_pointings == 0 || !gradients) fail(CHIP_ERR_INVALID, "an AW atlas needs at least one pointing gradient");
    for (uint64_t k = 0; k < 2ull * n_pointings; ++k)
        if (!std::isfinite(gradients[k])) fail(CHIP_ERR_INVALID, "pointing phase gradients must be finite");
    if (!(std::isfinite(w_increment) && w_increment >= 0.0))
        fail(CHIP_ERR_INVALID, "w_increment must be finite and non-negative");
    const uint64_t n_tables = static_cast<uint64_t>(n_cells) * n_pointings * 2;
    if (n_tables > kMaxTables) fail(CHIP_ERR_INVALID, "cells x pointings x 2 must be at most 8191 tap tables");
    const int64_t s = sampling;
    std::vector<AwCell> dc(n_cells);
    int32_t smax = 0;
    for (uint32_t k = 0; k < n_cells; ++k) {
        const chip_cf_cell &c = cells[k];
        if (c.shape_x == 0 || c.shape_y == 0 || c.shape_x > (1u << 20) || c.shape_y > (1u << 20))
            fail(CHIP_ERR_INVALID, "CF cell " + std::to_string(k) + " has an invalid shape");
        if (c.offset > n_pixels || c.shape_x * c.shape_y > n_pixels - c.offset)
            fail(CHIP_ERR_INVALID, "CF cell " + std::to_string(k) + " lies outside the pixel array");
        if (c.x_support > 1024 || c.y_support > 1024)
            fail(CHIP_ERR_INVALID, "CF cell " + std::to_string(k) + " support exceeds 1024");
        const int64_t cx = aw_origin(c.shape_x), cy = aw_origin(c.shape_y);
        const int64_t rx = static_cast<int64_t>(c.x_support) * s, ry = static_cast<int64_t>(c.y_support) * s;
        // AwProjector::new (gridder.rs:2210-2229)
        if (cx < rx || cy < ry || cx + rx >= static_cast<int64_t>(c.shape_x) ||
            cy + ry >= static_cast<int64_t>(c.shape_y))
            fail(CHIP_ERR_INVALID, "CF cell " + std::to_string(k) + ": support " + std::to_string(c.x_support) +
                                       "x" + std::to_string(c.y_support) + " at sampling " +
                                       std::to_string(sampling) + " is outside the kernel shape around its origin");
        dc[k] = AwCell{static_cast<int64_t>(c.shape_x), static_cast<int64_t>(c.shape_y),
                       static_cast<int32_t>(c.x_support), static_cast<int32_t>(c.y_support),
                       static_cast<int32_t>(cx), static_cast<int32_t>(cy), c.offset};
        smax = std::max<int32_t>(smax, static_cast<int32_t>(std::max(c.x_support, c.y_support)));
    }
    const int32_t R = static_cast<int32_t>((sampling + 1) / 2);
    const uint32_t n_off = static_cast<uint32_t>(2 * R + 1);
    const uint64_t stride = static_cast<uint64_t>(2 * smax + 1) * (2 * smax + 1);
    const uint64_t n_compact = n_tables * n_off * n_off * stride;
    if (n_compact * sizeof(float2) > (32ull << 30))
        fail(CHIP_ERR_INVALID, "AW tap tables would exceed 32 GiB");
    hipStream_t st = g.ctx->stream;

    // Normalisations per (cell, offset), on the host in the reference order
    // (iy outer, ix inner; plan_sample_detailed, gridder.rs:2276-2297).
    std::vector<float2> hp(n_pixels);
    CHIP_HIP(hipMemcpyAsync(hp.data(), pixels_dev, n_pixels * sizeof(float2), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    const uint64_t per_cell = static_cast<uint64_t>(n_off) * n_off;
    std::vector<float2> n32(n_cells * per_cell, make_float2(0.0f, 0.0f));
    std::vector<double2> n64(n_cells * per_cell, make_double2(0.0, 0.0));
    std::vector<double> mag(n_cells * per_cell, 0.0);
    for (uint32_t k = 0; k < n_cells; ++k) {
        const AwCell &c = dc[k];
        for (int ox = -R; ox <= R; ++ox)
            for (int oy = -R; oy <= R; ++oy) {
                const uint64_t e = k * per_cell + static_cast<uint64_t>(ox + R) * n_off + (oy + R);
                if (c.cx - c.xs * s + ox < 0 || c.cy - c.ys * s + oy < 0 || c.cx + c.xs * s + ox >= c.shape_x ||
                    c.cy + c.ys * s + oy >= c.shape_y)
                    continue;
                float nr = 0.0f, ni = 0.0f;
                double dr = 0.0, di = 0.0;
                for (int64_t iy = -c.ys; iy <= c.ys; ++iy) {
                    const int64_t ky = c.cy + iy * s + oy;
                    for (int64_t ix = -c.xs; ix <= c.xs; ++ix) {
                        const int64_t kx = c.cx + ix * s + ox;
                        const float2 t = hp[c.offset + static_cast<uint64_t>(kx) * c.shape_y + ky];
                        nr += t.x;
                        ni += t.y;
                        dr += static_cast<double>(t.x);
                        di += static_cast<double>(t.y);
                    }
                }
                n32[e] = make_float2(nr, ni);
                n64[e] = make_double2(dr, di);
                const bool ok = std::isfinite(nr) && std::isfinite(ni) && (nr != 0.0f || ni != 0.0f) &&
                                std::isfinite(dr) && std::isfinite(di) && (dr != 0.0 || di != 0.0);
                mag[e] = ok ? std::hypot(dr, di) : 0.0;
            }
    }
    // Normalisation per tap table (w sign applied) for the degrid division.
    std::vector<float2> tnorm(n_tables * per_cell);
    for (uint64_t t = 0; t < n_tables; ++t) {
        const uint64_t cell = (t >> 1) / n_pointings;
        for (uint64_t e = 0; e < per_cell; ++e) {
            float2 v = n32[cell * per_cell + e];
            if (t & 1) v.y = -v.y;
            tnorm[t * per_cell + e] = v;
        }
    }
    // Axis phasors per pointing (casa_aw_axis_phase, gridder.rs:2758-2760):
    // coordinates -half..half with half = smax * s + R.
    const int32_t ph_half = static_cast<int32_t>(smax * s + R);
    const uint32_t ph_len = static_cast<uint32_t>(2 * ph_half + 1);
    std::vector<float2> phx(static_cast<uint64_t>(n_pointings) * ph_len), phy(phx.size());
    for (uint32_t p = 0; p < n_pointings; ++p)
        for (int32_t c = -ph_half; c <= ph_half; ++c) {
            const double ax = static_cast<double>(c) * gradients[2 * p];
            const double ay = static_cast<double>(c) * gradients[2 * p + 1];
            phx[static_cast<uint64_t>(p) * ph_len + (c + ph_half)] =
                make_float2(static_cast<float>(std::cos(ax)), static_cast<float>(std::sin(ax)));
            phy[static_cast<uint64_t>(p) * ph_len + (c + ph_half)] =
                make_float2(static_cast<float>(std::cos(ay)), static_cast<float>(std::sin(ay)));
        }

    DevBuf d_pix(n_pixels * sizeof(float2));
    CHIP_HIP(hipMemcpyAsync(d_pix.p, pixels_dev, d_pix.bytes, hipMemcpyDeviceToDevice, st));
    auto upload = [&](DevBuf &d, const void *h, size_t bytes) {
        d.alloc(bytes);
        CHIP_HIP(hipMemcpyAsync(d.p, h, bytes, hipMemcpyHostToDevice, st));
    };
    DevBuf d_cell, d_n32, d_n64, d_mag, d_tnorm, d_phx, d_phy, d_sup;
    upload(d_cell, dc.data(), dc.size() * sizeof(AwCell));
    upload(d_n32, n32.data(), n32.size() * sizeof(float2));
    upload(d_n64, n64.data(), n64.size() * sizeof(double2));
    upload(d_mag, mag.data(), mag.size() * sizeof(double));
    upload(d_tnorm, tnorm.data(), tnorm.size() * sizeof(float2));
    upload(d_phx, phx.data(), phx.size() * sizeof(float2));
    upload(d_phy, phy.data(), phy.size() * sizeof(float2));
    std::vector<int32_t> sup(n_tables);
    for (uint64_t t = 0; t < n_tables; ++t) {
        const AwCell &c = dc[(t >> 1) / n_pointings];
        sup[t] = std::max(c.xs, c.ys);
    }
    upload(d_sup, sup.data(), sup.size() * sizeof(int32_t));
    DevBuf compact(n_compact * sizeof(float2));
    hipLaunchKernelGGL(k_aw_compact, dim3(nb(n_compact, 256)), dim3(256), 0, st, d_pix.as<float2>(),
                       d_cell.as<AwCell>(), n_pointings, static_cast<uint32_t>(n_tables), sampling, R,
                       static_cast<uint32_t>(stride), d_phx.as<float2>(), d_phy.as<float2>(), ph_len, ph_half,
                       compact.as<float2>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));

    g.wproject = true;
    g.aw = true;
    g.wk_chunk = 0;  // the atlas is always resident
    g.aw_cells = n_cells;
    g.aw_pointings = n_pointings;
    g.aw_w_incr = w_increment;
    g.aw_meta.assign(cells, cells + n_cells);
    g.aw_n_pixels = n_pixels;
    g.planes = static_cast<uint32_t>(n_tables);
    g.sampling = sampling;
    g.width = 0;
    g.max_support = static_cast<uint32_t>(smax);
    g.norm_radius = R;
    g.w_scale = w_increment;
    g.supports = sup;
    g.norm_table.clear();
    g.tap_stride = static_cast<uint32_t>(stride);
    g.d_compact = std::move(compact);
    g.d_supports = std::move(d_sup);
    g.d_aw_pixels = std::move(d_pix);
    g.d_aw_cell = std::move(d_cell);
    g.d_aw_n32 = std::move(d_n32);
    g.d_aw_n64 = std::move(d_n64);
    g.d_aw_mag = std::move(d_mag);
    g.d_aw_tnorm = std::move(d_tnorm);
    // image-plane corrections of the oversampled kernel, as W projection
    g.sinc = sinc_axis(std::max(g.gx, g.gy), sampling);
    g.sinc_x = sinc_axis(g.gx, sampling);
    g.sinc_y = sinc_axis(g.gy, sampling);
    upload(g.d_sinc, g.sinc.data(), g.sinc.size() * sizeof(float));
    upload(g.d_sinc_x, g.sinc_x.data(), g.sinc_x.size() * sizeof(float));
    upload(g.d_sinc_y, g.sinc_y.data(), g.sinc_y.size() * sizeof(float));
    CHIP_HIP(hipStreamSynchronize(st));
}

void attach_aw_airy(chip_gridder &g, const chip_airy_params &ap, uint32_t n_freq, const double *freqs,
                    uint32_t n_pointings, const double *gradients) {
    if (g.wproject) fail(CHIP_ERR_INVALID, "a W projection or AW atlas is already attached");
    if (!(std::isfinite(ap.dish_m) && ap.dish_m > 0.0 && std::isfinite(ap.blockage_m) && ap.blockage_m >= 0.0 &&
          ap.blockage_m < ap.dish_m))
        fail(CHIP_ERR_INVALID, "Airy dish / blockage diameters must satisfy 0 <= blockage < dish");
    if (n_freq == 0 || !freqs) fail(CHIP_ERR_INVALID, "at least one CF frequency is needed");
    for (uint32_t k = 0; k < n_freq; ++k)
        if (!(std::isfinite(freqs[k]) && freqs[k] > 0.0)) fail(CHIP_ERR_INVALID, "frequency must be positive");
    if (!std::isfinite(ap.pa_rad)) fail(CHIP_ERR_INVALID, "parallactic angle must be finite");
    if (ap.sampling < 2 || ap.sampling > 126 || ap.sampling % 2) fail(CHIP_ERR_INVALID, "sampling must be even, 2..126");
    if (ap.n_w == 0 || ap.n_w > 1023) fail(CHIP_ERR_INVALID, "n_w must be 1..1023");
    if (!(std::isfinite(ap.max_abs_w) && ap.max_abs_w >= 0.0)) fail(CHIP_ERR_INVALID, "max_abs_w must be finite");
    if (static_cast<uint64_t>(n_freq) * ap.n_w * n_pointings * 2 > kMaxTables)
        fail(CHIP_ERR_INVALID, "frequencies x W planes x pointings x 2 must be at most 8191 tap tables");
    hipStream_t st = g.ctx->stream;
    const chip_geometry &geo = g.geom;
    const uint32_t s = ap.sampling, planes = ap.n_w;
    const uint32_t conv = static_cast<uint32_t>(std::max(g.gx, g.gy));
    const uint32_t inner = std::max<uint32_t>(conv / s, 2u);
    const uint32_t quarter = conv / 2 - 1;
    if (quarter < 2) fail(CHIP_ERR_INVALID, "grid too small for AW projection");
    const std::vector<float> corr = gridder_correction_axis(inner);
    const std::vector<float> table = airy_table(ap.dish_m, ap.blockage_m);
    double w_scale = 1.0;
    if (planes > 1 && ap.max_abs_w > 0.0)
        w_scale = static_cast<double>((planes - 1) * (planes - 1)) / (1.05 * ap.max_abs_w);
    const int32_t R = static_cast<int32_t>((s + 1) / 2);
    AiryScreen a{};
    a.conv = conv;
    a.inner = inner;
    a.s0 = std::fabs(geo.cell_x) * s * static_cast<double>(g.gx) / conv;
    a.s1 = std::fabs(geo.cell_y) * s * static_cast<double>(g.gy) / conv;
    a.w_scale = w_scale;
    a.planes = planes;
    a.cos_pa = std::cos(ap.pa_rad);
    a.sin_pa = std::sin(ap.pa_rad);
    a.max_radius = airy_max_radius(ap.dish_m, ap.blockage_m);

    const uint64_t plane_bytes = static_cast<uint64_t>(conv) * conv * sizeof(float2);
    const uint32_t per_batch = static_cast<uint32_t>(
        std::min<uint64_t>(planes, std::max<uint64_t>(1, (4ull << 30) / plane_bytes)));
    DevBuf screens(plane_bytes * per_batch), d_corr(inner * sizeof(float)), d_table(table.size() * sizeof(float));
    DevBuf d_peak(sizeof(float)), d_mag(static_cast<uint64_t>(quarter) * per_batch * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(d_corr.p, corr.data(), d_corr.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(d_table.p, table.data(), d_table.bytes, hipMemcpyHostToDevice, st));
    const uint32_t n_cells = n_freq * planes;  // cell f * planes + k: frequency f, W plane k
    std::vector<uint32_t> support(n_cells), shape(n_cells);
    std::vector<DevBuf> cellbuf(n_cells);
    std::vector<float> mag(static_cast<uint64_t>(quarter) * per_batch);
    for (uint32_t f = 0; f < n_freq; ++f) {
        a.freq = freqs[f];
        for (uint32_t first = 0; first < planes; first += per_batch) {
            const uint32_t nbat = std::min(per_batch, planes - first);
            CHIP_HIP(hipMemsetAsync(screens.p, 0, plane_bytes * nbat, st));
            a.first = first;
            const uint64_t work = static_cast<uint64_t>(2 * (inner / 2)) * (2 * (inner / 2)) * nbat;
            hipLaunchKernelGGL(k_aw_screen, dim3(nb(work, 256)), dim3(256), 0, st, a, d_corr.as<float>(),
                               d_table.as<float>(), screens.as<float2>(), nbat);
            CHIP_HIP(hipGetLastError());
            hipfftHandle h = g.ctx->fft.get(conv, conv, nbat, HIPFFT_C2C, st);
            CHIP_FFT(hipfftExecC2C(h, screens.as<hipfftComplex>(), screens.as<hipfftComplex>(), HIPFFT_FORWARD));
            if (first == 0) {  // this frequency's plane-zero peak
                hipLaunchKernelGGL(k_aw_peak, dim3(1), dim3(64), 0, st, screens.as<float2>(), d_peak.as<float>());
                CHIP_HIP(hipGetLastError());
                float peak = 0.0f;
                CHIP_HIP(hipMemcpyAsync(&peak, d_peak.p, sizeof(float), hipMemcpyDeviceToHost, st));
                CHIP_HIP(hipStreamSynchronize(st));
                if (!(std::isfinite(peak) && peak > 0.0f))
                    fail(CHIP_ERR_NORMALIZATION, "AW plane-zero kernel peak is zero or non-finite");
            }
            hipLaunchKernelGGL(k_aw_axis_mag, dim3(nb(static_cast<uint64_t>(quarter) * nbat, 256)), dim3(256), 0,
                               st, screens.as<float2>(), conv, quarter, nbat, d_peak.as<float>(), d_mag.as<float>());
            CHIP_HIP(hipGetLastError());
            CHIP_HIP(hipMemcpyAsync(mag.data(), d_mag.p, static_cast<uint64_t>(quarter) * nbat * sizeof(float),
                                    hipMemcpyDeviceToHost, st));
            CHIP_HIP(hipStreamSynchronize(st));
            for (uint32_t b = 0; b < nbat; ++b) {
                const uint32_t k = f * planes + first + b;
                std::vector<float> m(mag.begin() + static_cast<uint64_t>(b) * quarter,
                                     mag.begin() + static_cast<uint64_t>(b + 1) * quarter);
                support[k] = find_w_support(m, conv, s);
                const uint64_t sh = 2 * (static_cast<uint64_t>(support[k]) * s + R + 1);
                if (sh > conv) fail(CHIP_ERR_INVALID, "AW kernel support exceeds the screen");
                shape[k] = static_cast<uint32_t>(sh);
                cellbuf[k].alloc(sh * sh * sizeof(float2));
                hipLaunchKernelGGL(k_aw_crop, dim3(nb(sh * sh, 256)), dim3(256), 0, st,
                                   screens.as<float2>() + static_cast<uint64_t>(b) * conv * conv, conv, shape[k],
                                   d_peak.as<float>(), cellbuf[k].as<float2>());
                CHIP_HIP(hipGetLastError());
            }
        }
    }
    screens.release();
    // pack the cells; each frequency's cells are divided by that frequency's
    // plane-zero integral at integer taps
    std::vector<chip_cf_cell> cells(n_cells);
    uint64_t total = 0;
    for (uint32_t k = 0; k < n_cells; ++k) {
        cells[k] = chip_cf_cell{shape[k], shape[k], support[k], support[k], total};
        total += static_cast<uint64_t>(shape[k]) * shape[k];
    }
    DevBuf pixels(total * sizeof(float2));
    for (uint32_t k = 0; k < n_cells; ++k)
        CHIP_HIP(hipMemcpyAsync(pixels.as<float2>() + cells[k].offset, cellbuf[k].p, cellbuf[k].bytes,
                                hipMemcpyDeviceToDevice, st));
    CHIP_HIP(hipStreamSynchronize(st));
    cellbuf.clear();
    for (uint32_t f = 0; f < n_freq; ++f) {
        const chip_cf_cell &c = cells[f * planes];
        std::vector<float2> c0(static_cast<uint64_t>(c.shape_x) * c.shape_y);
        CHIP_HIP(hipMemcpyAsync(c0.data(), pixels.as<float2>() + c.offset, c0.size() * sizeof(float2),
                                hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        float pb = 0.0f;
        {
            const int64_t S = c.x_support, o = static_cast<int64_t>(c.shape_x) / 2, n = static_cast<int64_t>(c.shape_x);
            for (int64_t iy = -S; iy <= S; ++iy)
                for (int64_t ix = -S; ix <= S; ++ix) pb += c0[(o + ix * s) * n + (o + iy * s)].x;
        }
        if (!(std::isfinite(pb) && pb > 0.0f))
            fail(CHIP_ERR_NORMALIZATION, "AW plane-zero kernel integral is non-positive");
        const uint64_t end = f + 1 < n_freq ? cells[(f + 1) * planes].offset : total;
        hipLaunchKernelGGL(k_aw_scale, dim3(nb(end - c.offset, 256)), dim3(256), 0, st, pixels.as<float2>() + c.offset,
                           end - c.offset, pb);
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
    attach_aw(g, n_cells, cells.data(), pixels.as<float>(), total, s, n_pointings, gradients, w_scale);
    g.aw_n_w = planes;
    g.aw_cf_freq.assign(freqs, freqs + n_freq);
}

// CF cell of every sample (rows x channels input): the CF frequency nearest
// to the sample's (freq = lambda_scale * c) times the W planes per frequency,
// plus the W plane round(sqrt(WIncr |w|)) clamped to those planes
// (select_key_for_sample, awprojection.rs:1127-1180, without the conjugate-
// beam and Mueller keys of a CASA cache).
__global__ void k_aw_select(Source s, double w_incr, uint32_t n_w, const double *cf_freq, uint32_t n_freq,
                            uint32_t *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= s.n_rows * s.n_chan) return;
    const uint64_t row = s.n_chan == 1 ? i : i / s.n_chan;
    const uint32_t chan = s.n_chan == 1 ? 0u : static_cast<uint32_t>(i - row * s.n_chan);
    const double ls = s.lambda_scale[chan];
    const double freq = ls * 299792458.0;
    uint32_t best = 0;
    double bd = fabs(cf_freq[0] - freq);
    for (uint32_t k = 1; k < n_freq; ++k) {
        const double d = fabs(cf_freq[k] - freq);
        if (d < bd) { bd = d; best = k; }
    }
    const double w = s.uvw_m[3 * row + 2] * ls;
    double r = (n_w > 1 && w_incr != 0.0) ? round(sqrt(w_incr * fabs(w))) : 0.0;
    const double top = static_cast<double>(n_w - 1);
    if (!(r >= 0.0)) r = 0.0;
    if (r > top) r = top;
    out[i] = best * n_w + static_cast<uint32_t>(r);
}

void aw_select_cells(chip_gridder &g, const chip_samples &cs, uint32_t *out) {
    if (!g.aw || g.aw_cf_freq.empty())
        fail(CHIP_ERR_INVALID, "no Airy CF atlas with CF frequencies attached");
    if (!cs.uvw_m_dev || !cs.lambda_scale_dev || cs.n_chan == 0)
        fail(CHIP_ERR_INVALID, "CF selection needs rows x channels input (uvw_m, lambda_scale)");
    if (!out) fail(CHIP_ERR_INVALID, "NULL output");
    const uint64_t n = cs.n_rows * cs.n_chan;
    if (n == 0) return;
    hipStream_t st = g.ctx->stream;
    DevBuf d_f(g.aw_cf_freq.size() * sizeof(double));
    CHIP_HIP(hipMemcpyAsync(d_f.p, g.aw_cf_freq.data(), d_f.bytes, hipMemcpyHostToDevice, st));
    const Source src = make_source(cs);
    hipLaunchKernelGGL(k_aw_select, dim3(nb(n, 256)), dim3(256), 0, st, src, g.aw_w_incr, g.aw_n_w,
                       d_f.as<double>(), static_cast<uint32_t>(g.aw_cf_freq.size()), out);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
}

// Every imaging cell of a CASA CF cache, in key order, as the gridder's AW
// atlas (AwProjector::new per selected cell, gridder.rs:2194-2232): one
// oversampling for all cells, the cache's WIncr, and the cache's selection
// axes kept for cf_select_cells.
void attach_aw_cache(chip_gridder &g, const CfCache &c, uint32_t n_pointings, const double *gradients) {
    if (c.keys.empty()) fail(CHIP_ERR_INVALID, "empty CF cache");
    const uint64_t sampling = c.imaging[0].sampling;
    for (const CfKernelMeta &m : c.imaging)
        if (m.sampling != sampling)
            fail(CHIP_ERR_INVALID, "CF cells of one atlas share one oversampling: " + m.path + " differs");
    if (sampling > 127) fail(CHIP_ERR_INVALID, "AW sampling must be 1..127");
    if (c.keys.size() > 0xffffffffull) fail(CHIP_ERR_INVALID, "too many CF cells");
    std::vector<chip_cf_cell> cells(c.keys.size());
    std::vector<float> host;
    for (size_t k = 0; k < c.keys.size(); ++k) {
        const CfKernelMeta &m = c.imaging[k];
        if (m.x_support > 1024 || m.y_support > 1024)
            fail(CHIP_ERR_INVALID, "CF cell " + m.path + " support exceeds 1024");
        std::vector<float> px = cf_read_kernel_pixels(m);
        cells[k] = chip_cf_cell{m.shape[0], m.shape[1], static_cast<uint32_t>(m.x_support),
                                static_cast<uint32_t>(m.y_support), host.size() / 2};
        host.insert(host.end(), px.begin(), px.end());
    }
    DevBuf pixels(host.size() * sizeof(float));
    hipStream_t st = g.ctx->stream;
    CHIP_HIP(hipMemcpyAsync(pixels.p, host.data(), pixels.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipStreamSynchronize(st));
    attach_aw(g, static_cast<uint32_t>(cells.size()), cells.data(), pixels.as<float>(), host.size() / 2,
              static_cast<uint32_t>(sampling), n_pointings, gradients, c.w_increment);
    g.aw_from_cache = true;
    g.aw_cache_axes = static_cast<const CfAxes &>(c);
}

// CASA CFBuffer selection of a CF cache cell per sample (select_key_for_sample
// and select_key_for_prediction_sample, awprojection.rs:1106-1300), the same
// FP64 steps as the host cf_select_key: conjugate-beam frequency map, Mueller
// element m or 15 - m by the sign of w, W plane round(sqrt(WIncr |w|)),
// nearest frequency and circularly nearest PA with ties to the smaller bit
// pattern.  Samples without a cell get 0xffffffff (rejected by the planner).
struct CfSelectArgs {
    const double *chan_freq;  // per channel, or NULL: lambda_scale * c
    double f_ref, pa_const;
    const double *pa_row;     // per row, or NULL: pa_const
    int32_t mueller, conjugate_beams, prediction;
    const double *freqs, *pas;
    const int32_t *muellers;
    const uint32_t *cell_of;
    uint32_t nf, nw, nm, np;
    double w_incr;
};

__device__ inline double cf_circular(double v) {
    double r = fmod(v + 180.0, 360.0);
    if (r < 0.0) r += 360.0;
    return r - 180.0;
}

__global__ __launch_bounds__(256) void k_cf_select(Source s, CfSelectArgs a, uint32_t *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= s.n_rows * s.n_chan) return;
    const uint64_t row = s.n_chan == 1 ? i : i / s.n_chan;
    const uint32_t chan = s.n_chan == 1 ? 0u : static_cast<uint32_t>(i - row * s.n_chan);
    const double ls = s.lambda_scale[chan];
    const double f = a.chan_freq ? a.chan_freq[chan] : ls * 299792458.0;
    const double w = s.uvw_m[3 * row + 2] * ls;
    const double pa = a.pa_row ? a.pa_row[row] : a.pa_const;
    uint32_t cell = 0xffffffffu;
    bool ok = isfinite(f) && f > 0.0 && isfinite(w) && isfinite(pa);
    double req = f;
    if (ok && !a.prediction) {
        ok = isfinite(a.f_ref) && a.f_ref > 0.0;
        if (ok && a.conjugate_beams) {
            const double rad = 2.0 * a.f_ref * a.f_ref - f * f;
            ok = isfinite(rad) && rad > 0.0;
            req = ok ? sqrt(rad) : 0.0;
        }
    }
    if (ok) {
        const bool direct = a.prediction ? !(w > 0.0) : (w > 0.0);
        const int64_t m = direct ? static_cast<int64_t>(a.mueller) : 15 - static_cast<int64_t>(a.mueller);
        uint32_t mi = a.nm;
        for (uint32_t k = 0; k < a.nm; ++k)
            if (a.muellers[k] == m) mi = k;
        if (mi < a.nm) {
            uint32_t wi = 0;
            if (a.nw > 1 && a.w_incr != 0.0) {
                double r = round(sqrt(a.w_incr * fabs(w)));
                r = fmin(fmax(r, 0.0), static_cast<double>(a.nw - 1));
                wi = static_cast<uint32_t>(r);
            }
            uint32_t fi = 0;
            for (uint32_t k = 1; k < a.nf; ++k) {
                const double x = fabs(a.freqs[k] - req), y = fabs(a.freqs[fi] - req);
                if (x < y || (x == y && static_cast<uint64_t>(__double_as_longlong(a.freqs[k])) <
                                            static_cast<uint64_t>(__double_as_longlong(a.freqs[fi]))))
                    fi = k;
            }
            uint32_t pi = 0;
            for (uint32_t k = 1; k < a.np; ++k) {
                const double x = fabs(cf_circular(a.pas[k] - pa)), y = fabs(cf_circular(a.pas[pi] - pa));
                if (x < y || (x == y && static_cast<uint64_t>(__double_as_longlong(a.pas[k])) <
                                            static_cast<uint64_t>(__double_as_longlong(a.pas[pi]))))
                    pi = k;
            }
            cell = a.cell_of[((static_cast<uint64_t>(fi) * a.nw + wi) * a.nm + mi) * a.np + pi];
        }
    }
    out[i] = cell;
}

void cf_select_cells(chip_gridder &g, const chip_samples &cs, const double *chan_freq_hz, double f_ref,
                     int32_t mueller, const double *pa_deg_dev, double pa_deg, int conjugate_beams, int prediction,
                     uint32_t *out) {
    if (!g.aw || !g.aw_from_cache) fail(CHIP_ERR_INVALID, "no CF-cache atlas attached");
    if (!cs.uvw_m_dev || !cs.lambda_scale_dev || cs.n_chan == 0)
        fail(CHIP_ERR_INVALID, "CF selection needs rows x channels input (uvw_m, lambda_scale)");
    if (!out) fail(CHIP_ERR_INVALID, "NULL output");
    const uint64_t n = cs.n_rows * cs.n_chan;
    if (n == 0) return;
    const CfAxes &x = g.aw_cache_axes;
    hipStream_t st = g.ctx->stream;
    const size_t nd = x.freqs.size() + x.pas.size() + (chan_freq_hz ? cs.n_chan : 0);
    std::vector<double> hd;
    hd.reserve(nd);
    hd.insert(hd.end(), x.freqs.begin(), x.freqs.end());
    hd.insert(hd.end(), x.pas.begin(), x.pas.end());
    if (chan_freq_hz) hd.insert(hd.end(), chan_freq_hz, chan_freq_hz + cs.n_chan);
    DevBuf d_d(nd * sizeof(double)), d_m(x.muellers.size() * sizeof(int32_t)),
        d_c(x.cell_of.size() * sizeof(uint32_t));
    CHIP_HIP(hipMemcpyAsync(d_d.p, hd.data(), d_d.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(d_m.p, x.muellers.data(), d_m.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(d_c.p, x.cell_of.data(), d_c.bytes, hipMemcpyHostToDevice, st));
    CfSelectArgs a{};
    a.freqs = d_d.as<double>();
    a.pas = a.freqs + x.freqs.size();
    a.chan_freq = chan_freq_hz ? a.pas + x.pas.size() : nullptr;
    a.f_ref = f_ref;
    a.pa_const = pa_deg;
    a.pa_row = pa_deg_dev;
    a.mueller = mueller;
    a.conjugate_beams = conjugate_beams != 0;
    a.prediction = prediction != 0;
    a.muellers = d_m.as<int32_t>();
    a.cell_of = d_c.as<uint32_t>();
    a.nf = static_cast<uint32_t>(x.freqs.size());
    a.nw = static_cast<uint32_t>(x.w_values.size());
    a.nm = static_cast<uint32_t>(x.muellers.size());
    a.np = static_cast<uint32_t>(x.pas.size());
    a.w_incr = x.w_increment;
    const Source src = make_source(cs);
    hipLaunchKernelGGL(k_cf_select, dim3(nb(n, 256)), dim3(256), 0, st, src, a, out);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
}

void aw_plan_raw(chip_gridder &g, const Source &src, int32_t *plan, float *n32, double *n64) {
    if (!g.aw) fail(CHIP_ERR_INVALID, "no AW atlas attached");
    const uint64_t n = src.n_rows * src.n_chan;
    if (n == 0) return;
    hipLaunchKernelGGL(k_aw_plan_raw, dim3(nb(n, 256)), dim3(256), 0, g.ctx->stream, plan_params(g, src), src,
                       g.d_aw_n32.as<float2>(), g.d_aw_n64.as<double2>(), plan, n32, n64);
    CHIP_HIP(hipGetLastError());
}

}  // namespace chip
