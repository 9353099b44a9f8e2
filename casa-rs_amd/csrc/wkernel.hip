// W-projection kernel construction on the GPU (WProjector::new,
// gridder.rs:3403-3586).
//
// Per plane k (w_k = k^2 / w_scale): an origin-wrapped conv x conv screen
// psi(l) psi(m) exp(2 pi i w (sqrt(1 - l^2 - m^2) - 1)) over the inner
// conv/sampling points (phase in FP64, cos/sin narrowed to f32), a forward
// FP32 FFT per plane (hipFFT batched over planes), the support from the 1e-3
// threshold on the two axes, a crop to support*sampling + R + 1, division by
// the plane-zero peak, zero padding to the common width, division by the
// plane-zero integral at integer taps and the per-(plane, offset)
// normalisation table.  The FFT runs in f32 like the reference's rustfft
// call, so kernel values match the reference to FFT rounding.
#include <cmath>

#include "device.hpp"

namespace chip {

struct ScreenParams {
    uint32_t conv, inner;
    double s0, s1;
    double w_scale;
    uint32_t planes;  // total planes (for the w of a plane)
    uint32_t first;   // first plane of this batch
};

__global__ void k_screen(ScreenParams sp, const float *corr, float2 *screens, uint32_t batch) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    // Points -inner/2 .. inner/2 - 1 on each axis (gridder.rs:3444-3447).
    const int half = static_cast<int>(sp.inner / 2);
    const uint64_t side = 2 * static_cast<uint64_t>(half);
    const uint64_t per = side * side;
    if (i >= per * batch) return;
    const uint32_t b = static_cast<uint32_t>(i / per);
    const uint64_t r = i - static_cast<uint64_t>(b) * per;
    const int iy = static_cast<int>(r / side) - half;
    const int ix = static_cast<int>(r % side) - half;
    const uint32_t plane = sp.first + b;
    const double w = sp.planes > 1 ? static_cast<double>(static_cast<uint64_t>(plane) * plane) / sp.w_scale : 0.0;
    const double m = sp.s1 * static_cast<double>(iy);
    const double msq = m * m;
    const double l = sp.s0 * static_cast<double>(ix);
    const double rsq = l * l + msq;
    if (rsq >= 1.0) return;
    const double phase = 6.283185307179586476925 * w * (sqrt(1.0 - rsq) - 1.0);
    const float corr_v = corr[ix + half] * corr[iy + half];
    const uint32_t x = ix >= 0 ? static_cast<uint32_t>(ix) : static_cast<uint32_t>(ix + static_cast<int>(sp.conv));
    const uint32_t y = iy >= 0 ? static_cast<uint32_t>(iy) : static_cast<uint32_t>(iy + static_cast<int>(sp.conv));
    const uint64_t cells = static_cast<uint64_t>(sp.conv) * sp.conv;
    screens[b * cells + static_cast<uint64_t>(x) * sp.conv + y] =
        make_float2(corr_v * static_cast<float>(cos(phase)), corr_v * static_cast<float>(sin(phase)));
}

// max(|T[c,0]|, |T[0,c]|) / peak for c < quarter, per plane in the batch.
__global__ void k_axis_mag(const float2 *t, uint32_t conv, uint32_t quarter, uint32_t batch,
                           const float *peak, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= quarter * batch) return;
    const uint32_t b = i / quarter, c = i - b * quarter;
    const uint64_t base = static_cast<uint64_t>(b) * conv * conv;
    const float p = *peak;
    const float2 a = t[base + static_cast<uint64_t>(c) * conv];  // (c, 0)
    const float2 d = t[base + c];                                // (0, c)
    const float ma = hypotf(a.x / p, a.y / p);
    const float md = hypotf(d.x / p, d.y / p);
    out[i] = fmaxf(ma, md);
}

__global__ void k_peak(const float2 *t, float *peak) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *peak = hypotf(t[0].x, t[0].y);
}

// Crop plane b of the batch to crop x crop, divided by the plane-zero peak,
// into the staging area (planes x stage_w x stage_w, zero padded).
__global__ void k_crop(const float2 *t, uint32_t conv, uint32_t crop, uint32_t stage_w,
                       float p, float2 *stage) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= crop * crop) return;
    const uint32_t x = i / crop, y = i - x * crop;
    const float2 v = t[static_cast<uint64_t>(x) * conv + y];
    stage[static_cast<uint64_t>(x) * stage_w + y] = make_float2(v.x / p, v.y / p);
}

__global__ void k_scale(float2 *k, uint64_t n, float pb) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 v = k[i];
    k[i] = make_float2(v.x / pb, v.y / pb);
}

// compute_w_project_sample_normalization (gridder.rs:3782-3799): sequential
// f32 sum of Re over taps, iy outer, ix inner, one thread per entry.
__global__ void k_norm_table(const float2 *kernels, const int32_t *supports, uint32_t planes,
                             uint32_t width, uint32_t sampling, int32_t R, float *table) {
#pragma clang fp contract(off)
    const int axis = 2 * R + 1;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= planes * axis * axis) return;
    const uint32_t plane = i / (axis * axis);
    const int rem = static_cast<int>(i - plane * axis * axis);
    const int ox = rem / axis - R, oy = rem % axis - R;
    const int S = supports[plane];
    const int s = static_cast<int>(sampling);
    const float2 *K = kernels + static_cast<uint64_t>(plane) * width * width;
    float sum = 0.0f;
    bool ok = true;
    for (int iy = -S; iy <= S && ok; ++iy) {
        const int ky = abs(iy * s + oy);
        for (int ix = -S; ix <= S; ++ix) {
            const int kx = abs(ix * s + ox);
            if (kx >= static_cast<int>(width) || ky >= static_cast<int>(width)) { ok = false; break; }
            sum += K[static_cast<uint64_t>(kx) * width + ky].x;
        }
    }
    table[i] = ok ? sum : 0.0f;
}

// Gather the taps of every (plane, off_x, off_y) into one contiguous run in
// the order the tap-parallel kernels walk them: f = (ix+S)*(2S+1) + (iy+S),
// value K[|ix*s + off_x|][|iy*s + off_y|]; zero beyond the plane's (2S+1)^2.
__global__ void k_compact(const float2 *kernels, const int32_t *supports, uint32_t planes,
                          uint32_t width, uint32_t sampling, int32_t R, uint32_t stride,
                          float2 *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint32_t n_off = static_cast<uint32_t>(2 * R + 1);
    const uint64_t total = static_cast<uint64_t>(planes) * n_off * n_off * stride;
    if (i >= total) return;
    const uint32_t f = static_cast<uint32_t>(i % stride);
    const uint64_t po = i / stride;
    const uint32_t plane = static_cast<uint32_t>(po / (n_off * n_off));
    const uint32_t off = static_cast<uint32_t>(po - static_cast<uint64_t>(plane) * n_off * n_off);
    const int ox = static_cast<int>(off / n_off) - R, oy = static_cast<int>(off % n_off) - R;
    const int S = supports[plane];
    const int side = 2 * S + 1;
    float2 v = make_float2(0.0f, 0.0f);
    if (static_cast<int>(f) < side * side) {
        const int ix = static_cast<int>(f) / side - S, iy = static_cast<int>(f) % side - S;
        const int s = static_cast<int>(sampling);
        const int kx = abs(ix * s + ox), ky = abs(iy * s + oy);
        if (kx < static_cast<int>(width) && ky < static_cast<int>(width))
            v = kernels[static_cast<uint64_t>(plane) * width * width + static_cast<uint64_t>(kx) * width + ky];
    }
    out[i] = v;
}

static unsigned nb(uint64_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

namespace {

// Screen of plane k and its forward FFT into `screen` (conv x conv).  One
// plane per transform, so that a plane's values do not depend on which
// planes are transformed with it.
void screen_fft(chip_ctx &ctx, const WkState &wk, uint32_t k, float2 *screen, hipStream_t st) {
    const uint64_t cells = static_cast<uint64_t>(wk.conv) * wk.conv;
    CHIP_HIP(hipMemsetAsync(screen, 0, cells * sizeof(float2), st));
    const ScreenParams sp{wk.conv, wk.inner, wk.s0, wk.s1, wk.w_scale, wk.planes, k};
    const uint64_t work = static_cast<uint64_t>(wk.inner) * wk.inner;
    hipLaunchKernelGGL(k_screen, dim3(nb(work, 256)), dim3(256), 0, st, sp, wk.d_corr.as<float>(), screen, 1u);
    CHIP_HIP(hipGetLastError());
    hipfftHandle h = ctx.fft.get(wk.conv, wk.conv, 1, HIPFFT_C2C, st);
    CHIP_FFT(hipfftExecC2C(h, reinterpret_cast<hipfftComplex *>(screen), reinterpret_cast<hipfftComplex *>(screen),
                           HIPFFT_FORWARD));
}

// Tables of planes [lo, hi): kernels (hi - lo) x width^2 (the cropped quarter
// planes / plane-zero peak / plane-zero integral, zero padded), norm (may be
// NULL) (hi - lo) x (2R+1)^2, compact (may be NULL) (hi - lo) x (2R+1)^2 x
// stride.  d_sup: supports of every plane.  screen: one conv^2 plane.
void wk_tables(chip_ctx &ctx, const WkState &wk, const int32_t *d_sup, uint32_t lo, uint32_t hi, float2 *screen,
               float2 *kernels, float *norm, float2 *compact, hipStream_t st) {
    const uint32_t n = hi - lo;
    const uint64_t kcells = static_cast<uint64_t>(wk.width) * wk.width;
    CHIP_HIP(hipMemsetAsync(kernels, 0, kcells * n * sizeof(float2), st));
    for (uint32_t k = lo; k < hi; ++k) {
        screen_fft(ctx, wk, k, screen, st);
        const uint32_t c = wk.crop[k];
        hipLaunchKernelGGL(k_crop, dim3(nb(static_cast<uint64_t>(c) * c, 256)), dim3(256), 0, st, screen, wk.conv,
                           c, wk.width, wk.peak0, kernels + (k - lo) * kcells);
        CHIP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_scale, dim3(nb(kcells * n, 256)), dim3(256), 0, st, kernels, kcells * n, wk.pb);
    CHIP_HIP(hipGetLastError());
    const uint32_t axis = static_cast<uint32_t>(2 * wk.R + 1);
    if (norm) {
        hipLaunchKernelGGL(k_norm_table, dim3(nb(static_cast<uint64_t>(n) * axis * axis, 128)), dim3(128), 0, st,
                           kernels, d_sup + lo, n, wk.width, wk.sampling, wk.R, norm);
        CHIP_HIP(hipGetLastError());
    }
    if (compact) {
        const uint64_t n_compact = static_cast<uint64_t>(n) * axis * axis * wk.stride;
        hipLaunchKernelGGL(k_compact, dim3(nb(n_compact, 256)), dim3(256), 0, st, kernels, d_sup + lo, n, wk.width,
                           wk.sampling, wk.R, wk.stride, compact);
        CHIP_HIP(hipGetLastError());
    }
}

uint64_t wk_plane_bytes(const WkState &wk) {
    const uint64_t axis = static_cast<uint64_t>(2 * wk.R + 1);
    return (static_cast<uint64_t>(wk.width) * wk.width + axis * axis * wk.stride) * sizeof(float2);
}

}  // namespace

void build_wproject(chip_gridder &g, double max_abs_w, uint32_t explicit_planes,
                    uint32_t sampling) {
    hipStream_t st = g.ctx->stream;
    const chip_geometry &geo = g.geom;
    uint32_t planes = explicit_planes ? explicit_planes
                                      : std::max<uint32_t>(suggested_w_plane_count(geo, max_abs_w), 1u);
    if (planes > 2047) fail(CHIP_ERR_INVALID, "at most 2047 W planes");
    const uint32_t s = sampling ? sampling : (planes > 1 ? 4u : 1u);  // gridder.rs:3413
    if (s > 127) fail(CHIP_ERR_INVALID, "W sampling must be at most 127");
    WkState wk;
    wk.max_abs_w = max_abs_w;
    wk.explicit_planes = explicit_planes;
    wk.sampling_arg = sampling;
    wk.planes = planes;
    wk.sampling = s;
    wk.conv = static_cast<uint32_t>(std::max(g.gx, g.gy));
    wk.inner = std::max<uint32_t>(wk.conv / s, 2u);
    wk.quarter = wk.conv / 2 - 1;
    if (wk.quarter < 2) fail(CHIP_ERR_INVALID, "grid too small for W projection");
    const std::vector<float> corr = gridder_correction_axis(wk.inner);
    wk.s0 = std::fabs(geo.cell_x) * s * static_cast<double>(g.gx) / wk.conv;
    wk.s1 = std::fabs(geo.cell_y) * s * static_cast<double>(g.gy) / wk.conv;
    const double maxw = w_max_uvw(geo, max_abs_w, explicit_planes != 0);
    wk.w_scale = 1.0;
    if (planes > 1 && std::isfinite(maxw) && maxw > 0.0)
        wk.w_scale = static_cast<double>((planes - 1) * (planes - 1)) / maxw;
    wk.R = static_cast<int32_t>((s + 1) / 2);
    const int32_t R = wk.R;
    const uint32_t conv = wk.conv, quarter = wk.quarter;
    wk.d_corr.alloc(wk.inner * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(wk.d_corr.p, corr.data(), wk.d_corr.bytes, hipMemcpyHostToDevice, st));

    // Pass 1, plane by plane: plane-zero peak, each plane's support from the
    // 1e-3 threshold on the two axes (gridder.rs:3832-3855), its crop, and
    // the plane-zero integral at integer taps (pb_sum, gridder.rs:3991-4002,
    // summed on the host in the reference order).
    const uint64_t plane_bytes = static_cast<uint64_t>(conv) * conv * sizeof(float2);
    DevBuf screen(plane_bytes);
    DevBuf d_peak(sizeof(float));
    DevBuf d_mag(static_cast<uint64_t>(quarter) * sizeof(float));
    std::vector<uint32_t> support(planes);
    wk.crop.assign(planes, 0);
    std::vector<float> mag(quarter);
    for (uint32_t k = 0; k < planes; ++k) {
        screen_fft(*g.ctx, wk, k, screen.as<float2>(), st);
        if (k == 0) {
            hipLaunchKernelGGL(k_peak, dim3(1), dim3(64), 0, st, screen.as<float2>(), d_peak.as<float>());
            CHIP_HIP(hipGetLastError());
            CHIP_HIP(hipMemcpyAsync(&wk.peak0, d_peak.p, sizeof(float), hipMemcpyDeviceToHost, st));
            CHIP_HIP(hipStreamSynchronize(st));
            if (!(wk.peak0 > 0.0f)) fail(CHIP_ERR_NORMALIZATION, "wproject plane-zero kernel peak is zero");
        }
        hipLaunchKernelGGL(k_axis_mag, dim3(nb(quarter, 256)), dim3(256), 0, st, screen.as<float2>(), conv, quarter,
                           1u, d_peak.as<float>(), d_mag.as<float>());
        CHIP_HIP(hipGetLastError());
        CHIP_HIP(hipMemcpyAsync(mag.data(), d_mag.p, static_cast<uint64_t>(quarter) * sizeof(float),
                                hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        support[k] = find_w_support(mag, conv, s);
        uint64_t cw = static_cast<uint64_t>(support[k]) * s + R + 1;
        cw = std::max<uint64_t>(std::min<uint64_t>(cw, quarter), 1);
        wk.crop[k] = static_cast<uint32_t>(cw);
        // Taps read |i*s + off| <= support*s + R; the crop must hold them.
        if (static_cast<uint64_t>(support[k]) * s + R + 1 > cw)
            fail(CHIP_ERR_INVALID, "W kernel support exceeds the stored quarter plane");
        if (k == 0) {
            DevBuf c0(cw * cw * sizeof(float2));
            hipLaunchKernelGGL(k_crop, dim3(nb(cw * cw, 256)), dim3(256), 0, st, screen.as<float2>(), conv,
                               static_cast<uint32_t>(cw), static_cast<uint32_t>(cw), wk.peak0, c0.as<float2>());
            CHIP_HIP(hipGetLastError());
            std::vector<float2> k0(cw * cw);
            CHIP_HIP(hipMemcpyAsync(k0.data(), c0.p, c0.bytes, hipMemcpyDeviceToHost, st));
            CHIP_HIP(hipStreamSynchronize(st));
            const int64_t S = support[0];
            float pb = 0.0f;
            for (int64_t iy = -S; iy <= S; ++iy) {
                const uint64_t ky = static_cast<uint64_t>(std::llabs(iy * static_cast<int64_t>(s)));
                for (int64_t ix = -S; ix <= S; ++ix) {
                    const uint64_t kx = static_cast<uint64_t>(std::llabs(ix * static_cast<int64_t>(s)));
                    pb += k0[kx * cw + ky].x;
                }
            }
            if (!(std::isfinite(pb) && pb > 0.0f))
                fail(CHIP_ERR_NORMALIZATION, "wproject plane-zero kernel integral is non-positive");
            wk.pb = pb;
        }
    }
    uint32_t width = 0, max_support = 0;
    for (uint32_t k = 0; k < planes; ++k) {
        width = std::max(width, wk.crop[k]);
        max_support = std::max(max_support, support[k]);
    }
    wk.width = width;
    wk.stride = (2 * max_support + 1) * (2 * max_support + 1);

    DevBuf d_sup(planes * sizeof(int32_t));
    std::vector<int32_t> sup32(support.begin(), support.end());
    CHIP_HIP(hipMemcpyAsync(d_sup.p, sup32.data(), d_sup.bytes, hipMemcpyHostToDevice, st));
    const uint32_t axis = static_cast<uint32_t>(2 * R + 1);
    DevBuf d_norm(static_cast<uint64_t>(planes) * axis * axis * sizeof(float));

    // Pass 2: the tables, resident or (above the budget) one chunk at a time
    // for the normalisation table only.
    const uint64_t per_plane = wk_plane_bytes(wk);
    const uint64_t resident = per_plane * planes;
    uint32_t chunk = 0;
    if (g.wk_budget && resident > g.wk_budget) {
        const uint64_t avail = g.wk_budget > plane_bytes ? g.wk_budget - plane_bytes : 0;
        chunk = static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>(avail / per_plane, 1), planes));
    }
    DevBuf kernels, compact;
    if (!chunk) {
        kernels.alloc(static_cast<uint64_t>(width) * width * planes * sizeof(float2));
        compact.alloc(static_cast<uint64_t>(planes) * axis * axis * wk.stride * sizeof(float2));
        wk_tables(*g.ctx, wk, d_sup.as<int32_t>(), 0, planes, screen.as<float2>(), kernels.as<float2>(),
                  d_norm.as<float>(), compact.as<float2>(), st);
    } else {
        DevBuf kc(static_cast<uint64_t>(width) * width * chunk * sizeof(float2));
        for (uint32_t lo = 0; lo < planes; lo += chunk) {
            const uint32_t hi = std::min(planes, lo + chunk);
            wk_tables(*g.ctx, wk, d_sup.as<int32_t>(), lo, hi, screen.as<float2>(), kc.as<float2>(),
                      d_norm.as<float>() + static_cast<uint64_t>(lo) * axis * axis, nullptr, st);
        }
        CHIP_HIP(hipStreamSynchronize(st));
    }
    std::vector<float> norm(static_cast<uint64_t>(planes) * axis * axis);
    CHIP_HIP(hipMemcpyAsync(norm.data(), d_norm.p, d_norm.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    screen.release();

    g.wproject = true;
    g.planes = planes;
    g.sampling = s;
    g.width = width;
    g.max_support = max_support;
    g.norm_radius = R;
    g.w_scale = wk.w_scale;
    g.supports = sup32;
    g.norm_table = norm;
    g.d_kernels = std::move(kernels);
    g.d_compact = std::move(compact);
    g.tap_stride = wk.stride;
    g.d_supports = std::move(d_sup);
    g.d_norm_table = std::move(d_norm);
    g.wk_chunk = chunk;
    g.wk = std::move(wk);
    g.sinc = sinc_axis(std::max(g.gx, g.gy), s);
    g.sinc_x = sinc_axis(g.gx, s);
    g.sinc_y = sinc_axis(g.gy, s);
    g.d_sinc.alloc(g.sinc.size() * sizeof(float));
    g.d_sinc_x.alloc(g.sinc_x.size() * sizeof(float));
    g.d_sinc_y.alloc(g.sinc_y.size() * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc.p, g.sinc.data(), g.d_sinc.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc_x.p, g.sinc_x.data(), g.d_sinc_x.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc_y.p, g.sinc_y.data(), g.d_sinc_y.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipStreamSynchronize(st));
}

// WProjector over tables the caller built (the reference's own projector:
// WProjector::flattened_kernel_weights, kernel_support, w_scale,
// gridder.rs:3640-3678): the quarter-plane kernels are uploaded as they are,
// the per-offset gathered taps are made on the device, and the per-(plane,
// offset) normalisation table is taken from the caller or, when NULL,
// computed on the device from the kernels (compute_w_project_sample_
// normalization, gridder.rs:3782-3799).
void attach_wproject_tables(chip_gridder &g, uint32_t planes, uint32_t sampling, double w_scale, uint32_t width,
                            const int32_t *supports, const float *kernels_host, const float *norm_host) {
    if (g.wproject) fail(CHIP_ERR_INVALID, "W projection already attached");
    if (planes == 0 || planes > 2047) fail(CHIP_ERR_INVALID, "planes must be 1..2047");
    if (sampling == 0 || sampling > 64) fail(CHIP_ERR_INVALID, "sampling must be 1..64");
    if (!(std::isfinite(w_scale) && w_scale > 0.0)) fail(CHIP_ERR_INVALID, "w_scale must be finite and positive");
    if (!supports || !kernels_host) fail(CHIP_ERR_INVALID, "NULL supports or kernels");
    if (width == 0 || width > 4096) fail(CHIP_ERR_INVALID, "kernel width must be 1..4096");
    const int32_t R = static_cast<int32_t>((sampling + 1) / 2);
    uint32_t max_support = 0;
    for (uint32_t k = 0; k < planes; ++k) {
        if (supports[k] <= 0 || supports[k] > 63) fail(CHIP_ERR_INVALID, "W supports must be 1..63");
        // taps read |i*s + off| <= support*s + R (gridder.rs:3686-3692)
        if (static_cast<uint64_t>(supports[k]) * sampling + R >= width)
            fail(CHIP_ERR_INVALID, "a W support reaches beyond the kernel width");
        max_support = std::max(max_support, static_cast<uint32_t>(supports[k]));
    }
    hipStream_t st = g.ctx->stream;
    const uint32_t axis = static_cast<uint32_t>(2 * R + 1);
    const uint32_t stride = (2 * max_support + 1) * (2 * max_support + 1);
    const uint64_t kcells = static_cast<uint64_t>(width) * width;
    DevBuf kernels(kcells * planes * sizeof(float2));
    DevBuf compact(static_cast<uint64_t>(planes) * axis * axis * stride * sizeof(float2));
    DevBuf d_sup(planes * sizeof(int32_t));
    DevBuf d_norm(static_cast<uint64_t>(planes) * axis * axis * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(kernels.p, kernels_host, kernels.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(d_sup.p, supports, d_sup.bytes, hipMemcpyHostToDevice, st));
    if (norm_host) {
        CHIP_HIP(hipMemcpyAsync(d_norm.p, norm_host, d_norm.bytes, hipMemcpyHostToDevice, st));
    } else {
        hipLaunchKernelGGL(k_norm_table, dim3(nb(static_cast<uint64_t>(planes) * axis * axis, 128)), dim3(128), 0, st,
                           kernels.as<float2>(), d_sup.as<int32_t>(), planes, width, sampling, R, d_norm.as<float>());
        CHIP_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_compact, dim3(nb(static_cast<uint64_t>(planes) * axis * axis * stride, 256)), dim3(256), 0, st,
                       kernels.as<float2>(), d_sup.as<int32_t>(), planes, width, sampling, R, stride,
                       compact.as<float2>());
    CHIP_HIP(hipGetLastError());
    std::vector<float> norm(static_cast<uint64_t>(planes) * axis * axis);
    CHIP_HIP(hipMemcpyAsync(norm.data(), d_norm.p, d_norm.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    g.wproject = true;
    g.planes = planes;
    g.sampling = sampling;
    g.width = width;
    g.max_support = max_support;
    g.norm_radius = R;
    g.w_scale = w_scale;
    g.supports.assign(supports, supports + planes);
    g.norm_table = norm;
    g.d_kernels = std::move(kernels);
    g.d_compact = std::move(compact);
    g.tap_stride = stride;
    g.d_supports = std::move(d_sup);
    g.d_norm_table = std::move(d_norm);
    g.wk_chunk = 0;
    g.wk_external = true;
    g.wk = WkState{};
    g.wk.planes = planes;
    g.wk.sampling = sampling;
    g.wk.R = R;
    g.wk.width = width;
    g.wk.stride = stride;
    g.wk.w_scale = w_scale;
    g.sinc = sinc_axis(std::max(g.gx, g.gy), sampling);
    g.sinc_x = sinc_axis(g.gx, sampling);
    g.sinc_y = sinc_axis(g.gy, sampling);
    g.d_sinc.alloc(g.sinc.size() * sizeof(float));
    g.d_sinc_x.alloc(g.sinc_x.size() * sizeof(float));
    g.d_sinc_y.alloc(g.sinc_y.size() * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc.p, g.sinc.data(), g.d_sinc.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc_x.p, g.sinc_x.data(), g.d_sinc_x.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(g.d_sinc_y.p, g.sinc_y.data(), g.d_sinc_y.bytes, hipMemcpyHostToDevice, st));
    CHIP_HIP(hipStreamSynchronize(st));
}

void wk_prepare(const chip_gridder &g, WkScratch &ws) {
    const WkState &wk = g.wk;
    const uint64_t axis = static_cast<uint64_t>(2 * wk.R + 1);
    ws.screen.alloc(static_cast<uint64_t>(wk.conv) * wk.conv * sizeof(float2));
    ws.kernels.alloc(static_cast<uint64_t>(wk.width) * wk.width * g.wk_chunk * sizeof(float2));
    ws.compact.alloc(axis * axis * wk.stride * g.wk_chunk * sizeof(float2));
}

void wk_fill(chip_gridder &g, WkScratch &ws, uint32_t lo, uint32_t hi, GridParams &gp) {
    if (hi <= lo || hi - lo > g.wk_chunk || hi > g.planes) fail(CHIP_ERR_INVALID, "bad W plane chunk");
    wk_tables(*g.ctx, g.wk, g.d_supports.as<int32_t>(), lo, hi, ws.screen.as<float2>(), ws.kernels.as<float2>(),
              nullptr, ws.compact.as<float2>(), g.ctx->stream);
    gp.kernels = ws.kernels.as<float2>();
    gp.compact = ws.compact.as<float2>();
    gp.plane_lo = lo;
    gp.plane_hi = hi;
}

void wk_all_kernels(const chip_gridder &g, float2 *kernels_dev) {
    hipStream_t st = g.ctx->stream;
    if (!g.wk_chunk) {
        CHIP_HIP(hipMemcpyAsync(kernels_dev, g.d_kernels.p, g.d_kernels.bytes, hipMemcpyDeviceToDevice, st));
    } else {
        DevBuf screen(static_cast<uint64_t>(g.wk.conv) * g.wk.conv * sizeof(float2));
        wk_tables(*g.ctx, g.wk, g.d_supports.as<int32_t>(), 0, g.planes, screen.as<float2>(), kernels_dev, nullptr,
                  nullptr, st);
    }
    CHIP_HIP(hipStreamSynchronize(st));
}

}  // namespace chip
