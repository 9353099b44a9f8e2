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
                       const float *peak, float2 *stage) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= crop * crop) return;
    const uint32_t x = i / crop, y = i - x * crop;
    const float p = *peak;
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

void build_wproject(chip_gridder &g, double max_abs_w, uint32_t explicit_planes,
                    uint32_t sampling) {
    hipStream_t st = g.ctx->stream;
    const chip_geometry &geo = g.geom;
    uint32_t planes = explicit_planes ? explicit_planes
                                      : std::max<uint32_t>(suggested_w_plane_count(geo, max_abs_w), 1u);
    if (planes > 2047) fail(CHIP_ERR_INVALID, "at most 2047 W planes");
    const uint32_t s = sampling ? sampling : (planes > 1 ? 4u : 1u);  // gridder.rs:3413
    if (s > 127) fail(CHIP_ERR_INVALID, "W sampling must be at most 127");
    const uint32_t conv = static_cast<uint32_t>(std::max(g.gx, g.gy));
    const uint32_t inner = std::max<uint32_t>(conv / s, 2u);
    const uint32_t quarter = conv / 2 - 1;
    if (quarter < 2) fail(CHIP_ERR_INVALID, "grid too small for W projection");
    const std::vector<float> corr = gridder_correction_axis(inner);
    const double s0 = std::fabs(geo.cell_x) * s * static_cast<double>(g.gx) / conv;
    const double s1 = std::fabs(geo.cell_y) * s * static_cast<double>(g.gy) / conv;
    const double maxw = w_max_uvw(geo, max_abs_w, explicit_planes != 0);
    double w_scale = 1.0;
    if (planes > 1 && std::isfinite(maxw) && maxw > 0.0)
        w_scale = static_cast<double>((planes - 1) * (planes - 1)) / maxw;
    const int32_t R = static_cast<int32_t>((s + 1) / 2);

    // Planes per FFT batch: keep the screens under ~4 GiB.
    const uint64_t plane_bytes = static_cast<uint64_t>(conv) * conv * sizeof(float2);
    uint32_t per_batch = static_cast<uint32_t>(std::max<uint64_t>(1, (4ull << 30) / plane_bytes));
    per_batch = std::min(per_batch, planes);
    DevBuf screens(plane_bytes * per_batch);
    DevBuf d_corr(inner * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(d_corr.p, corr.data(), d_corr.bytes, hipMemcpyHostToDevice, st));
    DevBuf d_peak(sizeof(float));
    DevBuf d_mag(static_cast<uint64_t>(quarter) * per_batch * sizeof(float));

    // Staging: each plane cropped to its own width, at most quarter.
    const uint32_t stage_w_max = quarter;
    std::vector<uint32_t> crop(planes), support(planes);
    std::vector<DevBuf> stage(planes);
    std::vector<float> mag(static_cast<uint64_t>(quarter) * per_batch);
    ScreenParams sp{conv, inner, s0, s1, w_scale, planes, 0};
    for (uint32_t first = 0; first < planes; first += per_batch) {
        const uint32_t nbat = std::min(per_batch, planes - first);
        CHIP_HIP(hipMemsetAsync(screens.p, 0, plane_bytes * nbat, st));
        sp.first = first;
        const uint64_t work = static_cast<uint64_t>(inner) * inner * nbat;
        hipLaunchKernelGGL(k_screen, dim3(nb(work, 256)), dim3(256), 0, st, sp, d_corr.as<float>(),
                           screens.as<float2>(), nbat);
        CHIP_HIP(hipGetLastError());
        hipfftHandle h = g.ctx->fft.get(conv, conv, nbat, HIPFFT_C2C, st);
        CHIP_FFT(hipfftExecC2C(h, screens.as<hipfftComplex>(), screens.as<hipfftComplex>(),
                               HIPFFT_FORWARD));
        if (first == 0) {
            hipLaunchKernelGGL(k_peak, dim3(1), dim3(64), 0, st, screens.as<float2>(), d_peak.as<float>());
            CHIP_HIP(hipGetLastError());
            float peak = 0.0f;
            CHIP_HIP(hipMemcpyAsync(&peak, d_peak.p, sizeof(float), hipMemcpyDeviceToHost, st));
            CHIP_HIP(hipStreamSynchronize(st));
            if (!(peak > 0.0f)) fail(CHIP_ERR_NORMALIZATION, "wproject plane-zero kernel peak is zero");
        }
        hipLaunchKernelGGL(k_axis_mag, dim3(nb(static_cast<uint64_t>(quarter) * nbat, 256)), dim3(256), 0,
                           st, screens.as<float2>(), conv, quarter, nbat, d_peak.as<float>(),
                           d_mag.as<float>());
        CHIP_HIP(hipGetLastError());
        CHIP_HIP(hipMemcpyAsync(mag.data(), d_mag.p, static_cast<uint64_t>(quarter) * nbat * sizeof(float),
                                hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        for (uint32_t b = 0; b < nbat; ++b) {
            const uint32_t k = first + b;
            std::vector<float> m(mag.begin() + static_cast<uint64_t>(b) * quarter,
                                 mag.begin() + static_cast<uint64_t>(b + 1) * quarter);
            support[k] = find_w_support(m, conv, s);
            uint64_t cw = static_cast<uint64_t>(support[k]) * s + R + 1;
            cw = std::max<uint64_t>(std::min<uint64_t>(cw, quarter), 1);
            crop[k] = static_cast<uint32_t>(cw);
            // Taps read |i*s + off| <= support*s + R; the crop must hold them.
            if (static_cast<uint64_t>(support[k]) * s + R + 1 > cw)
                fail(CHIP_ERR_INVALID, "W kernel support exceeds the stored quarter plane");
            stage[k].alloc(cw * cw * sizeof(float2));
            hipLaunchKernelGGL(k_crop, dim3(nb(cw * cw, 256)), dim3(256), 0, st,
                               screens.as<float2>() + static_cast<uint64_t>(b) * conv * conv, conv,
                               crop[k], crop[k], d_peak.as<float>(), stage[k].as<float2>());
            CHIP_HIP(hipGetLastError());
        }
    }
    (void)stage_w_max;
    screens.release();
    uint32_t width = 0, max_support = 0;
    for (uint32_t k = 0; k < planes; ++k) {
        width = std::max(width, crop[k]);
        max_support = std::max(max_support, support[k]);
    }
    const uint64_t kcells = static_cast<uint64_t>(width) * width;
    DevBuf kernels(kcells * planes * sizeof(float2));
    CHIP_HIP(hipMemsetAsync(kernels.p, 0, kernels.bytes, st));
    for (uint32_t k = 0; k < planes; ++k) {
        CHIP_HIP(hipMemcpy2DAsync(kernels.as<float2>() + k * kcells, width * sizeof(float2),
                                  stage[k].p, crop[k] * sizeof(float2), crop[k] * sizeof(float2),
                                  crop[k], hipMemcpyDeviceToDevice, st));
    }
    // pb_sum: plane-zero sum of Re at integer taps (gridder.rs:3991-4002),
    // summed on the host in the reference order from a copy of plane 0.
    std::vector<float2> k0(kcells);
    CHIP_HIP(hipMemcpyAsync(k0.data(), kernels.p, kcells * sizeof(float2), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    stage.clear();
    float pb = 0.0f;
    {
        const int64_t S = support[0];
        for (int64_t iy = -S; iy <= S; ++iy) {
            const uint64_t ky = static_cast<uint64_t>(std::llabs(iy * static_cast<int64_t>(s)));
            for (int64_t ix = -S; ix <= S; ++ix) {
                const uint64_t kx = static_cast<uint64_t>(std::llabs(ix * static_cast<int64_t>(s)));
                pb += k0[kx * width + ky].x;
            }
        }
    }
    if (!(std::isfinite(pb) && pb > 0.0f))
        fail(CHIP_ERR_NORMALIZATION, "wproject plane-zero kernel integral is non-positive");
    hipLaunchKernelGGL(k_scale, dim3(nb(kcells * planes, 256)), dim3(256), 0, st,
                       kernels.as<float2>(), kcells * planes, pb);
    CHIP_HIP(hipGetLastError());

    DevBuf d_sup(planes * sizeof(int32_t));
    std::vector<int32_t> sup32(support.begin(), support.end());
    CHIP_HIP(hipMemcpyAsync(d_sup.p, sup32.data(), d_sup.bytes, hipMemcpyHostToDevice, st));
    const uint32_t axis = static_cast<uint32_t>(2 * R + 1);
    DevBuf d_norm(static_cast<uint64_t>(planes) * axis * axis * sizeof(float));
    hipLaunchKernelGGL(k_norm_table, dim3(nb(static_cast<uint64_t>(planes) * axis * axis, 128)), dim3(128), 0,
                       st, kernels.as<float2>(), d_sup.as<int32_t>(), planes, width, s, R,
                       d_norm.as<float>());
    CHIP_HIP(hipGetLastError());
    std::vector<float> norm(static_cast<uint64_t>(planes) * axis * axis);
    CHIP_HIP(hipMemcpyAsync(norm.data(), d_norm.p, d_norm.bytes, hipMemcpyDeviceToHost, st));

    const uint32_t stride = (2 * max_support + 1) * (2 * max_support + 1);
    const uint64_t n_compact = static_cast<uint64_t>(planes) * axis * axis * stride;
    DevBuf compact(n_compact * sizeof(float2));
    hipLaunchKernelGGL(k_compact, dim3(nb(n_compact, 256)), dim3(256), 0, st, kernels.as<float2>(),
                       d_sup.as<int32_t>(), planes, width, s, R, stride, compact.as<float2>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));

    g.wproject = true;
    g.planes = planes;
    g.sampling = s;
    g.width = width;
    g.max_support = max_support;
    g.norm_radius = R;
    g.w_scale = w_scale;
    g.supports = sup32;
    g.norm_table = norm;
    g.d_kernels = std::move(kernels);
    g.d_compact = std::move(compact);
    g.tap_stride = stride;
    g.d_supports = std::move(d_sup);
    g.d_norm_table = std::move(d_norm);
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

}  // namespace chip
