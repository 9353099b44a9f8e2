// uv-grid <-> image transforms and the dirty-product finisher.
//
// Centered transforms (fft.rs:27-100, :630-700) on even grids are done with
// the checkerboard identity: for even N,
//   centered_ifft(G)[x] = (-1)^(x + N/2) * IFFT((-1)^k G[k])[x] / N
// (the same fact the reference uses in inverse_fft2_scale_centered_frequency_f64,
// fft.rs:820-839), so the shifts become sign flips fused into the pack and the
// crop kernels and rocFFT (through hipFFT) runs an unshifted in-place C2C.
#include <cmath>

#include "device.hpp"

namespace chip {

hipfftHandle FftCache::get(uint64_t gx, uint64_t gy, uint64_t batch, hipfftType type,
                           hipStream_t s) {
    auto key = std::make_tuple(gx, gy, batch, static_cast<int>(type));
    auto it = plans.find(key);
    if (it != plans.end()) {
        CHIP_FFT(hipfftSetStream(it->second, s));  // the context's stream may have changed
        return it->second;
    }
    hipfftHandle h;
    CHIP_FFT(hipfftCreate(&h));
    int n[2] = {static_cast<int>(gx), static_cast<int>(gy)};
    size_t work = 0;
    CHIP_FFT(hipfftMakePlanMany(h, 2, n, nullptr, 1, static_cast<int>(gx * gy), nullptr, 1,
                                static_cast<int>(gx * gy), type, static_cast<int>(batch), &work));
    CHIP_FFT(hipfftSetStream(h, s));
    plans[key] = h;
    return h;
}

void FftCache::clear() {
    for (auto &kv : plans) hipfftDestroy(kv.second);
    plans.clear();
}

template <typename T> struct C2;
template <> struct C2<float> { using t = float2; };
template <> struct C2<double> { using t = double2; };

// grid *= (-1)^(x+y), all batches.
template <typename T>
__global__ void k_checker(T *grid, uint64_t total, uint32_t gy, uint64_t cells) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t c = i % cells;
    const uint64_t x = c / gy, y = c - x * gy;
    if ((x + y) & 1) {
        grid[2 * i] = -grid[2 * i];
        grid[2 * i + 1] = -grid[2 * i + 1];
    }
}

// Post-transform sign of the centered transforms for a full grid:
// (-1)^(x + y + gx/2 + gy/2).  Used after the forward model transform.
template <typename T>
__global__ void k_checker_post(T *grid, uint64_t total, uint32_t gy, uint64_t cells, uint32_t base) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const uint64_t c = i % cells;
    const uint64_t x = c / gy, y = c - x * gy;
    if ((x + y + base) & 1) {
        grid[2 * i] = -grid[2 * i];
        grid[2 * i + 1] = -grid[2 * i + 1];
    }
}

struct FinishParams {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t base;  // gx/2 + gy/2
    uint32_t wproject;
    uint64_t cells;
    const float *corr_x, *corr_y, *sinc;
};

// Crop + correction of one product (gridder.rs:1646-1657 standard,
// :1685-1715 W-projection) followed by the first normalisation step:
//   standard: v / sumwt               (lib.rs:49080)
//   W:        2 v / sumwt             (lib.rs:71027)
// The raw transform is unnormalised; the 1/(gx*gy) scale of the centered
// inverse FFT is applied here in the transform's precision.
template <typename T>
__global__ void k_crop_correct(FinishParams fp, const T *raw, uint64_t batch, const float *nsum32,
                               float *out, unsigned int *peak_bits) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t npix = static_cast<uint64_t>(fp.nx) * fp.ny;
    float v = 0.0f;
    uint64_t b = 0;
    if (i < npix * batch) {
        b = i / npix;
        const uint64_t pix = i - b * npix;
        const uint32_t x = static_cast<uint32_t>(pix / fp.ny), y = static_cast<uint32_t>(pix % fp.ny);
        const uint32_t gxx = fp.blc_x + x, gyy = fp.blc_y + y;
        const uint64_t gi = b * fp.cells + static_cast<uint64_t>(gxx) * fp.gy + gyy;
        const T scale = T(1) / static_cast<T>(static_cast<uint64_t>(fp.gx) * fp.gy);
        T re = raw[2 * gi] * scale;
        if ((gxx + gyy + fp.base) & 1) re = -re;
        const float factor = fp.corr_x[gxx] * fp.corr_y[gyy];
        if (!fp.wproject) {
            v = static_cast<float>(re) * factor;
            v = v / nsum32[b];
        } else {
            const float sf = fp.sinc[gxx] * fp.sinc[gyy];
            v = fabsf(sf) > 1.0e-6f ? static_cast<float>(re) * (factor / sf) : 0.0f;
            v = 2.0f * v / nsum32[b];
        }
        out[i] = v;
    }
    if (peak_bits) {
        float m = fabsf(v);
        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        // All lanes of a wave share a batch index unless npix is not a
        // multiple of 64; reduce per lane group conservatively.
        const uint64_t b_lane0 = __shfl(b, 0, 64);
        bool same = true;
        for (int o = 32; o > 0; o >>= 1) same = same && (__shfl_xor(b, o, 64) == b_lane0);
        if (same) {
            if ((threadIdx.x & 63) == 0 && i < npix * batch) atomicMax(&peak_bits[b], __float_as_uint(m));
        } else if (i < npix * batch) {
            atomicMax(&peak_bits[b], __float_as_uint(fabsf(v)));
        }
    }
}

// Second normalisation step: / psf_peak (lib.rs:49087, :70885).
__global__ void k_div_peak(float *img, uint64_t npix, uint64_t batch, const float *peak) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= npix * batch) return;
    img[i] = img[i] / peak[i / npix];
}

static unsigned nblk(uint64_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

static FinishParams finish_params(const chip_gridder &g) {
    FinishParams fp{};
    fp.nx = static_cast<uint32_t>(g.geom.nx);
    fp.ny = static_cast<uint32_t>(g.geom.ny);
    fp.gx = static_cast<uint32_t>(g.gx);
    fp.gy = static_cast<uint32_t>(g.gy);
    fp.blc_x = static_cast<uint32_t>(g.blc_x);
    fp.blc_y = static_cast<uint32_t>(g.blc_y);
    fp.base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
    fp.wproject = g.wproject ? 1u : 0u;
    fp.cells = g.gx * g.gy;
    fp.corr_x = g.d_corr_x.as<float>();
    fp.corr_y = g.d_corr_y.as<float>();
    fp.sinc = g.wproject ? g.d_sinc.as<float>() : nullptr;
    return fp;
}

template <typename T>
static void inverse_unshifted(chip_gridder &g, T *grids, uint64_t batch) {
    hipStream_t st = g.ctx->stream;
    const uint64_t total = batch * g.gx * g.gy;
    hipLaunchKernelGGL(k_checker<T>, dim3(nblk(total, 256)), dim3(256), 0, st, grids, total,
                       static_cast<uint32_t>(g.gy), g.gx * g.gy);
    CHIP_HIP(hipGetLastError());
    if (sizeof(T) == 4) {
        hipfftHandle h = g.ctx->fft.get(g.gx, g.gy, batch, HIPFFT_C2C, st);
        CHIP_FFT(hipfftExecC2C(h, reinterpret_cast<hipfftComplex *>(grids),
                               reinterpret_cast<hipfftComplex *>(grids), HIPFFT_BACKWARD));
    } else {
        hipfftHandle h = g.ctx->fft.get(g.gx, g.gy, batch, HIPFFT_Z2Z, st);
        CHIP_FFT(hipfftExecZ2Z(h, reinterpret_cast<hipfftDoubleComplex *>(grids),
                               reinterpret_cast<hipfftDoubleComplex *>(grids), HIPFFT_BACKWARD));
    }
}

template <typename T>
static void finish_t(chip_gridder &g, uint64_t batch, T *psf, T *dirty, const double *nsumwt,
                     float *psf_out, float *res_out, float *peaks) {
    hipStream_t st = g.ctx->stream;
    const FinishParams fp = finish_params(g);
    const uint64_t npix = g.geom.nx * g.geom.ny;
    std::vector<float> ns(batch);
    for (uint64_t b = 0; b < batch; ++b) {
        if (!(nsumwt[b] > 0.0)) fail(CHIP_ERR_NO_SAMPLES, "normalization sum of weights is not positive");
        ns[b] = static_cast<float>(nsumwt[b]);
    }
    DevBuf d_ns(batch * sizeof(float)), d_peak(batch * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(d_ns.p, ns.data(), d_ns.bytes, hipMemcpyHostToDevice, st));
    if (psf) {
        if (!psf_out) fail(CHIP_ERR_INVALID, "PSF grids given without a PSF output");
        inverse_unshifted(g, psf, batch);
        CHIP_HIP(hipMemsetAsync(d_peak.p, 0, d_peak.bytes, st));
        hipLaunchKernelGGL(k_crop_correct<T>, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, fp,
                           psf, batch, d_ns.as<float>(), psf_out, d_peak.as<unsigned int>());
        CHIP_HIP(hipGetLastError());
        CHIP_HIP(hipMemcpyAsync(peaks, d_peak.p, d_peak.bytes, hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        for (uint64_t b = 0; b < batch; ++b)
            if (!(std::isfinite(peaks[b]) && peaks[b] > 0.0f))
                fail(CHIP_ERR_NORMALIZATION, "PSF peak is non-finite or zero");
        hipLaunchKernelGGL(k_div_peak, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, psf_out,
                           npix, batch, d_peak.as<float>());
        CHIP_HIP(hipGetLastError());
    } else {
        for (uint64_t b = 0; b < batch; ++b)
            if (!(std::isfinite(peaks[b]) && peaks[b] > 0.0f))
                fail(CHIP_ERR_NORMALIZATION, "PSF peak is non-finite or zero");
        CHIP_HIP(hipMemcpyAsync(d_peak.p, peaks, d_peak.bytes, hipMemcpyHostToDevice, st));
    }
    if (dirty) {
        if (!res_out) fail(CHIP_ERR_INVALID, "dirty grids given without a residual output");
        inverse_unshifted(g, dirty, batch);
        hipLaunchKernelGGL(k_crop_correct<T>, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, fp,
                           dirty, batch, d_ns.as<float>(), res_out, nullptr);
        CHIP_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_div_peak, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, res_out,
                           npix, batch, d_peak.as<float>());
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
}

void finish_products(chip_gridder &g, uint32_t precision, uint64_t batch, void *psf, void *dirty,
                     const double *nsumwt, float *psf_out, float *res_out, float *peaks) {
    if (g.gx % 2 || g.gy % 2) fail(CHIP_ERR_INVALID, "centered transforms need an even grid");
    if (batch == 0) return;
    if (precision == CHIP_F32)
        finish_t<float>(g, batch, static_cast<float *>(psf), static_cast<float *>(dirty), nsumwt,
                        psf_out, res_out, peaks);
    else if (precision == CHIP_F64)
        finish_t<double>(g, batch, static_cast<double *>(psf), static_cast<double *>(dirty), nsumwt,
                         psf_out, res_out, peaks);
    else
        fail(CHIP_ERR_INVALID, "unknown precision");
}

// Apodised model placed at the image origin, pre-multiplied by the
// checkerboard (gridder.rs:1554-1592).  Factor evaluated left to right in f32:
// cx*cy (standard) or cx*cy*sx*sy (W, per-axis sinc).
template <typename T>
__global__ void k_place_model(FinishParams fp, const float *model, uint64_t batch,
                              const float *sinc_x, const float *sinc_y, T *grid) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t npix = static_cast<uint64_t>(fp.nx) * fp.ny;
    if (i >= npix * batch) return;
    const uint64_t b = i / npix, pix = i - b * npix;
    const uint32_t x = static_cast<uint32_t>(pix / fp.ny), y = static_cast<uint32_t>(pix % fp.ny);
    const uint32_t gxx = fp.blc_x + x, gyy = fp.blc_y + y;
    float f = fp.corr_x[gxx] * fp.corr_y[gyy];
    if (fp.wproject) f = f * sinc_x[gxx] * sinc_y[gyy];
    T v = static_cast<T>(model[i] * f);
    if ((gxx + gyy) & 1) v = -v;
    const uint64_t gi = b * fp.cells + static_cast<uint64_t>(gxx) * fp.gy + gyy;
    grid[2 * gi] = v;
    grid[2 * gi + 1] = T(0);
}

template <typename T>
static void model_to_grid_t(chip_gridder &g, uint64_t batch, const float *model, T *grid) {
    hipStream_t st = g.ctx->stream;
    const FinishParams fp = finish_params(g);
    const uint64_t npix = g.geom.nx * g.geom.ny;
    const uint64_t total = batch * g.gx * g.gy;
    CHIP_HIP(hipMemsetAsync(grid, 0, total * 2 * sizeof(T), st));
    hipLaunchKernelGGL(k_place_model<T>, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, fp, model,
                       batch, g.wproject ? g.d_sinc_x.as<float>() : nullptr,
                       g.wproject ? g.d_sinc_y.as<float>() : nullptr, grid);
    CHIP_HIP(hipGetLastError());
    if (sizeof(T) == 4) {
        hipfftHandle h = g.ctx->fft.get(g.gx, g.gy, batch, HIPFFT_C2C, st);
        CHIP_FFT(hipfftExecC2C(h, reinterpret_cast<hipfftComplex *>(grid),
                               reinterpret_cast<hipfftComplex *>(grid), HIPFFT_FORWARD));
    } else {
        hipfftHandle h = g.ctx->fft.get(g.gx, g.gy, batch, HIPFFT_Z2Z, st);
        CHIP_FFT(hipfftExecZ2Z(h, reinterpret_cast<hipfftDoubleComplex *>(grid),
                               reinterpret_cast<hipfftDoubleComplex *>(grid), HIPFFT_FORWARD));
    }
    hipLaunchKernelGGL(k_checker_post<T>, dim3(nblk(total, 256)), dim3(256), 0, st, grid, total,
                       static_cast<uint32_t>(g.gy), g.gx * g.gy, fp.base);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
}

void model_to_grid(chip_gridder &g, uint32_t precision, uint64_t batch, const float *model,
                   void *grid_out) {
    if (g.gx % 2 || g.gy % 2) fail(CHIP_ERR_INVALID, "centered transforms need an even grid");
    if (!model || !grid_out) fail(CHIP_ERR_INVALID, "model and grid are required");
    if (batch == 0) return;
    if (precision == CHIP_F32)
        model_to_grid_t<float>(g, batch, model, static_cast<float *>(grid_out));
    else if (precision == CHIP_F64)
        model_to_grid_t<double>(g, batch, model, static_cast<double *>(grid_out));
    else
        fail(CHIP_ERR_INVALID, "unknown precision");
}

// Adjoint of model_to_grid: Re(factor * crop(F^H grid)), F^H the unnormalised
// centered inverse transform.
__global__ void k_crop_adjoint(FinishParams fp, const float *raw, uint64_t batch,
                               const float *sinc_x, const float *sinc_y, float *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t npix = static_cast<uint64_t>(fp.nx) * fp.ny;
    if (i >= npix * batch) return;
    const uint64_t b = i / npix, pix = i - b * npix;
    const uint32_t x = static_cast<uint32_t>(pix / fp.ny), y = static_cast<uint32_t>(pix % fp.ny);
    const uint32_t gxx = fp.blc_x + x, gyy = fp.blc_y + y;
    float f = fp.corr_x[gxx] * fp.corr_y[gyy];
    if (fp.wproject) f = f * sinc_x[gxx] * sinc_y[gyy];
    float re = raw[2 * (b * fp.cells + static_cast<uint64_t>(gxx) * fp.gy + gyy)];
    if ((gxx + gyy + fp.base) & 1) re = -re;
    out[i] = re * f;
}

void grid_to_model_adjoint(chip_gridder &g, uint64_t batch, float *grid, float *image_out) {
    if (g.gx % 2 || g.gy % 2) fail(CHIP_ERR_INVALID, "centered transforms need an even grid");
    if (batch == 0) return;
    hipStream_t st = g.ctx->stream;
    inverse_unshifted<float>(g, grid, batch);
    const FinishParams fp = finish_params(g);
    const uint64_t npix = g.geom.nx * g.geom.ny;
    hipLaunchKernelGGL(k_crop_adjoint, dim3(nblk(npix * batch, 256)), dim3(256), 0, st, fp, grid,
                       batch, g.wproject ? g.d_sinc_x.as<float>() : nullptr,
                       g.wproject ? g.d_sinc_y.as<float>() : nullptr, image_out);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
}

}  // namespace chip
