// Hogbom minor cycle on the GPU (run_hogbom_minor_cycle_cpu_with_progress,
// lib.rs:59748-59885), every image plane of a cube in one launch.
//
// One workgroup per plane runs the plane's whole minor cycle, so the
// residual never leaves the device between components.  A component costs
// one pass over the plane that subtracts the previous component's shifted
// PSF and, in the same pass, finds the next peak:
//   peak: largest |value| over unmasked pixels, ties to the smallest y-major
//         key y*nx + x (peak_location_masked_in_bounds_zero_tolerance,
//         lib.rs:74064-74115); the reference seeds its search with the first
//         unmasked pixel in x-major order, so a NaN there is the peak, and a
//         NaN anywhere else never is;
//   stop: minor_cycle_stop_reason (lib.rs:74653-74677) on |peak|;
//   update: component = gain * peak; model[peak] += component; residual -=
//         component * psf shifted to the peak over the overlap window, one
//         f32 product then the subtraction (subtract_shifted_kernel_in_window,
//         lib.rs:74805-74868), so planes are bit-identical to the CPU loop.
// A thread always owns the same pixels (row x, columns y = tid + k*threads),
// so passes need no barrier for the residual itself.
#include <cmath>

#include "device.hpp"

namespace chip {

namespace {

constexpr int kCleanThreads = 1024;

struct HogbomParams {
    uint32_t nx, ny, kx, ky;
    float gain, threshold;
    uint32_t budget;
};

// threshold_reached_with_tolerance, lib.rs:74628-74635
__device__ bool threshold_reached(float peak_abs, float threshold) {
    if (threshold <= 0.0f) return peak_abs <= threshold;
    const float tol = fmaxf(0.01f * threshold, 2.0e-8f);
    return peak_abs <= threshold || fabsf(peak_abs - threshold) <= tol;
}

// minor_cycle_stop_reason, lib.rs:74653-74677 (codes as chip_hogbom)
__device__ int minor_stop(float peak_abs, float threshold, float cyc, float nsig) {
    const bool floor_active = threshold > 0.0f && (0.01f * threshold) < 2.0e-8f && cyc <= threshold;
    if (peak_abs < threshold || (floor_active && threshold_reached(peak_abs, threshold))) return CHIP_CLEAN_GLOBAL;
    if (cyc >= nsig) return peak_abs < cyc ? CHIP_CLEAN_CYCLE : CHIP_CLEAN_BUDGET;
    if (nsig > threshold) return peak_abs < nsig ? CHIP_CLEAN_NSIGMA : CHIP_CLEAN_BUDGET;
    return CHIP_CLEAN_BUDGET;
}

// Candidate (|v|, y-major key); a < 0 = none.  b wins over a when larger,
// or equal with a smaller key.
struct Cand {
    float a;
    uint32_t key;
};
__device__ __forceinline__ bool better(const Cand &b, const Cand &a) {
    return b.a > a.a || (b.a == a.a && b.key < a.key);
}

__global__ __launch_bounds__(kCleanThreads) void k_hogbom(HogbomParams p, float *residual, const float *psf,
                                                         float *model, const uint8_t *mask, const float *cyc_thr,
                                                         const float *nsig_thr, uint32_t *updates, int32_t *stop) {
#pragma clang fp contract(off)
    __shared__ Cand wave_best[kCleanThreads / 64];
    __shared__ uint32_t s_seed;
    __shared__ int s_found;
    __shared__ float s_value;
    __shared__ uint32_t s_px, s_py;
    const uint64_t c = blockIdx.x;
    const uint64_t npix = static_cast<uint64_t>(p.nx) * p.ny;
    float *r = residual + c * npix;
    float *m = model + c * npix;
    const float *k = psf + c * static_cast<uint64_t>(p.kx) * p.ky;
    const uint8_t *mk = mask ? mask + c * npix : nullptr;
    const float cyc = cyc_thr[c], nsig = nsig_thr[c];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nth = blockDim.x;

    // seed of the reference's search: first unmasked pixel in x-major order
    if (tid == 0) s_seed = UINT32_MAX;
    __syncthreads();
    if (mk) {
        uint32_t first = UINT32_MAX;
        for (uint32_t x = 0; x < p.nx && first == UINT32_MAX; ++x)
            for (uint32_t y = tid; y < p.ny; y += nth)
                if (mk[static_cast<uint64_t>(x) * p.ny + y]) {
                    first = x * p.ny + y;
                    break;
                }
        if (first != UINT32_MAX) atomicMin(&s_seed, first);
    } else if (tid == 0) {
        s_seed = 0;
    }
    __syncthreads();
    const uint32_t seed = s_seed;

    // One pass: subtract comp * psf shifted to (px, py) where it overlaps
    // (when sub), then find the peak of the updated plane.
    auto pass = [&](bool sub, float comp, int64_t xmin, int64_t ymin) {
        int64_t x0 = 0, x1 = -1, y0 = 0, y1 = -1;
        if (sub) {
            const int64_t xmax = xmin + p.kx - 1, ymax = ymin + p.ky - 1;
            x0 = xmin > 0 ? xmin : 0;
            y0 = ymin > 0 ? ymin : 0;
            x1 = xmax < static_cast<int64_t>(p.nx) - 1 ? xmax : static_cast<int64_t>(p.nx) - 1;
            y1 = ymax < static_cast<int64_t>(p.ny) - 1 ? ymax : static_cast<int64_t>(p.ny) - 1;
        }
        Cand best{-1.0f, UINT32_MAX};
        for (uint32_t x = 0; x < p.nx; ++x) {
            float *row = r + static_cast<uint64_t>(x) * p.ny;
            const bool xin = static_cast<int64_t>(x) >= x0 && static_cast<int64_t>(x) <= x1;
            const float *krow = k + static_cast<uint64_t>(xin ? x - xmin : 0) * p.ky;
            for (uint32_t y = tid; y < p.ny; y += nth) {
                float v = row[y];
                if (xin && static_cast<int64_t>(y) >= y0 && static_cast<int64_t>(y) <= y1) {
                    const float pr = comp * krow[y - ymin];
                    v = v - pr;
                    row[y] = v;
                }
                if ((!mk || mk[static_cast<uint64_t>(x) * p.ny + y]) && !isnan(v)) {
                    const Cand cand{fabsf(v), y * p.nx + x};
                    if (better(cand, best)) best = cand;
                }
            }
        }
        // wavefront, then workgroup reduction of (|v|, key)
        for (int o = 32; o > 0; o >>= 1) {
            const Cand other{__shfl_xor(best.a, o, 64), static_cast<uint32_t>(__shfl_xor(static_cast<int>(best.key), o, 64))};
            if (better(other, best)) best = other;
        }
        if (lane == 0) wave_best[wave] = best;
        __syncthreads();
        if (tid == 0) {
            Cand b = wave_best[0];
            for (uint32_t w = 1; w < nth / 64; ++w)
                if (better(wave_best[w], b)) b = wave_best[w];
            int found = 0;
            if (seed != UINT32_MAX) {
                found = 1;
                const float sv = r[seed];
                if (isnan(sv)) {  // the seed stays the reference's best
                    s_px = seed / p.ny;
                    s_py = seed - s_px * p.ny;
                    s_value = sv;
                } else {
                    s_px = b.key % p.nx;
                    s_py = b.key / p.nx;
                    s_value = r[static_cast<uint64_t>(s_px) * p.ny + s_py];
                }
            }
            s_found = found;
        }
        __syncthreads();
    };

    pass(false, 0.0f, 0, 0);
    uint32_t done = 0;
    int why = CHIP_CLEAN_BUDGET;
    while (done < p.budget) {
        if (!s_found) {
            why = CHIP_CLEAN_NO_PIXELS;
            break;
        }
        const float pv = s_value;
        const uint32_t px = s_px, py = s_py;
        const int st = minor_stop(fabsf(pv), p.threshold, cyc, nsig);
        if (st != CHIP_CLEAN_BUDGET) {
            why = st;
            break;
        }
        const float comp = p.gain * pv;
        if (tid == 0) {
            float *mc = m + static_cast<uint64_t>(px) * p.ny + py;
            *mc = *mc + comp;
        }
        ++done;
        __syncthreads();  // every thread has read the peak before the pass overwrites it
        pass(true, comp, static_cast<int64_t>(px) - static_cast<int64_t>(p.kx / 2),
             static_cast<int64_t>(py) - static_cast<int64_t>(p.ky / 2));
    }
    if (tid == 0) {
        updates[c] = done;
        stop[c] = why;
    }
}

}  // namespace

void hogbom(chip_ctx &ctx, const chip_hogbom_params &hp, uint64_t n_planes, uint64_t nx, uint64_t ny,
            float *residual, uint64_t kx, uint64_t ky, const float *psf, float *model, const uint8_t *mask,
            const float *cycle_threshold, const float *nsigma_threshold, uint32_t *updates, int32_t *stop) {
    if (!residual || !psf || !model || !cycle_threshold || !nsigma_threshold)
        fail(CHIP_ERR_INVALID, "NULL argument");
    if (nx == 0 || ny == 0 || kx == 0 || ky == 0) fail(CHIP_ERR_INVALID, "empty image or PSF");
    if (nx * ny >= (1ull << 32) || kx * ky >= (1ull << 32)) fail(CHIP_ERR_INVALID, "plane too large");
    if (n_planes > 0x7fffffffull) fail(CHIP_ERR_INVALID, "too many planes");
    if (hp.budget > 0x7fffffffu) fail(CHIP_ERR_INVALID, "component budget too large");
    if (n_planes == 0) return;
    hipStream_t st = ctx.stream;
    DevBuf thr(2 * n_planes * sizeof(float)), out(n_planes * (sizeof(uint32_t) + sizeof(int32_t)));
    CHIP_HIP(hipMemcpyAsync(thr.p, cycle_threshold, n_planes * sizeof(float), hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(thr.as<float>() + n_planes, nsigma_threshold, n_planes * sizeof(float),
                            hipMemcpyHostToDevice, st));
    HogbomParams p{static_cast<uint32_t>(nx), static_cast<uint32_t>(ny), static_cast<uint32_t>(kx),
                   static_cast<uint32_t>(ky), hp.gain, hp.threshold, hp.budget};
    uint32_t *d_upd = out.as<uint32_t>();
    int32_t *d_stop = reinterpret_cast<int32_t *>(d_upd + n_planes);
    hipLaunchKernelGGL(k_hogbom, dim3(static_cast<unsigned>(n_planes)), dim3(kCleanThreads), 0, st, p, residual, psf,
                       model, mask, thr.as<float>(), thr.as<float>() + n_planes, d_upd, d_stop);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipMemcpyAsync(updates, d_upd, n_planes * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipMemcpyAsync(stop, d_stop, n_planes * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
}

}  // namespace chip
