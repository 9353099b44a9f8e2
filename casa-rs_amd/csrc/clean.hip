// Hogbom minor cycle on the GPU (run_hogbom_minor_cycle_cpu_with_progress,
// lib.rs:59748-59885), every image plane of a cube at once.
//
// A component is one launch over (plane, slice of rows) workgroups: each
// workgroup owns rows [x_lo, x_hi) of its plane, subtracts the previous
// component's shifted PSF there and, in the same pass, finds the slice's
// best candidate; the next launch reduces the slices' candidates of each
// plane (every workgroup of the plane the same way) to the peak.  So a
// single large plane spreads over the whole GPU, and the residual never
// leaves the device between components.
//   peak: largest |value| over unmasked pixels, ties to the smallest y-major
//         key y*nx + x (peak_location_masked_in_bounds_zero_tolerance,
//         lib.rs:74064-74115); the reference seeds its search with the first
//         unmasked pixel in x-major order, so a NaN there is the peak, and a
//         NaN anywhere else never is;
//   stop: minor_cycle_stop_reason (lib.rs:74653-74677) on |peak|;
//   update: component = gain * peak; model[peak] += component; residual -=
//         component * psf shifted to the peak over the overlap window, one
//         f32 product then the subtraction (subtract_shifted_kernel_in_window,
//         lib.rs:74805-74868).
// The order of candidates is total (|v|, then key), so the reduction order
// does not matter and planes are bit-identical to the CPU loop.  Candidates
// carry the signed value, and the slice holding the search seed records the
// seed's value, so no workgroup reads pixels another one is rewriting.
#include <cmath>

#include "device.hpp"

namespace chip {

namespace {

constexpr int kCleanThreads = 256;

struct HogbomParams {
    uint32_t nx, ny, kx, ky;
    float gain, threshold;
    uint32_t budget;
    uint32_t slices;  // workgroups (row slices) per plane
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

// Candidate (|v|, y-major key) with its signed value; a < 0 = none.  b wins
// over a when larger, or equal with a smaller key.
struct Cand {
    float a;
    uint32_t key;
    float v;
    uint32_t pad;
};
__device__ __forceinline__ bool better(const Cand &b, const Cand &a) {
    return b.a > a.a || (b.a == a.a && b.key < a.key);
}

// Per plane: the search seed and the outcome.
struct PlaneState {
    uint32_t seed;  // UINT32_MAX: no unmasked pixel
    uint32_t stopped, why, updates;
};

__device__ Cand block_best(Cand best, Cand *wave_best) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        const Cand other{__shfl_xor(best.a, o, 64), static_cast<uint32_t>(__shfl_xor(static_cast<int>(best.key), o, 64)),
                         __shfl_xor(best.v, o, 64), 0u};
        if (better(other, best)) best = other;
    }
    if (lane == 0) wave_best[wave] = best;
    __syncthreads();
    Cand b = wave_best[0];
    for (uint32_t w = 1; w < blockDim.x / 64; ++w)
        if (better(wave_best[w], b)) b = wave_best[w];
    __syncthreads();
    return b;
}

// Seed of the reference's search: the first unmasked pixel in x-major order,
// which is memory order.  Chunks end with two barriers around the read of
// s_seed, so every thread leaves the loop at the same chunk.
__global__ __launch_bounds__(kCleanThreads) void k_hog_init(HogbomParams p, const uint8_t *mask, PlaneState *state) {
    __shared__ uint32_t s_seed;
    const uint64_t c = blockIdx.x;
    const uint64_t npix = static_cast<uint64_t>(p.nx) * p.ny;
    const uint8_t *mk = mask ? mask + c * npix : nullptr;
    if (threadIdx.x == 0) s_seed = mk ? UINT32_MAX : 0u;
    __syncthreads();
    if (mk) {
        constexpr uint64_t kChunk = 16 * kCleanThreads;
        for (uint64_t base = 0; base < npix; base += kChunk) {
            const uint64_t end = base + kChunk < npix ? base + kChunk : npix;
            for (uint64_t i = base + threadIdx.x; i < end; i += blockDim.x)
                if (mk[i]) {
                    atomicMin(&s_seed, static_cast<uint32_t>(i));
                    break;
                }
            __syncthreads();
            const bool done = s_seed != UINT32_MAX;
            __syncthreads();
            if (done) break;
        }
    }
    if (threadIdx.x == 0) state[c] = PlaneState{s_seed, 0u, static_cast<uint32_t>(CHIP_CLEAN_BUDGET), 0u};
}

// Step k: k = 0 finds the first peak; k >= 1 decides component k - 1 (stop
// or subtract) from the candidates of step k - 1 and searches the next peak.
// Candidates of a plane: one per slice, then the search seed's value (in .v)
// recorded by the slice that holds the seed.  A slice the subtraction window
// misses keeps its candidates.
__global__ __launch_bounds__(kCleanThreads) void k_hog_step(HogbomParams p, uint32_t k, float *residual,
                                                           const float *psf, float *model, const uint8_t *mask,
                                                           const float *cyc_thr, const float *nsig_thr,
                                                           PlaneState *state, const Cand *cand_in, Cand *cand_out) {
#pragma clang fp contract(off)
    __shared__ Cand wave_best[kCleanThreads / 64];
    const uint64_t c = blockIdx.x / p.slices;
    const uint32_t w = blockIdx.x - static_cast<uint32_t>(c) * p.slices;
    const uint64_t cbase = c * (p.slices + 1);
    const uint32_t tid = threadIdx.x, nth = blockDim.x;
    PlaneState *ps = state + c;
    const uint64_t npix = static_cast<uint64_t>(p.nx) * p.ny;
    float *r = residual + c * npix;
    const float *kern = psf + c * static_cast<uint64_t>(p.kx) * p.ky;
    const uint8_t *mk = mask ? mask + c * npix : nullptr;
    const uint32_t seed = ps->seed;
    bool sub = false;
    float comp = 0.0f;
    int64_t xmin = 0, ymin = 0;
    if (k > 0) {
        if (ps->stopped) return;
        // the peak: the best candidate of the plane's slices
        Cand best{-1.0f, UINT32_MAX, 0.0f, 0u};
        for (uint32_t j = tid; j < p.slices; j += nth) {
            const Cand o = cand_in[cbase + j];
            if (better(o, best)) best = o;
        }
        best = block_best(best, wave_best);
        const bool found = seed != UINT32_MAX;
        uint32_t px = 0, py = 0;
        float pv = 0.0f;
        if (found) {
            const float sv = cand_in[cbase + p.slices].v;
            if (isnan(sv)) {  // the seed stays the reference's best
                px = seed / p.ny;
                py = seed - px * p.ny;
                pv = sv;
            } else {
                px = best.key % p.nx;
                py = best.key / p.nx;
                pv = best.v;
            }
        }
        int why = CHIP_CLEAN_BUDGET;
        bool stop = false;
        if (k - 1 >= p.budget) {
            stop = true;
        } else if (!found) {
            why = CHIP_CLEAN_NO_PIXELS;
            stop = true;
        } else {
            why = minor_stop(fabsf(pv), p.threshold, cyc_thr[c], nsig_thr[c]);
            stop = why != CHIP_CLEAN_BUDGET;
        }
        if (stop) {
            if (w == 0 && tid == 0) {
                ps->stopped = 1u;
                ps->why = static_cast<uint32_t>(why);
                ps->updates = k - 1;
            }
            return;
        }
        comp = p.gain * pv;
        if (w == 0 && tid == 0) {
            float *mc = model + c * npix + static_cast<uint64_t>(px) * p.ny + py;
            *mc = *mc + comp;
        }
        sub = true;
        xmin = static_cast<int64_t>(px) - static_cast<int64_t>(p.kx / 2);
        ymin = static_cast<int64_t>(py) - static_cast<int64_t>(p.ky / 2);
    }
    // this slice's rows: subtract comp * psf shifted to (px, py) where it
    // overlaps, then the slice's best candidate
    int64_t x0 = 0, x1 = -1, y0 = 0, y1 = -1;
    if (sub) {
        const int64_t xmax = xmin + p.kx - 1, ymax = ymin + p.ky - 1;
        x0 = xmin > 0 ? xmin : 0;
        y0 = ymin > 0 ? ymin : 0;
        x1 = xmax < static_cast<int64_t>(p.nx) - 1 ? xmax : static_cast<int64_t>(p.nx) - 1;
        y1 = ymax < static_cast<int64_t>(p.ny) - 1 ? ymax : static_cast<int64_t>(p.ny) - 1;
    }
    const uint32_t x_lo = static_cast<uint32_t>(static_cast<uint64_t>(w) * p.nx / p.slices);
    const uint32_t x_hi = static_cast<uint32_t>(static_cast<uint64_t>(w + 1) * p.nx / p.slices);
    const bool seed_here = seed != UINT32_MAX && seed / p.ny >= x_lo && seed / p.ny < x_hi;
    if (sub && (x1 < static_cast<int64_t>(x_lo) || x0 >= static_cast<int64_t>(x_hi))) {
        if (tid == 0) {
            cand_out[cbase + w] = cand_in[cbase + w];
            if (seed_here) cand_out[cbase + p.slices] = cand_in[cbase + p.slices];
        }
        return;
    }
    Cand best{-1.0f, UINT32_MAX, 0.0f, 0u};
    // four pixels per thread in flight: loads first, then the updates
    constexpr uint32_t kUnroll = 4;
    for (uint32_t x = x_lo; x < x_hi; ++x) {
        float *row = r + static_cast<uint64_t>(x) * p.ny;
        const uint8_t *mrow = mk ? mk + static_cast<uint64_t>(x) * p.ny : nullptr;
        const bool xin = static_cast<int64_t>(x) >= x0 && static_cast<int64_t>(x) <= x1;
        const float *krow = kern + static_cast<uint64_t>(xin ? x - xmin : 0) * p.ky;
        for (uint32_t yb = tid; yb < p.ny; yb += kUnroll * nth) {
            float v[kUnroll], kv[kUnroll];
            bool in[kUnroll], win[kUnroll], keep[kUnroll];
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t y = yb + j * nth;
                in[j] = y < p.ny;
                win[j] = in[j] && xin && static_cast<int64_t>(y) >= y0 && static_cast<int64_t>(y) <= y1;
                v[j] = in[j] ? row[y] : 0.0f;
                kv[j] = win[j] ? krow[y - ymin] : 0.0f;
                keep[j] = in[j] && (!mrow || mrow[y]);
            }
#pragma unroll
            for (uint32_t j = 0; j < kUnroll; ++j) {
                const uint32_t y = yb + j * nth;
                if (win[j]) {
                    const float pr = comp * kv[j];
                    v[j] = v[j] - pr;
                    row[y] = v[j];
                }
                if (keep[j] && !isnan(v[j])) {
                    const Cand cand{fabsf(v[j]), y * p.nx + x, v[j], 0u};
                    if (better(cand, best)) best = cand;
                }
            }
        }
    }
    best = block_best(best, wave_best);  // (its barriers also order this slice's writes)
    if (tid == 0) {
        cand_out[cbase + w] = best;
        if (seed_here) cand_out[cbase + p.slices] = Cand{0.0f, 0u, r[seed], 0u};
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
    if (n_planes > (1ull << 20)) fail(CHIP_ERR_INVALID, "at most 2^20 planes");
    if (hp.budget > 0x7fffffffu) fail(CHIP_ERR_INVALID, "component budget too large");
    if (n_planes == 0) return;
    hipStream_t st = ctx.stream;
    // about eight workgroups per CU in total, at most one per row
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx.device) != hipSuccess || cus <= 0)
        cus = 256;
    const uint64_t slices = std::max<uint64_t>(1, std::min<uint64_t>(nx, (8ull * cus + n_planes - 1) / n_planes));
    if (n_planes * slices * kCleanThreads >= (1ull << 32)) fail(CHIP_ERR_INVALID, "too many planes");
    HogbomParams p{static_cast<uint32_t>(nx), static_cast<uint32_t>(ny), static_cast<uint32_t>(kx),
                   static_cast<uint32_t>(ky), hp.gain, hp.threshold, hp.budget, static_cast<uint32_t>(slices)};
    DevBuf thr(2 * n_planes * sizeof(float)), state(n_planes * sizeof(PlaneState));
    DevBuf cand0(n_planes * (slices + 1) * sizeof(Cand)), cand1(n_planes * (slices + 1) * sizeof(Cand));
    CHIP_HIP(hipMemcpyAsync(thr.p, cycle_threshold, n_planes * sizeof(float), hipMemcpyHostToDevice, st));
    CHIP_HIP(hipMemcpyAsync(thr.as<float>() + n_planes, nsigma_threshold, n_planes * sizeof(float),
                            hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_hog_init, dim3(static_cast<unsigned>(n_planes)), dim3(kCleanThreads), 0, st, p, mask,
                       state.as<PlaneState>());
    CHIP_HIP(hipGetLastError());
    const dim3 grid(static_cast<unsigned>(n_planes * slices));
    std::vector<PlaneState> hs(n_planes);
    Cand *cin = cand1.as<Cand>(), *cout = cand0.as<Cand>();
    for (uint64_t k = 0; k <= static_cast<uint64_t>(hp.budget) + 1; ++k) {
        hipLaunchKernelGGL(k_hog_step, grid, dim3(kCleanThreads), 0, st, p, static_cast<uint32_t>(k), residual, psf,
                           model, mask, thr.as<float>(), thr.as<float>() + n_planes, state.as<PlaneState>(), cin,
                           cout);
        CHIP_HIP(hipGetLastError());
        std::swap(cin, cout);
        if (k % 16 == 15) {  // every plane stopped: the remaining launches would return at once
            CHIP_HIP(hipMemcpyAsync(hs.data(), state.p, state.bytes, hipMemcpyDeviceToHost, st));
            CHIP_HIP(hipStreamSynchronize(st));
            bool all = true;
            for (const PlaneState &q : hs) all = all && q.stopped;
            if (all) break;
        }
    }
    CHIP_HIP(hipMemcpyAsync(hs.data(), state.p, state.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    for (uint64_t c = 0; c < n_planes; ++c) {
        if (!hs[c].stopped) fail(CHIP_ERR_DEVICE, "minor cycle ended without a stop record");
        updates[c] = hs[c].updates;
        stop[c] = static_cast<int32_t>(hs[c].why);
    }
}

}  // namespace chip
