// Plain per-channel gridding and degridding (the binned path is in binned.hip).
//
// Plain path (chip_grid_plain): one launch per channel, one thread per
// sample, taps added straight to the global grid with FP64 (or FP32) atomics.
// This is the straightforward per-channel GPU port of the reference CPU loop
// and the FP64 baseline the binned path is measured against.
//
// Per-sample arithmetic follows the reference:
//   standard: tap = wx[tx]*wy[ty] in f32 (gridder.rs:1036-1050); PSF value is
//             the grid weight, dirty value V*weight, residual (V-model)*weight
//             (execution.rs:18306-18325, gridder.rs:1315-1396);
//   W:        cwt = K[|ix*s+off_x|][|iy*s+off_y|], conjugated when w > 0;
//             grid += value*cwt; degrid += conj(cwt)*grid (gridder.rs:3680-3747).
#include <mutex>

#include "device.hpp"

// Scratch of chip_grid_plain_noalloc: the per-channel FP64 sums
// (normalisation sum of weights, reported sum, gridded count) that
// chip_grid_plain keeps in a buffer it allocates per call.  Every GPU the
// library runs on holds one copy until the process ends; calls on one GPU
// share it (overlapping calls race).  Host code never passes this name to a
// kernel: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress, per device, at context creation.
static __device__ __align__(256) double scratch[3 * CHIP_NOALLOC_MAX_CHAN];

namespace chip {

namespace {
constexpr int kMaxDevices = 64;
double *g_scratch_dev[kMaxDevices] = {};
std::mutex g_scratch_mu;

__global__ __launch_bounds__(256) void k_zero_f64(double *p, uint64_t n) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0;
}
}  // namespace

// Device address of `scratch` on `device` (the current device), looked up
// once per device.
void noalloc_scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) fail(CHIP_ERR_DEVICE, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[device]) return;
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
        fail(CHIP_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_dev[device] = static_cast<double *>(ptr);
}

static double *noalloc_scratch(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    double *p = (device >= 0 && device < kMaxDevices) ? g_scratch_dev[device] : nullptr;
    if (!p) fail(CHIP_ERR_DEVICE, "scratch of the no-allocation path not set up on this device");
    return p;
}

// ---------------------------------------------------------------------------
// Plain per-channel kernels.
// ---------------------------------------------------------------------------
template <typename T> struct Cx;
template <> struct Cx<float> { using type = float2; };
template <> struct Cx<double> { using type = double2; };

template <typename T>
__device__ __forceinline__ void gadd(T *p, T v) {
    if (v != T(0)) unsafeAtomicAdd(p, v);
}

template <typename T, int MODE, bool WP>
__global__ __launch_bounds__(256) void k_plain(PlanParams pp, Source s, uint32_t chan,
                                               GridParams gp, const T *model, T *psf, T *acc,
                                               T *vis_out, double *sums) {
    __shared__ double red[3][4];
    const uint64_t row = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double nsum = 0.0, rsum = 0.0, cnt = 0.0;
    if (row < s.n_rows) {
        const uint64_t i = row * s.n_chan + chan;
        const SampleIn smp = load_sample(s, i);
        Plan pl;
        bool ok = sample_accepted(pp, smp, pl);
        // W tables recomputed per chunk of planes: samples of planes outside
        // [plane_lo, plane_hi) belong to another launch
        const bool other_chunk = WP && ok && (static_cast<uint32_t>(pl.plane) < gp.plane_lo ||
                                              static_cast<uint32_t>(pl.plane) >= gp.plane_hi);
        ok = ok && !other_chunk;
        const uint32_t flags = ok ? sample_flags(pp, smp) : 0u;
        if (ok && MODE == kModeResidual && !(flags & kResidOkBit)) ok = false;
        if (MODE == kModeDegrid && !ok && !other_chunk) vis_out[2 * i] = vis_out[2 * i + 1] = T(0);
        if (ok) {
            const float gwf = grid_weight(pp, smp);
            if (pp.aw) {
                nsum = static_cast<double>(smp.weight) * pl.mag;
                rsum = nsum;
            } else if (pp.wproject) {
                nsum = 2.0 * static_cast<double>(smp.weight) * static_cast<double>(pl.norm);
                rsum = static_cast<double>(smp.weight) * static_cast<double>(smp.sf);
            } else {
                nsum = static_cast<double>(gwf);
                rsum = nsum;
            }
            cnt = 1.0;
            const T w = static_cast<T>(gwf);
            const uint64_t base = static_cast<uint64_t>(chan) * gp.cells;
            if (!WP) {
                const float *wxr = gp.taps + pl.sx * kTaps;
                const float *wyr = gp.taps + pl.sy * kTaps;
                T pr = 0, pi = 0;
                if ((MODE == kModeResidual || MODE == kModeDegrid) && model) {
                    for (int tx = 0; tx < kTaps; ++tx) {
                        const uint64_t rowi = base + static_cast<uint64_t>(pl.x + tx) * gp.gy + pl.y;
                        for (int ty = 0; ty < kTaps; ++ty) {
                            const T t = static_cast<T>(wxr[tx] * wyr[ty]);
                            pr += model[2 * (rowi + ty)] * t;
                            pi += model[2 * (rowi + ty) + 1] * t;
                        }
                    }
                }
                if (MODE == kModeDegrid) {
                    vis_out[2 * i] = static_cast<T>(pr);
                    vis_out[2 * i + 1] = static_cast<T>(pi);
                } else {
                    T vr, vi;
                    if (MODE == kModeResidual) {
                        vr = (static_cast<T>(smp.vis.x) - pr) * w;
                        vi = (static_cast<T>(smp.vis.y) - pi) * w;
                    } else {
                        vr = static_cast<T>(smp.vis.x) * w;
                        vi = static_cast<T>(smp.vis.y) * w;
                    }
                    const bool dirty_on = MODE == kModeResidual || (MODE == kModeDirty && (flags & kVisOkBit));
                    for (int tx = 0; tx < kTaps; ++tx) {
                        const uint64_t rowi = base + static_cast<uint64_t>(pl.x + tx) * gp.gy + pl.y;
                        for (int ty = 0; ty < kTaps; ++ty) {
                            const T t = static_cast<T>(wxr[tx] * wyr[ty]);
                            const uint64_t ci = 2 * (rowi + ty);
                            if (MODE == kModePsf || MODE == kModeDirty) gadd(psf + ci, w * t);
                            if (dirty_on) {
                                gadd(acc + ci, vr * t);
                                gadd(acc + ci + 1, vi * t);
                            }
                        }
                    }
                }
            } else if (pp.aw) {
                // AWProject (gridder.rs:2318-2371): phased taps from the tap
                // table; contribution value*tap as Complex32 added to the T
                // grid; degrid sum of conj(tap)*grid in tap order divided by
                // conj(normalization).
#pragma clang fp contract(off)
                const int R = gp.norm_radius;
                const int n_off = static_cast<int>(gp.n_off);
                const int S = gp.supports[pl.plane];
                const int side = 2 * S + 1;
                const AwCell c = pp.aw_cell[static_cast<uint32_t>(pl.plane) / (2u * pp.aw_pointings)];
                const uint64_t oi = (static_cast<uint64_t>(pl.plane) * n_off + (pl.sx + R)) * n_off + (pl.sy + R);
                const float2 *tab = gp.compact + oi * gp.tap_stride;
                const float2 nm = gp.aw_norm[oi];
                T pr = 0, pi = 0;
                if ((MODE == kModeResidual || MODE == kModeDegrid) && model) {
                    for (int iy = -c.ys; iy <= c.ys; ++iy)
                        for (int ix = -c.xs; ix <= c.xs; ++ix) {
                            const float2 t = tab[(ix + S) * side + (iy + S)];
                            const T cr = static_cast<T>(t.x), ci = -static_cast<T>(t.y);
                            const uint64_t gi = 2 * (base + static_cast<uint64_t>(pl.x + ix) * gp.gy + (pl.y + iy));
                            const T gr = model[gi], gim = model[gi + 1];
                            const T ar = cr * gr - ci * gim;
                            const T ai = cr * gim + ci * gr;
                            pr += ar;
                            pi += ai;
                        }
                    div_conj<T>(pr, pi, static_cast<T>(nm.x), static_cast<T>(nm.y));
                }
                if (MODE == kModeDegrid) {
                    vis_out[2 * i] = pr;
                    vis_out[2 * i + 1] = pi;
                } else {
                    // value: PSF weight; dirty V*weight; residual (V - model)*weight, in f32
                    const float wf = gwf;
                    float vr, vi;
                    if (MODE == kModeResidual) {
                        const float rr = smp.vis.x - static_cast<float>(pr);
                        const float ri = smp.vis.y - static_cast<float>(pi);
                        vr = rr * wf;
                        vi = ri * wf;
                    } else {
                        vr = smp.vis.x * wf;
                        vi = smp.vis.y * wf;
                    }
                    const bool dirty_on = MODE == kModeResidual || (MODE == kModeDirty && (flags & kVisOkBit));
                    for (int iy = -c.ys; iy <= c.ys; ++iy)
                        for (int ix = -c.xs; ix <= c.xs; ++ix) {
                            const float2 t = tab[(ix + S) * side + (iy + S)];
                            const uint64_t gi = 2 * (base + static_cast<uint64_t>(pl.x + ix) * gp.gy + (pl.y + iy));
                            if (MODE == kModePsf || MODE == kModeDirty) {
                                const float cr = wf * t.x - 0.0f * t.y;
                                const float ci = wf * t.y + 0.0f * t.x;
                                gadd(psf + gi, static_cast<T>(cr));
                                gadd(psf + gi + 1, static_cast<T>(ci));
                            }
                            if (dirty_on) {
                                const float cr = vr * t.x - vi * t.y;
                                const float ci = vr * t.y + vi * t.x;
                                gadd(acc + gi, static_cast<T>(cr));
                                gadd(acc + gi + 1, static_cast<T>(ci));
                            }
                        }
                }
            } else {
                const int s_ = static_cast<int>(gp.sampling);
                const int W = static_cast<int>(gp.width);
                const int S = gp.supports[pl.plane];
                const float2 *K = gp.kernels + static_cast<uint64_t>(pl.plane - static_cast<int>(gp.plane_lo)) * W * W;
                const T csign = pl.conj ? T(-1) : T(1);
                T pr = 0, pi = 0;
                if ((MODE == kModeResidual || MODE == kModeDegrid) && model) {
                    for (int iy = -S; iy <= S; ++iy) {
                        const int ky = abs(iy * s_ + pl.sy);
                        for (int ix = -S; ix <= S; ++ix) {
                            const int kx = abs(ix * s_ + pl.sx);
                            const float2 c = K[kx * W + ky];
                            const T cr = c.x, ci = -(static_cast<T>(c.y) * csign);
                            const uint64_t gi = 2 * (base + static_cast<uint64_t>(pl.x + ix) * gp.gy + (pl.y + iy));
                            const T gr = model[gi], gim = model[gi + 1];
                            pr += cr * gr - ci * gim;
                            pi += cr * gim + ci * gr;
                        }
                    }
                }
                if (MODE == kModeDegrid) {
                    vis_out[2 * i] = static_cast<T>(pr);
                    vis_out[2 * i + 1] = static_cast<T>(pi);
                } else {
                    T vr, vi;
                    if (MODE == kModeResidual) {
                        vr = (static_cast<T>(smp.vis.x) - pr) * w;
                        vi = (static_cast<T>(smp.vis.y) - pi) * w;
                    } else {
                        vr = static_cast<T>(smp.vis.x) * w;
                        vi = static_cast<T>(smp.vis.y) * w;
                    }
                    for (int iy = -S; iy <= S; ++iy) {
                        const int ky = abs(iy * s_ + pl.sy);
                        for (int ix = -S; ix <= S; ++ix) {
                            const int kx = abs(ix * s_ + pl.sx);
                            const float2 c = K[kx * W + ky];
                            const T cr = c.x, ci = static_cast<T>(c.y) * csign;
                            const uint64_t gi = 2 * (base + static_cast<uint64_t>(pl.x + ix) * gp.gy + (pl.y + iy));
                            if (MODE == kModePsf || MODE == kModeDirty) {
                                gadd(psf + gi, w * cr);
                                gadd(psf + gi + 1, w * ci);
                            }
                            if (MODE != kModePsf) {
                                gadd(acc + gi, vr * cr - vi * ci);
                                gadd(acc + gi + 1, vr * ci + vi * cr);
                            }
                        }
                    }
                }
            }
        }
    }
    if (MODE == kModeDegrid || sums == nullptr) return;
    nsum = wave_sum(nsum);
    rsum = wave_sum(rsum);
    cnt = wave_sum(cnt);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) { red[0][wave] = nsum; red[1][wave] = rsum; red[2][wave] = cnt; }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0, b = 0, c = 0;
        for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) { a += red[0][k]; b += red[1][k]; c += red[2][k]; }
        if (c > 0) {
            atomicAdd(&sums[3 * chan], a);
            atomicAdd(&sums[3 * chan + 1], b);
            atomicAdd(&sums[3 * chan + 2], c);
        }
    }
}

template <typename T, int MODE>
static void launch_plain_t(chip_gridder &g, const Source &src, const void *model, void *psf,
                           void *acc, void *vis_out, double *sums) {
    GridParams gp = grid_params(g, 32, 0, 1);
    const PlanParams pp = plan_params(g, src);
    hipStream_t st = g.ctx->stream;
    const unsigned blocks = static_cast<unsigned>((src.n_rows + 255) / 256);
    if (blocks == 0) return;
    auto run = [&](const GridParams &p) {
        for (uint32_t c = 0; c < src.n_chan; ++c) {
            if (g.wproject)
                hipLaunchKernelGGL((k_plain<T, MODE, true>), dim3(blocks), dim3(256), 0, st, pp, src, c,
                                   p, static_cast<const T *>(model), static_cast<T *>(psf),
                                   static_cast<T *>(acc), static_cast<T *>(vis_out), sums);
            else
                hipLaunchKernelGGL((k_plain<T, MODE, false>), dim3(blocks), dim3(256), 0, st, pp, src,
                                   c, p, static_cast<const T *>(model), static_cast<T *>(psf),
                                   static_cast<T *>(acc), static_cast<T *>(vis_out), sums);
            CHIP_HIP(hipGetLastError());
        }
    };
    if (!g.wproject || g.aw || !g.wk_chunk) {
        run(gp);
        return;
    }
    // W tables recomputed on the fly: one pass per chunk of planes
    WkScratch ws;
    wk_prepare(g, ws);
    for (uint32_t lo = 0; lo < g.planes; lo += g.wk_chunk) {
        wk_fill(g, ws, lo, std::min(g.planes, lo + g.wk_chunk), gp);
        run(gp);
    }
    CHIP_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
}

template <typename T>
static void grid_plain_t(chip_gridder &g, const Source &src, uint32_t mode, const void *model,
                         void *psf, void *dirty, double *sums) {
    switch (mode) {
        case CHIP_MODE_PSF:
            if (!psf) fail(CHIP_ERR_INVALID, "PSF mode needs a PSF grid");
            launch_plain_t<T, kModePsf>(g, src, nullptr, psf, nullptr, nullptr, sums);
            break;
        case CHIP_MODE_DIRTY:
            if (!psf || !dirty) fail(CHIP_ERR_INVALID, "dirty mode needs PSF and dirty grids");
            launch_plain_t<T, kModeDirty>(g, src, nullptr, psf, dirty, nullptr, sums);
            break;
        case CHIP_MODE_RESIDUAL:
            if (!dirty) fail(CHIP_ERR_INVALID, "residual mode needs a residual grid");
            launch_plain_t<T, kModeResidual>(g, src, model, nullptr, dirty, nullptr, sums);
            break;
        default:
            fail(CHIP_ERR_INVALID, "unknown grid mode");
    }
}

void grid_plain(chip_gridder &g, const Source &src, uint32_t mode, uint32_t precision,
                const void *model, void *psf, void *dirty, chip_sums *sums_host) {
    hipStream_t st = g.ctx->stream;
    DevBuf sums(src.n_chan * 3 * sizeof(double));
    CHIP_HIP(hipMemsetAsync(sums.p, 0, sums.bytes, st));
    if (precision == CHIP_F64)
        grid_plain_t<double>(g, src, mode, model, psf, dirty, sums.as<double>());
    else if (precision == CHIP_F32)
        grid_plain_t<float>(g, src, mode, model, psf, dirty, sums.as<double>());
    else
        fail(CHIP_ERR_INVALID, "unknown precision");
    std::vector<double> hs(src.n_chan * 3);
    CHIP_HIP(hipMemcpyAsync(hs.data(), sums.p, sums.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    if (sums_host) {
        for (uint64_t c = 0; c < src.n_chan; ++c) {
            sums_host[c].normalization_sumwt = hs[3 * c];
            sums_host[c].reported_sumwt = hs[3 * c + 1];
            sums_host[c].gridded = static_cast<uint64_t>(hs[3 * c + 2]);
            sums_host[c].skipped = src.n_rows - sums_host[c].gridded;
        }
    }
}

// chip_grid_plain with its sums in the device's `scratch` (no allocation):
// the sums are cleared by a kernel (capturable), the grids are the caller's,
// and the sums come back through a fixed host buffer on the stack.
void grid_plain_noalloc(chip_gridder &g, const Source &src, uint32_t mode, uint32_t precision,
                        const void *model, void *psf, void *dirty, chip_sums *sums_host) {
    if (src.n_chan > CHIP_NOALLOC_MAX_CHAN)
        fail(CHIP_ERR_INVALID, "the no-allocation plain gridder takes at most " +
                                   std::to_string(CHIP_NOALLOC_MAX_CHAN) + " channels per call");
    if (g.wproject && !g.aw && g.wk_chunk)
        fail(CHIP_ERR_INVALID,
             "the no-allocation plain gridder needs resident W tables (recomputing them needs scratch)");
    if (precision != CHIP_F64 && precision != CHIP_F32) fail(CHIP_ERR_INVALID, "unknown precision");
    hipStream_t st = g.ctx->stream;
    double *sums = noalloc_scratch(g.device);
    const uint64_t n = 3 * src.n_chan;
    hipLaunchKernelGGL(k_zero_f64, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, st, sums, n);
    CHIP_HIP(hipGetLastError());
    if (precision == CHIP_F64)
        grid_plain_t<double>(g, src, mode, model, psf, dirty, sums);
    else
        grid_plain_t<float>(g, src, mode, model, psf, dirty, sums);
    constexpr uint64_t kChunk = 128;  // channels per copy into the stack buffer
    double hs[3 * kChunk];
    for (uint64_t c0 = 0; c0 < src.n_chan; c0 += kChunk) {
        const uint64_t nc = std::min<uint64_t>(kChunk, src.n_chan - c0);
        CHIP_HIP(hipMemcpyAsync(hs, sums + 3 * c0, 3 * nc * sizeof(double), hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        if (sums_host) {
            for (uint64_t k = 0; k < nc; ++k) {
                chip_sums &o = sums_host[c0 + k];
                o.normalization_sumwt = hs[3 * k];
                o.reported_sumwt = hs[3 * k + 1];
                o.gridded = static_cast<uint64_t>(hs[3 * k + 2]);
                o.skipped = src.n_rows - o.gridded;
            }
        }
    }
    CHIP_HIP(hipStreamSynchronize(st));
}

// ---------------------------------------------------------------------------
// Packed, pre-planned W samples (the reference's Metal seam input).
// ---------------------------------------------------------------------------
// WProjectMetalSample (lib.rs:71629-71635, packed by
// w_project_metal_sample_from_plan, :71708-71754): loc_x | loc_y << 16;
// plane (10 bits) | support << 10 (5 bits) | conjugate << 15 |
// (off_x + 128) << 16 | (off_y + 128) << 24; weight; visibility.
struct PackedWSample {
    uint32_t loc, pso;
    float weight, vr, vi;
};
static_assert(sizeof(PackedWSample) == 20, "the reference's packed sample is 20 bytes");

struct SeamParams {
    uint32_t gx, gy, width, sampling, planes, mode;
    uint64_t n;
    const float2 *kernels;
};

// One wavefront per sample, lanes over the taps (iy outer, ix inner): the
// arithmetic of W_PROJECT_METAL_SHADER (lib.rs:72656-72724) -- cwt =
// K[|ix s + off_x|][|iy s + off_y|], conjugated when flagged; prediction sum
// of conj(cwt) * model; PSF += weight * cwt; residual += weight * ((V - p) *
// cwt).  Samples whose footprint leaves the grid or the kernel are counted in
// *rejected and not gridded.
__global__ __launch_bounds__(256) void k_seam(SeamParams p, const PackedWSample *samples, const float2 *model,
                                              float *psf, float *res, unsigned int *rejected) {
    const uint64_t i = (static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (i >= p.n) return;
    const PackedWSample sm = samples[i];
    const int lx = static_cast<int>(sm.loc & 0xffffu), ly = static_cast<int>(sm.loc >> 16);
    const uint32_t plane = sm.pso & 0x3ffu;
    const int S = static_cast<int>((sm.pso >> 10) & 0x1fu);
    const bool conj = (sm.pso >> 15) & 1u;
    const int ox = static_cast<int>((sm.pso >> 16) & 0xffu) - 128, oy = static_cast<int>(sm.pso >> 24) - 128;
    const int s = static_cast<int>(p.sampling);
    const bool ok = plane < p.planes && lx - S >= 0 && ly - S >= 0 && lx + S < static_cast<int>(p.gx) &&
                    ly + S < static_cast<int>(p.gy) && S * s + (ox < 0 ? -ox : ox) < static_cast<int>(p.width) &&
                    S * s + (oy < 0 ? -oy : oy) < static_cast<int>(p.width);
    if (!ok) {
        if (lane == 0) atomicAdd(rejected, 1u);
        return;
    }
    const int side = 2 * S + 1, T = side * side;
    const float2 *K = p.kernels + static_cast<uint64_t>(plane) * p.width * p.width;
    auto tap = [&](int t, int &cell) {
        const int jy = t / side, jx = t - jy * side;
        const int iy = jy - S, ix = jx - S;
        const int kx = abs(ix * s + ox), ky = abs(iy * s + oy);
        float2 c = K[static_cast<uint64_t>(kx) * p.width + ky];
        if (conj) c.y = -c.y;
        cell = (lx + ix) * static_cast<int>(p.gy) + (ly + iy);
        return c;
    };
    float rr = sm.vr, ri = sm.vi;
    if (p.mode == CHIP_MODE_RESIDUAL && model) {
        float pr = 0.0f, pi = 0.0f;
        for (int t = lane; t < T; t += 64) {
            int cell;
            const float2 c = tap(t, cell);
            const float2 m = model[cell];
            pr += c.x * m.x + c.y * m.y;
            pi += c.x * m.y - c.y * m.x;
        }
        wave_sum_f2(pr, pi);
        rr = sm.vr - pr;
        ri = sm.vi - pi;
    }
    const bool wpsf = p.mode == CHIP_MODE_PSF || p.mode == CHIP_MODE_DIRTY;
    const bool wres = p.mode == CHIP_MODE_DIRTY || p.mode == CHIP_MODE_RESIDUAL;
    for (int t = lane; t < T; t += 64) {
        int cell;
        const float2 c = tap(t, cell);
        if (wpsf) {
            unsafeAtomicAdd(psf + 2 * static_cast<uint64_t>(cell), sm.weight * c.x);
            unsafeAtomicAdd(psf + 2 * static_cast<uint64_t>(cell) + 1, sm.weight * c.y);
        }
        if (wres) {
            unsafeAtomicAdd(res + 2 * static_cast<uint64_t>(cell), sm.weight * (rr * c.x - ri * c.y));
            unsafeAtomicAdd(res + 2 * static_cast<uint64_t>(cell) + 1, sm.weight * (rr * c.y + ri * c.x));
        }
    }
}

void grid_wproject_samples(chip_gridder &g, uint64_t n, const void *samples, const float *model, float *psf,
                           float *residual, uint32_t mode) {
    if (!g.wproject || g.aw) fail(CHIP_ERR_INVALID, "packed W samples need a W projection (not AW)");
    if (g.wk_chunk) fail(CHIP_ERR_INVALID, "packed W samples need resident W tables");
    if (mode > CHIP_MODE_RESIDUAL) fail(CHIP_ERR_INVALID, "unknown grid mode");
    if ((mode != CHIP_MODE_RESIDUAL && !psf) || (mode != CHIP_MODE_PSF && !residual))
        fail(CHIP_ERR_INVALID, "the mode's grids are required");
    if (n == 0) return;
    if (!samples) fail(CHIP_ERR_INVALID, "NULL samples");
    if (n > (1ull << 31)) fail(CHIP_ERR_INVALID, "at most 2^31 packed samples per call");
    SeamParams p{};
    p.gx = static_cast<uint32_t>(g.gx);
    p.gy = static_cast<uint32_t>(g.gy);
    p.width = g.width;
    p.sampling = g.sampling;
    p.planes = g.planes;
    p.mode = mode;
    p.n = n;
    p.kernels = g.d_kernels.as<float2>();
    hipStream_t st = g.ctx->stream;
    DevBuf rej(sizeof(unsigned int));
    CHIP_HIP(hipMemsetAsync(rej.p, 0, rej.bytes, st));
    hipLaunchKernelGGL(k_seam, dim3(static_cast<unsigned>((n * 64 + 255) / 256)), dim3(256), 0, st, p,
                       static_cast<const PackedWSample *>(samples), reinterpret_cast<const float2 *>(model), psf,
                       residual, rej.as<unsigned int>());
    CHIP_HIP(hipGetLastError());
    unsigned int bad = 0;
    CHIP_HIP(hipMemcpyAsync(&bad, rej.p, sizeof(bad), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    if (bad)
        fail(CHIP_ERR_INVALID, std::to_string(bad) + " packed samples lie outside the grid or the W kernels");
}

void degrid_plain(chip_gridder &g, const Source &src, uint32_t precision, const void *model,
                  void *out) {
    if (!model || !out) fail(CHIP_ERR_INVALID, "degrid needs a model grid and an output");
    if (precision == CHIP_F64)
        launch_plain_t<double, kModeDegrid>(g, src, model, nullptr, nullptr, out, nullptr);
    else if (precision == CHIP_F32)
        launch_plain_t<float, kModeDegrid>(g, src, model, nullptr, nullptr, out, nullptr);
    else
        fail(CHIP_ERR_INVALID, "unknown precision");
}

}  // namespace chip
