// Gridding and degridding kernels.
//
// Binned FP32 path (chip_vset_*): one workgroup per work item (a run of
// samples of one channel that fall in one 32x32 uv tile).  The workgroup
// accumulates the tile plus a halo of the largest half-support in LDS with
// LDS float atomics, then adds the non-zero LDS cells to the global grid with
// few global atomics.  Degridding reads a model tile staged in LDS.  Every
// lane owns whole samples, so no cross-lane reduction is needed and all 64
// lanes of a wavefront stay busy.  MFMA is not used: the work is a scatter.
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
#include "device.hpp"

namespace chip {

enum : int { kModePsf = 0, kModeDirty = 1, kModeResidual = 2, kModeAdjoint = 3, kModeDegrid = 4 };

struct GridParams {
    uint32_t gx, gy;
    uint32_t tile, halo, L;  // L = tile + 2*halo
    uint32_t tiles_y;
    uint32_t sampling, width;
    uint32_t tap_rows;       // standard tap-table rows (os + 1)
    uint64_t cells;          // gx * gy
    const float *taps;       // (os+1) x 7
    const float2 *kernels;   // planes x width x width
    const int32_t *supports;
};

static GridParams grid_params(const chip_gridder &g, uint32_t tile, uint32_t halo,
                              uint32_t tiles_y) {
    GridParams gp{};
    gp.gx = static_cast<uint32_t>(g.gx);
    gp.gy = static_cast<uint32_t>(g.gy);
    gp.tile = tile;
    gp.halo = halo;
    gp.L = tile + 2 * halo;
    gp.tiles_y = tiles_y;
    gp.sampling = g.sampling;
    gp.width = g.width;
    gp.tap_rows = g.os + 1;
    gp.cells = g.gx * g.gy;
    gp.taps = g.d_taps.as<float>();
    gp.kernels = g.wproject ? g.d_kernels.as<float2>() : nullptr;
    gp.supports = g.wproject ? g.d_supports.as<int32_t>() : nullptr;
    return gp;
}

template <int MODE>
struct LdsPlan {
    // number of float arrays of L*L cells
    static constexpr int arrays = MODE == kModePsf ? 2 : (MODE == kModeDirty ? 4 : (MODE == kModeResidual ? 4 : 2));
};

__device__ __forceinline__ void lds_add(float *a, int i, float v) { atomicAdd(a + i, v); }

// ---------------------------------------------------------------------------
// Binned kernel.
// ---------------------------------------------------------------------------
template <int MODE, bool WP>
__global__ __launch_bounds__(256) void k_binned(GridParams gp, const WorkItem *items,
                                                const uint2 *plan, const float2 *vis,
                                                const float *wt, const uint32_t *index,
                                                const float2 *values, const float2 *model,
                                                float *psf, float *acc_out, float2 *vis_out) {
    extern __shared__ float lds[];
    const WorkItem it = items[blockIdx.x];
    const int L = static_cast<int>(gp.L);
    const int cells = L * L;
    const int tap_floats = WP ? 0 : ((static_cast<int>(gp.tap_rows) * kTaps + 3) & ~3);
    float *taps = lds;
    float *a0 = lds + tap_floats;  // first array
    float *a1 = a0 + cells;
    float *a2 = a1 + cells;
    float *a3 = a2 + cells;
    constexpr int narr = LdsPlan<MODE>::arrays;
    // Array roles.
    float *P_re = nullptr, *P_im = nullptr, *A_re = nullptr, *A_im = nullptr, *M_re = nullptr,
          *M_im = nullptr;
    if (MODE == kModePsf) { P_re = a0; P_im = a1; }
    if (MODE == kModeDirty) { P_re = a0; P_im = a1; A_re = a2; A_im = a3; }
    if (MODE == kModeResidual) { M_re = a0; M_im = a1; A_re = a2; A_im = a3; }
    if (MODE == kModeAdjoint) { A_re = a0; A_im = a1; }
    if (MODE == kModeDegrid) { M_re = a0; M_im = a1; }

    const int tid = threadIdx.x;
    const int nth = blockDim.x;
    const uint32_t tile_x = it.tile / gp.tiles_y;
    const uint32_t tile_y = it.tile - tile_x * gp.tiles_y;
    const int x0 = static_cast<int>(tile_x * gp.tile) - static_cast<int>(gp.halo);
    const int y0 = static_cast<int>(tile_y * gp.tile) - static_cast<int>(gp.halo);
    const uint64_t chan_off = static_cast<uint64_t>(it.chan) * gp.cells;

    if (!WP)
        for (int i = tid; i < static_cast<int>(gp.tap_rows) * kTaps; i += nth) taps[i] = gp.taps[i];
    const bool has_model = (MODE == kModeResidual || MODE == kModeDegrid) && model != nullptr;
    for (int c = tid; c < cells; c += nth) {
        float m_re = 0.0f, m_im = 0.0f;
        if (has_model) {
            const int gxx = x0 + c / L, gyy = y0 + c % L;
            if (gxx >= 0 && gyy >= 0 && gxx < static_cast<int>(gp.gx) && gyy < static_cast<int>(gp.gy)) {
                const float2 m = model[chan_off + static_cast<uint64_t>(gxx) * gp.gy + gyy];
                m_re = m.x;
                m_im = m.y;
            }
        }
        if (MODE == kModeResidual || MODE == kModeDegrid) { M_re[c] = m_re; M_im[c] = m_im; }
        if (MODE == kModeResidual || MODE == kModeDirty || MODE == kModeAdjoint) { A_re[c] = 0.0f; A_im[c] = 0.0f; }
        if (MODE == kModePsf || MODE == kModeDirty) { P_re[c] = 0.0f; if (WP) P_im[c] = 0.0f; }
    }
    __syncthreads();

    const int s = static_cast<int>(gp.sampling);
    const int W = static_cast<int>(gp.width);
    for (uint32_t k = tid; k < it.count; k += nth) {
        const uint32_t p = it.start + k;
        const uint2 pk = plan[p];
        const int ax = static_cast<int>(pk.x & 0xffffu) - x0;
        const int ay = static_cast<int>(pk.x >> 16) - y0;
        const uint32_t flags = pk.y;
        float2 v = make_float2(0.0f, 0.0f);
        float w = 1.0f;
        if (MODE == kModeAdjoint) {
            v = values[index[p]];
        } else if (MODE != kModeDegrid) {
            v = vis[p];
            w = wt[p];
        }
        if (MODE == kModeResidual && !(flags & kResidOkBit)) continue;
        if (!WP) {
            const float *wxr = taps + (pk.y & 0xffu) * kTaps;
            const float *wyr = taps + ((pk.y >> 8) & 0xffu) * kTaps;
            float wx[kTaps], wy[kTaps];
#pragma unroll
            for (int t = 0; t < kTaps; ++t) { wx[t] = wxr[t]; wy[t] = wyr[t]; }
            float pr = 0.0f, pi = 0.0f;
            if (MODE == kModeResidual || MODE == kModeDegrid) {
#pragma unroll
                for (int tx = 0; tx < kTaps; ++tx) {
                    const int row = (ax + tx) * L + ay;
#pragma unroll
                    for (int ty = 0; ty < kTaps; ++ty) {
                        const float t = wx[tx] * wy[ty];
                        pr += M_re[row + ty] * t;
                        pi += M_im[row + ty] * t;
                    }
                }
            }
            if (MODE == kModeDegrid) {
                vis_out[index[p]] = make_float2(pr, pi);
                continue;
            }
            float vr, vi;
            if (MODE == kModeResidual) { vr = (v.x - pr) * w; vi = (v.y - pi) * w; }
            else if (MODE == kModeAdjoint) { vr = v.x; vi = v.y; }
            else { vr = v.x * w; vi = v.y * w; }
            const bool dirty_on = MODE == kModeResidual || MODE == kModeAdjoint ||
                                  (MODE == kModeDirty && (flags & kVisOkBit));
#pragma unroll
            for (int tx = 0; tx < kTaps; ++tx) {
                const int row = (ax + tx) * L + ay;
#pragma unroll
                for (int ty = 0; ty < kTaps; ++ty) {
                    const float t = wx[tx] * wy[ty];
                    if (MODE == kModePsf || MODE == kModeDirty) lds_add(P_re, row + ty, w * t);
                    if (dirty_on) {
                        lds_add(A_re, row + ty, vr * t);
                        lds_add(A_im, row + ty, vi * t);
                    }
                }
            }
        } else {
            const int ox = static_cast<int>(pk.y & 0xffu) - 128;
            const int oy = static_cast<int>((pk.y >> 8) & 0xffu) - 128;
            const int plane = static_cast<int>((pk.y >> 16) & 0x7ffu);
            const float csign = (flags & kConjBit) ? -1.0f : 1.0f;
            const int S = gp.supports[plane];
            const float2 *K = gp.kernels + static_cast<uint64_t>(plane) * W * W;
            float pr = 0.0f, pi = 0.0f;
            if (MODE == kModeResidual || MODE == kModeDegrid) {
                for (int iy = -S; iy <= S; ++iy) {
                    const int ky = abs(iy * s + oy);
                    for (int ix = -S; ix <= S; ++ix) {
                        const int kx = abs(ix * s + ox);
                        const float2 c = K[kx * W + ky];
                        const float cr = c.x, ci = -(c.y * csign);  // conj(cwt)
                        const int cell = (ax + ix) * L + (ay + iy);
                        const float gr = M_re[cell], gi = M_im[cell];
                        pr += cr * gr - ci * gi;
                        pi += cr * gi + ci * gr;
                    }
                }
            }
            if (MODE == kModeDegrid) {
                vis_out[index[p]] = make_float2(pr, pi);
                continue;
            }
            float vr, vi;
            if (MODE == kModeResidual) { vr = (v.x - pr) * w; vi = (v.y - pi) * w; }
            else if (MODE == kModeAdjoint) { vr = v.x; vi = v.y; }
            else { vr = v.x * w; vi = v.y * w; }
            for (int iy = -S; iy <= S; ++iy) {
                const int ky = abs(iy * s + oy);
                for (int ix = -S; ix <= S; ++ix) {
                    const int kx = abs(ix * s + ox);
                    const float2 c = K[kx * W + ky];
                    const float cr = c.x, ci = c.y * csign;
                    const int cell = (ax + ix) * L + (ay + iy);
                    if (MODE == kModePsf || MODE == kModeDirty) {
                        lds_add(P_re, cell, w * cr);
                        lds_add(P_im, cell, w * ci);
                    }
                    if (MODE != kModePsf) {
                        lds_add(A_re, cell, vr * cr - vi * ci);
                        lds_add(A_im, cell, vr * ci + vi * cr);
                    }
                }
            }
        }
    }
    if (MODE == kModeDegrid) return;
    __syncthreads();

    // Flush: add the non-zero cells of the LDS tile to the global grid.
    for (int c = tid; c < cells; c += nth) {
        const int gxx = x0 + c / L, gyy = y0 + c % L;
        if (gxx < 0 || gyy < 0 || gxx >= static_cast<int>(gp.gx) || gyy >= static_cast<int>(gp.gy))
            continue;
        const uint64_t gi = 2 * (chan_off + static_cast<uint64_t>(gxx) * gp.gy + gyy);
        if (MODE == kModePsf || MODE == kModeDirty) {
            const float r = P_re[c];
            if (r != 0.0f) unsafeAtomicAdd(psf + gi, r);
            if (WP) {
                const float i = P_im[c];
                if (i != 0.0f) unsafeAtomicAdd(psf + gi + 1, i);
            }
        }
        if (MODE != kModePsf) {
            const float r = A_re[c], i = A_im[c];
            if (r != 0.0f) unsafeAtomicAdd(acc_out + gi, r);
            if (i != 0.0f) unsafeAtomicAdd(acc_out + gi + 1, i);
        }
    }
    (void)narr;
}

template <int MODE, bool WP>
static void launch_binned_t(chip_vset &vs, const GridParams &gp, const float2 *values,
                            const float2 *model, float *psf, float *acc, float2 *vis_out,
                            hipStream_t st) {
    const int tap_floats = WP ? 0 : ((static_cast<int>(gp.tap_rows) * kTaps + 3) & ~3);
    const size_t lds = (static_cast<size_t>(tap_floats) +
                        static_cast<size_t>(LdsPlan<MODE>::arrays) * gp.L * gp.L) * sizeof(float);
    if (lds > 64 * 1024)
        fail(CHIP_ERR_INVALID, "kernel support too large for the LDS tile (halo " +
                                   std::to_string(gp.halo) + ")");
    if (vs.n_items == 0) return;
    // The kernel is bracketed by events on its own stream so callers can read
    // its device duration (chip_vset_last_kernel_ms).
    CHIP_HIP(hipEventRecord(vs.ev_start, st));
    hipLaunchKernelGGL((k_binned<MODE, WP>), dim3(static_cast<unsigned>(vs.n_items)), dim3(256),
                       lds, st, gp, vs.items.as<WorkItem>(), vs.plan.as<uint2>(),
                       vs.vis.as<float2>(), vs.wt.as<float>(), vs.index.as<uint32_t>(), values,
                       model, psf, acc, vis_out);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipEventRecord(vs.ev_stop, st));
    vs.timed = true;
}

template <int MODE>
static void launch_binned(chip_vset &vs, const float2 *values, const float2 *model,
                          float *psf, float *acc, float2 *vis_out) {
    const chip_gridder &g = *vs.g;
    const GridParams gp = grid_params(g, vs.tile, vs.halo, vs.tiles_y);
    if (vs.n_items > 0x7fffffffull) fail(CHIP_ERR_INVALID, "too many work items");
    if (g.wproject)
        launch_binned_t<MODE, true>(vs, gp, values, model, psf, acc, vis_out, g.ctx->stream);
    else
        launch_binned_t<MODE, false>(vs, gp, values, model, psf, acc, vis_out, g.ctx->stream);
}

void vset_grid(chip_vset &vs, uint32_t mode, const float *model, float *psf, float *dirty) {
    const float2 *m = reinterpret_cast<const float2 *>(model);
    switch (mode) {
        case CHIP_MODE_PSF:
            if (!psf) fail(CHIP_ERR_INVALID, "PSF mode needs a PSF grid");
            launch_binned<kModePsf>(vs, nullptr, nullptr, psf, nullptr, nullptr);
            break;
        case CHIP_MODE_DIRTY:
            if (!psf || !dirty) fail(CHIP_ERR_INVALID, "dirty mode needs PSF and dirty grids");
            launch_binned<kModeDirty>(vs, nullptr, nullptr, psf, dirty, nullptr);
            break;
        case CHIP_MODE_RESIDUAL:
            if (!dirty) fail(CHIP_ERR_INVALID, "residual mode needs a residual grid");
            launch_binned<kModeResidual>(vs, nullptr, m, nullptr, dirty, nullptr);
            break;
        default:
            fail(CHIP_ERR_INVALID, "unknown grid mode");
    }
}

void vset_degrid(chip_vset &vs, const float *model, float *vis_out) {
    if (!model || !vis_out) fail(CHIP_ERR_INVALID, "degrid needs a model grid and an output");
    CHIP_HIP(hipMemsetAsync(vis_out, 0, vs.n_samples * sizeof(float2), vs.g->ctx->stream));
    launch_binned<kModeDegrid>(vs, nullptr, reinterpret_cast<const float2 *>(model), nullptr,
                               nullptr, reinterpret_cast<float2 *>(vis_out));
}

void vset_grid_values(chip_vset &vs, const float *values, float *grid) {
    if (!values || !grid) fail(CHIP_ERR_INVALID, "adjoint gridding needs values and a grid");
    launch_binned<kModeAdjoint>(vs, reinterpret_cast<const float2 *>(values), nullptr, nullptr,
                                grid, nullptr);
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
                                               float2 *vis_out, double *sums) {
    __shared__ double red[3][4];
    const uint64_t row = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double nsum = 0.0, rsum = 0.0, cnt = 0.0;
    if (row < s.n_rows) {
        const uint64_t i = row * s.n_chan + chan;
        const SampleIn smp = load_sample(s, i);
        Plan pl;
        bool ok = sample_accepted(pp, smp, pl);
        const uint32_t flags = ok ? sample_flags(pp, smp) : 0u;
        if (ok && MODE == kModeResidual && !(flags & kResidOkBit)) ok = false;
        if (MODE == kModeDegrid && !ok) vis_out[i] = make_float2(0.0f, 0.0f);
        if (ok) {
            const float gwf = grid_weight(pp, smp);
            if (pp.wproject) {
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
                    vis_out[i] = make_float2(static_cast<float>(pr), static_cast<float>(pi));
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
            } else {
                const int s_ = static_cast<int>(gp.sampling);
                const int W = static_cast<int>(gp.width);
                const int S = gp.supports[pl.plane];
                const float2 *K = gp.kernels + static_cast<uint64_t>(pl.plane) * W * W;
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
                    vis_out[i] = make_float2(static_cast<float>(pr), static_cast<float>(pi));
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
                           void *acc, float *vis_out, double *sums) {
    const GridParams gp = grid_params(g, 32, 0, 1);
    const PlanParams pp = plan_params(g);
    hipStream_t st = g.ctx->stream;
    const unsigned blocks = static_cast<unsigned>((src.n_rows + 255) / 256);
    if (blocks == 0) return;
    for (uint32_t c = 0; c < src.n_chan; ++c) {
        if (g.wproject)
            hipLaunchKernelGGL((k_plain<T, MODE, true>), dim3(blocks), dim3(256), 0, st, pp, src, c,
                               gp, static_cast<const T *>(model), static_cast<T *>(psf),
                               static_cast<T *>(acc), reinterpret_cast<float2 *>(vis_out), sums);
        else
            hipLaunchKernelGGL((k_plain<T, MODE, false>), dim3(blocks), dim3(256), 0, st, pp, src,
                               c, gp, static_cast<const T *>(model), static_cast<T *>(psf),
                               static_cast<T *>(acc), reinterpret_cast<float2 *>(vis_out), sums);
        CHIP_HIP(hipGetLastError());
    }
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

void degrid_plain(chip_gridder &g, const Source &src, uint32_t precision, const void *model,
                  float *out) {
    if (!model || !out) fail(CHIP_ERR_INVALID, "degrid needs a model grid and an output");
    if (precision == CHIP_F64)
        launch_plain_t<double, kModeDegrid>(g, src, model, nullptr, nullptr, out, nullptr);
    else if (precision == CHIP_F32)
        launch_plain_t<float, kModeDegrid>(g, src, model, nullptr, nullptr, out, nullptr);
    else
        fail(CHIP_ERR_INVALID, "unknown precision");
}

}  // namespace chip
