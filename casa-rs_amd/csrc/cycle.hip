// The major-cycle residual refresh of every channel of a visibility set in
// one call (compute_residual_w_project_replay, lib.rs:70895-70963;
// compute_residual_standard_tiled, lib.rs:68862-68932):
//
//   model images -> apodise, centered forward FFT -> degrid, grid (V - p) w
//   -> centered inverse FFT -> crop, correct, / sumwt / psf peak.
//
// Everything stays in a workspace owned by the set.  The centered
// transforms cost no extra passes: the input checkerboard of the forward
// transform is folded into the placement of the model, its output
// checkerboard into the binned kernel's model-tile load, and the inverse
// transform's shifts into the crop (shift theorem: IFFT((-1)^k G)[x] =
// IFFT(G)[x + N/2], so centered_ifft(G)[x] = (-1)^(x + N/2) IFFT(G)[(x + N/2)
// mod N] / N).
//
// Two transform strategies:
//   fused  : the row / column FFT passes of xform.hip, with placement,
//            apodisation, Hermitian half-spectrum, crop and normalisation
//            folded into the passes (default whenever the grid factors into
//            2, 3, 5 and fits the LDS);
//   hipFFT : rocFFT 2-D real-to-complex / complex-to-real plans over
//            [chan][x][y] (the model is real; the residual image depends only
//            on the Hermitian part of the residual grid) with separate
//            placement and crop kernels.
#include <algorithm>
#include <cmath>

#include "device.hpp"
#include "xform.hpp"

namespace chip {

void vset_residual_with(chip_vset &vs, const GridParams &gp, const float2 *model, float *res, hipStream_t st,
                        bool timed);
void vset_split_items(chip_vset &vs, uint32_t gpi);
bool binned_two_groups_allowed(const chip_vset &vs);
void vset_rebin(chip_vset &vs, uint32_t tile, uint32_t order, bool force = false);

namespace {

struct CycleGeom {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t base;  // gx/2 + gy/2
    uint32_t wproject;
    uint64_t cstride;
    const float *corr_x, *corr_y;
    const float *sinc;
    const float *sinc_x, *sinc_y;
};

// Model (chan, x, y) -> real grid cell (blc + x, blc + y), times the
// apodisation (gridder.rs:1554-1592) and (-1)^(gx + gy).  Grid: (ceil(ny/256),
// nx, chan).
__global__ void k_cycle_place(CycleGeom cg, const float *__restrict__ model, float *grid) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t x = blockIdx.y;
    const uint32_t c = blockIdx.z;
    if (y >= cg.ny) return;
    const uint32_t gxx = cg.blc_x + x, gyy = cg.blc_y + y;
    float f = cg.corr_x[gxx] * cg.corr_y[gyy];
    if (cg.wproject) f = f * cg.sinc_x[gxx] * cg.sinc_y[gyy];
    float v = model[(static_cast<uint64_t>(c) * cg.nx + x) * cg.ny + y] * f;
    if ((gxx + gyy) & 1) v = -v;
    grid[c * cg.cstride + static_cast<uint64_t>(gxx) * cg.gy + gyy] = v;
}

// Real inverse transform of 2H -> residual image (gridder.rs:1646-1657
// standard, :1685-1715 W; normalisation lib.rs:68920-68932, :70960-70962).
__global__ void k_cycle_crop(CycleGeom cg, const float *__restrict__ raw, const float *__restrict__ nsum,
                             const float *__restrict__ peak, float *out) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t x = blockIdx.y;
    const uint32_t c = blockIdx.z;
    if (y >= cg.ny) return;
    const uint32_t gxx = cg.blc_x + x, gyy = cg.blc_y + y;
    uint32_t sx = gxx + cg.gx / 2, sy = gyy + cg.gy / 2;
    if (sx >= cg.gx) sx -= cg.gx;
    if (sy >= cg.gy) sy -= cg.gy;
    const float scale = 1.0f / static_cast<float>(static_cast<uint64_t>(cg.gx) * cg.gy);
    float re = (0.5f * raw[c * cg.cstride + static_cast<uint64_t>(sx) * cg.gy + sy]) * scale;
    if ((gxx + gyy + cg.base) & 1) re = -re;
    const float factor = cg.corr_x[gxx] * cg.corr_y[gyy];
    float v;
    if (!cg.wproject) {
        v = re * factor;
        v = v / nsum[c] / peak[c];
    } else {
        const float sf = cg.sinc[gxx] * cg.sinc[gyy];
        v = fabsf(sf) > 1.0e-6f ? re * (factor / sf) : 0.0f;
        v = 2.0f * v / nsum[c] / peak[c];
    }
    out[(static_cast<uint64_t>(c) * cg.nx + x) * cg.ny + y] = v;
}

// uv-coverage mask of the fused layout (XGeom::touch): one thread per work
// item marks the row chunks of the upper blocks that its tile and halo,
// rectangle [xa, xb] x [ya, yb] of the full grid, reads from the model grid
// or adds to the residual grid: cells y <= gy/2 in place, and cells y >= gy/2
// or y == 0 at their Hermitian mirror ((gx - x) mod gx, (gy - y) mod gy),
// where the binned kernel reads the model's conjugate and adds the residual's.
struct TouchGeom {
    uint32_t gx, gy, tile, halo, tiles_y, nyb_h, nxc, n_chan;
};

// t: the channel's chunk flags [nyb_h][nxc], followed at tb by its block flags [nyb_h]
struct TouchOut {
    uint8_t *t, *tb;
};

__device__ void touch_mark(const TouchOut &o, const TouchGeom &tg, uint32_t ylo, uint32_t yhi, uint32_t xlo,
                           uint32_t xhi) {
    for (uint32_t b = ylo / xf::kBlock; b <= yhi / xf::kBlock; ++b) {
        o.tb[b] = 1;
        for (uint32_t ch = xlo / xf::kTouchRows; ch <= xhi / xf::kTouchRows; ++ch) o.t[b * tg.nxc + ch] = 1;
    }
}

// rows x' = (gx - x) mod gx of x in [xa, xb]
__device__ void touch_mark_mirror(const TouchOut &t, const TouchGeom &tg, uint32_t ylo, uint32_t yhi, uint32_t xa,
                                  uint32_t xb) {
    if (xa == 0) {
        touch_mark(t, tg, ylo, yhi, 0, 0);
        if (xb >= 1) touch_mark(t, tg, ylo, yhi, tg.gx - xb, tg.gx - 1);
    } else {
        touch_mark(t, tg, ylo, yhi, tg.gx - xb, tg.gx - xa);
    }
}

__global__ void k_touch(const WorkItem *__restrict__ items, uint64_t n_items, TouchGeom tg, uint8_t *touch) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n_items) return;
    const WorkItem it = items[i];
    if (it.count == 0) return;  // padding item
    const uint32_t tx = it.tile / tg.tiles_y, ty = it.tile - tx * tg.tiles_y;
    const int64_t L = tg.tile + 2 * tg.halo;
    const int64_t x0 = static_cast<int64_t>(tx) * tg.tile - tg.halo, y0 = static_cast<int64_t>(ty) * tg.tile - tg.halo;
    const int64_t xa = x0 < 0 ? 0 : x0, ya = y0 < 0 ? 0 : y0;
    const int64_t xb = std::min<int64_t>(x0 + L - 1, tg.gx - 1), yb = std::min<int64_t>(y0 + L - 1, tg.gy - 1);
    if (xa > xb || ya > yb) return;
    const uint32_t chan = it.chan & 0xffffffu;  // (top 8 bits: the item's largest half-support)
    const TouchOut t{touch + static_cast<uint64_t>(chan) * tg.nyb_h * tg.nxc,
                     touch + static_cast<uint64_t>(tg.n_chan) * tg.nyb_h * tg.nxc +
                         static_cast<uint64_t>(chan) * tg.nyb_h};
    const int64_t hy = tg.gy / 2;
    const uint32_t uxa = static_cast<uint32_t>(xa), uxb = static_cast<uint32_t>(xb);
    if (ya <= hy) touch_mark(t, tg, static_cast<uint32_t>(ya), static_cast<uint32_t>(std::min(yb, hy)), uxa, uxb);
    if (yb >= hy) {
        const uint32_t ylo = static_cast<uint32_t>(tg.gy - yb), yhi = static_cast<uint32_t>(tg.gy - std::max(ya, hy));
        touch_mark_mirror(t, tg, ylo, yhi, uxa, uxb);
    }
    if (ya == 0) touch_mark_mirror(t, tg, 0, 0, uxa, uxb);
}

// Schedule autotuner of the binned residual kernel: for every uv tile size
// that fits the LDS with the set's halo (the size the set was built with,
// then 16, 32, 64) with groups ordered by anchor cell, and then for the
// fastest tile with groups ordered by plan, the stream is re-binned and
// every (threads per workgroup, groups per work item) candidate is launched
// on the major cycle's own grids (timed with the set's events); the fastest
// (tile, order, threads, groups) is kept, the earlier candidate on a near
// tie.  make_gp builds the launch parameters for the set's current tiling.
// The grids' contents do not matter: the residual grid is cleared again
// before the real pass.
template <class MakeGp>
void autotune_schedule(chip_vset &vs, MakeGp make_gp, const float2 *M, float *Rg, hipStream_t st) {
    static constexpr uint32_t kThreads[] = {256, 512};
    static constexpr uint32_t kGpi[] = {256, 512, 1024};
    const chip_gridder &g = *vs.g;
    std::vector<uint32_t> tiles{vs.tile};
    for (uint32_t t : {16u, 32u, 64u})
        if (t != vs.tile && binned_tile_lds(g, t, vs.halo) <= lds_limit(vs.device)) tiles.push_back(t);
    vs.tune_ms.clear();
    double best = -1.0;
    uint32_t best_tile = vs.tile, best_order = vs.order, best_t = vs.threads, best_g = vs.groups_per_item;
    // Tuning time is bounded: launches above kLong ms skip the warm-up and
    // are timed once, and no new candidate starts after kBudget ms of
    // launches (the candidates are ordered so that the likely best come first).
    constexpr double kLong = 20.0, kBudget = 3000.0;
    double spent = 0.0;
    for (uint32_t order : {0u, 1u}) {
        if (order == 1) tiles.assign(1, best_tile);  // plan order at the fastest tile only
        for (uint32_t tile : tiles) {
            if (spent > kBudget) break;
            vset_rebin(vs, tile, order);
            const GridParams gp = make_gp();
            for (uint32_t gpi : kGpi) {
                if (spent > kBudget) break;
                vset_split_items(vs, gpi);
                for (uint32_t t : kThreads) {
                    if (spent > kBudget) break;
                    vs.threads = t;
                    // a warm-up launch, then the fastest of up to three timed ones
                    // (one when the first is far behind the best so far)
                    float ms = -1.0f;
                    for (int rep = 0; rep < 4; ++rep) {
                        vset_residual_with(vs, gp, M, Rg, st, true);
                        CHIP_HIP(hipEventSynchronize(vs.ev_stop));
                        float m = 0.0f;
                        CHIP_HIP(hipEventElapsedTime(&m, vs.ev_start, vs.ev_stop));
                        spent += m;
                        if (rep == 0 && m < kLong) continue;
                        ms = ms < 0.0f ? m : std::min(ms, m);
                        if (m >= kLong || (best > 0.0 && ms > 1.25 * best)) break;
                    }
                    vs.tune_ms.insert(vs.tune_ms.end(), {static_cast<double>(tile), static_cast<double>(order),
                                                         static_cast<double>(t), static_cast<double>(gpi), ms});
                    // a later candidate must win by more than 1 % (timing noise)
                    if (best < 0.0 || ms < 0.99 * best) {
                        best = ms;
                        best_tile = tile;
                        best_order = order;
                        best_t = t;
                        best_g = gpi;
                    }
                }
            }
        }
    }
    vset_rebin(vs, best_tile, best_order);
    vset_split_items(vs, best_g);
    vs.threads = best_t;
    // two groups per wavefront, timed at the chosen uv tile, order and groups
    // per item with every thread count (recorded with order + 2: its best
    // thread count can differ from one group's); kept when it wins by more
    // than 1 %
    vs.tuned_wave_groups = 1;
    if (vs.wave_groups == 0 && binned_two_groups_allowed(vs) && spent <= kBudget) {
        const GridParams gp = make_gp();
        vs.tuned_wave_groups = 2;
        double best2 = -1.0;
        uint32_t best2_t = best_t;
        for (uint32_t t : kThreads) {
            vs.threads = t;
            float ms = -1.0f;
            for (int rep = 0; rep < 4; ++rep) {
                vset_residual_with(vs, gp, M, Rg, st, true);
                CHIP_HIP(hipEventSynchronize(vs.ev_stop));
                float m = 0.0f;
                CHIP_HIP(hipEventElapsedTime(&m, vs.ev_start, vs.ev_stop));
                if (rep == 0 && m < kLong) continue;
                ms = ms < 0.0f ? m : std::min(ms, m);
                if (m >= kLong) break;
            }
            vs.tune_ms.insert(vs.tune_ms.end(), {static_cast<double>(best_tile), static_cast<double>(best_order + 2),
                                                 static_cast<double>(t), static_cast<double>(best_g), ms});
            if (best2 < 0.0 || ms < best2) {
                best2 = ms;
                best2_t = t;
            }
        }
        if (best2 > 0.0 && best2 < 0.99 * best) {
            vs.threads = best2_t;
        } else {
            vs.tuned_wave_groups = 1;
            vs.threads = best_t;
        }
    }
    vs.schedule_tuned = true;
}

}  // namespace

// A captured major cycle (fused strategy): the two forward passes, the zero
// fill of the residual grid, the binned residual kernel and the two inverse
// passes, replayed with one hipGraphLaunch.  Captured on a stream of its own
// (the caller's stream may be the legacy default stream, which cannot be
// captured); replayed on the context's stream.  The model and residual
// pointers are baked in: a call with other pointers captures again.
struct CycleGraph {
    hipStream_t cap = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const float *model = nullptr;
    float *out = nullptr;
    void reset() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
    }
    ~CycleGraph() {
        reset();
        if (cap) (void)hipStreamDestroy(cap);
    }
};

struct CycleWorkspace {
    int strategy = -1;  // CHIP_CYCLE_HIPFFT or CHIP_CYCLE_FUSED once set up
    DevBuf model_grid, res_grid;
    DevBuf touch;                        // fused: uv-coverage mask of the work items
    uint64_t touch_version = ~0ull;      // items_version the mask was built for
    hipfftHandle r2c = nullptr, c2r = nullptr;
    xf::XPlan xp;
    double transform_ms = -1.0;
    CycleGraph graph;
    uint64_t replays = 0;  // major cycles run by graph replay
    // per-pass timing of eager fused cycles (chip_vset_cycle_passes): events
    // before, between and after the six passes, and the last cycle's ms
    hipEvent_t pev[7] = {};
    double pass_ms[6] = {-1.0, -1.0, -1.0, -1.0, -1.0, -1.0};
    ~CycleWorkspace() {
        if (r2c) hipfftDestroy(r2c);
        if (c2r) hipfftDestroy(c2r);
        for (hipEvent_t &e : pev)
            if (e) (void)hipEventDestroy(e);
    }
};

namespace {


void build_touch(const chip_vset &vs, CycleWorkspace &ws, hipStream_t st) {
    xf::XGeom &xg = ws.xp.xg;
    const uint64_t chunks = static_cast<uint64_t>(ws.xp.n_chan) * xg.nyb_h * xg.nxc;
    const uint64_t bytes = chunks + static_cast<uint64_t>(ws.xp.n_chan) * xg.nyb_h;
    if (ws.touch.bytes != bytes) ws.touch.alloc(bytes);
    CHIP_HIP(hipMemsetAsync(ws.touch.p, 0, bytes, st));
    if (vs.n_items) {
        const TouchGeom tg{xg.gx, xg.gy, vs.tile, vs.halo, vs.tiles_y, xg.nyb_h, xg.nxc, ws.xp.n_chan};
        hipLaunchKernelGGL(k_touch, dim3(static_cast<unsigned>((vs.n_items + 255) / 256)), dim3(256), 0, st,
                           vs.items.as<WorkItem>(), vs.n_items, tg, ws.touch.as<uint8_t>());
        CHIP_HIP(hipGetLastError());
    }
    xg.touch = ws.touch.as<uint8_t>();
    xg.touch_blk = xg.touch + chunks;
    ws.touch_version = vs.items_version;
}

// Forward + inverse transforms of the fused plan, in ms.  Writes scratch
// values into the workspace and `out` (overwritten afterwards).
double time_fused(chip_vset &vs, CycleWorkspace &ws, const float *model, float *out, hipStream_t st) {
    const xf::XPlan &p = ws.xp;
    float2 *M = ws.model_grid.as<float2>();
    float2 *R = ws.res_grid.as<float2>();
    const float *ones = vs.cycle_norm.as<float>();
    hipEvent_t a, b;
    CHIP_HIP(hipEventCreate(&a));
    CHIP_HIP(hipEventCreate(&b));
    xf::forward_model(p, model, R, M, st);  // warm-up
    CHIP_HIP(hipEventRecord(a, st));
    xf::forward_model(p, model, R, M, st);
    xf::inverse_residual(p, M, R, ones, ones, out, st);
    CHIP_HIP(hipEventRecord(b, st));
    CHIP_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    CHIP_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

void cycle_setup(chip_vset &vs, CycleWorkspace &ws, const float *model, float *out) {
    chip_gridder &g = *vs.g;
    hipStream_t st = g.ctx->stream;
    const uint64_t nc = vs.n_chan;
    bool fused = false;
    if (vs.cycle_strategy != CHIP_CYCLE_HIPFFT) {
        fused = xf::plan_transforms(g, nc, lds_limit(g.device), ws.xp);
        if (!fused && vs.cycle_strategy == CHIP_CYCLE_FUSED)
            fail(CHIP_ERR_INVALID,
                 "grid does not factor into 2, 3, 5 or exceeds one workgroup for fused transforms");
    }
    // hipFFT strategy: half-spectrum grids gx x (gy/2 + 1), whose memory also
    // holds the real gx x gy planes of the placement and of the inverse
    const uint64_t elements = fused ? xf::grid_elements(ws.xp) : nc * g.gx * (g.gy / 2 + 1);
    ws.model_grid.alloc(elements * sizeof(float2));
    ws.res_grid.alloc(elements * sizeof(float2));
    CHIP_HIP(hipMemsetAsync(ws.model_grid.p, 0, ws.model_grid.bytes, st));
    CHIP_HIP(hipMemsetAsync(ws.res_grid.p, 0, ws.res_grid.bytes, st));
    if (fused) {
        ws.strategy = CHIP_CYCLE_FUSED;
        build_touch(vs, ws, st);
        ws.transform_ms = time_fused(vs, ws, model, out, st);
        return;
    }
    if (g.gx * g.gy > 0x7fffffffull || nc > 0x7fffffffull)
        fail(CHIP_ERR_INVALID, "grid too large for a batched 2-D plan");
    // real-to-complex forward of the real model, complex-to-real inverse of
    // the Hermitian 2H: half the work of complex transforms
    int n[2] = {static_cast<int>(g.gx), static_cast<int>(g.gy)};
    int real_em[2] = {static_cast<int>(g.gx), static_cast<int>(g.gy)};
    int half_em[2] = {static_cast<int>(g.gx), static_cast<int>(g.gy / 2 + 1)};
    const int real_dist = static_cast<int>(g.gx * g.gy), half_dist = static_cast<int>(g.gx * (g.gy / 2 + 1));
    size_t work = 0;
    CHIP_FFT(hipfftCreate(&ws.r2c));
    CHIP_FFT(hipfftMakePlanMany(ws.r2c, 2, n, real_em, 1, real_dist, half_em, 1, half_dist, HIPFFT_R2C,
                                static_cast<int>(nc), &work));
    CHIP_FFT(hipfftCreate(&ws.c2r));
    CHIP_FFT(hipfftMakePlanMany(ws.c2r, 2, n, half_em, 1, half_dist, real_em, 1, real_dist, HIPFFT_C2R,
                                static_cast<int>(nc), &work));
    CHIP_FFT(hipfftSetStream(ws.r2c, st));
    CHIP_FFT(hipfftSetStream(ws.c2r, st));
    ws.strategy = CHIP_CYCLE_HIPFFT;
}

}  // namespace

void vset_major_cycle(chip_vset &vs, const float *model, const double *nsum, const float *peaks,
                      float *out) {
    chip_gridder &g = *vs.g;
    if (g.gx % 2 || g.gy % 2) fail(CHIP_ERR_INVALID, "centered transforms need an even grid");
    if (!model || !nsum || !peaks || !out) fail(CHIP_ERR_INVALID, "NULL argument");
    for (uint64_t c = 0; c < vs.n_chan; ++c) {
        if (!(nsum[c] > 0.0)) fail(CHIP_ERR_NO_SAMPLES, "normalization sum of weights is not positive");
        if (!(std::isfinite(peaks[c]) && peaks[c] > 0.0f))
            fail(CHIP_ERR_NORMALIZATION, "PSF peak is non-finite or zero");
    }
    if (vs.n_chan > 65535) fail(CHIP_ERR_INVALID, "more than 65535 channels in one major cycle");
    hipStream_t st = g.ctx->stream;
    const uint64_t nc = vs.n_chan;
    if (!vs.cycle_norm.p) {
        vs.cycle_norm.alloc(2 * nc * sizeof(float));
        const std::vector<float> ones(2 * nc, 1.0f);
        CHIP_HIP(hipMemcpyAsync(vs.cycle_norm.p, ones.data(), vs.cycle_norm.bytes, hipMemcpyHostToDevice, st));
        CHIP_HIP(hipStreamSynchronize(st));
    }
    if (!vs.cycle) {
        auto ws = std::make_shared<CycleWorkspace>();
        cycle_setup(vs, *ws, model, out);
        vs.cycle = ws;
    }
    CycleWorkspace &ws = *vs.cycle;
    if (ws.r2c) CHIP_FFT(hipfftSetStream(ws.r2c, st));  // the context's stream may have changed
    if (ws.c2r) CHIP_FFT(hipfftSetStream(ws.c2r, st));
    if (vs.cycle_graph && ws.strategy != CHIP_CYCLE_FUSED)
        fail(CHIP_ERR_INVALID, "graph replay of the major cycle needs the fused transforms");

    std::vector<float> host(2 * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        host[c] = static_cast<float>(nsum[c]);
        host[nc + c] = peaks[c];
    }
    CHIP_HIP(hipMemcpyAsync(vs.cycle_norm.p, host.data(), 2 * nc * sizeof(float), hipMemcpyHostToDevice, st));
    const float *d_nsum = vs.cycle_norm.as<float>();
    const float *d_peak = d_nsum + nc;

    float2 *M = ws.model_grid.as<float2>();
    float2 *Rg = ws.res_grid.as<float2>();
    const bool fused = ws.strategy == CHIP_CYCLE_FUSED;
    auto make_gp = [&]() {
        GridParams p = grid_params(g, vs.tile, vs.halo, vs.tiles_y);
        p.model_signed = 1;
        p.model_half = 1;     // both strategies store the model's columns y <= gy/2 only
        p.acc_hermitian = 1;  // and accumulate the residual as 2H in those columns
        if (fused) {
            p.paired = 1;
            p.cstride = ws.xp.xg.gstride;
            p.bstride = g.gx * xf::kBlock;
            p.lower = static_cast<uint64_t>(ws.xp.xg.nyb_h) * p.bstride;
        } else {
            p.xstride = g.gy / 2 + 1;  // hipFFT half spectra, x-major
            p.cstride = g.gx * (g.gy / 2 + 1);
        }
        return p;
    };

    if (fused) {
        if (!vs.schedule_fixed && !vs.schedule_tuned && vs.n_items) {
            autotune_schedule(vs, make_gp, M, reinterpret_cast<float *>(Rg), st);
            ws.graph.reset();  // the stream was re-binned
        }
        if (ws.touch_version != vs.items_version) {
            build_touch(vs, ws, st);
            ws.graph.reset();
        }
        const GridParams gp = make_gp();
        // the half spectrum borrows the residual grid, the inverse rows the model grid
        auto enqueue = [&](hipStream_t s, bool timed) {
            xf::forward_model(ws.xp, model, Rg, M, s);
            // the covered rows of the upper blocks (the residual is kept as 2H there)
            xf::zero_touched(ws.xp, Rg, s);
            vset_residual_with(vs, gp, M, reinterpret_cast<float *>(Rg), s, timed);
            xf::inverse_residual(ws.xp, Rg, M, d_nsum, d_peak, out, s);
        };
        if (!vs.cycle_graph && vs.cycle_passes) {
            for (hipEvent_t &e : ws.pev)
                if (!e) CHIP_HIP(hipEventCreate(&e));
            CHIP_HIP(hipEventRecord(ws.pev[0], st));
            xf::forward_model(ws.xp, model, Rg, M, st, ws.pev[1]);
            CHIP_HIP(hipEventRecord(ws.pev[2], st));
            xf::zero_touched(ws.xp, Rg, st);
            CHIP_HIP(hipEventRecord(ws.pev[3], st));
            vset_residual_with(vs, gp, M, reinterpret_cast<float *>(Rg), st, true);
            CHIP_HIP(hipEventRecord(ws.pev[4], st));
            xf::inverse_residual(ws.xp, Rg, M, d_nsum, d_peak, out, st, ws.pev[5]);
            CHIP_HIP(hipEventRecord(ws.pev[6], st));
            CHIP_HIP(hipEventSynchronize(ws.pev[6]));
            for (int k = 0; k < 6; ++k) {
                float ms = 0.0f;
                CHIP_HIP(hipEventElapsedTime(&ms, ws.pev[k], ws.pev[k + 1]));
                ws.pass_ms[k] = ms;
            }
        } else if (!vs.cycle_graph) {
            enqueue(st, true);
        } else {
            CycleGraph &cg = ws.graph;
            if (!cg.exec || cg.model != model || cg.out != out) {
                cg.reset();
                if (!cg.cap) CHIP_HIP(hipStreamCreateWithFlags(&cg.cap, hipStreamNonBlocking));
                CHIP_HIP(hipStreamBeginCapture(cg.cap, hipStreamCaptureModeRelaxed));
                try {
                    enqueue(cg.cap, false);
                } catch (...) {
                    hipGraph_t partial = nullptr;
                    (void)hipStreamEndCapture(cg.cap, &partial);
                    if (partial) (void)hipGraphDestroy(partial);
                    throw;
                }
                CHIP_HIP(hipStreamEndCapture(cg.cap, &cg.graph));
                CHIP_HIP(hipGraphInstantiate(&cg.exec, cg.graph, nullptr, nullptr, 0));
                cg.model = model;
                cg.out = out;
            }
            CHIP_HIP(hipGraphLaunch(cg.exec, st));
            ws.replays += 1;
        }
    } else {
        CycleGeom cg{};
        cg.nx = static_cast<uint32_t>(g.geom.nx);
        cg.ny = static_cast<uint32_t>(g.geom.ny);
        cg.gx = static_cast<uint32_t>(g.gx);
        cg.gy = static_cast<uint32_t>(g.gy);
        cg.blc_x = static_cast<uint32_t>(g.blc_x);
        cg.blc_y = static_cast<uint32_t>(g.blc_y);
        cg.base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
        cg.wproject = g.wproject ? 1u : 0u;
        cg.cstride = g.gx * g.gy;  // real planes
        cg.corr_x = g.d_corr_x.as<float>();
        cg.corr_y = g.d_corr_y.as<float>();
        cg.sinc = g.wproject ? g.d_sinc.as<float>() : nullptr;
        cg.sinc_x = g.wproject ? g.d_sinc_x.as<float>() : nullptr;
        cg.sinc_y = g.wproject ? g.d_sinc_y.as<float>() : nullptr;
        // real placement in the residual buffer -> R2C half spectrum in M;
        // residual 2H accumulated in the residual buffer -> C2R real planes in M
        float *Rr = reinterpret_cast<float *>(Rg), *Mr = reinterpret_cast<float *>(M);
        const uint64_t real_bytes = nc * g.gx * g.gy * sizeof(float);
        const uint64_t half_bytes = nc * g.gx * (g.gy / 2 + 1) * sizeof(float2);
        CHIP_HIP(hipMemsetAsync(Rr, 0, real_bytes, st));
        const dim3 img_grid((cg.ny + 255) / 256, cg.nx, static_cast<unsigned>(nc));
        hipLaunchKernelGGL(k_cycle_place, img_grid, dim3(256), 0, st, cg, model, Rr);
        CHIP_HIP(hipGetLastError());
        CHIP_FFT(hipfftExecR2C(ws.r2c, reinterpret_cast<hipfftReal *>(Rr), reinterpret_cast<hipfftComplex *>(M)));
        if (!vs.schedule_fixed && !vs.schedule_tuned && vs.n_items)
            autotune_schedule(vs, make_gp, M, reinterpret_cast<float *>(Rg), st);
        CHIP_HIP(hipMemsetAsync(Rg, 0, half_bytes, st));
        vset_residual_with(vs, make_gp(), M, reinterpret_cast<float *>(Rg), st, true);
        CHIP_FFT(hipfftExecC2R(ws.c2r, reinterpret_cast<hipfftComplex *>(Rg), reinterpret_cast<hipfftReal *>(Mr)));
        hipLaunchKernelGGL(k_cycle_crop, img_grid, dim3(256), 0, st, cg, Mr, d_nsum, d_peak, out);
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
}

void vset_cycle_graph(chip_vset &vs, int enable) {
    if (enable && vs.wk_chunk)
        fail(CHIP_ERR_INVALID, "W tables recomputed on the fly cannot be captured into a graph");
    vs.cycle_graph = enable != 0;
    if (vs.cycle) vs.cycle->graph.reset();
}

uint64_t vset_cycle_replays(const chip_vset &vs) { return vs.cycle ? vs.cycle->replays : 0; }

void vset_cycle_passes(chip_vset &vs, int enable, double *pass_ms) {
    if (enable == 0 || enable == 1) vs.cycle_passes = enable == 1;
    if (pass_ms)
        for (int k = 0; k < 6; ++k) pass_ms[k] = vs.cycle ? vs.cycle->pass_ms[k] : -1.0;
}

void vset_schedule(chip_vset &vs, uint32_t tile, uint32_t order, uint32_t threads, uint32_t gpi) {
    if (tile == 0 && threads == 0 && gpi == 0) {  // autotune again at the next major cycle
        vs.schedule_fixed = false;
        vs.schedule_tuned = false;
    } else {
        if (threads == 0 || threads > 512 || threads % 64)
            fail(CHIP_ERR_INVALID, "threads per workgroup must be a multiple of 64 up to 512");
        if (gpi == 0 || gpi > (1u << 20)) fail(CHIP_ERR_INVALID, "groups per work item must be in 1 .. 2^20");
        vset_rebin(vs, tile ? tile : vs.tile, order);
        vset_split_items(vs, gpi);
        vs.threads = threads;
        vs.schedule_fixed = true;
    }
    if (vs.cycle) vs.cycle->graph.reset();
}

void vset_cycle_strategy(chip_vset &vs, int strategy) {
    if (strategy != CHIP_CYCLE_AUTO && strategy != CHIP_CYCLE_HIPFFT && strategy != CHIP_CYCLE_FUSED)
        fail(CHIP_ERR_INVALID, "unknown major-cycle strategy");
    vs.cycle_strategy = strategy;
    vs.cycle.reset();  // the next call builds the workspace for the new strategy
}

void vset_cycle_info(const chip_vset &vs, int *strategy, uint32_t shape[3], double *transform_ms) {
    shape[0] = shape[1] = shape[2] = 0;
    if (!vs.cycle) {
        *strategy = -1;
        *transform_ms = -1.0;
        return;
    }
    *strategy = vs.cycle->strategy;
    *transform_ms = vs.cycle->transform_ms;
    if (vs.cycle->strategy == CHIP_CYCLE_FUSED) {
        shape[0] = 1u << vs.cycle->xp.logb_x;
        shape[1] = xf::kThreads;
        shape[2] = xf::kRowsPerGroup;
    }
}

}  // namespace chip
