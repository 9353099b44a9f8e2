// Binned FP32 gridding and degridding over the resident visibility stream.
//
// One workgroup per work item: a run of groups (samples with identical
// plans) of one channel whose anchor cells fall in one uv tile (32x32; 16x16
// or 8x8 when a large W support would not fit the LDS).  The workgroup owns
// an LDS tile of (tile + 2*halo)^2 cells (halo = largest half-support; above
// 64 KB the kernel asks for the device's full LDS per workgroup), accumulates
// into it, and finally adds the non-zero LDS cells
// to the global grid with global float atomics (few per tile).  Degridding
// reads a model tile staged in LDS.
//
// Work is tap-parallel: one wavefront handles one group at a time and its 64
// lanes own distinct taps, so the lanes of one LDS atomic always hit distinct
// cells, consecutive lanes hit consecutive cells of a row, and a group's W
// taps are one contiguous pre-gathered run (coalesced loads).  Each wavefront
// walks a contiguous run of groups; groups are sorted by anchor cell, so
// consecutive groups mostly share their footprint and the lanes keep the
// footprint's accumulators (and model values) in registers, touching LDS only
// when the anchor moves.  The degridded value is reduced across the wavefront
// with DPP lane swaps.  MFMA is not used: gridding is a scatter, not a dense
// contraction.
//
// Arithmetic per tap follows the reference (gridder.rs:1036-1050 standard,
// :3680-3747 W): tap = wx*wy (f32) or cwt = K[|ix*s+ox|][|iy*s+oy|] conjugated
// when w > 0; grid += value*tap; degrid += conj(cwt)*grid.  A group's value is
// the sum over its samples (sum w_i V_i; residual: sum w_i V_i - p sum w_i).
#include <climits>
#include <type_traits>

#include "device.hpp"

namespace chip {

GridParams grid_params(const chip_gridder &g, uint32_t tile, uint32_t halo, uint32_t tiles_y) {
    GridParams gp{};
    gp.gx = static_cast<uint32_t>(g.gx);
    gp.gy = static_cast<uint32_t>(g.gy);
    gp.tile = tile;
    gp.halo = halo;
    gp.L = tile + 2 * halo;
    gp.Ls = gp.L | 1u;
    gp.tiles_y = tiles_y;
    gp.sampling = g.sampling;
    gp.width = g.width;
    gp.tap_rows = g.os + 1;
    gp.n_off = static_cast<uint32_t>(2 * g.norm_radius + 1);
    gp.norm_radius = g.norm_radius;
    gp.tap_stride = g.tap_stride;
    gp.cells = g.gx * g.gy;
    gp.cstride = gp.cells;
    gp.xstride = g.gy;
    gp.bstride = 0;
    gp.lower = 0;
    gp.paired = 0;
    gp.hy = static_cast<uint32_t>(g.gy / 2);
    gp.model_signed = 0;
    gp.model_half = 0;
    gp.acc_hermitian = 0;
    gp.sign_base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
    gp.taps = g.d_taps.as<float>();
    gp.kernels = g.wproject ? g.d_kernels.as<float2>() : nullptr;
    gp.compact = g.wproject ? g.d_compact.as<float2>() : nullptr;
    gp.supports = g.wproject ? g.d_supports.as<int32_t>() : nullptr;
    gp.plane_lo = 0;
    gp.plane_hi = UINT32_MAX;
    gp.aw_norm = g.aw ? g.d_aw_tnorm.as<float2>() : nullptr;
    return gp;
}

// Number of L*Ls float arrays a mode keeps in LDS.
template <int MODE>
struct LdsArrays {
    static constexpr int n = (MODE == kModeDirty || MODE == kModeResidual) ? 4 : 2;
};

template <int MODE, bool WP>
static size_t binned_lds_bytes(const GridParams &gp) {
    const size_t tap_floats = WP ? 0 : ((static_cast<size_t>(gp.tap_rows) * kTaps + 3) & ~size_t(3));
    return (tap_floats + static_cast<size_t>(LdsArrays<MODE>::n) * gp.L * gp.Ls) * sizeof(float);
}

size_t binned_tile_lds(const chip_gridder &g, uint32_t tile, uint32_t halo) {
    const size_t L = tile + 2 * halo, Ls = L | 1u;
    const size_t tap_floats = g.wproject ? 0 : ((static_cast<size_t>(g.os + 1) * kTaps + 3) & ~size_t(3));
    return (tap_floats + 4 * L * Ls) * sizeof(float);
}

// Most threads per workgroup (8 wavefronts per work item); the schedule
// autotuner may launch fewer (any multiple of 64).
constexpr int kBinnedThreads = 512;
constexpr int kBig = 8;    // taps per lane per chunk of the large-footprint path (512 per chunk)

// z / conj(n) as z * conj(conj(n)) / |n|^2 with one hardware reciprocal of
// |n|^2 (about 1 ulp) in place of two divisions: the binned kernels are
// bound by instruction issue and their sums are not bit-reproducible anyway
// (atomic order); the plain gridders keep the reference's division.
__device__ __forceinline__ void div_conj_rcp(float &re, float &im, float nr, float ni) {
    const float c = nr, d = -ni;
    const float inv = __builtin_amdgcn_rcpf(c * c + d * d);
    const float a = re, b = im;
    re = (a * c + b * d) * inv;
    im = (b * c - a * d) * inv;
}

// Model value of grid cell (gxx, gyy) of one channel for an LDS tile: zero
// off the grid; read through the half-spectrum mirror (conjugated) and the
// checkerboard sign of the major cycle's grid layout.
__device__ __forceinline__ float2 tile_model(const GridParams &gp, const float2 *__restrict__ model,
                                             uint64_t chan_off, int gxx, int gyy) {
    float2 m = make_float2(0.0f, 0.0f);
    if (!model || gxx < 0 || gyy < 0 || gxx >= static_cast<int>(gp.gx) || gyy >= static_cast<int>(gp.gy)) return m;
    if (gp.model_half && gyy > static_cast<int>(gp.hy)) {
        const uint32_t mx = gxx == 0 ? 0u : gp.gx - static_cast<uint32_t>(gxx);
        m = model[chan_off + cell_offset(gp, mx, gp.gy - static_cast<uint32_t>(gyy))];
        m.y = -m.y;
    } else {
        m = model[chan_off + cell_offset(gp, gxx, gyy)];
    }
    if (gp.model_signed && ((gxx + gyy + gp.sign_base) & 1)) {
        m.x = -m.x;
        m.y = -m.y;
    }
    return m;
}

// Adds the non-zero cells of an LDS tile (origin (x0, y0), L x Ls cells) to
// the global grids with float atomics; with acc_hermitian the accumulator
// goes in as 2H (the cell and, for stored mirror columns, its conjugate at
// (-x, -y)).
template <bool kPsf, bool kAcc>
__device__ __forceinline__ void flush_tile(const GridParams &gp, uint64_t chan_off, int x0, int y0, int L, int Ls,
                                           const float *P_re, const float *P_im, const float *A_re,
                                           const float *A_im, float *psf, float *acc_out) {
    const int cells = L * Ls;
    for (int c = threadIdx.x; c < cells; c += blockDim.x) {
        const int lx = c / Ls, ly = c - lx * Ls;
        const int gxx = x0 + lx, gyy = y0 + ly;
        if (ly >= L || gxx < 0 || gyy < 0 || gxx >= static_cast<int>(gp.gx) || gyy >= static_cast<int>(gp.gy))
            continue;
        const uint64_t gi = 2 * (chan_off + cell_offset(gp, gxx, gyy));
        if (kPsf) {
            const float r = P_re[c], i = P_im[c];
            if (r != 0.0f) unsafeAtomicAdd(psf + gi, r);
            if (i != 0.0f) unsafeAtomicAdd(psf + gi + 1, i);
        }
        if (kAcc) {
            const float r = A_re[c], i = A_im[c];
            if (!gp.acc_hermitian) {
                if (r != 0.0f) unsafeAtomicAdd(acc_out + gi, r);
                if (i != 0.0f) unsafeAtomicAdd(acc_out + gi + 1, i);
            } else if (r != 0.0f || i != 0.0f) {
                const uint32_t hy = gp.hy, ux = static_cast<uint32_t>(gxx), uy = static_cast<uint32_t>(gyy);
                const uint32_t mx = ux == 0 ? 0u : gp.gx - ux;
                if (uy <= hy) {
                    unsafeAtomicAdd(acc_out + gi, r);
                    unsafeAtomicAdd(acc_out + gi + 1, i);
                }
                if (uy == 0 || uy >= hy) {  // conj v to (-x, -y) when that is a stored column
                    const uint64_t mi = 2 * (chan_off + cell_offset(gp, mx, uy == 0 ? 0u : gp.gy - uy));
                    unsafeAtomicAdd(acc_out + mi, r);
                    unsafeAtomicAdd(acc_out + mi + 1, -i);
                }
            }
        }
    }
}

// Register slots per lane for a footprint side M = 2 halo + 1 with 64 lanes:
// ceil(M / floor(64 / M)) rounded up to an instantiated count (1, 2, 4, 7,
// 16), 0 (LDS-atomic chunks) beyond 16 (halo > 15).
__host__ __device__ __forceinline__ int slots_for_halo(int halo) {
    const int m = 2 * halo + 1, p = 64 / m;
    if (p == 0) return 0;
    const int need = (m + p - 1) / p;
    return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : need <= 7 ? 7 : need <= 16 ? 16 : 0;
}

// Two groups per wavefront (GPW = 2): the lanes of each half-wavefront own
// the register slots of one group's footprint, so the per-group work that
// does not scale with the footprint (group record broadcast, footprint,
// slot bookkeeping, the degrid reduction) serves two groups at once.  Each
// half walks its own contiguous run of the wavefront's groups.  Used for
// halos up to 4 (a 9 x 9 footprint in 27 of 32 lanes with 3 slots each).
template <int MODE, bool WP, int KS, int GPW>
__global__ __launch_bounds__(kBinnedThreads) void k_binned(GridParams gp, const WorkItem *__restrict__ items,
                                                const uint4 *__restrict__ grec,
                                                const float4 *__restrict__ gsum,
                                                const uint2 *__restrict__ plan,
                                                const float2 *__restrict__ vis,
                                                const float *__restrict__ wt,
                                                const uint32_t *__restrict__ index,
                                                const float2 *__restrict__ values,
                                                const float2 *__restrict__ model, float *psf,
                                                float *acc_out, float2 *vis_out) {
    extern __shared__ float lds[];
    const WorkItem it = items[blockIdx.x];
    if (it.count == 0) return;  // padding of the XCD-aligned item order
    const int L = static_cast<int>(gp.L);
    const int Ls = static_cast<int>(gp.Ls);
    const int cells = L * Ls;
    const int tap_floats = WP ? 0 : ((static_cast<int>(gp.tap_rows) * kTaps + 3) & ~3);
    float *taps = lds;
    float *a0 = lds + tap_floats;
    float *a1 = a0 + cells;
    float *a2 = a1 + cells;
    float *a3 = a2 + cells;
    // the staged model is interleaved (re, im) over the first two arrays: one
    // 64-bit LDS read per cell
    float *P_re = nullptr, *P_im = nullptr, *A_re = nullptr, *A_im = nullptr;
    float2 *M2 = nullptr;
    if (MODE == kModePsf) { P_re = a0; P_im = a1; }
    if (MODE == kModeDirty) { P_re = a0; P_im = a1; A_re = a2; A_im = a3; }
    if (MODE == kModeResidual) { M2 = reinterpret_cast<float2 *>(a0); A_re = a2; A_im = a3; }
    if (MODE == kModeAdjoint) { A_re = a0; A_im = a1; }
    if (MODE == kModeDegrid) { M2 = reinterpret_cast<float2 *>(a0); }
    constexpr bool kModel = MODE == kModeResidual || MODE == kModeDegrid;
    constexpr bool kAcc = MODE == kModeResidual || MODE == kModeDirty || MODE == kModeAdjoint;
    constexpr bool kPsf = MODE == kModePsf || MODE == kModeDirty;
    // gridding modes read each group's precomputed sums instead of its samples
    constexpr bool kSums = MODE == kModePsf || MODE == kModeDirty || MODE == kModeResidual;

    static_assert(GPW == 1 || (GPW == 2 && KS > 0), "two groups per wavefront need register slots");
    constexpr int LW = 64 / GPW;  // lanes per group
    const int tid = threadIdx.x;
    const int nth = blockDim.x;
    const int lane = tid & 63;
    const int hl = lane & (LW - 1);             // lane within its group's lanes
    const bool upper = GPW == 2 && lane >= 32;  // the half that runs the second group run
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nwaves = nth >> 6;
    const uint32_t tile_x = it.tile / gp.tiles_y;
    const uint32_t tile_y = it.tile - tile_x * gp.tiles_y;
    const int x0 = static_cast<int>(tile_x * gp.tile) - static_cast<int>(gp.halo);
    const int y0 = static_cast<int>(tile_y * gp.tile) - static_cast<int>(gp.halo);
    const uint64_t chan_off = static_cast<uint64_t>(it.chan & 0xffffffu) * gp.cstride;  // top 8 bits: item halo

    if (!WP)
        for (int i = tid; i < static_cast<int>(gp.tap_rows) * kTaps; i += nth) taps[i] = gp.taps[i];
    for (int c = tid; c < cells; c += nth) {
        if (kModel) {
            const int lx = c / Ls, ly = c - lx * Ls;
            M2[c] = ly < L ? tile_model(gp, model, chan_off, x0 + lx, y0 + ly) : make_float2(0.0f, 0.0f);
        }
        if (kAcc) { A_re[c] = 0.0f; A_im[c] = 0.0f; }
        if (kPsf) { P_re[c] = 0.0f; P_im[c] = 0.0f; }
    }
    __syncthreads();

    // Register slots (KS > 0).  With M = 2 halo + 1 and P = floor(64 / M),
    // lane l < P M owns the LDS-tile cells (x, y) with y = l % M and
    // x = l / M + k P (modulo M), one x residue per slot k: every footprint
    // (side <= M) has exactly one cell per (lane, slot).  A slot accumulates
    // in registers while the taps of successive groups land on the same cell
    // (and keeps that cell's model value staged), and spills to the LDS tile
    // with one atomic per component when its cell changes.  Groups are sorted
    // by anchor cell, so consecutive footprints overlap and most slots keep
    // their cell (J. W. Romein, "An efficient work-distribution strategy for
    // gridding radio-telescope data on GPUs", ICS 2012).  A lane's y residue
    // is fixed, so per group the y offset is computed once and the x offsets
    // of its slots follow by steps of P.
    // The slot layout follows the work item's own largest half-support (W /
    // AW: the top 8 bits of its channel word), at most the launch's halo:
    // items are launched by support class, each class with the LDS tile and
    // register slots of its largest half-support.
    const int item_halo =
        WP ? min(static_cast<int>(it.chan >> 24), static_cast<int>(gp.halo)) : static_cast<int>(gp.halo);
    auto run = [&](auto ks_tag) {
    constexpr int KSV = decltype(ks_tag)::value;
    constexpr int NS = KSV > 0 ? KSV : 1;
    // Many slots (large supports): incremental slot offsets and fused
    // multiply-add chains issue fewer instructions; with one or two slots the
    // plain forms measured faster.
    constexpr bool kChains = NS > 2;
    // Slots keep their cell's model value in registers, except with 16 slots,
    // where those 32 registers cost a wavefront per SIMD and the model is read
    // from LDS at every group instead.
    constexpr bool kStageModel = NS < 16;
    const int Mside = 2 * item_halo + 1;
    const float invM = 1.0f / static_cast<float>(Mside);
    const int Pl = LW / Mside;
    const int lane_pg = static_cast<int>((static_cast<float>(hl) + 0.5f) * invM);  // hl / M
    const int lane_py = hl - lane_pg * Mside;
    uint32_t slot_mask = 0u;  // bit k: slot k owns an x residue
    int cell_k[NS];
    float accr[NS], acci[NS], pacr[NS], paci[NS], mre[NS], mim[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        if (KSV > 0 && lane_pg < Pl && lane_pg + k * Pl < Mside) slot_mask |= 1u << k;
        cell_k[k] = -1;
        accr[k] = acci[k] = pacr[k] = paci[k] = mre[k] = mim[k] = 0.0f;
    }
    auto spill = [&](int k) {
        if (cell_k[k] >= 0) {
            if (kPsf) {
                if (pacr[k] != 0.0f) atomicAdd(P_re + cell_k[k], pacr[k]);
                if (WP && paci[k] != 0.0f) atomicAdd(P_im + cell_k[k], paci[k]);
            }
            if (kAcc) {
                if (accr[k] != 0.0f) atomicAdd(A_re + cell_k[k], accr[k]);
                if (acci[k] != 0.0f) atomicAdd(A_im + cell_k[k], acci[k]);
            }
        }
        accr[k] = acci[k] = pacr[k] = paci[k] = 0.0f;
    };

    const uint32_t per = (it.count + nwaves - 1) / nwaves;
    const uint32_t g_begin = it.start + static_cast<uint32_t>(wave) * per;
    const uint32_t g_end = min(it.start + it.count, g_begin + per);
    // this lane's run of groups [h_begin, h_end) and the loop trip count
    uint32_t h_begin = g_begin, h_end = g_end;
    uint32_t n_iter = g_end > g_begin ? g_end - g_begin : 0u;
    if (GPW == 2) {
        const uint32_t n0 = (n_iter + 1) / 2;
        h_begin = g_begin + (upper ? n0 : 0u);
        h_end = upper ? g_end : g_begin + n0;
        n_iter = n0;
    }

    // Footprint of a group from its packed plan (wavefront-uniform).
    struct Foot {
        int side, bx, by;
        const float2 *ktap;
        const float *wxr, *wyr;
        float csign;
        float2 awn;  // AWProject normalisation: degridded values are divided by its conjugate
    };
    auto footprint = [&](const uint2 pk) {
        Foot f;
        const int ax = static_cast<int>(pk.x & 0xffffu) - x0;
        const int ay = static_cast<int>(pk.x >> 16) - y0;
        f.ktap = nullptr;
        f.wxr = f.wyr = nullptr;
        f.csign = 1.0f;
        f.awn = make_float2(1.0f, 0.0f);
        if (!WP) {
            f.side = kTaps;
            f.wxr = taps + (pk.y & 0xffu) * kTaps;
            f.wyr = taps + ((pk.y >> 8) & 0xffu) * kTaps;
            f.bx = ax;  // standard anchors are already x_start
            f.by = ay;
        } else {
            // the group record's fourth word: tap-table index (25 bits),
            // conjugation (bit 25), half-support (bits 26-31)
            const uint32_t po = pk.y & 0x1ffffffu;
            f.side = 2 * static_cast<int>(pk.y >> 26) + 1;
            const uint64_t local = po - gp.plane_lo * gp.n_off * gp.n_off;  // tables of this chunk
            f.ktap = gp.compact + local * gp.tap_stride;
            f.csign = ((pk.y >> 25) & 1u) ? -1.0f : 1.0f;
            if (kModel && gp.aw_norm) f.awn = gp.aw_norm[po];
            f.bx = ax - (f.side >> 1);
            f.by = ay - (f.side >> 1);
        }
        return f;
    };

    // Per-lane inputs of one group, loaded a group ahead: the lane's sample
    // (index s0 + lane) and, per register slot, the tap and LDS cell of the
    // slot's footprint cell (cell -1: the slot is outside the footprint).
    struct GroupData {
        float2 v;     // visibility (ADJOINT: the value)
        float wj;
        uint32_t fj;  // plan flags of the sample
        uint32_t idx; // original index (DEGRID)
        float2 tap[NS];
        int cell[NS];
    };
    auto load_group = [&](uint32_t s0, uint32_t s1, const Foot &f, GroupData &d, bool valid) {
        const uint32_t j = s0 + hl;
        d.v = make_float2(0.0f, 0.0f);
        d.wj = 0.0f;
        d.fj = 0u;
        d.idx = 0u;
        if (!kSums && j < s1) {
            if (MODE == kModeAdjoint) {
                d.v = values[index[j]];
            } else if (MODE == kModeDegrid) {
                d.idx = index[j];
            } else {
                d.fj = plan[j].y;
                d.wj = wt[j];
                d.v = vis[j];
            }
        }
        // (lane_py - by) mod M once, then x offsets (lane_pg + k P - bx) mod M
        // by steps of P; bx, by >= 0.
        const int qy = static_cast<int>((static_cast<float>(f.by - lane_py + Mside - 1) + 0.5f) * invM);
        const int jj = lane_py + qy * Mside - f.by;
        const int qx = static_cast<int>((static_cast<float>(f.bx - lane_pg + Mside - 1) + 0.5f) * invM);
        int i = lane_pg + qx * Mside - f.bx;
        const bool in_y = jj < f.side;
        const int cy = f.by + jj;
        // Per slot: the x offset advances by P and wraps past the footprint
        // side as min(i, i - M) in unsigned arithmetic (i < 2M), and the LDS
        // cell and the tap's byte offset from the group's table base follow
        // by one 24-bit multiply-add each from per-group bases (no selects,
        // no carried offsets).
        const uint32_t cbase = static_cast<uint32_t>(f.bx * Ls + cy);
        const uint32_t tbase = static_cast<uint32_t>(jj) * 8u;
        const uint32_t side8 = static_cast<uint32_t>(f.side) * 8u;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            d.cell[k] = -1;  // (the tap of a slot without a cell is never read)
            if (k > 0) {
                i += Pl;
                i = static_cast<int>(min(static_cast<uint32_t>(i), static_cast<uint32_t>(i - Mside)));
            }
            if (!((slot_mask >> k) & 1u) || !in_y || i >= f.side || !valid) continue;
            if (kChains) {
                const uint32_t ui = static_cast<uint32_t>(i);
                d.cell[k] = static_cast<int>(cbase + __umul24(ui, static_cast<uint32_t>(Ls)));
                d.tap[k] = WP ? *reinterpret_cast<const float2 *>(reinterpret_cast<const char *>(f.ktap) +
                                                                   (tbase + __umul24(ui, side8)))
                              : make_float2(f.wxr[i] * f.wyr[jj], 0.0f);
            } else {  // (few slots: plain index arithmetic)
                d.cell[k] = (f.bx + i) * Ls + cy;
                d.tap[k] = WP ? f.ktap[i * f.side + jj] : make_float2(f.wxr[i] * f.wyr[jj], 0.0f);
            }
        }
    };

    auto process = [&](uint32_t s0, uint32_t s1, const uint2 pk, const Foot &f, const GroupData &d,
                       const float4 gs) {
        const uint32_t count = s1 - s0;
        // Group sums:
        //   PSF/DIRTY: w = sum w_i (all samples), v = sum w_i V_i (finite V)
        //   RESIDUAL:  w = sum w_i, v = sum w_i V_i over usable samples
        //   (both precomputed per group when the stream was built)
        //   ADJOINT:   v = sum values_i, lane-parallel over the samples
        float vr = 0.0f, vi = 0.0f, w = 0.0f;
        if (kSums) {
            w = gs.x;
            vr = gs.y;
            vi = gs.z;
            if (GPW == 1 && MODE == kModeResidual && w == 0.0f) return;
        } else if (MODE != kModeDegrid) {
            auto add = [&](uint32_t fj, float wj, float2 vj) {
                if (MODE == kModeAdjoint) {
                    vr += vj.x;
                    vi += vj.y;
                } else if (MODE == kModeResidual) {
                    if (fj & kResidOkBit) {
                        w += wj;
                        vr += vj.x * wj;
                        vi += vj.y * wj;
                    }
                } else {
                    w += wj;
                    if (MODE == kModeDirty && (fj & kVisOkBit)) {
                        vr += vj.x * wj;
                        vi += vj.y * wj;
                    }
                }
            };
            if (static_cast<uint32_t>(hl) < count) add(d.fj, d.wj, d.v);
            for (uint32_t j = s0 + LW + hl; j < s1; j += LW) {
                if (MODE == kModeAdjoint) add(0u, 0.0f, values[index[j]]);
                else add(plan[j].y, wt[j], vis[j]);
            }
            if (GPW == 2) {
                half_sum_f2(vr, vi, upper);
                if (MODE != kModeAdjoint) w = half_sum_f(w, upper);
            } else if (count > 1) {
                wave_sum_f2(vr, vi);
                if (MODE != kModeAdjoint) w = wave_sum_f(w);
            } else {
                vr = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vr)));
                vi = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(vi)));
                w = __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(w)));
            }
            if (GPW == 1 && MODE == kModeResidual && w == 0.0f) return;
        }
        auto write_degrid = [&](float pr, float pi) {
            // every sample of the group has the same prediction
            if (static_cast<uint32_t>(hl) < count) vis_out[d.idx] = make_float2(pr, pi);
            for (uint32_t j = s0 + LW + hl; j < s1; j += LW) vis_out[index[j]] = make_float2(pr, pi);
        };
        const int side = f.side, T = side * side, bx = f.bx, by = f.by;
        const float2 *ktap = f.ktap;
        const float *wxr = f.wxr, *wyr = f.wyr;
        const float csign = f.csign;
        const float2 awn = f.awn;
        (void)pk;

        if (KSV == 0) {
            // Supports above the register slots: taps straight to LDS, in
            // chunks of 64 * kBig taps whose loads every lane issues together
            // and keeps in registers for both the degrid and the grid pass
            // when the footprint fits one chunk.
            const float inv = 1.0f / static_cast<float>(side);
            const int n_chunks = (T + 64 * kBig - 1) / (64 * kBig);
            float2 tv[kBig];
            auto load_chunk = [&](int ch) {
#pragma unroll
                for (int k = 0; k < kBig; ++k) {
                    const int f = ch * 64 * kBig + 64 * k + lane;
                    float2 c = make_float2(0.0f, 0.0f);
                    if (f < T) {
                        if (WP) {
                            c = ktap[f];
                        } else {
                            const int tx = static_cast<int>((static_cast<float>(f) + 0.5f) * inv);
                            c = make_float2(wxr[tx] * wyr[f - tx * side], 0.0f);
                        }
                    }
                    tv[k] = c;
                }
            };
            float pr = 0.0f, pi = 0.0f;
            if (kModel) {
                for (int ch = 0; ch < n_chunks; ++ch) {
                    load_chunk(ch);
#pragma unroll
                    for (int k = 0; k < kBig; ++k) {
                        const int f = ch * 64 * kBig + 64 * k + lane;
                        if (f < T) {
                            const int tx = static_cast<int>((static_cast<float>(f) + 0.5f) * inv);
                            const int cell = (bx + tx) * Ls + (by + f - tx * side);
                            const float cr = tv[k].x, ci = -(tv[k].y * csign);
                            const float2 mc = M2[cell];
                            const float mr = mc.x, mi = mc.y;
                            pr += cr * mr - ci * mi;
                            pi += cr * mi + ci * mr;
                        }
                    }
                }
                wave_sum_f2(pr, pi);
                if (WP && gp.aw_norm) {
                if (kChains) div_conj_rcp(pr, pi, awn.x, awn.y);
                else div_conj<float>(pr, pi, awn.x, awn.y);
            }
                if (MODE == kModeDegrid) {
                    write_degrid(pr, pi);
                    return;
                }
            }
            if (MODE == kModeResidual) { vr = vr - pr * w; vi = vi - pi * w; }
            for (int ch = 0; ch < n_chunks; ++ch) {
                if (!(kModel && n_chunks == 1)) load_chunk(ch);
#pragma unroll
                for (int k = 0; k < kBig; ++k) {
                    const int f = ch * 64 * kBig + 64 * k + lane;
                    if (f < T) {
                        const int tx = static_cast<int>((static_cast<float>(f) + 0.5f) * inv);
                        const int cell = (bx + tx) * Ls + (by + f - tx * side);
                        const float cr = tv[k].x, ci = tv[k].y * csign;
                        if (kPsf) {
                            atomicAdd(P_re + cell, w * cr);
                            if (WP) atomicAdd(P_im + cell, w * ci);
                        }
                        if (kAcc) {
                            atomicAdd(A_re + cell, vr * cr - vi * ci);
                            atomicAdd(A_im + cell, vr * ci + vi * cr);
                        }
                    }
                }
            }
            return;
        }

        // Register-slot path: move every slot whose cell is in this
        // footprint onto it, then accumulate.
        float cr[NS], ci[NS];
        float pr = 0.0f, pi = 0.0f;
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            cr[k] = ci[k] = 0.0f;
            const int cell = d.cell[k];
            if (cell < 0) continue;
            if (cell != cell_k[k]) {
                spill(k);
                cell_k[k] = cell;
                if (kModel && kStageModel) {
                    const float2 mc = M2[cell];
                    mre[k] = mc.x;
                    mim[k] = mc.y;
                }
            }
            cr[k] = d.tap[k].x;
            ci[k] = d.tap[k].y * csign;
            if (kModel && !kStageModel) {  // read at every group: registers for occupancy
                const float2 mc = M2[cell];
                mre[k] = mc.x;
                mim[k] = mc.y;
            }
            if (kModel) {
                // conj(cwt) * model (many slots: fused multiply-add chains,
                // fewer instructions; few slots: shorter dependency chains)
                if (kChains) {
                    pr = fmaf(cr[k], mre[k], pr);
                    pr = fmaf(ci[k], mim[k], pr);
                    pi = fmaf(cr[k], mim[k], pi);
                    pi = fmaf(-ci[k], mre[k], pi);
                } else {
                    pr += cr[k] * mre[k] + ci[k] * mim[k];
                    pi += cr[k] * mim[k] - ci[k] * mre[k];
                }
            }
        }
        if (kModel) {
            if (GPW == 2) half_sum_f2(pr, pi, upper);
            else wave_sum_f2(pr, pi);
            if (WP && gp.aw_norm) {
                if (kChains) div_conj_rcp(pr, pi, awn.x, awn.y);
                else div_conj<float>(pr, pi, awn.x, awn.y);
            }
            if (MODE == kModeDegrid) {
                write_degrid(pr, pi);
                return;
            }
            // sum_i (V_i - p) w_i = sum_i w_i V_i - p sum_i w_i
            vr = vr - pr * w;
            vi = vi - pi * w;
        }
#pragma unroll
        for (int k = 0; k < NS; ++k) {
            if (kPsf) {
                pacr[k] = fmaf(w, cr[k], pacr[k]);
                paci[k] = fmaf(w, ci[k], paci[k]);
            }
            if (kAcc && kChains) {  // value * tap
                accr[k] = fmaf(vr, cr[k], accr[k]);
                accr[k] = fmaf(-vi, ci[k], accr[k]);
                acci[k] = fmaf(vr, ci[k], acci[k]);
                acci[k] = fmaf(vi, cr[k], acci[k]);
            } else if (kAcc) {
                accr[k] += vr * cr[k] - vi * ci[k];
                acci[k] += vr * ci[k] + vi * cr[k];
            }
        }
    };

    // Group records (first sample, end, packed plan of the group) and, in the
    // gridding modes, the group's sums are read 64 at a time with coalesced
    // vector loads, lane l holding those of group base + l.  A group's
    // record moves to scalar registers with readlane, so its footprint is
    // scalar arithmetic and no scalar load waits behind the LDS traffic of the
    // group being processed.  Software pipeline: while group g is processed,
    // the taps and samples of group g + 1 are in flight.
    // GPW = 2: each half-wavefront holds a window of 32 records of its own
    // run; both halves advance one group per iteration, so their windows
    // move together, and a record reaches its half's lanes by one backward
    // permute (lane l reads lane (l & 32) + k).
    auto rec_load = [&](uint32_t wbase) {
        const uint32_t g = h_begin + wbase + static_cast<uint32_t>(hl);
        return g < h_end ? grec[g] : make_uint4(0u, 0u, 0u, 0u);
    };
    auto sum_load = [&](uint32_t wbase) {
        const uint32_t g = h_begin + wbase + static_cast<uint32_t>(hl);
        return (kSums && g < h_end) ? gsum[g] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    };
    const int perm_half = (lane & 32) << 2;  // byte address of the half's first lane
    auto bcast = [&](uint32_t x, uint32_t k) {
        if (GPW == 1) return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(k)));
        return static_cast<uint32_t>(
            __builtin_amdgcn_ds_bpermute(perm_half + static_cast<int>(k << 2), static_cast<int>(x)));
    };
    auto sum_at = [&](const float4 &sb, uint32_t k) {
        return make_float4(__uint_as_float(bcast(__float_as_uint(sb.x), k)),
                           __uint_as_float(bcast(__float_as_uint(sb.y), k)),
                           __uint_as_float(bcast(__float_as_uint(sb.z), k)), 0.0f);
    };
    auto rec_at = [&](const uint4 &rb, uint32_t k) {
        return make_uint4(bcast(rb.x, k), bcast(rb.y, k), bcast(rb.z, k), bcast(rb.w, k));
    };
    // one group in flight: its record, sums, footprint and per-lane inputs
    struct Stage {
        uint32_t s0, s1;
        uint2 pk;
        float4 g;
        Foot f;
        GroupData d;
    };
    uint32_t wbase = 0;
    uint4 rb = make_uint4(0u, 0u, 0u, 0u);
    float4 sb = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    auto fetch = [&](uint32_t t, Stage &x) {
        uint32_t k = t - wbase;
        if (k == static_cast<uint32_t>(LW)) {
            wbase += LW;
            rb = rec_load(wbase);
            sb = sum_load(wbase);
            k = 0;
        }
        const uint4 r = rec_at(rb, k);
        x.s0 = r.x;
        x.s1 = r.y;
        x.pk = make_uint2(r.z, r.w);
        x.g = kSums ? sum_at(sb, k) : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        x.f = footprint(x.pk);
        load_group(x.s0, x.s1, x.f, x.d, h_begin + t < h_end);
    };
    if (n_iter > 0) {
        rb = rec_load(0);
        sb = sum_load(0);
        Stage c;
        fetch(0, c);
        for (uint32_t t = 0; t < n_iter; ++t) {
            Stage n = c;
            if (t + 1 < n_iter) fetch(t + 1, n);
            process(c.s0, c.s1, c.pk, c.f, c.d, c.g);
            c = n;
        }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) spill(k);
    };
    run(std::integral_constant<int, KS>{});
    if (MODE == kModeDegrid) return;
    __syncthreads();

    // Flush the non-zero cells of the LDS tile to the global grid.
    flush_tile<kPsf, kAcc>(gp, chan_off, x0, y0, L, Ls, P_re, P_im, A_re, A_im, psf, acc_out);
}

// Residual refresh of work items whose largest half-support is 16 - 31
// (footprint sides M = 33 - 63).  One lane per footprint row leaves a single
// wavefront no room for register slots (up to 63 x residues per lane), and
// per-tap LDS atomics (the chunked path) bound the time, so the item's
// groups go through two phases:
//   1. predictions: the wavefronts split the groups; a group's prediction is
//      a chunked sum over its taps and the staged model (LDS reads only),
//      and its residual value sum w V - p sum w goes to an LDS array;
//   2. gridding: the wavefronts form sets of W = wset that split the x
//      residues (wavefront j of a set owns residues j S + s, s < S =
//      ceil(M / W) <= kWideSlots, lane l the row residue l), and each set
//      walks a contiguous run of the item's groups: every wavefront sees
//      each group of its run, accumulates its slots in registers and spills
//      a slot to the LDS tile (two atomics) only when the slot's cell
//      changes (Romein's work distribution with the residues spread over
//      wavefronts instead of lanes).
// The tile is then flushed as in k_binned.  Values of the taps and sums are
// those of the residual mode of k_binned.
constexpr int kWideSlots = 16;
constexpr int kWideSets = 2;

__global__ __launch_bounds__(kBinnedThreads) void k_binned_wide(GridParams gp, const WorkItem *__restrict__ items,
                                                               const uint4 *__restrict__ grec,
                                                               const float4 *__restrict__ gsum,
                                                               const float2 *__restrict__ model, float *acc_out,
                                                               uint32_t pred_cap, uint32_t wset) {
    extern __shared__ float lds[];
    const WorkItem it = items[blockIdx.x];
    if (it.count == 0 || it.count > pred_cap) return;  // padding of the XCD-aligned item order
    const int L = static_cast<int>(gp.L);
    const int Ls = static_cast<int>(gp.Ls);
    const int cells = L * Ls;
    float2 *M2 = reinterpret_cast<float2 *>(lds);
    float *A_re = lds + 2 * cells;
    float *A_im = A_re + cells;
    float2 *pred = reinterpret_cast<float2 *>(A_im + cells);
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int nwaves = static_cast<int>(blockDim.x) >> 6;
    const uint32_t tile_x = it.tile / gp.tiles_y;
    const uint32_t tile_y = it.tile - tile_x * gp.tiles_y;
    const int x0 = static_cast<int>(tile_x * gp.tile) - static_cast<int>(gp.halo);
    const int y0 = static_cast<int>(tile_y * gp.tile) - static_cast<int>(gp.halo);
    const uint64_t chan_off = static_cast<uint64_t>(it.chan & 0xffffffu) * gp.cstride;
    for (int c = tid; c < cells; c += static_cast<int>(blockDim.x)) {
        const int lx = c / Ls, ly = c - lx * Ls;
        M2[c] = ly < L ? tile_model(gp, model, chan_off, x0 + lx, y0 + ly) : make_float2(0.0f, 0.0f);
        A_re[c] = 0.0f;
        A_im[c] = 0.0f;
    }
    __syncthreads();

    struct Foot {
        int side, bx, by;
        const float2 *ktap;
        float csign;
        uint32_t po;
    };
    auto footprint = [&](uint32_t anchor, uint32_t word) {
        // the group record's fourth word: tap-table index (25 bits),
        // conjugation (bit 25), half-support (bits 26-31)
        Foot f;
        f.po = word & 0x1ffffffu;
        f.side = 2 * static_cast<int>(word >> 26) + 1;
        const uint64_t local = f.po - gp.plane_lo * gp.n_off * gp.n_off;
        f.ktap = gp.compact + local * gp.tap_stride;
        f.csign = ((word >> 25) & 1u) ? -1.0f : 1.0f;
        f.bx = static_cast<int>(anchor & 0xffffu) - x0 - (f.side >> 1);
        f.by = static_cast<int>(anchor >> 16) - y0 - (f.side >> 1);
        return f;
    };
    auto rl = [](uint32_t x, uint32_t k) {
        return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(x), static_cast<int>(k)));
    };

    // 1. predictions and residual values, groups split over the wavefronts
    {
        const uint32_t per = (it.count + nwaves - 1) / nwaves;
        const uint32_t b = it.start + static_cast<uint32_t>(wave) * per;
        const uint32_t e = min(it.start + it.count, b + per);
        for (uint32_t base = b; base < e; base += 64) {
            const uint32_t gl = base + static_cast<uint32_t>(lane);
            const uint4 rw = gl < e ? grec[gl] : make_uint4(0u, 0u, 0u, 0u);
            const float4 sw = gl < e ? gsum[gl] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            const uint32_t n = min(64u, e - base);
            for (uint32_t k = 0; k < n; ++k) {
                const float w = __uint_as_float(rl(__float_as_uint(sw.x), k));
                float2 v = make_float2(0.0f, 0.0f);
                if (w != 0.0f) {
                    const Foot f = footprint(rl(rw.z, k), rl(rw.w, k));
                    const int T = f.side * f.side;
                    const float inv = 1.0f / static_cast<float>(f.side);
                    float pr = 0.0f, pi = 0.0f;
                    for (int c0 = 0; c0 < T; c0 += 64 * kBig) {
                        float2 tv[kBig];
#pragma unroll
                        for (int q = 0; q < kBig; ++q) {
                            const int t = c0 + 64 * q + lane;
                            tv[q] = t < T ? f.ktap[t] : make_float2(0.0f, 0.0f);
                        }
#pragma unroll
                        for (int q = 0; q < kBig; ++q) {
                            const int t = c0 + 64 * q + lane;
                            if (t < T) {
                                const int tx = static_cast<int>((static_cast<float>(t) + 0.5f) * inv);
                                const float2 m = M2[(f.bx + tx) * Ls + (f.by + t - tx * f.side)];
                                const float cr = tv[q].x, ci = -(tv[q].y * f.csign);
                                pr += cr * m.x - ci * m.y;
                                pi += cr * m.y + ci * m.x;
                            }
                        }
                    }
                    wave_sum_f2(pr, pi);
                    if (gp.aw_norm) {
                        const float2 awn = gp.aw_norm[f.po];
                        div_conj<float>(pr, pi, awn.x, awn.y);
                    }
                    v = make_float2(__uint_as_float(rl(__float_as_uint(sw.y), k)) - pr * w,
                                    __uint_as_float(rl(__float_as_uint(sw.z), k)) - pi * w);
                }
                if (lane == 0) pred[base + k - it.start] = v;
            }
        }
    }
    __syncthreads();

    // 2. gridding: x residues split over the wavefronts of a set, groups over
    // the sets
    const int W = static_cast<int>(wset);
    const int sets = nwaves / W;
    const int set = wave / W;
    if (set < sets) {
        const int xw = wave - set * W;
        const int M = 2 * min(static_cast<int>(it.chan >> 24), static_cast<int>(gp.halo)) + 1;
        const float invM = 1.0f / static_cast<float>(M);
        // the item's M residues dealt evenly over the W wavefronts of a set
        // (at most kWideSlots each: W kWideSlots >= the launch's M >= M)
        const int Sw = (M + W - 1) / W;
        const int r0 = xw * Sw;  // first x residue of this wavefront
        const uint32_t per = (it.count + sets - 1) / sets;
        const uint32_t b = it.start + static_cast<uint32_t>(set) * per;
        const uint32_t e = min(it.start + it.count, b + per);
        int cell_k[kWideSlots];
        float accr[kWideSlots], acci[kWideSlots];
#pragma unroll
        for (int s = 0; s < kWideSlots; ++s) {
            cell_k[s] = -1;
            accr[s] = acci[s] = 0.0f;
        }
        auto spill = [&](int s) {
            if (cell_k[s] >= 0) {
                if (accr[s] != 0.0f) atomicAdd(A_re + cell_k[s], accr[s]);
                if (acci[s] != 0.0f) atomicAdd(A_im + cell_k[s], acci[s]);
            }
            accr[s] = acci[s] = 0.0f;
        };
        for (uint32_t base = b; base < e; base += 64) {
            const uint32_t gl = base + static_cast<uint32_t>(lane);
            const uint4 rw = gl < e ? grec[gl] : make_uint4(0u, 0u, 0u, 0u);
            const float2 vw = gl < e ? pred[gl - it.start] : make_float2(0.0f, 0.0f);
            const uint32_t n = min(64u, e - base);
            for (uint32_t k = 0; k < n; ++k) {
                const float vr = __uint_as_float(rl(__float_as_uint(vw.x), k));
                const float vi = __uint_as_float(rl(__float_as_uint(vw.y), k));
                if (vr == 0.0f && vi == 0.0f) continue;
                const Foot f = footprint(rl(rw.z, k), rl(rw.w, k));
                // this lane's row: y offset (lane - by) mod M; bx, by >= 0
                const int qy = static_cast<int>((static_cast<float>(f.by - lane + M - 1) + 0.5f) * invM);
                const int jj = lane + qy * M - f.by;
                const bool row_ok = lane < M && jj < f.side;
                const int cy = f.by + jj;
                // x offset of the first residue: (r0 - bx) mod M (wavefront-uniform)
                const int qx = static_cast<int>((static_cast<float>(f.bx - r0 + M - 1) + 0.5f) * invM);
                int i = __builtin_amdgcn_readfirstlane(r0 + qx * M - f.bx);
                const float sr = vr, si = vi * f.csign, tr = vr * f.csign;  // value times conj(tap) sign
                float2 tv[kWideSlots];
                int cell[kWideSlots];
#pragma unroll
                for (int s = 0; s < kWideSlots; ++s) {
                    const bool on = s < Sw && r0 + s < M && i < f.side;
                    cell[s] = (on && row_ok) ? (f.bx + i) * Ls + cy : -1;
                    tv[s] = (on && row_ok) ? f.ktap[i * f.side + jj] : make_float2(0.0f, 0.0f);
                    i = i + 1 == M ? 0 : i + 1;
                }
#pragma unroll
                for (int s = 0; s < kWideSlots; ++s) {
                    if (cell[s] < 0) continue;
                    if (cell[s] != cell_k[s]) {
                        spill(s);
                        cell_k[s] = cell[s];
                    }
                    // value * tap, the tap conjugated when csign = -1
                    accr[s] = fmaf(sr, tv[s].x, accr[s]);
                    accr[s] = fmaf(-si, tv[s].y, accr[s]);
                    acci[s] = fmaf(tr, tv[s].y, acci[s]);
                    acci[s] = fmaf(vi, tv[s].x, acci[s]);
                }
            }
        }
#pragma unroll
        for (int s = 0; s < kWideSlots; ++s) spill(s);
    }
    __syncthreads();
    flush_tile<false, true>(gp, chan_off, x0, y0, L, Ls, nullptr, nullptr, A_re, A_im, nullptr, acc_out);
}

// The stream is binned for the W-table residency the gridder had when the set
// was created.
static void check_wk_chunk(const chip_vset &vs) {
    const chip_gridder &g = *vs.g;
    const uint32_t now = (g.wproject && !g.aw) ? g.wk_chunk : 0u;
    if (now != vs.wk_chunk)
        fail(CHIP_ERR_INVALID, "the W kernel residency changed after the visibility set was created");
}

// One launch over the work items [first, first + count).
template <int MODE, bool WP, int KS, int GPW = 1>
static void launch_binned_k(chip_vset &vs, const GridParams &gp, const float2 *values, const float2 *model,
                            float *psf, float *acc, float2 *vis_out, size_t lds, hipStream_t st, uint64_t first,
                            uint64_t count) {
    if (lds > 64 * 1024)
        CHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_binned<MODE, WP, KS, GPW>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    const float4 *gsum = MODE == kModeResidual                         ? vs.gsum_r.as<float4>()
                         : (MODE == kModePsf || MODE == kModeDirty) ? vs.gsum_d.as<float4>()
                                                                     : nullptr;
    hipLaunchKernelGGL((k_binned<MODE, WP, KS, GPW>), dim3(static_cast<unsigned>(count)), dim3(vs.threads), lds,
                       st, gp, vs.items.as<WorkItem>() + first, vs.grec.as<uint4>(), gsum, vs.plan.as<uint2>(),
                       vs.vis.as<float2>(), vs.wt.as<float>(), vs.index.as<uint32_t>(), values, model, psf, acc,
                       vis_out);
    CHIP_HIP(hipGetLastError());
}

// Register slots per lane for a tile halo: the x residues per lane,
// ceil(M / floor(64 / M)) with M = 2 halo + 1, rounded up to an instantiated
// count; 0 (LDS-atomic chunks) beyond 16 (halo > 15).
static int register_slots(uint32_t halo) { return slots_for_halo(static_cast<int>(halo)); }

// Register slots per lane with two groups per wavefront (32 lanes per
// group): 1 - 3 for halos up to 4, 0 (not offered) above.
static int register_slots_half(uint32_t halo) {
    const uint32_t m = 2 * halo + 1, p = 32 / m;
    if (p == 0) return 0;
    const uint32_t need = (m + p - 1) / p;
    return need <= 3 ? static_cast<int>(need) : 0;
}

// Largest half-support of a support class (classes as support_class in
// plan.hip: up to 3, 4, 7, 10, 15, 22, above), at most the set's halo.
static uint32_t class_halo(const chip_vset &vs, uint32_t cls) {
    static constexpr uint32_t kCap[] = {3u, 4u, 7u, 10u, 15u, 22u};
    return (vs.n_classes > 1 && cls < 6) ? std::min(kCap[cls], vs.halo) : vs.halo;
}

static bool class_used(const chip_vset &vs, uint32_t cls) {
    const uint64_t n_chunks = vs.range_first.empty() ? 0 : (vs.range_first.size() - 1) / vs.n_classes;
    for (uint64_t c = 0; c < n_chunks; ++c)
        if (vs.range_first[c * vs.n_classes + cls + 1] > vs.range_first[c * vs.n_classes + cls]) return true;
    return false;
}

// Groups per wavefront for a launch with this halo: 1, or 2 where the halo
// allows it and it was asked for (wave_groups 2) or the autotuner chose it
// (wave_groups 0).
static int wave_groups_for(const chip_vset &vs, uint32_t halo) {
    if (register_slots_half(halo) == 0 || vs.wave_groups == 1) return 1;
    return (vs.wave_groups == 2 || vs.tuned_wave_groups == 2) ? 2 : 1;
}

// Two groups per wavefront in use for some class with work items.
int binned_wave_groups(const chip_vset &vs) {
    for (uint32_t k = 0; k < vs.n_classes; ++k)
        if (class_used(vs, k) && wave_groups_for(vs, class_halo(vs, k)) == 2) return 2;
    return 1;
}

bool binned_two_groups_allowed(const chip_vset &vs) {
    for (uint32_t k = 0; k < vs.n_classes; ++k)
        if (class_used(vs, k) && register_slots_half(class_halo(vs, k)) > 0) return true;
    return false;
}

// Residual range of half-supports 16 - 31: the two-phase kernel, when its
// LDS (tile, accumulators and one value per group) fits.
static bool launch_wide(chip_vset &vs, const GridParams &gp, const float2 *model, float *acc, hipStream_t st,
                        uint64_t first, uint64_t count) {
    if (gp.halo < 16 || gp.halo > 31) return false;
    const uint32_t cap = vs.groups_per_item;
    const size_t lds = 4 * static_cast<size_t>(gp.L) * gp.Ls * sizeof(float) + static_cast<size_t>(cap) * sizeof(float2);
    const uint32_t wset = (2 * gp.halo + 1 + kWideSlots - 1) / kWideSlots;
    const uint32_t threads = 64 * wset * kWideSets;
    if (lds > lds_limit(vs.device) || threads > static_cast<uint32_t>(kBinnedThreads)) return false;
    if (lds > 64 * 1024)
        CHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_binned_wide),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)));
    hipLaunchKernelGGL(k_binned_wide, dim3(static_cast<unsigned>(count)), dim3(threads), lds, st, gp,
                       vs.items.as<WorkItem>() + first, vs.grec.as<uint4>(), vs.gsum_r.as<float4>(), model, acc,
                       cap, wset);
    CHIP_HIP(hipGetLastError());
    return true;
}

// One range of work items with the kernel its halo needs.
template <int MODE, bool WP>
static void launch_range(chip_vset &vs, const GridParams &gp, const float2 *values, const float2 *model,
                         float *psf, float *acc, float2 *vis_out, hipStream_t st, uint64_t first, uint64_t count) {
    const size_t lds = binned_lds_bytes<MODE, WP>(gp);
    if (lds > lds_limit(vs.device))
        fail(CHIP_ERR_INVALID, "kernel support too large for the LDS tile (halo " + std::to_string(gp.halo) + ")");
    const int ks = register_slots(gp.halo);
    if constexpr (MODE == kModeResidual && WP) {
        if (ks == 0 && vs.wide && launch_wide(vs, gp, model, acc, st, first, count)) return;
    }
    const int ks2 = wave_groups_for(vs, gp.halo) == 2 ? register_slots_half(gp.halo) : 0;
#define CHIP_RANGE(KS_, GPW_) \
    launch_binned_k<MODE, WP, KS_, GPW_>(vs, gp, values, model, psf, acc, vis_out, lds, st, first, count)
    if (ks2 > 0) {  // two groups per wavefront
        if constexpr (!WP) {
            CHIP_RANGE(2, 2);
        } else {
            switch (ks2) {
                case 1: CHIP_RANGE(1, 2); break;
                case 2: CHIP_RANGE(2, 2); break;
                default: CHIP_RANGE(3, 2); break;
            }
        }
        return;
    }
    if constexpr (!WP) {
        if (ks != 1) fail(CHIP_ERR_INVALID, "standard gridder halo must be the 7x7 support");
        CHIP_RANGE(1, 1);
    } else {
        switch (ks) {
            case 1: CHIP_RANGE(1, 1); break;
            case 2: CHIP_RANGE(2, 1); break;
            case 4: CHIP_RANGE(4, 1); break;
            case 7: CHIP_RANGE(7, 1); break;
            case 16: CHIP_RANGE(16, 1); break;
            default: CHIP_RANGE(0, 1); break;
        }
    }
#undef CHIP_RANGE
}

// Every (plane chunk, support class) range of work items: each class with the
// LDS tile and register slots of its largest half-support.  Events on the
// launch stream bracket the launches (chip_vset_stats); with W tables
// recomputed on the fly they bracket the recomputation too.
template <int MODE, bool WP>
static void launch_binned_t(chip_vset &vs, const GridParams &gp, const float2 *values,
                            const float2 *model, float *psf, float *acc, float2 *vis_out,
                            hipStream_t st, bool timed = true) {
    const size_t lds = binned_lds_bytes<MODE, WP>(gp);
    if (lds > lds_limit(vs.device))
        fail(CHIP_ERR_INVALID, "kernel support too large for the LDS tile (halo " +
                                   std::to_string(gp.halo) + ")");
    if (vs.n_items == 0) return;
    if (vs.threads == 0 || vs.threads > kBinnedThreads || vs.threads % 64)
        fail(CHIP_ERR_INVALID, "binned kernel threads must be a multiple of 64 up to 512");
    auto launch_chunk = [&](const GridParams &p, uint64_t c) {
        for (uint32_t k = 0; k < vs.n_classes; ++k) {
            const uint64_t first = vs.range_first[c * vs.n_classes + k];
            const uint64_t end = vs.range_first[c * vs.n_classes + k + 1];
            if (end == first) continue;
            GridParams pc = p;
            pc.halo = class_halo(vs, k);
            pc.L = pc.tile + 2 * pc.halo;
            pc.Ls = pc.L | 1u;
            launch_range<MODE, WP>(vs, pc, values, model, psf, acc, vis_out, st, first, end - first);
        }
    };
    if (!timed) {  // captured into a graph: kernels only
        if (vs.wk_chunk) fail(CHIP_ERR_INVALID, "W tables recomputed on the fly cannot be captured");
        launch_chunk(gp, 0);
        vs.timed = false;
        return;
    }
    CHIP_HIP(hipEventRecord(vs.ev_start, st));
    if (!vs.wk_chunk) {
        launch_chunk(gp, 0);
    } else {
        chip_gridder &g = *vs.g;
        WkScratch ws;
        wk_prepare(g, ws);
        GridParams p = gp;
        for (size_t c = 0; c + 1 < vs.chunk_items.size(); ++c) {
            if (vs.chunk_items[c + 1] == vs.chunk_items[c]) continue;  // no samples in these planes
            const uint32_t lo = static_cast<uint32_t>(c) * vs.wk_chunk;
            wk_fill(g, ws, lo, std::min(g.planes, lo + vs.wk_chunk), p);
            launch_chunk(p, c);
        }
        CHIP_HIP(hipEventRecord(vs.ev_stop, st));
        CHIP_HIP(hipStreamSynchronize(st));  // the scratch is freed on return
        vs.timed = true;
        return;
    }
    CHIP_HIP(hipEventRecord(vs.ev_stop, st));
    vs.timed = true;
}

template <int MODE>
static void launch_binned(chip_vset &vs, const float2 *values, const float2 *model, float *psf,
                          float *acc, float2 *vis_out) {
    const chip_gridder &g = *vs.g;
    check_wk_chunk(vs);
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

// Residual pass of the major cycle over caller-laid-out grids (see GridParams
// cstride / paired / model_signed).
// (on stream st; timed = false while the major cycle is captured into a graph)
void vset_residual_with(chip_vset &vs, const GridParams &gp, const float2 *model, float *res, hipStream_t st,
                        bool timed) {
    check_wk_chunk(vs);
    if (vs.g->wproject)
        launch_binned_t<kModeResidual, true>(vs, gp, nullptr, model, nullptr, res, nullptr, st, timed);
    else
        launch_binned_t<kModeResidual, false>(vs, gp, nullptr, model, nullptr, res, nullptr, st, timed);
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

}  // namespace chip
