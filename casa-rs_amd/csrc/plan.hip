// FP64 planning on the GPU and the binned visibility stream.
//
// Planning restates StandardGridder::plan_positive_taps (gridder.rs:557-566)
// and WProjector::plan_sample (gridder.rs:3588-3624) bit for bit: the only
// operations are IEEE f64 division, addition, multiplication, sqrt and
// round-half-away-from-zero, each separately rounded.
//
// The resident stream replaces the reference's per-chunk sample packing
// (lib.rs:71708-71767): accepted samples are keyed by (channel, uv tile,
// anchor cell, plan) and sorted with an integer radix sort, so samples with
// identical plans form contiguous groups (the reference's Metal path merges
// adjacent identical plans, lib.rs:71646-71660; here every identical plan of
// a bin is grouped, and the group's samples are summed exactly as the
// per-sample sum would be: sum (V_i - p) w_i = sum w_i V_i - p sum w_i).
// Groups of one bin are split into work items of at most groups_per_item
// (1024 unless the schedule autotuner picks another) so that one launch
// covers every channel with balanced workgroups.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "device.hpp"

namespace chip {


PlanParams plan_params(const chip_gridder &g) {
    PlanParams p{};
    p.gx = g.gx; p.gy = g.gy; p.du = g.du; p.dv = g.dv;
    p.half_gx = static_cast<double>(g.gx) / 2.0;
    p.half_gy = static_cast<double>(g.gy) / 2.0;
    p.os = g.os;
    p.wproject = g.wproject ? 1u : 0u;
    p.sampling = g.sampling;
    p.planes = g.planes;
    p.w_scale = g.w_scale;
    p.norm_radius = g.norm_radius;
    p.supports = g.wproject ? g.d_supports.as<int32_t>() : nullptr;
    p.norm_table = g.wproject && !g.aw ? g.d_norm_table.as<float>() : nullptr;
    p.aw = g.aw ? 1u : 0u;
    p.aw_cells = g.aw_cells;
    p.aw_pointings = g.aw_pointings;
    p.aw_w_incr = g.aw_w_incr;
    p.aw_cell = g.aw ? g.d_aw_cell.as<AwCell>() : nullptr;
    p.aw_mag = g.aw ? g.d_aw_mag.as<double>() : nullptr;
    p.cf_cell = nullptr;
    p.pointing = nullptr;
    return p;
}

PlanParams plan_params(const chip_gridder &g, const Source &src) {
    PlanParams p = plan_params(g);
    p.cf_cell = src.cf_cell;
    p.pointing = src.pointing;
    return p;
}

void validate_source(const chip_samples &s) {
    if (s.n_chan == 0) fail(CHIP_ERR_INVALID, "n_chan must be at least 1");
    if (!s.vis_dev || !s.weight_dev) fail(CHIP_ERR_INVALID, "vis and weight are required");
    const bool per_sample = s.u_dev && s.v_dev && s.w_dev;
    const bool per_row = s.uvw_m_dev && s.lambda_scale_dev;
    if (per_sample == per_row)
        fail(CHIP_ERR_INVALID, "give either u/v/w per sample or uvw_m with lambda_scale");
    if (per_sample && s.n_chan != 1)
        fail(CHIP_ERR_INVALID, "per-sample u/v/w input requires n_chan == 1");
    if (s.n_rows * s.n_chan >= (1ull << 32) - 1)
        fail(CHIP_ERR_INVALID, "at most 2^32-2 samples per call");
}

Source make_source(const chip_samples &s) {
    Source r{};
    r.n_rows = s.n_rows; r.n_chan = s.n_chan;
    r.u = s.u_dev; r.v = s.v_dev; r.w = s.w_dev;
    r.uvw_m = s.uvw_m_dev; r.lambda_scale = s.lambda_scale_dev;
    r.vis = reinterpret_cast<const float2 *>(s.vis_dev);
    r.weight = s.weight_dev; r.sumwt_factor = s.sumwt_factor_dev;
    r.gridable = s.gridable_dev; r.flag = s.flag_dev;
    r.cf_cell = s.cf_cell_dev; r.pointing = s.pointing_dev;
    return r;
}

static unsigned blocks_for(uint64_t n, unsigned threads) {
    return static_cast<unsigned>((n + threads - 1) / threads);
}

// ---- raw plans (parity surface) -------------------------------------------
__global__ __launch_bounds__(256) void k_plan_raw(PlanParams p, Source s, int32_t *plan,
                                                  float *norm) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t n = s.n_rows * s.n_chan;
    if (i >= n) return;
    const SampleIn smp = load_sample(s, i);
    Plan pl;
    const bool ok = sample_accepted(p, smp, pl);
    int32_t *o = plan + 6 * i;
    if (ok) {
        o[0] = pl.x; o[1] = pl.y; o[2] = pl.sx; o[3] = pl.sy; o[4] = pl.plane; o[5] = pl.conj;
    } else {
        for (int k = 0; k < 6; ++k) o[k] = -1;
    }
    if (norm) norm[i] = ok ? pl.norm : __builtin_nanf("");
}

void launch_plan_raw(const chip_gridder &g, const Source &src, int32_t *plan, float *norm,
                     hipStream_t st) {
    const uint64_t n = src.n_rows * src.n_chan;
    if (n == 0) return;
    hipLaunchKernelGGL(k_plan_raw, dim3(blocks_for(n, 256)), dim3(256), 0, st, plan_params(g, src),
                       src, plan, norm);
    CHIP_HIP(hipGetLastError());
}

// ---- max |w| over usable samples (lib.rs:73642-73664) -----------------------
__global__ __launch_bounds__(256) void k_max_abs_w(Source s, unsigned long long *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t n = s.n_rows * s.n_chan;
    double m = 0.0;
    if (i < n) {
        const SampleIn smp = load_sample(s, i);
        const bool valid = smp.usable && finite_f(smp.weight) && smp.weight > 0.0f &&
                           finite_f(smp.sf) && smp.sf > 0.0f && finite_f(smp.vis.x) &&
                           finite_f(smp.vis.y) && isfinite(smp.u) && isfinite(smp.v) &&
                           isfinite(smp.w);
        if (valid) m = fabs(smp.w);
    }
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0 && m > 0.0)
        atomicMax(out, static_cast<unsigned long long>(__double_as_longlong(m)));
}

double launch_max_abs_w(const chip_gridder &g, const Source &src, hipStream_t st) {
    (void)g;
    const uint64_t n = src.n_rows * src.n_chan;
    if (n == 0) return 0.0;
    DevBuf d(sizeof(unsigned long long));
    CHIP_HIP(hipMemsetAsync(d.p, 0, d.bytes, st));
    hipLaunchKernelGGL(k_max_abs_w, dim3(blocks_for(n, 256)), dim3(256), 0, st, src,
                       d.as<unsigned long long>());
    CHIP_HIP(hipGetLastError());
    unsigned long long bits = 0;
    CHIP_HIP(hipMemcpyAsync(&bits, d.p, sizeof(bits), hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    double r;
    memcpy(&r, &bits, sizeof(r));
    return r;
}

// ---- binned stream ---------------------------------------------------------
// Sort key (64 bits, most significant first):
//   bin   = channel * n_tiles + uv tile of the anchor cell (with W tables
//           recomputed per chunk of planes, the plane chunk comes first)
//   cell  = anchor cell inside the tile (10 bits for 32x32 tiles)
//   plan  = the rest of the plan: tap-row selectors (standard) or
//           plane, conjugation and offsets (W)
// so that samples with identical plans become contiguous groups inside a
// bin.  Rejected samples get the invalid bin and sort last.  When the plan
// does not fit, plan_bits = 0 and every sample is its own group.
struct BinParams {
    uint32_t tile;
    uint32_t tiles_x, tiles_y;
    uint32_t n_tiles;
    uint32_t n_bins;        // chunks * n_chan * n_tiles (also the invalid bin)
    uint32_t n_chan;
    uint32_t chunk_planes;  // W tables recomputed per chunk of planes: bins
                            // are (plane / chunk_planes, channel, tile); 0: one chunk
    uint32_t plan_major;    // inside a bin, order groups by plan then anchor cell
                            // (consecutive groups share their tap table); 0:
                            // anchor cell then plan (consecutive footprints overlap)
    uint32_t cell_bits;     // log2(tile * tile)
    uint32_t plan_bits;
    uint32_t n_sel;         // selector values per axis
    uint32_t class_bits;    // W / AW: support class of the group's tap table
    const int32_t *supports;  // half-support per tap table (W / AW)
};

// Support class of a tap table: the binned kernel's register slots per lane
// (1, 2, 4, 7 for half-supports up to 3, 4, 7, 10), then LDS-atomic chunks
// with LDS tiles for half-supports up to 15, 22 and above.  Groups of one
// bin are ordered by class first, so that a work item holds one class and
// its launch runs with the fewest slots and the smallest LDS tile its
// footprints need.
__host__ __device__ __forceinline__ uint32_t support_class(int32_t support) {
    return support <= 3    ? 0u
           : support == 4  ? 1u
           : support <= 7  ? 2u
           : support <= 10 ? 3u
           : support <= 15 ? 4u
           : support <= 22 ? 5u
                           : 6u;
}
constexpr uint32_t kClassBits = 3;
constexpr uint32_t kSupportClasses = 7;

__device__ __forceinline__ uint64_t plan_code(const PlanParams &p, const BinParams &b, const Plan &pl) {
    if (p.wproject) {
        const uint64_t pc = static_cast<uint64_t>(pl.plane) * 2 + static_cast<uint64_t>(pl.conj);
        return (pc * b.n_sel + static_cast<uint64_t>(pl.sx + p.norm_radius)) * b.n_sel +
               static_cast<uint64_t>(pl.sy + p.norm_radius);
    }
    return static_cast<uint64_t>(pl.sx) * b.n_sel + static_cast<uint64_t>(pl.sy);
}

__global__ __launch_bounds__(256) void k_keys(PlanParams p, Source s, BinParams b,
                                              uint64_t *keys, uint32_t *vals,
                                              unsigned long long *n_valid) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t n = s.n_rows * s.n_chan;
    bool ok = false;
    if (i < n) {
        const SampleIn smp = load_sample(s, i);
        Plan pl;
        uint64_t key = static_cast<uint64_t>(b.n_bins) << (b.class_bits + b.cell_bits + b.plan_bits);
        if (sample_accepted(p, smp, pl)) {
            ok = true;
            const uint32_t ax = static_cast<uint32_t>(p.wproject ? pl.x : pl.x + kSupport);
            const uint32_t ay = static_cast<uint32_t>(p.wproject ? pl.y : pl.y + kSupport);
            const uint32_t tile = (ax / b.tile) * b.tiles_y + (ay / b.tile);
            uint64_t bin = static_cast<uint64_t>(smp.chan) * b.n_tiles + tile;
            if (b.chunk_planes)
                bin += static_cast<uint64_t>(static_cast<uint32_t>(pl.plane) / b.chunk_planes) * b.n_chan * b.n_tiles;
            const uint64_t cell = static_cast<uint64_t>((ax % b.tile) * b.tile + (ay % b.tile));
            key = (bin << (b.class_bits + b.cell_bits + b.plan_bits)) | (cell << b.plan_bits);
            if (b.class_bits)
                key |= static_cast<uint64_t>(support_class(b.supports[pl.plane])) << (b.cell_bits + b.plan_bits);
            if (b.plan_bits) key |= plan_code(p, b, pl);
        }
        keys[i] = key;
        vals[i] = static_cast<uint32_t>(i);
    }
    const unsigned long long c = wave_sum_u64(ok ? 1ull : 0ull);
    if ((threadIdx.x & 63) == 0 && c) atomicAdd(n_valid, c);
}

// Group heads: a sample starts a group when its key differs from the previous
// one (every sample when plans are not merged).
__global__ void k_group_flags(const uint64_t *keys, uint64_t n, bool merge, uint32_t *flags) {
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    flags[p] = (!merge || p == 0 || keys[p] != keys[p - 1]) ? 1u : 0u;
}

__global__ void k_group_starts(const uint64_t *keys, const uint32_t *flags, const uint32_t *gid,
                               uint64_t n, uint32_t low_bits, uint32_t *gstart, uint32_t *bin_groups) {
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n || !flags[p]) return;
    gstart[gid[p]] = static_cast<uint32_t>(p);
    atomicAdd(&bin_groups[keys[p] >> low_bits], 1u);
}

// One record per group: (first sample, end, packed plan), so that the binned
// kernel reads a group's bounds and plan with one load.
// W / AWProject: the fourth word is not the packed plan but what the binned
// kernel's footprint needs, so that it loads nothing else per group:
// tap-table index po = table * n_off^2 + (off_x + R) n_off + (off_y + R)
// (25 bits), conjugation (bit 25), half-support (bits 26-31).
constexpr uint32_t kRecPoBits = 25;

__global__ void k_group_records(const uint32_t *gstart, const uint2 *plan, uint32_t n_groups, const int32_t *supports,
                                uint32_t n_off, int32_t R, uint4 *rec) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const uint32_t s0 = gstart[g], s1 = gstart[g + 1];
    const uint2 pk = plan[s0];
    uint32_t w = pk.y;
    if (supports) {
        const int ox = static_cast<int>(pk.y & 0xffu) - 128, oy = static_cast<int>((pk.y >> 8) & 0xffu) - 128;
        const uint32_t table = packed_table(pk.y);
        const uint32_t po = (table * n_off + static_cast<uint32_t>(ox + R)) * n_off + static_cast<uint32_t>(oy + R);
        w = po | ((pk.y & kConjBit) ? (1u << kRecPoBits) : 0u) | (static_cast<uint32_t>(supports[table]) << 26);
    }
    rec[g] = make_uint4(s0, s1, pk.x, w);
}

// Per-group sums of what the gridding modes add, computed once in FP64 over
// the group's samples and rounded to FP32:
//   dsum = (sum w_i, sum w_i V_i over finite V)      PSF and DIRTY
//   rsum = (sum w_i, sum w_i V_i) over residual-usable samples   RESIDUAL
// so that a gridding pass reads one record per group instead of its samples.
__global__ void k_group_sums(const uint4 *grec, const uint2 *plan, const float2 *vis, const float *wt,
                             uint32_t n_groups, float4 *dsum, float4 *rsum) {
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const uint4 r = grec[g];
    double wa = 0.0, dr = 0.0, di = 0.0, wr = 0.0, rr = 0.0, ri = 0.0;
    for (uint32_t j = r.x; j < r.y; ++j) {
        const uint32_t f = plan[j].y;
        const double w = wt[j];
        const float2 v = vis[j];
        wa += w;
        if (f & kVisOkBit) {
            dr += static_cast<double>(v.x) * w;
            di += static_cast<double>(v.y) * w;
        }
        if (f & kResidOkBit) {
            wr += w;
            rr += static_cast<double>(v.x) * w;
            ri += static_cast<double>(v.y) * w;
        }
    }
    dsum[g] = make_float4(static_cast<float>(wa), static_cast<float>(dr), static_cast<float>(di), 0.0f);
    rsum[g] = make_float4(static_cast<float>(wr), static_cast<float>(rr), static_cast<float>(ri), 0.0f);
}

// Writes the packed stream in sorted order and accumulates per-channel sums
// (normalisation sumwt, reported sumwt, accepted count, tap count).  Sorted
// samples of one block span few channels, so sums are first reduced in LDS
// slots relative to the block's first channel.
constexpr int kSumSlots = 8;
constexpr int kSumFields = 4;

__global__ __launch_bounds__(256) void k_build_stream(PlanParams p, Source s, uint64_t n_valid,
                                                      const uint32_t *sorted_idx, uint2 *plan,
                                                      float2 *vis, float *wt, uint32_t *index,
                                                      double *sums /* n_chan x 4 */) {
    __shared__ double slot[kSumSlots][kSumFields];
    __shared__ uint32_t first_chan;
    const uint64_t pidx = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (threadIdx.x < kSumSlots * kSumFields) slot[threadIdx.x / kSumFields][threadIdx.x % kSumFields] = 0.0;
    if (threadIdx.x == 0) {
        const uint64_t p0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x;
        const uint64_t i0 = sorted_idx[p0];
        first_chan = s.n_chan == 1 ? 0u : static_cast<uint32_t>(i0 % s.n_chan);
    }
    __syncthreads();
    if (pidx < n_valid) {
        const uint32_t i = sorted_idx[pidx];
        const SampleIn smp = load_sample(s, i);
        Plan pl;
        const bool ok = sample_accepted(p, smp, pl);  // recomputed: identical to k_keys
        const float gw = grid_weight(p, smp);
        plan[pidx] = pack_plan(p, pl, ok ? sample_flags(p, smp) : 0u);
        vis[pidx] = smp.vis;
        wt[pidx] = gw;
        index[pidx] = i;
        double f[kSumFields];
        if (p.aw) {
            // AWProject: weight * |grid_normalization| (lib.rs:26535, one hand)
            f[0] = static_cast<double>(smp.weight) * pl.mag;
            f[1] = f[0];
            const AwCell c = p.aw_cell[static_cast<uint32_t>(pl.plane) / (2u * p.aw_pointings)];
            f[3] = static_cast<double>(2 * c.xs + 1) * static_cast<double>(2 * c.ys + 1);
        } else if (p.wproject) {
            // w_project_accumulation_for_sample (lib.rs:71460-71470)
            f[0] = 2.0 * static_cast<double>(smp.weight) * static_cast<double>(pl.norm);
            f[1] = static_cast<double>(smp.weight) * static_cast<double>(smp.sf);
            const double side = 2.0 * p.supports[pl.plane] + 1.0;
            f[3] = side * side;
        } else {
            // execution.rs:18306-18309
            f[0] = static_cast<double>(gw);
            f[1] = f[0];
            f[3] = static_cast<double>(kTaps * kTaps);
        }
        f[2] = 1.0;
        const uint32_t rel = smp.chan - first_chan;
        for (int k = 0; k < kSumFields; ++k) {
            if (rel < kSumSlots) atomicAdd(&slot[rel][k], f[k]);
            else atomicAdd(&sums[kSumFields * smp.chan + k], f[k]);
        }
    }
    __syncthreads();
    if (threadIdx.x < kSumSlots * kSumFields) {
        const int r = threadIdx.x / kSumFields, k = threadIdx.x % kSumFields;
        if (slot[r][2] > 0.0) atomicAdd(&sums[kSumFields * (first_chan + r) + k], slot[r][k]);
    }
}

// Work items per bin: runs of at most gpi groups of one bin.
__global__ void k_items_per_bin(const uint32_t *bin_groups, uint32_t n_bins, uint32_t gpi, uint32_t *items_per_bin) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bins) return;
    items_per_bin[k] = (bin_groups[k] + gpi - 1) / gpi;
}

// Work items of every bin; with W tables the item's channel word carries, in
// its top 8 bits, the largest half-support of the item's groups (the binned
// kernel sizes its register slots by it).
__global__ void k_write_items(const uint32_t *bin_groups, const uint32_t *bin_first_group,
                              const uint32_t *item_start, uint32_t n_bins, uint32_t n_tiles,
                              uint32_t n_chan, uint32_t gpi, const uint4 *grec, WorkItem *items) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bins) return;
    const uint32_t count = bin_groups[k];
    if (count == 0) return;
    const uint32_t ct = k / n_tiles;  // chunk * n_chan + channel
    const uint32_t chan = ct % n_chan;
    const uint32_t tile = k - ct * n_tiles;
    const uint32_t nit = (count + gpi - 1) / gpi;
    const uint32_t base = count / nit, extra = count % nit;  // spread evenly
    uint32_t pos = bin_first_group[k];
    for (uint32_t j = 0; j < nit; ++j) {
        const uint32_t c = base + (j < extra ? 1u : 0u);
        uint32_t halo = 0;
        if (grec)
            for (uint32_t q = pos; q < pos + c; ++q) halo = max(halo, grec[q].w >> 26);
        items[item_start[k] + j] = WorkItem{chan | (halo << 24), tile, pos, c};
        pos += c;
    }
}

static int bits_for(uint64_t v) {  // smallest b with v < 2^b
    int b = 0;
    while (b < 64 && (v >> b) != 0) ++b;
    return b;
}

static void exclusive_sum(const uint32_t *in, uint32_t *out, uint64_t n, hipStream_t st) {
    size_t bytes = 0;
    CHIP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, static_cast<int>(n), st));
    DevBuf tmp(std::max<size_t>(bytes, 16));
    CHIP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, static_cast<int>(n), st));
    CHIP_HIP(hipStreamSynchronize(st));
}

static uint32_t read_u32(const uint32_t *dev, hipStream_t st) {
    uint32_t v = 0;
    CHIP_HIP(hipMemcpyAsync(&v, dev, 4, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    return v;
}

// (Re)splits every bin's groups into work items of at most gpi groups and
// records the first item of each plane chunk.
void vset_split_items(chip_vset &vs, uint32_t gpi) {
    if (gpi == 0) fail(CHIP_ERR_INVALID, "groups per work item must be positive");
    hipStream_t st = vs.g->ctx->stream;
    vs.groups_per_item = gpi;
    if (vs.n_valid == 0 || vs.n_bins == 0) return;
    const uint32_t n_bins = vs.n_bins;
    DevBuf ipb((static_cast<uint64_t>(n_bins) + 1) * 4), item_start((static_cast<uint64_t>(n_bins) + 1) * 4);
    CHIP_HIP(hipMemsetAsync(ipb.p, 0, ipb.bytes, st));
    hipLaunchKernelGGL(k_items_per_bin, dim3(blocks_for(n_bins, 256)), dim3(256), 0, st,
                       vs.bin_groups.as<uint32_t>(), n_bins, gpi, ipb.as<uint32_t>());
    CHIP_HIP(hipGetLastError());
    exclusive_sum(ipb.as<uint32_t>(), item_start.as<uint32_t>(), n_bins + 1, st);
    std::vector<uint32_t> starts(static_cast<uint64_t>(n_bins) + 1);
    CHIP_HIP(hipMemcpyAsync(starts.data(), item_start.p, starts.size() * 4, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    const uint32_t n_items = starts[n_bins];
    vs.n_items = n_items;
    vs.items.alloc(std::max<uint64_t>(n_items, 1) * sizeof(WorkItem));
    hipLaunchKernelGGL(k_write_items, dim3(blocks_for(n_bins, 256)), dim3(256), 0, st,
                       vs.bin_groups.as<uint32_t>(), vs.bin_first.as<uint32_t>(),
                       item_start.as<uint32_t>(), n_bins, vs.n_tiles_all,
                       static_cast<uint32_t>(vs.n_chan), gpi, vs.g->wproject ? vs.grec.as<uint4>() : nullptr,
                       vs.items.as<WorkItem>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
    // Items ordered by (plane chunk, support class, bin): each (chunk, class)
    // range is launched with the LDS tile and register slots of its class
    // (range_first[chunk * n_classes + class] is its first item).  The class
    // of an item is that of its largest half-support (top 8 bits of its
    // channel word); the standard gridder has one class.
    const uint32_t n_cls = vs.g->wproject ? kSupportClasses : 1u;
    const uint64_t n_chunks = vs.chunk_items.size() - 1;
    const uint64_t per_chunk = static_cast<uint64_t>(vs.n_tiles_all) * vs.n_chan;
    std::vector<WorkItem> items(n_items);
    if (n_items) {
        CHIP_HIP(hipMemcpyAsync(items.data(), vs.items.p, n_items * sizeof(WorkItem), hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
    }
    auto class_of = [&](const WorkItem &w) { return n_cls == 1 ? 0u : support_class(static_cast<int32_t>(w.chan >> 24)); };
    // per (chunk, class): the items of every bin, in bin order
    std::vector<std::vector<uint32_t>> range_items(n_chunks * n_cls);
    for (uint64_t b = 0; b < n_bins; ++b)
        for (uint32_t k = starts[b]; k < starts[b + 1]; ++k)
            range_items[(b / per_chunk) * n_cls + class_of(items[k])].push_back(k);
    const bool xcd = vs.item_order == 1 && vs.sibling > 1 && n_chunks == 1;
    std::vector<WorkItem> order;
    order.reserve(n_items);
    vs.range_first.assign(n_chunks * n_cls + 1, 0u);
    for (uint64_t r = 0; r < range_items.size(); ++r) {
        vs.range_first[r] = static_cast<uint32_t>(order.size());
        const std::vector<uint32_t> &ri = range_items[r];
        if (!xcd || ri.empty()) {
            for (uint32_t k : ri) order.push_back(items[k]);
            continue;
        }
        // Sets of sibling work items: the items of one (sibling group, tile,
        // split) of the G planes of the group.  Set k goes to XCD k mod 8;
        // each XCD's queue holds its sets' items one after the other, and
        // the queues are interleaved so that dispatch position p (XCD p mod
        // 8, workgroups are dealt to the XCDs round robin) takes the next
        // item of queue p mod 8; shorter queues are padded with items of no
        // groups.
        constexpr uint32_t kXcd = 8;
        const uint32_t G = vs.sibling, n_tiles = vs.n_tiles_all;
        // this range's items of bin b: ri[bstart[b] .. bstart[b + 1]) (ri is in bin order)
        std::vector<uint32_t> bstart(static_cast<uint64_t>(n_bins) + 1, 0u);
        {
            uint64_t b = 0;
            for (uint32_t k : ri) {
                while (starts[b + 1] <= k) ++b;
                ++bstart[b + 1];
            }
            for (uint64_t q = 0; q < n_bins; ++q) bstart[q + 1] += bstart[q];
        }
        auto bin_count = [&](uint64_t b) { return static_cast<size_t>(bstart[b + 1] - bstart[b]); };
        std::vector<std::vector<WorkItem>> queue(kXcd);
        uint64_t set = 0;
        for (uint32_t grp = 0; grp * G < vs.n_chan; ++grp) {
            for (uint32_t t = 0; t < n_tiles; ++t) {
                size_t most = 0;
                for (uint32_t m = 0; m < G && grp * G + m < vs.n_chan; ++m)
                    most = std::max(most, bin_count(static_cast<uint64_t>(grp * G + m) * n_tiles + t));
                for (size_t q = 0; q < most; ++q, ++set) {
                    std::vector<WorkItem> &qu = queue[set % kXcd];
                    for (uint32_t m = 0; m < G && grp * G + m < vs.n_chan; ++m) {
                        const uint64_t b = static_cast<uint64_t>(grp * G + m) * n_tiles + t;
                        if (q < bin_count(b)) qu.push_back(items[ri[bstart[b] + q]]);
                    }
                }
            }
        }
        size_t longest = 0;
        for (const auto &q : queue) longest = std::max(longest, q.size());
        const size_t base = order.size();
        order.resize(base + longest * kXcd, WorkItem{0u, 0u, 0u, 0u});
        for (uint32_t x = 0; x < kXcd; ++x)
            for (size_t k = 0; k < queue[x].size(); ++k) order[base + k * kXcd + x] = queue[x][k];
    }
    vs.range_first.back() = static_cast<uint32_t>(order.size());
    vs.n_classes = n_cls;
    for (uint64_t c = 0; c <= n_chunks; ++c) vs.chunk_items[c] = vs.range_first[c * n_cls];
    vs.n_items = order.size();
    vs.items.alloc(std::max<uint64_t>(vs.n_items, 1) * sizeof(WorkItem));
    if (vs.n_items) {
        CHIP_HIP(hipMemcpyAsync(vs.items.p, order.data(), vs.n_items * sizeof(WorkItem), hipMemcpyHostToDevice, st));
        CHIP_HIP(hipStreamSynchronize(st));
    }
    vs.items_version += 1;
}

void vset_item_order(chip_vset &vs, uint32_t order) {
    if (order > 1) fail(CHIP_ERR_INVALID, "item order must be 0 or 1");
    vs.item_order = order;
    vset_split_items(vs, vs.groups_per_item);
}

namespace {

// Bins of a stream cut into tile x tile uv tiles.
struct BinSetup {
    BinParams b{};
    bool merge = false;
    int end_bit = 0;
    uint32_t n_bins = 0;
    uint64_t n_tiles = 0;
};

BinSetup bin_setup(chip_vset &vs, uint32_t tile) {
    chip_gridder &g = *vs.g;
    BinSetup s;
    vs.tile = tile;
    vs.tiles_x = static_cast<uint32_t>((g.gx + tile - 1) / tile);
    vs.tiles_y = static_cast<uint32_t>((g.gy + tile - 1) / tile);
    s.n_tiles = static_cast<uint64_t>(vs.tiles_x) * vs.tiles_y;
    const uint64_t n_chunks = vs.wk_chunk ? (g.planes + vs.wk_chunk - 1) / vs.wk_chunk : 1;
    const uint64_t n_bins64 = s.n_tiles * vs.n_chan * n_chunks;
    if (n_bins64 >= (1ull << 31)) fail(CHIP_ERR_INVALID, "too many (plane chunk, channel, tile) bins");
    vs.chunk_items.assign(n_chunks + 1, 0u);
    s.n_bins = static_cast<uint32_t>(n_bins64);
    BinParams &b = s.b;
    b.tile = tile;
    b.tiles_x = vs.tiles_x;
    b.tiles_y = vs.tiles_y;
    b.n_tiles = static_cast<uint32_t>(s.n_tiles);
    b.n_bins = s.n_bins;
    b.n_chan = static_cast<uint32_t>(vs.n_chan);
    b.chunk_planes = vs.wk_chunk;
    b.cell_bits = static_cast<uint32_t>(bits_for(static_cast<uint64_t>(tile) * tile - 1));
    // W tables recomputed per plane chunk keep one class (their launches
    // follow the chunks)
    b.class_bits = (g.wproject && !vs.wk_chunk) ? kClassBits : 0u;
    b.supports = g.wproject ? g.d_supports.as<int32_t>() : nullptr;
    b.n_sel = g.wproject ? static_cast<uint32_t>(2 * g.norm_radius + 1) : g.os + 1;
    const uint64_t max_code = g.wproject
        ? (static_cast<uint64_t>(g.planes) * 2 * b.n_sel) * b.n_sel - 1
        : static_cast<uint64_t>(b.n_sel) * b.n_sel - 1;
    const int bin_bits = bits_for(s.n_bins);  // includes the invalid bin
    const int plan_bits = bits_for(max_code);
    s.merge = bin_bits + static_cast<int>(b.class_bits + b.cell_bits) + plan_bits <= 64;
    b.plan_bits = s.merge ? static_cast<uint32_t>(plan_bits) : 0u;
    s.end_bit = bin_bits + static_cast<int>(b.class_bits + b.cell_bits + b.plan_bits);
    if (s.end_bit > 64) fail(CHIP_ERR_INVALID, "sort key exceeds 64 bits");
    return s;
}

// Radix sort of (keys, vals) by the low end_bit bits into (keys_s, vals_s).
void sort_pairs(DevBuf &keys, DevBuf &vals, DevBuf &keys_s, DevBuf &vals_s, uint64_t n, int end_bit, hipStream_t st) {
    size_t temp_bytes = 0;
    CHIP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.as<uint64_t>(), keys_s.as<uint64_t>(),
                                                vals.as<uint32_t>(), vals_s.as<uint32_t>(), static_cast<int>(n), 0,
                                                end_bit, st));
    DevBuf temp(std::max<size_t>(temp_bytes, 16));
    CHIP_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, keys.as<uint64_t>(), keys_s.as<uint64_t>(),
                                                vals.as<uint32_t>(), vals_s.as<uint32_t>(), static_cast<int>(n), 0,
                                                end_bit, st));
    CHIP_HIP(hipStreamSynchronize(st));
}

// Groups of identical plans, their records and sums, and the work items, from
// the stream sorted by keys_s.
void build_groups(chip_vset &vs, const BinSetup &s, const DevBuf &keys_s) {
    chip_gridder &g = *vs.g;
    hipStream_t st = g.ctx->stream;
    const uint64_t n_valid = vs.n_valid;
    const uint32_t n_bins = s.n_bins;
    DevBuf flags((n_valid + 1) * 4), gid((n_valid + 1) * 4);
    CHIP_HIP(hipMemsetAsync(flags.p, 0, flags.bytes, st));
    hipLaunchKernelGGL(k_group_flags, dim3(blocks_for(n_valid, 256)), dim3(256), 0, st,
                       keys_s.as<uint64_t>(), n_valid, s.merge, flags.as<uint32_t>());
    CHIP_HIP(hipGetLastError());
    exclusive_sum(flags.as<uint32_t>(), gid.as<uint32_t>(), n_valid + 1, st);
    const uint32_t n_groups = read_u32(gid.as<uint32_t>() + n_valid, st);
    vs.n_groups = n_groups;
    DevBuf gstart((static_cast<uint64_t>(n_groups) + 1) * 4);
    DevBuf bin_groups((static_cast<uint64_t>(n_bins) + 1) * 4);
    CHIP_HIP(hipMemsetAsync(bin_groups.p, 0, bin_groups.bytes, st));
    const uint32_t low_bits = s.b.class_bits + s.b.cell_bits + s.b.plan_bits;
    hipLaunchKernelGGL(k_group_starts, dim3(blocks_for(n_valid, 256)), dim3(256), 0, st,
                       keys_s.as<uint64_t>(), flags.as<uint32_t>(), gid.as<uint32_t>(), n_valid,
                       low_bits, gstart.as<uint32_t>(), bin_groups.as<uint32_t>());
    CHIP_HIP(hipGetLastError());
    const uint32_t nv32 = static_cast<uint32_t>(n_valid);
    CHIP_HIP(hipMemcpyAsync(gstart.as<uint32_t>() + n_groups, &nv32, 4, hipMemcpyHostToDevice, st));
    vs.grec.alloc(std::max<uint64_t>(n_groups, 1) * sizeof(uint4));
    if (n_groups) {
        const uint32_t n_off = static_cast<uint32_t>(2 * g.norm_radius + 1);
        if (g.wproject) {
            if (g.max_support > 63) fail(CHIP_ERR_INVALID, "W support above 63 for the binned stream");
            if (static_cast<uint64_t>(g.planes) * n_off * n_off >= (1ull << kRecPoBits))
                fail(CHIP_ERR_INVALID, "too many W tap tables for the binned stream");
        }
        hipLaunchKernelGGL(k_group_records, dim3(blocks_for(n_groups, 256)), dim3(256), 0, st,
                           gstart.as<uint32_t>(), vs.plan.as<uint2>(), n_groups,
                           g.wproject ? g.d_supports.as<int32_t>() : nullptr, n_off,
                           static_cast<int32_t>(g.norm_radius), vs.grec.as<uint4>());
        CHIP_HIP(hipGetLastError());
    }
    vs.gsum_d.alloc(std::max<uint64_t>(n_groups, 1) * sizeof(float4));
    vs.gsum_r.alloc(std::max<uint64_t>(n_groups, 1) * sizeof(float4));
    if (n_groups) {
        hipLaunchKernelGGL(k_group_sums, dim3(blocks_for(n_groups, 256)), dim3(256), 0, st, vs.grec.as<uint4>(),
                           vs.plan.as<uint2>(), vs.vis.as<float2>(), vs.wt.as<float>(), n_groups,
                           vs.gsum_d.as<float4>(), vs.gsum_r.as<float4>());
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
    gstart.release();
    flags.release();
    gid.release();

    // work items
    vs.bin_first.alloc((static_cast<uint64_t>(n_bins) + 1) * 4);
    exclusive_sum(bin_groups.as<uint32_t>(), vs.bin_first.as<uint32_t>(), n_bins + 1, st);
    vs.bin_groups = std::move(bin_groups);
    vs.n_bins = n_bins;
    vs.n_tiles_all = static_cast<uint32_t>(s.n_tiles);
    vset_split_items(vs, vs.groups_per_item);
}

// Sort keys of a stream already planned: (bin, anchor cell, plan) from the
// packed plan and the sample's original index (channel = index mod n_chan),
// the same key k_keys makes from the samples.
__device__ __forceinline__ uint64_t rekey_one(const uint2 pk, uint32_t chan, const BinParams &b, uint32_t wproject,
                                              int32_t R) {
    const uint32_t ax = (pk.x & 0xffffu) + (wproject ? 0u : static_cast<uint32_t>(kSupport));
    const uint32_t ay = (pk.x >> 16) + (wproject ? 0u : static_cast<uint32_t>(kSupport));
    const uint32_t plane = packed_table(pk.y);
    const uint32_t tile = (ax / b.tile) * b.tiles_y + (ay / b.tile);
    uint64_t bin = static_cast<uint64_t>(chan) * b.n_tiles + tile;
    if (b.chunk_planes) bin += static_cast<uint64_t>(plane / b.chunk_planes) * b.n_chan * b.n_tiles;
    const uint64_t cell = static_cast<uint64_t>((ax % b.tile) * b.tile + (ay % b.tile));
    uint64_t key = bin << (b.class_bits + b.cell_bits + b.plan_bits);
    if (b.class_bits) key |= static_cast<uint64_t>(support_class(b.supports[plane])) << (b.cell_bits + b.plan_bits);
    if (!b.plan_bits) {
        key |= cell;
    } else {
        uint64_t code;
        if (wproject) {
            const uint64_t pc = static_cast<uint64_t>(plane) * 2 + ((pk.y & kConjBit) ? 1u : 0u);
            const uint64_t sx = (pk.y & 0xffu) - 128u + static_cast<uint32_t>(R);
            const uint64_t sy = ((pk.y >> 8) & 0xffu) - 128u + static_cast<uint32_t>(R);
            code = (pc * b.n_sel + sx) * b.n_sel + sy;
        } else {
            code = static_cast<uint64_t>(pk.y & 0xffu) * b.n_sel + ((pk.y >> 8) & 0xffu);
        }
        key |= b.plan_major ? ((code << b.cell_bits) | cell) : ((cell << b.plan_bits) | code);
    }
    return key;
}

__global__ void k_rekey(const uint2 *plan, const uint32_t *index, uint64_t n, BinParams b, uint32_t wproject,
                        int32_t R, uint64_t *keys, uint32_t *vals) {
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint32_t chan = b.n_chan == 1 ? 0u : index[p] % b.n_chan;
    keys[p] = rekey_one(plan[p], chan, b, wproject, R);
    vals[p] = static_cast<uint32_t>(p);
}

// The same key per group, from the packed plan of its first sample.
__global__ void k_rekey_groups(const uint4 *grec, const uint2 *plan, const uint32_t *index, uint64_t n_groups,
                               BinParams b, uint32_t wproject, int32_t R, uint64_t *keys, uint32_t *vals) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    const uint32_t s0 = grec[g].x;
    const uint32_t chan = b.n_chan == 1 ? 0u : index[s0] % b.n_chan;
    keys[g] = rekey_one(plan[s0], chan, b, wproject, R);
    vals[g] = static_cast<uint32_t>(g);
}

__global__ void k_bin_counts(const uint64_t *keys_s, uint64_t n, uint32_t low_bits, uint32_t *bin_groups) {
    const uint64_t g = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n) return;
    atomicAdd(&bin_groups[keys_s[g] >> low_bits], 1u);
}

template <class T>
__global__ void k_gather(const T *src, const uint32_t *perm, uint64_t n, T *dst) {
    const uint64_t p = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[perm[p]];
}

template <class T>
void gather(DevBuf &buf, const DevBuf &perm, uint64_t n, hipStream_t st) {
    DevBuf out(std::max<uint64_t>(n, 1) * sizeof(T));
    hipLaunchKernelGGL(k_gather<T>, dim3(blocks_for(n, 256)), dim3(256), 0, st, buf.as<T>(), perm.as<uint32_t>(), n,
                       out.as<T>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
    buf = std::move(out);
}

}  // namespace

// Builds the resident stream for a vset: FP64 plans, 64-bit radix sort by
// (bin, anchor cell, plan), groups of identical plans, work items.
void vset_build(chip_vset &vs, const chip_samples &cs) {
    chip_gridder &g = *vs.g;
    hipStream_t st = g.ctx->stream;
    const Source src = make_source(cs);
    const uint64_t n = src.n_rows * src.n_chan;
    vs.n_samples = n;
    vs.n_chan = src.n_chan;
    vs.sibling = 1;
    if (src.n_chan >= (1u << 24)) fail(CHIP_ERR_INVALID, "a visibility set holds fewer than 2^24 channels");
    if (src.uvw_m && src.n_chan > 1) {
        // runs of planes with bit-identical lambda scales share their plans
        std::vector<double> ls(src.n_chan);
        CHIP_HIP(hipMemcpyAsync(ls.data(), src.lambda_scale, ls.size() * sizeof(double), hipMemcpyDeviceToHost, st));
        CHIP_HIP(hipStreamSynchronize(st));
        uint32_t run = 1;
        while (run < src.n_chan && ls[run] == ls[0]) ++run;
        bool regular = run > 1 && src.n_chan % run == 0;
        for (uint64_t c = 0; regular && c < src.n_chan; ++c) {
            const uint64_t first = c - c % run;  // first plane of c's run
            if (ls[c] != ls[first]) regular = false;
            if (c == first && c > 0 && ls[c] == ls[c - 1]) regular = false;  // runs of exactly `run`
        }
        if (regular) vs.sibling = run;
    }
    vs.item_order = vs.sibling > 1 ? 1u : 0u;  // sibling planes together on one XCD
    vs.halo = g.wproject ? g.max_support : static_cast<uint32_t>(kSupport);
    vs.wk_chunk = (g.wproject && !g.aw) ? g.wk_chunk : 0u;
    // 32x32 tiles; smaller tiles when a large W support would not leave the
    // (tile + 2 halo)^2 LDS tile inside one workgroup's LDS
    const size_t lds = lds_limit(vs.device);
    uint32_t tile = 32;
    while (tile > 8 && binned_tile_lds(g, tile, vs.halo) > lds) tile /= 2;
    if (binned_tile_lds(g, tile, vs.halo) > lds)
        fail(CHIP_ERR_INVALID, "W support " + std::to_string(vs.halo) + " too large for an LDS tile");
    const BinSetup bs = bin_setup(vs, tile);
    vs.sums.assign(src.n_chan, chip_sums{});
    CHIP_HIP(hipEventCreate(&vs.ev_start));
    CHIP_HIP(hipEventCreate(&vs.ev_stop));
    if (n == 0) return;

    DevBuf keys(n * 8), vals(n * 4), keys_s(n * 8), vals_s(n * 4);
    DevBuf d_nvalid(8);
    CHIP_HIP(hipMemsetAsync(d_nvalid.p, 0, 8, st));
    const PlanParams pp = plan_params(g, src);
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, pp, src, bs.b,
                       keys.as<uint64_t>(), vals.as<uint32_t>(), d_nvalid.as<unsigned long long>());
    CHIP_HIP(hipGetLastError());
    sort_pairs(keys, vals, keys_s, vals_s, n, bs.end_bit, st);
    keys.release();
    vals.release();
    unsigned long long nv_ll = 0;
    CHIP_HIP(hipMemcpyAsync(&nv_ll, d_nvalid.p, 8, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    const uint64_t n_valid = nv_ll;
    vs.n_valid = n_valid;

    // packed per-sample stream + per-channel sums
    const uint64_t nvb = std::max<uint64_t>(n_valid, 1);
    vs.plan.alloc(nvb * sizeof(uint2));
    vs.vis.alloc(nvb * sizeof(float2));
    vs.wt.alloc(nvb * sizeof(float));
    vs.index.alloc(nvb * sizeof(uint32_t));
    DevBuf sums(src.n_chan * kSumFields * sizeof(double));
    CHIP_HIP(hipMemsetAsync(sums.p, 0, sums.bytes, st));
    if (n_valid > 0) {
        hipLaunchKernelGGL(k_build_stream, dim3(blocks_for(n_valid, 256)), dim3(256), 0, st, pp,
                           src, n_valid, vals_s.as<uint32_t>(), vs.plan.as<uint2>(),
                           vs.vis.as<float2>(), vs.wt.as<float>(), vs.index.as<uint32_t>(),
                           sums.as<double>());
        CHIP_HIP(hipGetLastError());
    }
    std::vector<double> hs(src.n_chan * kSumFields);
    CHIP_HIP(hipMemcpyAsync(hs.data(), sums.p, sums.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    vs.taps = 0;
    for (uint64_t c = 0; c < src.n_chan; ++c) {
        const double *h = hs.data() + kSumFields * c;
        vs.sums[c].normalization_sumwt = h[0];
        vs.sums[c].reported_sumwt = h[1];
        vs.sums[c].gridded = static_cast<uint64_t>(h[2]);
        vs.sums[c].skipped = src.n_rows - vs.sums[c].gridded;
        vs.taps += static_cast<uint64_t>(h[3]);
    }
    vals_s.release();
    if (n_valid == 0) return;
    build_groups(vs, bs, keys_s);
}

// Cuts the stream into tile x tile uv tiles again, with the groups of a bin
// ordered by anchor cell (order 0) or by plan (order 1).  Group membership
// does not change (identical plans share their anchor, so they share a tile
// of any size); only the order of groups and their binning do, so the groups
// are sorted by their new keys (from the packed plan of their first sample)
// and their records and sums permuted, then bins and work items rebuilt.
// force: the samples are sorted again and the groups rebuilt from them.
void vset_rebin(chip_vset &vs, uint32_t tile, uint32_t order, bool force) {
    chip_gridder &g = *vs.g;
    if (order > 1) fail(CHIP_ERR_INVALID, "group order must be 0 (anchor cell first) or 1 (plan first)");
    if (tile == vs.tile && order == vs.order && !force) return;
    if (tile < 8 || tile > 256 || (tile & (tile - 1)))
        fail(CHIP_ERR_INVALID, "uv tile must be a power of two in 8 .. 256");
    if (binned_tile_lds(g, tile, vs.halo) > lds_limit(vs.device))
        fail(CHIP_ERR_INVALID, "uv tile of " + std::to_string(tile) + " does not fit the LDS with its halo");
    hipStream_t st = g.ctx->stream;
    BinSetup bs = bin_setup(vs, tile);
    bs.b.plan_major = order;
    vs.order = order;
    const uint64_t n = vs.n_valid;
    if (n == 0) return;
    if (vs.n_groups > 0 && !force) {
        // Groups keep their members (identical plans share their anchor, so
        // they share a tile of any size): sort the groups, not the samples,
        // by their new keys and permute their records and sums; each group's
        // samples stay where they are.
        const uint64_t ng = vs.n_groups;
        DevBuf keys(ng * 8), vals(ng * 4), keys_s(ng * 8), vals_s(ng * 4);
        hipLaunchKernelGGL(k_rekey_groups, dim3(blocks_for(ng, 256)), dim3(256), 0, st, vs.grec.as<uint4>(),
                           vs.plan.as<uint2>(), vs.index.as<uint32_t>(), ng, bs.b, g.wproject ? 1u : 0u,
                           static_cast<int32_t>(g.norm_radius), keys.as<uint64_t>(), vals.as<uint32_t>());
        CHIP_HIP(hipGetLastError());
        sort_pairs(keys, vals, keys_s, vals_s, ng, bs.end_bit, st);
        keys.release();
        vals.release();
        gather<uint4>(vs.grec, vals_s, ng, st);
        gather<float4>(vs.gsum_d, vals_s, ng, st);
        gather<float4>(vs.gsum_r, vals_s, ng, st);
        vals_s.release();
        const uint32_t n_bins = bs.n_bins;
        DevBuf bin_groups((static_cast<uint64_t>(n_bins) + 1) * 4);
        CHIP_HIP(hipMemsetAsync(bin_groups.p, 0, bin_groups.bytes, st));
        hipLaunchKernelGGL(k_bin_counts, dim3(blocks_for(ng, 256)), dim3(256), 0, st, keys_s.as<uint64_t>(), ng,
                           bs.b.class_bits + bs.b.cell_bits + bs.b.plan_bits, bin_groups.as<uint32_t>());
        CHIP_HIP(hipGetLastError());
        vs.bin_first.alloc((static_cast<uint64_t>(n_bins) + 1) * 4);
        exclusive_sum(bin_groups.as<uint32_t>(), vs.bin_first.as<uint32_t>(), n_bins + 1, st);
        vs.bin_groups = std::move(bin_groups);
        vs.n_bins = n_bins;
        vs.n_tiles_all = static_cast<uint32_t>(bs.n_tiles);
        vset_split_items(vs, vs.groups_per_item);
        return;
    }
    DevBuf keys(n * 8), vals(n * 4), keys_s(n * 8), vals_s(n * 4);
    hipLaunchKernelGGL(k_rekey, dim3(blocks_for(n, 256)), dim3(256), 0, st, vs.plan.as<uint2>(),
                       vs.index.as<uint32_t>(), n, bs.b, g.wproject ? 1u : 0u, static_cast<int32_t>(g.norm_radius),
                       keys.as<uint64_t>(), vals.as<uint32_t>());
    CHIP_HIP(hipGetLastError());
    sort_pairs(keys, vals, keys_s, vals_s, n, bs.end_bit, st);
    keys.release();
    vals.release();
    gather<uint2>(vs.plan, vals_s, n, st);
    gather<float2>(vs.vis, vals_s, n, st);
    gather<float>(vs.wt, vals_s, n, st);
    gather<uint32_t>(vs.index, vals_s, n, st);
    vals_s.release();
    build_groups(vs, bs, keys_s);
}

}  // namespace chip
