// FP64 planning on the GPU and the binned visibility stream.
//
// Planning restates StandardGridder::plan_positive_taps (gridder.rs:557-566)
// and WProjector::plan_sample (gridder.rs:3588-3624) bit for bit: the only
// operations are IEEE f64 division, addition, multiplication, sqrt and
// round-half-away-from-zero, each separately rounded.
//
// The resident stream replaces the reference's per-chunk sample packing
// (lib.rs:71708-71767): accepted samples are keyed by (channel, uv tile),
// sorted with an integer radix sort, and split into work items of at most
// kChunk samples so that one launch covers every channel with balanced
// workgroups.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "device.hpp"

namespace chip {

constexpr uint32_t kChunk = 2048;  // samples per work item (upper bound)

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
    p.norm_table = g.wproject ? g.d_norm_table.as<float>() : nullptr;
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
    hipLaunchKernelGGL(k_plan_raw, dim3(blocks_for(n, 256)), dim3(256), 0, st, plan_params(g),
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
struct BinParams {
    uint32_t tile;
    uint32_t tiles_x, tiles_y;
    uint32_t n_tiles;
    uint32_t invalid_key;  // n_chan * n_tiles
};

// Key = channel * n_tiles + tile of the anchor cell; rejected samples get
// invalid_key and sort to the end.  Also histograms the keys.
__global__ __launch_bounds__(256) void k_keys(PlanParams p, Source s, BinParams b,
                                              uint32_t *keys, uint32_t *vals, uint32_t *hist) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t n = s.n_rows * s.n_chan;
    if (i >= n) return;
    const SampleIn smp = load_sample(s, i);
    Plan pl;
    uint32_t key = b.invalid_key;
    if (sample_accepted(p, smp, pl)) {
        const uint32_t ax = static_cast<uint32_t>(p.wproject ? pl.x : pl.x + kSupport);
        const uint32_t ay = static_cast<uint32_t>(p.wproject ? pl.y : pl.y + kSupport);
        const uint32_t tile = (ax / b.tile) * b.tiles_y + (ay / b.tile);
        key = smp.chan * b.n_tiles + tile;
        atomicAdd(&hist[key], 1u);
    }
    keys[i] = key;
    vals[i] = static_cast<uint32_t>(i);
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
        if (p.wproject) {
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

// Number of work items per bin.
__global__ void k_items_per_bin(const uint32_t *hist, uint32_t n_bins, uint32_t *items_per_bin) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bins) return;
    items_per_bin[k] = (hist[k] + kChunk - 1) / kChunk;
}

__global__ void k_write_items(const uint32_t *hist, const uint32_t *bin_start,
                              const uint32_t *item_start, uint32_t n_bins, uint32_t n_tiles,
                              WorkItem *items) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n_bins) return;
    const uint32_t count = hist[k];
    if (count == 0) return;
    const uint32_t chan = k / n_tiles;
    const uint32_t tile = k - chan * n_tiles;
    const uint32_t nit = (count + kChunk - 1) / kChunk;
    // Spread the bin evenly over its items.
    const uint32_t base = count / nit, extra = count % nit;
    uint32_t pos = bin_start[k];
    for (uint32_t j = 0; j < nit; ++j) {
        const uint32_t c = base + (j < extra ? 1u : 0u);
        items[item_start[k] + j] = WorkItem{chan, tile, pos, c};
        pos += c;
    }
}

}  // namespace chip

using namespace chip;

// Builds the resident stream for a vset.  Defined here because it owns the
// sort; the C-ABI wrapper in api.cpp calls it.
namespace chip {

void vset_build(chip_vset &vs, const chip_samples &cs) {
    chip_gridder &g = *vs.g;
    hipStream_t st = g.ctx->stream;
    const Source src = make_source(cs);
    const uint64_t n = src.n_rows * src.n_chan;
    vs.n_samples = n;
    vs.n_chan = src.n_chan;
    vs.halo = g.wproject ? g.max_support : static_cast<uint32_t>(kSupport);
    vs.tile = 32;
    vs.tiles_x = static_cast<uint32_t>((g.gx + vs.tile - 1) / vs.tile);
    vs.tiles_y = static_cast<uint32_t>((g.gy + vs.tile - 1) / vs.tile);
    const uint64_t n_tiles = static_cast<uint64_t>(vs.tiles_x) * vs.tiles_y;
    const uint64_t n_bins64 = n_tiles * src.n_chan;
    if (n_bins64 >= (1ull << 31)) fail(CHIP_ERR_INVALID, "too many (channel, tile) bins");
    const uint32_t n_bins = static_cast<uint32_t>(n_bins64);
    vs.sums.assign(src.n_chan, chip_sums{});
    if (n == 0) return;

    BinParams b{vs.tile, vs.tiles_x, vs.tiles_y, static_cast<uint32_t>(n_tiles), n_bins};
    int end_bit = 1;
    while (end_bit < 32 && (1ull << end_bit) <= n_bins) ++end_bit;

    DevBuf keys(n * 4), vals(n * 4), keys_s(n * 4), vals_s(n * 4);
    DevBuf hist((static_cast<uint64_t>(n_bins) + 1) * 4);
    CHIP_HIP(hipMemsetAsync(hist.p, 0, hist.bytes, st));
    const PlanParams pp = plan_params(g);
    hipLaunchKernelGGL(k_keys, dim3(blocks_for(n, 256)), dim3(256), 0, st, pp, src, b,
                       keys.as<uint32_t>(), vals.as<uint32_t>(), hist.as<uint32_t>());
    CHIP_HIP(hipGetLastError());

    size_t temp_bytes = 0;
    CHIP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys.as<uint32_t>(),
                                                keys_s.as<uint32_t>(), vals.as<uint32_t>(),
                                                vals_s.as<uint32_t>(), static_cast<int>(n), 0,
                                                end_bit, st));
    DevBuf temp(std::max<size_t>(temp_bytes, 16));
    CHIP_HIP(hipcub::DeviceRadixSort::SortPairs(temp.p, temp_bytes, keys.as<uint32_t>(),
                                                keys_s.as<uint32_t>(), vals.as<uint32_t>(),
                                                vals_s.as<uint32_t>(), static_cast<int>(n), 0,
                                                end_bit, st));
    keys.release();
    vals.release();

    // Bin starts and item starts by exclusive scans of the histogram.
    DevBuf bin_start((static_cast<uint64_t>(n_bins) + 1) * 4);
    DevBuf ipb((static_cast<uint64_t>(n_bins) + 1) * 4), item_start((static_cast<uint64_t>(n_bins) + 1) * 4);
    size_t scan_bytes = 0;
    CHIP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, hist.as<uint32_t>(),
                                              bin_start.as<uint32_t>(), n_bins + 1, st));
    DevBuf scan_tmp(std::max<size_t>(scan_bytes, 16));
    CHIP_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, hist.as<uint32_t>(),
                                              bin_start.as<uint32_t>(), n_bins + 1, st));
    CHIP_HIP(hipMemsetAsync(ipb.p, 0, ipb.bytes, st));
    hipLaunchKernelGGL(k_items_per_bin, dim3(blocks_for(n_bins, 256)), dim3(256), 0, st,
                       hist.as<uint32_t>(), n_bins, ipb.as<uint32_t>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipcub::DeviceScan::ExclusiveSum(scan_tmp.p, scan_bytes, ipb.as<uint32_t>(),
                                              item_start.as<uint32_t>(), n_bins + 1, st));
    uint32_t n_valid = 0, n_items = 0;
    CHIP_HIP(hipMemcpyAsync(&n_valid, bin_start.as<uint32_t>() + n_bins, 4, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipMemcpyAsync(&n_items, item_start.as<uint32_t>() + n_bins, 4, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    vs.n_valid = n_valid;
    vs.n_items = n_items;

    vs.items.alloc(std::max<uint64_t>(n_items, 1) * sizeof(WorkItem));
    hipLaunchKernelGGL(k_write_items, dim3(blocks_for(n_bins, 256)), dim3(256), 0, st,
                       hist.as<uint32_t>(), bin_start.as<uint32_t>(), item_start.as<uint32_t>(),
                       n_bins, static_cast<uint32_t>(n_tiles), vs.items.as<WorkItem>());
    CHIP_HIP(hipGetLastError());

    const uint64_t nv = std::max<uint64_t>(n_valid, 1);
    vs.plan.alloc(nv * sizeof(uint2));
    vs.vis.alloc(nv * sizeof(float2));
    vs.wt.alloc(nv * sizeof(float));
    vs.index.alloc(nv * sizeof(uint32_t));
    DevBuf sums(src.n_chan * kSumFields * sizeof(double));
    CHIP_HIP(hipMemsetAsync(sums.p, 0, sums.bytes, st));
    if (n_valid > 0) {
        hipLaunchKernelGGL(k_build_stream, dim3(blocks_for(n_valid, 256)), dim3(256), 0, st, pp,
                           src, static_cast<uint64_t>(n_valid), vals_s.as<uint32_t>(),
                           vs.plan.as<uint2>(), vs.vis.as<float2>(), vs.wt.as<float>(),
                           vs.index.as<uint32_t>(), sums.as<double>());
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
    CHIP_HIP(hipEventCreate(&vs.ev_start));
    CHIP_HIP(hipEventCreate(&vs.ev_stop));
}

}  // namespace chip
