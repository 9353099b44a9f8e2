// Row-stream ingestion: host row blocks (uvw, DATA / FLAG [corr][chan],
// WEIGHT or WEIGHT_SPECTRUM, FLAG_ROW) into the rows x channels samples of
// one imaging plane, converted on the GPU behind double-buffered copies.
//
// Per (row, chan) the conversion follows the routed-sample builder
// (execution.rs:2168-2315): an explicit correlation is excluded when flagged
// and, unless non-finite visibilities may still add to the PSF, when
// non-finite; a collapsed pair is excluded when either hand is flagged, the
// collapsed visibility is non-finite or either weight is not finite and
// positive, and then weighs 0.5 (w0 + w1).  With a phase-centre rotation
// each hand is rotated before the collapse (casars-imager lib.rs:19207-19211).
//
// Staging: block k goes host -> pinned slot k % 2 (CPU copy) -> device slot
// k % 2 (async copy on a copy stream) -> conversion on the context's stream
// into the resident stream.  Reusing a slot waits for its previous copy
// (host side) and for its previous conversion (copy stream side), so at most
// two blocks are live, as in the reference's read-ahead contract
// (ARCHITECTURE.md:461-475).
#include <cmath>
#include <cstring>

#include "device.hpp"

namespace chip {

namespace {

struct IngestGeom {
    uint32_t n_corr, n_chan;
    int32_t pol, transform;
    uint32_t c0, c1;
    float sumwt_factor;
    int32_t allow_psf_only, spectrum, rotate;
    double uvrot[9], phrot[3];
    uint64_t n_rows;    // rows of this block
    uint64_t row0;      // first resident row of this block
    // byte offsets of the block sections in a staging slot
    uint64_t off_data, off_flag, off_weight, off_row_flag;
    int32_t has_row_flag;
};

// Staging layout of a block of `rows` rows: uvw, data, flag, weight, row flag,
// each section 256-byte aligned.
struct Layout {
    uint64_t data, flag, weight, row_flag, bytes;
};
Layout layout(const chip_ingest_params &p, uint64_t rows) {
    auto up = [](uint64_t b) { return (b + 255) & ~uint64_t(255); };
    const uint64_t per = static_cast<uint64_t>(p.n_corr) * p.n_chan;
    Layout l{};
    l.data = up(rows * 3 * sizeof(double));
    l.flag = l.data + up(rows * per * 2 * sizeof(float));
    l.weight = l.flag + up(rows * per);
    l.row_flag = l.weight + up(rows * (p.weight_spectrum ? per : p.n_corr) * sizeof(float));
    l.bytes = l.row_flag + up(rows);
    return l;
}

// StandardMfsPairCollapseTransform::collapse (types.rs:1593-1610): Float
// complex arithmetic, the imaginary variants as a full complex product.
__device__ __forceinline__ float2 collapse(int32_t t, float2 a, float2 b) {
#pragma clang fp contract(off)
    if (t == CHIP_PAIR_HALF_SUM) return make_float2((a.x + b.x) * 0.5f, (a.y + b.y) * 0.5f);
    const float2 d = make_float2(a.x - b.x, a.y - b.y);
    if (t == CHIP_PAIR_HALF_DIFFERENCE) return make_float2(d.x * 0.5f, d.y * 0.5f);
    const float im = t == CHIP_PAIR_POSITIVE_HALF_IMAG_DIFF ? 0.5f : -0.5f;
    return make_float2(d.x * 0.0f - d.y * im, d.x * im + d.y * 0.0f);
}

__device__ __forceinline__ bool finite2(float2 v) { return isfinite(v.x) && isfinite(v.y); }

// One thread per row: rotated uvw and the row's phase shift.
__global__ __launch_bounds__(256) void k_ingest_rows(IngestGeom g, const uint8_t *__restrict__ stage,
                                                     double *__restrict__ uvw_out, double *__restrict__ shift) {
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= g.n_rows) return;
    const double *uvw = reinterpret_cast<const double *>(stage) + 3 * r;
    const double x[3] = {uvw[0], uvw[1], uvw[2]};
    double o[3] = {x[0], x[1], x[2]}, s = 0.0;
    if (g.rotate) rotate_uvw_row(g.uvrot, g.phrot, x, o, s);
    double *u = uvw_out + 3 * (g.row0 + r);
    u[0] = o[0];
    u[1] = o[1];
    u[2] = o[2];
    shift[r] = s;
}

// One thread per (row, chan) of the block, channel fastest.
__global__ __launch_bounds__(256) void k_ingest_samples(IngestGeom g, const uint8_t *__restrict__ stage,
                                                        const double *__restrict__ shift,
                                                        const double *__restrict__ freq, float2 *__restrict__ vis,
                                                        float *__restrict__ wt, float *__restrict__ sf,
                                                        uint8_t *__restrict__ flag) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= g.n_rows * g.n_chan) return;
    const uint64_t r = i / g.n_chan, c = i - r * g.n_chan;
    const float2 *data = reinterpret_cast<const float2 *>(stage + g.off_data);
    const uint8_t *fl = stage + g.off_flag;
    const float *w = reinterpret_cast<const float *>(stage + g.off_weight);
    const uint64_t per = static_cast<uint64_t>(g.n_corr) * g.n_chan;
    auto at = [&](uint32_t corr) { return r * per + static_cast<uint64_t>(corr) * g.n_chan + c; };
    auto weight = [&](uint32_t corr) {
        return g.spectrum ? w[at(corr)] : w[r * g.n_corr + corr];
    };
    const double s = shift[r], f = freq[c];
    bool excluded = g.has_row_flag && stage[g.off_row_flag + r] != 0;
    float2 v = make_float2(0.0f, 0.0f);
    float wv = 0.0f;
    if (g.pol == CHIP_POL_EXPLICIT) {
        excluded = excluded || fl[at(g.c0)] != 0;
        v = data[at(g.c0)];
        if (!finite2(v) && !g.allow_psf_only) excluded = true;
        if (g.rotate) v = phase_rotate(v, s, f);
        wv = weight(g.c0);
    } else {
        excluded = excluded || fl[at(g.c0)] != 0 || fl[at(g.c1)] != 0;
        float2 a = data[at(g.c0)], b = data[at(g.c1)];
        if (g.rotate) {
            a = phase_rotate(a, s, f);
            b = phase_rotate(b, s, f);
        }
        v = collapse(g.transform, a, b);
        const float w0 = weight(g.c0), w1 = weight(g.c1);
        if (!finite2(v) || !(isfinite(w0) && w0 > 0.0f && isfinite(w1) && w1 > 0.0f)) excluded = true;
        wv = 0.5f * (w0 + w1);
    }
    const uint64_t o = (g.row0 + r) * g.n_chan + c;
    vis[o] = v;
    wt[o] = wv;
    sf[o] = g.sumwt_factor;
    flag[o] = excluded ? 1 : 0;
}

unsigned blocks_of(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

void ingest_destroy_slots(IngestSlot *slots, int n) {
    for (int k = 0; k < n; ++k) {
        IngestSlot &s = slots[k];
        if (s.copied) (void)hipEventSynchronize(s.copied);
        if (s.consumed) (void)hipEventSynchronize(s.consumed);
        if (s.host) (void)hipHostFree(s.host);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.consumed) (void)hipEventDestroy(s.consumed);
        s.host = nullptr;
        s.copied = s.consumed = nullptr;
        s.dev.release();
    }
}

void ingest_create(chip_ingest &ing, const chip_ingest_params &p, const double *freq_hz) {
    if (!freq_hz) fail(CHIP_ERR_INVALID, "NULL frequencies");
    if (p.n_corr == 0 || p.n_chan == 0) fail(CHIP_ERR_INVALID, "no correlations or channels");
    if (p.polarization != CHIP_POL_EXPLICIT && p.polarization != CHIP_POL_PAIR)
        fail(CHIP_ERR_INVALID, "unknown polarization");
    if (p.corr[0] >= p.n_corr || (p.polarization == CHIP_POL_PAIR && p.corr[1] >= p.n_corr))
        fail(CHIP_ERR_INVALID, "correlation index outside the rows");
    if (p.polarization == CHIP_POL_PAIR &&
        (p.transform < CHIP_PAIR_HALF_SUM || p.transform > CHIP_PAIR_NEGATIVE_HALF_IMAG_DIFF))
        fail(CHIP_ERR_INVALID, "unknown pair collapse");
    if (!(std::isfinite(p.sumwt_factor) && p.sumwt_factor > 0.0f))
        fail(CHIP_ERR_INVALID, "sumwt factor must be finite and positive");
    if (p.max_rows == 0 || p.block_rows == 0) fail(CHIP_ERR_INVALID, "no capacity");
    if (p.max_rows > (1ull << 36) || p.block_rows > p.max_rows ||
        p.max_rows * p.n_chan / p.n_chan != p.max_rows || p.n_chan > (1u << 24) || p.n_corr > 64)
        fail(CHIP_ERR_INVALID, "stream too large");
    ing.p = p;
    const uint64_t n = p.max_rows * p.n_chan;
    std::vector<double> ls(p.n_chan);
    for (uint32_t c = 0; c < p.n_chan; ++c) ls[c] = freq_hz[c] / 299792458.0;  // casars-imager lib.rs:557
    ing.freq.alloc(p.n_chan * sizeof(double));
    ing.lscale.alloc(p.n_chan * sizeof(double));
    CHIP_HIP(hipMemcpy(ing.freq.p, freq_hz, p.n_chan * sizeof(double), hipMemcpyHostToDevice));
    CHIP_HIP(hipMemcpy(ing.lscale.p, ls.data(), p.n_chan * sizeof(double), hipMemcpyHostToDevice));
    ing.uvw.alloc(p.max_rows * 3 * sizeof(double));
    ing.vis.alloc(n * sizeof(float2));
    ing.wt.alloc(n * sizeof(float));
    ing.sf.alloc(n * sizeof(float));
    ing.flag.alloc(n);
    const Layout l = layout(p, p.block_rows);
    for (IngestSlot &s : ing.slot) {
        CHIP_HIP(hipHostMalloc(&s.host, l.bytes, hipHostMallocDefault));
        s.dev.alloc(l.bytes + p.block_rows * sizeof(double));  // + the block's phase shifts
        CHIP_HIP(hipEventCreateWithFlags(&s.copied, hipEventDisableTiming));
        CHIP_HIP(hipEventCreateWithFlags(&s.consumed, hipEventDisableTiming));
    }
    CHIP_HIP(hipStreamCreateWithFlags(&ing.copy, hipStreamNonBlocking));
}

void ingest_push(chip_ingest &ing, uint64_t n_rows, const double *uvw, const float *data, const uint8_t *flag,
                 const float *weight, const uint8_t *row_flag) {
    const chip_ingest_params &p = ing.p;
    if (n_rows == 0) return;
    if (!uvw || !data || !flag || !weight) fail(CHIP_ERR_INVALID, "NULL block array");
    if (n_rows > p.block_rows) fail(CHIP_ERR_INVALID, "block larger than block_rows");
    if (ing.rows + n_rows > p.max_rows) fail(CHIP_ERR_INVALID, "the row stream would exceed max_rows");
    IngestSlot &s = ing.slot[ing.next];
    if (s.used) CHIP_HIP(hipEventSynchronize(s.copied));  // the pinned buffer is free again
    const Layout l = layout(p, p.block_rows);
    const uint64_t per = static_cast<uint64_t>(p.n_corr) * p.n_chan;
    uint8_t *h = static_cast<uint8_t *>(s.host);
    std::memcpy(h, uvw, n_rows * 3 * sizeof(double));
    std::memcpy(h + l.data, data, n_rows * per * 2 * sizeof(float));
    std::memcpy(h + l.flag, flag, n_rows * per);
    std::memcpy(h + l.weight, weight, n_rows * (p.weight_spectrum ? per : p.n_corr) * sizeof(float));
    if (row_flag) std::memcpy(h + l.row_flag, row_flag, n_rows);
    hipStream_t st = ing.ctx->stream;
    // the device slot is free once the previous conversion from it is done
    if (s.used) CHIP_HIP(hipStreamWaitEvent(ing.copy, s.consumed, 0));
    CHIP_HIP(hipMemcpyAsync(s.dev.p, s.host, l.bytes, hipMemcpyHostToDevice, ing.copy));
    CHIP_HIP(hipEventRecord(s.copied, ing.copy));
    CHIP_HIP(hipStreamWaitEvent(st, s.copied, 0));
    IngestGeom g{};
    g.n_corr = p.n_corr;
    g.n_chan = p.n_chan;
    g.pol = p.polarization;
    g.transform = p.transform;
    g.c0 = p.corr[0];
    g.c1 = p.corr[1];
    g.sumwt_factor = p.sumwt_factor;
    g.allow_psf_only = p.allow_psf_only;
    g.spectrum = p.weight_spectrum;
    g.rotate = p.rotate;
    for (int k = 0; k < 9; ++k) g.uvrot[k] = p.uvrot[k];
    for (int k = 0; k < 3; ++k) g.phrot[k] = p.phrot[k];
    g.n_rows = n_rows;
    g.row0 = ing.rows;
    g.off_data = l.data;
    g.off_flag = l.flag;
    g.off_weight = l.weight;
    g.off_row_flag = l.row_flag;
    g.has_row_flag = row_flag ? 1 : 0;
    const uint8_t *stage = s.dev.as<uint8_t>();
    double *shift = reinterpret_cast<double *>(s.dev.as<uint8_t>() + l.bytes);
    hipLaunchKernelGGL(k_ingest_rows, dim3(blocks_of(n_rows)), dim3(256), 0, st, g, stage, ing.uvw.as<double>(),
                       shift);
    CHIP_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_ingest_samples, dim3(blocks_of(n_rows * p.n_chan)), dim3(256), 0, st, g, stage, shift,
                       ing.freq.as<double>(), ing.vis.as<float2>(), ing.wt.as<float>(), ing.sf.as<float>(),
                       ing.flag.as<uint8_t>());
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipEventRecord(s.consumed, st));
    s.used = true;
    ing.rows += n_rows;
    ing.next ^= 1;
}

void ingest_wait(chip_ingest &ing) {
    for (IngestSlot &s : ing.slot)
        if (s.used) CHIP_HIP(hipEventSynchronize(s.consumed));
}

void ingest_samples(chip_ingest &ing, chip_samples &out) {
    ingest_wait(ing);
    out = chip_samples{};
    out.n_rows = ing.rows;
    out.n_chan = ing.p.n_chan;
    out.uvw_m_dev = ing.uvw.as<double>();
    out.lambda_scale_dev = ing.lscale.as<double>();
    out.vis_dev = ing.vis.as<float>();
    out.weight_dev = ing.wt.as<float>();
    out.sumwt_factor_dev = ing.sf.as<float>();
    out.flag_dev = ing.flag.as<uint8_t>();
}

void ingest_download(chip_ingest &ing, double *uvw, float *vis, float *weight, float *sumwt_factor, uint8_t *flag) {
    ingest_wait(ing);
    const uint64_t n = ing.rows * ing.p.n_chan;
    if (uvw) CHIP_HIP(hipMemcpy(uvw, ing.uvw.p, ing.rows * 3 * sizeof(double), hipMemcpyDeviceToHost));
    if (vis) CHIP_HIP(hipMemcpy(vis, ing.vis.p, n * sizeof(float2), hipMemcpyDeviceToHost));
    if (weight) CHIP_HIP(hipMemcpy(weight, ing.wt.p, n * sizeof(float), hipMemcpyDeviceToHost));
    if (sumwt_factor) CHIP_HIP(hipMemcpy(sumwt_factor, ing.sf.p, n * sizeof(float), hipMemcpyDeviceToHost));
    if (flag) CHIP_HIP(hipMemcpy(flag, ing.flag.p, n, hipMemcpyDeviceToHost));
}

}  // namespace chip

chip_ingest::~chip_ingest() {
    chip::ingest_destroy_slots(slot, 2);
    if (copy) (void)hipStreamDestroy(copy);
}
