// MODEL_DATA prediction of the standard MFS model predictor
// (StandardMfsModelPredictor::predict, lib.rs:3387-3403): per sample, the 7x7
// spheroid taps before normalisation (sample_taps_unnormalized,
// gridder.rs:1997-2023) gathered from the model grid and divided by their sum
// (degrid_sample_product_planned_sectdgrid, gridder.rs:1227-1266).
//
// One thread per sample: the FP64 coordinate and offset steps of the planner
// (bit-exact anchors and tap rows), then 49 gathered complex64 reads -- seven
// runs of 7 contiguous cells, one per x tap -- accumulated in the reference's
// order (x taps outer, y taps inner) with separately rounded products and
// sums, so that the prediction is bit-identical to the CPU restatement for
// the same model grid.  Degridding is a gather: no atomics, no LDS.
#include "device.hpp"

namespace chip {
namespace {

struct PredictParams {
    uint64_t gx, gy;
    double du, dv, half_gx, half_gy;
    uint32_t os;
    uint32_t per_chan;  // one model grid per channel (else one for all)
};

__global__ __launch_bounds__(256) void k_predict(PredictParams p, Source s, const float *raw,
                                                 const float2 *grids, float2 *out) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const uint64_t n = s.n_rows * s.n_chan;
    if (i >= n) return;
    const uint64_t row = s.n_chan == 1 ? i : i / s.n_chan;
    const uint32_t chan = s.n_chan == 1 ? 0u : static_cast<uint32_t>(i - row * s.n_chan);
    double u, v;
    if (s.uvw_m) {
        const double scale = s.lambda_scale[chan];
        u = s.uvw_m[3 * row] * scale;
        v = s.uvw_m[3 * row + 1] * scale;
    } else {
        u = s.u[i];
        v = s.v[i];
    }
    float2 r = make_float2(0.0f, 0.0f);
    // grid_coordinate_x / _y (gridder.rs:1877-1886)
    const double cx = u / p.du + p.half_gx;
    const double cy = -v / p.dv + p.half_gy;
    int32_t xs, xw, ys, yw;
    if (std_axis_span(cx, p.gx, p.os, xs, xw) && std_axis_span(cy, p.gy, p.os, ys, yw)) {
        const float *wx = raw + kTaps * xw;
        const float *wyp = raw + kTaps * yw;
        float wy[kTaps];
#pragma unroll
        for (int b = 0; b < kTaps; ++b) wy[b] = wyp[b];
        const float2 *G = grids + (p.per_chan ? static_cast<uint64_t>(chan) : 0ull) * p.gx * p.gy +
                          static_cast<uint64_t>(xs) * p.gy + static_cast<uint64_t>(ys);
        float vr = 0.0f, vi = 0.0f, norm = 0.0f;
#pragma unroll
        for (int a = 0; a < kTaps; ++a) {
            const float xa = wx[a];
            const float2 *line = G + static_cast<uint64_t>(a) * p.gy;
#pragma unroll
            for (int b = 0; b < kTaps; ++b) {
                const float weight = xa * wy[b];
                const float2 gv = line[b];
                vr += gv.x * weight;
                vi += gv.y * weight;
                norm += weight;
            }
        }
        if (norm > 0.0f && isfinite(norm)) {
            r.x = vr / norm;
            r.y = vi / norm;
        }
    }
    out[i] = r;
}

}  // namespace

void predict_model(chip_gridder &g, const chip_samples &cs, const float *grids, uint64_t n_grids, float *out) {
    if (g.wproject || g.aw)
        fail(CHIP_ERR_INVALID,
             "the model predictor degrids with the standard spheroidal taps; this gridder has W or AW kernels");
    if (cs.n_chan == 0) fail(CHIP_ERR_INVALID, "n_chan must be at least 1");
    const bool per_sample = cs.u_dev && cs.v_dev;
    const bool per_row = cs.uvw_m_dev && cs.lambda_scale_dev;
    if (per_sample == per_row) fail(CHIP_ERR_INVALID, "give either u/v per sample or uvw_m with lambda_scale");
    if (per_sample && cs.n_chan != 1) fail(CHIP_ERR_INVALID, "per-sample u/v input requires n_chan == 1");
    if (!grids || !out) fail(CHIP_ERR_INVALID, "NULL model grid or output");
    if (n_grids != 1 && n_grids != cs.n_chan) fail(CHIP_ERR_INVALID, "model grids must number 1 or n_chan");
    const uint64_t n = cs.n_rows * cs.n_chan;
    if (n == 0) return;
    if (n / cs.n_chan != cs.n_rows || n >= (1ull << 38)) fail(CHIP_ERR_INVALID, "too many samples");
    PredictParams p{};
    p.gx = g.gx;
    p.gy = g.gy;
    p.du = g.du;
    p.dv = g.dv;
    p.half_gx = static_cast<double>(g.gx) / 2.0;
    p.half_gy = static_cast<double>(g.gy) / 2.0;
    p.os = g.os;
    p.per_chan = n_grids > 1 ? 1u : 0u;
    Source src{};
    src.n_rows = cs.n_rows;
    src.n_chan = cs.n_chan;
    src.u = cs.u_dev;
    src.v = cs.v_dev;
    src.uvw_m = cs.uvw_m_dev;
    src.lambda_scale = cs.lambda_scale_dev;
    const uint64_t blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_predict, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, g.ctx->stream, p, src,
                       g.d_taps_raw.as<float>(), reinterpret_cast<const float2 *>(grids),
                       reinterpret_cast<float2 *>(out));
    CHIP_HIP(hipGetLastError());
}

}  // namespace chip
