// Uniform / Briggs density weighting on the GPU (weighting.rs:1180-1232,
// :1747-1856, :2501-2592; density cells gridder.rs:1888-1934).
//
//   1. every usable sample with a positive finite weight adds its weight to
//      its density cell and to the cell of (-u, -v);
//   2. per density plane: sum and sum of squares of the positive cells ->
//      f2 = (5 * 10^-robust)^2 / (sumsq / sum) (host, FP64);
//   3. every sample is reweighted from its own cell: w / d (uniform) or
//      w / (d * f2 + 1) (Briggs, separate Float operations).
//
// Density cells are integer work and bit-exact with the reference.  The
// density values are sums of weights: they are accumulated with FP64 atomics
// (order-independent for any realistic count) and rounded to Float once,
// where the reference adds Floats in sample order; the two agree exactly when
// the weights are small integers or halves, and to a few ulp otherwise.
#include <cmath>

#include "device.hpp"

namespace chip {

namespace {

struct DensityParams {
    uint64_t nx, ny;
    double center_x, center_y;  // n / 2
    double u_scale, v_scale;    // n * cell
    uint32_t convention;        // 0: VisImagingWeight (Float, truncate), 1: cube (Double, round)
    uint32_t per_plane;
};

__device__ __forceinline__ bool density_cell(const DensityParams &d, double u, double v, int64_t &ax,
                                             int64_t &ay) {
#pragma clang fp contract(off)
    double x, y;
    if (d.convention == 0) {
        const float uf = static_cast<float>(u), vf = static_cast<float>(v);
        const float xs = (-uf) * static_cast<float>(d.u_scale);
        const float ys = vf * static_cast<float>(d.v_scale);
        const float xf = xs + static_cast<float>(d.center_x);
        const float yf = ys + static_cast<float>(d.center_y);
        x = static_cast<double>(xf);
        y = static_cast<double>(yf);
    } else {
        x = u * d.u_scale + d.center_x;
        y = -v * d.v_scale + d.center_y;
    }
    if (!(isfinite(x) && isfinite(y))) return false;
    if (d.convention != 0) {
        x = round(x);
        y = round(y);
    }
    if (x >= 9.0e18 || x <= -9.0e18 || y >= 9.0e18 || y <= -9.0e18) return false;
    ax = static_cast<int64_t>(x);
    ay = static_cast<int64_t>(y);
    return ax > 0 && ay > 0 && ax < static_cast<int64_t>(d.nx) && ay < static_cast<int64_t>(d.ny);
}

__global__ __launch_bounds__(256) void k_density(DensityParams d, Source s, double *density) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= s.n_rows * s.n_chan) return;
    const SampleIn smp = load_sample(s, i);
    if (!smp.usable || !(finite_f(smp.weight) && smp.weight > 0.0f)) return;
    double *plane = density + (d.per_plane ? static_cast<uint64_t>(smp.chan) * d.nx * d.ny : 0);
    const double w = static_cast<double>(smp.weight);
    int64_t x, y;
    if (density_cell(d, smp.u, smp.v, x, y)) unsafeAtomicAdd(plane + x * d.ny + y, w);
    if (density_cell(d, -smp.u, -smp.v, x, y)) unsafeAtomicAdd(plane + x * d.ny + y, w);
}

// Round the FP64 sums to Float and reduce sum / sum of squares per plane.
__global__ __launch_bounds__(256) void k_density_finish(const double *density64, uint64_t cells,
                                                        float *density32, double *sums) {
    const uint32_t plane = blockIdx.y;
    double s = 0.0, q = 0.0;
    for (uint64_t c = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; c < cells;
         c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
        const float v = static_cast<float>(density64[plane * cells + c]);
        density32[plane * cells + c] = v;
        s += static_cast<double>(v);
        if (v > 0.0f) q += static_cast<double>(v) * static_cast<double>(v);
    }
    s = wave_sum(s);
    q = wave_sum(q);
    if ((threadIdx.x & 63) == 0) {
        unsafeAtomicAdd(sums + 2 * plane, s);
        unsafeAtomicAdd(sums + 2 * plane + 1, q);
    }
}

__global__ __launch_bounds__(256) void k_reweight(DensityParams d, Source s, uint32_t mode,
                                                  const float *density, const float *f2, float *out) {
#pragma clang fp contract(off)
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= s.n_rows * s.n_chan) return;
    const SampleIn smp = load_sample(s, i);
    const float w = smp.weight;
    if (!smp.usable) {
        out[i] = w;
        return;
    }
    const uint32_t p = d.per_plane ? smp.chan : 0u;
    int64_t x, y;
    if (!density_cell(d, smp.u, smp.v, x, y)) {
        out[i] = 0.0f;
        return;
    }
    const float dv = density[static_cast<uint64_t>(p) * d.nx * d.ny + x * d.ny + y];
    if (!(finite_f(w) && w > 0.0f && finite_f(dv) && dv > 0.0f)) {
        out[i] = 0.0f;
        return;
    }
    if (mode == CHIP_WEIGHT_UNIFORM) {
        out[i] = w / dv;
    } else {
        const float product = dv * f2[p];
        const float denominator = product + 1.0f;
        out[i] = w / denominator;
    }
}

unsigned blocks_of(uint64_t n, unsigned t) { return static_cast<unsigned>((n + t - 1) / t); }

}  // namespace

void density_weights(chip_gridder &g, const chip_samples &cs, uint32_t mode, float robust, uint32_t per_plane,
                     float *weight_out, float *f2_host, float *density_out) {
    if (mode != CHIP_WEIGHT_UNIFORM && mode != CHIP_WEIGHT_BRIGGS)
        fail(CHIP_ERR_INVALID, "weighting mode must be uniform or Briggs");
    if (!weight_out) fail(CHIP_ERR_INVALID, "weight output is required");
    if (!std::isfinite(robust)) fail(CHIP_ERR_INVALID, "robust must be finite");
    const Source src = make_source(cs);
    hipStream_t st = g.ctx->stream;
    const uint64_t n = src.n_rows * src.n_chan;
    DensityParams d{};
    d.nx = g.geom.nx;
    d.ny = g.geom.ny;
    d.center_x = static_cast<double>(g.geom.nx) / 2.0;
    d.center_y = static_cast<double>(g.geom.ny) / 2.0;
    d.u_scale = static_cast<double>(g.geom.nx) * g.geom.cell_x;
    d.v_scale = static_cast<double>(g.geom.ny) * g.geom.cell_y;
    d.convention = per_plane ? 1u : 0u;
    d.per_plane = per_plane ? 1u : 0u;
    const uint64_t planes = per_plane ? src.n_chan : 1;
    const uint64_t cells = d.nx * d.ny;
    if (planes > 65535) fail(CHIP_ERR_INVALID, "too many density planes");

    DevBuf d64(planes * cells * sizeof(double)), sums(2 * planes * sizeof(double));
    DevBuf own32;
    float *d32 = density_out;
    if (!d32) {
        own32.alloc(planes * cells * sizeof(float));
        d32 = own32.as<float>();
    }
    CHIP_HIP(hipMemsetAsync(d64.p, 0, d64.bytes, st));
    CHIP_HIP(hipMemsetAsync(sums.p, 0, sums.bytes, st));
    if (n) {
        hipLaunchKernelGGL(k_density, dim3(blocks_of(n, 256)), dim3(256), 0, st, d, src, d64.as<double>());
        CHIP_HIP(hipGetLastError());
    }
    const unsigned bx = static_cast<unsigned>(std::min<uint64_t>((cells + 255) / 256, 1024));
    hipLaunchKernelGGL(k_density_finish, dim3(bx, static_cast<unsigned>(planes)), dim3(256), 0, st,
                       d64.as<double>(), cells, d32, sums.as<double>());
    CHIP_HIP(hipGetLastError());
    std::vector<double> hs(2 * planes);
    CHIP_HIP(hipMemcpyAsync(hs.data(), sums.p, sums.bytes, hipMemcpyDeviceToHost, st));
    CHIP_HIP(hipStreamSynchronize(st));
    // f2 (weighting.rs:1211-1224): the density includes the conjugate samples
    std::vector<float> f2(planes, 1.0f);
    if (mode == CHIP_WEIGHT_BRIGGS) {
        const double scale = 5.0 * std::pow(10.0, -static_cast<double>(robust));
        for (uint64_t p = 0; p < planes; ++p) {
            const double sum = hs[2 * p], sumsq = hs[2 * p + 1];
            f2[p] = (sumsq > 0.0 && sum > 0.0) ? static_cast<float>(scale * scale / (sumsq / sum)) : 0.0f;
        }
    }
    DevBuf df2(planes * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(df2.p, f2.data(), df2.bytes, hipMemcpyHostToDevice, st));
    if (n) {
        hipLaunchKernelGGL(k_reweight, dim3(blocks_of(n, 256)), dim3(256), 0, st, d, src, mode, d32,
                           df2.as<float>(), weight_out);
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
    if (f2_host) std::copy(f2.begin(), f2.end(), f2_host);
}

}  // namespace chip
