// Phase-centre rotation of a row stream on the GPU, in FP64
// (casars-imager/src/lib.rs:47064-47069: uvw' = uvw * uvrot as a row vector,
// phase shift = phrot . uvw'; casa-imaging/src/lib.rs:248-259: per channel
// phase = -2 pi shift f / c, phasor (cos, sin) narrowed to Float, complex
// Float multiply).  The matrices come from chip_phase_rotation (host FP64).
#include <cmath>

#include "device.hpp"

namespace chip {

namespace {

struct RotParams {
    double uvrot[9];
    double phrot[3];
    uint64_t n_rows, n_chan;
};

// One thread per row; uvw_out may alias uvw.
__global__ __launch_bounds__(256) void k_rotate_uvw(RotParams p, const double *uvw, double *uvw_out,
                                                    double *shift) {
    const uint64_t r = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= p.n_rows) return;
    const double x[3] = {uvw[3 * r], uvw[3 * r + 1], uvw[3 * r + 2]};
    double o[3], s;
    rotate_uvw_row(p.uvrot, p.phrot, x, o, s);
    shift[r] = s;
    if (uvw_out) {
        uvw_out[3 * r] = o[0];
        uvw_out[3 * r + 1] = o[1];
        uvw_out[3 * r + 2] = o[2];
    }
}

// One thread per sample (row, chan), channel fastest.
__global__ __launch_bounds__(256) void k_rotate_vis(RotParams p, const double *shift, const double *freq,
                                                    float2 *vis) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= p.n_rows * p.n_chan) return;
    const uint64_t r = i / p.n_chan, c = i - r * p.n_chan;
    vis[i] = phase_rotate(vis[i], shift[r], freq[c]);
}

unsigned blocks_of(uint64_t n) { return static_cast<unsigned>((n + 255) / 256); }

}  // namespace

void rotate_rows(chip_ctx &ctx, uint64_t n_rows, uint64_t n_chan, const double uvrot[9], const double phrot[3],
                 const double *uvw, const double *freq, float *vis, double *uvw_out, double *shift_out) {
    if (!uvw || !uvrot || !phrot) fail(CHIP_ERR_INVALID, "uvw and the rotation are required");
    if (vis && !freq) fail(CHIP_ERR_INVALID, "rotating visibilities needs the channel frequencies");
    if (n_rows == 0) return;
    if (n_chan == 0 && vis) fail(CHIP_ERR_INVALID, "no channels");
    if (n_rows > (1ull << 40) || (vis && n_rows * n_chan / n_chan != n_rows))
        fail(CHIP_ERR_INVALID, "row stream too large");
    RotParams p{};
    for (int k = 0; k < 9; ++k) p.uvrot[k] = uvrot[k];
    for (int k = 0; k < 3; ++k) p.phrot[k] = phrot[k];
    p.n_rows = n_rows;
    p.n_chan = n_chan;
    hipStream_t st = ctx.stream;
    DevBuf own;
    double *shift = shift_out;
    if (!shift) {
        own.alloc(n_rows * sizeof(double));
        shift = own.as<double>();
    }
    hipLaunchKernelGGL(k_rotate_uvw, dim3(blocks_of(n_rows)), dim3(256), 0, st, p, uvw, uvw_out, shift);
    CHIP_HIP(hipGetLastError());
    if (vis) {
        hipLaunchKernelGGL(k_rotate_vis, dim3(blocks_of(n_rows * n_chan)), dim3(256), 0, st, p, shift, freq,
                           reinterpret_cast<float2 *>(vis));
        CHIP_HIP(hipGetLastError());
    }
    CHIP_HIP(hipStreamSynchronize(st));
}

}  // namespace chip
