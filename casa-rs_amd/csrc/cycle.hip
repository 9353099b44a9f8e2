// The major-cycle residual refresh of every channel of a visibility set in
// one call (compute_residual_w_project_replay, lib.rs:70895-70963;
// compute_residual_standard_tiled, lib.rs:68862-68932):
//
//   model images -> apodise, centered forward FFT -> degrid, grid (V - p) w
//   -> centered inverse FFT -> crop, correct, / sumwt / psf peak.
//
// Everything stays in a workspace owned by the set.  The centered transforms
// cost no extra passes: the input checkerboard of the forward transform is
// folded into the placement of the model, its output checkerboard into the
// binned kernel's model-tile load, and the inverse transform's shifts into
// the crop (shift theorem: IFFT((-1)^k G)[x] = IFFT(G)[x + N/2], so
// centered_ifft(G)[x] = (-1)^(x + N/2) IFFT(G)[(x + N/2) mod N] / N).
//
// The FFT strategy is chosen by timing on the first call: either rocFFT 2-D
// plans over the public [chan][x][y] layout, or two 1-D passes over an
// interleaved [x][chan][y] layout (a contiguous y pass, and an x pass whose
// batch runs over the contiguous (chan, y) plane), which needs no transposes.
#include <cmath>

#include "device.hpp"

namespace chip {

void vset_residual_with(chip_vset &vs, const GridParams &gp, const float2 *model, float *res);

struct CycleGeom {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t n_chan;
    uint32_t base;  // gx/2 + gy/2
    uint32_t wproject;
    uint64_t cstride, xstride;
    const float *corr_x, *corr_y;
    const float *sinc;            // W products: sinc of max(gx, gy) on both axes
    const float *sinc_x, *sinc_y; // W model apodisation: per-axis sinc
};

// Model (chan, x, y) -> grid cell (blc + x, blc + y), times the grid
// correction [x sinc] and (-1)^(gx + gy).  Launch grid: (ceil(ny/256), nx, chan).
__global__ void k_cycle_place(CycleGeom cg, const float *__restrict__ model, float2 *grid) {
    const uint32_t y = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t x = blockIdx.y;
    const uint32_t c = blockIdx.z;
    if (y >= cg.ny) return;
    const uint32_t gxx = cg.blc_x + x, gyy = cg.blc_y + y;
    float f = cg.corr_x[gxx] * cg.corr_y[gyy];
    if (cg.wproject) f = f * cg.sinc_x[gxx] * cg.sinc_y[gyy];
    float v = model[(static_cast<uint64_t>(c) * cg.nx + x) * cg.ny + y] * f;
    if ((gxx + gyy) & 1) v = -v;
    grid[c * cg.cstride + static_cast<uint64_t>(gxx) * cg.xstride + gyy] = make_float2(v, 0.0f);
}

// Inverse-transformed residual grid -> residual image (gridder.rs:1646-1657
// standard, :1685-1715 W; normalisation lib.rs:68920-68932, :70960-70962).
__global__ void k_cycle_crop(CycleGeom cg, const float2 *__restrict__ raw, const float *__restrict__ nsum,
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
    float re = raw[c * cg.cstride + static_cast<uint64_t>(sx) * cg.xstride + sy].x * scale;
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

static hipfftHandle make_plan_1d(int n, int stride, int dist, int batch, hipStream_t st) {
    hipfftHandle h;
    CHIP_FFT(hipfftCreate(&h));
    int nn[1] = {n};
    int emb[1] = {n};
    size_t work = 0;
    CHIP_FFT(hipfftMakePlanMany(h, 1, nn, emb, stride, dist, emb, stride, dist, HIPFFT_C2C, batch, &work));
    CHIP_FFT(hipfftSetStream(h, st));
    return h;
}

struct CycleWorkspace {
    bool ready = false;
    bool interleaved = false;
    DevBuf model_grid, res_grid;
    hipfftHandle p2d = nullptr;    // [chan][x][y]: 2-D, batch n_chan
    hipfftHandle py = nullptr;     // [x][chan][y]: y pass
    hipfftHandle px = nullptr;     // [x][chan][y]: x pass
    double tuned_ms_2d = -1.0, tuned_ms_1d = -1.0;
    ~CycleWorkspace() {
        if (p2d) hipfftDestroy(p2d);
        if (py) hipfftDestroy(py);
        if (px) hipfftDestroy(px);
    }
};

static void fft_exec(CycleWorkspace &ws, float2 *data, int dir) {
    hipfftComplex *d = reinterpret_cast<hipfftComplex *>(data);
    if (!ws.interleaved) {
        CHIP_FFT(hipfftExecC2C(ws.p2d, d, d, dir));
    } else {
        CHIP_FFT(hipfftExecC2C(ws.py, d, d, dir));
        CHIP_FFT(hipfftExecC2C(ws.px, d, d, dir));
    }
}

static double time_fft(CycleWorkspace &ws, float2 *data, hipStream_t st) {
    hipEvent_t a, b;
    CHIP_HIP(hipEventCreate(&a));
    CHIP_HIP(hipEventCreate(&b));
    fft_exec(ws, data, HIPFFT_FORWARD);  // warm-up (kernel compilation)
    CHIP_HIP(hipEventRecord(a, st));
    fft_exec(ws, data, HIPFFT_FORWARD);
    fft_exec(ws, data, HIPFFT_BACKWARD);
    CHIP_HIP(hipEventRecord(b, st));
    CHIP_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    CHIP_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return ms;
}

static void cycle_setup(chip_vset &vs, CycleWorkspace &ws) {
    chip_gridder &g = *vs.g;
    hipStream_t st = g.ctx->stream;
    const uint64_t cells = vs.n_chan * g.gx * g.gy;
    if (vs.n_chan * g.gy > 0x7fffffffull || cells > 0x7fffffffull * 2)
        fail(CHIP_ERR_INVALID, "cube too large for one major-cycle workspace");
    ws.model_grid.alloc(cells * sizeof(float2));
    ws.res_grid.alloc(cells * sizeof(float2));
    CHIP_HIP(hipMemsetAsync(ws.model_grid.p, 0, ws.model_grid.bytes, st));
    // 2-D plans over [chan][x][y]
    {
        CHIP_FFT(hipfftCreate(&ws.p2d));
        int n[2] = {static_cast<int>(g.gx), static_cast<int>(g.gy)};
        size_t work = 0;
        CHIP_FFT(hipfftMakePlanMany(ws.p2d, 2, n, nullptr, 1, static_cast<int>(g.gx * g.gy), nullptr, 1,
                                    static_cast<int>(g.gx * g.gy), HIPFFT_C2C, static_cast<int>(vs.n_chan),
                                    &work));
        CHIP_FFT(hipfftSetStream(ws.p2d, st));
    }
    // 1-D passes over [x][chan][y]
    const int gx = static_cast<int>(g.gx), gy = static_cast<int>(g.gy), nc = static_cast<int>(vs.n_chan);
    ws.py = make_plan_1d(gy, 1, gy, gx * nc, st);
    ws.px = make_plan_1d(gx, nc * gy, 1, nc * gy, st);
    float2 *d = ws.model_grid.as<float2>();
    ws.interleaved = false;
    ws.tuned_ms_2d = time_fft(ws, d, st);
    ws.interleaved = true;
    ws.tuned_ms_1d = time_fft(ws, d, st);
    if (vs.cycle_layout < 0) ws.interleaved = ws.tuned_ms_1d < ws.tuned_ms_2d;
    else ws.interleaved = vs.cycle_layout == 1;
    if (ws.interleaved) { hipfftDestroy(ws.p2d); ws.p2d = nullptr; }
    else { hipfftDestroy(ws.py); hipfftDestroy(ws.px); ws.py = ws.px = nullptr; }
    ws.ready = true;
}

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
    hipStream_t st = g.ctx->stream;
    if (!vs.cycle) vs.cycle = std::make_shared<CycleWorkspace>();
    CycleWorkspace &ws = *vs.cycle;
    if (!ws.ready) cycle_setup(vs, ws);

    const uint64_t nc = vs.n_chan;
    CycleGeom cg{};
    cg.nx = static_cast<uint32_t>(g.geom.nx);
    cg.ny = static_cast<uint32_t>(g.geom.ny);
    cg.gx = static_cast<uint32_t>(g.gx);
    cg.gy = static_cast<uint32_t>(g.gy);
    cg.blc_x = static_cast<uint32_t>(g.blc_x);
    cg.blc_y = static_cast<uint32_t>(g.blc_y);
    cg.n_chan = static_cast<uint32_t>(nc);
    cg.base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
    cg.wproject = g.wproject ? 1u : 0u;
    cg.cstride = ws.interleaved ? g.gy : g.gx * g.gy;
    cg.xstride = ws.interleaved ? nc * g.gy : g.gy;
    cg.corr_x = g.d_corr_x.as<float>();
    cg.corr_y = g.d_corr_y.as<float>();
    cg.sinc = g.wproject ? g.d_sinc.as<float>() : nullptr;
    cg.sinc_x = g.wproject ? g.d_sinc_x.as<float>() : nullptr;
    cg.sinc_y = g.wproject ? g.d_sinc_y.as<float>() : nullptr;

    std::vector<float> host(2 * nc);
    for (uint64_t c = 0; c < nc; ++c) {
        host[c] = static_cast<float>(nsum[c]);
        host[nc + c] = peaks[c];
    }
    if (!vs.cycle_norm.p) vs.cycle_norm.alloc(2 * nc * sizeof(float));
    CHIP_HIP(hipMemcpyAsync(vs.cycle_norm.p, host.data(), 2 * nc * sizeof(float), hipMemcpyHostToDevice, st));

    float2 *M = ws.model_grid.as<float2>();
    float2 *Rg = ws.res_grid.as<float2>();
    CHIP_HIP(hipMemsetAsync(M, 0, ws.model_grid.bytes, st));
    CHIP_HIP(hipMemsetAsync(Rg, 0, ws.res_grid.bytes, st));
    const dim3 img_grid((cg.ny + 255) / 256, cg.nx, static_cast<unsigned>(nc));
    hipLaunchKernelGGL(k_cycle_place, img_grid, dim3(256), 0, st, cg, model, M);
    CHIP_HIP(hipGetLastError());
    fft_exec(ws, M, HIPFFT_FORWARD);

    GridParams gp = grid_params(g, vs.tile, vs.halo, vs.tiles_y);
    gp.cstride = cg.cstride;
    gp.xstride = cg.xstride;
    gp.model_signed = 1;
    vset_residual_with(vs, gp, M, reinterpret_cast<float *>(Rg));

    fft_exec(ws, Rg, HIPFFT_BACKWARD);
    hipLaunchKernelGGL(k_cycle_crop, img_grid, dim3(256), 0, st, cg, Rg, vs.cycle_norm.as<float>(),
                       vs.cycle_norm.as<float>() + nc, out);
    CHIP_HIP(hipGetLastError());
    CHIP_HIP(hipStreamSynchronize(st));
}

void vset_cycle_layout(chip_vset &vs, int layout) {
    if (layout < -1 || layout > 1) fail(CHIP_ERR_INVALID, "layout must be -1, 0 or 1");
    vs.cycle_layout = layout;
    vs.cycle.reset();  // the next call builds the workspace for the new layout
}

void vset_cycle_info(const chip_vset &vs, int *interleaved, double *ms_2d, double *ms_1d) {
    if (!vs.cycle || !vs.cycle->ready) {
        *interleaved = -1;
        *ms_2d = *ms_1d = -1.0;
        return;
    }
    *interleaved = vs.cycle->interleaved ? 1 : 0;
    *ms_2d = vs.cycle->tuned_ms_2d;
    *ms_1d = vs.cycle->tuned_ms_1d;
}

}  // namespace chip
