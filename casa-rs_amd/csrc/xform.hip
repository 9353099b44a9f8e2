// Image <-> uv-grid transforms of the major cycle with the pre- and
// post-processing fused into the FFT passes.
//
// Forward (centered_fft2(apodize_model(model)), gridder.rs:1554-1592,
// fft.rs:34): the model is real, so the 2-D transform is Hermitian,
// B[-kx, -ky] = conj(B[kx, ky]).
//   rows pass   : for each image row x, the apodised, checkerboard-signed
//                 model row zero-padded to gy points -> FFT along y; only
//                 ky in [0, gy/2] is kept (half spectrum, nx rows).
//   column pass : for each block of C columns ky <= gy/2, rows outside the
//                 image are zero -> FFT along x; writes column ky and its
//                 Hermitian mirror -ky, so the full grid is produced from
//                 half of the columns.
// Inverse (centered inverse FFT + crop + correction, gridder.rs:1646-1715):
//   column pass : for each block of C columns, IFFT along x; keeps only the
//                 nx rows the crop needs (shift theorem: output row
//                 (blc_x + x + gx/2) mod gx).
//   rows pass   : IFFT along y of those rows; the crop, 1/(gx gy) scale,
//                 sign (-1)^(x + y + gx/2 + gy/2), grid correction and
//                 normalisation are applied on the way out.
//
// Each 1-D transform is a mixed-radix (2, 3, 4, 5) Stockham autosort FFT
// held in LDS (one row, or a tile of gx x C complex values for the column
// passes so that global accesses are C*8-byte row segments).  Every stage
// reads its butterfly inputs into registers, synchronises, and writes the
// outputs back, so one LDS buffer suffices.  Stockham stage with radix r
// after sub-transforms of length ns: butterfly j reads elements j + t*n/r,
// multiplies input t by exp(-+2 pi i k t / (ns r)) with k = j mod ns, and
// writes output t to (j - k) r + k + t ns.  Twiddles are read from a table
// exp(-2 pi i m / n) rounded once from FP64.
#include <cmath>

#include "device.hpp"
#include "xform.hpp"

namespace chip {
namespace xf {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// x (-i) for the forward transform, x (+i) for the inverse.
template <bool INV> __device__ __forceinline__ float2 rot(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// In-place DFT of R points (forward: exp(-2 pi i jk / R); inverse: +).
template <int R, bool INV> __device__ __forceinline__ void butterfly(float2 *a);

template <> __device__ __forceinline__ void butterfly<2, false>(float2 *a) {
    const float2 t = a[0];
    a[0] = cadd(t, a[1]);
    a[1] = csub(t, a[1]);
}
template <> __device__ __forceinline__ void butterfly<2, true>(float2 *a) { butterfly<2, false>(a); }

template <bool INV> __device__ __forceinline__ void bf3(float2 *a) {
    constexpr float s = 0.86602540378443864676f;  // sin(2 pi / 3)
    const float2 b = cadd(a[1], a[2]);
    const float2 d = csub(a[1], a[2]);
    const float2 t = csub(a[0], cscale(b, 0.5f));
    const float2 u = rot<INV>(cscale(d, s));
    a[0] = cadd(a[0], b);
    a[1] = cadd(t, u);
    a[2] = csub(t, u);
}
template <> __device__ __forceinline__ void butterfly<3, false>(float2 *a) { bf3<false>(a); }
template <> __device__ __forceinline__ void butterfly<3, true>(float2 *a) { bf3<true>(a); }

template <bool INV> __device__ __forceinline__ void bf4(float2 *a) {
    const float2 t0 = cadd(a[0], a[2]);
    const float2 t1 = csub(a[0], a[2]);
    const float2 t2 = cadd(a[1], a[3]);
    const float2 t3 = rot<INV>(csub(a[1], a[3]));
    a[0] = cadd(t0, t2);
    a[2] = csub(t0, t2);
    a[1] = cadd(t1, t3);
    a[3] = csub(t1, t3);
}
template <> __device__ __forceinline__ void butterfly<4, false>(float2 *a) { bf4<false>(a); }
template <> __device__ __forceinline__ void butterfly<4, true>(float2 *a) { bf4<true>(a); }

template <bool INV> __device__ __forceinline__ void bf5(float2 *a) {
    constexpr float c1 = 0.30901699437494742410f;   // cos(2 pi / 5)
    constexpr float c2 = -0.80901699437494742410f;  // cos(4 pi / 5)
    constexpr float s1 = 0.95105651629515357212f;   // sin(2 pi / 5)
    constexpr float s2 = 0.58778525229247312917f;   // sin(4 pi / 5)
    const float2 b1 = cadd(a[1], a[4]), b2 = cadd(a[2], a[3]);
    const float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
    const float2 p1 = make_float2(a[0].x + c1 * b1.x + c2 * b2.x, a[0].y + c1 * b1.y + c2 * b2.y);
    const float2 p2 = make_float2(a[0].x + c2 * b1.x + c1 * b2.x, a[0].y + c2 * b1.y + c1 * b2.y);
    const float2 q1 = rot<INV>(make_float2(s1 * d1.x + s2 * d2.x, s1 * d1.y + s2 * d2.y));
    const float2 q2 = rot<INV>(make_float2(s2 * d1.x - s1 * d2.x, s2 * d1.y - s1 * d2.y));
    a[0] = cadd(a[0], cadd(b1, b2));
    a[1] = cadd(p1, q1);
    a[4] = csub(p1, q1);
    a[2] = cadd(p2, q2);
    a[3] = csub(p2, q2);
}
template <> __device__ __forceinline__ void butterfly<5, false>(float2 *a) { bf5<false>(a); }
template <> __device__ __forceinline__ void butterfly<5, true>(float2 *a) { bf5<true>(a); }

// One Stockham stage over a tile of C = 2^logc interleaved transforms:
// element i of transform c is buf[(i << logc) + c].  At most E elements per
// thread (n * C <= E * T).
template <int R, bool INV, int T, int E>
__device__ __forceinline__ void stage(float2 *buf, uint32_t n, uint32_t logc, uint32_t ns,
                                      const float2 *__restrict__ tw) {
    constexpr int MB = (E + R - 1) / R;
    const uint32_t nb = n / R;
    const uint32_t total = nb << logc;
    const uint32_t cmask = (1u << logc) - 1u;
    const uint32_t step = n / (ns * R);
    float2 v[MB][R];
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const uint32_t bi = threadIdx.x + b * T;
        if (bi < total) {
            const uint32_t c = bi & cmask, j = bi >> logc;
#pragma unroll
            for (int t = 0; t < R; ++t) v[b][t] = buf[((j + t * nb) << logc) + c];
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const uint32_t bi = threadIdx.x + b * T;
        if (bi < total) {
            const uint32_t c = bi & cmask, j = bi >> logc;
            const uint32_t k = j % ns;
            if (k != 0) {
#pragma unroll
                for (int t = 1; t < R; ++t) {
                    float2 w = tw[k * t * step];
                    if (INV) w.y = -w.y;
                    v[b][t] = cmul(v[b][t], w);
                }
            }
            butterfly<R, INV>(v[b]);
            const uint32_t base = (j - k) * R + k;
#pragma unroll
            for (int t = 0; t < R; ++t) buf[((base + t * ns) << logc) + c] = v[b][t];
        }
    }
    __syncthreads();
}

template <bool INV, int T, int E>
__device__ __forceinline__ void run_fft(float2 *buf, const Axis &ax, uint32_t logc) {
    uint32_t ns = 1;
    for (uint32_t s = 0; s < ax.n_stages; ++s) {
        const uint32_t r = ax.radix[s];
        if (r == 5) stage<5, INV, T, E>(buf, ax.n, logc, ns, ax.tw);
        else if (r == 4) stage<4, INV, T, E>(buf, ax.n, logc, ns, ax.tw);
        else if (r == 3) stage<3, INV, T, E>(buf, ax.n, logc, ns, ax.tw);
        else stage<2, INV, T, E>(buf, ax.n, logc, ns, ax.tw);
        ns *= r;
    }
}

// ---------------------------------------------------------------- kernels
// Launch grids: rows passes (nx, n_chan); column passes (column blocks, n_chan).

template <int T, int E>
__global__ __launch_bounds__(T) void k_fwd_rows(XGeom xg, Axis ay, const float *__restrict__ model,
                                                float2 *__restrict__ half) {
    extern __shared__ float2 buf[];
    const uint32_t x = blockIdx.x, c = blockIdx.y;
    const uint32_t gxx = xg.blc_x + x;
    const float *row = model + (static_cast<uint64_t>(c) * xg.nx + x) * xg.ny;
    const float cx = xg.corr_x[gxx];
    const float sx = xg.wproject ? xg.sinc_x[gxx] : 1.0f;
    for (uint32_t i = threadIdx.x; i < xg.gy; i += T) {
        float v = 0.0f;
        if (i >= xg.blc_y && i < xg.blc_y + xg.ny) {
            // factor evaluated left to right as the reference: cx*cy [*sx*sy]
            float f = cx * xg.corr_y[i];
            if (xg.wproject) f = f * sx * xg.sinc_y[i];
            v = row[i - xg.blc_y] * f;
            if ((gxx + i) & 1) v = -v;
        }
        buf[i] = make_float2(v, 0.0f);
    }
    __syncthreads();
    run_fft<false, T, E>(buf, ay, 0);
    float2 *out = half + (static_cast<uint64_t>(c) * xg.nx + x) * xg.half_pitch;
    for (uint32_t k = threadIdx.x; k <= xg.gy / 2; k += T) out[k] = buf[k];
}

template <int T, int E>
__global__ __launch_bounds__(T) void k_fwd_cols(XGeom xg, Axis ax, const float2 *__restrict__ half,
                                                float2 *__restrict__ grid) {
    extern __shared__ float2 buf[];
    const uint32_t logc = xg.logc, cmask = (1u << logc) - 1u;
    const uint32_t y0 = blockIdx.x << logc, c = blockIdx.y;
    const uint32_t hy = xg.gy / 2;
    const float2 *src = half + static_cast<uint64_t>(c) * xg.nx * xg.half_pitch;
    const uint32_t tile = xg.gx << logc;
    for (uint32_t e = threadIdx.x; e < tile; e += T) {
        const uint32_t i = e >> logc, ky = y0 + (e & cmask);
        float2 v = make_float2(0.0f, 0.0f);
        if (ky <= hy && i >= xg.blc_x && i < xg.blc_x + xg.nx)
            v = src[static_cast<uint64_t>(i - xg.blc_x) * xg.half_pitch + ky];
        buf[e] = v;
    }
    __syncthreads();
    run_fft<false, T, E>(buf, ax, logc);
    float2 *dst = grid + static_cast<uint64_t>(c) * xg.cstride;
    for (uint32_t e = threadIdx.x; e < tile; e += T) {
        const uint32_t kx = e >> logc, ky = y0 + (e & cmask);
        if (ky > hy) continue;
        const float2 g = buf[e];
        dst[static_cast<uint64_t>(kx) * xg.gy + ky] = g;
        if (ky > 0 && ky < hy) {
            const uint32_t mx = kx == 0 ? 0u : xg.gx - kx;
            dst[static_cast<uint64_t>(mx) * xg.gy + (xg.gy - ky)] = make_float2(g.x, -g.y);
        }
    }
}

template <int T, int E>
__global__ __launch_bounds__(T) void k_inv_cols(XGeom xg, Axis ax, const float2 *__restrict__ grid,
                                                float2 *__restrict__ rows) {
    extern __shared__ float2 buf[];
    const uint32_t logc = xg.logc, cmask = (1u << logc) - 1u;
    const uint32_t y0 = blockIdx.x << logc, c = blockIdx.y;
    const float2 *src = grid + static_cast<uint64_t>(c) * xg.cstride;
    const uint32_t tile = xg.gx << logc;
    for (uint32_t e = threadIdx.x; e < tile; e += T) {
        const uint32_t i = e >> logc, ky = y0 + (e & cmask);
        buf[e] = ky < xg.gy ? src[static_cast<uint64_t>(i) * xg.gy + ky] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    run_fft<true, T, E>(buf, ax, logc);
    float2 *dst = rows + static_cast<uint64_t>(c) * xg.nx * xg.gy;
    const uint32_t out_n = xg.nx << logc;
    for (uint32_t e = threadIdx.x; e < out_n; e += T) {
        const uint32_t x = e >> logc, cc = e & cmask, ky = y0 + cc;
        if (ky >= xg.gy) continue;
        uint32_t i = xg.blc_x + x + xg.gx / 2;
        if (i >= xg.gx) i -= xg.gx;
        dst[static_cast<uint64_t>(x) * xg.gy + ky] = buf[(i << logc) + cc];
    }
}

template <int T, int E>
__global__ __launch_bounds__(T) void k_inv_rows(XGeom xg, Axis ay, const float2 *__restrict__ rows,
                                                const float *__restrict__ nsum,
                                                const float *__restrict__ peak, float *__restrict__ out) {
    extern __shared__ float2 buf[];
    const uint32_t x = blockIdx.x, c = blockIdx.y;
    const float2 *src = rows + (static_cast<uint64_t>(c) * xg.nx + x) * xg.gy;
    for (uint32_t i = threadIdx.x; i < xg.gy; i += T) buf[i] = src[i];
    __syncthreads();
    run_fft<true, T, E>(buf, ay, 0);
    const uint32_t gxx = xg.blc_x + x;
    const float scale = 1.0f / static_cast<float>(static_cast<uint64_t>(xg.gx) * xg.gy);
    const float ns = nsum[c], pk = peak[c];
    float *o = out + (static_cast<uint64_t>(c) * xg.nx + x) * xg.ny;
    for (uint32_t y = threadIdx.x; y < xg.ny; y += T) {
        const uint32_t gyy = xg.blc_y + y;
        uint32_t i = gyy + xg.gy / 2;
        if (i >= xg.gy) i -= xg.gy;
        float re = buf[i].x * scale;
        if ((gxx + gyy + xg.base) & 1) re = -re;
        const float factor = xg.corr_x[gxx] * xg.corr_y[gyy];
        float v;
        if (!xg.wproject) {
            v = re * factor;          // gridder.rs:1646-1657
            v = v / ns / pk;          // lib.rs:68928-68929
        } else {
            const float sf = xg.sinc[gxx] * xg.sinc[gyy];
            v = fabsf(sf) > 1.0e-6f ? re * (factor / sf) : 0.0f;  // gridder.rs:1685-1715
            v = 2.0f * v / ns / pk;                                // lib.rs:70960-70961
        }
        o[y] = v;
    }
}

// ---------------------------------------------------------------- host side

bool factor_axis(uint32_t n, Axis &ax) {
    ax.n = n;
    ax.n_stages = 0;
    const uint32_t order[4] = {5, 3, 4, 2};
    for (uint32_t p : order)
        while (n % p == 0) {
            if (ax.n_stages >= kMaxStages) return false;
            ax.radix[ax.n_stages++] = p;
            n /= p;
        }
    return n == 1;
}

std::vector<float2> twiddles(uint32_t n) {
    std::vector<float2> t(n);
    for (uint32_t m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * static_cast<double>(m) / static_cast<double>(n);
        t[m] = make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
    }
    return t;
}

namespace {

// Rows: 256 threads, up to 40 elements each (gy <= 10240).  Columns: 1024
// threads, up to 16 elements each (gx * C <= 16384, 128 KB of LDS); more
// per thread would spill registers.
constexpr int kRowT = 256;
constexpr int kColT = 1024;

int row_bucket(uint64_t per_thread) {
    if (per_thread <= 8) return 8;
    if (per_thread <= 16) return 16;
    if (per_thread <= 24) return 24;
    if (per_thread <= 40) return 40;
    return 0;
}

int col_bucket(uint64_t per_thread) {
    if (per_thread <= 8) return 8;
    if (per_thread <= 12) return 12;
    if (per_thread <= 16) return 16;
    return 0;
}

template <class K> void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        CHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

#define CHIP_XF_CASE(KERNEL, T, EV, GDIM, LDS, ST, ...)                                           \
    case EV: {                                                                                     \
        allow_lds(KERNEL<T, EV>, LDS);                                                             \
        hipLaunchKernelGGL((KERNEL<T, EV>), GDIM, dim3(T), LDS, ST, __VA_ARGS__);                   \
        break;                                                                                     \
    }

#define CHIP_XF_ROWS(E, KERNEL, GDIM, LDS, ST, ...)                                               \
    do {                                                                                           \
        switch (E) {                                                                               \
            CHIP_XF_CASE(KERNEL, kRowT, 8, GDIM, LDS, ST, __VA_ARGS__)                             \
            CHIP_XF_CASE(KERNEL, kRowT, 16, GDIM, LDS, ST, __VA_ARGS__)                            \
            CHIP_XF_CASE(KERNEL, kRowT, 24, GDIM, LDS, ST, __VA_ARGS__)                            \
            CHIP_XF_CASE(KERNEL, kRowT, 40, GDIM, LDS, ST, __VA_ARGS__)                            \
            default: fail(CHIP_ERR_INVALID, "row transform too long");                            \
        }                                                                                          \
        CHIP_HIP(hipGetLastError());                                                               \
    } while (0)

#define CHIP_XF_COLS(E, KERNEL, GDIM, LDS, ST, ...)                                               \
    do {                                                                                           \
        switch (E) {                                                                               \
            CHIP_XF_CASE(KERNEL, kColT, 8, GDIM, LDS, ST, __VA_ARGS__)                             \
            CHIP_XF_CASE(KERNEL, kColT, 12, GDIM, LDS, ST, __VA_ARGS__)                            \
            CHIP_XF_CASE(KERNEL, kColT, 16, GDIM, LDS, ST, __VA_ARGS__)                            \
            default: fail(CHIP_ERR_INVALID, "column transform tile too large");                   \
        }                                                                                          \
        CHIP_HIP(hipGetLastError());                                                               \
    } while (0)

}  // namespace

bool plan_transforms(const chip_gridder &g, uint64_t n_chan, size_t lds_limit, uint32_t logc_max,
                     XPlan &p) {
    if (g.gx % 2 || g.gy % 2) return false;
    if (n_chan == 0 || n_chan > 65535 || g.geom.nx > 0x7fffffffu) return false;
    if (!factor_axis(static_cast<uint32_t>(g.gx), p.ax) || !factor_axis(static_cast<uint32_t>(g.gy), p.ay))
        return false;
    p.row_e = row_bucket((g.gy + kRowT - 1) / kRowT);
    if (!p.row_e || g.gy * sizeof(float2) > lds_limit) return false;
    // widest column block that fits the LDS and the register budget
    p.logc = 0;
    bool ok = false;
    for (int lc = static_cast<int>(logc_max); lc >= 0; --lc) {
        const uint64_t tile = g.gx << lc;
        if (tile * sizeof(float2) <= lds_limit && tile <= 16ull * kColT) {
            p.logc = static_cast<uint32_t>(lc);
            ok = true;
            break;
        }
    }
    if (!ok) return false;
    p.col_e = col_bucket(((g.gx << p.logc) + kColT - 1) / kColT);
    p.xg.nx = static_cast<uint32_t>(g.geom.nx);
    p.xg.ny = static_cast<uint32_t>(g.geom.ny);
    p.xg.gx = static_cast<uint32_t>(g.gx);
    p.xg.gy = static_cast<uint32_t>(g.gy);
    p.xg.blc_x = static_cast<uint32_t>(g.blc_x);
    p.xg.blc_y = static_cast<uint32_t>(g.blc_y);
    p.xg.base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
    p.xg.wproject = g.wproject ? 1u : 0u;
    p.xg.half_pitch = static_cast<uint32_t>(((g.gy / 2 + 1) + 7) / 8 * 8);
    // the half spectrum and the inverse rows borrow the grid buffers
    if (g.geom.nx * p.xg.half_pitch > g.gx * g.gy) return false;
    p.xg.logc = p.logc;
    p.xg.cstride = g.gx * g.gy;
    p.xg.corr_x = g.d_corr_x.as<float>();
    p.xg.corr_y = g.d_corr_y.as<float>();
    p.xg.sinc = g.wproject ? g.d_sinc.as<float>() : nullptr;
    p.xg.sinc_x = g.wproject ? g.d_sinc_x.as<float>() : nullptr;
    p.xg.sinc_y = g.wproject ? g.d_sinc_y.as<float>() : nullptr;
    p.n_chan = static_cast<uint32_t>(n_chan);
    std::vector<float2> tx = twiddles(p.ax.n), ty = twiddles(p.ay.n);
    p.tw_x.alloc(tx.size() * sizeof(float2));
    p.tw_y.alloc(ty.size() * sizeof(float2));
    CHIP_HIP(hipMemcpy(p.tw_x.p, tx.data(), p.tw_x.bytes, hipMemcpyHostToDevice));
    CHIP_HIP(hipMemcpy(p.tw_y.p, ty.data(), p.tw_y.bytes, hipMemcpyHostToDevice));
    p.ax.tw = p.tw_x.as<float2>();
    p.ay.tw = p.tw_y.as<float2>();
    return true;
}

void forward_model(const XPlan &p, const float *model, float2 *half, float2 *grid, hipStream_t st) {
    const dim3 rows(p.xg.nx, p.n_chan);
    const size_t row_lds = p.xg.gy * sizeof(float2);
    CHIP_XF_ROWS(p.row_e, k_fwd_rows, rows, row_lds, st, p.xg, p.ay, model, half);
    const uint32_t cols = p.xg.gy / 2 + 1;
    const dim3 blocks((cols + (1u << p.logc) - 1) >> p.logc, p.n_chan);
    const size_t col_lds = (static_cast<size_t>(p.xg.gx) << p.logc) * sizeof(float2);
    CHIP_XF_COLS(p.col_e, k_fwd_cols, blocks, col_lds, st, p.xg, p.ax, half, grid);
}

void inverse_residual(const XPlan &p, const float2 *grid, float2 *rows_buf, const float *nsum,
                      const float *peak, float *out, hipStream_t st) {
    const dim3 blocks((p.xg.gy + (1u << p.logc) - 1) >> p.logc, p.n_chan);
    const size_t col_lds = (static_cast<size_t>(p.xg.gx) << p.logc) * sizeof(float2);
    CHIP_XF_COLS(p.col_e, k_inv_cols, blocks, col_lds, st, p.xg, p.ax, grid, rows_buf);
    const dim3 rows(p.xg.nx, p.n_chan);
    const size_t row_lds = p.xg.gy * sizeof(float2);
    CHIP_XF_ROWS(p.row_e, k_inv_rows, rows, row_lds, st, p.xg, p.ay, rows_buf, nsum, peak, out);
}

}  // namespace xf
}  // namespace chip
