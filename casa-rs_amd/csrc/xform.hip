// Image <-> uv-grid transforms of the major cycle with the pre- and
// post-processing fused into the FFT passes.
//
// Forward (centered_fft2(apodize_model(model)), gridder.rs:1554-1592,
// fft.rs:34): the model is real, so the 2-D transform is Hermitian,
// B[-kx, -ky] = conj(B[kx, ky]).
//   rows pass   : 8 image rows per workgroup; the apodised,
//                 checkerboard-signed model rows zero-padded to gy points,
//                 two real rows per complex FFT along y, split into their
//                 spectra; only ky in [0, gy/2] is kept (half spectrum).
//   column pass : one block of 8 columns ky <= gy/2 per workgroup; rows
//                 outside the image are zero -> FFT along x; writes column
//                 ky only (the binned kernel reads the columns ky > gy/2 as
//                 conjugate mirrors).
// Inverse (centered inverse FFT + crop + correction, gridder.rs:1646-1715):
// only Re of the inverse is used, which is the inverse of the Hermitian part
// H = (G[k] + conj G[-k]) / 2 of the residual grid G.
//   column pass : one block of 8 columns ky <= gy/2 per workgroup of 2H
//                 (accumulated so by the binned kernel), IFFT along x; keeps only
//                 the nx rows the crop needs (shift theorem: output row
//                 (blc_x + x + gx/2) mod gx).
//   rows pass   : 8 rows per workgroup, each Hermitian in ky, two rows per
//                 complex IFFT along y (real and imaginary outputs); the
//                 crop, 1/(gx gy) scale, sign (-1)^(x + y + gx/2 + gy/2),
//                 grid correction and normalisation are applied on the way
//                 out.
//
// A workgroup holds 8 (columns) or 4 (row pairs) transforms in LDS, 64
// threads per transform, and runs a mixed-radix Stockham autosort FFT whose
// stages pair two radices (25 = 5 x 5, 20, 16, 15, 12, 10, 9, 8, 6, or one
// radix 2 .. 5 left over), so a length-1250 transform takes three passes
// through LDS (25, 25, 2) instead of five.
// Stage with radix R = r1 r2 after sub-transforms of length ns: butterfly j
// reads elements j + t n/R, multiplies input t by exp(-+2 pi i k t / (ns R))
// with k = j mod ns, takes their R-point DFT and writes output s to
// (j - k) R + k + s ns.  The R-point DFT runs in registers as r1 DFTs of
// r2 points, the constant twiddles exp(-+2 pi i t1 s2 / R), and r2 DFTs of
// r1 points (t = t1 + r1 t2, s = s2 + r2 s1); the stage twiddle of input t
// is applied as exp(-+2 pi i k r1 t2 / (ns R)) times exp(-+2 pi i k t1 /
// (ns R)), both rounded once from FP64 and read from a per-stage table in
// global memory (a few KB, cache resident), which keeps LDS for the data.
// Every stage reads its inputs into registers, synchronises, and writes its
// outputs back, so one LDS buffer suffices.  Column tiles are interleaved in
// LDS (element i of transform c at 8 i + c) to match the paired grid layout;
// row tiles are row-major (c n + i).
#include <cmath>
#include <type_traits>

#include "device.hpp"
#include "xform.hpp"

namespace chip {
namespace xf {

// Word 3 of a raw buffer resource on gfx9-family GPUs (no format, no
// swizzle), and an offset beyond any buffer this file builds (loads there
// return 0).
constexpr int kBufferFlags = 0x00020000;
constexpr uint32_t kBufferOut = 0x80000000u;

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

// exp(-2 pi i m / R) for m < R <= 32, evaluated at compile time (Taylor
// series in FP64 on [-pi, pi], exact at multiples of a quarter turn).
constexpr double kPi = 3.14159265358979323846;
constexpr double taylor_sin(double x) {
    double term = x, sum = x;
    for (int i = 1; i < 30; ++i) {
        term *= -x * x / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double taylor_cos(double x) {
    double term = 1.0, sum = 1.0;
    for (int i = 1; i < 30; ++i) {
        term *= -x * x / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
struct Roots {
    float re[32], im[32];
};
constexpr Roots make_roots(int R) {
    Roots r{};
    for (int m = 0; m < R; ++m) {
        if ((4 * m) % R == 0) {  // 1, -i, -1, +i
            const int q = 4 * m / R;
            r.re[m] = q == 0 ? 1.0f : q == 2 ? -1.0f : 0.0f;
            r.im[m] = q == 1 ? -1.0f : q == 3 ? 1.0f : 0.0f;
            continue;
        }
        double a = -2.0 * kPi * m / R;
        if (a < -kPi) a += 2.0 * kPi;
        r.re[m] = static_cast<float>(taylor_cos(a));
        r.im[m] = static_cast<float>(taylor_sin(a));
    }
    return r;
}

template <bool INV> __device__ __forceinline__ float2 conj_if(float2 w) { return INV ? make_float2(w.x, -w.y) : w; }

// One Stockham stage of radix R1 * R2.  Thread (c, jt) handles butterflies
// j = jt + b * LANES of transform c; MB bounds the butterflies per thread.
// tw: this stage's twiddle table (read only when ns > 1); nb = n / R comes
// from the stage table, so that addresses are not hoisted out of the stage
// loop for every radix at once.
template <int R1, int R2, bool INV, int MB, bool ILV, int LANES, int LOGB>
__device__ __forceinline__ void stage(float2 *buf, const float2 *__restrict__ tw, uint32_t nb, uint32_t pitch,
                                      uint32_t ns, float inv_ns, uint32_t c, uint32_t jt) {
    constexpr int R = R1 * R2, M = (R1 - 1) + (R2 - 1);
    constexpr Roots kr = make_roots(R);
    const bool twiddled = ns > 1;
    float2 v[MB][R], w[MB][M];
    uint32_t qv[MB], kv[MB];
    // twiddle loads first: they travel while the LDS reads and the barrier run
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const uint32_t j = jt + b * LANES;
        // j / ns, exact for n < 2^22: the fraction of (j + 1/2) / ns is at
        // least 1/(2 ns) away from an integer, the float error far smaller
        qv[b] = static_cast<uint32_t>((static_cast<float>(j) + 0.5f) * inv_ns);
        kv[b] = j - qv[b] * ns;
        if (twiddled && j < nb) {
            const float2 *t = tw + static_cast<size_t>(kv[b]) * M;
#pragma unroll
            for (int m = 0; m < M; ++m) w[b][m] = t[m];
        }
    }
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const uint32_t j = jt + b * LANES;
        if (j < nb) {
#pragma unroll
            for (int t = 0; t < R; ++t) {
                const uint32_t i = j + t * nb;
                v[b][t] = buf[ILV ? (i << LOGB) + c : c * pitch + i];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < MB; ++b) {
        const uint32_t j = jt + b * LANES;
        if (j >= nb) continue;
        if (twiddled) {
#pragma unroll
            for (int t2 = 1; t2 < R2; ++t2) {
                const float2 wt = conj_if<INV>(w[b][t2 - 1]);
#pragma unroll
                for (int t1 = 0; t1 < R1; ++t1) v[b][t1 + R1 * t2] = cmul(v[b][t1 + R1 * t2], wt);
            }
#pragma unroll
            for (int t1 = 1; t1 < R1; ++t1) {
                const float2 wt = conj_if<INV>(w[b][R2 - 1 + t1 - 1]);
#pragma unroll
                for (int t2 = 0; t2 < R2; ++t2) v[b][t1 + R1 * t2] = cmul(v[b][t1 + R1 * t2], wt);
            }
        }
        if constexpr (R2 > 1) {
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) {
                float2 a[R2];
#pragma unroll
                for (int t2 = 0; t2 < R2; ++t2) a[t2] = v[b][t1 + R1 * t2];
                butterfly<R2, INV>(a);
#pragma unroll
                for (int s2 = 0; s2 < R2; ++s2) v[b][t1 + R1 * s2] = a[s2];
            }
#pragma unroll
            for (int t1 = 1; t1 < R1; ++t1)
#pragma unroll
                for (int s2 = 1; s2 < R2; ++s2) {
                    const int m = (t1 * s2) % R;
                    v[b][t1 + R1 * s2] = cmul(v[b][t1 + R1 * s2], make_float2(kr.re[m], INV ? -kr.im[m] : kr.im[m]));
                }
        }
        const uint32_t base = qv[b] * ns * R + kv[b];
#pragma unroll
        for (int s2 = 0; s2 < R2; ++s2) {
            float2 a[R1];
#pragma unroll
            for (int t1 = 0; t1 < R1; ++t1) a[t1] = v[b][t1 + R1 * s2];
            butterfly<R1, INV>(a);
#pragma unroll
            for (int s1 = 0; s1 < R1; ++s1) {
                const uint32_t i = base + (s2 + R2 * s1) * ns;
                buf[ILV ? (i << LOGB) + c : c * pitch + i] = a[s1];
            }
        }
    }
    __syncthreads();
}

// E: elements per thread per transform (n <= LANES E); interleaved tiles
// hold 2^LOGB transforms (element i of transform c at (i << LOGB) + c).
template <bool INV, int E, bool ILV, int LANES, int LOGB = kLogBlock>
__device__ __forceinline__ void run_fft(float2 *buf, const Axis &ax, uint32_t pitch, uint32_t c, uint32_t jt) {
    for (uint32_t s = 0; s < ax.n_stages; ++s) {
        const uint32_t ns = ax.ns[s];
        const float inv = ax.inv_ns[s];
        const float2 *tw = ax.tw + ax.tw_off[s];
        switch (ax.r1[s] * 8 + ax.r2[s]) {
#define CHIP_XF_STAGE(A, B)                                                                          \
    case A * 8 + B:                                                                                  \
        stage<A, B, INV, (E + A * B - 1) / (A * B), ILV, LANES, LOGB>(buf, tw, ax.nb[s], pitch, ns, inv, c, jt); \
        break;
            CHIP_XF_STAGE(5, 5)
            CHIP_XF_STAGE(5, 3)
            CHIP_XF_STAGE(5, 4)
            CHIP_XF_STAGE(5, 2)
            CHIP_XF_STAGE(3, 3)
            CHIP_XF_STAGE(3, 4)
            CHIP_XF_STAGE(3, 2)
            CHIP_XF_STAGE(4, 4)
            CHIP_XF_STAGE(4, 2)
            CHIP_XF_STAGE(5, 1)
            CHIP_XF_STAGE(3, 1)
            CHIP_XF_STAGE(4, 1)
            CHIP_XF_STAGE(2, 1)
#undef CHIP_XF_STAGE
            default: break;
        }
    }
}

// ---------------------------------------------------------------- kernels
// Rows passes: 8 rows per workgroup of 4 LANES threads, packed as 4 complex
// transforms of two real rows each (row 2p real part, row 2p + 1 imaginary
// part), row-major in LDS; grid (nxp / 8, n_chan).  Column passes: one block
// of 8 columns (4 for axes above 2048 points) per workgroup of 512 threads,
// interleaved in LDS; grid (nyb_h x columns-per-block split,
// n_chan).  LDS holds only the data (two column workgroups per CU up to
// length 1280); waves_per_eu states the occupancy the LDS allows, so the
// compiler may give a thread up to 128 VGPRs.

// Coverage mask of one upper block as kRowWords 64-bit words (row chunk ch
// at bit ch % 64 of word ch / 64), gathered by wavefront ballots.
struct RowMask {
    // one named word per 64 row chunks (kRowWords = 4): named members, not an
    // array, so that a lane-dependent choice of word stays in registers
    uint64_t w0, w1, w2, w3;
};
static_assert(kRowWords == 4, "RowMask holds four words");
__device__ __forceinline__ RowMask row_mask(const XGeom &xg, uint32_t c, uint32_t yb) {
    if (!xg.touch) return RowMask{~0ull, ~0ull, ~0ull, ~0ull};
    const uint8_t *t = xg.touch + (static_cast<uint64_t>(c) * xg.nyb_h + yb) * xg.nxc;
    const uint32_t lane = threadIdx.x & 63u;
    auto word = [&](uint32_t q) {
        const uint32_t i = lane + 64 * q;
        const uint8_t f = t[i < xg.nxc ? i : 0];
        return static_cast<uint64_t>(__ballot(i < xg.nxc && f != 0));
    };
    return RowMask{word(0), word(1), word(2), word(3)};
}
__device__ __forceinline__ bool row_touched(const RowMask &m, uint32_t x) {
    const uint32_t ch = x / kTouchRows;
    const uint32_t q = ch >> 6;
    const uint64_t w = q == 0 ? m.w0 : q == 1 ? m.w1 : q == 2 ? m.w2 : m.w3;
    return ch < 64 * kRowWords && ((w >> (ch & 63u)) & 1ull) != 0;
}

// Block flags of one channel (touch_blk) as kBlockWords 64-bit words,
// gathered by wavefront ballots; every block without a mask.
struct BlockMask {
    uint64_t w0, w1, w2, w3, w4;  // (named, as RowMask)
};
static_assert(kBlockWords == 5, "BlockMask holds five words");
__device__ __forceinline__ BlockMask block_mask(const XGeom &xg, uint32_t c) {
    if (!xg.touch_blk) return BlockMask{~0ull, ~0ull, ~0ull, ~0ull, ~0ull};
    const uint8_t *t = xg.touch_blk + static_cast<uint64_t>(c) * xg.nyb_h;
    const uint32_t lane = threadIdx.x & 63u;
    auto word = [&](uint32_t q) {
        const uint32_t i = lane + 64 * q;
        const uint8_t f = t[i < xg.nyb_h ? i : 0];
        return static_cast<uint64_t>(__ballot(i < xg.nyb_h && f != 0));
    };
    return BlockMask{word(0), word(1), word(2), word(3), word(4)};
}
__device__ __forceinline__ bool block_live(const BlockMask &m, uint32_t yb) {
    const uint32_t q = yb >> 6;
    const uint64_t w = q == 0 ? m.w0 : q == 1 ? m.w1 : q == 2 ? m.w2 : q == 3 ? m.w3 : m.w4;
    return q < kBlockWords && ((w >> (yb & 63u)) & 1ull) != 0;
}

__device__ __forceinline__ uint64_t half_at(const XGeom &xg, uint32_t yb, uint32_t x, uint32_t lane) {
    return static_cast<uint64_t>(yb) * xg.nxp * kBlock + x * kBlock + lane;
}

template <int E, int LANES>
__global__ __launch_bounds__(kPairs * LANES) __attribute__((amdgpu_waves_per_eu(E <= 20 && LANES == 64 ? 4 : 2))) void k_fwd_rows(
    XGeom xg, Axis ay, const float *__restrict__ model, float2 *__restrict__ half) {
    extern __shared__ float2 buf[];
    constexpr uint32_t NT = kPairs * LANES;
    const uint32_t n = xg.gy;
    const uint32_t tid = threadIdx.x, p = tid / LANES, jt = tid % LANES;
    const uint32_t x0 = blockIdx.x * kRowsPerGroup, c = blockIdx.y;
    // z = a + i b for rows a = x0 + 2p, b = a + 1 (apodised, signed, zero-padded)
    float f[2], sxv[2];
    const float *row[2];
    bool live[2];
    uint32_t gxx[2];
    for (int q = 0; q < 2; ++q) {
        const uint32_t x = x0 + 2 * p + q;
        live[q] = x < xg.nx;
        gxx[q] = xg.blc_x + x;
        row[q] = model + (static_cast<uint64_t>(c) * xg.nx + x) * xg.ny;
        f[q] = live[q] ? xg.corr_x[gxx[q]] : 0.0f;
        sxv[q] = live[q] && xg.wproject ? xg.sinc_x[gxx[q]] : 1.0f;
    }
    // Two phases per chunk of CH elements: every global load of the chunk is
    // issued before the first use, so they are in flight together.
    constexpr int CH = E < 16 ? E : 16;
#pragma unroll
    for (int t0 = 0; t0 < E; t0 += CH) {
        float mv[CH][2], cyv[CH], syv[CH];
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const uint32_t i = jt + static_cast<uint32_t>(t0 + u) * LANES;
            const bool in = t0 + u < E && i < n && i >= xg.blc_y && i < xg.blc_y + xg.ny;
            cyv[u] = in ? xg.corr_y[i] : 0.0f;
            syv[u] = in && xg.wproject ? xg.sinc_y[i] : 1.0f;
            for (int q = 0; q < 2; ++q) mv[u][q] = in && live[q] ? row[q][i - xg.blc_y] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < CH; ++u) {
            const uint32_t i = jt + static_cast<uint32_t>(t0 + u) * LANES;
            if (t0 + u >= E || i >= n) break;
            float v[2] = {0.0f, 0.0f};
            if (i >= xg.blc_y && i < xg.blc_y + xg.ny) {
                for (int q = 0; q < 2; ++q) {
                    if (!live[q]) continue;
                    // factor evaluated left to right as the reference: cx*cy [*sx*sy]
                    float fac = f[q] * cyv[u];
                    if (xg.wproject) fac = fac * sxv[q] * syv[u];
                    float val = mv[u][q] * fac;
                    if ((gxx[q] + i) & 1) val = -val;
                    v[q] = val;
                }
            }
            buf[p * n + i] = make_float2(v[0], v[1]);
        }
    }
    __syncthreads();
    run_fft<false, E, false, LANES>(buf, ay, n, p, jt);
    // A[k] = (Z[k] + conj Z[-k]) / 2, B[k] = (Z[k] - conj Z[-k]) / 2i, k <= gy/2,
    // into the half spectrum [yb][x][8]: 8 rows x 8 lanes contiguous per block
    // (blocks of columns the uv coverage misses are not stored: the column
    // pass skips them)
    float2 *h = half + static_cast<uint64_t>(c) * xg.hstride;
    const uint32_t hy = n / 2;
    // the block flags straight from the cached flag bytes (a mask word chosen
    // per lane would sit in scratch memory)
    const uint8_t *tb = xg.touch_blk ? xg.touch_blk + static_cast<uint64_t>(c) * xg.nyb_h : nullptr;
    for (uint32_t e = tid; e < xg.nyb_h * (kPairs * kBlock); e += NT) {
        const uint32_t yb = e / (kPairs * kBlock), pp = (e / kBlock) % kPairs, lane = e % kBlock;
        if (tb && !tb[yb]) continue;
        const uint32_t k = yb * kBlock + lane;
        float2 A = make_float2(0.0f, 0.0f), B = A;
        if (k <= hy) {
            const float2 z = buf[pp * n + k], zm = buf[pp * n + (k == 0 ? 0 : n - k)];
            A = make_float2(0.5f * (z.x + zm.x), 0.5f * (z.y - zm.y));
            B = make_float2(0.5f * (z.y + zm.y), 0.5f * (zm.x - z.x));
        }
        const uint32_t xa = x0 + 2 * pp;
        h[half_at(xg, yb, xa, lane)] = A;
        h[half_at(xg, yb, xa + 1, lane)] = B;
    }
}

// Column passes: a workgroup holds BLK = 2^LOGB columns of an 8-column block
// (all 8, or half of them for long axes: workgroup blockIdx.x is block
// blockIdx.x / SUB, columns col0 .. col0 + BLK - 1 of it), interleaved in LDS.
template <int E, int LANES, int LOGB>
__global__ __launch_bounds__(LANES << LOGB) __attribute__((amdgpu_waves_per_eu(E <= 20 && LANES == 64 ? 4 : 2))) void k_fwd_cols(
    XGeom xg, Axis ax, const float2 *__restrict__ half, float2 *__restrict__ grid) {
    extern __shared__ float2 buf[];
    constexpr uint32_t BLK = 1u << LOGB, NT = LANES << LOGB, SUB = kBlock / BLK;
    const uint32_t n = xg.gx;
    const uint32_t tid = threadIdx.x, cc = tid % BLK, jt = tid / BLK;
    const uint32_t yb = blockIdx.x / SUB, col0 = (blockIdx.x % SUB) * BLK, c = blockIdx.y;
    if (xg.touch_blk && !xg.touch_blk[static_cast<uint64_t>(c) * xg.nyb_h + yb]) return;  // outside the uv coverage
    const float2 *h = half + static_cast<uint64_t>(c) * xg.hstride + static_cast<uint64_t>(yb) * xg.nxp * kBlock;
    float2 hv[E];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT, i = e >> LOGB;
        hv[t] = (e < n * BLK && i >= xg.blc_x && i < xg.blc_x + xg.nx)
                    ? h[(i - xg.blc_x) * kBlock + col0 + (e & (BLK - 1))]
                    : make_float2(0.0f, 0.0f);
    }
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT;
        if (e < n * BLK) buf[e] = hv[t];
    }
    __syncthreads();
    run_fft<false, E, true, LANES, LOGB>(buf, ax, 0, cc, jt);
    // column ky into upper block yb.  The columns ky > gy/2 are not stored:
    // the transform of a real image is Hermitian, so the binned kernel reads
    // G[kx, ky] as conj G[-kx, gy - ky] from the upper blocks.  Rows outside
    // the uv coverage are not stored either: the binned kernel never reads them.
    const uint64_t bstride = static_cast<uint64_t>(n) * kBlock;
    float2 *up = grid + static_cast<uint64_t>(c) * xg.gstride + yb * bstride;
    const uint32_t hy = xg.gy / 2;
    // row chunks from the cached flag bytes (as in the rows pass)
    const uint8_t *tr = xg.touch ? xg.touch + (static_cast<uint64_t>(c) * xg.nyb_h + yb) * xg.nxc : nullptr;
    for (uint32_t e = tid; e < n * BLK; e += NT) {
        const uint32_t x = e >> LOGB, lane = col0 + (e & (BLK - 1));
        const uint32_t ky = yb * kBlock + lane;
        if (ky <= hy && (!tr || tr[x / kTouchRows])) up[x * kBlock + lane] = buf[e];
    }
}

template <int E, int LANES, int LOGB>
__global__ __launch_bounds__(LANES << LOGB) __attribute__((amdgpu_waves_per_eu(E <= 20 && LANES == 64 ? 4 : 2))) void k_inv_cols(
    XGeom xg, Axis ax, const float2 *__restrict__ grid, float2 *__restrict__ rows) {
    extern __shared__ float2 buf[];
    constexpr uint32_t BLK = 1u << LOGB, NT = LANES << LOGB, SUB = kBlock / BLK;
    const uint32_t n = xg.gx;
    const uint32_t tid = threadIdx.x, cc = tid % BLK, jt = tid / BLK;
    const uint32_t yb = blockIdx.x / SUB, col0 = (blockIdx.x % SUB) * BLK, c = blockIdx.y;
    if (xg.touch_blk && !xg.touch_blk[static_cast<uint64_t>(c) * xg.nyb_h + yb]) return;  // outside the uv coverage
    // Hermitian part H[kx, ky] = (G[kx, ky] + conj G[-kx, -ky]) / 2 of the
    // columns ky <= gy/2, accumulated as 2H by the binned kernel; its inverse
    // transform is real and equals Re(centered inverse of G) after the shifts.
    const uint64_t bstride = static_cast<uint64_t>(n) * kBlock;
    const float2 *up = grid + static_cast<uint64_t>(c) * xg.gstride + yb * bstride;
    const uint32_t hy = xg.gy / 2;
    // Buffer loads, issued unconditionally so that they are all in flight
    // together (a conditional load is compiled as a branch that waits for
    // it): an element outside the block, a column ky > gy/2 or a row outside
    // the uv coverage (zero residual) gets an offset beyond the buffer's
    // size, for which the load returns 0 without a memory access.
    const RowMask rm = row_mask(xg, c, yb);
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2 *>(up), 0, static_cast<int>(n * kBlock * sizeof(float2)), kBufferFlags);
    float2 gv[E];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT;
        const uint32_t x = e >> LOGB, lane = col0 + (e & (BLK - 1));
        const uint32_t ky = yb * kBlock + lane;
        const bool live = e < n * BLK && ky <= hy && row_touched(rm, x);
        gv[t] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                               rs, live ? (x * kBlock + lane) * static_cast<uint32_t>(sizeof(float2))
                                                        : kBufferOut,
                                               0, 0));
    }
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT;
        if (e < n * BLK) buf[e] = make_float2(0.5f * gv[t].x, 0.5f * gv[t].y);
    }
    __syncthreads();
    run_fft<true, E, true, LANES, LOGB>(buf, ax, 0, cc, jt);
    float2 *o = rows + static_cast<uint64_t>(c) * xg.hstride + static_cast<uint64_t>(yb) * xg.nxp * kBlock;
    for (uint32_t e = tid; e < xg.nx * BLK; e += NT) {
        const uint32_t x = e >> LOGB, lane = e & (BLK - 1);
        uint32_t i = xg.blc_x + x + n / 2;
        if (i >= n) i -= n;
        o[x * kBlock + col0 + lane] = buf[(i << LOGB) + lane];
    }
}

template <int E, int LANES>
__global__ __launch_bounds__(kPairs * LANES) __attribute__((amdgpu_waves_per_eu(E <= 20 && LANES == 64 ? 4 : 2))) void k_inv_rows(
    XGeom xg, Axis ay, const float2 *__restrict__ rows, const float *__restrict__ nsum,
    const float *__restrict__ peak, float *__restrict__ out) {
    extern __shared__ float2 buf[];
    constexpr uint32_t NT = kPairs * LANES;
    const uint32_t n = xg.gy;
    const uint32_t tid = threadIdx.x, p = tid / LANES, jt = tid % LANES;
    const uint32_t x0 = blockIdx.x * kRowsPerGroup, c = blockIdx.y;
    // Rows a = x0 + 2p and b = a + 1 are Hermitian in ky: Z[k] = A[k] + i B[k]
    // over the full length (A[-k] = conj A[k]), whose inverse is a + i b.
    const float2 *h = rows + static_cast<uint64_t>(c) * xg.hstride;
    const uint32_t hy = n / 2;
    const uint32_t rows_live = xg.nx - x0;  // rows of this group inside the image (>= 1)
    // elements per thread: ceil(nyb_h * 32 / NT) <= E / 2 + 1 (checked when
    // the transforms are planned)
    constexpr int EL = E / 2 + 1;
    const uint32_t total = xg.nyb_h * (kPairs * kBlock);
    // Buffer loads (see k_inv_cols); the blocks of columns the uv coverage
    // misses were not transformed and read as zero.
    const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<float2 *>(h), 0, static_cast<int>(xg.hstride * sizeof(float2)), kBufferFlags);
    const BlockMask bm = block_mask(xg, c);
    float2 av[EL], bv[EL];
#pragma unroll
    for (int t = 0; t < EL; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT;
        const uint32_t yb = e / (kPairs * kBlock), pp = (e / kBlock) % kPairs, lane = e % kBlock;
        const uint32_t k = yb * kBlock + lane, ra = 2 * pp;
        const bool live = e < total && k <= hy && block_live(bm, yb);
        const uint32_t oa = static_cast<uint32_t>(half_at(xg, yb, x0 + ra, lane) * sizeof(float2));
        av[t] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                               rs, live && ra < rows_live ? oa : kBufferOut, 0, 0));
        bv[t] = __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(
                                               rs, live && ra + 1 < rows_live ? oa + kBlock * sizeof(float2) : kBufferOut,
                                               0, 0));
    }
#pragma unroll
    for (int t = 0; t < EL; ++t) {
        const uint32_t e = tid + static_cast<uint32_t>(t) * NT;
        const uint32_t yb = e / (kPairs * kBlock), pp = (e / kBlock) % kPairs, lane = e % kBlock;
        const uint32_t k = yb * kBlock + lane;
        if (e >= total || k > hy) continue;
        const float2 A = av[t], B = bv[t];
        buf[pp * n + k] = make_float2(A.x - B.y, A.y + B.x);
        if (k > 0 && k < hy) buf[pp * n + n - k] = make_float2(A.x + B.y, B.x - A.y);
    }
    __syncthreads();
    run_fft<true, E, false, LANES>(buf, ay, n, p, jt);
    // Output v = Re * scale * sign * (cx cy) / ns / pk (gridder.rs:1646-1657,
    // lib.rs:68928-68929), with W projection v = 2 Re * scale * sign * (cx cy)
    // / (sx sy) / ns / pk where |sx sy| > 1e-6, else 0 (gridder.rs:1685-1715,
    // lib.rs:70960-70961).  The divisions are folded into per-row, per-column
    // and per-channel factors (cx / sx, cy / sy, gain / (ns pk)), so a pixel
    // costs a few products; this moves the result by a few ulp.
    const float scale = 1.0f / static_cast<float>(static_cast<uint64_t>(xg.gx) * xg.gy);
    const float chan_gain = (xg.wproject ? 2.0f : 1.0f) / nsum[c] / peak[c];
    // y factors of this thread's columns, loaded together (ny <= n <= E LANES)
    float ryv[E], syv[E];
#pragma unroll
    for (int t = 0; t < E; ++t) {
        const uint32_t y = jt + static_cast<uint32_t>(t) * LANES, gyy = xg.blc_y + y;
        const float cy = y < xg.ny ? xg.corr_y[gyy] : 0.0f;
        syv[t] = y < xg.ny && xg.wproject ? xg.sinc[gyy] : 1.0f;
        ryv[t] = xg.wproject ? cy / syv[t] : cy;
    }
    for (int q = 0; q < 2; ++q) {
        const uint32_t x = x0 + 2 * p + q;
        if (x >= xg.nx) break;
        const uint32_t gxx = xg.blc_x + x;
        const float sincx = xg.wproject ? xg.sinc[gxx] : 1.0f;
        const float rx = xg.wproject ? xg.corr_x[gxx] / sincx : xg.corr_x[gxx];
        const float row_gain = scale * rx * chan_gain;
        float *o = out + (static_cast<uint64_t>(c) * xg.nx + x) * xg.ny;
#pragma unroll
        for (int t = 0; t < E; ++t) {
            const uint32_t y = jt + static_cast<uint32_t>(t) * LANES;
            if (y >= xg.ny) break;
            const uint32_t gyy = xg.blc_y + y;
            uint32_t i = gyy + n / 2;
            if (i >= n) i -= n;
            const float2 z = buf[p * n + i];
            float v = (q == 0 ? z.x : z.y) * row_gain * ryv[t];
            if ((gxx + gyy + xg.base) & 1) v = -v;
            if (xg.wproject && !(fabsf(sincx * syv[t]) > 1.0e-6f)) v = 0.0f;
            o[y] = v;
        }
    }
}

// Zero fill of the covered rows of the upper blocks: grid (nyb_h, n_chan),
// 256 threads; a row of a block is 8 float2 = 4 float4.
__global__ __launch_bounds__(256) void k_zero_touched(XGeom xg, float4 *__restrict__ grid) {
    const uint32_t yb = blockIdx.x, c = blockIdx.y;
    // row chunks from the cached flag bytes (as in the forward passes)
    const uint8_t *tr = xg.touch ? xg.touch + (static_cast<uint64_t>(c) * xg.nyb_h + yb) * xg.nxc : nullptr;
    float4 *blk = grid + (static_cast<uint64_t>(c) * xg.gstride + static_cast<uint64_t>(yb) * xg.gx * kBlock) / 2;
    const uint32_t n4 = xg.gx * (kBlock / 2);
    for (uint32_t q = threadIdx.x; q < n4; q += 256)
        if (!tr || tr[(q >> 2) / kTouchRows]) blk[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
}

// ---------------------------------------------------------------- host side

bool factor_axis(uint32_t n, Axis &ax, std::vector<float2> &table) {
    ax.n = n;
    ax.n_stages = 0;
    table.clear();
    if (n == 0 || n >= (1u << 22)) return false;
    // radices 5, 3, 4 (as many as divide), then at most one 2; consecutive
    // radices pair into one stage
    std::vector<uint32_t> f;
    uint32_t rest = n;
    for (uint32_t p : {5u, 3u, 4u, 2u})
        while (rest % p == 0) {
            f.push_back(p);
            rest /= p;
        }
    if (rest != 1) return false;
    uint32_t ns = 1;
    for (size_t i = 0; i < f.size(); i += 2) {
        if (ax.n_stages >= kMaxStages) return false;
        const uint32_t s = ax.n_stages++;
        const uint32_t r1 = f[i], r2 = i + 1 < f.size() ? f[i + 1] : 1u, R = r1 * r2;
        ax.r1[s] = r1;
        ax.r2[s] = r2;
        ax.ns[s] = ns;
        ax.nb[s] = n / R;
        ax.inv_ns[s] = 1.0f / static_cast<float>(ns);
        ax.tw_off[s] = static_cast<uint32_t>(table.size());
        if (ns > 1) {
            const uint64_t N = static_cast<uint64_t>(ns) * R;
            auto root = [&](uint64_t m) {  // exp(-2 pi i m / N)
                const double a = -2.0 * M_PI * static_cast<double>(m % N) / static_cast<double>(N);
                return make_float2(static_cast<float>(std::cos(a)), static_cast<float>(std::sin(a)));
            };
            for (uint32_t k = 0; k < ns; ++k) {
                for (uint32_t t2 = 1; t2 < r2; ++t2) table.push_back(root(static_cast<uint64_t>(k) * r1 * t2));
                for (uint32_t t1 = 1; t1 < r1; ++t1) table.push_back(root(static_cast<uint64_t>(k) * t1));
            }
        }
        ns *= R;
    }
    if (table.empty()) table.push_back(make_float2(1.0f, 0.0f));  // keep the table pointer valid
    return true;
}

namespace {

// Per-thread element budget of an axis of n points with `lanes` threads per
// transform (the instantiated budgets; 0: none fits).
int e_bucket(uint32_t n, uint32_t lanes) {
    const uint32_t per = (n + lanes - 1) / lanes;
    if (lanes == kLanes) {
        for (int e : {4, 8, 12, 16, 20, 24, 32})
            if (per <= static_cast<uint32_t>(e)) return e;
    } else {
        for (int e : {16, 20, 24, 32})
            if (per <= static_cast<uint32_t>(e)) return e;
    }
    return 0;
}

// rows tile: 4 transforms; column tile: 2^logb interleaved transforms
size_t row_lds(uint32_t n) { return static_cast<size_t>(kPairs) * n * sizeof(float2); }
size_t col_lds(uint32_t n, uint32_t logb) { return (static_cast<size_t>(1) << logb) * n * sizeof(float2); }

template <class K> void allow_lds(K kernel, size_t bytes) {
    if (bytes > 64 * 1024)
        CHIP_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

template <int V> using IC = std::integral_constant<int, V>;

// f(E, LANES) with the axis's element budget and lanes per transform as
// compile-time constants: the instantiated combinations only.
template <class F> void dispatch_axis(int e, uint32_t lanes, F f) {
    if (lanes == kLanes) {
        switch (e) {
            case 4: f(IC<4>{}, IC<kLanes>{}); return;
            case 8: f(IC<8>{}, IC<kLanes>{}); return;
            case 12: f(IC<12>{}, IC<kLanes>{}); return;
            case 16: f(IC<16>{}, IC<kLanes>{}); return;
            case 20: f(IC<20>{}, IC<kLanes>{}); return;
            case 24: f(IC<24>{}, IC<kLanes>{}); return;
            case 32: f(IC<32>{}, IC<kLanes>{}); return;
            default: break;
        }
    } else if (lanes == kLongLanes) {
        switch (e) {
            case 16: f(IC<16>{}, IC<kLongLanes>{}); return;
            case 20: f(IC<20>{}, IC<kLongLanes>{}); return;
            case 24: f(IC<24>{}, IC<kLongLanes>{}); return;
            case 32: f(IC<32>{}, IC<kLongLanes>{}); return;
            default: break;
        }
    }
    fail(CHIP_ERR_INVALID, "transform too long for one workgroup");
}

#define CHIP_XF_GO(KEXPR, GDIM, NT, LDS, ST, ...)                                                 \
    do {                                                                                           \
        allow_lds(KEXPR, LDS);                                                                     \
        hipLaunchKernelGGL(KEXPR, GDIM, dim3(NT), LDS, ST, __VA_ARGS__);                           \
        CHIP_HIP(hipGetLastError());                                                               \
    } while (0)

}  // namespace

bool plan_transforms(const chip_gridder &g, uint64_t n_chan, size_t lds_limit, XPlan &p) {
    if (g.gx % 2 || g.gy % 2) return false;
    if (n_chan == 0 || n_chan > 65535 || g.gx > 0xffffffu || g.gy > 0xffffffu) return false;
    std::vector<float2> tx, ty;
    if (!factor_axis(static_cast<uint32_t>(g.gx), p.ax, tx) || !factor_axis(static_cast<uint32_t>(g.gy), p.ay, ty))
        return false;
    // axes up to 1536 points (at most 24 elements per thread): 64 threads per
    // transform, 8 columns per column workgroup; longer axes, up to 4096: 128
    // threads, 4 columns (2048 points: 1.83 against 2.20 ms for the forward
    // and inverse transforms of 64 planes; 1250 points: 2.86 against 2.25 ms
    // for 256 planes, so those keep 64)
    p.lanes_x = p.ax.n > kLanes * 24 ? kLongLanes : kLanes;
    p.lanes_y = p.ay.n > kLanes * 24 ? kLongLanes : kLanes;
    p.logb_x = p.lanes_x == kLanes ? kLogBlock : kLongLogBlock;
    p.e_x = e_bucket(p.ax.n, p.lanes_x);
    p.e_y = e_bucket(p.ay.n, p.lanes_y);
    if (!p.e_x || !p.e_y || col_lds(p.ax.n, p.logb_x) > lds_limit || row_lds(p.ay.n) > lds_limit) return false;
    XGeom &xg = p.xg;
    xg.nx = static_cast<uint32_t>(g.geom.nx);
    xg.ny = static_cast<uint32_t>(g.geom.ny);
    xg.gx = static_cast<uint32_t>(g.gx);
    xg.gy = static_cast<uint32_t>(g.gy);
    xg.blc_x = static_cast<uint32_t>(g.blc_x);
    xg.blc_y = static_cast<uint32_t>(g.blc_y);
    xg.base = static_cast<uint32_t>(g.gx / 2 + g.gy / 2);
    xg.wproject = g.wproject ? 1u : 0u;
    xg.nxp = (xg.nx + kRowsPerGroup - 1) / kRowsPerGroup * kRowsPerGroup;
    xg.nyb_h = (xg.gy / 2 + 1 + kBlock - 1) / kBlock;
    // the half spectra borrow the grid buffers
    if (xg.nxp > 2 * xg.gx) return false;
    xg.gstride = 2ull * xg.nyb_h * xg.gx * kBlock;
    xg.hstride = static_cast<uint64_t>(xg.nyb_h) * xg.nxp * kBlock;
    xg.corr_x = g.d_corr_x.as<float>();
    xg.corr_y = g.d_corr_y.as<float>();
    xg.sinc = g.wproject ? g.d_sinc.as<float>() : nullptr;
    xg.sinc_x = g.wproject ? g.d_sinc_x.as<float>() : nullptr;
    xg.sinc_y = g.wproject ? g.d_sinc_y.as<float>() : nullptr;
    xg.touch = nullptr;
    xg.touch_blk = nullptr;
    xg.nxc = (xg.gx + kTouchRows - 1) / kTouchRows;
    if (xg.nxc > 64 * kRowWords || xg.nyb_h > 64 * kBlockWords) return false;
    // the inverse rows pass loads its half-spectrum elements in E / 2 + 1 steps
    const uint64_t rows_nt = static_cast<uint64_t>(kPairs) * p.lanes_y;
    if ((static_cast<uint64_t>(xg.nyb_h) * kPairs * kBlock + rows_nt - 1) / rows_nt >
        static_cast<uint64_t>(p.e_y / 2 + 1))
        return false;
    p.n_chan = static_cast<uint32_t>(n_chan);
    p.tw_x.alloc(tx.size() * sizeof(float2));
    p.tw_y.alloc(ty.size() * sizeof(float2));
    CHIP_HIP(hipMemcpy(p.tw_x.p, tx.data(), p.tw_x.bytes, hipMemcpyHostToDevice));
    CHIP_HIP(hipMemcpy(p.tw_y.p, ty.data(), p.tw_y.bytes, hipMemcpyHostToDevice));
    p.ax.tw = p.tw_x.as<float2>();
    p.ay.tw = p.tw_y.as<float2>();
    return true;
}

void forward_model(const XPlan &p, const float *model, float2 *half, float2 *grid, hipStream_t st,
                   hipEvent_t between) {
    const dim3 rows(p.xg.nxp / kRowsPerGroup, p.n_chan), cols(p.xg.nyb_h * (kBlock >> p.logb_x), p.n_chan);
    const size_t rl = row_lds(p.ay.n), cl = col_lds(p.ax.n, p.logb_x);
    dispatch_axis(p.e_y, p.lanes_y, [&](auto E, auto L) {
        constexpr int e = decltype(E)::value, l = decltype(L)::value;
        CHIP_XF_GO((k_fwd_rows<e, l>), rows, kPairs * l, rl, st, p.xg, p.ay, model, half);
    });
    if (between) CHIP_HIP(hipEventRecord(between, st));
    dispatch_axis(p.e_x, p.lanes_x, [&](auto E, auto L) {
        constexpr int e = decltype(E)::value, l = decltype(L)::value;
        constexpr int b = l == kLanes ? kLogBlock : kLongLogBlock;
        CHIP_XF_GO((k_fwd_cols<e, l, b>), cols, l << b, cl, st, p.xg, p.ax, half, grid);
    });
}

void inverse_residual(const XPlan &p, const float2 *grid, float2 *rows_buf, const float *nsum,
                      const float *peak, float *out, hipStream_t st, hipEvent_t between) {
    const dim3 rows(p.xg.nxp / kRowsPerGroup, p.n_chan), cols(p.xg.nyb_h * (kBlock >> p.logb_x), p.n_chan);
    const size_t rl = row_lds(p.ay.n), cl = col_lds(p.ax.n, p.logb_x);
    dispatch_axis(p.e_x, p.lanes_x, [&](auto E, auto L) {
        constexpr int e = decltype(E)::value, l = decltype(L)::value;
        constexpr int b = l == kLanes ? kLogBlock : kLongLogBlock;
        CHIP_XF_GO((k_inv_cols<e, l, b>), cols, l << b, cl, st, p.xg, p.ax, grid, rows_buf);
    });
    if (between) CHIP_HIP(hipEventRecord(between, st));
    dispatch_axis(p.e_y, p.lanes_y, [&](auto E, auto L) {
        constexpr int e = decltype(E)::value, l = decltype(L)::value;
        CHIP_XF_GO((k_inv_rows<e, l>), rows, kPairs * l, rl, st, p.xg, p.ay, rows_buf, nsum, peak, out);
    });
}

void zero_touched(const XPlan &p, float2 *grid, hipStream_t st) {
    hipLaunchKernelGGL(k_zero_touched, dim3(p.xg.nyb_h, p.n_chan), dim3(256), 0, st, p.xg,
                       reinterpret_cast<float4 *>(grid));
    CHIP_HIP(hipGetLastError());
}

}  // namespace xf
}  // namespace chip
