// Host-side tables of the gridding path: spheroidal function, padding,
// correction axes, sinc axis, 7-tap weights, W plane count.  These run once
// per gridder on the CPU (a few thousand values); the per-sample work is on
// the GPU.  Each formula keeps the reference's precision and operation order
// because the tap weights and corrections enter the results bit for bit.
#include <array>
#include <cmath>

#include "internal.hpp"

// Every multiply and add is rounded separately, as in the reference.
#pragma clang fp contract(off)

namespace chip {

// Schwab's rational approximation of the m=6, alpha=1 prolate spheroidal
// function (gridder.rs:4230-4268).  Powers are formed by repeated squaring,
// the evaluation order of integer powers in the reference build.
double grdsf(double nu) {
    static const double P[2][5] = {
        {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
        {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
    static const double Q[2][3] = {{1.0, 8.212018e-1, 2.078043e-1},
                                   {1.0, 9.599102e-1, 2.918724e-1}};
    if (!(nu >= 0.0 && nu <= 1.0)) return 0.0;
    const int part = nu < 0.75 ? 0 : 1;
    const double end = part == 0 ? 0.75 : 1.0;
    const double d = nu * nu - end * end;
    const double d2 = d * d;
    const double powers[5] = {1.0, d, d2, d * d2, d2 * d2};
    double num = 0.0, den = 0.0;
    for (int k = 0; k < 5; ++k) num = num + P[part][k] * powers[k];
    for (int k = 0; k < 3; ++k) den = den + Q[part][k] * powers[k];
    return den == 0.0 ? 0.0 : num / den;
}

// (1 - nu^2) psi(nu), nu = distance / support (gridder.rs:4200-4209).
float spheroidal_kernel(double distance, double support) {
    if (!(std::isfinite(distance) && distance <= support)) return 0.0f;
    const double nu = distance / support;
    if (nu > 1.0) return 0.0f;
    return static_cast<float>((1.0 - nu * nu) * grdsf(nu));
}

// gridder.rs:4152-4156
uint64_t padded_len(uint64_t n, double padding) {
    const double f = std::floor(padding * static_cast<double>(n) - 0.5);
    uint64_t p = f <= 0.0 ? 0 : static_cast<uint64_t>(f);
    if (p < n) p = n;
    return p % 2 == 0 ? p : p + 1;
}

// gridder.rs:4166-4173
bool is_composite(uint64_t n) {
    for (uint64_t f : {2ull, 3ull, 5ull})
        while (n > 1 && n % f == 0) n /= f;
    return n == 1;
}

// gridder.rs:4158-4164
uint64_t composite_padded_len(uint64_t n, double padding) {
    uint64_t p = padded_len(n, padding);
    while (!is_composite(p)) p += 2;
    return p;
}

static double axis_nu(uint64_t i, uint64_t n) {
    const double c = static_cast<double>(n) / 2.0;
    double nu = std::fabs(static_cast<double>(i) - c) / c;
    return nu < 0.0 ? 0.0 : (nu > 1.0 ? 1.0 : nu);
}

// 1/psi or 0 (gridder.rs:4175-4188)
std::vector<float> correction_axis(uint64_t n) {
    std::vector<float> out(n);
    for (uint64_t i = 0; i < n; ++i) {
        const double v = grdsf(axis_nu(i, n));
        out[i] = v > 1.0e-6 ? static_cast<float>(1.0 / v) : 0.0f;
    }
    return out;
}

// psi (gridder.rs:4190-4198)
std::vector<float> gridder_correction_axis(uint64_t n) {
    std::vector<float> out(n);
    for (uint64_t i = 0; i < n; ++i) out[i] = static_cast<float>(grdsf(axis_nu(i, n)));
    return out;
}

// Float sinc of the sampling offset (gridder.rs:4211-4228).
std::vector<float> sinc_axis(uint64_t n, uint64_t sampling) {
    std::vector<float> out(n);
    const int64_t half = static_cast<int64_t>(n / 2);
    for (uint64_t i = 0; i < n; ++i) {
        const float offset = static_cast<float>(static_cast<int64_t>(i) - half);
        const float den = static_cast<float>(n) * static_cast<float>(sampling);
        const float arg = static_cast<float>(3.14159265358979323846 * static_cast<double>(offset) /
                                             static_cast<double>(den));
        out[i] = i == n / 2 ? 1.0f : std::sin(arg) / arg;
    }
    return out;
}

// Spheroid table (gridder.rs:407-416) folded into per-offset 7-tap weights
// normalised to sum 1 (gridder.rs:2026-2051).  Row r is offset r - os/2.
std::vector<float> tap_weights(uint32_t os) {
    const uint64_t len = static_cast<uint64_t>(os) * (kSupport + 1);
    std::vector<float> table(len, 0.0f);
    for (uint64_t i = 0; i < static_cast<uint64_t>(os) * kSupport; ++i) {
        const double d = static_cast<double>(i) / (static_cast<double>(kSupport) * os);
        table[i] = spheroidal_kernel(d * kSupport, static_cast<double>(kSupport));
    }
    const int64_t half = os / 2;
    std::vector<float> rows;
    rows.reserve((2 * half + 1) * kTaps);
    for (int64_t off = -half; off <= half; ++off) {
        float w[kTaps];
        float norm = 0.0f;
        for (int t = 0; t < kTaps; ++t) {
            const int64_t lookup = std::llabs(static_cast<int64_t>(t - kSupport) * os + off);
            w[t] = static_cast<uint64_t>(lookup) < len ? table[lookup] : 0.0f;
            norm += w[t];
        }
        if (norm > 0.0f)
            for (float &x : w) x /= norm;
        rows.insert(rows.end(), w, w + kTaps);
    }
    return rows;
}

// The same per-offset 7-tap rows before their normalisation: the weights
// sample_taps_unnormalized looks up (gridder.rs:1997-2023).  Row r is offset
// r - os/2.
std::vector<float> tap_weights_raw(uint32_t os) {
    const uint64_t len = static_cast<uint64_t>(os) * (kSupport + 1);
    std::vector<float> table(len, 0.0f);
    for (uint64_t i = 0; i < static_cast<uint64_t>(os) * kSupport; ++i) {
        const double d = static_cast<double>(i) / (static_cast<double>(kSupport) * os);
        table[i] = spheroidal_kernel(d * kSupport, static_cast<double>(kSupport));
    }
    const int64_t half = os / 2;
    std::vector<float> rows;
    rows.reserve((2 * half + 1) * kTaps);
    for (int64_t off = -half; off <= half; ++off)
        for (int t = 0; t < kTaps; ++t) {
            const int64_t lookup = std::llabs(static_cast<int64_t>(t - kSupport) * os + off);
            rows.push_back(static_cast<uint64_t>(lookup) < len ? table[lookup] : 0.0f);
        }
    return rows;
}

// gridder.rs:3822-3830
uint32_t suggested_w_plane_count(const chip_geometry &g, double max_abs_w) {
    if (!(std::isfinite(max_abs_w) && max_abs_w > 0.0)) return 1;
    const double half = std::max(g.nx * std::fabs(g.cell_x), g.ny * std::fabs(g.cell_y)) / 2.0;
    const double v = 1.05 * max_abs_w * std::fabs(std::sin(half));
    return v < 1.0 ? 0u : static_cast<uint32_t>(v);
}

// gridder.rs:3808-3820
double w_max_uvw(const chip_geometry &g, double max_abs_w, bool explicit_planes) {
    return explicit_planes ? 0.25 / std::fabs(g.cell_x) : 1.05 * max_abs_w;
}

// find_w_project_support_from_transformed (gridder.rs:3832-3855).
// axis_mag_over_peak[c] = max(|T[c,0]/peak|, |T[0,c]/peak|) for c < conv/2 - 1.
uint32_t find_w_support(const std::vector<float> &mag, uint32_t conv, uint32_t sampling) {
    const uint32_t quarter = conv / 2 - 1;
    int64_t trial = -1;
    for (uint32_t c = quarter - 1; c >= 1; --c) {
        if (mag[c] > 1.0e-3f) { trial = c; break; }
    }
    const int64_t max_support = static_cast<int64_t>(quarter / sampling) - 1;
    const int64_t cap = std::max<int64_t>(max_support, 1);
    if (trial < 0) return static_cast<uint32_t>(cap);
    const float q = static_cast<float>(trial) / static_cast<float>(sampling) + 0.5f;
    const int64_t sup = static_cast<int64_t>(std::floor(q)) + 1;
    return static_cast<uint32_t>(std::min(sup, cap));
}



// ---- phase-centre rotation (casars-imager/src/lib.rs:48030-48159) ----

namespace {

using M3 = std::array<std::array<double, 3>, 3>;

M3 mat3_mul(const M3 &a, const M3 &b) {
    M3 t{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) t[r][c] = a[r][0] * b[0][c] + a[r][1] * b[1][c] + a[r][2] * b[2][c];
    return t;
}

// gridft_axis_rotation: about X (axis 0) or Z (axis 1)
M3 axis_rotation(double angle, int axis) {
    const double s = std::sin(angle), c = std::cos(angle);
    if (axis == 0) return M3{{{1.0, 0.0, 0.0}, {0.0, c, -s}, {0.0, s, c}}};
    return M3{{{c, -s, 0.0}, {s, c, 0.0}, {0.0, 0.0, 1.0}}};
}

// gridft_euler_rotation of two operations from the identity
M3 euler2(double a1, int ax1, double a2, int ax2) {
    M3 m{{{1.0, 0.0, 0.0}, {0.0, 1.0, 0.0}, {0.0, 0.0, 1.0}}};
    m = mat3_mul(m, axis_rotation(a1, ax1));
    return mat3_mul(m, axis_rotation(a2, ax2));
}

constexpr double kHalfPi = 1.57079632679489661923;

M3 target_rotation(const double tgt[2]) { return euler2(kHalfPi - tgt[1], 0, -(tgt[0] - kHalfPi), 1); }

}  // namespace

void phase_rotation(const double src[2], const double tgt[2], double uvrot[9], double phrot[3]) {
    const M3 rot1 = euler2(-(kHalfPi - src[0]), 1, src[1] - kHalfPi, 0);
    const M3 rot3 = target_rotation(tgt);
    const M3 p = mat3_mul(rot3, rot1);
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) uvrot[3 * r + c] = p[c][r];  // transpose
    const double tc[3] = {std::cos(tgt[1]) * std::cos(tgt[0]), std::cos(tgt[1]) * std::sin(tgt[0]),
                          std::sin(tgt[1])};
    const double sc[3] = {std::cos(src[1]) * std::cos(src[0]), std::cos(src[1]) * std::sin(src[0]),
                          std::sin(src[1])};
    const double d[3] = {tc[0] - sc[0], tc[1] - sc[1], tc[2] - sc[2]};
    for (int r = 0; r < 3; ++r) phrot[r] = rot3[r][0] * d[0] + rot3[r][1] * d[1] + rot3[r][2] * d[2];
}

}  // namespace chip
