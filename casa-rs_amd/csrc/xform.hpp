// Fused image <-> uv-grid transforms of the major cycle (xform.hip).
#pragma once

#include "internal.hpp"

namespace chip {
namespace xf {

constexpr uint32_t kMaxStages = 16;

// One transform axis: length n = product of radix[0..n_stages) (2, 3, 4, 5).
struct Axis {
    uint32_t n;
    uint32_t n_stages;
    uint32_t radix[kMaxStages];
    const float2 *tw;  // n twiddles exp(-2 pi i m / n)
};

struct XGeom {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t base;        // gx/2 + gy/2
    uint32_t wproject;
    uint32_t half_pitch;  // row pitch of the half spectrum (>= gy/2 + 1)
    uint32_t logc;        // log2 of the columns per column-pass workgroup
    uint64_t cstride;     // gx * gy
    const float *corr_x, *corr_y;
    const float *sinc;            // W products: sinc of max(gx, gy) on both axes
    const float *sinc_x, *sinc_y; // W model apodisation: per-axis sinc
};

struct XPlan {
    Axis ax{}, ay{};
    XGeom xg{};
    uint32_t n_chan = 0;
    uint32_t logc = 0;
    int row_e = 0, col_e = 0;
    DevBuf tw_x, tw_y;
};

bool factor_axis(uint32_t n, Axis &ax);
std::vector<float2> twiddles(uint32_t n);

// False when the grid does not factor into 2, 3, 5 or a transform does not
// fit the LDS limit; logc_max bounds the column block width (2^logc_max).
bool plan_transforms(const chip_gridder &g, uint64_t n_chan, size_t lds_limit, uint32_t logc_max,
                     XPlan &p);

// model [chan][nx][ny] -> half spectrum (n_chan x nx x half_pitch) -> grid
// [chan][gx][gy], the raw (unshifted) FFT of the checkerboard-signed,
// apodised, zero-padded model.
void forward_model(const XPlan &p, const float *model, float2 *half, float2 *grid, hipStream_t st);

// residual grid [chan][gx][gy] -> rows (n_chan x nx x gy) -> residual image
// [chan][nx][ny], corrected and normalised (nsum, peak: n_chan device floats).
void inverse_residual(const XPlan &p, const float2 *grid, float2 *rows, const float *nsum,
                      const float *peak, float *out, hipStream_t st);

}  // namespace xf
}  // namespace chip
