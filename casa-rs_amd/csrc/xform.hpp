// Fused image <-> uv-grid transforms of the major cycle (xform.hip).
//
// Grid layout of the major-cycle workspace ("column blocked"): cell (x, y)
// of channel c is element
//     c * gstride + (y / 8) * (gx * 8) + x * 8 + (y % 8),
// i.e. blocks of 8 adjacent y columns, each block stored x-major.  A column
// FFT over one block then reads and writes one contiguous run, and a rows
// pass over 8 adjacent x rows moves 8 x 8 contiguous values per y block.
#pragma once

#include "internal.hpp"

namespace chip {
namespace xf {

constexpr uint32_t kMaxStages = 16;
constexpr uint32_t kBlock = 8;     // y columns per block = transforms per workgroup
constexpr uint32_t kLogBlock = 3;
constexpr uint32_t kThreads = 1024;
constexpr uint32_t kLanes = kThreads / kBlock;  // threads per transform

// One transform axis: length n = product of radix[0..n_stages) (2, 3, 4, 5),
// with the per-stage constants of the Stockham recursion.
struct Axis {
    uint32_t n;
    uint32_t n_stages;
    uint32_t radix[kMaxStages];
    uint32_t ns[kMaxStages];      // product of the radices before stage s
    uint32_t step[kMaxStages];    // n / (ns * radix): twiddle index stride
    float inv_ns[kMaxStages];     // 1 / ns, for the exact float quotient j / ns
    const float2 *tw;             // n twiddles exp(-2 pi i m / n)
};

struct XGeom {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t base;      // gx/2 + gy/2
    uint32_t wproject;
    uint32_t nxp;       // nx rounded up to a multiple of 8
    uint32_t nyb;       // ceil(gy / 8): y blocks of a grid
    uint32_t nyb_h;     // ceil((gy/2 + 1) / 8): y blocks of the half spectrum
    uint32_t _pad;
    uint64_t gstride;   // elements per channel of a blocked grid: nyb * gx * 8
    uint64_t hstride;   // elements per channel of the half spectrum / inverse rows: nyb * nxp * 8
    const float *corr_x, *corr_y;
    const float *sinc;            // W products: sinc of max(gx, gy) on both axes
    const float *sinc_x, *sinc_y; // W model apodisation: per-axis sinc
};

struct XPlan {
    Axis ax{}, ay{};
    XGeom xg{};
    uint32_t n_chan = 0;
    int e_x = 0, e_y = 0;  // per-thread element budgets of the x and y transforms
    DevBuf tw_x, tw_y;
};

bool factor_axis(uint32_t n, Axis &ax);
std::vector<float2> twiddles(uint32_t n);

// False when a grid length does not factor into 2, 3, 5, a transform does
// not fit the LDS limit or the per-thread register budget, or the image
// rows do not fit the workspace.
bool plan_transforms(const chip_gridder &g, uint64_t n_chan, size_t lds_limit, XPlan &p);

// Elements of one blocked grid cube (all channels).
inline uint64_t grid_elements(const XPlan &p) { return static_cast<uint64_t>(p.n_chan) * p.xg.gstride; }

// model [chan][nx][ny] -> half spectrum (in `half`) -> blocked grid: the raw
// (unshifted) FFT of the checkerboard-signed, apodised, zero-padded model.
void forward_model(const XPlan &p, const float *model, float2 *half, float2 *grid, hipStream_t st);

// blocked residual grid -> inverse rows (in `rows`) -> residual image
// [chan][nx][ny], corrected and normalised (nsum, peak: n_chan device floats).
void inverse_residual(const XPlan &p, const float2 *grid, float2 *rows, const float *nsum,
                      const float *peak, float *out, hipStream_t st);

}  // namespace xf
}  // namespace chip
