// Fused image <-> uv-grid transforms of the major cycle (xform.hip).
//
// Grid layout of the major-cycle workspace ("mirror-paired blocks"): with
// hy = gy/2, column y <= hy lives in upper block y / 8 at lane y % 8, and
// column y > hy is stored by its mirror index m = gy - y in lower block m / 8
// at lane m % 8; each block holds its 8 lanes for every x, x-major:
//     c * gstride + block * (gx * 8) + x * 8 + lane,
// lower blocks following the nyb_h upper ones.  A column and its Hermitian
// partner -y then sit at the same lane of paired blocks, so the forward pass
// writes a transformed column and its mirror contiguously, and the inverse
// pass forms the Hermitian part of a residual column from two contiguous
// reads.
#pragma once

#include "internal.hpp"

namespace chip {
namespace xf {

constexpr uint32_t kMaxStages = 16;
constexpr uint32_t kBlock = 8;     // y columns per block = transforms per workgroup
constexpr uint32_t kLogBlock = 3;
constexpr uint32_t kThreads = 512;               // column-pass workgroup: 8 transforms
constexpr uint32_t kLanes = kThreads / kBlock;  // threads per transform
constexpr uint32_t kRowsPerGroup = 8;            // rows-pass workgroup: 8 rows as 4 transforms
constexpr uint32_t kPairs = kRowsPerGroup / 2;
constexpr uint32_t kRowThreads = kPairs * kLanes;
constexpr uint32_t kTouchRows = 16;              // x rows per chunk of the uv-coverage mask
constexpr uint32_t kRowWords = 4;                // 64-bit words of a block's row-chunk mask: gx <= 4096
constexpr uint32_t kBlockWords = 5;              // 64-bit words of a channel's block mask: nyb_h <= 320
// Axes longer than 2048 points (up to 4096): 128 threads per transform, and
// column workgroups of 4 columns (half an 8-column block) so that the tile
// fits the LDS.
constexpr uint32_t kLongLanes = 128;
constexpr uint32_t kLongLogBlock = 2;

// One transform axis: length n = product over the stages of r1[s] * r2[s]
// (r1 in 2..5, r2 in 1..5), with the per-stage constants of the Stockham
// recursion.  A stage of radix r1 * r2 runs two radix passes in registers.
struct Axis {
    uint32_t n;
    uint32_t n_stages;
    uint32_t r1[kMaxStages];
    uint32_t r2[kMaxStages];
    uint32_t ns[kMaxStages];      // product of the radices before stage s
    uint32_t nb[kMaxStages];      // butterflies of stage s: n / (r1 r2)
    uint32_t tw_off[kMaxStages];  // first twiddle of stage s in tw
    float inv_ns[kMaxStages];     // 1 / ns, for the exact float quotient j / ns
    // Per stage with ns > 1 and W = exp(-2 pi i / (ns r1 r2)), for each k < ns:
    // W^(k r1 t2) for t2 = 1 .. r2-1, then W^(k t1) for t1 = 1 .. r1-1, each
    // rounded once from FP64.
    const float2 *tw;
};

struct XGeom {
    uint32_t nx, ny, gx, gy, blc_x, blc_y;
    uint32_t base;      // gx/2 + gy/2
    uint32_t wproject;
    uint32_t nxp;       // nx rounded up to a multiple of 8
    uint32_t nyb_h;     // ceil((gy/2 + 1) / 8): upper (= lower) blocks of a grid
    uint64_t gstride;   // elements per channel of a paired grid: 2 * nyb_h * gx * 8
    uint64_t hstride;   // elements per channel of a half spectrum [nyb_h][nxp][8]
    const float *corr_x, *corr_y;
    const float *sinc;            // W products: sinc of max(gx, gy) on both axes
    const float *sinc_x, *sinc_y; // W model apodisation: per-axis sinc
    // uv coverage: touch[(c * nyb_h + yb) * nxc + x / kTouchRows] != 0 when the
    // binned kernel may read the model or add to the residual in rows
    // [x, x + kTouchRows) of upper block yb; nullptr = every row.  The column
    // passes store (forward) and load (inverse) only those rows: the model
    // grid elsewhere is never read, the residual grid elsewhere is zero.
    // touch_blk[c * nyb_h + yb] != 0 when any row of block yb is covered: the
    // passes skip the other blocks (their columns are never read, their
    // residual is zero).
    const uint8_t *touch, *touch_blk;
    uint32_t nxc;       // ceil(gx / kTouchRows) <= 128
};

struct XPlan {
    Axis ax{}, ay{};
    XGeom xg{};
    uint32_t n_chan = 0;
    int e_x = 0, e_y = 0;  // per-thread element budgets of the x and y transforms
    uint32_t lanes_x = kLanes, lanes_y = kLanes;  // threads per transform of each axis
    uint32_t logb_x = kLogBlock;                   // log2 columns per column-pass workgroup
    DevBuf tw_x, tw_y;
};

// Stages of a length-n axis and its twiddle table (false when n does not
// factor into 2, 3, 5 or has too many stages).
bool factor_axis(uint32_t n, Axis &ax, std::vector<float2> &table);

// False when a grid length does not factor into 2, 3, 5, a transform does
// not fit the LDS limit or the per-thread register budget, or the image
// rows do not fit the workspace.
bool plan_transforms(const chip_gridder &g, uint64_t n_chan, size_t lds_limit, XPlan &p);

// Elements of one blocked grid cube (all channels).
inline uint64_t grid_elements(const XPlan &p) { return static_cast<uint64_t>(p.n_chan) * p.xg.gstride; }

// model [chan][nx][ny] -> half spectrum (in `half`) -> paired grid: the raw
// (unshifted) FFT of the checkerboard-signed, apodised, zero-padded model.
// `between` (may be null) is recorded on st between the two passes.
void forward_model(const XPlan &p, const float *model, float2 *half, float2 *grid, hipStream_t st,
                   hipEvent_t between = nullptr);

// paired residual grid -> Hermitian half of the x-inverse (in `rows`) ->
// residual image [chan][nx][ny], corrected and normalised (nsum, peak: n_chan
// device floats).
void inverse_residual(const XPlan &p, const float2 *grid, float2 *rows, const float *nsum,
                      const float *peak, float *out, hipStream_t st,
                      hipEvent_t between = nullptr);

// Zeroes the upper blocks of a paired grid cube in the rows the coverage
// mask marks (every row without a mask).
void zero_touched(const XPlan &p, float2 *grid, hipStream_t st);

}  // namespace xf
}  // namespace chip
