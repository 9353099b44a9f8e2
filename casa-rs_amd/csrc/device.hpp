// Device-side helpers shared by the HIP translation units: the FP64 planner,
// sample loading and the packed stream format.
#pragma once

#include <hip/hip_runtime.h>

#include "internal.hpp"

namespace chip {

// Packed plan of one accepted sample (8 bytes).
//   x: bits 0-15  x anchor (standard: first tap column x_start; W: loc_x)
//      bits 16-31 y anchor (standard: y_start; W: loc_y)
//   y: bits 0-7   x selector (standard: tap row wix; W: off_x + 128)
//      bits 8-15  y selector
//      bits 16-26 W plane / tap table, low 11 bits
//      bit 27     conjugate kernel (w > 0)
//      bit 28     visibility finite (contributes to dirty grids)
//      bit 29     usable for residual refresh
//      bits 30-31 W plane / tap table, high 2 bits (tables up to 8191)
constexpr uint32_t kConjBit = 1u << 27;
constexpr uint32_t kVisOkBit = 1u << 28;
constexpr uint32_t kResidOkBit = 1u << 29;
constexpr uint32_t kMaxTables = 8191;

__host__ __device__ __forceinline__ uint32_t packed_table(uint32_t y) {
    return ((y >> 16) & 0x7ffu) | ((y >> 30) << 11);
}

struct Plan {
    int32_t x, y, sx, sy, plane, conj;
    float norm;
    double mag;  // AWProject: |grid_normalization|
};

// One work item: a run of groups (identical plans) of one (channel, tile).
struct WorkItem {
    uint32_t chan;   // channel (low 24 bits); W / AW: the largest half-support of the groups (top 8 bits)
    uint32_t tile;   // tile_x * tiles_y + tile_y
    uint32_t start;  // first group
    uint32_t count;  // groups
};

struct SampleIn {
    double u, v, w;
    float weight, sf;
    float2 vis;
    bool usable;  // gridable and not flagged
    uint32_t chan;
    uint64_t index, row;
};

__device__ __forceinline__ SampleIn load_sample(const Source &s, uint64_t i) {
    SampleIn r;
    const uint64_t row = s.n_chan == 1 ? i : i / s.n_chan;
    r.chan = s.n_chan == 1 ? 0u : static_cast<uint32_t>(i - row * s.n_chan);
    r.index = i;
    r.row = row;
    if (s.uvw_m) {
        const double scale = s.lambda_scale[r.chan];
        r.u = s.uvw_m[3 * row] * scale;
        r.v = s.uvw_m[3 * row + 1] * scale;
        r.w = s.uvw_m[3 * row + 2] * scale;
    } else {
        r.u = s.u[i];
        r.v = s.v[i];
        r.w = s.w[i];
    }
    r.weight = s.weight[i];
    r.sf = s.sumwt_factor ? s.sumwt_factor[i] : 1.0f;
    r.vis = s.vis[i];
    r.usable = (!s.gridable || s.gridable[i]) && (!s.flag || !s.flag[i]);
    return r;
}

__device__ __forceinline__ bool finite_f(float x) { return isfinite(x); }

// sample_tap_span (gridder.rs:1958-1974) for one axis.  Every step is a
// separately rounded IEEE operation (no contraction), with round() = half
// away from zero as f64::round.
__device__ __forceinline__ bool std_axis_span(double coord, uint64_t size, uint32_t os,
                                              int32_t &start, int32_t &widx) {
#pragma clang fp contract(off)
    if (!isfinite(coord)) return false;
    if (coord > 4.0e18 || coord < -4.0e18) return false;
    const double a = round(coord);
    const int64_t anchor = static_cast<int64_t>(a);
    const int64_t off = static_cast<int64_t>(round((static_cast<double>(anchor) - coord) *
                                                   static_cast<double>(os)));
    const int64_t s = anchor - kSupport, e = anchor + kSupport;
    if (s < 0 || e >= static_cast<int64_t>(size)) return false;
    const int64_t half = os / 2;
    const int64_t idx = off + half;
    if (idx < 0 || idx > 2 * half) return false;
    start = static_cast<int32_t>(s);
    widx = static_cast<int32_t>(idx);
    return true;
}

// Plan one sample (StandardGridder::plan_positive_taps, gridder.rs:557-566 or
// WProjector::plan_sample, gridder.rs:3588-3624).  Validity filters are
// applied by the caller.
__device__ __forceinline__ bool plan_sample(const PlanParams &p, double u, double v, double w,
                                            Plan &out) {
#pragma clang fp contract(off)
    if (!p.wproject) {
        // grid_coordinate_x / _y (gridder.rs:1877-1886)
        const double cx = u / p.du + p.half_gx;
        const double cy = -v / p.dv + p.half_gy;
        int32_t xs, xw, ys, yw;
        if (!std_axis_span(cx, p.gx, p.os, xs, xw)) return false;
        if (!std_axis_span(cy, p.gy, p.os, ys, yw)) return false;
        out.x = xs; out.y = ys; out.sx = xw; out.sy = yw;
        out.plane = 0; out.conj = 0; out.norm = 1.0f;
        return true;
    }
    const double pos_x = u / p.du + p.half_gx;
    const double pos_y = -v / p.dv + p.half_gy;
    if (!(isfinite(pos_x) && isfinite(pos_y) && isfinite(w))) return false;
    if (pos_x > 4.0e18 || pos_x < -4.0e18 || pos_y > 4.0e18 || pos_y < -4.0e18) return false;
    const int64_t lx = static_cast<int64_t>(round(pos_x));
    const int64_t ly = static_cast<int64_t>(round(pos_y));
    const double s = static_cast<double>(p.sampling);
    const int64_t ox = static_cast<int64_t>(round((static_cast<double>(lx) - pos_x) * s));
    const int64_t oy = static_cast<int64_t>(round((static_cast<double>(ly) - pos_y) * s));
    // w_plane_index (gridder.rs:3749-3757)
    uint32_t plane = 0;
    if (p.planes > 1) {
        double c = p.w_scale * fabs(w);
        c = sqrt(c < 0.0 ? 0.0 : c);
        double r = round(c);
        const double top = static_cast<double>(p.planes - 1);
        r = r < 0.0 ? 0.0 : (r > top ? top : r);
        plane = static_cast<uint32_t>(r);
    }
    const int64_t S = p.supports[plane];
    if (S <= 0 || lx - S < 0 || ly - S < 0 || lx + S >= static_cast<int64_t>(p.gx) ||
        ly + S >= static_cast<int64_t>(p.gy))
        return false;
    const int64_t R = p.norm_radius;
    // |off| <= sampling/2 <= R always holds because |loc - pos| <= 1/2.
    if (ox > R || ox < -R || oy > R || oy < -R) return false;
    const int64_t axis = 2 * R + 1;
    out.norm = p.norm_table[(plane * axis + (ox + R)) * axis + (oy + R)];
    out.x = static_cast<int32_t>(lx);
    out.y = static_cast<int32_t>(ly);
    out.sx = static_cast<int32_t>(ox);
    out.sy = static_cast<int32_t>(oy);
    out.plane = static_cast<int32_t>(plane);
    out.conj = w > 0.0 ? 1 : 0;
    return true;
}

// awproject_sample_geometry (gridder.rs:2699-2748): anchor, sub-cell offset,
// grid and cell bounds.  Returns a CHIP_AW_* code.
__device__ __forceinline__ int aw_geometry(const PlanParams &p, const AwCell &c, double u, double v, double w,
                                           int64_t &lx, int64_t &ly, int64_t &ox, int64_t &oy) {
#pragma clang fp contract(off)
    const double pos_x = u / p.du + p.half_gx;
    const double pos_y = -v / p.dv + p.half_gy;
    if (!(isfinite(pos_x) && isfinite(pos_y) && isfinite(w))) return CHIP_AW_NON_FINITE_COORDINATE;
    if (pos_x > 4.0e18 || pos_x < -4.0e18 || pos_y > 4.0e18 || pos_y < -4.0e18) return CHIP_AW_OUTSIDE_GRID;
    lx = static_cast<int64_t>(round(pos_x));
    ly = static_cast<int64_t>(round(pos_y));
    const double s = static_cast<double>(p.sampling);
    ox = static_cast<int64_t>(round((static_cast<double>(lx) - pos_x) * s));
    oy = static_cast<int64_t>(round((static_cast<double>(ly) - pos_y) * s));
    if (lx - c.xs < 0 || ly - c.ys < 0 || lx + c.xs >= static_cast<int64_t>(p.gx) ||
        ly + c.ys >= static_cast<int64_t>(p.gy))
        return CHIP_AW_OUTSIDE_GRID;
    const int64_t si = p.sampling;
    if (c.cx - c.xs * si + ox < 0 || c.cy - c.ys * si + oy < 0 || c.cx + c.xs * si + ox >= c.shape_x ||
        c.cy + c.ys * si + oy >= c.shape_y)
        return CHIP_AW_KERNEL_INDEX_OUTSIDE_CELL;
    return CHIP_AW_OK;
}

// The sample's CF cell: the caller's selection, or the W plane
// round(sqrt(WIncr |w|)) clamped to the cells (awprojection.rs:1239-1246).
__device__ __forceinline__ uint32_t aw_cell_of(const PlanParams &p, const SampleIn &s) {
    if (p.cf_cell) return p.cf_cell[s.index];
    if (p.aw_cells <= 1 || p.aw_w_incr == 0.0) return 0u;
    double r = round(sqrt(p.aw_w_incr * fabs(s.w)));
    const double top = static_cast<double>(p.aw_cells - 1);
    if (!(r >= 0.0)) r = 0.0;
    if (r > top) r = top;
    return static_cast<uint32_t>(r);
}

// AWProject plan of one sample: geometry, then the (cell, offset)
// normalisation's validity (plan_sample_detailed, gridder.rs:2253-2311).
// The tap table folds in the pointing and the w sign, so conj stays 0.
__device__ __forceinline__ int aw_plan_sample(const PlanParams &p, const SampleIn &s, Plan &out) {
    const uint32_t cell = aw_cell_of(p, s);
    const uint32_t pt = p.pointing ? p.pointing[s.row] : 0u;
    if (cell >= p.aw_cells || pt >= p.aw_pointings) return CHIP_AW_NO_CELL;
    const AwCell c = p.aw_cell[cell];
    int64_t lx, ly, ox, oy;
    const int code = aw_geometry(p, c, s.u, s.v, s.w, lx, ly, ox, oy);
    if (code != CHIP_AW_OK) return code;
    const int64_t R = p.norm_radius, axis = 2 * R + 1;
    if (ox > R || ox < -R || oy > R || oy < -R) return CHIP_AW_KERNEL_INDEX_OUTSIDE_CELL;
    const double mag = p.aw_mag[(static_cast<int64_t>(cell) * axis + (ox + R)) * axis + (oy + R)];
    if (!(mag > 0.0)) return CHIP_AW_INVALID_NORMALIZATION;
    out.x = static_cast<int32_t>(lx);
    out.y = static_cast<int32_t>(ly);
    out.sx = static_cast<int32_t>(ox);
    out.sy = static_cast<int32_t>(oy);
    out.plane = static_cast<int32_t>((cell * p.aw_pointings + pt) * 2u + (s.w > 0.0 ? 1u : 0u));
    out.conj = 0;
    out.mag = mag;
    out.norm = static_cast<float>(mag);
    return CHIP_AW_OK;
}

// Validity filters ahead of planning.
//  standard: execution.rs:18270-18300 (gridable, weight and sumwt_factor
//            finite and positive);
//  W:        raw_w_project_sample_is_valid, lib.rs:73666-73676.
__device__ __forceinline__ bool sample_accepted(const PlanParams &p, const SampleIn &s,
                                                Plan &plan) {
    if (!s.usable) return false;
    if (!(finite_f(s.weight) && s.weight > 0.0f && finite_f(s.sf) && s.sf > 0.0f)) return false;
    if (p.wproject) {
        if (!(finite_f(s.vis.x) && finite_f(s.vis.y) && isfinite(s.u) && isfinite(s.v) &&
              isfinite(s.w)))
            return false;
    }
    if (p.aw) return aw_plan_sample(p, s, plan) == CHIP_AW_OK;
    return plan_sample(p, s.u, s.v, s.w, plan);
}

__device__ __forceinline__ uint2 pack_plan(const PlanParams &p, const Plan &pl, uint32_t flags) {
    uint2 r;
    r.x = static_cast<uint32_t>(pl.x) | (static_cast<uint32_t>(pl.y) << 16);
    uint32_t sx, sy;
    if (p.wproject) {
        sx = static_cast<uint32_t>(pl.sx + 128);
        sy = static_cast<uint32_t>(pl.sy + 128);
    } else {
        sx = static_cast<uint32_t>(pl.sx);
        sy = static_cast<uint32_t>(pl.sy);
    }
    const uint32_t table = static_cast<uint32_t>(pl.plane);
    r.y = (sx & 0xffu) | ((sy & 0xffu) << 8) | ((table & 0x7ffu) << 16) | ((table >> 11) << 30) |
          (pl.conj ? kConjBit : 0u) | flags;
    return r;
}

// Weight a sample carries into the grids: weight * sumwt_factor (f32) for the
// standard gridder (execution.rs:18306), weight alone for W-projection
// (lib.rs:71499-71505).
__device__ __forceinline__ float grid_weight(const PlanParams &p, const SampleIn &s) {
    return p.wproject ? s.weight : s.weight * s.sf;
}

__device__ __forceinline__ uint32_t sample_flags(const PlanParams &p, const SampleIn &s) {
    const bool vis_ok = finite_f(s.vis.x) && finite_f(s.vis.y);
    uint32_t f = vis_ok ? kVisOkBit : 0u;
    const float gw = grid_weight(p, s);
    if (vis_ok && finite_f(gw) && gw > 0.0f) f |= kResidOkBit;
    return f;
}

enum : int { kModePsf = 0, kModeDirty = 1, kModeResidual = 2, kModeAdjoint = 3, kModeDegrid = 4 };

// Kernel-side view of a gridder.
struct GridParams {
    uint32_t gx, gy;
    uint32_t tile, halo;
    uint32_t L;              // LDS tile edge: tile + 2*halo
    uint32_t Ls;             // LDS row stride (L rounded up to odd, fewer bank conflicts)
    uint32_t tiles_y;
    uint32_t sampling, width;
    uint32_t tap_rows;       // standard tap-table rows (os + 1)
    uint32_t n_off;          // W offsets per axis, 2R + 1
    int32_t norm_radius;     // R
    uint32_t tap_stride;     // W compact taps per (plane, offset): (2*max_support+1)^2
    uint64_t cells;          // gx * gy
    // Grid addressing: cell (chan, x, y) is at chan*cstride + cell_offset(x, y):
    // public layout x*xstride + y (cstride = gx*gy, xstride = gy); the
    // major-cycle workspace uses mirror-paired blocks (paired = 1): column
    // y <= hy at (y/8)*bstride + 8x + y%8, column y > hy by m = gy - y at
    // lower + (m/8)*bstride + 8x + m%8, with bstride = 8*gx.
    uint64_t cstride;
    uint64_t xstride;
    uint64_t bstride;
    uint64_t lower;
    uint32_t paired;
    uint32_t hy;             // gy / 2
    // Major cycle: the model grid is the raw forward transform of the
    // checkerboarded model, so (-1)^(x + y + sign_base) is applied on load.
    uint32_t model_signed;
    uint32_t sign_base;      // gx/2 + gy/2
    // Major cycle: only the columns y <= hy of the (Hermitian) model grid are
    // stored; G[x, y > hy] = conj G[(gx - x) mod gx, gy - y].
    uint32_t model_half;
    // Major cycle: the residual is accumulated as 2H, H = (G[k] + conj G[-k]) / 2
    // its Hermitian part, in the columns y <= hy only: a cell (x, y > hy) adds
    // conj v to (-x, gy - y); a cell of column 0 or hy adds v to itself and
    // conj v to (-x, y).
    uint32_t acc_hermitian;
    const float *taps;       // (os+1) x 7
    const float2 *kernels;   // planes x width x width
    const float2 *compact;   // planes x n_off^2 x tap_stride, tap order (ix, iy) with iy fastest
    const int32_t *supports;
    // Planes [plane_lo, plane_hi) are present in kernels / compact, which
    // start at plane_lo (W tables recomputed per chunk of planes); 0 and
    // UINT32_MAX when every table is resident.
    uint32_t plane_lo, plane_hi;
    // AWProject: per (table, offset) Complex32 normalisation with the table's
    // w sign; degridded values are divided by its conjugate.  NULL otherwise.
    const float2 *aw_norm;
};

// num-complex 0.4 division z / conj(n): ((a c + b d), (b c - a d)) / (c^2 + d^2)
// with c = n.re, d = -n.im (degrid_sample_planned, gridder.rs:2359-2371).
template <typename T>
__device__ __forceinline__ void div_conj(T &re, T &im, T nr, T ni) {
#pragma clang fp contract(off)
    const T c = nr, d = -ni;
    const T ns = c * c + d * d;
    const T a = re, b = im;
    const T r = a * c + b * d;
    const T i = b * c - a * d;
    re = r / ns;
    im = i / ns;
}

__device__ __forceinline__ uint64_t cell_offset(const GridParams &gp, uint32_t x, uint32_t y) {
    if (!gp.paired) return static_cast<uint64_t>(x) * gp.xstride + y;
    if (y <= gp.hy) return static_cast<uint64_t>(y >> 3) * gp.bstride + x * 8u + (y & 7u);
    const uint32_t m = 2 * gp.hy - y;
    return gp.lower + static_cast<uint64_t>(m >> 3) * gp.bstride + x * 8u + (m & 7u);
}

GridParams grid_params(const chip_gridder &g, uint32_t tile, uint32_t halo, uint32_t tiles_y);

// W tables recomputed on the fly (g.wk_chunk != 0): scratch for one chunk of
// planes, then per chunk wk_fill recomputes planes [lo, hi) into it and points
// gp's kernels / compact / plane range at them.
struct WkScratch {
    DevBuf screen, kernels, compact;
};
void wk_prepare(const chip_gridder &g, WkScratch &ws);
void wk_fill(chip_gridder &g, WkScratch &ws, uint32_t lo, uint32_t hi, GridParams &gp);
// Kernels (planes x width x width complex64) of every plane into kernels_dev,
// recomputed when they are not resident.
void wk_all_kernels(const chip_gridder &g, float2 *kernels_dev);

// Wavefront-wide sum of a float (64 lanes, all active), returned as a
// wavefront-uniform value.  DPP lane swaps inside each row of 16 (xor 1, xor 2,
// half-row mirror, row mirror) leave the row sum in every lane of the row;
// two row broadcasts then carry the row sums into the last row, whose lane
// 63 holds the total.
#define CHIP_DPP_ADD(v, ctrl) \
    ((v) + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), (ctrl), 0xF, 0xF, true)))
// row_bcast15 / row_bcast31 into the rows of row_mask; the other rows add
// their own value (their result is not used: only lane 63 is read)
#define CHIP_DPP_ADD_ROWS(v, ctrl, rows) \
    ((v) + __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), (ctrl), (rows), 0xF, false)))
__device__ __forceinline__ float wave_sum_f(float v) {
    v = CHIP_DPP_ADD(v, 0xB1);             // quad_perm [1,0,3,2]
    v = CHIP_DPP_ADD(v, 0x4E);             // quad_perm [2,3,0,1]
    v = CHIP_DPP_ADD(v, 0x141);            // row_half_mirror
    v = CHIP_DPP_ADD(v, 0x140);            // row_mirror: every lane holds its row's sum
    v = CHIP_DPP_ADD_ROWS(v, 0x142, 0xA);  // row_bcast15: rows 1, 3 hold rows 0+1, 2+3
    v = CHIP_DPP_ADD_ROWS(v, 0x143, 0x8);  // row_bcast31: row 3 holds the total
    // (lanes of rows 0 - 2 now hold partial values; lane 63 is broadcast)
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Two wavefront sums with their steps interleaved (no DPP hazard waits
// between the dependent steps of one sum).
__device__ __forceinline__ void wave_sum_f2(float &a, float &b) {
    a = CHIP_DPP_ADD(a, 0xB1);
    b = CHIP_DPP_ADD(b, 0xB1);
    a = CHIP_DPP_ADD(a, 0x4E);
    b = CHIP_DPP_ADD(b, 0x4E);
    a = CHIP_DPP_ADD(a, 0x141);
    b = CHIP_DPP_ADD(b, 0x141);
    a = CHIP_DPP_ADD(a, 0x140);
    b = CHIP_DPP_ADD(b, 0x140);
    a = CHIP_DPP_ADD_ROWS(a, 0x142, 0xA);
    b = CHIP_DPP_ADD_ROWS(b, 0x142, 0xA);
    a = CHIP_DPP_ADD_ROWS(a, 0x143, 0x8);
    b = CHIP_DPP_ADD_ROWS(b, 0x143, 0x8);
    a = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(a), 63));
    b = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(b), 63));
}

// Sums over each half-wavefront (lanes 0-31 and 32-63), returned to every
// lane of that half: row sums, then v_permlane16_swap of the row sums with
// themselves, which leaves rows (0, 0, 2, 2) in one register and (1, 1, 3, 3)
// in the other, so one add gives rows 0 + 1 and 2 + 3.  (Same additions, in
// the same order, as row_bcast15: bit-identical sums.)
__device__ __forceinline__ float pair_rows(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}
__device__ __forceinline__ void half_sum_f2(float &a, float &b, bool upper) {
    (void)upper;
    a = CHIP_DPP_ADD(a, 0xB1);
    b = CHIP_DPP_ADD(b, 0xB1);
    a = CHIP_DPP_ADD(a, 0x4E);
    b = CHIP_DPP_ADD(b, 0x4E);
    a = CHIP_DPP_ADD(a, 0x141);
    b = CHIP_DPP_ADD(b, 0x141);
    a = CHIP_DPP_ADD(a, 0x140);
    b = CHIP_DPP_ADD(b, 0x140);
    a = pair_rows(a);
    b = pair_rows(b);
}

__device__ __forceinline__ float half_sum_f(float v, bool upper) {
    (void)upper;
    v = CHIP_DPP_ADD(v, 0xB1);
    v = CHIP_DPP_ADD(v, 0x4E);
    v = CHIP_DPP_ADD(v, 0x141);
    v = CHIP_DPP_ADD(v, 0x140);
    return pair_rows(v);
}

// Wavefront-wide sum of a double (64 lanes).
__device__ __forceinline__ double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Phase-centre rotation of one row (casars-imager/src/lib.rs:47064-47069):
// uvw' = uvw * uvrot as a row vector, shift = phrot . uvw', FP64 in this
// order without contraction.
__device__ __forceinline__ void rotate_uvw_row(const double *uvrot, const double *phrot, const double x[3],
                                               double o[3], double &shift) {
#pragma clang fp contract(off)
    for (int k = 0; k < 3; ++k) o[k] = x[0] * uvrot[k] + x[1] * uvrot[3 + k] + x[2] * uvrot[6 + k];
    shift = phrot[0] * o[0] + phrot[1] * o[1] + phrot[2] * o[2];
}

// phase_rotate_visibility (casa-imaging lib.rs:248-259): phi = -2 pi shift
// f / c in FP64, phasor (cos, sin) narrowed to Float, Float complex product;
// unchanged when shift or f is 0.
__device__ __forceinline__ float2 phase_rotate(float2 v, double shift, double f) {
#pragma clang fp contract(off)
    if (shift == 0.0 || f == 0.0) return v;
    const double phase = -6.28318530717958647692 * shift * f / 299792458.0;
    const float pr = static_cast<float>(cos(phase)), pi = static_cast<float>(sin(phase));
    return make_float2(v.x * pr - v.y * pi, v.x * pi + v.y * pr);
}

}  // namespace chip
