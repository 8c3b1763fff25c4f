// fk_fused_math.hpp -- per-cell arithmetic of the FP32 fused kernels, shared by the tiled kernels
// (fk_kernels_fused.hip) and the row-rolling kernels (fk_kernels_roll.hip).  Both families evaluate
// every cell with these expressions, under the same contraction rule (multiply-adds fused only
// inside one expression), so a cell's values do not depend on which kernel computes it.
// Reference formulas: src/cuda/fk_grad_Et_mono_PARAMETERS.cu (G) lap_func G:244-440, lapT_func_dev
// G:871-1158, jacT G:1160-1184, hermite3 / hermite4_1 G:1217-1241.
#pragma once
#pragma clang fp contract(on)

#include "fk_internal.hpp"

namespace fk {

// Per-cell stencil weights decoded once from the cell code.  The mirror Laplacian (a missing
// neighbour takes the opposite one's value; a cell never misses both neighbours on one axis) is
// sum_w = [a present](1+[b missing]) a + [b present](1+[a missing]) b per axis: weight 0 for a missing
// neighbour, 2 for its mirror partner, so the values held at hole and out-of-grid cells never enter
// an active cell's stencil.  L^T uses its 0/1/2 weights directly.
struct LapW {
    float t, b, l, r;   // weights of the top (i-1), bottom (i+1), left (j-1), right (j+1) values
    float gh;           // 2 Du/h * ghost_lr
};

__device__ inline LapW lap_weights(const Model32 &m, uint32_t c)
{
    LapW w;
    w.t = (c & C_TOP_MISS) ? 0.0f : ((c & C_BOT_MISS) ? 2.0f : 1.0f);
    w.b = (c & C_BOT_MISS) ? 0.0f : ((c & C_TOP_MISS) ? 2.0f : 1.0f);
    w.l = (c & C_LEFT_MISS) ? 0.0f : ((c & C_RIGHT_MISS) ? 2.0f : 1.0f);
    w.r = (c & C_RIGHT_MISS) ? 0.0f : ((c & C_LEFT_MISS) ? 2.0f : 1.0f);
    w.gh = m.ghc * (float)code_gh(c);
    return w;
}

__device__ inline LapW lapT_weights(uint32_t c)
{
    LapW w;
    w.t = (float)((c >> C_WT_SHIFT) & 3u);
    w.b = (float)((c >> C_WB_SHIFT) & 3u);
    w.l = (float)((c >> C_WL_SHIFT) & 3u);
    w.r = (float)((c >> C_WR_SHIFT) & 3u);
    w.gh = 0.0f;
    return w;
}

// sum of weighted neighbours minus 4 x centre (grid units)
__device__ inline float stencil5(const LapW &w, float top, float bot, float lft, float rgt, float ctr)
{
    return fmaf(w.t, top, fmaf(w.b, bot, fmaf(w.l, lft, fmaf(w.r, rgt, -4.0f * ctr))));
}

// f(y) + E gh from the reaction terms (fu, fv, fw) and the three stencils (s0, s1, s2)
__device__ inline void rhs_combine(const Model32 &m, const LapW &w, float E, float fu, float fv, float fw, float s0,
                                   float s1, float s2, float *f)
{
    f[0] = fmaf(m.lc[0], s0, fmaf(E, w.gh, fu));
    f[1] = fmaf(m.lc[1], s1, fv);
    f[2] = fmaf(m.lc[2], s2, fw);
}

// f(y) + E gh at an active cell from its value and its four neighbours' values, per variable
__device__ inline void rhs_vals(const Model32 &m, const LapW &w, const float *yc, const float *top, const float *bot,
                                const float *lft, const float *rgt, float E, float *f)
{
    Gates32 g = gates32(m, yc[0]);
    float fu, fv, fw;
    fion32(m, g, yc[0], yc[1], yc[2], fu, fv, fw);
    rhs_combine(m, w, E, fu, fv, fw, stencil5(w, top[0], bot[0], lft[0], rgt[0], yc[0]),
                stencil5(w, top[1], bot[1], lft[1], rgt[1], yc[1]), stencil5(w, top[2], bot[2], lft[2], rgt[2], yc[2]),
                f);
}

__device__ inline void jtz(const Jac32 &J, const float *z, float *o)
{
    o[0] = J.a00 * z[0] + J.a01 * z[1] + J.a02 * z[2];
    o[1] = J.a10 * z[0] + J.a11 * z[1];
    o[2] = J.a20 * z[0] + J.a22 * z[2];
}

// cubic Hermite at s = 1/3 and the quartic midpoint part at s = 1/2 (G:1217-1241)
template <typename T>
__device__ inline T h3(T x0, T x1, T m0, T m1)
{
    constexpr double s = 1.0 / 3.0;
    return T((s - 1.0) * (s - 1.0) * (2.0 * s + 1.0)) * x0 + T(s * (s - 1.0) * (s - 1.0)) * m0 +
           T(s * s * (3.0 - 2.0 * s)) * x1 + T(s * s * (s - 1.0)) * m1;
}
template <typename T>
__device__ inline T h4p(T x0, T x1, T m0, T m1)
{
    constexpr double s = 0.5;
    return T((1.0 + 2.0 * s + 9.0 * s * s) * (s - 1.0) * (s - 1.0)) * x0 + T(s * (s - 1.0) * (s - 1.0)) * m0 +
           T(-s * s * (6.0 - 16.0 * s + 9.0 * s * s)) * x1 + T((s - 1.0) * (9.0 * s - 5.0) * s * s / 4.0) * m1;
}

// Sum over the 64 lanes of a wavefront, valid in lane 0 (fixed shuffle tree).  The z.gh partial of
// an adjoint tile is the sum, in row order from 0.0, of these per-row sums, in every kernel family.
__device__ inline double wave_sum64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// The contribution of one cell to its row's z.gh sum (G:1457-1483): 0 unless active with gh != 0.
__device__ inline double zgh_term(const Model32 &m, bool act, uint32_t c, float zu)
{
    const int gh = code_gh(c);
    return (act && gh != 0) ? (double)zu * ((double)m.ghc * (double)gh) : 0.0;
}

}  // namespace fk
