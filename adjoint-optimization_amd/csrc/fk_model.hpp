// fk_model.hpp -- Fenton-Karma model on the device: cell codes, reaction, its transposed
// Jacobian, mirror Laplacian and its transpose, for FP64 (reference formulation with tanh) and
// FP32 (the same functions written with exp2-based logistic sigmoids).
//
// Reference: src/cuda/fk_grad_Et_mono_PARAMETERS.cu (G): F_ion G:30-46, lap_func G:244-440,
// lapT_func_dev G:871-1158, jacT G:1160-1184.  The reference marks holes with NaN in the state;
// here every cell carries a 16-bit code (built once on the host from the hole mask) and hole
// cells hold zeros, so the kernels never depend on NaN propagation.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fk {

// ---- cell code bits
enum : uint32_t {
    C_ACTIVE = 1u << 0,
    C_TOP_MISS = 1u << 1,    // cell (i-1, j) missing: the mirror rule uses (i+1, j) twice
    C_BOT_MISS = 1u << 2,    // cell (i+1, j) missing
    C_LEFT_MISS = 1u << 3,   // cell (i, j-1) missing
    C_RIGHT_MISS = 1u << 4,  // cell (i, j+1) missing
    C_WT_SHIFT = 5,          // 2-bit weights of L^T: top, bottom, left, right
    C_WB_SHIFT = 7,
    C_WL_SHIFT = 9,
    C_WR_SHIFT = 11,
    C_GH_SHIFT = 13,         // ghost_lr + 1 in {0, 1, 2}
};

__host__ __device__ inline int code_gh(uint32_t c) { return (int)((c >> C_GH_SHIFT) & 3u) - 1; }

// ---- per-instance constants (filled on the host)
struct Model64 {
    double u_c, u_v, tau_d, tau_v1_minus, tau_v2_minus, tau_v_plus, tau_0, tau_r, tau_si,
        u_csi, k_fk, tau_w_minus, tau_w_plus, ksm;
    double lc[3];   // D_v / h^2
    double ghc;     // 2 Du / h   (gh = ghc * ghost_lr, G:1588-1595)
};

struct Model32 {
    float kq, bq;        // exp2(kq*u + bq) = exp(-2 ksm (u - u_v))
    float cpq;           // exp(2 ksm (u_c - u_v)):  e_p = exp(-2 ksm (u - u_c)) = e_q * cpq
    float ks, bs;        // exp2(ks*u + bs) = exp(-2 k (u - u_csi))
    float two_ksm, two_k;
    float u_c, one_plus_uc;
    float i_tau_d, i_tau_0, i_tau_r, i_tau_si, i_tau_vp, i_tau_wm, i_tau_wp;
    float d_tau_w;       // 1/tau_w+ - 1/tau_w-
    float tau_v2m, dtv;  // tau_v- = tau_v2m + q*dtv
    float lc[3];
    float ghc;
};

// =====================================================================================
// FP64: the reference's formulas (tanh-based smoothed steps)
// =====================================================================================
__device__ inline void fion64(const Model64 &m, double u, double v, double w, double &fu, double &fv,
                              double &fw)
{
    double p = 0.5 * (1 + tanh(m.ksm * (u - m.u_c)));
    double q = 0.5 * (1 + tanh(m.ksm * (u - m.u_v)));
    double j_fi = -v * p * (1 - u) * (u - m.u_c) / m.tau_d;
    double j_so = u * (1 - p) / m.tau_0 + p / m.tau_r;
    double tvm = q * m.tau_v1_minus + (1 - q) * m.tau_v2_minus;
    double j_si = -w * (1 + tanh(m.k_fk * (u - m.u_csi))) / (2 * m.tau_si);
    fu = -(j_fi + j_so + j_si);
    fv = ((1 - p) * (1 - v) / tvm) - (p * v) / m.tau_v_plus;
    fw = ((1 - p) * (1 - w) / m.tau_w_minus) - (p * w) / m.tau_w_plus;
}

// (J_F(y)^T z): the seven non-zero entries of dF_b/dy_a contracted with z.
__device__ inline void jacTz64(const Model64 &m, double u, double v, double w, double zu, double zv, double zw,
                               double &ou, double &ov, double &ow)
{
    double tp = tanh(m.ksm * (u - m.u_c));
    double tq = tanh(m.ksm * (u - m.u_v));
    double ts = tanh(m.k_fk * (u - m.u_csi));
    double p = 0.5 * tp + 0.5, q = 0.5 * tq + 0.5;
    double dp = 0.5 * m.ksm * (1 - tp * tp);
    double dq = 0.5 * m.ksm * (1 - tq * tq);
    double tvm = q * m.tau_v1_minus + (1 - q) * m.tau_v2_minus;
    double dtvm = dq * (m.tau_v1_minus - m.tau_v2_minus);
    double g = (1 - u) * (u - m.u_c);
    double dg = (1 - u) - (u - m.u_c);
    double a00 = v * (dp * g + p * dg) / m.tau_d - ((1 - p) - u * dp) / m.tau_0 - dp / m.tau_r
               + w * m.k_fk * (1 - ts * ts) / (2 * m.tau_si);
    double a10 = p * g / m.tau_d;
    double a20 = (1 + ts) / (2 * m.tau_si);
    double a01 = -dp * (1 - v) / tvm - (1 - p) * (1 - v) * dtvm / (tvm * tvm) - dp * v / m.tau_v_plus;
    double a11 = -(1 - p) / tvm - p / m.tau_v_plus;
    double a02 = -dp * (1 - w) / m.tau_w_minus - dp * w / m.tau_w_plus;
    double a22 = -(1 - p) / m.tau_w_minus - p / m.tau_w_plus;
    ou = a00 * zu + a01 * zv + a02 * zw;
    ov = a10 * zu + a11 * zv;
    ow = a20 * zu + a22 * zw;
}

// =====================================================================================
// FP32: p = sigma(2 ksm (u-u_c)), q = sigma(2 ksm (u-u_v)), (1+tanh(k(u-u_csi)))/2 = sigma(2k(u-u_csi)),
// written with e_q = exp(-2 ksm (u-u_v)), e_p = e_q exp(2 ksm (u_c-u_v)), e_s = exp(-2k (u-u_csi)):
//   p = 1/A, q = 1/B, sigma_s = 1/C with A = 1+e_p, B = 1+e_q, C = 1+e_s, and
//   1/tau_v- = 1/(tau_v2 + q (tau_v1-tau_v2)) = B/Dv with Dv = tau_v2 B + (tau_v1-tau_v2).
// One reciprocal r = 1/(A B C Dv) gives all four.  The exponents are capped at 2^24 so the
// product stays below 2^127 (the gates then bottom out at 2^-24, FP32 rounding next to 1).
// 1-p = e_p p (and alike) avoids the cancellation of 1 - p near p = 1.
// =====================================================================================
struct Gates32 {
    float p, omp, q, omq, s, oms, itvm;   // p, 1-p, q, 1-q, sigma_s, 1-sigma_s, 1/tau_v-
};

__device__ inline float fast_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

__device__ inline Gates32 gates32(const Model32 &m, float u)
{
    Gates32 g;
    float eq = fast_exp2(fminf(fmaf(m.kq, u, m.bq), 24.0f));
    float es = fast_exp2(fminf(fmaf(m.ks, u, m.bs), 24.0f));
    float ep = eq * m.cpq;
    float A = 1.0f + ep, B = 1.0f + eq, C = 1.0f + es;
    float Dv = fmaf(m.tau_v2m, B, m.dtv);
    float AB = A * B, ABC = AB * C;
    float r = fast_rcp(ABC * Dv);
    float rD = r * Dv;                       // 1 / (A B C)
    g.p = rD * (B * C);
    g.q = rD * (A * C);
    g.s = rD * AB;
    g.itvm = r * (ABC * B);                  // B / Dv
    g.omp = ep * g.p;
    g.omq = eq * g.q;
    g.oms = es * g.s;
    return g;
}

// The reaction written with the fewest operations (FP32 only; the FP64 path keeps the reference's
// form):  (1-u)(u-u_c) = u (1+u_c-u) - u_c,  J_fi + J_so terms grouped by p, and
// (1-p)(1-v)/tau_v- - p v/tau_v+ = a - v (a + b) with a = (1-p)/tau_v-, b = p/tau_v+ (same for w).
__device__ inline void fion32(const Model32 &m, const Gates32 &g, float u, float v, float w, float &fu,
                              float &fv, float &fw)
{
    const float gg = u * (m.one_plus_uc - u) - m.u_c;
    const float a = fmaf(v * gg, m.i_tau_d, -m.i_tau_r);
    const float b = fmaf(w * g.s, m.i_tau_si, -(u * g.omp) * m.i_tau_0);
    fu = fmaf(g.p, a, b);
    const float av = g.omp * g.itvm, bv = g.p * m.i_tau_vp;
    fv = fmaf(-v, av + bv, av);
    const float aw = g.omp * m.i_tau_wm, bw = g.p * m.i_tau_wp;
    fw = fmaf(-w, aw + bw, aw);
}

// Transposed Jacobian entries (seven non-zeros) at one state.
struct Jac32 {
    float a00, a01, a02, a10, a11, a20, a22;
};

// Same grouping as fion32: d(1/tau_v-)/du = -(1/tau_v-)^2 dq (tau_v1- - tau_v2-) enters a01 as
// -(1-v)(1/tau_v-)(dp + (1-p)(1/tau_v-) dq dtv).
__device__ inline Jac32 jac32(const Model32 &m, const Gates32 &g, float u, float v, float w)
{
    Jac32 J;
    const float dp = m.two_ksm * g.p * g.omp;
    const float dq = m.two_ksm * g.q * g.omq;
    const float ds = m.two_k * g.s * g.oms;
    const float gg = u * (m.one_plus_uc - u) - m.u_c;
    const float dg = m.one_plus_uc - 2.0f * u;
    const float av = g.omp * g.itvm, bv = g.p * m.i_tau_vp;
    J.a00 = fmaf(v * m.i_tau_d, fmaf(dp, gg, g.p * dg),
                 fmaf(fmaf(u, dp, -g.omp), m.i_tau_0, fmaf(w * ds, m.i_tau_si, -dp * m.i_tau_r)));
    J.a10 = g.p * gg * m.i_tau_d;
    J.a20 = g.s * m.i_tau_si;
    J.a01 = -fmaf((1.0f - v) * g.itvm, fmaf(av * dq, m.dtv, dp), dp * v * m.i_tau_vp);
    J.a11 = -(av + bv);
    J.a02 = -dp * fmaf(w, m.d_tau_w, m.i_tau_wm);
    J.a22 = -fmaf(g.omp, m.i_tau_wm, g.p * m.i_tau_wp);
    return J;
}

// =====================================================================================
// Laplacian pieces on values already gathered (used by all kernels)
// =====================================================================================
// Mirror Laplacian at an active cell: missing neighbour -> opposite neighbour (G:320-421).
template <typename T>
__device__ inline T lap_val(uint32_t c, T ctr, T top, T bot, T lft, T rgt, T lc)
{
    T a = (c & C_TOP_MISS) ? bot : top;
    T b = (c & C_BOT_MISS) ? top : bot;
    T l = (c & C_LEFT_MISS) ? rgt : lft;
    T r = (c & C_RIGHT_MISS) ? lft : rgt;
    return lc * (a + b + l + r - T(4) * ctr);
}

// Transposed Laplacian at an active cell: weights 0/1/2 per neighbour (G:1004-1116).  Missing
// neighbours hold 0 in the state, so weight 0 times that value is exactly 0.
template <typename T>
__device__ inline T lapT_val(uint32_t c, T ctr, T top, T bot, T lft, T rgt, T lc)
{
    T wt = T((c >> C_WT_SHIFT) & 3u), wb = T((c >> C_WB_SHIFT) & 3u);
    T wl = T((c >> C_WL_SHIFT) & 3u), wr = T((c >> C_WR_SHIFT) & 3u);
    return lc * (wt * top + wb * bot + wl * lft + wr * rgt - T(4) * ctr);
}

}  // namespace fk
