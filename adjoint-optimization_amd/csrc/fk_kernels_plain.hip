// fk_kernels_plain.hip -- one kernel per RK stage ("plain" path), templated on the state type.
//
// This is the reference's kernel structure (G = src/cuda/fk_grad_Et_mono_PARAMETERS.cu):
//   forward   fk_lap_ion_rk4_1 / _23 / _4          G:460-563
//   adjoint   interp_midpoint_1 / _2, fk_adj_rk4_1..4   G:1244-1301, G:1349-1427
//   reduction zgh_dot_partial                       G:1457-1483
// laid out for CDNA: a 64x4 block maps one wavefront to 64 contiguous cells of a row (the
// reference puts threadIdx.x on the strided axis), holes are codes instead of NaNs, the state
// stays in HBM for the whole evaluation, and the per-block FP64 partial sums go to a
// [step][block] table reduced once at the end instead of a host round-trip per step.
// With T = double this is the FP64 baseline path.
#include "fk_internal.hpp"

namespace fk {

// ---------------------------------------------------------------------- helpers
__device__ inline void react(const Model64 &m, double u, double v, double w, double *f)
{
    fion64(m, u, v, w, f[0], f[1], f[2]);
}
__device__ inline void react(const Model32 &m, float u, float v, float w, float *f)
{
    Gates32 g = gates32(m, u);
    fion32(m, g, u, v, w, f[0], f[1], f[2]);
}
__device__ inline void adj_react(const Model64 &m, const double *y, const double *z, double *o)
{
    jacTz64(m, y[0], y[1], y[2], z[0], z[1], z[2], o[0], o[1], o[2]);
}
__device__ inline void adj_react(const Model32 &m, const float *y, const float *z, float *o)
{
    Gates32 g = gates32(m, y[0]);
    Jac32 J = jac32(m, g, y[0], y[1], y[2]);
    o[0] = J.a00 * z[0] + J.a01 * z[1] + J.a02 * z[2];
    o[1] = J.a10 * z[0] + J.a11 * z[1];
    o[2] = J.a20 * z[0] + J.a22 * z[2];
}
__device__ inline const double *lc_of(const Model64 &m) { return m.lc; }
__device__ inline const float *lc_of(const Model32 &m) { return m.lc; }
__device__ inline double ghc_of(const Model64 &m) { return m.ghc; }
__device__ inline float ghc_of(const Model32 &m) { return m.ghc; }

struct CellPos {
    int i, j, k;
    bool in;
};

__device__ inline CellPos cell_pos(const GridArgs &g)
{
    CellPos p;
    p.j = blockIdx.x * PLAIN_BX + threadIdx.x;
    p.i = blockIdx.y * PLAIN_BY + threadIdx.y;
    p.in = p.i < g.rows;
    p.k = p.i * g.pitch + p.j;
    return p;
}

// RHS f(y) + E gh at an active cell (G:1187-1215): mirror neighbours chosen by index, so no
// load leaves the grid.
template <typename T, typename M>
__device__ inline void rhs_at(const GridArgs &g, const M &m, const T *__restrict__ Y, int k, uint32_t c, T E, T *f)
{
    const int n = g.n, P = g.pitch;
    int kt = (c & C_TOP_MISS) ? k + P : k - P;
    int kb = (c & C_BOT_MISS) ? k - P : k + P;
    int kl = (c & C_LEFT_MISS) ? k + 1 : k - 1;
    int kr = (c & C_RIGHT_MISS) ? k - 1 : k + 1;
    T y[3], l[3];
    const T *lc = lc_of(m);
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const T *p = Y + s * n;
        y[s] = p[k];
        l[s] = lc[s] * (p[kt] + p[kb] + p[kl] + p[kr] - T(4) * y[s]);
    }
    react(m, y[0], y[1], y[2], f);
    f[0] = l[0] + f[0] + E * (ghc_of(m) * T(code_gh(c)));
    f[1] = l[1] + f[1];
    f[2] = l[2] + f[2];
}

// Adjoint field fz = L^T z + J_F(y)^T z at an active cell (G:1303-1347).
template <typename T, typename M>
__device__ inline void adjf_at(const GridArgs &g, const M &m, const T *__restrict__ Z, const T *__restrict__ Yj,
                               int k, uint32_t c, T *fz, T *zc)
{
    const int n = g.n, P = g.pitch;
    const T *lc = lc_of(m);
    unsigned wt = (c >> C_WT_SHIFT) & 3u, wb = (c >> C_WB_SHIFT) & 3u;
    unsigned wl = (c >> C_WL_SHIFT) & 3u, wr = (c >> C_WR_SHIFT) & 3u;
    T y[3], lt[3], o[3];
#pragma unroll
    for (int s = 0; s < 3; s++) {
        const T *p = Z + s * n;
        zc[s] = p[k];
        T vt = wt ? p[k - P] : T(0), vb = wb ? p[k + P] : T(0);
        T vl = wl ? p[k - 1] : T(0), vr = wr ? p[k + 1] : T(0);
        lt[s] = lc[s] * (T(wt) * vt + T(wb) * vb + T(wl) * vl + T(wr) * vr - T(4) * zc[s]);
        y[s] = Yj[s * n + k];
    }
    adj_react(m, y, zc, o);
#pragma unroll
    for (int s = 0; s < 3; s++) fz[s] = lt[s] + o[s];
}

// ---------------------------------------------------------------------- forward stages
// S=1: U = y + dt/2 k1, du = dt/6 k1                       (G:460-493)
// S=2: Ut = y + dt/2 k2, du += dt/3 k2; S=3: U = y + dt k3  (G:495-538)
// S=4: y_out = y + du + dt/6 k4                             (G:540-563)
template <typename T, typename M, int S>
__global__ __launch_bounds__(256) void k_fwd_stage(GridArgs g, M m, const T *__restrict__ Yin,
                                                   const T *__restrict__ y, T *__restrict__ Yout,
                                                   T *__restrict__ du, const double *__restrict__ Ef, int step,
                                                   T dt)
{
    CellPos p = cell_pos(g);
    if (!p.in) return;
    const int n = g.n, k = p.k;
    uint32_t c = g.code[k];
    if (!(c & C_ACTIVE)) {
#pragma unroll
        for (int s = 0; s < 3; s++) {
            Yout[s * n + k] = T(0);
            if (S < 4) du[s * n + k] = T(0);
        }
        return;
    }
    T E = (T)Ef[3 * step + (S == 1 ? 0 : (S == 4 ? 2 : 1))];
    T f[3];
    rhs_at<T>(g, m, Yin, k, c, E, f);
#pragma unroll
    for (int s = 0; s < 3; s++) {
        int q = s * n + k;
        if (S == 1) {
            Yout[q] = y[q] + dt * T(0.5) * f[s];
            du[q] = T(1.0 / 6.0) * dt * f[s];
        } else if (S == 2) {
            Yout[q] = y[q] + dt * T(0.5) * f[s];
            du[q] = du[q] + T(1.0 / 3.0) * dt * f[s];
        } else if (S == 3) {
            Yout[q] = y[q] + dt * f[s];
            du[q] = du[q] + T(1.0 / 3.0) * dt * f[s];
        } else {
            Yout[q] = y[q] + du[q] + dt * T(1.0 / 6.0) * f[s];
        }
    }
}

// ---------------------------------------------------------------------- adjoint
// Hermite reconstruction over [t_{i-1}, t_i] (G:1217-1273): y(t_{i-1}+dt/3) and the partial
// quartic value at the midpoint, from the end states and the end slopes dt*f.
template <typename T>
__device__ inline T herm3(T x0, T x1, T m0, T m1)
{
    constexpr double s = 1.0 / 3.0;
    return T((s - 1.0) * (s - 1.0) * (2.0 * s + 1.0)) * x0 + T(s * (s - 1.0) * (s - 1.0)) * m0 +
           T(s * s * (3.0 - 2.0 * s)) * x1 + T(s * s * (s - 1.0)) * m1;
}
template <typename T>
__device__ inline T herm4_part(T x0, T x1, T m0, T m1)
{
    constexpr double s = 0.5;
    return T((1.0 + 2.0 * s + 9.0 * s * s) * (s - 1.0) * (s - 1.0)) * x0 + T(s * (s - 1.0) * (s - 1.0)) * m0 +
           T(-s * s * (6.0 - 16.0 * s + 9.0 * s * s)) * x1 + T((s - 1.0) * (9.0 * s - 5.0) * s * s / 4.0) * m1;
}
constexpr double kHermMid = 27.0 * 0.5 * 0.5 * (0.5 - 1.0) * (0.5 - 1.0) / 4.0;   // 27/64 (G:1295)

template <typename T, typename M>
__global__ __launch_bounds__(256) void k_adj_interp1(GridArgs g, M m, const T *__restrict__ yA,
                                                     const T *__restrict__ yB, T *__restrict__ y3,
                                                     T *__restrict__ y4, const double *__restrict__ Ea, int row,
                                                     T dt)
{
    CellPos p = cell_pos(g);
    if (!p.in) return;
    const int n = g.n, k = p.k;
    uint32_t c = g.code[k];
    if (!(c & C_ACTIVE)) {
#pragma unroll
        for (int s = 0; s < 3; s++) { y3[s * n + k] = T(0); y4[s * n + k] = T(0); }
        return;
    }
    T f0[3], f1[3];
    rhs_at<T>(g, m, yA, k, c, (T)Ea[3 * row], f0);
    rhs_at<T>(g, m, yB, k, c, (T)Ea[3 * row + 1], f1);
#pragma unroll
    for (int s = 0; s < 3; s++) {
        T a = yA[s * n + k], b = yB[s * n + k];
        T m0 = f0[s] * dt, m1 = f1[s] * dt;
        y3[s * n + k] = herm3(a, b, m0, m1);
        y4[s * n + k] = herm4_part(a, b, m0, m1);
    }
}

template <typename T, typename M>
__global__ __launch_bounds__(256) void k_adj_interp2(GridArgs g, M m, const T *__restrict__ y3,
                                                     const T *__restrict__ y4, T *__restrict__ yh,
                                                     const double *__restrict__ Ea, int row, T dt)
{
    CellPos p = cell_pos(g);
    if (!p.in) return;
    const int n = g.n, k = p.k;
    uint32_t c = g.code[k];
    if (!(c & C_ACTIVE)) {
#pragma unroll
        for (int s = 0; s < 3; s++) yh[s * n + k] = T(0);
        return;
    }
    T f[3];
    rhs_at<T>(g, m, y3, k, c, (T)Ea[3 * row + 2], f);
#pragma unroll
    for (int s = 0; s < 3; s++) yh[s * n + k] = y4[s * n + k] + T(kHermMid) * (f[s] * dt);
}

// S=1: Z = z + dt/2 fz(z; y_i), dz = dt/6 fz        S=2: Zt = z + dt/2 fz(Z; yh), dz += dt/3 fz
// S=3: Z = z + dt fz(Zt; yh), dz += dt/3 fz         S=4: z_out = z + dz + dt/6 fz(Z; y_{i-1})
template <typename T, typename M, int S>
__global__ __launch_bounds__(256) void k_adj_stage(GridArgs g, M m, const T *__restrict__ Zin,
                                                   const T *__restrict__ Yj, const T *__restrict__ z,
                                                   T *__restrict__ Zout, T *__restrict__ dz, T dt)
{
    CellPos p = cell_pos(g);
    if (!p.in) return;
    const int n = g.n, k = p.k;
    uint32_t c = g.code[k];
    if (!(c & C_ACTIVE)) {
#pragma unroll
        for (int s = 0; s < 3; s++) {
            Zout[s * n + k] = T(0);
            if (S < 4) dz[s * n + k] = T(0);
        }
        return;
    }
    T f[3], zc[3];
    adjf_at<T>(g, m, Zin, Yj, k, c, f, zc);
#pragma unroll
    for (int s = 0; s < 3; s++) {
        int q = s * n + k;
        if (S == 1) {
            Zout[q] = z[q] + T(0.5) * dt * f[s];
            dz[q] = T(1.0 / 6.0) * dt * f[s];
        } else if (S == 2) {
            Zout[q] = z[q] + T(0.5) * dt * f[s];
            dz[q] = dz[q] + T(2.0 / 6.0) * dt * f[s];
        } else if (S == 3) {
            Zout[q] = z[q] + dt * f[s];
            dz[q] = dz[q] + T(2.0 / 6.0) * dt * f[s];
        } else {
            Zout[q] = z[q] + dz[q] + T(1.0 / 6.0) * dt * f[s];
        }
    }
}

// Per-block FP64 partial of sum_k z_u[k] gh[k] (G:1457-1483), fixed reduction order.
template <typename T, typename M>
__global__ __launch_bounds__(256) void k_zgh_partial(GridArgs g, M m, const T *__restrict__ z,
                                                     double *__restrict__ out)
{
    CellPos p = cell_pos(g);
    double v = 0.0;
    if (p.in) {
        uint32_t c = g.code[p.k];
        int gh = code_gh(c);
        if ((c & C_ACTIVE) && gh != 0) v = (double)z[p.k] * ((double)ghc_of(m) * (double)gh);
    }
    double s = block_sum_256(v, threadIdx.y * PLAIN_BX + threadIdx.x);
    if (threadIdx.x == 0 && threadIdx.y == 0) out[blockIdx.y * gridDim.x + blockIdx.x] = s;
}

// ---------------------------------------------------------------------- launchers
template <typename T, typename M>
void plain_forward_step(const PlainCtx<T, M> &c, const T *y_in, T *y_out, int step)
{
    dim3 blk(PLAIN_BX, PLAIN_BY), grd(c.g.pitch / PLAIN_BX, (c.g.rows + PLAIN_BY - 1) / PLAIN_BY);
    T dt = (T)c.dt;
    T *U = c.w0, *Ut = c.w1, *du = c.w2;
    hipLaunchKernelGGL((k_fwd_stage<T, M, 1>), grd, blk, 0, c.stream, c.g, c.m, y_in, y_in, U, du, c.Ef, step, dt);
    hipLaunchKernelGGL((k_fwd_stage<T, M, 2>), grd, blk, 0, c.stream, c.g, c.m, U, y_in, Ut, du, c.Ef, step, dt);
    hipLaunchKernelGGL((k_fwd_stage<T, M, 3>), grd, blk, 0, c.stream, c.g, c.m, Ut, y_in, U, du, c.Ef, step, dt);
    hipLaunchKernelGGL((k_fwd_stage<T, M, 4>), grd, blk, 0, c.stream, c.g, c.m, U, y_in, y_out, du, c.Ef, step, dt);
}

template <typename T, typename M>
void plain_adjoint_step(const PlainCtx<T, M> &c, const T *yA, const T *yB, const T *z_in, T *z_out, int row,
                        double *partial_row)
{
    dim3 blk(PLAIN_BX, PLAIN_BY), grd(c.g.pitch / PLAIN_BX, (c.g.rows + PLAIN_BY - 1) / PLAIN_BY);
    T dt = (T)c.dt;
    T *y3 = c.w0, *y4 = c.w1, *yh = c.w2, *Z = c.w3, *Zt = c.w4, *dz = c.w5;
    hipLaunchKernelGGL((k_adj_interp1<T, M>), grd, blk, 0, c.stream, c.g, c.m, yA, yB, y3, y4, c.Ea, row, dt);
    hipLaunchKernelGGL((k_adj_interp2<T, M>), grd, blk, 0, c.stream, c.g, c.m, y3, y4, yh, c.Ea, row, dt);
    hipLaunchKernelGGL((k_adj_stage<T, M, 1>), grd, blk, 0, c.stream, c.g, c.m, z_in, yB, z_in, Z, dz, dt);
    hipLaunchKernelGGL((k_adj_stage<T, M, 2>), grd, blk, 0, c.stream, c.g, c.m, Z, yh, z_in, Zt, dz, dt);
    hipLaunchKernelGGL((k_adj_stage<T, M, 3>), grd, blk, 0, c.stream, c.g, c.m, Zt, yh, z_in, Z, dz, dt);
    hipLaunchKernelGGL((k_adj_stage<T, M, 4>), grd, blk, 0, c.stream, c.g, c.m, Z, yA, z_in, z_out, dz, dt);
    hipLaunchKernelGGL((k_zgh_partial<T, M>), grd, blk, 0, c.stream, c.g, c.m, z_out, partial_row);
}

template <typename T, typename M>
void plain_zgh(const PlainCtx<T, M> &c, const T *z, double *partial_row)
{
    dim3 blk(PLAIN_BX, PLAIN_BY), grd(c.g.pitch / PLAIN_BX, (c.g.rows + PLAIN_BY - 1) / PLAIN_BY);
    hipLaunchKernelGGL((k_zgh_partial<T, M>), grd, blk, 0, c.stream, c.g, c.m, z, partial_row);
}

template void plain_forward_step<double, Model64>(const PlainCtx<double, Model64> &, const double *, double *, int);
template void plain_adjoint_step<double, Model64>(const PlainCtx<double, Model64> &, const double *, const double *,
                                                  const double *, double *, int, double *);
template void plain_zgh<double, Model64>(const PlainCtx<double, Model64> &, const double *, double *);
template void plain_forward_step<float, Model32>(const PlainCtx<float, Model32> &, const float *, float *, int);
template void plain_adjoint_step<float, Model32>(const PlainCtx<float, Model32> &, const float *, const float *,
                                                 const float *, float *, int, double *);
template void plain_zgh<float, Model32>(const PlainCtx<float, Model32> &, const float *, double *);

}  // namespace fk
