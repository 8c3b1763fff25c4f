// fk_kernels_roll.hip -- FP32 fused time-step kernels for large launches, rolling down row strips.
//
// The tiled kernels (fk_kernels_fused.hip) give each workgroup a 64 x TY tile and recompute a
// 3-cell halo around it in every step: about a third of their arithmetic, plus the index work of
// the halo shells.  Here a workgroup holds one lane per column of an instance's row (the workgroup
// is as wide as the instance's pitch, 64 NW lanes) and walks down a strip of Hs rows, one row per
// iteration, with the four RK stages skewed by one row each: in iteration t the forward kernel
// evaluates stage 1 on row t, stage 2 on row t-1, stage 3 on row t-2 and stage 4 on row t-3 (the
// adjoint kernel likewise, with the Hermite midpoint state on the stage-1 row).  A stage needs the
// previous stage on its own row (left and right neighbours), on the row above (computed one
// iteration earlier) and on the row below (computed earlier in the same iteration by the same
// lane), so nothing is recomputed across columns and only 3 + 2 + 1 rows per strip end are
// evaluated twice.  Row values that neighbouring lanes read go through LDS in two-row rings (row
// parity); one barrier per iteration orders them.  A lane reads the row above from the ring slot
// that the current row's values overwrite afterwards, in its own column only, so two slots suffice.
//
// Every cell's arithmetic is the tiled kernels' (fk_fused_math.hpp, same expressions, same fused
// multiply-adds), so states, slopes and adjoint fields are bit for bit those of the tiled kernels,
// and the z.gh partial of each adjoint tile (the row sums of wave_sum64 added in row order) is too.
// Reference: src/cuda/fk_grad_Et_mono_PARAMETERS.cu (G): RK4 stages G:460-563, Hermite midpoint
// G:1244-1301, adjoint stages G:1349-1427, z.gh G:1457-1483.
#pragma clang fp contract(on)

#include "fk_internal.hpp"
#include "fk_tile_map.hpp"
#include "fk_fused_math.hpp"

namespace fk {

// Two-row LDS ring of three planes, each row P values between two zero pads (the values of the
// missing neighbours at the grid's left and right edges).
#define FK_LDS __attribute__((address_space(3)))
template <int P>
struct Ring2 {
    static constexpr int PP = P + 2;
    static constexpr int FLOATS = 2 * 3 * PP;
    FK_LDS float *b;
    __device__ FK_LDS float &at(int par, int v, int c) const { return b[(par * 3 + v) * PP + 1 + c]; }
};

// Lap weights of a row cell packed one per byte (top, bottom, left, right; 1 or 2), 0 for an
// inactive cell: one word per row instead of four registers, unpacked by byte conversions.
__device__ inline uint32_t pack_lap(uint32_t c)
{
    if (!(c & C_ACTIVE)) return 0u;
    return 0x01010101u + ((c >> 2) & 1u) + (((c >> 1) & 1u) << 8) + (((c >> 4) & 1u) << 16) + (((c >> 3) & 1u) << 24);
}
// L^T weights packed one per byte (0, 1 or 2), 0 for an inactive cell (an active cell has a
// neighbour on each axis, so its word is not 0).
__device__ inline uint32_t pack_lapT(uint32_t c)
{
    if (!(c & C_ACTIVE)) return 0u;
    return ((c >> C_WT_SHIFT) & 3u) | (((c >> C_WB_SHIFT) & 3u) << 8) | (((c >> C_WL_SHIFT) & 3u) << 16) |
           (((c >> C_WR_SHIFT) & 3u) << 24);
}
__device__ inline LapW unpack_w(uint32_t w, float gh)
{
    LapW r;
    r.t = (float)(w & 0xffu);
    r.b = (float)((w >> 8) & 0xffu);
    r.l = (float)((w >> 16) & 0xffu);
    r.r = (float)(w >> 24);
    r.gh = gh;
    return r;
}

// Row loads: wave-uniform row base (scalar registers) plus the lane's column.
__device__ inline float ld_row(const float *plane, int row, int P, int c)
{
    return ld_at(gmem(uniform_ptr(plane + (size_t)row * P)), (uint32_t)c);
}
__device__ inline void st_row(float *plane, int row, int P, int c, float v)
{
    st_at(gmem(uniform_ptr(plane + (size_t)row * P)), (uint32_t)c, v);
}

// ------------------------------------------------------------------------------ forward
// State of one lane of k_roll_fwd.  Register rings hold the rows a lane still needs (state y and RK
// sum acc of rows t-3..t+1, stage inputs Y2..Y4 of the rows being stencilled), indexed by
// (row - first row) mod the ring size, which is a compile-time value in the unrolled loop.
constexpr int RF_U = 12;   // iterations per unrolled round (multiple of the ring sizes 2, 4 and 6)

template <int P>
struct RollFwd {
    Model32 m;
    float dt, a12, a3, b1, b23, b4, E1, E23, E4;
    int rows, n, c, y0, y1, tb, mode;
    const float *yin;
    float *yout, *kout, *kout2;
    const uint16_t *code;
    Ring2<P> Ly, L2, L3, L4;
    float y[6][3];        // state, rows t-3 .. t+1
    uint32_t w[6];        // packed lap weights, rows t-3 .. t+1
    float gh[6];
    float acc[4][3];      // RK sums, rows t-3 .. t
    float Y2[4][3], Y3[4][3], Y4[4][3];
    float yn[3];          // arriving row (t+1 at iteration t)
    uint32_t cn;

    __device__ void load(int row)
    {
        if ((unsigned)row < (unsigned)rows) {
#pragma unroll
            for (int v = 0; v < 3; v++) yn[v] = ld_row(yin + (size_t)v * n, row, P, c);
            cn = ld_at(gmem(uniform_ptr(code + (size_t)row * P)), (uint32_t)c);
        } else {
            yn[0] = yn[1] = yn[2] = 0.0f;
            cn = 0u;
        }
    }
};

// One iteration: the arriving row t+1 enters the rings, the loads of row t+2 are issued, then
// stage 1 on row t, stage 2 on t-1, stage 3 on t-2, stage 4 on t-3.  PH = (t - tb) mod RF_U.
template <int P, int PH>
__device__ __forceinline__ void roll_fwd_iter(RollFwd<P> &F, int t)
{
    const int c = F.c;
    constexpr int q0 = PH % 6, qa = (PH + 1) % 6, qm1 = (PH + 5) % 6, qm2 = (PH + 4) % 6, qm3 = (PH + 3) % 6;
    constexpr int r0 = PH % 4, rm1 = (PH + 3) % 4, rm2 = (PH + 2) % 4, rm3 = (PH + 1) % 4, rm4 = PH % 4;
    constexpr int p0 = PH & 1, p1 = (PH + 1) & 1;   // parity of rows t (and t-2, t-4), t+1 (t-1, t-3)
    // the arriving row t+1: ring, weights, and its values for the neighbours' stencils (its ring slot
    // held row t-1, whose left/right values were read in the previous iteration)
#pragma unroll
    for (int v = 0; v < 3; v++) {
        F.y[qa][v] = F.yn[v];
        F.Ly.at(p1, v, c) = F.yn[v];
    }
    F.w[qa] = pack_lap(F.cn);
    F.gh[qa] = F.m.ghc * (float)code_gh(F.cn);
    F.load(t + 2);

    // stage 1 on row t
    {
        float f[3] = {0.0f, 0.0f, 0.0f};
        if ((unsigned)t < (unsigned)F.rows && F.w[q0] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.Ly.at(p0, v, c - 1); rt[v] = F.Ly.at(p0, v, c + 1); }
            rhs_vals(F.m, unpack_w(F.w[q0], F.gh[q0]), F.y[q0], F.y[qm1], F.y[qa], lf, rt, F.E1, f);
        }
        if (F.kout && t >= F.y0 && t < F.y1) {
#pragma unroll
            for (int v = 0; v < 3; v++) st_row(F.kout + (size_t)v * F.n, t, P, c, f[v]);
            if (F.kout2)
#pragma unroll
                for (int v = 0; v < 3; v++) st_row(F.kout2 + (size_t)v * F.n, t, P, c, f[v]);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y2[r0][v] = F.y[q0][v] + F.a12 * f[v];
            F.acc[r0][v] = F.b1 * f[v];
        }
    }
    // stage 2 on row t-1: Y2 of rows t-2 (above), t-1 (left, right) from the ring, t (below) in registers
    if (t >= F.tb + 2) {
        const int row = t - 1;
        float f[3] = {0.0f, 0.0f, 0.0f};
        if ((unsigned)row < (unsigned)F.rows && F.w[qm1] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L2.at(p1, v, c - 1); rt[v] = F.L2.at(p1, v, c + 1); }
            rhs_vals(F.m, unpack_w(F.w[qm1], F.gh[qm1]), F.Y2[rm1], F.Y2[rm2], F.Y2[r0], lf, rt, F.E23, f);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y3[rm1][v] = F.y[qm1][v] + F.a12 * f[v];
            F.acc[rm1][v] = F.acc[rm1][v] + F.b23 * f[v];
        }
    }
    // stage 3 on row t-2
    if (t >= F.tb + 4) {
        const int row = t - 2;
        float f[3] = {0.0f, 0.0f, 0.0f};
        if ((unsigned)row < (unsigned)F.rows && F.w[qm2] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L3.at(p0, v, c - 1); rt[v] = F.L3.at(p0, v, c + 1); }
            rhs_vals(F.m, unpack_w(F.w[qm2], F.gh[qm2]), F.Y3[rm2], F.Y3[rm3], F.Y3[rm1], lf, rt, F.E23, f);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y4[rm2][v] = F.y[qm2][v] + F.a3 * f[v];
            F.acc[rm2][v] = F.acc[rm2][v] + F.b23 * f[v];
        }
    }
    // stage 4 on row t-3 and the new state
    if (t >= F.tb + 6) {
        const int row = t - 3;
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (F.w[qm3] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L4.at(p1, v, c - 1); rt[v] = F.L4.at(p1, v, c + 1); }
            rhs_vals(F.m, unpack_w(F.w[qm3], F.gh[qm3]), F.Y4[rm3], F.Y4[rm4], F.Y4[rm2], lf, rt, F.E4, f);
        }
        const bool act = F.w[qm3] != 0u;
#pragma unroll
        for (int v = 0; v < 3; v++)
            st_row(F.yout + (size_t)v * F.n, row, P, c, act ? F.y[qm3][v] + (F.acc[rm3][v] + F.b4 * f[v]) : 0.0f);
    }
    // this iteration's stage values for the neighbours' stencils in the next one: each ring slot
    // held the row two above, read above in this lane's own column only
#pragma unroll
    for (int v = 0; v < 3; v++) {
        F.L2.at(p0, v, c) = F.Y2[r0][v];
        if (t >= F.tb + 2) F.L3.at(p1, v, c) = F.Y3[rm1][v];
        if (t >= F.tb + 4) F.L4.at(p0, v, c) = F.Y4[rm2][v];
    }
    __syncthreads();
}

// Forward-kind step (MODE_FWD, MODE_REC, MODE_KEEP) of one strip of Hs rows per workgroup; one lane
// per column of a pitch-P instance.  Inputs and outputs as k_fused_fwd with one step per launch.
template <int NW>
__global__ __launch_bounds__(64 * NW, 4) void k_roll_fwd(const FusedSlot *__restrict__ slots, Cursors cur,
                                                        LaunchMap map, int Hs)
{
    constexpr int P = 64 * NW;
    extern __shared__ float lds_raw[];
    FK_LDS float *lds = (FK_LDS float *)lds_raw;
    int slot, strip;
    launch_tile_index(blockIdx.x, map, slot, strip);
    const SlotCursor sc = cur.c[slot];
    if (sc.mode != MODE_FWD && sc.mode != MODE_KEEP && sc.mode != MODE_REC) return;   // uniform per block
    const FusedSlot &S = slots[slot];
    RollFwd<P> F;
    F.mode = sc.mode;
    const int step = sc.step, seg_a = sc.seg_a;
    const size_t st = (size_t)3 * S.n;
    F.kout = nullptr;
    F.kout2 = nullptr;
    if (F.mode == MODE_FWD) {
        F.yin = (step % S.K == 0) ? S.ck + (size_t)(step / S.K) * st : S.ring + (size_t)(step & 1) * st;
        const int nx = step + 1;
        F.yout = (nx % S.K == 0 || nx == S.nt) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                                : S.ring + (size_t)(nx & 1) * st;
    } else {
        const int j = step - seg_a, segi = seg_a / S.K, buf = seg_buf(S, segi), nx = step + 1;
        float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
        F.yin = (j == 0) ? S.ck + (size_t)segi * st : segb + (size_t)j * st;
        F.yout = (F.mode == MODE_KEEP && (nx % S.K == 0 || nx == S.nt)) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                                                       : segb + (size_t)(j + 1) * st;
        F.kout = S.kseg + ((size_t)buf * S.K + j) * st;
        if (F.mode == MODE_REC && j == 0 && segi < S.keep0) F.kout2 = S.kfirst + (size_t)(segi % 3) * st;
    }
    F.E1 = (float)S.Ef[3 * step];
    F.E23 = (float)S.Ef[3 * step + 1];
    F.E4 = (float)S.Ef[3 * step + 2];
    F.m = S.m;
    F.dt = S.dt;
    // the tiled kernels' stage coefficients
    F.a12 = 0.5f * F.dt;
    F.a3 = F.dt;
    F.b1 = F.dt * (1.0f / 6.0f);
    F.b23 = F.dt * (1.0f / 3.0f);
    F.b4 = F.dt * (1.0f / 6.0f);
    F.rows = S.rows;
    F.n = S.n;
    F.code = S.code;
    F.c = threadIdx.x;
    F.y0 = strip * Hs;
    F.y1 = min(F.y0 + Hs, S.rows);
    F.tb = F.y0 - 3;
    F.Ly.b = lds;
    F.L2.b = lds + Ring2<P>::FLOATS;
    F.L3.b = lds + 2 * Ring2<P>::FLOATS;
    F.L4.b = lds + 3 * Ring2<P>::FLOATS;
    // zero pads: 4 rings x 2 rows x 3 planes, two pads each
    for (int q = F.c; q < 48; q += P) lds[(q >> 1) * Ring2<P>::PP + ((q & 1) ? Ring2<P>::PP - 1 : 0)] = 0.0f;
    // rows tb-1 and tb before the first iteration (row tb also for its neighbours), row tb+1 in flight
    F.load(F.tb - 1);
    constexpr int q_m1 = 5, q_0 = 0;   // ring slots of rows tb-1 and tb
#pragma unroll
    for (int v = 0; v < 3; v++) F.y[q_m1][v] = F.yn[v];
    F.w[q_m1] = pack_lap(F.cn);
    F.gh[q_m1] = F.m.ghc * (float)code_gh(F.cn);
    F.load(F.tb);
#pragma unroll
    for (int v = 0; v < 3; v++) {
        F.y[q_0][v] = F.yn[v];
        F.Ly.at(0, v, F.c) = F.yn[v];
    }
    F.w[q_0] = pack_lap(F.cn);
    F.gh[q_0] = F.m.ghc * (float)code_gh(F.cn);
    F.load(F.tb + 1);
    __syncthreads();
    const int te = F.y1 + 3;
    for (int t0 = F.tb; t0 < te; t0 += RF_U) {
        roll_fwd_iter<P, 0>(F, t0);
        if (t0 + 1 >= te) break;
        roll_fwd_iter<P, 1>(F, t0 + 1);
        if (t0 + 2 >= te) break;
        roll_fwd_iter<P, 2>(F, t0 + 2);
        if (t0 + 3 >= te) break;
        roll_fwd_iter<P, 3>(F, t0 + 3);
        if (t0 + 4 >= te) break;
        roll_fwd_iter<P, 4>(F, t0 + 4);
        if (t0 + 5 >= te) break;
        roll_fwd_iter<P, 5>(F, t0 + 5);
        if (t0 + 6 >= te) break;
        roll_fwd_iter<P, 6>(F, t0 + 6);
        if (t0 + 7 >= te) break;
        roll_fwd_iter<P, 7>(F, t0 + 7);
        if (t0 + 8 >= te) break;
        roll_fwd_iter<P, 8>(F, t0 + 8);
        if (t0 + 9 >= te) break;
        roll_fwd_iter<P, 9>(F, t0 + 9);
        if (t0 + 10 >= te) break;
        roll_fwd_iter<P, 10>(F, t0 + 10);
        if (t0 + 11 >= te) break;
        roll_fwd_iter<P, 11>(F, t0 + 11);
    }
}


// ------------------------------------------------------------------------------ adjoint
// State of one lane of k_roll_adj.  Register rings: z and the RK sum dz of rows t-3..t+1, raw cell
// codes of rows t-3..t (slot (row - first row) mod 6), the quartic midpoint part y4 and the Jacobian
// at the midpoint state J(yh) of two rows (mod 2).  Stage values that the next stage stencils
// (y(t_{i-1} + dt/3), Z1, Z2, Z3) live in two-row LDS rings; the row above comes from the slot
// that the current row's value overwrites right after (own column), the row below from registers.
constexpr int RA_U = 6;   // iterations per unrolled round (multiple of the ring sizes 2 and 6)

template <int P>
struct RollAdj {
    Model32 m;
    float dt, E3, dEB;
    int rows, n, c, y0, y1, tb, tyA, ntiles, row_i;
    const float *zin, *yA, *yB, *kA, *kB;
    float *zout;
    double *partial;
    const uint16_t *code;
    Ring2<P> Xz, X13, X1, X2, X3;
    float z[6][3];        // adjoint state, rows t-3 .. t+1
    float dz[6][3];       // RK sums, rows t-3 .. t
    uint32_t cd[6];       // cell codes, rows t-3 .. t
    float y4[2][3];       // quartic midpoint part, rows t-1, t
    Jac32 JH[2];          // J(yh), rows t-2, t-1
    double tacc;          // z.gh sum of the current tile (lane 0 of each wavefront)
    // arriving: z of row t+1, and y_{i-1}, y_i, their slopes and the code of row t
    float zn[3], an[3], bn[3], kan[3], kbn[3];
    uint32_t cn;
    float ar[3];          // y_{i-1} of row t-3 for stage 4

    __device__ void load(int row)   // the arrivals of iteration `row`
    {
        if ((unsigned)(row + 1) < (unsigned)rows) {
#pragma unroll
            for (int v = 0; v < 3; v++) zn[v] = ld_row(zin + (size_t)v * n, row + 1, P, c);
        } else {
            zn[0] = zn[1] = zn[2] = 0.0f;
        }
        if ((unsigned)row < (unsigned)rows) {
#pragma unroll
            for (int v = 0; v < 3; v++) {
                an[v] = ld_row(yA + (size_t)v * n, row, P, c);
                bn[v] = ld_row(yB + (size_t)v * n, row, P, c);
                kan[v] = ld_row(kA + (size_t)v * n, row, P, c);
                kbn[v] = ld_row(kB + (size_t)v * n, row, P, c);
            }
            cn = ld_at(gmem(uniform_ptr(code + (size_t)row * P)), (uint32_t)c);
        } else {
#pragma unroll
            for (int v = 0; v < 3; v++) an[v] = bn[v] = kan[v] = kbn[v] = 0.0f;
            cn = 0u;
        }
    }
};

template <int P, int PH>
__device__ __forceinline__ void roll_adj_iter(RollAdj<P> &A, int t)
{
    const int c = A.c;
    const Model32 &m = A.m;
    const float dt = A.dt;
    constexpr int q0 = PH % 6, qa = (PH + 1) % 6, qm1 = (PH + 5) % 6, qm2 = (PH + 4) % 6, qm3 = (PH + 3) % 6;
    constexpr int h0 = PH & 1, h1 = (PH + 1) & 1;   // mod-2 slots of rows t (t-2) and t-1; also LDS parities
    constexpr int p0 = h0, p1 = h1;
    // y_{i-1} of row t-3 for stage 4 (in the grid whenever stage 4 runs)
    if (t >= A.tb + 6)
#pragma unroll
        for (int v = 0; v < 3; v++) A.ar[v] = ld_row(A.yA + (size_t)v * A.n, t - 3, P, c);
    // arrivals: z of row t+1 (its ring slot held row t-1, whose neighbours' values were read in the
    // previous iteration), the code of row t
    const float zup[3] = {A.z[qm1][0], A.z[qm1][1], A.z[qm1][2]};
#pragma unroll
    for (int v = 0; v < 3; v++) {
        A.z[qa][v] = A.zn[v];
        A.Xz.at(p1, v, c) = A.zn[v];
    }
    A.cd[q0] = A.cn;
    const uint32_t cc0 = A.cn;
    const bool act0 = (cc0 & C_ACTIVE) != 0;

    // Hermite midpoint on row t: y(t_{i-1} + dt/3) and the quartic part, from the stored end slopes
    float y3[3] = {0.0f, 0.0f, 0.0f};
    if (act0) {
        const float gh = m.ghc * (float)code_gh(cc0);
#pragma unroll
        for (int v = 0; v < 3; v++) {
            float a = A.an[v], b = A.bn[v];
            float fb = A.kbn[v];
            if (v == 0) fb += A.dEB * gh;
            float m0 = A.kan[v] * dt, m1 = fb * dt;
            y3[v] = h3<float>(a, b, m0, m1);
            A.y4[h0][v] = h4p<float>(a, b, m0, m1);
        }
    }
    // stage 1 on row t with J(y_i)
    float Z1[3] = {0.0f, 0.0f, 0.0f};
    {
        const LapW tw = lapT_weights(cc0);
        float lt[3];
#pragma unroll
        for (int v = 0; v < 3; v++)
            lt[v] = m.lc[v] * stencil5(tw, zup[v], A.z[qa][v], A.Xz.at(p0, v, c - 1), A.Xz.at(p0, v, c + 1), A.z[q0][v]);
        A.dz[q0][0] = A.dz[q0][1] = A.dz[q0][2] = 0.0f;
        if (act0) {
            Gates32 g = gates32(m, A.bn[0]);
            Jac32 J = jac32(m, g, A.bn[0], A.bn[1], A.bn[2]);
            float o[3];
            jtz(J, A.z[q0], o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Z1[v] = A.z[q0][v] + 0.5f * dt * fz;
                A.dz[q0][v] = (1.0f / 6.0f) * dt * fz;
            }
        }
    }
    A.load(t + 1);
    // stage 2 on row t-1: midpoint state yh = y4 + 27/64 dt f(y(t_{i-1} + dt/3)), J(yh)
    float Z2[3] = {0.0f, 0.0f, 0.0f};
    if (t >= A.tb + 2) {
        const uint32_t cc = A.cd[qm1];
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X1.at(p1, v, c);
            lt[v] = m.lc[v] * stencil5(tw, A.X1.at(p0, v, c), Z1[v], A.X1.at(p1, v, c - 1), A.X1.at(p1, v, c + 1), zc[v]);
        }
        if (cc & C_ACTIVE) {
            float yc[3], top[3], lf[3], rt[3], f3[3], yh[3], o[3];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                yc[v] = A.X13.at(p1, v, c);
                top[v] = A.X13.at(p0, v, c);
                lf[v] = A.X13.at(p1, v, c - 1);
                rt[v] = A.X13.at(p1, v, c + 1);
            }
            rhs_vals(m, lap_weights(m, cc), yc, top, y3, lf, rt, A.E3, f3);
#pragma unroll
            for (int v = 0; v < 3; v++) yh[v] = A.y4[h1][v] + (float)(27.0 / 64.0) * (f3[v] * dt);
            Gates32 g = gates32(m, yh[0]);
            A.JH[h1] = jac32(m, g, yh[0], yh[1], yh[2]);
            jtz(A.JH[h1], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Z2[v] = A.z[qm1][v] + 0.5f * dt * fz;
                A.dz[qm1][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
    }
    // this row's y(t_{i-1} + dt/3) and Z1 over the row two above (read above, own column)
#pragma unroll
    for (int v = 0; v < 3; v++) {
        A.X13.at(p0, v, c) = y3[v];
        A.X1.at(p0, v, c) = Z1[v];
    }
    // stage 3 on row t-2 with J(yh) of that row
    float Z3[3] = {0.0f, 0.0f, 0.0f};
    if (t >= A.tb + 4) {
        const uint32_t cc = A.cd[qm2];
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X2.at(p0, v, c);
            lt[v] = m.lc[v] * stencil5(tw, A.X2.at(p1, v, c), Z2[v], A.X2.at(p0, v, c - 1), A.X2.at(p0, v, c + 1), zc[v]);
        }
        if (cc & C_ACTIVE) {
            float o[3];
            jtz(A.JH[h0], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Z3[v] = A.z[qm2][v] + dt * fz;
                A.dz[qm2][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
    }
    if (t >= A.tb + 2)
#pragma unroll
        for (int v = 0; v < 3; v++) A.X2.at(p1, v, c) = Z2[v];
    // stage 4 on row t-3 with J(y_{i-1}); the new z and the row's z.gh sum
    if (t >= A.tb + 6) {
        const int row = t - 3;
        const uint32_t cc = A.cd[qm3];
        const bool act = (cc & C_ACTIVE) != 0;
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3], zo[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X3.at(p1, v, c);
            lt[v] = m.lc[v] * stencil5(tw, A.X3.at(p0, v, c), Z3[v], A.X3.at(p1, v, c - 1), A.X3.at(p1, v, c + 1), zc[v]);
        }
        if (act) {
            Gates32 g = gates32(m, A.ar[0]);
            Jac32 J = jac32(m, g, A.ar[0], A.ar[1], A.ar[2]);
            float o[3];
            jtz(J, zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) zo[v] = A.z[qm3][v] + (A.dz[qm3][v] + (1.0f / 6.0f) * dt * (lt[v] + o[v]));
        }
#pragma unroll
        for (int v = 0; v < 3; v++) st_row(A.zout + (size_t)v * A.n, row, P, c, zo[v]);
        const double rs = wave_sum64(zgh_term(m, act, cc, zo[0]));
        if ((c & 63) == 0) {
            const int j = (row - A.y0) % A.tyA;
            if (j == 0) A.tacc = 0.0;
            A.tacc += rs;
            if (j == A.tyA - 1 || row == A.rows - 1)
                gmem(A.partial)[(size_t)A.row_i * A.ntiles + (size_t)(row / A.tyA) * (P / 64) + (c >> 6)] = A.tacc;
        }
    }
    if (t >= A.tb + 4)
#pragma unroll
        for (int v = 0; v < 3; v++) A.X3.at(p0, v, c) = Z3[v];
    __syncthreads();
}

// Adjoint step of one strip of Hs rows per workgroup (Hs a multiple of the instance's adjoint tile
// rows); inputs and outputs as k_fused_adj.
template <int NW>
__global__ __launch_bounds__(64 * NW, 4) void k_roll_adj(const FusedSlot *__restrict__ slots, Cursors cur,
                                                        LaunchMap map, int Hs)
{
    constexpr int P = 64 * NW;
    extern __shared__ float lds_raw[];
    FK_LDS float *lds = (FK_LDS float *)lds_raw;
    int slot, strip;
    launch_tile_index(blockIdx.x, map, slot, strip);
    const SlotCursor sc = cur.c[slot];
    if (sc.mode != MODE_ADJ) return;   // uniform per block
    const FusedSlot &S = slots[slot];
    RollAdj<P> A;
    const int i = sc.step, seg_a = sc.seg_a;
    const size_t st = (size_t)3 * S.n;
    const int segi = seg_a / S.K;
    const int b_end = (seg_a + S.K < S.nt) ? seg_a + S.K : S.nt;
    const int buf = seg_buf(S, segi);
    const float *kbuf = S.kseg + (size_t)buf * S.K * st;
    const float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
    A.yA = (i - 1 == seg_a) ? S.ck + (size_t)segi * st : segb + (size_t)(i - 1 - seg_a) * st;
    A.yB = (i == b_end) ? S.ck + (size_t)(segi + 1) * st : segb + (size_t)(i - seg_a) * st;
    A.kA = kbuf + (size_t)(i - 1 - seg_a) * st;
    const float *knext = (b_end == S.nt) ? S.kT
                         : (segi + 1 >= S.keep0 ? S.kseg + (size_t)seg_buf(S, segi + 1) * S.K * st
                                                : S.kfirst + (size_t)((segi + 1) % 3) * st);
    A.kB = (i < b_end) ? kbuf + (size_t)(i - seg_a) * st : knext;
    A.zin = S.zb + (size_t)((S.nt - i) & 1) * st;
    A.zout = S.zb + (size_t)((S.nt - i + 1) & 1) * st;
    A.m = S.m;
    A.dt = S.dt;
    A.row_i = i - 1;
    A.E3 = (float)S.Ea[3 * A.row_i + 2];
    A.dEB = (i < S.nt) ? (float)(S.Ea[3 * A.row_i + 1] - S.Ef[3 * i]) : 0.0f;
    A.rows = S.rows;
    A.n = S.n;
    A.code = S.code;
    A.partial = S.partial;
    A.ntiles = S.ntiles;
    A.tyA = S.atile_rows;
    A.c = threadIdx.x;
    A.y0 = strip * Hs;
    A.y1 = min(A.y0 + Hs, S.rows);
    A.tb = A.y0 - 3;
    A.tacc = 0.0;
    A.Xz.b = lds;
    A.X13.b = lds + Ring2<P>::FLOATS;
    A.X1.b = lds + 2 * Ring2<P>::FLOATS;
    A.X2.b = lds + 3 * Ring2<P>::FLOATS;
    A.X3.b = lds + 4 * Ring2<P>::FLOATS;
    // zero pads: 5 rings x 2 rows x 3 planes, two pads each
    for (int q = A.c; q < 60; q += P) lds[(q >> 1) * Ring2<P>::PP + ((q & 1) ? Ring2<P>::PP - 1 : 0)] = 0.0f;
    // z of rows tb-1 and tb (row tb also for its neighbours' stencils), the arrivals of iteration tb
    A.load(A.tb - 2);   // z of row tb-1
#pragma unroll
    for (int v = 0; v < 3; v++) A.z[5][v] = A.zn[v];
    A.load(A.tb - 1);   // z of row tb
#pragma unroll
    for (int v = 0; v < 3; v++) {
        A.z[0][v] = A.zn[v];
        A.Xz.at(0, v, A.c) = A.zn[v];
    }
    A.load(A.tb);
    __syncthreads();
    const int te = A.y1 + 3;
    for (int t0 = A.tb; t0 < te; t0 += RA_U) {
        roll_adj_iter<P, 0>(A, t0);
        if (t0 + 1 >= te) break;
        roll_adj_iter<P, 1>(A, t0 + 1);
        if (t0 + 2 >= te) break;
        roll_adj_iter<P, 2>(A, t0 + 2);
        if (t0 + 3 >= te) break;
        roll_adj_iter<P, 3>(A, t0 + 3);
        if (t0 + 4 >= te) break;
        roll_adj_iter<P, 4>(A, t0 + 4);
        if (t0 + 5 >= te) break;
        roll_adj_iter<P, 5>(A, t0 + 5);
    }
}

// ------------------------------------------------------------------------------ launchers
// Dynamic LDS of the two kernels: 4 (forward) and 5 (adjoint) two-row rings of three planes.
template <int NW>
constexpr size_t roll_lds_bytes(int rings) { return sizeof(float) * (size_t)rings * Ring2<64 * NW>::FLOATS; }

template <int NW>
void roll_launch_nw(bool adj, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, int Hs,
                    hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map)), block(64 * NW);
    if (adj) {
        const size_t lds = roll_lds_bytes<NW>(5);
        static const bool attr = hipFuncSetAttribute((const void *)k_roll_adj<NW>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
        (void)attr;
        hipLaunchKernelGGL(k_roll_adj<NW>, grid, block, lds, stream, slots_dev, cur, map, Hs);
    } else {
        const size_t lds = roll_lds_bytes<NW>(4);
        static const bool attr = hipFuncSetAttribute((const void *)k_roll_fwd<NW>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
        (void)attr;
        hipLaunchKernelGGL(k_roll_fwd<NW>, grid, block, lds, stream, slots_dev, cur, map, Hs);
    }
}

bool roll_pitch_supported(int pitch)
{
    return pitch == 256 || pitch == 384 || pitch == 512 || pitch == 768 || pitch == 1024;
}

void roll_launch(bool adj, int pitch, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, int Hs,
                 hipStream_t stream)
{
    if (launch_map_tiles(map) == 0) return;
    switch (pitch) {
    case 256: roll_launch_nw<4>(adj, slots_dev, map, cur, Hs, stream); break;
    case 384: roll_launch_nw<6>(adj, slots_dev, map, cur, Hs, stream); break;
    case 512: roll_launch_nw<8>(adj, slots_dev, map, cur, Hs, stream); break;
    case 768: roll_launch_nw<12>(adj, slots_dev, map, cur, Hs, stream); break;
    case 1024: roll_launch_nw<16>(adj, slots_dev, map, cur, Hs, stream); break;
    default: break;
    }
}

}  // namespace fk
