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
// Odd strips roll upwards (the same code with the rolling row order mirrored): the halo rows two
// neighbouring strips both read then reach both of them at about the same time, at the start or at
// the end of their runs, so the second read is served by L2 (the neighbours run on one XCD).
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

// Two-row LDS rings of three planes, interleaved by column: the values of column c (every ring,
// both row parities, the three planes) are CS consecutive floats (CS odd: lanes reading the same
// slot of consecutive columns hit distinct banks), so one base register per lane addresses every
// ring value and its left and right neighbours with immediate offsets.  Columns -1 and P are zero
// pads (the values of the missing neighbours at the grid's left and right edges).
#define FK_LDS __attribute__((address_space(3)))
template <int P, int NRINGS>
struct Rings {
    static constexpr int CS = 6 * NRINGS + 1;
    static constexpr int FLOATS = (P + 2) * CS;
    FK_LDS float *b;   // this lane's column: lds + (c + 1) * CS
    __device__ FK_LDS float &at(int ring, int par, int v, int dc) const { return b[dc * CS + ring * 6 + par * 3 + v]; }
};

// Model constants that meet a second scalar operand in a fused multiply-add, held in vector
// registers for the whole kernel (a VOP3 instruction reads at most one scalar register; the
// compiler would otherwise copy them again for every cell evaluation).  The values are unchanged.
__device__ inline float pin_v(float x)
{
    asm volatile("" : "+v"(x));
    return x;
}
__device__ inline void pin_model(Model32 &m)
{
    m.bq = pin_v(m.bq);
    m.bs = pin_v(m.bs);
    m.dtv = pin_v(m.dtv);
    m.i_tau_r = pin_v(m.i_tau_r);
}

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

// Row accesses through buffer resources: a per-lane 32-bit byte offset (the column, a VGPR held
// for the whole kernel) and a wave-uniform byte offset of the row and plane (a scalar register),
// so an access costs no vector address arithmetic.  Every access is in bounds (rows are checked
// before, columns are the pitch), so the resource's range is the whole 31-bit offset space.
__device__ inline __amdgpu_buffer_rsrc_t buf_rsrc(const void *p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), 0, 0x7fffffff, 0x00020000);
}
// three planes (n cells apart) of row `row` of a state-shaped buffer; c4 = 4 x column
template <int P>
__device__ inline void ld3(__amdgpu_buffer_rsrc_t r, int row, int n, uint32_t c4, float *out)
{
#pragma unroll
    for (int v = 0; v < 3; v++)
        out[v] = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, c4, (v * n + row * P) * 4, 0));
}
template <int P>
__device__ inline void st3(__amdgpu_buffer_rsrc_t r, int row, int n, uint32_t c4, const float *val)
{
#pragma unroll
    for (int v = 0; v < 3; v++)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, val[v]), r, c4, (v * n + row * P) * 4, 0);
}
template <int P>
__device__ inline uint32_t ld_code(__amdgpu_buffer_rsrc_t r, int row, uint32_t c2)
{
    return (uint32_t)__builtin_amdgcn_raw_buffer_load_b16(r, c2, row * P * 2, 0);
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
    __amdgpu_buffer_rsrc_t yin, yout, kout, kout2, code;
    bool has_k, has_k2;      // slopes stored (MODE_REC, MODE_KEEP), and a second copy (first step of a segment)
    Rings<P, 4> L;        // rings 0..3: state y, stage inputs Y2, Y3, Y4
    uint32_t c4, c2;      // byte offsets of the lane's column in a float and in a code plane
    float y[6][3];        // state, rows t-3 .. t+1
    uint32_t w[6];        // packed lap weights, rows t-3 .. t+1
    float gh[6];
    float acc[4][3];      // RK sums, rows t-3 .. t
    float Y2[4][3], Y3[4][3], Y4[4][3];
    float yn[3];          // arriving row (t+1 at iteration t)
    uint32_t cn;
    int flip;             // y0 + y1 - 1: grid row of rolling row t in a strip rolled upwards

    // grid row of rolling row t: strips roll down (DIR = 1) or up (DIR = -1)
    template <int DIR>
    __device__ int prow(int t) const { return DIR > 0 ? t : flip - t; }

    template <int DIR>
    __device__ void load(int t)
    {
        const int row = prow<DIR>(t);
        if ((unsigned)row < (unsigned)rows) {
            ld3<P>(yin, row, n, c4, yn);
            cn = ld_code<P>(code, row, c2);
        } else {
            yn[0] = yn[1] = yn[2] = 0.0f;
            cn = 0u;
        }
    }
};

// One iteration: the arriving row t+1 enters the rings, the loads of row t+2 are issued, then
// stage 1 on row t, stage 2 on t-1, stage 3 on t-2, stage 4 on t-3.  PH = (t - tb) mod RF_U.
// Rows are rolling rows; DIR = -1 rolls upwards (the rolling row before is the grid row below,
// so the stencils take their top and bottom values from the rolling rows in that order).
template <int P, int PH, int DIR>
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
        F.L.at(0, p1, v, 0) = F.yn[v];
    }
    F.w[qa] = pack_lap(F.cn);
    F.gh[qa] = F.m.ghc * (float)code_gh(F.cn);
    F.template load<DIR>(t + 2);

    // stage 1 on row t
    {
        float f[3] = {0.0f, 0.0f, 0.0f};
        const int row = F.template prow<DIR>(t);
        if ((unsigned)row < (unsigned)F.rows && F.w[q0] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L.at(0, p0, v, -1); rt[v] = F.L.at(0, p0, v, 1); }
            rhs_vals(F.m, unpack_w(F.w[q0], F.gh[q0]), F.y[q0], DIR > 0 ? F.y[qm1] : F.y[qa],
                     DIR > 0 ? F.y[qa] : F.y[qm1], lf, rt, F.E1, f);
        }
        if (F.has_k && t >= F.y0 && t < F.y1) {
            st3<P>(F.kout, row, F.n, F.c4, f);
            if (F.has_k2) st3<P>(F.kout2, row, F.n, F.c4, f);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y2[r0][v] = F.y[q0][v] + F.a12 * f[v];
            F.acc[r0][v] = F.b1 * f[v];
        }
    }
    // stage 2 on row t-1: Y2 of rows t-2 (above), t-1 (left, right) from the ring, t (below) in registers
    if (t >= F.tb + 2) {
        const int row = F.template prow<DIR>(t - 1);
        float f[3] = {0.0f, 0.0f, 0.0f};
        if ((unsigned)row < (unsigned)F.rows && F.w[qm1] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L.at(1, p1, v, -1); rt[v] = F.L.at(1, p1, v, 1); }
            rhs_vals(F.m, unpack_w(F.w[qm1], F.gh[qm1]), F.Y2[rm1], DIR > 0 ? F.Y2[rm2] : F.Y2[r0],
                     DIR > 0 ? F.Y2[r0] : F.Y2[rm2], lf, rt, F.E23, f);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y3[rm1][v] = F.y[qm1][v] + F.a12 * f[v];
            F.acc[rm1][v] = F.acc[rm1][v] + F.b23 * f[v];
        }
    }
    // stage 3 on row t-2
    if (t >= F.tb + 4) {
        const int row = F.template prow<DIR>(t - 2);
        float f[3] = {0.0f, 0.0f, 0.0f};
        if ((unsigned)row < (unsigned)F.rows && F.w[qm2] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L.at(2, p0, v, -1); rt[v] = F.L.at(2, p0, v, 1); }
            rhs_vals(F.m, unpack_w(F.w[qm2], F.gh[qm2]), F.Y3[rm2], DIR > 0 ? F.Y3[rm3] : F.Y3[rm1],
                     DIR > 0 ? F.Y3[rm1] : F.Y3[rm3], lf, rt, F.E23, f);
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            F.Y4[rm2][v] = F.y[qm2][v] + F.a3 * f[v];
            F.acc[rm2][v] = F.acc[rm2][v] + F.b23 * f[v];
        }
    }
    // stage 4 on row t-3 and the new state
    if (t >= F.tb + 6) {
        const int row = F.template prow<DIR>(t - 3);
        float f[3] = {0.0f, 0.0f, 0.0f};
        if (F.w[qm3] != 0u) {
            float lf[3], rt[3];
#pragma unroll
            for (int v = 0; v < 3; v++) { lf[v] = F.L.at(3, p1, v, -1); rt[v] = F.L.at(3, p1, v, 1); }
            rhs_vals(F.m, unpack_w(F.w[qm3], F.gh[qm3]), F.Y4[rm3], DIR > 0 ? F.Y4[rm4] : F.Y4[rm2],
                     DIR > 0 ? F.Y4[rm2] : F.Y4[rm4], lf, rt, F.E4, f);
        }
        const bool act = F.w[qm3] != 0u;
        float yo[3];
#pragma unroll
        for (int v = 0; v < 3; v++) yo[v] = act ? F.y[qm3][v] + (F.acc[rm3][v] + F.b4 * f[v]) : 0.0f;
        st3<P>(F.yout, row, F.n, F.c4, yo);
    }
    // this iteration's stage values for the neighbours' stencils in the next one: each ring slot
    // held the row two above, read above in this lane's own column only
#pragma unroll
    for (int v = 0; v < 3; v++) {
        F.L.at(1, p0, v, 0) = F.Y2[r0][v];
        if (t >= F.tb + 2) F.L.at(2, p1, v, 0) = F.Y3[rm1][v];
        if (t >= F.tb + 4) F.L.at(3, p0, v, 0) = F.Y4[rm2][v];
    }
    __syncthreads();
}

template <int P, int DIR>
__device__ __forceinline__ void roll_fwd_run(RollFwd<P> &F)
{
    // rows tb-1 and tb before the first iteration (row tb also for its neighbours), row tb+1 in flight
    F.template load<DIR>(F.tb - 1);
    constexpr int q_m1 = 5, q_0 = 0;   // ring slots of rows tb-1 and tb
#pragma unroll
    for (int v = 0; v < 3; v++) F.y[q_m1][v] = F.yn[v];
    F.w[q_m1] = pack_lap(F.cn);
    F.gh[q_m1] = F.m.ghc * (float)code_gh(F.cn);
    F.template load<DIR>(F.tb);
#pragma unroll
    for (int v = 0; v < 3; v++) {
        F.y[q_0][v] = F.yn[v];
        F.L.at(0, 0, v, 0) = F.yn[v];
    }
    F.w[q_0] = pack_lap(F.cn);
    F.gh[q_0] = F.m.ghc * (float)code_gh(F.cn);
    F.template load<DIR>(F.tb + 1);
    __syncthreads();
    const int te = F.y1 + 3;
    for (int t0 = F.tb; t0 < te; t0 += RF_U) {
        roll_fwd_iter<P, 0, DIR>(F, t0);
        if (t0 + 1 >= te) break;
        roll_fwd_iter<P, 1, DIR>(F, t0 + 1);
        if (t0 + 2 >= te) break;
        roll_fwd_iter<P, 2, DIR>(F, t0 + 2);
        if (t0 + 3 >= te) break;
        roll_fwd_iter<P, 3, DIR>(F, t0 + 3);
        if (t0 + 4 >= te) break;
        roll_fwd_iter<P, 4, DIR>(F, t0 + 4);
        if (t0 + 5 >= te) break;
        roll_fwd_iter<P, 5, DIR>(F, t0 + 5);
        if (t0 + 6 >= te) break;
        roll_fwd_iter<P, 6, DIR>(F, t0 + 6);
        if (t0 + 7 >= te) break;
        roll_fwd_iter<P, 7, DIR>(F, t0 + 7);
        if (t0 + 8 >= te) break;
        roll_fwd_iter<P, 8, DIR>(F, t0 + 8);
        if (t0 + 9 >= te) break;
        roll_fwd_iter<P, 9, DIR>(F, t0 + 9);
        if (t0 + 10 >= te) break;
        roll_fwd_iter<P, 10, DIR>(F, t0 + 10);
        if (t0 + 11 >= te) break;
        roll_fwd_iter<P, 11, DIR>(F, t0 + 11);
    }
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
    const float *yin;
    float *yout, *kout = S.kT, *kout2 = S.kT;   // kT: a valid address for the unused resources
    F.has_k = F.has_k2 = false;
    if (F.mode == MODE_FWD) {
        yin = (step % S.K == 0) ? S.ck + (size_t)(step / S.K) * st : S.ring + (size_t)(step & 1) * st;
        const int nx = step + 1;
        yout = (nx % S.K == 0 || nx == S.nt) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                              : S.ring + (size_t)(nx & 1) * st;
    } else {
        const int j = step - seg_a, segi = seg_a / S.K, buf = seg_buf(S, segi), nx = step + 1;
        float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
        yin = (j == 0) ? S.ck + (size_t)segi * st : segb + (size_t)j * st;
        yout = (F.mode == MODE_KEEP && (nx % S.K == 0 || nx == S.nt)) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                                                     : segb + (size_t)(j + 1) * st;
        kout = S.kseg + ((size_t)buf * S.K + j) * st;
        F.has_k = true;
        if (F.mode == MODE_REC && j == 0 && segi < S.keep0) {
            kout2 = S.kfirst + (size_t)(segi % 3) * st;
            F.has_k2 = true;
        }
    }
    F.yin = buf_rsrc(yin);
    F.yout = buf_rsrc(yout);
    F.kout = buf_rsrc(kout);
    F.kout2 = buf_rsrc(kout2);
    F.code = buf_rsrc(S.code);
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
    F.c = threadIdx.x;
    F.c4 = (uint32_t)F.c * 4u;
    F.c2 = (uint32_t)F.c * 2u;
    F.y0 = strip * Hs;
    F.y1 = min(F.y0 + Hs, S.rows);
    F.tb = F.y0 - 3;
    F.L.b = lds + (F.c + 1) * Rings<P, 4>::CS;
    pin_model(F.m);
    // zero pads: columns -1 and P
    constexpr int CSF = Rings<P, 4>::CS;
    for (int q = F.c; q < 2 * CSF; q += P) lds[q < CSF ? q : (P + 1) * CSF + q - CSF] = 0.0f;
    F.flip = F.y0 + F.y1 - 1;
    // odd strips roll upwards: the rows two neighbouring strips both read are then read by both at
    // about the same time, and the second read finds them in L2
    if (strip & 1) roll_fwd_run<P, -1>(F);
    else roll_fwd_run<P, 1>(F);
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
    int rows, n, c, y0, y1, tb, ntiles, row_i;
    __amdgpu_buffer_rsrc_t zin, yA, yB, kA, kB, zout, code;
    double *partial;
    Rings<P, 5> X;        // rings 0..4: z, y(t_{i-1} + dt/3), Z1, Z2, Z3
    uint32_t c4, c2;      // byte offsets of the lane's column in a float and in a code plane
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
    int flip;             // y0 + y1 - 1: grid row of rolling row t in a strip rolled upwards
    FK_LDS double *prs;   // this wavefront's row sums of the current partial block (upward strips)

    // grid row of rolling row t: strips roll down (DIR = 1) or up (DIR = -1)
    template <int DIR>
    __device__ int prow(int t) const { return DIR > 0 ? t : flip - t; }

    template <int DIR>
    __device__ void load(int t)   // the arrivals of iteration t: z of rolling row t+1, the rest of row t
    {
        const int rz = prow<DIR>(t + 1), row = prow<DIR>(t);
        if ((unsigned)rz < (unsigned)rows) {
            ld3<P>(zin, rz, n, c4, zn);
        } else {
            zn[0] = zn[1] = zn[2] = 0.0f;
        }
        if ((unsigned)row < (unsigned)rows) {
            ld3<P>(yA, row, n, c4, an);
            ld3<P>(yB, row, n, c4, bn);
            ld3<P>(kA, row, n, c4, kan);
            ld3<P>(kB, row, n, c4, kbn);
            cn = ld_code<P>(code, row, c2);
        } else {
#pragma unroll
            for (int v = 0; v < 3; v++) an[v] = bn[v] = kan[v] = kbn[v] = 0.0f;
            cn = 0u;
        }
    }
};

// Rows are rolling rows (t-1 is the row rolled before t); DIR = -1 rolls a strip upwards, so the
// rolling row before is the grid row below: every stencil takes its top and bottom values from
// the rolling rows in that order, with the same expressions as rolling down.
template <int P, int PH, int DIR>
__device__ __forceinline__ void roll_adj_iter(RollAdj<P> &A, int t)
{
    const int c = A.c;
    const Model32 &m = A.m;
    const float dt = A.dt;
    constexpr int q0 = PH % 6, qa = (PH + 1) % 6, qm1 = (PH + 5) % 6, qm2 = (PH + 4) % 6, qm3 = (PH + 3) % 6;
    constexpr int h0 = PH & 1, h1 = (PH + 1) & 1;   // mod-2 slots of rows t (t-2) and t-1; also LDS parities
    constexpr int p0 = h0, p1 = h1;
    // y_{i-1} of row t-3 for stage 4 (in the grid whenever stage 4 runs)
    if (t >= A.tb + 6) ld3<P>(A.yA, A.template prow<DIR>(t - 3), A.n, A.c4, A.ar);
    // arrivals: z of row t+1 (its ring slot held row t-1, whose neighbours' values were read in the
    // previous iteration), the code of row t
    const float zup[3] = {A.z[qm1][0], A.z[qm1][1], A.z[qm1][2]};
#pragma unroll
    for (int v = 0; v < 3; v++) {
        A.z[qa][v] = A.zn[v];
        A.X.at(0, p1, v, 0) = A.zn[v];
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
            lt[v] = m.lc[v] * stencil5(tw, DIR > 0 ? zup[v] : A.z[qa][v], DIR > 0 ? A.z[qa][v] : zup[v],
                                       A.X.at(0, p0, v, -1), A.X.at(0, p0, v, 1), A.z[q0][v]);
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
    A.template load<DIR>(t + 1);
    // stage 2 on row t-1: midpoint state yh = y4 + 27/64 dt f(y(t_{i-1} + dt/3)), J(yh)
    float Z2[3] = {0.0f, 0.0f, 0.0f};
    if (t >= A.tb + 2) {
        const uint32_t cc = A.cd[qm1];
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X.at(2, p1, v, 0);
            const float za = A.X.at(2, p0, v, 0);
            lt[v] = m.lc[v] * stencil5(tw, DIR > 0 ? za : Z1[v], DIR > 0 ? Z1[v] : za, A.X.at(2, p1, v, -1),
                                       A.X.at(2, p1, v, 1), zc[v]);
        }
        if (cc & C_ACTIVE) {
            float yc[3], top[3], lf[3], rt[3], f3[3], yh[3], o[3];
#pragma unroll
            for (int v = 0; v < 3; v++) {
                yc[v] = A.X.at(1, p1, v, 0);
                top[v] = A.X.at(1, p0, v, 0);
                lf[v] = A.X.at(1, p1, v, -1);
                rt[v] = A.X.at(1, p1, v, 1);
            }
            rhs_vals(m, lap_weights(m, cc), yc, DIR > 0 ? top : y3, DIR > 0 ? y3 : top, lf, rt, A.E3, f3);
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
        A.X.at(1, p0, v, 0) = y3[v];
        A.X.at(2, p0, v, 0) = Z1[v];
    }
    // stage 3 on row t-2 with J(yh) of that row
    float Z3[3] = {0.0f, 0.0f, 0.0f};
    if (t >= A.tb + 4) {
        const uint32_t cc = A.cd[qm2];
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3];
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X.at(3, p0, v, 0);
            const float za = A.X.at(3, p1, v, 0);
            lt[v] = m.lc[v] * stencil5(tw, DIR > 0 ? za : Z2[v], DIR > 0 ? Z2[v] : za, A.X.at(3, p0, v, -1),
                                       A.X.at(3, p0, v, 1), zc[v]);
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
        for (int v = 0; v < 3; v++) A.X.at(3, p1, v, 0) = Z2[v];
    // stage 4 on row t-3 with J(y_{i-1}); the new z and the row's z.gh sum
    if (t >= A.tb + 6) {
        const int row = A.template prow<DIR>(t - 3);
        const uint32_t cc = A.cd[qm3];
        const bool act = (cc & C_ACTIVE) != 0;
        const LapW tw = lapT_weights(cc);
        float lt[3], zc[3], zo[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int v = 0; v < 3; v++) {
            zc[v] = A.X.at(4, p1, v, 0);
            const float za = A.X.at(4, p0, v, 0);
            lt[v] = m.lc[v] * stencil5(tw, DIR > 0 ? za : Z3[v], DIR > 0 ? Z3[v] : za, A.X.at(4, p1, v, -1),
                                       A.X.at(4, p1, v, 1), zc[v]);
        }
        if (act) {
            Gates32 g = gates32(m, A.ar[0]);
            Jac32 J = jac32(m, g, A.ar[0], A.ar[1], A.ar[2]);
            float o[3];
            jtz(J, zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) zo[v] = A.z[qm3][v] + (A.dz[qm3][v] + (1.0f / 6.0f) * dt * (lt[v] + o[v]));
        }
        st3<P>(A.zout, row, A.n, A.c4, zo);
        // rows of a wavefront without ghost cells (most of them) add exactly 0.0: no reduction
        const double term = zgh_term(m, act, cc, zo[0]);
        const double rs = __builtin_amdgcn_ballot_w64(term != 0.0) ? wave_sum64(term) : 0.0;
        if ((c & 63) == 0) {
            const int j = row % PART_ROWS;
            auto *dst = gmem(A.partial) + (size_t)A.row_i * A.ntiles + (size_t)(row / PART_ROWS) * (P / 64) + (c >> 6);
            if (DIR > 0) {
                if (j == 0) A.tacc = 0.0;
                A.tacc += rs;
                if (j == PART_ROWS - 1 || row == A.rows - 1) *dst = A.tacc;
            } else {
                // rows of a block arrive from the last one up: keep them, add in row order at row 0
                A.prs[j] = rs;
                if (j == 0) {
                    const int nr = min(PART_ROWS, A.rows - row);
                    double s = 0.0;
                    for (int q = 0; q < nr; q++) s += A.prs[q];
                    *dst = s;
                }
            }
        }
    }
    if (t >= A.tb + 4)
#pragma unroll
        for (int v = 0; v < 3; v++) A.X.at(4, p0, v, 0) = Z3[v];
    __syncthreads();
}

template <int P, int DIR>
__device__ __forceinline__ void roll_adj_run(RollAdj<P> &A)
{
    // z of rows tb-1 and tb (row tb also for its neighbours' stencils), the arrivals of iteration tb
    A.template load<DIR>(A.tb - 2);   // z of row tb-1
#pragma unroll
    for (int v = 0; v < 3; v++) A.z[5][v] = A.zn[v];
    A.template load<DIR>(A.tb - 1);   // z of row tb
#pragma unroll
    for (int v = 0; v < 3; v++) {
        A.z[0][v] = A.zn[v];
        A.X.at(0, 0, v, 0) = A.zn[v];
    }
    A.template load<DIR>(A.tb);
    __syncthreads();
    const int te = A.y1 + 3;
    for (int t0 = A.tb; t0 < te; t0 += RA_U) {
        roll_adj_iter<P, 0, DIR>(A, t0);
        if (t0 + 1 >= te) break;
        roll_adj_iter<P, 1, DIR>(A, t0 + 1);
        if (t0 + 2 >= te) break;
        roll_adj_iter<P, 2, DIR>(A, t0 + 2);
        if (t0 + 3 >= te) break;
        roll_adj_iter<P, 3, DIR>(A, t0 + 3);
        if (t0 + 4 >= te) break;
        roll_adj_iter<P, 4, DIR>(A, t0 + 4);
        if (t0 + 5 >= te) break;
        roll_adj_iter<P, 5, DIR>(A, t0 + 5);
    }
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
    A.yA = buf_rsrc((i - 1 == seg_a) ? S.ck + (size_t)segi * st : segb + (size_t)(i - 1 - seg_a) * st);
    A.yB = buf_rsrc((i == b_end) ? S.ck + (size_t)(segi + 1) * st : segb + (size_t)(i - seg_a) * st);
    A.kA = buf_rsrc(kbuf + (size_t)(i - 1 - seg_a) * st);
    const float *knext = (b_end == S.nt) ? S.kT
                         : (segi + 1 >= S.keep0 ? S.kseg + (size_t)seg_buf(S, segi + 1) * S.K * st
                                                : S.kfirst + (size_t)((segi + 1) % 3) * st);
    A.kB = buf_rsrc((i < b_end) ? kbuf + (size_t)(i - seg_a) * st : knext);
    A.zin = buf_rsrc(S.zb + (size_t)(sc.zpar & 1) * st);
    A.zout = buf_rsrc(S.zb + (size_t)((sc.zpar & 1) ^ 1) * st);
    A.m = S.m;
    A.dt = S.dt;
    A.row_i = i - 1;
    A.E3 = (float)S.Ea[3 * A.row_i + 2];
    A.dEB = (i < S.nt) ? (float)(S.Ea[3 * A.row_i + 1] - S.Ef[3 * i]) : 0.0f;
    A.rows = S.rows;
    A.n = S.n;
    A.code = buf_rsrc(S.code);
    A.partial = S.partial;
    A.ntiles = S.ntiles;
    A.c = threadIdx.x;
    A.c4 = (uint32_t)A.c * 4u;
    A.c2 = (uint32_t)A.c * 2u;
    A.y0 = strip * Hs;
    A.y1 = min(A.y0 + Hs, S.rows);
    A.tb = A.y0 - 3;
    A.tacc = 0.0;
    A.X.b = lds + (A.c + 1) * Rings<P, 5>::CS;
    pin_model(A.m);
    // zero pads: columns -1 and P
    constexpr int CSA = Rings<P, 5>::CS;
    for (int q = A.c; q < 2 * CSA; q += P) lds[q < CSA ? q : (P + 1) * CSA + q - CSA] = 0.0f;
    A.flip = A.y0 + A.y1 - 1;
    A.prs = (FK_LDS double *)(lds + Rings<P, 5>::FLOATS) + (A.c >> 6) * PART_ROWS;
    // odd strips roll upwards: the rows two neighbouring strips both read (each one's halo) are then
    // read by both at about the same time, at the start or at the end of their runs, and the second
    // read finds them in L2
    if (strip & 1) roll_adj_run<P, -1>(A);
    else roll_adj_run<P, 1>(A);
}

// ------------------------------------------------------------------------------ launchers
// Dynamic LDS of the two kernels: 4 (forward) and 5 (adjoint) two-row rings of three planes; the
// adjoint also keeps PART_ROWS row sums per wavefront (FLOATS is even: the sums are 8-byte aligned).
template <int NW, int NR>
constexpr size_t roll_lds_bytes()
{
    return sizeof(float) * (size_t)Rings<64 * NW, NR>::FLOATS + (NR == 5 ? sizeof(double) * NW * PART_ROWS : 0);
}

template <int NW>
void roll_launch_nw(bool adj, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, int Hs,
                    hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map)), block(64 * NW);
    if (adj) {
        const size_t lds = roll_lds_bytes<NW, 5>();
        static const bool attr = hipFuncSetAttribute((const void *)k_roll_adj<NW>,
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) == hipSuccess;
        (void)attr;
        hipLaunchKernelGGL(k_roll_adj<NW>, grid, block, lds, stream, slots_dev, cur, map, Hs);
    } else {
        const size_t lds = roll_lds_bytes<NW, 4>();
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
