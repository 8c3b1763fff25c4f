// fk_kernels_fused.hip -- FP32 fused time-step kernels over the tiles of one or many instances.
//
// One launch advances every listed tile by one full time step (the forward sweep of small launches by
// two, see k_fused_fwd's D):
//   forward  : the four RK4 stages of G:460-563 (G = src/cuda/fk_grad_Et_mono_PARAMETERS.cu)
//   adjoint  : Hermite midpoint (G:1244-1301), the four adjoint RK4 stages (G:1349-1427) and the
//              z.gh partial (G:1457-1483)
// A 64 x TY output tile is staged in LDS with a 4-cell halo (one cell per RK stage); stage k is
// evaluated on the tile grown by 4-k cells, so the tile needs no neighbour communication inside a
// step and HBM sees one read of the inputs and one write of the outputs per step.
// Work split: thread t owns the tile cells t + NT*s (one wavefront = one 64-cell row) and halo
// cells.  Stage k is needed at the cells within Manhattan distance 4-k of the tile (a stage-k value
// reads the stage-(k-1) values of its four neighbours), so the halo cells computed are the shells
// m = 1, 2, 3 of cells at distance m: four straight bands beside the tile edges plus m-1 cells per
// corner (the square rings' far corner cells are never read).  The shells are numbered in one
// sequence (shell 1 first) and dealt out NT at a time, every other pass in reverse thread order, so
// each wavefront carries about the same share of halo work (a shell-m cell takes part in stages
// 1..4-m).
// The ownership is the same in every stage, so per-cell values (state, RK increments, Jacobians)
// stay in registers across the stages; LDS only holds the values neighbours read.
// Multiply-adds are fused only inside one source expression (as written), never across
// statements: a cell's arithmetic is then the same whichever slot (tile or halo shell) computes it,
// so results do not depend on the tile configuration.
#pragma clang fp contract(on)

#include "fk_internal.hpp"
#include "fk_tile_map.hpp"
#include "fk_fused_math.hpp"

namespace fk {

// sum of weighted neighbours minus 4 x centre (grid units) from the LDS tile
template <int RY, int RX>
__device__ inline float stencil(const float (*s)[RY][RX], int v, int ly, int lx, const LapW &w, float ctr)
{
    return stencil5(w, s[v][ly - 1][lx], s[v][ly + 1][lx], s[v][ly][lx - 1], s[v][ly][lx + 1], ctr);
}

// f(y) + E gh at an active cell from the LDS tile
template <int RY, int RX>
__device__ inline void rhs_lds(const Model32 &m, const float (*s)[RY][RX], const LapW &w, int ly, int lx, float E,
                               float *f, float *yc)
{
    // the reaction first, then the stencils straight from LDS (fewer values live at once than
    // gathering the neighbours first; the same expressions as rhs_vals)
#pragma unroll
    for (int v = 0; v < 3; v++) yc[v] = s[v][ly][lx];
    Gates32 g = gates32(m, yc[0]);
    float fu, fv, fw;
    fion32(m, g, yc[0], yc[1], yc[2], fu, fv, fw);
    rhs_combine(m, w, E, fu, fv, fw, stencil<RY, RX>(s, 0, ly, lx, w, yc[0]), stencil<RY, RX>(s, 1, ly, lx, w, yc[1]),
                stencil<RY, RX>(s, 2, ly, lx, w, yc[2]), f);
}

// L^T z at an active cell from the LDS tile (missing neighbours have weight 0)
template <int RY, int RX>
__device__ inline void lapT_lds(const Model32 &m, const float (*s)[RY][RX], const LapW &w, int ly, int lx, float *lt,
                                float *zc)
{
#pragma unroll
    for (int v = 0; v < 3; v++) {
        zc[v] = s[v][ly][lx];
        lt[v] = m.lc[v] * stencil<RY, RX>(s, v, ly, lx, w, zc[v]);
    }
}

// Launch tile of this workgroup (XCD-aware order: launch_tile_index).
struct TileRef {
    int slot, x0, y0, tile;   // instance slot, tile origin (cells), tile index within the instance
};

template <int TY>
__device__ inline TileRef launch_tile(const FusedSlot *__restrict__ slots, const LaunchMap &map)
{
    TileRef t;
    launch_tile_index(blockIdx.x, map, t.slot, t.tile);
    const int ntx = slots[t.slot].pitch / FUSED_TX;
    t.x0 = (t.tile % ntx) * FUSED_TX;
    t.y0 = (t.tile / ntx) * TY;
    return t;
}

// Region (tile + halo of H cells) into LDS, by RegionMap: lanes 0..63 load the first 64 cells of a
// row (coalesced), then the last 2H cells of 64/(2H) rows per pass.  Cells outside the grid read as
// 0 with code 0 (inactive).  All loads are issued before the first LDS store
// (their addresses are clamped into the grid, so every load is unconditional and in bounds).
template <class G>
__device__ inline void load_region(const FusedSlot &S, const float *__restrict__ src, int x0, int y0, int tid,
                                   float (*dst)[G::RY][G::RXP], uint16_t (*sc)[G::RXP])
{
    constexpr int RY = G::RY, NE = RegionMap<G>::NE;
    const int lane = tid & 63, wave = tid >> 6;
    const int gx0 = x0 - G::H, gy0 = y0 - G::H;
    const auto *g = gmem(src);
    const auto *gc = gmem(S.code);
    int lyv[NE], lxv[NE];
    bool in[NE];
    float val[NE][3];
    uint16_t cv[NE];
#pragma unroll
    for (int e = 0; e < NE; e++) {
        RegionMap<G>::pos(e, lane, wave, lyv[e], lxv[e]);
        const int gy = gy0 + lyv[e], gx = gx0 + lxv[e];
        in[e] = lyv[e] < RY && (unsigned)gy < (unsigned)S.rows && (unsigned)gx < (unsigned)S.pitch;
        const uint32_t k = in[e] ? gy * S.pitch + gx : 0;
        cv[e] = ld_at(gc, k);
#pragma unroll
        for (int v = 0; v < 3; v++) val[e][v] = ld_at(uniform_ptr(g + (size_t)v * S.n), k);
    }
#pragma unroll
    for (int e = 0; e < NE; e++) {
        if (lyv[e] >= RY) continue;
        if (sc) sc[lyv[e]][lxv[e]] = in[e] ? cv[e] : (uint16_t)0;
#pragma unroll
        for (int v = 0; v < 3; v++) dst[v][lyv[e]][lxv[e]] = in[e] ? val[e][v] : 0.0f;
    }
}

// ------------------------------------------------------------------------------ forward
// WPE: the occupancy asked of the register allocator (waves per SIMD; 1 = no request).  D: time
// steps per launch (D > 1 only for the forward sweep, MODE_FWD and MODE_KEEP): the step-(st+1) state
// of every cell within distance 4(D-1-st) of the tile is formed in the launch, so D steps read and
// write HBM once (MODE_KEEP also stores the tile's intermediate states and every step's slope).
template <int TY, int NT, int WPE, int D = 1>
__global__ __launch_bounds__(NT, WPE) void k_fused_fwd(const FusedSlot *__restrict__ slots, Cursors cur, LaunchMap map)
{
    using G = Tile<TY, NT, D>;
    constexpr int RX = G::RXP, RY = G::RY, NS = G::NSLOT;
    constexpr int NA = D == 1 ? G::TS : NS;   // slots holding RK increments (cells that finish a step)
    __shared__ float sY[3][RY][RX];
    __shared__ uint16_t sC[RY][RX];

    const TileRef t = launch_tile<TY>(slots, map);
    const SlotCursor sc = cur.c[t.slot];
    if (sc.mode != MODE_FWD && sc.mode != MODE_KEEP && (D > 1 || (sc.mode != MODE_REC && sc.mode != MODE_RHS)))
        return;   // uniform per block
    const int mode = sc.mode, step = sc.step, seg_a = sc.seg_a;
    const FusedSlot &S = slots[t.slot];
    const int tid = threadIdx.x;
    const size_t st = (size_t)3 * S.n;
    const float *yin;
    float *yout;
    float *kout = nullptr;      // stage-1 slopes of the tile (MODE_REC, MODE_KEEP: one per step; MODE_RHS)
    float *kout2 = nullptr;     // second copy of the first slope of a segment (MODE_REC, j = 0)
    float *ymid = nullptr;      // MODE_KEEP with D > 1: the tile's states after the inner steps
    float E1, E23, E4;
    if (mode == MODE_FWD) {
        // between checkpoints the sweep alternates between the two ring states, launch by launch
        yin = (step % S.K == 0) ? S.ck + (size_t)(step / S.K) * st : S.ring + (size_t)((step / D) & 1) * st;
        int nx = step + D;
        yout = (nx % S.K == 0 || nx == S.nt) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                             : S.ring + (size_t)((nx / D) & 1) * st;
    } else if (mode == MODE_REC || mode == MODE_KEEP) {
        // with two buffers (seg_par 1) recomputed segments alternate between them, so the
        // recompute of segment s-1 may run while the adjoint steps of segment s read segment s's
        // buffers; the first slope of each recomputed segment is also kept by segment index mod 3
        // for the adjoint step that ends at it (the last one of the previous segment), which reads
        // a kept segment's first slope from that segment's own buffer
        // (D > 1: the D steps stay inside one segment, the host keeps the step and K multiples of D)
        const int j = step - seg_a, segi = seg_a / S.K, buf = seg_buf(S, segi), nx = step + D;
        float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
        yin = (j == 0) ? S.ck + (size_t)segi * st : segb + (size_t)j * st;
        yout = (mode == MODE_KEEP && (nx % S.K == 0 || nx == S.nt)) ? S.ck + (size_t)((nx + S.K - 1) / S.K) * st
                                                                   : segb + (size_t)(j + D) * st;
        kout = S.kseg + ((size_t)buf * S.K + j) * st;
        if (D > 1) ymid = segb + (size_t)(j + 1) * st;
        if (mode == MODE_REC && j == 0 && segi < S.keep0) kout2 = S.kfirst + (size_t)(segi % 3) * st;
    } else {
        yin = S.ck + (size_t)(S.nck - 1) * st;
        yout = nullptr;
        kout = S.kT;
    }
    if (mode == MODE_RHS) {
        E1 = E23 = E4 = (float)S.Ea[3 * (S.nt - 1) + 1];
    } else {
        E1 = (float)S.Ef[3 * step]; E23 = (float)S.Ef[3 * step + 1]; E4 = (float)S.Ef[3 * step + 2];
    }
    const Model32 m = S.m;
    const float dt = S.dt;

    load_region<G>(S, yin, t.x0, t.y0, tid, sY, sC);
    __syncthreads();

    int ly[NS], lx[NS], ls[NS];
    bool own[NS];
    uint32_t cc[NS];
    LapW lw[NS];
    float y[NS][3], acc[NA][3], f[NS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        own[sl] = slot_cell<G>(sl, tid, ly[sl], lx[sl], ls[sl]);
        if (!own[sl]) { ly[sl] = G::H; lx[sl] = G::H; }
        cc[sl] = sC[ly[sl]][lx[sl]];
        lw[sl] = lap_weights(m, cc[sl]);
#pragma unroll
        for (int v = 0; v < 3; v++) y[sl][v] = sY[v][ly[sl]][lx[sl]];
    }

#pragma unroll
    for (int ss = 0; ss < D; ss++) {
        if (ss > 0) {
            E1 = (float)S.Ef[3 * (step + ss)];
            E23 = (float)S.Ef[3 * (step + ss) + 1];
            E4 = (float)S.Ef[3 * (step + ss) + 2];
        }
#pragma unroll
        for (int k = 1; k <= 4; k++) {
            const int gs = 4 * ss + k;   // stage of the launch
            const float E = (k == 1) ? E1 : (k == 4 ? E4 : E23);
#pragma unroll
            for (int sl = 0; sl < NS; sl++) {
                if (!slot_in_stage<G>(sl, gs)) continue;
                f[sl][0] = f[sl][1] = f[sl][2] = 0.0f;
                if (gs <= ls[sl] && (cc[sl] & C_ACTIVE)) {
                    float yc[3];
                    rhs_lds<RY, RX>(m, sY, lw[sl], ly[sl], lx[sl], E, f[sl], yc);
                }
            }
            if (k == 1 && kout) {
                // stage-1 slope f(y_j) of the tile, read back by the adjoint step
                float *ko = kout + (size_t)ss * st;
#pragma unroll
                for (int sl = 0; sl < G::TS; sl++) {
                    int gy = t.y0 + ly[sl] - G::H, gx = t.x0 + lx[sl] - G::H;
                    if (gy >= S.rows) continue;
                    int kk = gy * S.pitch + gx;
#pragma unroll
                    for (int v = 0; v < 3; v++) st_at(uniform_ptr(gmem(ko) + (size_t)v * S.n), kk, f[sl][v]);
                    if (kout2)
#pragma unroll
                        for (int v = 0; v < 3; v++) st_at(uniform_ptr(gmem(kout2) + (size_t)v * S.n), kk, f[sl][v]);
                }
                if (mode == MODE_RHS) return;   // uniform over the block: no barrier is left behind
            }
            if (k < 4) {
                const float a = (k == 3) ? dt : 0.5f * dt;
                const float b = (k == 1) ? dt * (1.0f / 6.0f) : dt * (1.0f / 3.0f);
                __syncthreads();
#pragma unroll
                for (int sl = 0; sl < NS; sl++) {
                    if (!slot_in_stage<G>(sl, gs) || gs > ls[sl]) continue;
#pragma unroll
                    for (int v = 0; v < 3; v++) sY[v][ly[sl]][lx[sl]] = y[sl][v] + a * f[sl][v];
                    if (sl < NA && slot_in_stage<G>(sl, 4 * ss + 4)) {
#pragma unroll
                        for (int v = 0; v < 3; v++) acc[sl][v] = (k == 1) ? b * f[sl][v] : acc[sl][v] + b * f[sl][v];
                    }
                }
                __syncthreads();
            } else if (ss + 1 < D) {
                // end of an inner step: the new state of every cell that took part in all four of
                // its stages, in registers and in LDS for the next step's stencils
                const float b4 = dt * (1.0f / 6.0f);
                __syncthreads();
#pragma unroll
                for (int sl = 0; sl < NA; sl++) {
                    if (!slot_in_stage<G>(sl, gs) || gs > ls[sl]) continue;
#pragma unroll
                    for (int v = 0; v < 3; v++) {
                        y[sl][v] = y[sl][v] + (acc[sl][v] + b4 * f[sl][v]);
                        sY[v][ly[sl]][lx[sl]] = y[sl][v];
                    }
                }
                if (ymid) {
                    // the tile's state after this inner step, for the adjoint (kept segments)
                    float *ym = ymid + (size_t)ss * st;
#pragma unroll
                    for (int sl = 0; sl < G::TS; sl++) {
                        int gy = t.y0 + ly[sl] - G::H, gx = t.x0 + lx[sl] - G::H;
                        if (gy >= S.rows) continue;
                        int kk = gy * S.pitch + gx;
#pragma unroll
                        for (int v = 0; v < 3; v++)
                            st_at(uniform_ptr(gmem(ym) + (size_t)v * S.n), kk, (cc[sl] & C_ACTIVE) ? y[sl][v] : 0.0f);
                    }
                }
                __syncthreads();
            }
        }
    }
    const float b4 = dt * (1.0f / 6.0f);
#pragma unroll
    for (int sl = 0; sl < G::TS; sl++) {
        int gy = t.y0 + ly[sl] - G::H, gx = t.x0 + lx[sl] - G::H;
        if (gy >= S.rows) continue;
        int kk = gy * S.pitch + gx;
#pragma unroll
        for (int v = 0; v < 3; v++)
            st_at(uniform_ptr(gmem(yout) + (size_t)v * S.n), kk,
                  (cc[sl] & C_ACTIVE) ? y[sl][v] + (acc[sl][v] + b4 * f[sl][v]) : 0.0f);
    }
}

// ------------------------------------------------------------------------------ adjoint
template <int TY, int NT, int WPE>
__global__ __launch_bounds__(NT, WPE) void k_fused_adj(const FusedSlot *__restrict__ slots, Cursors cur, LaunchMap map)
{
    using G = Tile<TY, NT>;
    constexpr int RX = G::RXP, RY = G::RY, NS = G::NSLOT, TS = G::TS;
    // Each stage writes its values to a buffer that no stage of the same phase reads, so one
    // barrier per stage suffices: z -> sA; y(t_{i-1} + dt/3) -> sB; Z_1 -> sX; Z_2 -> sA (z is read
    // only in stage 1); Z_3 -> sX (Z_1 is read only in stage 2).
    __shared__ float sA[3][RY][RX];
    __shared__ float sB[3][RY][RX];
    __shared__ float sX[3][RY][RX];
    __shared__ uint16_t sC[RY][RX];

    const TileRef t = launch_tile<TY>(slots, map);
    const SlotCursor sc = cur.c[t.slot];
    if (sc.mode != MODE_ADJ) return;   // uniform per block
    const int i = sc.step, seg_a = sc.seg_a;
    const FusedSlot &S = slots[t.slot];
    const int tid = threadIdx.x;
    const size_t st = (size_t)3 * S.n;
    const int segi = seg_a / S.K;
    const int b_end = (seg_a + S.K < S.nt) ? seg_a + S.K : S.nt;
    const int buf = seg_buf(S, segi);
    const float *kbuf = S.kseg + (size_t)buf * S.K * st;
    const float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
    const auto *yA = gmem((i - 1 == seg_a) ? S.ck + (size_t)segi * st : segb + (size_t)(i - 1 - seg_a) * st);
    const auto *yB = gmem((i == b_end) ? S.ck + (size_t)(segi + 1) * st : segb + (size_t)(i - seg_a) * st);
    const auto *kA = gmem(kbuf + (size_t)(i - 1 - seg_a) * st);
    // f(y_i): slope of the next recomputed step, the first slope of the later segment (from its own
    // buffer when it is kept), or f(y_nt)
    const float *knext = (b_end == S.nt) ? S.kT
                         : (segi + 1 >= S.keep0 ? S.kseg + (size_t)seg_buf(S, segi + 1) * S.K * st
                                                : S.kfirst + (size_t)((segi + 1) % 3) * st);
    const auto *kB = gmem((i < b_end) ? kbuf + (size_t)(i - seg_a) * st : knext);
    const float *zin = S.zb + (size_t)(sc.zpar & 1) * st;
    auto *zout = gmem(S.zb + (size_t)((sc.zpar & 1) ^ 1) * st);
    const Model32 m = S.m;
    const float dt = S.dt;
    const int row = i - 1;
    const float E3 = (float)S.Ea[3 * row + 2];
    // the stored slope at y_i used the forward field of step i; the reconstruction needs the field
    // of this interval (they differ at a piecewise-constant jump): f_u += (E_a - E_f) gh
    const float dEB = (i < S.nt) ? (float)(S.Ea[3 * row + 1] - S.Ef[3 * i]) : 0.0f;

    int ly[NS], lx[NS], kk[NS], ls[NS];
    bool own[NS], act[NS];
    uint32_t cc[NS];
    LapW tw[NS];
    // phase-1a inputs (y_{i-1}, y_i and their slopes) are requested before the region load; the
    // index is clamped into the grid and the values of cells outside it are never used
    float pa[NS][3], pb[NS][3], pka[NS][3], pkb[NS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        own[sl] = slot_cell<G>(sl, tid, ly[sl], lx[sl], ls[sl]);
        if (!own[sl]) { ly[sl] = G::H; lx[sl] = G::H; }
        kk[sl] = (t.y0 + ly[sl] - G::H) * S.pitch + (t.x0 + lx[sl] - G::H);
        const int q = min(max(kk[sl], 0), S.n - 1);
#pragma unroll
        for (int v = 0; v < 3; v++) {
            pa[sl][v] = yA[v * S.n + q];
            pb[sl][v] = yB[v * S.n + q];
            pka[sl][v] = kA[v * S.n + q];
            pkb[sl][v] = kB[v * S.n + q];
        }
    }

    load_region<G>(S, zin, t.x0, t.y0, tid, sA, sC);
    __syncthreads();

#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        cc[sl] = sC[ly[sl]][lx[sl]];
        act[sl] = own[sl] && (cc[sl] & C_ACTIVE);
        tw[sl] = lapT_weights(cc[sl]);
    }

    // phase 1a: y(t_{i-1}+dt/3) and the partial midpoint value on R1 from the stored end slopes
    float y4[NS][3], yb[NS][3], ya[TS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        float y3[3] = {0.0f, 0.0f, 0.0f};
        if (act[sl]) {
            const float gh = m.ghc * (float)code_gh(cc[sl]);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float a = pa[sl][v], b = pb[sl][v];
                float fb = pkb[sl][v];
                if (v == 0) fb += dEB * gh;
                float m0 = pka[sl][v] * dt, m1 = fb * dt;
                y3[v] = h3<float>(a, b, m0, m1);
                y4[sl][v] = h4p<float>(a, b, m0, m1);
                yb[sl][v] = b;
                if (sl < TS) ya[sl][v] = a;
            }
        }
        if (own[sl])
#pragma unroll
            for (int v = 0; v < 3; v++) sB[v][ly[sl]][lx[sl]] = y3[v];
    }

    // phase 1b: adjoint stage 1 on R1 with J(y_i)
    float z[NS][3], dz[TS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        float lt[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        z[sl][0] = z[sl][1] = z[sl][2] = 0.0f;
        if (sl < TS) dz[sl][0] = dz[sl][1] = dz[sl][2] = 0.0f;
        if (!own[sl]) continue;
        lapT_lds<RY, RX>(m, sA, tw[sl], ly[sl], lx[sl], lt, z[sl]);
        if (act[sl]) {
            Gates32 g = gates32(m, yb[sl][0]);
            Jac32 J = jac32(m, g, yb[sl][0], yb[sl][1], yb[sl][2]);
            jtz(J, z[sl], o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (sl < TS) dz[sl][v] = (1.0f / 6.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 2: midpoint state yh, J(yh), adjoint stage 2 on R2
    Jac32 Jh[NS];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 2) || ls[sl] < 2) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, tw[sl], ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            float f3[3], yc[3], yh[3];
            rhs_lds<RY, RX>(m, sB, lap_weights(m, cc[sl]), ly[sl], lx[sl], E3, f3, yc);
#pragma unroll
            for (int v = 0; v < 3; v++) yh[v] = y4[sl][v] + (float)(27.0 / 64.0) * (f3[v] * dt);
            Gates32 g = gates32(m, yh[0]);
            Jh[sl] = jac32(m, g, yh[0], yh[1], yh[2]);
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (sl < TS) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sA[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 3: adjoint stage 3 on R3 with J(yh)
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 3) || ls[sl] < 3) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sA, tw[sl], ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + dt * fz;
                if (sl < TS) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 4: adjoint stage 4 on the tile with J(y_{i-1}); new z and its z.gh partial: each wavefront
    // holds one tile row per slot; the row sums (wave_sum64) are added in row order from 0.0
    __shared__ double sP[TY];
#pragma unroll
    for (int sl = 0; sl < TS; sl++) {
        float lt[3], zc[3], o[3], zn[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, tw[sl], ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            Gates32 g = gates32(m, ya[sl][0]);
            Jac32 J = jac32(m, g, ya[sl][0], ya[sl][1], ya[sl][2]);
            jtz(J, zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) zn[v] = z[sl][v] + (dz[sl][v] + (1.0f / 6.0f) * dt * (lt[v] + o[v]));
        }
        int gy = t.y0 + ly[sl] - G::H;
        const bool in = gy < S.rows;
        if (in) {
#pragma unroll
            for (int v = 0; v < 3; v++) zout[v * S.n + kk[sl]] = zn[v];
        }
        const double rs = wave_sum64(zgh_term(m, act[sl] && in, cc[sl], zn[0]));
        if ((tid & 63) == 0) sP[ly[sl] - G::H] = rs;
    }
    __syncthreads();
    // one partial per 64 x PART_ROWS block of the tile (blocks below the grid are not written)
    static_assert(TY % PART_ROWS == 0, "tiles hold whole partial blocks");
    if (tid < TY / PART_ROWS) {
        const int brow = t.y0 / PART_ROWS + tid;
        if (brow * PART_ROWS < S.rows) {
            double s = 0.0;
            for (int j = 0; j < PART_ROWS; j++) s += sP[tid * PART_ROWS + j];
            gmem(S.partial)[(size_t)row * S.ntiles + (size_t)brow * (S.pitch / FUSED_TX) + t.x0 / FUSED_TX] = s;
        }
    }
}

// ------------------------------------------------------------------------------ adjoint, two steps
// Two adjoint steps per launch (MODE_ADJ2: i and then i - 1, both inside one segment) on a 64 x TY
// tile with an 8-cell halo, for launches too small to fill the GPU: stages 1..4 (step i) run on the
// cells within distance 7, 6, 5, 4 of the tile and stages 5..8 (step i - 1) within 3, 2, 1, 0, so the
// intermediate adjoint state z_{i-1} never leaves LDS.  J(y_{i-1}) of step i's stage 4 is the
// Jacobian of step i - 1's stage 1 and stays in registers.  Every cell's arithmetic is the one-step
// kernel's, expression for expression, and each step writes its z.gh partial per 64 x 8 block, so
// results equal two one-step launches bit for bit.
template <int TY, int NT, int WPE>
__global__ __launch_bounds__(NT, WPE) void k_fused_adj2(const FusedSlot *__restrict__ slots, Cursors cur, LaunchMap map)
{
    using G = Tile<TY, NT, 2>;
    constexpr int RX = G::RXP, RY = G::RY, NS = G::NSLOT, TS = G::TS;
    // z_i -> sA; y(t_{i-1} + dt/3) -> sB; Z_1 -> sX; Z_2 -> sA; Z_3 -> sX; z_{i-1} -> sA;
    // y(t_{i-2} + dt/3) -> sB; Z_1' -> sX; Z_2' -> sA; Z_3' -> sX (one barrier per stage)
    __shared__ float sA[3][RY][RX];
    __shared__ float sB[3][RY][RX];
    __shared__ float sX[3][RY][RX];
    __shared__ uint16_t sC[RY][RX];
    __shared__ double sP[TY];

    const TileRef t = launch_tile<TY>(slots, map);
    const SlotCursor sc = cur.c[t.slot];
    if (sc.mode != MODE_ADJ2) return;   // uniform per block
    const int i = sc.step, seg_a = sc.seg_a;   // steps i and i - 1 > seg_a
    const FusedSlot &S = slots[t.slot];
    const int tid = threadIdx.x;
    const size_t st = (size_t)3 * S.n;
    const int segi = seg_a / S.K;
    const int b_end = (seg_a + S.K < S.nt) ? seg_a + S.K : S.nt;
    const int buf = seg_buf(S, segi);
    const float *kbuf = S.kseg + (size_t)buf * S.K * st;
    const float *segb = S.seg + (size_t)buf * (S.K + 1) * st;
    // y_{i-2}, y_{i-1}, y_i and their slopes (y_{i-1} is inside the segment: i - 1 > seg_a)
    const auto *yA2 = gmem((i - 2 == seg_a) ? S.ck + (size_t)segi * st : segb + (size_t)(i - 2 - seg_a) * st);
    const auto *yA = gmem(segb + (size_t)(i - 1 - seg_a) * st);
    const auto *yB = gmem((i == b_end) ? S.ck + (size_t)(segi + 1) * st : segb + (size_t)(i - seg_a) * st);
    const auto *kA2 = gmem(kbuf + (size_t)(i - 2 - seg_a) * st);
    const auto *kA = gmem(kbuf + (size_t)(i - 1 - seg_a) * st);
    const float *knext = (b_end == S.nt) ? S.kT
                         : (segi + 1 >= S.keep0 ? S.kseg + (size_t)seg_buf(S, segi + 1) * S.K * st
                                                : S.kfirst + (size_t)((segi + 1) % 3) * st);
    const auto *kB = gmem((i < b_end) ? kbuf + (size_t)(i - seg_a) * st : knext);
    const float *zin = S.zb + (size_t)(sc.zpar & 1) * st;
    auto *zout = gmem(S.zb + (size_t)((sc.zpar & 1) ^ 1) * st);
    const Model32 m = S.m;
    const float dt = S.dt;
    // step i: interval (i-1, i); step i - 1: interval (i-2, i-1) (the stored end slope f(y_{i-1})
    // used the forward field of step i - 1)
    const float E3a = (float)S.Ea[3 * (i - 1) + 2];
    const float dEBa = (i < S.nt) ? (float)(S.Ea[3 * (i - 1) + 1] - S.Ef[3 * i]) : 0.0f;
    const float E3b = (float)S.Ea[3 * (i - 2) + 2];
    const float dEBb = (float)(S.Ea[3 * (i - 2) + 1] - S.Ef[3 * (i - 1)]);

    int ly[NS], lx[NS], ls[NS];
    uint32_t q[NS];
    bool own[NS], act[NS];
    uint32_t cc[NS];
    float pa[NS][3], pb[NS][3], pka[NS][3], pkb[NS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        own[sl] = slot_cell<G>(sl, tid, ly[sl], lx[sl], ls[sl]);
        if (!own[sl]) { ly[sl] = G::H; lx[sl] = G::H; }
        const int kk = (t.y0 + ly[sl] - G::H) * S.pitch + (t.x0 + lx[sl] - G::H);
        q[sl] = (uint32_t)min(max(kk, 0), S.n - 1);
#pragma unroll
        for (int v = 0; v < 3; v++) {
            pa[sl][v] = yA[v * S.n + q[sl]];
            pb[sl][v] = yB[v * S.n + q[sl]];
            pka[sl][v] = kA[v * S.n + q[sl]];
            pkb[sl][v] = kB[v * S.n + q[sl]];
        }
    }

    load_region<G>(S, zin, t.x0, t.y0, tid, sA, sC);
    __syncthreads();

#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        cc[sl] = sC[ly[sl]][lx[sl]];
        act[sl] = own[sl] && (cc[sl] & C_ACTIVE);
    }

    // ---- step i
    // phase 1a: y(t_{i-1}+dt/3) and the partial midpoint value from the stored end slopes
    float y4[NS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        float y3[3] = {0.0f, 0.0f, 0.0f};
        if (act[sl]) {
            const float gh = m.ghc * (float)code_gh(cc[sl]);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float a = pa[sl][v], b = pb[sl][v];
                float fb = pkb[sl][v];
                if (v == 0) fb += dEBa * gh;
                float m0 = pka[sl][v] * dt, m1 = fb * dt;
                y3[v] = h3<float>(a, b, m0, m1);
                y4[sl][v] = h4p<float>(a, b, m0, m1);
            }
        }
        if (own[sl])
#pragma unroll
            for (int v = 0; v < 3; v++) sB[v][ly[sl]][lx[sl]] = y3[v];
    }

    // phase 1b: stage 1 with J(y_i)
    float z[NS][3], dz[NS][3];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        float lt[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        z[sl][0] = z[sl][1] = z[sl][2] = 0.0f;
        if (slot_in_stage<G>(sl, 4)) dz[sl][0] = dz[sl][1] = dz[sl][2] = 0.0f;
        if (!own[sl]) continue;
        lapT_lds<RY, RX>(m, sA, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, z[sl]);
        if (act[sl]) {
            Gates32 g = gates32(m, pb[sl][0]);
            Jac32 J = jac32(m, g, pb[sl][0], pb[sl][1], pb[sl][2]);
            jtz(J, z[sl], o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (slot_in_stage<G>(sl, 4)) dz[sl][v] = (1.0f / 6.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 2: midpoint state yh, J(yh), stage 2
    Jac32 Jh[NS];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 2) || ls[sl] < 2) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            float f3[3], yc[3], yh[3];
            rhs_lds<RY, RX>(m, sB, lap_weights(m, cc[sl]), ly[sl], lx[sl], E3a, f3, yc);
#pragma unroll
            for (int v = 0; v < 3; v++) yh[v] = y4[sl][v] + (float)(27.0 / 64.0) * (f3[v] * dt);
            Gates32 g = gates32(m, yh[0]);
            Jh[sl] = jac32(m, g, yh[0], yh[1], yh[2]);
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (slot_in_stage<G>(sl, 4)) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sA[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 3: stage 3 with J(yh)
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 3) || ls[sl] < 3) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sA, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + dt * fz;
                if (slot_in_stage<G>(sl, 4)) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 4: stage 4 with J(y_{i-1}) (kept for step i - 1's stage 1); z_{i-1} into LDS and, for the
    // tile's rows, the z.gh sums of step i (row i - 1 of the partials)
    Jac32 Jm[NS];
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 4) || ls[sl] < 4) continue;
        float lt[3], zc[3], o[3], zn[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            float ya[3];   // y_{i-1}, again from memory (fewer registers held across the stages)
#pragma unroll
            for (int v = 0; v < 3; v++) ya[v] = yA[v * S.n + q[sl]];
            Gates32 g = gates32(m, ya[0]);
            Jm[sl] = jac32(m, g, ya[0], ya[1], ya[2]);
            jtz(Jm[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) zn[v] = z[sl][v] + (dz[sl][v] + (1.0f / 6.0f) * dt * (lt[v] + o[v]));
        }
#pragma unroll
        for (int v = 0; v < 3; v++) {
            sA[v][ly[sl]][lx[sl]] = zn[v];
            z[sl][v] = zn[v];
        }
        if (sl < TS) {
            const bool in = t.y0 + ly[sl] - G::H < S.rows;
            const double rs = wave_sum64(zgh_term(m, act[sl] && in, cc[sl], zn[0]));
            if ((tid & 63) == 0) sP[ly[sl] - G::H] = rs;
        }
    }
    __syncthreads();
    static_assert(TY % PART_ROWS == 0, "tiles hold whole partial blocks");
    if (tid < TY / PART_ROWS) {
        const int brow = t.y0 / PART_ROWS + tid;
        if (brow * PART_ROWS < S.rows) {
            double s = 0.0;
            for (int j = 0; j < PART_ROWS; j++) s += sP[tid * PART_ROWS + j];
            gmem(S.partial)[(size_t)(i - 1) * S.ntiles + (size_t)brow * (S.pitch / FUSED_TX) + t.x0 / FUSED_TX] = s;
        }
    }

    // ---- step i - 1
    // phase 5a: y(t_{i-2}+dt/3) and the partial midpoint value, on the cells of stages 5 and 6
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 5)) continue;
        float y3[3] = {0.0f, 0.0f, 0.0f};
        if (act[sl] && ls[sl] >= 5) {
            const float gh = m.ghc * (float)code_gh(cc[sl]);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float a = yA2[v * S.n + q[sl]], b = yA[v * S.n + q[sl]];
                float fb = kA[v * S.n + q[sl]];
                if (v == 0) fb += dEBb * gh;
                float m0 = kA2[v * S.n + q[sl]] * dt, m1 = fb * dt;
                y3[v] = h3<float>(a, b, m0, m1);
                y4[sl][v] = h4p<float>(a, b, m0, m1);
            }
        }
        if (own[sl] && ls[sl] >= 5)
#pragma unroll
            for (int v = 0; v < 3; v++) sB[v][ly[sl]][lx[sl]] = y3[v];
    }

    // phase 5b: stage 5 with J(y_{i-1})
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 5) || ls[sl] < 5) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        if (sl < TS) dz[sl][0] = dz[sl][1] = dz[sl][2] = 0.0f;
        lapT_lds<RY, RX>(m, sA, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            jtz(Jm[sl], z[sl], o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (sl < TS) dz[sl][v] = (1.0f / 6.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 6: midpoint state, J, stage 6
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 6) || ls[sl] < 6) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            float f3[3], yc[3], yh[3];
            rhs_lds<RY, RX>(m, sB, lap_weights(m, cc[sl]), ly[sl], lx[sl], E3b, f3, yc);
#pragma unroll
            for (int v = 0; v < 3; v++) yh[v] = y4[sl][v] + (float)(27.0 / 64.0) * (f3[v] * dt);
            Gates32 g = gates32(m, yh[0]);
            Jh[sl] = jac32(m, g, yh[0], yh[1], yh[2]);
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + 0.5f * dt * fz;
                if (sl < TS) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sA[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 7: stage 7
#pragma unroll
    for (int sl = 0; sl < NS; sl++) {
        if (!slot_in_stage<G>(sl, 7) || ls[sl] < 7) continue;
        float lt[3], zc[3], o[3], Zk[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sA, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            jtz(Jh[sl], zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) {
                float fz = lt[v] + o[v];
                Zk[v] = z[sl][v] + dt * fz;
                if (sl < TS) dz[sl][v] += (1.0f / 3.0f) * dt * fz;
            }
        }
#pragma unroll
        for (int v = 0; v < 3; v++) sX[v][ly[sl]][lx[sl]] = Zk[v];
    }
    __syncthreads();

    // phase 8: stage 8 on the tile with J(y_{i-2}); z_{i-2} and the z.gh partial of step i - 1
#pragma unroll
    for (int sl = 0; sl < TS; sl++) {
        float lt[3], zc[3], o[3], zn[3] = {0.0f, 0.0f, 0.0f};
        lapT_lds<RY, RX>(m, sX, lapT_weights(cc[sl]), ly[sl], lx[sl], lt, zc);
        if (act[sl]) {
            float ya[3];   // y_{i-2}
#pragma unroll
            for (int v = 0; v < 3; v++) ya[v] = yA2[v * S.n + q[sl]];
            Gates32 g = gates32(m, ya[0]);
            Jac32 J = jac32(m, g, ya[0], ya[1], ya[2]);
            jtz(J, zc, o);
#pragma unroll
            for (int v = 0; v < 3; v++) zn[v] = z[sl][v] + (dz[sl][v] + (1.0f / 6.0f) * dt * (lt[v] + o[v]));
        }
        const int gy = t.y0 + ly[sl] - G::H;
        const bool in = gy < S.rows;
        if (in) {
            const uint32_t kk = (uint32_t)(gy * S.pitch + t.x0 + lx[sl] - G::H);
#pragma unroll
            for (int v = 0; v < 3; v++) zout[v * S.n + kk] = zn[v];
        }
        const double rs = wave_sum64(zgh_term(m, act[sl] && in, cc[sl], zn[0]));
        if ((tid & 63) == 0) sP[ly[sl] - G::H] = rs;
    }
    __syncthreads();
    if (tid < TY / PART_ROWS) {
        const int brow = t.y0 / PART_ROWS + tid;
        if (brow * PART_ROWS < S.rows) {
            double s = 0.0;
            for (int j = 0; j < PART_ROWS; j++) s += sP[tid * PART_ROWS + j];
            gmem(S.partial)[(size_t)(i - 2) * S.ntiles + (size_t)brow * (S.pitch / FUSED_TX) + t.x0 / FUSED_TX] = s;
        }
    }
}

// ------------------------------------------------------------------------------ launchers
// Tile configurations: id -> (rows, threads).  1: 64x16 / 256, 2: 64x32 / 256, 3: 64x32 / 512,
// 4: 64x16 / 512, 5: 64x32 / 1024, 6: 64x8 / 512, 7: 64x8 / 256.  The occupancy requests (third
// template argument) are measured choices: the 64x16/512 forward step at 8 waves per SIMD (64
// VGPRs, two spilled dwords) runs 8% faster than at the 6 its 67 VGPRs allow; the 64x8/512 adjoint
// step fits 6 waves without spilling.  Requests that force larger spills (64x16/512 adjoint at 6
// waves: 190 us instead of 115 us per 16 x 512^2 launch) are not made; at 5 waves it would spill
// ten dwords and still hold two workgroups per CU (its 66.6 KB of LDS allow two).
void fused_forward_launch(int cfg, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur,
                          hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map));
    if (grid.x == 0) return;
    switch (cfg) {
    case 2: hipLaunchKernelGGL((k_fused_fwd<32, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    case 3: hipLaunchKernelGGL((k_fused_fwd<32, 512, 1>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 4: hipLaunchKernelGGL((k_fused_fwd<16, 512, 8>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 5: hipLaunchKernelGGL((k_fused_fwd<32, 1024, 6>), grid, dim3(1024), 0, stream, slots_dev, cur, map); break;
    case 6: hipLaunchKernelGGL((k_fused_fwd<8, 512, 1>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 7: hipLaunchKernelGGL((k_fused_fwd<8, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    default: hipLaunchKernelGGL((k_fused_fwd<16, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    }
}

// Two forward-sweep steps per launch (MODE_FWD and MODE_KEEP) on 64x8 tiles of 512 threads with an 8-cell
// halo: for launches too small to fill the GPU, where a launch costs more than its arithmetic.
void fused_forward2_launch(const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map));
    if (grid.x == 0) return;
    hipLaunchKernelGGL((k_fused_fwd<8, 512, 1, 2>), grid, dim3(512), 0, stream, slots_dev, cur, map);
}

// Two adjoint steps per launch (MODE_ADJ2) on 64x8 tiles of 512 threads with an 8-cell halo: for
// launches too small to fill the GPU, as the forward sweep's two-step launches.
void fused_adjoint2_launch(const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map));
    if (grid.x == 0) return;
    hipLaunchKernelGGL((k_fused_adj2<8, 512, 4>), grid, dim3(512), 0, stream, slots_dev, cur, map);
}

void fused_adjoint_launch(int cfg, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur,
                          hipStream_t stream)
{
    const dim3 grid(launch_map_tiles(map));
    if (grid.x == 0) return;
    switch (cfg) {
    case 2: hipLaunchKernelGGL((k_fused_adj<32, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    case 3: hipLaunchKernelGGL((k_fused_adj<32, 512, 1>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 4: hipLaunchKernelGGL((k_fused_adj<16, 512, 1>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 5: hipLaunchKernelGGL((k_fused_adj<32, 1024, 1>), grid, dim3(1024), 0, stream, slots_dev, cur, map); break;
    case 6: hipLaunchKernelGGL((k_fused_adj<8, 512, 6>), grid, dim3(512), 0, stream, slots_dev, cur, map); break;
    case 7: hipLaunchKernelGGL((k_fused_adj<8, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    default: hipLaunchKernelGGL((k_fused_adj<16, 256, 1>), grid, dim3(256), 0, stream, slots_dev, cur, map); break;
    }
}

}  // namespace fk
