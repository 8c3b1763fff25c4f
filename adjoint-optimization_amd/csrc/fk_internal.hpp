// fk_internal.hpp -- shared declarations of the kernels and the host engine (not installed).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fk_model.hpp"

namespace fk {

constexpr int PLAIN_BX = 64;   // one wavefront per 64 contiguous cells of a row
constexpr int PLAIN_BY = 4;

struct GridArgs {
    int rows, pitch, n;        // n = rows * pitch; pitch is ncols rounded up to 64 (extra columns are holes)
    const uint16_t *code;
};

// Deterministic FP64 sum over an NT-thread block (wavefront shuffles, then the NT/64 wavefront
// partials added in order); the result is valid in thread 0.
template <int NT>
__device__ inline double block_sum(double v, int tid)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    __shared__ double part[NT / 64];
    if ((tid & 63) == 0) part[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0) {
#pragma unroll
        for (int w = 0; w < NT / 64; w++) s += part[w];
    }
    return s;
}
__device__ inline double block_sum_256(double v, int tid) { return block_sum<256>(v, tid); }

// Global-memory views of pointers that reach a kernel through a struct in memory.  Such pointers
// are generic (flat) to the compiler: flat accesses may alias LDS, so every LDS store orders them
// and each flat load also counts against the LDS wait counter.  Casting to the global address
// space lets loads batch across LDS traffic and use a scalar base with a 32-bit vector offset.
#define FK_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ inline const FK_GLOBAL T *gmem(const T *p) { return (const FK_GLOBAL T *)p; }
template <typename T>
__device__ inline FK_GLOBAL T *gmem(T *p) { return (FK_GLOBAL T *)p; }

// p[i] for a wave-uniform base p and an index whose byte offset fits in 32 bits (instance creation
// keeps all three planes of a state under 2^31 bytes): the address is the scalar base plus a
// zero-extended 32-bit vector offset, so an access costs no 64-bit address arithmetic per lane.
// The same wave-uniform pointer, held in scalar registers and opaque to the compiler's address
// reassociation: plane bases formed with it stay scalar, so ld_at/st_at on them keep the
// "scalar base + vector offset" form instead of 64-bit per-lane additions.
template <typename T>
__device__ inline T *uniform_ptr(T *p)
{
    uint64_t a = (uint64_t)p;
    asm("" : "+s"(a));
    return (T *)a;
}
template <typename T>
__device__ inline T ld_at(const FK_GLOBAL T *p, uint32_t i)
{
    return *(const FK_GLOBAL T *)((const FK_GLOBAL char *)p + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ inline void st_at(FK_GLOBAL T *p, uint32_t i, T v)
{
    *(FK_GLOBAL T *)((FK_GLOBAL char *)p + i * (uint32_t)sizeof(T)) = v;
}

// ---- plain path context
template <typename T, typename M>
struct PlainCtx {
    GridArgs g;
    M m;
    double dt;
    const double *Ef, *Ea;          // stage-time field tables, 3 per step
    T *w0, *w1, *w2, *w3, *w4, *w5; // scratch planes (3n each)
    hipStream_t stream;
};

template <typename T, typename M>
void plain_forward_step(const PlainCtx<T, M> &c, const T *y_in, T *y_out, int step);
template <typename T, typename M>
void plain_adjoint_step(const PlainCtx<T, M> &c, const T *yA, const T *yB, const T *z_in, T *z_out, int row,
                        double *partial_row);
template <typename T, typename M>
void plain_zgh(const PlainCtx<T, M> &c, const T *z, double *partial_row);

// ---- fused FP32 path: one launch per time step over the tiles of one or more instances
constexpr int FUSED_TX = 64;    // tile width: one wavefront row
constexpr int FUSED_HALO = 4;   // one cell per RK stage
// Rows of the cell blocks that own one z.gh partial each (64 x PART_ROWS cells, numbered row-major): every
// adjoint kernel adds a block's per-row wavefront sums in row order from 0.0, whatever its tile.
constexpr int PART_ROWS = 8;
// tile configurations (rows x threads): 1 = 16 x 256, 2 = 32 x 256, 3 = 32 x 512, 4 = 16 x 512, 5 = 32 x 1024,
// 6 = 8 x 512, 7 = 8 x 256
constexpr int TILE_16_256 = 1, TILE_32_256 = 2, TILE_32_512 = 3, TILE_16_512 = 4, TILE_32_1024 = 5, TILE_8_512 = 6,
              TILE_8_256 = 7;
constexpr int NUM_TILE_CFG = 8;
__host__ __device__ constexpr int tile_rows_of(int cfg)
{
    return (cfg == TILE_16_256 || cfg == TILE_16_512) ? 16 : ((cfg == TILE_8_512 || cfg == TILE_8_256) ? 8 : 32);
}

struct FusedSlot {
    int rows, pitch, n;
    int nt, K, nck;             // steps, checkpoint interval, checkpoints
    int ntiles;                 // z.gh partial blocks of this instance (64 x PART_ROWS cells each; partial row length)
    int seg_par;                // 1: seg and kseg hold two buffers, chosen by segment parity; 0: one
    int keep0;                  // first kept segment (segments keep0.. have buffers of their own,
                                // filled by the forward sweep); the segment count when none is kept
    const uint16_t *code;
    float *ck;                  // nck states: ck[c] = state at step min(c*K, nt)
    float *seg;                 // (seg_par+1+kept) x (K+1) states of recomputed and kept segments
                                // (state 0 unused)
    float *ring;                // 2 states for the forward sweep between checkpoints
    float *zb;                  // 2 adjoint states (ping-pong)
    float *kseg;                // (seg_par+1+kept) x K slopes f(y_j) of recomputed and kept steps
    float *kfirst;              // 3 slopes: f(y_a) of the segment starting at a, by segment index mod 3
    float *kT;                  // slope f(y_nt) at the final time with the adjoint's field value
    const double *Ef, *Ea;      // stage field tables, 3 per step
    double *partial;            // (nt) rows x ntiles: zgh partials of adjoint steps
    Model32 m;
    float dt;
};


// Segment buffer of segment segi: a kept segment's own, else a recompute buffer by parity.
__host__ __device__ inline int seg_buf(const FusedSlot &S, int segi)
{
    return segi >= S.keep0 ? S.seg_par + 1 + (segi - S.keep0) : (segi & S.seg_par);
}

// MODE_FWD: forward sweep (checkpoints + ring); MODE_REC: segment recompute, also stores the
// stage-1 slopes for the adjoint's Hermite reconstruction; MODE_KEEP: forward-sweep step of a kept
// segment (states and slopes into the segment's own buffers as MODE_REC, the segment's end state to
// its checkpoint); MODE_RHS: only f(y_nt) -> kT; MODE_ADJ: adjoint step; MODE_ADJ2: two adjoint steps
// (i and i - 1, inside one segment); MODE_IDLE: the slot's tiles do nothing in this launch.
enum FusedMode { MODE_IDLE = -1, MODE_FWD = 0, MODE_REC = 1, MODE_RHS = 2, MODE_ADJ = 3, MODE_KEEP = 4, MODE_ADJ2 = 5 };

// Per-slot work of one launch, passed by value (1 KiB of kernel arguments): slots of one launch
// may be at different steps and phases (continuous batching).  step is the forward step index
// (MODE_FWD/REC) or the adjoint interval end i (MODE_ADJ); seg_a is the first step of the segment.
constexpr int MAX_SLOTS = 64;
struct SlotCursor {
    int mode, step, seg_a;
    int zpar;   // adjoint launches: the ping-pong half of zb holding the input z (the output goes to the other)
};
struct Cursors {
    SlotCursor c[MAX_SLOTS];
};

// The tiles of one launch, passed by value: the launch's slots in order and the running count of
// their tiles (tile t of slot[i] is launch tile first[i] + t; an instance's tiles are numbered
// row-major, 64 x TY cells each).  first[count] is the grid size.  Only slots with work in the
// launch are listed, so no workgroup is spent on an idle instance.
struct LaunchMap {
    int count;
    int first[MAX_SLOTS + 1];
    int slot[MAX_SLOTS];
};
inline void launch_map_clear(LaunchMap &m)
{
    m.count = 0;
    m.first[0] = 0;
}
inline void launch_map_add(LaunchMap &m, int slot, int tiles)
{
    m.slot[m.count] = slot;
    m.first[m.count + 1] = m.first[m.count] + tiles;
    m.count++;
}
inline int launch_map_tiles(const LaunchMap &m) { return m.first[m.count]; }
// tiles of a rows x pitch instance in tile configuration cfg
inline int tiles_of(int rows, int pitch, int cfg) { return (pitch / FUSED_TX) * ((rows + tile_rows_of(cfg) - 1) / tile_rows_of(cfg)); }

// Launchers (stream-ordered, no synchronisation; nothing is launched for an empty map); cfg is a
// tile configuration.  Tiles whose slot cursor is not of the kernel's kind return at once.
void fused_forward_launch(int cfg, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur,
                          hipStream_t stream);
void fused_forward2_launch(const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, hipStream_t stream);
void fused_adjoint_launch(int cfg, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur,
                          hipStream_t stream);
void fused_adjoint2_launch(const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, hipStream_t stream);
// Row-rolling kernels (fk_kernels_roll.hip) for instances of one pitch (256, 384, 512, 768 or 1024):
// one workgroup per strip of Hs rows (launch tiles are strips), Hs a multiple of every member's
// adjoint tile rows; same results as the tiled kernels bit for bit.
bool roll_pitch_supported(int pitch);
void roll_launch(bool adjoint, int pitch, const FusedSlot *slots_dev, const LaunchMap &map, const Cursors &cur, int Hs,
                 hipStream_t stream);

// ---- common small kernels (fk_kernels_common.hip)
// z (terminal adjoint, T planes) and L0 partials from y(T): CSR S over packed indices or s_diag*I.
template <typename T>
void objective_launch(int n_num, const int64_t *rowptr, const int32_t *colind, const double *val, double s_diag,
                      const int32_t *cell_of, int n, const T *y, const double *gamma3, T *z, double *L0_part,
                      hipStream_t stream);
int objective_blocks(int n_num);
// out[r] = sum_{b < nb} part[r*nb + b] in increasing b, r < nrows_out.
void sum_rows_launch(const double *part, int nb, int nrows_out, double *out, hipStream_t stream);
template <typename T>
void pack_launch(const T *planes, const int32_t *cell_of, int n_num, int n, double *packed, hipStream_t stream);
template <typename T>
void unpack_launch(const double *packed, const int32_t *cell_of, int n_num, int n, T *planes, hipStream_t stream);
template <typename T>
void stimulate_launch(T *planes, const uint8_t *mask, int n, T value, hipStream_t stream);

}  // namespace fk
