// fk_tile_map.hpp -- index maps of the fused kernels, shared by the kernels and a host-side check
// (tests/tile_map_check.cpp): which cells a thread owns (tile cells and halo shells), which LDS cell
// each lane loads, and which launch tile a workgroup takes.
#pragma once
#include "fk_internal.hpp"

namespace fk {

// A 64 x TY tile advanced by D time steps per launch: 4D RK stages, a halo of 4D cells.
template <int TY, int NT_, int D = 1>
struct Tile {
    static constexpr int TX = FUSED_TX, H = FUSED_HALO * D, NT = NT_;
    static constexpr int LAST = 4 * D;               // stages of a launch; the tile's cells take part in all
    static constexpr int NSH = LAST - 1;             // halo shells computed (shell m: stages 1..LAST-m)
    static constexpr int RX = TX + 2 * H, RY = TY + 2 * H;
    static constexpr int RXP = RX + 1;               // odd LDS row pitch: column walks hit distinct banks
    static constexpr int TS = TX * TY / NT;          // tile cells per thread
    static_assert(TX * TY % NT == 0, "tile must split evenly");
    // cells at Manhattan distance m from the tile: 2 TX + 2 TY in the bands, m - 1 per corner
    __host__ __device__ static constexpr int shell_count(int m) { return 2 * (TX + TY) + 4 * (m - 1); }
    static constexpr int RC = NSH * 2 * (TX + TY) + 2 * NSH * (NSH - 1);   // halo cells computed
    static constexpr int NRS = (RC + NT - 1) / NT;                         // halo slots per thread
    static constexpr int NSLOT = TS + NRS;
    // shell of the r-th halo cell in the shell-1, shell-2, ... sequence
    __host__ __device__ static constexpr int depth_of(int r)
    {
        int m = 1;
        while (m < NSH && r >= shell_count(m)) r -= shell_count(m++);
        return m;
    }
    // halo slot j holds sequence numbers NT*j .. NT*(j+1)-1: its smallest shell
    __host__ __device__ static constexpr int min_depth(int j) { return depth_of(NT * j); }
};

// Cell of slot `sl` of thread `tid` (LDS coordinates) and the last RK stage it takes part in
// (LAST for tile cells, LAST-m for a shell-m cell); returns false if the thread owns no cell in
// that slot.
template <class G>
__host__ __device__ inline bool slot_cell(int sl, int tid, int &ly, int &lx, int &last_stage)
{
    constexpr int TX = G::TX, TY = G::RY - 2 * G::H, H = G::H;
    if (sl < G::TS) {
        int lin = tid + G::NT * sl;
        ly = H + lin / TX;
        lx = H + lin % TX;
        last_stage = G::LAST;
        return true;
    }
    const int j = sl - G::TS;
    int e = (j & 1) ? G::NT * (j + 1) - 1 - tid : G::NT * j + tid;
    last_stage = 0;
    if (e >= G::RC) return false;
    int m = 1;
#pragma unroll
    for (int mm = 1; mm < G::NSH; mm++)
        if (m == mm && e >= G::shell_count(mm)) {
            e -= G::shell_count(mm);
            m = mm + 1;
        }
    last_stage = G::LAST - m;
    if (e < 2 * TX) {                      // bands above and below the tile
        const bool below = e >= TX;
        ly = below ? H + TY - 1 + m : H - m;
        lx = H + (below ? e - TX : e);
    } else if (e < 2 * (TX + TY)) {        // bands left and right of the tile
        const int e2 = e - 2 * TX;
        const bool right = e2 >= TY;
        ly = H + (right ? e2 - TY : e2);
        lx = right ? H + TX - 1 + m : H - m;
    } else {                               // corner q, offsets a (rows) and m - a (columns)
        const int c = e - 2 * (TX + TY);
        const int q = c / (m - 1), a = c - q * (m - 1) + 1;
        ly = (q & 2) ? H + TY - 1 + a : H - a;
        lx = (q & 1) ? H + TX - 1 + (m - a) : H - (m - a);
    }
    return true;
}

// Some cell of slot sl may take part in stage k (1..LAST): tile slots always, a halo slot if its
// smallest shell m has m <= LAST-k (which of its cells do is decided per thread).
template <class G>
__host__ __device__ constexpr bool slot_in_stage(int sl, int k)
{
    return sl < G::TS || G::min_depth(sl - G::TS) <= G::LAST - k;
}

// LDS region (tile + halo of H cells) load map: each wavefront takes whole rows; pass e < NR loads
// the first 64 cells of row wave + e*NW (lane = column), the tail passes the last 2H cells of
// 64/(2H) rows per wavefront.  Positions with ly >= RY are not loaded.
template <class G>
struct RegionMap {
    static constexpr int NW = G::NT / 64;
    static constexpr int TW = G::RX - 64, RPW = 64 / TW;     // tail width, tail rows per wavefront pass
    static_assert(TW * RPW == 64, "the tail width must divide a wavefront");
    static constexpr int NR = (G::RY + NW - 1) / NW;         // passes over the first 64 cells of a row
    static constexpr int NTL = (G::RY + RPW * NW - 1) / (RPW * NW);   // tail passes
    static constexpr int NE = NR + NTL;
    __host__ __device__ static inline void pos(int e, int lane, int wave, int &ly, int &lx)
    {
        if (e < NR) { ly = wave + e * NW; lx = lane; }
        else { ly = (e - NR) * RPW * NW + wave * RPW + lane / TW; lx = 64 + lane % TW; }
    }
};

// Launch tile of workgroup b.  Workgroups are dealt round-robin over the GPU's 8 XCDs (block b runs
// on XCD b % 8, each XCD with its own L2); XCD x takes a contiguous run of launch tiles, so tiles
// that share halo lines share an L2.  The map is a bijection for any grid size.  Returns the slot
// and the tile index within that slot's instance.
__host__ __device__ inline void launch_tile_index(int b, const LaunchMap &map, int &slot, int &tile)
{
    const int n = map.first[map.count];
    const int x = b & 7;
    int lin = b >> 3;
    for (int q = 0; q < x; q++) lin += (n - q + 7) >> 3;
    int lo = 0, hi = map.count - 1;   // last i with first[i] <= lin
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (map.first[mid] <= lin) lo = mid;
        else hi = mid - 1;
    }
    slot = map.slot[lo];
    tile = lin - map.first[lo];
}

// Cut n members with the given cell counts into k contiguous, non-empty runs of about equal cells
// (1 <= k <= n): member q goes to the current run unless its larger half lies beyond the run's share
// of the total, keeping one member for each later run.  end[j] is one past the last member of run j.
inline void split_runs(const int64_t *cells, int n, int k, int *end)
{
    int64_t total = 0;
    for (int q = 0; q < n; q++) total += cells[q];
    int q = 0;
    int64_t acc = 0;
    for (int j = 0; j < k; j++) {
        const int64_t share = total * (j + 1) / k;
        const int last = n - (k - 1 - j);
        const int first = q;
        while (q < last) {
            if (q > first && acc + cells[q] / 2 > share) break;
            acc += cells[q++];
        }
        end[j] = q;
    }
    end[k - 1] = n;
}

}  // namespace fk
