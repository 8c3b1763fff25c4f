// tile_map_check.cpp -- host-side check of the fused kernels' index maps (csrc/fk_tile_map.hpp),
// the same functions the kernels call, for every tile configuration and time-blocking depth:
//   * ownership: every cell within Manhattan distance 4D-1 of the tile is owned by exactly one
//     (thread, slot), with last stage LAST - distance, and no other cell is;
//   * bounds: every owned cell's stencil (+-1 in each direction) stays inside the LDS region;
//   * closure: stage s at a cell reads its four neighbours' stage s-1 (or the state after the
//     previous inner step), and every such neighbour computes it;
//   * slot predicates: a slot holding a cell of stage s passes slot_in_stage(sl, s) (the kernels
//     skip whole slots on it), and a slot with a cell finishing an inner step keeps its increments;
//   * the region load map covers the LDS region exactly once;
//   * the XCD-aware launch map is a bijection from workgroups onto (slot, tile) for many maps;
//   * the batch split cuts members into contiguous, non-empty runs of about equal cells.
// Prints "tile maps ok" and exits 0 when everything holds; prints the failures otherwise.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "fk_tile_map.hpp"

using namespace fk;

static int fails = 0;
#define CHECK(cond, ...)                                                  \
    do {                                                                  \
        if (!(cond)) {                                                    \
            if (fails++ < 20) {                                           \
                std::fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
                std::fprintf(stderr, __VA_ARGS__);                        \
                std::fprintf(stderr, "\n");                               \
            }                                                             \
        }                                                                 \
    } while (0)

template <int TY, int NT, int D>
static void check_tile()
{
    using G = Tile<TY, NT, D>;
    constexpr int RY = G::RY, RX = G::RX, NS = G::NSLOT, H = G::H;
    std::vector<int> stage(RY * RX, -1);   // last stage of the owned cell, -1 if none
    std::vector<int> slot_max(NS, 0);
    for (int sl = 0; sl < NS; sl++)
        for (int tid = 0; tid < NT; tid++) {
            int ly, lx, ls;
            if (!slot_cell<G>(sl, tid, ly, lx, ls)) continue;
            CHECK(ly >= 1 && ly < RY - 1 && lx >= 1 && lx < RX - 1, "64x%d/%d D%d: cell (%d,%d) at the region edge",
                  TY, NT, D, ly, lx);
            if (ly < 0 || ly >= RY || lx < 0 || lx >= RX) continue;
            CHECK(stage[ly * RX + lx] < 0, "64x%d/%d D%d: cell (%d,%d) owned twice", TY, NT, D, ly, lx);
            stage[ly * RX + lx] = ls;
            if (ls > slot_max[sl]) slot_max[sl] = ls;
        }
    for (int y = 0; y < RY; y++)
        for (int x = 0; x < RX; x++) {
            int dy = H - y > 0 ? H - y : (y - (H + TY - 1) > 0 ? y - (H + TY - 1) : 0);
            int dx = H - x > 0 ? H - x : (x - (H + G::TX - 1) > 0 ? x - (H + G::TX - 1) : 0);
            const int d = dy + dx, want = d <= G::NSH ? G::LAST - d : -1;
            CHECK(stage[y * RX + x] == want, "64x%d/%d D%d: cell (%d,%d) stage %d, want %d", TY, NT, D, y, x,
                  stage[y * RX + x], want);
        }
    for (int y = 1; y < RY - 1; y++)
        for (int x = 1; x < RX - 1; x++) {
            const int ls = stage[y * RX + x];
            for (int s = 2; s <= ls; s++) {
                const int nb[4] = {(y - 1) * RX + x, (y + 1) * RX + x, y * RX + x - 1, y * RX + x + 1};
                for (int q = 0; q < 4; q++)
                    CHECK(stage[nb[q]] >= s - 1, "64x%d/%d D%d: stage %d at (%d,%d) reads an uncomputed neighbour",
                          TY, NT, D, s, y, x);
            }
        }
    for (int sl = 0; sl < NS; sl++) {
        for (int s = 1; s <= G::LAST; s++)
            if (slot_max[sl] >= s)
                CHECK(slot_in_stage<G>(sl, s), "64x%d/%d D%d: slot %d skips stage %d", TY, NT, D, sl, s);
        // a slot with a cell finishing inner step k (stage 4k+4) must pass that stage's predicate
        // and, at depth 1, be a tile slot (only tile slots keep increments there)
        for (int k = 0; k < D; k++)
            if (slot_max[sl] >= 4 * (k + 1)) {
                CHECK(slot_in_stage<G>(sl, 4 * k + 4), "64x%d/%d D%d: slot %d finishing step %d", TY, NT, D, sl, k);
                CHECK(D > 1 || sl < G::TS, "64x%d/%d: halo slot %d finishes a step", TY, NT, sl);
            }
    }
    std::vector<int> loaded(RY * RX, 0);
    for (int tid = 0; tid < NT; tid++)
        for (int e = 0; e < RegionMap<G>::NE; e++) {
            int ly, lx;
            RegionMap<G>::pos(e, tid & 63, tid >> 6, ly, lx);
            if (ly >= RY) continue;
            CHECK(ly >= 0 && lx >= 0 && lx < RX, "64x%d/%d D%d: load position (%d,%d)", TY, NT, D, ly, lx);
            if (ly >= 0 && lx >= 0 && lx < RX) loaded[ly * RX + lx]++;
        }
    for (int c = 0; c < RY * RX; c++)
        CHECK(loaded[c] == 1, "64x%d/%d D%d: region cell %d loaded %d times", TY, NT, D, c, loaded[c]);
}

static void check_launch_maps()
{
    unsigned seed = 12345;
    auto rnd = [&](int n) { seed = seed * 1103515245u + 12345u; return (int)((seed >> 8) % (unsigned)n); };
    for (int trial = 0; trial < 400; trial++) {
        LaunchMap map;
        launch_map_clear(map);
        const int count = 1 + rnd(trial < 200 ? 5 : MAX_SLOTS);
        std::vector<int> tiles(count);
        for (int i = 0; i < count; i++) {
            tiles[i] = 1 + rnd(trial % 3 == 0 ? 9 : 700);
            launch_map_add(map, (count - 1 - i) % MAX_SLOTS, tiles[i]);   // slots listed in any order
        }
        const int n = launch_map_tiles(map);
        std::vector<std::vector<int>> hit(MAX_SLOTS);
        for (int i = 0; i < count; i++) hit[map.slot[i]].assign(tiles[i], 0);
        for (int b = 0; b < n; b++) {
            int slot, tile;
            launch_tile_index(b, map, slot, tile);
            const bool ok = slot >= 0 && slot < MAX_SLOTS && tile >= 0 && tile < (int)hit[slot].size();
            CHECK(ok, "launch map %d: block %d -> slot %d tile %d", trial, b, slot, tile);
            if (ok) hit[slot][tile]++;
        }
        for (int s = 0; s < MAX_SLOTS; s++)
            for (size_t t = 0; t < hit[s].size(); t++)
                CHECK(hit[s][t] == 1, "launch map %d: slot %d tile %zu taken %d times", trial, s, t, hit[s][t]);
    }
}

// split_runs: contiguous, non-empty, covering runs; with two runs, or four members or more per run,
// every run ends within one member of its share of the cells
static void check_split_runs()
{
    unsigned seed = 777;
    auto rnd = [&](int n) { seed = seed * 1103515245u + 12345u; return (int)((seed >> 8) % (unsigned)n); };
    for (int trial = 0; trial < 2000; trial++) {
        const int n = 1 + rnd(64), k = 1 + rnd(n < 8 ? n : 8);
        std::vector<int64_t> cells(n);
        int64_t total = 0, biggest = 0;
        for (auto &c : cells) {
            c = trial % 4 == 0 ? 262144 : 4096 * (1 + rnd(256));
            total += c;
            biggest = c > biggest ? c : biggest;
        }
        std::vector<int> end(k);
        split_runs(cells.data(), n, k, end.data());
        int prev = 0;
        int64_t acc = 0;
        for (int j = 0; j < k; j++) {
            CHECK(end[j] > prev, "split %d: run %d of %d empty (n %d)", trial, j, k, n);
            for (int q = prev; q < end[j]; q++) acc += cells[q];
            const int64_t share = total * (j + 1) / k;
            // with few members per run a run may have to take one member whatever its size
            if (k <= 2 || n >= 4 * k)
                CHECK(acc - share <= biggest && share - acc <= biggest, "split %d: run %d ends %lld cells from its share",
                      trial, j, (long long)(acc - share));
            prev = end[j];
        }
        CHECK(end[k - 1] == n, "split %d: runs cover %d of %d members", trial, end[k - 1], n);
        if (trial % 4 == 0 && n % k == 0)
            for (int j = 0; j < k; j++)
                CHECK(end[j] == n / k * (j + 1), "split %d: equal members, run %d ends at %d", trial, j, end[j]);
    }
}

int main()
{
    check_tile<16, 256, 1>();
    check_tile<32, 256, 1>();
    check_tile<32, 512, 1>();
    check_tile<16, 512, 1>();
    check_tile<32, 1024, 1>();
    check_tile<8, 512, 1>();
    check_tile<8, 256, 1>();
    check_tile<8, 512, 2>();
    check_launch_maps();
    check_split_runs();
    if (fails) {
        std::fprintf(stderr, "%d failures\n", fails);
        return 1;
    }
    std::printf("tile maps ok\n");
    return 0;
}
