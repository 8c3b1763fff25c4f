// fk_engine.hip -- host side of libfkadj: problem set-up, stage-time field tables, the forward /
// recompute / adjoint sweeps over device-resident checkpoints, batched evaluation and the C ABI
// declared in include/fkadj.h.
//
// Reference host routine: fk_grad_Et_comp, src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1569-1938 (G).
// Differences by design: the trajectory never leaves HBM (the reference copies every state to
// host memory and back, G:1721-1723, G:1871-1872), it is kept as checkpoints every K steps and
// recomputed one segment at a time (or, within a caller's keep_bytes, kept whole for the last
// segments), the per-step z.gh partials stay on the device until one fixed-order reduction at the
// end, and errors are returned instead of exit().  Many instances share each launch (groups); a
// large batch call runs as two groups from two host threads on two streams, and a continuous-
// batching server admits requests by their buffers' bytes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "../../include/fkadj.h"
#include "fk_internal.hpp"
#include "fk_tile_map.hpp"

using namespace fk;

namespace {

thread_local char g_errmsg[256];

#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) {                                                                        \
            snprintf(g_errmsg, sizeof g_errmsg, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                     __LINE__);                                                                        \
            return FKADJ_ERR_DEVICE;                                                                   \
        }                                                                                              \
    } while (0)

#define CHK(expr)                      \
    do {                               \
        int rc_ = (expr);              \
        if (rc_ != FKADJ_OK) return rc_; \
    } while (0)

std::atomic<int> g_profiling{0};
std::mutex g_timing_mu;
fkadj_timing g_timing{};

// ------------------------------------------------------------------------ spline
// Second derivatives of the not-a-knot cubic spline through equally or unequally spaced knots
// (the reference uses Burkardt's spline_cubic_set with ibc = 3, G:1681, G:2223).  The two
// not-a-knot conditions are folded into the first and last interior equations and the
// remaining tridiagonal system is solved by elimination.
bool spline_not_a_knot(int n, const double *t, const double *y, std::vector<double> &ypp)
{
    if (n < 4) return false;
    int m = n - 2;
    std::vector<double> lo(m), di(m), up(m), rh(m);
    for (int r = 0; r < m; r++) {
        int i = r + 1;
        double h0 = t[i] - t[i - 1], h1 = t[i + 1] - t[i];
        lo[r] = h0 / 6.0;
        di[r] = (h0 + h1) / 3.0;
        up[r] = h1 / 6.0;
        rh[r] = (y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0;
    }
    double h0 = t[1] - t[0], h1 = t[2] - t[1];
    di[0] += lo[0] * (h0 + h1) / h1;
    up[0] -= lo[0] * h0 / h1;
    double ha = t[n - 2] - t[n - 3], hb = t[n - 1] - t[n - 2];
    di[m - 1] += up[m - 1] * (ha + hb) / ha;
    lo[m - 1] -= up[m - 1] * hb / ha;
    for (int r = 1; r < m; r++) {
        double f = lo[r] / di[r - 1];
        di[r] -= f * up[r - 1];
        rh[r] -= f * rh[r - 1];
    }
    ypp.assign(n, 0.0);
    ypp[m] = rh[m - 1] / di[m - 1];
    for (int r = m - 2; r >= 0; r--) ypp[r + 1] = (rh[r] - up[r] * ypp[r + 2]) / di[r];
    ypp[0] = ((h0 + h1) * ypp[1] - h0 * ypp[2]) / h1;
    ypp[n - 1] = ((ha + hb) * ypp[n - 2] - hb * ypp[n - 3]) / ha;
    return true;
}

double spline_eval(const double *t, const double *y, const std::vector<double> &ypp, int lo, double tv)
{
    int hi = lo + 1;
    double h = t[hi] - t[lo], a = tv - t[lo], b = t[hi] - tv;
    return (ypp[lo] * b * b * b + ypp[hi] * a * a * a) / (6.0 * h) + (y[lo] / h - ypp[lo] * h / 6.0) * b +
           (y[hi] / h - ypp[hi] * h / 6.0) * a;
}

// Active cells of an active_index grid (entries > 0), or -1 for a grid too small to hold a problem.
// The drop-in entry points compare it with the caller's n_num before anything reads the n_num-sized
// arrays (gh_num, S), whose lengths the caller sized by its own n_num (G:1759-1762).
int64_t count_active(const int32_t *active_index, int nrows, int ncols)
{
    if (nrows < 1 || ncols < 1) return -1;
    int64_t n = 0;
    for (int64_t k = 0; k < (int64_t)nrows * ncols; k++) n += active_index[k] > 0;
    return n;
}

bool all_finite(const double *a, size_t n)
{
    for (size_t i = 0; i < n; i++)
        if (!std::isfinite(a[i])) return false;
    return true;
}

}  // namespace

// ============================================================================ instance
// An instance is the resident problem: geometry codes, objective matrix, model constants and the
// initial state.  Everything one gradient evaluation writes lives in an EvalCtx, so several
// evaluations of one instance (a server's requests) can be in flight at once.
// Entry points that allocate host memory turn a failed allocation (or any other C++ exception)
// into an error code: no exception crosses the C ABI.
#define FKADJ_ENTRY try {
#define FKADJ_EXIT                                \
    }                                             \
    catch (const std::bad_alloc &) {              \
        return FKADJ_ERR_NOMEM;                   \
    }                                             \
    catch (...) {                                 \
        return FKADJ_ERR_DEVICE;                  \
    }

struct EvalCtx;

struct fkadj_instance {
    fkadj_problem p{};
    int rows = 0, cols = 0, pitch = 0, n = 0, n_num = 0;
    int nt = 0, K = 0, nck = 0, ncontrols = 0;
    int prec = FKADJ_F64;
    size_t esz = 8;        // bytes per state element
    size_t st_bytes = 0;   // bytes of one 3-plane state
    Model64 m64{};
    Model32 m32{};
    std::vector<int32_t> cell_of;
    bool state_set = false;
    bool lazy_ws = false;    // the default evaluation context is allocated on first use
    int busy = 0;            // server requests in flight on this instance
    int plans = 0;           // batch plans holding this instance

    uint16_t *d_code = nullptr;
    int32_t *d_cell_of = nullptr;
    int64_t *d_rowptr = nullptr;
    int32_t *d_colind = nullptr;
    double *d_val = nullptr;
    double *d_gamma = nullptr;
    char *d_y0 = nullptr;
    double *d_packed = nullptr;
    uint8_t *d_mask = nullptr;
    hipStream_t stream = nullptr;
    int nb_plain = 0;     // plain blocks (zgh partial row length of the plain path)
    int nb_part = 0;      // partial row length actually used by the adjoint sweep
    int fwd_cfg = 0, adj_cfg = TILE_16_512;   // fused tile configurations (fwd 0: per group)
    int fwd_depth = 0, adj_depth = 0;         // time steps per tiled launch (0: built-in rule)
    int roll_rows = 0;                        // strip rows of the row-rolling kernels (0: library setting)
    int keep = 0;         // FP32: last segments the default context keeps from the forward sweep
    int device = 0;       // GPU the instance lives on (current device at create time)

    EvalCtx *ctx = nullptr;   // default context of fkadj_grad / fkadj_forward / batches

    char *state(char *base, int idx) const { return base + (size_t)idx * st_bytes; }
};

struct EvalCtx {
    fkadj_instance *I = nullptr;
    double *d_Ef = nullptr, *d_Ea = nullptr;          // stage-time field tables, 3 per step
    double *d_part = nullptr, *d_partT = nullptr;     // z.gh partials: adjoint steps, final time
    double *d_zgh = nullptr, *d_L0part = nullptr, *d_L0 = nullptr;
    char *d_ck = nullptr, *d_seg = nullptr, *d_ring = nullptr, *d_z = nullptr, *d_work = nullptr;
    char *d_kseg = nullptr, *d_kT = nullptr;          // FP32: stored slopes of the recompute sweep
    char *d_kfirst = nullptr;                         // FP32: first slope of each segment, mod 3
    int seg_copies = 1;                               // FP32: segment state / slope buffers (1 or 2)
    int keep = 0;                                     // FP32: kept segments (buffers of their own)
    FusedSlot slot_h{};
    struct FusedGroupDev *group = nullptr;            // this context alone (FP32)
    std::vector<double> Ef, Ea;                       // host tables of the last evaluation
    int64_t bytes = 0;
};

// A set of FP32 evaluations advanced together: one forward launch per tile configuration and one
// adjoint launch per time step cover all members.  A single evaluation is a group of one.
struct FusedGroupDev {
    std::vector<EvalCtx *> members;
    FusedSlot *d_slots = nullptr;
    // the adjoint steps of segment s run on `side` while the caller's stream recomputes segment
    // s-1; ev_rec[p] / ev_adj[p]: recompute / adjoint of the last segment of parity p done
    hipStream_t side = nullptr;
    hipEvent_t ev_rec[2] = {}, ev_adj[2] = {};
    // one launch map per kernel configuration (index = tile config id); slot k = member k
    LaunchMap fwd[NUM_TILE_CFG], adj[NUM_TILE_CFG];
    // forward sweep two steps per launch (64x8 tiles, all members) instead of one; the adjoint
    // steps likewise (two steps of one segment per launch, 64x8 tiles), where adj_depth is 2
    bool depth2 = false;
    LaunchMap fwd2;
    int adj_depth = 1;
    LaunchMap adj2;
    // segments from keep0 on are kept by every member: the forward sweep fills them and the
    // backward pass does not recompute them
    int keep0 = 0;
    // a group of a split batch runs on a batch lane stream (from a host thread of its own), forked
    // from and joined to the caller's stream through ev_lane
    hipEvent_t ev_lane = nullptr;
    // the members of each pitch that the row-rolling kernels support, when the group holds 2M cells
    // or more of that pitch, run them for their forward-kind and adjoint steps (one launch per such
    // pitch, one workgroup per strip of roll_hs[r] rows); the other members run the tiled kernels
    // (fwd_rest, adj_rest).  The same results either way.  f(y_nt) (MODE_RHS) always takes the tiled
    // maps of all members.
    static constexpr int kMaxRoll = 5;
    int nroll = 0;
    int roll_hs[kMaxRoll] = {};
    int roll_pitch[kMaxRoll] = {};
    LaunchMap rollmap[kMaxRoll];
    LaunchMap fwd_rest[NUM_TILE_CFG], adj_rest[NUM_TILE_CFG];
};

namespace {

template <typename T>
T *as(char *p) { return reinterpret_cast<T *>(p); }

void group_free(FusedGroupDev *g);
int group_build(const std::vector<EvalCtx *> &members, hipStream_t s, FusedGroupDev **out, int concurrent = 1);

void fill_models(fkadj_instance *I)
{
    const fkadj_model &md = I->p.model;
    Model64 &a = I->m64;
    a.u_c = md.u_c; a.u_v = md.u_v; a.tau_d = md.tau_d; a.tau_v1_minus = md.tau_v1_minus;
    a.tau_v2_minus = md.tau_v2_minus; a.tau_v_plus = md.tau_v_plus; a.tau_0 = md.tau_0; a.tau_r = md.tau_r;
    a.tau_si = md.tau_si; a.u_csi = md.u_csi; a.k_fk = md.k_fk; a.tau_w_minus = md.tau_w_minus;
    a.tau_w_plus = md.tau_w_plus; a.ksm = md.ksm;
    a.lc[0] = md.Du / (md.h * md.h);
    a.lc[1] = md.Dv / (md.h * md.h);
    a.lc[2] = md.Dw / (md.h * md.h);
    a.ghc = md.Du * 2 / md.h;

    Model32 &b = I->m32;
    const double L2E = 1.4426950408889634;
    b.kq = (float)(-2.0 * md.ksm * L2E);
    b.bq = (float)(2.0 * md.ksm * md.u_v * L2E);
    b.cpq = (float)std::exp(2.0 * md.ksm * (md.u_c - md.u_v));
    b.ks = (float)(-2.0 * md.k_fk * L2E);
    b.bs = (float)(2.0 * md.k_fk * md.u_csi * L2E);
    b.two_ksm = (float)(2.0 * md.ksm);
    b.two_k = (float)(2.0 * md.k_fk);
    b.u_c = (float)md.u_c;
    b.one_plus_uc = (float)(1.0 + md.u_c);
    b.i_tau_d = (float)(1.0 / md.tau_d);
    b.i_tau_0 = (float)(1.0 / md.tau_0);
    b.i_tau_r = (float)(1.0 / md.tau_r);
    b.i_tau_si = (float)(1.0 / md.tau_si);
    b.i_tau_vp = (float)(1.0 / md.tau_v_plus);
    b.i_tau_wm = (float)(1.0 / md.tau_w_minus);
    b.i_tau_wp = (float)(1.0 / md.tau_w_plus);
    b.d_tau_w = (float)(1.0 / md.tau_w_plus - 1.0 / md.tau_w_minus);
    b.tau_v2m = (float)md.tau_v2_minus;
    b.dtv = (float)(md.tau_v1_minus - md.tau_v2_minus);
    for (int v = 0; v < 3; v++) b.lc[v] = (float)a.lc[v];
    b.ghc = (float)a.ghc;
}

// Built-in tile table, from measurements on MI355X (grids 128^2..1024^2 batched to about 4M cells
// per launch, and single instances of 256^2..1024^2):
//  * batched launches: 64x16 cells per 512-thread workgroup (two tile cells per thread) is within
//    1% of the best configuration for both kernels at every size (64x32/512 is 35% slower for the
//    adjoint step: register pressure);
//  * launches of fewer than 1M cells leave the GPU part empty with 64x16 tiles: 64x8 cells per
//    512 threads doubles the workgroups (single 256^2 instance: forward and adjoint steps 20% faster).
// The forward tile does not change any result, so it follows the launch's size.  The adjoint tile
// regroups the z.gh partial sums, so it is fixed per instance (by its own size): batched or alone,
// an instance gives the same result.
// The adjoint tile only regroups work: every adjoint kernel writes one z.gh partial per 64 x 8 cell block
// (PART_ROWS), so it follows the launch's size as the forward tile does.
// Measured: a lone 256^2 evaluation (T = 100 ms) takes 77 ms with 64x8 tiles and 91 ms with 64x16, a lone
// 512^2 one (T = 200 ms) 234 and 230 ms.
int default_adj_cfg(int64_t launch_cells) { return launch_cells <= ((int64_t)1 << 17) ? TILE_8_512 : TILE_16_512; }
int default_fwd_cfg(int64_t launch_cells) { return launch_cells < ((int64_t)1 << 20) ? TILE_8_512 : TILE_16_512; }

int validate_and_codes(const fkadj_problem *p, int pitch, std::vector<uint16_t> &code, std::vector<int32_t> &cell_of,
                       int &n_num)
{
    const int rows = p->nrows, cols = p->ncols;
    n_num = 0;
    for (int k = 0; k < rows * cols; k++)
        if (p->active_index[k] > 0) n_num++;
    if (n_num == 0) return FKADJ_ERR_GEOMETRY;
    cell_of.assign(n_num, -1);
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) {
            int a = p->active_index[i * cols + j];
            if (a < 0 || a > n_num) return FKADJ_ERR_GEOMETRY;
            if (a > 0) {
                if (cell_of[a - 1] != -1) return FKADJ_ERR_GEOMETRY;
                cell_of[a - 1] = i * pitch + j;
            }
        }
    auto act = [&](int i, int j) -> bool {
        return i >= 0 && i < rows && j >= 0 && j < cols && p->active_index[i * cols + j] > 0;
    };
    code.assign((size_t)rows * pitch, 0);
    for (int i = 0; i < rows; i++)
        for (int j = 0; j < cols; j++) {
            int a = p->active_index[i * cols + j];
            if (a <= 0) continue;
            uint32_t c = C_ACTIVE;
            bool t = act(i - 1, j), b = act(i + 1, j), l = act(i, j - 1), r = act(i, j + 1);
            if ((!t && !b) || (!l && !r)) return FKADJ_ERR_GEOMETRY;   // mirror rule undefined (NaN in G)
            if (!t) c |= C_TOP_MISS;
            if (!b) c |= C_BOT_MISS;
            if (!l) c |= C_LEFT_MISS;
            if (!r) c |= C_RIGHT_MISS;
            auto w = [&](int di, int dj) -> uint32_t {
                if (!act(i + di, j + dj)) return 0u;
                return act(i + 2 * di, j + 2 * dj) ? 1u : 2u;
            };
            c |= w(-1, 0) << C_WT_SHIFT;
            c |= w(1, 0) << C_WB_SHIFT;
            c |= w(0, -1) << C_WL_SHIFT;
            c |= w(0, 1) << C_WR_SHIFT;
            int gh;
            if (p->gh_num) {
                double g = p->gh_num[a - 1];
                if (g == 1.0) gh = 1;
                else if (g == -1.0) gh = -1;
                else if (g == 0.0) gh = 0;
                else return FKADJ_ERR_ARG;
            } else {
                gh = (int)!b - (int)!t;   // [i+1 missing] - [i-1 missing]
            }
            c |= (uint32_t)(gh + 1) << C_GH_SHIFT;
            code[(size_t)i * pitch + j] = (uint16_t)c;
        }
    return FKADJ_OK;
}

template <typename T>
PlainCtx<T, typename std::conditional<std::is_same<T, double>::value, Model64, Model32>::type>
plain_ctx(fkadj_instance *I, EvalCtx *C)
{
    PlainCtx<T, typename std::conditional<std::is_same<T, double>::value, Model64, Model32>::type> c;
    c.g = GridArgs{I->rows, I->pitch, I->n, I->d_code};
    if constexpr (std::is_same<T, double>::value) c.m = I->m64; else c.m = I->m32;
    c.dt = I->p.dt;
    c.Ef = C->d_Ef;
    c.Ea = C->d_Ea;
    T *w = as<T>(C->d_work);
    size_t s = (size_t)3 * I->n;
    c.w0 = w; c.w1 = w + s; c.w2 = w + 2 * s; c.w3 = w + 3 * s; c.w4 = w + 4 * s; c.w5 = w + 5 * s;
    c.stream = I->stream;
    return c;
}

// Location of forward state `step` during the forward sweep (checkpoint or ring buffer).
char *fwd_loc(fkadj_instance *I, EvalCtx *C, int step)
{
    if (step % I->K == 0 || step == I->nt) return I->state(C->d_ck, (step + I->K - 1) / I->K);
    return I->state(C->d_ring, step & 1);
}

// Stage-time field tables (see fkadj.h for the two pulse modes), uploaded on stream s.
int build_tables(fkadj_instance *I, EvalCtx *C, const double *E, hipStream_t s)
{
    const int nt = I->nt;
    C->Ef.assign((size_t)3 * nt, 0.0);
    C->Ea.assign((size_t)3 * nt, 0.0);
    if (I->p.pulse_mode == FKADJ_PULSE_KNOTS) {
        std::vector<double> t(nt + 1), ypp;
        for (int i = 0; i <= nt; i++) t[i] = i * I->p.dt;
        if (!spline_not_a_knot(nt + 1, t.data(), E, ypp)) return FKADJ_ERR_ARG;
        for (int i = 0; i < nt; i++) {
            C->Ef[3 * i] = E[i];
            C->Ef[3 * i + 1] = spline_eval(t.data(), E, ypp, i, (i + 0.5) * I->p.dt);        // G:1692-1694
            C->Ef[3 * i + 2] = E[i + 1];
        }
        for (int i = 1; i <= nt; i++) {
            C->Ea[3 * (i - 1)] = E[i - 1];
            C->Ea[3 * (i - 1) + 1] = E[i];
            C->Ea[3 * (i - 1) + 2] = spline_eval(t.data(), E, ypp, i - 1, (i - 1 + 1.0 / 3.0) * I->p.dt);  // G:1877
        }
    } else {
        const int P = nt / I->p.n_amp;
        for (int i = 0; i < nt; i++) {
            double a = E[i / P];
            C->Ef[3 * i] = C->Ef[3 * i + 1] = C->Ef[3 * i + 2] = a;
            C->Ea[3 * i] = C->Ea[3 * i + 1] = C->Ea[3 * i + 2] = a;
        }
    }
    HIPCHK(hipMemcpyAsync(C->d_Ef, C->Ef.data(), sizeof(double) * 3 * nt, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(C->d_Ea, C->Ea.data(), sizeof(double) * 3 * nt, hipMemcpyHostToDevice, s));
    return FKADJ_OK;
}

void assemble(const fkadj_instance *I, const double *E, const double *zgh, double L0, double *grad, double *Lc,
              double *L0o, double *LEo)
{
    const int nt = I->nt;
    const double dt = I->p.dt, alpha = I->p.alpha;
    double LE;
    if (I->p.pulse_mode == FKADJ_PULSE_KNOTS) {
        double e2 = 0.0;
        for (int i = 0; i <= nt; i++) e2 += E[i] * E[i];
        LE = dt * e2;                                                 // G:1803
        if (grad)
            for (int i = 0; i <= nt; i++) grad[i] = alpha * E[i] - zgh[i];   // G:1924-1926
    } else {
        const int P = nt / I->p.n_amp;
        const double Tp = P * dt;
        double e2 = 0.0;
        for (int k = 0; k < I->p.n_amp; k++) e2 += E[k] * E[k];
        LE = Tp * e2;
        if (grad)
            for (int k = 0; k < I->p.n_amp; k++) {
                double q = 0.0;
                for (int s = k * P; s < (k + 1) * P; s++) q += 0.5 * (zgh[s] + zgh[s + 1]);
                grad[k] = alpha * E[k] * Tp - dt * q;
            }
    }
    if (L0o) *L0o = L0;
    if (LEo) *LEo = LE;
    if (Lc) *Lc = 0.5 * L0 + 0.5 * alpha * LE;   // G:1804
}

struct EventPool {
    std::vector<hipEvent_t> ev;
    size_t used = 0;
    hipEvent_t get()
    {
        if (used == ev.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            ev.push_back(e);
        }
        return ev[used++];
    }
    void reset() { used = 0; }
    ~EventPool()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

// Launch timing: a window of kWindow consecutive launches of one kind (forward-kind or adjoint) on one
// stream, opened once every kSampleEvery launches of that kind, is bracketed by two events; the mean
// launch duration is the windows' elapsed time over their launches.  A window spans back-to-back
// launches only, so the gap between the opening event and the first kernel is shared by eight
// launches instead of charged to one; a launch of the other kind or a phase boundary (cut) drops
// the open window.
constexpr int kWindow = 8;
constexpr int kSampleEvery = 64;

struct Sampler {
    bool on;
    EventPool pool;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fwd, adj;
    int64_t nf = 0, na = 0;
    hipEvent_t open = nullptr;   // the open window's first event
    int open_kind = 0, open_n = 0;
    explicit Sampler(bool o) : on(o) {}
    void cut() { open = nullptr; }
    template <typename F>
    void launch(int kind, hipStream_t s, F f)
    {
        int64_t &cnt = kind ? na : nf;
        if (open && open_kind != kind) open = nullptr;
        if (on && !open && cnt % kSampleEvery == 0) {
            open = pool.get();
            if (open && hipEventRecord(open, s) != hipSuccess) open = nullptr;
            open_kind = kind;
            open_n = 0;
        }
        f();
        cnt++;
        if (open && ++open_n == kWindow) {
            hipEvent_t b = pool.get();
            if (b && hipEventRecord(b, s) == hipSuccess) (kind ? adj : fwd).push_back({open, b});
            open = nullptr;
        }
    }
    template <typename F>
    void fwd_launch(hipStream_t s, F f) { launch(0, s, f); }
    template <typename F>
    void adj_launch(hipStream_t s, F f) { launch(1, s, f); }
    static double mean(const std::vector<std::pair<hipEvent_t, hipEvent_t>> &v)
    {
        if (v.empty()) return 0.0;
        double s = 0.0;
        for (auto &p : v) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, p.first, p.second);
            s += ms;
        }
        return s / (v.size() * (double)kWindow);
    }
};

// ------------------------------------------------------------------------ evaluation contexts
// The FP32 backward pass overlaps the recompute of segment s-1 with the adjoint of segment s (two
// streams, two segment buffers) when its group has fewer cells than this; larger groups fill the
// GPU with one stream and need one segment buffer.
constexpr int64_t kTwoStreamCells = (int64_t)2 << 20;
// Forward sweeps of groups below this many cells take two steps per launch.
constexpr int64_t kDepth2Cells = (int64_t)1 << 20;

// Segment state / slope buffers of an FP32 context: two only where an overlapped backward pass can
// happen, i.e. for an instance's own context (it may join a group) smaller than the threshold.  A
// server request's context and a context of an instance at or above the threshold take one.
int seg_copies_of(int prec, int64_t cells, bool own_context)
{
    return prec == FKADJ_F32 && own_context && cells < kTwoStreamCells ? 2 : 1;
}

// States held by an evaluation context.  FP64: checkpoints, one segment, ring, adjoint pair, 6 stage
// planes; FP32: checkpoints, c recompute segments and `keep` kept ones, ring, adjoint pair, K stored
// slopes per segment buffer, f(y_nt), 3 first slopes.
int64_t ctx_states(int prec, int64_t nck, int64_t K, int c, int64_t keep = 0)
{
    return prec == FKADJ_F64 ? nck + (K + 1) + 2 + 2 + 6 : nck + (c + keep) * (2 * K + 1) + 2 + 2 + 1 + 3;
}

// Segments an FP32 instance keeps from the forward sweep within keep_bytes (each: K+1 states and
// K slopes), at most all of them.
int kept_segments(int prec, int64_t keep_bytes, int64_t st_bytes, int K, int nt)
{
    if (prec != FKADJ_F32 || keep_bytes <= 0) return 0;
    const int64_t nseg = (nt + K - 1) / K;
    return (int)std::min<int64_t>(nseg, keep_bytes / ((2 * (int64_t)K + 1) * st_bytes));
}

// Device bytes of one evaluation context (tables, partials, checkpoints, segment states, ring,
// adjoint ping-pong, FP64 stage scratch or FP32 stored slopes) with c recompute segment buffers and
// `keep` kept segments.
int64_t ctx_bytes(const fkadj_instance *I, int c, int keep = 0)
{
    const int64_t st = (int64_t)I->st_bytes;
    const int64_t states = ctx_states(I->prec, I->nck, I->K, c, keep);
    int64_t tables = 8 * (int64_t)(6 * I->nt + (int64_t)I->nt * I->nb_part + I->nb_plain + (I->nt + 1) +
                                   (I->n_num + 255) / 256 + 1);
    return states * st + tables;
}

void ctx_free(EvalCtx *C)
{
    if (!C) return;
    for (char **q : {&C->d_ck, &C->d_seg, &C->d_ring, &C->d_z, &C->d_work, &C->d_kseg, &C->d_kT, &C->d_kfirst})
        if (*q) (void)hipFree(*q);
    for (double **q : {&C->d_Ef, &C->d_Ea, &C->d_part, &C->d_partT, &C->d_zgh, &C->d_L0part, &C->d_L0})
        if (*q) (void)hipFree(*q);
    group_free(C->group);
    delete C;
}

// The FP32 slot record of a context whose buffers are in place.
void ctx_fill_slot(EvalCtx *C)
{
    const fkadj_instance *I = C->I;
    FusedSlot &S = C->slot_h;
    S.rows = I->rows; S.pitch = I->pitch; S.n = I->n;
    S.nt = I->nt; S.K = I->K; S.nck = I->nck;
    S.ntiles = I->nb_part;
    S.seg_par = C->seg_copies - 1;
    S.keep0 = (I->nt + I->K - 1) / I->K - C->keep;
    S.code = I->d_code;
    S.ck = as<float>(C->d_ck); S.seg = as<float>(C->d_seg); S.ring = as<float>(C->d_ring); S.zb = as<float>(C->d_z);
    S.kseg = as<float>(C->d_kseg); S.kT = as<float>(C->d_kT); S.kfirst = as<float>(C->d_kfirst);
    S.Ef = C->d_Ef; S.Ea = C->d_Ea; S.partial = C->d_part;
    S.m = I->m32; S.dt = (float)I->p.dt;
}

// Allocates a zeroed context on stream s (hole cells read 0 in every state buffer).  An FP32
// context also gets its own one-member group (launch maps, a second stream) unless it is only
// driven through a server (self_group false).
int ctx_create(fkadj_instance *I, hipStream_t s, EvalCtx **out, bool self_group = true)
{
    auto *C = new (std::nothrow) EvalCtx();
    if (!C) return FKADJ_ERR_NOMEM;
    C->I = I;
    C->seg_copies = seg_copies_of(I->prec, (int64_t)I->rows * I->pitch, self_group);
    C->keep = self_group ? I->keep : 0;   // a server request recomputes every segment
    const int nbuf = C->seg_copies + C->keep;
    const size_t st = I->st_bytes;
    auto alc = [&](char *&q, size_t bytes) -> bool {
        if (hipMalloc((void **)&q, bytes) != hipSuccess) { q = nullptr; return false; }
        return hipMemsetAsync(q, 0, bytes, s) == hipSuccess;
    };
    auto ald = [&](double *&q, size_t count) -> bool {
        return hipMalloc((void **)&q, sizeof(double) * std::max<size_t>(count, 1)) == hipSuccess;
    };
    const int nt = I->nt;
    bool ok = ald(C->d_Ef, 3 * (size_t)nt) && ald(C->d_Ea, 3 * (size_t)nt) && ald(C->d_part, (size_t)nt * I->nb_part) &&
              ald(C->d_partT, I->nb_plain) && ald(C->d_zgh, nt + 1) && ald(C->d_L0part, objective_blocks(I->n_num)) &&
              ald(C->d_L0, 1) && alc(C->d_ck, st * I->nck) &&
              alc(C->d_seg, st * (I->K + 1) * nbuf) &&
              alc(C->d_ring, st * 2) && alc(C->d_z, st * 2);
    if (ok && I->prec == FKADJ_F64) ok = alc(C->d_work, st * 6);
    if (ok && I->prec == FKADJ_F32)
        ok = alc(C->d_kseg, st * nbuf * I->K) && alc(C->d_kT, st) && alc(C->d_kfirst, st * 3);
    if (ok && I->prec == FKADJ_F32) {
        ctx_fill_slot(C);
        if (self_group) {
            std::vector<EvalCtx *> self{C};
            ok = group_build(self, s, &C->group) == FKADJ_OK;
        }
    }
    // an instance's own context is used from other streams too (batch lanes): wait for the zero
    // fills; a server context is only used on the server's stream, where they are ordered already
    if (self_group) ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) {
        if (!self_group) (void)hipStreamSynchronize(s);
        ctx_free(C);
        return FKADJ_ERR_NOMEM;
    }
    C->bytes = ctx_bytes(I, C->seg_copies, C->keep);
    *out = C;
    return FKADJ_OK;
}

int ensure_default_ctx(fkadj_instance *I)
{
    if (I->ctx) return FKADJ_OK;
    return ctx_create(I, I->stream, &I->ctx);
}

// ------------------------------------------------------------------------ one evaluation
// Forward sweep with checkpoints, objective and terminal adjoint, then per segment (last to
// first) recompute + adjoint steps, on the caller's stream (the FP32 backward pass may also use the
// group's side stream, joined before the final sums).
template <typename T>
void terminal(EvalCtx *C, hipStream_t s)
{
    fkadj_instance *I = C->I;
    auto pc = plain_ctx<T>(I, C);
    pc.stream = s;
    T *yT = as<T>(I->state(C->d_ck, I->nck - 1));
    T *z0 = as<T>(I->state(C->d_z, 0));
    objective_launch<T>(I->n_num, I->d_rowptr, I->d_colind, I->d_val, I->p.s_diag, I->d_cell_of, I->n, yT,
                        I->d_gamma, z0, C->d_L0part, s);
    sum_rows_launch(C->d_L0part, objective_blocks(I->n_num), 1, C->d_L0, s);
    plain_zgh(pc, z0, C->d_partT);
    sum_rows_launch(C->d_partT, I->nb_plain, 1, C->d_zgh + I->nt, s);
}

int run_plain(EvalCtx *C, Sampler &smp)
{
    using T = double;
    fkadj_instance *I = C->I;
    const int nt = I->nt, K = I->K;
    hipStream_t s = I->stream;
    auto pc = plain_ctx<T>(I, C);
    HIPCHK(hipMemcpyAsync(I->state(C->d_ck, 0), I->d_y0, I->st_bytes, hipMemcpyDeviceToDevice, s));
    for (int i = 0; i < nt; i++) {
        T *yin = as<T>(fwd_loc(I, C, i)), *yout = as<T>(fwd_loc(I, C, i + 1));
        smp.fwd_launch(s, [&] { plain_forward_step(pc, yin, yout, i); });
    }
    HIPCHK(hipGetLastError());
    smp.cut();
    terminal<T>(C, s);
    const int nseg = (nt + K - 1) / K;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int a = sg * K, b = std::min(a + K, nt);
        smp.cut();
        for (int j = 0; j + 1 < b - a; j++) {
            T *yin = as<T>(j == 0 ? I->state(C->d_ck, sg) : I->state(C->d_seg, j));
            T *yout = as<T>(I->state(C->d_seg, j + 1));
            smp.fwd_launch(s, [&] { plain_forward_step(pc, yin, yout, a + j); });
        }
        for (int i = b; i > a; i--) {
            T *yA = as<T>(i - 1 == a ? I->state(C->d_ck, sg) : I->state(C->d_seg, i - 1 - a));
            T *yB = as<T>(i == b ? I->state(C->d_ck, sg + 1) : I->state(C->d_seg, i - a));
            T *zin = as<T>(I->state(C->d_z, (nt - i) & 1));
            T *zout = as<T>(I->state(C->d_z, (nt - i + 1) & 1));
            double *prow = C->d_part + (size_t)(i - 1) * I->nb_part;
            smp.adj_launch(s, [&] { plain_adjoint_step(pc, yA, yB, zin, zout, i - 1, prow); });
        }
        HIPCHK(hipGetLastError());
    }
    sum_rows_launch(C->d_part, I->nb_part, nt, C->d_zgh, s);
    HIPCHK(hipGetLastError());
    return FKADJ_OK;
}

void group_free(FusedGroupDev *g)
{
    if (!g) return;
    if (g->side) (void)hipStreamSynchronize(g->side);
    if (g->d_slots) (void)hipFree(g->d_slots);
    for (int p = 0; p < 2; p++) {
        if (g->ev_rec[p]) (void)hipEventDestroy(g->ev_rec[p]);
        if (g->ev_adj[p]) (void)hipEventDestroy(g->ev_adj[p]);
    }
    if (g->ev_lane) (void)hipEventDestroy(g->ev_lane);
    if (g->side) (void)hipStreamDestroy(g->side);

    delete g;
}

int fwd_cfg_of(const fkadj_instance *I, int64_t launch_cells)
{
    return I->fwd_cfg > 0 ? I->fwd_cfg : default_fwd_cfg(launch_cells);
}
int adj_cfg_of(const fkadj_instance *I, int64_t launch_cells)
{
    return I->adj_cfg > 0 ? I->adj_cfg : default_adj_cfg(launch_cells);
}

// Row-rolling kernels: 0 never, 1 for groups of 2M cells or more (default), 2 for every group that
// can use them; strip rows (a multiple of 8 and of every member's adjoint tile rows).
std::atomic<int> g_roll_mode{1};
std::atomic<int> g_roll_hs{0};   // strip rows of the rolling launches; 0: chosen per launch (auto_strip_rows)
constexpr int64_t kRollCells = (int64_t)2 << 20;

// Compute units of the current device (queried once per device).
int device_cus()
{
    static std::mutex mu;
    static std::map<int, int> cus;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 256;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cus.find(dev);
    if (it != cus.end()) return it->second;
    int n = 0;
    if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n < 1) n = 256;
    cus[dev] = n;
    return n;
}

// Strip rows of a rolling launch over instances of these heights at this pitch, when neither the
// members nor fkadj_set_rolling fix them: the tallest of 64, 32 and 16 rows whose strips still give
// the launch its share of the workgroups the GPU holds at once (CUs x workgroups of this width per
// CU, at the kernels' 16 waves per CU, divided by the launches that run side by side), else 16.
// Taller strips evaluate fewer rows twice at the strip ends (6 per strip) and re-read fewer; a
// launch with fewer strips than its share leaves CUs idle.  Sixteen 512^2 instances in two groups
// get 16 rows; a group of ten 1024^2 instances gets 64 and one of ten 768^2 instances 32 (64 uneven
// grids of 256^2..1024^2, T = 150 ms: 8.06 -> 7.41 s per batched evaluation).  Results do not
// depend on it.
int auto_strip_rows(const std::vector<int> &rows, int pitch, int concurrent)
{
    const int per_cu = std::max(1, 16 / (pitch / 64));
    const int64_t share = (int64_t)device_cus() * per_cu / std::max(1, concurrent);
    for (int hs : {64, 32, 16}) {
        int64_t strips = 0;
        for (int r : rows) strips += (r + hs - 1) / hs;
        if (strips >= share) return hs;
    }
    return 16;
}

// Host part of a group: members, launch maps, depth and kept-segment choice; the slot records go to
// `slots` in member order (no device work).  `concurrent`: groups of the same call that run side by
// side (the share of the GPU a rolling launch should fill).
void group_maps(FusedGroupDev *g, const std::vector<EvalCtx *> &members, std::vector<FusedSlot> &slots,
                int concurrent = 1)
{
    g->members = members;
    slots.clear();
    for (int c = 0; c < NUM_TILE_CFG; c++) {
        launch_map_clear(g->fwd[c]);
        launch_map_clear(g->adj[c]);
    }
    int64_t cells = 0;
    for (auto *C : members) cells += (int64_t)C->I->rows * C->I->pitch;
    launch_map_clear(g->fwd2);
    launch_map_clear(g->adj2);
    // Two forward steps per launch where one launch costs more than its arithmetic: launches of
    // fewer than 1M cells (those the built-in rule gives 64x8 tiles), members on the built-in
    // forward tile, and a step count and checkpoint interval that pair up.
    // The members' own choice (fkadj_problem.fwd_depth), else the built-in rule; all members must
    // agree.  Two adjoint steps per launch only where the members ask for it (fkadj_problem.adj_depth
    // = 2): its 8-cell halo on 64x8 tiles doubles the stage work of a step, which costs more than
    // the launch and the Jacobian it saves (lone 512^2 evaluation, T = 200 ms: 270 ms against 231 ms
    // with one step per launch; 256^2, T = 100 ms: 78 against 77 ms).
    const fkadj_instance *I0 = members[0]->I;
    const bool small = cells < kDepth2Cells;
    g->depth2 = (I0->fwd_depth == 2 || (I0->fwd_depth == 0 && small && I0->fwd_cfg == 0)) && I0->nt % 2 == 0 &&
                I0->K % 2 == 0;
    g->adj_depth = I0->adj_depth == 2 ? 2 : 1;
    g->keep0 = 0;
    for (size_t k = 0; k < members.size(); k++) {
        fkadj_instance *I = members[k]->I;
        slots.push_back(members[k]->slot_h);
        g->keep0 = std::max(g->keep0, members[k]->slot_h.keep0);
        const int fc = fwd_cfg_of(I, cells);
        launch_map_add(g->fwd[fc], (int)k, tiles_of(I->rows, I->pitch, fc));
        const int ac = adj_cfg_of(I, cells);
        launch_map_add(g->adj[ac], (int)k, tiles_of(I->rows, I->pitch, ac));
        launch_map_add(g->fwd2, (int)k, tiles_of(I->rows, I->pitch, TILE_8_512));
        launch_map_add(g->adj2, (int)k, tiles_of(I->rows, I->pitch, TILE_8_512));
        if (I->fwd_depth != I0->fwd_depth || I->fwd_cfg != I0->fwd_cfg) g->depth2 = false;
        if (I->adj_depth != I0->adj_depth) g->adj_depth = 1;
    }
    // row-rolling kernels per pitch they support: every such pitch (mode 2) or those with 2M cells
    // or more in the group (mode 1); strips of hs rows (a multiple of the partial block height):
    // the members' own strip rows (fkadj_problem.roll_rows) where all members of the pitch ask for
    // the same, else the library setting, else (setting 0) auto_strip_rows
    const int rmode = g_roll_mode.load(), hs = g_roll_hs.load();
    std::map<int, int64_t> pitch_cells;
    std::map<int, int> pitch_rows;   // -1: the members disagree
    std::map<int, std::vector<int>> pitch_heights;
    for (auto *C : members) {
        pitch_cells[C->I->pitch] += (int64_t)C->I->rows * C->I->pitch;
        pitch_heights[C->I->pitch].push_back(C->I->rows);
        auto it = pitch_rows.find(C->I->pitch);
        if (it == pitch_rows.end()) pitch_rows[C->I->pitch] = C->I->roll_rows;
        else if (it->second != C->I->roll_rows) it->second = -1;
    }
    g->nroll = 0;
    for (auto &kv : pitch_cells)
        if (roll_pitch_supported(kv.first) && (rmode == 2 || (rmode == 1 && kv.second >= kRollCells)) &&
            g->nroll < FusedGroupDev::kMaxRoll) {
            g->roll_pitch[g->nroll] = kv.first;
            g->roll_hs[g->nroll] = pitch_rows[kv.first] > 0 ? pitch_rows[kv.first]
                                   : hs > 0                 ? hs
                                                            : auto_strip_rows(pitch_heights[kv.first], kv.first, concurrent);
            launch_map_clear(g->rollmap[g->nroll]);
            g->nroll++;
        }
    for (int c = 0; c < NUM_TILE_CFG; c++) {
        launch_map_clear(g->fwd_rest[c]);
        launch_map_clear(g->adj_rest[c]);
    }
    for (size_t k = 0; k < members.size(); k++) {
        const fkadj_instance *I = members[k]->I;
        int r = 0;
        while (r < g->nroll && g->roll_pitch[r] != I->pitch) r++;
        if (r < g->nroll) {
            launch_map_add(g->rollmap[r], (int)k, (I->rows + g->roll_hs[r] - 1) / g->roll_hs[r]);
        } else {
            const int fc = fwd_cfg_of(I, cells), ac = adj_cfg_of(I, cells);
            launch_map_add(g->fwd_rest[fc], (int)k, tiles_of(I->rows, I->pitch, fc));
            launch_map_add(g->adj_rest[ac], (int)k, tiles_of(I->rows, I->pitch, ac));
        }
    }
    if (g->nroll) g->depth2 = false;
    if (g->nroll) g->adj_depth = 1;
}

int group_build(const std::vector<EvalCtx *> &members, hipStream_t s, FusedGroupDev **out, int concurrent)
{
    if (members.empty() || members.size() > (size_t)MAX_SLOTS) return FKADJ_ERR_ARG;
    auto *g = new (std::nothrow) FusedGroupDev();
    if (!g) return FKADJ_ERR_NOMEM;
    std::vector<FusedSlot> slots;
    group_maps(g, members, slots, concurrent);
    bool ok = hipMalloc((void **)&g->d_slots, sizeof(FusedSlot) * slots.size()) == hipSuccess &&
              hipMemcpyAsync(g->d_slots, slots.data(), sizeof(FusedSlot) * slots.size(), hipMemcpyHostToDevice, s) == hipSuccess;
    ok = ok && hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&g->ev_lane, hipEventDisableTiming) == hipSuccess;
    for (int p = 0; p < 2 && ok; p++)
        ok = hipEventCreateWithFlags(&g->ev_rec[p], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&g->ev_adj[p], hipEventDisableTiming) == hipSuccess;
    ok = ok && hipStreamSynchronize(s) == hipSuccess;
    if (!ok) { group_free(g); return FKADJ_ERR_NOMEM; }
    *out = g;
    return FKADJ_OK;
}

// Groups a batch call of 2M cells or more is split into (fkadj_set_batch_groups).
std::atomic<int> g_batch_groups{2};

// Streams of the groups of split batch calls, per device: created back to back, once, so that they
// sit on different hardware queues (a stream gets one of the device's few queues when it is created,
// and two streams on one queue run one after the other).
constexpr int kLanes = 4;
std::mutex g_lanes_mu;
std::map<int, std::vector<hipStream_t>> g_lanes;

int batch_lanes(std::vector<hipStream_t> &out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_lanes_mu);
    auto &v = g_lanes[dev];
    while (v.size() < (size_t)kLanes) {
        hipStream_t st = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        v.push_back(st);
    }
    out = v;
    return FKADJ_OK;
}

void groups_free(std::vector<FusedGroupDev *> &gs)
{
    for (auto *g : gs) group_free(g);
    gs.clear();
}

// The members of a batch call as groups, each run from its own host thread on its own stream: a
// batch of 2M cells or more (its backward pass runs on one stream) is cut into up to max_groups
// contiguous runs of members of about equal cells; a smaller one stays one group.
int groups_build(const std::vector<EvalCtx *> &members_in, hipStream_t s, int max_groups,
                 std::vector<FusedGroupDev *> &out)
{
    out.clear();
    // members of one pitch next to each other (stable), so that the groups hold few pitches with
    // many cells each, which the row-rolling kernels take; results do not depend on the grouping
    std::vector<EvalCtx *> members(members_in);
    std::stable_sort(members.begin(), members.end(),
                     [](const EvalCtx *a, const EvalCtx *b) { return a->I->pitch < b->I->pitch; });
    const int n = (int)members.size();
    std::vector<int64_t> cells(n);
    int64_t total = 0;
    for (int q = 0; q < n; q++) total += cells[q] = (int64_t)members[q]->I->rows * members[q]->I->pitch;
    const int k = total >= kTwoStreamCells ? std::max(1, std::min(max_groups, n)) : 1;
    std::vector<int> end(k);
    split_runs(cells.data(), n, k, end.data());
    std::vector<std::vector<EvalCtx *>> parts(k);
    for (int j = 0, q = 0; j < k; j++)
        for (; q < end[j]; q++) parts[j].push_back(members[q]);
    for (auto &part : parts) {
        FusedGroupDev *g = nullptr;
        const int rc = group_build(part, s, &g, k);
        if (rc) { groups_free(out); return rc; }
        out.push_back(g);
    }
    return FKADJ_OK;
}

Cursors uniform_cursors(int count, int mode, int step, int seg_a, int zpar = 0)
{
    Cursors c;
    for (int k = 0; k < MAX_SLOTS; k++) c.c[k] = SlotCursor{k < count ? mode : MODE_IDLE, step, seg_a, zpar};
    return c;
}

void group_forward(const FusedGroupDev *g, int mode, int step, int seg_a, hipStream_t s)
{
    const Cursors c = uniform_cursors((int)g->members.size(), mode, step, seg_a);
    if (g->nroll && mode != MODE_RHS) {
        for (int k = 1; k < NUM_TILE_CFG; k++) fused_forward_launch(k, g->d_slots, g->fwd_rest[k], c, s);
        for (int r = 0; r < g->nroll; r++)
            roll_launch(false, g->roll_pitch[r], g->d_slots, g->rollmap[r], c, g->roll_hs[r], s);
        return;
    }
    for (int k = 1; k < NUM_TILE_CFG; k++) fused_forward_launch(k, g->d_slots, g->fwd[k], c, s);
}

void group_forward2(const FusedGroupDev *g, int mode, int step, int seg_a, hipStream_t s)
{
    const Cursors c = uniform_cursors((int)g->members.size(), mode, step, seg_a);
    fused_forward2_launch(g->d_slots, g->fwd2, c, s);
}

// One adjoint step (i) of every member, z read from half zpar of the members' zb, written to the other.
void group_adjoint(const FusedGroupDev *g, int i, int seg_a, int zpar, hipStream_t s)
{
    const Cursors c = uniform_cursors((int)g->members.size(), MODE_ADJ, i, seg_a, zpar);
    if (g->nroll) {
        for (int k = 1; k < NUM_TILE_CFG; k++) fused_adjoint_launch(k, g->d_slots, g->adj_rest[k], c, s);
        for (int r = 0; r < g->nroll; r++)
            roll_launch(true, g->roll_pitch[r], g->d_slots, g->rollmap[r], c, g->roll_hs[r], s);
        return;
    }
    for (int k = 1; k < NUM_TILE_CFG; k++) fused_adjoint_launch(k, g->d_slots, g->adj[k], c, s);
}

// Adjoint steps i and i - 1 (both inside the segment starting at seg_a) of every member in one launch.
void group_adjoint2(const FusedGroupDev *g, int i, int seg_a, int zpar, hipStream_t s)
{
    const Cursors c = uniform_cursors((int)g->members.size(), MODE_ADJ2, i, seg_a, zpar);
    fused_adjoint2_launch(g->d_slots, g->adj2, c, s);
}

int run_fused(const FusedGroupDev *g, hipStream_t s, Sampler &smp)
{
    fkadj_instance *I0 = g->members[0]->I;
    const int nt = I0->nt, K = I0->K;
    for (auto *C : g->members)
        HIPCHK(hipMemcpyAsync(C->I->state(C->d_ck, 0), C->I->d_y0, C->I->st_bytes, hipMemcpyDeviceToDevice, s));
    // steps of the kept segments (from keep0 on) store their states and slopes as they go; keep0*K
    // is a checkpoint, and even where two steps run per launch (K even, so two steps never straddle
    // a segment boundary)
    const int kept_from = std::min(g->keep0 * K, nt);
    if (g->depth2) {
        for (int i = 0; i < kept_from; i += 2) smp.fwd_launch(s, [&] { group_forward2(g, MODE_FWD, i, 0, s); });
        for (int i = kept_from; i < nt; i += 2)
            smp.fwd_launch(s, [&] { group_forward2(g, MODE_KEEP, i, i / K * K, s); });
    } else {
        for (int i = 0; i < kept_from; i++) smp.fwd_launch(s, [&] { group_forward(g, MODE_FWD, i, 0, s); });
        for (int i = kept_from; i < nt; i++)
            smp.fwd_launch(s, [&] { group_forward(g, MODE_KEEP, i, i / K * K, s); });
    }
    HIPCHK(hipGetLastError());
    smp.cut();
    for (auto *C : g->members) terminal<float>(C, s);
    group_forward(g, MODE_RHS, nt - 1, 0, s);   // f(y_nt) for the last adjoint step
    // Backward pass, two streams: the recompute of segment s-1 (this stream) runs while the adjoint
    // steps of segment s run on the side stream.  Segments alternate between two state / slope
    // buffers, so the recompute of segment s-2 waits for the adjoint of segment s.  Launches of
    // fewer than 2M cells leave the GPU part idle and gain from the overlap (single 512^2 instance:
    // evaluation 325 -> 263 ms; 256^2: 238 -> 181 ms); a full GPU does not (16 x 512^2: 1% slower),
    // so larger groups, and groups with a member of one segment buffer (only contexts of 2M cells
    // or more have one), keep both sweeps on one stream (the events then order nothing new).
    const int nseg = (nt + K - 1) / K;
    int64_t cells = 0;
    bool two_buffers = true;
    for (auto *C : g->members) {
        cells += (int64_t)C->I->rows * C->I->pitch;
        two_buffers = two_buffers && C->seg_copies == 2;
    }
    hipStream_t side = cells < kTwoStreamCells && two_buffers ? g->side : s;
    int zpar = 0;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int a = sg * K, b = std::min(a + K, nt), p = sg & 1;
        if (sg + 2 <= nseg - 1) HIPCHK(hipStreamWaitEvent(s, g->ev_adj[p], 0));
        smp.cut();
        // all b-a steps are replayed: the last one only to store the slope f(y_{b-1}); kept
        // segments are not (a member that keeps more segments than the group recomputes some of
        // its kept ones, into their own buffers, with the same values)
        if (sg < g->keep0)
            for (int j = 0; j < b - a; j++) smp.fwd_launch(s, [&] { group_forward(g, MODE_REC, a + j, a, s); });
        HIPCHK(hipEventRecord(g->ev_rec[p], s));
        HIPCHK(hipStreamWaitEvent(side, g->ev_rec[p], 0));
        smp.cut();
        // the adjoint state alternates between the two halves of zb launch by launch (z_nt in half 0)
        for (int i = b; i > a;) {
            if (g->adj_depth == 2 && i - 2 >= a) {
                smp.adj_launch(side, [&] { group_adjoint2(g, i, a, zpar, side); });
                i -= 2;
            } else {
                smp.adj_launch(side, [&] { group_adjoint(g, i, a, zpar, side); });
                i -= 1;
            }
            zpar ^= 1;
        }
        HIPCHK(hipEventRecord(g->ev_adj[p], side));
        HIPCHK(hipGetLastError());
    }
    HIPCHK(hipStreamWaitEvent(s, g->ev_adj[0], 0));   // segment 0's adjoint: the last side-stream work
    for (auto *C : g->members) sum_rows_launch(C->d_part, C->I->nb_part, nt, C->d_zgh, s);
    HIPCHK(hipGetLastError());
    return FKADJ_OK;
}

// Timing of the last evaluation from the samplers of its groups (launch counts summed, sampled
// durations averaged over all groups' samples).
void record_timing(const std::vector<Sampler *> &smps, hipEvent_t e0, hipEvent_t e1, int64_t cell_steps)
{
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    std::vector<std::pair<hipEvent_t, hipEvent_t>> fwd, adj;
    int64_t nf = 0, na = 0;
    for (auto *smp : smps) {
        fwd.insert(fwd.end(), smp->fwd.begin(), smp->fwd.end());
        adj.insert(adj.end(), smp->adj.begin(), smp->adj.end());
        nf += smp->nf;
        na += smp->na;
    }
    const double fms = Sampler::mean(fwd), ams = Sampler::mean(adj);
    std::lock_guard<std::mutex> lk(g_timing_mu);
    g_timing.eval_ms = ms;
    g_timing.fwd_kernel_ms = fms;
    g_timing.adj_kernel_ms = ams;
    g_timing.fwd_launches = nf;
    g_timing.adj_launches = na;
    g_timing.cell_steps = cell_steps;
}
void record_timing(Sampler &smp, hipEvent_t e0, hipEvent_t e1, int64_t cell_steps)
{
    record_timing(std::vector<Sampler *>{&smp}, e0, e1, cell_steps);
}

// run_fused of one group, for a host thread of its own: nothing escapes the thread, and a device
// error keeps the thread's message
struct GroupRun {
    int rc = FKADJ_OK;
    char err[sizeof g_errmsg] = {0};
};
void run_group(FusedGroupDev *g, hipStream_t st, Sampler *smp, GroupRun *out)
{
    try {
        out->rc = run_fused(g, st, *smp);
    } catch (const std::bad_alloc &) {
        out->rc = FKADJ_ERR_NOMEM;
    } catch (...) {
        out->rc = FKADJ_ERR_DEVICE;
    }
    if (out->rc == FKADJ_ERR_DEVICE) memcpy(out->err, g_errmsg, sizeof out->err);
}

// z.gh (nt+1 values, into the host buffer zgh) and L0 of a finished evaluation, then the gradient.
int fetch_into_and_assemble(EvalCtx *C, hipStream_t s, double *zgh, const double *E, double *grad, double *Lc,
                            double *L0, double *LE)
{
    fkadj_instance *I = C->I;
    double L0v = 0.0;
    HIPCHK(hipMemcpyAsync(zgh, C->d_zgh, sizeof(double) * (I->nt + 1), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(&L0v, C->d_L0, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (!all_finite(zgh, (size_t)I->nt + 1) || !std::isfinite(L0v)) return FKADJ_ERR_NONFINITE;
    assemble(I, E, zgh, L0v, grad, Lc, L0, LE);
    return FKADJ_OK;
}

int fetch_and_assemble(EvalCtx *C, hipStream_t s, const double *E, double *grad, double *Lc, double *L0,
                       double *LE)
{
    std::vector<double> zgh(C->I->nt + 1);
    return fetch_into_and_assemble(C, s, zgh.data(), E, grad, Lc, L0, LE);
}

int eval_one(fkadj_instance *I, const double *E, double *grad, double *Lc, double *L0, double *LE)
{
    if (!I->state_set) return FKADJ_ERR_STATE;
    if (!all_finite(E, I->ncontrols)) return FKADJ_ERR_NONFINITE;
    CHK(ensure_default_ctx(I));
    EvalCtx *C = I->ctx;
    CHK(build_tables(I, C, E, I->stream));
    Sampler smp(g_profiling.load() != 0);
    hipEvent_t e0 = smp.pool.get(), e1 = smp.pool.get();
    HIPCHK(hipEventRecord(e0, I->stream));
    int rc = (I->prec == FKADJ_F64) ? run_plain(C, smp) : run_fused(C->group, I->stream, smp);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, I->stream));
    CHK(fetch_and_assemble(C, I->stream, E, grad, Lc, L0, LE));
    record_timing(smp, e0, e1, (int64_t)I->rows * I->pitch * I->nt);
    return FKADJ_OK;
}

template <typename T>
void dfree(T *&p)
{
    if (p) (void)hipFree((void *)p);
    p = nullptr;
}

}  // namespace

// ============================================================================ C ABI
extern "C" {

const char *fkadj_strerror(int code)
{
    switch (code) {
    case FKADJ_OK: return "ok";
    case FKADJ_ERR_ARG: return "invalid argument";
    case FKADJ_ERR_GEOMETRY: return "invalid grid geometry (both axial neighbours of an active cell missing, "
                                    "bad active_index, or S size mismatch)";
    case FKADJ_ERR_NOMEM: return "device memory exhausted or over budget";
    case FKADJ_ERR_DEVICE: return g_errmsg[0] ? g_errmsg : "HIP runtime error";
    case FKADJ_ERR_NONFINITE: return "non-finite input or result";
    case FKADJ_ERR_STATE: return "call out of order (no state set)";
    case FKADJ_ERR_CANCELLED: return "request cancelled";
    case FKADJ_ERR_BUSY: return "result not ready";
    case FKADJ_ERR_NOTFOUND: return "unknown request";
    case FKADJ_ERR_SCRATCH: return g_errmsg[0] ? g_errmsg : "evaluation larger than the no-allocation array";
    default: return "unknown error";
    }
}

void fkadj_model_reference(fkadj_model *m)
{
    // src/cuda/fk_parameters_00_05_alpha_0p8_dx_0p035.h:2-24
    m->u_c = 0.13; m->u_v = 0.04; m->tau_d = 0.388; m->tau_v1_minus = 7.6; m->tau_v2_minus = 8.8;
    m->tau_v_plus = 3.324; m->tau_0 = 8.2; m->tau_r = 33.264; m->tau_si = 29.0; m->u_csi = 0.54; m->k_fk = 15.0;
    m->tau_w_minus = 68.0; m->tau_w_plus = 399.99999999999994; m->Du = 1.0e-3; m->Dv = 1.0e-5; m->Dw = 1.0e-5;
    m->ksm = 100.0; m->h = 0.035;
}

int fkadj_param_string(const fkadj_model *m, char *buf, size_t cap)
{
    if (!m) return FKADJ_ERR_ARG;
    char tmp[1024];
    int len = snprintf(tmp, sizeof tmp,
                       "// Fenton-Karma parameters (libfkadj)\nu_c %.17g\nu_v %.17g\ntau_d %.17g\ntau_v1_minus %.17g\n"
                       "tau_v2_minus %.17g\ntau_v_plus %.17g\ntau_0 %.17g\ntau_r %.17g\ntau_si %.17g\nu_csi %.17g\n"
                       "k_fk %.17g\ntau_w_minus %.17g\ntau_w_plus %.17g\nDu %.17g\nDv %.17g\nDw %.17g\nksm %.17g\n"
                       "h %.17g\nnvar 3\n",
                       m->u_c, m->u_v, m->tau_d, m->tau_v1_minus, m->tau_v2_minus, m->tau_v_plus, m->tau_0, m->tau_r,
                       m->tau_si, m->u_csi, m->k_fk, m->tau_w_minus, m->tau_w_plus, m->Du, m->Dv, m->Dw, m->ksm, m->h);
    if (buf && cap > 0) {
        size_t c = std::min((size_t)len, cap - 1);
        memcpy(buf, tmp, c);
        buf[c] = '\0';
    }
    return len;
}

int64_t fkadj_footprint_bytes(const fkadj_problem *p)
{
    if (!p || p->nrows < 1 || p->ncols < 1 || p->nt < 1) return -1;
    const int pitch = (p->ncols + 63) / 64 * 64;
    const int64_t n = (int64_t)p->nrows * pitch;
    const int K = p->ckpt > 0 ? p->ckpt : 50;
    const int64_t esz = p->precision == FKADJ_F32 ? 4 : 8;
    const int64_t st = 3 * n * esz;
    const int64_t nck = (p->nt + K - 1) / K + 1;
    const int64_t states = 1 + ctx_states(p->precision, nck, K, seg_copies_of(p->precision, n, true),
                                          kept_segments(p->precision, p->keep_bytes, st, K, p->nt));
    int64_t nb = p->precision == FKADJ_F64 ? (pitch / 64) * ((p->nrows + 3) / 4) : (pitch / 64) * ((p->nrows + PART_ROWS - 1) / PART_ROWS);
    int64_t extra = n * 3 /*code+mask*/ + 16 * (int64_t)p->nt * 3 + 8 * (int64_t)p->nt * nb + 8 * (p->nt + 1) +
                    8 * 3 * n;
    return states * st + extra;
}

int fkadj_mem_info(int64_t *free_bytes, int64_t *total_bytes)
{
    if (!free_bytes || !total_bytes) return FKADJ_ERR_ARG;
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
        snprintf(g_errmsg, sizeof g_errmsg, "hipMemGetInfo failed (no GPU?)");
        return FKADJ_ERR_DEVICE;
    }
    *free_bytes = (int64_t)fr;
    *total_bytes = (int64_t)tot;
    return FKADJ_OK;
}

int fkadj_create(const fkadj_problem *p, fkadj_instance **out)
{
    FKADJ_ENTRY
    if (!p || !out || !p->active_index) return FKADJ_ERR_ARG;
    *out = nullptr;
    if (p->nrows < 4 || p->ncols < 4 || p->nt < 1 || !(p->dt > 0) || !std::isfinite(p->dt)) return FKADJ_ERR_ARG;
    if (p->precision != FKADJ_F64 && p->precision != FKADJ_F32) return FKADJ_ERR_ARG;
    // kernels index a state with 32-bit offsets (bytes for the FP32 buffer loads)
    if ((int64_t)p->nrows * ((p->ncols + 63) / 64 * 64) * 3 * (p->precision == FKADJ_F32 ? 4 : 8) >= (int64_t)1 << 31)
        return FKADJ_ERR_ARG;
    if (p->fwd_tile < 0 || p->fwd_tile >= NUM_TILE_CFG || p->adj_tile < 0 || p->adj_tile >= NUM_TILE_CFG ||
        p->keep_bytes < 0 || p->fwd_depth < 0 || p->fwd_depth > 2 || p->adj_depth < 0 || p->adj_depth > 2 ||
        p->roll_rows < 0 || p->roll_rows > 256 || p->roll_rows % 8 != 0)
        return FKADJ_ERR_ARG;
    if (p->pulse_mode == FKADJ_PULSE_KNOTS) {
        if (p->nt + 1 < 4) return FKADJ_ERR_ARG;
    } else if (p->pulse_mode == FKADJ_PULSE_PIECEWISE) {
        if (p->n_amp < 1 || p->nt % p->n_amp != 0) return FKADJ_ERR_ARG;
    } else {
        return FKADJ_ERR_ARG;
    }
    auto *I = new (std::nothrow) fkadj_instance();
    if (!I) return FKADJ_ERR_NOMEM;
    I->p = *p;
    I->rows = p->nrows;
    I->cols = p->ncols;
    I->pitch = (p->ncols + 63) / 64 * 64;
    I->n = I->rows * I->pitch;
    I->nt = p->nt;
    I->K = p->ckpt > 0 ? p->ckpt : 50;
    I->nck = (I->nt + I->K - 1) / I->K + 1;
    I->prec = p->precision;
    I->esz = I->prec == FKADJ_F32 ? 4 : 8;
    I->st_bytes = (size_t)3 * I->n * I->esz;
    I->ncontrols = p->pulse_mode == FKADJ_PULSE_KNOTS ? p->nt + 1 : p->n_amp;
    I->keep = kept_segments(I->prec, p->keep_bytes, (int64_t)I->st_bytes, I->K, I->nt);
    fill_models(I);

    std::vector<uint16_t> code;
    int rc = validate_and_codes(p, I->pitch, code, I->cell_of, I->n_num);
    if (rc) { delete I; return rc; }
    I->p.active_index = nullptr;
    I->p.gh_num = nullptr;

    // objective matrix: CSC over packed indices -> CSR (rows keep increasing column order)
    std::vector<int64_t> rowptr;
    std::vector<int32_t> colind;
    std::vector<double> val;
    if (p->S_colptr) {
        if (!p->S_rowind || !p->S_val) { delete I; return FKADJ_ERR_ARG; }
        const int nn = I->n_num;
        int64_t nnz = p->S_colptr[nn];
        if (p->S_colptr[0] != 0 || nnz < 0) { delete I; return FKADJ_ERR_GEOMETRY; }
        rowptr.assign(nn + 1, 0);
        for (int c = 0; c < nn; c++) {
            if (p->S_colptr[c + 1] < p->S_colptr[c]) { delete I; return FKADJ_ERR_GEOMETRY; }
            for (int64_t q = p->S_colptr[c]; q < p->S_colptr[c + 1]; q++) {
                int64_t r = p->S_rowind[q];
                if (r < 0 || r >= nn) { delete I; return FKADJ_ERR_GEOMETRY; }
                rowptr[r + 1]++;
            }
        }
        for (int r = 0; r < nn; r++) rowptr[r + 1] += rowptr[r];
        colind.resize(nnz);
        val.resize(nnz);
        std::vector<int64_t> fillp(rowptr.begin(), rowptr.end() - 1);
        for (int c = 0; c < nn; c++)
            for (int64_t q = p->S_colptr[c]; q < p->S_colptr[c + 1]; q++) {
                int64_t r = p->S_rowind[q];
                colind[fillp[r]] = c;
                val[fillp[r]++] = p->S_val[q];
            }
        if (!all_finite(val.data(), val.size())) { delete I; return FKADJ_ERR_NONFINITE; }
    }
    I->p.S_colptr = nullptr;
    I->p.S_rowind = nullptr;
    I->p.S_val = nullptr;

    auto fail = [&](int code_) {
        fkadj_destroy(I);
        return code_;
    };
#define DMALLOC(ptr, bytes)                                                  \
    do {                                                                     \
        if (hipMalloc((void **)&(ptr), (bytes)) != hipSuccess) return fail(FKADJ_ERR_NOMEM); \
    } while (0)

    if (hipGetDevice(&I->device) != hipSuccess || hipStreamCreateWithFlags(&I->stream, hipStreamNonBlocking) != hipSuccess) {
        snprintf(g_errmsg, sizeof g_errmsg, "hipStreamCreate failed (no GPU?)");
        delete I;
        return FKADJ_ERR_DEVICE;
    }
    const size_t st = I->st_bytes;
    I->nb_plain = (I->pitch / PLAIN_BX) * ((I->rows + PLAIN_BY - 1) / PLAIN_BY);
    if (I->prec == FKADJ_F32) {
        I->fwd_cfg = p->fwd_tile;   // 0: chosen per group from its total size
        I->adj_cfg = p->adj_tile;   // 0: chosen per launch from its total size
        I->fwd_depth = p->fwd_depth;
        I->adj_depth = p->adj_depth;
        I->roll_rows = p->roll_rows;
        I->nb_part = (I->pitch / FUSED_TX) * ((I->rows + PART_ROWS - 1) / PART_ROWS);   // 64 x 8 partial blocks
    } else {
        I->nb_part = I->nb_plain;
    }
    DMALLOC(I->d_code, sizeof(uint16_t) * I->n);
    DMALLOC(I->d_cell_of, sizeof(int32_t) * I->n_num);
    DMALLOC(I->d_gamma, sizeof(double) * 3);
    DMALLOC(I->d_y0, st);
    DMALLOC(I->d_packed, sizeof(double) * 3 * I->n_num);
    DMALLOC(I->d_mask, I->n);
    if (!rowptr.empty()) {
        DMALLOC(I->d_rowptr, sizeof(int64_t) * rowptr.size());
        DMALLOC(I->d_colind, sizeof(int32_t) * std::max<size_t>(1, colind.size()));
        DMALLOC(I->d_val, sizeof(double) * std::max<size_t>(1, val.size()));
    }
    // hole cells of every state buffer stay zero
    if (hipMemsetAsync(I->d_y0, 0, st, I->stream) != hipSuccess) return fail(FKADJ_ERR_DEVICE);
    bool ok = hipMemcpyAsync(I->d_code, code.data(), sizeof(uint16_t) * I->n, hipMemcpyHostToDevice, I->stream) == hipSuccess &&
              hipMemcpyAsync(I->d_cell_of, I->cell_of.data(), sizeof(int32_t) * I->n_num, hipMemcpyHostToDevice, I->stream) == hipSuccess &&
              hipMemcpyAsync(I->d_gamma, p->gamma, sizeof(double) * 3, hipMemcpyHostToDevice, I->stream) == hipSuccess;
    if (ok && !rowptr.empty())
        ok = hipMemcpyAsync(I->d_rowptr, rowptr.data(), sizeof(int64_t) * rowptr.size(), hipMemcpyHostToDevice, I->stream) == hipSuccess &&
             (colind.empty() || (hipMemcpyAsync(I->d_colind, colind.data(), sizeof(int32_t) * colind.size(), hipMemcpyHostToDevice, I->stream) == hipSuccess &&
                                 hipMemcpyAsync(I->d_val, val.data(), sizeof(double) * val.size(), hipMemcpyHostToDevice, I->stream) == hipSuccess));
    if (!ok) return fail(FKADJ_ERR_DEVICE);
    I->lazy_ws = p->lazy_workspace != 0;
    if (!I->lazy_ws) {
        int wrc = ensure_default_ctx(I);
        if (wrc) return fail(wrc);
    }
#undef DMALLOC
    if (hipStreamSynchronize(I->stream) != hipSuccess) return fail(FKADJ_ERR_DEVICE);
    *out = I;
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_destroy(fkadj_instance *I)
{
    if (!I) return FKADJ_ERR_ARG;
    if (I->busy || I->plans) return FKADJ_ERR_BUSY;
    if (I->stream) (void)hipStreamSynchronize(I->stream);
    dfree(I->d_code); dfree(I->d_cell_of); dfree(I->d_rowptr); dfree(I->d_colind); dfree(I->d_val); dfree(I->d_gamma);
    ctx_free(I->ctx);
    I->ctx = nullptr;
    dfree(I->d_y0); dfree(I->d_packed); dfree(I->d_mask);
    if (I->stream) (void)hipStreamDestroy(I->stream);
    delete I;
    return FKADJ_OK;
}

int fkadj_n_active(const fkadj_instance *I) { return I ? I->n_num : FKADJ_ERR_ARG; }
int64_t fkadj_request_bytes(const fkadj_instance *I)
{
    return I ? ctx_bytes(I, seg_copies_of(I->prec, (int64_t)I->rows * I->pitch, false)) : FKADJ_ERR_ARG;
}
int fkadj_n_controls(const fkadj_instance *I) { return I ? I->ncontrols : FKADJ_ERR_ARG; }

int fkadj_set_state(fkadj_instance *I, const double *y)
{
    if (!I || !y) return FKADJ_ERR_ARG;
    if (I->busy) return FKADJ_ERR_BUSY;   // served requests still read the resident state
    if (!all_finite(y, (size_t)3 * I->n_num)) return FKADJ_ERR_NONFINITE;
    HIPCHK(hipMemcpyAsync(I->d_packed, y, sizeof(double) * 3 * I->n_num, hipMemcpyHostToDevice, I->stream));
    if (I->prec == FKADJ_F64)
        unpack_launch<double>(I->d_packed, I->d_cell_of, I->n_num, I->n, as<double>(I->d_y0), I->stream);
    else
        unpack_launch<float>(I->d_packed, I->d_cell_of, I->n_num, I->n, as<float>(I->d_y0), I->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(I->stream));
    I->state_set = true;
    return FKADJ_OK;
}

static int pack_state(fkadj_instance *I, const char *planes, double *y)
{
    if (I->prec == FKADJ_F64)
        pack_launch<double>(reinterpret_cast<const double *>(planes), I->d_cell_of, I->n_num, I->n, I->d_packed, I->stream);
    else
        pack_launch<float>(reinterpret_cast<const float *>(planes), I->d_cell_of, I->n_num, I->n, I->d_packed, I->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(y, I->d_packed, sizeof(double) * 3 * I->n_num, hipMemcpyDeviceToHost, I->stream));
    HIPCHK(hipStreamSynchronize(I->stream));
    return FKADJ_OK;
}

// Forward solve from the resident state over nsteps with field E; snapshot ks (ks < max_snaps)
// of the packed state after step (ks+1)*snap_every (F:794-813), the packed final state, and with
// advance the final state as the new resident state.
static int forward_run(fkadj_instance *I, const double *E, int nsteps, int snap_every, int max_snaps, double *snaps,
                       double *y_final, int advance)
{
    if (!I || !E || nsteps < 0 || nsteps > I->nt || snap_every < 0 || max_snaps < 0) return FKADJ_ERR_ARG;
    if (max_snaps > 0 && (!snaps || snap_every < 1)) return FKADJ_ERR_ARG;
    if (!I->state_set) return FKADJ_ERR_STATE;
    if (I->busy) return FKADJ_ERR_BUSY;
    if (!all_finite(E, I->ncontrols)) return FKADJ_ERR_NONFINITE;
    CHK(ensure_default_ctx(I));
    EvalCtx *C = I->ctx;
    hipStream_t s = I->stream;
    CHK(build_tables(I, C, E, s));
    HIPCHK(hipMemcpyAsync(I->state(C->d_ck, 0), I->d_y0, I->st_bytes, hipMemcpyDeviceToDevice, s));
    int taken = 0;
    for (int i = 0; i < nsteps; i++) {
        if (I->prec == FKADJ_F32) {
            group_forward(C->group, MODE_FWD, i, 0, s);
        } else {
            auto pc = plain_ctx<double>(I, C);
            plain_forward_step(pc, as<double>(fwd_loc(I, C, i)), as<double>(fwd_loc(I, C, i + 1)), i);
        }
        HIPCHK(hipGetLastError());
        if (snap_every > 0 && (i + 1) % snap_every == 0 && taken < max_snaps) {
            CHK(pack_state(I, fwd_loc(I, C, i + 1), snaps + (size_t)taken * 3 * I->n_num));
            taken++;
        }
    }
    char *fin = fwd_loc(I, C, nsteps);
    if (y_final) CHK(pack_state(I, fin, y_final));
    if (advance) HIPCHK(hipMemcpyAsync(I->d_y0, fin, I->st_bytes, hipMemcpyDeviceToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return FKADJ_OK;
}

int fkadj_get_state(fkadj_instance *I, double *y)
{
    if (!I || !y) return FKADJ_ERR_ARG;
    if (!I->state_set) return FKADJ_ERR_STATE;
    return pack_state(I, I->d_y0, y);
}

int fkadj_grad(fkadj_instance *I, const double *E, double *grad, double *Lc, double *L0, double *LE)
{
    FKADJ_ENTRY
    if (!I || !E) return FKADJ_ERR_ARG;
    return eval_one(I, E, grad, Lc, L0, LE);
    FKADJ_EXIT
}

int fkadj_forward(fkadj_instance *I, const double *E, int nsteps, int snap_every, double *snaps, double *y_final,
                  int advance)
{
    FKADJ_ENTRY
    if (!I || nsteps < 0 || snap_every < 0) return FKADJ_ERR_ARG;
    const int nsnap = snap_every > 0 ? fkadj_snapshot_count(nsteps, snap_every) : 0;
    return forward_run(I, E, nsteps, snap_every, nsnap, snaps, y_final, advance);
    FKADJ_EXIT
}

int fkadj_snapshot_count(int nsteps, int snap_every)
{
    if (nsteps < 0 || snap_every < 1) return FKADJ_ERR_ARG;
    // FM:7-25: nt/nstep - 1 when nstep divides nt (the final state is returned separately),
    // floor(nt/nstep) otherwise; its relative integer test fails at nt = 0, which gives 0
    return nsteps > 0 && nsteps % snap_every == 0 ? nsteps / snap_every - 1 : nsteps / snap_every;
}

int fkadj_forward_ref(double dt, const double *Et, int nt, const double *y_ini, const int32_t *active_index, int nrows,
                      int ncols, int n_num, const double *gh_num, int nstep, int ns, double *ys, double *yf,
                      char *param_str, size_t param_str_cap)
{
    FKADJ_ENTRY
    if (!Et || !y_ini || !active_index || !gh_num || !yf || nstep < 1 || ns < 0 || nt < 1) return FKADJ_ERR_ARG;
    if (ns > nt / nstep || (ns > 0 && !ys)) return FKADJ_ERR_ARG;
    if (nrows < 4 || ncols < 4 || n_num < 1) return FKADJ_ERR_ARG;
    if (count_active(active_index, nrows, ncols) != n_num) return FKADJ_ERR_GEOMETRY;   // before gh_num is read
    fkadj_problem p{};
    p.nrows = nrows;
    p.ncols = ncols;
    p.active_index = active_index;
    p.gh_num = gh_num;
    fkadj_model_reference(&p.model);
    p.dt = dt;
    p.nt = nt;
    p.pulse_mode = FKADJ_PULSE_KNOTS;
    p.gamma[0] = 1.0;
    p.precision = FKADJ_F64;
    p.lazy_workspace = 1;
    fkadj_instance *I = nullptr;
    int rc = fkadj_create(&p, &I);
    if (rc) return rc;
    if (I->n_num != n_num) { fkadj_destroy(I); return FKADJ_ERR_GEOMETRY; }
    const size_t n = (size_t)nrows * ncols, np = (size_t)3 * n_num;
    std::vector<double> packed(np * std::max(ns, 1)), fin(np);
    rc = fkadj_set_state(I, y_ini);
    if (!rc) rc = forward_run(I, Et, nt, nstep, ns, packed.data(), fin.data(), 0);
    fkadj_destroy(I);
    if (rc) return rc;
    // the reference's output layout: three full planes per state (row-major i*ncols + j), NaN at holes
    auto unpack = [&](const double *src, double *dst) {
        for (size_t k = 0; k < n; k++) {
            const int a = active_index[k];
            for (int v = 0; v < 3; v++) dst[v * n + k] = a > 0 ? src[(size_t)v * n_num + a - 1] : std::nan("");
        }
    };
    for (int ks = 0; ks < ns; ks++) unpack(packed.data() + (size_t)ks * np, ys + (size_t)ks * 3 * n);
    unpack(fin.data(), yf);
    if (param_str && param_str_cap) fkadj_param_string(&p.model, param_str, param_str_cap);
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_batch_forward(int count, fkadj_instance *const *insts, const double *const *E, int nsteps)
{
    FKADJ_ENTRY
    if (count < 1 || count > MAX_SLOTS || !insts || !E || nsteps < 0) return FKADJ_ERR_ARG;
    for (int k = 0; k < count; k++) {
        fkadj_instance *I = insts[k];
        if (!I || !E[k] || I->prec != FKADJ_F32 || I->nt != insts[0]->nt || I->K != insts[0]->K || nsteps > I->nt)
            return FKADJ_ERR_ARG;
        for (int q = 0; q < k; q++)
            if (insts[q] == I) return FKADJ_ERR_ARG;
        if (!I->state_set) return FKADJ_ERR_STATE;
        if (I->busy) return FKADJ_ERR_BUSY;
        if (!all_finite(E[k], I->ncontrols)) return FKADJ_ERR_NONFINITE;
    }
    hipStream_t s = insts[0]->stream;
    std::vector<EvalCtx *> members;
    for (int k = 0; k < count; k++) {
        CHK(ensure_default_ctx(insts[k]));
        members.push_back(insts[k]->ctx);
        CHK(build_tables(insts[k], insts[k]->ctx, E[k], s));
    }
    FusedGroupDev *g = nullptr;
    CHK(group_build(members, s, &g));
    int rc = FKADJ_OK;
    for (auto *C : members)
        if (hipMemcpyAsync(C->I->state(C->d_ck, 0), C->I->d_y0, C->I->st_bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
            rc = FKADJ_ERR_DEVICE;
    for (int i = 0; i < nsteps && !rc; i++) group_forward(g, MODE_FWD, i, 0, s);
    if (!rc && hipGetLastError() != hipSuccess) rc = FKADJ_ERR_DEVICE;
    for (auto *C : members)
        if (!rc && hipMemcpyAsync(C->I->d_y0, fwd_loc(C->I, C, nsteps), C->I->st_bytes, hipMemcpyDeviceToDevice, s) !=
                       hipSuccess)
            rc = FKADJ_ERR_DEVICE;
    if (hipStreamSynchronize(s) != hipSuccess && !rc) rc = FKADJ_ERR_DEVICE;
    group_free(g);
    return rc;
    FKADJ_EXIT
}

int fkadj_stimulate(fkadj_instance *I, const uint8_t *cell_mask, double value)
{
    FKADJ_ENTRY
    if (!I || !cell_mask || !std::isfinite(value)) return FKADJ_ERR_ARG;
    if (I->busy) return FKADJ_ERR_BUSY;
    if (!I->state_set) return FKADJ_ERR_STATE;
    std::vector<uint8_t> m((size_t)I->n, 0);
    for (int r = 0; r < I->n_num; r++) {
        int k = I->cell_of[r];
        int i = k / I->pitch, j = k % I->pitch;
        m[k] = cell_mask[(size_t)i * I->cols + j] ? 1 : 0;
    }
    HIPCHK(hipMemcpyAsync(I->d_mask, m.data(), m.size(), hipMemcpyHostToDevice, I->stream));
    if (I->prec == FKADJ_F64)
        stimulate_launch<double>(as<double>(I->d_y0), I->d_mask, I->n, value, I->stream);
    else
        stimulate_launch<float>(as<float>(I->d_y0), I->d_mask, I->n, (float)value, I->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(I->stream));
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_grad_ref(double dt, const double *Et, int nt, const double *y_ini, const int32_t *active_index, int nrows,
                   int ncols, int n_num, const double *gh_num, const double *gamma, const int64_t *S_colptr,
                   const int64_t *S_rowind, const double *S_val, double alpha, double *L0, double *LE, double *Lcost,
                   double *gradE, char *param_str, size_t param_str_cap)
{
    FKADJ_ENTRY
    if (!Et || !y_ini || !active_index || !gh_num || !gamma || !S_colptr || !S_rowind || !S_val || !gradE)
        return FKADJ_ERR_ARG;
    if (nrows < 4 || ncols < 4 || n_num < 1) return FKADJ_ERR_ARG;
    // S is n_num x n_num and gh_num has n_num values: a caller whose n_num differs from the grid's
    // active cells gets an error before either is read (G:1759-1762)
    if (count_active(active_index, nrows, ncols) != n_num) return FKADJ_ERR_GEOMETRY;
    fkadj_problem p{};
    p.nrows = nrows;
    p.ncols = ncols;
    p.active_index = active_index;
    p.gh_num = gh_num;
    fkadj_model_reference(&p.model);
    p.dt = dt;
    p.nt = nt;
    p.pulse_mode = FKADJ_PULSE_KNOTS;
    p.alpha = alpha;
    for (int v = 0; v < 3; v++) p.gamma[v] = gamma[v];
    p.S_colptr = S_colptr;
    p.S_rowind = S_rowind;
    p.S_val = S_val;
    p.precision = FKADJ_F64;
    p.ckpt = 50;
    fkadj_instance *I = nullptr;
    int rc = fkadj_create(&p, &I);
    if (rc) return rc;
    if (I->n_num != n_num) { fkadj_destroy(I); return FKADJ_ERR_GEOMETRY; }   // G:1759-1762
    rc = fkadj_set_state(I, y_ini);
    if (!rc) rc = fkadj_grad(I, Et, gradE, Lcost, L0, LE);
    fkadj_destroy(I);
    if (!rc && param_str && param_str_cap) fkadj_param_string(&p.model, param_str, param_str_cap);
    return rc;
    FKADJ_EXIT
}

int fkadj_set_batch_groups(int n)
{
    if (n < 1 || n > 8) return FKADJ_ERR_ARG;
    return g_batch_groups.exchange(n);
}

int fkadj_set_rolling(int mode, int strip_rows)
{
    if (mode < 0 || mode > 2 || strip_rows < 0 || strip_rows > 256 || strip_rows % 8 != 0) return FKADJ_ERR_ARG;
    g_roll_hs.store(strip_rows);
    return g_roll_mode.exchange(mode);
}

int fkadj_set_profiling(int on)
{
    g_profiling.store(on ? 1 : 0);
    return FKADJ_OK;
}

int fkadj_last_timing(fkadj_timing *t)
{
    if (!t) return FKADJ_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_timing_mu);
    *t = g_timing;
    return FKADJ_OK;
}

int fkadj_batch_grad(int count, fkadj_instance *const *insts, const double *const *E, double *const *grad,
                     double *Lcost, double *L0, double *LE)
{
    FKADJ_ENTRY
    if (count < 1 || count > MAX_SLOTS || !insts || !E) return FKADJ_ERR_ARG;
    for (int k = 0; k < count; k++) {
        fkadj_instance *I = insts[k];
        if (!I || !E[k]) return FKADJ_ERR_ARG;
        if (I->prec != FKADJ_F32 || I->nt != insts[0]->nt || I->K != insts[0]->K) return FKADJ_ERR_ARG;
        for (int q = 0; q < k; q++)
            if (insts[q] == I) return FKADJ_ERR_ARG;
        if (!I->state_set) return FKADJ_ERR_STATE;
        if (I->busy) return FKADJ_ERR_BUSY;
        if (!all_finite(E[k], I->ncontrols)) return FKADJ_ERR_NONFINITE;
    }
    hipStream_t s = insts[0]->stream;
    std::vector<EvalCtx *> members;
    for (int k = 0; k < count; k++) {
        CHK(ensure_default_ctx(insts[k]));
        members.push_back(insts[k]->ctx);
        CHK(build_tables(insts[k], insts[k]->ctx, E[k], s));
    }
    // one sampler per possible group, made before any device resource so that nothing leaks if
    // host memory runs out
    const int max_groups = std::max(1, std::min(g_batch_groups.load(), count));
    std::vector<std::unique_ptr<Sampler>> smp;
    std::vector<Sampler *> smps;
    for (int j = 0; j < max_groups; j++) {
        smp.push_back(std::make_unique<Sampler>(g_profiling.load() != 0));
        smps.push_back(smp.back().get());
    }
    std::vector<FusedGroupDev *> gs;
    CHK(groups_build(members, s, max_groups, gs));
    const size_t ng = gs.size();   // at most max_groups
    smps.resize(ng);
    hipEvent_t e0 = smp[0]->pool.get(), e1 = smp[0]->pool.get();
    int rc = hipEventRecord(e0, s) == hipSuccess ? FKADJ_OK : FKADJ_ERR_DEVICE;
    // a split batch: every group starts on a lane stream of its own when the caller's stream gets
    // here, each issued from a host thread of its own (one thread issuing every group's launches
    // would be the bottleneck), and the caller's stream waits for all of them at the end
    std::vector<hipStream_t> lane(ng, s);
    if (ng > 1 && !rc) {
        std::vector<hipStream_t> lanes;
        rc = batch_lanes(lanes);
        for (size_t j = 0; j < ng && !rc; j++) {
            lane[j] = lanes[j % lanes.size()];
            if (hipEventRecord(gs[j]->ev_lane, s) != hipSuccess || hipStreamWaitEvent(lane[j], gs[j]->ev_lane, 0) != hipSuccess)
                rc = FKADJ_ERR_DEVICE;
        }
    }
    if (!rc) {
        std::vector<GroupRun> runs(ng);
        std::vector<std::thread> threads;
        try {
            for (size_t j = 1; j < ng; j++) threads.emplace_back(run_group, gs[j], lane[j], smps[j], &runs[j]);
        } catch (...) {
            runs[0].rc = FKADJ_ERR_NOMEM;   // a group without its thread: the call fails after the others
        }
        if (!runs[0].rc) run_group(gs[0], lane[0], smps[0], &runs[0]);
        for (auto &t : threads) t.join();
        for (size_t j = 0; j < ng && !rc; j++) {
            rc = runs[j].rc;
            if (rc == FKADJ_ERR_DEVICE) memcpy(g_errmsg, runs[j].err, sizeof g_errmsg);
        }
    }
    for (size_t j = 0; j < ng && ng > 1; j++)
        if (hipEventRecord(gs[j]->ev_lane, lane[j]) != hipSuccess || hipStreamWaitEvent(s, gs[j]->ev_lane, 0) != hipSuccess)
            rc = rc ? rc : FKADJ_ERR_DEVICE;
    if (!rc && hipEventRecord(e1, s) != hipSuccess) rc = FKADJ_ERR_DEVICE;
    int64_t cells = 0;
    for (int k = 0; k < count && !rc; k++) {
        rc = fetch_and_assemble(insts[k]->ctx, s, E[k], grad ? grad[k] : nullptr, Lcost ? Lcost + k : nullptr,
                                L0 ? L0 + k : nullptr, LE ? LE + k : nullptr);
        cells += (int64_t)insts[k]->rows * insts[k]->pitch * insts[k]->nt;
    }
    if (!rc) record_timing(smps, e0, e1, cells);
    (void)hipStreamSynchronize(s);
    groups_free(gs);
    return rc;
    FKADJ_EXIT
}

}  // extern "C"

// ============================================================================ no-allocation evaluation
// fkadj_grad_noalloc: the evaluation of fkadj_grad with every evaluation buffer (field tables,
// partials, checkpoints, one segment buffer, ring, adjoint pair, stored slopes or FP64 stage
// planes, and the slot record) carved out of one static device array instead of hipMalloc'd
// memory, for callers that cannot allocate device memory.  Each GPU holds its own copy of the array
// at its own address, taken once per device with hipGetSymbolAddress; host code never passes the
// array's name to the device.  Calls on one GPU are serialised by a per-device mutex and each call
// waits for its work, so they never overlap there; the array keeps the last call's contents, and
// every call clears the part it uses before its first kernel.
static __device__ __align__(256) char scratch[FKADJ_NOALLOC_SCRATCH_BYTES];

namespace {

struct NoAllocDev {
    std::mutex mu;                     // serialises calls on this device
    char *ptr = nullptr;               // device address of `scratch` on this device
    hipEvent_t ev[5] = {};             // the group's ev_rec[2], ev_adj[2], ev_lane
    EvalCtx ctx;                       // buffers point into the array; host tables reused
    FusedGroupDev group;               // the context alone, on the instance's stream
    std::vector<FusedSlot> slots;
    std::vector<double> zgh;           // host staging of z.gh
};
constexpr int kNoAllocDevices = 64;
NoAllocDev g_noalloc[kNoAllocDevices];

// Set-up on the current device (first call there): the array's device address and the events.
int noalloc_setup(NoAllocDev &d)
{
    if (d.ptr) return FKADJ_OK;
    char *p = nullptr;
    HIPCHK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    for (auto &e : d.ev)
        if (!e) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    d.ptr = p;
    return FKADJ_OK;
}

// Carves the context of instance I (one segment buffer, nothing kept) out of [base, base + cap):
// returns the bytes used (256-byte aligned pieces), or -1 if they exceed cap (then nothing is set
// beyond that point).  The slot record takes the first piece.
int64_t noalloc_layout(fkadj_instance *I, char *base, int64_t cap, EvalCtx *C, FusedSlot **d_slot)
{
    int64_t off = 0;
    auto take = [&](int64_t bytes) -> char * {
        char *p = base + off;
        off += (std::max<int64_t>(bytes, 1) + 255) / 256 * 256;
        return off <= cap ? p : nullptr;
    };
    const int64_t st = (int64_t)I->st_bytes, nt = I->nt;
    *d_slot = reinterpret_cast<FusedSlot *>(take(sizeof(FusedSlot)));
    C->d_Ef = as<double>(take(8 * 3 * nt));
    C->d_Ea = as<double>(take(8 * 3 * nt));
    C->d_part = as<double>(take(8 * nt * I->nb_part));
    C->d_partT = as<double>(take(8 * (int64_t)I->nb_plain));
    C->d_zgh = as<double>(take(8 * (nt + 1)));
    C->d_L0part = as<double>(take(8 * (int64_t)objective_blocks(I->n_num)));
    C->d_L0 = as<double>(take(8));
    C->d_ck = take(st * I->nck);
    C->d_seg = take(st * (I->K + 1));
    C->d_ring = take(2 * st);
    C->d_z = take(2 * st);
    C->d_work = I->prec == FKADJ_F64 ? take(6 * st) : nullptr;
    C->d_kseg = I->prec == FKADJ_F32 ? take(st * I->K) : nullptr;
    C->d_kT = I->prec == FKADJ_F32 ? take(st) : nullptr;
    C->d_kfirst = I->prec == FKADJ_F32 ? take(3 * st) : nullptr;
    return off <= cap ? off : -1;
}

int noalloc_eval(fkadj_instance *I, const double *E, double *grad, double *Lc, double *L0, double *LE)
{
    if (I->device < 0 || I->device >= kNoAllocDevices) return FKADJ_ERR_ARG;
    NoAllocDev &d = g_noalloc[I->device];
    std::lock_guard<std::mutex> lk(d.mu);
    CHK(noalloc_setup(d));
    EvalCtx &C = d.ctx;
    C.I = I;
    C.seg_copies = 1;
    C.keep = 0;
    C.group = nullptr;
    FusedSlot *d_slot = nullptr;
    const int64_t used = noalloc_layout(I, d.ptr, (int64_t)FKADJ_NOALLOC_SCRATCH_BYTES, &C, &d_slot);
    if (used < 0) {
        snprintf(g_errmsg, sizeof g_errmsg, "evaluation needs %lld bytes, the no-allocation array holds %lld",
                 (long long)noalloc_layout(I, nullptr, INT64_MAX, &C, &d_slot),
                 (long long)FKADJ_NOALLOC_SCRATCH_BYTES);
        return FKADJ_ERR_SCRATCH;
    }
    hipStream_t s = I->stream;
    // the previous call's contents are still there: hole cells of every state must read 0, and the
    // objective writes the terminal adjoint only at active cells
    HIPCHK(hipMemsetAsync(d.ptr, 0, (size_t)used, s));
    CHK(build_tables(I, &C, E, s));
    Sampler smp(false);
    int rc;
    if (I->prec == FKADJ_F64) {
        rc = run_plain(&C, smp);
    } else {
        ctx_fill_slot(&C);
        FusedGroupDev &g = d.group;
        group_maps(&g, std::vector<EvalCtx *>{&C}, d.slots);
        g.d_slots = d_slot;
        g.side = s;   // one segment buffer: recompute and adjoint stay on one stream
        g.ev_rec[0] = d.ev[0]; g.ev_rec[1] = d.ev[1];
        g.ev_adj[0] = d.ev[2]; g.ev_adj[1] = d.ev[3];
        g.ev_lane = d.ev[4];
        // pageable source: the copy is complete when the call returns (the caller's buffer is ours)
        HIPCHK(hipMemcpyAsync(d_slot, d.slots.data(), sizeof(FusedSlot), hipMemcpyHostToDevice, s));
        rc = run_fused(&g, s, smp);
    }
    if (rc) return rc;
    d.zgh.resize((size_t)I->nt + 1);
    return fetch_into_and_assemble(&C, s, d.zgh.data(), E, grad, Lc, L0, LE);
}

}  // namespace

extern "C" int fkadj_grad_noalloc(fkadj_instance *I, const double *E, double *grad, double *Lc, double *L0,
                                  double *LE)
{
    FKADJ_ENTRY
    if (!I || !E) return FKADJ_ERR_ARG;
    if (!I->state_set) return FKADJ_ERR_STATE;
    if (!all_finite(E, I->ncontrols)) return FKADJ_ERR_NONFINITE;
    int cur = 0;
    HIPCHK(hipGetDevice(&cur));
    if (cur != I->device) HIPCHK(hipSetDevice(I->device));
    const int rc = noalloc_eval(I, E, grad, Lc, L0, LE);
    if (cur != I->device) (void)hipSetDevice(cur);
    return rc;
    FKADJ_EXIT
}

// ============================================================================ batch plans
// A captured HIP graph of one whole evaluation of a fixed group.  The graph holds kernel and
// device-to-device copy nodes only (no memset nodes); it reads the field tables and the resident
// states at replay time, so a replay after new tables (or a new state) computes the new gradient.
// A plan of 2M cells or more holds its members as groups, like a batch call, with one graph per
// group (each captured alone on the caller's stream); a replay launches the graphs side by side on
// the batch streams.
struct fkadj_plan {
    std::vector<fkadj_instance *> insts;
    std::vector<FusedGroupDev *> gs;
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    hipStream_t stream = nullptr;
};

namespace {

int batch_validate(int count, fkadj_instance *const *insts, const double *const *E)
{
    if (count < 1 || count > MAX_SLOTS || !insts) return FKADJ_ERR_ARG;
    for (int k = 0; k < count; k++) {
        fkadj_instance *I = insts[k];
        if (!I || (E && !E[k])) return FKADJ_ERR_ARG;
        if (I->prec != FKADJ_F32 || I->nt != insts[0]->nt || I->K != insts[0]->K) return FKADJ_ERR_ARG;
        for (int q = 0; q < k; q++)
            if (insts[q] == I) return FKADJ_ERR_ARG;
        if (!I->state_set) return FKADJ_ERR_STATE;
        if (I->busy) return FKADJ_ERR_BUSY;
        if (E && !all_finite(E[k], I->ncontrols)) return FKADJ_ERR_NONFINITE;
    }
    return FKADJ_OK;
}

void plan_free(fkadj_plan *P)
{
    if (!P) return;
    if (P->stream) (void)hipStreamSynchronize(P->stream);
    for (auto x : P->execs) (void)hipGraphExecDestroy(x);
    for (auto g : P->graphs) (void)hipGraphDestroy(g);
    groups_free(P->gs);
    for (auto *I : P->insts) I->plans--;
    delete P;
}

}  // namespace

extern "C" {

int fkadj_plan_create(int count, fkadj_instance *const *insts, fkadj_plan **out)
{
    FKADJ_ENTRY
    if (!out) return FKADJ_ERR_ARG;
    *out = nullptr;
    CHK(batch_validate(count, insts, nullptr));
    auto *P = new (std::nothrow) fkadj_plan();
    if (!P) return FKADJ_ERR_NOMEM;
    P->stream = insts[0]->stream;
    std::vector<EvalCtx *> members;
    for (int k = 0; k < count; k++) {
        P->insts.push_back(insts[k]);
        insts[k]->plans++;
        int rc = ensure_default_ctx(insts[k]);
        if (rc) { plan_free(P); return rc; }
        members.push_back(insts[k]->ctx);
    }
    int rc = groups_build(members, P->stream, std::max(1, std::min(g_batch_groups.load(), count)), P->gs);
    if (rc) { plan_free(P); return rc; }
    for (auto *g : P->gs) {
        Sampler off(false);
        if (hipStreamBeginCapture(P->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            plan_free(P);
            return FKADJ_ERR_DEVICE;
        }
        rc = run_fused(g, P->stream, off);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(P->stream, &graph);
        if (graph) P->graphs.push_back(graph);
        hipGraphExec_t exec = nullptr;
        if (rc == FKADJ_OK && e != hipSuccess) rc = FKADJ_ERR_DEVICE;
        if (rc == FKADJ_OK && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
            rc = FKADJ_ERR_DEVICE;
        if (exec) P->execs.push_back(exec);
        if (rc) {
            (void)hipGetLastError();
            plan_free(P);
            return rc;
        }
    }
    *out = P;
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_plan_grad(fkadj_plan *P, const double *const *E, double *const *grad, double *Lcost, double *L0,
                    double *LE)
{
    FKADJ_ENTRY
    if (!P || !E) return FKADJ_ERR_ARG;
    const int count = (int)P->insts.size();
    CHK(batch_validate(count, P->insts.data(), E));
    hipStream_t s = P->stream;
    for (int k = 0; k < count; k++) CHK(build_tables(P->insts[k], P->insts[k]->ctx, E[k], s));
    Sampler smp(false);
    hipEvent_t e0 = smp.pool.get(), e1 = smp.pool.get();
    HIPCHK(hipEventRecord(e0, s));
    if (P->execs.size() == 1) {
        HIPCHK(hipGraphLaunch(P->execs[0], s));
    } else {
        // each group's graph on a batch stream of its own, forked from and joined to s
        std::vector<hipStream_t> lanes;
        CHK(batch_lanes(lanes));
        const size_t ng = P->execs.size();
        for (size_t j = 0; j < ng; j++) {   // every fork before any join, so the graphs run together
            HIPCHK(hipEventRecord(P->gs[j]->ev_lane, s));
            HIPCHK(hipStreamWaitEvent(lanes[j % lanes.size()], P->gs[j]->ev_lane, 0));
        }
        // each graph launched from a host thread of its own: a launch submits its nodes one by one, so
        // from one thread the second group's graph would start only after the first one's nodes
        std::vector<hipError_t> lrc(ng, hipSuccess);
        std::vector<std::thread> threads;
        try {
            for (size_t j = 1; j < ng; j++)
                threads.emplace_back([&, j] { lrc[j] = hipGraphLaunch(P->execs[j], lanes[j % lanes.size()]); });
        } catch (...) {
            for (size_t j = threads.size() + 1; j < ng; j++) lrc[j] = hipGraphLaunch(P->execs[j], lanes[j % lanes.size()]);
        }
        lrc[0] = hipGraphLaunch(P->execs[0], lanes[0]);
        for (auto &t : threads) t.join();
        for (size_t j = 0; j < ng; j++) HIPCHK(lrc[j]);
        for (size_t j = 0; j < ng; j++) {
            HIPCHK(hipEventRecord(P->gs[j]->ev_lane, lanes[j % lanes.size()]));
            HIPCHK(hipStreamWaitEvent(s, P->gs[j]->ev_lane, 0));
        }
    }
    HIPCHK(hipEventRecord(e1, s));
    int64_t cells = 0;
    for (int k = 0; k < count; k++) {
        fkadj_instance *I = P->insts[k];
        CHK(fetch_and_assemble(I->ctx, s, E[k], grad ? grad[k] : nullptr, Lcost ? Lcost + k : nullptr,
                               L0 ? L0 + k : nullptr, LE ? LE + k : nullptr));
        cells += (int64_t)I->rows * I->pitch * I->nt;
    }
    record_timing(smp, e0, e1, cells);
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_plan_destroy(fkadj_plan *P)
{
    if (!P) return FKADJ_ERR_ARG;
    plan_free(P);
    return FKADJ_OK;
}

}  // extern "C"

// ============================================================================ serving
// Continuous batching on one stream.  Every admitted request owns a slot; each round issues one
// forward-kind launch (slots in their forward or recompute phase) and one adjoint launch (slots in
// an adjoint phase), each slot at its own step, through per-slot cursors.  A request's phase
// machine: forward steps 0..nt-1 -> objective, terminal z.gh, f(y_nt) -> for each segment from the
// last: recompute (b-a steps) then adjoint (i = b..a+1) -> fixed-order sums and an async copy of
// z.gh and L0 to pinned host memory.  Slots free and new requests join at round boundaries.  The
// per-request arithmetic is that of fkadj_grad on the same instance, so every result equals the
// request evaluated alone.
namespace {

enum { RQ_QUEUED, RQ_RUNNING, RQ_FINISHING, RQ_DONE, RQ_CANCELLED, RQ_FAILED };
enum { PH_FWD, PH_REC, PH_ADJ };
constexpr int kRoundsInFlight = 8;

// Evaluation contexts of server requests depend only on the instance's shape: a finished request's
// context goes to a pool keyed by that shape and the next request of the same shape takes it (its
// slot record is re-pointed at the new instance).  Pooled contexts count against the budget and are
// freed when a request of another shape needs the room.
struct CtxShape {
    int prec, rows, pitch, nt, K, nb_part, nb_plain, n_num;
    bool operator<(const CtxShape &o) const
    {
        const int a[8] = {prec, rows, pitch, nt, K, nb_part, nb_plain, n_num};
        const int b[8] = {o.prec, o.rows, o.pitch, o.nt, o.K, o.nb_part, o.nb_plain, o.n_num};
        return std::lexicographical_compare(a, a + 8, b, b + 8);
    }
};
CtxShape shape_of(const fkadj_instance *I)
{
    return CtxShape{I->prec, I->rows, I->pitch, I->nt, I->K, I->nb_part, I->nb_plain, I->n_num};
}

struct ServReq {
    int64_t id = 0;
    fkadj_instance *I = nullptr;
    EvalCtx *C = nullptr;        // this request's buffers, taken at admission (pooled by shape)
    std::vector<double> E;
    int state = RQ_QUEUED;
    int rc = FKADJ_OK;
    int slot = -1;
    int phase = PH_FWD, step = 0, seg = 0;
    int zpar = 0;                // half of zb holding the adjoint state (z_nt in half 0)
    bool cancel = false;
    int64_t bytes = 0;
    double *h_out = nullptr;     // pinned: zgh[nt+1], L0
    hipEvent_t done = nullptr;
    std::vector<double> grad;
    double Lc = 0, L0 = 0, LE = 0;
};

}  // namespace

struct fkadj_server {
    int64_t budget = 0, used = 0;
    int max_slots = 0;
    hipStream_t stream = nullptr;
    FusedSlot *d_slots = nullptr;
    FusedSlot *h_slots = nullptr;          // pinned staging of the slot records
    hipEvent_t slot_ev[MAX_SLOTS] = {};    // completion of each slot record's last upload
    std::vector<ServReq *> slot_req;
    std::deque<ServReq *> queue;
    std::vector<ServReq *> finishing;
    std::map<int64_t, ServReq *> reqs;
    int64_t next_id = 1;
    hipEvent_t round_ev[kRoundsInFlight] = {};
    int64_t rounds = 0;
    // free evaluation contexts by shape (their bytes are part of `used`), pinned result buffers by
    // length, completion events
    std::map<CtxShape, std::vector<EvalCtx *>> ctx_pool;
    std::map<int, std::vector<double *>> out_pool;
    std::vector<hipEvent_t> ev_pool;
    // failure injection for tests: the request with this id reports a device error when its result
    // is fetched, as a failed completion would (environment FKADJ_TEST_FAIL_REQUEST at creation)
    int64_t fail_id = 0;
};

namespace {

// A slot's record is uploaded in stream order when a request takes the slot: kernels of the
// previous occupant, enqueued earlier, still read the old record.  The staging copy is reused only
// after its previous upload has completed.
int server_upload_slot(fkadj_server *S, int k, const FusedSlot &rec)
{
    HIPCHK(hipEventSynchronize(S->slot_ev[k]));
    S->h_slots[k] = rec;
    HIPCHK(hipMemcpyAsync(S->d_slots + k, S->h_slots + k, sizeof(FusedSlot), hipMemcpyHostToDevice, S->stream));
    HIPCHK(hipEventRecord(S->slot_ev[k], S->stream));
    return FKADJ_OK;
}

// Retire a running request from its slot: results (or nothing, if cancelled) are copied out
// asynchronously; the slot is free for the next round.
int server_finish(fkadj_server *S, ServReq *r, bool cancelled)
{
    fkadj_instance *I = r->I;
    EvalCtx *C = r->C;
    if (!cancelled) {
        sum_rows_launch(C->d_part, I->nb_part, I->nt, C->d_zgh, S->stream);
        HIPCHK(hipMemcpyAsync(r->h_out, C->d_zgh, sizeof(double) * (I->nt + 1), hipMemcpyDeviceToHost, S->stream));
        HIPCHK(hipMemcpyAsync(r->h_out + I->nt + 1, C->d_L0, sizeof(double), hipMemcpyDeviceToHost, S->stream));
    } else {
        r->rc = FKADJ_ERR_CANCELLED;
    }
    HIPCHK(hipEventRecord(r->done, S->stream));
    S->slot_req[r->slot] = nullptr;
    r->slot = -1;
    r->state = RQ_FINISHING;
    S->finishing.push_back(r);
    return FKADJ_OK;
}

// Returns a finished request's context, pinned buffer and event to the pools (the context's bytes
// stay counted in `used` while pooled).
void server_release(fkadj_server *S, ServReq *r)
{
    if (r->C) S->ctx_pool[shape_of(r->I)].push_back(r->C);
    r->C = nullptr;
    if (r->h_out) S->out_pool[r->I->nt + 2].push_back(r->h_out);
    r->h_out = nullptr;
    if (r->done) S->ev_pool.push_back(r->done);
    r->done = nullptr;
}

// Frees pooled contexts (largest first) until `need` more bytes fit the budget; false if they do not.
bool server_make_room(fkadj_server *S, int64_t need)
{
    while (S->used + need > S->budget) {
        EvalCtx *victim = nullptr;
        std::vector<EvalCtx *> *from = nullptr;
        for (auto &kv : S->ctx_pool)
            for (auto *C : kv.second)
                if (!victim || C->bytes > victim->bytes) { victim = C; from = &kv.second; }
        if (!victim) return false;
        from->erase(std::find(from->begin(), from->end(), victim));
        S->used -= victim->bytes;
        (void)hipStreamSynchronize(S->stream);   // its last request's kernels may still read it
        ctx_free(victim);
    }
    return true;
}

// Completed requests: assemble results and return their context and buffers to the pools.
int server_reap(fkadj_server *S, bool wait)
{
    int n = 0;
    for (size_t k = 0; k < S->finishing.size();) {
        ServReq *r = S->finishing[k];
        hipError_t q = wait ? hipEventSynchronize(r->done) : hipEventQuery(r->done);
        if (q == hipErrorNotReady) { k++; continue; }
        if (q == hipSuccess && r->id == S->fail_id) {
            q = hipErrorLaunchFailure;
            snprintf(g_errmsg, sizeof g_errmsg, "injected completion failure of request %lld", (long long)r->id);
        }
        if (q != hipSuccess && r->rc == FKADJ_OK) r->rc = FKADJ_ERR_DEVICE;
        fkadj_instance *I = r->I;
        if (r->rc == FKADJ_OK) {
            const double *zgh = r->h_out;
            const double L0v = r->h_out[I->nt + 1];
            if (!all_finite(zgh, I->nt + 1) || !std::isfinite(L0v)) {
                r->rc = FKADJ_ERR_NONFINITE;
            } else {
                r->grad.resize(I->ncontrols);
                assemble(I, r->E.data(), zgh, L0v, r->grad.data(), &r->Lc, &r->L0, &r->LE);
            }
        }
        r->state = r->rc == FKADJ_OK ? RQ_DONE : (r->rc == FKADJ_ERR_CANCELLED ? RQ_CANCELLED : RQ_FAILED);
        server_release(S, r);
        I->busy--;
        S->finishing.erase(S->finishing.begin() + k);
        n++;
    }
    return n;
}

// Admit queued requests in FIFO order while slots and budget allow.  Each request gets its own
// evaluation context, so several requests on one instance run side by side.
int server_admit(fkadj_server *S)
{
    for (size_t q = 0; q < S->queue.size();) {
        int free_slot = -1;
        for (int k = 0; k < S->max_slots; k++)
            if (!S->slot_req[k]) { free_slot = k; break; }
        if (free_slot < 0) return FKADJ_OK;
        ServReq *r = S->queue[q];
        fkadj_instance *I = r->I;
        const int64_t bytes = fkadj_request_bytes(I);
        if (bytes > S->budget) {   // can never fit: fail it instead of blocking the queue
            r->rc = FKADJ_ERR_NOMEM;
            r->state = RQ_FAILED;
            I->busy--;
            S->queue.erase(S->queue.begin() + q);
            continue;
        }
        // a pooled context of this shape (already counted in `used`), else room for a new one
        auto &pool = S->ctx_pool[shape_of(I)];
        int rc = FKADJ_OK;
        if (!pool.empty()) {
            r->C = pool.back();
            pool.pop_back();
            r->C->I = I;
            ctx_fill_slot(r->C);
            // the objective writes the terminal adjoint at active cells only: hole cells of this
            // instance must read 0, whatever the previous instance left there
            if (hipMemsetAsync(r->C->d_z, 0, 2 * I->st_bytes, S->stream) != hipSuccess) rc = FKADJ_ERR_DEVICE;
        } else {
            if (!server_make_room(S, bytes)) return FKADJ_OK;   // wait for running requests to finish
            rc = ctx_create(I, S->stream, &r->C, false);
            if (rc == FKADJ_OK) S->used += bytes;
        }
        if (rc == FKADJ_OK) rc = build_tables(I, r->C, r->E.data(), S->stream);
        if (rc == FKADJ_OK) {
            auto &op = S->out_pool[I->nt + 2];
            if (!op.empty()) {
                r->h_out = op.back();
                op.pop_back();
            } else if (hipHostMalloc((void **)&r->h_out, sizeof(double) * (I->nt + 2)) != hipSuccess) {
                rc = FKADJ_ERR_NOMEM;
            }
        }
        if (rc == FKADJ_OK) {
            if (!S->ev_pool.empty()) {
                r->done = S->ev_pool.back();
                S->ev_pool.pop_back();
            } else if (hipEventCreateWithFlags(&r->done, hipEventDisableTiming) != hipSuccess) {
                rc = FKADJ_ERR_DEVICE;
            }
        }
        if (rc == FKADJ_OK &&
            hipMemcpyAsync(I->state(r->C->d_ck, 0), I->d_y0, I->st_bytes, hipMemcpyDeviceToDevice, S->stream) != hipSuccess)
            rc = FKADJ_ERR_DEVICE;
        if (rc == FKADJ_OK) rc = server_upload_slot(S, free_slot, r->C->slot_h);
        S->queue.erase(S->queue.begin() + q);
        r->bytes = bytes;
        if (rc != FKADJ_OK) {
            r->rc = rc;
            r->state = RQ_FAILED;
            I->busy--;
            server_release(S, r);   // a context it took stays pooled (and counted)
            continue;
        }
        r->slot = free_slot;
        r->state = RQ_RUNNING;
        r->phase = PH_FWD;
        r->step = 0;
        S->slot_req[free_slot] = r;
    }
    return FKADJ_OK;
}

// One round: one forward-kind launch over the slots in a forward or recompute phase and one
// adjoint launch over the slots in an adjoint phase (per tile configuration; each launch lists only
// its slots), then every slot's phase machine advances by one step.
int server_round(fkadj_server *S)
{
    Cursors fc, ac, rc_rhs;
    LaunchMap fm[NUM_TILE_CFG], am[NUM_TILE_CFG], rm[NUM_TILE_CFG];
    for (int c = 0; c < NUM_TILE_CFG; c++) {
        launch_map_clear(fm[c]);
        launch_map_clear(am[c]);
        launch_map_clear(rm[c]);
    }
    for (int k = 0; k < MAX_SLOTS; k++) fc.c[k] = ac.c[k] = rc_rhs.c[k] = SlotCursor{MODE_IDLE, 0, 0, 0};
    int64_t fcells = 0, acells = 0;   // cells of this round's forward-kind and adjoint launches
    std::map<int, int64_t> fpc, apc;  // the same per pitch
    std::map<int, std::vector<int>> fph, aph;   // heights of the instances per pitch
    for (int k = 0; k < S->max_slots; k++)
        if (S->slot_req[k]) {
            const fkadj_instance *I = S->slot_req[k]->I;
            const int64_t cells = (int64_t)I->rows * I->pitch;
            const bool fw = S->slot_req[k]->phase != PH_ADJ;
            (fw ? fcells : acells) += cells;
            (fw ? fpc : apc)[I->pitch] += cells;
            (fw ? fph : aph)[I->pitch].push_back(I->rows);
        }
    // the slots of a pitch that the row-rolling kernels support take them when that pitch alone has
    // 4M cells or more in the launch (one launch per such pitch; the same results as the tiled
    // kernels); the other slots share one tiled launch.  A rolling launch fills the GPU from 4M cells
    // on (256 CUs x 16 waves x 1024 cells per wave of a 16-row strip); a smaller one would leave
    // part of it idle while the round's launches run one after the other on the server's stream.
    const int rmode = g_roll_mode.load(), hs = g_roll_hs.load();
    auto roll_for = [&](const std::map<int, int64_t> &pc, int pitch) {
        if (!roll_pitch_supported(pitch)) return false;
        return rmode == 2 || (rmode == 1 && pc.at(pitch) >= 2 * kRollCells);
    };
    // strip rows per pitch and launch kind: the library setting, else auto_strip_rows (one launch at
    // a time on the server's stream)
    std::map<int, int> fhs, ahs;
    for (auto &kv : fph) fhs[kv.first] = hs > 0 ? hs : auto_strip_rows(kv.second, kv.first, 1);
    for (auto &kv : aph) ahs[kv.first] = hs > 0 ? hs : auto_strip_rows(kv.second, kv.first, 1);
    std::map<int, LaunchMap> rfm, ram;
    auto rmap = [](std::map<int, LaunchMap> &mm, int pitch) -> LaunchMap & {
        auto it = mm.find(pitch);
        if (it == mm.end()) {
            it = mm.emplace(pitch, LaunchMap{}).first;
            launch_map_clear(it->second);
        }
        return it->second;
    };
    for (int k = 0; k < S->max_slots; k++) {
        ServReq *r = S->slot_req[k];
        if (!r) continue;
        const fkadj_instance *I = r->I;
        const int K = I->K, f = fwd_cfg_of(I, fcells);
        if (r->phase == PH_ADJ) {
            ac.c[k] = SlotCursor{MODE_ADJ, r->step, r->seg * K, r->zpar};
            const int ac = adj_cfg_of(I, acells);
            if (roll_for(apc, I->pitch))
                launch_map_add(rmap(ram, I->pitch), k, (I->rows + ahs[I->pitch] - 1) / ahs[I->pitch]);
            else
                launch_map_add(am[ac], k, tiles_of(I->rows, I->pitch, ac));
        } else {
            fc.c[k] = r->phase == PH_FWD ? SlotCursor{MODE_FWD, r->step, 0, 0}
                                         : SlotCursor{MODE_REC, r->seg * K + r->step, r->seg * K, 0};
            if (roll_for(fpc, I->pitch))
                launch_map_add(rmap(rfm, I->pitch), k, (I->rows + fhs[I->pitch] - 1) / fhs[I->pitch]);
            else
                launch_map_add(fm[f], k, tiles_of(I->rows, I->pitch, f));
        }
    }
    for (int c = 1; c < NUM_TILE_CFG; c++) fused_forward_launch(c, S->d_slots, fm[c], fc, S->stream);
    for (auto &kv : rfm) roll_launch(false, kv.first, S->d_slots, kv.second, fc, fhs[kv.first], S->stream);
    for (int c = 1; c < NUM_TILE_CFG; c++) fused_adjoint_launch(c, S->d_slots, am[c], ac, S->stream);
    for (auto &kv : ram) roll_launch(true, kv.first, S->d_slots, kv.second, ac, ahs[kv.first], S->stream);
    HIPCHK(hipGetLastError());
    for (int k = 0; k < S->max_slots; k++) {
        ServReq *r = S->slot_req[k];
        if (!r) continue;
        fkadj_instance *I = r->I;
        const int nt = I->nt, K = I->K;
        if (r->phase == PH_FWD) {
            if (++r->step == nt) {
                terminal<float>(r->C, S->stream);               // objective, z(T), z.gh at T
                rc_rhs.c[k] = SlotCursor{MODE_RHS, nt - 1, 0, 0};
                const int fr = fwd_cfg_of(I, (int64_t)I->rows * I->pitch);
                launch_map_add(rm[fr], k, tiles_of(I->rows, I->pitch, fr));
                r->phase = PH_REC;
                r->seg = (nt + K - 1) / K - 1;
                r->step = 0;
                r->zpar = 0;
            }
        } else if (r->phase == PH_REC) {
            const int a = r->seg * K, b = std::min(a + K, nt);
            if (++r->step == b - a) { r->phase = PH_ADJ; r->step = b; }
        } else {
            const int a = r->seg * K;
            r->zpar ^= 1;
            if (--r->step == a) {
                if (--r->seg < 0) CHK(server_finish(S, r, false));
                else { r->phase = PH_REC; r->step = 0; }
            }
        }
    }
    for (int c = 1; c < NUM_TILE_CFG; c++) fused_forward_launch(c, S->d_slots, rm[c], rc_rhs, S->stream);
    HIPCHK(hipGetLastError());
    // cancellations take effect at the round boundary
    for (int k = 0; k < S->max_slots; k++) {
        ServReq *r = S->slot_req[k];
        if (r && r->cancel) CHK(server_finish(S, r, true));
    }
    return FKADJ_OK;
}

}  // namespace

extern "C" {

int fkadj_server_create(int64_t budget_bytes, int max_slots, fkadj_server **out)
{
    FKADJ_ENTRY
    if (!out || budget_bytes <= 0 || max_slots < 1 || max_slots > MAX_SLOTS) return FKADJ_ERR_ARG;
    auto *S = new (std::nothrow) fkadj_server();
    if (!S) return FKADJ_ERR_NOMEM;
    S->budget = budget_bytes;
    S->max_slots = max_slots;
    if (const char *f = getenv("FKADJ_TEST_FAIL_REQUEST")) S->fail_id = atoll(f);
    S->slot_req.assign(max_slots, nullptr);
    bool ok = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void **)&S->d_slots, sizeof(FusedSlot) * max_slots) == hipSuccess &&
              hipHostMalloc((void **)&S->h_slots, sizeof(FusedSlot) * max_slots, hipHostMallocDefault) == hipSuccess;
    for (int k = 0; k < kRoundsInFlight && ok; k++)
        ok = hipEventCreateWithFlags(&S->round_ev[k], hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < max_slots && ok; k++)   // recorded once, so the first upload need not wait
        ok = hipEventCreateWithFlags(&S->slot_ev[k], hipEventDisableTiming) == hipSuccess &&
             hipEventRecord(S->slot_ev[k], S->stream) == hipSuccess;
    if (!ok) {
        fkadj_server_destroy(S);
        return FKADJ_ERR_DEVICE;
    }
    *out = S;
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_server_destroy(fkadj_server *S)
{
    if (!S) return FKADJ_ERR_ARG;
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    for (int k = 0; k < S->max_slots; k++)
        if (S->slot_req[k]) (void)server_finish(S, S->slot_req[k], true);
    (void)server_reap(S, true);
    for (auto &kv : S->reqs) {
        ServReq *r = kv.second;
        if (r->state == RQ_QUEUED) r->I->busy--;
        ctx_free(r->C);
        if (r->h_out) (void)hipHostFree(r->h_out);
        if (r->done) (void)hipEventDestroy(r->done);
        delete r;
    }
    for (auto &kv : S->ctx_pool)
        for (auto *C : kv.second) ctx_free(C);
    for (auto &kv : S->out_pool)
        for (auto *h : kv.second) (void)hipHostFree(h);
    for (auto e : S->ev_pool) (void)hipEventDestroy(e);
    if (S->d_slots) (void)hipFree(S->d_slots);
    if (S->h_slots) (void)hipHostFree(S->h_slots);
    for (int k = 0; k < kRoundsInFlight; k++)
        if (S->round_ev[k]) (void)hipEventDestroy(S->round_ev[k]);
    for (int k = 0; k < MAX_SLOTS; k++)
        if (S->slot_ev[k]) (void)hipEventDestroy(S->slot_ev[k]);
    if (S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return FKADJ_OK;
}

int fkadj_server_submit(fkadj_server *S, fkadj_instance *I, const double *E, int64_t *request_id)
{
    FKADJ_ENTRY
    if (!S || !I || !E || !request_id) return FKADJ_ERR_ARG;
    if (I->prec != FKADJ_F32) return FKADJ_ERR_ARG;
    if (!I->state_set) return FKADJ_ERR_STATE;
    if (!all_finite(E, I->ncontrols)) return FKADJ_ERR_NONFINITE;
    auto *r = new (std::nothrow) ServReq();
    if (!r) return FKADJ_ERR_NOMEM;
    r->id = S->next_id++;
    r->I = I;
    r->E.assign(E, E + I->ncontrols);
    S->reqs[r->id] = r;
    S->queue.push_back(r);
    I->busy++;   // released when the request finishes, fails or is cancelled
    *request_id = r->id;
    return FKADJ_OK;
    FKADJ_EXIT
}

int fkadj_server_cancel(fkadj_server *S, int64_t id)
{
    if (!S) return FKADJ_ERR_ARG;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return FKADJ_ERR_NOTFOUND;
    ServReq *r = it->second;
    if (r->state == RQ_QUEUED) {
        for (size_t q = 0; q < S->queue.size(); q++)
            if (S->queue[q] == r) { S->queue.erase(S->queue.begin() + q); break; }
        r->rc = FKADJ_ERR_CANCELLED;
        r->state = RQ_CANCELLED;
        r->I->busy--;
    } else if (r->state == RQ_RUNNING) {
        r->cancel = true;
    }
    return FKADJ_OK;
}

int fkadj_server_run(fkadj_server *S, int max_rounds)
{
    FKADJ_ENTRY
    if (!S || max_rounds < 0) return FKADJ_ERR_ARG;
    int completed = 0;
    for (int round = 0; max_rounds == 0 || round < max_rounds; round++) {
        completed += server_reap(S, false);
        CHK(server_admit(S));
        bool running = false;
        for (auto *r : S->slot_req) running |= (r != nullptr);
        if (!running) {
            if (!S->finishing.empty()) {
                completed += server_reap(S, true);
                continue;
            }
            break;
        }
        if (S->rounds >= kRoundsInFlight) HIPCHK(hipEventSynchronize(S->round_ev[S->rounds % kRoundsInFlight]));
        CHK(server_round(S));
        HIPCHK(hipEventRecord(S->round_ev[S->rounds % kRoundsInFlight], S->stream));
        S->rounds++;
        if (max_rounds == 0 && completed > 0) break;
    }
    completed += server_reap(S, false);
    return completed;
    FKADJ_EXIT
}

int fkadj_server_poll(fkadj_server *S, int64_t id, double *grad, double *Lcost, double *L0, double *LE)
{
    if (!S) return FKADJ_ERR_ARG;
    (void)server_reap(S, false);
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return FKADJ_ERR_NOTFOUND;
    ServReq *r = it->second;
    if (r->state == RQ_QUEUED || r->state == RQ_RUNNING || r->state == RQ_FINISHING) return FKADJ_ERR_BUSY;
    int rc = r->rc;
    if (rc == FKADJ_OK) {
        if (grad) std::copy(r->grad.begin(), r->grad.end(), grad);
        if (Lcost) *Lcost = r->Lc;
        if (L0) *L0 = r->L0;
        if (LE) *LE = r->LE;
    }
    if (r->done) (void)hipEventDestroy(r->done);
    delete r;
    S->reqs.erase(it);
    return rc;
}

int fkadj_server_pending(const fkadj_server *S)
{
    if (!S) return FKADJ_ERR_ARG;
    int n = 0;
    for (auto &kv : S->reqs) {
        int st = kv.second->state;
        n += (st == RQ_QUEUED || st == RQ_RUNNING || st == RQ_FINISHING);
    }
    return n;
}

}  // extern "C"
