/*
 * fkadj.h -- C ABI of the MI355X Fenton-Karma adjoint-gradient library (libfkadj.so).
 *
 * Plain C: pointers, sizes and error codes; no C++ or PyTorch types cross this boundary.
 * All host arrays are caller-owned; the library copies what it needs.  Every entry point
 * returns FKADJ_OK (0) or a negative error code and never exits the process (the reference
 * aborts the MATLAB process on any error, src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1488-1512).
 *
 * Reference interfaces replaced (paths relative to the reference checkout):
 *   fk_grad_Et_comp          src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1569-1572  -> fkadj_grad_ref
 *   mexFunction (gradient)   src/cuda/fk_grad_Et_param_mex.cpp:11-92           -> a MEX gateway over
 *                                                                                 fkadj_grad_ref
 *   fk_forward_Et_comp       src/cuda/fk_forward_Et_mono_PARAMETERS.cu:646     -> fkadj_forward_ref
 *                                                                                 (fkadj_forward on a handle)
 *   compute_ns               src/cuda/fk_forward_Et_param_mex.cpp:7-25         -> fkadj_snapshot_count
 *   compile-time constants   src/cuda/fk_parameters_00_05_alpha_0p8_dx_0p035.h:2-24 -> fkadj_model
 *   parameters_string        src/cuda/fk_parameters_string_00_05_alpha_0p8_dx_0p035.cpp:1-32
 *                                                                              -> fkadj_param_string
 * Grid convention (src/cuda/fk_grad_Et_mono_PARAMETERS.cu:25-28): cell k = i*ncols + j, j contiguous.
 * active_index[k] is the 1-based packed index of cell k, 0 for a hole; packed state vectors are
 * [u; v; w], each n_num long, in increasing k (src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1533-1551).
 */
#ifndef FKADJ_H
#define FKADJ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- error codes */
#define FKADJ_OK             0
#define FKADJ_ERR_ARG       -1  /* bad argument (null pointer, size, mode) */
#define FKADJ_ERR_GEOMETRY  -2  /* grid/mask invalid: a cell with both axial neighbours missing,
                                   S size differs from the active-cell count (reference: exit) */
#define FKADJ_ERR_NOMEM     -3  /* device allocation failed or admission over the memory budget */
#define FKADJ_ERR_DEVICE    -4  /* HIP runtime error (no GPU, launch failure, ...) */
#define FKADJ_ERR_NONFINITE -5  /* NaN/Inf in an input or a result */
#define FKADJ_ERR_STATE     -6  /* call out of order (e.g. gradient before a state was set) */
#define FKADJ_ERR_CANCELLED -7  /* request cancelled before completion */
#define FKADJ_ERR_BUSY      -8  /* no result ready yet (serving poll) */
#define FKADJ_ERR_NOTFOUND  -9  /* unknown request id */
#define FKADJ_ERR_SCRATCH  -10  /* evaluation needs more than the no-allocation array holds */

const char *fkadj_strerror(int code);

/* ---------------------------------------------------------------- model */
/* FK constants, diffusion coefficients (cm^2/ms), smoothing ksm and mesh width h (cm).
 * Field order and values of fkadj_model_reference() are those of
 * src/cuda/fk_parameters_00_05_alpha_0p8_dx_0p035.h:2-24 (h = 0.035 there, compile-time). */
typedef struct fkadj_model {
    double u_c, u_v, tau_d, tau_v1_minus, tau_v2_minus, tau_v_plus, tau_0, tau_r, tau_si,
        u_csi, k_fk, tau_w_minus, tau_w_plus, Du, Dv, Dw, ksm, h;
} fkadj_model;

void fkadj_model_reference(fkadj_model *m);

/* Provenance text: the model constants in effect (replaces the compiled-in parameters_string,
 * returned by the reference as a callee-allocated, unterminated buffer, G:1932-1935).
 * Writes at most cap bytes including the terminating NUL; returns the full length. */
int fkadj_param_string(const fkadj_model *m, char *buf, size_t cap);

/* ---------------------------------------------------------------- reference-shaped entry */
/* Drop-in for fk_grad_Et_comp (src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1569-1938):
 *   same inputs and outputs; the sparse S is plain CSC over packed indices (n_num x n_num,
 *   int64 column pointers and row indices) instead of a MATLAB mxArray; gh_num is the per-active
 *   cell ghost_lr (double, as MATLAB passes it); the provenance string goes into a caller buffer.
 *   E(t) is given at nt+1 knots t_i = i*dt and read at the stage times through a not-a-knot cubic
 *   spline, as in the reference.  Runs the FP64 GPU path with the forward trajectory kept in HBM
 *   as checkpoints every 50 steps (recomputed per segment in the backward sweep).
 * Outputs: gradE[nt+1] = alpha*E_i - z_u(t_i).gh, L0 = sum_v gamma_v y_v^T S y_v at T,
 *   LE = dt*sum E_i^2, Lcost = L0/2 + alpha*LE/2. */
int fkadj_grad_ref(double dt, const double *Et, int nt, const double *y_ini, const int32_t *active_index,
                   int nrows, int ncols, int n_num, const double *gh_num, const double *gamma,
                   const int64_t *S_colptr, const int64_t *S_rowind, const double *S_val, double alpha,
                   double *L0, double *LE, double *Lcost, double *gradE, char *param_str, size_t param_str_cap);

/* ---------------------------------------------------------------- problems and instances */
#define FKADJ_F64 0            /* FP64 state, one kernel per RK stage ("plain" path) */
#define FKADJ_F32 1            /* FP32 state, fused step kernels, FP64 reductions */

#define FKADJ_PULSE_KNOTS     0  /* E given at nt+1 knots, spline at sub-step times (reference) */
#define FKADJ_PULSE_PIECEWISE 1  /* n_amp piecewise-constant amplitudes over equal pieces of T;
                                    E is constant over each time step; the gradient returned is
                                    dJ/dA_k (trapezoid quadrature of alpha*E - z.gh per step) */

typedef struct fkadj_problem {
    int nrows, ncols;                 /* grid; >= 4 each way, one state (3 planes) under 2 GiB */
    const int32_t *active_index;      /* nrows*ncols */
    const double *gh_num;             /* n_num values of ghost_lr (-1, 0, 1); NULL: derived from the mask */
    fkadj_model model;
    double dt;                        /* ms */
    int nt;                           /* time steps; T = nt*dt */
    int pulse_mode;                   /* FKADJ_PULSE_* */
    int n_amp;                        /* PIECEWISE: number of amplitudes, nt % n_amp == 0 */
    double alpha;                     /* energy weight: Lcost = L0/2 + alpha*LE/2 */
    double gamma[3];                  /* objective weights of u, v, w */
    /* objective matrix over packed indices: CSC (n_num x n_num), or S = s_diag*I when S_colptr==NULL */
    const int64_t *S_colptr;
    const int64_t *S_rowind;
    const double *S_val;
    double s_diag;
    int precision;                    /* FKADJ_F64 / FKADJ_F32 */
    int ckpt;                         /* forward checkpoint interval K (steps); 0 -> 50 */
    int fwd_tile;                     /* FP32 forward tile: 1 = 64x16 cells / 256 threads, 2 = 64x32 / 256,
                                         3 = 64x32 / 512, 4 = 64x16 / 512, 5 = 64x32 / 1024,
                                         6 = 64x8 / 512, 7 = 64x8 / 256;
                                         0 -> built-in choice by the cells of the launch, with
                                         two forward-sweep steps per launch under 1M cells (the
                                         forward tile and depth never change a result) */
    int adj_tile;                     /* FP32 adjoint tile, same ids; 0 -> built-in table */
    int lazy_workspace;               /* 1: the evaluation buffers of fkadj_grad / fkadj_forward /
                                         batch calls are allocated on first use (a server gives each
                                         request its own, so an instance used only through a server
                                         holds just its resident data) */
    int64_t keep_bytes;               /* FP32: device bytes the instance may spend on keeping the last
                                         checkpoint segments of the forward trajectory (K+1 states and
                                         K stage-1 slopes each); the forward sweep of fkadj_grad and the
                                         batch calls fills them and the backward pass recomputes only
                                         the segments before them.  Results are unchanged.  0: none
                                         (every segment is recomputed); FP64 and server requests keep
                                         none */
    int fwd_depth;                    /* FP32 forward-sweep time steps per launch on the tiled kernels:
                                         1 or 2 (two needs nt and ckpt even); 0 -> built-in (2 for
                                         launches under 1M cells on the built-in forward tile) */
    int adj_depth;                    /* FP32 adjoint time steps per launch on the tiled kernels: 1 or
                                         2 (two steps of one segment per launch, 64x8 tiles); 0 ->
                                         built-in (1).  The members of a batch group take 2 only if
                                         all ask for it.  Neither depth changes a result */
    int roll_rows;                    /* FP32: strip rows of the row-rolling kernels when this instance
                                         runs on them (a multiple of 8, 8..256): one workgroup per
                                         strip; the members of one pitch in a batch group use it when
                                         all ask for the same.  0 -> the library setting
                                         (fkadj_set_rolling).  Never changes a result */
} fkadj_problem;

typedef struct fkadj_instance fkadj_instance;

/* Uploads geometry, codes and objective matrix; allocates the device state, checkpoint and
 * segment buffers.  The initial state must be set before the first gradient. */
int fkadj_create(const fkadj_problem *p, fkadj_instance **out);
int fkadj_destroy(fkadj_instance *inst);

/* Device bytes held by an instance with this problem (for admission by memory footprint). */
int64_t fkadj_footprint_bytes(const fkadj_problem *p);
/* Device bytes one server request on this instance holds while it is admitted (its evaluation
 * buffers), the figure the server checks against its budget; negative on a bad argument.  No
 * reference counterpart. */
int64_t fkadj_request_bytes(const fkadj_instance *inst);
/* Free and total device memory of the current GPU (bytes), for sizing keep_bytes and server
 * budgets.  No reference counterpart. */
int fkadj_mem_info(int64_t *free_bytes, int64_t *total_bytes);

/* y_packed: 3*n_num values [u; v; w].  Stays resident until replaced. */
int fkadj_set_state(fkadj_instance *inst, const double *y_packed);
int fkadj_get_state(fkadj_instance *inst, double *y_packed);
int fkadj_n_active(const fkadj_instance *inst);
int fkadj_n_controls(const fkadj_instance *inst); /* nt+1 (KNOTS) or n_amp (PIECEWISE) */

/* One gradient evaluation from the resident initial state.
 * E: fkadj_n_controls() values; grad: same length.  Any of the scalar outputs may be NULL. */
int fkadj_grad(fkadj_instance *inst, const double *E, double *grad, double *Lcost, double *L0, double *LE);

/* Forward solve from the resident state over nsteps (<= nt) with field E (same layout as for
 * fkadj_grad).  With advance != 0 the final state replaces the resident state.
 * snap_every > 0: snapshots of the packed state after steps snap_every, 2*snap_every, ...
 * (fkadj_snapshot_count(nsteps, snap_every) of them; the final state is not among them) go to
 * snaps, 3*n_num values each.  y_final (may be NULL): packed final state. */
int fkadj_forward(fkadj_instance *inst, const double *E, int nsteps, int snap_every, double *snaps,
                  double *y_final, int advance);
/* Number of snapshots a forward run of nsteps keeps every snap_every steps, as compute_ns
 * (src/cuda/fk_forward_Et_param_mex.cpp:7-25): nsteps/snap_every - 1 when snap_every divides
 * nsteps (the final state is returned separately), floor(nsteps/snap_every) otherwise. */
int fkadj_snapshot_count(int nsteps, int snap_every);
/* Drop-in for fk_forward_Et_comp (src/cuda/fk_forward_Et_mono_PARAMETERS.cu:646-849; gateway
 * src/cuda/fk_forward_Et_param_mex.cpp:40-121), FP64 GPU path, knot-sampled field with the
 * not-a-knot spline.  Same arguments and output layout: ys holds ns states after steps nstep,
 * 2*nstep, ... (ns <= nt/nstep, normally fkadj_snapshot_count(nt, nstep)) and yf the final state,
 * each three full planes of nrows*ncols values (u, v, w; cell k = i*ncols + j) with NaN at holes.
 * Differences: an error code instead of exit(), the provenance string in a caller buffer. */
int fkadj_forward_ref(double dt, const double *Et, int nt, const double *y_ini, const int32_t *active_index,
                      int nrows, int ncols, int n_num, const double *gh_num, int nstep, int ns, double *ys,
                      double *yf, char *param_str, size_t param_str_cap);

/* Stimulus helper for building initial states: u = value on the cells where mask[k] != 0. */
int fkadj_stimulate(fkadj_instance *inst, const uint8_t *cell_mask, double value);

/* ---------------------------------------------------------------- batched evaluation */
/* Gradients of `count` (<= 64) FP32 instances (any grid sizes, equal nt and ckpt) computed together:
 * every forward, recompute and adjoint step of all instances is one kernel launch over a
 * launch over the tiles of all instances.  Results equal the instances' own fkadj_grad bit for bit. */
int fkadj_batch_grad(int count, fkadj_instance *const *insts, const double *const *E, double *const *grad,
                     double *Lcost, double *L0, double *LE);

/* Forward solve of `count` FP32 instances (equal nt and ckpt) over nsteps in shared launches;
 * each final state replaces the instance's resident state (used to prepare initial states). */
int fkadj_batch_forward(int count, fkadj_instance *const *insts, const double *const *E, int nsteps);

/* A batch plan: the launch lists of fkadj_batch_grad for a fixed set of instances and one HIP
 * graph holding a whole evaluation (forward sweep, objective, recompute and adjoint segments,
 * sums), captured once and replayed per call.  Only the field tables move per call (host to
 * device, before the replay).  The plan uses the instances' own evaluation buffers: while it
 * exists the instances cannot be destroyed (FKADJ_ERR_BUSY), and calls on them must not overlap
 * a plan call.  Results equal fkadj_batch_grad (and fkadj_grad) bit for bit.  No reference
 * counterpart. */
typedef struct fkadj_plan fkadj_plan;
int fkadj_plan_create(int count, fkadj_instance *const *insts, fkadj_plan **out);
int fkadj_plan_grad(fkadj_plan *plan, const double *const *E, double *const *grad, double *Lcost, double *L0,
                    double *LE);
int fkadj_plan_destroy(fkadj_plan *plan);

/* ---------------------------------------------------------------- no-allocation evaluation */
/* Bytes of the static device array behind fkadj_grad_noalloc (one per GPU the library runs on, held
 * until the process ends). */
#define FKADJ_NOALLOC_SCRATCH_BYTES ((size_t)256 << 20)

/* fkadj_grad with no device allocation: no counterpart in the reference.  For callers that cannot
 * call hipMalloc (tight serving loops, harnesses that forbid allocation).  Same arguments, results
 * and errors as fkadj_grad, and the same gradient bit for bit; the evaluation buffers (field
 * tables, z.gh partials, checkpoints every ckpt steps, one segment of K+1 states, ring and adjoint
 * states, FP32 stored slopes or FP64 stage planes) live in a static device array of
 * FKADJ_NOALLOC_SCRATCH_BYTES per GPU instead of memory the library allocates.  An evaluation that
 * needs more (about (nt/ckpt + 2 ckpt + 9) states of 3*nrows*pitch values plus 8*nt*(6 + tiles)
 * bytes of tables; pitch = ncols rounded up to 64) returns FKADJ_ERR_SCRATCH before any device
 * work.  The instance keeps its resident data (created with lazy_workspace = 1 it holds nothing
 * else); nothing is kept across calls: every call clears the part of the array it uses.  Calls on
 * one GPU share its array: the library serialises them (each waits for the one before), so calls
 * from several host threads do not race but do not overlap either.  Uses the instance's stream;
 * the first call on a GPU takes the array's address there and creates the few events the sweep
 * orders its streams with. */
int fkadj_grad_noalloc(fkadj_instance *inst, const double *E, double *grad, double *Lcost, double *L0,
                       double *LE);

/* ---------------------------------------------------------------- timing */
typedef struct fkadj_timing {
    double eval_ms;          /* last gradient evaluation, device time between HIP events */
    double fwd_kernel_ms;    /* mean duration of one forward-step launch (all tiles of the call) */
    double adj_kernel_ms;    /* mean duration of one adjoint-step launch */
    int64_t fwd_launches, adj_launches;
    int64_t cell_steps;      /* active+padding cells x steps processed per sweep */
} fkadj_timing;

/* Enables launch timing with HIP events on the next calls (adds small overhead); 0 disables.  One
 * window of 8 back-to-back launches of one kind (forward-kind or adjoint) in every 64 is bracketed by
 * two events on its own stream; the mean launch duration is the windows' time over their launches.
 * The groups of a split batch call (fkadj_set_batch_groups) run at the same time, so their launches
 * overlap; set 1 group to time kernels alone. */
int fkadj_set_profiling(int on);
int fkadj_last_timing(fkadj_timing *t);
/* fkadj_batch_grad calls over 2M cells or more run their members as up to n groups (contiguous
 * runs of about equal cells), each issued from a host thread of its own on one of the library's
 * batch streams (created once per device, back to back, so that they sit on different hardware
 * queues), so one group's kernels fill the GPU while another's drain (16 x 512^2: 4-7% faster with
 * 2; 3 and 4 measured slower).  Plans made while n > 1 capture one graph per group and replay them
 * side by side on those streams (16 x 512^2: 9% faster).  Results do not depend on n; concurrent
 * split calls from several host threads share the batch streams.  Process-wide, 1..8, default 2;
 * returns the previous value.  No reference counterpart. */
int fkadj_set_batch_groups(int n);

/* In an FP32 group (a fkadj_grad, a batch call or plan, or one group of a split batch), the members
 * of a pitch of 256, 384, 512, 768 or 1024 cells can run their forward-kind and adjoint steps with
 * row-rolling kernels: one workgroup per strip of strip_rows rows, as wide as the pitch, the four RK
 * stages of a step skewed by one row, so no halo is recomputed across columns; one rolling launch per
 * such pitch, the other members in the tiled launches of the same step.  Results equal the tiled
 * kernels' bit for bit.  mode 0: never; 1 (default): the pitches with 2M cells or more in the group,
 * and in a server round the pitches with 4M cells or more; 2: every such pitch.  strip_rows: a
 * multiple of 8 (8..256), or 0 (default): chosen per launch, the tallest of 64, 32 and 16 rows whose
 * strips still give the launch its share of the workgroups the GPU holds at once (the groups of a
 * split batch share it), else 16.  Applies to groups built after the call (plans keep their
 * kernels).  Process-wide; returns the previous mode.  No reference counterpart. */
int fkadj_set_rolling(int mode, int strip_rows);

/* ---------------------------------------------------------------- serving */
/* Continuous-batching gradient server on one GPU (one host thread drives it).  A request is
 * admitted when a slot is free and its evaluation buffers (checkpoints, segment states, slopes,
 * adjoint states, tables: fkadj_request_bytes of its instance) fit the remaining budget; a finished
 * request's buffers are kept for the next request of the same shape (counted against the budget,
 * freed when a request of another shape needs the room).  Admitted
 * requests advance together, one forward-kind and one adjoint launch per round over all slots,
 * each slot at its own step and phase; new requests join and finished ones leave at round
 * boundaries.  Each result equals the same request evaluated alone (fkadj_grad on the instance)
 * bit for bit.  No reference counterpart (the reference evaluates one gradient per synchronous
 * MEX call). */
typedef struct fkadj_server fkadj_server;

int fkadj_server_create(int64_t budget_bytes, int max_slots /* <= 64 */, fkadj_server **out);
int fkadj_server_destroy(fkadj_server *s);
/* Queue a gradient request on an FP32 instance whose state is set; E is copied.  Several requests
 * on one instance run side by side (each has its own buffers); while any is queued or running,
 * fkadj_set_state, fkadj_stimulate, fkadj_forward and fkadj_destroy on it return FKADJ_ERR_BUSY. */
int fkadj_server_submit(fkadj_server *s, fkadj_instance *inst, const double *E, int64_t *request_id);
/* A queued request is dropped; a running one leaves at the next round boundary (others unaffected). */
int fkadj_server_cancel(fkadj_server *s, int64_t request_id);
/* Issue up to max_rounds rounds (0: until at least one request completes or nothing is left).
 * Returns the number of requests that completed during the call (>= 0) or an error code. */
int fkadj_server_run(fkadj_server *s, int max_rounds);
/* Result of a finished request (then forgotten): FKADJ_ERR_BUSY while queued or running,
 * FKADJ_ERR_CANCELLED, FKADJ_ERR_NOMEM (footprint above the whole budget), ... */
int fkadj_server_poll(fkadj_server *s, int64_t request_id, double *grad, double *Lcost, double *L0, double *LE);
int fkadj_server_pending(const fkadj_server *s); /* queued + running + finishing requests */

#ifdef __cplusplus
}
#endif
#endif /* FKADJ_H */
