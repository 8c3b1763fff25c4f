/* c_abi_client.c -- a plain C99 caller of libfkadj (include/fkadj.h), no Python involved.
 *
 * Checks, on a small synthetic problem (48 x 80 cells, one 3x3 hole, piecewise pulse of 4
 * amplitudes, 40 steps):
 *   1. the FP64 gradient dJ/dA_k against central finite differences of J with alpha = 0, so the
 *      whole gradient comes from the adjoint solve (relative 1e-3; measured 4e-4, the method's own
 *      second-order mismatch: the functional gradient at the knots integrated over each piece by the
 *      trapezoid rule; a wrong sign, factor or time index fails by orders of magnitude);
 *   2. FP32 against FP64 (relative 1e-4);
 *   3. a batch of two FP32 instances, a batch plan (HIP graph) and a server, each equal to the
 *      instances' own results bit for bit;
 *   4. the forward drop-in (fkadj_forward_ref): floor(40/15) = 2 snapshots in the reference's
 *      full-grid layout with NaN at the hole, equal to a forward run on an FP64 instance.
 * Prints "c abi ok" and exits 0 on success; prints the failed check and exits 1 otherwise.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "fkadj.h"

#define NR 48
#define NC 80
#define NT 40
#define NAMP 4

static int fails = 0;

#define CHECK(cond, ...)                                   \
    do {                                                   \
        if (!(cond)) {                                     \
            fprintf(stderr, "FAILED %s:%d: ", __FILE__, __LINE__); \
            fprintf(stderr, __VA_ARGS__);                  \
            fprintf(stderr, "\n");                         \
            fails++;                                       \
        }                                                  \
    } while (0)

#define CALL(expr)                                                                   \
    do {                                                                             \
        int rc_ = (expr);                                                            \
        if (rc_ != FKADJ_OK) {                                                       \
            fprintf(stderr, "FAILED %s -> %d (%s)\n", #expr, rc_, fkadj_strerror(rc_)); \
            exit(1);                                                                 \
        }                                                                            \
    } while (0)

static int32_t active_index[NR * NC];
static double y0v[3 * NR * NC];
static int n_num = 0;

static void make_problem(fkadj_problem *p, int precision, double shift)
{
    int i, j, k = 0;
    for (i = 0; i < NR; i++)
        for (j = 0; j < NC; j++) {
            int hole = (i >= 20 && i < 23 && j >= 30 && j < 33);
            active_index[i * NC + j] = hole ? 0 : ++k;
        }
    n_num = k;
    for (i = 0; i < NR; i++)
        for (j = 0; j < NC; j++) {
            int a = active_index[i * NC + j];
            double front = 0.3 * NC + 6.0 * sin(0.2 * i + shift);
            double u = 1.0 / (1.0 + exp((j - front) / 1.5));
            if (!a) continue;
            y0v[a - 1] = u;
            y0v[n_num + a - 1] = 1.0 - 0.9 * u;
            y0v[2 * n_num + a - 1] = 0.95 - 0.3 * u;
        }
    memset(p, 0, sizeof *p);
    p->nrows = NR;
    p->ncols = NC;
    p->active_index = active_index;
    fkadj_model_reference(&p->model);
    p->model.h = 0.025;
    p->dt = 0.02;
    p->nt = NT;
    p->pulse_mode = FKADJ_PULSE_PIECEWISE;
    p->n_amp = NAMP;
    p->alpha = 2e-3;
    p->gamma[0] = 1.0;
    p->s_diag = 2.0 / n_num;
    p->precision = precision;
    p->ckpt = 8;
}

static double rel_diff(const double *a, const double *b, int n)
{
    double num = 0, den = 0;
    int i;
    for (i = 0; i < n; i++) {
        num += (a[i] - b[i]) * (a[i] - b[i]);
        den += b[i] * b[i];
    }
    return sqrt(num / den);
}

int main(void)
{
    fkadj_problem p;
    fkadj_instance *f64 = NULL, *f32a = NULL, *f32b = NULL;
    double A[NAMP] = {3.0, 7.5, 1.0, 5.0}, B[NAMP] = {6.0, 2.0, 4.5, 8.0};
    double g64[NAMP], g32[NAMP], gB[NAMP], Lc, L0, LE, worst = 0.0;
    int k;

    /* 1. FP64 gradient against central differences of J, with alpha = 0 so that the whole
     *    gradient comes from the adjoint solve (the field-energy term is exact either way) */
    make_problem(&p, FKADJ_F64, 0.0);
    p.alpha = 0.0;
    CALL(fkadj_create(&p, &f64));
    CALL(fkadj_set_state(f64, y0v));
    CALL(fkadj_grad(f64, A, g64, &Lc, &L0, &LE));
    CHECK(Lc == 0.5 * L0, "Lcost = L0/2 for alpha = 0");
    for (k = 0; k < NAMP; k++) {
        double Ap[NAMP], Am[NAMP], Jp, Jm, h = 1e-3 * (1.0 + fabs(A[k])), fd;
        memcpy(Ap, A, sizeof A);
        memcpy(Am, A, sizeof A);
        Ap[k] += h;
        Am[k] -= h;
        CALL(fkadj_grad(f64, Ap, NULL, &Jp, NULL, NULL));
        CALL(fkadj_grad(f64, Am, NULL, &Jm, NULL, NULL));
        fd = (Jp - Jm) / (2 * h);
        if (fabs(fd - g64[k]) > worst * fabs(g64[k])) worst = fabs(fd - g64[k]) / fabs(g64[k]);
        CHECK(fabs(fd - g64[k]) <= 1e-3 * fabs(g64[k]), "dJ/dA_%d: adjoint %.9g, central difference %.9g", k,
              g64[k], fd);
    }
    CALL(fkadj_destroy(f64));
    make_problem(&p, FKADJ_F64, 0.0);     /* the FP64 reference of check 2 (alpha as for FP32) */
    CALL(fkadj_create(&p, &f64));
    CALL(fkadj_set_state(f64, y0v));
    CALL(fkadj_grad(f64, A, g64, NULL, NULL, NULL));

    /* 2. FP32 against FP64 */
    make_problem(&p, FKADJ_F32, 0.0);
    CALL(fkadj_create(&p, &f32a));
    CALL(fkadj_set_state(f32a, y0v));
    CALL(fkadj_grad(f32a, A, g32, NULL, NULL, NULL));
    CHECK(rel_diff(g32, g64, NAMP) < 1e-4, "FP32 vs FP64 gradient %.3g", rel_diff(g32, g64, NAMP));

    /* 3. batch, plan and server against the instances alone */
    make_problem(&p, FKADJ_F32, 1.3);
    CALL(fkadj_create(&p, &f32b));
    CALL(fkadj_set_state(f32b, y0v));
    CALL(fkadj_grad(f32b, B, gB, NULL, NULL, NULL));
    {
        fkadj_instance *insts[2];
        const double *E[2];
        double gb0[NAMP], gb1[NAMP], gp0[NAMP], gp1[NAMP], gs[NAMP], Lb[2];
        double *G[2];
        fkadj_plan *plan = NULL;
        fkadj_server *srv = NULL;
        int64_t rid;
        int done = 0, guard = 0;
        insts[0] = f32a;
        insts[1] = f32b;
        E[0] = A;
        E[1] = B;
        G[0] = gb0;
        G[1] = gb1;
        CALL(fkadj_batch_grad(2, insts, E, G, Lb, NULL, NULL));
        CHECK(memcmp(gb0, g32, sizeof g32) == 0 && memcmp(gb1, gB, sizeof gB) == 0, "batch equals alone");
        CALL(fkadj_plan_create(2, insts, &plan));
        G[0] = gp0;
        G[1] = gp1;
        CALL(fkadj_plan_grad(plan, E, G, NULL, NULL, NULL));
        CHECK(memcmp(gp0, g32, sizeof g32) == 0 && memcmp(gp1, gB, sizeof gB) == 0, "plan equals alone");
        CHECK(fkadj_destroy(f32a) == FKADJ_ERR_BUSY, "an instance held by a plan is not destroyed");
        CALL(fkadj_plan_destroy(plan));
        CALL(fkadj_server_create((int64_t)1 << 32, 4, &srv));
        CALL(fkadj_server_submit(srv, f32b, B, &rid));
        while (!done && guard++ < 100000) {
            int n = fkadj_server_run(srv, 0);
            CHECK(n >= 0, "server run %d", n);
            if (n < 0) break;
            done = fkadj_server_pending(srv) == 0;
        }
        CALL(fkadj_server_poll(srv, rid, gs, NULL, NULL, NULL));
        CHECK(memcmp(gs, gB, sizeof gB) == 0, "served equals alone");
        CALL(fkadj_server_destroy(srv));
    }
    CALL(fkadj_destroy(f32b));
    CALL(fkadj_destroy(f32a));
    CALL(fkadj_destroy(f64));

    /* 3b. kept trajectory segments (a budget for 3 of the 5 segments) change no result; the footprint
     *     counts them and fits the free device memory */
    {
        int64_t fr = 0, tot = 0, plain_bytes, kept_bytes;
        fkadj_instance *f32k = NULL;
        double gk[NAMP];
        const int64_t seg = (int64_t)(2 * 8 + 1) * 3 * 4 * NR * 128;   /* K = 8, pitch 128 */
        make_problem(&p, FKADJ_F32, 0.0);
        plain_bytes = fkadj_footprint_bytes(&p);
        p.keep_bytes = 3 * seg + seg / 2;
        kept_bytes = fkadj_footprint_bytes(&p);
        CHECK(kept_bytes - plain_bytes == 3 * seg, "footprint of 3 kept segments %lld", (long long)(kept_bytes - plain_bytes));
        CALL(fkadj_mem_info(&fr, &tot));
        CHECK(0 < fr && fr <= tot && kept_bytes < fr, "free %lld of %lld bytes", (long long)fr, (long long)tot);
        CALL(fkadj_create(&p, &f32k));
        CALL(fkadj_set_state(f32k, y0v));
        CALL(fkadj_grad(f32k, A, gk, NULL, NULL, NULL));
        CHECK(memcmp(gk, g32, sizeof g32) == 0, "kept segments change the gradient");
        CALL(fkadj_destroy(f32k));
    }

    /* 4. forward drop-in against a forward run on a handle */
    {
        static double ys[2 * 3 * NR * NC], yf[3 * NR * NC], snaps[2 * 3 * NR * NC], fin[3 * NR * NC];
        static double Et[NT + 1], gh[NR * NC];
        const int nstep = 15, ns = fkadj_snapshot_count(NT, nstep);
        char pstr[2048];
        int i, v, s2, bad = 0;
        CHECK(ns == 2, "snapshot count %d", ns);
        for (i = 0; i <= NT; i++) Et[i] = 1.5 + sin(0.15 * i);
        for (i = 0; i < n_num; i++) gh[i] = (double)(i % 3 - 1);   /* ghost_lr in {-1, 0, 1} */
        make_problem(&p, FKADJ_F64, 0.4);
        p.pulse_mode = FKADJ_PULSE_KNOTS;
        p.n_amp = 0;
        p.model.h = 0.035;              /* fkadj_forward_ref runs the reference's compile-time model */
        p.gh_num = gh;
        CALL(fkadj_forward_ref(0.02, Et, NT, y0v, active_index, NR, NC, n_num, gh, nstep, ns, ys, yf, pstr,
                               sizeof pstr));
        CALL(fkadj_create(&p, &f64));
        CALL(fkadj_set_state(f64, y0v));
        CALL(fkadj_forward(f64, Et, NT, nstep, snaps, fin, 0));
        CALL(fkadj_destroy(f64));
        for (s2 = 0; s2 <= ns; s2++) {
            const double *full = s2 < ns ? ys + (size_t)s2 * 3 * NR * NC : yf;
            const double *packed = s2 < ns ? snaps + (size_t)s2 * 3 * n_num : fin;
            for (v = 0; v < 3; v++)
                for (i = 0; i < NR * NC; i++) {
                    const double x = full[(size_t)v * NR * NC + i];
                    const int a = active_index[i];
                    if (a == 0 ? !isnan(x) : x != packed[(size_t)v * n_num + a - 1]) bad++;
                }
        }
        CHECK(bad == 0, "forward drop-in differs from the handle run at %d values", bad);
        CHECK(strstr(pstr, "tau_w_plus") != NULL, "provenance string");
    }
    if (fails) return 1;
    printf("c abi ok: adjoint vs central differences max rel %.2e; dJ/dA = %.6g %.6g %.6g %.6g\n", worst, g64[0],
           g64[1], g64[2], g64[3]);
    return 0;
}
