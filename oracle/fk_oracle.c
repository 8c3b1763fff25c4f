/*
 * fk_oracle.c -- FP64 CPU restatement of the reference's Fenton-Karma gradient evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the tests, the smoke check and the
 * CPU-baseline leg of the benchmark call it to check or time the HIP library.  No product
 * code path links, loads or calls it.
 *
 * It restates, in plain C with OpenMP over grid rows, the numerics of the reference
 * `fk_grad_Et_comp` (src/cuda/fk_grad_Et_mono_PARAMETERS.cu:1569-1938, cited as G:<line>):
 *   - FK reaction                              G:30-46
 *   - 5-point mirror Laplacian, NaN holes      G:244-440
 *   - RK4 split into four stages               G:460-563, driver G:1689-1724
 *   - not-a-knot cubic spline for E(t)         G:2223 (set), G:2520 (evaluate)
 *   - terminal adjoint z = -gamma*S*y, L0/LE   G:1741-1808
 *   - Hermite midpoint reconstruction          G:1217-1301
 *   - transposed Laplacian, reaction Jacobian^T G:871-1184
 *   - adjoint RK4 stages                       G:1303-1427, driver G:1866-1908
 *   - z.gh reduction (NaN terms skipped)       G:1457-1483
 *   - gradE = alpha*E - zgh                    G:1924-1926
 * Holes are NaN in the state planes exactly as in the reference.  The forward trajectory is
 * kept as checkpoints every `ckpt` steps and recomputed per segment; the recomputation is the
 * same code in the same order, so the result equals a store-all run bit for bit.
 *
 * Precision probe (not part of the reference's algorithm): with round_state != 0, orc_grad rounds the
 * state to FP32 after every forward and recompute step, so that a test can measure what FP32 storage of
 * the trajectory alone, with FP64 arithmetic everywhere else, does to the gradient.
 *
 * Grid convention: cell k = i*ncols + j (j contiguous); nx = nrows, ny = ncols (G:25-28).
 * Build: gcc -O2 -fopenmp -ffp-contract=off -fPIC -shared (no fast-math: NaN semantics matter).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct {
    double u_c, u_v, tau_d, tau_v1_minus, tau_v2_minus, tau_v_plus, tau_0, tau_r, tau_si,
        u_csi, k_fk, tau_w_minus, tau_w_plus, Du, Dv, Dw, ksm, h;
} orc_model;

/* Constants of src/cuda/fk_parameters_00_05_alpha_0p8_dx_0p035.h:2-24 (h is a run-time value here). */
/* Host threads used by the OpenMP loops below (the CPU-baseline timing runs all cores and one). */
void orc_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int orc_max_threads(void) { return omp_get_max_threads(); }

void orc_model_reference(orc_model *m)
{
    m->u_c = 0.13; m->u_v = 0.04; m->tau_d = 0.388; m->tau_v1_minus = 7.6; m->tau_v2_minus = 8.8;
    m->tau_v_plus = 3.324; m->tau_0 = 8.2; m->tau_r = 33.264; m->tau_si = 29.0; m->u_csi = 0.54;
    m->k_fk = 15.0; m->tau_w_minus = 68.0; m->tau_w_plus = 399.99999999999994;
    m->Du = 1.0e-3; m->Dv = 1.0e-5; m->Dw = 1.0e-5; m->ksm = 100.0; m->h = 0.035;
}

/* ---------------------------------------------------------------- reaction (G:30-46) */
void orc_fion(const orc_model *m, double u, double v, double w, double *out)
{
    double p = 0.5 * (1 + tanh(m->ksm * (u - m->u_c)));
    double q = 0.5 * (1 + tanh(m->ksm * (u - m->u_v)));
    double j_fi = -v * p * (1 - u) * (u - m->u_c) / m->tau_d;
    double j_so = u * (1 - p) / m->tau_0 + p / m->tau_r;
    double tvm = q * m->tau_v1_minus + (1 - q) * m->tau_v2_minus;
    double j_si = -w * (1 + tanh(m->k_fk * (u - m->u_csi))) / (2 * m->tau_si);
    out[0] = -(j_fi + j_so + j_si);
    out[1] = ((1 - p) * (1 - v) / tvm) - (p * v) / m->tau_v_plus;
    out[2] = ((1 - p) * (1 - w) / m->tau_w_minus) - (p * w) / m->tau_w_plus;
}

/* Transposed reaction Jacobian, out[3*a + b] = dF_b / dy_a (layout of G:1160-1184).
 * Written from the derivative of orc_fion; pinned against its finite differences. */
void orc_jacT(const orc_model *m, const double *y, double *out)
{
    double u = y[0], v = y[1], w = y[2];
    double tp = tanh(m->ksm * (u - m->u_c));
    double tq = tanh(m->ksm * (u - m->u_v));
    double ts = tanh(m->k_fk * (u - m->u_csi));
    double p = 0.5 * tp + 0.5, q = 0.5 * tq + 0.5;
    double dp = 0.5 * m->ksm * (1 - tp * tp);       /* dp/du */
    double dq = 0.5 * m->ksm * (1 - tq * tq);       /* dq/du */
    double tvm = q * m->tau_v1_minus + (1 - q) * m->tau_v2_minus;
    double dtvm = dq * (m->tau_v1_minus - m->tau_v2_minus);
    double g = (1 - u) * (u - m->u_c);               /* J_fi = -v p g / tau_d */
    double dg = (1 - u) - (u - m->u_c);
    /* F_u = v p g/tau_d - u(1-p)/tau_0 - p/tau_r + w(1+ts)/(2 tau_si) */
    out[0] = v * (dp * g + p * dg) / m->tau_d - ((1 - p) - u * dp) / m->tau_0 - dp / m->tau_r
           + w * m->k_fk * (1 - ts * ts) / (2 * m->tau_si);
    out[3] = p * g / m->tau_d;
    out[6] = (1 + ts) / (2 * m->tau_si);
    /* F_v = (1-p)(1-v)/tvm - p v / tau_v_plus */
    out[1] = -dp * (1 - v) / tvm - (1 - p) * (1 - v) * dtvm / (tvm * tvm) - dp * v / m->tau_v_plus;
    out[4] = -(1 - p) / tvm - p / m->tau_v_plus;
    out[7] = 0.0;
    /* F_w = (1-p)(1-w)/tau_w_minus - p w / tau_w_plus */
    out[2] = -dp * (1 - w) / m->tau_w_minus - dp * w / m->tau_w_plus;
    out[5] = 0.0;
    out[8] = -(1 - p) / m->tau_w_minus - p / m->tau_w_plus;
}

/* ------------------------------------------------------- Laplacian (G:244-440) */
/* One cell: neighbour values with the mirror rule; a NaN (hole) or a wall on one side takes
 * the value of the other side.  NaN at holes. */
static void lap_cell(const orc_model *m, int nx, int ny, const double *y, int i, int j, double *out)
{
    size_t n = (size_t)nx * ny, k = (size_t)i * ny + j;
    double c[3], a[3], b[3], l[3], r[3];
    if (isnan(y[k])) { out[0] = out[1] = out[2] = NAN; return; }
    for (int s = 0; s < 3; s++) {
        const double *p = y + s * n;
        c[s] = p[k];
        if (i == 0) { b[s] = p[k + ny]; a[s] = b[s]; }
        else if (i == nx - 1) { a[s] = p[k - ny]; b[s] = a[s]; }
        else { a[s] = p[k - ny]; b[s] = p[k + ny]; }
        if (j == 0) { r[s] = p[k + 1]; l[s] = r[s]; }
        else if (j == ny - 1) { l[s] = p[k - 1]; r[s] = l[s]; }
        else { l[s] = p[k - 1]; r[s] = p[k + 1]; }
    }
    /* interior mirror (the u plane decides, as in G:356-369, G:399-412) */
    if (i > 0 && i < nx - 1) {
        if (isnan(a[0])) { a[0] = b[0]; a[1] = b[1]; a[2] = b[2]; }
        else if (isnan(b[0])) { b[0] = a[0]; b[1] = a[1]; b[2] = a[2]; }
    }
    if (j > 0 && j < ny - 1) {
        if (isnan(l[0])) { l[0] = r[0]; l[1] = r[1]; l[2] = r[2]; }
        else if (isnan(r[0])) { r[0] = l[0]; r[1] = l[1]; r[2] = l[2]; }
    }
    double hh = m->h * m->h;
    out[0] = m->Du / hh * (a[0] + b[0] + l[0] + r[0] - 4.0 * c[0]);
    out[1] = m->Dv / hh * (a[1] + b[1] + l[1] + r[1] - 4.0 * c[1]);
    out[2] = m->Dw / hh * (a[2] + b[2] + l[2] + r[2] - 4.0 * c[2]);
}

void orc_lap(const orc_model *m, int nx, int ny, const double *y, double *out)
{
    size_t n = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double o[3];
            size_t k = (size_t)i * ny + j;
            lap_cell(m, nx, ny, y, i, j, o);
            out[k] = o[0]; out[k + n] = o[1]; out[k + 2 * n] = o[2];
        }
}

/* ------------------------------------------ transposed Laplacian (G:871-1158) */
/* Neighbour weight: 0 if missing, 2 if the neighbour mirrors onto this cell (its own far
 * neighbour on the same axis is missing), else 1.  NaN at holes. */
static double lapT_weight(const double *u, int nx, int ny, int i, int j, int di, int dj)
{
    int i1 = i + di, j1 = j + dj, i2 = i + 2 * di, j2 = j + 2 * dj;
    if (i1 < 0 || i1 >= nx || j1 < 0 || j1 >= ny) return 0.0;
    if (isnan(u[(size_t)i1 * ny + j1])) return 0.0;
    if (i2 < 0 || i2 >= nx || j2 < 0 || j2 >= ny) return 2.0;
    if (isnan(u[(size_t)i2 * ny + j2])) return 2.0;
    return 1.0;
}

static void lapT_cell(const orc_model *m, int nx, int ny, const double *z, int i, int j, double *out)
{
    size_t n = (size_t)nx * ny, k = (size_t)i * ny + j;
    if (isnan(z[k])) { out[0] = out[1] = out[2] = NAN; return; }
    double pt = lapT_weight(z, nx, ny, i, j, -1, 0), pb = lapT_weight(z, nx, ny, i, j, 1, 0);
    double pl = lapT_weight(z, nx, ny, i, j, 0, -1), pr = lapT_weight(z, nx, ny, i, j, 0, 1);
    double hh = m->h * m->h, D[3] = {m->Du, m->Dv, m->Dw};
    for (int s = 0; s < 3; s++) {
        const double *p = z + s * n;
        double vt = pt != 0.0 ? p[k - ny] : 0.0, vb = pb != 0.0 ? p[k + ny] : 0.0;
        double vl = pl != 0.0 ? p[k - 1] : 0.0, vr = pr != 0.0 ? p[k + 1] : 0.0;
        out[s] = D[s] / hh * (pt * vt + pb * vb + pl * vl + pr * vr - 4.0 * p[k]);
    }
}

void orc_lapT(const orc_model *m, int nx, int ny, const double *z, double *out)
{
    size_t n = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double o[3];
            size_t k = (size_t)i * ny + j;
            lapT_cell(m, nx, ny, z, i, j, o);
            out[k] = o[0]; out[k + n] = o[1]; out[k + 2 * n] = o[2];
        }
}

/* --------------------------------------------- right-hand side (G:1187-1215) */
static void rhs_cell(const orc_model *m, int nx, int ny, const double *y, const double *gh, double E,
                     int i, int j, double *f)
{
    size_t n = (size_t)nx * ny, k = (size_t)i * ny + j;
    double l[3], r[3];
    lap_cell(m, nx, ny, y, i, j, l);
    orc_fion(m, y[k], y[k + n], y[k + 2 * n], r);
    f[0] = l[0] + r[0] + E * gh[k];
    f[1] = l[1] + r[1];
    f[2] = l[2] + r[2];
}

void orc_rhs(const orc_model *m, int nx, int ny, const double *y, const double *gh, double E, double *f)
{
    size_t n = (size_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double o[3];
            size_t k = (size_t)i * ny + j;
            rhs_cell(m, nx, ny, y, gh, E, i, j, o);
            f[k] = o[0]; f[k + n] = o[1]; f[k + 2 * n] = o[2];
        }
}

/* Adjoint vector field fz = L^T z + J_F(y)^T z, NaN at holes (G:1303-1347). */
static void adjf_cell(const orc_model *m, int nx, int ny, const double *z, const double *y, int i, int j,
                      double *fz)
{
    size_t n = (size_t)nx * ny, k = (size_t)i * ny + j;
    double lt[3], yy[3], JT[9];
    lapT_cell(m, nx, ny, z, i, j, lt);
    if (isnan(z[k])) { fz[0] = fz[1] = fz[2] = NAN; return; }
    yy[0] = y[k]; yy[1] = y[k + n]; yy[2] = y[k + 2 * n];
    orc_jacT(m, yy, JT);
    for (int a = 0; a < 3; a++) {
        double s = 0.0;
        for (int b = 0; b < 3; b++) s += JT[3 * a + b] * z[k + b * n];
        fz[a] = lt[a] + s;
    }
}

/* ------------------------------------------------- spline (G:2223, G:2520) */
/* Second derivatives of the not-a-knot cubic spline through (t, y), n >= 4.  The two
 * not-a-knot rows (third derivative continuous at t[1] and t[n-2]) are folded into the first
 * and last interior rows, the interior tridiagonal system is solved by forward elimination
 * and back substitution, and the end values are recovered from the not-a-knot relations. */
int orc_spline_set(int n, const double *t, const double *y, double *ypp)
{
    if (n < 4) return -1;
    int m = n - 2;   /* unknowns ypp[1..n-2] */
    double *lo = malloc(sizeof(double) * m), *di = malloc(sizeof(double) * m);
    double *up = malloc(sizeof(double) * m), *rh = malloc(sizeof(double) * m);
    if (!lo || !di || !up || !rh) { free(lo); free(di); free(up); free(rh); return -2; }
    for (int r = 0; r < m; r++) {
        int i = r + 1;
        double h0 = t[i] - t[i - 1], h1 = t[i + 1] - t[i];
        lo[r] = h0 / 6.0; di[r] = (h0 + h1) / 3.0; up[r] = h1 / 6.0;
        rh[r] = (y[i + 1] - y[i]) / h1 - (y[i] - y[i - 1]) / h0;
    }
    /* ypp0 = ((h0+h1) ypp1 - h0 ypp2) / h1 */
    {
        double h0 = t[1] - t[0], h1 = t[2] - t[1];
        di[0] += lo[0] * (h0 + h1) / h1;
        up[0] -= lo[0] * h0 / h1;
        lo[0] = 0.0;
    }
    /* ypp[n-1] = ((ha+hb) ypp[n-2] - hb ypp[n-3]) / ha, ha = t[n-2]-t[n-3], hb = t[n-1]-t[n-2] */
    {
        double ha = t[n - 2] - t[n - 3], hb = t[n - 1] - t[n - 2];
        di[m - 1] += up[m - 1] * (ha + hb) / ha;
        lo[m - 1] -= up[m - 1] * hb / ha;
        up[m - 1] = 0.0;
    }
    for (int r = 1; r < m; r++) {
        double f = lo[r] / di[r - 1];
        di[r] -= f * up[r - 1];
        rh[r] -= f * rh[r - 1];
    }
    ypp[m] = rh[m - 1] / di[m - 1];
    for (int r = m - 2; r >= 0; r--) ypp[r + 1] = (rh[r] - up[r] * ypp[r + 2]) / di[r];
    {
        double h0 = t[1] - t[0], h1 = t[2] - t[1];
        ypp[0] = ((h0 + h1) * ypp[1] - h0 * ypp[2]) / h1;
        double ha = t[n - 2] - t[n - 3], hb = t[n - 1] - t[n - 2];
        ypp[n - 1] = ((ha + hb) * ypp[n - 2] - hb * ypp[n - 3]) / ha;
    }
    free(lo); free(di); free(up); free(rh);
    return 0;
}

/* Value of the spline at tv (interval found by bisection; extrapolates outside). */
double orc_spline_val(int n, const double *t, const double *y, const double *ypp, double tv)
{
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) { int mid = (lo + hi) / 2; if (tv >= t[mid]) lo = mid; else hi = mid; }
    double h = t[hi] - t[lo], a = tv - t[lo], b = t[hi] - tv;
    return (ypp[lo] * b * b * b + ypp[hi] * a * a * a) / (6.0 * h)
         + (y[lo] / h - ypp[lo] * h / 6.0) * b + (y[hi] / h - ypp[hi] * h / 6.0) * a;
}

/* Stage-time field tables of the reference (knot mode):
 *   forward step i (t_i -> t_{i+1}):  Ef[3i..3i+2] = (E_i, E(t_{i+1/2}), E_{i+1})   G:1692-1715
 *   adjoint step over [t_{i-1}, t_i] (stored at row i-1):
 *                        Ea[3(i-1)..] = (E_{i-1}, E_i, E(t_{i-1} + dt/3))           G:1874-1879 */
int orc_knot_tables(double dt, int nt, const double *Et, double *Ef, double *Ea)
{
    int n = nt + 1;
    double *t = malloc(sizeof(double) * n), *ypp = malloc(sizeof(double) * n);
    if (!t || !ypp) { free(t); free(ypp); return -2; }
    for (int i = 0; i < n; i++) t[i] = i * dt;
    int rc = orc_spline_set(n, t, Et, ypp);
    if (rc == 0) {
        for (int i = 0; i < nt; i++) {
            Ef[3 * i] = Et[i];
            Ef[3 * i + 1] = orc_spline_val(n, t, Et, ypp, (i + 0.5) * dt);
            Ef[3 * i + 2] = Et[i + 1];
        }
        for (int i = 1; i <= nt; i++) {
            Ea[3 * (i - 1)] = Et[i - 1];
            Ea[3 * (i - 1) + 1] = Et[i];
            Ea[3 * (i - 1) + 2] = orc_spline_val(n, t, Et, ypp, (i - 1 + 1.0 / 3.0) * dt);
        }
    }
    free(t); free(ypp);
    return rc;
}

/* ------------------------------------------------------ one RK4 step (G:460-563) */
typedef struct { double *U, *Ut, *du; } fwd_ws;

static void rk4_step(const orc_model *m, int nx, int ny, double dt, const double *gh, const double *E3,
                     double *y, fwd_ws *ws)
{
    size_t n = (size_t)nx * ny;
    double *U = ws->U, *Ut = ws->Ut, *du = ws->du;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double f[3]; size_t k = (size_t)i * ny + j;
            rhs_cell(m, nx, ny, y, gh, E3[0], i, j, f);
            for (int s = 0; s < 3; s++) {
                U[k + s * n] = y[k + s * n] + dt * 0.5 * f[s];
                du[k + s * n] = 1.0 / 6.0 * dt * f[s];
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double f[3]; size_t k = (size_t)i * ny + j;
            rhs_cell(m, nx, ny, U, gh, E3[1], i, j, f);
            for (int s = 0; s < 3; s++) {
                Ut[k + s * n] = y[k + s * n] + dt * 0.5 * f[s];
                du[k + s * n] = du[k + s * n] + 1.0 / 3.0 * dt * f[s];
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double f[3]; size_t k = (size_t)i * ny + j;
            rhs_cell(m, nx, ny, Ut, gh, E3[1], i, j, f);
            for (int s = 0; s < 3; s++) {
                U[k + s * n] = y[k + s * n] + dt * 1.0 * f[s];
                du[k + s * n] = du[k + s * n] + 1.0 / 3.0 * dt * f[s];
            }
        }
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; i++)
        for (int j = 0; j < ny; j++) {
            double f[3]; size_t k = (size_t)i * ny + j;
            rhs_cell(m, nx, ny, U, gh, E3[2], i, j, f);
            for (int s = 0; s < 3; s++)
                y[k + s * n] = y[k + s * n] + du[k + s * n] + dt * 1.0 / 6.0 * f[s];
        }
}

/* Forward solve of nt steps in place on y (3 NaN-padded planes); Ef as in orc_knot_tables. */
int orc_forward(const orc_model *m, int nx, int ny, double dt, int nt, const double *Ef, const double *gh,
                double *y)
{
    size_t n3 = (size_t)3 * nx * ny;
    fwd_ws ws;
    ws.U = malloc(sizeof(double) * n3); ws.Ut = malloc(sizeof(double) * n3); ws.du = malloc(sizeof(double) * n3);
    if (!ws.U || !ws.Ut || !ws.du) { free(ws.U); free(ws.Ut); free(ws.du); return -2; }
    for (int i = 0; i < nt; i++) rk4_step(m, nx, ny, dt, gh, Ef + 3 * i, y, &ws);
    free(ws.U); free(ws.Ut); free(ws.du);
    return 0;
}

/* ---------------------------------------------- Hermite midpoint (G:1217-1301) */
static double herm3(double x0, double x1, double m0, double m1)
{
    /* cubic Hermite basis at s = 1/3 */
    const double s = 1.0 / 3.0;
    return (s - 1.0) * (s - 1.0) * (2.0 * s + 1.0) * x0 + s * (s - 1.0) * (s - 1.0) * m0
         + s * s * (3.0 - 2.0 * s) * x1 + s * s * (s - 1.0) * m1;
}

static double herm4_part(double x0, double x1, double m0, double m1)
{
    /* quartic through (x0, m0) at 0, (x1, m1) at 1 and a slope at 1/3, evaluated at s = 1/2,
     * without the slope-at-1/3 term (added by the caller with weight 27 s^2 (s-1)^2 / 4). */
    const double s = 0.5;
    return (1.0 + 2.0 * s + 9.0 * s * s) * (s - 1.0) * (s - 1.0) * x0 + s * (s - 1.0) * (s - 1.0) * m0
         - s * s * (6.0 - 16.0 * s + 9.0 * s * s) * x1 + (s - 1.0) * (9.0 * s - 5.0) * s * s / 4.0 * m1;
}

/* --------------------------------------------- gradient evaluation (G:1569-1938) */
/* Inputs: NaN-padded y0 (3 planes), gh (NaN at holes), tables Ef/Ea (3*nt each), packed-index
 * CSC matrix S (n_num x n_num) with active_index mapping, gamma[3].
 * Outputs: zgh[nt+1], L0 (= sum_v gamma_v y_v^T S y_v at T), yT (final state, 3 planes, may be NULL).
 * ckpt: forward checkpoint interval (>= 1). */
/* State values to the nearest FP32 (NaN holes stay NaN): the precision probe of orc_grad. */
static void round_to_f32(double *p, size_t n)
{
    for (size_t k = 0; k < n; k++) p[k] = (double)(float)p[k];
}

int orc_grad(const orc_model *m, int nx, int ny, double dt, int nt, const double *Ef, const double *Ea,
             const double *gh, const double *y0, const int32_t *active_index, int n_num,
             const int64_t *S_colptr, const int64_t *S_rowind, const double *S_val, const double *gamma,
             int ckpt, double *zgh, double *L0_out, double *yT, int round_state)
{
    size_t n = (size_t)nx * ny, n3 = 3 * n;
    if (ckpt < 1) ckpt = 1;
    int nck = (nt + ckpt - 1) / ckpt + 1;
    double *ck = malloc(sizeof(double) * n3 * nck);
    double *seg = malloc(sizeof(double) * n3 * (ckpt + 1));
    double *y = malloc(sizeof(double) * n3);
    fwd_ws ws;
    ws.U = malloc(sizeof(double) * n3); ws.Ut = malloc(sizeof(double) * n3); ws.du = malloc(sizeof(double) * n3);
    double *z = malloc(sizeof(double) * n3), *Z = malloc(sizeof(double) * n3), *Zt = malloc(sizeof(double) * n3);
    double *dz = malloc(sizeof(double) * n3), *y3 = malloc(sizeof(double) * n3), *y4 = malloc(sizeof(double) * n3);
    double *yh = malloc(sizeof(double) * n3);
    double *un = malloc(sizeof(double) * (size_t)n_num), *zn = malloc(sizeof(double) * (size_t)n_num);
    int rc = 0;
    if (!ck || !seg || !y || !ws.U || !ws.Ut || !ws.du || !z || !Z || !Zt || !dz || !y3 || !y4 || !yh || !un || !zn) {
        rc = -2; goto done;
    }

    /* forward with checkpoints */
    memcpy(y, y0, sizeof(double) * n3);
    if (round_state) round_to_f32(y, n3);
    memcpy(ck, y, sizeof(double) * n3);
    for (int i = 0; i < nt; i++) {
        rk4_step(m, nx, ny, dt, gh, Ef + 3 * i, y, &ws);
        if (round_state) round_to_f32(y, n3);
        if ((i + 1) % ckpt == 0) memcpy(ck + n3 * ((i + 1) / ckpt), y, sizeof(double) * n3);
    }
    if (yT) memcpy(yT, y, sizeof(double) * n3);

    /* objective and terminal adjoint (G:1741-1838) */
    double L0 = 0.0;
    for (size_t v = 0; v < 3; v++) {
        for (size_t k = 0; k < n; k++) if (active_index[k] > 0) un[active_index[k] - 1] = y[v * n + k];
        for (int r = 0; r < n_num; r++) zn[r] = 0.0;
        for (int c = 0; c < n_num; c++)
            for (int64_t q = S_colptr[c]; q < S_colptr[c + 1]; q++) zn[S_rowind[q]] += S_val[q] * un[c];
        double g = -gamma[v];
        for (int r = 0; r < n_num; r++) zn[r] *= g;
        double d = 0.0;
        for (int r = 0; r < n_num; r++) d += un[r] * zn[r];
        L0 -= d;
        for (size_t k = 0; k < n; k++) z[v * n + k] = active_index[k] > 0 ? zn[active_index[k] - 1] : NAN;
    }
    *L0_out = L0;
    {
        double s = 0.0;
        for (size_t k = 0; k < n; k++) { double t = z[k] * gh[k]; if (!isnan(t)) s += t; }
        zgh[nt] = s;
    }

    const double c27 = 27.0 * 0.5 * 0.5 * (0.5 - 1.0) * (0.5 - 1.0) / 4.0;
    /* backward, one checkpoint segment at a time */
    for (int sgi = (nt - 1) / ckpt; sgi >= 0; sgi--) {
        int a = sgi * ckpt, b = a + ckpt < nt ? a + ckpt : nt;
        memcpy(seg, ck + n3 * sgi, sizeof(double) * n3);
        for (int i = a; i < b; i++) {
            memcpy(seg + n3 * (i - a + 1), seg + n3 * (i - a), sizeof(double) * n3);
            rk4_step(m, nx, ny, dt, gh, Ef + 3 * i, seg + n3 * (i - a + 1), &ws);
            if (round_state) round_to_f32(seg + n3 * (i - a + 1), n3);
        }
        for (int i = b; i > a; i--) {
            const double *yA = seg + n3 * (i - 1 - a);   /* y(t_{i-1}) */
            const double *yB = seg + n3 * (i - a);       /* y(t_i)     */
            const double *E = Ea + 3 * (i - 1);
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c;
                    double f0[3], f1[3];
                    rhs_cell(m, nx, ny, yA, gh, E[0], r, c, f0);
                    rhs_cell(m, nx, ny, yB, gh, E[1], r, c, f1);
                    for (int s = 0; s < 3; s++) {
                        double m0 = f0[s] * dt, m1 = f1[s] * dt;
                        y3[k + s * n] = herm3(yA[k + s * n], yB[k + s * n], m0, m1);
                        y4[k + s * n] = herm4_part(yA[k + s * n], yB[k + s * n], m0, m1);
                    }
                }
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c;
                    double f2[3];
                    rhs_cell(m, nx, ny, y3, gh, E[2], r, c, f2);
                    for (int s = 0; s < 3; s++) yh[k + s * n] = y4[k + s * n] + c27 * (f2[s] * dt);
                }
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c; double f[3];
                    adjf_cell(m, nx, ny, z, yB, r, c, f);
                    for (int s = 0; s < 3; s++) {
                        Z[k + s * n] = z[k + s * n] + 0.5 * dt * f[s];
                        dz[k + s * n] = 1.0 / 6.0 * dt * f[s];
                    }
                }
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c; double f[3];
                    adjf_cell(m, nx, ny, Z, yh, r, c, f);
                    for (int s = 0; s < 3; s++) {
                        Zt[k + s * n] = z[k + s * n] + 0.5 * dt * f[s];
                        dz[k + s * n] = dz[k + s * n] + 2.0 / 6.0 * dt * f[s];
                    }
                }
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c; double f[3];
                    adjf_cell(m, nx, ny, Zt, yh, r, c, f);
                    for (int s = 0; s < 3; s++) {
                        Z[k + s * n] = z[k + s * n] + dt * f[s];
                        dz[k + s * n] = dz[k + s * n] + 2.0 / 6.0 * dt * f[s];
                    }
                }
#pragma omp parallel for schedule(static)
            for (int r = 0; r < nx; r++)
                for (int c = 0; c < ny; c++) {
                    size_t k = (size_t)r * ny + c; double f[3];
                    adjf_cell(m, nx, ny, Z, yA, r, c, f);
                    for (int s = 0; s < 3; s++) Zt[k + s * n] = z[k + s * n] + dz[k + s * n] + 1.0 / 6.0 * dt * f[s];
                }
            double *tmp = z; z = Zt; Zt = tmp;
            double s = 0.0;
            for (size_t k = 0; k < n; k++) { double t = z[k] * gh[k]; if (!isnan(t)) s += t; }
            zgh[i - 1] = s;
        }
    }
done:
    free(ck); free(seg); free(y); free(ws.U); free(ws.Ut); free(ws.du);
    free(z); free(Z); free(Zt); free(dz); free(y3); free(y4); free(yh); free(un); free(zn);
    return rc;
}

/* Packed [u;v;w] (3*n_num) -> NaN-padded planes (G:1533-1542). */
void orc_num2nan(const double *packed, const int32_t *active_index, int n, int n_num, double *planes)
{
    for (int v = 0; v < 3; v++)
        for (int k = 0; k < n; k++) {
            int a = active_index[k];
            planes[(size_t)v * n + k] = a > 0 ? packed[(size_t)v * n_num + a - 1] : NAN;
        }
}

/* Mirror of fk_grad_Et_comp (G:1569-1938) in knot mode: returns gradE[nt+1], L0, LE, Lcost. */
int orc_grad_ref(double dt, const double *Et, int nt, const double *y_ini, const int32_t *active_index,
                 int nrows, int ncols, int n_num, const double *gh_num, const double *gamma,
                 const int64_t *S_colptr, const int64_t *S_rowind, const double *S_val, double alpha,
                 const orc_model *model, int ckpt, double *L0, double *LE, double *Lcost, double *gradE)
{
    orc_model mref;
    if (!model) { orc_model_reference(&mref); model = &mref; }
    int n = nrows * ncols;
    double *y = malloc(sizeof(double) * 3 * (size_t)n), *gh = malloc(sizeof(double) * (size_t)n);
    double *Ef = malloc(sizeof(double) * 3 * (size_t)nt), *Ea = malloc(sizeof(double) * 3 * (size_t)nt);
    double *zgh = malloc(sizeof(double) * (size_t)(nt + 1));
    int rc = 0;
    if (!y || !gh || !Ef || !Ea || !zgh) { rc = -2; goto out; }
    orc_num2nan(y_ini, active_index, n, n_num, y);
    for (int k = 0; k < n; k++) {
        int a = active_index[k];
        gh[k] = a > 0 ? model->Du * 2 / model->h * gh_num[a - 1] : NAN;
    }
    rc = orc_knot_tables(dt, nt, Et, Ef, Ea);
    if (rc) goto out;
    rc = orc_grad(model, nrows, ncols, dt, nt, Ef, Ea, gh, y, active_index, n_num, S_colptr, S_rowind, S_val,
                  gamma, ckpt, zgh, L0, NULL, 0);
    if (rc) goto out;
    double e2 = 0.0;
    for (int i = 0; i <= nt; i++) e2 += Et[i] * Et[i];
    *LE = dt * e2;
    *Lcost = 0.5 * (*L0) + 0.5 * alpha * (*LE);
    for (int i = 0; i <= nt; i++) gradE[i] = alpha * Et[i] - zgh[i];
out:
    free(y); free(gh); free(Ef); free(Ea); free(zgh);
    return rc;
}
