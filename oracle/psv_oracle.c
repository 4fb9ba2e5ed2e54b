/*
 * psv_oracle.c -- CPU restatement of the reference P-SV FWI hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Only tests/, the smoke check and
 * the benchmark's CPU-baseline leg load this library; the product path never
 * does.  Build: oracle/Makefile (gcc, -ffp-contract=off so that no fused
 * multiply-add changes the reference's rounding).
 */
#include "psv_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>

#define PSVO_CAT_(a, b) a##b
#define PSVO_CAT(a, b) PSVO_CAT_(a, b)

/* PI as the reference defines it (CUDA_week2/globvar.cuh:18). */
#define PSVO_PI 3.14159265

/* Argument validation (the reference has none; an out-of-range receiver or
 * window there is undefined behaviour, see CUDA_week2/fd_fwi.cu:37-51). */
static int psvo_check_seis(const psvo_cfg* c, const psvo_seis* s)
{
    if (!s || s->n <= 0) return 0;
    if (!s->x || !s->z || !s->comp || !s->shot) return -20;
    for (int i = 0; i < s->n; i++) {
        if (s->x[i] < 0 || s->x[i] >= c->nxt || s->z[i] < 0 || s->z[i] >= c->nzt) return -21;
    }
    return 0;
}

/* Offset of the bottom CPML strip into the z-type profiles and memory
 * arrays.  CUDA_week2/forward_kernel_PSV_GPU.cu:750-755 uses
 * nzt - 2 ppad - 1 with a free surface, which only fits the 2(npml+1)-row
 * arrays for fpad == 1 (the reference only runs fdorder == 2); the form
 * nzt - 2 ppad + fpad - 2 equals it at fpad == 1 and stays in range at
 * fpad == 2. */
static int psvo_pnz(const psvo_cfg* c)
{
    if (c->fsurf) return c->nzt - 2 * c->ppad + c->fpad - 2;
    return c->nzt - 2 * c->ppad + c->fpad - 1;
}

static int psvo_check_cfg(const psvo_cfg* c, const psvo_seis* src, const psvo_seis* rec)
{
    if (!c) return -1;
    if (c->fdorder != 2 && c->fdorder != 4) return -2;
    if (c->fpad != c->fdorder / 2) return -3;
    if (c->npml < 0 || c->ppad != c->npml + c->fpad) return -4;
    if (c->nt < 1 || c->nzt < 2 * c->fpad + 1 || c->nxt < 2 * c->fpad + 1) return -5;
    if (c->npml > 0) {
        /* every strip index must land inside the 2(npml+1) profile entries */
        const int tp = 2 * (c->npml + 1);
        const int pnx = c->nxt - 2 * c->ppad + c->fpad - 1;
        const int pnz = psvo_pnz(c);
        if (c->nxt - c->fpad - 1 - pnx >= tp || c->nxt - c->ppad - 1 - pnx < 0) return -8;
        if (c->nzt - c->fpad - 1 - pnz >= tp || c->nzt - c->ppad - 1 - pnz < 0) return -8;
        if (c->nxt < 2 * c->ppad || c->nzt < c->ppad + c->fpad) return -8;
    }
    if (c->fwinv) {
        if (c->fwi_dt < 1 || c->fwi_dx < 1 || c->fwi_dz < 1) return -6;
        if (c->fwi_x1 < 0 || c->fwi_z1 < 0 || c->fwi_x2 > c->nxt || c->fwi_z2 > c->nzt ||
            c->fwi_x2 < c->fwi_x1 || c->fwi_z2 < c->fwi_z1) return -7;
    }
    int st = psvo_check_seis(c, src);
    if (st) return st;
    return psvo_check_seis(c, rec);
}

/* ------------------------------------------------------------------ */
/* Holberg coefficient table, rows = maximum relative group-velocity
 * error class 0..4 (0 = Taylor), columns = FD order 2..12; entry 0 of
 * each row is the points-per-wavelength figure.  The numbers are data
 * the scheme is defined by (CUDA_week2/holbergcoeff.cu:10-52). */
static const double psvo_hc_table[5][6][7] = {
    {
        {23.0, 1.0, 0.0, 0.0, 0.0, 0.0, 0.0},
        {8.0, 9.0 / 8.0, -1.0 / 24.0, 0.0, 0.0, 0.0, 0.0},
        {6.0, 75.0 / 64.0, -25.0 / 384.0, 3.0 / 640.0, 0.0, 0.0, 0.0},
        {5.0, 1225.0 / 1024.0, -245.0 / 3072.0, 49.0 / 5120.0, -5.0 / 7168.0, 0.0, 0.0},
        {5.0, 19845.0 / 16384.0, -735.0 / 8192.0, 567.0 / 40960.0, -405.0 / 229376.0,
         35.0 / 294912.0, 0.0},
        {4.0, 160083.0 / 131072.0, -12705.0 / 131072.0, 22869.0 / 1310720.0,
         -5445.0 / 1835008.0, 847.0 / 2359296.0, -63.0 / 2883584.0},
    },
    {
        {49.7, 1.0010, 0.0, 0.0, 0.0, 0.0, 0.0},
        {8.32, 1.1382, -0.046414, 0.0, 0.0, 0.0, 0.0},
        {4.77, 1.1965, -0.078804, 0.0081781, 0.0, 0.0, 0.0},
        {3.69, 1.2257, -0.099537, 0.018063, -0.0026274, 0.0, 0.0},
        {3.19, 1.2415, -0.11231, 0.026191, -0.0064682, 0.001191, 0.0},
        {2.91, 1.2508, -0.12034, 0.032131, -0.010142, 0.0029857, -0.00066667},
    },
    {
        {22.2, 1.0050, 0.0, 0.0, 0.0, 0.0, 0.0},
        {5.65, 1.1534, -0.052806, 0.0, 0.0, 0.0, 0.0},
        {3.74, 1.2111, -0.088313, 0.011768, 0.0, 0.0, 0.0},
        {3.11, 1.2367, -0.10815, 0.023113, -0.0046905, 0.0, 0.0},
        {2.80, 1.2496, -0.11921, 0.031130, -0.0093272, 0.0025161, 0.0},
        {2.62, 1.2568, -0.12573, 0.036423, -0.013132, 0.0047484, -0.0015979},
    },
    {
        {15.8, 1.0100, 0.0, 0.0, 0.0, 0.0, 0.0},
        {4.80, 1.1640, -0.057991, 0.0, 0.0, 0.0, 0.0},
        {3.39, 1.2192, -0.094070, 0.014608, 0.0, 0.0, 0.0},
        {2.90, 1.2422, -0.11269, 0.026140, -0.0064054, 0.0, 0.0},
        {2.65, 1.2534, -0.12257, 0.033755, -0.011081, 0.0036784, 0.0},
        {2.51, 1.2596, -0.12825, 0.038550, -0.014763, 0.0058619, -0.0024538},
    },
    {
        {9.16, 1.0300, 0.0, 0.0, 0.0, 0.0, 0.0},
        {3.47, 1.1876, -0.072518, 0.0, 0.0, 0.0, 0.0},
        {2.91, 1.2341, -0.10569, 0.022589, 0.0, 0.0, 0.0},
        {2.61, 1.2516, -0.12085, 0.032236, -0.011459, 0.0, 0.0},
        {2.45, 1.2596, -0.12829, 0.038533, -0.014681, 0.0072580, 0.0},
        {2.36, 1.2640, -0.13239, 0.042217, -0.017803, 0.0081959, -0.0051848},
    },
};

/* CUDA_week2/holbergcoeff.cu:6-65 (hc sized 7 here; the reference writes
 * 7 values into a 6-element member, CUDA_week2/fd_wave_model.cuh:33). */
void psvo_holbergcoeff(int fdorder, int maxrelerr, double hc[7])
{
    const int nl = fdorder / 2 - 1;
    for (int i = 0; i < 7; i++) hc[i] = psvo_hc_table[maxrelerr][nl][i];
}

/* CPML profiles: CUDA_week2/fd_cpml.cu:59-192.  Vectors have length
 * 2(npml+1): indices 0..npml are the left/top side, npml+1..2npml+1 the
 * right/bottom side.  (The reference's initialisation loop also writes
 * index 2npml+2, one past its allocation; that element is never read.) */
void psvo_cpml_pro(int npml, double dh, double fpml, double npower, double k_max_pml,
                   double rcoef, double damp_v_pml, double dt,
                   double* a, double* b, double* K,
                   double* a_half, double* b_half, double* K_half)
{
    const double alpha_max = 2.0 * PSVO_PI * (fpml / 2.0);
    const double thickness = npml * dh;
    const double d0 = -(npower + 1) * damp_v_pml * log(rcoef) / (2.0 * thickness);
    const int tp = 2 * (npml + 1);

    for (int i = 0; i < tp; i++) {
        a[i] = 0.0; a_half[i] = 0.0;
        b[i] = 1.0; b_half[i] = 1.0;
        K[i] = 1.0; K_half[i] = 1.0;
    }
    for (int i = 0; i < tp; i++) {
        const int left = (i <= npml);
        double d = 0.0, d_half = 0.0, alpha = 0.0, alpha_half = 0.0, xn;
        const double xval = dh * (double)i;
        double absc;

        if (left) absc = thickness - xval;
        else absc = xval - dh * (npml + 1);
        if (absc >= 0) {
            xn = absc / thickness;
            d = d0 * pow(xn, npower);
            K[i] = 1.0 + (k_max_pml - 1.0) * pow(xn, npower);
            alpha = alpha_max * (1.0 - xn);
        }
        if (left) absc = thickness - (xval + dh / 2.0);
        else absc = (xval + dh / 2.0) - dh * (npml + 1);
        if (absc >= 0) {
            xn = absc / thickness;
            d_half = d0 * pow(xn, npower);
            K_half[i] = 1.0 + (k_max_pml - 1.0) * pow(xn, npower);
            alpha_half = alpha_max * (1.0 - xn);
        }
        if (alpha < 0.0) alpha = 0.0;
        if (alpha_half < 0.0) alpha_half = 0.0;

        b[i] = exp(-(d / K[i] + alpha) * dt);
        b_half[i] = exp(-(d_half / K_half[i] + alpha_half) * dt);
        if (fabs(d) > 1.0e-6) a[i] = d * (b[i] - 1.0) / (K[i] * (d + K[i] * alpha));
        if (fabs(d_half) > 1.0e-6)
            a_half[i] = d_half * (b_half[i] - 1.0) / (K_half[i] * (d_half + K_half[i] * alpha_half));
    }
}

/* CUDA_week2/fd_medium.cu:54-59 */
void psvo_velocity_to_lame(double vp, double vs, double rho, double* lam, double* mu)
{
    *mu = pow(vs, 2) * rho;
    *lam = pow(vp, 2) * rho - 2.0 * *mu;
}

/* CUDA_week2/fd_medium.cu:128-159.  dimz = nzt-1, dimx = nxt-1; mu and
 * rho are [dimz+1][dimx+1].  The air guard sets rho_zp twice, as the
 * reference does. */
void psvo_average_parameters(int dimz, int dimx, const double* mu, const double* rho,
                             double* mu_zx, double* rho_zp, double* rho_xp)
{
    const size_t s = (size_t)dimx + 1;
    for (int ix = 0; ix < dimx; ix++) {
        for (int iz = 0; iz < dimz; iz++) {
            const size_t p = (size_t)iz * s + ix;
            const size_t o = (size_t)iz * dimx + ix;
            mu_zx[o] = 4.0 / ((1.0 / mu[p]) + (1.0 / mu[p + 1]) + (1.0 / mu[p + s]) +
                              (1.0 / mu[p + s + 1]));
            if (mu[p] == 0.0 || mu[p + 1] == 0.0 || mu[p + s] == 0.0 || mu[p + s + 1] == 0.0)
                mu_zx[o] = 0.0;
            rho_zp[o] = 1.0 / (0.5 * (rho[p] + rho[p + s]));
            rho_xp[o] = 1.0 / (0.5 * (rho[p] + rho[p + 1]));
            if (rho[p] < 1e-4 && rho[p + s] < 1e-4) rho_zp[o] = 0.0;
            if (rho[p] < 1e-4 && rho[p + 1] < 1e-4) rho_zp[o] = 0.0;
        }
    }
}

/* Constant model with a lambda x 1.3 disc at the grid centre, integer
 * radius test: CUDA_week2/fd_medium.cu:77-110. */
void psvo_stagger_over_grid_modify(int dimz, int dimx, double scalar_rho,
                                   double scalar_lam, double scalar_mu,
                                   double* rho, double* lam, double* mu)
{
    for (int ix = 0; ix < dimx; ix++) {
        for (int iz = 0; iz < dimz; iz++) {
            const size_t p = (size_t)iz * dimx + ix;
            rho[p] = scalar_rho;
            lam[p] = scalar_lam;
            mu[p] = scalar_mu;
            const int dz = iz - dimz / 2, dx = ix - dimx / 2;
            if (dz * dz + dx * dx < dimz * dimz / 49) lam[p] = 1.3 * scalar_lam;
        }
    }
}

/* Ricker wavelet: CUDA_week2/fd_seismogram.cu:166-186 (shape 1 only). */
void psvo_create_wavelet(double* signal, int nt, double dt, double amp, double fc,
                         double ts, int shape)
{
    const double fci = 1.0 / fc;
    if (shape != 1) return;
    for (int it = 0; it < nt; it++) {
        const double t = it * dt;
        const double tau = PSVO_PI * (t - 1.5 * fci - ts) / (1.5 * fci);
        signal[it] = amp * (1.0 - 2.0 * tau * tau) * exp(-2.0 * tau * tau);
    }
}

/* CUDA_week2/fd_wave_model.cu:972-1002: misfit and in-place residual. */
int psvo_l2_adjoint_sources(int nr, int nt, double dt, double* signal,
                            const double* signal_meas, double* misfit)
{
    double m = 0.0;
    for (int ir = 0; ir < nr; ir++) {
        for (int it = 0; it < nt; it++) {
            const size_t p = (size_t)ir * nt + it;
            const double diff = signal[p] - signal_meas[p];
            m += 0.5 * dt * diff * diff;
            signal[p] = diff;
            if (fabs(diff) > 1.0e+10) { *misfit = m; return 1; }
        }
    }
    *misfit = m;
    return 0;
}

/* CUDA_week2/fd_wave_model.cu:1004-1052.  The loops run over full-grid
 * indices [z1, nfz) x [x1, nfx) with stride (fdz, fdx) and index the
 * [nfz][nfx] gradients and the [nzt][nxt] energies with the same (iz, ix),
 * exactly as the reference does. */
void psvo_scale_gradients_energy(const double* Ws, const double* Wr, int nxt,
                                 double C_rho, double C_lam, double C_mu,
                                 int nfz, int nfx, int z1, int x1, int fdz, int fdx,
                                 double* grad_lam, double* grad_mu, double* grad_rho)
{
    const double eps = 0.005;
    double wmax = 0.0;
    double* W = (double*)calloc((size_t)nfz * nfx, sizeof(double));
    if (!W) return;
    for (int iz = z1; iz < nfz; iz += fdz)
        for (int ix = x1; ix < nfx; ix += fdx) {
            const size_t g = (size_t)iz * nfx + ix, p = (size_t)iz * nxt + ix;
            W[g] = sqrt(Ws[p] * Wr[p]);
            if (W[g] > wmax) wmax = W[g];
        }
    for (int iz = z1; iz < nfz; iz += fdz)
        for (int ix = x1; ix < nfx; ix += fdx) W[(size_t)iz * nfx + ix] += eps * wmax;
    for (int iz = z1; iz < nfz; iz += fdz)
        for (int ix = x1; ix < nfx; ix += fdx) {
            const size_t g = (size_t)iz * nfx + ix;
            grad_lam[g] = grad_lam[g] / (W[g] * C_lam * C_lam);
            grad_mu[g] = grad_mu[g] / (W[g] * C_mu * C_mu);
            grad_rho[g] = grad_rho[g] / (W[g] * C_rho * C_rho);
        }
    free(W);
}

/* CUDA_week2/fd_medium.cu:253-277 */
void psvo_medium_average(const double* lam, const double* mu, const double* rho,
                         int dimz, int dimx, double* C_lam, double* C_mu, double* C_rho)
{
    double l = 0.0, m = 0.0, r = 0.0;
    for (int iz = 0; iz < dimz; iz++)
        for (int ix = 0; ix < dimx; ix++) {
            const size_t p = (size_t)iz * dimx + ix;
            l += lam[p]; m += mu[p]; r += rho[p];
        }
    *C_lam = l / (dimz * dimx);
    *C_mu = m / (dimz * dimx);
    *C_rho = r / (dimz * dimx);
}

/* CUDA_week2/fd_medium.cu:161-240, including the signed grad_max quirk
 * (:194) and the full-grid/decimated index mix.  mat, mat_new are
 * [nzt][nxt] (row stride nxt), grad is [nfz][nfx] (row stride nfx). */
void psvo_fwi_grad_update(double* mat, double* mat_new, const double* grad, int nxt,
                          double step_length, int dimz, int dimx, int z1, int x1,
                          int fwi_dz, int fwi_dx, int override_old, int mat_type, int nfx)
{
    double mat_max = 0.0, grad_max = 0.0, hi = 0.0, lo = 0.0;
    if (mat_type == 1) { hi = 3000; lo = 1.5; }
    else if (mat_type == 2) { hi = 48.0e+9; lo = 0; }
    else if (mat_type == 3) { hi = 27.0e+9; lo = 0; }
    for (int iz = z1; iz < dimz; iz += fwi_dz)
        for (int ix = x1; ix < dimx; ix += fwi_dx) {
            const size_t p = (size_t)iz * nxt + ix, g = (size_t)iz * nfx + ix;
            if (mat[p] > mat_max) mat_max = mat[p];
            if (fabs(grad[g]) > grad_max) grad_max = grad[g];
        }
    const double sn = step_length * (mat_max / grad_max);
    for (int iz = z1; iz < dimz; iz += fwi_dz)
        for (int ix = x1; ix < dimx; ix += fwi_dx) {
            const size_t p = (size_t)iz * nxt + ix, g = (size_t)iz * nfx + ix;
            mat_new[p] = mat[p] + sn * grad[g];
            if (mat_new[p] < lo) mat_new[p] = lo;
            if (mat_new[p] > hi) mat_new[p] = hi;
            if (mat_new[p] < 0.0) mat_new[p] = mat[p];
        }
    if (override_old)
        for (int iz = z1; iz < dimz; iz += fwi_dz)
            for (int ix = x1; ix < dimx; ix += fwi_dx) {
                const size_t p = (size_t)iz * nxt + ix;
                mat[p] = mat_new[p];
            }
}

/* Vertex of the parabola through (sl[i], L2[i]), i = 0..2: the quantity
 * calc_opt_step computes (CUDA_week2/fwi_step_length_estimation.cu:226-271).
 * The reference's 3x3 Gauss solve (CUDA_week2/solvelin.cu:6-41) reads and
 * writes past its arrays during back substitution; here the fit is solved
 * exactly with Newton divided differences in FP64. */
double psvo_calc_opt_step(const double L2[3], const double sl[3])
{
    const double d01 = (L2[1] - L2[0]) / (sl[1] - sl[0]);
    const double d02 = (L2[2] - L2[0]) / (sl[2] - sl[0]);
    const double x2 = (d02 - d01) / (sl[2] - sl[1]);
    const double x1 = d01 - x2 * (sl[0] + sl[1]);
    return -x1 / (2.0 * x2);
}

/* Chain rule from the coefficient arrays to lam, mu, rho (not in the
 * reference): mu_zx and rho_zp / rho_xp are the averages of
 * CUDA_week2/fd_medium.cu:128-159 (psvo_average_parameters), so
 *   d mu_zx[o] / d mu[q]  = mu_zx[o]^2 / (4 mu[q]^2) for its 4 corners q,
 *   d rho_zp[o] / d rho[q] = -rho_zp[o]^2 / 2 for (iz, ix) and (iz + 1, ix),
 *   d rho_xp[o] / d rho[q] = -rho_xp[o]^2 / 2 for (iz, ix) and (iz, ix + 1);
 * an average the guards set to zero has zero derivative.  Contributions to
 * a cell are added in the order o = (iz, ix), (iz, ix - 1), (iz - 1, ix),
 * (iz - 1, ix - 1). */
void psvo_chain_gradient(int nzt, int nxt, const double* mu, const double* mu_zx,
                         const double* rho_zp, const double* rho_xp,
                         const double* g_lam, const double* g_mu, const double* g_mu_zx,
                         const double* g_rho_zp, const double* g_rho_xp,
                         double* out_lam, double* out_mu, double* out_rho)
{
    const int ax = nxt - 1;
    for (int iz = 0; iz < nzt; iz++) {
        for (int ix = 0; ix < nxt; ix++) {
            const size_t q = (size_t)iz * nxt + ix;
            double gm = g_mu[q], gr = 0.0;
            for (int k = 0; k < 4; k++) {
                const int oz = iz - (k >> 1), ox = ix - (k & 1);
                if (oz < 0 || ox < 0 || oz >= nzt - 1 || ox >= nxt - 1) continue;
                const size_t o = (size_t)oz * ax + ox;
                if (mu_zx[o] != 0.0) gm += g_mu_zx[o] * (mu_zx[o] * mu_zx[o] / (4.0 * mu[q] * mu[q]));
                if (ox == ix) gr += g_rho_zp[o] * (-0.5 * rho_zp[o] * rho_zp[o]);   /* rows iz, iz + 1 */
                if (oz == iz) gr += g_rho_xp[o] * (-0.5 * rho_xp[o] * rho_xp[o]);   /* columns ix, ix + 1 */
            }
            out_lam[q] = g_lam[q];
            out_mu[q] = gm;
            out_rho[q] = gr;
        }
    }
}

/* ------------------------------------------------------------------ */
#define REAL double
#define SUF _d
#include "psv_oracle_prop.inc"
#include "psv_oracle_discrete.inc"
#undef REAL
#undef SUF

#define REAL float
#define SUF _f
#include "psv_oracle_prop.inc"
#include "psv_oracle_discrete.inc"
#undef REAL
#undef SUF
