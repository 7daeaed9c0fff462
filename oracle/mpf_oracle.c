/*
 * mpf_oracle.c -- CPU oracle for the multi-phase-field time step.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path never calls it.
 *
 * A plain-C FP64 restatement of the reference kernels, function by function,
 * in the reference's own dense layout and arithmetic order:
 *   3D: src/gpu_kernels_3d.py   (x, y periodic; z mirror)
 *   2D: src/gpu_kernels.py      (x periodic; y mirror)
 * Layout: phi[g][l][m][k] (3D) / phi[g][l][m] (2D), C order, last index
 * fastest; mf[t][l][m][k], nf[l][m][k]; temp dense [l][m][k].
 *
 * FP64 contract ("pure FP64", faithful = 0): every value is an IEEE double
 * and every operation is evaluated in the order the reference writes it; the
 * reference's float32 casts (facet/A12/A13 returns, lap_sl/b_sl local
 * arrays, phi_new stores) are not applied.  With faithful = 1 (3D update) the
 * values are rounded to float32 at exactly those points, as the reference's
 * float32 arrays and float32-typed returns do (src/gpu_kernels_3d.py:257-261,
 * 296-300, 311, 337-340, 383, 553-554, 661, 671, 678); the arithmetic in
 * between stays FP64, as Numba types it.  With use_libm=1 the transcendentals are the C
 * library's (what Python's math module calls); with use_libm=0 they are the
 * shared routines of include/mpf_math.h, which the GPU FP64 solver also uses.
 * Build with -ffp-contract=off (oracle/Makefile).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "mpf_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int nx, ny, nz, N;
    double dx, dt, T_melt, Sf, eps0_sl, w0_sl;
    double a0, delta_a, mu_a, p_round, g2_floor, ksi, theta_c_rad;
    int use_libm;
    int faithful; /* 1: float32 rounding at the reference's truncation points (3D) */
} orc_params;

/* Round to float32 in faithful mode. */
static double f32p(const orc_params *p, double v) { return p->faithful ? (double)(float)v : v; }

typedef struct {
    const orc_params *p;
    const double *phi;
    const double *n111; /* [N][8][3] */
} ctx_t;

static double o_tan(const orc_params *p, double x) { return p->use_libm ? tan(x) : mpf_tan(x); }
static double o_acos(const orc_params *p, double x) { return p->use_libm ? acos(x) : mpf_acos(x); }
static double o_tanh(const orc_params *p, double x) { return p->use_libm ? tanh(x) : mpf_tanh(x); }

#define PHI3(c, g, l, m, k) \
    ((c)->phi[((((size_t)(g) * (c)->p->nx + (l)) * (c)->p->ny + (m)) * (c)->p->nz) + (k)])
#define PHI2(c, g, l, m) \
    ((c)->phi[(((size_t)(g) * (c)->p->nx + (l)) * (c)->p->ny) + (m)])
#define N111(c, s, t, d) ((c)->n111[((size_t)(s) * 8 + (t)) * 3 + (d)])

/* ---- anisotropy functions: src/gpu_kernels_3d.py:24-55 ------------------ */

/* calc_a_from_cos, src/gpu_kernels_3d.py:24-30 */
static double calc_a_from_cos(const orc_params *p, double cost)
{
    double c2 = cost * cost;
    double C = sqrt(c2 + p->p_round * p->p_round);
    double S = sqrt(mpf_pymax(1.0 - c2, 0.0) + p->p_round * p->p_round);
    return p->mu_a * (1.0 + p->delta_a * (C + o_tan(p, p->a0) * S));
}

/* calc_b_from_cos, src/gpu_kernels_3d.py:33-55 */
static double calc_b_from_cos(const orc_params *p, double best_cost)
{
    double c = best_cost;
    if (c > 1.0)
        c = 1.0;
    else if (c < -1.0)
        c = -1.0;
    double theta = o_acos(p, c);
    if (theta >= p->theta_c_rad) return 1.0;
    double x = theta / p->theta_c_rad;
    double eps = 1.0e-6;
    double y = 0.5 * M_PI * x;
    if (y > 0.5 * M_PI - eps) y = 0.5 * M_PI - eps;
    if (y < eps) y = eps;
    double t = o_tan(p, y);
    return p->ksi + (1.0 - p->ksi) * t * o_tanh(p, 1.0 / t);
}

/* ---- 3D boundary helpers: src/gpu_kernels_3d.py:58-91 ------------------- */
static int idx_xp(int l, int n) { return l < n - 1 ? l + 1 : 0; }
static int idx_xm(int l, int n) { return l > 0 ? l - 1 : n - 1; }
static int idx_zp(int k, int n) { return k < n - 1 ? k + 1 : n - 1; }
static int idx_zm(int k, int n) { return k > 0 ? k - 1 : 0; }

/* grad_phi_xyz, src/gpu_kernels_3d.py:94-107 */
static void grad3(const ctx_t *c, int gid, int l, int m, int k, double *gx, double *gy, double *gz)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_xp(m, p->ny), mm = idx_xm(m, p->ny);
    int kp = idx_zp(k, p->nz), km = idx_zm(k, p->nz);
    *gx = (PHI3(c, gid, lp, m, k) - PHI3(c, gid, lm, m, k)) / (2.0 * p->dx);
    *gy = (PHI3(c, gid, l, mp, k) - PHI3(c, gid, l, mm, k)) / (2.0 * p->dx);
    *gz = (PHI3(c, gid, l, m, kp) - PHI3(c, gid, l, m, km)) / (2.0 * p->dx);
}

/* best_cos_from_grad_3d, src/gpu_kernels_3d.py:110-134 */
static double best_cos3(const ctx_t *c, double gx, double gy, double gz, int s)
{
    double g2 = gx * gx + gy * gy + gz * gz;
    if (g2 < c->p->g2_floor) return 0.0;
    double inv_g = 1.0 / sqrt(g2);
    double nxn = gx * inv_g, nyn = gy * inv_g, nzn = gz * inv_g;
    double best = 0.0;
    for (int t = 0; t < 8; ++t) {
        double v = fabs(nxn * N111(c, s, t, 0) + nyn * N111(c, s, t, 1) + nzn * N111(c, s, t, 2));
        if (v > best) best = v;
    }
    if (best > 1.0) best = 1.0;
    return best;
}

/* eps2_at_cell_from_liquid_3d, src/gpu_kernels_3d.py:137-146 */
static double eps2_at3(const ctx_t *c, int l, int m, int k, int s)
{
    double gx, gy, gz;
    grad3(c, 0, l, m, k, &gx, &gy, &gz);
    double bc = best_cos3(c, gx, gy, gz, s);
    double a = calc_a_from_cos(c->p, bc);
    double e = c->p->eps0_sl * a;
    return e * e;
}

/* aniso_term1_solid_3d, src/gpu_kernels_3d.py:149-204 */
static double aniso_term1_3(const ctx_t *c, int l, int m, int k, int s)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_xp(m, p->ny), mm = idx_xm(m, p->ny);
    int kp = idx_zp(k, p->nz), km = idx_zm(k, p->nz);
    double e_c = eps2_at3(c, l, m, k, s);
    double e_xp = eps2_at3(c, lp, m, k, s);
    double e_xm = eps2_at3(c, lm, m, k, s);
    double e_yp = eps2_at3(c, l, mp, k, s);
    double e_ym = eps2_at3(c, l, mm, k, s);
    double e_zp = eps2_at3(c, l, m, kp, s);
    double e_zm = eps2_at3(c, l, m, km, s);
    double pc = PHI3(c, 0, l, m, k);
    double pxp = PHI3(c, 0, lp, m, k), pxm = PHI3(c, 0, lm, m, k);
    double pyp = PHI3(c, 0, l, mp, k), pym = PHI3(c, 0, l, mm, k);
    double pzp = PHI3(c, 0, l, m, kp), pzm = PHI3(c, 0, l, m, km);
    double Fx_p = (e_c + e_xp) * (pxp - pc) / (2.0 * p->dx);
    double Fx_m = (e_c + e_xm) * (pc - pxm) / (2.0 * p->dx);
    double Fy_p = (e_c + e_yp) * (pyp - pc) / (2.0 * p->dx);
    double Fy_m = (e_c + e_ym) * (pc - pym) / (2.0 * p->dx);
    double Fz_p = (e_c + e_zp) * (pzp - pc) / (2.0 * p->dx);
    double Fz_m = (e_c + e_zm) * (pc - pzm) / (2.0 * p->dx);
    return (Fx_p - Fx_m) / p->dx + (Fy_p - Fy_m) / p->dx + (Fz_p - Fz_m) / p->dx;
}

/* d2_phi_xyz, src/gpu_kernels_3d.py:207-248.  out: xx, yy, zz, xy, xz, yz */
static void d2_3(const ctx_t *c, int gid, int l, int m, int k, double *o)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_xp(m, p->ny), mm = idx_xm(m, p->ny);
    int kp = idx_zp(k, p->nz), km = idx_zm(k, p->nz);
    double cc = PHI3(c, gid, l, m, k);
    double xp = PHI3(c, gid, lp, m, k), xm = PHI3(c, gid, lm, m, k);
    double yp = PHI3(c, gid, l, mp, k), ym = PHI3(c, gid, l, mm, k);
    double zp = PHI3(c, gid, l, m, kp), zm = PHI3(c, gid, l, m, km);
    double inv_dx2 = 1.0 / (p->dx * p->dx);
    o[0] = (xp - 2.0 * cc + xm) * inv_dx2;
    o[1] = (yp - 2.0 * cc + ym) * inv_dx2;
    o[2] = (zp - 2.0 * cc + zm) * inv_dx2;
    double xpy = PHI3(c, gid, lp, mp, k), xpm = PHI3(c, gid, lp, mm, k);
    double xmy = PHI3(c, gid, lm, mp, k), xmm = PHI3(c, gid, lm, mm, k);
    o[3] = (xpy - xpm - xmy + xmm) * (0.25 * inv_dx2);
    double xpz = PHI3(c, gid, lp, m, kp), xpzm = PHI3(c, gid, lp, m, km);
    double xmz = PHI3(c, gid, lm, m, kp), xmzm = PHI3(c, gid, lm, m, km);
    o[4] = (xpz - xpzm - xmz + xmzm) * (0.25 * inv_dx2);
    double ypz = PHI3(c, gid, l, mp, kp), ypzm = PHI3(c, gid, l, mp, km);
    double ymz = PHI3(c, gid, l, mm, kp), ymzm = PHI3(c, gid, l, mm, km);
    o[5] = (ypz - ypzm - ymz + ymzm) * (0.25 * inv_dx2);
}

/* facet_cos_and_nxyz_from_grad, src/gpu_kernels_3d.py:251-301.
 * out: signed cos, |cos|, nx, ny, nz */
static void facet3(const ctx_t *c, double phix, double phiy, double phiz, int s, double *o)
{
    double q = phix * phix + phiy * phiy + phiz * phiz;
    if (q < c->p->g2_floor) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 1.0; o[3] = 0.0; o[4] = 0.0;
        return;
    }
    double inv_g = 1.0 / sqrt(q);
    double gx = phix * inv_g, gy = phiy * inv_g, gz = phiz * inv_g;
    const orc_params *p = c->p;
    double best_abs = 0.0, best_signed = 0.0, bnx = 1.0, bny = 0.0, bnz = 0.0;
    for (int t = 0; t < 8; ++t) {
        double ux = N111(c, s, t, 0), uy = N111(c, s, t, 1), uz = N111(c, s, t, 2);
        double dot = gx * ux + gy * uy + gz * uz;
        double ad = fabs(dot);
        if (ad > best_abs) {
            best_abs = ad; best_signed = dot; bnx = ux; bny = uy; bnz = uz;
        }
    }
    if (best_abs > 1.0) best_abs = 1.0;
    if (best_signed > 1.0) best_signed = 1.0;
    if (best_signed < -1.0) best_signed = -1.0;
    o[0] = f32p(p, best_signed); o[1] = f32p(p, best_abs);
    o[2] = f32p(p, bnx); o[3] = f32p(p, bny); o[4] = f32p(p, bnz);
}

/* da_dphixyz_A12, src/gpu_kernels_3d.py:304-341.  out: a, da/dphix, da/dphiy, da/dphiz */
static void A12_3(const ctx_t *c, int l, int m, int k, int s, double *o)
{
    const orc_params *p = c->p;
    double phix, phiy, phiz;
    grad3(c, 0, l, m, k, &phix, &phiy, &phiz);
    double q = phix * phix + phiy * phiy + phiz * phiz;
    if (q < p->g2_floor) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; o[3] = 0.0;
        return;
    }
    double f[5];
    facet3(c, phix, phiy, phiz, s, f);
    double a_val = calc_a_from_cos(p, f[1]);
    double inv_g = 1.0 / sqrt(q);
    double inv_g3 = 1.0 / (q * sqrt(q));
    double udotg = phix * f[2] + phiy * f[3] + phiz * f[4];
    double dcx = f[2] * inv_g - udotg * phix * inv_g3;
    double dcy = f[3] * inv_g - udotg * phiy * inv_g3;
    double dcz = f[4] * inv_g - udotg * phiz * inv_g3;
    double cs = f[0];
    double C = sqrt(cs * cs + p->p_round * p->p_round);
    double S = sqrt(mpf_pymax(1.0 - cs * cs, 0.0) + p->p_round * p->p_round);
    double coef = (cs / C) - o_tan(p, p->a0) * (cs / S);
    o[0] = f32p(p, a_val);
    o[1] = f32p(p, p->mu_a * p->delta_a * coef * dcx);
    o[2] = f32p(p, p->mu_a * p->delta_a * coef * dcy);
    o[3] = f32p(p, p->mu_a * p->delta_a * coef * dcz);
}

/* A13_divergence_3d, src/gpu_kernels_3d.py:344-383 */
static void A13_3(const ctx_t *c, int l, int m, int k, int s, double *o)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_xp(m, p->ny), mm = idx_xm(m, p->ny);
    int kp = idx_zp(k, p->nz), km = idx_zm(k, p->nz);
    double a[4], b[4];
    A12_3(c, lp, m, k, s, a); A12_3(c, lm, m, k, s, b);
    o[0] = f32p(p, (a[1] - b[1]) * (0.5 / p->dx));
    A12_3(c, l, mp, k, s, a); A12_3(c, l, mm, k, s, b);
    o[1] = f32p(p, (a[2] - b[2]) * (0.5 / p->dx));
    A12_3(c, l, m, kp, s, a); A12_3(c, l, m, km, s, b);
    o[2] = f32p(p, (a[3] - b[3]) * (0.5 / p->dx));
}

/* torque_A11_3d, src/gpu_kernels_3d.py:386-441 */
static double torque3(const ctx_t *c, int l, int m, int k, int s)
{
    const orc_params *p = c->p;
    double phix, phiy, phiz;
    grad3(c, 0, l, m, k, &phix, &phiy, &phiz);
    double q = phix * phix + phiy * phiy + phiz * phiz;
    if (q < p->g2_floor) return 0.0;
    double d2[6];
    d2_3(c, 0, l, m, k, d2);
    double phixx = d2[0], phiyy = d2[1], phizz = d2[2], phixy = d2[3], phixz = d2[4], phiyz = d2[5];
    double q_x = 2.0 * (phix * phixx + phiy * phixy + phiz * phixz);
    double q_y = 2.0 * (phix * phixy + phiy * phiyy + phiz * phiyz);
    double q_z = 2.0 * (phix * phixz + phiy * phiyz + phiz * phizz);
    double a12[4];
    A12_3(c, l, m, k, s, a12);
    double f[5];
    facet3(c, phix, phiy, phiz, s, f);
    double inv_g = 1.0 / sqrt(q);
    double inv_g3 = 1.0 / (q * sqrt(q));
    double udotg = phix * f[2] + phiy * f[3] + phiz * f[4];
    double dcx = f[2] * inv_g - udotg * phix * inv_g3;
    double dcy = f[3] * inv_g - udotg * phiy * inv_g3;
    double dcz = f[4] * inv_g - udotg * phiz * inv_g3;
    double c_x = dcx * phixx + dcy * phixy + dcz * phixz;
    double c_y = dcx * phixy + dcy * phiyy + dcz * phiyz;
    double c_z = dcx * phixz + dcy * phiyz + dcz * phizz;
    double cs = f[0];
    double C = sqrt(cs * cs + p->p_round * p->p_round);
    double S = sqrt(mpf_pymax(1.0 - cs * cs, 0.0) + p->p_round * p->p_round);
    double coef = (cs / C) - o_tan(p, p->a0) * (cs / S);
    double da_dx = p->mu_a * p->delta_a * coef * c_x;
    double da_dy = p->mu_a * p->delta_a * coef * c_y;
    double da_dz = p->mu_a * p->delta_a * coef * c_z;
    double a13[3];
    A13_3(c, l, m, k, s, a13);
    double a_val = a12[0];
    double Ex = (da_dx * a12[1] * q) + (a_val * a13[0] * q) + (a_val * a12[1] * q_x);
    double Ey = (da_dy * a12[2] * q) + (a_val * a13[1] * q) + (a_val * a12[2] * q_y);
    double Ez = (da_dz * a12[3] * q) + (a_val * a13[2] * q) + (a_val * a12[3] * q_z);
    return (p->eps0_sl * p->eps0_sl) * (Ex + Ey + Ez);
}

/* laplacian_3d, src/gpu_kernels_3d.py:444-459 (also the inline form :620-625) */
static double lap3(const ctx_t *c, int gid, int l, int m, int k, double inv_dx2)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_xp(m, p->ny), mm = idx_xm(m, p->ny);
    int kp = idx_zp(k, p->nz), km = idx_zm(k, p->nz);
    return (PHI3(c, gid, lp, m, k) + PHI3(c, gid, lm, m, k) + PHI3(c, gid, l, mp, k) +
            PHI3(c, gid, l, mm, k) + PHI3(c, gid, l, m, kp) + PHI3(c, gid, l, m, km) -
            6.0 * PHI3(c, gid, l, m, k)) * inv_dx2;
}

/* ---- 3D kernels ---------------------------------------------------------- */

/* kernel_update_nfmf_3d_checked, src/gpu_kernels_3d.py:489-527.
 * mf has shape [depth][nx][ny][nz]; writes only the first min(n, kmax) ids. */
void orc_apt3(const double *phi, int32_t *mf, int32_t *nf, int32_t *status,
              int depth, int kmax, int nx, int ny, int nz, int N)
{
    orc_params pp;
    memset(&pp, 0, sizeof(pp));
    pp.nx = nx; pp.ny = ny; pp.nz = nz; pp.N = N;
    ctx_t cx = {&pp, phi, NULL};
    const ctx_t *c = &cx;
    size_t ncell = (size_t)nx * ny * nz;
    int st0 = status[0], st1 = status[1];
#pragma omp parallel for collapse(2) reduction(max : st0, st1) schedule(static)
    for (int l = 0; l < nx; ++l)
        for (int m = 0; m < ny; ++m)
            for (int k = 0; k < nz; ++k) {
                int lp = idx_xp(l, nx), lm = idx_xm(l, nx);
                int mp = idx_xp(m, ny), mm = idx_xm(m, ny);
                int kp = idx_zp(k, nz), km = idx_zm(k, nz);
                size_t cell = ((size_t)l * ny + m) * nz + k;
                int n = 0;
                for (int i = 0; i < N; ++i) {
                    double v = PHI3(c, i, l, m, k);
                    if ((v > 0.0) ||
                        ((v == 0.0) && ((PHI3(c, i, lp, m, k) > 0.0) || (PHI3(c, i, lm, m, k) > 0.0) ||
                                        (PHI3(c, i, l, mp, k) > 0.0) || (PHI3(c, i, l, mm, k) > 0.0) ||
                                        (PHI3(c, i, l, m, kp) > 0.0) || (PHI3(c, i, l, m, km) > 0.0)))) {
                        n += 1;
                        if (n <= kmax && n <= depth) mf[(size_t)(n - 1) * ncell + cell] = i;
                    }
                }
                if (n > kmax) {
                    if (st0 < 1) st0 = 1;
                    nf[cell] = kmax;
                } else {
                    nf[cell] = n;
                }
                if (n > st1) st1 = n;
            }
    status[0] = st0;
    status[1] = st1;
}

/* Per-solid cached terms of one cell, src/gpu_kernels_3d.py:558-577 and
 * the switchable variant :706-733. */
static void cached_terms3(const ctx_t *c, int l, int m, int k, int gid, int aniso, int torque,
                          double gxl, double gyl, double gzl, double lap_liq,
                          double *lap_sl, double *b_sl)
{
    const orc_params *p = c->p;
    if (gid == 0) {
        *b_sl = 1.0;
        *lap_sl = 0.0;
        return;
    }
    if (aniso) {
        double lap_val = aniso_term1_3(c, l, m, k, gid);
        if (torque) lap_val += torque3(c, l, m, k, gid);
        double bc = best_cos3(c, gxl, gyl, gzl, gid);
        *b_sl = f32p(p, calc_b_from_cos(p, bc));
        *lap_sl = f32p(p, lap_val);
    } else {
        *lap_sl = f32p(p, (p->eps0_sl * p->eps0_sl) * lap_liq);
        *b_sl = 1.0;
    }
}

/* kernel_update_phasefield_active_3d(_switchable),
 * src/gpu_kernels_3d.py:530-680 and :683-830.  temp is dense [nx][ny][nz]. */
void orc_update3(const double *phi, double *phi_new, const double *temp,
                 const int32_t *mf, const int32_t *nf, int depth,
                 const double *wij, const double *aij, const double *mij, const double *n111,
                 const orc_params *p, int aniso, int torque)
{
    ctx_t cx = {p, phi, n111};
    const ctx_t *c = &cx;
    int nx = p->nx, ny = p->ny, nz = p->nz, N = p->N;
    size_t ncell = (size_t)nx * ny * nz;
    (void)depth;
#pragma omp parallel for collapse(2) schedule(dynamic, 1)
    for (int l = 0; l < nx; ++l)
        for (int m = 0; m < ny; ++m)
            for (int k = 0; k < nz; ++k) {
                double lap_sl[64], b_sl[64];
                int ids[64];
                int lp = idx_xp(l, nx), lm = idx_xm(l, nx);
                int mp = idx_xp(m, ny), mm = idx_xm(m, ny);
                int kp = idx_zp(k, nz), km = idx_zm(k, nz);
                size_t cell = ((size_t)l * ny + m) * nz + k;
                double inv_dx2 = 1.0 / (p->dx * p->dx);
                double Tcur = temp[cell];
                int n_act = nf[cell];
                for (int t = 0; t < n_act; ++t) ids[t] = mf[(size_t)t * ncell + cell];

                double gxl, gyl, gzl;
                grad3(c, 0, l, m, k, &gxl, &gyl, &gzl);
                double lap_liq = lap3(c, 0, l, m, k, inv_dx2);
                for (int t = 0; t < n_act; ++t)
                    cached_terms3(c, l, m, k, ids[t], aniso, torque, gxl, gyl, gzl, lap_liq,
                                  &lap_sl[t], &b_sl[t]);

                /* dominant solid, :579-595 / :735-751 */
                int i_s = -1;
                double maxv = -1.0;
                for (int t = 0; t < n_act; ++t) {
                    int gid = ids[t];
                    if (gid == 0) continue;
                    double v = PHI3(c, gid, l, m, k);
                    if (v > maxv) {
                        maxv = v;
                        i_s = gid;
                    }
                }
                double w_sl = p->w0_sl;
                if (aniso && i_s != -1) {
                    double sx, sy, sz;
                    grad3(c, i_s, l, m, k, &sx, &sy, &sz);
                    double cmax = best_cos3(c, sx, sy, sz, i_s);
                    double a_loc = calc_a_from_cos(p, cmax);
                    w_sl = p->w0_sl * (a_loc * a_loc);
                }

                for (int g = 0; g < N; ++g) phi_new[(size_t)g * ncell + cell] = 0.0;

                for (int t1 = 0; t1 < n_act; ++t1) {
                    int i = ids[t1];
                    double dpi = 0.0;
                    for (int t2 = 0; t2 < n_act; ++t2) {
                        int j = ids[t2];
                        if (i == j) continue;
                        double driving_force = 0.0;
                        if (i != 0 && j == 0)
                            driving_force = -p->Sf * (Tcur - p->T_melt);
                        else if (i == 0 && j != 0)
                            driving_force = p->Sf * (Tcur - p->T_melt);
                        double ppp = 0.0;
                        for (int t3 = 0; t3 < n_act; ++t3) {
                            int pk = ids[t3];
                            double phi_k = PHI3(c, pk, l, m, k);
                            double lap_k = (PHI3(c, pk, lp, m, k) + PHI3(c, pk, lm, m, k) +
                                            PHI3(c, pk, l, mp, k) + PHI3(c, pk, l, mm, k) +
                                            PHI3(c, pk, l, m, kp) + PHI3(c, pk, l, m, km) -
                                            6.0 * phi_k) * inv_dx2;
                            double wik = wij[(size_t)i * N + pk];
                            double wjk = wij[(size_t)j * N + pk];
                            if (pk == 0) {
                                if (i != 0) wik = w_sl;
                                if (j != 0) wjk = w_sl;
                            }
                            double term1 = (wik - wjk) * phi_k;
                            double term2 = 0.0;
                            if (j == 0 && i != 0) {
                                if (pk == i) term2 = 0.5 * lap_sl[t1];
                            } else if (i == 0 && j != 0) {
                                if (pk == j) term2 = -0.5 * lap_sl[t2];
                            } else {
                                double aik = aij[(size_t)i * N + pk], ajk = aij[(size_t)j * N + pk];
                                double eps2ik = aik * aik;
                                double eps2jk = ajk * ajk;
                                term2 = 0.5 * (eps2ik - eps2jk) * lap_k;
                            }
                            ppp += term1 + term2;
                        }
                        double phii_phij = PHI3(c, i, l, m, k) * PHI3(c, j, l, m, k);
                        double term_force = (8.0 / 3.1415926535) * sqrt(mpf_pymax(phii_phij, 0.0)) * driving_force;
                        double mij_eff = mij[(size_t)i * N + j];
                        if (i != 0 && j == 0)
                            mij_eff = mij[(size_t)i * N + j] * b_sl[t1];
                        else if (i == 0 && j != 0)
                            mij_eff = mij[(size_t)i * N + j] * b_sl[t2];
                        dpi -= 2.0 * mij_eff / (double)n_act * (ppp - term_force);
                    }
                    phi_new[(size_t)i * ncell + cell] = f32p(p, PHI3(c, i, l, m, k) + p->dt * dpi);
                }

                double s = 0.0;
                for (int t = 0; t < n_act; ++t) {
                    size_t o = (size_t)ids[t] * ncell + cell;
                    double v = phi_new[o];
                    if (v < 0.0) v = 0.0;
                    if (v > 1.0) v = 1.0;
                    phi_new[o] = v;
                    s += v;
                }
                if (s > 1e-20) {
                    double invs = 1.0 / s;
                    for (int t = 0; t < n_act; ++t) {
                        size_t o = (size_t)ids[t] * ncell + cell;
                        phi_new[o] = f32p(p, phi_new[o] * invs);
                    }
                } else {
                    phi_new[cell] = 1.0;
                }
            }
}

/* kernel_update_temp_3d / kernel_update_temp, src/gpu_kernels_3d.py:833-838,
 * src/gpu_kernels.py:520-525.  temp_f32 reproduces a float32 temperature
 * array (tests/verification_3d.py:817): the subtraction is done in float. */
void orc_temp(double *temp, size_t ncell, double cooling_rate, int temp_f32)
{
    for (size_t i = 0; i < ncell; ++i) {
        if (temp_f32) {
            float t = (float)temp[i];
            float cr = (float)cooling_rate;
            t = t - cr;
            temp[i] = (double)t;
        } else {
            temp[i] = temp[i] - cooling_rate;
        }
    }
}

/* Diagnostic probe: the per-solid terms of one 3D cell (for golden vectors).
 * out[0] aniso_term1, out[1] torque, out[2] b_sl, out[3..5] A13,
 * out[6..9] A12 at the cell, out[10] best_cos of the liquid gradient. */
void orc_probe3(const double *phi, const double *n111, const orc_params *p,
                int l, int m, int k, int s, double *out)
{
    ctx_t cx = {p, phi, n111};
    double gx, gy, gz;
    out[0] = aniso_term1_3(&cx, l, m, k, s);
    out[1] = torque3(&cx, l, m, k, s);
    grad3(&cx, 0, l, m, k, &gx, &gy, &gz);
    out[10] = best_cos3(&cx, gx, gy, gz, s);
    out[2] = calc_b_from_cos(p, out[10]);
    A13_3(&cx, l, m, k, s, out + 3);
    A12_3(&cx, l, m, k, s, out + 6);
}

double orc_calc_a(const orc_params *p, double c) { return calc_a_from_cos(p, c); }
double orc_calc_b(const orc_params *p, double c) { return calc_b_from_cos(p, c); }

/* ---- 2D kernels: src/gpu_kernels.py (x periodic, y mirror) -------------- */

/* grad_phi_xy, src/gpu_kernels.py:108-115 */
static void grad2(const ctx_t *c, int gid, int l, int m, double *gx, double *gy)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_zp(m, p->ny), mm = idx_zm(m, p->ny);
    *gx = (PHI2(c, gid, lp, m) - PHI2(c, gid, lm, m)) / (2.0 * p->dx);
    *gy = (PHI2(c, gid, l, mp) - PHI2(c, gid, l, mm)) / (2.0 * p->dx);
}

/* best_cos_from_grad, src/gpu_kernels.py:120-136 */
static double best_cos2(const ctx_t *c, double gx, double gy, int s)
{
    double g2 = gx * gx + gy * gy;
    if (g2 < c->p->g2_floor) return 0.0;
    double inv_g = 1.0 / sqrt(g2);
    double nxn = gx * inv_g, nyn = gy * inv_g;
    double best = 0.0;
    for (int t = 0; t < 8; ++t) {
        double v = fabs(nxn * N111(c, s, t, 0) + nyn * N111(c, s, t, 1));
        if (v > best) best = v;
    }
    if (best > 1.0) best = 1.0;
    return best;
}

/* eps2_at_cell_from_liquid, src/gpu_kernels.py:141-150 */
static double eps2_at2(const ctx_t *c, int l, int m, int s)
{
    double gx, gy;
    grad2(c, 0, l, m, &gx, &gy);
    double a = calc_a_from_cos(c->p, best_cos2(c, gx, gy, s));
    double e = c->p->eps0_sl * a;
    return e * e;
}

/* aniso_term1_solid, src/gpu_kernels.py:155-179 */
static double aniso_term1_2(const ctx_t *c, int l, int m, int s)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_zp(m, p->ny), mm = idx_zm(m, p->ny);
    double e_c = eps2_at2(c, l, m, s);
    double e_xp = eps2_at2(c, lp, m, s), e_xm = eps2_at2(c, lm, m, s);
    double e_yp = eps2_at2(c, l, mp, s), e_ym = eps2_at2(c, l, mm, s);
    double pc = PHI2(c, 0, l, m);
    double pxp = PHI2(c, 0, lp, m), pxm = PHI2(c, 0, lm, m);
    double pyp = PHI2(c, 0, l, mp), pym = PHI2(c, 0, l, mm);
    double Fx_p = (e_c + e_xp) * (pxp - pc) / (2.0 * p->dx);
    double Fx_m = (e_c + e_xm) * (pc - pxm) / (2.0 * p->dx);
    double Fy_p = (e_c + e_yp) * (pyp - pc) / (2.0 * p->dx);
    double Fy_m = (e_c + e_ym) * (pc - pym) / (2.0 * p->dx);
    return (Fx_p - Fx_m) / p->dx + (Fy_p - Fy_m) / p->dx;
}

/* d2_phi_xy, src/gpu_kernels.py:184-202.  out: xx, yy, xy */
static void d2_2(const ctx_t *c, int gid, int l, int m, double *o)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_zp(m, p->ny), mm = idx_zm(m, p->ny);
    double cc = PHI2(c, gid, l, m);
    double xp = PHI2(c, gid, lp, m), xm = PHI2(c, gid, lm, m);
    double yp = PHI2(c, gid, l, mp), ym = PHI2(c, gid, l, mm);
    double inv_dx2 = 1.0 / (p->dx * p->dx);
    o[0] = (xp - 2.0 * cc + xm) * inv_dx2;
    o[1] = (yp - 2.0 * cc + ym) * inv_dx2;
    double xpy = PHI2(c, gid, lp, mp), xpm = PHI2(c, gid, lp, mm);
    double xmy = PHI2(c, gid, lm, mp), xmm = PHI2(c, gid, lm, mm);
    o[2] = (xpy - xpm - xmy + xmm) * (0.25 * inv_dx2);
}

/* facet_cos_and_nxy_from_grad, src/gpu_kernels.py:207-232.  out: cs, |c|, nx, ny */
static void facet2(const ctx_t *c, double phix, double phiy, int s, double *o)
{
    double q = phix * phix + phiy * phiy;
    if (q < c->p->g2_floor) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 1.0; o[3] = 0.0;
        return;
    }
    double inv_g = 1.0 / sqrt(q);
    double gx = phix * inv_g, gy = phiy * inv_g;
    double best_abs = 0.0, best_signed = 0.0, bnx = 1.0, bny = 0.0;
    for (int t = 0; t < 8; ++t) {
        double ux = N111(c, s, t, 0), uy = N111(c, s, t, 1);
        double dot = gx * ux + gy * uy;
        double ad = fabs(dot);
        if (ad > best_abs) {
            best_abs = ad; best_signed = dot; bnx = ux; bny = uy;
        }
    }
    if (best_abs > 1.0) best_abs = 1.0;
    if (best_signed > 1.0) best_signed = 1.0;
    if (best_signed < -1.0) best_signed = -1.0;
    o[0] = best_signed; o[1] = best_abs; o[2] = bnx; o[3] = bny;
}

/* da_dphixy_A12, src/gpu_kernels.py:237-264.  out: a, da/dphix, da/dphiy */
static void A12_2(const ctx_t *c, int l, int m, int s, double *o)
{
    const orc_params *p = c->p;
    double phix, phiy;
    grad2(c, 0, l, m, &phix, &phiy);
    double q = phix * phix + phiy * phiy;
    if (q < p->g2_floor) {
        o[0] = 0.0; o[1] = 0.0; o[2] = 0.0;
        return;
    }
    double f[4];
    facet2(c, phix, phiy, s, f);
    double a_val = calc_a_from_cos(p, f[1]);
    double g = sqrt(q);
    double inv_g = 1.0 / g;
    double inv_g3 = 1.0 / (q * g);
    double udotg = phix * f[2] + phiy * f[3];
    double dcx = f[2] * inv_g - udotg * phix * inv_g3;
    double dcy = f[3] * inv_g - udotg * phiy * inv_g3;
    double cs = f[0];
    double C = sqrt(cs * cs + p->p_round * p->p_round);
    double S = sqrt(mpf_pymax(1.0 - cs * cs, 0.0) + p->p_round * p->p_round);
    double coef = (cs / C) - o_tan(p, p->a0) * (cs / S);
    o[0] = a_val;
    o[1] = p->mu_a * p->delta_a * coef * dcx;
    o[2] = p->mu_a * p->delta_a * coef * dcy;
}

/* d_dx_da_dphix_and_d_dy_da_dphiy_A13, src/gpu_kernels.py:269-285 */
static void A13_2(const ctx_t *c, int l, int m, int s, double *o)
{
    const orc_params *p = c->p;
    int lp = idx_xp(l, p->nx), lm = idx_xm(l, p->nx);
    int mp = idx_zp(m, p->ny), mm = idx_zm(m, p->ny);
    double a[3], b[3];
    A12_2(c, lp, m, s, a); A12_2(c, lm, m, s, b);
    o[0] = (a[1] - b[1]) * (0.5 / p->dx);
    A12_2(c, l, mp, s, a); A12_2(c, l, mm, s, b);
    o[1] = (a[2] - b[2]) * (0.5 / p->dx);
}

/* torque_A11, src/gpu_kernels.py:290-332 */
static double torque2(const ctx_t *c, int l, int m, int s)
{
    const orc_params *p = c->p;
    double phix, phiy;
    grad2(c, 0, l, m, &phix, &phiy);
    double q = phix * phix + phiy * phiy;
    if (q < p->g2_floor) return 0.0;
    double d2[3];
    d2_2(c, 0, l, m, d2);
    double phixx = d2[0], phiyy = d2[1], phixy = d2[2];
    double q_x = 2.0 * (phix * phixx + phiy * phixy);
    double q_y = 2.0 * (phix * phixy + phiy * phiyy);
    double a12[3];
    A12_2(c, l, m, s, a12);
    double f[4];
    facet2(c, phix, phiy, s, f);
    double g = sqrt(q);
    double inv_g = 1.0 / g;
    double inv_g3 = 1.0 / (q * g);
    double udotg = phix * f[2] + phiy * f[3];
    double dcx = f[2] * inv_g - udotg * phix * inv_g3;
    double dcy = f[3] * inv_g - udotg * phiy * inv_g3;
    double c_x = dcx * phixx + dcy * phixy;
    double c_y = dcx * phixy + dcy * phiyy;
    double cs = f[0];
    double C = sqrt(cs * cs + p->p_round * p->p_round);
    double S = sqrt(mpf_pymax(1.0 - cs * cs, 0.0) + p->p_round * p->p_round);
    double coef = (cs / C) - o_tan(p, p->a0) * (cs / S);
    double da_dx = p->mu_a * p->delta_a * coef * c_x;
    double da_dy = p->mu_a * p->delta_a * coef * c_y;
    double a13[2];
    A13_2(c, l, m, s, a13);
    double a_val = a12[0];
    double Ex = (da_dx * a12[1] * q) + (a_val * a13[0] * q) + (a_val * a12[1] * q_x);
    double Ey = (da_dy * a12[2] * q) + (a_val * a13[1] * q) + (a_val * a12[2] * q_y);
    return (p->eps0_sl * p->eps0_sl) * (Ex + Ey);
}

/* kernel_update_nfmf, src/gpu_kernels.py:337-356, with the checked clamp of
 * the 3D kernel (src/gpu_kernels_3d.py:518-527) so that overflow is defined. */
void orc_apt2(const double *phi, int32_t *mf, int32_t *nf, int32_t *status,
              int depth, int kmax, int nx, int ny, int N)
{
    orc_params pp;
    memset(&pp, 0, sizeof(pp));
    pp.nx = nx; pp.ny = ny; pp.nz = 1; pp.N = N;
    ctx_t cx = {&pp, phi, NULL};
    const ctx_t *c = &cx;
    size_t ncell = (size_t)nx * ny;
    int st0 = status[0], st1 = status[1];
    for (int l = 0; l < nx; ++l)
        for (int m = 0; m < ny; ++m) {
            int lp = idx_xp(l, nx), lm = idx_xm(l, nx);
            int mp = idx_zp(m, ny), mm = idx_zm(m, ny);
            size_t cell = (size_t)l * ny + m;
            int n = 0;
            for (int i = 0; i < N; ++i) {
                double v = PHI2(c, i, l, m);
                if ((v > 0.0) || ((v == 0.0) && ((PHI2(c, i, lp, m) > 0.0) || (PHI2(c, i, lm, m) > 0.0) ||
                                                 (PHI2(c, i, l, mp) > 0.0) || (PHI2(c, i, l, mm) > 0.0)))) {
                    n += 1;
                    if (n <= kmax && n <= depth) mf[(size_t)(n - 1) * ncell + cell] = i;
                }
            }
            if (n > kmax) {
                if (st0 < 1) st0 = 1;
                nf[cell] = kmax;
            } else {
                nf[cell] = n;
            }
            if (n > st1) st1 = n;
        }
    status[0] = st0;
    status[1] = st1;
}

/* kernel_update_phasefield_active, src/gpu_kernels.py:361-515 */
void orc_update2(const double *phi, double *phi_new, const double *temp,
                 const int32_t *mf, const int32_t *nf, int depth,
                 const double *wij, const double *aij, const double *mij, const double *n111,
                 const orc_params *p)
{
    ctx_t cx = {p, phi, n111};
    const ctx_t *c = &cx;
    int nx = p->nx, ny = p->ny, N = p->N;
    size_t ncell = (size_t)nx * ny;
    (void)depth;
    for (int l = 0; l < nx; ++l)
        for (int m = 0; m < ny; ++m) {
            double lap_sl[64], b_sl[64];
            int ids[64];
            int lp = idx_xp(l, nx), lm = idx_xm(l, nx);
            int mp = idx_zp(m, ny), mm = idx_zm(m, ny);
            size_t cell = (size_t)l * ny + m;
            double inv_dx2 = 1.0 / (p->dx * p->dx);
            double Tcur = temp[cell];
            int n_act = nf[cell];
            for (int t = 0; t < n_act; ++t) ids[t] = mf[(size_t)t * ncell + cell];
            for (int t = 0; t < n_act; ++t) {
                int gid = ids[t];
                if (gid == 0) {
                    b_sl[t] = 1.0;
                    lap_sl[t] = 0.0;
                } else {
                    lap_sl[t] = aniso_term1_2(c, l, m, gid);
                    lap_sl[t] += torque2(c, l, m, gid);
                    double gx, gy;
                    grad2(c, 0, l, m, &gx, &gy);
                    b_sl[t] = calc_b_from_cos(p, best_cos2(c, gx, gy, gid));
                }
            }
            /* dominant solid, src/gpu_kernels.py:410-436 */
            int i_s = 1;
            double maxv = 0.0;
            for (int t = 0; t < n_act; ++t) {
                int gid = ids[t];
                if (gid == 0) continue;
                double v = PHI2(c, gid, l, m);
                if (v > maxv) {
                    maxv = v;
                    i_s = gid;
                }
            }
            double phix_ = (PHI2(c, i_s, lp, m) - PHI2(c, i_s, lm, m)) * (0.5 / p->dx);
            double phiy_ = (PHI2(c, i_s, l, mp) - PHI2(c, i_s, l, mm)) * (0.5 / p->dx);
            double gn2 = phix_ * phix_ + phiy_ * phiy_;
            double cmax = 0.0;
            if (gn2 >= p->g2_floor) {
                double inv_gn = 1.0 / sqrt(gn2);
                double nx_ = phix_ * inv_gn, ny_ = phiy_ * inv_gn;
                for (int tt = 0; tt < 8; ++tt) {
                    double v = fabs(nx_ * N111(c, i_s, tt, 0) + ny_ * N111(c, i_s, tt, 1));
                    if (v > cmax) cmax = v;
                }
                if (cmax > 1.0) cmax = 1.0;
            }
            double a_loc = calc_a_from_cos(p, cmax);
            double w_sl = p->w0_sl * (a_loc * a_loc);

            for (int g = 0; g < N; ++g) phi_new[(size_t)g * ncell + cell] = 0.0;
            for (int t1 = 0; t1 < n_act; ++t1) {
                int i = ids[t1];
                double dpi = 0.0;
                for (int t2 = 0; t2 < n_act; ++t2) {
                    int j = ids[t2];
                    if (i == j) continue;
                    double driving_force = 0.0;
                    if (i != 0 && j == 0)
                        driving_force = -p->Sf * (Tcur - p->T_melt);
                    else if (i == 0 && j != 0)
                        driving_force = p->Sf * (Tcur - p->T_melt);
                    double ppp = 0.0;
                    for (int t3 = 0; t3 < n_act; ++t3) {
                        int kk = ids[t3];
                        double lap_k = (PHI2(c, kk, lp, m) + PHI2(c, kk, lm, m) + PHI2(c, kk, l, mp) +
                                        PHI2(c, kk, l, mm) - 4.0 * PHI2(c, kk, l, m)) * inv_dx2;
                        double wik = wij[(size_t)i * N + kk], wjk = wij[(size_t)j * N + kk];
                        if (kk == 0) {
                            if (i != 0) wik = w_sl;
                            if (j != 0) wjk = w_sl;
                        }
                        double term1 = (wik - wjk) * PHI2(c, kk, l, m);
                        double term2 = 0.0;
                        if (j == 0 && i != 0) {
                            if (kk == i) term2 = 0.5 * lap_sl[t1];
                        } else if (i == 0 && j != 0) {
                            if (kk == j) term2 = -0.5 * lap_sl[t2];
                        } else {
                            double aik = aij[(size_t)i * N + kk], ajk = aij[(size_t)j * N + kk];
                            double eps2ik = aik * aik;
                            double eps2jk = ajk * ajk;
                            term2 = 0.5 * (eps2ik - eps2jk) * lap_k;
                        }
                        ppp += term1 + term2;
                    }
                    double phii_phij = PHI2(c, i, l, m) * PHI2(c, j, l, m);
                    double term_force = (8.0 / 3.1415926535) * sqrt(mpf_pymax(phii_phij, 0.0)) * driving_force;
                    double mij_eff = mij[(size_t)i * N + j];
                    if (i != 0 && j == 0)
                        mij_eff = mij[(size_t)i * N + j] * b_sl[t1];
                    else if (i == 0 && j != 0)
                        mij_eff = mij[(size_t)i * N + j] * b_sl[t2];
                    dpi -= 2.0 * mij_eff / (double)n_act * (ppp - term_force);
                }
                phi_new[(size_t)i * ncell + cell] = PHI2(c, i, l, m) + p->dt * dpi;
            }
            double s = 0.0;
            for (int t = 0; t < n_act; ++t) {
                size_t o = (size_t)ids[t] * ncell + cell;
                double v = phi_new[o];
                if (v < 0.0) v = 0.0;
                if (v > 1.0) v = 1.0;
                phi_new[o] = v;
                s += v;
            }
            if (s > 1e-20) {
                double invs = 1.0 / s;
                for (int t = 0; t < n_act; ++t) phi_new[(size_t)ids[t] * ncell + cell] *= invs;
            } else {
                phi_new[cell] = 1.0;
            }
        }
}

/* The verification time loop, tests/verification_3d.py:830-905, on the CPU:
 * temperature -> checked APT -> (halt on KMAX overflow) -> update -> swap.
 * phi (in/out) [N][...] ; temp (in/out) dense.  dim = 3 or 2.
 * Returns the number of completed updates; on an overflow with halt_on_kmax
 * the state is the one the overflowing step started from (temp already
 * decremented), matching the reference loop that raises before the update.
 * status[0] |= overflow seen, status[1] = max active count seen. */
int orc_run(int dim, double *phi, double *temp, int nsteps, int kmax,
            const double *wij, const double *aij, const double *mij, const double *n111,
            const orc_params *p, int aniso, int torque, double cooling_rate, int temp_f32,
            int halt_on_kmax, int32_t *status)
{
    size_t ncell = (size_t)p->nx * p->ny * (dim == 3 ? p->nz : 1);
    size_t nphi = ncell * p->N;
    int depth = kmax;
    double *buf = (double *)malloc(nphi * sizeof(double));
    int32_t *mf = (int32_t *)calloc((size_t)depth * ncell, sizeof(int32_t));
    int32_t *nf = (int32_t *)calloc(ncell, sizeof(int32_t));
    if (!buf || !mf || !nf) {
        free(buf); free(mf); free(nf);
        return -1;
    }
    double *cur = phi, *nxt = buf;
    int done = 0;
    for (int step = 1; step <= nsteps; ++step) {
        orc_temp(temp, ncell, cooling_rate, temp_f32);
        int32_t st[2] = {0, 0};
        if (dim == 3)
            orc_apt3(cur, mf, nf, st, depth, kmax, p->nx, p->ny, p->nz, p->N);
        else
            orc_apt2(cur, mf, nf, st, depth, kmax, p->nx, p->ny, p->N);
        if (st[0]) status[0] = 1;
        if (st[1] > status[1]) status[1] = st[1];
        if (st[0] && halt_on_kmax) break;
        if (dim == 3)
            orc_update3(cur, nxt, temp, mf, nf, depth, wij, aij, mij, n111, p, aniso, torque);
        else
            orc_update2(cur, nxt, temp, mf, nf, depth, wij, aij, mij, n111, p);
        double *t = cur; cur = nxt; nxt = t;
        done = step;
    }
    if (cur != phi) memcpy(phi, cur, nphi * sizeof(double));
    free(buf); free(mf); free(nf);
    return done;
}

/* Threads the OpenMP loops of the oracle use. */
int orc_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* The shared transcendentals, exported so the tests can compare them with
 * the C library directly. */
double orc_mpf_tan(double x) { return mpf_tan(x); }
double orc_mpf_acos(double x) { return mpf_acos(x); }
double orc_mpf_tanh(double x) { return mpf_tanh(x); }
