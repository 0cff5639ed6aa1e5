/*
 * steps_oracle.c -- CPU restatement of the StePS direct-summation force path.
 *
 * TEST INFRASTRUCTURE ONLY (see steps_oracle.h): the checker for the HIP
 * product, never linked into it.  Built with -O2 -fopenmp and without
 * -ffast-math so that results are reproducible; the timed reference-faithful
 * baseline lives in steps_oracle_faithful.c.
 *
 * Citations are StePS/src/<file>:<line> in the reference repository.
 */
#include "steps_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <omp.h>

#define ORC_PI 3.14159265358979323846264338327950288419716939937510 /* global_variables.h:16 */
#define ORC_SQRTPI 1.7724538509055160272981674833411                /* ewald_space.cc:23 */
#define ORC_RADIAL_EXTENT 2.25                                       /* global_variables.h:39 */

/* ------------------------------------------------------------------------ */
/* softening law                                                             */
/* ------------------------------------------------------------------------ */

/* forces.cc:52-87 -- cubic-spline softened force coefficient, CPU form
 * (pow-based constants). */
double oracle_force_softening(double r, double beta)
{
    double betap2 = beta * 0.5;
    double wij;
    if (r >= beta) {
        wij = pow(r, -3);
    } else if (r > betap2 && r < beta) {
        double c0 = -32.0 / (3.0 * pow(beta, 6));
        double c1 = 38.4 / pow(beta, 5);
        double c2 = -48.0 / pow(beta, 4);
        double c3 = 64.0 / (3.0 * pow(beta, 3));
        double c4 = -1.0 / 15.0;
        wij = c0 * pow(r, 3) + c1 * pow(r, 2) + c2 * r + c3 + c4 / pow(r, 3);
    } else {
        double c0 = 32.0 / pow(beta, 6);
        double c1 = -38.4 / pow(beta, 5);
        double c2 = 32.0 / (3.0 * pow(beta, 3));
        wij = c0 * pow(r, 3) + c1 * pow(r, 2) + c2;
    }
    return wij;
}

/* forces_cuda.cu:466-519 -- same law, GPU form (1/(r*r*r), products of beta). */
double oracle_force_softening_gpu(double r, double beta)
{
    const double half_beta = beta * 0.5;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double inv_r3 = 1.0 / r3;
    if (r >= beta)
        return inv_r3;
    const double b2 = beta * beta, b3 = b2 * beta, b4 = b2 * b2, b5 = b4 * beta, b6 = b3 * b3;
    if (r > half_beta) {
        const double c0 = -32.0 / (3.0 * b6), c1 = 38.4 / b5, c2 = -48.0 / b4;
        const double c3 = 64.0 / (3.0 * b3), c4 = -1.0 / 15.0;
        return c0 * r3 + c1 * r2 + c2 * r + c3 + c4 * inv_r3;
    }
    const double c0 = 32.0 / b6, c1 = -38.4 / b5, c2 = (32.0 / 3.0) / b3;
    return c0 * r3 + c1 * r2 + c2;
}

/* ------------------------------------------------------------------------ */
/* T3 Ewald correction table                                                 */
/* ------------------------------------------------------------------------ */

/* ewald_space.cc:118-205 -- integer image vectors with |n|^2 < R^2, each
 * coordinate in (-R, R), emitted in the reference's enumeration order. */
static int ewald_space_impl(double R, int *idx)
{
    int l = 0;
    for (int i = 0; i < R; i++)
        for (int j = 0; j < R; j++)
            for (int k = 0; k < R; k++) {
                if (!((double)(i * i + j * j + k * k) < R * R))
                    continue;
                int n2 = i * i + j * j + k * k;
                int v[8][3];
                int c = 0;
                v[c][0] = i;  v[c][1] = j;  v[c][2] = k;  c++;
                if (i != 0)           { v[c][0] = -i; v[c][1] = j;  v[c][2] = k;  c++; }
                if (j != 0)           { v[c][0] = i;  v[c][1] = -j; v[c][2] = k;  c++; }
                if (k != 0)           { v[c][0] = i;  v[c][1] = j;  v[c][2] = -k; c++; }
                if (j != 0 && k != 0) { v[c][0] = i;  v[c][1] = -j; v[c][2] = -k; c++; }
                if (i != 0 && k != 0) { v[c][0] = -i; v[c][1] = j;  v[c][2] = -k; c++; }
                if (i != 0 && j != 0) { v[c][0] = -i; v[c][1] = -j; v[c][2] = k;  c++; }
                if (i != 0 && j != 0 && k != 0) { v[c][0] = -i; v[c][1] = -j; v[c][2] = -k; c++; }
                for (int q = 0; q < c; q++) {
                    if (idx) {
                        idx[4 * l + 0] = v[q][0];
                        idx[4 * l + 1] = v[q][1];
                        idx[4 * l + 2] = v[q][2];
                        idx[4 * l + 3] = n2;
                    }
                    l++;
                }
            }
    return l;
}

int oracle_ewald_space_count(double R) { return ewald_space_impl(R, NULL); }
int oracle_ewald_space(double R, int *idx) { return ewald_space_impl(R, idx); }

/* ewald_space.cc:210-286 -- D(r) = (F_real + F_rec) - F_newton per unit G m_i m_j. */
static void t3_correction_idx(const double r[3], double L, double alpha, int nreal,
                              const int *ridx, int nrec, const int *kidx, double rel_cut,
                              double rec_cut, double D[3])
{
    double Fr[3] = {0, 0, 0}, Fk[3] = {0, 0, 0};
    const double V = L * L * L;
    const double two_alpha_over_sqrtpi = 2.0 * alpha / sqrt(ORC_PI);
    const double fourpi_over_V = (4.0 * ORC_PI) / V;
    const double relcutL2 = (rel_cut * L) * (rel_cut * L);
    for (int m = 0; m < nreal; ++m) {
        double Rv[3] = {r[0] + (double)ridx[4 * m] * L, r[1] + (double)ridx[4 * m + 1] * L,
                        r[2] + (double)ridx[4 * m + 2] * L};
        double R2 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        if (R2 == 0.0 || R2 > relcutL2)
            continue;
        double R = sqrt(R2);
        double invR = 1.0 / R;
        double invR3 = invR * invR * invR;
        double bracket = erfc(alpha * R) + two_alpha_over_sqrtpi * R * exp(-(alpha * alpha) * R2);
        double coeff = invR3 * bracket;
        Fr[0] += -Rv[0] * coeff;
        Fr[1] += -Rv[1] * coeff;
        Fr[2] += -Rv[2] * coeff;
    }
    for (int m = 0; m < nrec; ++m) {
        int h = kidx[4 * m], k = kidx[4 * m + 1], l = kidx[4 * m + 2];
        if (h == 0 && k == 0 && l == 0)
            continue;
        double factor = 2.0 * ORC_PI / L;
        double kv[3] = {factor * (double)h, factor * (double)k, factor * (double)l};
        double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (k2 == 0.0 || k2 > rec_cut * rec_cut)
            continue;
        double damp = exp(-k2 / (4.0 * alpha * alpha));
        double coeff = fourpi_over_V * damp / k2;
        double s = sin(kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2]);
        Fk[0] += -kv[0] * coeff * s;
        Fk[1] += -kv[1] * coeff * s;
        Fk[2] += -kv[2] * coeff * s;
    }
    double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (r2 > 0.0) {
        double rn = sqrt(r2);
        double invr3 = 1.0 / (rn * rn * rn);
        double Fn[3] = {-r[0] * invr3, -r[1] * invr3, -r[2] * invr3};
        D[0] = (Fr[0] + Fk[0]) - Fn[0];
        D[1] = (Fr[1] + Fk[1]) - Fn[1];
        D[2] = (Fr[2] + Fk[2]) - Fn[2];
    } else {
        D[0] = D[1] = D[2] = 0.0;
    }
}

void oracle_t3_ewald_correction(const double r[3], double L, double alpha, double rel_cut,
                                double rec_cut, double D[3])
{
    /* main.cc:467-468: index sets ewald_space(rel_cut+1), ewald_space(rec_cut+2) */
    int nreal = oracle_ewald_space_count(rel_cut + 1.0);
    int nrec = oracle_ewald_space_count(rec_cut + 2.0);
    int *ridx = (int *)malloc(sizeof(int) * 4 * (size_t)nreal);
    int *kidx = (int *)malloc(sizeof(int) * 4 * (size_t)nrec);
    oracle_ewald_space(rel_cut + 1.0, ridx);
    oracle_ewald_space(rec_cut + 2.0, kidx);
    t3_correction_idx(r, L, alpha, nreal, ridx, nrec, kidx, rel_cut, rec_cut, D);
    free(ridx);
    free(kidx);
}

/* ewald_space.cc:63-114 -- the 24 proper rotations of O, in the reference order. */
typedef struct { int perm[3]; int sign[3]; } orc_rot;
static void build_o_rotations(orc_rot R[24])
{
    static const int pe[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    static const int po[3][3] = {{0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    static const int sp[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    static const int sn[4][3] = {{-1, -1, -1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
    int n = 0;
    for (int p = 0; p < 3; ++p)
        for (int s = 0; s < 4; ++s, ++n)
            for (int k = 0; k < 3; ++k) { R[n].perm[k] = pe[p][k]; R[n].sign[k] = sp[s][k]; }
    for (int p = 0; p < 3; ++p)
        for (int s = 0; s < 4; ++s, ++n)
            for (int k = 0; k < 3; ++k) { R[n].perm[k] = po[p][k]; R[n].sign[k] = sn[s][k]; }
}

static inline size_t t3_off(int Ng, int ix, int iy, int iz, int c)
{
    /* ewald_space.cc:46-49 */
    return ((size_t)((ix * Ng + iy) * Ng + iz) * 3u) + (size_t)c;
}

/* ewald_space.cc:288-383 -- wedge x>=y>=z>=0 evaluated, grid filled by O_h. */
void oracle_t3_ewald_table(int Ng, double L, double alpha, double rel_cut, double rec_cut,
                           double *T)
{
    int nreal = oracle_ewald_space_count(rel_cut + 1.0);
    int nrec = oracle_ewald_space_count(rec_cut + 2.0);
    int *ridx = (int *)malloc(sizeof(int) * 4 * (size_t)nreal);
    int *kidx = (int *)malloc(sizeof(int) * 4 * (size_t)nrec);
    oracle_ewald_space(rel_cut + 1.0, ridx);
    oracle_ewald_space(rec_cut + 2.0, kidx);
    memset(T, 0, (size_t)Ng * Ng * Ng * 3 * sizeof(double));
    orc_rot R[24];
    build_o_rotations(R);
    const double gs = L / (double)Ng;
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = Ng / 2; i < Ng; ++i)
        for (int j = Ng / 2; j <= i; ++j)
            for (int k = Ng / 2; k <= j; ++k) {
                double r[3] = {((double)i + 0.5) * gs - L / 2.0, ((double)j + 0.5) * gs - L / 2.0,
                               ((double)k + 0.5) * gs - L / 2.0};
                double Dw[3];
                t3_correction_idx(r, L, alpha, nreal, ridx, nrec, kidx, rel_cut, rec_cut, Dw);
                for (int c = 0; c < 3; ++c)
                    T[t3_off(Ng, i, j, k, c)] = Dw[c];
                int src[3] = {i, j, k};
                for (int q = 0; q < 24; ++q) {
                    int dst[3];
                    double Dr[3];
                    for (int d = 0; d < 3; ++d) {
                        int sc = R[q].perm[d];
                        int v = src[sc];
                        if (R[q].sign[d] < 0)
                            v = (Ng - 1) - v;
                        dst[d] = v;
                        Dr[d] = R[q].sign[d] * Dw[sc];
                    }
                    for (int c = 0; c < 3; ++c)
                        T[t3_off(Ng, dst[0], dst[1], dst[2], c)] = Dr[c];
                    int inv[3] = {(Ng - 1) - dst[0], (Ng - 1) - dst[1], (Ng - 1) - dst[2]};
                    for (int c = 0; c < 3; ++c)
                        T[t3_off(Ng, inv[0], inv[1], inv[2], c)] = -Dr[c];
                }
            }
    free(ridx);
    free(kidx);
}

/* main.cc:422-448 -- table parameters for IS_PERIODIC = 2, 3, >=4. */
void oracle_t3_params(int is_periodic, double L, double *alpha, double *rel_cut,
                      double *rec_cut, int *ngrid)
{
    *alpha = 2.0 / L;
    if (is_periodic == 2) { *rel_cut = 2.6; *rec_cut = 8.0; *ngrid = 63; }
    else if (is_periodic == 3) { *rel_cut = 3.6; *rec_cut = 10.0; *ngrid = 127; }
    else { *rel_cut = 4.6; *rec_cut = 12.0; *ngrid = 255; }
}

static inline int imodp(int i, int n)
{
    int r = i % n; /* forces_cuda.cu:87-92 */
    return (r < 0) ? (r + n) : r;
}

/* forces_cuda.cu:93-101 */
static inline void map_to_centered_grid(double r, double L, int Ng, int *i0, double *fx)
{
    double gs = L / (double)Ng;
    double u = (r + L * 0.5) / gs - 0.5;
    double uf = floor(u);
    *i0 = imodp((int)uf, Ng);
    *fx = u - uf;
}

/* forces_cuda.cu:103-111 (GPU) == ewald_space.cc:440-449 (CPU order 4): Catmull-Rom */
static inline void cubic_weights(double t, double w[4])
{
    double t2 = t * t, t3 = t * t2;
    w[0] = -0.5 * t3 + t2 - 0.5 * t;
    w[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
    w[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
    w[3] = 0.5 * t3 - 0.5 * t2;
}

/* ewald_space.cc:409-438 -- 1-D Lagrange weights on a uniform stencil */
static void lagrange_weights(double fx, int stencil, double *w)
{
    int start = -(stencil / 2 - 1);
    for (int j = 0; j < stencil; j++) {
        w[j] = 1.0;
        double xj = (double)(start + j);
        for (int k = 0; k < stencil; k++) {
            if (k == j) continue;
            double xk = (double)(start + k);
            w[j] *= (fx - xk) / (xj - xk);
        }
    }
}

/* forces_cuda.cu:113-167 (order 4, Catmull-Rom: the GPU path) and
 * ewald_space.cc:452-523 (order 2: trilinear, the CPU default build). */
void oracle_t3_interpolate(int Ng, double L, const double *T, double dx, double dy, double dz,
                           double D[3], int order)
{
    int ix0, iy0, iz0;
    double fx, fy, fz;
    map_to_centered_grid(dx, L, Ng, &ix0, &fx);
    map_to_centered_grid(dy, L, Ng, &iy0, &fy);
    map_to_centered_grid(dz, L, Ng, &iz0, &fz);
    double wx[8], wy[8], wz[8];
    int so;
    if (order == 4) {
        cubic_weights(fx, wx);
        cubic_weights(fy, wy);
        cubic_weights(fz, wz);
        so = 4;
    } else {
        if (order < 2 || order % 2 != 0) order = 2;
        so = order > 8 ? 8 : order;
        lagrange_weights(fx, so, wx);
        lagrange_weights(fy, so, wy);
        lagrange_weights(fz, so, wz);
    }
    double s[3] = {0, 0, 0};
    int off = -(so / 2 - 1);
    for (int i = 0; i < so; i++) {
        int ix = imodp(ix0 + off + i, Ng);
        for (int j = 0; j < so; j++) {
            int iy = imodp(iy0 + off + j, Ng);
            double wxy = wx[i] * wy[j];
            for (int k = 0; k < so; k++) {
                int iz = imodp(iz0 + off + k, Ng);
                double w = wxy * wz[k];
                size_t id = t3_off(Ng, ix, iy, iz, 0);
                s[0] += w * T[id];
                s[1] += w * T[id + 1];
                s[2] += w * T[id + 2];
            }
        }
    }
    D[0] = s[0];
    D[1] = s[1];
    D[2] = s[2];
}

/* ------------------------------------------------------------------------ */
/* S1R2 Ewald correction table (Tornberg 1P sum)                             */
/* ------------------------------------------------------------------------ */

/* ewald_space.cc:539-561 -- the reference's erfcx fit (A&S 7.1.26 coefficients). */
double oracle_fast_erfcx(double x)
{
    if (x < 0) return 2.0 * exp(x * x) - oracle_fast_erfcx(-x);
    if (x == 0) return 1.0;
    if (x > 1e10) return 0.564189583547756286 / x;
    double t = 1.0 / (1.0 + 0.3275911 * x);
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                 a5 = 1.061405429;
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
}

/* ewald_space.cc:564-595 -- exp(+-km*rho) * erfc(km/(2 alpha) +- alpha*rho) */
static double exp_erfc_product(double km, double rho, double alpha, int term_p)
{
    double a = km / (2.0 * alpha);
    double ar = alpha * rho;
    double arg = term_p ? (a + ar) : (a - ar);
    double common = -(a * a + ar * ar);
    if (arg > 5.0)
        return exp(common) * oracle_fast_erfcx(arg);
    else if (arg < -5.0)
        return 2.0 * exp(-km * rho) - exp(common) * oracle_fast_erfcx(-arg);
    double shift = term_p ? (km * rho) : (-km * rho);
    return exp(shift) * erfc(arg);
}

/* ewald_space.cc:599-608 */
static double wrap_dz(double dz, double Lz)
{
    double half = 0.5 * Lz;
    if (dz > half) dz -= Lz * floor((dz + half) / Lz);
    if (dz <= -half) dz -= Lz * floor((dz - half) / Lz);
    return dz;
}

/* ewald_space.cc:618-752 -- S1xR2 periodic force per unit G m_i m_j.
 * rspace=0: Ewald (default); rspace=1: PERIODIC_Z_RSPACELOOKUP direct image sum. */
void oracle_s1r2_force_pair(double dx, double dy, double dz, double Lz, double alpha, int nmax,
                            int mmax, int rspace, double F[3])
{
    dz = wrap_dz(dz, Lz);
    double rho2 = dx * dx + dy * dy;
    if (!rspace) {
        double rho = sqrt(rho2);
        double Fxr = 0, Fyr = 0, Fzr = 0;
        for (int n = -nmax; n <= nmax; ++n) {
            double dzn = dz + (double)n * Lz;
            double r2 = rho2 + dzn * dzn;
            if (r2 < 1e-18) continue;
            double r = sqrt(r2);
            double invr3 = 1.0 / (r2 * r);
            double ar = alpha * r;
            double coeff = (erfc(ar) + (2.0 / ORC_SQRTPI) * ar * exp(-ar * ar)) * invr3;
            Fxr -= dx * coeff;
            Fyr -= dy * coeff;
            Fzr -= dzn * coeff;
        }
        double Frk = 0, Fzk = 0, invL = 1.0 / Lz;
        for (int m = 1; m <= mmax; ++m) {
            double km = (2.0 * ORC_PI * m) * invL;
            double tp = exp_erfc_product(km, rho, alpha, 1);
            double tm = exp_erfc_product(km, rho, alpha, 0);
            double B = tp + tm;
            double dB = km * (tp - tm);
            Frk += (2.0 * invL) * cos(km * dz) * dB;
            Fzk -= (2.0 * invL) * km * sin(km * dz) * B;
        }
        if (rho > 1e-12) {
            double Fr0 = -(2.0 * invL / rho) * (1.0 - exp(-alpha * alpha * rho2));
            Frk += Fr0;
        }
        double Fxk = (rho > 0) ? (Frk * dx / rho) : 0;
        double Fyk = (rho > 0) ? (Frk * dy / rho) : 0;
        F[0] = Fxr + Fxk;
        F[1] = Fyr + Fyk;
        F[2] = Fzr + Fzk;
    } else {
        double Fx = 0, Fy = 0, Fz = 0;
        double r2 = rho2 + dz * dz;
        if (r2 >= 1e-18) {
            double r = sqrt(r2), invr3 = 1.0 / (r2 * r);
            Fx -= dx * invr3; Fy -= dy * invr3; Fz -= dz * invr3;
        }
        for (int n = 1; n <= nmax; ++n)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                double dzn = dz + (double)(sgn * n) * Lz;
                r2 = rho2 + dzn * dzn;
                if (r2 < 1e-18) continue;
                double r = sqrt(r2), invr3 = 1.0 / (r2 * r);
                Fx -= dx * invr3; Fy -= dy * invr3; Fz -= dzn * invr3;
            }
        F[0] = Fx; F[1] = Fy; F[2] = Fz;
    }
}

/* ewald_space.cc:754-798 */
void oracle_s1r2_table(int Nrho, int Nz, double rho_max, double Lz, double alpha, int nmax,
                       int mmax, int rspace, double *T)
{
    double drho = rho_max / (double)(Nrho - 1 > 1 ? Nrho - 1 : 1);
    double dzg = Lz / (double)Nz;
#pragma omp parallel for schedule(dynamic, 4)
    for (int ir = 0; ir < Nrho; ++ir) {
        double rho = (double)ir * drho;
        for (int iz = 0; iz < Nz; ++iz) {
            double z = ((double)iz + 0.5) * dzg - 0.5 * Lz;
            double Fp[3];
            oracle_s1r2_force_pair(rho, 0.0, z, Lz, alpha, nmax, mmax, rspace, Fp);
            double dzw = wrap_dz(z, Lz);
            double r2 = rho * rho + dzw * dzw;
            double Frn = 0, Fzn = 0;
            if (r2 > 0.0) {
                double r = sqrt(r2), invr3 = 1.0 / (r * r * r);
                Frn = -rho * invr3;
                Fzn = -dzw * invr3;
            }
            T[((size_t)ir * Nz + iz) * 2 + 0] = Fp[0] - Frn;
            T[((size_t)ir * Nz + iz) * 2 + 1] = Fp[2] - Fzn;
        }
    }
}

/* main.cc:574-628 -- S1R2 table parameters (rspace: PERIODIC_Z_RSPACELOOKUP) */
void oracle_s1r2_params(int is_periodic, double L, double Rsim, int rspace, double *alpha,
                        int *rel_cut, int *rec_cut, int *nz, int *nrho)
{
    if (!rspace) {
        if (is_periodic == 2) { *rel_cut = 4; *rec_cut = 10; *alpha = 0.787875 / L; *nz = 128; }
        else if (is_periodic == 3) { *rel_cut = 5; *rec_cut = 12; *alpha = 0.71805 / L; *nz = 256; }
        else { *rel_cut = is_periodic + 2; *rec_cut = is_periodic + 9; *alpha = 0.6642 / L; *nz = 512; }
    } else {
        *alpha = 0.0;
        *rec_cut = 0;
        if (is_periodic == 2) { *rel_cut = 1000; *nz = 128; }
        else if (is_periodic == 3) { *rel_cut = 10000; *nz = 256; }
        else { *rel_cut = 100000; *nz = 512; }
    }
    *nrho = (int)floor(((double)*nz) * ORC_RADIAL_EXTENT * Rsim / L);
}

/* forces_cuda.cu:188-223 NGP, :225-283 CIC, :285-418 TSC, :420-455 Cartesian map */
void oracle_s1r2_interpolate(const double *T, int Nrho, int Nz, double rho_max, double Lz,
                             double dx, double dy, double dz, int order, double D[3])
{
    double rho = sqrt(dx * dx + dy * dy);
    double Dr = 0, Dzv = 0;
    const double half = 0.5 * Lz;
    const double drho = rho_max / (double)(Nrho - 1 > 1 ? Nrho - 1 : 1);
    const double dzg = Lz / (double)Nz;
#define T2(ir, iz, c) T[((size_t)(ir) * (size_t)Nz + (size_t)(iz)) * 2u + (size_t)(c)]
    if (order == 0) {
        double rc = rho;
        if (rho < 0) rc = 0;
        if (rho > rho_max) rc = rho_max;
        double ur = (drho > 0) ? (rc / drho) : 0;
        double uz = (dz + half) / dzg - 0.5;
        int ir = (int)floor(ur + 0.5);
        int iz = (int)floor(uz + 0.5);
        if (ir < 0) ir = 0;
        if (ir > Nrho - 1) ir = Nrho - 1;
        iz = imodp(iz, Nz);
        Dr = T2(ir, iz, 0);
        Dzv = T2(ir, iz, 1);
    } else if (order == 2) {
        double ur;
        if (rho < 0) ur = 0;
        else if (rho > rho_max) ur = (double)(Nrho - 1);
        else ur = rho / drho;
        int ir0 = (int)floor(ur);
        double fr = ur - (double)ir0;
        if (ir0 < 0) { ir0 = 0; fr = 0; }
        if (ir0 > Nrho - 2) { ir0 = (Nrho - 2 > 0) ? Nrho - 2 : 0; fr = 1.0; }
        int ir1 = ir0 + 1;
        double uz = (dz + half) / dzg - 0.5;
        int iz0 = (int)floor(uz);
        double fz = uz - (double)iz0;
        iz0 = imodp(iz0, Nz);
        int iz1 = imodp(iz0 + 1, Nz);
        double w00 = (1.0 - fr); w00 *= (1.0 - fz);
        double w10 = fr; w10 *= (1.0 - fz);
        double w01 = (1.0 - fr); w01 *= fz;
        double w11 = fr * fz;
        Dr = w00 * T2(ir0, iz0, 0) + w10 * T2(ir1, iz0, 0) + w01 * T2(ir0, iz1, 0) + w11 * T2(ir1, iz1, 0);
        Dzv = w00 * T2(ir0, iz0, 1) + w10 * T2(ir1, iz0, 1) + w01 * T2(ir0, iz1, 1) + w11 * T2(ir1, iz1, 1);
    } else {
        double ur;
        if (rho < 0) ur = 0;
        else if (rho > rho_max) ur = (double)(Nrho - 1);
        else ur = rho / drho;
        double uz = (dz + half) / dzg - 0.5;
        int jr = (int)floor(ur + 0.5), jz = (int)floor(uz + 0.5);
        double sr = ur - (double)jr, sz = uz - (double)jz;
        double wr[3] = {0.5 * (0.5 - sr) * (0.5 - sr), 0.75 - sr * sr, 0.5 * (0.5 + sr) * (0.5 + sr)};
        double wz[3] = {0.5 * (0.5 - sz) * (0.5 - sz), 0.75 - sz * sz, 0.5 * (0.5 + sz) * (0.5 + sz)};
        int ir0 = jr - 1; if (ir0 < 0) ir0 = 0;
        int ir1 = jr; if (ir1 < 0) ir1 = 0; if (ir1 > Nrho - 1) ir1 = Nrho - 1;
        int ir2 = jr + 1; if (ir2 > Nrho - 1) ir2 = Nrho - 1;
        int irs[3] = {ir0, ir1, ir2};
        int izs[3] = {imodp(jz - 1, Nz), imodp(jz, Nz), imodp(jz + 1, Nz)};
        for (int row = 0; row < 3; ++row) {
            double w0 = wr[0] * wz[row], w1 = wr[1] * wz[row], w2 = wr[2] * wz[row];
            Dr += w0 * T2(irs[0], izs[row], 0) + w1 * T2(irs[1], izs[row], 0) + w2 * T2(irs[2], izs[row], 0);
            Dzv += w0 * T2(irs[0], izs[row], 1) + w1 * T2(irs[1], izs[row], 1) + w2 * T2(irs[2], izs[row], 1);
        }
    }
#undef T2
    double ex = (rho > 0) ? dx / rho : 0.0;
    double ey = (rho > 0) ? dy / rho : 0.0;
    D[0] = Dr * ex;
    D[1] = Dr * ey;
    D[2] = Dzv;
}

/* ------------------------------------------------------------------------ */
/* cylindrical radial table                                                  */
/* ------------------------------------------------------------------------ */

/* utils.cc:24-37 */
static double lin_interp(double X, double X1, double Y1, double X2, double Y2)
{
    double A = (Y2 - Y1) / (X2 - X1);
    double B = Y1 - A * X1;
    return A * X + B;
}

/* utils.cc:162-228 -- trapezoid integral of the finite-cylinder radial pull ratio */
void oracle_cylindrical_force_table(double *T, double R, double Lz, int size, int acc)
{
    double step = R / (double)acc;
    T[0] = 0.0;
    for (int i = 1; i < size; i++) {
        double a = (i == size - 1) ? R : R / (double)(size - 1) * i;
        double total = 0;
        for (int j = 1; j <= acc; j++) {
            double ip, in;
            if (a >= R) {
                a = R;
                double Y = (j - 1) * step;
                ip = -2 * log((pow((pow(Lz, 2) + 2 * pow(R, 2) + 2 * R * pow((pow(R, 2) - pow(Y, 2)), 0.5)), 0.5) + Lz) /
                              (pow((pow(Lz, 2) + 2 * pow(R, 2) - 2 * R * pow((pow(R, 2) - pow(Y, 2)), 0.5)), 0.5) + Lz));
                Y = j * step;
                in = -2 * log((pow((pow(Lz, 2) + 2 * pow(R, 2) + 2 * R * pow((pow(R, 2) - pow(Y, 2)), 0.5)), 0.5) + Lz) /
                              (pow((pow(Lz, 2) + 2 * pow(R, 2) - 2 * R * pow((pow(R, 2) - pow(Y, 2)), 0.5)), 0.5) + Lz));
            } else {
                double Y = (j - 1) * step;
                double s = pow((pow(R, 2) - pow(Y, 2)), 0.5);
                double f1 = -2 * log((pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) + 2 * a * s), 0.5) + Lz) /
                                     (pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) - 2 * a * s), 0.5) + Lz));
                double f2 = log((pow(R, 2) + pow(a, 2) + 2 * a * s) / (pow(R, 2) + pow(a, 2) - 2 * a * s));
                ip = f1 + f2;
                Y = j * step;
                s = pow((pow(R, 2) - pow(Y, 2)), 0.5);
                f1 = -2 * log((pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) + 2 * a * s), 0.5) + Lz) /
                              (pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) - 2 * a * s), 0.5) + Lz));
                f2 = log((pow(R, 2) + pow(a, 2) + 2 * a * s) / (pow(R, 2) + pow(a, 2) - 2 * a * s));
                in = f1 + f2;
            }
            total += step * (ip + in) * 0.5;
        }
        if (a >= R)
            T[i] = 2 * (total + ORC_PI * R) / (2 * ORC_PI * R);
        else
            T[i] = 2 * total / (2 * ORC_PI * a);
    }
    if (size >= 3)
        T[0] = lin_interp(0, 2 * R / (double)size, T[1], 4 * R / (double)size, T[2]);
}

/* forces_cuda.cu:59-70 (linear only on the GPU; forces.cc:279-316 order 1 on CPU) */
double oracle_cylindrical_force_correction(double r, double R, const double *T, int size)
{
    double step = R / (double)size;
    int i = (int)floor(r / R * (size - 1));
    double c = T[size - 1];
    if (i < size - 1)
        c = lin_interp(r, step * i, T[i], step * (i + 1), T[i + 1]);
    return c;
}

/* utils.cc:39-56: Lagrange interpolation through three / four nodes */
static double quad_interp(double X, double X1, double Y1, double X2, double Y2, double X3, double Y3)
{
    double L1 = (X - X2) * (X - X3) / ((X1 - X2) * (X1 - X3));
    double L2 = (X - X1) * (X - X3) / ((X2 - X1) * (X2 - X3));
    double L3 = (X - X1) * (X - X2) / ((X3 - X1) * (X3 - X2));
    return Y1 * L1 + Y2 * L2 + Y3 * L3;
}

static double cubic_interp(double X, double X1, double Y1, double X2, double Y2, double X3, double Y3,
                           double X4, double Y4)
{
    double L1 = (X - X2) * (X - X3) * (X - X4) / ((X1 - X2) * (X1 - X3) * (X1 - X4));
    double L2 = (X - X1) * (X - X3) * (X - X4) / ((X2 - X1) * (X2 - X3) * (X2 - X4));
    double L3 = (X - X1) * (X - X2) * (X - X4) / ((X3 - X1) * (X3 - X2) * (X3 - X4));
    double L4 = (X - X1) * (X - X2) * (X - X3) / ((X4 - X1) * (X4 - X2) * (X4 - X3));
    return Y1 * L1 + Y2 * L2 + Y3 * L3 + Y4 * L4;
}

/* forces.cc:279-316: the CPU correction of order 1, 2 or 3 (nodes assumed at
 * R/size * i, as on the GPU).  Order 2 shifts its three nodes down by one at
 * i == size-2; order 3 shifts its four nodes down by one at i == size-3 and
 * returns the last entry from i == size-2 on.  Every force path of the
 * reference calls order 1 (forces.cc:1136,1177,1296,1388). */
double oracle_cylindrical_force_correction_order(double r, double R, const double *T, int size, int order)
{
    double step = R / (double)size;
    int i = (int)floor(r / R * (size - 1));
    double c = T[size - 1];
    if (order == 1) {
        if (i < size - 1) c = lin_interp(r, step * i, T[i], step * (i + 1), T[i + 1]);
    } else if (order == 2) {
        if (i < size - 2)
            c = quad_interp(r, step * i, T[i], step * (i + 1), T[i + 1], step * (i + 2), T[i + 2]);
        else if (i == size - 2)
            c = quad_interp(r, step * (i - 1), T[i - 1], step * i, T[i], step * (i + 1), T[i + 1]);
    } else if (order == 3) {
        if (i < size - 3)
            c = cubic_interp(r, step * i, T[i], step * (i + 1), T[i + 1], step * (i + 2), T[i + 2],
                             step * (i + 3), T[i + 3]);
        else if (i == size - 3)
            c = cubic_interp(r, step * (i - 1), T[i - 1], step * i, T[i], step * (i + 1), T[i + 1],
                             step * (i + 2), T[i + 2]);
    }
    return c;
}

/* ------------------------------------------------------------------------ */
/* direct sums                                                               */
/* ------------------------------------------------------------------------ */

static void pair_row(const oracle_params *p, int N, const double *x, const double *M,
                     const double *h, int i, double out[3])
{
    const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2], hi = h[i];
    const double L = p->L;
    const double DE = p->H0 * p->H0 * p->Omega_lambda; /* forces_cuda.cu:887 */
    double fx = 0, fy = 0, fz = 0;
    if (p->topology == ORC_R3) {
        /* forces_cuda.cu:531-544 / forces.cc:538-556 */
        for (int j = 0; j < N; j++) {
            double beta = hi + h[j];
            double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            double w = M[j] * oracle_force_softening(r, beta);
            fx += w * dx; fy += w * dy; fz += w * dz;
        }
        /* forces_cuda.cu:545-556 */
        if (p->cosmology == 1 && p->comoving == 1) {
            fx += p->mass_in_unit_sphere * xi;
            fy += p->mass_in_unit_sphere * yi;
            fz += p->mass_in_unit_sphere * zi;
        } else if (p->cosmology == 1 && p->comoving == 0) {
            fx += DE * xi; fy += DE * yi; fz += DE * zi;
        }
    } else if (p->topology == ORC_T3) {
        /* forces_cuda.cu:575-643 */
        for (int j = 0; j < N; j++) {
            double beta = hi + h[j];
            double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
            if (fabs(dx) > 0.5 * L) dx = dx - L * dx / fabs(dx);
            if (fabs(dy) > 0.5 * L) dy = dy - L * dy / fabs(dy);
            if (fabs(dz) > 0.5 * L) dz = dz - L * dz / fabs(dz);
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            if (p->is_periodic >= 2) {
                double w = oracle_force_softening(r, beta), D[3];
                oracle_t3_interpolate(p->t3_ngrid, L, p->t3_table, dx, dy, dz, D, p->interp_order);
                fx += M[j] * (w * dx - D[0]);
                fy += M[j] * (w * dy - D[1]);
                fz += M[j] * (w * dz - D[2]);
            } else {
                double w = M[j] * oracle_force_softening(r, beta);
                fx += w * dx; fy += w * dy; fz += w * dz;
            }
        }
    } else {
        if (p->is_periodic >= 2 && !p->s1r2_lookup) {
            /* forces_cuda.cu:711-759 (no-lookup image sum) */
            int emax = p->is_periodic + 1;
            double ecut = (((double)emax) - 0.4) * L;
            for (int j = 0; j < N; j++) {
                double beta = hi + h[j];
                double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
                for (int m = -emax; m < emax + 1; m++) {
                    double dze = dz + ((double)m) * L;
                    double r = sqrt(dx * dx + dy * dy + dze * dze);
                    int inside = p->cut_inclusive ? (fabs(dze) <= ecut) : (fabs(dze) < ecut);
                    if (inside) {
                        double w = M[j] * oracle_force_softening(r, beta);
                        fx += w * dx; fy += w * dy; fz += w * dze;
                    }
                }
            }
        } else {
            /* forces_cuda.cu:772-863 */
            for (int j = 0; j < N; j++) {
                double beta = hi + h[j];
                double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
                if (dz > 0.5 * L) dz -= L;
                else if (dz < -0.5 * L) dz += L;
                double r = sqrt(dx * dx + dy * dy + dz * dz);
                if (p->is_periodic >= 2) {
                    double w = oracle_force_softening(r, beta), D[3];
                    oracle_s1r2_interpolate(p->s1r2_table, p->s1r2_nrho, p->s1r2_nz,
                                            ORC_RADIAL_EXTENT * p->Rsim, L, dx, dy, dz,
                                            p->interp_order, D);
                    fx += M[j] * (w * dx - D[0]);
                    fy += M[j] * (w * dy - D[1]);
                    fz += M[j] * (w * dz - D[2]);
                } else {
                    double w = M[j] * oracle_force_softening(r, beta);
                    fx += w * dx; fy += w * dy; fz += w * dz;
                }
            }
        }
        /* external term in x,y only: lookup with IS_PERIODIC>=2 is uncorrected
         * (forces_cuda.cu:848-857); otherwise scaled by the radial table
         * (forces_cuda.cu:697-707, :741-751) */
        if (p->cosmology == 1 && p->comoving == 1) {
            double corr = 1.0;
            if (p->is_periodic == 1 || !p->s1r2_lookup) {
                double rxy = sqrt(xi * xi + yi * yi);
                corr = oracle_cylindrical_force_correction(rxy, p->Rsim, p->radial_table,
                                                           p->radial_table_size);
            }
            fx += p->mass_in_unit_sphere * xi * corr;
            fy += p->mass_in_unit_sphere * yi * corr;
        } else if (p->cosmology == 1 && p->comoving == 0) {
            fx += DE * xi;
            fy += DE * yi;
        }
    }
    out[0] = fx;
    out[1] = fy;
    out[2] = fz;
}

static int check_args(const oracle_params *p, int N, const double *x, const double *M,
                      const double *h)
{
    if (!p || N < 0 || (N > 0 && (!x || !M || !h))) return -1;
    if (p->topology == ORC_T3 && p->is_periodic >= 2 && (!p->t3_table || p->t3_ngrid < 1)) return -2;
    if (p->topology == ORC_S1R2 && p->is_periodic >= 2 && p->s1r2_lookup &&
        (!p->s1r2_table || p->s1r2_nrho < 1 || p->s1r2_nz < 1)) return -2;
    if (p->topology == ORC_S1R2 && p->cosmology == 1 && p->comoving == 1 &&
        (p->is_periodic == 1 || !p->s1r2_lookup) && (!p->radial_table || p->radial_table_size < 2))
        return -2;
    return 0;
}

int oracle_forces(const oracle_params *p, int N, const double *x, const double *M,
                  const double *h, double *F, int id_min, int id_max)
{
    int rc = check_args(p, N, x, M, h);
    if (rc) return rc;
    if (id_min < 0 || id_max >= N || id_max < id_min - 1) return -3;
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = id_min; i <= id_max; i++)
        pair_row(p, N, x, M, h, i, F + 3 * (size_t)(i - id_min));
    return 0;
}

int oracle_forces_rows(const oracle_params *p, int N, const double *x, const double *M,
                       const double *h, double *F, const int *rows, int nrows)
{
    int rc = check_args(p, N, x, M, h);
    if (rc) return rc;
    for (int k = 0; k < nrows; k++)
        if (rows[k] < 0 || rows[k] >= N) return -3;
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nrows; k++)
        pair_row(p, N, x, M, h, rows[k], F + 3 * (size_t)k);
    return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

/* The reference's single-precision GPU build (USE_SINGLE_PRECISION,
 * global_variables.h:26-32, set by Template-LinuxICC-Makefile:24) for R3:
 * ForceKernel (forces_cuda.cu:522-563) with REAL = float.  Positions, masses
 * and softening lengths are stored as float, and every operation, the
 * softening law (forces_cuda.cu:466-519) and the accumulation over
 * j = 0..N-1 run in float with no compensation.  A statistical stand-in for
 * that build, not a bitwise one: the fused multiply-adds nvcc contracts by
 * default are modelled as fmaf in r^2 and in the accumulation; sqrt and the
 * divisions are IEEE float. */
static float softening_f32(float r, float beta)
{
    const float r3 = (r * r) * r;
    const float inv_r3 = 1.0f / r3;
    if (r >= beta) return inv_r3;
    const float b2 = beta * beta, b3 = b2 * beta, b4 = b2 * b2, b5 = b4 * beta, b6 = b3 * b3;
    if (r > 0.5f * beta)
        return (-32.0f / (3.0f * b6)) * r3 + (38.4f / b5) * (r * r) + (-48.0f / b4) * r +
               64.0f / (3.0f * b3) + (float)(-1.0 / 15.0) * inv_r3;
    return (32.0f / b6) * r3 + (-38.4f / b5) * (r * r) + (float)(32.0 / 3.0) / b3;
}

int oracle_forces_r3_fp32_rows(const oracle_params *p, int N, const double *x, const double *M,
                               const double *h, double *F, const int *rows, int nrows)
{
    int rc = check_args(p, N, x, M, h);
    if (rc) return rc;
    if (p->topology != ORC_R3) return -4;
    for (int k = 0; k < nrows; k++)
        if (rows[k] < 0 || rows[k] >= N) return -3;
    const float mius = (float)p->mass_in_unit_sphere;
    const float DE = (float)(p->H0 * p->H0 * p->Omega_lambda);
#pragma omp parallel for schedule(dynamic, 4)
    for (int k = 0; k < nrows; k++) {
        const int i = rows[k];
        const float xi = (float)x[3 * i], yi = (float)x[3 * i + 1], zi = (float)x[3 * i + 2];
        const float hi = (float)h[i];
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        for (int j = 0; j < N; j++) {
            const float dx = (float)x[3 * j] - xi, dy = (float)x[3 * j + 1] - yi,
                        dz = (float)x[3 * j + 2] - zi;
            const float r = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
            const float w = (float)M[j] * softening_f32(r, hi + (float)h[j]);
            fx = fmaf(w, dx, fx);
            fy = fmaf(w, dy, fy);
            fz = fmaf(w, dz, fz);
        }
        if (p->cosmology == 1 && p->comoving == 1) {
            fx = fmaf(mius, xi, fx);
            fy = fmaf(mius, yi, fy);
            fz = fmaf(mius, zi, fz);
        } else if (p->cosmology == 1 && p->comoving == 0) {
            fx = fmaf(DE, xi, fx);
            fy = fmaf(DE, yi, fy);
            fz = fmaf(DE, zi, fz);
        }
        F[3 * (size_t)k] = fx;
        F[3 * (size_t)k + 1] = fy;
        F[3 * (size_t)k + 2] = fz;
    }
    return 0;
}

/* long-double brute force of the R3 sum (external term included). */
void oracle_forces_r3_longdouble(const oracle_params *p, int N, const double *x,
                                 const double *M, const double *h, long double *F)
{
    for (int i = 0; i < N; i++) {
        long double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < N; j++) {
            long double dx = (long double)x[3 * j] - x[3 * i];
            long double dy = (long double)x[3 * j + 1] - x[3 * i + 1];
            long double dz = (long double)x[3 * j + 2] - x[3 * i + 2];
            long double r = sqrtl(dx * dx + dy * dy + dz * dz);
            long double b = (long double)h[i] + (long double)h[j];
            long double w;
            if (r >= b) w = 1.0L / (r * r * r);
            else {
                long double u = r / b;
                if (u > 0.5L)
                    w = (-32.0L / 3.0L * u * u * u + 38.4L * u * u - 48.0L * u + 64.0L / 3.0L -
                         1.0L / (15.0L * u * u * u)) / (b * b * b);
                else
                    w = (32.0L * u * u * u - 38.4L * u * u + 32.0L / 3.0L) / (b * b * b);
            }
            w *= (long double)M[j];
            fx += w * dx; fy += w * dy; fz += w * dz;
        }
        if (p->cosmology == 1 && p->comoving == 1) {
            fx += (long double)p->mass_in_unit_sphere * x[3 * i];
            fy += (long double)p->mass_in_unit_sphere * x[3 * i + 1];
            fz += (long double)p->mass_in_unit_sphere * x[3 * i + 2];
        } else if (p->cosmology == 1 && p->comoving == 0) {
            long double DE = (long double)p->H0 * p->H0 * p->Omega_lambda;
            fx += DE * x[3 * i]; fy += DE * x[3 * i + 1]; fz += DE * x[3 * i + 2];
        }
        F[3 * i] = fx; F[3 * i + 1] = fy; F[3 * i + 2] = fz;
    }
}

/* ------------------------------------------------------------------------ */
/* softening set-up, cosmology, KDK                                          */
/* ------------------------------------------------------------------------ */

/* utils.cc:59-82 */
void oracle_softening_lengths(const double *M, int N, double R0, double *h, double *M_min,
                              double *rho_part)
{
    double mmin = M[0];
    for (int i = 0; i < N; i++)
        if (mmin > M[i]) mmin = M[i];
    double rp = mmin / (4.0 * ORC_PI * pow(R0, 3.0) / 3.0);
    double cb = 3.0 / rp / (4.0 * ORC_PI);
    for (int i = 0; i < N; i++)
        h[i] = cbrt(M[i] * cb);
    *M_min = mmin;
    *rho_part = rp;
}

/* friedmann_solver.cc:100-159 (COSMOPARAM 0, RK4 in a) */
double oracle_friedmann_step(const oracle_cosmo *c, double a0, double h)
{
    const double Om = c->Omega_m, Or = c->Omega_r, OL = c->Omega_lambda, Ok = c->Omega_k, H0 = c->H0;
    double b = a0;
    int collapse = H0 > 0 ? 0 : 1;
    if (fabs(Ok) < 1e-9) {
        double j = Om * pow(b, -3.0) + Or * pow(b, -4.0) + OL;
        double k1 = b * H0 * sqrt(Om * pow(b, -3.0) + Or * pow(b, -4.0) + OL);
        double l = Om * pow((b + h * k1 / 2.0), -3.0) + Or * pow((b + h * k1 / 2), -4.0) + OL;
        double k2 = (b + h * k1 / 2.0) * H0 * sqrt(Om * pow((b + h * k1 / 2.0), -3.0) + Or * pow((b + h * k1 / 2), -4.0) + OL);
        double m = Om * pow((b + h * k2 / 2.0), -3.0) + Or * pow((b + h * k2 / 2), -4.0) + OL;
        double k3 = (b + h * k2 / 2.0) * H0 * sqrt(Om * pow((b + h * k2 / 2.0), -3.0) + Or * pow((b + h * k2 / 2), -4.0) + OL);
        double n = Om * pow((b + h * k3), -3.0) + Or * pow((b + h * k3), -4.0) + OL;
        double k4 = (b + h * k3) * H0 * sqrt(Om * pow((b + h * k3), -3.0) + Or * pow((b + h * k3), -4.0) + OL);
        double K = h * (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
        if (j < 0 || l < 0 || m < 0 || n < 0) b = -1;
        else b += K;
        (void)collapse;
    } else {
        double j = Om * pow(b, -3.0) + Or * pow(b, -4.0) + OL + Ok * pow(b, -2.0);
        double k1 = b * H0 * sqrt(fabs(j));
        double bb = b + h * k1 / 2.0;
        double l = Om * pow(bb, -3.0) + Or * pow(bb, -4.0) + OL + Ok * pow(bb, -2.0);
        double k2 = bb * H0 * sqrt(fabs(l));
        bb = b + h * k2 / 2.0;
        double m = Om * pow(bb, -3.0) + Or * pow(bb, -4.0) + OL + Ok * pow(bb, -2.0);
        double k3 = bb * H0 * sqrt(fabs(m));
        bb = b + h * k3;
        double n = Om * pow(bb, -3.0) + Or * pow(bb, -4.0) + OL + Ok * pow(bb, -2.0);
        double k4 = bb * H0 * sqrt(fabs(n));
        if (j < 0 && l < 0 && m < 0 && n < 0) collapse = 1;
        double K = h * (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
        b = (collapse == 0) ? b + K : b - K;
    }
    return b;
}

/* friedmann_solver.cc:161-164 */
double oracle_hubble(const oracle_cosmo *c, double a)
{
    return c->H0 * sqrt(c->Omega_m * pow(a, -3) + c->Omega_r * pow(a, -4) + c->Omega_lambda +
                        c->Omega_k * pow(a, -2));
}

static void wrap_positions(const oracle_params *p, int N, double *x, int else_branch)
{
    /* step.cc:151-180 (else-if form) and step.cc:40-67 (two ifs) */
    if (p->topology == ORC_T3) {
        for (int i = 0; i < 3 * N; i++) {
            if (x[i] < 0) x[i] += p->L;
            else if (else_branch && x[i] >= p->L) x[i] -= p->L;
            if (!else_branch && x[i] >= p->L) x[i] -= p->L;
        }
    } else if (p->topology == ORC_S1R2) {
        for (int i = 0; i < N; i++) {
            double *z = &x[3 * i + 2];
            if (*z < 0) *z += p->L;
            else if (else_branch && *z >= p->L) *z -= p->L;
            if (!else_branch && *z >= p->L) *z -= p->L;
        }
    }
}

/* step.cc:35-98 -- errmax for the initial step length (wraps x like the reference) */
double oracle_initial_errmax(const oracle_params *p, oracle_step_state *s, int N, double *x,
                             const double *v, const double *F, const double *M)
{
    wrap_positions(p, N, x, 0);
    double cb = 3.0 / s->rho_part / (4.0 * ORC_PI);
    double errmax = 0;
    double ainv3 = pow(s->a, -3.0);
    for (int i = 0; i < N; i++) {
        double A[3];
        for (int k = 0; k < 3; k++)
            A[k] = s->G * F[3 * i + k] * ainv3 - 2.0 * s->hubble * v[3 * i + k];
        double err = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) / cbrt(M[i] * cb);
        if (err > errmax) errmax = err;
    }
    return errmax;
}

/* step.cc:100-312 -- kick(h/2), drift(h), wrap, force, a <- RK4, kick(h/2), errmax */
double oracle_kdk_step(const oracle_params *p, const oracle_cosmo *c, oracle_step_state *s,
                       int N, double *x, double *v, double *F, const double *M,
                       const double *hsoft)
{
    const double h = s->h;
    double ainv3 = pow(s->a, -3.0);
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            double A = s->G * F[3 * i + k] * ainv3 - 2.0 * s->hubble * v[3 * i + k];
            v[3 * i + k] += A * (h / 2.0);
            x[3 * i + k] = x[3 * i + k] + v[3 * i + k] * h;
        }
    wrap_positions(p, N, x, 1);
    oracle_forces(p, N, x, M, hsoft, F, 0, N - 1);
    if (p->cosmology == 1 && p->comoving == 1) {
        s->a = oracle_friedmann_step(c, s->a, h);
        s->hubble = oracle_hubble(c, s->a);
    }
    ainv3 = pow(s->a, -3.0);
    double cb = 3.0 / s->rho_part / (4.0 * ORC_PI);
    double errmax = 0;
    for (int i = 0; i < N; i++) {
        double A[3];
        for (int k = 0; k < 3; k++) {
            A[k] = s->G * F[3 * i + k] * ainv3 - 2.0 * s->hubble * v[3 * i + k];
            v[3 * i + k] += A[k] * (h / 2.0);
        }
        double err = sqrt(A[0] * A[0] + A[1] * A[1] + A[2] * A[2]) / cbrt(M[i] * cb);
        if (err > errmax) errmax = err;
    }
    return errmax;
}

/* ------------------------------------------------------------------------ */
/* Barnes-Hut octree (forces.cc:97-187, 321-506)                            */
/* ------------------------------------------------------------------------ */
typedef struct {
    double cx, cy, cz, size;   /* cube centre and side */
    double mass, com[3];
    int pidx;                  /* particle of a leaf, -1 for internal or empty */
    int child[8];              /* -1: none */
} orc_bh_node;

typedef struct {
    orc_bh_node *n;
    int count, cap;
} orc_bh_tree;

static int orc_bh_new(orc_bh_tree *t, double cx, double cy, double cz, double size)
{
    if (t->count == t->cap) {
        int cap = t->cap ? 2 * t->cap : 1024;
        orc_bh_node *m = (orc_bh_node *)realloc(t->n, sizeof(orc_bh_node) * (size_t)cap);
        if (!m) return -1;
        t->n = m;
        t->cap = cap;
    }
    orc_bh_node *q = &t->n[t->count];
    q->cx = cx; q->cy = cy; q->cz = cz; q->size = size;
    q->mass = 0; q->com[0] = q->com[1] = q->com[2] = 0;
    q->pidx = -1;
    for (int k = 0; k < 8; ++k) q->child[k] = -1;
    return t->count++;
}

/* forces.cc:121-128: x > cx sets bit 0, y > cy bit 1, z > cz bit 2 */
static int orc_bh_octant(const orc_bh_node *q, const double *x, int i)
{
    return (x[3 * i] > q->cx ? 1 : 0) | (x[3 * i + 1] > q->cy ? 2 : 0) | (x[3 * i + 2] > q->cz ? 4 : 0);
}

/* forces.cc:130-174 (recursive insertion; the incremental centre of mass) */
static int orc_bh_insert(orc_bh_tree *t, int node, const double *x, const double *M, int i, int depth)
{
    if (depth > 64) return -5;
    orc_bh_node *q = &t->n[node];
    if (q->mass == 0 && q->pidx == -1) {
        q->pidx = i;
        q->mass = M[i];
        q->com[0] = x[3 * i]; q->com[1] = x[3 * i + 1]; q->com[2] = x[3 * i + 2];
        return 0;
    }
    if (q->pidx != -1) {
        int existing = q->pidx;
        q->pidx = -1;
        for (int j = 0; j < 8; ++j) {
            q = &t->n[node];
            double off = q->size / 4;
            int c = orc_bh_new(t, q->cx + ((j & 1) ? off : -off), q->cy + ((j & 2) ? off : -off),
                               q->cz + ((j & 4) ? off : -off), q->size / 2);
            if (c < 0) return -1;
            t->n[node].child[j] = c;
        }
        q = &t->n[node];
        int rc = orc_bh_insert(t, q->child[orc_bh_octant(q, x, existing)], x, M, existing, depth + 1);
        if (rc) return rc;
    }
    q = &t->n[node];
    int rc = orc_bh_insert(t, q->child[orc_bh_octant(q, x, i)], x, M, i, depth + 1);
    if (rc) return rc;
    q = &t->n[node];
    double total = q->mass + M[i];
    for (int a = 0; a < 3; ++a) q->com[a] = (q->com[a] * q->mass + x[3 * i + a] * M[i]) / total;
    q->mass = total;
    return 0;
}

/* The pair term of one traversal: the topology's image choice and table
 * correction (forces.cc:373-398 R3; :584-652 T3 Ewald and nearest image;
 * :903-981 S1R2 image sum, Ewald table and nearest z-image). */
typedef struct orc_bh_ctx {
    const oracle_params *p;
    const orc_bh_tree *t;
    const double *x, *h;
    double theta, pr, M_min;
    int mode;            /* 0 R3, 1 T3 nearest, 2 T3 Ewald, 3 S1R2 nearest, 4 S1R2 table, 5 S1R2 images */
    double zi_image;     /* mode 5: the row's z plus m L */
    double cut;          /* mode 5: ewald_cut * L */
    long long cnt[2];
} orc_bh_ctx;

static void orc_bh_force(orc_bh_ctx *c, int node, int i, double *f)
{
    if (node < 0) return;
    const orc_bh_node *q = &c->t->n[node];
    if (q->mass == 0 || q->pidx == i) return;
    const double *x = c->x;
    const double L = c->p->L;
    double dx = q->com[0] - x[3 * i], dy = q->com[1] - x[3 * i + 1];
    double dz = q->com[2] - (c->mode == 5 ? c->zi_image : x[3 * i + 2]);
    if (c->mode == 1 || c->mode == 2) { /* forces.cc:594-599 */
        if (fabs(dx) > 0.5 * L) dx = dx - L * dx / fabs(dx);
        if (fabs(dy) > 0.5 * L) dy = dy - L * dy / fabs(dy);
    }
    if (c->mode >= 1 && c->mode <= 4) /* forces.cc:600-601, 941-942 */
        if (fabs(dz) > 0.5 * L) dz = dz - L * dz / fabs(dz);
    double dist = sqrt(dx * dx + dy * dy + dz * dz);
    c->cnt[0]++;
    if (q->pidx != -1 || q->size / dist < c->theta) {
        if (c->mode == 5 && fabs(dz) > c->cut) return; /* forces.cc:915 */
        double beta = cbrt(q->mass / c->M_min) * c->pr + c->h[i];
        double w = oracle_force_softening(dist, beta);
        double D[3] = {0, 0, 0};
        if (c->mode == 2)
            oracle_t3_interpolate(c->p->t3_ngrid, L, c->p->t3_table, dx, dy, dz, D, c->p->interp_order);
        else if (c->mode == 4)
            oracle_s1r2_interpolate(c->p->s1r2_table, c->p->s1r2_nrho, c->p->s1r2_nz, ORC_RADIAL_EXTENT * c->p->Rsim,
                                    L, dx, dy, dz, c->p->interp_order, D);
        if (c->mode == 2 || c->mode == 4) { /* forces.cc:607-610, 948-951 */
            f[0] += q->mass * (w * dx - D[0]);
            f[1] += q->mass * (w * dy - D[1]);
            f[2] += q->mass * (w * dz - D[2]);
        } else {
            w *= q->mass;
            f[0] += w * dx; f[1] += w * dy; f[2] += w * dz;
        }
        c->cnt[1]++;
    } else {
        for (int j = 0; j < 8; ++j) orc_bh_force(c, q->child[j], i, f);
    }
}

/* Node masses and centres of mass summed from the children, bottom-up: the
 * running averages of forces.cc:164-170 replaced by sums over the same tree
 * (oracle_bh_opts.com_from_children). */
static void orc_bh_com_from_children(orc_bh_tree *t, int node)
{
    orc_bh_node *q = &t->n[node];
    if (q->pidx >= 0) return;
    double sm = 0, sx = 0, sy = 0, sz = 0;
    for (int j = 0; j < 8; ++j) {
        const int c = q->child[j];
        if (c < 0) continue;
        orc_bh_com_from_children(t, c);
        const orc_bh_node *k = &t->n[c];
        sx += k->mass * k->com[0];
        sy += k->mass * k->com[1];
        sz += k->mass * k->com[2];
        sm += k->mass;
    }
    if (sm > 0) {
        q->mass = sm;
        q->com[0] = sx / sm;
        q->com[1] = sy / sm;
        q->com[2] = sz / sm;
    }
}

static long long orc_bh_counts[2];
static double orc_bh_root[4];

void oracle_bh_last_counts(long long *visited, long long *accepted)
{
    *visited = orc_bh_counts[0];
    *accepted = orc_bh_counts[1];
}

void oracle_bh_last_root(double out[4])
{
    for (int k = 0; k < 4; ++k) out[k] = orc_bh_root[k];
}

/* utils.cc:24-37 */
static double orc_lin(double X, double X1, double Y1, double X2, double Y2)
{
    double A = (Y2 - Y1) / (X2 - X1);
    double B = Y1 - A * X1;
    return A * X + B;
}

/* forces.cc:182-187 */
static double orc_bin_center(double R, int b, int n) { return (((double)b) + 0.5) * (R / (double)n); }

/* forces.cc:189-245, ORDER 1 */
double oracle_radial_bh_correction(double r, double R, const double *t, int n)
{
    if (r < 0) return 0;
    if (r > R) return orc_lin(r, orc_bin_center(R, n - 2, n), t[n - 2], orc_bin_center(R, n - 1, n), t[n - 1]);
    int b1 = (int)floor(r / (R / (double)n));
    if (b1 > n - 1) b1 = n - 1; /* r within rounding of R */
    double c1 = orc_bin_center(R, b1, n);
    if (c1 <= r) {
        int b2;
        if (b1 < n - 1) {
            b2 = b1 + 1;
        } else {
            b1 -= 1;
            c1 = orc_bin_center(R, b1, n);
            b2 = b1 + 1;
        }
        return orc_lin(r, c1, t[b1], orc_bin_center(R, b2, n), t[b2]);
    }
    if (b1 > 1) return orc_lin(r, orc_bin_center(R, b1 - 1, n), t[b1 - 1], c1, t[b1]);
    return orc_lin(r, 0.0, 0.0, c1, t[b1]);
}

/* utils.cc:96-156 */
void oracle_radial_bh_table_accumulate(int topology, double Rsim, int size, int N, const double *x,
                                       const double *F, double *table, int *count)
{
    double *loc = (double *)calloc((size_t)size, sizeof(double));
    int *cnt = (int *)calloc((size_t)size, sizeof(int));
    if (!loc || !cnt) {
        free(loc);
        free(cnt);
        return;
    }
    for (int i = 0; i < N; i++) {
        double r, Fr;
        if (topology != ORC_S1R2) {
            r = sqrt(pow(x[3 * i], 2) + pow(x[3 * i + 1], 2) + pow(x[3 * i + 2], 2));
            Fr = (F[3 * i] * x[3 * i] + F[3 * i + 1] * x[3 * i + 1] + F[3 * i + 2] * x[3 * i + 2]) / r;
        } else {
            r = sqrt(pow(x[3 * i], 2) + pow(x[3 * i + 1], 2));
            Fr = (F[3 * i] * x[3 * i] + F[3 * i + 1] * x[3 * i + 1]) / r;
        }
        int b = (int)floor(r / (Rsim / (double)size));
        if (b >= size) b = size - 1;
        loc[b] += Fr;
        cnt[b]++;
    }
    for (int b = 0; b < size; b++) {
        table[b] += loc[b];
        count[b] += cnt[b];
    }
    free(loc);
    free(cnt);
}

/* forces.cc:324-356 (Rodrigues' formula about the unit vector a) */
static void orc_rotate3(double *v, const double *a, double c, double s)
{
    double X[3] = {v[0], v[1], v[2]};
    double cr[3] = {a[1] * X[2] - a[2] * X[1], a[2] * X[0] - a[0] * X[2], a[0] * X[1] - a[1] * X[0]};
    double dot = a[0] * X[0] + a[1] * X[1] + a[2] * X[2];
    for (int k = 0; k < 3; ++k) v[k] = X[k] * c + cr[k] * s + a[k] * dot * (1.0 - c);
}

/* forces.cc:884-899 */
static void orc_rotate2(double *v, double c, double s)
{
    double r0 = v[0] * c - v[1] * s, r1 = v[0] * s + v[1] * c;
    v[0] = r0;
    v[1] = r1;
}

/* forces.cc:670-682 */
static double orc_shift_wrap(double v, double d, double L)
{
    v += d;
    if (v < 0) v = v + L;
    else if (v >= L) v = v - L;
    return v;
}

int oracle_forces_bh(const oracle_params *p, double theta, double particle_radius, int N, const double *x,
                     const double *M, const double *h, double *F, int id_min, int id_max)
{
    return oracle_forces_bh_ex(p, theta, particle_radius, NULL, N, x, M, h, F, id_min, id_max);
}

int oracle_forces_bh_ex(const oracle_params *p, double theta, double particle_radius, const oracle_bh_opts *o,
                        int N, const double *x_in, const double *M, const double *h, double *F, int id_min,
                        int id_max)
{
    if (!p || N <= 0 || id_min < 0 || id_max >= N || id_max < id_min - 1) return -1;
    const int rnd = o && o->randomize;
    const double *rbh = (o && o->radial_bh_table && p->topology != ORC_T3 && p->cosmology == 1 && p->comoving == 1)
                            ? o->radial_bh_table
                            : NULL;
    int mode;
    if (p->topology == ORC_R3) mode = 0;
    else if (p->topology == ORC_T3) mode = p->is_periodic >= 2 ? 2 : 1;
    else if (p->topology == ORC_S1R2) mode = p->is_periodic >= 2 ? (p->s1r2_lookup ? 4 : 5) : 3;
    else return -4;
    if (mode == 2 && !p->t3_table) return -3;
    if (mode == 4 && !p->s1r2_table) return -3;
    if ((mode == 3 || mode == 5) && p->cosmology == 1 && p->comoving == 1 && !p->radial_table) return -3;
    double rmax = 0, M_min = M[0];
    for (int i = 0; i < N; ++i) {
        double r = (p->topology == ORC_S1R2) ? sqrt(pow(x_in[3 * i], 2) + pow(x_in[3 * i + 1], 2)) /* forces.cc:1018 */
                                             : sqrt(pow(x_in[3 * i], 2) + pow(x_in[3 * i + 1], 2) + pow(x_in[3 * i + 2], 2));
        if (r > rmax) rmax = r;
        if (M_min > M[i]) M_min = M[i]; /* utils.cc:62-67 */
    }
    /* the positions in the octree's frame (a copy: the reference moves x in
     * place and back) */
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    if (!x) return -1;
    memcpy(x, x_in, sizeof(double) * 3 * (size_t)N);
    double axis[3] = {0, 0, 0}, angle = 0;
    double c0[3] = {0, 0, 0}, side;
    if (rnd && p->topology == ORC_R3) { /* forces.cc:420-434, 358-369 */
        for (int k = 0; k < 3; k++) c0[k] = (o->u[k] - 0.5) * 2.0 * rmax;
        side = 4.00004 * rmax;
        angle = o->u[3] * ORC_PI;
        double th = o->u[4] * 2.0 * M_PI, z = o->u[5] * 2.0 - 1.0, r = sqrt(1.0 - z * z);
        axis[0] = r * cos(th);
        axis[1] = r * sin(th);
        axis[2] = z;
        double c = cos(angle), s = sin(angle);
        for (int i = 0; i < N; i++) orc_rotate3(x + 3 * i, axis, c, s);
    } else if (rnd && p->topology == ORC_T3) { /* forces.cc:662-685 */
        double sh[3];
        for (int k = 0; k < 3; k++) sh[k] = (o->u[k] - 0.5) * p->L;
        for (int i = 0; i < N; i++)
            for (int k = 0; k < 3; k++) x[3 * i + k] = orc_shift_wrap(x[3 * i + k], sh[k], p->L);
        side = p->L;
        c0[0] = c0[1] = c0[2] = 0.50 * p->L;
    } else if (rnd) { /* forces.cc:1026-1073 */
        angle = o->u[0] * 2.0 * ORC_PI;
        c0[0] = (o->u[1] - 0.5) * 2 * rmax;
        c0[1] = (o->u[2] - 0.5) * 2 * rmax;
        double shz = (o->u[3] - 0.5) * p->L;
        double c = cos(angle), s = sin(angle);
        for (int i = 0; i < N; i++) {
            orc_rotate2(x + 3 * i, c, s);
            x[3 * i + 2] = orc_shift_wrap(x[3 * i + 2], shz, p->L);
        }
        side = (4 * rmax < p->L) ? p->L : 4.0 * rmax;
        c0[2] = 0.50 * side;
    } else if (p->topology == ORC_R3) { /* the root cube: forces.cc:437-438 (R3), :690 (T3), :1047-1058 (S1R2) */
        side = 2.00002 * rmax;
    } else if (p->topology == ORC_T3) {
        side = p->L;
        c0[0] = c0[1] = c0[2] = 0.50 * p->L;
    } else {
        side = (2 * rmax < p->L) ? p->L : 2.0 * rmax;
        c0[2] = 0.50 * side;
    }
    orc_bh_root[0] = c0[0];
    orc_bh_root[1] = c0[1];
    orc_bh_root[2] = c0[2];
    orc_bh_root[3] = side;
    orc_bh_tree t = {NULL, 0, 0};
    int rc = orc_bh_new(&t, c0[0], c0[1], c0[2], side) < 0 ? -1 : 0;
    for (int i = 0; i < N && !rc; ++i) rc = orc_bh_insert(&t, 0, x, M, i, 0);
    if (!rc && o && o->com_from_children) orc_bh_com_from_children(&t, 0);
    if (rc) {
        free(t.n);
        free(x);
        return rc;
    }
    const double cb = cos(-angle), sb = sin(-angle);
    const double DE = p->H0 * p->H0 * p->Omega_lambda;
    const int ewald_max = p->is_periodic + 1;          /* main.cc:1270 */
    const double cut = ((double)ewald_max - 0.4) * p->L; /* main.cc:1271, forces.cc:1100 */
    long long visited = 0, accepted = 0;
    int kk;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : visited, accepted)
    for (kk = id_min; kk <= id_max; ++kk) {
        orc_bh_ctx c = {p, &t, x, h, theta, particle_radius, M_min, mode, 0.0, cut, {0, 0}};
        double f[3] = {0, 0, 0};
        if (mode == 5) { /* forces.cc:1108-1118: one traversal per z image, summed in image order */
            for (int m = -ewald_max; m < ewald_max + 1; m++) {
                double g[3] = {0, 0, 0};
                c.zi_image = x[3 * kk + 2] + ((double)m) * p->L;
                orc_bh_force(&c, 0, kk, g);
                for (int a = 0; a < 3; ++a) f[a] += g[a];
            }
        } else {
            orc_bh_force(&c, 0, kk, f);
        }
        visited += c.cnt[0];
        accepted += c.cnt[1];
        const double *xi = x + 3 * kk;
        if (p->topology == ORC_R3 && p->cosmology == 1) { /* forces.cc:474-491 */
            double k = p->comoving == 1 ? p->mass_in_unit_sphere : DE;
            for (int a = 0; a < 3; ++a) f[a] += k * xi[a];
            if (rbh) { /* forces.cc:479-485 */
                double r = sqrt(pow(xi[0], 2) + pow(xi[1], 2) + pow(xi[2], 2));
                double c = oracle_radial_bh_correction(r, p->Rsim, rbh, o->radial_bh_table_size);
                for (int a = 0; a < 3; ++a) f[a] -= xi[a] * c / r;
            }
        } else if (p->topology == ORC_S1R2 && p->cosmology == 1) { /* forces.cc:1127-1150, 1170-1190 */
            double k = DE;
            double rxy = sqrt(pow(xi[0], 2) + pow(xi[1], 2));
            if (p->comoving == 1) {
                k = p->mass_in_unit_sphere;
                if (mode != 4)
                    k *= oracle_cylindrical_force_correction_order(rxy, p->Rsim, p->radial_table,
                                                                   p->radial_table_size, 1);
            }
            f[0] += k * xi[0];
            f[1] += k * xi[1];
            if (rbh) { /* forces.cc:1143-1147, 1180-1184 */
                double c = oracle_radial_bh_correction(rxy, p->Rsim, rbh, o->radial_bh_table_size);
                f[0] -= xi[0] * c / rxy;
                f[1] -= xi[1] * c / rxy;
            }
        }
        if (rnd && p->topology == ORC_R3) orc_rotate3(f, axis, cb, sb); /* forces.cc:502-503 */
        else if (rnd && p->topology == ORC_S1R2) orc_rotate2(f, cb, sb); /* forces.cc:1210 */
        for (int a = 0; a < 3; ++a) F[3 * (kk - id_min) + a] = f[a];
    }
    orc_bh_counts[0] = visited;
    orc_bh_counts[1] = accepted;
    free(t.n);
    free(x);
    return 0;
}
