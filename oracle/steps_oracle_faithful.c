/*
 * steps_oracle_faithful.c -- the timed CPU baseline.
 *
 * TEST INFRASTRUCTURE ONLY (see steps_oracle.h).  This translation unit is
 * compiled with the reference's own CPU flags, -O3 -ffast-math -fopenmp
 * (StePS/Template-LinuxGCC-Makefile:81), and restates the reference's OpenMP
 * direct sum including its costs: schedule(dynamic,chunk) with
 * chunk=(rows)/threads, pow()/sqrt(), the branchy softening function and three
 * `omp atomic` adds per pair (forces.cc:510-577, :776-876, :1221-1404).
 * The _fast_ variant is the same mathematics without atomics and with
 * register accumulation, reported beside it for context.
 */
#include "steps_oracle.h"

#include <math.h>
#include <omp.h>

#define ORC_RADIAL_EXTENT 2.25

/* forces.cc:52-87 */
static double softening_ff(double r, double beta)
{
    double betap2 = beta * 0.5;
    double wij;
    if (r >= beta) {
        wij = pow(r, -3);
    } else if (r > betap2 && r < beta) {
        wij = -32.0 / (3.0 * pow(beta, 6)) * pow(r, 3) + 38.4 / pow(beta, 5) * pow(r, 2) +
              -48.0 / pow(beta, 4) * r + 64.0 / (3.0 * pow(beta, 3)) + (-1.0 / 15.0) / pow(r, 3);
    } else {
        wij = 32.0 / pow(beta, 6) * pow(r, 3) + -38.4 / pow(beta, 5) * pow(r, 2) +
              32.0 / (3.0 * pow(beta, 3));
    }
    return wij;
}

int oracle_forces_faithful_rows(const oracle_params *p, int N, const double *x,
                                const double *M, const double *h, double *F,
                                const int *rows, int nrows)
{
    if (!p || N < 0 || nrows < 0) return -1;
    const double L = p->L;
    const double DE = p->H0 * p->H0 * p->Omega_lambda;
    for (int k = 0; k < 3 * nrows; k++) F[k] = 0; /* forces.cc:520-526 */
    int chunk = nrows / omp_get_max_threads();
    if (chunk < 1) chunk = 1;
    int kk;
#pragma omp parallel for schedule(dynamic, chunk)
    for (kk = 0; kk < nrows; kk++) {
        int i = rows[kk];
        for (int j = 0; j < N; j++) {
            double beta = h[i] + h[j];
            double dx = x[3 * j] - x[3 * i], dy = x[3 * j + 1] - x[3 * i + 1],
                   dz = x[3 * j + 2] - x[3 * i + 2];
            double fx, fy, fz;
            if (p->topology == ORC_R3) {
                double r = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
                double w = M[j] * softening_ff(r, beta);
                fx = w * dx; fy = w * dy; fz = w * dz;
            } else if (p->topology == ORC_T3) {
                if (fabs(dx) > 0.5 * L) dx = dx - L * dx / fabs(dx);
                if (fabs(dy) > 0.5 * L) dy = dy - L * dy / fabs(dy);
                if (fabs(dz) > 0.5 * L) dz = dz - L * dz / fabs(dz);
                double r = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
                if (p->is_periodic >= 2) {
                    double w = softening_ff(r, beta), D[3];
                    oracle_t3_interpolate(p->t3_ngrid, L, p->t3_table, dx, dy, dz, D,
                                          p->interp_order);
                    fx = M[j] * (w * dx - D[0]);
                    fy = M[j] * (w * dy - D[1]);
                    fz = M[j] * (w * dz - D[2]);
                } else {
                    double w = M[j] * softening_ff(r, beta);
                    fx = w * dx; fy = w * dy; fz = w * dz;
                }
            } else {
                if (fabs(dz) > 0.5 * L) dz = dz - L * dz / fabs(dz);
                double r = sqrt(pow(dx, 2) + pow(dy, 2) + pow(dz, 2));
                if (p->is_periodic >= 2) {
                    double w = softening_ff(r, beta), D[3];
                    oracle_s1r2_interpolate(p->s1r2_table, p->s1r2_nrho, p->s1r2_nz,
                                            ORC_RADIAL_EXTENT * p->Rsim, L, dx, dy, dz,
                                            p->interp_order, D);
                    fx = M[j] * (w * dx - D[0]);
                    fy = M[j] * (w * dy - D[1]);
                    fz = M[j] * (w * dz - D[2]);
                } else {
                    double w = M[j] * softening_ff(r, beta);
                    fx = w * dx; fy = w * dy; fz = w * dz;
                }
            }
#pragma omp atomic
            F[3 * kk] += fx;
#pragma omp atomic
            F[3 * kk + 1] += fy;
#pragma omp atomic
            F[3 * kk + 2] += fz;
        }
        if (p->cosmology == 1 && p->topology != ORC_T3) {
            double c = p->comoving == 1 ? p->mass_in_unit_sphere : DE;
            F[3 * kk] += c * x[3 * i];
            F[3 * kk + 1] += c * x[3 * i + 1];
            if (p->topology == ORC_R3) F[3 * kk + 2] += c * x[3 * i + 2];
        }
    }
    return 0;
}

/* Same R3 mathematics, arranged for the compiler: no atomics, one row per
 * iteration accumulated in registers. */
int oracle_forces_fast_rows(const oracle_params *p, int N, const double *x, const double *M,
                            const double *h, double *F, const int *rows, int nrows)
{
    if (!p || p->topology != ORC_R3) return -1;
    const double DE = p->H0 * p->H0 * p->Omega_lambda;
    int kk;
#pragma omp parallel for schedule(dynamic, 4)
    for (kk = 0; kk < nrows; kk++) {
        int i = rows[kk];
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2], hi = h[i];
        double fx = 0, fy = 0, fz = 0;
        for (int j = 0; j < N; j++) {
            double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi, dz = x[3 * j + 2] - zi;
            double r2 = dx * dx + dy * dy + dz * dz;
            double beta = hi + h[j];
            double w;
            if (r2 >= beta * beta) {
                double r = sqrt(r2);
                w = 1.0 / (r2 * r);
            } else {
                w = softening_ff(sqrt(r2), beta);
            }
            w *= M[j];
            fx += w * dx; fy += w * dy; fz += w * dz;
        }
        if (p->cosmology == 1) {
            double c = p->comoving == 1 ? p->mass_in_unit_sphere : DE;
            fx += c * xi; fy += c * yi; fz += c * zi;
        }
        F[3 * kk] = fx; F[3 * kk + 1] = fy; F[3 * kk + 2] = fz;
    }
    return 0;
}
