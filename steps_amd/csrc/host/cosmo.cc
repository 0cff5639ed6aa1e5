// cosmo.cc -- the expansion history of a StePS run (friedmann_solver.cc): the
// Hubble parameter, one fourth-order Runge-Kutta step of da/dt = a H(a), and
// the cosmic time of the starting scale factor, for the four COSMOPARAM
// builds: 0 LCDM, 1 wCDM (constant w0), 2 w0-wa (w = w0 + wa (1 - a)), -1 a
// tabulated history (t, a, H) read from EXPANSION_FILE and interpolated.
#include "host_common.h"

#include <cmath>
#include <cstdio>

namespace {

// The dark-energy density in units of its present value: a^(-3(1+w0)) for
// wCDM, a^(-3(1+w0+wa)) exp(-3 wa (1-a)) for w0-wa (friedmann_solver.cc:
// 183-194, 259-272), 1 for LCDM (Omega_lambda enters as it is).
double de_term(const steps_expansion *e, double b) {
    if (e->cosmoparam == 1) return e->Omega_lambda * std::pow(b, -3.0 * (1.0 + e->w0));
    if (e->cosmoparam == 2)
        return e->Omega_lambda * std::pow(b, -3.0 * (1.0 + e->wa + e->w0)) * std::exp(-3.0 * e->wa * (1.0 - b));
    return e->Omega_lambda;
}

// E(b)^2 without (flat) and with the curvature term
double e2_flat(const steps_expansion *e, double b) {
    return e->Omega_m * std::pow(b, -3.0) + e->Omega_r * std::pow(b, -4.0) + de_term(e, b);
}
double e2_curved(const steps_expansion *e, double b) {
    return e2_flat(e, b) + e->Omega_k * std::pow(b, -2.0);
}

// utils.cc:24-56
double lin(double X, double X1, double Y1, double X2, double Y2) {
    const double A = (Y2 - Y1) / (X2 - X1);
    const double B = Y1 - A * X1;
    return A * X + B;
}
double quad(double X, double X1, double Y1, double X2, double Y2, double X3, double Y3) {
    const double L1 = (X - X2) * (X - X3) / ((X1 - X2) * (X1 - X3));
    const double L2 = (X - X1) * (X - X3) / ((X2 - X1) * (X2 - X3));
    const double L3 = (X - X1) * (X - X2) / ((X3 - X1) * (X3 - X2));
    return Y1 * L1 + Y2 * L2 + Y3 * L3;
}
double cubic(double X, double X1, double Y1, double X2, double Y2, double X3, double Y3, double X4, double Y4) {
    const double L1 = (X - X2) * (X - X3) * (X - X4) / ((X1 - X2) * (X1 - X3) * (X1 - X4));
    const double L2 = (X - X1) * (X - X3) * (X - X4) / ((X2 - X1) * (X2 - X3) * (X2 - X4));
    const double L3 = (X - X1) * (X - X2) * (X - X4) / ((X3 - X1) * (X3 - X2) * (X3 - X4));
    const double L4 = (X - X1) * (X - X2) * (X - X3) / ((X4 - X1) * (X4 - X2) * (X4 - X3));
    return Y1 * L1 + Y2 * L2 + Y3 * L3 + Y4 * L4;
}

// Column c of row r of the tabulated history (t, a, H).
double tab(const steps_expansion *e, int64_t r, int c) { return e->table[3 * r + c]; }

// Interpolation of column cy against column cx around row i (the first row
// whose cx value passes the argument), order 1, 2 or 3 with the reference's
// index clamps (friedmann_solver.cc:343-379, 381-415).
double tab_interp(const steps_expansion *e, int64_t i, int cx, int cy, double X) {
    const int64_t n = e->rows;
    switch (e->order) {
    case 1:
        if (i < 1) i = 1;
        else if (i >= n) i = n - 1;
        return lin(X, tab(e, i - 1, cx), tab(e, i - 1, cy), tab(e, i, cx), tab(e, i, cy));
    case 2:
        if (i < 2) i = 2;
        else if (i >= n) i = n - 1;
        return quad(X, tab(e, i - 2, cx), tab(e, i - 2, cy), tab(e, i - 1, cx), tab(e, i - 1, cy), tab(e, i, cx),
                    tab(e, i, cy));
    default:  // 3
        if (i < 2) i = 2;
        else if (i >= n - 1) i = n - 2;
        return cubic(X, tab(e, i - 2, cx), tab(e, i - 2, cy), tab(e, i - 1, cx), tab(e, i - 1, cy), tab(e, i, cx),
                     tab(e, i, cy), tab(e, i + 1, cx), tab(e, i + 1, cy));
    }
}

}  // namespace

extern "C" {

// friedmann_solver.cc:161-164, 237-240, 322-325, 381-415
double steps_hubble(const steps_expansion *e, double a) {
    if (e->cosmoparam == -1) {
        int64_t i = e->index - 2;
        do {
            ++i;
            if (i == e->rows) {
                --i;
                break;
            }
        } while (tab(e, i, 1) <= a);
        return tab_interp(e, i, 1, 2, a);
    }
    return e->H0 * std::sqrt(e2_curved(e, a));
}

// friedmann_solver.cc:100-159 (and its wCDM / w0-wa twins): one classical RK4
// step of da/dt = a H(a).  Flat models (|Omega_k| < 1e-9) leave out the
// curvature term and return -1 when a stage's E^2 is negative; curved ones
// integrate sqrt|E^2| and run backwards once every stage is negative (a
// recollapse; also when H0 <= 0).  The tabulated history interpolates a at
// the time T at the end of the step from the rows around T - h
// (friedmann_solver.cc:343-379), and moves the history's cursor.
double steps_friedmann_step(steps_expansion *e, double a0, double h, double T) {
    if (e->cosmoparam == -1) {
        int64_t i = e->index - 2;
        const double T_prev = T - h;
        do {
            ++i;
            if (i == e->rows) {
                --i;
                break;
            }
        } while (tab(e, i, 0) <= T_prev);
        // the reference's case clamps rewrite the cursor too
        const int64_t lo = (e->order == 1) ? 1 : 2;
        const int64_t hi = (e->order == 3) ? e->rows - 2 : e->rows - 1;
        if (i < lo) i = lo;
        else if (i > hi) i = hi;
        e->index = i;
        return tab_interp(e, i, 0, 1, T);
    }
    const double H0 = e->H0;
    int collapse = H0 > 0 ? 0 : 1;
    const double b = a0;
    if (std::fabs(e->Omega_k) < 1e-9) {
        auto k = [&](double x) { return x * H0 * std::sqrt(e2_flat(e, x)); };
        const double j = e2_flat(e, b);
        const double k1 = k(b);
        const double l = e2_flat(e, b + h * k1 / 2.0);
        const double k2 = k(b + h * k1 / 2.0);
        const double m = e2_flat(e, b + h * k2 / 2.0);
        const double k3 = k(b + h * k2 / 2.0);
        const double n = e2_flat(e, b + h * k3);
        const double k4 = k(b + h * k3);
        const double K = h * (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
        if (j < 0 || l < 0 || m < 0 || n < 0) return -1;
        return b + K;
    }
    const double j = e2_curved(e, b);
    const double k1 = b * H0 * std::sqrt(std::fabs(j));
    const double l = e2_curved(e, b + h * k1 / 2.0);
    const double k2 = (b + h * k1 / 2.0) * H0 * std::sqrt(std::fabs(l));
    const double m = e2_curved(e, b + h * k2 / 2.0);
    const double k3 = (b + h * k2 / 2.0) * H0 * std::sqrt(std::fabs(m));
    const double n = e2_curved(e, b + h * k3);
    const double k4 = (b + h * k3) * H0 * std::sqrt(std::fabs(n));
    if (j < 0 && l < 0 && m < 0 && n < 0) collapse = 1;
    const double K = h * (k1 + k2 * 2.0 + k3 * 2.0 + k4) / 6.0;
    return collapse == 0 ? b + K : b - K;
}

// friedmann_solver.cc:33-95: the time since the big bang at a_start, by
// integrating backwards from a_start with step h until a <= 0, extrapolating
// the last two positive steps to a = 0 as a power law (a ~ t^(2/3) when
// matter dominates the last step, t^(1/2) under radiation), and repeating
// with h halved until two estimates agree to 1e-10 (at most 20 times).  The
// tabulated history takes the time of the row before the first one past
// a_start: the reference's interpolation cases fall through to that default
// (friedmann_solver.cc:438-467).
double steps_friedmann_start(steps_expansion *e, double a_start, double h) {
    if (e->cosmoparam == -1) {
        int64_t i = 0;
        do {
            ++i;
        } while (i < e->rows - 1 && tab(e, i, 1) <= a_start);
        if (i < 2) i = 2;  // the clamps of the cases the reference falls through
        e->index = i;
        return tab(e, i - 1, 0);
    }
    int it = 0;
    double hv = h, t = 0.0, t_tmp = 1.0;
    while (std::fabs(t - t_tmp) > 1e-10 && it < 20) {
        t_tmp = t;
        t = 0.0;
        double b = a_start, b0 = b, b1 = b;
        while (0 < b) {
            b1 = b0;
            b0 = b;
            b = steps_friedmann_step(e, b, -hv, 0.0);
            t -= hv;
        }
        const double Hb = steps_hubble(e, b0);
        const double Om_b = e->Omega_m * std::pow(b0, -3) * std::pow(e->H0 / Hb, 2);
        const double Or_b = e->Omega_r * std::pow(b0, -4) * std::pow(e->H0 / Hb, 2);
        const double p = (Om_b > Or_b) ? std::pow(b1 / b0, 1.5) : std::pow(b1 / b0, 2.0);
        const double t2 = t + 2 * hv, t1 = t + hv;
        t = (p * t1 - t2) / (p - 1);
        hv *= 0.5;
        ++it;
    }
    return -1.0 * t;
}

// read_expansion_history (inputoutput.cc:64-97): whitespace-separated rows
// t [Gy], a, H [km/s/Mpc], converted to internal units.  With table == NULL
// only the row count is returned (rows are counted as read_ascii_ic counts
// particles: one per newline).
int64_t steps_read_expansion_history(const char *path, double *table, int64_t cap) {
    const int64_t n = steps_ascii_count_particles(path);
    if (n < 0 || !table) return n;
    if (n > cap) return host_fail(STEPS_ERR_ARG, "expansion history larger than the buffer");
    std::FILE *f = std::fopen(path, "r");
    if (!f) return host_fail(STEPS_ERR_ARG, std::string("cannot open ") + path);
    for (int64_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c) {
            double v = 0.0;
            if (std::fscanf(f, "%lf", &v) != 1) {
                std::fclose(f);
                return host_fail(STEPS_ERR_ARG, std::string("short expansion history ") + path);
            }
            table[3 * i + c] = (c == 0) ? v / STEPS_UNIT_T : (c == 2 ? v / STEPS_UNIT_V : v);
        }
    std::fclose(f);
    return n;
}

}  // extern "C"
