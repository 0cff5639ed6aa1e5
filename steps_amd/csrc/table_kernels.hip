// table_kernels.hip -- Ewald and radial correction tables built on the device.
//
// The reference builds these on the host at start-up (T3: ewald_space.cc:288-383,
// OpenMP over a 1/48 wedge; S1R2: ewald_space.cc:754-798, serial; radial:
// utils.cc:162-228, serial).  Every table node is independent, so each one is a
// GPU thread here; the results are returned in host memory because the force
// descriptor takes host tables (the reference's globals are host arrays).
#include "internal.h"

#include <map>
#include <mutex>
#include <tuple>

#include <cmath>
#include <cstring>
#include <vector>

// No FMA contraction in this file: the reference builds its tables on the host
// without fused multiply-adds, and the grid positions must come out the same --
// fused, (31 + 0.5)*(L/63) - L/2 is ~1e-14 instead of exactly 0 and the T3
// correction at that node becomes the difference of two ~1e28 forces.
#pragma clang fp contract(off)

namespace steps {

namespace {

constexpr double kSqrtPi = 1.7724538509055160272981674833411;  // ewald_space.cc:23

// --------------------------------------------------------------------------
// T3
// --------------------------------------------------------------------------

// The 24 proper rotations of the cube group O as (permutation, signs), in the
// order ewald_space.cc:63-114 builds them (even permutations with an even
// number of sign flips, then odd permutations with an odd number); the fill
// order matters only where a node is its own image.  Computed from q in
// registers, on host and device alike.
__host__ __device__ inline void o_rotation(int q, int perm[3], int sign[3]) {
    const int pe[3][3] = {{0, 1, 2}, {1, 2, 0}, {2, 0, 1}};
    const int po[3][3] = {{0, 2, 1}, {2, 1, 0}, {1, 0, 2}};
    const int sp[4][3] = {{1, 1, 1}, {1, -1, -1}, {-1, 1, -1}, {-1, -1, 1}};
    const int sn[4][3] = {{-1, -1, -1}, {-1, 1, 1}, {1, -1, 1}, {1, 1, -1}};
    const int p = (q % 12) / 4, s = q % 4;
    for (int k = 0; k < 3; ++k) {
        perm[k] = q < 12 ? pe[p][k] : po[p][k];
        sign[k] = q < 12 ? sp[s][k] : sn[s][k];
    }
}

// Integer image vectors n with |n_c| < R and |n|^2 < R^2 in the order of
// ewald_space.cc:118-205 (first octant point, then its sign variants).
std::vector<int4> image_vectors(double R) {
    std::vector<int4> v;
    for (int i = 0; i < R; i++)
        for (int j = 0; j < R; j++)
            for (int k = 0; k < R; k++) {
                if (!((double)(i * i + j * j + k * k) < R * R)) continue;
                const int n2 = i * i + j * j + k * k;
                v.push_back(make_int4(i, j, k, n2));
                if (i) v.push_back(make_int4(-i, j, k, n2));
                if (j) v.push_back(make_int4(i, -j, k, n2));
                if (k) v.push_back(make_int4(i, j, -k, n2));
                if (j && k) v.push_back(make_int4(i, -j, -k, n2));
                if (i && k) v.push_back(make_int4(-i, j, -k, n2));
                if (i && j) v.push_back(make_int4(-i, -j, k, n2));
                if (i && j && k) v.push_back(make_int4(-i, -j, -k, n2));
            }
    return v;
}

__host__ __device__ inline size_t t3_off(int Ng, int ix, int iy, int iz, int c) {
    return ((size_t)((ix * Ng + iy) * Ng + iz) * 3u) + (size_t)c;
}

// ewald_space.cc:210-286: D(r) = (F_real + F_rec) - F_newton at one separation.
__host__ __device__ inline void t3_node_correction(const double r[3], const int4 *ridx, int nreal,
                                                   const int4 *kidx, int nrec, double L, double alpha,
                                                   double rel_cut, double rec_cut, double D[3]) {
    double Fr[3] = {0, 0, 0}, Fk[3] = {0, 0, 0};
    const double V = L * L * L;
    const double two_alpha_over_sqrtpi = 2.0 * alpha / sqrt(kPi);
    const double fourpi_over_V = (4.0 * kPi) / V;
    const double relcutL2 = (rel_cut * L) * (rel_cut * L);
    for (int m = 0; m < nreal; ++m) {
        const int4 n = ridx[m];
        const double Rv[3] = {r[0] + (double)n.x * L, r[1] + (double)n.y * L, r[2] + (double)n.z * L};
        const double R2 = Rv[0] * Rv[0] + Rv[1] * Rv[1] + Rv[2] * Rv[2];
        if (R2 == 0.0 || R2 > relcutL2) continue;
        const double R = sqrt(R2);
        const double invR = 1.0 / R;
        const double invR3 = invR * invR * invR;
        const double bracket = erfc(alpha * R) + two_alpha_over_sqrtpi * R * exp(-(alpha * alpha) * R2);
        const double coeff = invR3 * bracket;
        Fr[0] += -Rv[0] * coeff;
        Fr[1] += -Rv[1] * coeff;
        Fr[2] += -Rv[2] * coeff;
    }
    const double factor = 2.0 * kPi / L;
    for (int m = 0; m < nrec; ++m) {
        const int4 n = kidx[m];
        if (n.x == 0 && n.y == 0 && n.z == 0) continue;
        const double kv[3] = {factor * (double)n.x, factor * (double)n.y, factor * (double)n.z};
        const double k2 = kv[0] * kv[0] + kv[1] * kv[1] + kv[2] * kv[2];
        if (k2 == 0.0 || k2 > rec_cut * rec_cut) continue;
        const double damp = exp(-k2 / (4.0 * alpha * alpha));
        const double coeff = fourpi_over_V * damp / k2;
        const double s = sin(kv[0] * r[0] + kv[1] * r[1] + kv[2] * r[2]);
        Fk[0] += -kv[0] * coeff * s;
        Fk[1] += -kv[1] * coeff * s;
        Fk[2] += -kv[2] * coeff * s;
    }
    D[0] = D[1] = D[2] = 0.0;
    const double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
    if (r2 > 0.0) {
        const double rn = sqrt(r2);
        const double invr3 = 1.0 / (rn * rn * rn);
        for (int c = 0; c < 3; ++c) D[c] = (Fr[c] + Fk[c]) - (-r[c] * invr3);
    }
}

// One wedge node per thread, then the O_h fill of ewald_space.cc:322-378.
__global__ void t3_wedge_kernel(const int4 *__restrict__ wedge, int nw, const int4 *__restrict__ ridx,
                                int nreal, const int4 *__restrict__ kidx, int nrec, double L,
                                double alpha, double rel_cut, double rec_cut, int Ng,
                                double *__restrict__ T) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nw) return;
    const int4 w = wedge[t];
    const double gs = L / (double)Ng;
    const double r[3] = {((double)w.x + 0.5) * gs - L / 2.0, ((double)w.y + 0.5) * gs - L / 2.0,
                         ((double)w.z + 0.5) * gs - L / 2.0};
    double D[3];
    t3_node_correction(r, ridx, nreal, kidx, nrec, L, alpha, rel_cut, rec_cut, D);
    for (int c = 0; c < 3; ++c) T[t3_off(Ng, w.x, w.y, w.z, c)] = D[c];
    const int src[3] = {w.x, w.y, w.z};
    for (int q = 0; q < 24; ++q) {
        int perm[3], sign[3];
        o_rotation(q, perm, sign);
        int dst[3];
        double Dr[3];
        for (int d = 0; d < 3; ++d) {
            int v = src[perm[d]];
            if (sign[d] < 0) v = (Ng - 1) - v;
            dst[d] = v;
            Dr[d] = sign[d] * D[perm[d]];
        }
        for (int c = 0; c < 3; ++c) T[t3_off(Ng, dst[0], dst[1], dst[2], c)] = Dr[c];
        const int inv[3] = {(Ng - 1) - dst[0], (Ng - 1) - dst[1], (Ng - 1) - dst[2]};
        for (int c = 0; c < 3; ++c) T[t3_off(Ng, inv[0], inv[1], inv[2], c)] = -Dr[c];
    }
}

// --------------------------------------------------------------------------
// S1R2
// --------------------------------------------------------------------------

// ewald_space.cc:539-561 (the reference's erfcx fit, A&S 7.1.26 coefficients);
// the x < 0 recursion unrolled once.
__device__ double fit_erfcx_pos(double x) {
    if (x == 0.0) return 1.0;
    if (x > 1e10) return 0.564189583547756286 / x;
    const double t = 1.0 / (1.0 + 0.3275911 * x);
    const double a1 = 0.254829592, a2 = -0.284496736, a3 = 1.421413741, a4 = -1.453152027,
                 a5 = 1.061405429;
    return t * (a1 + t * (a2 + t * (a3 + t * (a4 + t * a5))));
}
__device__ double fit_erfcx(double x) {
    if (x < 0) return 2.0 * exp(x * x) - fit_erfcx_pos(-x);
    return fit_erfcx_pos(x);
}

// ewald_space.cc:564-595
__device__ double exp_erfc_product(double km, double rho, double alpha, bool term_p) {
    const double a = km / (2.0 * alpha);
    const double ar = alpha * rho;
    const double arg = term_p ? (a + ar) : (a - ar);
    const double common = -(a * a + ar * ar);
    if (arg > 5.0) return exp(common) * fit_erfcx(arg);
    if (arg < -5.0) return 2.0 * exp(-km * rho) - exp(common) * fit_erfcx(-arg);
    const double shift = term_p ? (km * rho) : (-km * rho);
    return exp(shift) * erfc(arg);
}

// ewald_space.cc:599-608
__device__ double wrap_dz(double dz, double Lz) {
    const double half = 0.5 * Lz;
    if (dz > half) dz -= Lz * floor((dz + half) / Lz);
    if (dz <= -half) dz -= Lz * floor((dz - half) / Lz);
    return dz;
}

// ewald_space.cc:618-752 (dy = 0 on the table grid) and :754-798
__global__ void s1r2_table_kernel(int Nrho, int Nz, double rho_max, double Lz, double alpha, int nmax,
                                  int mmax, int rspace, double *__restrict__ T) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nrho * Nz) return;
    const int ir = (int)(t / Nz), iz = (int)(t % Nz);
    const double drho = rho_max / (double)max(1, Nrho - 1);
    const double dzg = Lz / (double)Nz;
    const double rho0 = (double)ir * drho;
    const double z = ((double)iz + 0.5) * dzg - 0.5 * Lz;
    // periodic force at separation (rho0, 0, z)
    const double dx = rho0, dy = 0.0;
    const double dz = wrap_dz(z, Lz);
    const double rho2 = dx * dx + dy * dy;
    double Fx = 0, Fz = 0;
    if (!rspace) {
        const double rho = sqrt(rho2);
        double Fxr = 0, Fzr = 0;
        for (int n = -nmax; n <= nmax; ++n) {
            const double dzn = dz + (double)n * Lz;
            const double r2 = rho2 + dzn * dzn;
            if (r2 < 1e-18) continue;
            const double r = sqrt(r2);
            const double invr3 = 1.0 / (r2 * r);
            const double ar = alpha * r;
            const double coeff = (erfc(ar) + (2.0 / kSqrtPi) * ar * exp(-ar * ar)) * invr3;
            Fxr -= dx * coeff;
            Fzr -= dzn * coeff;
        }
        double Frk = 0, Fzk = 0;
        const double invL = 1.0 / Lz;
        for (int m = 1; m <= mmax; ++m) {
            const double km = (2.0 * kPi * m) * invL;
            const double tp = exp_erfc_product(km, rho, alpha, true);
            const double tm = exp_erfc_product(km, rho, alpha, false);
            const double B = tp + tm;
            const double dB = km * (tp - tm);
            Frk += (2.0 * invL) * cos(km * dz) * dB;
            Fzk -= (2.0 * invL) * km * sin(km * dz) * B;
        }
        if (rho > 1e-12) Frk += -(2.0 * invL / rho) * (1.0 - exp(-alpha * alpha * rho2));
        const double Fxk = (rho > 0) ? (Frk * dx / rho) : 0.0;
        Fx = Fxr + Fxk;
        Fz = Fzr + Fzk;
    } else {
        double r2 = rho2 + dz * dz;
        if (r2 >= 1e-18) {
            const double r = sqrt(r2), invr3 = 1.0 / (r2 * r);
            Fx -= dx * invr3;
            Fz -= dz * invr3;
        }
        for (int n = 1; n <= nmax; ++n)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const double dzn = dz + (double)(sgn * n) * Lz;
                r2 = rho2 + dzn * dzn;
                if (r2 < 1e-18) continue;
                const double r = sqrt(r2), invr3 = 1.0 / (r2 * r);
                Fx -= dx * invr3;
                Fz -= dzn * invr3;
            }
    }
    // minus the nearest-image Newtonian force
    const double dzw = wrap_dz(z, Lz);
    const double r2 = rho0 * rho0 + dzw * dzw;
    double Frn = 0, Fzn = 0;
    if (r2 > 0.0) {
        const double r = sqrt(r2), invr3 = 1.0 / (r * r * r);
        Frn = -rho0 * invr3;
        Fzn = -dzw * invr3;
    }
    T[((size_t)ir * Nz + iz) * 2 + 0] = Fx - Frn;
    T[((size_t)ir * Nz + iz) * 2 + 1] = Fz - Fzn;
}

// --------------------------------------------------------------------------
// radial (finite cylinder) table, utils.cc:162-228, one node per thread
// --------------------------------------------------------------------------
__device__ double cyl_in(double R, double Lz, double a, double Y) {
    const double s = pow((pow(R, 2) - pow(Y, 2)), 0.5);
    const double f1 = -2 * log((pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) + 2 * a * s), 0.5) + Lz) /
                               (pow((pow(Lz, 2) + pow(R, 2) + pow(a, 2) - 2 * a * s), 0.5) + Lz));
    const double f2 = log((pow(R, 2) + pow(a, 2) + 2 * a * s) / (pow(R, 2) + pow(a, 2) - 2 * a * s));
    return f1 + f2;
}
__device__ double cyl_edge(double R, double Lz, double Y) {
    const double s = pow((pow(R, 2) - pow(Y, 2)), 0.5);
    return -2 * log((pow((pow(Lz, 2) + 2 * pow(R, 2) + 2 * R * s), 0.5) + Lz) /
                    (pow((pow(Lz, 2) + 2 * pow(R, 2) - 2 * R * s), 0.5) + Lz));
}

__global__ void radial_table_kernel(double R, double Lz, int size, int acc, double *__restrict__ T) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    if (i >= size) return;
    const double step = R / (double)acc;
    const double a = (i == size - 1) ? R : R / (double)(size - 1) * i;
    double total = 0;
    for (int j = 1; j <= acc; j++) {
        double ip, in;
        if (a >= R) {
            ip = cyl_edge(R, Lz, (j - 1) * step);
            in = cyl_edge(R, Lz, j * step);
        } else {
            ip = cyl_in(R, Lz, a, (j - 1) * step);
            in = cyl_in(R, Lz, a, j * step);
        }
        total += step * (ip + in) * 0.5;
    }
    T[i] = (a >= R) ? 2 * (total + kPi * R) / (2 * kPi * R) : 2 * total / (2 * kPi * a);
}

struct DevMem {
    void *p = nullptr;
    ~DevMem() {
        if (p) (void)hipFree(p);
    }
};

}  // namespace

int build_t3_table_uncached(int is_periodic, double L, double *table, int *ngrid) {
    // main.cc:422-448
    if (is_periodic < 2) return fail(STEPS_ERR_ARG, "T3 Ewald table needs IS_PERIODIC >= 2");
    if (!(L > 0)) return fail(STEPS_ERR_ARG, "T3 table needs L > 0");
    const double alpha = 2.0 / L;
    double rel_cut, rec_cut;
    int Ng;
    if (is_periodic == 2) { rel_cut = 2.6; rec_cut = 8.0; Ng = 63; }
    else if (is_periodic == 3) { rel_cut = 3.6; rec_cut = 10.0; Ng = 127; }
    else { rel_cut = 4.6; rec_cut = 12.0; Ng = 255; }
    if (ngrid) *ngrid = Ng;
    if (!table) return STEPS_OK;
    const std::vector<int4> ridx = image_vectors(rel_cut + 1.0);  // main.cc:467
    const std::vector<int4> kidx = image_vectors(rec_cut + 2.0);  // main.cc:468
    std::vector<int4> wedge;
    for (int i = Ng / 2; i < Ng; ++i)
        for (int j = Ng / 2; j <= i; ++j)
            for (int k = Ng / 2; k <= j; ++k) wedge.push_back(make_int4(i, j, k, 0));
    const size_t nT = (size_t)Ng * Ng * Ng * 3;
    DevMem dT, dr, dk, dw;
    hipError_t e;
    if ((e = hipMalloc(&dT.p, nT * sizeof(double))) != hipSuccess ||
        (e = hipMalloc(&dr.p, ridx.size() * sizeof(int4))) != hipSuccess ||
        (e = hipMalloc(&dk.p, kidx.size() * sizeof(int4))) != hipSuccess ||
        (e = hipMalloc(&dw.p, wedge.size() * sizeof(int4))) != hipSuccess)
        return hip_fail(e, "T3 table alloc");
    if ((e = hipMemcpy(dr.p, ridx.data(), ridx.size() * sizeof(int4), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dk.p, kidx.data(), kidx.size() * sizeof(int4), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(dw.p, wedge.data(), wedge.size() * sizeof(int4), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(dT.p, 0, nT * sizeof(double))) != hipSuccess)
        return hip_fail(e, "T3 table upload");
    const int nw = (int)wedge.size();
    hipLaunchKernelGGL(t3_wedge_kernel, dim3((nw + 63) / 64), dim3(64), 0, 0, (const int4 *)dw.p, nw,
                       (const int4 *)dr.p, (int)ridx.size(), (const int4 *)dk.p, (int)kidx.size(), L, alpha,
                       rel_cut, rec_cut, Ng, (double *)dT.p);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "T3 table kernel");
    if ((e = hipMemcpy(table, dT.p, nT * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "T3 table download");
    return STEPS_OK;
}

int build_s1r2_table_uncached(int is_periodic, double L, double Rsim, double *table, int *nrho, int *nz) {
    // main.cc:569-602 (Tornberg 1P parameters)
    if (is_periodic < 2) return fail(STEPS_ERR_ARG, "S1R2 Ewald table needs IS_PERIODIC >= 2");
    if (!(L > 0) || !(Rsim > 0)) return fail(STEPS_ERR_ARG, "S1R2 table needs L > 0 and Rsim > 0");
    int rel_cut, rec_cut, Nz;
    double alpha;
    if (is_periodic == 2) { rel_cut = 4; rec_cut = 10; alpha = 0.787875 / L; Nz = 128; }
    else if (is_periodic == 3) { rel_cut = 5; rec_cut = 12; alpha = 0.71805 / L; Nz = 256; }
    else { rel_cut = is_periodic + 2; rec_cut = is_periodic + 9; alpha = 0.6642 / L; Nz = 512; }
    const int Nrho = (int)floor(((double)Nz) * kRadialExtent * Rsim / L);
    if (nrho) *nrho = Nrho;
    if (nz) *nz = Nz;
    if (!table) return STEPS_OK;
    if (Nrho < 1) return fail(STEPS_ERR_ARG, "S1R2 table: Rsim too small for the grid");
    const size_t n = (size_t)Nrho * Nz * 2;
    DevMem dT;
    hipError_t e;
    if ((e = hipMalloc(&dT.p, n * sizeof(double))) != hipSuccess) return hip_fail(e, "S1R2 table alloc");
    const int64_t nodes = (int64_t)Nrho * Nz;
    hipLaunchKernelGGL(s1r2_table_kernel, dim3((unsigned)((nodes + 127) / 128)), dim3(128), 0, 0, Nrho, Nz,
                       kRadialExtent * Rsim, L, alpha, rel_cut, rec_cut, 0, (double *)dT.p);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "S1R2 table kernel");
    if ((e = hipMemcpy(table, dT.p, n * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "S1R2 table download");
    return STEPS_OK;
}

int build_radial_table_uncached(double R, double Lz, int size, int accuracy, double *table) {
    if (!table || size < 1 || accuracy < 1 || !(R > 0) || !(Lz > 0))
        return fail(STEPS_ERR_ARG, "bad radial table arguments");
    DevMem dT;
    hipError_t e;
    if ((e = hipMalloc(&dT.p, (size_t)size * sizeof(double))) != hipSuccess) return hip_fail(e, "radial alloc");
    if ((e = hipMemset(dT.p, 0, (size_t)size * sizeof(double))) != hipSuccess) return hip_fail(e, "radial memset");
    if (size > 1) {
        hipLaunchKernelGGL(radial_table_kernel, dim3((size - 1 + 63) / 64), dim3(64), 0, 0, R, Lz, size, accuracy,
                           (double *)dT.p);
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "radial kernel");
    }
    if ((e = hipMemcpy(table, dT.p, (size_t)size * sizeof(double), hipMemcpyDeviceToHost)) != hipSuccess)
        return hip_fail(e, "radial download");
    // utils.cc:221-227: node 0 extrapolated linearly from nodes 1 and 2
    if (size >= 3) {
        const double X1 = 2 * R / (double)size, Y1 = table[1], X2 = 4 * R / (double)size, Y2 = table[2];
        const double A = (Y2 - Y1) / (X2 - X1);
        const double B = Y1 - A * X1;
        table[0] = A * 0.0 + B;
    }
    return STEPS_OK;
}

// ---------------------------------------------------------------------------
// Built tables are kept, keyed by their kind and parameters, so that a second
// request with the same parameters (another simulation of an ensemble with
// the same box, the size query and the fill of one caller) is a host copy.
// The cache holds at most kTableCacheBytes; beyond that it starts over.
// ---------------------------------------------------------------------------
namespace {

struct TableKey {
    int kind, is_periodic, size, accuracy;
    uint64_t a, b;  // bit patterns of the two length parameters
    bool operator<(const TableKey &o) const {
        return std::tie(kind, is_periodic, size, accuracy, a, b) <
               std::tie(o.kind, o.is_periodic, o.size, o.accuracy, o.a, o.b);
    }
};

struct CachedTable {
    std::vector<double> data;
    int d0 = 0, d1 = 0;
};

constexpr size_t kTableCacheBytes = (size_t)2 << 30;
std::mutex g_table_mu;
std::map<TableKey, CachedTable> g_tables;
size_t g_table_bytes = 0;

uint64_t bits(double x) {
    uint64_t u;
    std::memcpy(&u, &x, sizeof u);
    return u;
}

const CachedTable *cache_find(const TableKey &k) {
    auto it = g_tables.find(k);
    return it == g_tables.end() ? nullptr : &it->second;
}

void cache_put(const TableKey &k, CachedTable t) {
    const size_t b = t.data.size() * sizeof(double);
    if (b > kTableCacheBytes) return;
    if (g_table_bytes + b > kTableCacheBytes) {
        g_tables.clear();
        g_table_bytes = 0;
    }
    g_table_bytes += b;
    g_tables[k] = std::move(t);
}

}  // namespace

int build_t3_table(int is_periodic, double L, double *table, int *ngrid) {
    int ng = 0;
    int rc = build_t3_table_uncached(is_periodic, L, nullptr, &ng);  // validates, sizes
    if (ngrid) *ngrid = ng;
    if (rc || !table) return rc;
    const TableKey key{1, is_periodic, ng, 0, bits(L), 0};
    std::lock_guard<std::mutex> lk(g_table_mu);
    if (const CachedTable *c = cache_find(key)) {
        std::memcpy(table, c->data.data(), c->data.size() * sizeof(double));
        return STEPS_OK;
    }
    CachedTable t;
    t.data.resize((size_t)ng * ng * ng * 3);
    if ((rc = build_t3_table_uncached(is_periodic, L, t.data.data(), &ng))) return rc;
    std::memcpy(table, t.data.data(), t.data.size() * sizeof(double));
    cache_put(key, std::move(t));
    return STEPS_OK;
}

int build_s1r2_table(int is_periodic, double L, double Rsim, double *table, int *nrho, int *nz) {
    int a = 0, b = 0;
    int rc = build_s1r2_table_uncached(is_periodic, L, Rsim, nullptr, &a, &b);
    if (nrho) *nrho = a;
    if (nz) *nz = b;
    if (rc || !table) return rc;
    const TableKey key{2, is_periodic, a, b, bits(L), bits(Rsim)};
    std::lock_guard<std::mutex> lk(g_table_mu);
    if (const CachedTable *c = cache_find(key)) {
        std::memcpy(table, c->data.data(), c->data.size() * sizeof(double));
        return STEPS_OK;
    }
    CachedTable t;
    t.data.resize((size_t)std::max(a, 0) * b * 2);
    if ((rc = build_s1r2_table_uncached(is_periodic, L, Rsim, t.data.data(), &a, &b))) return rc;
    std::memcpy(table, t.data.data(), t.data.size() * sizeof(double));
    cache_put(key, std::move(t));
    return STEPS_OK;
}

int build_radial_table(double R, double Lz, int size, int accuracy, double *table) {
    if (!table || size < 1 || accuracy < 1 || !(R > 0) || !(Lz > 0))
        return fail(STEPS_ERR_ARG, "bad radial table arguments");
    const TableKey key{3, 0, size, accuracy, bits(R), bits(Lz)};
    std::lock_guard<std::mutex> lk(g_table_mu);
    if (const CachedTable *c = cache_find(key)) {
        std::memcpy(table, c->data.data(), c->data.size() * sizeof(double));
        return STEPS_OK;
    }
    CachedTable t;
    t.data.resize((size_t)size);
    int rc = build_radial_table_uncached(R, Lz, size, accuracy, t.data.data());
    if (rc) return rc;
    std::memcpy(table, t.data.data(), t.data.size() * sizeof(double));
    cache_put(key, std::move(t));
    return STEPS_OK;
}

}  // namespace steps
