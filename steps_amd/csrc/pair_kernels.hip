// pair_kernels.hip -- direct O(N^2) pair-force kernels for gfx950 (MI355X).
//
// Work decomposition (one force evaluation = rows id_min..id_max against all
// N sources, self pair included as in forces_cuda.cu:531 and forces.cc:538):
//   * a workgroup of 256 lanes (4 wave64) owns kBlockRows = 256*NI target rows,
//     NI rows per lane held in VGPRs together with their FP64 accumulators;
//   * blockIdx.y selects one of `splits` contiguous source ranges, so small N
//     still fills 256 CUs; every (row-block, split) writes its FP64 partial
//     sums to partials[split][3][nrows] and a second kernel adds the splits in
//     fixed order (deterministic, no atomics) and adds the external term;
//   * the source particle j is wave-uniform, so its record (x,y,z,m) is
//     fetched with scalar loads into SGPRs and reaches all 64 lanes as a VALU
//     scalar operand -- the broadcast costs no VALU or LDS instruction;
//   * sources are processed in tiles of 64; a tiny pre-pass stores each tile's
//     largest softening length, so one v_cmp against (h_i + h_max,tile)^2
//     replaces the per-pair h_i + h_j test, and the exact cubic-spline branch
//     (r < h_i + h_j, forces_cuda.cu:466-519, the reference's own formula)
//     runs only when some lane of the wave is that close;
//   * 1/r^3 uses v_rsq_f64 (2^-24 accurate on gfx950) plus one first-order
//     Newton correction; the R3 and nearest-image FP64 kernels read source
//     copies whose mass slot holds -1.5 m, which folds the mass into that
//     correction (mr3_scaled: 5 VALU instead of 6).
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace steps {

namespace {

struct KScalars {
    int64_t N;
    int64_t id_min;
    int64_t nrows;
    int64_t jchunk;
    double L, halfL;
    // T3 table: grid nodes per axis, first x/y node of the device layout
    int ng, t3_lo;
    // S1R2 table
    int nrho, nz;
    double rho_max;
    // S1R2 image sum
    int images;
    double image_cut;
    // Newton-correction constant, passed in so that it lives in a register
    // for the whole loop instead of being rematerialised per pair
    double c15;
    // -5/3, the constant of mr3_scaled
    double cm53;
    // reciprocals of the table grid spacings: T3 Ng/L, S1R2 1/drho and Nz/Lz
    double inv_gs, inv_drho, inv_dzg;
    double last_r;  // S1R2: index of the last rho node, Nrho - 1 (Nrho >= 2)
    // FP32 copies for the mixed-precision S1R2 pair
    float Lf, halfLf, rho_maxf, inv_drhof, inv_dzgf, last_rf;
    float inv_gsf;  // FP32 copy of inv_gs for the mixed-precision T3 pair
    float wrap_epsf;  // 1e-6 L: the band around |d| = L/2 where an FP32 displacement may wrap otherwise
};

// Everything one evaluation's pair kernel reads: rows (caller order),
// sources (the same arrays for FP64; spatially ordered copies for the mixed
// path, which jrel and the tiles describe), tables, partial sums, scalars.
struct PairArgs {
    const double4 *irec;
    const double *isoft;
    const double4 *jrec;
    const double *soft;
    const float4 *jrel;
    const double4 *tiles;
    const double4 *tile_ext;  // half extents of the tiles' bounding boxes (sorted sources)
    const double4 *sub;       // centres of the 8-source sub-tiles (mixed S1R2 CIC / TSC)
    const double *table;
    const float2 *tab32;      // FP32 S1R2 table (mixed)
    const float4 *t3tab32;    // FP32 T3 table, (Dx, Dy, Dz, 0) per node (mixed)
    const uint32_t *keys;     // pencil keys of the sorted particles (T3 Ewald), else null
    double *part;
    KScalars k;
};

// One evaluation inside a packed launch: its arguments, its row-block count
// and the first flat workgroup index it owns.
struct PackSeg {
    PairArgs a;
    int nbx;
    int blk_begin;
};

// ---------------------------------------------------------------------------
// FP64 building blocks
// ---------------------------------------------------------------------------

// Copy a wave-uniform constant into a VGPR once, so that an FMA taking it and
// another SGPR constant (VOP3 reads at most one SGPR on gfx9) does not
// re-materialise it with a v_mov in every pair.
__device__ __forceinline__ double in_vgpr(double x) {
    double r;
    asm("v_mov_b64 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// r2^(-3/2): v_rsq_f64 seed y (relative error |d| <= 5.2e-8 on gfx950,
// probed), e = 1 - r2*y^2 ~= -2d, y^3 * (1 - e)^(-3/2) ~= y^3 * (1 + 1.5 e).
// The dropped terms, 1.875 e^2 + ..., bound the relative error by 2.1e-14
// (one-sided: the result is never too large); one correction term instead of
// two saves an FMA per pair (3.5% of the R3 pair kernel, measured).
// r2 == 0 gives NaN: the caller's softening branch owns it.
__device__ __forceinline__ double inv_r3(double r2, double c15) {
    const double y = __builtin_amdgcn_rsq(r2);
    const double t = y * y;
    const double e = fma(-r2, t, 1.0);
    return (t * y) * fma(e, c15, 1.0);
}

// m r2^(-3/2) for a source whose record carries mp = -1.5 m (the FP64 R3 and
// nearest-image kernels' source copies, tiles_kernel): with y = rsq(r2) and
// t = y^2, mp y^3 (r2 t - 5/3) = m y^3 (2.5 - 1.5 r2 t) = m y^3 (1 + 1.5 e),
// e = 1 - r2 t, which is inv_r3's first-order correction times m, in 5 VALU
// (y m', t (y m'), r2 t - 5/3, the product, and t) instead of inv_r3's 5 plus
// the mass multiply.  r2 t - 5/3 lies near -2/3, so its rounding adds 1 ulp.
__device__ __forceinline__ double mr3_scaled(double r2, double mp, double cm53) {
    const double y = __builtin_amdgcn_rsq(r2);
    const double t = y * y;
    const double b = t * (y * mp);
    return b * fma(r2, t, cm53);
}

// The mass of a source whose record carries -1.5 m (within 2 ulp of m); only
// the rare pairs inside the softening length need it.
__device__ __forceinline__ double unscaled_mass(double mp) { return mp * (-2.0 / 3.0); }

// forces_cuda.cu:466-519, written out as the reference does (beta powers and
// divisions); only reached for r < beta, so its cost does not matter.
__device__ __forceinline__ double softening_exact(double r, double beta) {
    const double half_beta = beta * 0.5;
    const double r2 = r * r;
    const double r3 = r2 * r;
    const double inv_r3v = 1.0 / r3;
    if (r >= beta) return inv_r3v;
    const double b2 = beta * beta, b3 = b2 * beta, b4 = b2 * b2, b5 = b4 * beta, b6 = b3 * b3;
    if (r > half_beta) {
        const double c0 = -32.0 / (3.0 * b6), c1 = 38.4 / b5, c2 = -48.0 / b4;
        const double c3 = 64.0 / (3.0 * b3), c4 = -1.0 / 15.0;
        return c0 * r3 + c1 * r2 + c2 * r + c3 + c4 * inv_r3v;
    }
    const double c0 = 32.0 / b6, c1 = -38.4 / b5, c2 = (32.0 / 3.0) / b3;
    return c0 * r3 + c1 * r2 + c2;
}

__device__ __forceinline__ int imodp(int i, int n) {
    int r = i % n;  // forces_cuda.cu:87-92
    return (r < 0) ? (r + n) : r;
}

// imodp(i, n) without an integer division for i in [-n, 2n), which covers
// every index a nearest-image displacement produces; other values (positions
// outside the box) take the general remainder in a branch waves rarely enter.
__device__ __forceinline__ int wrap_index(int i, int n) {
    int r = (i < 0) ? i + n : ((i >= n) ? i - n : i);
    if ((unsigned)r >= (unsigned)n) r = imodp(i, n);
    return r;
}

// One period step for an index known to lie in [-n, 2n).
__device__ __forceinline__ int wrap_once(int i, int n) {
    return (i < 0) ? i + n : ((i >= n) ? i - n : i);
}

// x^(-1/2) for x > 0: v_rsq_f64 seed and a first-order Newton correction,
// y * (1 - e)^(-1/2) ~= y * (1 + e/2) with e = 1 - x*y^2; the dropped
// 3e^2/8 bounds the relative error by 4e-15.
__device__ __forceinline__ double rsqrt_f64(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x, y * y, 1.0);
    return fma(y * e, 0.5, y);
}

// Eight-byte aligned pair of doubles: a 16-byte load at any 8-byte boundary.
typedef double double2u __attribute__((ext_vector_type(2), aligned(8)));

// Catmull-Rom weights and first node of one axis (forces_cuda.cu:93-111):
// u in grid units from the first cell centre, nodes i0-1 .. i0+2.
__device__ __forceinline__ int t3_axis(double d, const KScalars &k, double w[4]) {
    const double u = (d + k.L * 0.5) * k.inv_gs - 0.5;
    const double uf = floor(u);
    const double t = u - uf;
    const double t2 = t * t, t3 = t * t2;
    w[0] = -0.5 * t3 + t2 - 0.5 * t;
    w[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
    w[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * t;
    w[3] = 0.5 * t3 - 0.5 * t2;
    return wrap_index((int)uf, k.ng);
}

// forces_cuda.cu:93-167: tricubic Catmull-Rom interpolation of the T3 table
// (cell-centred grid, periodic index wrap), read from the device layout of
// t3_device_layout: x and y nodes lo..Ng-1, a node i < lo read at its mirror
// Ng-1-i with its own component negated (the sign rides on that axis'
// weight, so x-negated and y-negated values never need a select), and each
// stencil's four z nodes as twelve contiguous doubles (six 16-byte loads).
// The contraction runs z, then y, then x: Sum_i wx_i Sum_j wy_j Sum_k wz_k T.
// The reference divides by the grid spacing; multiplying by its reciprocal
// moves u by at most an ulp, and the interpolant is continuous in u.
struct T3Stencil {
    double wx[4], wy[4], wz[4], wxs[4], wys[4];
    unsigned mx[4], my[4];  // layout x / y node of each stencil plane
    int iz0;                // first z node + 1 = layout z index of the run
};

__device__ __forceinline__ void t3_stencil(const KScalars &k, double dx, double dy, double dz, T3Stencil &s) {
    const int Ng = k.ng, lo = k.t3_lo;
    const int ix0 = t3_axis(dx, k, s.wx);
    const int iy0 = t3_axis(dy, k, s.wy);
    s.iz0 = t3_axis(dz, k, s.wz);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int jx = wrap_once(ix0 - 1 + q, Ng);  // Ng >= 3
        const int jy = wrap_once(iy0 - 1 + q, Ng);
        s.mx[q] = (unsigned)((jx >= lo) ? jx - lo : Ng - 1 - lo - jx);
        s.my[q] = (unsigned)((jy >= lo) ? jy - lo : Ng - 1 - lo - jy);
        s.wxs[q] = (jx >= lo) ? s.wx[q] : -s.wx[q];
        s.wys[q] = (jy >= lo) ? s.wy[q] : -s.wy[q];
    }
}

// The staged contraction over the 16 (x, y) rows; load(i, j, v) fills v with
// the stencil's twelve z-run values of table row (mx[i], my[j]), node-major
// (v[3q + c] = component c of z node q).
template <class Load>
__device__ __forceinline__ void t3_contract(const T3Stencil &s, Load load, double &Dx, double &Dy, double &Dz) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double v[12];
            load(i, j, v);
            double t0 = s.wz[0] * v[0], t1 = s.wz[0] * v[1], t2 = s.wz[0] * v[2];
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                t0 = fma(s.wz[q], v[3 * q], t0);
                t1 = fma(s.wz[q], v[3 * q + 1], t1);
                t2 = fma(s.wz[q], v[3 * q + 2], t2);
            }
            u0 = fma(s.wy[j], t0, u0);
            u1 = fma(s.wys[j], t1, u1);
            u2 = fma(s.wy[j], t2, u2);
        }
        s0 = fma(s.wxs[i], u0, s0);
        s1 = fma(s.wx[i], u1, s1);
        s2 = fma(s.wx[i], u2, s2);
    }
    Dx = s0;
    Dy = s1;
    Dz = s2;
}

// The same contraction reading the LDS window: (x, y) of the four z nodes as
// aligned 16-byte pairs, and the z components of six nodes from the even
// start a = iz0 & ~1 as three aligned 16-byte pairs (an 8-byte-aligned pair
// read costs ~2.8x an aligned one in LDS, probed).  wz5 carries the four
// weights at the nodes' places and 0 at the extra one, so each FMA chain is
// the four-node chain (a zero-weight term adds a signed zero at most).
template <class Load>
__device__ __forceinline__ void t3_contract_win(const T3Stencil &s, const double wz5[5], Load load, double &Dx,
                                                double &Dy, double &Dz) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double u0 = 0.0, u1 = 0.0, u2 = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double vxy[8], vz[6];
            load(i, j, vxy, vz);
            double t0 = s.wz[0] * vxy[0], t1 = s.wz[0] * vxy[1], t2 = wz5[0] * vz[0];
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                t0 = fma(s.wz[q], vxy[2 * q], t0);
                t1 = fma(s.wz[q], vxy[2 * q + 1], t1);
            }
#pragma unroll
            for (int q = 1; q < 5; ++q) t2 = fma(wz5[q], vz[q], t2);
            u0 = fma(s.wy[j], t0, u0);
            u1 = fma(s.wys[j], t1, u1);
            u2 = fma(s.wy[j], t2, u2);
        }
        s0 = fma(s.wxs[i], u0, s0);
        s1 = fma(s.wx[i], u1, s1);
        s2 = fma(s.wx[i], u2, s2);
    }
    Dx = s0;
    Dy = s1;
    Dz = s2;
}

// Twelve contiguous doubles (a z run of the device layout) as six 16-byte loads.
__device__ __forceinline__ void t3_load_aos(const double *p, double *v) {
    const double2u *r = (const double2u *)p;
#pragma unroll
    for (int h = 0; h < 6; ++h) {
        const double2u a = r[h];
        v[2 * h] = a.x;
        v[2 * h + 1] = a.y;
    }
}

__device__ __forceinline__ void t3_interp(const double *__restrict__ T, const KScalars &k, double dx,
                                          double dy, double dz, double &Dx, double &Dy, double &Dz) {
    T3Stencil s;
    t3_stencil(k, dx, dy, dz, s);
    const unsigned NY = (unsigned)(k.ng - k.t3_lo), NZ = (unsigned)(k.ng + 3);
    // z node iz0-1+q sits at layout index iz0+q (one ghost below)
    t3_contract(
        s, [&](int i, int j, double *v) { t3_load_aos(T + ((s.mx[i] * NY + s.my[j]) * NZ + (unsigned)s.iz0) * 3u, v); },
        Dx, Dy, Dz);
}

// S1R2 table in the device layout of s1r2_device_layout: one row per rho
// node -1 .. Nrho (the ghost rows repeat the edge nodes, which is the
// reference's clamp of the rho index), each row the z nodes -1 .. Nz + 1
// (periodic ghosts), each node (D_rho, D_z) as 16 bytes.  Node (ir, iz) sits
// at ((ir + 1) * (Nz + 3) + iz + 1) * 16 bytes; s1r2_row(T, k, ir) + iz + 1
// points at it.
// Row offsets use the full-rate 24-bit multiply (v_mul_u32_u24; the 32-bit
// v_mul_lo_u32 issues at a quarter of the rate): make_plan keeps Nrho + 2 and
// Nz + 3 below 2^24, and the 2^28-node cap keeps the products in 32 bits.
__device__ __forceinline__ const double2 *s1r2_row(const double *__restrict__ T, const KScalars &k, int ir) {
    return (const double2 *)((const char *)T + __umul24((unsigned)(ir + 1), (unsigned)(k.nz + 3)) * 16u);
}

// A z index in [lo_ok, hi_ok] needs no wrap (the ghost rows cover it); any
// other value (a position outside the box) takes the general periodic
// remainder, so every read stays inside the table.
__device__ __forceinline__ int s1r2_zfix(int iz, int lo_ok, int hi_ok, int Nz) {
    if (iz < lo_ok || iz > hi_ok) iz = imodp(iz, Nz);
    return iz;
}

// forces_cuda.cu:188-455: NGP / CIC / TSC interpolation of (D_rho, D_z) at
// (rho, z), rho clamped to [0, rho_max], z periodic.  Returns q = D_rho / rho
// (0 on the axis) and D_z: the caller forms D_x = q dx, D_y = q dy.
// Divisions by the grid spacings become multiplications by their reciprocals
// and rho, 1/rho come from one reciprocal square root (each within an ulp or
// two of the reference's sqrt and divides; the interpolants are continuous).
// The interpolation split into its index math and a gather over rho rows:
// row(ir) returns the z run of rho node ir (device layout: z nodes -1 .. Nz+1),
// from the global table or from an LDS window of rows; s1r2_rows gives the
// rho nodes [lo, hi] a pair's stencil reads.
template <int ORDER>
__device__ __forceinline__ void s1r2_rows(const KScalars &k, double rho, int &lo, int &hi) {
    if (ORDER == 0) {
        lo = hi = (int)floor(fmin(rho, k.rho_max) * k.inv_drho + 0.5);
    } else if (ORDER == 2) {
        lo = (int)floor(fmin(rho * k.inv_drho, k.last_r));
        hi = lo + 1;
    } else {
        const int jr = (int)floor(fmin(rho * k.inv_drho, k.last_r) + 0.5);
        lo = jr - 1;
        hi = jr + 1;
    }
}

template <int ORDER, class Row>
__device__ __forceinline__ void s1r2_gather(Row row, const KScalars &k, double rho, double uz, double &Dr,
                                            double &Dzv) {
    const int Nz = k.nz;
    if (ORDER == 0) {
        const int ir = (int)floor(fmin(rho, k.rho_max) * k.inv_drho + 0.5);  // in [0, Nrho-1]
        const int iz = s1r2_zfix((int)floor(uz + 0.5), 0, Nz, Nz);
        const double2 t = row(ir)[iz + 1];
        Dr = t.x;
        Dzv = t.y;
    } else if (ORDER == 2) {
        // rho beyond rho_max clamps to the last node (forces_cuda.cu:225-239);
        // at ur = Nrho-1 the upper node is the ghost copy of the last one
        const double ur = fmin(rho * k.inv_drho, k.last_r);
        const double urf = floor(ur);
        const double fr = ur - urf;
        const int ir0 = (int)urf;
        const double uzf = floor(uz);
        const double fz = uz - uzf;
        const int iz0 = s1r2_zfix((int)uzf, -1, Nz - 1, Nz);
        const double2 *r0 = row(ir0) + (iz0 + 1);  // z nodes iz0, iz0 + 1
        const double2 *r1 = row(ir0 + 1) + (iz0 + 1);
        const double2 t00 = r0[0], t01 = r0[1], t10 = r1[0], t11 = r1[1];
        // bilinear as two lerps in rho and one in z (same interpolant as the
        // reference's four weights, fewer operations)
        const double a0 = fma(fr, t10.x - t00.x, t00.x), b0 = fma(fr, t11.x - t01.x, t01.x);
        const double a1 = fma(fr, t10.y - t00.y, t00.y), b1 = fma(fr, t11.y - t01.y, t01.y);
        Dr = fma(fz, b0 - a0, a0);
        Dzv = fma(fz, b1 - a1, a1);
    } else {
        const double ur = fmin(rho * k.inv_drho, k.last_r);
        const double urf = floor(ur + 0.5), uzf = floor(uz + 0.5);
        const int jr = (int)urf;  // in [0, Nrho-1]; nodes jr-1 .. jr+1 (ghosts at the ends)
        const int jz = s1r2_zfix((int)uzf, 0, Nz, Nz);
        const double sr = ur - urf, sz = uz - uzf;
        const double wr[3] = {0.5 * (0.5 - sr) * (0.5 - sr), 0.75 - sr * sr,
                              0.5 * (0.5 + sr) * (0.5 + sr)};
        const double wz[3] = {0.5 * (0.5 - sz) * (0.5 - sz), 0.75 - sz * sz,
                              0.5 * (0.5 + sz) * (0.5 + sz)};
        Dr = 0.0;
        Dzv = 0.0;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double2 *rr = row(jr - 1 + c) + jz;  // rho node jr-1+c, z nodes jz-1 .. jz+1
            const double w0 = wr[c] * wz[0], w1 = wr[c] * wz[1], w2 = wr[c] * wz[2];
            const double2 a = rr[0], b = rr[1], d = rr[2];
            Dr += w0 * a.x + w1 * b.x + w2 * d.x;
            Dzv += w0 * a.y + w1 * b.y + w2 * d.y;
        }
    }
}

// rho = |(dx, dy)|, its reciprocal, and the z grid coordinate of dz.
__device__ __forceinline__ void s1r2_coords(const KScalars &k, double dx, double dy, double dz, double &rho,
                                            double &inv_rho, double &uz) {
    const double rxy2 = fma(dx, dx, dy * dy);
    inv_rho = (rxy2 > 0.0) ? rsqrt_f64(rxy2) : 0.0;
    rho = rxy2 * inv_rho;
    uz = fma(dz + k.halfL, k.inv_dzg, -0.5);
}

template <int ORDER>
__device__ __forceinline__ void s1r2_interp(const double *__restrict__ T, const KScalars &k, double dx, double dy,
                                            double dz, double &q, double &Dz) {
    double rho, inv_rho, uz, Dr, Dzv;
    s1r2_coords(k, dx, dy, dz, rho, inv_rho, uz);
    s1r2_gather<ORDER>([&](int ir) { return s1r2_row(T, k, ir); }, k, rho, uz, Dr, Dzv);
    q = Dr * inv_rho;
    Dz = Dzv;
}

// Softened coefficient for a pair, with the tile bound B = (h_i + h_max,tile)^2:
// only pairs with r2 < B can be inside h_i + h_j, so the exact test (and the
// softening_exact evaluation) runs in a branch that almost no wave takes.
__device__ __forceinline__ double pair_w_tile(double r2, double B, double hi,
                                              const double *__restrict__ soft, int64_t j,
                                              const KScalars &k) {
    double w = inv_r3(r2, k.c15);
    if (r2 < B) {
        const double beta = hi + soft[j];
        if (r2 < beta * beta) w = softening_exact(sqrt(r2), beta);
    }
    return w;
}

// pair_w_tile times the source mass, for a source carrying mp = -1.5 m.
__device__ __forceinline__ double pair_s_tile(double r2, double B, double hi, double mp,
                                              const double *__restrict__ soft, int64_t j, const KScalars &k) {
    double s = mr3_scaled(r2, mp, k.cm53);
    if (r2 < B) {
        const double beta = hi + soft[j];
        if (r2 < beta * beta) s = unscaled_mass(mp) * softening_exact(sqrt(r2), beta);
    }
    return s;
}

// An LDS copy of the S1R2 table rows for rho nodes w0 .. w0 + nw - 1 (each
// the full z run of the device layout), valid when ok.
struct S1R2Win {
    const double2 *win;
    int w0, nw;
    bool ok;
};

// One pair (i = lane row, j = uniform source) for variant V, FP64.  WIN: an
// S1R2 lookup pair reads its rho rows from the window when they all lie in
// it, else from the global table (the same values either way).  MS: the
// source record carries -1.5 m instead of m (R3 and nearest-image variants;
// the FP64 kernel's source copies, mr3_scaled).
// One nearest-image wrap, d - L sign(d) when |d| > L/2 (forces_cuda.cu:
// 622-627, 833-836; positions in [0, L) need no second one), as a multiply
// by 0 or 1 in the FMA rather than two 64-bit selects: fma(-1, c, d) is the
// subtraction d - c, fma(-0, c, d) is d.
__device__ __forceinline__ double wrap_half(double d, const KScalars &k) {
    return fma(fabs(d) > k.halfL ? -1.0 : 0.0, copysign(k.L, d), d);
}

template <int V, bool WIN = false, bool MS = false>
__device__ __forceinline__ void pair_f64(const KScalars &k, const double *__restrict__ table,
                                         const double *__restrict__ soft, const int64_t j,
                                         const double4 pj, const double xi, const double yi,
                                         const double zi, const double hi, const double B,
                                         double &fx, double &fy, double &fz, const S1R2Win sw = S1R2Win{}) {
    double dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    if (V == PV_R3) {
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        const double s = MS ? pair_s_tile(r2, B, hi, pj.w, soft, j, k) : pj.w * pair_w_tile(r2, B, hi, soft, j, k);
        fx = fma(s, dx, fx);
        fy = fma(s, dy, fy);
        fz = fma(s, dz, fz);
    } else if (V == PV_T3_NEAREST || V == PV_T3_EWALD) {
        // forces_cuda.cu:622-627: d - L*d/|d| when |d| > L/2 (d/|d| is exactly +-1)
        dx = wrap_half(dx, k);
        dy = wrap_half(dy, k);
        dz = wrap_half(dz, k);
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        if (V == PV_T3_NEAREST) {
            const double s = MS ? pair_s_tile(r2, B, hi, pj.w, soft, j, k) : pj.w * pair_w_tile(r2, B, hi, soft, j, k);
            fx = fma(s, dx, fx);
            fy = fma(s, dy, fy);
            fz = fma(s, dz, fz);
        } else {
            const double w = pair_w_tile(r2, B, hi, soft, j, k);
            double Dx, Dy, Dz;
            t3_interp(table, k, dx, dy, dz, Dx, Dy, Dz);
            fx = fma(pj.w, fma(w, dx, -Dx), fx);
            fy = fma(pj.w, fma(w, dy, -Dy), fy);
            fz = fma(pj.w, fma(w, dz, -Dz), fz);
        }
    } else if (V == PV_S1R2_IMAGES) {
        // forces_cuda.cu:723-737: raw dz, images m in [-(P+1), P+1], cut |dz+mL| < (P+0.6)L
        const double rxy2 = fma(dx, dx, dy * dy);
        double sx = 0.0, sz = 0.0;
        for (int m = -k.images; m <= k.images; ++m) {
            const double dze = dz + (double)m * k.L;
            if (fabs(dze) < k.image_cut) {
                const double r2 = fma(dze, dze, rxy2);
                const double s = pj.w * pair_w_tile(r2, B, hi, soft, j, k);
                sx += s;
                sz = fma(s, dze, sz);
            }
        }
        fx = fma(sx, dx, fx);
        fy = fma(sx, dy, fy);
        fz += sz;
    } else {
        // forces_cuda.cu:833-836: nearest z-image only
        dz = wrap_half(dz, k);
        const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
        if (V == PV_S1R2_NEAREST) {
            const double s = MS ? pair_s_tile(r2, B, hi, pj.w, soft, j, k) : pj.w * pair_w_tile(r2, B, hi, soft, j, k);
            fx = fma(s, dx, fx);
            fy = fma(s, dy, fy);
            fz = fma(s, dz, fz);
        } else {
            const double w = pair_w_tile(r2, B, hi, soft, j, k);
            // m (w d - D) with D_x = q dx, D_y = q dy: x and y share m (w - q)
            constexpr int ORD = (V == PV_S1R2_NGP) ? 0 : ((V == PV_S1R2_CIC) ? 2 : 4);
            double rho, inv_rho, uz, Dr, Dzv;
            s1r2_coords(k, dx, dy, dz, rho, inv_rho, uz);
            bool in = false;
            if (WIN && sw.ok) {
                int lo, hi;
                s1r2_rows<ORD>(k, rho, lo, hi);
                in = lo >= sw.w0 && hi < sw.w0 + sw.nw;
            }
            if (in) {
                const unsigned nz3 = (unsigned)(k.nz + 3);
                s1r2_gather<ORD>([&](int ir) { return sw.win + __umul24((unsigned)(ir - sw.w0), nz3); }, k, rho, uz, Dr,
                                 Dzv);
            } else {
                s1r2_gather<ORD>([&](int ir) { return s1r2_row(table, k, ir); }, k, rho, uz, Dr, Dzv);
            }
            const double q = Dr * inv_rho, Dz = Dzv;
            const double sxy = pj.w * (w - q);
            fx = fma(sxy, dx, fx);
            fy = fma(sxy, dy, fy);
            fz = fma(pj.w, fma(w, dz, -Dz), fz);
        }
    }
}

// ---------------------------------------------------------------------------
// FP64 pair kernel: sources in tiles of kTile (64); per tile each lane forms
// B = (h_i + h_max,tile)^2 once, then UNR sources at a time come in through
// scalar loads and are applied to the NI rows of the lane.
// ---------------------------------------------------------------------------
// R3 pair of a tile that no row of the wave can reach within its softening
// bound: the plain r^-3 law, no per-pair test (pj.w = -1.5 m).
__device__ __forceinline__ void pair_r3_far(const KScalars &k, const double4 pj, const double xi, const double yi,
                                            const double zi, double &fx, double &fy, double &fz) {
    const double dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double s = mr3_scaled(r2, pj.w, k.cm53);
    fx = fma(s, dx, fx);
    fy = fma(s, dy, fy);
    fz = fma(s, dz, fz);
}

// The nearest-image test (|d| > L/2 -> d -/+ L, forces_cuda.cu:622-627 and
// :833-836) for every source of a tile at once: the displacements from x to
// the tile (centre c, half extent e on this axis) lie in [c - e - x, c + e - x].
// All inside (-L/2, L/2): shift 0; all above L/2: every one wraps to d - L,
// shift -L; all below -L/2: shift +L (margins of rel * L/2 cover the rounding
// of the displacements: 1e-12 for FP64 ones, 1e-5 for FP32 ones).  False when
// the tile straddles a wrap.  d + shift is the reference's d -/+ L to the bit.
__device__ __forceinline__ bool tile_shift(double x, double c, double e, const KScalars &k, double &s,
                                           double rel = 1e-12) {
    const double m = k.halfL * rel;
    s = 0.0;
    if (fabs(x - c) + e < k.halfL - m) return true;
    if (c - e - x > k.halfL + m) {
        s = -k.L;
        return true;
    }
    if (c + e - x < -k.halfL - m) {
        s = k.L;
        return true;
    }
    return false;
}

// x moved by a period towards c when it lies more than half a period away
// (positions inside the box need at most one step).
__device__ __forceinline__ double near_image(double x, double c, const KScalars &k) {
    const double d = x - c;
    return (d > k.halfL) ? x - k.L : ((d < -k.halfL) ? x + k.L : x);
}

// The R3 pair of a far tile with per-axis nearest-image shifts.
__device__ __forceinline__ void pair_far_shift(const KScalars &k, const double4 pj, const double xi, const double yi,
                                               const double zi, const double sx, const double sy, const double sz,
                                               double &fx, double &fy, double &fz) {
    const double dx = (pj.x - xi) + sx, dy = (pj.y - yi) + sy, dz = (pj.z - zi) + sz;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    const double s = mr3_scaled(r2, pj.w, k.cm53);  // pj.w = -1.5 m
    fx = fma(s, dx, fx);
    fy = fma(s, dy, fy);
    fz = fma(s, dz, fz);
}

// Squared distance from a point to an axis-aligned box (centre c, half
// extents e); every source of the tile is at least this far from the point.
__device__ __forceinline__ double box_dist2(double x, double y, double z, const double4 &c, const double4 &e) {
    const double gx = fmax(fabs(x - c.x) - e.x, 0.0);
    const double gy = fmax(fabs(y - c.y) - e.y, 0.0);
    const double gz = fmax(fabs(z - c.z) - e.z, 0.0);
    return fma(gx, gx, fma(gy, gy, gz * gz));
}

// S1R2 CIC / TSC kernels: LDS for the rho rows of one source tile (39 KB:
// four workgroups per CU; 19 rows of a 128-cell z run), filled from the
// global table at each tile; a tile whose rows do not fit reads the global
// table.  The table loads otherwise co-limit these kernels on the vector-L1
// path (4 and 9 16-byte loads per pair).  At N = 524288 (cylinder, sources in
// input order): CIC 345 -> 394 and TSC 192 -> 251 G pairs/s, bitwise the same
// sums; windows of 52 KB (3 workgroups) and 26 KB (6) were slower (360 / 250,
// 390 / 225), and NGP, one load per pair, lost 15% and keeps the global reads.
constexpr int kS1R2WinNodes = 2496;

template <int V, int NI, int UNR>
__device__ __forceinline__ void body_f64(const PairArgs &a, const int bx, const int split) {
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const double4 *__restrict__ tiles = a.tiles;
    const double4 *__restrict__ ext = a.tile_ext;
    const double *__restrict__ table = a.table;
    double *__restrict__ part = a.part;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * (kBlock * NI);
    const int64_t j0 = (int64_t)split * k.jchunk;  // multiple of kTile
    const int64_t j1 = min(k.N, j0 + k.jchunk);

    double xi[NI], yi[NI], zi[NI], hi[NI], fx[NI], fy[NI], fz[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = min(row_base + q * kBlock + tid, k.nrows - 1);
        const double4 p = irec[k.id_min + r];
        xi[q] = p.x;
        yi[q] = p.y;
        zi[q] = p.z;
        hi[q] = isoft[k.id_min + r];
        fx[q] = fy[q] = fz[q] = 0.0;
    }

    // S1R2 lookup: per source tile, the rho rows every pair of the workgroup
    // can read, staged in LDS (see kS1R2WinNodes); range slots alternate
    // between tiles so that one barrier closes the atomics and the previous
    // tile's reads of the window
    constexpr bool SWIN = V == PV_S1R2_CIC || V == PV_S1R2_TSC;
    // sources carry -1.5 m (scaled_sources, mr3_scaled)
    constexpr bool MSC = V == PV_R3 || V == PV_T3_NEAREST || V == PV_S1R2_NEAREST;
    __shared__ __align__(16) double2 s1win[SWIN ? kS1R2WinNodes : 1];
    __shared__ int s1rng[4];
    if (SWIN) {
        if (tid < 4) s1rng[tid] = (tid & 1) ? INT32_MIN : INT32_MAX;
        __syncthreads();
    }

    // loop bounds in 32 bits, relative to the tile: the loop tests stay on the
    // SALU (a 64-bit signed compare would take two VALU instructions per trip)
    const int ntiles = (int)((j1 - j0 + kTile - 1) / kTile);
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t t0 = j0 + (int64_t)ti * kTile;
        const double4 tc = tiles[t0 / kTile];
        const double hmax = tc.w;
        double B[NI];
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            const double b = hi[q] + hmax;
            B[q] = b * b;
        }
        const int cnt = (ti + 1 < ntiles) ? kTile : (int)(j1 - t0);  // sources in this tile
        const double4 *__restrict__ tj = jrec + t0;
        int u0 = 0;
        S1R2Win sw{s1win, 0, 0, false};
        if (SWIN) {
            // rho of every pair with this tile lies between the distances from
            // the row to the near and far sides of the tile's xy box; the
            // stencils read rho nodes floor(rho / drho) - 1 .. floor + 2
            const double4 te = ext[t0 / kTile];
            const double ax = fabs(xi[0] - tc.x), ay = fabs(yi[0] - tc.y);
            const double gx = fmax(ax - te.x, 0.0), gy = fmax(ay - te.y, 0.0);
            const double ex = ax + te.x, ey = ay + te.y;
            const int lo = (int)floor(fmin(sqrt(fma(gx, gx, gy * gy)) * k.inv_drho, k.last_r)) - 1;
            const int hi = (int)floor(fmin(sqrt(fma(ex, ex, ey * ey)) * k.inv_drho, k.last_r)) + 2;
            int *rg = s1rng + 2 * (ti & 1);
            atomicMin(&rg[0], lo);
            atomicMax(&rg[1], hi);
            __syncthreads();  // ranges complete; every wave is done with the previous window
            const int w0 = max(rg[0], -1), w1 = min(rg[1], k.nrho);  // ghost rows -1 and Nrho
            if (tid < 2) s1rng[2 * ((ti + 1) & 1) + tid] = tid ? INT32_MIN : INT32_MAX;  // next tile's slot
            const unsigned nz3 = (unsigned)(k.nz + 3);
            sw.w0 = w0;
            sw.nw = w1 - w0 + 1;
            sw.ok = (int64_t)sw.nw * nz3 <= kS1R2WinNodes;
            if (sw.ok) {
                const double2 *src = s1r2_row(table, k, w0);  // rows w0 .. w1 are contiguous
                const unsigned nn = (unsigned)sw.nw * nz3;
                for (unsigned o = tid; o < nn; o += kBlock) s1win[o] = src[o];
            }
            __syncthreads();  // the window is complete
        }
        if (V == PV_R3 || V == PV_T3_NEAREST || V == PV_S1R2_NEAREST) {
            // Spatially ordered sources: when the tile's bounding box lies
            // outside every lane's softening bound (B widened by 1e-9 for
            // rounding) and, for the periodic variants, every lane sees the
            // whole tile on one side of each wrap (tile_shift), its pairs are
            // the R3 pair, with fixed per-lane shifts where the tile wraps:
            // same arithmetic, no per-pair tests
            const double4 te = ext[t0 / kTile];
            bool far = true, shifted = false;
            double sx[NI], sy[NI], sz[NI];
#pragma unroll
            for (int q = 0; q < NI; ++q) {
                sx[q] = sy[q] = sz[q] = 0.0;
                if (V != PV_R3) {
                    far = far && tile_shift(zi[q], tc.z, te.z, k, sz[q]);
                    if (V == PV_T3_NEAREST)
                        far = far && tile_shift(xi[q], tc.x, te.x, k, sx[q]) && tile_shift(yi[q], tc.y, te.y, k, sy[q]);
                    shifted = shifted || sx[q] != 0.0 || sy[q] != 0.0 || sz[q] != 0.0;
                }
                far = far && box_dist2(xi[q] - sx[q], yi[q] - sy[q], zi[q] - sz[q], tc, te) > B[q] * (1.0 + 1e-9);
            }
            if (__all(far)) {
                if (V != PV_R3 && __any(shifted)) {
                    for (; u0 + UNR <= cnt; u0 += UNR) {
                        double4 pj[UNR];
#pragma unroll
                        for (int u = 0; u < UNR; ++u) pj[u] = tj[u0 + u];
#pragma unroll
                        for (int u = 0; u < UNR; ++u)
#pragma unroll
                            for (int q = 0; q < NI; ++q)
                                pair_far_shift(k, pj[u], xi[q], yi[q], zi[q], sx[q], sy[q], sz[q], fx[q], fy[q], fz[q]);
                    }
                    for (; u0 < cnt; ++u0) {
                        const double4 pj = tj[u0];
#pragma unroll
                        for (int q = 0; q < NI; ++q)
                            pair_far_shift(k, pj, xi[q], yi[q], zi[q], sx[q], sy[q], sz[q], fx[q], fy[q], fz[q]);
                    }
                    continue;
                }
                for (; u0 + UNR <= cnt; u0 += UNR) {
                    double4 pj[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; ++u) pj[u] = tj[u0 + u];
#pragma unroll
                    for (int u = 0; u < UNR; ++u)
#pragma unroll
                        for (int q = 0; q < NI; ++q) pair_r3_far(k, pj[u], xi[q], yi[q], zi[q], fx[q], fy[q], fz[q]);
                }
                for (; u0 < cnt; ++u0) {
                    const double4 pj = tj[u0];
#pragma unroll
                    for (int q = 0; q < NI; ++q) pair_r3_far(k, pj, xi[q], yi[q], zi[q], fx[q], fy[q], fz[q]);
                }
                continue;
            }
        }
        for (; u0 + UNR <= cnt; u0 += UNR) {
            double4 pj[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) pj[u] = tj[u0 + u];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int q = 0; q < NI; ++q)
                    pair_f64<V, SWIN, MSC>(k, table, soft, t0 + u0 + u, pj[u], xi[q], yi[q], zi[q], hi[q], B[q], fx[q],
                                      fy[q], fz[q], sw);
        }
        for (; u0 < cnt; ++u0) {
            const double4 pj = tj[u0];
#pragma unroll
            for (int q = 0; q < NI; ++q)
                pair_f64<V, SWIN, MSC>(k, table, soft, t0 + u0, pj, xi[q], yi[q], zi[q], hi[q], B[q], fx[q], fy[q], fz[q],
                                  sw);
        }
    }

#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = row_base + q * kBlock + tid;
        if (r < k.nrows) {
            double *out = part + (size_t)split * 3u * (size_t)k.nrows;
            out[r] = fx[q];
            out[k.nrows + r] = fy[q];
            out[2 * k.nrows + r] = fz[q];
        }
    }
}

// ---------------------------------------------------------------------------
// mixed-precision pair kernel (R3): FP32 pair arithmetic on positions
// recentred on each 64-source tile (offsets from the tile's bounding-box
// centre, prepared once per evaluation), FP32 sums inside a tile, FP64 across
// tiles; pairs inside the softening radius are recomputed in FP64 from the
// FP64 records, so close-pair forces keep full precision.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void mixed_pair(const float4 pj, const float xr, const float yr,
                                           const float zr, const float B, bool &near, float &ax,
                                           float &ay, float &az) {
    const float dx = pj.x - xr, dy = pj.y - yr, dz = pj.z - zr;
    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
    const float y = __builtin_amdgcn_rsqf(r2);
    const float s = (y * y) * (y * pj.w);
    if (r2 < B) {
        near = true;
    } else {
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
    }
}

// Rows come from (irec, isoft) in the caller's order; sources from the
// spatially ordered copies (jrec, soft) that jrel and the tiles describe.
template <int NI, int UNR>
__device__ __forceinline__ void body_mixed_r3(const PairArgs &a, const int bx, const int split) {
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const float4 *__restrict__ jrel = a.jrel;
    const double4 *__restrict__ tiles = a.tiles;
    double *__restrict__ part = a.part;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * (kBlock * NI);
    const int64_t j0 = (int64_t)split * k.jchunk;  // multiple of kTile
    const int64_t j1 = min(k.N, j0 + k.jchunk);

    double xi[NI], yi[NI], zi[NI], hi[NI], fx[NI], fy[NI], fz[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = min(row_base + q * kBlock + tid, k.nrows - 1);
        const double4 p = irec[k.id_min + r];
        xi[q] = p.x;
        yi[q] = p.y;
        zi[q] = p.z;
        hi[q] = isoft[k.id_min + r];
        fx[q] = fy[q] = fz[q] = 0.0;
    }

    const int ntiles = (int)((j1 - j0 + kTile - 1) / kTile);  // 32-bit loop bounds: SALU tests
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t t0 = j0 + (int64_t)ti * kTile;
        const double4 c = tiles[t0 / kTile];
        float xr[NI], yr[NI], zr[NI], B[NI], ax[NI], ay[NI], az[NI];
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            xr[q] = (float)(xi[q] - c.x);
            yr[q] = (float)(yi[q] - c.y);
            zr[q] = (float)(zi[q] - c.z);
            // conservative FP32 bound: (h_i + h_max)^2 widened by 1e-5 so that
            // FP32 rounding of r2 cannot hide a pair inside the radius
            const float b = (float)((hi[q] + c.w) * (1.0 + 1e-5));
            B[q] = b * b;
            ax[q] = ay[q] = az[q] = 0.0f;
        }
        const int64_t tend = min(j1, t0 + kTile);
        const int cnt = (int)(tend - t0);
        const float4 *__restrict__ tr = jrel + t0;
        bool near = false;
        int j = 0;
        for (; j + UNR <= cnt; j += UNR) {
            float4 pj[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) pj[u] = tr[j + u];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int q = 0; q < NI; ++q) mixed_pair(pj[u], xr[q], yr[q], zr[q], B[q], near, ax[q], ay[q], az[q]);
        }
        for (; j < cnt; ++j) {
            const float4 pj = tr[j];
#pragma unroll
            for (int q = 0; q < NI; ++q) mixed_pair(pj, xr[q], yr[q], zr[q], B[q], near, ax[q], ay[q], az[q]);
        }
        // rare: some lane had a pair within its bound -> redo those pairs in FP64
        if (__builtin_expect(near, 0)) {
            for (int64_t jj = t0; jj < tend; ++jj) {
                const float4 pr = jrel[jj];
                const double4 pd = jrec[jj];
#pragma unroll
                for (int q = 0; q < NI; ++q) {
                    const float dx = pr.x - xr[q], dy = pr.y - yr[q], dz = pr.z - zr[q];
                    const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
                    if (r2 < B[q]) {
                        const double ddx = pd.x - xi[q], ddy = pd.y - yi[q], ddz = pd.z - zi[q];
                        const double rr2 = fma(ddx, ddx, fma(ddy, ddy, ddz * ddz));
                        double w = inv_r3(rr2, k.c15);
                        const double beta = hi[q] + soft[jj];
                        if (rr2 < beta * beta) w = softening_exact(sqrt(rr2), beta);
                        const double s = pd.w * w;
                        fx[q] = fma(s, ddx, fx[q]);
                        fy[q] = fma(s, ddy, fy[q]);
                        fz[q] = fma(s, ddz, fz[q]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            fx[q] += (double)ax[q];
            fy[q] += (double)ay[q];
            fz[q] += (double)az[q];
        }
    }

#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = row_base + q * kBlock + tid;
        if (r < k.nrows) {
            double *out = part + (size_t)split * 3u * (size_t)k.nrows;
            out[r] = fx[q];
            out[k.nrows + r] = fy[q];
            out[2 * k.nrows + r] = fz[q];
        }
    }
}

// ---------------------------------------------------------------------------
// mixed-precision S1xR2 kernel (nearest z-image, optionally with the tabulated
// Ewald correction): the R3 scheme -- FP32 pair arithmetic on tile-recentred
// positions, FP32 sums inside a tile, FP64 across tiles, FP64 redo of the
// pairs inside the softening bound -- with the z-image wrap in FP32 and the
// correction (D_rho, D_z) interpolated in FP32 from an FP32 copy of the table
// (forces_cuda.cu:819-863 with :188-455).
// ---------------------------------------------------------------------------
constexpr int s1r2_order(int V) { return V == PV_S1R2_CIC ? 2 : 4; }

// FP32 copy of the device layout (s1r2_row): same rows, 8-byte nodes.
__device__ __forceinline__ const float2 *s1r2_row32(const float2 *__restrict__ T, const KScalars &k, int ir) {
    return (const float2 *)((const char *)T + __umul24((unsigned)(ir + 1), (unsigned)(k.nz + 3)) * 8u);
}

// s1r2_interp in FP32 (CIC, TSC): returns q = D_rho / rho and D_z.
// An LDS copy of the FP32 S1R2 table rows for rho nodes w0 .. w0 + nw - 1.
struct S1R2Win32 {
    const float2 *win;
    int w0, nw;
    bool ok;
};

// WIN: the rows come from the window when the stencil's rho nodes all lie in
// it, else from the global table (the same values either way).
template <int ORDER, bool WIN = false>
__device__ __forceinline__ void s1r2_interp_f32(const float2 *__restrict__ T, const KScalars &k, float dx,
                                                float dy, float dz, float &q, float &Dz,
                                                const S1R2Win32 sw = S1R2Win32{}) {
    const int Nz = k.nz;
    const float rxy2 = fmaf(dx, dx, dy * dy);
    const float inv_rho = (rxy2 > 0.0f) ? __builtin_amdgcn_rsqf(rxy2) : 0.0f;
    const float rho = rxy2 * inv_rho;
    const float uz = fmaf(dz + k.halfLf, k.inv_dzgf, -0.5f);
    const float lastr = k.last_rf;
    float Dr, Dzv;
    static_assert(ORDER == 2 || ORDER == 4, "FP32 interpolation is CIC or TSC");
    auto gather = [&](auto rowf) {
    if (ORDER == 2) {
        const float ur = fminf(rho * k.inv_drhof, lastr);
        const float urf = floorf(ur);
        const float fr = ur - urf;
        const int ir0 = (int)urf;
        const float uzf = floorf(uz);
        const float fz = uz - uzf;
        const int iz0 = s1r2_zfix((int)uzf, -1, Nz - 1, Nz);
        const float2 *r0 = rowf(ir0) + (iz0 + 1);
        const float2 *r1 = rowf(ir0 + 1) + (iz0 + 1);
        const float2 t00 = r0[0], t01 = r0[1], t10 = r1[0], t11 = r1[1];
        const float a0 = fmaf(fr, t10.x - t00.x, t00.x), b0 = fmaf(fr, t11.x - t01.x, t01.x);
        const float a1 = fmaf(fr, t10.y - t00.y, t00.y), b1 = fmaf(fr, t11.y - t01.y, t01.y);
        Dr = fmaf(fz, b0 - a0, a0);
        Dzv = fmaf(fz, b1 - a1, a1);
    } else {
        const float ur = fminf(rho * k.inv_drhof, lastr);
        const float urf = floorf(ur + 0.5f), uzf = floorf(uz + 0.5f);
        const int jr = (int)urf;
        const int jz = s1r2_zfix((int)uzf, 0, Nz, Nz);
        const float sr = ur - urf, sz = uz - uzf;
        const float wr[3] = {0.5f * (0.5f - sr) * (0.5f - sr), 0.75f - sr * sr, 0.5f * (0.5f + sr) * (0.5f + sr)};
        const float wz[3] = {0.5f * (0.5f - sz) * (0.5f - sz), 0.75f - sz * sz, 0.5f * (0.5f + sz) * (0.5f + sz)};
        Dr = 0.0f;
        Dzv = 0.0f;
#pragma unroll
        for (int row = 0; row < 3; ++row)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float2 t = rowf(jr - 1 + c)[jz + row];
                const float w = wr[c] * wz[row];
                Dr = fmaf(w, t.x, Dr);
                Dzv = fmaf(w, t.y, Dzv);
            }
    }
    };
    bool in = false;
    if (WIN && sw.ok) {
        const float ur = fminf(rho * k.inv_drhof, lastr);
        const int lo = (ORDER == 2) ? (int)floorf(ur) : (int)floorf(ur + 0.5f) - 1;
        in = lo >= sw.w0 && lo + (ORDER == 2 ? 1 : 2) < sw.w0 + sw.nw;
    }
    if (in) {
        const unsigned nz3 = (unsigned)(k.nz + 3);
        gather([&](int ir) { return sw.win + __umul24((unsigned)(ir - sw.w0), nz3); });
    } else {
        gather([&](int ir) { return s1r2_row32(T, k, ir); });
    }
    q = Dr * inv_rho;
    Dz = Dzv;
}

// FP32 displacement with the nearest z-image (forces_cuda.cu:833-836) and r^2;
// T3 nearest image: the nearest image on every axis (forces_cuda.cu:622-627).
// bnd: a wrapped axis lies within 1e-6 L of L/2, where the FP32 displacement
// (off by a few ulp of L) may take the other image than the FP64 one: such a
// pair is left to the FP64 redo with the close pairs, so every pair takes the
// reference's image.  (On the S1R2 N = 524288 cylinder about 4% of the rows
// had one pair in that band, whose flipped image set the p99 error.)
template <int V>
__device__ __forceinline__ float mixed_r2_z(const KScalars &k, const float4 pj, float xr, float yr, float zr,
                                            float &dx, float &dy, float &dz, bool &bnd) {
    dx = pj.x - xr;
    dy = pj.y - yr;
    dz = pj.z - zr;
    bnd = fabsf(fabsf(dz) - k.halfLf) < k.wrap_epsf;
    if (V == PV_T3_NEAREST)
        bnd = bnd || fabsf(fabsf(dx) - k.halfLf) < k.wrap_epsf || fabsf(fabsf(dy) - k.halfLf) < k.wrap_epsf;
    if (V == PV_T3_NEAREST) {
        if (dx > k.halfLf) dx -= k.Lf;
        else if (dx < -k.halfLf) dx += k.Lf;
        if (dy > k.halfLf) dy -= k.Lf;
        else if (dy < -k.halfLf) dy += k.Lf;
    }
    if (dz > k.halfLf) dz -= k.Lf;
    else if (dz < -k.halfLf) dz += k.Lf;
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

template <int V, bool WIN = false>
__device__ __forceinline__ void mixed_pair_z(const KScalars &k, const float2 *__restrict__ T, const float4 pj,
                                             float xr, float yr, float zr, float B, bool &near, float &ax,
                                             float &ay, float &az, const S1R2Win32 sw = S1R2Win32{}) {
    float dx, dy, dz;
    bool bnd;
    const float r2 = mixed_r2_z<V>(k, pj, xr, yr, zr, dx, dy, dz, bnd);
    const bool nr = r2 < B || bnd;  // left to the FP64 redo (r2 may be 0 here)
    near = near || nr;
    const float y = __builtin_amdgcn_rsqf(r2);
    const float y3 = nr ? 0.0f : (y * y) * y;
    const float m = nr ? 0.0f : pj.w;
    if (V == PV_S1R2_NEAREST || V == PV_T3_NEAREST) {
        const float s = y3 * m;
        ax = fmaf(s, dx, ax);
        ay = fmaf(s, dy, ay);
        az = fmaf(s, dz, az);
    } else {
        float q, Dz;
        s1r2_interp_f32<s1r2_order(V), WIN>(T, k, dx, dy, dz, q, Dz, sw);
        const float sxy = m * (y3 - q);
        ax = fmaf(sxy, dx, ax);
        ay = fmaf(sxy, dy, ay);
        az = fmaf(m, fmaf(y3, dz, -Dz), az);
    }
}

// FP32 pairs of a far tile (no pair within the softening bound, no straddled
// wrap) with the nearest image (z only for S1R2, every axis for T3):
// mixed_pair_z's arithmetic without the per-pair wrap and bound tests, the
// wrap applied as a per-lane shift.  With FLUSH the FP32 sums are added to
// the FP64 accumulators f* after every UNR sources: on the T3 nearest-image
// lattice at 64^3 this took the p99 error against FP64 from 1.09e-5 to
// 9.7e-6 for 1% of the speed (on the S1R2 cylinder it changed nothing there
// and cost 4%, so S1R2 keeps one FP32 sum per tile).
template <bool SHIFT, bool XY, bool FLUSH, int NI, int UNR>
__device__ __forceinline__ void mixed_far_tile_z(const float4 *__restrict__ tr, const int cnt, const float *xr,
                                                 const float *yr, const float *zr, const float *sx,
                                                 const float *sy, const float *sz, float *ax, float *ay,
                                                 float *az, double *fx, double *fy, double *fz) {
    auto pair = [&](const float4 pj, const int q) {
        const float dx = (SHIFT && XY) ? (pj.x - xr[q]) + sx[q] : pj.x - xr[q];
        const float dy = (SHIFT && XY) ? (pj.y - yr[q]) + sy[q] : pj.y - yr[q];
        const float dz = SHIFT ? (pj.z - zr[q]) + sz[q] : pj.z - zr[q];
        const float r2 = fmaf(dx, dx, fmaf(dy, dy, dz * dz));
        const float y = __builtin_amdgcn_rsqf(r2);
        const float s = ((y * y) * y) * pj.w;
        ax[q] = fmaf(s, dx, ax[q]);
        ay[q] = fmaf(s, dy, ay[q]);
        az[q] = fmaf(s, dz, az[q]);
    };
    int j = 0;
    for (; j + UNR <= cnt; j += UNR) {
        float4 pj[UNR];
#pragma unroll
        for (int u = 0; u < UNR; ++u) pj[u] = tr[j + u];
#pragma unroll
        for (int u = 0; u < UNR; ++u)
#pragma unroll
            for (int q = 0; q < NI; ++q) pair(pj[u], q);
        if (FLUSH) {
#pragma unroll
            for (int q = 0; q < NI; ++q) {
                fx[q] += (double)ax[q];
                fy[q] += (double)ay[q];
                fz[q] += (double)az[q];
                ax[q] = ay[q] = az[q] = 0.0f;
            }
        }
    }
    for (; j < cnt; ++j) {
        const float4 pj = tr[j];
#pragma unroll
        for (int q = 0; q < NI; ++q) pair(pj, q);
    }
}

template <int V, int NI, int UNR>
__device__ __forceinline__ void body_mixed_z(const PairArgs &a, const int bx, const int split) {
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const float4 *__restrict__ jrel = a.jrel;
    const double4 *__restrict__ tiles = a.tiles;
    const double *__restrict__ table = a.table;
    const float2 *__restrict__ tab32 = a.tab32;
    const double4 *__restrict__ ext = a.tile_ext;
    double *__restrict__ part = a.part;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * (kBlock * NI);
    const int64_t j0 = (int64_t)split * k.jchunk;  // multiple of kTile
    const int64_t j1 = min(k.N, j0 + k.jchunk);

    double xi[NI], yi[NI], zi[NI], hi[NI], fx[NI], fy[NI], fz[NI];
#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = min(row_base + q * kBlock + tid, k.nrows - 1);
        const double4 p = irec[k.id_min + r];
        xi[q] = p.x;
        yi[q] = p.y;
        zi[q] = p.z;
        hi[q] = isoft[k.id_min + r];
        fx[q] = fy[q] = fz[q] = 0.0;
    }

    // TSC: the rho rows of each source tile in LDS, as in body_f64 (FP32
    // nodes: 4992 of them, 39 KB, up to 38 z runs of 131), with the sources
    // in input order (xy-compact tiles on column-ordered input).  At N =
    // 524288: 300 -> 431 G pairs/s, pair-sum p99 error 1.5e-5 -> 2.0e-5
    // (comoving term included).  CIC kept Morton tiles and global reads:
    // with the window it ran at 582 against 681.
    constexpr bool SWIN = V == PV_S1R2_TSC;
    constexpr bool SUBT = V == PV_S1R2_CIC || V == PV_S1R2_TSC;  // sub_tiles(): jrel per 8-source sub-tile
    __shared__ __align__(16) float2 s1win[SWIN ? 2 * kS1R2WinNodes : 1];
    __shared__ int s1rng[4];
    if (SWIN) {
        if (tid < 4) s1rng[tid] = (tid & 1) ? INT32_MIN : INT32_MAX;
        __syncthreads();
    }

    const int ntiles = (int)((j1 - j0 + kTile - 1) / kTile);  // 32-bit loop bounds: SALU tests
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t t0 = j0 + (int64_t)ti * kTile;
        const double4 c = tiles[t0 / kTile];
        // each row's image nearest the tile centre on the periodic axes, so
        // that a close pair across the box edge rounds like any close pair
        // (the per-pair wrap below then fires only for distant pairs)
        double xw[NI], yw[NI], zw[NI];
        float xr[NI], yr[NI], zr[NI], B[NI], ax[NI], ay[NI], az[NI];
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            xw[q] = xi[q];
            yw[q] = yi[q];
            zw[q] = zi[q];
            if (V == PV_T3_NEAREST) {
                xw[q] = near_image(xw[q], c.x, k);
                yw[q] = near_image(yw[q], c.y, k);
            }
            zw[q] = near_image(zw[q], c.z, k);
            xr[q] = (float)(xw[q] - c.x);
            yr[q] = (float)(yw[q] - c.y);
            zr[q] = (float)(zw[q] - c.z);
            const float b = (float)((hi[q] + c.w) * (1.0 + 1e-5));
            B[q] = b * b;
            ax[q] = ay[q] = az[q] = 0.0f;
        }
        const int64_t tend = min(j1, t0 + kTile);
        const int cnt = (int)(tend - t0);
        const float4 *__restrict__ tr = jrel + t0;
        S1R2Win32 sw{s1win, 0, 0, false};
        if (SWIN) {
            static_assert(!SWIN || NI == 1, "one row per lane");
            const double4 te = ext[t0 / kTile];
            const double axd = fabs(xi[0] - c.x), ayd = fabs(yi[0] - c.y);
            const double gx = fmax(axd - te.x, 0.0), gy = fmax(ayd - te.y, 0.0);
            const double ex = axd + te.x, ey = ayd + te.y;
            const int lo = (int)floor(fmin(sqrt(fma(gx, gx, gy * gy)) * k.inv_drho, k.last_r)) - 1;
            const int hi = (int)floor(fmin(sqrt(fma(ex, ex, ey * ey)) * k.inv_drho, k.last_r)) + 2;
            int *rg = s1rng + 2 * (ti & 1);
            atomicMin(&rg[0], lo);
            atomicMax(&rg[1], hi);
            __syncthreads();  // ranges complete; every wave is done with the previous window
            const int w0 = max(rg[0], -1), w1 = min(rg[1], k.nrho);
            if (tid < 2) s1rng[2 * ((ti + 1) & 1) + tid] = tid ? INT32_MIN : INT32_MAX;  // next tile's slot
            const unsigned nz3 = (unsigned)(k.nz + 3);
            sw.w0 = w0;
            sw.nw = w1 - w0 + 1;
            sw.ok = (int64_t)sw.nw * nz3 <= 2 * kS1R2WinNodes;
            if (sw.ok) {
                const float2 *src = s1r2_row32(tab32, k, w0);
                const unsigned nn = (unsigned)sw.nw * nz3;
                for (unsigned o = tid; o < nn; o += kBlock) s1win[o] = src[o];
            }
            __syncthreads();  // the window is complete
        }
        if (V == PV_S1R2_NEAREST || V == PV_T3_NEAREST) {
            // Far tile (decided in FP64, with margins far above the FP32
            // rounding of the displacements): every lane sees the whole tile
            // on one side of each wrap and outside 1.001 x its bound
            constexpr bool XY = V == PV_T3_NEAREST;
            const double4 te = ext[t0 / kTile];
            bool far = true, shifted = false;
            float sx[NI], sy[NI], sz[NI];
#pragma unroll
            for (int q = 0; q < NI; ++q) {
                double hx = 0.0, hy = 0.0, hz = 0.0;
                far = far && tile_shift(zw[q], c.z, te.z, k, hz, 1e-5);
                if (XY)
                    far = far && tile_shift(xw[q], c.x, te.x, k, hx, 1e-5) && tile_shift(yw[q], c.y, te.y, k, hy, 1e-5);
                far = far && box_dist2(xw[q] - hx, yw[q] - hy, zw[q] - hz, c, te) > (double)B[q] * (1.0 + 1e-3);
                sx[q] = (float)hx;  // 0, -L or +L: (float)(-L) is -Lf
                sy[q] = (float)hy;
                sz[q] = (float)hz;
                shifted = shifted || hx != 0.0 || hy != 0.0 || hz != 0.0;
            }
            if (__all(far)) {
                if (__any(shifted))
                    mixed_far_tile_z<true, XY, XY, NI, UNR>(tr, cnt, xr, yr, zr, sx, sy, sz, ax, ay, az, fx, fy, fz);
                else
                    mixed_far_tile_z<false, XY, XY, NI, UNR>(tr, cnt, xr, yr, zr, sx, sy, sz, ax, ay, az, fx, fy, fz);
#pragma unroll
                for (int q = 0; q < NI; ++q) {
                    fx[q] += (double)ax[q];
                    fy[q] += (double)ay[q];
                    fz[q] += (double)az[q];
                }
                continue;
            }
        }
        if (SUBT) {
            // S1R2 CIC / TSC: per sub-tile of kSub sources, the row's FP32
            // offset from the sub-tile's centre (its z image nearest that
            // centre) and FP32 sums over kSub pairs, added to the FP64
            // accumulators after each sub-tile.  Against whole-tile offsets
            // and sums this cut the error against FP64 at N = 524288 with the
            // comoving term (CIC p99 1.5e-5 -> 8.9e-6, TSC 2.0e-5 -> 9.2e-6;
            // the TSC tiles span half the box in z, so their offsets carried
            // ulp(Lz/4)), for 4-7% of the speed (1.76x / 1.64x FP64).
            for (int s0 = 0; s0 < cnt; s0 += kSub) {
                const int scnt = min(kSub, cnt - s0);
                const double4 cs = a.sub[(t0 + s0) / kSub];
                float xs[NI], ys[NI], zs[NI];
#pragma unroll
                for (int q = 0; q < NI; ++q) {
                    xs[q] = (float)(xi[q] - cs.x);
                    ys[q] = (float)(yi[q] - cs.y);
                    zs[q] = (float)(near_image(zi[q], cs.z, k) - cs.z);
                    ax[q] = ay[q] = az[q] = 0.0f;
                }
                const float4 *__restrict__ ts = tr + s0;
                bool near = false;
                if (scnt == kSub) {
#pragma unroll
                    for (int j = 0; j < kSub; j += UNR) {
                        float4 pj[UNR];
#pragma unroll
                        for (int u = 0; u < UNR; ++u) pj[u] = ts[j + u];
#pragma unroll
                        for (int u = 0; u < UNR; ++u)
#pragma unroll
                            for (int q = 0; q < NI; ++q)
                                mixed_pair_z<V, SWIN>(k, tab32, pj[u], xs[q], ys[q], zs[q], B[q], near, ax[q], ay[q],
                                                      az[q], sw);
                    }
                } else {
                    for (int j = 0; j < scnt; ++j) {
                        const float4 pj = ts[j];
#pragma unroll
                        for (int q = 0; q < NI; ++q)
                            mixed_pair_z<V, SWIN>(k, tab32, pj, xs[q], ys[q], zs[q], B[q], near, ax[q], ay[q], az[q],
                                                  sw);
                    }
                }
                if (__builtin_expect(near, 0)) {
                    for (int jj = 0; jj < scnt; ++jj) {
                        const float4 pr = ts[jj];
#pragma unroll
                        for (int q = 0; q < NI; ++q) {
                            float dx, dy, dz;
                            bool bnd;
                            if (mixed_r2_z<V>(k, pr, xs[q], ys[q], zs[q], dx, dy, dz, bnd) < B[q] || bnd) {
                                const double b = hi[q] + c.w;
                                const int64_t jg = t0 + s0 + jj;
                                pair_f64<V>(k, table, soft, jg, jrec[jg], xi[q], yi[q], zi[q], hi[q], b * b, fx[q],
                                            fy[q], fz[q]);
                            }
                        }
                    }
                }
#pragma unroll
                for (int q = 0; q < NI; ++q) {
                    fx[q] += (double)ax[q];
                    fy[q] += (double)ay[q];
                    fz[q] += (double)az[q];
                }
            }
            continue;
        }
        bool near = false;
        int j = 0;
        for (; j + UNR <= cnt; j += UNR) {
            float4 pj[UNR];
#pragma unroll
            for (int u = 0; u < UNR; ++u) pj[u] = tr[j + u];
#pragma unroll
            for (int u = 0; u < UNR; ++u)
#pragma unroll
                for (int q = 0; q < NI; ++q)
                    mixed_pair_z<V, SWIN>(k, tab32, pj[u], xr[q], yr[q], zr[q], B[q], near, ax[q], ay[q], az[q], sw);
        }
        for (; j < cnt; ++j) {
            const float4 pj = tr[j];
#pragma unroll
            for (int q = 0; q < NI; ++q)
                mixed_pair_z<V, SWIN>(k, tab32, pj, xr[q], yr[q], zr[q], B[q], near, ax[q], ay[q], az[q], sw);
        }
        if (__builtin_expect(near, 0)) {
            for (int64_t jj = t0; jj < tend; ++jj) {
                const float4 pr = jrel[jj];
#pragma unroll
                for (int q = 0; q < NI; ++q) {
                    float dx, dy, dz;
                    bool bnd;
                    if (mixed_r2_z<V>(k, pr, xr[q], yr[q], zr[q], dx, dy, dz, bnd) < B[q] || bnd) {
                        const double b = hi[q] + c.w;
                        pair_f64<V>(k, table, soft, jj, jrec[jj], xi[q], yi[q], zi[q], hi[q], b * b, fx[q],
                                    fy[q], fz[q]);
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < NI; ++q) {
            fx[q] += (double)ax[q];
            fy[q] += (double)ay[q];
            fz[q] += (double)az[q];
        }
    }

#pragma unroll
    for (int q = 0; q < NI; ++q) {
        const int64_t r = row_base + q * kBlock + tid;
        if (r < k.nrows) {
            double *out = part + (size_t)split * 3u * (size_t)k.nrows;
            out[r] = fx[q];
            out[k.nrows + r] = fy[q];
            out[2 * k.nrows + r] = fz[q];
        }
    }
}

// ---------------------------------------------------------------------------
// kernels: one evaluation per launch (grid row-blocks x splits), or several
// evaluations of one variant packed into one launch (flat grid, each
// workgroup finds its evaluation among at most kMaxPack segments)
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// T3 Ewald pair kernel with a table window in LDS.  Rows and sources are in z
// pencils (order_kernels.hip: key = cell column (x, y), then z), so a
// workgroup's 256 rows cover a few adjacent cell columns and one run of
// sources covers one column.  The table rows that any of their stencils can
// reach then form a small (x, y) window of full z runs (5 x 9 rows of the
// folded 63^3 table: 70 KB; serpentine pencils keep the workgroups that
// cross a column change to 6 x 8).  At each change of source column the workgroup
// copies that window into LDS between two barriers (the source loop is the
// same for every wave, so both barriers are reached by all); each pair then
// reads its 16 stencil rows from LDS, which feeds 128 B per clock against the
// 64 B per clock of the vector-L1 path the global reads take.  A lane whose
// stencil leaves the window (a coordinate rounded across a cell edge) reads
// that pair from global memory, and a window that does not fit (workgroups
// whose rows wrap around the box) leaves the whole run on the global path, so
// every read stays in bounds.  Same arithmetic on the same values as the
// global kernel: the sums are bitwise equal.
// ---------------------------------------------------------------------------
constexpr int kT3WinDoubles = 10080;  // 78.75 KB of LDS per workgroup: two workgroups per CU
                                      // (a 6 x 8 window of 63^3 rows)
// A window row holds the (x, y) components of its NZ z nodes as 16-byte pairs
// (2 NZ doubles), then their z components (NZ doubles and two zeros, so that
// six nodes from an even start stay inside); rows start on 16-byte
// boundaries.  A stencil row is then seven aligned 16-byte reads
// (t3_contract_win).  Earlier forms: four aligned reads for (x, y) and two
// 8-byte-aligned pair reads for z, 30.2 G pairs/s.  Measured at 64^3 (row-major
// pencils): node-major rows (24 bytes a node, six unaligned pair reads) 21.3
// G pairs/s, with bank-conflict stalls on 57% of LDS cycles; three component
// runs, 20.9; this layout, 26.7.  Rows padded to 256 bytes (so that a change
// of row keeps the bank pattern) fit fewer windows: 27.9 against 30.2 with
// serpentine pencils, and on the current kernel 30.3 against 34.4 (rows at
// odd multiples of 128 bytes: 32.4).
__host__ __device__ constexpr unsigned t3_win_stride(unsigned nz3) { return (nz3 + 1u) & ~1u; }

// Layout nodes one axis of the window needs: sources in cell column s, rows
// in columns [bmin, bmax].  d / h lies in (s - b - 1, s - b + 1) (mod Ng), so
// the stencils' unfolded nodes lie in [s - bmax + (Ng-1)/2 - 2,
// s - bmin + Ng/2 + 2] (mod Ng); their images in the folded layout are a
// contiguous range [f0, f0 + nf).  False when the range is too wide to be one.
__device__ __forceinline__ bool t3_win_axis(int s, int bmin, int bmax, int Ng, int lo, int &f0, int &nf) {
    const int nlo = (s - bmax) + (Ng - 1) / 2 - 2;
    const int nhi = (s - bmin) + Ng / 2 + 2;
    if (nhi - nlo >= Ng / 2) return false;
    int mn = Ng, mxv = -1;
    for (int n = nlo; n <= nhi; ++n) {
        const int jj = imodp(n, Ng);
        const int m = (jj >= lo) ? jj - lo : Ng - 1 - lo - jj;
        mn = min(mn, m);
        mxv = max(mxv, m);
    }
    f0 = mn;
    nf = mxv - mn + 1;
    return true;
}

// Cell column y of a pencil key (order_kernels.hip reverses y in odd x columns).
__device__ __forceinline__ int pencil_y(uint32_t key, int Ng) {
    const int by = (int)((key >> 10) & 1023u);
    return ((key >> 20) & 1u) ? Ng - 1 - by : by;
}

__device__ __forceinline__ void body_t3_win(const PairArgs &a, const int bx, const int split) {
    __shared__ __align__(16) double win[kT3WinDoubles];
    __shared__ int brange[4];
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const double4 *__restrict__ tiles = a.tiles;
    const double *__restrict__ table = a.table;
    const uint32_t *__restrict__ keys = a.keys;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * kBlock;
    const int64_t j0 = (int64_t)split * k.jchunk;  // multiple of kTile
    const int64_t j1 = min(k.N, j0 + k.jchunk);
    // rows are the sorted particles (id_min = 0, nrows = N)
    const int64_t r = min(row_base + tid, k.nrows - 1);
    const double4 p = irec[r];
    const double xi = p.x, yi = p.y, zi = p.z, hi = isoft[r];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const uint32_t kr = keys[r];
    if (tid == 0) {
        brange[0] = INT32_MAX;
        brange[1] = -1;
        brange[2] = INT32_MAX;
        brange[3] = -1;
    }
    __syncthreads();
    const int rbx = (int)(kr >> 20), rby = pencil_y(kr, k.ng);
    atomicMin(&brange[0], rbx);
    atomicMax(&brange[1], rbx);
    atomicMin(&brange[2], rby);
    atomicMax(&brange[3], rby);
    __syncthreads();
    const int bxmin = brange[0], bxmax = brange[1], bymin = brange[2], bymax = brange[3];
    const int Ng = k.ng, lo = k.t3_lo;
    const unsigned NY = (unsigned)(Ng - lo), NZ = (unsigned)(Ng + 3), NZ3 = 3u * NZ, WS = t3_win_stride(NZ3 + 2u);

    int cur = -1;  // cell column of the sources the window was built for
    bool wok = false;
    int fx0 = 0, nxw = 0, fy0 = 0, nyw = 0;
    const int ntiles = (int)((j1 - j0 + kTile - 1) / kTile);  // 32-bit loop bounds: SALU tests
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t t0 = j0 + (int64_t)ti * kTile;
        const double b = hi + tiles[t0 / kTile].w;
        const double B = b * b;
        const int cnt = (int)min((int64_t)kTile, j1 - t0);
        for (int u = 0; u < cnt; ++u) {
            const int64_t j = t0 + u;
            const uint32_t kj = keys[j];
            const int pk = (int)(kj >> 10);
            if (pk != cur) {
                cur = pk;
                wok = t3_win_axis(pk >> 10, bxmin, bxmax, Ng, lo, fx0, nxw) &&
                      t3_win_axis(pencil_y(kj, Ng), bymin, bymax, Ng, lo, fy0, nyw) &&
                      (int64_t)nxw * nyw * WS <= kT3WinDoubles;
                __syncthreads();  // every wave is done with the previous window
                if (wok) {
                    for (int wa = 0; wa < nxw; ++wa)
                        for (int wc = 0; wc < nyw; ++wc) {
                            const double *src = table + ((unsigned)(fx0 + wa) * NY + (unsigned)(fy0 + wc)) * NZ3;
                            double *dst = win + (unsigned)(wa * nyw + wc) * WS;
                            for (unsigned o = tid; o < NZ3 + 2u; o += kBlock) {
                                if (o < NZ3) {
                                    const unsigned z = o / 3u, c = o - 3u * z;
                                    dst[c < 2u ? 2u * z + c : 2u * NZ + z] = src[o];
                                } else {
                                    dst[o] = 0.0;  // the two pad entries after the z run
                                }
                            }
                        }
                }
                __syncthreads();  // the window is complete
            }
            const double4 pj = jrec[j];
            double dx = pj.x - xi, dy = pj.y - yi, dz = pj.z - zi;
            // forces_cuda.cu:622-627: d - L*d/|d| when |d| > L/2
            if (fabs(dx) > k.halfL) dx -= copysign(k.L, dx);
            if (fabs(dy) > k.halfL) dy -= copysign(k.L, dy);
            if (fabs(dz) > k.halfL) dz -= copysign(k.L, dz);
            const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
            const double w = pair_w_tile(r2, B, hi, soft, j, k);
            T3Stencil st;
            t3_stencil(k, dx, dy, dz, st);
            bool in = wok;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                in = in && (st.mx[q] - (unsigned)fx0 < (unsigned)nxw) && (st.my[q] - (unsigned)fy0 < (unsigned)nyw);
            double Dx, Dy, Dz;
            if (in) {
                const bool odd = (st.iz0 & 1) != 0;
                const double wz5[5] = {odd ? 0.0 : st.wz[0], odd ? st.wz[0] : st.wz[1], odd ? st.wz[1] : st.wz[2],
                                       odd ? st.wz[2] : st.wz[3], odd ? st.wz[3] : 0.0};
                const unsigned za = 2u * NZ + ((unsigned)st.iz0 & ~1u);
                t3_contract_win(
                    st, wz5,
                    [&](int i, int jj, double *vxy, double *vz) {
                        const double *w0 =
                            win + ((st.mx[i] - (unsigned)fx0) * (unsigned)nyw + (st.my[jj] - (unsigned)fy0)) * WS;
                        const double2 *xy = (const double2 *)w0 + st.iz0;  // 16-byte aligned (x, y) nodes
                        const double2 *zz = (const double2 *)(w0 + za);    // 16-byte aligned z pairs
#pragma unroll
                        for (int q = 0; q < 4; ++q) {
                            const double2 a = xy[q];
                            vxy[2 * q] = a.x;
                            vxy[2 * q + 1] = a.y;
                        }
#pragma unroll
                        for (int h = 0; h < 3; ++h) {
                            const double2 b = zz[h];
                            vz[2 * h] = b.x;
                            vz[2 * h + 1] = b.y;
                        }
                    },
                    Dx, Dy, Dz);
            }
            else
                t3_contract(
                    st,
                    [&](int i, int jj, double *v) {
                        t3_load_aos(table + ((st.mx[i] * NY + st.my[jj]) * NZ + (unsigned)st.iz0) * 3u, v);
                    },
                    Dx, Dy, Dz);
            fx = fma(pj.w, fma(w, dx, -Dx), fx);
            fy = fma(pj.w, fma(w, dy, -Dy), fy);
            fz = fma(pj.w, fma(w, dz, -Dz), fz);
        }
    }
    const int64_t rr = row_base + tid;
    if (rr < k.nrows) {
        double *out = a.part + (size_t)split * 3u * (size_t)k.nrows;
        out[rr] = fx;
        out[k.nrows + rr] = fy;
        out[2 * k.nrows + rr] = fz;
    }
}

// ---------------------------------------------------------------------------
// T3 Ewald pair kernel for tables larger than an XCD's L2 (the folded 127^3
// table of IS_PERIODIC = 3 is 12.8 MB against 4 MB of L2 per XCD), with rows
// and sources in z pencils.  The source columns (table cell columns in x)
// are dealt out to the eight XCDs by the table x nodes their pairs read: for
// the row block's first x column c0, XCD k takes the source columns whose
// displacement column (c - c0 + Ng/2) mod Ng lies in [k Ng/8, (k+1) Ng/8), so
// the pairs an XCD evaluates read one slab of the table (~20 x nodes of 127:
// ~4 MB), which its L2 can hold, instead of all of it.  Workgroups are
// dispatched to the XCDs round robin, so workgroup id = 8 * row block + k
// runs on XCD k; its sums are split k of the row block's partial sums (eight
// splits, reduced in fixed order).  Every source column falls in exactly one
// slab, so each pair is evaluated once; the pair arithmetic is pair_f64's.
// ---------------------------------------------------------------------------
constexpr int kXcds = 8;

// first index of the sorted keys at or above v (scalar binary search)
__device__ __forceinline__ int64_t keys_lower_bound(const uint32_t *__restrict__ keys, int64_t n, uint32_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (keys[mid] < v) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__device__ __forceinline__ void body_t3_slab(const PairArgs &a, const int id) {
    const int bx = id / kXcds, slab = id % kXcds;
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const double4 *__restrict__ tiles = a.tiles;
    const double *__restrict__ table = a.table;
    const uint32_t *__restrict__ keys = a.keys;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * kBlock;
    // rows are the sorted particles (id_min = 0, nrows = N)
    const int64_t r = min(row_base + tid, k.nrows - 1);
    const double4 p = irec[r];
    const double xi = p.x, yi = p.y, zi = p.z, hi = isoft[r];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int Ng = k.ng;
    const int c0 = (int)(keys[row_base] >> 20);  // x column of the block's first row
    const int lo_k = slab * Ng / kXcds, hi_k = (slab + 1) * Ng / kXcds;
    const int cs = imodp(c0 - Ng / 2 + lo_k, Ng), ncol = hi_k - lo_k;
    for (int part = 0; part < 2; ++part) {
        // source columns [ca, cb): the slab's columns, in two runs where they wrap
        const int ca = part == 0 ? cs : 0;
        const int cb = part == 0 ? min(Ng, cs + ncol) : cs + ncol - Ng;
        if (cb <= ca) continue;
        const int64_t j0 = keys_lower_bound(keys, k.N, (uint32_t)ca << 20);
        const int64_t j1 = (cb >= Ng) ? k.N : keys_lower_bound(keys, k.N, (uint32_t)cb << 20);
        for (int64_t t0 = (j0 / kTile) * kTile; t0 < j1; t0 += kTile) {
            const double b = hi + tiles[t0 / kTile].w;
            const double B = b * b;
            const int u0 = (int)(max(j0, t0) - t0), u1 = (int)(min(j1, t0 + kTile) - t0);
            const double4 *__restrict__ tj = jrec + t0;
            for (int u = u0; u < u1; ++u)
                pair_f64<PV_T3_EWALD>(k, table, soft, t0 + u, tj[u], xi, yi, zi, hi, B, fx, fy, fz);
        }
    }
    const int64_t rr = row_base + tid;
    if (rr < k.nrows) {
        double *out = a.part + (size_t)slab * 3u * (size_t)k.nrows;
        out[rr] = fx;
        out[k.nrows + rr] = fy;
        out[2 * k.nrows + rr] = fz;
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision T3 Ewald pair kernel.  Each displacement is formed and
// wrapped in FP64 (forces_cuda.cu:620-627) and rounded to FP32 once: the
// sources are in z pencils, whose 64-source tiles span the box in z, so
// tile-recentred FP32 positions would carry errors of ulp(L/2) on close
// pairs.  r^2, r^-3 and the tricubic correction are FP32, the correction read
// from an FP32 copy of the device table with one 16-byte (Dx, Dy, Dz, 0) word
// per node, so a stencil row is four aligned 16-byte reads.  Pairs within
// 1.00001 x the softening bound are zeroed in FP32 and redone in FP64 (FP64
// table); FP32 sums run over one 64-source tile and are added to FP64
// accumulators.  WIN: rows and sources in pencils and the table window of
// body_t3_win staged in LDS (16 bytes a node: at most 3264 nodes, 51 KB,
// three workgroups per CU); a lane whose stencil leaves the window reads the
// global FP32 table, so both forms give the same sums bitwise.  With every
// workgroup on the window it is bound by LDS bank conflicts like the FP64
// kernel (33 against 34 G pairs/s at 64^3); a third of the workgroups on the
// global table takes it to 38-39.
// ---------------------------------------------------------------------------
constexpr int kT3WinNodes32 = 3264;
// Row blocks with bx % 3 < 2 use the window and the others read the global
// FP32 table, so that the LDS and the vector-L1 path both carry taps: at 64^3
// 38-39 G pairs/s, against 33.1 with every block on the window (LDS-bound) and
// 26.0 with none (vector-L1-bound).  Shares of 4/7, 5/7 and 7/9 measured 38.0,
// 37.7 and 37.3; shares over an even count (1/2, 3/4) left the global blocks
// on a few XCDs and ran at 26.  Rows padded to 256 bytes ran at 31.5.
constexpr unsigned kT3WinNum32 = 2u, kT3WinDen32 = 3u;

struct T3Stencil32 {
    float wx[4], wy[4], wz[4], wxs[4], wys[4];
    unsigned mx[4], my[4];
    int iz0;
};

// t3_axis in FP32 (forces_cuda.cu:93-111).
__device__ __forceinline__ int t3_axis_f32(float d, const KScalars &k, float w[4]) {
    const float u = (d + k.halfLf) * k.inv_gsf - 0.5f;
    const float uf = floorf(u);
    const float t = u - uf;
    const float t2 = t * t, t3 = t * t2;
    w[0] = -0.5f * t3 + t2 - 0.5f * t;
    w[1] = 1.5f * t3 - 2.5f * t2 + 1.0f;
    w[2] = -1.5f * t3 + 2.0f * t2 + 0.5f * t;
    w[3] = 0.5f * t3 - 0.5f * t2;
    return wrap_index((int)uf, k.ng);
}

__device__ __forceinline__ void t3_stencil_f32(const KScalars &k, float dx, float dy, float dz, T3Stencil32 &s) {
    const int Ng = k.ng, lo = k.t3_lo;
    const int ix0 = t3_axis_f32(dx, k, s.wx);
    const int iy0 = t3_axis_f32(dy, k, s.wy);
    s.iz0 = t3_axis_f32(dz, k, s.wz);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int jx = wrap_once(ix0 - 1 + q, Ng);
        const int jy = wrap_once(iy0 - 1 + q, Ng);
        s.mx[q] = (unsigned)((jx >= lo) ? jx - lo : Ng - 1 - lo - jx);
        s.my[q] = (unsigned)((jy >= lo) ? jy - lo : Ng - 1 - lo - jy);
        s.wxs[q] = (jx >= lo) ? s.wx[q] : -s.wx[q];
        s.wys[q] = (jy >= lo) ? s.wy[q] : -s.wy[q];
    }
}

// t3_contract in FP32 over 16-byte nodes: load(i, j, n) fills n[q] with z
// node iz0 - 1 + q of table row (mx[i], my[j]).
template <class Load>
__device__ __forceinline__ void t3_contract_f32(const T3Stencil32 &s, Load load, float &Dx, float &Dy, float &Dz) {
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            float4 n[4];
            load(i, j, n);
            float t0 = s.wz[0] * n[0].x, t1 = s.wz[0] * n[0].y, t2 = s.wz[0] * n[0].z;
#pragma unroll
            for (int q = 1; q < 4; ++q) {
                t0 = fmaf(s.wz[q], n[q].x, t0);
                t1 = fmaf(s.wz[q], n[q].y, t1);
                t2 = fmaf(s.wz[q], n[q].z, t2);
            }
            u0 = fmaf(s.wy[j], t0, u0);
            u1 = fmaf(s.wys[j], t1, u1);
            u2 = fmaf(s.wy[j], t2, u2);
        }
        s0 = fmaf(s.wxs[i], u0, s0);
        s1 = fmaf(s.wx[i], u1, s1);
        s2 = fmaf(s.wx[i], u2, s2);
    }
    Dx = s0;
    Dy = s1;
    Dz = s2;
}

// Nearest-image displacement in FP64, rounded to FP32; returns the FP32 r^2.
// The main loop and the FP64 redo both call it, so they take the same
// near-pair decisions.
__device__ __forceinline__ float t3_mixed_disp(const KScalars &k, const double4 &pj, double xi, double yi, double zi,
                                               float &dx, float &dy, float &dz) {
    double ex = pj.x - xi, ey = pj.y - yi, ez = pj.z - zi;
    if (fabs(ex) > k.halfL) ex -= copysign(k.L, ex);
    if (fabs(ey) > k.halfL) ey -= copysign(k.L, ey);
    if (fabs(ez) > k.halfL) ez -= copysign(k.L, ez);
    dx = (float)ex;
    dy = (float)ey;
    dz = (float)ez;
    return fmaf(dx, dx, fmaf(dy, dy, dz * dz));
}

template <bool WIN>
__device__ __forceinline__ void body_t3_mixed(const PairArgs &a, const int bx, const int split) {
    __shared__ __align__(16) float4 win[WIN ? kT3WinNodes32 : 1];
    __shared__ int brange[4];
    const double4 *__restrict__ irec = a.irec;
    const double *__restrict__ isoft = a.isoft;
    const double4 *__restrict__ jrec = a.jrec;
    const double *__restrict__ soft = a.soft;
    const double4 *__restrict__ tiles = a.tiles;
    const double *__restrict__ table = a.table;
    const float4 *__restrict__ T32 = a.t3tab32;
    const uint32_t *__restrict__ keys = a.keys;
    KScalars k = a.k;
    k.c15 = in_vgpr(a.k.c15);
    const int tid = threadIdx.x;
    const int64_t row_base = (int64_t)bx * kBlock;
    const int64_t j0 = (int64_t)split * k.jchunk;  // multiple of kTile
    const int64_t j1 = min(k.N, j0 + k.jchunk);
    const int64_t r = min(row_base + tid, k.nrows - 1);
    const double4 p = irec[k.id_min + r];
    const double xi = p.x, yi = p.y, zi = p.z, hi = isoft[k.id_min + r];
    double fx = 0.0, fy = 0.0, fz = 0.0;
    const int Ng = k.ng, lo = k.t3_lo;
    const unsigned NY = (unsigned)(Ng - lo), NZ = (unsigned)(Ng + 3);
    int bxmin = 0, bxmax = 0, bymin = 0, bymax = 0;
    const bool win_share = (unsigned)bx % kT3WinDen32 < kT3WinNum32;
    if (WIN) {
        // rows are the sorted particles (id_min = 0, nrows = N)
        if (tid == 0) {
            brange[0] = INT32_MAX;
            brange[1] = -1;
            brange[2] = INT32_MAX;
            brange[3] = -1;
        }
        __syncthreads();
        const uint32_t kr = keys[r];
        const int rbx = (int)(kr >> 20), rby = pencil_y(kr, Ng);
        atomicMin(&brange[0], rbx);
        atomicMax(&brange[1], rbx);
        atomicMin(&brange[2], rby);
        atomicMax(&brange[3], rby);
        __syncthreads();
        bxmin = brange[0];
        bxmax = brange[1];
        bymin = brange[2];
        bymax = brange[3];
    }

    int cur = -1;  // cell column of the sources the window was built for
    bool wok = false;
    int fx0 = 0, nxw = 0, fy0 = 0, nyw = 0;
    const int ntiles = (int)((j1 - j0 + kTile - 1) / kTile);  // 32-bit loop bounds: SALU tests
    for (int ti = 0; ti < ntiles; ++ti) {
        const int64_t t0 = j0 + (int64_t)ti * kTile;
        const double hb = hi + tiles[t0 / kTile].w;
        const float bf = (float)(hb * (1.0 + 1e-5));
        const float B = bf * bf;
        const int cnt = (int)min((int64_t)kTile, j1 - t0);
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        bool near = false;
        for (int u = 0; u < cnt; ++u) {
            const int64_t j = t0 + u;
            if (WIN) {
                const uint32_t kj = keys[j];
                const int pk = (int)(kj >> 10);
                if (pk != cur) {  // the same for every wave: both barriers are reached by all
                    cur = pk;
                    wok = win_share && t3_win_axis(pk >> 10, bxmin, bxmax, Ng, lo, fx0, nxw) &&
                          t3_win_axis(pencil_y(kj, Ng), bymin, bymax, Ng, lo, fy0, nyw) &&
                          (int64_t)nxw * nyw * NZ <= kT3WinNodes32;
                    __syncthreads();  // every wave is done with the previous window
                    if (wok) {
                        const unsigned nn = (unsigned)(nxw * nyw) * NZ;
                        for (unsigned o = tid; o < nn; o += kBlock) {
                            const unsigned row = o / NZ, z = o - row * NZ;
                            const unsigned wa = row / (unsigned)nyw, wc = row - wa * (unsigned)nyw;
                            win[o] = T32[(((unsigned)fx0 + wa) * NY + (unsigned)fy0 + wc) * NZ + z];
                        }
                    }
                    __syncthreads();  // the window is complete
                }
            }
            const double4 pj = jrec[j];
            float dx, dy, dz;
            const float r2 = t3_mixed_disp(k, pj, xi, yi, zi, dx, dy, dz);
            const bool nr = r2 < B;  // left to the FP64 redo (r2 may be 0 here)
            near = near || nr;
            const float y = __builtin_amdgcn_rsqf(r2);
            const float y3 = nr ? 0.0f : (y * y) * y;
            const float m = nr ? 0.0f : (float)pj.w;
            T3Stencil32 st;
            t3_stencil_f32(k, dx, dy, dz, st);
            bool in = WIN && wok;
#pragma unroll
            for (int q = 0; q < 4; ++q)
                in = in && (st.mx[q] - (unsigned)fx0 < (unsigned)nxw) && (st.my[q] - (unsigned)fy0 < (unsigned)nyw);
            float Dx, Dy, Dz;
            if (in)
                t3_contract_f32(
                    st,
                    [&](int i, int jj, float4 *n) {
                        const float4 *w0 =
                            win + ((st.mx[i] - (unsigned)fx0) * (unsigned)nyw + (st.my[jj] - (unsigned)fy0)) * NZ +
                            (unsigned)st.iz0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) n[q] = w0[q];
                    },
                    Dx, Dy, Dz);
            else
                t3_contract_f32(
                    st,
                    [&](int i, int jj, float4 *n) {
                        const float4 *g = T32 + (st.mx[i] * NY + st.my[jj]) * NZ + (unsigned)st.iz0;
#pragma unroll
                        for (int q = 0; q < 4; ++q) n[q] = g[q];
                    },
                    Dx, Dy, Dz);
            ax = fmaf(m, fmaf(y3, dx, -Dx), ax);
            ay = fmaf(m, fmaf(y3, dy, -Dy), ay);
            az = fmaf(m, fmaf(y3, dz, -Dz), az);
        }
        if (__builtin_expect(near, 0)) {
            for (int u = 0; u < cnt; ++u) {
                const int64_t jj = t0 + u;
                const double4 pj = jrec[jj];
                float dx, dy, dz;
                if (t3_mixed_disp(k, pj, xi, yi, zi, dx, dy, dz) < B)
                    pair_f64<PV_T3_EWALD>(k, table, soft, jj, pj, xi, yi, zi, hi, hb * hb, fx, fy, fz);
            }
        }
        fx += (double)ax;
        fy += (double)ay;
        fz += (double)az;
    }
    const int64_t rr = row_base + tid;
    if (rr < k.nrows) {
        double *out = a.part + (size_t)split * 3u * (size_t)k.nrows;
        out[rr] = fx;
        out[k.nrows + rr] = fy;
        out[2 * k.nrows + rr] = fz;
    }
}

enum BodyKind : int { BK_F64 = 0, BK_MIXED_R3 = 1, BK_MIXED_Z = 2, BK_T3_WIN = 3, BK_T3_MIXED = 4, BK_T3_WIN_MIXED = 5,
                      BK_T3_SLAB = 6 };

template <int BK, int V, int NI, int UNR>
__device__ __forceinline__ void run_body(const PairArgs &a, int bx, int by) {
    if (BK == BK_F64) body_f64<V, NI, UNR>(a, bx, by);
    else if (BK == BK_T3_WIN) body_t3_win(a, bx, by);
    else if (BK == BK_T3_SLAB) body_t3_slab(a, bx);
    else if (BK == BK_T3_MIXED) body_t3_mixed<false>(a, bx, by);
    else if (BK == BK_T3_WIN_MIXED) body_t3_mixed<true>(a, bx, by);
    else if (BK == BK_MIXED_R3) body_mixed_r3<NI, UNR>(a, bx, by);
    else body_mixed_z<V, NI, UNR>(a, bx, by);
}

// The mixed R3 kernel is held to 80 VGPRs (6 waves per SIMD; no spills),
// 1% faster than the 84 the compiler picks on its own (measured at N = 2^21).
// The mixed S1R2 nearest-image kernel is held to 128 (4 waves): 6% faster
// than at its own 132 VGPRs and 3 waves (N = 524288); the mixed T3 one (one
// row per lane) to 80 (6 waves; 52 bytes of spills, still 5% faster than
// at 4 waves).  The
// mixed T3 window kernel is held to 168 (3 waves, which its 51 KB of LDS
// allows).
template <int BK, int V, int NI, int UNR>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(
    (BK == BK_MIXED_R3 || (BK == BK_F64 && (V == PV_T3_NEAREST || V == PV_S1R2_NEAREST)))
        ? 6
        : ((BK == BK_MIXED_Z && (V == PV_S1R2_NEAREST || V == PV_T3_NEAREST)) ? (V == PV_T3_NEAREST ? 6 : 4)
                                                                                : (BK == BK_T3_WIN_MIXED ? 3 : 1)))))
void pair_kernel(const PairArgs a) {
    run_body<BK, V, NI, UNR>(a, blockIdx.x, blockIdx.y);
}

template <int BK, int V, int NI, int UNR>
__global__ __launch_bounds__(kBlock) void pair_kernel_packed(const PackSeg *__restrict__ segs, int nseg) {
    const int b = blockIdx.x;
    int s = 0;
    for (int q = 1; q < nseg; ++q)
        if (segs[q].blk_begin <= b) s = q;
    const PackSeg &g = segs[s];
    const int local = b - g.blk_begin;
    run_body<BK, V, NI, UNR>(g.a, local % g.nbx, local / g.nbx);
}

__global__ void table_f32_kernel(const double *__restrict__ T, float2 *__restrict__ T32, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) T32[i] = make_float2((float)T[2 * i], (float)T[2 * i + 1]);
}

// T3 device layout (three doubles a node) to 16-byte FP32 nodes (Dx, Dy, Dz, 0).
__global__ void t3_f32_kernel(const double *__restrict__ T, float4 *__restrict__ T32, int64_t nodes) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nodes) T32[i] = make_float4((float)T[3 * i], (float)T[3 * i + 1], (float)T[3 * i + 2], 0.0f);
}

// ---------------------------------------------------------------------------
// fixed-order reduction of the splits + external term, AoS output
// ---------------------------------------------------------------------------
__device__ __forceinline__ double radial_correction(double r, double R, const double *T,
                                                    int size) {
    // forces_cuda.cu:41-70 (linear interpolation; nodes assumed at R/size*i)
    const double step = R / (double)size;
    const int i = (int)floor(r / R * (size - 1));
    double c = T[size - 1];
    if (i < size - 1) {
        const double X1 = step * i, Y1 = T[i], X2 = step * (i + 1), Y2 = T[i + 1];
        const double A = (Y2 - Y1) / (X2 - X1);
        const double B = Y1 - A * X1;
        c = A * r + B;
    }
    return c;
}

// perm (or null): partial-sum row r belongs to input row perm[r] (rows taken
// in spatial order; id_min is 0 then).
__global__ __launch_bounds__(256) void reduce_kernel(const double *__restrict__ part,
                                                     const double4 *__restrict__ jrec,
                                                     const uint32_t *__restrict__ perm,
                                                     const double *__restrict__ radial,
                                                     double *__restrict__ F, int64_t nrows,
                                                     int64_t id_min, int splits, int ext_mode,
                                                     double ext_coeff, double Rsim,
                                                     int radial_size) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nrows) return;
    const int64_t r = perm ? (int64_t)perm[k] : k;  // output row
    double f[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < splits; ++s) {
        const double *p = part + (size_t)s * 3u * (size_t)nrows;
        f[0] += p[k];
        f[1] += p[nrows + k];
        f[2] += p[2 * nrows + k];
    }
    if (ext_mode == 3) {
        // forces_cuda.cu:698-701: mass_in_unit_sphere * x * correction(r_xy), x and y only
        const double4 x = jrec[id_min + r];
        const double corr = radial_correction(sqrt(x.x * x.x + x.y * x.y), Rsim, radial, radial_size);
        f[0] += ext_coeff * x.x * corr;
        f[1] += ext_coeff * x.y * corr;
    } else if (ext_mode != 0) {
        // forces_cuda.cu:545-556 (mode 1, xyz) and :848-857 (mode 2, xy)
        const double4 x = jrec[id_min + r];
        f[0] += ext_coeff * x.x;
        f[1] += ext_coeff * x.y;
        if (ext_mode == 1) f[2] += ext_coeff * x.z;
    }
    F[3 * r + 0] = f[0];
    F[3 * r + 1] = f[1];
    F[3 * r + 2] = f[2];
}

// ---------------------------------------------------------------------------
// layout kernels
// ---------------------------------------------------------------------------
__global__ void pack_kernel(const double *__restrict__ x, const double *__restrict__ M,
                            const double *__restrict__ h, double4 *__restrict__ jrec,
                            double *__restrict__ soft, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    jrec[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], M[i]);
    soft[i] = h[i];
}

__global__ void pack_positions_kernel(const double *__restrict__ x, double4 *__restrict__ jrec,
                                      int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = jrec[i].w;
    jrec[i] = make_double4(x[3 * i], x[3 * i + 1], x[3 * i + 2], m);
}

__global__ void unpack_positions_kernel(const double4 *__restrict__ jrec, double *__restrict__ x,
                                        int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = jrec[i];
    x[3 * i] = p.x;
    x[3 * i + 1] = p.y;
    x[3 * i + 2] = p.z;
}

// One wave64 per source tile: the tile's bounding-box centre and largest
// softening length (double4 (cx, cy, cz, h_max)); with MIXED also the FP32
// offsets of each source from that centre plus its mass.  A non-null `scaled`
// receives each source as (x, y, z, -1.5 m), the records mr3_scaled reads; it
// may be jrec itself (each lane reads its record before it stores, and the
// padding lanes that re-read the first record of a tile store nothing).
template <bool MIXED, bool SUB = false>
__global__ __launch_bounds__(kTile) void tiles_kernel(const double4 *jrec,
                                                     const double *__restrict__ soft,
                                                     double4 *__restrict__ tiles,
                                                     float4 *__restrict__ jrel, double4 *__restrict__ ext,
                                                     double4 *scaled, double4 *__restrict__ sub, int64_t n) {
    const int64_t t = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t i = t * kTile + lane;
    const bool ok = i < n;
    // padding lanes repeat the first source of their sub-tile, or of the tile
    // (t*kTile < n) when the whole sub-tile is padding
    const int64_t g0 = i - (lane % kSub);
    const int64_t src = ok ? i : (g0 < n ? g0 : t * kTile);
    const double4 p = jrec[src];
    double lo[3] = {p.x, p.y, p.z}, hi[3] = {p.x, p.y, p.z};
    double hm = soft[src];
#pragma unroll
    for (int m = kTile / 2; m > 0; m >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = fmin(lo[a], __shfl_xor(lo[a], m, kTile));
            hi[a] = fmax(hi[a], __shfl_xor(hi[a], m, kTile));
        }
        hm = fmax(hm, __shfl_xor(hm, m, kTile));
    }
    const double cx = 0.5 * (lo[0] + hi[0]), cy = 0.5 * (lo[1] + hi[1]), cz = 0.5 * (lo[2] + hi[2]);
    if (lane == 0) tiles[t] = make_double4(cx, cy, cz, hm);
    // half extents, rounded up so that centre +- extent contains every source
    if (ext && lane == 0)
        ext[t] = make_double4((hi[0] - lo[0]) * (0.5 + 1e-15), (hi[1] - lo[1]) * (0.5 + 1e-15),
                              (hi[2] - lo[2]) * (0.5 + 1e-15), 0.0);
    if (MIXED && !SUB && ok)
        jrel[i] = make_float4((float)(p.x - cx), (float)(p.y - cy), (float)(p.z - cz), (float)p.w);
    if (SUB) {
        // the bounding-box centre of each group of kSub consecutive sources,
        // and each source's FP32 offset from its group's centre
        double sl[3] = {p.x, p.y, p.z}, sh[3] = {p.x, p.y, p.z};
#pragma unroll
        for (int m = kSub / 2; m > 0; m >>= 1)
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                sl[a] = fmin(sl[a], __shfl_xor(sl[a], m, kTile));
                sh[a] = fmax(sh[a], __shfl_xor(sh[a], m, kTile));
            }
        const double sx = 0.5 * (sl[0] + sh[0]), sy = 0.5 * (sl[1] + sh[1]), sz = 0.5 * (sl[2] + sh[2]);
        if (ok && lane % kSub == 0) sub[i / kSub] = make_double4(sx, sy, sz, 0.0);
        if (ok) jrel[i] = make_float4((float)(p.x - sx), (float)(p.y - sy), (float)(p.z - sz), (float)p.w);
    }
    if (scaled && ok) scaled[i] = make_double4(p.x, p.y, p.z, -1.5 * p.w);
}

// Sources per inner-loop step (scalar-loaded together), chosen by measurement
// on the N = 2^21 R3 workload together with the rows per lane: FP64 1 row x 4
// sources (80 VGPRs, 6 waves per SIMD), mixed 2 rows x 8 sources.
constexpr int unroll_of(int v, int precision) {
    return precision == STEPS_PREC_MIXED ? ((v == PV_R3 || v == PV_S1R2_NEAREST || v == PV_T3_NEAREST) ? 8 : 2)
                                         : ((v == PV_R3 || v == PV_T3_NEAREST || v == PV_S1R2_NEAREST) ? 4 : 1);
}

template <int BK, int V>
hipError_t launch_pair_t(const PairArgs *a, const PackSeg *segs, int nseg, unsigned gx, unsigned gy,
                         hipStream_t st) {
    constexpr int prec = (BK == BK_F64 || BK == BK_T3_WIN || BK == BK_T3_SLAB) ? STEPS_PREC_FP64 : STEPS_PREC_MIXED;
    constexpr int NI = rows_per_lane(V, prec);
    constexpr int UNR = unroll_of(V, prec);
    if (segs)
        hipLaunchKernelGGL((pair_kernel_packed<BK, V, NI, UNR>), dim3(gx), dim3(kBlock), 0, st, segs, nseg);
    else
        hipLaunchKernelGGL((pair_kernel<BK, V, NI, UNR>), dim3(gx, gy), dim3(kBlock), 0, st, *a);
    return hipGetLastError();
}

// Runtime (variant, precision) -> kernel instantiation; one evaluation (a)
// on a gx x gy grid, or nseg packed evaluations (segs, device) on gx blocks.
hipError_t launch_pair(int variant, int precision, const PairArgs *a, const PackSeg *segs, int nseg, unsigned gx,
                       unsigned gy, hipStream_t st) {
    if (precision == STEPS_PREC_MIXED) {
        switch (variant) {
        case PV_R3: return launch_pair_t<BK_MIXED_R3, PV_R3>(a, segs, nseg, gx, gy, st);
        case PV_T3_NEAREST: return launch_pair_t<BK_MIXED_Z, PV_T3_NEAREST>(a, segs, nseg, gx, gy, st);
        case PV_T3_EWALD: return launch_pair_t<BK_T3_MIXED, PV_T3_EWALD>(a, segs, nseg, gx, gy, st);
        case PV_S1R2_NEAREST: return launch_pair_t<BK_MIXED_Z, PV_S1R2_NEAREST>(a, segs, nseg, gx, gy, st);
        case PV_S1R2_CIC: return launch_pair_t<BK_MIXED_Z, PV_S1R2_CIC>(a, segs, nseg, gx, gy, st);
        case PV_S1R2_TSC: return launch_pair_t<BK_MIXED_Z, PV_S1R2_TSC>(a, segs, nseg, gx, gy, st);
        default: return hipErrorInvalidValue;
        }
    }
    switch (variant) {
    case PV_R3: return launch_pair_t<BK_F64, PV_R3>(a, segs, nseg, gx, gy, st);
    case PV_T3_NEAREST: return launch_pair_t<BK_F64, PV_T3_NEAREST>(a, segs, nseg, gx, gy, st);
    case PV_T3_EWALD: return launch_pair_t<BK_F64, PV_T3_EWALD>(a, segs, nseg, gx, gy, st);
    case PV_S1R2_NEAREST: return launch_pair_t<BK_F64, PV_S1R2_NEAREST>(a, segs, nseg, gx, gy, st);
    case PV_S1R2_NGP: return launch_pair_t<BK_F64, PV_S1R2_NGP>(a, segs, nseg, gx, gy, st);
    case PV_S1R2_CIC: return launch_pair_t<BK_F64, PV_S1R2_CIC>(a, segs, nseg, gx, gy, st);
    case PV_S1R2_TSC: return launch_pair_t<BK_F64, PV_S1R2_TSC>(a, segs, nseg, gx, gy, st);
    case PV_S1R2_IMAGES: return launch_pair_t<BK_F64, PV_S1R2_IMAGES>(a, segs, nseg, gx, gy, st);
    default: return hipErrorInvalidValue;
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

bool is_mixed(const EvalPlan &p) { return p.precision == STEPS_PREC_MIXED && mixed_supported(p.variant); }

// The mixed kernels read spatially ordered sources (Morton order; their FP32
// offsets need compact tiles) except the S1R2 TSC one, whose table window
// needs tiles compact in xy (input order), and the R3 one below N = 65536; the R3 and
// nearest-image FP64 kernels from N = 65536 on (their far-tile test gains 7%
// for R3 at N = 2^21); the T3 Ewald kernel sources in z pencils.  Below
// N = 65536 the ~16 launches of the ordering cost more than the pair kernels
// save: the mixed R3 force at N = 16384 took 0.32-0.34 ms with it against
// 0.24-0.25 ms for FP64.
constexpr int64_t kR3SortMinN = 65536;
bool sorted_sources(const EvalPlan &p) {
    const bool far_tiles = p.variant == PV_R3 || p.variant == PV_T3_NEAREST || p.variant == PV_S1R2_NEAREST;
    const bool mixed_sorted = is_mixed(p) && p.variant != PV_S1R2_TSC && !(p.variant == PV_R3 && p.N < kR3SortMinN);
    return mixed_sorted || (far_tiles && p.N >= kR3SortMinN) || p.variant == PV_T3_EWALD;
}
bool pencil_order(const EvalPlan &p) { return p.variant == PV_T3_EWALD; }
// The mixed S1R2 CIC / TSC kernels take their FP32 source offsets relative to
// the centres of 8-source sub-tiles (tiles_kernel<true, true>).
bool sub_tiles(const EvalPlan &p) {
    return is_mixed(p) && (p.variant == PV_S1R2_CIC || p.variant == PV_S1R2_TSC);
}
size_t sub_bytes(const EvalPlan &p) {
    return align256((size_t)std::max<int64_t>((p.N + kSub - 1) / kSub, 1) * sizeof(double4));
}
// The FP64 R3 and nearest-image kernels read source records whose mass slot
// holds -1.5 m (mr3_scaled): the sorted copy, scaled in place, or a scaled
// copy of the input-order records.
bool scaled_sources(const EvalPlan &p) {
    return !is_mixed(p) && (p.variant == PV_R3 || p.variant == PV_T3_NEAREST || p.variant == PV_S1R2_NEAREST);
}
// The T3 Ewald kernel also takes its rows in that order when it evaluates
// every particle, and so do the nearest-image FP64 kernels (Morton order: a
// wave's rows then sit in one small region, so whole tiles pass the far /
// no-wrap test; in input order a wave of a lattice spans the box); the
// reduction scatters the sums back to input order.
bool sorted_rows(const EvalPlan &p) {
    const bool nearest = p.variant == PV_T3_NEAREST || p.variant == PV_S1R2_NEAREST;
    return (pencil_order(p) || nearest) && sorted_sources(p) && p.id_min == 0 && p.nrows == p.N;
}
// The LDS-window T3 kernel needs rows in pencils and room for a 5 x 5 window
// of full z runs (Ng <= 127); packed launches keep the global kernel.  On
// finer tables a workgroup's 256 rows span many more table columns, its
// windows rarely fit and it ran slower than the global kernel (127^3 table
// on the 64^3 lattice: 5.67 against 5.87 G pairs/s), so it serves Ng <= 100.
bool t3_window(const EvalPlan &p) {
    return p.variant == PV_T3_EWALD && sorted_rows(p) && !is_mixed(p) && p.t3_ng <= 100 &&
           (int64_t)t3_win_stride(3u * (unsigned)(p.t3_ng + 3) + 2u) * 25 <= kT3WinDoubles;
}
// The same for the mixed T3 kernel (16-byte nodes; Ng <= 127).
bool t3_window_mixed(const EvalPlan &p) {
    return p.variant == PV_T3_EWALD && sorted_rows(p) && is_mixed(p) &&
           (int64_t)(p.t3_ng + 3) * 25 <= kT3WinNodes32;
}

// The XCD-slab T3 kernel (body_t3_slab) takes the tables the window does not
// serve, up to Ng = 160 (a slab of the folded table: Ng/8 + 4 x nodes of
// (Ng - Ng/2) y nodes x (Ng + 3) z nodes x 24 bytes, 4.4 MB at Ng = 127), for
// full evaluations of at least 4096 particles (eight non-empty splits).
}  // namespace
bool t3_slab_plan(const EvalPlan &p) {
    return p.variant == PV_T3_EWALD && p.precision == STEPS_PREC_FP64 && p.t3_ng > 100 && p.t3_ng <= 160 &&
           p.id_min == 0 && p.nrows == p.N && p.N >= 4096;
}
namespace {
bool t3_slab(const EvalPlan &p) { return t3_slab_plan(p) && sorted_rows(p) && p.splits == kXcds; }

size_t s1r2_f32_entries(const EvalPlan &p) {
    return (is_mixed(p) && is_s1r2_lookup(p.variant)) ? s1r2_device_nodes(p.s1r2_nrho, p.s1r2_nz) : 0;
}
// Nodes of the FP32 T3 table of the mixed kernel (one float4 each), sized for
// the unfolded layout so that the scratch size does not depend on the table.
size_t t3_f32_nodes(const EvalPlan &p) {
    return (is_mixed(p) && p.variant == PV_T3_EWALD) ? t3_device_entries(p.t3_ng, 0) / 3 : 0;
}
// Bytes of the FP32 table copy of a mixed table kernel (S1R2 or T3).
size_t table32_bytes(const EvalPlan &p) {
    return std::max(s1r2_f32_entries(p) * sizeof(float2), t3_f32_nodes(p) * sizeof(float4));
}

// Pieces of an evaluation's scratch buffer, in the order tile_extra_bytes
// counts them.
struct ScratchView {
    double *part;
    double4 *tiles;
    double4 *tile_ext;
    double4 *sub;
    float4 *jrel;
    float2 *tab32;
    double4 *jrec_s;
    double *soft_s;
    void *order;
};

ScratchView scratch_view(const EvalPlan &p, double *partials) {
    ScratchView v;
    const int64_t nt = (p.N + kTile - 1) / kTile;
    const size_t n = (size_t)std::max<int64_t>(p.N, 1);
    const size_t tb = align256((size_t)std::max<int64_t>(nt, 1) * sizeof(double4));
    v.part = partials;
    v.tiles = (double4 *)((char *)partials + align256(partial_bytes(p)));
    v.tile_ext = (double4 *)((char *)v.tiles + tb);
    v.sub = (double4 *)((char *)v.tile_ext + tb);
    v.jrel = (float4 *)((char *)v.sub + sub_bytes(p));
    v.tab32 = (float2 *)((char *)v.jrel + align256(n * sizeof(float4)));
    v.jrec_s = (double4 *)((char *)v.tab32 + align256(table32_bytes(p)));
    v.soft_s = (double *)((char *)v.jrec_s + align256(n * sizeof(double4)));
    v.order = (char *)v.soft_s + align256(n * sizeof(double));
    return v;
}

KScalars make_scalars(const EvalPlan &p) {
    KScalars k{};
    k.N = p.N;
    k.id_min = p.id_min;
    k.nrows = p.nrows;
    k.jchunk = p.jchunk;
    k.L = p.L;
    k.halfL = 0.5 * p.L;
    k.ng = p.t3_ng;
    k.t3_lo = p.t3_lo;
    k.nrho = p.s1r2_nrho;
    k.nz = p.s1r2_nz;
    k.rho_max = kRadialExtent * p.Rsim;
    k.images = p.images;
    k.image_cut = (((double)p.images) - 0.4) * p.L;
    k.c15 = 1.5;
    k.cm53 = -5.0 / 3.0;
    k.inv_gs = (p.t3_ng > 0 && p.L > 0) ? 1.0 / (p.L / (double)p.t3_ng) : 0.0;
    const double drho = k.rho_max / (double)std::max(1, p.s1r2_nrho - 1);
    k.inv_drho = (drho > 0.0) ? 1.0 / drho : 0.0;
    k.inv_dzg = (p.s1r2_nz > 0 && p.L > 0) ? 1.0 / (p.L / (double)p.s1r2_nz) : 0.0;
    k.last_r = (double)std::max(1, p.s1r2_nrho - 1);
    k.Lf = (float)p.L;
    k.halfLf = (float)(0.5 * p.L);
    k.wrap_epsf = (float)(1e-6 * p.L);
    k.rho_maxf = (float)k.rho_max;
    k.inv_drhof = (float)k.inv_drho;
    k.inv_dzgf = (float)k.inv_dzg;
    k.last_rf = (float)k.last_r;
    k.inv_gsf = (float)k.inv_gs;
    return k;
}

PairArgs make_args(const EvalPlan &p, const DeviceParticles &dp, const ScratchView &v) {
    PairArgs a{};
    const bool sorted = sorted_sources(p);
    a.irec = sorted_rows(p) ? v.jrec_s : dp.jrec;
    a.isoft = sorted_rows(p) ? v.soft_s : dp.soft;
    a.jrec = (sorted || scaled_sources(p)) ? v.jrec_s : dp.jrec;
    a.soft = sorted ? v.soft_s : dp.soft;
    a.jrel = v.jrel;
    a.tiles = v.tiles;
    a.tile_ext = v.tile_ext;
    a.sub = v.sub;
    a.table = (p.variant == PV_T3_EWALD) ? p.t3 : (is_s1r2_lookup(p.variant) ? p.s1r2 : nullptr);
    a.tab32 = v.tab32;
    a.t3tab32 = (const float4 *)v.tab32;  // the same scratch region: one mixed table per evaluation
    a.keys = (sorted && pencil_order(p)) ? spatial_order_keys(v.order, p.N) : nullptr;
    a.part = v.part;
    a.k = make_scalars(p);
    return a;
}

unsigned row_blocks(const EvalPlan &p) {
    const int64_t rpb = (int64_t)kBlock * rows_per_lane(p.variant, p.precision);
    return (unsigned)((p.nrows + rpb - 1) / rpb);
}

bool has_pairs(const EvalPlan &p) { return p.nrows > 0 && p.N > 0; }

}  // namespace

// The reference table is antisymmetric in each axis (ewald_space.cc:288-383
// fills it from one O_h wedge): the node mirrored in x, Ng-1-ix, holds
// (-Tx, Ty, Tz), likewise in y.  Components that vanish by symmetry (Tx on
// the plane x = 0, Ty on y = 0) hold rounding noise instead of zero
// (|T| ~ 1e-22 against max|T| ~ 4e-4 for the 63^3 table); two such entries
// count as equal when both are below 2^-52 max|T|.  The plane a mirror maps
// onto itself (odd Ng) is read directly, never negated.  The fold is used
// only when every node passes, so a folded read differs from the stored
// value by at most 2^-51 max|T|, and only in a component that vanishes by
// symmetry.
static bool t3_antisymmetric_xy(const double *T, int ng) {
    const size_t n = (size_t)ng;
    double tmax = 0.0;
    for (size_t i = 0; i < n * n * n * 3u; ++i) tmax = std::max(tmax, std::fabs(T[i]));
    if (!(tmax < INFINITY)) return false;  // NaN or infinite entries: no fold
    const double eps0 = std::ldexp(tmax, -52);
    auto same = [eps0](double b, double a) {
        return b == a || (std::fabs(a) <= eps0 && std::fabs(b) <= eps0);
    };
    auto at = [&](size_t ix, size_t iy, size_t iz) { return T + ((ix * n + iy) * n + iz) * 3u; };
    for (size_t ix = 0; ix < n; ++ix)
        for (size_t iy = 0; iy < n; ++iy)
            for (size_t iz = 0; iz < n; ++iz) {
                const double *a = at(ix, iy, iz), *bx = at(n - 1 - ix, iy, iz), *by = at(ix, n - 1 - iy, iz);
                if (!((2 * ix + 1 == n || same(bx[0], -a[0])) && same(bx[1], a[1]) && same(bx[2], a[2])))
                    return false;
                if (!(same(by[0], a[0]) && (2 * iy + 1 == n || same(by[1], -a[1])) && same(by[2], a[2])))
                    return false;
            }
    return true;
}

int t3_layout_lo(const double *T, int ng) { return t3_antisymmetric_xy(T, ng) ? ng / 2 : 0; }

int t3_device_layout(const double *T, int ng, std::vector<double> &out) {
    const int lo = t3_layout_lo(T, ng);
    const size_t n = (size_t)ng, nx = (size_t)(ng - lo), nz = n + 3;
    out.resize(t3_device_entries(ng, lo));
    for (size_t mx = 0; mx < nx; ++mx)
        for (size_t my = 0; my < nx; ++my) {
            const double *src = T + (((size_t)lo + mx) * n + (size_t)lo + my) * n * 3u;
            double *dst = out.data() + (mx * nx + my) * nz * 3u;
            for (size_t z = 0; z < nz; ++z) {
                const size_t iz = (z + n - 1) % n;  // layout z = node z + 1, periodic
                for (int c = 0; c < 3; ++c) dst[z * 3 + c] = src[iz * 3 + c];
            }
        }
    return lo;
}

void s1r2_device_layout(const double *T, int nrho, int nz, std::vector<double> &out) {
    const size_t nr = (size_t)nrho + 2, nzl = (size_t)nz + 3;
    out.resize(nr * nzl * 2u);
    for (size_t r = 0; r < nr; ++r) {
        const size_t ir = std::min<size_t>(r == 0 ? 0 : r - 1, (size_t)nrho - 1);  // clamped edges
        for (size_t z = 0; z < nzl; ++z) {
            const size_t iz = (z + (size_t)nz - 1) % (size_t)nz;  // entry z holds z node z - 1, periodic
            const double *src = T + (ir * (size_t)nz + iz) * 2u;  // reference layout [(ir*Nz + iz)*2 + c]
            out[(r * nzl + z) * 2u] = src[0];
            out[(r * nzl + z) * 2u + 1] = src[1];
        }
    }
}

size_t partial_bytes(const EvalPlan &p) {
    return (size_t)p.splits * 3u * (size_t)std::max<int64_t>(p.nrows, 1) * sizeof(double);
}

// Scratch after the partial sums (same buffer, 256-byte aligned pieces, in
// ScratchView order): tiles and their half extents [ceil(N/64)] double4; for
// spatially ordered sources (mixed path, R3) also jrel [N] float4, the FP32
// S1R2 table [nrho*nz] float2, the ordered copies (jrec_s double4 [N],
// soft_s double [N]) and the ordering's own scratch.
size_t tile_extra_bytes(const EvalPlan &p) {
    const int64_t nt = (p.N + kTile - 1) / kTile;
    const size_t n = (size_t)std::max<int64_t>(p.N, 1);
    size_t b = 2 * align256((size_t)std::max<int64_t>(nt, 1) * sizeof(double4)) + sub_bytes(p);
    if (sorted_sources(p) || is_mixed(p) || scaled_sources(p)) {  // FP32 records relative to the tiles, FP32 table
        b += align256(n * sizeof(float4));
        b += align256(table32_bytes(p));
    }
    if (sorted_sources(p) || scaled_sources(p)) b += align256(n * sizeof(double4)) + align256(n * sizeof(double));
    if (sorted_sources(p)) {
        b += align256(order_scratch_bytes(p.N));
    }
    return b + 256;
}

hipError_t launch_forces_prepare(const EvalPlan &p, const DeviceParticles &dp, double *partials,
                                 hipStream_t st) {
    if (p.bh) return hipErrorInvalidValue;  // launch_forces runs the octree whole
    if (!has_pairs(p)) return hipSuccess;
    const ScratchView v = scratch_view(p, partials);
    const unsigned nt = (unsigned)((p.N + kTile - 1) / kTile);
    hipError_t e;
    const int64_t n32 = (int64_t)s1r2_f32_entries(p);
    if (n32 > 0) {
        hipLaunchKernelGGL(table_f32_kernel, dim3((unsigned)((n32 + 255) / 256)), dim3(256), 0, st, p.s1r2,
                           v.tab32, n32);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (t3_f32_nodes(p) > 0) {
        const int64_t nodes = (int64_t)(t3_device_entries(p.t3_ng, p.t3_lo) / 3);  // the layout in use
        hipLaunchKernelGGL(t3_f32_kernel, dim3((unsigned)((nodes + 255) / 256)), dim3(256), 0, st, p.t3,
                           (float4 *)v.tab32, nodes);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    if (sorted_sources(p)) {
        if ((e = launch_spatial_order(dp.jrec, dp.soft, p.N, v.order, v.jrec_s, v.soft_s, st,
                                      pencil_order(p) ? p.L : 0.0, pencil_order(p) ? p.t3_ng : 0)) != hipSuccess)
            return e;
        double4 *scaled = scaled_sources(p) ? v.jrec_s : nullptr;  // in place
        if (sub_tiles(p))
            hipLaunchKernelGGL((tiles_kernel<true, true>), dim3(nt), dim3(kTile), 0, st, v.jrec_s, v.soft_s, v.tiles,
                               v.jrel, v.tile_ext, scaled, v.sub, p.N);
        else if (is_mixed(p))
            hipLaunchKernelGGL((tiles_kernel<true>), dim3(nt), dim3(kTile), 0, st, v.jrec_s, v.soft_s, v.tiles,
                               v.jrel, v.tile_ext, scaled, v.sub, p.N);
        else
            hipLaunchKernelGGL((tiles_kernel<false>), dim3(nt), dim3(kTile), 0, st, v.jrec_s, v.soft_s, v.tiles,
                               v.jrel, v.tile_ext, scaled, v.sub, p.N);
    } else {
        // tile extents too: the R3 kernel's far-tile test reads them (on unsorted
        // tiles it rarely passes, and the per-pair loop runs), and so do the
        // S1R2 table windows; the mixed TSC kernel takes its FP32 records
        // relative to input-order tiles
        double4 *scaled = scaled_sources(p) ? v.jrec_s : nullptr;
        if (sub_tiles(p))
            hipLaunchKernelGGL((tiles_kernel<true, true>), dim3(nt), dim3(kTile), 0, st, dp.jrec, dp.soft, v.tiles,
                               v.jrel, v.tile_ext, scaled, v.sub, p.N);
        else if (is_mixed(p))
            hipLaunchKernelGGL((tiles_kernel<true>), dim3(nt), dim3(kTile), 0, st, dp.jrec, dp.soft, v.tiles, v.jrel,
                               v.tile_ext, scaled, v.sub, p.N);
        else
            hipLaunchKernelGGL((tiles_kernel<false>), dim3(nt), dim3(kTile), 0, st, dp.jrec, dp.soft, v.tiles,
                               v.jrel, v.tile_ext, scaled, v.sub, p.N);
    }
    return hipGetLastError();
}

hipError_t launch_forces_pairs(const EvalPlan &p, const DeviceParticles &dp, double *partials, hipStream_t st) {
    if (!has_pairs(p)) return hipSuccess;
    const PairArgs a = make_args(p, dp, scratch_view(p, partials));
    if (t3_window(p)) return launch_pair_t<BK_T3_WIN, PV_T3_EWALD>(&a, nullptr, 0, row_blocks(p), (unsigned)p.splits, st);
    if (t3_slab(p)) return launch_pair_t<BK_T3_SLAB, PV_T3_EWALD>(&a, nullptr, 0, row_blocks(p) * kXcds, 1, st);
    if (t3_window_mixed(p))
        return launch_pair_t<BK_T3_WIN_MIXED, PV_T3_EWALD>(&a, nullptr, 0, row_blocks(p), (unsigned)p.splits, st);
    return launch_pair(p.variant, is_mixed(p) ? STEPS_PREC_MIXED : STEPS_PREC_FP64, &a, nullptr, 0, row_blocks(p),
                       (unsigned)p.splits, st);
}

hipError_t launch_forces_finish(const EvalPlan &p, const DeviceParticles &dp, double *partials, double *F,
                                hipStream_t st) {
    if (p.nrows <= 0) return hipSuccess;
    const int splits = (p.N > 0) ? p.splits : 0;
    const uint32_t *perm =
        (sorted_rows(p) && has_pairs(p)) ? spatial_order_perm(scratch_view(p, partials).order, p.N) : nullptr;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((p.nrows + 255) / 256)), dim3(256), 0, st, partials,
                       dp.jrec, perm, p.radial, F, p.nrows, p.id_min, splits, p.ext_mode, p.ext_coeff, p.Rsim,
                       p.radial_size);
    return hipGetLastError();
}

hipError_t launch_forces(const EvalPlan &p, const DeviceParticles &dp, double *partials, double *F, hipStream_t st,
                         hipEvent_t ev_begin, hipEvent_t ev_pair_end, hipEvent_t ev_end) {
    hipError_t e;
    if (p.bh) {  // tree build and traversal between the first two events
        if (ev_begin && (e = hipEventRecord(ev_begin, st)) != hipSuccess) return e;
        if ((e = launch_bh_forces(p, dp, partials, F, st, nullptr)) != hipSuccess) return e;
        if (ev_pair_end && (e = hipEventRecord(ev_pair_end, st)) != hipSuccess) return e;
        if (ev_end && (e = hipEventRecord(ev_end, st)) != hipSuccess) return e;
        return hipSuccess;
    }
    if ((e = launch_forces_prepare(p, dp, partials, st)) != hipSuccess) return e;
    if (ev_begin && (e = hipEventRecord(ev_begin, st)) != hipSuccess) return e;
    if ((e = launch_forces_pairs(p, dp, partials, st)) != hipSuccess) return e;
    if (ev_pair_end && (e = hipEventRecord(ev_pair_end, st)) != hipSuccess) return e;
    if ((e = launch_forces_finish(p, dp, partials, F, st)) != hipSuccess) return e;
    if (ev_end && (e = hipEventRecord(ev_end, st)) != hipSuccess) return e;
    return hipSuccess;
}

size_t pack_desc_bytes() { return sizeof(PackSeg); }

int64_t pack_describe(const EvalPlan &p, const DeviceParticles &dp, double *partials, void *host_descs, int idx,
                      int64_t blk_begin) {
    PackSeg g;
    g.a = make_args(p, dp, scratch_view(p, partials));
    g.nbx = std::max(1u, row_blocks(p));
    g.blk_begin = (int)blk_begin;
    ((PackSeg *)host_descs)[idx] = g;
    return has_pairs(p) ? (int64_t)g.nbx * p.splits : 0;  // an empty segment owns no workgroup
}

hipError_t launch_pair_packed(int variant, int precision, const void *dev_descs, int nseg, int64_t total_blocks,
                              hipStream_t st) {
    if (nseg <= 0 || total_blocks <= 0) return hipSuccess;
    if (nseg > kMaxPack || total_blocks > INT32_MAX) return hipErrorInvalidValue;
    return launch_pair(variant, precision, nullptr, (const PackSeg *)dev_descs, nseg, (unsigned)total_blocks, 1,
                       st);
}

hipError_t launch_pack(const double *x, const double *M, const double *soft, int64_t n,
                       DeviceParticles &dp, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, x, M,
                       soft, dp.jrec, dp.soft, n);
    return hipGetLastError();
}

hipError_t launch_pack_positions(const double *x, int64_t n, DeviceParticles &dp,
                                 hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(pack_positions_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st,
                       x, dp.jrec, n);
    return hipGetLastError();
}

hipError_t launch_unpack_positions(const DeviceParticles &dp, double *x, hipStream_t st) {
    if (dp.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(unpack_positions_kernel, dim3((unsigned)((dp.n + 255) / 256)), dim3(256), 0,
                       st, dp.jrec, x, dp.n);
    return hipGetLastError();
}

}  // namespace steps
