// bh_kernels.hip -- Barnes-Hut octree forces on the device: the reference's
// USE_BH engine (forces.cc:97-187, 373-506 R3, 582-772 T3, 881-1217 S1R2),
// which StePS runs on the CPU only (main.cc:249-252 refuses USE_BH with
// CUDA), with its RANDOMIZE_BH frames and radial correction table.
//
// The reference builds a pointer octree by inserting the particles one by
// one into a root cube of side 2.00002 max|x_i| centred on the origin: a cube
// holding two or more particles is split into eight at its centre (a particle
// goes to the upper half of an axis when its coordinate is strictly above the
// centre), so the tree's internal nodes are exactly the cubes that hold two
// or more particles, at every level, and its leaves the particles.  Here the
// same tree is built in parallel:
//   * each particle's 42 octant digits are computed with the reference's own
//     centre arithmetic (c +- side/4, side/2 per level) into two 63-bit keys;
//   * the particles are sorted by the first key (LSD radix), then each run
//     of equal first keys (particles closer than 2^-21 of the root side,
//     rare) by the second: a stable sort by the 126-bit key; adjacent sorted
//     particles share lcp[p]
//     leading digits, so the cubes that start at particle p are those of
//     levels lcp[p-1]+1 .. lcp[p] and, followed by p's leaf, they list the
//     tree in pre-order with children in octant order (the reference's loop
//     order); a cube of level l starting at p ends at the last b >= p whose
//     key shares l digits with p's (an exponential, then binary, search over
//     the sorted keys: 2 log2 of the cube's particle count steps);
//   * a cube whose particles all lie in one octant (a chain: one non-empty
//     child, the same particles a level down) becomes one item with its
//     deepest cube: the cubes starting at p are then only the levels lcp[p],
//     lcp[e1], lcp[e2], ... above lcp[p-1], e_k the end of the k-th (each
//     level the value of lcp that ends the one below).  The chain's cubes
//     share particles, mass, centre of mass and softening, and their sides
//     halve exactly, so side / dist < theta holds at some cube of the chain
//     exactly when it holds at the deepest: the reference takes the chain's
//     particles as one interaction or opens the deepest cube, as the item
//     does.  Every item then has two or more children, and the N leaves have
//     at most N - 1 items above them;
//   * every item gets a skip link: the item after its subtree;
//   * node masses and centres of mass are summed bottom-up, level by level.
// The traversal makes the reference's decision per (row, node): a leaf, or a
// node with side / dist < theta, is taken at its centre of mass with
// softening length cbrt(mass / M_min) * ParticleRadi + h_i; any other node
// is opened; the row's own leaf and massless nodes are skipped.  The lanes of
// a wave (rows in key order, so neighbours in space) walk the pre-order list
// together, stackless by the skip links, each lane with its own decisions
// (bh_force_kernel), so the set of interactions is the reference's and only
// the rounding of the centres of mass and of the sums differs.
//
// Particles closer than 2^-42 of the root side share one level-42 node whose
// children are their leaves (the reference splits further, which changes the
// force only for rows within a few 2^-42 root sides of such a group; at
// 2^-21, 22 of the compactified R^3 particles at N = 2^21 would share).
// The item arrays hold 2N items and cannot overflow.
#include "internal.h"

#include <algorithm>
#include <cmath>

namespace steps {
namespace {

constexpr int kKeyLevels = 21;             // octant digits in one 63-bit key
constexpr int kLevels = 2 * kKeyLevels;  // octree levels below the root resolved by the two keys
constexpr int kRedBlocks = 256;
constexpr int kScanItems = 1024;  // counts per workgroup of the item scan

size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

int64_t scan_blocks(int64_t n) { return std::max<int64_t>(1, (n + kScanItems - 1) / kScanItems); }

// The frame the randomised octree (RANDOMIZE_BH) is built in, with the
// reference's operations: R3 turns the positions by Rodrigues' formula about
// a random axis (forces.cc:324-356, 422-434), T3 shifts them by a random
// vector and wraps them into [0, L) (forces.cc:662-685), S1R2 turns them
// about the z axis and shifts and wraps z (forces.cc:884-899, 1026-1073).
// The root cube moves with it (bh_reduce_final) and each row's force is
// turned back at the end (forces.cc:500-503, 1207-1210).
struct BhFrame {
    int kind = 0;                   // 0 none, 1 R3 rotation, 2 T3 shift, 3 S1R2 rotation and z shift
    double c = 1.0, s = 0.0;        // cos, sin of the angle
    double cb = 1.0, sb = 0.0;      // cos, sin of minus the angle: forces turned back
    double ax = 0, ay = 0, az = 0;  // R3 axis
    double sx = 0, sy = 0, sz = 0;  // T3 shift; S1R2 z shift in sz
    double L = 0;
    double cu0 = 0.5, cu1 = 0.5, cu2 = 0.5;  // draws placing the root centre (R3 x, y, z; S1R2 x, y)
};

// forces.cc:324-356: rotated = X cos + (Y x X) sin + Y (Y . X)(1 - cos)
__device__ __forceinline__ void bh_rodrigues(double &x, double &y, double &z, double ax, double ay, double az,
                                             double c, double s) {
#pragma clang fp contract(off)
    const double cx = ay * z - az * y, cy = az * x - ax * z, cz = ax * y - ay * x;
    const double dot = ax * x + ay * y + az * z;
    const double rx = x * c + cx * s + ax * dot * (1.0 - c);
    const double ry = y * c + cy * s + ay * dot * (1.0 - c);
    const double rz = z * c + cz * s + az * dot * (1.0 - c);
    x = rx;
    y = ry;
    z = rz;
}

// forces.cc:884-899: the 2-D rotation in the x-y plane
__device__ __forceinline__ void bh_rotate_xy(double &x, double &y, double c, double s) {
#pragma clang fp contract(off)
    const double rx = x * c - y * s, ry = x * s + y * c;
    x = rx;
    y = ry;
}

// forces.cc:670-682, 1061-1071: shift, then one periodic wrap
__device__ __forceinline__ double bh_shift_wrap(double v, double d, double L) {
    v += d;
    if (v < 0) v = v + L;
    else if (v >= L) v = v - L;
    return v;
}

__device__ __forceinline__ double4 bh_to_frame(const BhFrame &fr, double4 p) {
    if (fr.kind == 1) {
        bh_rodrigues(p.x, p.y, p.z, fr.ax, fr.ay, fr.az, fr.c, fr.s);
    } else if (fr.kind == 2) {
        p.x = bh_shift_wrap(p.x, fr.sx, fr.L);
        p.y = bh_shift_wrap(p.y, fr.sy, fr.L);
        p.z = bh_shift_wrap(p.z, fr.sz, fr.L);
    } else if (fr.kind == 3) {
        bh_rotate_xy(p.x, p.y, fr.c, fr.s);
        p.z = bh_shift_wrap(p.z, fr.sz, fr.L);
    }
    return p;
}

// max |x_i| (R3) or max sqrt(x_i^2 + y_i^2) (S1R2), and min m_i, in two steps
// (per-block partials, then one block that also places the root cube)
__global__ __launch_bounds__(256) void bh_reduce_partial(const double4 *__restrict__ rec, int64_t n, bool xy,
                                                         double *__restrict__ part) {
    __shared__ double rr[256], mm[256];
    double r = 0.0, m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double4 p = rec[i];
        {
#pragma clang fp contract(off)
            // forces.cc:410-418 and 1016-1023: the radius without fused operations
            r = fmax(r, sqrt(xy ? p.x * p.x + p.y * p.y : p.x * p.x + p.y * p.y + p.z * p.z));
        }
        m = fmin(m, p.w);
    }
    const int t = threadIdx.x;
    rr[t] = r;
    mm[t] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            rr[t] = fmax(rr[t], rr[t + s]);
            mm[t] = fmin(mm[t], mm[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[2 * blockIdx.x] = rr[0];
        part[2 * blockIdx.x + 1] = mm[0];
    }
}

// out: root side, M_min (utils.cc:62-67), root centre (x, y, z).  The root
// cube of each topology: R3 side 2.00002 max|x| at the origin (forces.cc:
// 437-438); T3 the box, side L at L/2 (forces.cc:690); S1R2 side L, or twice
// the largest cylindrical radius when that is not below L, centred on the
// axis at z = side/2 (forces.cc:1047-1058).  Randomised: R3 side 4.00004
// max|x| centred at (u - 0.5) 2 max|x| per axis (forces.cc:424-428); S1R2
// side L unless 4 max r_xy >= L (then 4 max r_xy), centred at
// ((u1, u2) - 0.5) 2 max r_xy and z = side/2 (forces.cc:1030-1045, 1058).
__global__ __launch_bounds__(256) void bh_reduce_final(const double *__restrict__ part, int nb, int topology,
                                                       double L, BhFrame fr, double *__restrict__ out) {
    __shared__ double rr[256], mm[256];
    const int t = threadIdx.x;
    double r = 0.0, m = INFINITY;
    for (int b = t; b < nb; b += 256) {
        r = fmax(r, part[2 * b]);
        m = fmin(m, part[2 * b + 1]);
    }
    rr[t] = r;
    mm[t] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (t < s) {
            rr[t] = fmax(rr[t], rr[t + s]);
            mm[t] = fmin(mm[t], mm[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        double side, cx = 0.0, cy = 0.0, cz = 0.0;
        const double R = rr[0];
        if (topology == STEPS_TOPO_T3) {
            side = L;
            cx = cy = cz = 0.50 * L;
        } else if (topology == STEPS_TOPO_S1R2) {
            if (fr.kind == 3) {
                side = (4 * R < L) ? L : 4.0 * R;
                cx = (fr.cu0 - 0.5) * 2 * R;
                cy = (fr.cu1 - 0.5) * 2 * R;
            } else {
                side = (2 * R < L) ? L : 2.0 * R;
            }
            cz = 0.50 * side;
        } else if (fr.kind == 1) {
            side = 4.00004 * R;
            cx = (fr.cu0 - 0.5) * 2.0 * R;
            cy = (fr.cu1 - 0.5) * 2.0 * R;
            cz = (fr.cu2 - 0.5) * 2.0 * R;
        } else {
            side = 2.00002 * R;
        }
        out[0] = side;
        out[1] = mm[0];
        out[2] = cx;
        out[3] = cy;
        out[4] = cz;
    }
}

// The 42 octant digits of each particle, most significant first: at each
// level the reference's test x > cx (bit 0), y > cy (bit 1), z > cz (bit 2)
// against the cube centre, then the child's centre c +- side/4 and side/2
// (forces.cc:121-128, 147-155), in the same floating-point operations.
// Levels 1-21 go into the first key (its high and low halves k0, k1),
// 22-42 into the second (k2).
__global__ __launch_bounds__(256) void bh_keys_kernel(const double4 *__restrict__ rec, int64_t n, BhFrame fr,
                                                      const double *__restrict__ red, uint32_t *__restrict__ k0,
                                                      uint32_t *__restrict__ k1, uint64_t *__restrict__ k2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = bh_to_frame(fr, rec[i]);
    double cx = red[2], cy = red[3], cz = red[4], side = red[0];
    uint64_t key[2] = {0, 0};
#pragma unroll
    for (int h = 0; h < 2; ++h)
        for (int l = 0; l < kKeyLevels; ++l) {
            const bool bx = p.x > cx, by = p.y > cy, bz = p.z > cz;
            key[h] = (key[h] << 3) | (bx ? 1u : 0u) | (by ? 2u : 0u) | (bz ? 4u : 0u);
            const double off = side / 4;
            cx = cx + (bx ? off : -off);
            cy = cy + (by ? off : -off);
            cz = cz + (bz ? off : -off);
            side = side / 2;
        }
    k0[i] = (uint32_t)(key[0] >> 32);
    k1[i] = (uint32_t)key[0];
    k2[i] = key[1];
}

// The sort by the first key completed to a stable sort by both: a particle
// whose first key is shared moves to the start of its run plus its rank in
// the run by (second key, position), and its record is written there again
// (the gather placed the run's records in first-key order).  A run of s
// particles costs s^2 comparisons: with all N particles inside one 2^-21
// cube (a compact cluster away from the R3 origin) an evaluation took 38 ms
// at N = 2^17 against 2.5 ms for the same N spread out.
__global__ __launch_bounds__(256) void bh_refine_kernel(const double4 *__restrict__ rec, BhFrame fr,
                                                        const double *__restrict__ soft,
                                                        const uint64_t *__restrict__ k2,
                                                        const uint64_t *__restrict__ skey,
                                                        const uint32_t *__restrict__ perm, int64_t n,
                                                        double4 *__restrict__ srec, double *__restrict__ ssoft,
                                                        uint64_t *__restrict__ skey2, uint32_t *__restrict__ sidx) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const uint64_t k = skey[p];
    const bool shared = (p > 0 && skey[p - 1] == k) || (p + 1 < n && skey[p + 1] == k);
    if (!shared) return;
    int64_t s = p, e = p;
    while (s > 0 && skey[s - 1] == k) --s;
    while (e + 1 < n && skey[e + 1] == k) ++e;
    const uint32_t me = perm[p];
    const uint64_t m = k2[me];
    int64_t rank = 0;
    for (int64_t q = s; q <= e; ++q) {
        const uint64_t o = k2[perm[q]];
        rank += (o < m || (o == m && q < p)) ? 1 : 0;
    }
    const int64_t d = s + rank;
    srec[d] = bh_to_frame(fr, rec[me]);
    ssoft[d] = soft[me];
    skey2[d] = m;
    sidx[d] = me;
}

// sorted records, in the octree's frame
__global__ __launch_bounds__(256) void bh_gather_kernel(const double4 *__restrict__ rec, BhFrame fr,
                                                        const double *__restrict__ soft,
                                                        const uint32_t *__restrict__ k0,
                                                        const uint32_t *__restrict__ k1,
                                                        const uint64_t *__restrict__ k2,
                                                        const uint32_t *__restrict__ perm, int64_t n,
                                                        double4 *__restrict__ srec, double *__restrict__ ssoft,
                                                        uint64_t *__restrict__ skey, uint64_t *__restrict__ skey2,
                                                        uint32_t *__restrict__ sidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    srec[i] = bh_to_frame(fr, rec[s]);
    ssoft[i] = soft[s];
    skey[i] = ((uint64_t)k0[s] << 32) | k1[s];
    skey2[i] = k2[s];
    sidx[i] = s;
}

// lcp[p]: the leading octant digits sorted particles p and p+1 share
// (0..42); -1 after the last particle.
__global__ __launch_bounds__(256) void bh_lcp_kernel(const uint64_t *__restrict__ skey,
                                                     const uint64_t *__restrict__ skey2, int64_t n,
                                                     int8_t *__restrict__ lcp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int v = -1;
    if (p + 1 < n) {
        const uint64_t x = skey[p] ^ skey[p + 1];  // keys use bits 0..62
        if (x != 0) {
            v = (__clzll((long long)x) - 1) / 3;
        } else {
            const uint64_t x2 = skey2[p] ^ skey2[p + 1];
            v = kKeyLevels + ((x2 == 0) ? kKeyLevels : (__clzll((long long)x2) - 1) / 3);
        }
    }
    lcp[p] = (int8_t)v;
}

// The leading octant digits sorted particles a and b share (0..42).
__device__ __forceinline__ int bh_shared_digits(const uint64_t *__restrict__ skey,
                                                const uint64_t *__restrict__ skey2, int64_t a, int64_t b) {
    const uint64_t x = skey[a] ^ skey[b];  // keys use bits 0..62
    if (x != 0) return (__clzll((long long)x) - 1) / 3;
    const uint64_t x2 = skey2[a] ^ skey2[b];
    return kKeyLevels + ((x2 == 0) ? kKeyLevels : (__clzll((long long)x2) - 1) / 3);
}

// The last particle of the cube of level l that starts at p, searched from a
// particle b >= p known to lie in it: the shared digits with p do not grow
// along the sorted keys, so an exponential search past b, then a binary
// search back, finds the last particle sharing l digits with p.
__device__ __forceinline__ int64_t bh_cube_end(const uint64_t *__restrict__ skey, const uint64_t *__restrict__ skey2,
                                               int64_t n, int64_t p, int64_t b, int l) {
    int64_t step = 1;
    while (b + step < n && bh_shared_digits(skey, skey2, p, b + step) >= l) {
        b += step;
        step <<= 1;
    }
    for (step >>= 1; step > 0; step >>= 1)
        if (b + step < n && bh_shared_digits(skey, skey2, p, b + step) >= l) b += step;
    return b;
}

// Items that start at sorted particle p: the leaf, and one per level of
// lcp[p], lcp[e1], lcp[e2], ... above lcp[p-1] (e_k the end of the k-th;
// the levels strictly decrease, so at most 43 steps).
__device__ __forceinline__ uint32_t bh_items_at(const int8_t *__restrict__ lcp, const uint64_t *__restrict__ skey,
                                                const uint64_t *__restrict__ skey2, int64_t n, int64_t p) {
    const int lp = (p > 0) ? lcp[p - 1] : -1;
    uint32_t c = 1;
    int64_t b = p;
    for (int v = lcp[p]; v > lp; v = lcp[b]) {
        ++c;
        b = bh_cube_end(skey, skey2, n, p, b, v);
    }
    return c;
}

// Items (internal nodes + the leaf) that start at each sorted particle, and
// their exclusive scan within each workgroup of kScanItems particles.
__global__ __launch_bounds__(256) void bh_count_scan_kernel(const int8_t *__restrict__ lcp,
                                                            const uint64_t *__restrict__ skey,
                                                            const uint64_t *__restrict__ skey2, int64_t n,
                                                            uint32_t *__restrict__ base,
                                                            uint32_t *__restrict__ bsum) {
    __shared__ uint32_t s[256];
    constexpr int kPer = kScanItems / 256;
    const int t = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * kScanItems + (int64_t)t * kPer;
    uint32_t c[kPer], sum = 0;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const int64_t p = p0 + q;
        c[q] = (p < n) ? bh_items_at(lcp, skey, skey2, n, p) : 0u;
        sum += c[q];
    }
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        if (p0 + q < n) base[p0 + q] = run;
        run += c[q];
    }
    if (t == 255) bsum[blockIdx.x] = s[255];
}

// Exclusive scan of the workgroup totals (one workgroup); info[0] = items.
__global__ __launch_bounds__(1024) void bh_scan_top_kernel(uint32_t *__restrict__ bsum, int64_t nb,
                                                           int *__restrict__ info) {
    __shared__ uint32_t s[1024];
    const int t = threadIdx.x;
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b = (int64_t)t * per, e = min(nb, b + per);
    uint32_t sum = 0;
    for (int64_t i = b; i < e; ++i) sum += bsum[i];
    s[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const uint32_t v = (t >= off) ? s[t - off] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    uint32_t run = s[t] - sum;
    for (int64_t i = b; i < e; ++i) {
        const uint32_t v = bsum[i];
        bsum[i] = run;
        run += v;
    }
    if (t == 1023) info[0] = (int)s[1023];
}

__global__ __launch_bounds__(256) void bh_scan_add_kernel(uint32_t *__restrict__ base, int64_t n,
                                                          const uint32_t *__restrict__ bsum) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) base[p] += bsum[p / kScanItems];
}

// One item of the pre-order list, 64 bytes (one scalar load of 16 dwords):
// centre of mass and mass; the cube's side (0 for a leaf) and side^2/theta^2; the
// mass part of the softening length, cbrt(mass / M_min) * ParticleRadi; the
// skip link (the item after the subtree) and the particle (input index) of a
// leaf, -1 for an internal node, kBhMassless for an item of zero mass (which
// every row skips: an integer test on the scalar unit, where a test of the
// mass would take a vector compare per item).
constexpr int kBhMassless = -2;
struct __align__(64) BhNode {
    double x, y, z, m;
    double side, msoft;
    double open2;  // side^2 / theta^2: the R3 walk opens the cube at r^2 <= open2
    int next, pidx;
};
static_assert(sizeof(BhNode) == 64, "one 64-byte record per item");

// Every item of the pre-order list: the internal nodes starting at p
// (shallowest first), then p's leaf.  level[] holds an internal node's level
// (its deepest cube), -1 for a leaf.
__global__ __launch_bounds__(256) void bh_fill_kernel(const int8_t *__restrict__ lcp,
                                                      const uint64_t *__restrict__ skey,
                                                      const uint64_t *__restrict__ skey2,
                                                      const uint32_t *__restrict__ base,
                                                      const double4 *__restrict__ srec,
                                                      const uint32_t *__restrict__ sidx, int64_t n,
                                                      const double *__restrict__ red, double pr,
                                                      double inv_theta2, const int *__restrict__ info,
                                                      BhNode *__restrict__ nodes, int8_t *__restrict__ level) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int total = info[0];
    const int lp = (p > 0) ? lcp[p - 1] : -1;
    const double root = red[0], mmin = red[1];
    const int leaf = ((p + 1 < n) ? (int)base[p + 1] : total) - 1;
    int item = leaf - 1;  // the deepest internal node starting at p, then up
    int64_t b = p;
    for (int v = lcp[p]; v > lp; v = lcp[b], --item) {
        b = bh_cube_end(skey, skey2, n, p, b, v);  // the node's last particle
        double side = root;
        for (int q = 0; q < v; ++q) side = side / 2;  // forces.cc:155, level by level
        BhNode &o = nodes[item];
        o.side = side;
        o.open2 = (side * side) * inv_theta2;
        o.next = (b + 1 < n) ? (int)base[b + 1] : total;
        o.pidx = -1;
        level[item] = (int8_t)v;
    }
    const double4 r = srec[p];
    BhNode o;
    o.x = r.x;
    o.y = r.y;
    o.z = r.z;
    o.m = r.w;
    o.side = 0.0;
    o.open2 = 0.0;
    o.msoft = cbrt(r.w / mmin) * pr;
    o.next = (p + 1 < n) ? (int)base[p + 1] : total;
    o.pidx = (r.w == 0.0) ? kBhMassless : (int)sidx[p];
    nodes[leaf] = o;
    level[leaf] = -1;
}

// The internal nodes of one level: mass and centre of mass from their direct
// children (deeper levels are done), and the mass part of the softening.
__global__ __launch_bounds__(256) void bh_level_kernel(int lev, const int *__restrict__ info,
                                                       const double *__restrict__ red, double pr,
                                                       BhNode *__restrict__ nodes, const int8_t *__restrict__ level) {
    const int total = info[0];
    for (int it = blockIdx.x * blockDim.x + threadIdx.x; it < total; it += gridDim.x * blockDim.x) {
        if (level[it] != lev) continue;
        const int end = nodes[it].next;
        double sx = 0, sy = 0, sz = 0, sm = 0;
        for (int c = it + 1; c < end; c = nodes[c].next) {
            const BhNode &q = nodes[c];
            if (q.m == 0.0) continue;  // a massless child adds nothing (and has no centre of mass)
            sx += q.m * q.x;
            sy += q.m * q.y;
            sz += q.m * q.z;
            sm += q.m;
        }
        BhNode &o = nodes[it];
        o.x = sx / sm;
        o.y = sy / sm;
        o.z = sz / sm;
        o.m = sm;
        o.msoft = cbrt(sm / red[1]) * pr;
        if (sm == 0.0) o.pidx = kBhMassless;
    }
}

// forces.cc:52-87 (force_softening): r^-3 outside beta, the cubic spline inside
__device__ __forceinline__ double bh_soft(double r, double beta) {
    const double r2 = r * r, r3 = r2 * r;
    if (r >= beta) return 1.0 / r3;
    const double b2 = beta * beta, b3 = b2 * beta, b4 = b2 * b2, b5 = b4 * beta, b6 = b3 * b3;
    if (r > beta * 0.5)
        return -32.0 / (3.0 * b6) * r3 + 38.4 / b5 * r2 + -48.0 / b4 * r + 64.0 / (3.0 * b3) + (-1.0 / 15.0) / r3;
    return 32.0 / b6 * r3 + -38.4 / b5 * r2 + 32.0 / (3.0 * b3);
}

// (The periodic walks.  The R3 walk needs no 1/r: it tests side^2 / theta^2 < r^2
// and r^2 >= beta^2 and forms m r^-3 with one correction of the seed.)
// y = 1/r and r from r^2: the v_rsq_f64 seed y0 (relative error |d| <= 5.2e-8
// on gfx950) refined to second order, y0 (1 + e/2 + 3e^2/8) with
// e = 1 - r^2 y0^2, relative error below 1e-21 before rounding, so y and r
// are within 2 ulp of 1/sqrt(r^2) and sqrt(r^2).  The reference's tests
// side / r < theta and r >= beta become side * y < theta and r >= beta; 1/r^3
// on the r >= beta branch is y^3.  r^2 = 0 (coincident particles) gives NaN
// for both: the opening test fails, so such a node is opened as the
// reference's side / 0 = infinity opens it, and a leaf falls to the softened
// branch, which takes r = 0 there.
__device__ __forceinline__ void bh_rsq(double r2, double &y, double &r) {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double e = fma(-r2, y0 * y0, 1.0);
    y = fma(y0, e * fma(e, 0.375, 0.5), y0);
    r = r2 * y;
}

// Pair terms of the traversal, one per topology mode (forces.cc:373-398 R3;
// :584-652 T3 Ewald and nearest image; :903-981 S1R2 image sum, Ewald table
// and nearest z-image).
enum BhMode : int { BH_R3 = 0, BH_T3_NEAR = 1, BH_T3_EWALD = 2, BH_S1R2_NEAR = 3, BH_S1R2_TABLE = 4, BH_S1R2_IMAGES = 5 };

struct BhTables {
    const double *t3;  // T3 table, device layout (t3_device_layout)
    int ng, lo;
    const double *s1r2;  // S1R2 table, device layout (s1r2_device_layout)
    int nrho, nz;
    double rho_max;
    const double *radial;  // cylindrical radial table
    int radial_size;
    double Rsim;
    const double *rbh;  // radial octree correction table, or null
    int rbh_size;
};

__device__ __forceinline__ int bh_imodp(int i, int n) {
    const int r = i % n;
    return (r < 0) ? r + n : r;
}

// ewald_space.cc:398-449: the left cell centre and the offset from it; 1-D
// Lagrange weights of an even stencil (Catmull-Rom for 4).
__device__ __forceinline__ int bh_t3_axis(double r, double L, int ng, double &f) {
    const double gs = L / (double)ng;
    const double u = (r + L * 0.5) / gs - 0.5;
    const double uf = floor(u);
    f = u - uf;
    return bh_imodp((int)uf, ng);
}

template <int ORD> __device__ __forceinline__ void bh_t3_weights(double f, double *w) {
    if (ORD == 4) {  // ewald_space.cc:440-449
        const double t2 = f * f, t3 = f * t2;
        w[0] = -0.5 * t3 + t2 - 0.5 * f;
        w[1] = 1.5 * t3 - 2.5 * t2 + 1.0;
        w[2] = -1.5 * t3 + 2.0 * t2 + 0.5 * f;
        w[3] = 0.5 * t3 - 0.5 * t2;
    } else {  // ewald_space.cc:409-438
        constexpr int start = -(ORD / 2 - 1);
#pragma unroll
        for (int j = 0; j < ORD; ++j) {
            w[j] = 1.0;
#pragma unroll
            for (int k = 0; k < ORD; ++k)
                if (k != j) w[j] *= (f - (double)(start + k)) / (double)(j - k);
        }
    }
}

// ewald_space.cc:452-523: the T3 correction of the CPU build, Lagrange order
// ORD (Catmull-Rom for 4) on the cell-centred grid, summed x, then y, then z
// node by node.  The table is the device layout: x and y nodes lo..Ng-1, a
// node i < lo read at its mirror Ng-1-i with that axis' component negated;
// z nodes at layout index iz + 1.
template <int ORD>
__device__ __forceinline__ void bh_t3_correction(const BhTables &tb, double L, double dx, double dy, double dz,
                                                 double &Dx, double &Dy, double &Dz) {
    const int ng = tb.ng, lo = tb.lo;
    double fx, fy, fz, wx[ORD], wy[ORD], wz[ORD];
    const int ix0 = bh_t3_axis(dx, L, ng, fx), iy0 = bh_t3_axis(dy, L, ng, fy), iz0 = bh_t3_axis(dz, L, ng, fz);
    bh_t3_weights<ORD>(fx, wx);
    bh_t3_weights<ORD>(fy, wy);
    bh_t3_weights<ORD>(fz, wz);
    constexpr int off = -(ORD / 2 - 1);
    const size_t NY = (size_t)(ng - lo), NZ = (size_t)(ng + 3);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int i = 0; i < ORD; ++i) {
        const int jx = bh_imodp(ix0 + off + i, ng);
        const size_t mx = (size_t)(jx >= lo ? jx - lo : ng - 1 - lo - jx);
        const double sx = jx >= lo ? 1.0 : -1.0;
        for (int j = 0; j < ORD; ++j) {
            const int jy = bh_imodp(iy0 + off + j, ng);
            const size_t my = (size_t)(jy >= lo ? jy - lo : ng - 1 - lo - jy);
            const double sy = jy >= lo ? 1.0 : -1.0;
            const double wxy = wx[i] * wy[j];
            const double *row = tb.t3 + (mx * NY + my) * NZ * 3u;
            for (int k = 0; k < ORD; ++k) {
                const int iz = bh_imodp(iz0 + off + k, ng);
                const double w = wxy * wz[k];
                const double *v = row + (size_t)(iz + 1) * 3u;
                s0 += w * (sx * v[0]);
                s1 += w * (sy * v[1]);
                s2 += w * v[2];
            }
        }
    }
    Dx = s0;
    Dy = s1;
    Dz = s2;
}

// ewald_space.cc:803-1050: the S1R2 correction, NGP / CIC / TSC (ORD 0 / 2 /
// 4) in (rho, z), z periodic, rho clamped, then rotated to Cartesian.  The
// table is the device layout: node (ir, iz) at ((ir + 1) (Nz + 3) + iz + 1) * 2.
template <int ORD>
__device__ __forceinline__ void bh_s1r2_correction(const BhTables &tb, double Lz, double dx, double dy, double dz,
                                                   double &Dx, double &Dy, double &Dz) {
    const int Nrho = tb.nrho, Nz = tb.nz;
    const double rho_max = tb.rho_max;
    const double rho = sqrt(dx * dx + dy * dy);
    const double half = 0.5 * Lz;
    const double drho = rho_max / (double)max(1, Nrho - 1);
    const double dzg = Lz / (double)Nz;
    const size_t NZ3 = (size_t)(Nz + 3);
    auto node = [&](int ir, int iz, int c) { return tb.s1r2[(((size_t)(ir + 1)) * NZ3 + (size_t)(iz + 1)) * 2u + c]; };
    double Dr = 0.0, Dzv = 0.0;
    if (ORD == 0) {
        const double rc = fmin(fmax(rho, 0.0), rho_max);
        const double ur = (drho > 0.0) ? (rc / drho) : 0.0;
        const double uz = (dz + half) / dzg - 0.5;
        int ir = (int)floor(ur + 0.5);
        const int iz = bh_imodp((int)floor(uz + 0.5), Nz);
        ir = min(max(ir, 0), Nrho - 1);
        Dr = node(ir, iz, 0);
        Dzv = node(ir, iz, 1);
    } else if (ORD == 2) {
        const double rc = fmax(0.0, fmin(rho, rho_max));
        const double ur = (drho > 0.0) ? rc / drho : 0.0;
        int ir0 = (int)floor(ur);
        double fr = ur - (double)ir0;
        if (ir0 < 0) {
            ir0 = 0;
            fr = 0.0;
        }
        if (ir0 > Nrho - 2) {
            ir0 = max(0, Nrho - 2);
            fr = 1.0;
        }
        const int ir1 = ir0 + 1;
        const double uz = (dz + half) / dzg - 0.5;
        int iz0 = (int)floor(uz);
        const double fz = uz - (double)iz0;
        iz0 = bh_imodp(iz0, Nz);
        const int iz1 = bh_imodp(iz0 + 1, Nz);
        double w00 = 1.0 - fr;
        w00 *= 1.0 - fz;
        double w10 = fr;
        w10 *= 1.0 - fz;
        double w01 = 1.0 - fr;
        w01 *= fz;
        const double w11 = fr * fz;
        Dr = w00 * node(ir0, iz0, 0) + w10 * node(ir1, iz0, 0) + w01 * node(ir0, iz1, 0) + w11 * node(ir1, iz1, 0);
        Dzv = w00 * node(ir0, iz0, 1) + w10 * node(ir1, iz0, 1) + w01 * node(ir0, iz1, 1) + w11 * node(ir1, iz1, 1);
    } else {
        const double rc = fmin(fmax(rho, 0.0), rho_max);
        const double ur = (drho > 0.0) ? (rc / drho) : 0.0;
        const double uz = (dz + half) / dzg - 0.5;
        const int jr = (int)floor(ur + 0.5), jz = (int)floor(uz + 0.5);
        const double sr = ur - (double)jr, sz = uz - (double)jz;
        const double wr[3] = {0.5 * (0.5 - sr) * (0.5 - sr), 0.75 - sr * sr, 0.5 * (0.5 + sr) * (0.5 + sr)};
        const double wz[3] = {0.5 * (0.5 - sz) * (0.5 - sz), 0.75 - sz * sz, 0.5 * (0.5 + sz) * (0.5 + sz)};
        const int irs[3] = {max(jr - 1, 0), min(max(jr, 0), Nrho - 1), min(jr + 1, Nrho - 1)};
        for (int row = 0; row < 3; ++row) {
            const int iz = bh_imodp(jz - 1 + row, Nz);
            const double w0 = wr[0] * wz[row], w1 = wr[1] * wz[row], w2 = wr[2] * wz[row];
            Dr += w0 * node(irs[0], iz, 0) + w1 * node(irs[1], iz, 0) + w2 * node(irs[2], iz, 0);
            Dzv += w0 * node(irs[0], iz, 1) + w1 * node(irs[1], iz, 1) + w2 * node(irs[2], iz, 1);
        }
    }
    const double ex = (rho > 0.0) ? dx / rho : 0.0;
    const double ey = (rho > 0.0) ? dy / rho : 0.0;
    Dx = Dr * ex;
    Dy = Dr * ey;
    Dz = Dzv;
}

// forces.cc:276-289 with order 1: linear interpolation of the radial table,
// nodes taken at R / size * i (the reference's convention).
__device__ __forceinline__ double bh_cyl_correction(double r, double R, const double *T, int size) {
    const double step = R / (double)size;
    const int i = (int)floor(r / R * (size - 1));
    double c = T[size - 1];
    if (i < size - 1) {
        const double x1 = step * i, y1 = T[i], x2 = step * (i + 1), y2 = T[i + 1];
        const double A = (y2 - y1) / (x2 - x1);  // utils.cc:24-37
        const double B = y1 - A * x1;
        c = A * r + B;
    }
    return c;
}

// utils.cc:24-37
__device__ __forceinline__ double bh_lin(double X, double X1, double Y1, double X2, double Y2) {
#pragma clang fp contract(off)
    const double A = (Y2 - Y1) / (X2 - X1);
    const double B = Y1 - A * X1;
    return A * X + B;
}

// forces.cc:182-245 with RADIAL_BH_FORCE_ORDER 1 (forces.cc:24): the table's
// values sit at the bin centres (b + 0.5) R / n and are interpolated
// linearly: beyond R along the last two bins; at or above the centre of bin
// b towards bin b + 1 (the last two bins in the last bin); below the centre
// of bin b >= 2 towards bin b - 1; below the centre of bin 0 or 1 along the
// line from the origin to that centre.  r within rounding of R, whose bin
// index is n, is taken in the last bin (the reference reads entry n there);
// a NaN radius takes no correction (and indexes nothing).
__device__ __forceinline__ double bh_radial_bh(double r, double R, const double *t, int n) {
#pragma clang fp contract(off)
    if (!(r >= 0.0)) return 0.0;
    const double w = R / (double)n;
    if (r > R) return bh_lin(r, ((double)(n - 2) + 0.5) * w, t[n - 2], ((double)(n - 1) + 0.5) * w, t[n - 1]);
    int b1 = (int)floor(r / w);
    if (b1 > n - 1) b1 = n - 1;
    double c1 = ((double)b1 + 0.5) * w;
    if (c1 <= r) {
        if (b1 >= n - 1) {
            b1 -= 1;
            c1 = ((double)b1 + 0.5) * w;
        }
        return bh_lin(r, c1, t[b1], ((double)(b1 + 1) + 0.5) * w, t[b1 + 1]);
    }
    if (b1 > 1) return bh_lin(r, ((double)(b1 - 1) + 0.5) * w, t[b1 - 1], c1, t[b1]);
    return bh_lin(r, 0.0, 0.0, c1, t[b1]);
}

// One walk of the pre-order list for the lanes of a wave, each lane with its
// own row (xi, yi, zi, hi, row index i) and decisions, accumulating into f.
//
// The wave walks the pre-order list together: the item `it` is the same for
// every lane (a scalar load of its 64-byte record), and each lane makes its
// own decision there -- take it (a leaf, or side / dist < theta), skip it
// (the row's own leaf, a massless node), or open it.  A lane that takes or
// skips an item waits until the wave reaches that item's skip link; the wave
// moves to the next item in pre-order when any lane opened the current one,
// else to its skip link (every waiting lane waits for an item at or beyond
// it, since it took an ancestor).  So each lane sees exactly the items of its
// own traversal, in the same order: the interactions and their summation
// order are those of a lane walking alone.
template <int MODE, int ORD>
__device__ __forceinline__ void bh_walk(const BhNode *__restrict__ nodes, int total, bool mine, int i, double xi,
                                        double yi, double zi, double hi, double theta, double L, double cut,
                                        const BhTables &tb, double *f) {
    int wait = mine ? 0 : INT32_MAX;  // the next item this lane takes part in
    int it = 0;
    if constexpr (MODE == BH_R3) {
        // The R3 walk with selects in place of branches: the item's test and
        // interaction are formed for every lane and kept where the lane takes
        // part (the scalar unit, not the vector unit, bounds this walk, and
        // each branch costs it exec-mask work); only the softened branch stays.
        // side / r < theta as side^2 / theta^2 < r^2, r >= beta as r^2 >= beta^2
        // (the same decisions to rounding), and m r^-3 from the v_rsq_f64 seed y
        // with one correction, y^3 (1 + 1.5 e), e = 1 - r^2 y^2: within 2.1e-14
        // (one-sided).
        while (it < total) {
            const BhNode q = nodes[it];
            const bool active = wait <= it && q.pidx != i && q.pidx != kBhMassless;
            const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            const bool take = q.pidx >= 0 || q.open2 < r2;
            const double beta = q.msoft + hi;
            const double y0 = __builtin_amdgcn_rsq(r2);
            const double t = y0 * y0;
            const double e = fma(-r2, t, 1.0);
            double wm = (t * (y0 * q.m)) * fma(e, 1.5, 1.0);
            const bool soft = active && take && !(r2 >= beta * beta);
            if (soft) wm = q.m * bh_soft(sqrt(r2), beta);
            wm = (active && take) ? wm : 0.0;
            f[0] = fma(wm, dx, f[0]);
            f[1] = fma(wm, dy, f[1]);
            f[2] = fma(wm, dz, f[2]);
            const bool open = active && !take;
            if (wait <= it && !open) wait = q.next;  // took or skipped it: wait for its skip link
            it = __ballot(open) ? it + 1 : q.next;
        }
        return;
    }
    if constexpr (MODE == BH_T3_NEAR || MODE == BH_S1R2_NEAR) {
        // The same select walk for the nearest-image modes (2.5-3% faster at
        // the 64^3 lattice and the N = 524288 cylinder); the second-order 1/r
        // stays.  The table modes keep their branches: there the select walk
        // made every lane gather table taps (S1R2 CIC 12.3 against 11.0 ms).
        while (it < total) {
            const BhNode q = nodes[it];
            const bool active = wait <= it && q.pidx != i && q.pidx != kBhMassless;
            double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
            if (MODE == BH_T3_NEAR) {  // forces.cc:594-599
                if (fabs(dx) > 0.5 * L) dx -= copysign(L, dx);
                if (fabs(dy) > 0.5 * L) dy -= copysign(L, dy);
            }
            if (fabs(dz) > 0.5 * L) dz -= copysign(L, dz);  // forces.cc:600-601, 941-942
            const double r2 = dx * dx + dy * dy + dz * dz;
            double y, dist;
            bh_rsq(r2, y, dist);
            const bool take = q.pidx >= 0 || q.side * y < theta;
            const bool use = active && take;
            const double beta = q.msoft + hi;
            double w = (y * y) * y;
            if (use && !(dist >= beta)) w = bh_soft(r2 > 0.0 ? dist : 0.0, beta);
            const double wm = use ? q.m * w : 0.0;  // selected: an unused lane's w may be NaN
            f[0] += wm * dx;
            f[1] += wm * dy;
            f[2] += wm * dz;
            const bool open = active && !take;
            if (wait <= it && !open) wait = q.next;  // took or skipped it: wait for its skip link
            it = __ballot(open) ? it + 1 : q.next;
        }
        return;
    }
    while (it < total) {
        const BhNode q = nodes[it];
        bool open = false;
        if (wait <= it) {
            if (q.pidx == i || q.pidx == kBhMassless) {  // the row's own leaf; a massless item
                wait = q.next;
            } else {
                double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                if (MODE == BH_T3_NEAR || MODE == BH_T3_EWALD) {  // forces.cc:594-599
                    if (fabs(dx) > 0.5 * L) dx -= copysign(L, dx);
                    if (fabs(dy) > 0.5 * L) dy -= copysign(L, dy);
                }
                if (MODE >= BH_T3_NEAR && MODE <= BH_S1R2_TABLE)  // forces.cc:600-601, 941-942
                    if (fabs(dz) > 0.5 * L) dz -= copysign(L, dz);
                const double r2 = dx * dx + dy * dy + dz * dz;
                double y, dist;
                bh_rsq(r2, y, dist);
                if (q.pidx >= 0 || q.side * y < theta) {
                    wait = q.next;
                    if (MODE != BH_S1R2_IMAGES || fabs(dz) <= cut) {  // forces.cc:915
                        const double beta = q.msoft + hi;
                        const double w = dist >= beta ? (y * y) * y : bh_soft(r2 > 0.0 ? dist : 0.0, beta);
                        if constexpr (MODE == BH_T3_EWALD || MODE == BH_S1R2_TABLE) {  // forces.cc:606-610, 947-951
                            double Dx, Dy, Dz;
                            if constexpr (MODE == BH_T3_EWALD)
                                bh_t3_correction<ORD>(tb, L, dx, dy, dz, Dx, Dy, Dz);
                            else
                                bh_s1r2_correction<ORD>(tb, L, dx, dy, dz, Dx, Dy, Dz);
                            f[0] += q.m * (w * dx - Dx);
                            f[1] += q.m * (w * dy - Dy);
                            f[2] += q.m * (w * dz - Dz);
                        } else {
                            const double wm = q.m * w;
                            f[0] += wm * dx;
                            f[1] += wm * dy;
                            f[2] += wm * dz;
                        }
                    }
                } else {
                    open = true;  // its first child follows in pre-order
                }
            }
        }
        it = __ballot(open) ? it + 1 : q.next;
    }
}

// One row per lane, rows in key order: the traversal of the row's topology
// mode, then its external term (forces.cc:474-491 R3; none for T3;
// forces.cc:1127-1150, 1170-1190 S1R2 in the x and y components only, times
// the cylindrical correction unless the Ewald table is in use), less the
// radial octree correction x c(r) / r when its table is given (forces.cc:
// 479-485, 1143-1147, 1180-1184), all in the octree's frame; a randomised
// rotation is then undone (forces.cc:500-503, 1207-1210).  The S1R2
// image sum walks the tree once per z image m = -(P+1) .. P+1 with the row
// moved by m L and no wrap, and adds the walks in that order (forces.cc:
// 1108-1118).
template <int MODE, int ORD>
__global__ __launch_bounds__(256) void bh_force_kernel(const double4 *__restrict__ srec,
                                                       const double *__restrict__ ssoft,
                                                       const uint32_t *__restrict__ sidx, int64_t n,
                                                       const int *__restrict__ info,
                                                       const BhNode *__restrict__ nodes, double theta, double L,
                                                       int images, BhTables tb, BhFrame fr, int ext_mode,
                                                       double ext_coeff, int64_t id_min, int64_t id_max,
                                                       double *__restrict__ F) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sc = min(s, n - 1);
    const int i = (int)sidx[sc];
    const bool mine = s < n && i >= id_min && i <= id_max;
    const double4 me = srec[sc];
    const double hi = ssoft[sc];
    double f[3] = {0.0, 0.0, 0.0};
    const int total = info[0];
    if (MODE == BH_S1R2_IMAGES) {
        const double cut = ((double)images - 0.4) * L;  // main.cc:1270-1271, forces.cc:1100
        for (int m = -images; m < images + 1; ++m) {
            double g[3] = {0.0, 0.0, 0.0};
            bh_walk<MODE, ORD>(nodes, total, mine, i, me.x, me.y, me.z + (double)m * L, hi, theta, L, cut, tb, g);
            f[0] += g[0];
            f[1] += g[1];
            f[2] += g[2];
        }
    } else {
        bh_walk<MODE, ORD>(nodes, total, mine, i, me.x, me.y, me.z, hi, theta, L, 0.0, tb, f);
    }
    if (!mine) return;
    if (ext_mode == 1) {
        f[0] += ext_coeff * me.x;
        f[1] += ext_coeff * me.y;
        f[2] += ext_coeff * me.z;
    } else if (ext_mode == 2 || ext_mode == 3) {
        double k = ext_coeff;
        if (ext_mode == 3) k *= bh_cyl_correction(sqrt(me.x * me.x + me.y * me.y), tb.Rsim, tb.radial, tb.radial_size);
        f[0] += k * me.x;
        f[1] += k * me.y;
    }
    if (tb.rbh) {
#pragma clang fp contract(off)
        if (ext_mode == 1) {
            const double r = sqrt(me.x * me.x + me.y * me.y + me.z * me.z);
            const double c = bh_radial_bh(r, tb.Rsim, tb.rbh, tb.rbh_size);
            f[0] -= me.x * c / r;
            f[1] -= me.y * c / r;
            f[2] -= me.z * c / r;
        } else {
            const double r = sqrt(me.x * me.x + me.y * me.y);
            const double c = bh_radial_bh(r, tb.Rsim, tb.rbh, tb.rbh_size);
            f[0] -= me.x * c / r;
            f[1] -= me.y * c / r;
        }
    }
    if (fr.kind == 1) bh_rodrigues(f[0], f[1], f[2], fr.ax, fr.ay, fr.az, fr.cb, fr.sb);
    else if (fr.kind == 3) bh_rotate_xy(f[0], f[1], fr.cb, fr.sb);
    double *o = F + 3 * (i - id_min);
    o[0] = f[0];
    o[1] = f[1];
    o[2] = f[2];
}

struct BhView {
    double *red;  // kRedBlocks partials (max radius, min m), then root side, M_min, root centre
    uint32_t *k0, *k1;  // high and low halves of the first key
    uint64_t *k2;       // the second key
    uint32_t *base, *bsum;
    double4 *srec;
    double *ssoft;
    uint64_t *skey, *skey2;
    uint32_t *sidx;
    int8_t *lcp;
    int *info;
    BhNode *nodes;
    int8_t *level;
    void *sort;
};

// the pieces of the scratch buffer, 256-byte aligned, in this order
template <class Take> void bh_pieces(int64_t n, Take &&take) {
    const size_t nn = (size_t)std::max<int64_t>(n, 1), cap = (size_t)bh_item_capacity(n);
    take((size_t)(2 * kRedBlocks + 5) * 8);  // red
    take(nn * 4);                             // k0
    take(nn * 4);                             // k1
    take(nn * 8);                             // k2
    take(nn * 4);                             // base
    take((size_t)scan_blocks(n) * 4);         // bsum
    take(nn * 32);                            // srec
    take(nn * 8);                             // ssoft
    take(nn * 8);                             // skey
    take(nn * 8);                             // skey2
    take(nn * 4);                             // sidx
    take(nn);                                 // lcp
    take(16);                                 // info
    take(cap * sizeof(BhNode));               // nodes
    take(cap);                                // level
    take(sort_u64_scratch_bytes(n));          // sort
}

BhView bh_view(void *scratch, int64_t n) {
    char *base = (char *)scratch;
    char *ptr[16];
    int k = 0;
    size_t off = 0;
    bh_pieces(n, [&](size_t b) {
        ptr[k++] = base + off;
        off += a256(b);
    });
    BhView v;
    v.red = (double *)ptr[0];
    v.k0 = (uint32_t *)ptr[1];
    v.k1 = (uint32_t *)ptr[2];
    v.k2 = (uint64_t *)ptr[3];
    v.base = (uint32_t *)ptr[4];
    v.bsum = (uint32_t *)ptr[5];
    v.srec = (double4 *)ptr[6];
    v.ssoft = (double *)ptr[7];
    v.skey = (uint64_t *)ptr[8];
    v.skey2 = (uint64_t *)ptr[9];
    v.sidx = (uint32_t *)ptr[10];
    v.lcp = (int8_t *)ptr[11];
    v.info = (int *)ptr[12];
    v.nodes = (BhNode *)ptr[13];
    v.level = (int8_t *)ptr[14];
    v.sort = ptr[15];
    return v;
}

// The randomised frame from the draws u0..u5, with the reference's host
// arithmetic (forces.cc:358-369, 422-429, 664-667, 1027-1033): R3 angle
// u3 pi about the unit vector of (theta = u4 2 pi, z = u5 2 - 1); T3 shift
// (u - 0.5) L; S1R2 angle u0 2 pi and z shift (u3 - 0.5) L.  The forces are
// turned back with cos and sin of minus the angle (forces.cc:502-503, 1210).
BhFrame bh_frame(const EvalPlan &p, int topology) {
    BhFrame fr;
    if (!p.bh_randomize) return fr;
    const double pi = 3.14159265358979323846264338327950288419716939937510;
    const double *u = p.bh_u;
    fr.L = p.L;
    if (topology == STEPS_TOPO_R3) {
        const double angle = u[3] * pi;
        const double theta = u[4] * 2.0 * M_PI;
        const double z = u[5] * 2.0 - 1.0;
        const double r = std::sqrt(1.0 - z * z);
        fr.kind = 1;
        fr.ax = r * std::cos(theta);
        fr.ay = r * std::sin(theta);
        fr.az = z;
        fr.c = std::cos(angle);
        fr.s = std::sin(angle);
        fr.cb = std::cos(-angle);
        fr.sb = std::sin(-angle);
        fr.cu0 = u[0];
        fr.cu1 = u[1];
        fr.cu2 = u[2];
    } else if (topology == STEPS_TOPO_T3) {
        fr.kind = 2;
        fr.sx = (u[0] - 0.5) * p.L;
        fr.sy = (u[1] - 0.5) * p.L;
        fr.sz = (u[2] - 0.5) * p.L;
    } else {
        const double angle = u[0] * 2.0 * pi;
        fr.kind = 3;
        fr.c = std::cos(angle);
        fr.s = std::sin(angle);
        fr.cb = std::cos(-angle);
        fr.sb = std::sin(-angle);
        fr.cu0 = u[1];
        fr.cu1 = u[2];
        fr.sz = (u[3] - 0.5) * p.L;
    }
    return fr;
}

}  // namespace

int64_t bh_item_capacity(int64_t n) {
    const int64_t nn = std::max<int64_t>(n, 1);
    return 2 * nn + 64;  // N leaves and at most N - 1 internal nodes
}

size_t bh_scratch_bytes(int64_t n) {
    size_t b = 0;
    bh_pieces(n, [&](size_t s) { b += a256(s); });
    return b + 256;
}

hipError_t launch_bh_forces(const EvalPlan &p, const DeviceParticles &dp, void *scratch, double *F,
                            hipStream_t st, hipEvent_t ev_tree_end) {
    const int64_t n = p.N;
    if (n <= 0 || p.nrows <= 0) return hipSuccess;
    if (bh_item_capacity(n) > INT32_MAX) return hipErrorInvalidValue;
    const BhView v = bh_view(scratch, n);
    const unsigned ng = (unsigned)((n + 255) / 256);
    const int nb = (int)std::min<int64_t>(kRedBlocks, (n + 255) / 256);
    hipError_t e;
    double *red = v.red + 2 * kRedBlocks;  // root side, M_min, root centre
    const int topology = (p.variant == PV_R3)                                     ? STEPS_TOPO_R3
                         : (p.variant == PV_T3_NEAREST || p.variant == PV_T3_EWALD) ? STEPS_TOPO_T3
                                                                                    : STEPS_TOPO_S1R2;
    const BhFrame fr = bh_frame(p, topology);
    hipLaunchKernelGGL(bh_reduce_partial, dim3(nb), dim3(256), 0, st, dp.jrec, n, topology == STEPS_TOPO_S1R2, v.red);
    hipLaunchKernelGGL(bh_reduce_final, dim3(1), dim3(256), 0, st, v.red, nb, topology, p.L, fr, red);
    hipLaunchKernelGGL(bh_keys_kernel, dim3(ng), dim3(256), 0, st, dp.jrec, n, fr, red, v.k0, v.k1, v.k2);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t *perm1 = nullptr;
    if ((e = launch_sort_u64(v.k0, v.k1, n, v.sort, &perm1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(bh_gather_kernel, dim3(ng), dim3(256), 0, st, dp.jrec, fr, dp.soft, v.k0, v.k1, v.k2, perm1, n,
                       v.srec, v.ssoft, v.skey, v.skey2, v.sidx);
    hipLaunchKernelGGL(bh_refine_kernel, dim3(ng), dim3(256), 0, st, dp.jrec, fr, dp.soft, v.k2, v.skey, perm1, n,
                       v.srec, v.ssoft, v.skey2, v.sidx);
    hipLaunchKernelGGL(bh_lcp_kernel, dim3(ng), dim3(256), 0, st, v.skey, v.skey2, n, v.lcp);
    const int64_t sb = scan_blocks(n);
    hipLaunchKernelGGL(bh_count_scan_kernel, dim3((unsigned)sb), dim3(256), 0, st, v.lcp, v.skey, v.skey2, n, v.base,
                       v.bsum);
    hipLaunchKernelGGL(bh_scan_top_kernel, dim3(1), dim3(1024), 0, st, v.bsum, sb, v.info);
    hipLaunchKernelGGL(bh_scan_add_kernel, dim3(ng), dim3(256), 0, st, v.base, n, v.bsum);
    hipLaunchKernelGGL(bh_fill_kernel, dim3(ng), dim3(256), 0, st, v.lcp, v.skey, v.skey2, v.base, v.srec, v.sidx, n,
                       red, p.bh_particle_radius, 1.0 / (p.bh_theta * p.bh_theta), v.info, v.nodes, v.level);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    // the items of one level are few next to the grid: a grid-stride loop over
    // the items the tree has (info[0], known only on the device)
    const unsigned gi = (unsigned)std::min<int64_t>(4096, (bh_item_capacity(n) + 255) / 256);
    for (int l = kLevels; l >= 0; --l)
        hipLaunchKernelGGL(bh_level_kernel, dim3(gi), dim3(256), 0, st, l, v.info, red, p.bh_particle_radius,
                           v.nodes, v.level);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_tree_end && (e = hipEventRecord(ev_tree_end, st)) != hipSuccess) return e;
    BhTables tb{p.t3, p.t3_ng, p.t3_lo, p.s1r2, p.s1r2_nrho, p.s1r2_nz, kRadialExtent * p.Rsim, p.radial,
                p.radial_size, p.Rsim, p.bh_radial_size > 0 ? p.bh_radial : nullptr, p.bh_radial_size};
    const int64_t id_max = p.id_min + p.nrows - 1;
#define STEPS_BH_LAUNCH(MODE, ORD)                                                                                  \
    hipLaunchKernelGGL((bh_force_kernel<MODE, ORD>), dim3(ng), dim3(256), 0, st, v.srec, v.ssoft, v.sidx, n, v.info, \
                       v.nodes, p.bh_theta, p.L, p.images, tb, fr, p.ext_mode, p.ext_coeff, p.id_min, id_max, F)
    switch (p.variant) {
    case PV_R3: STEPS_BH_LAUNCH(BH_R3, 0); break;
    case PV_T3_NEAREST: STEPS_BH_LAUNCH(BH_T3_NEAR, 0); break;
    case PV_T3_EWALD:
        switch (p.bh_order) {
        case 4: STEPS_BH_LAUNCH(BH_T3_EWALD, 4); break;
        case 6: STEPS_BH_LAUNCH(BH_T3_EWALD, 6); break;
        case 8: STEPS_BH_LAUNCH(BH_T3_EWALD, 8); break;
        default: STEPS_BH_LAUNCH(BH_T3_EWALD, 2); break;
        }
        break;
    case PV_S1R2_NEAREST: STEPS_BH_LAUNCH(BH_S1R2_NEAR, 0); break;
    case PV_S1R2_NGP: STEPS_BH_LAUNCH(BH_S1R2_TABLE, 0); break;
    case PV_S1R2_CIC: STEPS_BH_LAUNCH(BH_S1R2_TABLE, 2); break;
    case PV_S1R2_TSC: STEPS_BH_LAUNCH(BH_S1R2_TABLE, 4); break;
    case PV_S1R2_IMAGES: STEPS_BH_LAUNCH(BH_S1R2_IMAGES, 0); break;
    default: return hipErrorInvalidValue;
    }
#undef STEPS_BH_LAUNCH
    return hipGetLastError();
}

}  // namespace steps
