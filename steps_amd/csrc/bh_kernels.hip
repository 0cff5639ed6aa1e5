// bh_kernels.hip -- Barnes-Hut octree forces on the device: the reference's
// USE_BH engine for the free (R3) boundary (forces.cc:97-187, 378-506), which
// StePS runs on the CPU only (main.cc:249-252 refuses USE_BH with CUDA).
//
// The reference builds a pointer octree by inserting the particles one by
// one into a root cube of side 2.00002 max|x_i| centred on the origin: a cube
// holding two or more particles is split into eight at its centre (a particle
// goes to the upper half of an axis when its coordinate is strictly above the
// centre), so the tree's internal nodes are exactly the cubes that hold two
// or more particles, at every level, and its leaves the particles.  Here the
// same tree is built in parallel:
//   * each particle's 21 octant digits are computed with the reference's own
//     centre arithmetic (c +- side/4, side/2 per level) into a 63-bit key;
//   * the keys are sorted; adjacent sorted particles share lcp[p] leading
//     digits, so the internal nodes that start at particle p are the cubes of
//     levels lcp[p-1]+1 .. lcp[p] and, followed by p's leaf, they list the
//     tree in pre-order with children in octant order (the reference's loop
//     order); a node of level l starting at p ends at the first b >= p with
//     lcp[b] < l (binary lifting over a range-minimum table of lcp);
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
// Particles closer than 2^-21 of the root side share one level-21 node whose
// children are their leaves (the reference splits further, which changes the
// force only for rows within a few 2^-21 root sides of such a group).
// At each of the 22 levels the cubes holding two or more particles are
// disjoint, so there are at most 22 * floor(N/2) internal nodes: the item
// arrays are sized for N + 11 N and cannot overflow.
#include "internal.h"

#include <algorithm>

namespace steps {
namespace {

constexpr int kLevels = 21;  // octree levels below the root resolved by a 63-bit key
constexpr int kRedBlocks = 256;
constexpr int kScanItems = 1024;  // counts per workgroup of the item scan

size_t a256(size_t b) { return (b + 255) & ~(size_t)255; }

// range-minimum table rows: spans 2^0 .. 2^(rows-1) >= n
int table_rows(int64_t n) {
    int r = 1;
    while (((int64_t)1 << (r - 1)) < n) ++r;
    return r;
}

int64_t scan_blocks(int64_t n) { return std::max<int64_t>(1, (n + kScanItems - 1) / kScanItems); }

// max |x_i| and min m_i, in two steps (per-block partials, then one block)
__global__ __launch_bounds__(256) void bh_reduce_partial(const double4 *__restrict__ rec, int64_t n,
                                                         double *__restrict__ part) {
    __shared__ double rr[256], mm[256];
    double r = 0.0, m = INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double4 p = rec[i];
        {
#pragma clang fp contract(off)
            // forces.cc:410-418: sqrt(x^2 + y^2 + z^2), no fused operations
            r = fmax(r, sqrt(p.x * p.x + p.y * p.y + p.z * p.z));
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

__global__ __launch_bounds__(256) void bh_reduce_final(const double *__restrict__ part, int nb,
                                                       double *__restrict__ out) {
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
        out[0] = 2.00002 * rr[0];  // root side (forces.cc:437-438)
        out[1] = mm[0];            // M_min (utils.cc:62-67)
    }
}

// The 21 octant digits of each particle, most significant first: at each
// level the reference's test x > cx (bit 0), y > cy (bit 1), z > cz (bit 2)
// against the cube centre, then the child's centre c +- side/4 and side/2
// (forces.cc:121-128, 147-155), in the same floating-point operations.
__global__ __launch_bounds__(256) void bh_keys_kernel(const double4 *__restrict__ rec, int64_t n,
                                                      const double *__restrict__ red, uint32_t *__restrict__ khi,
                                                      uint32_t *__restrict__ klo) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double4 p = rec[i];
    double cx = 0.0, cy = 0.0, cz = 0.0, side = red[0];
    uint64_t key = 0;
    for (int l = 0; l < kLevels; ++l) {
        const bool bx = p.x > cx, by = p.y > cy, bz = p.z > cz;
        key = (key << 3) | (bx ? 1u : 0u) | (by ? 2u : 0u) | (bz ? 4u : 0u);
        const double off = side / 4;
        cx = cx + (bx ? off : -off);
        cy = cy + (by ? off : -off);
        cz = cz + (bz ? off : -off);
        side = side / 2;
    }
    khi[i] = (uint32_t)(key >> 32);
    klo[i] = (uint32_t)key;
}

__global__ __launch_bounds__(256) void bh_gather_kernel(const double4 *__restrict__ rec,
                                                        const double *__restrict__ soft,
                                                        const uint32_t *__restrict__ khi,
                                                        const uint32_t *__restrict__ klo,
                                                        const uint32_t *__restrict__ perm, int64_t n,
                                                        double4 *__restrict__ srec, double *__restrict__ ssoft,
                                                        uint64_t *__restrict__ skey, uint32_t *__restrict__ sidx) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t s = perm[i];
    srec[i] = rec[s];
    ssoft[i] = soft[s];
    skey[i] = ((uint64_t)khi[s] << 32) | klo[s];
    sidx[i] = s;
}

// lcp[p]: the leading octant digits sorted particles p and p+1 share
// (0..21); -1 after the last particle.  Row 0 of the range-minimum table.
__global__ __launch_bounds__(256) void bh_lcp_kernel(const uint64_t *__restrict__ skey, int64_t n,
                                                     int8_t *__restrict__ lcp) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    int v = -1;
    if (p + 1 < n) {
        const uint64_t x = skey[p] ^ skey[p + 1];
        v = (x == 0) ? kLevels : (__clzll((long long)x) - 1) / 3;  // keys use bits 0..62
    }
    lcp[p] = (int8_t)v;
}

// row k of the range-minimum table: min(lcp[j .. j + 2^k - 1]), the range
// clamped at the last particle (whose lcp is -1)
__global__ __launch_bounds__(256) void bh_table_kernel(int8_t *__restrict__ tab, int64_t n, int k) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int8_t *prev = tab + (size_t)(k - 1) * n;
    const int64_t j2 = min(j + ((int64_t)1 << (k - 1)), n - 1);
    tab[(size_t)k * n + j] = min(prev[j], prev[j2]);
}

// Items (internal nodes + the leaf) that start at each sorted particle, and
// their exclusive scan within each workgroup of kScanItems particles.
__global__ __launch_bounds__(256) void bh_count_scan_kernel(const int8_t *__restrict__ lcp, int64_t n,
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
        c[q] = 0;
        if (p < n) {
            const int lp = (p > 0) ? lcp[p - 1] : -1, lc = lcp[p];
            c[q] = (uint32_t)(max(0, lc - lp) + 1);
        }
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
// centre of mass and mass; the cube's side (0 for a leaf); the mass part of
// the softening length, cbrt(mass / M_min) * ParticleRadi; the skip link
// (the item after the subtree) and the particle (input index) of a leaf, or
// -1 for an internal node.
struct __align__(64) BhNode {
    double x, y, z, m;
    double side, msoft;
    int next, pidx;
    int pad[2];
};
static_assert(sizeof(BhNode) == 64, "one 64-byte record per item");

// Every item of the pre-order list: the internal nodes starting at p, then
// p's leaf.  level[] holds an internal node's level, -1 for a leaf.
__global__ __launch_bounds__(256) void bh_fill_kernel(const int8_t *__restrict__ tab, int rows,
                                                      const uint32_t *__restrict__ base,
                                                      const double4 *__restrict__ srec,
                                                      const uint32_t *__restrict__ sidx, int64_t n,
                                                      const double *__restrict__ red, double pr,
                                                      const int *__restrict__ info, BhNode *__restrict__ nodes,
                                                      int8_t *__restrict__ level) {
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int total = info[0];
    const int8_t *lcp = tab;
    const int lp = (p > 0) ? lcp[p - 1] : -1, lc = lcp[p];
    const double root = red[0], mmin = red[1];
    int item = (int)base[p];
    for (int l = lp + 1; l <= lc; ++l, ++item) {
        // the node's last particle: the first j >= p with lcp[j] < l
        int64_t j = p;
        for (int k = rows - 1; k >= 0; --k)
            if (tab[(size_t)k * n + j] >= l) j = min(j + ((int64_t)1 << k), n - 1);
        double side = root;
        for (int q = 0; q < l; ++q) side = side / 2;  // forces.cc:155, level by level
        BhNode &o = nodes[item];
        o.side = side;
        o.next = (j + 1 < n) ? (int)base[j + 1] : total;
        o.pidx = -1;
        level[item] = (int8_t)l;
    }
    const double4 r = srec[p];
    BhNode o;
    o.x = r.x;
    o.y = r.y;
    o.z = r.z;
    o.m = r.w;
    o.side = 0.0;
    o.msoft = cbrt(r.w / mmin) * pr;
    o.next = (p + 1 < n) ? (int)base[p + 1] : total;
    o.pidx = (int)sidx[p];
    o.pad[0] = o.pad[1] = 0;
    nodes[item] = o;
    level[item] = -1;
}

// The internal nodes of one level: mass and centre of mass from their direct
// children (deeper levels are done), and the mass part of the softening.
__global__ __launch_bounds__(256) void bh_level_kernel(int lev, const int *__restrict__ info,
                                                       const double *__restrict__ red, double pr,
                                                       BhNode *__restrict__ nodes, const int8_t *__restrict__ level) {
    const int64_t it = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= info[0] || level[it] != lev) return;
    const int end = nodes[it].next;
    double sx = 0, sy = 0, sz = 0, sm = 0;
    for (int c = (int)it + 1; c < end; c = nodes[c].next) {
        const BhNode &q = nodes[c];
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

// One row per lane, rows in key order: the reference's traversal
// (forces.cc:378-399), then the external term (forces.cc:474-491).
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
__global__ __launch_bounds__(256) void bh_force_kernel(const double4 *__restrict__ srec,
                                                       const double *__restrict__ ssoft,
                                                       const uint32_t *__restrict__ sidx, int64_t n,
                                                       const int *__restrict__ info,
                                                       const BhNode *__restrict__ nodes, double theta, int ext_mode,
                                                       double ext_coeff, int64_t id_min, int64_t id_max,
                                                       double *__restrict__ F) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t sc = min(s, n - 1);
    const int i = (int)sidx[sc];
    const bool mine = s < n && i >= id_min && i <= id_max;
    const double4 me = srec[sc];
    const double hi = ssoft[sc];
    double fx = 0, fy = 0, fz = 0;
    const int total = info[0];
    int wait = mine ? 0 : INT32_MAX;  // the next item this lane takes part in
    int it = 0;
    while (it < total) {
        const BhNode q = nodes[it];
        bool open = false;
        if (wait <= it) {
            if (q.pidx == i || q.m == 0.0) {  // the row's own leaf; a massless node
                wait = q.next;
            } else {
                const double dx = q.x - me.x, dy = q.y - me.y, dz = q.z - me.z;
                const double dist = sqrt(dx * dx + dy * dy + dz * dz);
                if (q.pidx >= 0 || q.side / dist < theta) {
                    const double w = q.m * bh_soft(dist, q.msoft + hi);
                    fx += w * dx;
                    fy += w * dy;
                    fz += w * dz;
                    wait = q.next;
                } else {
                    open = true;  // its first child follows in pre-order
                }
            }
        }
        it = __ballot(open) ? it + 1 : q.next;
    }
    if (!mine) return;
    if (ext_mode == 1) {
        fx += ext_coeff * me.x;
        fy += ext_coeff * me.y;
        fz += ext_coeff * me.z;
    }
    double *o = F + 3 * (i - id_min);
    o[0] = fx;
    o[1] = fy;
    o[2] = fz;
}

struct BhView {
    double *red;  // kRedBlocks partials (max |x|, min m), then root side and M_min
    uint32_t *khi, *klo, *base, *bsum;
    double4 *srec;
    double *ssoft;
    uint64_t *skey;
    uint32_t *sidx;
    int8_t *tab;
    int *info;
    BhNode *nodes;
    int8_t *level;
    void *sort;
};

// the pieces of the scratch buffer, 256-byte aligned, in this order
template <class Take> void bh_pieces(int64_t n, Take &&take) {
    const size_t nn = (size_t)std::max<int64_t>(n, 1), cap = (size_t)bh_item_capacity(n);
    take((size_t)(2 * kRedBlocks + 2) * 8);  // red
    take(nn * 4);                             // khi
    take(nn * 4);                             // klo
    take(nn * 4);                             // base
    take((size_t)scan_blocks(n) * 4);         // bsum
    take(nn * 32);                            // srec
    take(nn * 8);                             // ssoft
    take(nn * 8);                             // skey
    take(nn * 4);                             // sidx
    take(nn * (size_t)table_rows(n));         // tab
    take(16);                                 // info
    take(cap * sizeof(BhNode));               // nodes
    take(cap);                                // level
    take(sort_u64_scratch_bytes(n));          // sort
}

BhView bh_view(void *scratch, int64_t n) {
    char *base = (char *)scratch;
    char *ptr[14];
    int k = 0;
    size_t off = 0;
    bh_pieces(n, [&](size_t b) {
        ptr[k++] = base + off;
        off += a256(b);
    });
    BhView v;
    v.red = (double *)ptr[0];
    v.khi = (uint32_t *)ptr[1];
    v.klo = (uint32_t *)ptr[2];
    v.base = (uint32_t *)ptr[3];
    v.bsum = (uint32_t *)ptr[4];
    v.srec = (double4 *)ptr[5];
    v.ssoft = (double *)ptr[6];
    v.skey = (uint64_t *)ptr[7];
    v.sidx = (uint32_t *)ptr[8];
    v.tab = (int8_t *)ptr[9];
    v.info = (int *)ptr[10];
    v.nodes = (BhNode *)ptr[11];
    v.level = (int8_t *)ptr[12];
    v.sort = ptr[13];
    return v;
}

}  // namespace

int64_t bh_item_capacity(int64_t n) {
    const int64_t nn = std::max<int64_t>(n, 1);
    return nn + (int64_t)(kLevels + 1) * (nn / 2) + 64;
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
    const int rows = table_rows(n);
    hipError_t e;
    double *red = v.red + 2 * kRedBlocks;  // root side, M_min
    hipLaunchKernelGGL(bh_reduce_partial, dim3(nb), dim3(256), 0, st, dp.jrec, n, v.red);
    hipLaunchKernelGGL(bh_reduce_final, dim3(1), dim3(256), 0, st, v.red, nb, red);
    hipLaunchKernelGGL(bh_keys_kernel, dim3(ng), dim3(256), 0, st, dp.jrec, n, red, v.khi, v.klo);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const uint32_t *perm = nullptr;
    if ((e = launch_sort_u64(v.khi, v.klo, n, v.sort, &perm, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(bh_gather_kernel, dim3(ng), dim3(256), 0, st, dp.jrec, dp.soft, v.khi, v.klo, perm, n, v.srec,
                       v.ssoft, v.skey, v.sidx);
    hipLaunchKernelGGL(bh_lcp_kernel, dim3(ng), dim3(256), 0, st, v.skey, n, v.tab);
    for (int k = 1; k < rows; ++k) hipLaunchKernelGGL(bh_table_kernel, dim3(ng), dim3(256), 0, st, v.tab, n, k);
    const int64_t sb = scan_blocks(n);
    hipLaunchKernelGGL(bh_count_scan_kernel, dim3((unsigned)sb), dim3(256), 0, st, v.tab, n, v.base, v.bsum);
    hipLaunchKernelGGL(bh_scan_top_kernel, dim3(1), dim3(1024), 0, st, v.bsum, sb, v.info);
    hipLaunchKernelGGL(bh_scan_add_kernel, dim3(ng), dim3(256), 0, st, v.base, n, v.bsum);
    hipLaunchKernelGGL(bh_fill_kernel, dim3(ng), dim3(256), 0, st, v.tab, rows, v.base, v.srec, v.sidx, n, red,
                       p.bh_particle_radius, v.info, v.nodes, v.level);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const unsigned gi = (unsigned)((bh_item_capacity(n) + 255) / 256);
    for (int l = kLevels; l >= 0; --l)
        hipLaunchKernelGGL(bh_level_kernel, dim3(gi), dim3(256), 0, st, l, v.info, red, p.bh_particle_radius,
                           v.nodes, v.level);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev_tree_end && (e = hipEventRecord(ev_tree_end, st)) != hipSuccess) return e;
    hipLaunchKernelGGL(bh_force_kernel, dim3(ng), dim3(256), 0, st, v.srec, v.ssoft, v.sidx, n, v.info, v.nodes,
                       p.bh_theta, p.ext_mode, p.ext_coeff, p.id_min, p.id_min + p.nrows - 1, F);
    return hipGetLastError();
}

}  // namespace steps
