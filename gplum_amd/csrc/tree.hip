// tree.hip -- octree build, moments, i-groups and interaction-list walk on the GPU.
//
// FDPS (the reference's tree library) is not vendored; the construction rules are
// the ones stated in oracle/orc_tree.c, implemented here level-synchronously:
//   keys    : 63-bit Morton keys of the particle bbox cube, radix-sorted (stable, so
//             ties keep the previous order);
//   nodes   : breadth-first, two kernels per level (child count, child write) with
//             the level ranges kept on the device;
//   moments : bottom-up per level, one thread per node, FP64, FDPS MomentQuadrupole
//             accumulation order (src/particle.h:1649-1715);
//   groups  : maximal nodes with <= n_group_limit particles, in particle order;
//   walk    : one wave64 per i-group, breadth-first queue in global memory, 64 nodes
//             per step; opened leaves -> EP list, unopened nodes -> SP list, children
//             of opened internal nodes appended in order (ballot + wave prefix sums),
//             so the lists equal the sequential FIFO walk; one pass, each workgroup
//             writing its groups' lists into a region of its own.
#include <hipcub/hipcub.hpp>

#include <cstring>

#include "tree.hpp"

namespace gplum {

constexpr int MAXLEVEL = 21;

// ---------------------------------------------------------------- helpers

__device__ inline uint64_t spread3(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

__device__ inline int digit_at(uint64_t key, int level) {
    return (int)((key >> (3 * (MAXLEVEL - 1 - level))) & 7ULL);
}

// Exclusive prefix sum over the 64 lanes of a wave; *total receives the wave sum.
__device__ inline int wave_excl_scan(int v, int *total) {
    const int lane = __lane_id();
    int x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    *total = __shfl(x, 63);
    return x - v;
}

// ------------------------------------------------------------ bbox + keys

__global__ void k_bbox_partial(const double *__restrict__ pos, int n, double *__restrict__ part) {
    __shared__ double s[6][256];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        for (int k = 0; k < 3; k++) {
            const double x = pos[3 * i + k];
            mn[k] = x < mn[k] ? x : mn[k];
            mx[k] = x > mx[k] ? x : mx[k];
        }
    }
    for (int k = 0; k < 3; k++) { s[k][threadIdx.x] = mn[k]; s[3 + k][threadIdx.x] = mx[k]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            for (int k = 0; k < 3; k++) {
                const double a = s[k][threadIdx.x + w], b = s[3 + k][threadIdx.x + w];
                if (a < s[k][threadIdx.x]) s[k][threadIdx.x] = a;
                if (b > s[3 + k][threadIdx.x]) s[3 + k][threadIdx.x] = b;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; k++) part[6 * blockIdx.x + k] = s[k][0];
}

// one block of 256 threads: min/max over the per-block partials (exact in any order)
__global__ __launch_bounds__(256) void k_bbox_final(const double *__restrict__ part, int nb, double *__restrict__ os) {
    __shared__ double s[6][256];
    const int t = threadIdx.x;
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int b = t; b < nb; b += 256)
        for (int k = 0; k < 3; k++) {
            if (part[6 * b + k] < mn[k]) mn[k] = part[6 * b + k];
            if (part[6 * b + 3 + k] > mx[k]) mx[k] = part[6 * b + 3 + k];
        }
    for (int k = 0; k < 3; k++) { s[k][t] = mn[k]; s[3 + k][t] = mx[k]; }
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w)
            for (int k = 0; k < 3; k++) {
                if (s[k][t + w] < s[k][t]) s[k][t] = s[k][t + w];
                if (s[3 + k][t + w] > s[3 + k][t]) s[3 + k][t] = s[3 + k][t + w];
            }
        __syncthreads();
    }
    if (t != 0) return;
    for (int k = 0; k < 3; k++) { mn[k] = s[k][0]; mx[k] = s[3 + k][0]; }
    double ext = 0.;
    for (int k = 0; k < 3; k++) {
        const double e = mx[k] - mn[k];
        if (e > ext) ext = e;
    }
    if (ext == 0.) ext = 1.;
    const double side = ext * 1.0000001;
    os[0] = mn[0]; os[1] = mn[1]; os[2] = mn[2];
    os[3] = side;
    os[4] = 2097152.0 / side;
}

__device__ inline uint64_t coord21(double x, double o, double scale) {
    const double t = (x - o) * scale;
    long long c = (long long)floor(t);
    if (c < 0) c = 0;
    if (c > 2097151) c = 2097151;
    return (uint64_t)c;
}

__global__ void k_keys(const double *__restrict__ pos, int n, const double *__restrict__ os,
                       uint64_t *__restrict__ key, int32_t *__restrict__ idx) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double scale = os[4];
    const uint64_t cx = coord21(pos[3 * i], os[0], scale);
    const uint64_t cy = coord21(pos[3 * i + 1], os[1], scale);
    const uint64_t cz = coord21(pos[3 * i + 2], os[2], scale);
    key[i] = (spread3(cx) << 2) | (spread3(cy) << 1) | spread3(cz);
    idx[i] = i;
}

// ------------------------------------------------------------------ nodes

__global__ void k_root(int n, int32_t *start, int32_t *count, int32_t *level, int32_t *parent) {
    start[0] = 0; count[0] = n; level[0] = 0; parent[0] = -1;
}

// Level kernels: child count per node, then child write (children in digit order,
// appended after the level).  The level's node range is read from lvl[L], lvl[L+1]
// (the previous level's write kernel writes it), so the levels run back to back
// without a host round trip.  LEVEL_GRID workgroups each own one contiguous chunk of
// the level: the count kernel sums its chunk's children, and the write kernel turns
// the chunk sums into its chunk's base and scans its chunk in tiles.
constexpr int LEVEL_GRID = 256, LEVEL_BLOCK = 256;
constexpr int LVL_ERR = MAXLEVEL + 4;       // lvl[]: ranges 0..MAXLEVEL+2, flag, totals

// First index in [s, e) whose digit at level lv is >= d, for d = 1..7 (bnd[0] = s,
// bnd[8] = e): seven binary searches advanced in lockstep, so their loads are in flight
// together (a node of the top levels spans up to all N keys).
__device__ inline void digit_bounds(const uint64_t *__restrict__ key, int s, int e, int lv, int bnd[9]) {
    int lo[7], hi[7];
#pragma unroll
    for (int d = 0; d < 7; d++) { lo[d] = s; hi[d] = e; }
    for (int span = e - s; span > 0; span >>= 1) {
#pragma unroll
        for (int d = 0; d < 7; d++) {
            if (lo[d] < hi[d]) {
                const int mid = (lo[d] + hi[d]) >> 1;
                if (digit_at(key[mid], lv) <= d) lo[d] = mid + 1; else hi[d] = mid;
            }
        }
    }
    bnd[0] = s;
#pragma unroll
    for (int d = 0; d < 7; d++) bnd[d + 1] = lo[d];
    bnd[8] = e;
}

__device__ inline void level_chunk(const int32_t *__restrict__ lvl, int L, int &lb, int &m, int &c0, int &c1) {
    lb = lvl[L];
    m = lvl[L + 1] - lb;
    const int chunk = (m + LEVEL_GRID - 1) / LEVEL_GRID;
    c0 = min(m, (int)blockIdx.x * chunk);
    c1 = min(m, c0 + chunk);
}

// block-wide sum (LEVEL_BLOCK threads), result in every thread
__device__ inline int block_sum(int v, int *s_w) {
    int tot = 0;
    (void)wave_excl_scan(v, &tot);
    if (__lane_id() == 0) s_w[threadIdx.x >> 6] = tot;
    __syncthreads();
    int all = 0;
    for (int w = 0; w < LEVEL_BLOCK / 64; w++) all += s_w[w];
    __syncthreads();
    return all;
}

// Nodes of at least coop_min keys (Tree::coop_min, 2^16 by default: the top levels,
// where a workgroup's chunk holds one node) find their digit bounds with the whole
// workgroup: each round samples 255 cut points of each of the seven search intervals at
// once and keeps the piece holding the first key of a larger digit, so a node of N keys
// takes log_256 N rounds of parallel loads instead of log_2 N dependent ones.  Nodes of
// at least coop_min / 256 keys in a chunk of at most 32 nodes take one wave each
// (log_64 N rounds); the others keep one thread each (their searches run side by side).

// bnd8[d] = first index in [s, e) whose digit at level lv is > d (d = 0..6), bnd8[7] = e:
// the same bounds as digit_bounds (bnd[d + 1]).  Every thread of the workgroup calls it.
__device__ void digit_bounds_wg(const uint64_t *__restrict__ key, int s, int e, int lv, int *s_lo, int *s_hi,
                                int *s_first, int bnd8[8]) {
    const int t = (int)threadIdx.x;
    if (t < 7) { s_lo[t] = s; s_hi[t] = e; }
    __syncthreads();
    for (;;) {
        bool open = false;
        for (int d = 0; d < 7; d++) open = open || (s_hi[d] > s_lo[d]);
        if (!open) break;            // the same in every thread (read after a barrier)
        int dg[7];
#pragma unroll
        for (int d = 0; d < 7; d++) {          // the seven loads first, in flight together
            const int lo = s_lo[d], hi = s_hi[d];
            const int q = lo + (int)(((long long)(hi - lo) * (t + 1)) >> 8);   // cut point t, in [lo, hi)
            dg[d] = (t < LEVEL_BLOCK - 1 && hi > lo) ? digit_at(key[q], lv) : -1;
        }
        // the first cut point with a larger digit: a ballot per wave, then the smallest
        // over the waves (no atomics: the predicate rises with t)
#pragma unroll
        for (int d = 0; d < 7; d++) {
            const unsigned long long m = __ballot(dg[d] > d);
            if (__lane_id() == 0) s_first[d * (LEVEL_BLOCK / 64) + (t >> 6)] = m ? (t & ~63) + __ffsll((long long)m) - 1 : LEVEL_BLOCK - 1;
        }
        __syncthreads();
        if (t < 7) {
            const int lo = s_lo[t], hi = s_hi[t];
            if (hi > lo) {
                int f = LEVEL_BLOCK - 1;
                for (int w = 0; w < LEVEL_BLOCK / 64; w++) f = min(f, s_first[t * (LEVEL_BLOCK / 64) + w]);
                const long long len = hi - lo;
                // cut points q_k = lo + (len (k + 1)) >> 8; the first index with a larger
                // digit lies in (q_{f-1}, q_f] (f = 255: none among the cut points)
                const int nlo = (f == 0) ? lo : lo + (int)((len * f) >> 8) + 1;
                const int nhi = (f == LEVEL_BLOCK - 1) ? hi : lo + (int)((len * (f + 1)) >> 8);
                s_lo[t] = nlo;
                s_hi[t] = nhi;
            }
        }
        __syncthreads();
    }
    for (int d = 0; d < 7; d++) bnd8[d] = s_lo[d];
    bnd8[7] = e;
    __syncthreads();
}

// The same bounds found by one wave: each round samples 63 cut points of each open
// interval (lane t: cut point t) and keeps the piece holding the first key of a larger
// digit (a ballot), so a node of N keys takes log_64 N rounds.  Every lane calls it
// with the same node and receives the same bounds.
__device__ void digit_bounds_wave(const uint64_t *__restrict__ key, int s, int e, int lv, int bnd8[8]) {
    const int t = __lane_id();
    int lo[7], hi[7];
#pragma unroll
    for (int d = 0; d < 7; d++) { lo[d] = s; hi[d] = e; }
    for (;;) {
        bool open = false;
#pragma unroll
        for (int d = 0; d < 7; d++) open = open || (hi[d] > lo[d]);
        if (!open) break;                  // wave-uniform
        int dg[7];
#pragma unroll
        for (int d = 0; d < 7; d++) {
            const long long len = hi[d] - lo[d];
            const int q = lo[d] + (int)((len * (t + 1)) >> 6);    // in [lo, hi) for t < 63
            dg[d] = (t < 63 && len > 0) ? digit_at(key[q], lv) : -1;
        }
#pragma unroll
        for (int d = 0; d < 7; d++) {
            if (hi[d] <= lo[d]) continue;
            const unsigned long long m = __ballot(dg[d] > d);
            const int f = m ? __ffsll((long long)m) - 1 : 63;     // 63: none of the cut points
            const long long len = hi[d] - lo[d];
            // cut points q_k = lo + (len (k + 1)) >> 6; the first index with a larger
            // digit lies in (q_{f-1}, q_f]
            const int nlo = (f == 0) ? lo[d] : lo[d] + (int)((len * f) >> 6) + 1;
            const int nhi = (f == 63) ? hi[d] : lo[d] + (int)((len * (f + 1)) >> 6);
            lo[d] = nlo;
            hi[d] = nhi;
        }
    }
#pragma unroll
    for (int d = 0; d < 7; d++) bnd8[d] = lo[d];
    bnd8[7] = e;
}

__global__ __launch_bounds__(LEVEL_BLOCK) void k_count_children_lv(
    const uint64_t *__restrict__ key, const int32_t *__restrict__ lvl, int L, int n_leaf,
    const int32_t *__restrict__ start, const int32_t *__restrict__ count, const int32_t *__restrict__ level,
    int32_t *__restrict__ nchild, int32_t *__restrict__ part, int32_t *__restrict__ bnd_out, int coop_min,
    int wave_min) {
    __shared__ int s_w[LEVEL_BLOCK / 64];
    __shared__ int s_big[LEVEL_BLOCK], s_nbig, s_lo[7], s_hi[7], s_first[7 * (LEVEL_BLOCK / 64)];
    __shared__ int s_mid[LEVEL_BLOCK], s_nmid;
    int lb, m, c0, c1;
    level_chunk(lvl, L, lb, m, c0, c1);
    if (m <= 0) return;
    int sum = 0;
    for (int r0 = c0; r0 < c1; r0 += LEVEL_BLOCK) {
        if (threadIdx.x == 0) { s_nbig = 0; s_nmid = 0; }
        __syncthreads();
        const int r = r0 + (int)threadIdx.x;
        if (r < c1) {
            const int c = lb + r;
            const int s = start[c], cnt = count[c], lv = level[c];
            int nc = 0;
            if (cnt > n_leaf && lv < MAXLEVEL) {
                if (cnt >= coop_min) {
                    s_big[atomicAdd(&s_nbig, 1)] = r;   // the workgroup searches it below
                    nc = -1;
                } else if (cnt >= wave_min && c1 - c0 <= 32) {
                    // a wave searches it below (few nodes per workgroup: otherwise one
                    // thread per node, all side by side, takes fewer rounds)
                    s_mid[atomicAdd(&s_nmid, 1)] = r;
                    nc = -1;
                } else {
                    int bnd[9];
                    digit_bounds(key, s, s + cnt, lv, bnd);
#pragma unroll
                    for (int d = 0; d < 8; d++) {
                        nc += bnd[d + 1] > bnd[d] ? 1 : 0;
                        bnd_out[8 * (size_t)r + d] = bnd[d + 1];
                    }
                }
            }
            if (nc >= 0) {
                nchild[r] = nc;
                sum += nc;
            }
        }
        __syncthreads();
        const int nmid = s_nmid;
        for (int k = (int)(threadIdx.x >> 6); k < nmid; k += LEVEL_BLOCK / 64) {
            const int rb = s_mid[k];
            const int c = lb + rb;
            const int s = start[c];
            int bnd8[8];
            digit_bounds_wave(key, s, s + count[c], level[c], bnd8);
            if (__lane_id() == 0) {
                int nc = 0, prev = s;
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    nc += bnd8[d] > prev ? 1 : 0;
                    prev = bnd8[d];
                    bnd_out[8 * (size_t)rb + d] = bnd8[d];
                }
                nchild[rb] = nc;
                sum += nc;
            }
        }
        const int nbig = s_nbig;
        for (int k = 0; k < nbig; k++) {
            const int rb = s_big[k];
            const int c = lb + rb;
            const int s = start[c];
            int bnd8[8];
            digit_bounds_wg(key, s, s + count[c], level[c], s_lo, s_hi, s_first, bnd8);
            if (threadIdx.x == 0) {
                int nc = 0, prev = s;
#pragma unroll
                for (int d = 0; d < 8; d++) {
                    nc += bnd8[d] > prev ? 1 : 0;
                    prev = bnd8[d];
                    bnd_out[8 * (size_t)rb + d] = bnd8[d];
                }
                nchild[rb] = nc;
                sum += nc;
            }
        }
        __syncthreads();   // s_nbig read by every thread before the next tile resets it
    }
    const int all = block_sum(sum, s_w);
    if (threadIdx.x == 0) part[blockIdx.x] = all;
}

// The chunk sums of the count kernel give each workgroup its base (the sum over the
// chunks before it) and the level's total; workgroup 0 publishes the next level's
// range lvl[L+2] = lvl[L+1] + total, or the overflow flag lvl[LVL_ERR] past cap (then
// nothing is written).
__global__ __launch_bounds__(LEVEL_BLOCK) void k_write_children_lv(
    const uint64_t *__restrict__ key, int32_t *__restrict__ lvl, int L, const int32_t *__restrict__ part,
    int cap, int32_t *__restrict__ start, int32_t *__restrict__ count, int32_t *__restrict__ level,
    int32_t *__restrict__ parent, int32_t *__restrict__ child, int32_t *__restrict__ nchild_out,
    const int32_t *__restrict__ nchild, const int32_t *__restrict__ bnd_in) {
    static_assert(LEVEL_GRID == LEVEL_BLOCK, "one chunk sum per thread");
    __shared__ int s_w[LEVEL_BLOCK / 64];
    int lb, m, c0, c1;
    level_chunk(lvl, L, lb, m, c0, c1);
    if (m <= 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) lvl[L + 2] = lvl[L + 1];   // empty: so is the next
        return;
    }
    const int next_begin = lvl[L + 1];
    const int ps = part[threadIdx.x];
    const int before = block_sum((int)threadIdx.x < (int)blockIdx.x ? ps : 0, s_w);
    const int total = block_sum(ps, s_w);
    const bool over = next_begin + total > cap;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        if (over) { lvl[LVL_ERR] = 1; lvl[L + 2] = next_begin; }
        else lvl[L + 2] = next_begin + total;
    }
    if (over) return;
    int carry = before;
    for (int r0 = c0; r0 < c1; r0 += LEVEL_BLOCK) {      // tiles of the chunk, in order
        const int r = r0 + (int)threadIdx.x;
        const int nc = (r < c1) ? nchild[r] : 0;
        int tot = 0;
        const int ex = wave_excl_scan(nc, &tot);
        if (__lane_id() == 0) s_w[threadIdx.x >> 6] = tot;
        __syncthreads();
        int base = carry, all = 0;
        for (int w = 0; w < LEVEL_BLOCK / 64; w++) {
            if (w < (int)(threadIdx.x >> 6)) base += s_w[w];
            all += s_w[w];
        }
        __syncthreads();
        carry += all;
        if (r >= c1) continue;
        const int c = lb + r;
        nchild_out[c] = nc;
        const int first = next_begin + base + ex;
        child[c] = nc ? first : -1;
        if (!nc) continue;
        // the digit bounds found by the count kernel
        const int s = start[c], lv = level[c];
        int bnd[9];
        bnd[0] = s;
#pragma unroll
        for (int d = 0; d < 8; d++) bnd[d + 1] = bnd_in[8 * (size_t)r + d];
        int w = first;
#pragma unroll
        for (int d = 0; d < 8; d++) {
            if (bnd[d + 1] > bnd[d]) {          // non-empty children in digit order
                start[w] = bnd[d]; count[w] = bnd[d + 1] - bnd[d]; level[w] = lv + 1; parent[w] = c;
                w++;
            }
        }
    }
}

__global__ void k_lvl_init(int32_t *lvl) {
    if (threadIdx.x < LVL_ERR + 1) lvl[threadIdx.x] = (threadIdx.x == 1) ? 1 : 0;
}

// ---------------------------------------------------------------- moments

// The levels up to L_end have been built (launched before the host knows how many levels
// the tree has) and the tree is whole: no node overflow, and no nodes at level L_end,
// whose children would not have been built.  Kernels after the level build check this
// on the device and do nothing otherwise; the host reads the ranges back with the walk's
// results and finishes the tree then.
__device__ inline bool tree_whole(const int32_t *__restrict__ lvl, int L_end) {
    return !lvl[LVL_ERR] && !(L_end <= MAXLEVEL && lvl[L_end + 1] > lvl[L_end]);
}

// Moments of node c: mass, centre of mass, quadrupole about it, bounding box and largest
// search radius, from its particles (a leaf) or its children (in child order).
__device__ inline void node_moments(int c, const double *__restrict__ pos, const double *__restrict__ mass,
                                    const double *__restrict__ rsearch, const int32_t *__restrict__ start,
                                    const int32_t *__restrict__ count, const int32_t *__restrict__ child,
                                    const int32_t *__restrict__ nchild, double *__restrict__ nm,
                                    double *__restrict__ ncom, double *__restrict__ nq, double *__restrict__ bmin,
                                    double *__restrict__ bmax, double *__restrict__ nrs) {
    double m = 0., mp[3] = {0., 0., 0.}, q[6] = {0., 0., 0., 0., 0., 0.};
    double bmn[3], bmx[3], rs = 0., com[3];
    const int s = start[c], cnt = count[c];
    if (nchild[c] == 0) {
        for (int k = 0; k < 3; k++) { bmn[k] = pos[3 * s + k]; bmx[k] = pos[3 * s + k]; }
        for (int a = s; a < s + cnt; a++) {
            const double ma = mass[a];
            m += ma;
            for (int k = 0; k < 3; k++) {
                const double x = pos[3 * a + k];
                mp[k] += ma * x;
                if (x < bmn[k]) bmn[k] = x;
                if (x > bmx[k]) bmx[k] = x;
            }
            if (rsearch[a] > rs) rs = rsearch[a];
        }
        for (int k = 0; k < 3; k++) com[k] = (m != 0.) ? mp[k] / m : 0.;
        for (int a = s; a < s + cnt; a++) {
            const double ma = mass[a];
            const double dx = pos[3 * a] - com[0], dy = pos[3 * a + 1] - com[1], dz = pos[3 * a + 2] - com[2];
            const double cx = ma * dx, cy = ma * dy, cz = ma * dz;
            q[0] += cx * dx; q[1] += cy * dy; q[2] += cz * dz;
            q[3] += cx * dy; q[4] += cx * dz; q[5] += cy * dz;
        }
    } else {
        const int c0 = child[c], nc = nchild[c];
        for (int k = 0; k < 3; k++) { bmn[k] = bmin[3 * c0 + k]; bmx[k] = bmax[3 * c0 + k]; }
        for (int ch = c0; ch < c0 + nc; ch++) {
            m += nm[ch];
            for (int k = 0; k < 3; k++) {
                mp[k] += nm[ch] * ncom[3 * ch + k];
                if (bmin[3 * ch + k] < bmn[k]) bmn[k] = bmin[3 * ch + k];
                if (bmax[3 * ch + k] > bmx[k]) bmx[k] = bmax[3 * ch + k];
            }
            if (nrs[ch] > rs) rs = nrs[ch];
        }
        for (int k = 0; k < 3; k++) com[k] = (m != 0.) ? mp[k] / m : 0.;
        for (int ch = c0; ch < c0 + nc; ch++) {
            const double mt = nm[ch];
            const double dx = ncom[3 * ch] - com[0], dy = ncom[3 * ch + 1] - com[1], dz = ncom[3 * ch + 2] - com[2];
            const double cx = mt * dx, cy = mt * dy, cz = mt * dz;
            const double *qc = &nq[6 * ch];
            q[0] += cx * dx + qc[0]; q[1] += cy * dy + qc[1]; q[2] += cz * dz + qc[2];
            q[3] += cx * dy + qc[3]; q[4] += cx * dz + qc[4]; q[5] += cy * dz + qc[5];
        }
    }
    nm[c] = m;
    for (int k = 0; k < 3; k++) { ncom[3 * c + k] = com[k]; bmin[3 * c + k] = bmn[k]; bmax[3 * c + k] = bmx[k]; }
    for (int k = 0; k < 6; k++) nq[6 * c + k] = q[k];
    nrs[c] = rs;
}

// one level's moments (its range read from lvl), the levels launched deepest first
__global__ void k_moments(int L, int L_end, const int32_t *__restrict__ lvl, const double *__restrict__ pos,
                          const double *__restrict__ mass, const double *__restrict__ rsearch,
                          const int32_t *__restrict__ start, const int32_t *__restrict__ count,
                          const int32_t *__restrict__ child, const int32_t *__restrict__ nchild,
                          double *__restrict__ nm, double *__restrict__ ncom, double *__restrict__ nq,
                          double *__restrict__ bmin, double *__restrict__ bmax, double *__restrict__ nrs) {
    if (!tree_whole(lvl, L_end)) return;
    const int le = lvl[L + 1];
    for (int c = lvl[L] + blockIdx.x * blockDim.x + threadIdx.x; c < le; c += gridDim.x * blockDim.x)
        node_moments(c, pos, mass, rsearch, start, count, child, nchild, nm, ncom, nq, bmin, bmax, nrs);
}

// ----------------------------------------------------------------- groups

// the node count lvl[L_end] is read on the device; no flags (so no groups) unless the
// tree is whole
__global__ void k_group_flags(int L_end, const int32_t *__restrict__ lvl, int G, const int32_t *__restrict__ start,
                              const int32_t *__restrict__ count, const int32_t *__restrict__ nchild,
                              const int32_t *__restrict__ parent, int32_t *__restrict__ flag) {
    if (!tree_whole(lvl, L_end)) return;
    const int n_node = lvl[L_end];
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < n_node; c += gridDim.x * blockDim.x) {
        const int cnt = count[c];
        const bool cand = (cnt <= G) || (nchild[c] == 0);
        const bool top = (parent[c] < 0) || (count[parent[c]] > G);
        if (!(cand && top)) continue;
        for (int off = 0; off < cnt; off += G) flag[start[c] + off] = c + 1;
    }
}

__global__ void k_group_write(int n, int G, const int32_t *__restrict__ flag,
                              const int32_t *__restrict__ scan, const int32_t *__restrict__ start,
                              const int32_t *__restrict__ count, int32_t *__restrict__ gnode,
                              int32_t *__restrict__ gstart, int32_t *__restrict__ gcount) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= n) return;
    const int f = flag[a];
    if (!f) return;
    const int c = f - 1;
    const int g = scan[a];
    gnode[g] = c;
    gstart[g] = a;
    const int rem = start[c] + count[c] - a;
    gcount[g] = rem < G ? rem : G;
}

// the number of groups (last exclusive offset + last flag), left on the device for the walk
__global__ void k_group_total(int n, const int32_t *__restrict__ flag01, const int32_t *__restrict__ scan,
                              int32_t *__restrict__ ng) {
    if (threadIdx.x == 0) *ng = n > 0 ? scan[n - 1] + flag01[n - 1] : 0;
}

__global__ void k_flag01(int n, const int32_t *__restrict__ flag, int32_t *__restrict__ out) {
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < n) out[a] = flag[a] ? 1 : 0;
}

// ------------------------------------------------------------------- walk

// One pass: workgroup b owns the list regions [b*cap_q, (b+1)*cap_q) of the three
// lists (q: EP near, EP far, SP) and writes the lists of its groups one after another
// into them, so a group's lists are contiguous, in the order of the sequential FIFO
// walk, and placed with no atomics and no second pass; the soft kernel reads them
// through the per-group offsets and counts only.  use[3b+q] receives the entries the
// workgroup wrote.  err bits: 1 queue, 4 region too small.
__global__ __launch_bounds__(64) void k_walk(
    const int32_t *__restrict__ ng_dev, const int32_t *__restrict__ gstart, const int32_t *__restrict__ gcount,
    const double *__restrict__ pos, const double *__restrict__ rsearch, const double *__restrict__ os,
    double theta2, const int32_t *__restrict__ nstart, const int32_t *__restrict__ ncount,
    const int32_t *__restrict__ nchild_first, const int32_t *__restrict__ nchild,
    const int32_t *__restrict__ nlevel, const double *__restrict__ ncom, const double *__restrict__ bmin,
    const double *__restrict__ bmax, const double *__restrict__ nrs, int32_t *__restrict__ queue_all,
    int qcap, int32_t *__restrict__ ep_cnt, int32_t *__restrict__ epf_cnt, int32_t *__restrict__ sp_cnt,
    int32_t *__restrict__ ep_off, int32_t *__restrict__ epf_off, int32_t *__restrict__ sp_off,
    int32_t *__restrict__ ep_list, int32_t *__restrict__ epf_list, int32_t *__restrict__ sp_list,
    int cap_ep, int cap_epf, int cap_sp, int32_t *__restrict__ use, int32_t *__restrict__ err) {
    const int lane = __lane_id();
    const int n_group = *ng_dev;
    int32_t *queue = queue_all + (size_t)blockIdx.x * qcap;
    const int64_t r_ep = (int64_t)blockIdx.x * cap_ep, r_epf = (int64_t)blockIdx.x * cap_epf,
                  r_sp = (int64_t)blockIdx.x * cap_sp;
    int u_ep = 0, u_epf = 0, u_sp = 0;      // entries of the region used so far
    const double side = os[3];
    int fail = 0;
    for (int g = blockIdx.x; g < n_group && !fail; g += gridDim.x) {
        const int s = gstart[g], cnt = gcount[g];
        // group bbox and max r_search (wave reduction; exact)
        double gmn[3] = {1e300, 1e300, 1e300}, gmx[3] = {-1e300, -1e300, -1e300}, grs = 0.;
        for (int a = s + lane; a < s + cnt; a += 64) {
            for (int k = 0; k < 3; k++) {
                const double x = pos[3 * a + k];
                gmn[k] = x < gmn[k] ? x : gmn[k];
                gmx[k] = x > gmx[k] ? x : gmx[k];
            }
            grs = rsearch[a] > grs ? rsearch[a] : grs;
        }
        for (int d = 32; d > 0; d >>= 1) {
            for (int k = 0; k < 3; k++) {
                const double a = __shfl_xor(gmn[k], d), b = __shfl_xor(gmx[k], d);
                gmn[k] = a < gmn[k] ? a : gmn[k];
                gmx[k] = b > gmx[k] ? b : gmx[k];
            }
            const double r = __shfl_xor(grs, d);
            grs = r > grs ? r : grs;
        }
        int head = 0, tail = 1;
        if (lane == 0) queue[0] = 0;
        int n_ep = 0, n_epf = 0, n_sp = 0;
        int32_t *const d_ep = ep_list + r_ep + u_ep;
        int32_t *const d_epf = epf_list + r_epf + u_epf;
        int32_t *const d_sp = sp_list + r_sp + u_sp;
        while (head < tail) {
            __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
            __syncthreads();
            const int chunk = min(64, tail - head);
            int c = -1;
            bool open = false, leaf = false, near = false;
            int nc = 0, cc = 0;
            if (lane < chunk) {
                c = queue[head + lane];
                double d2 = 0., gap2 = 0.;
                for (int k = 0; k < 3; k++) {
                    const double cm = ncom[3 * c + k];
                    double dd = 0.;
                    if (cm < gmn[k]) dd = gmn[k] - cm;
                    else if (cm > gmx[k]) dd = cm - gmx[k];
                    d2 += dd * dd;
                    double gg = 0.;
                    if (bmax[3 * c + k] < gmn[k]) gg = gmn[k] - bmax[3 * c + k];
                    else if (bmin[3 * c + k] > gmx[k]) gg = bmin[3 * c + k] - gmx[k];
                    gap2 += gg * gg;
                }
                const double w = side / (double)(1ULL << nlevel[c]);
                const double rsm = 1.1 * (grs > nrs[c] ? grs : nrs[c]);
                near = gap2 < rsm * rsm;
                open = (w * w > theta2 * d2) || near;
                nc = nchild[c];
                leaf = nc == 0;
                cc = ncount[c];
            }
            const bool f_ep = (lane < chunk) && open && leaf;
            const bool f_epn = f_ep && near;
            const bool f_epf = f_ep && !near;
            const bool f_in = (lane < chunk) && open && !leaf;
            const bool f_sp = (lane < chunk) && !open;
            int t_ep, t_epf, t_sp, t_ch;
            const int o_ep = wave_excl_scan(f_epn ? cc : 0, &t_ep);
            const int o_epf = wave_excl_scan(f_epf ? cc : 0, &t_epf);
            const int o_sp = wave_excl_scan(f_sp ? 1 : 0, &t_sp);
            const int o_ch = wave_excl_scan(f_in ? nc : 0, &t_ch);
            // the totals are wave-uniform, so every lane leaves the loop together
            if (tail + t_ch > qcap) { fail = 1; break; }
            if (u_ep + n_ep + t_ep > cap_ep || u_epf + n_epf + t_epf > cap_epf || u_sp + n_sp + t_sp > cap_sp) {
                fail = 4;
                break;
            }
            if (f_in) {
                const int c0 = nchild_first[c];
                for (int k = 0; k < nc; k++) queue[tail + o_ch + k] = c0 + k;
            }
            if (f_ep) {
                const int st = nstart[c];
                int32_t *dst = f_epn ? d_ep + n_ep + o_ep : d_epf + n_epf + o_epf;
                for (int k = 0; k < cc; k++) dst[k] = st + k;
            }
            if (f_sp) d_sp[n_sp + o_sp] = c;
            n_ep += t_ep;
            n_epf += t_epf;
            n_sp += t_sp;
            tail += t_ch;
            head += chunk;
        }
        if (!fail && lane == 0) {
            ep_off[g] = (int32_t)(r_ep + u_ep);
            epf_off[g] = (int32_t)(r_epf + u_epf);
            sp_off[g] = (int32_t)(r_sp + u_sp);
            ep_cnt[g] = n_ep;
            epf_cnt[g] = n_epf;
            sp_cnt[g] = n_sp;
        }
        u_ep += n_ep;
        u_epf += n_epf;
        u_sp += n_sp;
        __syncthreads();                    // the next group reuses the queue
    }
    if (lane == 0) {
        if (fail) atomicOr(err, fail);
        use[3 * blockIdx.x] = u_ep;
        use[3 * blockIdx.x + 1] = u_epf;
        use[3 * blockIdx.x + 2] = u_sp;
    }
}

// one workgroup: per list the largest use of a region (out[0..2]) and the total
// (out[3..5])
__global__ __launch_bounds__(1024) void k_walk_use(int nblk, const int32_t *__restrict__ use,
                                                   long long *__restrict__ out) {
    __shared__ long long s[6][1024];
    long long mx[3] = {0, 0, 0}, sm[3] = {0, 0, 0};
    for (int b = threadIdx.x; b < nblk; b += 1024)
        for (int q = 0; q < 3; q++) {
            const long long u = use[3 * b + q];
            mx[q] = u > mx[q] ? u : mx[q];
            sm[q] += u;
        }
    for (int q = 0; q < 3; q++) {
        s[q][threadIdx.x] = mx[q];
        s[3 + q][threadIdx.x] = sm[q];
    }
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 3; q++) {
                const long long a = s[q][threadIdx.x + w];
                if (a > s[q][threadIdx.x]) s[q][threadIdx.x] = a;
                s[3 + q][threadIdx.x] += s[3 + q][threadIdx.x + w];
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[threadIdx.x] = s[threadIdx.x][0];
}

// ------------------------------------------------------------- host side

template <class T>
static int excl_scan(Tree &t, const T *in, T *out, int n, hipStream_t st) {
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, st));
    GPLUM_ENSURE(t.scan_tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceScan::ExclusiveSum(t.scan_tmp.p, bytes, in, out, n, st));
    return GPLUM_OK;
}

int tree_sort_keys(Tree &t, const double *d_pos, int n, hipStream_t st) {
    const int nb = std::min(1024, std::max(1, ceil_div(n, 256)));
    GPLUM_ENSURE(t.bbox, 6 * nb + 8);
    GPLUM_ENSURE(t.origin_scale, 8);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nb), dim3(256), 0, st, d_pos, n, t.bbox.p);
    hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(256), 0, st, t.bbox.p, nb, t.origin_scale.p);
    GPLUM_ENSURE(t.key, n);
    GPLUM_ENSURE(t.key_alt, n);
    GPLUM_ENSURE(t.idx, n);
    GPLUM_ENSURE(t.perm, n);
    hipLaunchKernelGGL(k_keys, dim3(ceil_div(n, 256)), dim3(256), 0, st, d_pos, n, t.origin_scale.p,
                       t.key_alt.p, t.idx.p);
    size_t bytes = 0;
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, t.key_alt.p, t.key.p, t.idx.p,
                                                  t.perm.p, n, 0, 63, st));
    GPLUM_ENSURE(t.scan_tmp, bytes + 16);
    GPLUM_HIP(hipcub::DeviceRadixSort::SortPairs(t.scan_tmp.p, bytes, t.key_alt.p, t.key.p, t.idx.p,
                                                  t.perm.p, n, 0, 63, st));
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

// Level kernels for levels [L0, L1) (ranges kept on the device; a level past the last one
// is empty and its kernels return at once).
static void launch_levels(Tree &t, const gplum_params &p, int L0, int L1, int cap, hipStream_t st) {
    for (int L = L0; L < L1; L++) {
        hipLaunchKernelGGL(k_count_children_lv, dim3(LEVEL_GRID), dim3(LEVEL_BLOCK), 0, st, t.key.p, t.lvl.p, L,
                           p.n_leaf_limit, t.node_start.p, t.node_count.p, t.node_level.p, t.grp_flag.p,
                           t.lvl_part.p, t.lvl_bnd.p, t.coop_min, std::max(1, t.coop_min >> 8));
        hipLaunchKernelGGL(k_write_children_lv, dim3(LEVEL_GRID), dim3(LEVEL_BLOCK), 0, st, t.key.p, t.lvl.p, L,
                           t.lvl_part.p, cap, t.node_start.p, t.node_count.p, t.node_level.p, t.node_parent.p,
                           t.node_child.p, t.node_nchild.p, t.grp_flag.p, t.lvl_bnd.p);
    }
}

// Moments (levels L_end-1 .. 0) and i-groups of a tree built up to level L_end, all
// guarded on the device by tree_whole.  Launch widths follow the previous tree's level
// sizes (the kernels loop over whatever the level holds).
static int launch_moments_groups(Tree &t, const gplum_params &p, int L_end, hipStream_t st) {
    const int n = t.src_n;
    for (int L = L_end - 1; L >= 0; L--) {
        int hint = ceil_div(n, 4);
        if (L + 1 < (int)t.level_begin.size()) {
            const int sz = t.level_begin[L + 1] - t.level_begin[L];
            hint = sz + sz / 4 + 64;
        }
        hipLaunchKernelGGL(k_moments, dim3(std::max(1, ceil_div(hint, 128))), dim3(128), 0, st, L, L_end, t.lvl.p,
                           t.src_pos, t.src_mass, t.src_rs, t.node_start.p, t.node_count.p, t.node_child.p,
                           t.node_nchild.p, t.node_mass.p, t.node_com.p, t.node_quad.p, t.node_bmin.p,
                           t.node_bmax.p, t.node_rs.p);
    }
    GPLUM_HIP(hipMemsetAsync(t.grp_flag.p, 0, sizeof(int32_t) * n, st));
    hipLaunchKernelGGL(k_group_flags, dim3(std::max(1, ceil_div(n, 256))), dim3(256), 0, st, L_end, t.lvl.p,
                       p.n_group_limit, t.node_start.p, t.node_count.p, t.node_nchild.p, t.node_parent.p,
                       t.grp_flag.p);
    hipLaunchKernelGGL(k_flag01, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, t.grp_flag.p, t.nchild_scan.p);
    int rc = excl_scan(t, t.nchild_scan.p, t.grp_scan.p, n, st);
    if (rc) return rc;
    // the group count stays on the device: the walk reads it there and reads it back with
    // its own results (at most one group per particle until then)
    hipLaunchKernelGGL(k_group_total, dim3(1), dim3(64), 0, st, n, t.nchild_scan.p, t.grp_scan.p, t.ng_dev.p);
    hipLaunchKernelGGL(k_group_write, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, p.n_group_limit,
                       t.grp_flag.p, t.grp_scan.p, t.node_start.p, t.node_count.p, t.group_node.p,
                       t.group_start.p, t.group_count.p);
    GPLUM_HIP(hipGetLastError());
    return GPLUM_OK;
}

int tree_build_nodes(Tree &t, const gplum_params &p, const double *d_pos, const double *d_mass,
                     const double *d_rsearch, int n, hipStream_t st) {
    const int cap = 2 * n + 64;
    GPLUM_ENSURE(t.node_start, cap);
    GPLUM_ENSURE(t.node_count, cap);
    GPLUM_ENSURE(t.node_child, cap);
    GPLUM_ENSURE(t.node_nchild, cap);
    GPLUM_ENSURE(t.node_level, cap);
    GPLUM_ENSURE(t.node_parent, cap);
    GPLUM_ENSURE(t.nchild_scan, std::max(n, cap));
    GPLUM_ENSURE(t.grp_flag, cap);   // per-level child counts, then per-particle group flags
    GPLUM_ENSURE(t.grp_scan, n);
    // the node count is known only on the device until the walk's read-back: node arrays
    // at the node capacity
    GPLUM_ENSURE(t.node_mass, cap);
    GPLUM_ENSURE(t.node_com, 3 * (size_t)cap);
    GPLUM_ENSURE(t.node_quad, 6 * (size_t)cap);
    GPLUM_ENSURE(t.node_bmin, 3 * (size_t)cap);
    GPLUM_ENSURE(t.node_bmax, 3 * (size_t)cap);
    GPLUM_ENSURE(t.node_rs, cap);
    GPLUM_ENSURE(t.ng_dev, 1);
    GPLUM_ENSURE(t.group_node, std::max(n, 1));
    GPLUM_ENSURE(t.group_start, std::max(n, 1));
    GPLUM_ENSURE(t.group_count, std::max(n, 1));
    hipLaunchKernelGGL(k_root, dim3(1), dim3(1), 0, st, n, t.node_start.p, t.node_count.p,
                       t.node_level.p, t.node_parent.p);
    GPLUM_ENSURE(t.lvl, LVL_ERR + 5);
    hipLaunchKernelGGL(k_lvl_init, dim3(1), dim3(64), 0, st, t.lvl.p);
    GPLUM_ENSURE(t.lvl_part, LEVEL_GRID);
    GPLUM_ENSURE(t.lvl_bnd, 8 * ((size_t)n + 64));   // a level holds at most n nodes
    // levels up to one past the previous tree's depth (their kernels cost launches even
    // when empty); a deeper tree is finished after the walk's read-back (tree_walk)
    t.src_pos = d_pos;
    t.src_mass = d_mass;
    t.src_rs = d_rsearch;
    t.src_n = n;
    t.cap = cap;
    t.L_end = t.n_level > 0 ? std::min(MAXLEVEL + 1, t.n_level + 2) : MAXLEVEL + 1;
    launch_levels(t, p, 0, t.L_end, cap, st);
    t.n_part = n;
    t.n_group = -1;
    return launch_moments_groups(t, p, t.L_end, st);
}

int tree_walk(Tree &t, const gplum_params &p, const double *d_pos, const double *d_rsearch,
              hipStream_t st) {
    // the group count is on the device (at most one group per particle)
    const int ng_max = std::max(t.n_part, 1);
    GPLUM_ENSURE(t.ep_cnt, ng_max + 1);
    GPLUM_ENSURE(t.epf_cnt, ng_max + 1);
    GPLUM_ENSURE(t.sp_cnt, ng_max + 1);
    GPLUM_ENSURE(t.ep_off, ng_max + 1);
    GPLUM_ENSURE(t.epf_off, ng_max + 1);
    GPLUM_ENSURE(t.sp_off, ng_max + 1);
    const int nblk = std::min(ng_max, 4096);
    GPLUM_ENSURE(t.walk_use, 3 * (size_t)nblk);
    GPLUM_ENSURE(t.walk_err, 2);
    GPLUM_ENSURE(t.walk_stat, 6);
    GPLUM_HIP(t.walk_host.ensure(8 * sizeof(long long) + (LVL_ERR + 1) * sizeof(int32_t)));
    const double theta2 = p.theta * p.theta;
    for (int attempt = 0; attempt < 12; attempt++) {
        for (int q = 0; q < 3; q++)
            if ((int64_t)nblk * t.region_cap[q] > INT32_MAX)
                return fail(GPLUM_E_OVERFLOW, "interaction lists exceed 2^31 entries");
        GPLUM_ENSURE(t.ep_list, (size_t)nblk * t.region_cap[0]);
        GPLUM_ENSURE(t.epf_list, (size_t)nblk * t.region_cap[1]);
        GPLUM_ENSURE(t.sp_list, (size_t)nblk * t.region_cap[2]);
        GPLUM_ENSURE(t.walk_queue, (size_t)nblk * t.queue_cap);
        GPLUM_HIP(hipMemsetAsync(t.walk_err.p, 0, sizeof(int32_t), st));
        hipLaunchKernelGGL(k_walk, dim3(nblk), dim3(64), 0, st, t.ng_dev.p, t.group_start.p, t.group_count.p, d_pos,
                           d_rsearch, t.origin_scale.p, theta2, t.node_start.p, t.node_count.p,
                           t.node_child.p, t.node_nchild.p, t.node_level.p, t.node_com.p, t.node_bmin.p,
                           t.node_bmax.p, t.node_rs.p, t.walk_queue.p, t.queue_cap, t.ep_cnt.p, t.epf_cnt.p,
                           t.sp_cnt.p, t.ep_off.p, t.epf_off.p, t.sp_off.p, t.ep_list.p, t.epf_list.p,
                           t.sp_list.p, t.region_cap[0], t.region_cap[1], t.region_cap[2], t.walk_use.p,
                           t.walk_err.p);
        hipLaunchKernelGGL(k_walk_use, dim3(1), dim3(1024), 0, st, nblk, t.walk_use.p, t.walk_stat.p);
        long long *h = t.walk_host.as<long long>();
        int32_t *lv = reinterpret_cast<int32_t *>(h + 8);
        GPLUM_HIP(hipMemcpyAsync(h, t.walk_stat.p, 6 * sizeof(long long), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(h + 6, t.walk_err.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(h + 7, t.ng_dev.p, sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipMemcpyAsync(lv, t.lvl.p, (LVL_ERR + 1) * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        GPLUM_HIP(hipStreamSynchronize(st));
        GPLUM_HIP(hipGetLastError());
        // the tree first: the kernels after the level build did nothing unless it is whole
        if (lv[LVL_ERR]) return fail(GPLUM_E_OVERFLOW, "tree node capacity exceeded");
        if (t.L_end <= MAXLEVEL && lv[t.L_end + 1] > lv[t.L_end]) {
            // level L_end holds nodes: a deeper tree than before; its remaining levels,
            // then moments, groups and the walk again
            launch_levels(t, p, t.L_end, MAXLEVEL + 1, t.cap, st);
            t.L_end = MAXLEVEL + 1;
            int rc = launch_moments_groups(t, p, t.L_end, st);
            if (rc) return rc;
            continue;
        }
        if (t.n_group < 0) {
            t.level_begin.assign(1, 0);
            for (int L = 1; L <= MAXLEVEL + 2 && lv[L] > lv[L - 1]; L++) t.level_begin.push_back(lv[L]);
            t.n_level = (int)t.level_begin.size() - 1;
            t.n_node = t.level_begin.back();
        }
        int32_t err, ng;
        std::memcpy(&err, h + 6, sizeof(int32_t));
        std::memcpy(&ng, h + 7, sizeof(int32_t));
        t.n_group = ng;
        if (err & 1) t.queue_cap *= 4;
        if (err & 4)    // the use of a workgroup that stopped early is a lower bound
            for (int q = 0; q < 3; q++)
                t.region_cap[q] = (int)std::min<int64_t>(std::max<int64_t>(2 * (int64_t)t.region_cap[q], h[q] + 64),
                                                         INT32_MAX);
        if (!err) {
            t.n_ep_total = h[3];
            t.n_epf_total = h[4];
            t.n_sp_total = h[5];
            // regions for the next walk: the largest use and an eighth more
            for (int q = 0; q < 3; q++) {
                const int64_t want = h[q] + h[q] / 8 + 64;
                if (want > t.region_cap[q] || want < t.region_cap[q] / 2) t.region_cap[q] = (int)std::min<int64_t>(want, INT32_MAX);
            }
            return GPLUM_OK;
        }
        if (t.queue_cap > (1 << 26)) break;
    }
    return fail(GPLUM_E_OVERFLOW, "tree walk queue or list regions overflow");
}

}  // namespace gplum
