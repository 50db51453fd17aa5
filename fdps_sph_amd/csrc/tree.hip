// tree.hip -- on-device replacement of FDPS's tree / i-group / interaction-list
// construction (FDPS v5.0c, used at src/main.cpp:79-91,139-146; not vendored).
//
// Particles are sorted by (system, 63-bit Morton key).  For each adjacent
// pair (k-1, k) the particle count of the smallest octree cell holding both
// is computed by two binary searches on the sorted keys; a node boundary at
// threshold T sits exactly where that count exceeds T.  Leaves use
// T = n_leaf, tree level l uses n_leaf * 8^l and the top level holds one node
// per system.  Every such node is an octree cell (or a run of equal keys),
// nodes nest level by level, and node moments are summed sequentially in index
// order so that the CPU oracle reproduces them bit for bit.  i-groups pack
// consecutive leaves greedily (up to n_group particles) inside each level-2
// node, so a 64-lane wavefront holds ~57 compact i-particles on average
// instead of the ~25 of one octree cell.  This file is compiled with
// -ffp-contract=off.
//
// Every count the build finds (nodes per level, groups) stays on the device
// (Tree::info): the host sizes launches and buffers by capacities it chose
// from earlier builds, and kernels past a count exit.  A build whose counts
// exceed the capacities sets an error bit and leaves no groups, so nothing
// downstream runs on it; the host reads the bits when it waits anyway, grows
// the capacities and builds again.  No step of a build waits for the host,
// so it can be recorded in a HIP graph.
#include "internal.hpp"

namespace sphgpu {
namespace {

constexpr int kChunk = 4096;       // bbox partial-reduction chunk
constexpr int kListSlots = 8192;      // most persistent list-builder workgroups (32 waves per CU)
// default: 16 waves per CU.  A walk runs beside force kernels (second stream,
// or the same queue of a recorded step); with all 8192 it took the slots the
// density passes needed (1M step, recorded: FP64 7.33 -> 7.22 ms, mixed 5.35
// -> 5.24 ms; 2048 waves: 7.40 / 5.56 ms)
constexpr int kListWavesDefault = 4096;
constexpr int64_t kFrontier0 = 4096;  // initial per-workgroup BFS frontier capacity (doubled on overflow)

__device__ __forceinline__ uint64_t expand_bits(uint64_t v) {
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

__device__ __forceinline__ uint64_t quantise(double x, double lo, double scale, uint64_t qmax) {
    const double f = (x - lo) * scale;
    uint64_t q = (f > 0) ? (uint64_t) f : 0;
    return q > qmax ? qmax : q;
}

// ---- per-system bounding cube
__global__ void k_bbox_partial(const int32_t *sys_start, int nchunk, const double *x, const double *y,
                               const double *z, double *part) {
    const int s = blockIdx.y, c = blockIdx.x;
    const int b = sys_start[s] + c * kChunk;
    const int e = min(sys_start[s + 1], b + kChunk);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = b + threadIdx.x; p < e; p += blockDim.x) {
        const double v[3] = {x[p], y[p], z[p]};
        for (int k = 0; k < 3; ++k) { lo[k] = dmin(lo[k], v[k]); hi[k] = dmax(hi[k], v[k]); }
    }
    __shared__ double sh[6][256];
    for (int k = 0; k < 3; ++k) { sh[k][threadIdx.x] = lo[k]; sh[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) {
            for (int k = 0; k < 3; ++k) {
                sh[k][threadIdx.x] = dmin(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
                sh[3 + k][threadIdx.x] = dmax(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[((int64_t) s * nchunk + c) * 6 + k] = sh[k][0];
}

__global__ void k_bbox_final(int nchunk, int bits, const double *part, double *lo_out, double *scale_out) {
    const int s = blockIdx.x;
    __shared__ double sh[6][256];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = threadIdx.x; c < nchunk; c += blockDim.x) {
        const double *q = part + ((int64_t) s * nchunk + c) * 6;
        for (int k = 0; k < 3; ++k) { v[k] = dmin(v[k], q[k]); v[3 + k] = dmax(v[3 + k], q[3 + k]); }
    }
    for (int k = 0; k < 6; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) {
            for (int k = 0; k < 3; ++k) {
                sh[k][threadIdx.x] = dmin(sh[k][threadIdx.x], sh[k][threadIdx.x + o]);
                sh[3 + k][threadIdx.x] = dmax(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + o]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        double sd = 0;
        for (int k = 0; k < 3; ++k) sd = dmax(sd, sh[3 + k][0] - sh[k][0]);
        const double side = (sd > 0) ? sd : 1.0;
        for (int k = 0; k < 3; ++k) lo_out[3 * s + k] = sh[k][0];
        scale_out[s] = (double) (1ULL << bits) / side;
    }
}

__global__ void k_keys(int64_t n, const double *x, const double *y, const double *z, const int32_t *sys, int bits,
                       const double *lo, const double *scale, uint64_t *key) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const int s = sys[p];
    const uint64_t qmax = (1ULL << bits) - 1;
    const double sc = scale[s];
    const uint64_t ix = quantise(x[p], lo[3 * s + 0], sc, qmax);
    const uint64_t iy = quantise(y[p], lo[3 * s + 1], sc, qmax);
    const uint64_t iz = quantise(z[p], lo[3 * s + 2], sc, qmax);
    key[p] = ((uint64_t) s << (3 * bits)) | (expand_bits(ix) << 2) | (expand_bits(iy) << 1) | expand_bits(iz);
}

struct PermArgs {
    const int32_t *perm;
    const double *src[13];
    double *dst[13];
    const int32_t *srci[2];
    int32_t *dsti[2];
    const int64_t *srcl;
    int64_t *dstl;
};

__global__ void k_permute(int64_t n, PermArgs a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int32_t q = a.perm[k];
#pragma unroll
    for (int f = 0; f < 13; ++f) a.dst[f][k] = a.src[f][q];
    a.dsti[0][k] = a.srci[0][q];
    a.dsti[1][k] = a.srci[1][q];
    a.dstl[k] = a.srcl[q];
}

__device__ __forceinline__ int64_t lb_u64(const uint64_t *a, int64_t n, uint64_t v) {
    int64_t first = 0, len = n;
    while (len > 0) {
        const int64_t half = len >> 1;
        if (a[first + half] < v) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}
__device__ __forceinline__ int64_t ub_u64(const uint64_t *a, int64_t n, uint64_t v) {
    int64_t first = 0, len = n;
    while (len > 0) {
        const int64_t half = len >> 1;
        if (!(v < a[first + half])) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}
__device__ __forceinline__ int lb_i32(const int32_t *a, int n, int32_t v) {
    int first = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (a[first + half] < v) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

// count of the smallest octree cell holding particles k-1 and k
__global__ void k_lca(int64_t n, const uint64_t *key, int bits, int sysbits, int64_t *cnt) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (k == 0) { cnt[0] = INT64_MAX; return; }
    const uint64_t a = key[k - 1], b = key[k];
    if ((a >> (3 * bits)) != (b >> (3 * bits))) { cnt[k] = INT64_MAX; return; }
    const uint64_t xo = a ^ b;
    const int p = (xo == 0) ? 63 : (__clzll((long long) xo) - 1);
    const int unused = 63 - sysbits - 3 * bits;
    int lev = (p - unused - sysbits) / 3;
    if (lev > bits) lev = bits;
    const int shift = 3 * (bits - lev);
    const uint64_t mask = (shift >= 64) ? ~0ULL : ((1ULL << shift) - 1);
    const uint64_t clo = a & ~mask, chi = clo | mask;
    cnt[k] = ub_u64(key, n, chi) - lb_u64(key, n, clo);
}

// Node starts of every level in one selection.  Particle k starts a node of
// level l when cnt[k] > T_l, T_l = n_leaf 8^l below the top level and
// INT64_MAX - 1 at the top level L-1 (only the starts of the systems pass).
// The thresholds grow with l, so the levels are nested: k starts a node of
// every level up to its start level, and of none when that is -1.
__device__ __forceinline__ int start_level(int64_t c, int L, int64_t n_leaf) {
    if (c == INT64_MAX) return L - 1;
    int lv = -1;
    int64_t T = n_leaf;
    for (int l = 0; l < L - 1; ++l) {
        if (c > T) lv = l;
        T *= 8;
    }
    return lv;
}
constexpr int kSelThreads = 256, kSelItems = 8, kSelTile = kSelThreads * kSelItems;

// per block of kSelTile particles: the node starts of each level
__global__ __launch_bounds__(kSelThreads) void k_sel_count(int64_t n, const int64_t *cnt, int L, int64_t n_leaf,
                                                           int32_t *bcount) {
    __shared__ int hist[kMaxLevels];
    const int tid = threadIdx.x;
    if (tid < kMaxLevels) hist[tid] = 0;
    __syncthreads();
    const int64_t base = (int64_t) blockIdx.x * kSelTile + (int64_t) tid * kSelItems;
    for (int i = 0; i < kSelItems; ++i) {
        const int64_t k = base + i;
        if (k < n) {
            const int lv = start_level(cnt[k], L, n_leaf);
            if (lv >= 0) atomicAdd(&hist[lv], 1);
        }
    }
    __syncthreads();
    if (tid < L) {
        int c = 0;
        for (int m = tid; m < L; ++m) c += hist[m];  // starts of level tid: start level >= tid
        bcount[(int64_t) blockIdx.x * kMaxLevels + tid] = c;
    }
}

// exclusive scan of the block counts of every level (one workgroup); the
// level totals to info[kInfoRawCnt + l]
__global__ __launch_bounds__(kSelThreads) void k_sel_scan(int nb, int L, int32_t *bcount, int32_t *info) {
    __shared__ int buf[kSelThreads];
    const int tid = threadIdx.x;
    for (int l = 0; l < L; ++l) {
        int carry = 0;
        for (int c0 = 0; c0 < nb; c0 += kSelThreads) {
            const int b = c0 + tid;
            const int v = (b < nb) ? bcount[(int64_t) b * kMaxLevels + l] : 0;
            buf[tid] = v;
            __syncthreads();
            for (int o = 1; o < kSelThreads; o <<= 1) {
                const int t = (tid >= o) ? buf[tid - o] : 0;
                __syncthreads();
                buf[tid] += t;
                __syncthreads();
            }
            if (b < nb) bcount[(int64_t) b * kMaxLevels + l] = carry + buf[tid] - v;
            carry += buf[kSelThreads - 1];
            __syncthreads();
        }
        if (tid == 0) info[kInfoRawCnt + l] = carry;
    }
}

// writes the node starts of level l, in particle order, to out[l n + ...]
__global__ __launch_bounds__(kSelThreads) void k_sel_scatter(int64_t n, const int64_t *cnt, int L, int64_t n_leaf,
                                                             const int32_t *boff, int32_t *out) {
    __shared__ int wsum[kSelThreads / kWave];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), w = tid / kWave;
    const int64_t base = (int64_t) blockIdx.x * kSelTile + (int64_t) tid * kSelItems;
    int lv[kSelItems];
#pragma unroll
    for (int i = 0; i < kSelItems; ++i) lv[i] = (base + i < n) ? start_level(cnt[base + i], L, n_leaf) : -1;
    for (int l = 0; l < L; ++l) {
        int c = 0;
#pragma unroll
        for (int i = 0; i < kSelItems; ++i) c += (lv[i] >= l) ? 1 : 0;
        int incl = c;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += t;
        }
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int pos = boff[(int64_t) blockIdx.x * kMaxLevels + l] + incl - c;
        for (int v = 0; v < w; ++v) pos += wsum[v];
        int32_t *o = out + (int64_t) l * n;
#pragma unroll
        for (int i = 0; i < kSelItems; ++i)
            if (lv[i] >= l) o[pos++] = (int32_t) (base + i);
        __syncthreads();
    }
}

// node capacity of each level, passed by value
struct LevelCaps {
    int32_t cap[kMaxLevels];
};

// Level layout from the counts found: when every level fits its capacity, the
// counts and their exclusive prefix (the flattened node array's level
// offsets); otherwise every count 0 and error bit 1.
__global__ void k_level_layout(int L, LevelCaps caps, int32_t *info) {
    if (threadIdx.x != 0) return;
    bool ok = true;
    for (int l = 0; l < L; ++l) ok = ok && info[kInfoRawCnt + l] <= caps.cap[l];
    int32_t off = 0;
    for (int l = 0; l < L; ++l) {
        const int32_t c = ok ? info[kInfoRawCnt + l] : 0;
        info[kInfoOff + l] = off;
        info[kInfoCnt + l] = c;
        off += c;
    }
    info[kInfoOff + L] = off;
    info[kInfoErr] = ok ? 0 : 1;
    info[kInfoG] = 0;
}

// node starts of every level into the flattened node array (level = blockIdx.y)
__global__ void k_copy_levels(int64_t n, const int32_t *info, const int32_t *sel, int32_t *nstart) {
    const int l = blockIdx.y;
    const int64_t cnt_l = info[kInfoCnt + l], off_l = info[kInfoOff + l];
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < cnt_l; q += (int64_t) gridDim.x * blockDim.x)
        nstart[off_l + q] = sel[(int64_t) l * n + q];
}

// node ranges and children of every level
__global__ void k_node_ranges(int64_t n, const int32_t *info, int32_t *nstart, int32_t *nend, int32_t *c0,
                              int32_t *c1) {
    const int l = blockIdx.y;
    const int64_t cnt_l = info[kInfoCnt + l], off_l = info[kInfoOff + l];
    for (int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; q < cnt_l; q += (int64_t) gridDim.x * blockDim.x) {
        const int64_t id = off_l + q;
        const int32_t st = nstart[id];
        const int32_t en = (q + 1 < cnt_l) ? nstart[id + 1] : (int32_t) n;
        nend[id] = en;
        if (l == 0) {
            c0[id] = -1;
            c1[id] = -1;
        } else {
            const int64_t off_below = info[kInfoOff + l - 1];
            const int32_t *below = nstart + off_below;
            const int cnt_below = info[kInfoCnt + l - 1];
            c0[id] = (int32_t) (off_below + lb_i32(below, cnt_below, st));
            c1[id] = (int32_t) (off_below + lb_i32(below, cnt_below, en));
        }
    }
}

struct MomArgs {
    const int32_t *nstart, *nend, *c0, *c1;
    double *mass, *mx, *my, *mz, *comx, *comy, *comz, *lox, *loy, *loz, *hix, *hiy, *hiz, *maxH;
    double4 *rec;  // (com, mass) records
};

__global__ void k_moments_leaf(const int32_t *info, MomArgs a, const double *x, const double *y, const double *z,
                               const double *m, const double *h) {
    const int64_t id = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= info[kInfoCnt]) return;
    double M = 0, sx = 0, sy = 0, sz = 0, mh = 0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int p = a.nstart[id]; p < a.nend[id]; ++p) {
        const double mp = m[p];
        const double c[3] = {x[p], y[p], z[p]};
        M += mp;
        sx += mp * c[0];
        sy += mp * c[1];
        sz += mp * c[2];
        for (int k = 0; k < 3; ++k) { lo[k] = dmin(lo[k], c[k]); hi[k] = dmax(hi[k], c[k]); }
        mh = dmax(mh, kSupport * h[p]);
    }
    a.mass[id] = M;
    a.mx[id] = sx; a.my[id] = sy; a.mz[id] = sz;
    a.comx[id] = sx / M; a.comy[id] = sy / M; a.comz[id] = sz / M;
    a.rec[id] = make_double4(sx / M, sy / M, sz / M, M);
    a.lox[id] = lo[0]; a.loy[id] = lo[1]; a.loz[id] = lo[2];
    a.hix[id] = hi[0]; a.hiy[id] = hi[1]; a.hiz[id] = hi[2];
    a.maxH[id] = mh;
}

__global__ void k_moments_level(int l, const int32_t *info, MomArgs a) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= info[kInfoCnt + l]) return;
    const int64_t id = info[kInfoOff + l] + q;
    double M = 0, sx = 0, sy = 0, sz = 0, mh = 0;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int c = a.c0[id]; c < a.c1[id]; ++c) {
        M += a.mass[c];
        sx += a.mx[c];
        sy += a.my[c];
        sz += a.mz[c];
        lo[0] = dmin(lo[0], a.lox[c]); lo[1] = dmin(lo[1], a.loy[c]); lo[2] = dmin(lo[2], a.loz[c]);
        hi[0] = dmax(hi[0], a.hix[c]); hi[1] = dmax(hi[1], a.hiy[c]); hi[2] = dmax(hi[2], a.hiz[c]);
        mh = dmax(mh, a.maxH[c]);
    }
    a.mass[id] = M;
    a.mx[id] = sx; a.my[id] = sy; a.mz[id] = sz;
    a.comx[id] = sx / M; a.comy[id] = sy / M; a.comz[id] = sz / M;
    a.rec[id] = make_double4(sx / M, sy / M, sz / M, M);
    a.lox[id] = lo[0]; a.loy[id] = lo[1]; a.loz[id] = lo[2];
    a.hix[id] = hi[0]; a.hiy[id] = hi[1]; a.hiz[id] = hi[2];
    a.maxH[id] = mh;
}

__global__ void k_maxH_leaf(const int32_t *info, const int32_t *nstart, const int32_t *nend, const double *h,
                            double *maxH) {
    const int64_t id = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= info[kInfoCnt]) return;
    double mh = 0;
    for (int p = nstart[id]; p < nend[id]; ++p) mh = dmax(mh, kSupport * h[p]);
    maxH[id] = mh;
}

__global__ void k_maxH_level(int l, const int32_t *info, const int32_t *c0, const int32_t *c1, double *maxH) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= info[kInfoCnt + l]) return;
    const int64_t id = info[kInfoOff + l] + q;
    double mh = 0;
    for (int c = c0[id]; c < c1[id]; ++c) mh = dmax(mh, maxH[c]);
    maxH[id] = mh;
}

__global__ void k_sys_root(int l, const int32_t *info, const int32_t *nstart, const int32_t *sys, int32_t *sys_root) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= info[kInfoCnt + l]) return;
    const int64_t id = info[kInfoOff + l] + q;
    sys_root[sys[nstart[id]]] = (int32_t) id;
}

// i-groups: inside node q of the packing level, consecutive leaves are packed
// greedily while a group stays within n_group particles.  The node's leaves
// are found by descending `depth` levels along its first and its last child
// (every node has at least one child; leaves are the nodes of level 0).
template <bool WRITE>
__device__ __forceinline__ int pack_node(const int32_t *nstart, const int32_t *nend, const int32_t *c0,
                                         const int32_t *c1, int node, int depth, int n_group, int32_t *out) {
    int a = node, last = node;
    for (int l = 0; l < depth; ++l) {
        a = c0[a];
        last = c1[last] - 1;
    }
    const int b = last + 1;
    int run = 0, ng = 0;
#pragma unroll 8
    for (int c = a; c < b; ++c) {
        const int sz = nend[c] - nstart[c];
        if (c == a || run + sz > n_group) {
            if (WRITE) out[ng] = nstart[c];
            ++ng;
            run = 0;
        }
        run += sz;
    }
    return ng;
}

// groups of each node of level l
__global__ void k_group_count(int l, const int32_t *info, const int32_t *nstart, const int32_t *nend,
                              const int32_t *c0, const int32_t *c1, int n_group, int32_t *gcnt) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= info[kInfoCnt + l]) return;
    gcnt[q] = pack_node<false>(nstart, nend, c0, c1, (int) (info[kInfoOff + l] + q), l, n_group, nullptr);
}

// exclusive scan of the per-node group counts in place (one workgroup); the
// total to info[kInfoRawG]
constexpr int kScanThreads = 1024;
__global__ __launch_bounds__(kScanThreads) void k_group_scan(int l, int32_t *info, int32_t *gcnt) {
    __shared__ int wsum[kScanThreads / kWave];
    const int tid = threadIdx.x, w = tid / kWave, lane = tid % kWave;
    const int cnt = info[kInfoCnt + l];
    int carry = 0;
    for (int c0 = 0; c0 < cnt; c0 += kScanThreads) {
        const int q = c0 + tid;
        const int v = (q < cnt) ? gcnt[q] : 0;
        int incl = v;
#pragma unroll
        for (int o = 1; o < kWave; o <<= 1) {
            const int t = __shfl_up(incl, o, kWave);
            if (lane >= o) incl += t;
        }
        if (lane == kWave - 1) wsum[w] = incl;
        __syncthreads();
        int pre = carry, chunk = 0;
        for (int u = 0; u < kScanThreads / kWave; ++u) {
            pre += (u < w) ? wsum[u] : 0;
            chunk += wsum[u];
        }
        if (q < cnt) gcnt[q] = pre + incl - v;
        carry += chunk;
        __syncthreads();
    }
    if (tid == 0) info[kInfoRawG] = carry;
}

// the group count (0 and error bit 2 beyond the capacity; 0 after a node
// overflow) and the end of the last group
__global__ void k_group_layout(int32_t *info, int64_t gcap, int64_t n, int32_t *gstart) {
    if (threadIdx.x != 0) return;
    const int32_t raw = info[kInfoRawG];
    if (raw > gcap) info[kInfoErr] |= 2;
    const int32_t G = (info[kInfoErr] == 0) ? raw : 0;
    info[kInfoG] = G;
    gstart[G] = (int32_t) n;
}

// each node's group starts at its offset
__global__ void k_group_write(int l, const int32_t *info, const int32_t *nstart, const int32_t *nend,
                              const int32_t *c0, const int32_t *c1, int n_group, const int32_t *goff,
                              int32_t *gstart) {
    const int64_t q = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= info[kInfoCnt + l] || info[kInfoG] == 0) return;
    pack_node<true>(nstart, nend, c0, c1, (int) (info[kInfoOff + l] + q), l, n_group, gstart + goff[q]);
}

__global__ void k_fill_i32(int64_t n, int32_t *a, int32_t v) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = v;
}
__global__ void k_iota(int64_t n, int32_t *a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = (int32_t) k;
}

// ------------------------------------------------------------------ lists
struct ListArgs {
    const int32_t *gcount;  // device: the group count of the tree build
    const int32_t *gstart, *sys;
    const double *x, *y, *z, *h;
    int nlevels;
    const int32_t *sys_root, *nstart, *nend, *c0, *c1;
    const double *lox, *loy, *loz, *hix, *hiy, *hiz, *comx, *comy, *comz, *maxH;
    int32_t *out, *cnt;
    int64_t cap;
    int32_t *sout, *scnt;
    int64_t scap;
    int32_t *frontier;
    int64_t fcap;
    int flds;  // frontier entries in LDS (<= kFrontLds)
    double theta2;
    double hfac;    // gather search radius = 2.5 h hfac (hfac > 1 makes a reusable superset list)
    int32_t *queue; // 8 per-XCD group counters, zero at launch
    double *sbox;   // optional, per group: the search box the list was built for (lo[3], hi[3])
    int32_t *err;   // overflow bits: 1 list, 2 monopole list, 4 frontier
    const int32_t *need;  // optional device flag: build only when it is nonzero
    int32_t *builds;      // optional device counter of the constructions done
};

// Decides on the device whether this launch builds at all (the flag is read by
// every wave, so no host round trip decides it); counts the constructions.
__device__ __forceinline__ bool list_launch_wanted(const ListArgs &a) {
    if (a.need && *a.need == 0) return false;
    if (a.builds && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(a.builds, 1);
    return true;
}

__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, kWave));
    return v;
}

struct GBox {
    double slo[3], shi[3], glo[3], ghi[3];
    double r2;  // gather lists: (largest search radius)^2 (1 + 1e-12), the reach beyond the group's box
};

// (a group of more than 64 particles is read 64 at a time)
__device__ __forceinline__ void group_box(const ListArgs &a, int i0, int ni, GBox &b) {
    const int lane = threadIdx.x;
    double glo[3] = {INFINITY, INFINITY, INFINITY}, ghi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double slo[3] = {INFINITY, INFINITY, INFINITY}, shi[3] = {-INFINITY, -INFINITY, -INFINITY};
    double R = 0.0;
    for (int base = 0; base < ni; base += kWave) {
        if (base + lane >= ni) continue;
        const int p = i0 + base + lane;
        const double c[3] = {a.x[p], a.y[p], a.z[p]};
        const double H = a.h ? kSupport * a.h[p] * a.hfac : 0.0;  // the gravity walk reads no h
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            glo[k] = dmin(glo[k], c[k]);
            ghi[k] = dmax(ghi[k], c[k]);
            slo[k] = dmin(slo[k], c[k] - H);
            shi[k] = dmax(shi[k], c[k] + H);
        }
        R = dmax(R, H);
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        b.glo[k] = wave_min(glo[k]);
        b.ghi[k] = wave_max(ghi[k]);
        b.slo[k] = wave_min(slo[k]);
        b.shi[k] = wave_max(shi[k]);
    }
    R = wave_max(R);
    b.r2 = (R * R) * (1.0 + 1.0e-12);
}

// squared distance from a point to the group's box (0 inside)
__device__ __forceinline__ double box_dist2(const GBox &b, const double c[3]) {
    double d2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double d = (c[k] < b.glo[k]) ? (b.glo[k] - c[k]) : ((c[k] > b.ghi[k]) ? (c[k] - b.ghi[k]) : 0.0);
        d2 = d2 + d * d;
    }
    return d2;
}

template <int KIND>
__device__ __forceinline__ bool node_test(const ListArgs &a, int id, const GBox &b) {
    const double lo[3] = {a.lox[id], a.loy[id], a.loz[id]}, hi[3] = {a.hix[id], a.hiy[id], a.hiz[id]};
    bool ov = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ov = ov & (lo[k] <= b.shi[k]) & (hi[k] >= b.slo[k]);
    if (KIND == 0) {  // and within the search radius of the group's box
        double d2 = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            const double d = (lo[k] > b.ghi[k]) ? (lo[k] - b.ghi[k]) : ((hi[k] < b.glo[k]) ? (b.glo[k] - hi[k]) : 0.0);
            d2 = d2 + d * d;
        }
        return ov && d2 <= b.r2;
    }
    if (ov) return ov;
    const double mh = a.maxH[id];
    bool ov2 = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ov2 = ov2 & (lo[k] - mh <= b.ghi[k]) & (hi[k] + mh >= b.glo[k]);
    return ov2;
}

template <int KIND>
__device__ __forceinline__ bool part_test(const ListArgs &a, int p, const GBox &b) {
    const double c[3] = {a.x[p], a.y[p], a.z[p]};
    bool in = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) in = in & (b.slo[k] <= c[k]) & (c[k] <= b.shi[k]);
    if (KIND == 0) return in && box_dist2(b, c) <= b.r2;
    if (in) return in;
    const double H = kSupport * a.h[p];
    bool ov = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ov = ov & (c[k] - H <= b.ghi[k]) & (c[k] + H >= b.glo[k]);
    return ov;
}

// monopole acceptance (see the CPU oracle for the same rule)
__device__ __forceinline__ bool accept_node(const ListArgs &a, int id, const GBox &b) {
    const double lo[3] = {a.lox[id], a.loy[id], a.loz[id]}, hi[3] = {a.hix[id], a.hiy[id], a.hiz[id]};
    bool ov = true;
#pragma unroll
    for (int k = 0; k < 3; ++k) ov = ov & (lo[k] <= b.ghi[k]) & (hi[k] >= b.glo[k]);
    if (ov) return false;
    const double c[3] = {a.comx[id], a.comy[id], a.comz[id]};
    double d[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) d[k] = (c[k] < b.glo[k]) ? (b.glo[k] - c[k]) : ((c[k] > b.ghi[k]) ? (c[k] - b.ghi[k]) : 0.0);
    const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double sz = hi[0] - lo[0];
    sz = dmax(sz, hi[1] - lo[1]);
    sz = dmax(sz, hi[2] - lo[2]);
    return (sz * sz) < (a.theta2 * d2);
}

// The groups are split into 8 contiguous ranges, one per XCD (workgroup b
// runs on XCD b mod 8), so that neighbouring groups -- which walk the same
// nodes and read the same particles -- share an L2.  Inside its range each
// wave takes the next group from the XCD's counter: walk costs vary a lot
// between groups, and a static split leaves a long tail of busy waves.
struct GroupQueue {
    int beg, end;
    int32_t *ctr;
    __device__ __forceinline__ int next() const {
        int k = 0;
        if (threadIdx.x == 0) k = atomicAdd(ctr, 1);
        k = __shfl(k, 0, kWave);
        return beg + k;
    }
};
__device__ __forceinline__ GroupQueue xcd_queue(int ngroups, int32_t *ctr8) {
    const int x = blockIdx.x & 7;
    const int q = ngroups >> 3, r = ngroups & 7;
    GroupQueue Q;
    Q.beg = x * q + (x < r ? x : r);
    Q.end = Q.beg + q + (x < r ? 1 : 0);
    Q.ctr = ctr8 + x;
    return Q;
}

// A BFS frontier of one wave: its first n (<= kFrontLds) entries live in LDS,
// the rest in the workgroup's slice of global scratch (fcap entries).  Most
// frontiers fit in LDS, so a level's node ids are neither written back to L2
// nor fetched from it before the node data they index can be loaded.
// The two parts carry their address spaces in their types: through generic
// pointers the compiler selects one of them per lane and issues flat loads and
// stores, which wait on both the LDS and the vector-memory counters.
typedef __attribute__((address_space(3))) int32_t lds_i32;
typedef __attribute__((address_space(1))) int32_t glb_i32;
struct Front {
    lds_i32 *s;  // LDS part
    glb_i32 *g;  // global part
    int n;       // entries in the LDS part
    __device__ __forceinline__ int32_t get(int i) const { return i < n ? s[i] : g[i - n]; }
    __device__ __forceinline__ void put(int i, int32_t v) const {
        if (i < n) s[i] = v;
        else g[i - n] = v;
    }
};
__device__ __forceinline__ int64_t front_cap(const ListArgs &a) { return a.flds + a.fcap; }

// expand the frontier of one level: children of the nodes that pass `test`
template <class Test>
__device__ __forceinline__ int expand_level(const ListArgs &a, const Front &fa, int nf, const Front &fb, Test test) {
    const int lane = threadIdx.x;
    const int64_t fcap = front_cap(a);
    int nn = 0;
    for (int base = 0; base < nf; base += kWave) {
        const int idx = base + lane;
        int nc = 0, cfirst = 0;
        if (idx < nf) {
            const int node = fa.get(idx);
            const int32_t c0 = a.c0[node], c1 = a.c1[node];  // loaded beside the node's box, not after its test
            if (test(node)) { cfirst = c0; nc = c1 - c0; }
        }
        const int incl = wave_incl_scan(nc);
        const int tot = __shfl(incl, kWave - 1, kWave);
        const int pos0 = nn + incl - nc;
        for (int c = 0; c < nc; ++c)
            if (pos0 + c < fcap) fb.put(pos0 + c, cfirst + c);
        nn += tot;
    }
    if (nn > fcap) {
        if (lane == 0) atomicOr(a.err, 4);
        nn = (int) fcap;
    }
    __syncthreads();
    return nn;
}

constexpr int kEmitUnroll = 4;

// emit the particles of one chunk of leaves (one per lane: particles first ..
// first + sz - 1, sz = 0 for none) that pass part_ok, in leaf order then
// index order
template <class PartOk>
__device__ __forceinline__ int emit_chunk(const ListArgs &a, int first, int sz, int32_t *out, int64_t cap, int count,
                                          PartOk part_ok, int *s_pref, int *s_first) {
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ULL : ((~0ULL) >> (64 - lane));
    const int incl = wave_incl_scan(sz);
    const int tot = __shfl(incl, kWave - 1, kWave);
    s_pref[lane] = incl - sz;
    s_first[lane] = first;
    if (lane == kWave - 1) s_pref[kWave] = incl;
    __syncthreads();
    // kEmitUnroll rounds of 64 slots at a time: their particle tests' loads
    // are in flight together
    for (int s0 = 0; s0 < tot; s0 += kEmitUnroll * kWave) {
        int p[kEmitUnroll];
        bool pass[kEmitUnroll];
#pragma unroll
        for (int u = 0; u < kEmitUnroll; ++u) {
            const int slot = s0 + u * kWave + lane;
            p[u] = -1;
            if (slot < tot) {
                int lo = 0, hi = kWave;
                while (hi - lo > 1) {
                    const int mid = (lo + hi) >> 1;
                    if (s_pref[mid] <= slot) lo = mid;
                    else hi = mid;
                }
                p[u] = s_first[lo] + (slot - s_pref[lo]);
            }
        }
#pragma unroll
        for (int u = 0; u < kEmitUnroll; ++u) pass[u] = (p[u] >= 0) && part_ok(p[u]);
#pragma unroll
        for (int u = 0; u < kEmitUnroll; ++u) {
            const uint64_t m = __ballot(pass[u]);
            if (pass[u]) {
                const int pos = count + __popcll(m & lt);
                if (pos < cap) out[pos] = p[u];
            }
            count += __popcll(m);
        }
    }
    __syncthreads();
    return count;
}

// emit the particles of the leaves fa[0..nf) that pass leaf_ok, filtered by
// part_ok, in leaf order then index order
template <class LeafOk, class PartOk>
__device__ __forceinline__ int emit_particles(const ListArgs &a, const Front &fa, int nf, int32_t *out, int64_t cap,
                                              int count, LeafOk leaf_ok, PartOk part_ok, int *s_pref, int *s_first) {
    const int lane = threadIdx.x;
    for (int base = 0; base < nf; base += kWave) {
        const int idx = base + lane;
        int sz = 0, first = 0;
        if (idx < nf) {
            const int leaf = fa.get(idx);
            const int32_t st = a.nstart[leaf], en = a.nend[leaf];  // loaded beside the leaf's box
            if (leaf_ok(leaf)) { first = st; sz = en - st; }
        }
        count = emit_chunk(a, first, sz, out, cap, count, part_ok, s_pref, s_first);
    }
    return count;
}

template <int KIND>
__global__ __launch_bounds__(kWave) void k_list_sph(ListArgs a) {
    if (!list_launch_wanted(a)) return;
    const int lane = threadIdx.x;
    int32_t *fa0 = a.frontier + (int64_t) blockIdx.x * 2 * a.fcap;
    int32_t *fb0 = fa0 + a.fcap;
    __shared__ int s_pref[kWave + 1];
    __shared__ int s_first[kWave];
    __shared__ int32_t s_front[2 * kFrontLds];
    const GroupQueue Q = xcd_queue(*a.gcount, a.queue);
    for (int g = Q.next(); g < Q.end; g = Q.next()) {
        const int i0 = a.gstart[g], ni = a.gstart[g + 1] - i0;
        GBox b;
        group_box(a, i0, ni, b);
        Front fa{(lds_i32 *) s_front, (glb_i32 *) fa0, a.flds};
        Front fb{(lds_i32 *) (s_front + kFrontLds), (glb_i32 *) fb0, a.flds};
        if (lane == 0) fa.put(0, a.sys_root[a.sys[i0]]);
        if (a.sbox && lane < 6) {  // selects, not an index by lane: the box stays in registers
            const double v = (lane == 0) ? b.slo[0] : (lane == 1) ? b.slo[1] : (lane == 2) ? b.slo[2]
                           : (lane == 3) ? b.shi[0] : (lane == 4) ? b.shi[1] : b.shi[2];
            a.sbox[(int64_t) g * 6 + lane] = v;
        }
        __syncthreads();
        int nf = 1;
        auto test = [&](int node) { return node_test<KIND>(a, node, b); };
        for (int l = a.nlevels - 1; l >= 1; --l) {
            nf = expand_level(a, fa, nf, fb, test);
            const Front t = fa; fa = fb; fb = t;
        }
        int32_t *out = a.out + (int64_t) g * a.cap;
        auto ptest = [&](int p) { return part_test<KIND>(a, p, b); };
        const int count = emit_particles(a, fa, nf, out, a.cap, 0, test, ptest, s_pref, s_first);
        if (lane == 0) {
            a.cnt[g] = (int32_t) min((int64_t) count, a.cap);  // an overflowed list stays an in-bounds prefix
            if (count > a.cap) atomicOr(a.err, 1);
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kWave) void k_list_grav(ListArgs a) {
    if (!list_launch_wanted(a)) return;
    const int lane = threadIdx.x;
    const uint64_t lt = (lane == 0) ? 0ULL : ((~0ULL) >> (64 - lane));
    int32_t *fa0 = a.frontier + (int64_t) blockIdx.x * 2 * a.fcap;
    int32_t *fb0 = fa0 + a.fcap;
    __shared__ int s_pref[kWave + 1];
    __shared__ int s_first[kWave];
    __shared__ int32_t s_front[2 * kFrontLds];
    const int64_t fcap = front_cap(a);
    const GroupQueue Q = xcd_queue(*a.gcount, a.queue);
    for (int g = Q.next(); g < Q.end; g = Q.next()) {
        const int i0 = a.gstart[g], ni = a.gstart[g + 1] - i0;
        GBox b;
        group_box(a, i0, ni, b);
        Front fa{(lds_i32 *) s_front, (glb_i32 *) fa0, a.flds};
        Front fb{(lds_i32 *) (s_front + kFrontLds), (glb_i32 *) fb0, a.flds};
        if (lane == 0) fa.put(0, a.sys_root[a.sys[i0]]);
        __syncthreads();
        int nf = 1, cj = 0, cs = 0;
        int32_t *jout = a.out + (int64_t) g * a.cap;
        int32_t *sout = a.sout + (int64_t) g * a.scap;
        auto all = [&](int) { return true; };
        for (int l = a.nlevels - 1; l >= 0; --l) {
            // one pass over the frontier: the accepted nodes become monopoles
            // (in frontier order), the others are opened -- their children
            // form the next frontier, or at the leaf level their particles
            // are emitted
            int nn = 0;
            for (int base = 0; base < nf; base += kWave) {
                const int idx = base + lane;
                const int node = (idx < nf) ? fa.get(idx) : -1;
                const bool acc = (node >= 0) && accept_node(a, node, b);
                const uint64_t m = __ballot(acc);
                if (acc) {
                    const int pos = cs + __popcll(m & lt);
                    if (pos < a.scap) sout[pos] = node;
                }
                cs += __popcll(m);
                const bool open = (node >= 0) && !acc;
                if (l >= 1) {
                    int nc = 0, cfirst = 0;
                    if (node >= 0) {  // loaded beside the node's data, not after its test
                        const int32_t c0 = a.c0[node], c1 = a.c1[node];
                        if (open) { cfirst = c0; nc = c1 - c0; }
                    }
                    const int incl = wave_incl_scan(nc);
                    const int pos0 = nn + incl - nc;
                    for (int c = 0; c < nc; ++c)
                        if (pos0 + c < fcap) fb.put(pos0 + c, cfirst + c);
                    nn += __shfl(incl, kWave - 1, kWave);
                } else {
                    int sz = 0, first = 0;
                    if (node >= 0) {
                        const int32_t st = a.nstart[node], en = a.nend[node];
                        if (open) { first = st; sz = en - st; }
                    }
                    cj = emit_chunk(a, first, sz, jout, a.cap, cj, all, s_pref, s_first);
                }
            }
            if (l >= 1) {
                if (nn > fcap) {
                    if (lane == 0) atomicOr(a.err, 4);
                    nn = (int) fcap;
                }
                __syncthreads();
                nf = nn;
                const Front t = fa; fa = fb; fb = t;
            }
        }
        if (lane == 0) {
            a.cnt[g] = (int32_t) min((int64_t) cj, a.cap);  // overflowed lists stay in-bounds prefixes
            a.scnt[g] = (int32_t) min((int64_t) cs, a.scap);
            if (cj > a.cap) atomicOr(a.err, 1);
            if (cs > a.scap) atomicOr(a.err, 2);
        }
        __syncthreads();
    }
}

__device__ __forceinline__ float wave_maxf(float v) { return dpp_wave_maxf(v); }

// Which entries of a tile lie inside the cut of some lane: the dot form of
// mfma_dot_mask_f with the roles of the operands swapped -- rows are the
// group's particles, columns the tile's entries -- so that each lane holds
// one entry's column and the OR over the particles is a minimum over the
// lane's own accumulators (and its partner half's, through one ballot).
// Lane l supplies A[l%32][l/32] (particle 32 ib + l%32: its own when its half
// is ib, else its partner's) and B[l/32][l%32] (entry 32 jb + l%32); register
// r of lane half h holds row 8(r/4) + 4h + r%4 of the lane's column.  K = 6
// in three steps: particle side (-2 o_x, -2 o_y | -2 o_z, 1 | -c, 0), entry
// side (o_x, o_y | o_z, w | 1, 0).  Returns the wave-uniform mask of entries
// < nt with t - c < 0 for some lane (negc = -c; +inf admits nothing).
__device__ __forceinline__ uint64_t mfma_dot_any_(const float4 *q, int nt, float tx, float ty, float tz, float negc,
                                                  float px, float py, float pz, float pc) {
    const int lane = threadIdx.x, h = lane >> 5, m = lane & 31;
    uint64_t any = 0;
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const float4 o = q[32 * jb + m];
        const float b0 = h ? o.y : o.x, b1 = h ? o.w : o.z, b2 = h ? 0.f : 1.f;
        float mn = INFINITY;
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
            const bool own = (ib == h);  // particle 32 ib + m is this lane's own, else its partner's
            const float a0 = h ? (own ? ty : py) : (own ? tx : px);
            const float a1 = h ? 1.f : (own ? tz : pz);
            const float a2 = h ? 0.f : (own ? negc : pc);
            f32x16_ c = f32x16_{};
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, b2, c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) mn = fminf(mn, c[r]);
        }
        const uint64_t b = __ballot(mn < 0.f);
        any |= (uint64_t) ((uint32_t) b | (uint32_t) (b >> 32)) << (32 * jb);
    }
    if (nt < kWave) any &= (1ULL << nt) - 1;
    return any;
}

// Compaction of the gather lists.  A gather list holds every particle inside
// the group's search box, most of it beyond every support (the box's corners
// and the margin on h): the density and derivative kernels stage, mask and
// skip those entries in every pass.  This keeps, in list order, the entries j
// with |x_j - x_i| <= R_g for some i of the group, R_g = hfac max_i 2.5 h_i
// (h as at construction), and records R_g; while every 2.5 h_i <= R_g
// (checked with the box in k_list_box_check) the kept entries include every
// particle inside each i's support, in the same order as before.  The test is
// the matrix-core dot form of the interaction kernels with the lane threshold
// of R_g (dot_cut_: room for the FP32 offsets' rounding), so it admits a
// superset of the FP64 condition; an entry is kept when some lane's test
// admits it (mfma_dot_any_).  In place: chunk c reads entries [64c, 64c + 64) into
// registers before it writes at most as many entries below 64c + 64, and the
// chunks after it start at 64c + 64.
// 8 waves per SIMD: 62 registers, no spills (the compiler's own choice, 106
// registers at 4 waves, measured 212 against 197 us at 1M)
__global__ __launch_bounds__(kWave, 8) void k_list_compact(const int32_t *gcount, const int32_t *gstart,
                                                        const double *x, const double *y, const double *z,
                                                        const double *h, const double4 *pos, double hfac, int32_t *idx,
                                                        int32_t *cnt, int64_t cap, double *crad, const int32_t *need) {
    if (need && *need == 0) return;
    const int G = *gcount;
    if ((int) blockIdx.x >= G) return;
    const int g = xcd_remap((int) blockIdx.x, G);
    const int lane = threadIdx.x;
    const int i0 = gstart[g], ni = gstart[g + 1] - i0;
    const bool act = lane < ni;
    const int p = i0 + (act ? lane : 0);
    const double xi = x[p], yi = y[p], zi = z[p];
    const double R = hfac * wave_max(act ? kSupport * h[p] : 0.0);
    if (lane == 0) crad[g] = R;
    const double cx = 0.5 * (wave_min(act ? xi : INFINITY) + wave_max(act ? xi : -INFINITY));
    const double cy = 0.5 * (wave_min(act ? yi : INFINITY) + wave_max(act ? yi : -INFINITY));
    const double cz = 0.5 * (wave_min(act ? zi : INFINITY) + wave_max(act ? zi : -INFINITY));
    const float oxi = (float) (xi - cx), oyi = (float) (yi - cy), ozi = (float) (zi - cz);
    const float wi = fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)), n1i = norm1_(oxi, oyi, ozi);
    const float Rc = (float) R * 1.0001f;
    const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    int32_t *list = idx + (int64_t) g * cap;
    const int nj = cnt[g];
    const uint64_t lt = (1ULL << lane) - 1;
    __shared__ float4 so[kWave];
    // software-pipelined chunks: the record of chunk c+1 and the list index of
    // chunk c+2 are loaded before chunk c is written back (its writes stay below
    // 64c + 64); a lane past the end loads particle 0
    auto load = [&](int q) { return pos ? pos[q] : make_double4(x[q], y[q], z[q], 0.0); };
    int q = (lane < nj) ? list[lane] : 0;
    double4 r = load(q);
    int qn = (kWave + lane < nj) ? list[kWave + lane] : 0;
    int count = 0;
    for (int base = 0; base < nj; base += kWave) {
        const int k = base + lane;
        float n1 = 0.f;
        if (k < nj) {
            const float ox = (float) (r.x - cx), oy = (float) (r.y - cy), oz = (float) (r.z - cz);
            so[lane] = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            n1 = norm1_(ox, oy, oz);
        }
        const int qk = q;
        if (base + kWave < nj) {
            q = qn;
            r = load(q);
            qn = (base + 2 * kWave + lane < nj) ? list[base + 2 * kWave + lane] : 0;
        }
        const float n1max = wave_maxf(n1);  // every lane takes part: outside the condition
        const float negc = act ? -dot_cut_(Rc, wi, n1i, n1max) : INFINITY;
        const float pnegc = xor32_f(negc);
        __syncthreads();
        const uint64_t kept = mfma_dot_any_(so, min(kWave, nj - base), tx, ty, tz, negc, ptx, pty, ptz, pnegc);
        const bool keep = (k < nj) && ((kept >> lane) & 1);
        const uint64_t b = __ballot(keep);
        if (keep) list[count + __popcll(b & lt)] = qk;
        count += __popcll(b);
        __syncthreads();
    }
    if (lane == 0) cnt[g] = count;
}

// mfma_dot_any_ with a threshold per entry instead of per particle: the last
// K pair is (w_i, 1) on the particle side against (1, -Q_j) on the entry side,
// so each accumulator holds |o_j - o_i|^2 - Q_j (w = |o|^2).  Returns the
// wave-uniform mask of entries < nt with |o_j - o_i|^2 < Q_j for some lane
// (wi = +inf for a lane without a particle admits nothing).  qs[jj] = -Q_j.
__device__ __forceinline__ uint64_t mfma_dot_any_q_(const float4 *q, const float *qs, int nt, float tx, float ty,
                                                    float tz, float wi, float px, float py, float pz, float pw) {
    const int lane = threadIdx.x, h = lane >> 5, m = lane & 31;
    uint64_t any = 0;
#pragma unroll
    for (int jb = 0; jb < 2; ++jb) {
        const float4 o = q[32 * jb + m];
        const float b0 = h ? o.y : o.x, b1 = h ? o.w : o.z, b2 = h ? qs[32 * jb + m] : 1.f;
        float mn = INFINITY;
#pragma unroll
        for (int ib = 0; ib < 2; ++ib) {
            const bool own = (ib == h);  // particle 32 ib + m is this lane's own, else its partner's
            const float a0 = h ? (own ? ty : py) : (own ? tx : px);
            const float a1 = h ? 1.f : (own ? tz : pz);
            const float a2 = h ? 1.f : (own ? wi : pw);
            f32x16_ c = f32x16_{};
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, c, 0, 0, 0);
            c = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, b2, c, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 16; ++r) mn = fminf(mn, c[r]);
        }
        const uint64_t b = __ballot(mn < 0.f);
        any |= (uint64_t) ((uint32_t) b | (uint32_t) (b >> 32)) << (32 * jb);
    }
    if (nt < kWave) any &= (1ULL << nt) - 1;
    return any;
}

// Compaction of the symmetric lists for the hydro force terms.  A symmetric
// list holds every particle whose support or the group's reaches the group's
// tree cell; at 1M about 40 % of it lies beyond both supports of every pair
// with the group (CPU count on the impact), staged, masked and skipped by the
// hydro kernel.  This writes to a second list, in list order, the entries j
// with |x_j - x_i| < max(max_i H_i, H_j) for some i of the group (H = 2.5 h,
// with the hydro kernel's 1e-4 room), a superset of the pairs inside either
// support, so the force sums take the same nonzero terms.  The threshold is
// the entry's (dot_cut_j_: room for the FP32 offsets' rounding, with the
// group's largest |o_i|_1).  Out of place: the dt kernel reads the whole
// list meanwhile, since v_sig_max runs over every entry (src/force.h:122-124).
// Groups of at most 64 particles; pos: packed records or null (x, y, z read).
__global__ __launch_bounds__(kWave, 8) void k_list_compact_sym(const int32_t *gcount, const int32_t *gstart,
                                                            const double *x, const double *y, const double *z,
                                                            const double *h, const double4 *pos,
                                                            const int32_t *idx, const int32_t *cnt, int64_t cap,
                                                            int32_t *oidx, int32_t *ocnt) {
    const int G = *gcount;
    if ((int) blockIdx.x >= G) return;
    const int g = xcd_remap((int) blockIdx.x, G);
    const int lane = threadIdx.x;
    const int i0 = gstart[g], ni = gstart[g + 1] - i0;
    const bool act = lane < ni;
    const int p = i0 + (act ? lane : 0);
    const double xi = x[p], yi = y[p], zi = z[p];
    const float Hmax = wave_maxf(act ? (float) (kSupport * h[p]) * 1.0001f : 0.f);
    const double cx = 0.5 * (wave_min(act ? xi : INFINITY) + wave_max(act ? xi : -INFINITY));
    const double cy = 0.5 * (wave_min(act ? yi : INFINITY) + wave_max(act ? yi : -INFINITY));
    const double cz = 0.5 * (wave_min(act ? zi : INFINITY) + wave_max(act ? zi : -INFINITY));
    const float oxi = (float) (xi - cx), oyi = (float) (yi - cy), ozi = (float) (zi - cz);
    const float wi = act ? fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)) : INFINITY;
    const float n1max = wave_maxf(act ? norm1_(oxi, oyi, ozi) : 0.f);
    const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    const float pwi = __shfl_xor(wi, 32, kWave);
    const int32_t *list = idx + (int64_t) g * cap;
    int32_t *out = oidx + (int64_t) g * cap;
    const int nj = cnt[g];
    const uint64_t lt = (1ULL << lane) - 1;
    __shared__ float4 so[kWave];
    __shared__ float sq[kWave];
    // software-pipelined: the records of chunk c+1 and the list index of chunk
    // c+2 are in flight while chunk c is tested; a lane past the end loads
    // particle 0
    auto load = [&](int q) { return pos ? pos[q] : make_double4(x[q], y[q], z[q], 0.0); };
    int q = (lane < nj) ? list[lane] : 0;
    double4 r = load(q);
    double hq = h[q];
    int qn = (kWave + lane < nj) ? list[kWave + lane] : 0;
    int count = 0;
    for (int base = 0; base < nj; base += kWave) {
        const int k = base + lane;
        if (k < nj) {
            const float ox = (float) (r.x - cx), oy = (float) (r.y - cy), oz = (float) (r.z - cz);
            so[lane] = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            sq[lane] = -dot_cut_j_(fmaxf(Hmax, (float) (kSupport * hq) * 1.0001f), norm1_(ox, oy, oz), n1max);
        }
        const int qk = q;
        if (base + kWave < nj) {
            q = qn;
            r = load(q);
            hq = h[q];
            qn = (base + 2 * kWave + lane < nj) ? list[base + 2 * kWave + lane] : 0;
        }
        __syncthreads();
        const uint64_t kept = mfma_dot_any_q_(so, sq, min(kWave, nj - base), tx, ty, tz, wi, ptx, pty, ptz, pwi);
        const bool keep = (k < nj) && ((kept >> lane) & 1);
        const uint64_t b = __ballot(keep);
        if (keep) out[count + __popcll(b & lt)] = qk;
        count += __popcll(b);
        __syncthreads();
    }
    if (lane == 0) ocnt[g] = count;
}

// The same for groups of more than 64 particles: the group's radius and box
// come from all its particles, read 64 at a time, and an entry is kept when
// some particle of any of those waves admits it.
__global__ __launch_bounds__(kWave) void k_list_compact_multi(const int32_t *gcount, const int32_t *gstart,
                                                              const double *x, const double *y, const double *z,
                                                              const double *h, const double4 *pos, double hfac,
                                                              int32_t *idx, int32_t *cnt, int64_t cap, double *crad,
                                                              const int32_t *need) {
    if (need && *need == 0) return;
    const int G = *gcount;
    if ((int) blockIdx.x >= G) return;
    const int g = xcd_remap((int) blockIdx.x, G);
    const int lane = threadIdx.x;
    const int i0 = gstart[g], ni = gstart[g + 1] - i0;
    double Hm = 0.0, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < ni; b += kWave) {
        if (b + lane >= ni) continue;
        const int p = i0 + b + lane;
        Hm = dmax(Hm, kSupport * h[p]);
        lo[0] = dmin(lo[0], x[p]); hi[0] = dmax(hi[0], x[p]);
        lo[1] = dmin(lo[1], y[p]); hi[1] = dmax(hi[1], y[p]);
        lo[2] = dmin(lo[2], z[p]); hi[2] = dmax(hi[2], z[p]);
    }
    const double R = hfac * wave_max(Hm);
    if (lane == 0) crad[g] = R;
    const double cx = 0.5 * (wave_min(lo[0]) + wave_max(hi[0]));
    const double cy = 0.5 * (wave_min(lo[1]) + wave_max(hi[1]));
    const double cz = 0.5 * (wave_min(lo[2]) + wave_max(hi[2]));
    const float Rc = (float) R * 1.0001f;
    int32_t *list = idx + (int64_t) g * cap;
    const int nj = cnt[g];
    const uint64_t lt = (1ULL << lane) - 1;
    __shared__ float4 so[kWave];
    int count = 0;
    for (int base = 0; base < nj; base += kWave) {
        const int k = base + lane;
        float n1 = 0.f;
        int qk = 0;
        if (k < nj) {
            qk = list[k];
            const double4 r = pos ? pos[qk] : make_double4(x[qk], y[qk], z[qk], 0.0);
            const float ox = (float) (r.x - cx), oy = (float) (r.y - cy), oz = (float) (r.z - cz);
            so[lane] = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            n1 = norm1_(ox, oy, oz);
        }
        const float n1max = wave_maxf(n1);  // every lane takes part: outside the condition
        __syncthreads();
        uint64_t kept = 0;
        for (int b = 0; b < ni; b += kWave) {  // the group's particles, one wave at a time
            const bool act = b + lane < ni;
            const int p = i0 + b + (act ? lane : 0);
            const float oxi = (float) (x[p] - cx), oyi = (float) (y[p] - cy), ozi = (float) (z[p] - cz);
            const float wi = fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)), n1i = norm1_(oxi, oyi, ozi);
            const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
            const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave),
                        ptz = __shfl_xor(tz, 32, kWave);
            const float negc = act ? -dot_cut_(Rc, wi, n1i, n1max) : INFINITY;
            const float pnegc = xor32_f(negc);
            kept |= mfma_dot_any_(so, min(kWave, nj - base), tx, ty, tz, negc, ptx, pty, ptz, pnegc);
        }
        const bool keep = (k < nj) && ((kept >> lane) & 1);
        const uint64_t bk = __ballot(keep);
        __syncthreads();  // every lane has read its entry before any is written over
        if (keep) list[count + __popcll(bk & lt)] = qk;
        count += __popcll(bk);
    }
    if (lane == 0) cnt[g] = count;
}

inline int nblk(int64_t n, int b) { return (int) ((n + b - 1) / b); }

}  // namespace

// ------------------------------------------------------------------ host side
static void compute_sys_boxes(Tree &t, const TreeParams &tp, int64_t n, const PArrays &P, hipStream_t s) {
    (void) n;
    const std::vector<int32_t> &ss = t.sys_start_h;  // system offsets never change after init
    if ((int) ss.size() != tp.nsys + 1) throw Error(-2, "tree: system offsets not set");
    int maxn = 0;
    for (int q = 0; q < tp.nsys; ++q) maxn = std::max(maxn, ss[q + 1] - ss[q]);
    const int nchunk = std::max(1, (maxn + kChunk - 1) / kChunk);
    t.part.ensure((size_t) tp.nsys * nchunk * 6);
    t.sys_lo.ensure(3 * tp.nsys);
    t.sys_scale.ensure(tp.nsys);
    hipLaunchKernelGGL(k_bbox_partial, dim3(nchunk, tp.nsys), dim3(256), 0, s, t.sys_start.get(), nchunk, P.x, P.y,
                       P.z, t.part.get());
    SPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_bbox_final, dim3(tp.nsys), dim3(256), 0, s, nchunk, tp.bits, t.part.get(), t.sys_lo.get(),
                       t.sys_scale.get());
    SPH_LAUNCH_CHECK();
}

// capacity for a count found: 1/16 and 16 more, so that the counts that creep
// up from step to step stay inside it for many steps
static int32_t cap_for(int64_t c) { return (int32_t) std::min<int64_t>(INT32_MAX / 2, c + c / 16 + 16); }

// the device counts into t.hinfo (the stream must then drain before they are read)
static void fetch_info(Tree &t, hipStream_t s) {
    int32_t *h = t.hinfo.ensure(kInfoWords);
    SPH_HIP(hipMemcpyAsync(h, t.info.get(), sizeof(int32_t) * kInfoWords, hipMemcpyDeviceToHost, s));
}

static void build_levels(Tree &t, const TreeParams &tp, int64_t n, const PArrays &P, hipStream_t s) {
    t.n = n;
    const int L = tp.nlevels;
    t.nlevels = L;
    const bool discover = !t.caps_known;
    int32_t *info = t.info.ensure(kInfoWords);
    t.cnt.ensure(n);
    hipLaunchKernelGGL(k_lca, dim3(nblk(n, 256)), dim3(256), 0, s, n, t.key.get(), tp.bits, tp.sysbits, t.cnt.get());
    SPH_LAUNCH_CHECK();
    t.sel_out.ensure((size_t) L * n);
    {  // node starts of all levels: per-block counts, their scan, the scatter
        const int nb = nblk(n, kSelTile);
        int32_t *blk = t.sel_blk.ensure((size_t) std::max(nb, 1) * kMaxLevels);
        if (nb > 0) {
            hipLaunchKernelGGL(k_sel_count, dim3(nb), dim3(kSelThreads), 0, s, n, t.cnt.get(), L,
                               (int64_t) tp.n_leaf, blk);
            SPH_LAUNCH_CHECK();
        }
        hipLaunchKernelGGL(k_sel_scan, dim3(1), dim3(kSelThreads), 0, s, nb, L, blk, info);
        SPH_LAUNCH_CHECK();
        if (nb > 0) {
            hipLaunchKernelGGL(k_sel_scatter, dim3(nb), dim3(kSelThreads), 0, s, n, t.cnt.get(), L,
                               (int64_t) tp.n_leaf, blk, t.sel_out.get());
            SPH_LAUNCH_CHECK();
        }
    }
    if (discover) {  // first build: wait for the counts and size the capacities by them
        fetch_info(t, s);
        SPH_HIP(hipStreamSynchronize(s));
        for (int l = 0; l < L; ++l) t.cap_lv[l] = cap_for(t.hinfo.p[kInfoRawCnt + l]);
    }
    LevelCaps caps{};
    int64_t total = 0, cmax = 1;
    for (int l = 0; l < L; ++l) {
        caps.cap[l] = t.cap_lv[l];
        total += t.cap_lv[l];
        cmax = std::max<int64_t>(cmax, t.cap_lv[l]);
    }
    t.cap_nodes = total;
    hipLaunchKernelGGL(k_level_layout, dim3(1), dim3(64), 0, s, L, caps, info);
    SPH_LAUNCH_CHECK();
    const size_t nn = (size_t) std::max<int64_t>(total, 1);
    t.nstart.ensure(nn); t.nend.ensure(nn); t.c0.ensure(nn); t.c1.ensure(nn);
    t.mass.ensure(nn); t.mx.ensure(nn); t.my.ensure(nn); t.mz.ensure(nn);
    t.comx.ensure(nn); t.comy.ensure(nn); t.comz.ensure(nn);
    t.lox.ensure(nn); t.loy.ensure(nn); t.loz.ensure(nn);
    t.hix.ensure(nn); t.hiy.ensure(nn); t.hiz.ensure(nn); t.maxH.ensure(nn); t.nrec.ensure(nn);
    {  // one launch per job for all levels
        const dim3 grid((unsigned) std::min<int64_t>(nblk(cmax, 256), 1024), (unsigned) L);
        hipLaunchKernelGGL(k_copy_levels, grid, dim3(256), 0, s, n, info, t.sel_out.get(), t.nstart.get());
        SPH_LAUNCH_CHECK();
        hipLaunchKernelGGL(k_node_ranges, grid, dim3(256), 0, s, n, info, t.nstart.get(), t.nend.get(), t.c0.get(),
                           t.c1.get());
        SPH_LAUNCH_CHECK();
    }
    MomArgs ma{t.nstart.get(), t.nend.get(), t.c0.get(), t.c1.get(), t.mass.get(), t.mx.get(), t.my.get(), t.mz.get(),
               t.comx.get(), t.comy.get(), t.comz.get(), t.lox.get(), t.loy.get(), t.loz.get(), t.hix.get(),
               t.hiy.get(), t.hiz.get(), t.maxH.get(), t.nrec.get()};
    hipLaunchKernelGGL(k_moments_leaf, dim3(nblk(t.cap_lv[0], 128)), dim3(128), 0, s, info, ma, P.x, P.y, P.z, P.m,
                       P.h);
    SPH_LAUNCH_CHECK();
    for (int l = 1; l < L; ++l) {
        hipLaunchKernelGGL(k_moments_level, dim3(nblk(t.cap_lv[l], 128)), dim3(128), 0, s, l, info, ma);
        SPH_LAUNCH_CHECK();
    }
    t.sys_root.ensure(tp.nsys);
    hipLaunchKernelGGL(k_fill_i32, dim3(nblk(tp.nsys, 64)), dim3(64), 0, s, (int64_t) tp.nsys, t.sys_root.get(), -1);
    SPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sys_root, dim3(nblk(t.cap_lv[L - 1], 64)), dim3(64), 0, s, L - 1, info, t.nstart.get(), P.sys,
                       t.sys_root.get());
    SPH_LAUNCH_CHECK();
    // i-groups: leaves packed greedily inside each level-2 node (or the top level)
    const int pl = std::min(2, L - 1);
    int32_t *gcnt = t.gcnt.ensure((size_t) t.cap_lv[pl] + 1);
    hipLaunchKernelGGL(k_group_count, dim3(nblk(t.cap_lv[pl], 64)), dim3(64), 0, s, pl, info, t.nstart.get(),
                       t.nend.get(), t.c0.get(), t.c1.get(), tp.n_group, gcnt);
    SPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_group_scan, dim3(1), dim3(kScanThreads), 0, s, pl, info, gcnt);
    SPH_LAUNCH_CHECK();
    if (discover) {  // first build: wait for the group count
        fetch_info(t, s);
        SPH_HIP(hipStreamSynchronize(s));
        t.gcap = cap_for(t.hinfo.p[kInfoRawG]);
    }
    t.gstart.ensure((size_t) t.gcap + 1);
    hipLaunchKernelGGL(k_group_layout, dim3(1), dim3(64), 0, s, info, t.gcap, n, t.gstart.get());
    SPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_group_write, dim3(nblk(t.cap_lv[pl], 64)), dim3(64), 0, s, pl, info, t.nstart.get(),
                       t.nend.get(), t.c0.get(), t.c1.get(), tp.n_group, gcnt, t.gstart.get());
    SPH_LAUNCH_CHECK();
    t.caps_known = true;
}

static void compute_keys(Tree &t, const TreeParams &tp, int64_t n, const PArrays &P, hipStream_t s) {
    compute_sys_boxes(t, tp, n, P, s);
    t.key.ensure(n);
    t.key_alt.ensure(n);
    hipLaunchKernelGGL(k_keys, dim3(nblk(n, 256)), dim3(256), 0, s, n, P.x, P.y, P.z, P.sys, tp.bits, t.sys_lo.get(),
                       t.sys_scale.get(), t.key.get());
    SPH_LAUNCH_CHECK();
}

void tree_sort_and_build(Tree &t, const TreeParams &tp, int64_t n, PArrays &P, PArrays &Palt, hipStream_t s) {
    compute_keys(t, tp, n, P, s);
    t.perm.ensure(n);
    t.vtmp.ensure(n);
    const int nb = radix_sort_tiles(n);
    t.rs_bh.ensure((size_t) std::max(nb, 1) * 256);
    t.rs_tot.ensure(256);
    uint64_t *kres = nullptr;
    int32_t *vres = nullptr;
    radix_sort_pairs(n, 3 * tp.bits + tp.sysbits, t.key.get(), t.key_alt.get(), t.perm.get(), t.vtmp.get(),
                     t.rs_bh.get(), t.rs_tot.get(), &kres, &vres, s);
    if (kres != t.key.get()) {
        std::swap(t.key.p, t.key_alt.p);
        std::swap(t.key.n, t.key_alt.n);
    }
    if (vres != t.perm.get()) {
        std::swap(t.perm.p, t.vtmp.p);
        std::swap(t.perm.n, t.vtmp.n);
    }
    // carry the fields that survive into the next pass of the step
    PermArgs pa;
    pa.perm = t.perm.get();
    const double *src[13] = {P.x, P.y, P.z, P.vx, P.vy, P.vz, P.vhx, P.vhy, P.vhz, P.uh, P.u, P.m, P.h};
    double *dst[13] = {Palt.x, Palt.y, Palt.z, Palt.vx, Palt.vy, Palt.vz, Palt.vhx, Palt.vhy, Palt.vhz, Palt.uh, Palt.u,
                       Palt.m, Palt.h};
    for (int f = 0; f < 13; ++f) { pa.src[f] = src[f]; pa.dst[f] = dst[f]; }
    pa.srci[0] = P.mat; pa.dsti[0] = Palt.mat;
    pa.srci[1] = P.sys; pa.dsti[1] = Palt.sys;
    pa.srcl = P.id; pa.dstl = Palt.id;
    hipLaunchKernelGGL(k_permute, dim3(nblk(n, 256)), dim3(256), 0, s, n, pa);
    SPH_LAUNCH_CHECK();
    std::swap(P, Palt);
    build_levels(t, tp, n, P, s);
}

void tree_build_only(Tree &t, const TreeParams &tp, int64_t n, const PArrays &P, hipStream_t s) {
    compute_keys(t, tp, n, P, s);
    // keys of an already sorted state are sorted; no permutation
    build_levels(t, tp, n, P, s);
}

void tree_update_maxH(Tree &t, const double *h, hipStream_t s) {
    const int L = t.nlevels;
    const int32_t *info = t.info.get();
    hipLaunchKernelGGL(k_maxH_leaf, dim3(nblk(t.cap_lv[0], 128)), dim3(128), 0, s, info, t.nstart.get(), t.nend.get(),
                       h, t.maxH.get());
    SPH_LAUNCH_CHECK();
    for (int l = 1; l < L; ++l) {
        hipLaunchKernelGGL(k_maxH_level, dim3(nblk(t.cap_lv[l], 128)), dim3(128), 0, s, l, info, t.c0.get(),
                           t.c1.get(), t.maxH.get());
        SPH_LAUNCH_CHECK();
    }
}

void tree_apply_info(Tree &t, const int32_t *h) {
    const int L = t.nlevels;
    for (int l = 0; l <= L; ++l) t.level_off[l] = h[kInfoOff + l];
    t.nnodes = h[kInfoOff + L];
    t.ngroups = h[kInfoG];
}

bool tree_grow_caps(Tree &t, const int32_t *h) {
    bool grown = false;
    for (int l = 0; l < t.nlevels; ++l) {
        const int64_t c = h[kInfoRawCnt + l];
        if (c + c / 64 + 8 > t.cap_lv[l]) {  // over, or close to, the capacity
            t.cap_lv[l] = cap_for(c);
            grown = true;
        }
    }
    const int64_t g = h[kInfoRawG];
    if ((h[kInfoErr] & 1) == 0 && g + g / 64 + 8 > t.gcap) {  // (after a node overflow the group count is 0)
        t.gcap = cap_for(g);
        grown = true;
    }
    return grown;
}

void tree_sync_counts(Tree &t, hipStream_t s) {
    fetch_info(t, s);
    SPH_HIP(hipStreamSynchronize(s));
    if (t.hinfo.p[kInfoErr]) throw Error(-2, "tree: counts beyond the capacities of the last build");
    tree_apply_info(t, t.hinfo.p);
}

static ListArgs list_args(const Tree &t, const TreeParams &tp, const PArrays &P, Lists &L) {
    ListArgs a;
    a.gcount = t.info.get() + kInfoG;
    a.gstart = t.gstart.get();
    a.sys = P.sys;
    a.x = P.x; a.y = P.y; a.z = P.z; a.h = P.h;
    a.nlevels = t.nlevels;
    a.sys_root = t.sys_root.get();
    a.nstart = t.nstart.get(); a.nend = t.nend.get(); a.c0 = t.c0.get(); a.c1 = t.c1.get();
    a.lox = t.lox.get(); a.loy = t.loy.get(); a.loz = t.loz.get();
    a.hix = t.hix.get(); a.hiy = t.hiy.get(); a.hiz = t.hiz.get();
    a.comx = t.comx.get(); a.comy = t.comy.get(); a.comz = t.comz.get();
    a.maxH = t.maxH.get();
    a.out = L.idx.get(); a.cnt = L.cnt.get(); a.cap = L.cap;
    a.sout = L.sidx.get(); a.scnt = L.scnt.get(); a.scap = L.scap;
    a.frontier = L.frontier.get();
    a.fcap = L.fcap;
    a.flds = std::min(std::max(L.flds, 1), kFrontLds);
    a.theta2 = tp.theta * tp.theta;
    a.hfac = 1.0;
    a.sbox = nullptr;
    a.err = L.err.get();
    a.builds = L.err.get() + 1;
    a.queue = L.err.get() + 2;
    a.need = nullptr;
    return a;
}

static int list_grid(int64_t ngroups, int waves) {
    int64_t b = std::min<int64_t>(ngroups, waves > 0 ? std::min(waves, kListSlots) : kListWavesDefault);
    b = (b + 7) & ~7LL;
    return (int) std::max<int64_t>(8, b);
}

namespace {
struct FillRanges {
    int32_t *p[3];
    int n[3];
};
__global__ void k_fill_ranges(FillRanges f) {
    for (int k = 0; k < 3; ++k)
        for (int i = threadIdx.x; i < f.n[k]; i += blockDim.x) f.p[k][i] = 0;
}
}  // namespace

void lists_reset_all(Lists &a, Lists &b, int32_t *flag, int nflag, hipStream_t s) {
    FillRanges f{{a.err.ensure(kListWords), b.err.ensure(kListWords), flag}, {kListWords, kListWords, flag ? nflag : 0}};
    hipLaunchKernelGGL(k_fill_ranges, dim3(1), dim3(64), 0, s, f);
    SPH_LAUNCH_CHECK();
    for (Lists *L : {&a, &b}) {
        int32_t *h = L->hflag.ensure(4);
        for (int k = 0; k < 4; ++k) h[k] = 0;
    }
}

void lists_reset(Lists &L, hipStream_t s) {
    L.err.ensure(kListWords);
    hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int64_t) 2, L.err.get(), 0);
    SPH_LAUNCH_CHECK();
    int32_t *h = L.hflag.ensure(4);
    for (int k = 0; k < 4; ++k) h[k] = 0;
}

void enqueue_lists(const Tree &t, const TreeParams &tp, const PArrays &P, int kind, Lists &L, double hfac,
                   hipStream_t s, const int32_t *need, bool cleared) {
    const int64_t G = t.gcap;
    if (kind != 0) hfac = 1.0;  // symmetric lists are exact: v_sig_max runs over the whole list
    L.idx.ensure((size_t) G * L.cap);
    L.cnt.ensure(G);
    if (kind == 2) {
        L.sidx.ensure((size_t) G * L.scap);
        L.scnt.ensure(G);
    }
    if (L.fcap <= 0) L.fcap = kFrontier0;
    const int grid = list_grid(G, L.waves);
    L.frontier.ensure((size_t) grid * 2 * L.fcap);
    L.err.ensure(kListWords);  // error bits, construction counter, the 8 per-XCD group counters
    if (!cleared) {
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(64), 0, s, (int64_t) 8, L.err.get() + 2, 0);
        SPH_LAUNCH_CHECK();
    }
    ListArgs a = list_args(t, tp, P, L);
    a.need = need;
    if (kind == 2) a.h = nullptr;  // geometry only: h may be in flux on another stream
    if (kind != 0) a.builds = nullptr;  // the counter counts gather-list constructions
    if (kind == 0) {  // gather lists remember their search boxes (reuse check)
        L.sbox.ensure((size_t) G * 6);
        a.sbox = L.sbox.get();
        a.hfac = hfac;
    }
    if (kind == 0) hipLaunchKernelGGL(k_list_sph<0>, dim3(grid), dim3(kWave), 0, s, a);
    else if (kind == 1) hipLaunchKernelGGL(k_list_sph<1>, dim3(grid), dim3(kWave), 0, s, a);
    else hipLaunchKernelGGL(k_list_grav, dim3(grid), dim3(kWave), 0, s, a);
    SPH_LAUNCH_CHECK();
    L.kind = kind;
}

void enqueue_gather_compact(const Tree &t, const PArrays &P, Lists &L, double hfac, hipStream_t s,
                            const int32_t *need, const double4 *pos) {
    if (L.kind != 0) throw Error(-1, "enqueue_gather_compact: the lists held are not gather lists");
    const int64_t G = t.gcap;
    L.crad.ensure((size_t) std::max<int64_t>(G, 1));
    if (G <= 0) return;
    hipLaunchKernelGGL(t.max_group > kWave ? k_list_compact_multi : k_list_compact, dim3((unsigned) G), dim3(kWave), 0,
                       s, t.info.get() + kInfoG, t.gstart.get(), P.x, P.y, P.z, P.h, pos, hfac, L.idx.get(),
                       L.cnt.get(), L.cap, L.crad.get(), need);
    SPH_LAUNCH_CHECK();
}

void enqueue_sym_compact(const Tree &t, const PArrays &P, const Lists &L, const double4 *pos, int32_t *oidx,
                         int32_t *ocnt, hipStream_t s) {
    if (L.kind != 1) throw Error(-1, "enqueue_sym_compact: the lists held are not symmetric lists");
    if (t.max_group > kWave) throw Error(-1, "enqueue_sym_compact: groups of more than 64 particles");
    const int64_t G = t.gcap;
    if (G <= 0) return;
    hipLaunchKernelGGL(k_list_compact_sym, dim3((unsigned) G), dim3(kWave), 0, s, t.info.get() + kInfoG,
                       t.gstart.get(), P.x, P.y, P.z, P.h, pos, L.idx.get(), L.cnt.get(), L.cap, oidx, ocnt);
    SPH_LAUNCH_CHECK();
}

void lists_fetch(Lists &L, hipStream_t s, int slot) {
    int32_t *h = L.hflag.ensure(4);
    SPH_HIP(hipMemcpyAsync(h + 2 * slot, L.err.get(), sizeof(int32_t) * 2, hipMemcpyDeviceToHost, s));
}

bool lists_grow_bits(Lists &L, int32_t err) {
    if (err & 4) L.fcap *= 2;
    if (err & 1) L.cap *= 2;
    if (err & 2) L.scap *= 2;
    if (err) L.kind = -1;
    return err == 0;
}

bool lists_grow(Lists &L) { return lists_grow_bits(L, L.hflag.p[0] | L.hflag.p[2]); }

// synchronous forms: build, wait, grow the capacities and build again until the lists fit
static bool build_lists_sync(const Tree &t, const TreeParams &tp, const PArrays &P, int kind, Lists &L, double hfac,
                             hipStream_t s) {
    for (int attempt = 0; attempt < 12; ++attempt) {
        lists_reset(L, s);
        enqueue_lists(t, tp, P, kind, L, hfac, s);
        lists_fetch(L, s);
        SPH_HIP(hipStreamSynchronize(s));
        if (lists_grow(L)) return true;
    }
    throw Error(-2, "interaction list capacity could not be grown enough");
}

bool build_lists_sph(const Tree &t, const TreeParams &tp, const PArrays &P, int kind, Lists &L, hipStream_t s,
                     double hfac) {
    return build_lists_sync(t, tp, P, kind, L, hfac, s);
}

bool build_lists_grav(const Tree &t, const TreeParams &tp, const PArrays &P, Lists &L, hipStream_t s) {
    return build_lists_sync(t, tp, P, 2, L, 1.0, s);
}

}  // namespace sphgpu
