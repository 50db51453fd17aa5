// internal.hpp -- shared declarations of the gfx950 SPH library (not part of
// the public C-ABI; see include/fdps_sph_gpu.h for that).
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace sphgpu {

// ------------------------------------------------------------------ errors
struct Error : public std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define SPH_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::sphgpu::Error(-4, std::string(#call) + ": " + hipGetErrorString(e_) + " at " + \
                                          __FILE__ + ":" + std::to_string(__LINE__));              \
    } while (0)

#define SPH_LAUNCH_CHECK() SPH_HIP(hipGetLastError())

// ------------------------------------------------------------------ device buffer (grow-only)
// Counts every allocation of the buffers below: a recorded step (HIP graph)
// holds buffer addresses and is replayed only while this is unchanged.
inline std::atomic<uint64_t> g_alloc_epoch{0};
inline uint64_t alloc_epoch() { return g_alloc_epoch.load(); }

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void) hipFree(p);
        p = nullptr;
        n = 0;
    }
    // The first allocation is exact; a buffer that has to grow takes a quarter
    // more than asked, so that sizes which creep up from step to step (node,
    // group and list counts) reallocate rarely -- hipFree waits for the device.
    T *ensure(size_t count) {
        if (count <= n && p) return p;
        size_t c = count ? count : 1;
        if (p) c = std::max(c, n + n / 4);
        release();
        SPH_HIP(hipMalloc(&p, c * sizeof(T)));
        n = c;
        g_alloc_epoch.fetch_add(1);
        return p;
    }
    T *get() const { return p; }
};

template <class T>
struct PinnedBuf {
    T *p = nullptr;
    size_t n = 0;
    // hipHostMallocCoherent for words that kernels write and the host reads
    // after the stream drains
    unsigned flags = hipHostMallocDefault;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() {
        if (p) (void) hipHostFree(p);
    }
    T *ensure(size_t count) {
        if (count <= n && p) return p;
        if (p) (void) hipHostFree(p);
        p = nullptr;
        size_t c = count ? count : 1;
        SPH_HIP(hipHostMalloc(&p, c * sizeof(T), flags));
        n = c;
        g_alloc_epoch.fetch_add(1);
        return p;
    }
};

// ------------------------------------------------------------------ constants
// src/param.h:9-17, src/kernel.h:36-38, src/force.h:171-172, src/mathfunc.h:4
constexpr double kSMTH = 1.2;
constexpr double kCFL = 0.3;
constexpr double kAV = 1.0;
constexpr double kSupport = 2.5;
constexpr double kG = 6.67e-11;
constexpr double kPi = 3.14159265358979323846;
// relative widening of the squared support radius in the exact skip test: a
// pair with r^2 above (2.5h)^2 (1 + 1e-10) has W = gradW = 0 exactly
constexpr double kCutWiden = 1.0 + 1.0e-10;

constexpr int kWave = 64;
constexpr int kMaxLevels = 24;

// ------------------------------------------------------------------ interaction-list reference
// Group g's j-list is idx[begin(g) .. begin(g)+cnt[g]).  off == nullptr means
// fixed capacity lists: begin(g) = g * cap.
struct ListRef {
    const int32_t *idx;
    const int64_t *off;
    const int32_t *cnt;
    int64_t cap;
};

// Particle SoA arrays of the device-resident simulation.
struct PArrays {
    double *x, *y, *z, *vx, *vy, *vz, *ax, *ay, *az, *gx, *gy, *gz, *pot, *m, *h, *rho, *u, *p, *cs, *divv,
        *rotx, *roty, *rotz, *gs, *bal, *du, *dti, *vhx, *vhy, *vhz, *uh;
    int32_t *mat, *sys, *ngb;
    int64_t *id;
};

// ------------------------------------------------------------------ interaction kernel arguments
struct DensArgs {
    int ngroups;
    const int32_t *gstart;
    const double *xi, *yi, *zi, *mi, *hi;
    const double *mj, *xj, *yj, *zj;
    ListRef L;
    double *rho_out, *h_out;
    int32_t *ngb_out;
    // if set: the j side packed as (x, y, z, m), one 32-byte record per particle
    const double4 *jpos = nullptr;
    // persistent scheduling: 8 per-XCD group counters, cleared before the launch
    // (null: one workgroup per group)
    int32_t *sched = nullptr;
    // if set: the group count is read here (a tree build's device count, at
    // most ngroups, which then only sizes the launch)
    const int32_t *gcount = nullptr;
    // work items per group: a group of up to 64 * chunks particles runs as
    // `chunks` one-wave items sharing its list (ngroups then counts items)
    int chunks = 1;
    // if flag is set: after the pass, whether each group's gather list still
    // covers the new supports (gather_cover_check against the list's search
    // box sbox and, if given, its compaction radius crad); *flag |= 1 if not
    const double *sbox = nullptr, *crad = nullptr;
    int32_t *flag = nullptr;
};

struct DrvtArgs {
    int ngroups;
    const int32_t *gstart;
    const double *xi, *yi, *zi, *vxi, *vyi, *vzi, *hi, *rhoi;
    const double *csi;  // non-null: also write the Balsara switch (bal)
    const double *mj, *xj, *yj, *zj, *vxj, *vyj, *vzj;
    ListRef L;
    double *divv, *rotx, *roty, *rotz, *gs, *bal;
    // if set: the j side packed as (x, y, z, m) and (vx, vy, vz, 0)
    const double4 *jpos = nullptr, *jvel = nullptr;
    // persistent scheduling: 8 per-XCD group counters, cleared before the launch
    // (null: one workgroup per group)
    int32_t *sched = nullptr;
    // if set: the group count is read here (a tree build's device count, at
    // most ngroups, which then only sizes the launch)
    const int32_t *gcount = nullptr;
    // work items per group: a group of up to 64 * chunks particles runs as
    // `chunks` one-wave items sharing its list (ngroups then counts items)
    int chunks = 1;
};

// The j side of the hydro pass as one 128-byte record per particle (one
// cache line per list entry instead of 13 gathers from separate arrays):
// A = f P / rho^2, H = 2.5 h, 1/H, the kernel normalisation and the gradient's
// regularisation reg = 1e-6 h are formed by the same FP64 operations the
// kernel would apply to the separate fields.
struct alignas(16) HydroRec {
    double x, y, z, vx, vy, vz, m, rho, cs, bal, gs, A, reg, H, norm, invH;
};
static_assert(sizeof(HydroRec) == 128, "one cache line per record");
// The precision form's hydro j record: position and velocity in FP64 (the
// kernel forms their offsets from the group's centres), the other fields
// converted to FP32 once per particle from the same FP64 values as HydroRec's.
struct alignas(16) HydroRecF {
    double x, y, z, vx, vy, vz;
    float4 k;  // rho, Balsara switch, 1/H, H
    float4 w;  // kernel normalisation, 1e-6 h, grad-h term f, mass
    float4 c;  // sound speed, A = f P / rho^2, 0, 0
};
static_assert(sizeof(HydroRecF) == 96, "record stride");
// The dt part's j record: position, velocity and sound speed, half a cache
// line (the 128-byte hydro record would be fetched whole for 56 bytes of it).
struct alignas(16) VsigRec {
    double x, y, z, vx, vy, vz, cs, pad;
};
static_assert(sizeof(VsigRec) == 64, "record stride");

struct HydroArgs {
    int ngroups;
    const int32_t *gstart;
    const double *xi, *yi, *zi, *vxi, *vyi, *vzi, *hi, *rhoi, *pi, *csi, *gsi, *bali;
    const double *xj, *yj, *zj, *vxj, *vyj, *vzj, *hj, *rhoj, *pj, *csj, *gsj, *balj, *mj;
    ListRef L;
    double *ax, *ay, *az, *du, *dt;
    const HydroRec *jrec = nullptr;  // if set: the j side is read from here (list indices index it)
    const HydroRecF *jrecf = nullptr;  // the precision form's records, read in preference to jrec
    const VsigRec *vrec = nullptr;     // the dt kernel's j records (launch_vsig reads only these)
    // persistent scheduling: 8 per-XCD group counters, cleared before the launch
    // (null: one workgroup per group)
    int32_t *sched = nullptr;
    // if set: the group count is read here (a tree build's device count, at
    // most ngroups, which then only sizes the launch)
    const int32_t *gcount = nullptr;
    // work items per group: a group of up to 64 * chunks particles runs as
    // `chunks` one-wave items sharing its list (ngroups then counts items)
    int chunks = 1;
    // 1: the hydro kernel also forms v_sig_max over the list and writes dt;
    // 0: launch_vsig does that (a kernel of its own, which may run beside
    // other work); the hydro kernel then writes acc and du only
    int pass1 = 1;
};

struct GravArgs {
    int ngroups;
    const int32_t *gstart;
    const double *xi, *yi, *zi;
    const double *eps2i;  // if null, eps2 = h*h from hi
    const double *hi;
    const double *xj, *yj, *zj, *mj;  // particles (EPJ)
    ListRef Lj;
    const double *xs, *ys, *zs, *ms;  // monopoles (SPJ); may be null
    ListRef Ls;
    double *ax, *ay, *az, *pot;
    // if set: the particles' (x, y, z, m) and the monopoles' (x, y, z, m) as
    // one 32-byte record each, read instead of the separate arrays
    const double4 *jrec = nullptr, *srec = nullptr;
    // persistent scheduling: 8 per-XCD group counters, cleared before the launch
    // (null: one workgroup per group)
    int32_t *sched = nullptr;
    // if set: the group count is read here (a tree build's device count, at
    // most ngroups, which then only sizes the launch)
    const int32_t *gcount = nullptr;
    // work items per group: a group of up to 64 * chunks particles runs as
    // `chunks` one-wave items sharing its list (ngroups then counts items)
    int chunks = 1;
};

// Pack the j records of the density and derivative passes for particles
// [0, n): (x, y, z, m) and (vx, vy, vz, 0); positions, masses and velocities
// are fixed during a force evaluation.
// Persistent launches of the interaction kernels: the 8 per-XCD group
// counters cleared by a small kernel on the launch's stream, and the grid of
// workgroups the GPU holds at once for `kernel` (occupancy x compute units),
// a multiple of 8, at most the group count rounded up to 8.
void launch_sched_reset(int32_t *sched, hipStream_t s);
int persistent_grid(const void *kernel, int ngroups);
void launch_pack_kin(int64_t n, const double *x, const double *y, const double *z, const double *m, const double *vx,
                     const double *vy, const double *vz, double4 *pos, double4 *vel, hipStream_t s,
                     const double *h = nullptr, double *hcopy = nullptr);  // and h to hcopy, if given
// Pack the hydro records of particles [0, n) from the SoA fields.
void launch_pack_hydro(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *m, const double *rho, const double *cs,
                       const double *bal, const double *gs, const double *p, const double *h, HydroRec *out,
                       hipStream_t s, VsigRec *vout = nullptr);
// The same particles' HydroRecF records (precision form).
void launch_pack_hydro_f(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz, const double *m, const double *rho, const double *cs,
                         const double *bal, const double *gs, const double *p, const double *h, HydroRecF *out,
                         hipStream_t s, VsigRec *vout = nullptr);

// FP64 (plain) interaction kernels -- kernels_fp64.hip
void launch_dens_fp64(const DensArgs &a, hipStream_t s);
void launch_drvt_fp64(const DrvtArgs &a, hipStream_t s);
void launch_hydro_fp64(const HydroArgs &a, hipStream_t s);
// dt from v_sig_max over each group's list in FP64 (both precision forms),
// the j side from the dt records a.vrec (packed by launch_pack_hydro[_f])
void launch_vsig(const HydroArgs &a, hipStream_t s);
void launch_grav_fp64(const GravArgs &a, hipStream_t s);

// Mixed-precision interaction kernels -- kernels_mixed.hip
void launch_dens_mixed(const DensArgs &a, hipStream_t s);
void launch_drvt_mixed(const DrvtArgs &a, hipStream_t s);
void launch_hydro_mixed(const HydroArgs &a, hipStream_t s);
void launch_grav_mixed(const GravArgs &a, hipStream_t s);

// ------------------------------------------------------------------ EoS -- eos.hip
struct MatDev {
    int kind;             // 0: rho-u table, 1: analytic Tillotson (till)
    double till[10];      // rho0, u0, u_iv, u_cv, A, B, a, b, alpha, beta (src/EoS.h:446 order)
    int nr, nu;
    const double *rho;    // nr
    const double *u;      // nu
    const double *plane;  // [6][nr][nu]
    const double *full_rho, *full_S, *full_u;  // row-major cell fields (nr*nu)
    int sorted_rho;       // full_rho non-decreasing (binary search allowed)
};
void launch_eos_pressure(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                         int nmat, double *p, double *cs, hipStream_t s);
void launch_eos_property(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                         int nmat, int prop, double *out, hipStream_t s);
void launch_eos_internal_energy(int64_t n, const int32_t *mat, const double *rho, const double *S,
                                const MatDev *mats, int nmat, int abs_int, double *out, hipStream_t s);

// ------------------------------------------------------------------ tree -- tree.hip
struct TreeParams {
    int n_leaf, n_group, nsys, bits, sysbits, nlevels;
    double theta;
};

// Device-side counts of a tree build (Tree::info, int32 words): the node
// count of each level and the levels' offsets in the flattened node array
// (every count 0 when some level exceeds its capacity), the group count (0
// when the nodes or the groups exceed their capacities), error bits (1 nodes,
// 2 groups beyond capacity) and the counts as found.
constexpr int kInfoCnt = 0;
constexpr int kInfoOff = kInfoCnt + kMaxLevels;  // kMaxLevels + 1 words
constexpr int kInfoG = kInfoOff + kMaxLevels + 1;
constexpr int kInfoErr = kInfoG + 1;
constexpr int kInfoRawCnt = kInfoErr + 1;
constexpr int kInfoRawG = kInfoRawCnt + kMaxLevels;
constexpr int kInfoWords = kInfoRawG + 1;

struct Tree {
    // per-system
    DevBuf<int32_t> sys_start;  // nsys+1 (host-known, fixed)
    std::vector<int32_t> sys_start_h;  // the same offsets on the host
    DevBuf<double> sys_lo, sys_scale;  // 3*nsys, nsys
    DevBuf<double> part;           // bbox partials
    DevBuf<int32_t> sys_root;      // nsys
    // keys and sort
    DevBuf<uint64_t> key, key_alt;
    DevBuf<int32_t> perm, vtmp;    // the sorted permutation; the sort's other value buffer
    DevBuf<int32_t> rs_bh, rs_tot; // the sort's per-tile digit counts and digit totals
    DevBuf<int64_t> cnt;
    DevBuf<int32_t> sel_out;
    DevBuf<int32_t> sel_blk;  // per-block node-start counts of every level, then their offsets
    // counts of the last build (see kInfo*) and their host copy
    DevBuf<int32_t> info;
    PinnedBuf<int32_t> hinfo;
    // capacities the launches and buffers are sized by; chosen from the counts
    // of the first build (which waits for them) and grown when counts come near
    bool caps_known = false;
    int32_t cap_lv[kMaxLevels] = {0};
    int64_t cap_nodes = 0;
    int64_t gcap = 0;
    int max_group = 64;  // the most particles a group holds (n_group)
    // nodes (levels flattened, level 0 = leaves); the counts on the host are
    // those of the last tree_apply_info
    int nlevels = 0;
    int64_t level_off[kMaxLevels + 1] = {0};
    int64_t nnodes = 0;
    DevBuf<int32_t> nstart, nend, c0, c1;
    DevBuf<double> mass, mx, my, mz, comx, comy, comz, lox, loy, loz, hix, hiy, hiz, maxH;
    DevBuf<double4> nrec;  // (com x, y, z, mass) of each node, the gravity kernels' monopole record
    // i-groups
    int64_t ngroups = 0;
    DevBuf<int32_t> gstart;
    DevBuf<int32_t> gcnt;  // groups of each packing-level node, then their offsets
    int64_t n = 0;  // particles of the last build
};

// Stable radix sort of n (key, value) pairs on the low `keybits` key bits:
// keys from k0, values the input indices; k1 and v0, v1 are the other
// buffers, bh holds radix_sort_tiles(n) * 256 counts and tot 256.  The sorted
// keys and values end in *kres and *vres (one of k0/k1, v0/v1).
int radix_sort_tiles(int64_t n);
void radix_sort_pairs(int64_t n, int keybits, uint64_t *k0, uint64_t *k1, int32_t *v0, int32_t *v1, int32_t *bh,
                      int32_t *tot, uint64_t **kres, int32_t **vres, hipStream_t s);

void tree_sort_and_build(Tree &t, const TreeParams &tp, int64_t n, PArrays &P, PArrays &Palt, hipStream_t s);
void tree_build_only(Tree &t, const TreeParams &tp, int64_t n, const PArrays &P, hipStream_t s);
void tree_update_maxH(Tree &t, const double *h, hipStream_t s);
// Host copies of the counts from a host copy of Tree::info.
void tree_apply_info(Tree &t, const int32_t *info_host);
// Grows the capacities that the counts of info_host exceed or come close to;
// true if any grew.
bool tree_grow_caps(Tree &t, const int32_t *info_host);
// Waits for the stream, reads the counts and applies them (an error if the
// last build exceeded its capacities).
void tree_sync_counts(Tree &t, hipStream_t s);

// BFS frontier entries a list-walk wave keeps in LDS (two buffers of this many)
constexpr int kFrontLds = 512;
struct Lists {
    DevBuf<int32_t> idx, cnt;
    DevBuf<int32_t> sidx, scnt;
    int64_t cap = 0, scap = 0;
    DevBuf<int32_t> frontier;
    int64_t fcap = 0;        // per-workgroup BFS frontier capacity in global memory
    int flds = kFrontLds;    // frontier entries kept in LDS before the global part
    int waves = 0;           // persistent waves of a walk (0: the default, 4096)
    DevBuf<int32_t> err;       // overflow bits, construction counter, per-XCD group counters
    DevBuf<int32_t> flag;      // per density pass: 1 when some group outgrew its gather list
    PinnedBuf<int32_t> hflag;  // host copy of the overflow bits and the construction counter
    DevBuf<double> sbox;     // gather lists: per-group search box (lo[3], hi[3])
    DevBuf<double> crad;     // compacted gather lists: per-group radius R_g they cover
    int kind = -1;           // kind of the lists held (-1 none, 0 gather, 1 symmetric, 2 gravity)
};
// Search-radius factor of the gather lists that the density passes reuse:
// 8 % on h covers the change of h over the density iterations of nearly every
// step at 1M (1.02 constructions per force evaluation against 1.39 with 3 %,
// for lists 2 % longer).
constexpr double kGatherMargin = 1.08;
constexpr int kListWords = 10;  // Lists::err: overflow bits, construction counter, 8 group counters
// kind 0 gather, 1 symmetric, 2 gravity.  Gather lists use the search radius
// 2.5 h hfac (hfac >= 1: a superset that stays usable while h grows);
// symmetric lists are always exact.
//
// Asynchronous form: lists_reset clears the overflow bits and the construction
// counter; enqueue_lists builds without waiting -- when `need` is given, only
// if that device flag is nonzero -- and ORs overflows into the bits (the lists
// are then incomplete); lists_fetch copies bits and counter to slot 0 or 1 of
// L.hflag (4 ints), and after the stream has drained lists_grow enlarges the
// capacities the bits of either slot name and returns true when there were none.
void lists_reset(Lists &L, hipStream_t s);
// lists_reset of two lists and nflag zero words of flag in one launch, which
// also clears the lists' group counters: the next walk of each may then pass
// `cleared` to enqueue_lists, which otherwise clears them itself
void lists_reset_all(Lists &a, Lists &b, int32_t *flag, int nflag, hipStream_t s);
void enqueue_lists(const Tree &t, const TreeParams &tp, const PArrays &P, int kind, Lists &L, double hfac,
                   hipStream_t s, const int32_t *need = nullptr, bool cleared = false);
void lists_fetch(Lists &L, hipStream_t s, int slot = 0);
bool lists_grow(Lists &L);
bool lists_grow_bits(Lists &L, int32_t err);
// Compacts each group's gather list in place to the entries within R_g of
// some particle of the group, R_g = hfac * max_i 2.5 h_i (h as now), keeping
// their order; when `need` is given, only if that device flag is nonzero (as
// the construction it follows).  pos: optional packed (x, y, z, m) records.
void enqueue_gather_compact(const Tree &t, const PArrays &P, Lists &L, double hfac, hipStream_t s,
                            const int32_t *need = nullptr, const double4 *pos = nullptr);
// Writes to (oidx, ocnt), capacity L.cap per group, each group's symmetric
// list reduced in order to the entries within max(max_i 2.5 h_i, 2.5 h_j) of
// some particle of the group: every pair inside either support is kept.
// Groups of at most 64 particles; pos: packed (x, y, z, m) records or null.
void enqueue_sym_compact(const Tree &t, const PArrays &P, const Lists &L, const double4 *pos, int32_t *oidx,
                         int32_t *ocnt, hipStream_t s);
// Synchronous forms: capacities grow until the lists fit.
bool build_lists_sph(const Tree &t, const TreeParams &tp, const PArrays &P, int kind, Lists &L, hipStream_t s,
                     double hfac = 1.0);
bool build_lists_grav(const Tree &t, const TreeParams &tp, const PArrays &P, Lists &L, hipStream_t s);

// ------------------------------------------------------------------ integrator -- integrate.hip
// Error words a step's last kernels read before they change the state (the
// final kick, the time step): any nonzero word means that the force
// evaluation ran out of a capacity and is to be run again, so they leave the
// state as it was.  Null words are not read.
struct StepGuard {
    const int32_t *w[3] = {nullptr, nullptr, nullptr};
    __device__ __forceinline__ bool failed() const {
        int32_t e = 0;
        for (int k = 0; k < 3; ++k)
            if (w[k]) e |= *w[k];
        return e != 0;
    }
};
void launch_init_h(int64_t n, PArrays P, hipStream_t s);
// kick and drift of every particle, and each system's time advanced by its dt
void launch_kick_drift(int64_t n, PArrays P, const double *dt_sys, int nsys, double *time_sys, hipStream_t s);
void launch_finalize(int64_t n, PArrays P, const double *dt_sys, const double *time_sys, double until,
                     double damping, StepGuard guard, hipStream_t s);
void launch_ext_force(int64_t n, PArrays P, const double *time_sys, double until, hipStream_t s);
void launch_dt_reduce(int nsys, const int32_t *sys_start_host, const int32_t *sys_start, int64_t n,
                      const double *dti, double *part, double *dt_sys, StepGuard guard, hipStream_t s);

// ------------------------------------------------------------------ device helpers
__device__ __forceinline__ int xcd_remap(int b, int n) {
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of groups so that neighbouring groups share its L2.
    const int q = n >> 3, r = n & 7;
    const int x = b & 7, k = b >> 3;
    return x * q + (x < r ? x : r) + k;
}

// The group count of an interaction launch: a device count when the args
// carry one (the grid is then sized by a capacity), else the launch's own.
template <class A>
__device__ __forceinline__ int group_count(const A &a) { return a.gcount ? *a.gcount * a.chunks : a.ngroups; }
// The group of work item k: a group of more than 64 particles is processed as
// `chunks` items of one wave each (items g*chunks .. g*chunks + chunks - 1),
// which share the group's list.
template <class A>
__device__ __forceinline__ int group_of(const A &a, int k) { return a.chunks == 1 ? k : k / a.chunks; }

// Work distribution of the interaction kernels over the i-groups.  Without a
// scheduler each workgroup takes one group through xcd_remap.  With one (8
// per-XCD counters, zero at the launch) a fixed grid of persistent workgroups,
// a multiple of 8 so that each keeps one XCD, loops: a workgroup takes the
// next group of its own XCD's contiguous range (the ranges of xcd_remap), and
// when that is exhausted the next group of another XCD's range.
struct GroupIter {
    int32_t *sched;
    int n, x;
    __device__ GroupIter(int32_t *s, int ngroups) : sched(s), n(ngroups), x((int) (blockIdx.x & 7)) {}
    __device__ int first() {
        if (!sched) return (int) blockIdx.x < n ? xcd_remap((int) blockIdx.x, n) : -1;
        return next();
    }
    __device__ int next() {
        if (!sched) return -1;
        const int q = n >> 3, r = n & 7;
        for (int k = 0; k < 8; ++k) {
            const int X = (x + k) & 7;
            const int lo = X * q + (X < r ? X : r), cnt = q + (X < r ? 1 : 0);
            int t = 0;
            if (threadIdx.x == 0) t = atomicAdd(&sched[X], 1);
            t = __shfl(t, 0, kWave);
            if (t < cnt) return lo + t;
        }
        return -1;
    }
};

// Tiling of a group's j-list: T = ceil(nj / 64) tiles, tile t holding the
// entries t, t + T, t + 2T, ... (lane l loads entry t + T l).  Interleaving
// spreads every tile over the whole search region, so each lane finds about
// the same number of in-support j per tile and the per-lane pair loops stay
// balanced (consecutive entries are one compact Morton chunk, close to some
// lanes and far from others).  Each lane still visits every entry once, so the
// sums run over the same terms, associated in tile order.
struct Tiles {
    int nj, T;
    __device__ __forceinline__ explicit Tiles(int n) : nj(n), T((n + kWave - 1) / kWave) {}
    __device__ __forceinline__ int entry(int t, int lane) const { return t + T * lane; }
    __device__ __forceinline__ int count(int t) const { return (nj - t + T - 1) / T; }
};

// Cross-lane operations on the VALU's data-parallel-primitive and lane-swap
// paths instead of an LDS permute (a ds_bpermute round trip per step).  The
// reductions take six steps -- swaps within quads (1,0,3,2 and 2,3,0,1), row
// rotations by 4 and 8 (every lane of a 16-lane row then holds its row's
// result), lane 15's value into rows 1 and 3, lane 31's into rows 2 and 3 --
// after which lane 63 holds the wave's, read back as a uniform value.  Lanes
// of rows a step does not write keep their value, which min and max absorb.
// Every lane of the wave must take part (uniform control flow: a lane
// switched off would neither pass its value on nor receive the result).
// Lane layouts checked on the device against the plain loop.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i_(int v) { return __builtin_amdgcn_update_dpp(v, v, CTRL, ROWS, 0xF, false); }
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_f_(float v) { return __int_as_float(dpp_i_<CTRL, ROWS>(__float_as_int(v))); }
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d_(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = dpp_i_<CTRL, ROWS>((int) b), hi = dpp_i_<CTRL, ROWS>((int) (b >> 32));
    return __longlong_as_double((long long) (((unsigned long long) (unsigned) hi << 32) | (unsigned) lo));
}
__device__ __forceinline__ float lane63_f_(float v) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63)); }
__device__ __forceinline__ double lane63_d_(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int) b, 63), hi = __builtin_amdgcn_readlane((int) (b >> 32), 63);
    return __longlong_as_double((long long) (((unsigned long long) (unsigned) hi << 32) | (unsigned) lo));
}
#define SPHGPU_WAVE_REDUCE_(v, OP, DPP)  \
    v = OP(v, DPP<0xB1, 0xF>(v));        \
    v = OP(v, DPP<0x4E, 0xF>(v));        \
    v = OP(v, DPP<0x124, 0xF>(v));       \
    v = OP(v, DPP<0x128, 0xF>(v));       \
    v = OP(v, DPP<0x142, 0xA>(v));       \
    v = OP(v, DPP<0x143, 0xC>(v))
__device__ __forceinline__ double min_d_(double a, double b) { return (b < a) ? b : a; }  // std::min
__device__ __forceinline__ double max_d_(double a, double b) { return (a < b) ? b : a; }  // std::max
__device__ __forceinline__ float dpp_wave_maxf(float v) {
    SPHGPU_WAVE_REDUCE_(v, fmaxf, dpp_f_);
    return lane63_f_(v);
}
__device__ __forceinline__ float dpp_wave_minf(float v) {
    SPHGPU_WAVE_REDUCE_(v, fminf, dpp_f_);
    return lane63_f_(v);
}
__device__ __forceinline__ double dpp_wave_min(double v) {
    SPHGPU_WAVE_REDUCE_(v, min_d_, dpp_d_);
    return lane63_d_(v);
}
__device__ __forceinline__ double dpp_wave_max(double v) {
    SPHGPU_WAVE_REDUCE_(v, max_d_, dpp_d_);
    return lane63_d_(v);
}
#undef SPHGPU_WAVE_REDUCE_
// The value of lane l ^ 32, through one swap of the two wave halves.
__device__ __forceinline__ int xor32_i(int v) {
    const auto r = __builtin_amdgcn_permlane32_swap(v, v, false, false);
    return (threadIdx.x & 32) ? (int) r[0] : (int) r[1];
}
__device__ __forceinline__ float xor32_f(float v) { return __int_as_float(xor32_i(__float_as_int(v))); }

// Dot-form mask threshold of one lane for one tile (see dot_mask): the
// in-support test |x_j - x_i|^2 <= H^2 evaluated through FP32 offsets needs
// room for the offsets' rounding (radius + 2^-22 (|o_i|_1 + N)) and for the
// FMA chain's (2^-19 (|o_i|_1^2 + N^2)), N = largest |o_j|_1 of the tile.
__device__ __forceinline__ float dot_cut_(float Hc, float wi, float ni, float nmax) {
    const float r = Hc + 2.384185791015625e-7f * (ni + nmax);
    return fmaf(r, r, -wi) + 1.9073486328125e-6f * (ni * ni + nmax * nmax);
}
// The same room for j's support tested from any lane of the group: Q_j for
// mfma_dot_mask_sym with n1max the group's largest |o_i|_1.
__device__ __forceinline__ float dot_cut_j_(float Hcj, float nj, float n1max) {
    const float R = Hcj + 2.384185791015625e-7f * (n1max + nj);
    return fmaf(R, R, 1.9073486328125e-6f * (n1max * n1max + nj * nj));
}
__device__ __forceinline__ float norm1_(float ox, float oy, float oz) { return fabsf(ox) + fabsf(oy) + fabsf(oz); }

// Support mask of one j-tile in dot form.  q[jj] = (o.x, o.y, o.z, w) holds
// FP32 offsets o_j of the tile's entries from the group centre and
// w = |o_j|^2; with t = (-2 o_i) . o_j + w = |o_j - o_i|^2 - |o_i|^2, bit jj
// of the result is the sign of t - c (set: t < c, taken as inside the cut;
// c = cut^2 - |o_i|^2, c = -inf admits nothing).  Each test is three FMAs, a
// subtraction and one funnel shift that moves the sign bit into the mask.
__device__ __forceinline__ uint32_t dot_bit_(uint32_t m, const float4 &q, float tx, float ty, float tz, float c) {
    float t = fmaf(tx, q.x, q.w);
    t = fmaf(ty, q.y, t);
    t = fmaf(tz, q.z, t);
    return __builtin_amdgcn_alignbit(m, __float_as_uint(t - c), 31);
}
__device__ __forceinline__ uint64_t dot_mask(const float4 *q, int nt, float tx, float ty, float tz, float c) {
    uint32_t lo = 0, hi = 0;
    if (nt == kWave) {  // full tile: unrolled, constant LDS offsets
#pragma unroll
        for (int jj = kWave - 1; jj >= 32; --jj) hi = dot_bit_(hi, q[jj], tx, ty, tz, c);
#pragma unroll
        for (int jj = 31; jj >= 0; --jj) lo = dot_bit_(lo, q[jj], tx, ty, tz, c);
    } else {
        for (int jj = nt - 1; jj >= 32; --jj) hi = dot_bit_(hi, q[jj], tx, ty, tz, c);
        for (int jj = (nt < 32 ? nt : 32) - 1; jj >= 0; --jj) lo = dot_bit_(lo, q[jj], tx, ty, tz, c);
    }
    return ((uint64_t) hi << 32) | lo;
}

// The same mask from the matrix cores.  t_ij - c_i is the bilinear form
// (o_j, w_j, 1, 0) . (-2 o_i, 1, -c_i, 0): twelve v_mfma_f32_32x32x2f32 per
// 64 x 64 tile (four 32 x 32 blocks, K = 6 in three steps) instead of 64 x 3
// FMAs per lane, then one sign-bit shift per pair.  Operand layout of the
// instruction: lane l supplies A[l%32][l/32] and B[l/32][l%32]; the result
// holds column l%32, register r row 8(r/4) + 4(l/32) + r%4.  Column n of
// column block ib is particle 32 ib + n: lane n supplies its even components
// and lane n + 32 its odd ones, each lane giving its own particle's values
// when its half is ib and its partner's (lane l^32) otherwise.  Rows j are fed
// tile entries in the order that makes register r of lane half h the entry
// 32 jb + 16 h + r, so each lane's 16 signs of a block are 16 consecutive
// entries; the blocks of ib = h are this lane's particle, those of the other
// ib its partner's, and lanes l and l^32 swap the latter.
// (tx, ty, tz) = -2 o_i and negc = -c_i of this lane, (px, py, pz, pc) the
// same of lane l^32; entries >= nt are cut off.  The rounding of the matrix
// sums (a few ulp of max(H^2, |o|^2)) is of the order the thresholds of
// dot_cut_ already allow for the FMA chain.
typedef float f32x16_ __attribute__((ext_vector_type(16)));
__device__ __forceinline__ uint32_t sign_bits16_(const f32x16_ &c) {
    uint32_t w = 0;
#pragma unroll
    for (int r = 15; r >= 0; --r) w = __builtin_amdgcn_alignbit(w, __float_as_uint(c[r]), 31);
    return w;
}
// q(e) gives entry e's (o_x, o_y, o_z, w); the array form follows.
template <class QF>
__device__ __forceinline__ uint64_t mfma_dot_mask_f(QF q, int nt, float tx, float ty, float tz, float negc, float px,
                                                    float py, float pz, float pc) {
    const int lane = threadIdx.x, h = lane >> 5, m = lane & 31;
    // j operands: half 0 supplies k = 0, 2, 4 (o_x, o_z, 1), half 1 k = 1, 3, 5 (o_y, w, 0)
    const int e = 16 * ((m >> 2) & 1) + 4 * (m >> 3) + (m & 3);
    const float4 q0 = q(e), q1 = q(32 + e);
    const float a00 = h ? q0.y : q0.x, a01 = h ? q0.w : q0.z, a02 = h ? 0.f : 1.f;
    const float a10 = h ? q1.y : q1.x, a11 = h ? q1.w : q1.z, a12 = a02;
    // i operands: half 0 supplies k = 0, 2, 4 (-2 o_x, -2 o_z, -c), half 1 k = 1, 3, 5 (-2 o_y, 1, 0),
    // of particle n (block 0) or 32 + n (block 1)
    const float b00 = h ? py : tx, b01 = h ? 1.f : tz, b02 = h ? 0.f : negc;
    const float b10 = h ? ty : px, b11 = h ? 1.f : pz, b12 = h ? 0.f : pc;
    f32x16_ c;
    c = f32x16_{};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a00, b00, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a01, b01, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a02, b02, c, 0, 0, 0);
    const uint32_t w00 = sign_bits16_(c);  // entries of row block 0, particles of column block 0
    c = f32x16_{};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a10, b00, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a11, b01, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a12, b02, c, 0, 0, 0);
    const uint32_t w10 = sign_bits16_(c);
    c = f32x16_{};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a00, b10, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a01, b11, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a02, b12, c, 0, 0, 0);
    const uint32_t w01 = sign_bits16_(c);
    c = f32x16_{};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a10, b10, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a11, b11, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a12, b12, c, 0, 0, 0);
    const uint32_t w11 = sign_bits16_(c);
    const uint32_t own0 = h ? w01 : w00, own1 = h ? w11 : w10;  // column block h: this lane's particle
    const uint32_t par0 = h ? w00 : w01, par1 = h ? w10 : w11;  // the other: the partner's
    const uint32_t recv = (uint32_t) __shfl_xor((int) (par0 | (par1 << 16)), 32, kWave);
    const int so_ = 16 * h, sp_ = 16 - so_;
    const uint32_t lo = (own0 << so_) | ((recv & 0xffffu) << sp_);
    const uint32_t hi = (own1 << so_) | ((recv >> 16) << sp_);
    uint64_t mask = ((uint64_t) hi << 32) | lo;
    if (nt < kWave) mask &= (1ULL << nt) - 1;
    return mask;
}
__device__ __forceinline__ uint64_t mfma_dot_mask(const float4 *q, int nt, float tx, float ty, float tz, float negc,
                                                  float px, float py, float pz, float pc) {
    return mfma_dot_mask_f([q](int e) { return q[e]; }, nt, tx, ty, tz, negc, px, py, pz, pc);
}


// Symmetric form from the matrix cores (inside either support): the bit is
// set when t <= c_i or t + w_i <= Q_j, t = |o_j|^2 - 2 o_i.o_j as above, c_i
// the lane's own threshold, w_i = |o_i|^2 and Q_j the threshold of j's support
// (qs[jj], see dot_cut_j_).  Rows (o_j, w_j, 1, -Q_j) against the columns
// (-2 o_i, 1, -c_i, 0) give C1 = t - c_i; adding the columns (d_i, 1) of the
// last pair to C1 in place gives C2 = C1 + d_i - Q_j = t + w_i - Q_j with
// d_i = w_i + c_i.  Four MFMAs per 32 x 32 block (sixteen per tile) into one
// set of 16 accumulators, and a bit is set when either result's sign is.  The
// chained sum adds one rounding of the magnitudes already covered by the
// thresholds' room.  Layout and operands as in mfma_dot_mask; qs[jj] = Q_j,
// d = d_i of this lane (0 where negc = +inf, which admits nothing: C2 stays
// +inf) and pd that of lane l^32.
__device__ __forceinline__ uint32_t sym_block_(float a0, float a1, float a2, float b0, float b1, float b2, float b3) {
    f32x16_ c = f32x16_{};
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, c, 0, 0, 0);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, b2, c, 0, 0, 0);
    const uint32_t own = sign_bits16_(c);
    c = __builtin_amdgcn_mfma_f32_32x32x2f32(a2, b3, c, 0, 0, 0);
    return own | sign_bits16_(c);
}
// q(e) and qs(e) give entry e's (o_x, o_y, o_z, w) and Q.
template <class QF, class QSF>
__device__ __forceinline__ uint64_t mfma_dot_mask_sym(QF q, QSF qs, int nt, float tx, float ty, float tz, float negc,
                                                      float d, float px, float py, float pz, float pc, float pd) {
    const int lane = threadIdx.x, h = lane >> 5, m = lane & 31;
    // j operands: half 0 supplies k = 0, 2, 4 (o_x, o_z, 1), half 1 k = 1, 3, 5 (o_y, w_j, 0 for C1 and
    // -Q_j for C2: C1's k = 5 column entry is 0, so the row may carry -Q_j throughout)
    const int e = 16 * ((m >> 2) & 1) + 4 * (m >> 3) + (m & 3);
    const float4 q0 = q(e), q1 = q(32 + e);
    const float a00 = h ? q0.y : q0.x, a01 = h ? q0.w : q0.z, a02 = h ? -qs(e) : 1.f;
    const float a10 = h ? q1.y : q1.x, a11 = h ? q1.w : q1.z, a12 = h ? -qs(32 + e) : 1.f;
    // i operands: half 0 supplies k = 0, 2 (-2 o_x, -2 o_z) and k = 4 (-c_i; then d_i), half 1 k = 1, 3
    // (-2 o_y, 1) and k = 5 (0; then 1)
    const float b00 = h ? py : tx, b01 = h ? 1.f : tz, b02 = h ? 0.f : negc, b03 = h ? 1.f : d;
    const float b10 = h ? ty : px, b11 = h ? 1.f : pz, b12 = h ? 0.f : pc, b13 = h ? 1.f : pd;
    // one block at a time (scheduling barriers), the form measured in the
    // hydro kernels; interleaving the four blocks saves no registers there
    const uint32_t w00 = sym_block_(a00, a01, a02, b00, b01, b02, b03);
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t w10 = sym_block_(a10, a11, a12, b00, b01, b02, b03);
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t w01 = sym_block_(a00, a01, a02, b10, b11, b12, b13);
    __builtin_amdgcn_sched_barrier(0);
    const uint32_t w11 = sym_block_(a10, a11, a12, b10, b11, b12, b13);
    const uint32_t own0 = h ? w01 : w00, own1 = h ? w11 : w10;
    const uint32_t par0 = h ? w00 : w01, par1 = h ? w10 : w11;
    const uint32_t recv = (uint32_t) __shfl_xor((int) (par0 | (par1 << 16)), 32, kWave);
    const int so_ = 16 * h, sp_ = 16 - so_;
    const uint32_t lo = (own0 << so_) | ((recv & 0xffffu) << sp_);
    const uint32_t hi = (own1 << so_) | ((recv >> 16) << sp_);
    uint64_t mask = ((uint64_t) hi << 32) | lo;
    if (nt < kWave) mask &= (1ULL << nt) - 1;
    return mask;
}

// Whether group g's gather list still holds every particle within 2.5 h of
// each of its particles, h as just updated by a density pass: the group's
// search box (particle +- 2.5 h) must lie inside the box sbox[g] the list was
// built for, and for a compacted list every 2.5 h must stay within its radius
// crad[g] (see k_list_compact).  Sets *flag when not, so that the list is
// rebuilt before it is used again.  The box's extremes lie outside exactly
// when some particle's do, so each lane tests its own particle and one ballot
// decides.  Every lane of the wave takes part; `act` marks the lanes that hold
// a particle.
__device__ __forceinline__ void gather_cover_check(const double *sbox, const double *crad, int32_t *flag, int g,
                                                   bool act, double x, double y, double z, double h) {
    const double H = kSupport * h;
    const double *b = sbox + (int64_t) g * 6;
    bool bad = (x - H < b[0]) || (y - H < b[1]) || (z - H < b[2]) || (x + H > b[3]) || (y + H > b[4]) ||
               (z + H > b[5]);
    if (crad) bad = bad || (H > crad[g]);
    if (__ballot(act && bad) != 0 && threadIdx.x == 0) atomicOr(flag, 1);
}

__device__ __forceinline__ int64_t list_begin(const ListRef &L, int g) {
    return L.off ? L.off[g] : (int64_t) g * L.cap;
}

template <class T>
__device__ __forceinline__ T dmax(T a, T b) { return (a < b) ? b : a; }  // std::max
template <class T>
__device__ __forceinline__ T dmin(T a, T b) { return (b < a) ? b : a; }  // std::min

}  // namespace sphgpu
