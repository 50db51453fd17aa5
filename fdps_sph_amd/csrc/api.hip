// api.hip -- the extern "C" entry points of include/fdps_sph_gpu.h.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>

#include "../../include/fdps_sph_gpu.h"
#include "internal.hpp"

using namespace sphgpu;

// ------------------------------------------------------------------ fixed scratch (no-alloc EoS entry)
// Staging of sphgpu_eos_pressure_noalloc: material ids, density and energy in,
// pressure and sound speed out, for at most SPHGPU_NOALLOC_MAX_N particles, each
// field at a 256-byte aligned offset.  Every GPU the library runs on holds its
// own copy until the process ends.  In host code `scratch` names only a host
// stand-in: the device address comes from hipGetSymbolAddress, looked up once
// per device when a context is created (bind_scratch) and used everywhere else.
constexpr int64_t kNoAllocMaxN = SPHGPU_NOALLOC_MAX_N;
constexpr size_t kNoAllocAlign = 256;
constexpr size_t kNoAllocBytes = (size_t) kNoAllocMaxN * (sizeof(int32_t) + 4 * sizeof(double)) + 5 * kNoAllocAlign;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

thread_local std::string g_last_error;

template <class F>
int32_t guard(F &&f) {
    try {
        f();
        return SPHGPU_OK;
    } catch (const Error &e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc &) {
        g_last_error = "host allocation failed";
        return SPHGPU_ERR_HIP;
    } catch (const std::exception &e) {
        g_last_error = e.what();
        return SPHGPU_ERR_ARG;
    }
}

inline int nblk(int64_t n, int b) { return (int) ((n + b - 1) / b); }

// ------------------------------------------------------------------ AoS <-> SoA
constexpr int kMaxFields = 16;
struct UnpackArgs {
    const double *raw;
    int stride, nf;
    int fidx[kMaxFields];
    double *dst[kMaxFields];
};
__global__ void k_unpack(int64_t n, UnpackArgs a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double *r = a.raw + k * a.stride;
    for (int f = 0; f < a.nf; ++f) a.dst[f][k] = r[a.fidx[f]];
}
struct PackArgs {
    double *raw;
    int stride, nf;
    const double *src[kMaxFields];  // null -> fill value
    double fill[kMaxFields];
};
__global__ void k_pack(int64_t n, PackArgs a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    double *r = a.raw + k * a.stride;
    for (int f = 0; f < a.nf; ++f) r[f] = a.src[f] ? a.src[f][k] : a.fill[f];
}
__global__ void k_iota(int64_t n, int32_t *a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = (int32_t) k;
}

// sums of n_i, n_j, n_s and n_i*(n_j+n_s) over the groups of one launch
__global__ void k_list_counts(const int32_t *gcount, const int32_t *gstart, const int32_t *cnt, const int32_t *scnt,
                              long long *acc) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    long long v[4] = {0, 0, 0, 0};
    if (g < *gcount) {
        const long long ni = gstart[g + 1] - gstart[g];
        const long long nj = cnt[g];
        const long long ns = scnt ? scnt[g] : 0;
        v[0] = ni;
        v[1] = nj;
        v[2] = ns;
        v[3] = ni * (nj + ns);
    }
    __shared__ long long sh[4][256];
    for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o)
            for (int k = 0; k < 4; ++k) sh[k][threadIdx.x] += sh[k][threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 4; ++k) atomicAdd((unsigned long long *) &acc[k], (unsigned long long) sh[k][0]);
}

// ------------------------------------------------------------------ batches
enum Kind { KIND_DENS = 0, KIND_DRVT, KIND_HYDRO, KIND_GRAV };

struct Batch {
    int kind = -1;
    bool pending = false;
    int n_walk = 0;
    std::vector<int32_t> n_epi;
    int64_t total_i = 0;
    int res_stride = 0;
    DevBuf<uint8_t> d_in;
    DevBuf<double> soa;
    DevBuf<double> out;
    DevBuf<double> d_res;
    // group tables, borrowed from the staged copy in d_in
    int32_t *gstart = nullptr, *jcnt = nullptr, *scnt = nullptr;
    int64_t *joff = nullptr, *soff = nullptr;
    PinnedBuf<uint8_t> h_in;
    PinnedBuf<double> h_out;
    hipEvent_t ev = nullptr;
    ~Batch() {
        if (ev) (void) hipEventDestroy(ev);
    }
};

struct MatHost {
    int kind = 0;  // 0: rho-u table, 1: analytic Tillotson
    double till[10] = {};
    int nr = 0, nu = 0;
    DevBuf<double> rho, u, plane, full;
    bool sorted = false;
};

constexpr int kTags = 64;
constexpr int kStages = 8;

}  // namespace

// ------------------------------------------------------------------ simulation state
constexpr int kStatKinds = 6;

// one recorded step (the graph of one parity of the particle buffer sets)
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<int64_t> sig;  // the signature it was recorded with (step_sig)
    StepGraph() = default;
    StepGraph(const StepGraph &) = delete;
    StepGraph &operator=(const StepGraph &) = delete;
    ~StepGraph() { reset(); }
    void reset() {
        if (exec) (void) hipGraphExecDestroy(exec);
        exec = nullptr;
        sig.clear();
    }
};

struct SimState {
    sphgpu_sim_params prm{};
    TreeParams tp{};
    int64_t n = 0;
    int nsys = 1;
    std::vector<int32_t> sys_start_h;
    DevBuf<double> f[2][31];
    DevBuf<int32_t> mat[2], sys[2], ngb[2];
    DevBuf<int64_t> id[2];
    PArrays P{}, Palt{};
    Tree tree;
    Lists lsph, lgrav;
    DevBuf<int32_t> sched;  // per-XCD group counters of persistent force-kernel launches
    DevBuf<double> dt_sys, time_sys, dt_part;
    hipEvent_t ev[kStages + 1] = {};
    hipEvent_t xev[6] = {};  // cross-stream order of a force evaluation (no timing)
    double stage_ms[kStages] = {0};
    bool ready = false;
    // per-kernel statistics (density x4, derivative, hydro, gravity per step)
    bool stats = false;
    // kinds: 0 density, 1 derivative, 2 hydro, 3 gravity, 4 density pass run in
    // FP64 by the precision form, 5 v_sig maximum (dt)
    hipEvent_t kev[8][2] = {};
    int kkind[8] = {0, 0, 0, 0, 1, 2, 3, 5};
    double kms[kStatKinds] = {0};
    int64_t klaunch[kStatKinds] = {0};
    DevBuf<long long> kcount;  // [kStatKinds][4]: n_i, n_j, n_s, pairs
    DevBuf<long long> kcount_try;  // the same for the force evaluation in progress
    DevBuf<double> hsave;      // h at the start of a force evaluation (for its redo)
    DevBuf<HydroRec> hrec;     // j records of the hydro pass
    DevBuf<HydroRecF> hrecf;   // the same in the precision form
    DevBuf<VsigRec> vrec;      // j records of the dt kernel (vsig_split)
    DevBuf<int32_t> hidx, hcnt; // symmetric lists compacted for the hydro force terms (hydro_compact)
    DevBuf<double4> kpos, kvel; // j records of the density and derivative passes
    int64_t gather_builds = 0; // gather-list constructions since init
    int64_t force_evals = 0;   // force evaluations completed since init
    int kslot = 0;             // statistics slots used by the evaluation in progress
    PinnedBuf<int32_t> stash;  // error words and counts of the last evaluation (kSt*)
    // recorded steps, one per parity of the particle buffer sets (prm.graph)
    StepGraph graphs[2];
    std::vector<int64_t> direct_sig;  // signature of the last step enqueued directly
    int64_t graph_steps = 0, graph_captures = 0;
    ~SimState() {
        graphs[0].reset();
        graphs[1].reset();
        for (auto &e : ev)
            if (e) (void) hipEventDestroy(e);
        for (auto &e : xev)
            if (e) (void) hipEventDestroy(e);
        for (auto &p : kev)
            for (auto &e : p)
                if (e) (void) hipEventDestroy(e);
    }
};

struct sphgpu_ctx {
    int device = 0;
    int precision = SPHGPU_PREC_FP64;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // list walks that overlap the force kernels (default priority)
    MatHost mats[SPHGPU_MAX_MATERIALS];
    DevBuf<MatDev> d_mats;
    DevBuf<int32_t> iota;
    int64_t iota_n = 0;
    Batch batches[kTags];
    std::unique_ptr<SimState> sim;
    // EoS staging
    DevBuf<double> e_in1, e_in2, e_out1, e_out2;
    DevBuf<int32_t> e_mat;
    void *scratch = nullptr;  // this device's address of the fixed no-alloc scratch
};

namespace {

void set_device(sphgpu_ctx *c) { SPH_HIP(hipSetDevice(c->device)); }

void check_ctx(sphgpu_ctx *c) {
    if (!c) throw Error(SPHGPU_ERR_ARG, "null context");
    set_device(c);
}

// device address of `scratch` on each device ordinal, looked up once
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch[kMaxDevices] = {};

void *bind_scratch(int device) {  // `device` is current
    if (device < 0 || device >= kMaxDevices) throw Error(SPHGPU_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            throw Error(SPHGPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    return g_scratch[device];
}

void upload_mats(sphgpu_ctx *c) {
    MatDev h[SPHGPU_MAX_MATERIALS];
    for (int m = 0; m < SPHGPU_MAX_MATERIALS; ++m) {
        const MatHost &t = c->mats[m];
        h[m].kind = t.kind;
        for (int k = 0; k < 10; ++k) h[m].till[k] = t.till[k];
        h[m].nr = t.nr;
        h[m].nu = t.nu;
        h[m].rho = t.rho.get();
        h[m].u = t.u.get();
        h[m].plane = t.plane.get();
        const int64_t N = (int64_t) t.nr * t.nu;
        h[m].full_rho = t.full.get();
        h[m].full_S = t.full.get() ? t.full.get() + N : nullptr;
        h[m].full_u = t.full.get() ? t.full.get() + 2 * N : nullptr;
        h[m].sorted_rho = t.sorted ? 1 : 0;
    }
    c->d_mats.ensure(SPHGPU_MAX_MATERIALS);
    SPH_HIP(hipMemcpy(c->d_mats.get(), h, sizeof(h), hipMemcpyHostToDevice));
}

void set_table(sphgpu_ctx *c, int material, int nr, int nu, const double *cells) {
    if (material < 0 || material >= SPHGPU_MAX_MATERIALS) throw Error(SPHGPU_ERR_ARG, "material id out of range");
    if (nr < 1 || nu < 1 || !cells) throw Error(SPHGPU_ERR_EOS, "empty EoS table");
    const int64_t N = (int64_t) nr * nu;
    std::vector<double> rho(nr), u(nu), plane(6 * N), full(3 * N);
    for (int r = 0; r < nr; ++r) rho[r] = cells[((int64_t) r * nu) * 6 + 0];  // column 0, field 0
    for (int q = 0; q < nu; ++q) u[q] = cells[(int64_t) q * 6 + 1];            // row 0, field 1
    for (int f = 0; f < 6; ++f)
        for (int64_t k = 0; k < N; ++k) plane[f * N + k] = cells[k * 6 + f];
    bool sorted = true;
    for (int64_t k = 0; k < N; ++k) {
        full[k] = cells[k * 6 + 0];
        full[N + k] = cells[k * 6 + 5];
        full[2 * N + k] = cells[k * 6 + 1];
        if (k > 0 && full[k] < full[k - 1]) sorted = false;
    }
    MatHost &t = c->mats[material];
    t.kind = 0;
    t.nr = nr;
    t.nu = nu;
    t.sorted = sorted;
    SPH_HIP(hipMemcpy(t.rho.ensure(nr), rho.data(), sizeof(double) * nr, hipMemcpyHostToDevice));
    SPH_HIP(hipMemcpy(t.u.ensure(nu), u.data(), sizeof(double) * nu, hipMemcpyHostToDevice));
    SPH_HIP(hipMemcpy(t.plane.ensure(6 * N), plane.data(), sizeof(double) * 6 * N, hipMemcpyHostToDevice));
    SPH_HIP(hipMemcpy(t.full.ensure(3 * N), full.data(), sizeof(double) * 3 * N, hipMemcpyHostToDevice));
    upload_mats(c);
}

// Table reader of src/EoS.h:290-385 (header "# nrow ncol", '#' lines skipped,
// one cell of up to 6 numbers per line).
bool mat_loaded(const sphgpu_ctx *c, int m) {
    return m >= 0 && m < SPHGPU_MAX_MATERIALS && (c->mats[m].kind == 1 || c->mats[m].nr > 0);
}

void set_tillotson(sphgpu_ctx *c, int material, const double *prm) {
    if (material < 0 || material >= SPHGPU_MAX_MATERIALS) throw Error(SPHGPU_ERR_ARG, "material id out of range");
    if (!prm) throw Error(SPHGPU_ERR_ARG, "null Tillotson parameters");
    for (int k = 0; k < 10; ++k)
        if (!std::isfinite(prm[k])) throw Error(SPHGPU_ERR_EOS, "non-finite Tillotson parameter");
    if (!(prm[0] > 0 && prm[1] > 0 && prm[2] >= 0 && prm[3] > prm[2]))
        throw Error(SPHGPU_ERR_EOS, "Tillotson parameters need rho0 > 0, u0 > 0 and 0 <= u_iv < u_cv");
    MatHost &t = c->mats[material];
    t.kind = 1;
    for (int k = 0; k < 10; ++k) t.till[k] = prm[k];
    t.nr = t.nu = 0;
    t.sorted = false;
    upload_mats(c);
}

void load_table_file(sphgpu_ctx *c, int material, const char *path) {
    if (!path) throw Error(SPHGPU_ERR_ARG, "null path");
    FILE *fp = fopen(path, "r");
    if (!fp) throw Error(SPHGPU_ERR_EOS, std::string("cannot open EoS table ") + path);
    char line[4096];
    int nr = 0, nc = 0;
    bool header = false;
    std::vector<double> cells;
    int64_t ncell = 0;
    while (fgets(line, sizeof line, fp)) {
        if (!header) {
            char tag[64];
            if (sscanf(line, "%63s %d %d", tag, &nr, &nc) != 3 || nr <= 0 || nc <= 0) {
                fclose(fp);
                throw Error(SPHGPU_ERR_EOS, std::string("bad EoS table header in ") + path);
            }
            header = true;
            cells.assign((size_t) nr * nc * 6, 0.0);
            continue;
        }
        if (line[0] == '#') continue;
        char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\n' || *p == '\r' || *p == '\0') continue;
        if (ncell >= (int64_t) nr * nc) break;
        double *cell = cells.data() + ncell * 6;
        for (int f = 0; f < 6; ++f) {
            char *end;
            const double v = strtod(p, &end);
            if (end == p) break;
            cell[f] = v;
            p = end;
        }
        ++ncell;
    }
    fclose(fp);
    if (!header || ncell != (int64_t) nr * nc)
        throw Error(SPHGPU_ERR_EOS, std::string("EoS table ") + path + " has fewer cells than its header");
    set_table(c, material, nr, nc, cells.data());
}

const int32_t *ensure_iota(sphgpu_ctx *c, int64_t n) {
    if (n > c->iota_n) {
        c->iota.ensure(n);
        hipLaunchKernelGGL(k_iota, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, c->iota.get());
        SPH_LAUNCH_CHECK();
        c->iota_n = n;
    }
    return c->iota.get();
}

// i-groups of at most 64 particles per walk; each shares the walk's j range
struct GroupPlan {
    std::vector<int32_t> gstart;
    std::vector<int64_t> joff, soff;
    std::vector<int32_t> jcnt, scnt;
};

GroupPlan plan_groups(int n_walk, const int32_t *n_epi, const int32_t *n_epj, const int32_t *n_spj) {
    GroupPlan g;
    int64_t i0 = 0, j0 = 0, s0 = 0;
    g.gstart.push_back(0);
    for (int w = 0; w < n_walk; ++w) {
        const int ni = n_epi[w], nj = n_epj[w], ns = n_spj ? n_spj[w] : 0;
        for (int k = 0; k < ni; k += kWave) {
            const int take = std::min(kWave, ni - k);
            g.gstart.push_back((int32_t) (i0 + k + take));
            g.joff.push_back(j0);
            g.jcnt.push_back(nj);
            g.soff.push_back(s0);
            g.scnt.push_back(ns);
        }
        i0 += ni;
        j0 += nj;
        s0 += ns;
    }
    return g;
}

template <class T>
void check_walks(int n_walk, const T *const arr[], const int32_t n[], const char *what) {
    if (n_walk < 0) throw Error(SPHGPU_ERR_ARG, "negative n_walk");
    if (n_walk > 0 && (!arr || !n)) throw Error(SPHGPU_ERR_ARG, std::string("null ") + what + " arrays");
    for (int w = 0; w < n_walk; ++w) {
        if (n[w] < 0) throw Error(SPHGPU_ERR_ARG, std::string("negative ") + what + " count");
        if (n[w] > 0 && !arr[w]) throw Error(SPHGPU_ERR_ARG, std::string("null ") + what + " pointer");
    }
}

int64_t sum_counts(int n_walk, const int32_t *n) {
    int64_t s = 0;
    for (int w = 0; w < n_walk; ++w) s += n[w];
    if (s > INT32_MAX) throw Error(SPHGPU_ERR_CAPACITY, "batch exceeds 2^31 particles");
    return s;
}

// Stage the walks' AoS arrays (and the group plan) into the batch's device buffers.
struct Staged {
    const double *epi, *epj, *spj;
    int64_t ni, nj, ns;
};

template <class EI, class EJ, class SJ>
Staged stage(sphgpu_ctx *c, Batch &b, int n_walk, const EI *const epi[], const int32_t n_epi[], const EJ *const epj[],
             const int32_t n_epj[], const SJ *const spj[], const int32_t n_spj[], const GroupPlan &gp) {
    Staged st;
    st.ni = sum_counts(n_walk, n_epi);
    st.nj = sum_counts(n_walk, n_epj);
    st.ns = spj ? sum_counts(n_walk, n_spj) : 0;
    const size_t bi = st.ni * sizeof(EI), bj = st.nj * sizeof(EJ), bs = st.ns * sizeof(SJ);
    const size_t ng = gp.jcnt.size();
    const size_t bg = (ng + 1) * sizeof(int32_t) + ng * (2 * sizeof(int64_t) + 2 * sizeof(int32_t));
    const size_t total = bi + bj + bs + bg + 64;
    // the previous batch on this tag must be finished before its staging is reused
    if (b.ev) SPH_HIP(hipEventSynchronize(b.ev));
    uint8_t *h = b.h_in.ensure(total);
    size_t o = 0;
    for (int w = 0; w < n_walk; ++w) {
        memcpy(h + o, epi[w], n_epi[w] * sizeof(EI));
        o += n_epi[w] * sizeof(EI);
    }
    for (int w = 0; w < n_walk; ++w) {
        memcpy(h + o, epj[w], n_epj[w] * sizeof(EJ));
        o += n_epj[w] * sizeof(EJ);
    }
    if (spj)
        for (int w = 0; w < n_walk; ++w) {
            if (n_spj[w] > 0) memcpy(h + o, spj[w], n_spj[w] * sizeof(SJ));
            o += n_spj[w] * sizeof(SJ);
        }
    o = (o + 7) & ~size_t(7);
    const size_t og = o;
    memcpy(h + o, gp.gstart.data(), (ng + 1) * sizeof(int32_t));
    o += (ng + 1) * sizeof(int32_t);
    o = (o + 7) & ~size_t(7);
    const size_t o_joff = o;
    memcpy(h + o, gp.joff.data(), ng * sizeof(int64_t));
    o += ng * sizeof(int64_t);
    const size_t o_soff = o;
    memcpy(h + o, gp.soff.data(), ng * sizeof(int64_t));
    o += ng * sizeof(int64_t);
    const size_t o_jcnt = o;
    memcpy(h + o, gp.jcnt.data(), ng * sizeof(int32_t));
    o += ng * sizeof(int32_t);
    const size_t o_scnt = o;
    memcpy(h + o, gp.scnt.data(), ng * sizeof(int32_t));
    o += ng * sizeof(int32_t);
    uint8_t *d = b.d_in.ensure(o);
    SPH_HIP(hipMemcpyAsync(d, h, o, hipMemcpyHostToDevice, c->stream));
    st.epi = (const double *) d;
    st.epj = (const double *) (d + bi);
    st.spj = (const double *) (d + bi + bj);
    // group tables are used in place from the staging copy
    b.gstart = (int32_t *) (d + og);
    b.joff = (int64_t *) (d + o_joff);
    b.soff = (int64_t *) (d + o_soff);
    b.jcnt = (int32_t *) (d + o_jcnt);
    b.scnt = (int32_t *) (d + o_scnt);
    return st;
}

void unpack(sphgpu_ctx *c, const double *raw, int64_t n, int stride, std::initializer_list<int> fields, double *soa,
            int64_t soa_stride, std::vector<double *> &out) {
    UnpackArgs a{};
    a.raw = raw;
    a.stride = stride;
    a.nf = 0;
    out.clear();
    for (int f : fields) {
        a.fidx[a.nf] = f;
        a.dst[a.nf] = soa + (int64_t) a.nf * soa_stride;
        out.push_back(a.dst[a.nf]);
        ++a.nf;
    }
    if (n > 0) {
        hipLaunchKernelGGL(k_unpack, dim3(nblk(n, 256)), dim3(256), 0, c->stream, n, a);
        SPH_LAUNCH_CHECK();
    }
}

void finish_batch(sphgpu_ctx *c, Batch &b, int kind, int n_walk, const int32_t n_epi[], int64_t ni,
                  std::initializer_list<const double *> res_fields, std::initializer_list<double> fills) {
    PackArgs pa{};
    pa.nf = 0;
    auto fi = fills.begin();
    for (const double *f : res_fields) {
        pa.src[pa.nf] = f;
        pa.fill[pa.nf] = *fi++;
        ++pa.nf;
    }
    pa.stride = pa.nf;
    pa.raw = b.d_res.ensure(std::max<int64_t>(1, ni * pa.nf));
    if (ni > 0) {
        hipLaunchKernelGGL(k_pack, dim3(nblk(ni, 256)), dim3(256), 0, c->stream, ni, pa);
        SPH_LAUNCH_CHECK();
    }
    double *h = b.h_out.ensure(std::max<int64_t>(1, ni * pa.nf));
    if (ni > 0) SPH_HIP(hipMemcpyAsync(h, pa.raw, sizeof(double) * ni * pa.nf, hipMemcpyDeviceToHost, c->stream));
    if (!b.ev) SPH_HIP(hipEventCreateWithFlags(&b.ev, hipEventDisableTiming));
    SPH_HIP(hipEventRecord(b.ev, c->stream));
    b.kind = kind;
    b.pending = true;
    b.n_walk = n_walk;
    b.n_epi.assign(n_epi, n_epi + n_walk);
    b.total_i = ni;
    b.res_stride = pa.nf;
}

Batch &batch_for(sphgpu_ctx *c, int tag) {
    if (tag < 0 || tag >= kTags) throw Error(SPHGPU_ERR_TAG, "tag out of range [0, 64)");
    return c->batches[tag];
}

template <class R>
void retrieve(sphgpu_ctx *c, int kind, int tag, int n_walk, const int32_t n_epi[], R *const res[]) {
    check_ctx(c);
    Batch &b = batch_for(c, tag);
    if (!b.pending || b.kind != kind) throw Error(SPHGPU_ERR_TAG, "no batch of this kind dispatched on this tag");
    if (n_walk != b.n_walk) throw Error(SPHGPU_ERR_ARG, "n_walk differs from the dispatched batch");
    if (n_walk > 0 && !n_epi) throw Error(SPHGPU_ERR_ARG, "null n_epi array");
    for (int w = 0; w < n_walk; ++w)
        if (n_epi[w] != b.n_epi[w]) throw Error(SPHGPU_ERR_ARG, "n_epi differs from the dispatched batch");
    if (n_walk > 0 && !res) throw Error(SPHGPU_ERR_ARG, "null res array");
    SPH_HIP(hipEventSynchronize(b.ev));
    static_assert(sizeof(R) % sizeof(double) == 0, "result layout");
    const int stride = (int) (sizeof(R) / sizeof(double));
    if (stride != b.res_stride) throw Error(SPHGPU_ERR_STATE, "result layout mismatch");
    int64_t o = 0;
    for (int w = 0; w < n_walk; ++w) {
        if (n_epi[w] > 0) {
            if (!res[w]) throw Error(SPHGPU_ERR_ARG, "null res pointer");
            memcpy(res[w], b.h_out.p + o * stride, sizeof(R) * n_epi[w]);
        }
        o += n_epi[w];
    }
    b.pending = false;
}

ListRef list_ref(const int32_t *idx, const int64_t *off, const int32_t *cnt) { return ListRef{idx, off, cnt, 0}; }

// ------------------------------------------------------------------ simulation helpers
__global__ void k_fill_ll(int64_t n, long long *a) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) a[k] = 0;
}
__global__ void k_add_ll(int64_t n, const long long *src, long long *dst) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] += src[k];
}

void sim_bind(SimState &S) {
    auto bind = [&](int k, PArrays &P) {
        double **fp[31] = {&P.x, &P.y, &P.z, &P.vx, &P.vy, &P.vz, &P.ax, &P.ay, &P.az, &P.gx, &P.gy,
                           &P.gz, &P.pot, &P.m, &P.h, &P.rho, &P.u, &P.p, &P.cs, &P.divv, &P.rotx, &P.roty,
                           &P.rotz, &P.gs, &P.bal, &P.du, &P.dti, &P.vhx, &P.vhy, &P.vhz, &P.uh};
        for (int f = 0; f < 31; ++f) *fp[f] = S.f[k][f].ensure(S.n);
        P.mat = S.mat[k].ensure(S.n);
        P.sys = S.sys[k].ensure(S.n);
        P.ngb = S.ngb[k].ensure(S.n);
        P.id = S.id[k].ensure(S.n);
    };
    bind(0, S.P);
    bind(1, S.Palt);
}

// The state a recorded step depends on beyond its buffers' contents: the
// allocation epoch (every buffer address), precision, kernel options and
// every capacity.  A recorded step is replayed only while this is unchanged.
std::vector<int64_t> step_sig(const sphgpu_ctx *c, const SimState &S) {
    std::vector<int64_t> v;
    v.push_back((int64_t) alloc_epoch());
    v.push_back(c->precision);
    v.push_back(S.prm.persistent);
    v.push_back(S.prm.eos_density_fp64);
    v.push_back(S.prm.vsig_split);
    v.push_back(S.prm.hydro_compact);
    v.push_back(S.tree.gcap);
    for (int l = 0; l < S.tree.nlevels; ++l) v.push_back(S.tree.cap_lv[l]);
    for (const Lists *L : {&S.lsph, &S.lgrav}) {
        v.push_back(L->cap);
        v.push_back(L->scap);
        v.push_back(L->fcap);
        v.push_back(L->flds);
        v.push_back(L->waves);
    }
    return v;
}

// Words the host reads after a force evaluation (SimState::stash): the
// gather lists' overflow bits and construction counter (taken before the
// symmetric lists reuse their buffers), the symmetric and the gravity lists'
// overflow words, and the tree's counts (kInfo*).
constexpr int kStGather = 0, kStSph = 2, kStGrav = 4, kStInfo = 6, kStWords = kStInfo + kInfoWords;

// Copies up to four word ranges, one after another, into host memory: an
// ordinary kernel, so that a recorded step carries it.
struct StashArgs {
    const int32_t *src[4];
    int n[4];
    int32_t *dst;
};
__global__ void k_stash(StashArgs a) {
    int o = 0;
    for (int k = 0; k < 4; ++k) {
        if (a.src[k])
            for (int i = threadIdx.x; i < a.n[k]; i += blockDim.x) a.dst[o + i] = a.src[k][i];
        o += a.n[k];
    }
}
void launch_stash(StashArgs a, hipStream_t s) {
    hipLaunchKernelGGL(k_stash, dim3(1), dim3(64), 0, s, a);
    SPH_LAUNCH_CHECK();
}

__global__ void k_copy_d(int64_t n, const double *src, double *dst) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = src[k];
}
void launch_copy_d(int64_t n, const double *src, double *dst, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_copy_d, dim3(nblk(n, 256)), dim3(256), 0, s, n, src, dst);
    SPH_LAUNCH_CHECK();
}

// The error words the last kernels of a step check (StepGuard): the tree's
// and both list kinds' overflow bits.
StepGuard step_guard(SimState &S) {
    StepGuard g;
    g.w[0] = S.tree.info.get() + kInfoErr;
    g.w[1] = S.lsph.err.get();
    g.w[2] = S.lgrav.err.get();
    return g;
}

// One force evaluation (src/main.cpp:125-146), enqueued without any wait for
// the host: the sort and tree build (launches sized by capacities, counts on
// the device), the lists (each ORs capacity overflows into its error bits;
// whether a gather list must be rebuilt is decided on the device), the
// density passes, EoS, derivative, gravity and hydro, and finally the error
// words and counts copied to S.stash.  `restore_h` first puts back the h that
// a failed evaluation of the same state started from (the only input an
// evaluation overwrites: the density passes update h); the sort of the
// sorted state is then the identity.
//
// Two list walks run on a second stream, beside force kernels they do not feed:
// the walks wait on dependent loads and leave the VALU idle, the kernels are
// VALU-bound.  The gravity lists need only the tree (the walk reads no h), so
// they are built while the density passes run (after the first gather lists,
// which the first pass waits for); the symmetric lists of the hydro
// pass need the final h and replace the gather lists, so they are built after
// the derivative pass, while the gravity kernel runs.  With kernel statistics
// on, everything runs in this order on one stream, so that each kernel's timing
// is its own.
void forces_enqueue(sphgpu_ctx *c, SimState &S, bool timed, bool restore_h) {
    hipStream_t s = c->stream;
    hipStream_t s2 = S.stats ? s : c->stream2;
    auto signal = [&](hipEvent_t e, hipStream_t from) {  // order later work on the other stream after `from`'s so far
        if (s2 != s) SPH_HIP(hipEventRecord(e, from));
    };
    auto await = [&](hipStream_t on, hipEvent_t e) {
        if (s2 != s) SPH_HIP(hipStreamWaitEvent(on, e, 0));
    };
    const int64_t n = S.n;
    const bool mixed = c->precision == SPHGPU_PREC_MIXED;
    auto rec_ev = [&](int k) {
        if (timed) SPH_HIP(hipEventRecord(S.ev[k], s));
    };
    double *hsave = S.hsave.ensure(n);
    if (restore_h) launch_copy_d(n, hsave, S.P.h, s);
    tree_sort_and_build(S.tree, S.tp, n, S.P, S.Palt, s);
    rec_ev(2);
    PArrays &P = S.P;
    const int G = (int) S.tree.gcap;  // launches are sized by the capacity, the kernels read the count
    // a group of more than 64 particles runs as one work item per wave of them
    const int chunks = (S.tp.n_group + kWave - 1) / kWave;
    const int32_t *gs = S.tree.gstart.get();
    const int32_t *gcount = S.tree.info.get() + kInfoG;
    double4 *kpos = S.kpos.ensure(n), *kvel = S.kvel.ensure(n);  // j records of the density / derivative passes
    launch_pack_kin(n, P.x, P.y, P.z, P.m, P.vx, P.vy, P.vz, kpos, kvel, s, P.h, hsave);  // h saved beside
    Lists &L = S.lsph, &Lg = S.lgrav;
    int32_t *stash = S.stash.p;
    // persistent launches of the force kernels named by the bitmask (1 density,
    // 2 derivative, 4 hydro, 8 gravity); their counters live on the main stream,
    // where every force kernel runs, one launch after another
    int32_t *sched_buf = S.prm.persistent ? S.sched.ensure(8) : nullptr;
    auto sched_for = [&](int bit) { return (S.prm.persistent & bit) ? sched_buf : nullptr; };
    auto lref = [&]() { return ListRef{L.idx.get(), nullptr, L.cnt.get(), L.cap}; };
    S.kslot = 0;
    if (S.stats) {
        S.kcount_try.ensure(4 * kStatKinds);
        hipLaunchKernelGGL(k_fill_ll, dim3(1), dim3(64), 0, s, (int64_t) 4 * kStatKinds, S.kcount_try.get());
        SPH_LAUNCH_CHECK();
    }
    auto kbegin = [&](const Lists *Lp, bool spj, int kind, const int32_t *cnt = nullptr) {
        if (!S.stats) return;
        S.kkind[S.kslot] = kind;
        hipLaunchKernelGGL(k_list_counts, dim3(nblk(G, 256)), dim3(256), 0, s, gcount, gs, cnt ? cnt : Lp->cnt.get(),
                           spj ? Lp->scnt.get() : nullptr, S.kcount_try.get() + 4 * kind);
        SPH_LAUNCH_CHECK();
        SPH_HIP(hipEventRecord(S.kev[S.kslot][0], s));
    };
    auto kend = [&]() {
        if (!S.stats) return;
        SPH_HIP(hipEventRecord(S.kev[S.kslot][1], s));
        ++S.kslot;
    };
    // per density pass: set when some group's gather list no longer covers the
    // supports of the h that pass computed (the list is then rebuilt first);
    // cleared with both lists' words in one launch
    int32_t *gflag = L.flag.ensure(4);
    lists_reset_all(L, Lg, gflag, 4, s);
    enqueue_lists(S.tree, S.tp, P, 0, L, kGatherMargin, s, nullptr, true);
    enqueue_gather_compact(S.tree, P, L, kGatherMargin, s, nullptr, kpos);
    // the gravity walk beside the density passes (started beside the gather
    // walk instead, the step measured the same: two walks together take as
    // long as one after the other)
    signal(S.xev[0], s);  // tree, h restored, counters cleared, first gather lists
    await(s2, S.xev[0]);
    enqueue_lists(S.tree, S.tp, P, 2, Lg, 1.0, s2, nullptr, true);
    signal(S.xev[1], s2);  // gravity lists
    // 4 density passes, each over lists that cover the current h
    // (src/main.cpp:138-140).  One gather list, built with a margin on h,
    // serves every pass (and the derivative pass) while each group's
    // search box still fits inside the box it was built for; each pass checks
    // that for the h it computes.
    for (int loop = 0; loop <= 3; ++loop) {
        if (loop > 0) {
            enqueue_lists(S.tree, S.tp, P, 0, L, kGatherMargin, s, gflag + loop - 1);
            enqueue_gather_compact(S.tree, P, L, kGatherMargin, s, gflag + loop - 1, kpos);
        }
        DensArgs a{G, gs, P.x, P.y, P.z, P.m, P.h, P.m, P.x, P.y, P.z, lref(), P.rho, P.h, P.ngb, kpos};
        a.sched = sched_for(1);
        a.gcount = gcount;
        a.ngroups = G * chunks;
        a.chunks = chunks;
        a.sbox = L.sbox.get();
        a.crad = L.crad.get();
        a.flag = gflag + loop;
        // the density of the last pass feeds the EoS; the precision form
        // keeps it FP64 unless asked otherwise (the earlier passes only steer h)
        const bool d64 = !mixed || (loop == 3 && S.prm.eos_density_fp64);
        kbegin(&L, false, (mixed && d64) ? 4 : 0);
        if (d64) launch_dens_fp64(a, s);
        else launch_dens_mixed(a, s);
        kend();
    }
    rec_ev(3);
    // h is final: the tree's largest supports for the symmetric lists, beside
    // the EoS and the derivative pass
    signal(S.xev[4], s);
    await(s2, S.xev[4]);
    tree_update_maxH(S.tree, P.h, s2);
    launch_eos_pressure(n, P.mat, P.rho, P.u, c->d_mats.get(), SPHGPU_MAX_MATERIALS, P.p, P.cs, s);
    rec_ev(4);
    enqueue_lists(S.tree, S.tp, P, 0, L, kGatherMargin, s, gflag + 3);
    enqueue_gather_compact(S.tree, P, L, kGatherMargin, s, gflag + 3, kpos);
    {
        DrvtArgs a{G, gs, P.x, P.y, P.z, P.vx, P.vy, P.vz, P.h, P.rho, P.cs, P.m, P.x, P.y, P.z, P.vx, P.vy, P.vz,
                   lref(), P.divv, P.rotx, P.roty, P.rotz, P.gs, P.bal, kpos, kvel};
        a.sched = sched_for(2);
        a.gcount = gcount;
        a.ngroups = G * chunks;
        a.chunks = chunks;
        kbegin(&L, false, 1);
        if (mixed) launch_drvt_mixed(a, s);
        else launch_drvt_fp64(a, s);
        kend();
    }
    rec_ev(5);
    // the hydro pass's j records, packed here rather than on the second stream:
    // there they would wait behind the gravity kernel's workgroups
    // the dt part of the hydro pass (v_sig_max over each list, FP64) as a kernel
    // of its own beside the gravity kernel, when asked for (vsig_split bit of
    // the precision form); the hydro kernel then forms the force terms only
    const bool split = (S.prm.vsig_split >> (mixed ? 1 : 0)) & 1;
    HydroRec *hrec = nullptr;
    HydroRecF *hrecf = nullptr;
    VsigRec *vrec = split ? S.vrec.ensure(n) : nullptr;
    if (mixed) {
        hrecf = S.hrecf.ensure(n);
        launch_pack_hydro_f(n, P.x, P.y, P.z, P.vx, P.vy, P.vz, P.m, P.rho, P.cs, P.bal, P.gs, P.p, P.h, hrecf, s,
                            vrec);
    } else {
        hrec = S.hrec.ensure(n);
        launch_pack_hydro(n, P.x, P.y, P.z, P.vx, P.vy, P.vz, P.m, P.rho, P.cs, P.bal, P.gs, P.p, P.h, hrec, s, vrec);
    }
    signal(S.xev[2], s);  // the derivative pass is done with the gather lists; records packed
    await(s2, S.xev[2]);
    // gather-list bits and construction counter, before the symmetric lists reuse L
    launch_stash(StashArgs{{L.err.get(), nullptr, nullptr, nullptr}, {2, 0, 0, 0}, stash + kStGather}, s2);
    enqueue_lists(S.tree, S.tp, P, 1, L, 1.0, s2);
    // with dt in a kernel of its own, the force terms read the symmetric lists
    // compacted to the entries within reach of either support; dt still runs
    // over the whole lists, after the compaction on this stream, beside the
    // gravity and hydro kernels
    const bool hcompact = split && S.prm.hydro_compact && S.tree.max_group <= kWave;
    int32_t *hidx = nullptr, *hcnt = nullptr;
    if (hcompact) {
        hidx = S.hidx.ensure((size_t) G * (size_t) L.cap);
        hcnt = S.hcnt.ensure((size_t) std::max(G, 1));
        enqueue_sym_compact(S.tree, P, L, kpos, hidx, hcnt, s2);
    }
    signal(S.xev[3], s2);  // symmetric lists (and their compacted form)
    auto hydro_args = [&]() {
        HydroArgs a{G, gs, P.x, P.y, P.z, P.vx, P.vy, P.vz, P.h, P.rho, P.p, P.cs, P.gs, P.bal,
                    P.x, P.y, P.z, P.vx, P.vy, P.vz, P.h, P.rho, P.p, P.cs, P.gs, P.bal, P.m,
                    lref(), P.ax, P.ay, P.az, P.du, P.dti, hrec};
        a.jrecf = hrecf;
        a.gcount = gcount;
        a.ngroups = G * chunks;
        a.chunks = chunks;
        a.pass1 = split ? 0 : 1;
        return a;
    };
    if (split) {
        kbegin(&L, false, 5);
        HydroArgs va = hydro_args();
        va.vrec = vrec;
        launch_vsig(va, s2);
        kend();
        signal(S.xev[5], s2);  // dt
    }
    await(s, S.xev[1]);
    {
        const Tree &t = S.tree;
        GravArgs a{G, gs, P.x, P.y, P.z, nullptr, P.h, P.x, P.y, P.z, P.m,
                   ListRef{Lg.idx.get(), nullptr, Lg.cnt.get(), Lg.cap},
                   t.comx.get(), t.comy.get(), t.comz.get(), t.mass.get(),
                   ListRef{Lg.sidx.get(), nullptr, Lg.scnt.get(), Lg.scap},
                   P.gx, P.gy, P.gz, P.pot};
        a.sched = sched_for(8);
        a.gcount = gcount;
        a.ngroups = G * chunks;
        a.chunks = chunks;
        a.jrec = kpos;
        a.srec = t.nrec.get();
        kbegin(&Lg, true, 3);
        if (mixed) launch_grav_mixed(a, s);
        else launch_grav_fp64(a, s);
        kend();
    }
    rec_ev(6);
    {
        await(s, S.xev[3]);
        HydroArgs a = hydro_args();
        if (hcompact) a.L = ListRef{hidx, nullptr, hcnt, L.cap};
        a.sched = sched_for(4);
        kbegin(&L, false, 2, hcompact ? hcnt : nullptr);
        if (mixed) launch_hydro_mixed(a, s);
        else launch_hydro_fp64(a, s);
        kend();
    }
    if (split) await(s, S.xev[5]);
    rec_ev(7);
    launch_stash(StashArgs{{L.err.get(), Lg.err.get(), S.tree.info.get(), nullptr}, {2, 2, kInfoWords, 0},
                           stash + kStSph},
                 s);
    S.direct_sig = step_sig(c, S);
}

// After the stream has drained: reads the stash of the evaluation just run,
// grows every capacity it exceeded (and the tree's that it came close to) and
// returns true when nothing overflowed.  On success the tree's counts are
// applied on the host and the evaluation is counted.
bool forces_check(SimState &S) {
    const int32_t *h = S.stash.p;
    const int32_t *info = h + kStInfo;
    tree_grow_caps(S.tree, info);
    if (info[kInfoErr]) return false;  // no groups were formed: the lists hold nothing to learn from
    const bool ok_s = lists_grow_bits(S.lsph, h[kStGather] | h[kStSph]);
    const bool ok_g = lists_grow_bits(S.lgrav, h[kStGrav]);
    if (!(ok_s && ok_g)) return false;
    tree_apply_info(S.tree, info);
    S.gather_builds += h[kStGather + 1];
    ++S.force_evals;
    return true;
}

// kernel statistics of a successful evaluation (statistics on)
void forces_stats(sphgpu_ctx *c, SimState &S) {
    if (!S.stats) return;
    hipStream_t s = c->stream;
    hipLaunchKernelGGL(k_add_ll, dim3(1), dim3(64), 0, s, (int64_t) 4 * kStatKinds, S.kcount_try.get(),
                       S.kcount.get());
    SPH_LAUNCH_CHECK();
    for (int k = 0; k < S.kslot; ++k) {
        float ms = 0;
        SPH_HIP(hipEventElapsedTime(&ms, S.kev[k][0], S.kev[k][1]));
        S.kms[S.kkind[k]] += ms;
        S.klaunch[S.kkind[k]] += 1;
    }
}

// A force evaluation run to completion: enqueued, waited for once, and run
// again with grown capacities while something overflows.  `restore_h`: the
// state is that of a failed evaluation (see forces_enqueue).
void sim_forces(sphgpu_ctx *c, SimState &S, bool timed, bool restore_h = false) {
    for (int attempt = 0;; ++attempt) {
        if (attempt >= 12) throw Error(-2, "interaction list capacity could not be grown enough");
        forces_enqueue(c, S, timed, restore_h || attempt > 0);
        SPH_HIP(hipStreamSynchronize(c->stream));  // the one wait of the evaluation
        if (forces_check(S)) break;
    }
    forces_stats(c, S);
}

void sim_dt(sphgpu_ctx *c, SimState &S) {
    launch_dt_reduce(S.nsys, S.sys_start_h.data(), S.tree.sys_start.get(), S.n, S.P.dti, S.dt_part.get(),
                     S.dt_sys.get(), step_guard(S), c->stream);
}

int sim_parity(const SimState &S) { return S.P.x == S.f[0][0].get() ? 0 : 1; }

// the work of one step (src/main.cpp:125-173) on the context's streams
void step_enqueue(sphgpu_ctx *c, SimState &S, bool timed) {
    hipStream_t s = c->stream;
    if (timed) SPH_HIP(hipEventRecord(S.ev[0], s));
    launch_kick_drift(S.n, S.P, S.dt_sys.get(), S.nsys, S.time_sys.get(), s);
    if (timed) SPH_HIP(hipEventRecord(S.ev[1], s));
    forces_enqueue(c, S, timed, false);
}

void step_finish(sphgpu_ctx *c, SimState &S, bool timed) {
    hipStream_t s = c->stream;
    launch_finalize(S.n, S.P, S.dt_sys.get(), S.time_sys.get(), S.prm.ext_force_until, S.prm.damping, step_guard(S),
                    s);
    sim_dt(c, S);
    if (timed) SPH_HIP(hipEventRecord(S.ev[8], s));
}

// Records one step of the current parity into a graph.  The capture runs no
// work; the host-side exchange of the particle buffer sets that the sort makes
// is undone afterwards (the launch of the graph makes it).  False when the
// step would have allocated (its buffers are then not those of the graph).
bool capture_step(sphgpu_ctx *c, SimState &S, StepGraph &g, const std::vector<int64_t> &sig) {
    hipStream_t s = c->stream;
    const PArrays P0 = S.P, A0 = S.Palt;
    const uint64_t ep = alloc_epoch();
    hipGraph_t graph = nullptr;
    SPH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    try {
        step_enqueue(c, S, false);
        step_finish(c, S, false);
    } catch (...) {
        hipGraph_t tmp = nullptr;
        (void) hipStreamEndCapture(s, &tmp);
        if (tmp) (void) hipGraphDestroy(tmp);
        S.P = P0;
        S.Palt = A0;
        throw;
    }
    SPH_HIP(hipStreamEndCapture(s, &graph));
    S.P = P0;
    S.Palt = A0;
    if (alloc_epoch() != ep) {
        (void) hipGraphDestroy(graph);
        return false;
    }
    hipGraphExec_t exec = nullptr;
    const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void) hipGraphDestroy(graph);
    SPH_HIP(e);
    g.reset();
    g.exec = exec;
    g.sig = sig;
    ++S.graph_captures;
    return true;
}

// One step as a replay of its recorded graph, when the simulation records its
// steps and the graph of this parity is (or can now be) recorded with the
// current signature; false when the step must run directly.  A recorded step
// is captured only after a direct step has run with the same signature, so
// that every buffer it uses is allocated.  A replay whose evaluation overflowed
// left the state as the evaluation found it (StepGuard); the evaluation is
// then run directly with grown capacities, and the step finished.
bool step_graph(sphgpu_ctx *c, SimState &S) {
    if (!S.prm.graph || S.stats) return false;
    const std::vector<int64_t> sig = step_sig(c, S);
    StepGraph &g = S.graphs[sim_parity(S)];
    if (g.exec && g.sig != sig) g.reset();
    if (!g.exec) {
        if (sig != S.direct_sig) return false;
        if (!capture_step(c, S, g, sig)) return false;
    }
    hipStream_t s = c->stream;
    SPH_HIP(hipGraphLaunch(g.exec, s));
    std::swap(S.P, S.Palt);  // the recorded sort exchanged the buffer sets
    SPH_HIP(hipStreamSynchronize(s));
    ++S.graph_steps;
    if (!forces_check(S)) {
        sim_forces(c, S, false, true);
        step_finish(c, S, false);
    }
    return true;
}

SimState &sim_of(sphgpu_ctx *c) {
    if (!c->sim || !c->sim->ready) throw Error(SPHGPU_ERR_STATE, "simulation not initialised");
    return *c->sim;
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

const char *sphgpu_last_error(void) { return g_last_error.c_str(); }

int32_t sphgpu_version(void) { return 100; }

int32_t sphgpu_create(sphgpu_ctx **ctx, int32_t device, int32_t precision) {
    return guard([&] {
        if (!ctx) throw Error(SPHGPU_ERR_ARG, "null ctx pointer");
        if (precision != SPHGPU_PREC_FP64 && precision != SPHGPU_PREC_MIXED)
            throw Error(SPHGPU_ERR_ARG, "unknown precision");
        int ndev = 0;
        SPH_HIP(hipGetDeviceCount(&ndev));
        if (device < 0 || device >= ndev) throw Error(SPHGPU_ERR_ARG, "device index out of range");
        std::unique_ptr<sphgpu_ctx> c(new sphgpu_ctx());
        c->device = device;
        c->precision = precision;
        SPH_HIP(hipSetDevice(device));
        c->scratch = bind_scratch(device);
        SPH_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        SPH_HIP(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
        upload_mats(c.get());
        *ctx = c.release();
    });
}

int32_t sphgpu_set_precision(sphgpu_ctx *c, int32_t precision) {
    return guard([&] {
        check_ctx(c);
        if (precision != SPHGPU_PREC_FP64 && precision != SPHGPU_PREC_MIXED)
            throw Error(SPHGPU_ERR_ARG, "unknown precision");
        for (const Batch &b : c->batches)
            if (b.pending) throw Error(SPHGPU_ERR_STATE, "a multiwalk batch is pending; retrieve it first");
        // work already queued keeps the precision it was launched with
        SPH_HIP(hipStreamSynchronize(c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream2));
        c->precision = precision;
    });
}

int32_t sphgpu_destroy(sphgpu_ctx *ctx) {
    return guard([&] {
        if (!ctx) return;
        (void) hipSetDevice(ctx->device);
        (void) hipStreamSynchronize(ctx->stream);
        if (ctx->stream2) (void) hipStreamSynchronize(ctx->stream2);
        ctx->sim.reset();
        hipStream_t s = ctx->stream, s2 = ctx->stream2;
        delete ctx;
        (void) hipStreamDestroy(s);
        if (s2) (void) hipStreamDestroy(s2);
    });
}

// ---------------------------------------------------------------- multiwalk: density
int32_t sphgpu_dens_dispatch(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const sphgpu_epi_dens *const epi[],
                             const int32_t n_epi[], const sphgpu_epj_dens *const epj[], const int32_t n_epj[]) {
    return guard([&] {
        check_ctx(c);
        check_walks(n_walk, epi, n_epi, "epi");
        check_walks(n_walk, epj, n_epj, "epj");
        Batch &b = batch_for(c, tag);
        const GroupPlan gp = plan_groups(n_walk, n_epi, n_epj, nullptr);
        const Staged st = stage<sphgpu_epi_dens, sphgpu_epj_dens, sphgpu_spj_monopole>(
            c, b, n_walk, epi, n_epi, epj, n_epj, nullptr, nullptr, gp);
        const int64_t ni = st.ni, nj = st.nj;
        double *soa = b.soa.ensure(5 * std::max<int64_t>(ni, 1) + 4 * std::max<int64_t>(nj, 1));
        std::vector<double *> fi, fj;
        unpack(c, st.epi, ni, 6, {0, 1, 2, 3, 4}, soa, std::max<int64_t>(ni, 1), fi);
        unpack(c, st.epj, nj, 5, {0, 1, 2, 3}, soa + 5 * std::max<int64_t>(ni, 1), std::max<int64_t>(nj, 1), fj);
        double *out = b.out.ensure(2 * std::max<int64_t>(ni, 1));
        const int32_t *iota = ensure_iota(c, std::max<int64_t>(nj, 1));
        DensArgs a{(int) gp.jcnt.size(), b.gstart, fi[0], fi[1], fi[2], fi[3], fi[4], fj[0], fj[1], fj[2], fj[3],
                   list_ref(iota, b.joff, b.jcnt), out, out + std::max<int64_t>(ni, 1), nullptr};
        if (c->precision == SPHGPU_PREC_MIXED) launch_dens_mixed(a, c->stream);
        else launch_dens_fp64(a, c->stream);
        finish_batch(c, b, KIND_DENS, n_walk, n_epi, ni, {out, out + std::max<int64_t>(ni, 1)}, {0.0, 0.0});
    });
}

int32_t sphgpu_dens_retrieve(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_dens *const res[]) {
    return guard([&] { retrieve(c, KIND_DENS, tag, n_walk, n_epi, res); });
}

// ---------------------------------------------------------------- multiwalk: derivative
int32_t sphgpu_drvt_dispatch(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const sphgpu_epi_drvt *const epi[],
                             const int32_t n_epi[], const sphgpu_epj_drvt *const epj[], const int32_t n_epj[]) {
    return guard([&] {
        check_ctx(c);
        check_walks(n_walk, epi, n_epi, "epi");
        check_walks(n_walk, epj, n_epj, "epj");
        Batch &b = batch_for(c, tag);
        const GroupPlan gp = plan_groups(n_walk, n_epi, n_epj, nullptr);
        const Staged st = stage<sphgpu_epi_drvt, sphgpu_epj_drvt, sphgpu_spj_monopole>(
            c, b, n_walk, epi, n_epi, epj, n_epj, nullptr, nullptr, gp);
        const int64_t ni = std::max<int64_t>(st.ni, 1), nj = std::max<int64_t>(st.nj, 1);
        double *soa = b.soa.ensure(8 * ni + 7 * nj);
        std::vector<double *> fi, fj;
        unpack(c, st.epi, st.ni, 8, {0, 1, 2, 3, 4, 5, 6, 7}, soa, ni, fi);
        unpack(c, st.epj, st.nj, 8, {0, 1, 2, 3, 4, 5, 6}, soa + 8 * ni, nj, fj);
        double *out = b.out.ensure(6 * ni);
        const int32_t *iota = ensure_iota(c, nj);
        DrvtArgs a{(int) gp.jcnt.size(), b.gstart, fi[0], fi[1], fi[2], fi[3], fi[4], fi[5], fi[6], fi[7], nullptr,
                   fj[0], fj[1], fj[2], fj[3], fj[4], fj[5], fj[6], list_ref(iota, b.joff, b.jcnt),
                   out, out + ni, out + 2 * ni, out + 3 * ni, out + 4 * ni, out + 5 * ni};
        if (c->precision == SPHGPU_PREC_MIXED) launch_drvt_mixed(a, c->stream);
        else launch_drvt_fp64(a, c->stream);
        finish_batch(c, b, KIND_DRVT, n_walk, n_epi, st.ni, {out, out + ni, out + 2 * ni, out + 3 * ni, out + 4 * ni},
                     {0, 0, 0, 0, 0});
    });
}

int32_t sphgpu_drvt_retrieve(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_drvt *const res[]) {
    return guard([&] { retrieve(c, KIND_DRVT, tag, n_walk, n_epi, res); });
}

// ---------------------------------------------------------------- multiwalk: hydro
int32_t sphgpu_hydro_dispatch(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const sphgpu_epi_hydro *const epi[],
                              const int32_t n_epi[], const sphgpu_epj_hydro *const epj[], const int32_t n_epj[]) {
    return guard([&] {
        check_ctx(c);
        check_walks(n_walk, epi, n_epi, "epi");
        check_walks(n_walk, epj, n_epj, "epj");
        Batch &b = batch_for(c, tag);
        const GroupPlan gp = plan_groups(n_walk, n_epi, n_epj, nullptr);
        const Staged st = stage<sphgpu_epi_hydro, sphgpu_epj_hydro, sphgpu_spj_monopole>(
            c, b, n_walk, epi, n_epi, epj, n_epj, nullptr, nullptr, gp);
        const int64_t ni = std::max<int64_t>(st.ni, 1), nj = std::max<int64_t>(st.nj, 1);
        double *soa = b.soa.ensure(12 * ni + 13 * nj);
        std::vector<double *> fi, fj;
        // EPI::Hydro: pos 0-2 vel 3-5 smth 6 dens 7 pres 8 snds 9 grad_smth 10 Bal 11
        unpack(c, st.epi, st.ni, 14, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}, soa, ni, fi);
        // EPJ::Hydro: pos 0-2 vel 3-5 dens 6 mass 7 smth 8 pres 9 grad_smth 10 snds 11 Bal 12
        unpack(c, st.epj, st.nj, 15, {0, 1, 2, 3, 4, 5, 8, 6, 9, 11, 10, 12, 7}, soa + 12 * ni, nj, fj);
        double *out = b.out.ensure(5 * ni);
        const int32_t *iota = ensure_iota(c, nj);
        HydroArgs a{(int) gp.jcnt.size(), b.gstart,
                    fi[0], fi[1], fi[2], fi[3], fi[4], fi[5], fi[6], fi[7], fi[8], fi[9], fi[10], fi[11],
                    fj[0], fj[1], fj[2], fj[3], fj[4], fj[5], fj[6], fj[7], fj[8], fj[9], fj[10], fj[11], fj[12],
                    list_ref(iota, b.joff, b.jcnt), out, out + ni, out + 2 * ni, out + 3 * ni, out + 4 * ni};
        if (c->precision == SPHGPU_PREC_MIXED) launch_hydro_mixed(a, c->stream);
        else launch_hydro_fp64(a, c->stream);
        finish_batch(c, b, KIND_HYDRO, n_walk, n_epi, st.ni, {out, out + ni, out + 2 * ni, out + 3 * ni, out + 4 * ni},
                     {0, 0, 0, 0, 0});
    });
}

int32_t sphgpu_hydro_retrieve(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                              sphgpu_res_hydro *const res[]) {
    return guard([&] { retrieve(c, KIND_HYDRO, tag, n_walk, n_epi, res); });
}

// ---------------------------------------------------------------- multiwalk: gravity
int32_t sphgpu_grav_dispatch(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const sphgpu_epi_grav *const epi[],
                             const int32_t n_epi[], const sphgpu_epj_grav *const epj[], const int32_t n_epj[],
                             const sphgpu_spj_monopole *const spj[], const int32_t n_spj[]) {
    return guard([&] {
        check_ctx(c);
        check_walks(n_walk, epi, n_epi, "epi");
        check_walks(n_walk, epj, n_epj, "epj");
        if ((spj == nullptr) != (n_spj == nullptr)) throw Error(SPHGPU_ERR_ARG, "spj and n_spj must both be given");
        if (spj) check_walks(n_walk, spj, n_spj, "spj");
        Batch &b = batch_for(c, tag);
        const GroupPlan gp = plan_groups(n_walk, n_epi, n_epj, n_spj);
        const Staged st = stage<sphgpu_epi_grav, sphgpu_epj_grav, sphgpu_spj_monopole>(c, b, n_walk, epi, n_epi, epj,
                                                                                      n_epj, spj, n_spj, gp);
        const int64_t ni = std::max<int64_t>(st.ni, 1), nj = std::max<int64_t>(st.nj, 1),
                      ns = std::max<int64_t>(st.ns, 1);
        double *soa = b.soa.ensure(4 * ni + 4 * nj + 4 * ns);
        std::vector<double *> fi, fj, fs;
        unpack(c, st.epi, st.ni, 5, {0, 1, 2, 3}, soa, ni, fi);
        unpack(c, st.epj, st.nj, 5, {0, 1, 2, 3}, soa + 4 * ni, nj, fj);
        if (spj) unpack(c, st.spj, st.ns, 4, {1, 2, 3, 0}, soa + 4 * ni + 4 * nj, ns, fs);
        double *out = b.out.ensure(4 * ni);
        const int32_t *iota = ensure_iota(c, std::max(nj, ns));
        GravArgs a{(int) gp.jcnt.size(), b.gstart, fi[0], fi[1], fi[2], fi[3], nullptr,
                   fj[0], fj[1], fj[2], fj[3], list_ref(iota, b.joff, b.jcnt),
                   spj ? fs[0] : nullptr, spj ? fs[1] : nullptr, spj ? fs[2] : nullptr, spj ? fs[3] : nullptr,
                   list_ref(iota, b.soff, b.scnt), out, out + ni, out + 2 * ni, out + 3 * ni};
        if (c->precision == SPHGPU_PREC_MIXED) launch_grav_mixed(a, c->stream);
        else launch_grav_fp64(a, c->stream);
        finish_batch(c, b, KIND_GRAV, n_walk, n_epi, st.ni, {out, out + ni, out + 2 * ni, out + 3 * ni, nullptr},
                     {0, 0, 0, 0, 1.0e+30});
    });
}

int32_t sphgpu_grav_retrieve(sphgpu_ctx *c, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_grav *const res[]) {
    return guard([&] { retrieve(c, KIND_GRAV, tag, n_walk, n_epi, res); });
}

// ---------------------------------------------------------------- EoS
int32_t sphgpu_eos_load_table(sphgpu_ctx *c, int32_t material, const char *path) {
    return guard([&] {
        check_ctx(c);
        load_table_file(c, material, path);
    });
}

int32_t sphgpu_eos_set_tillotson(sphgpu_ctx *c, int32_t material, const double *params) {
    return guard([&] {
        check_ctx(c);
        set_tillotson(c, material, params);
    });
}

int32_t sphgpu_eos_set_table(sphgpu_ctx *c, int32_t material, int32_t nrho, int32_t nu, const double *cells) {
    return guard([&] {
        check_ctx(c);
        set_table(c, material, nrho, nu, cells);
    });
}

namespace {
void eos_io(sphgpu_ctx *c, int64_t n, const int32_t *material, const double *a, const double *b) {
    if (n < 0) throw Error(SPHGPU_ERR_ARG, "negative n");
    if (n > 0 && (!material || !a || !b)) throw Error(SPHGPU_ERR_ARG, "null input array");
    for (int64_t i = 0; i < n; ++i) {
        const int m = material[i];
        if (!mat_loaded(c, m))
            throw Error(SPHGPU_ERR_EOS, "material id " + std::to_string(m) + " has no EoS table");
    }
    SPH_HIP(hipMemcpyAsync(c->e_mat.ensure(n), material, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
    SPH_HIP(hipMemcpyAsync(c->e_in1.ensure(n), a, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    SPH_HIP(hipMemcpyAsync(c->e_in2.ensure(n), b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    c->e_out1.ensure(n);
    c->e_out2.ensure(n);
}
}  // namespace

int32_t sphgpu_eos_pressure(sphgpu_ctx *c, int64_t n, const int32_t *material, const double *dens, const double *eng,
                            double *pres, double *snds) {
    return guard([&] {
        check_ctx(c);
        if (n > 0 && (!pres || !snds)) throw Error(SPHGPU_ERR_ARG, "null output array");
        eos_io(c, n, material, dens, eng);
        launch_eos_pressure(n, c->e_mat.get(), c->e_in1.get(), c->e_in2.get(), c->d_mats.get(), SPHGPU_MAX_MATERIALS,
                            c->e_out1.get(), c->e_out2.get(), c->stream);
        SPH_HIP(hipMemcpyAsync(pres, c->e_out1.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipMemcpyAsync(snds, c->e_out2.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
    });
}

int32_t sphgpu_eos_pressure_noalloc(sphgpu_ctx *c, int64_t n, const int32_t *material, const double *dens,
                                    const double *eng, double *pres, double *snds) {
    return guard([&] {
        check_ctx(c);
        if (n < 0) throw Error(SPHGPU_ERR_ARG, "negative n");
        if (n > kNoAllocMaxN)
            throw Error(SPHGPU_ERR_CAPACITY, "n exceeds the fixed scratch of SPHGPU_NOALLOC_MAX_N (1048576) particles; "
                                             "sphgpu_eos_pressure takes any n");
        if (n == 0) return;
        if (!material || !dens || !eng || !pres || !snds) throw Error(SPHGPU_ERR_ARG, "null array");
        for (int64_t i = 0; i < n; ++i) {
            const int m = material[i];
            if (!mat_loaded(c, m))
                throw Error(SPHGPU_ERR_EOS, "material id " + std::to_string(m) + " has no EoS table");
        }
        if (!c->scratch) throw Error(SPHGPU_ERR_STATE, "context has no scratch address");
        // field offsets inside this device's scratch (every output is written
        // before it is read back, so nothing needs clearing between calls)
        auto up = [](size_t b) { return (b + kNoAllocAlign - 1) & ~(kNoAllocAlign - 1); };
        unsigned char *base = (unsigned char *) c->scratch;
        const size_t o_mat = 0, o_rho = up(o_mat + sizeof(int32_t) * n), o_u = up(o_rho + sizeof(double) * n),
                     o_p = up(o_u + sizeof(double) * n), o_c = up(o_p + sizeof(double) * n);
        if (o_c + sizeof(double) * n > kNoAllocBytes) throw Error(SPHGPU_ERR_CAPACITY, "scratch layout overflow");
        int32_t *d_mat = (int32_t *) (base + o_mat);
        double *d_rho = (double *) (base + o_rho), *d_u = (double *) (base + o_u);
        double *d_p = (double *) (base + o_p), *d_c = (double *) (base + o_c);
        SPH_HIP(hipMemcpyAsync(d_mat, material, sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
        SPH_HIP(hipMemcpyAsync(d_rho, dens, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        SPH_HIP(hipMemcpyAsync(d_u, eng, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
        launch_eos_pressure(n, d_mat, d_rho, d_u, c->d_mats.get(), SPHGPU_MAX_MATERIALS, d_p, d_c, c->stream);
        SPH_HIP(hipMemcpyAsync(pres, d_p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipMemcpyAsync(snds, d_c, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
    });
}

int32_t sphgpu_eos_property(sphgpu_ctx *c, int64_t n, const int32_t *material, const double *dens, const double *eng,
                            int32_t prop, double *out) {
    return guard([&] {
        check_ctx(c);
        if (prop < 0 || prop > 5) throw Error(SPHGPU_ERR_ARG, "property index out of range");
        if (n > 0 && !out) throw Error(SPHGPU_ERR_ARG, "null output array");
        eos_io(c, n, material, dens, eng);
        launch_eos_property(n, c->e_mat.get(), c->e_in1.get(), c->e_in2.get(), c->d_mats.get(), SPHGPU_MAX_MATERIALS,
                            prop, c->e_out1.get(), c->stream);
        SPH_HIP(hipMemcpyAsync(out, c->e_out1.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
    });
}

int32_t sphgpu_eos_internal_energy(sphgpu_ctx *c, int64_t n, const int32_t *material, const double *dens,
                                   const double *ent, int32_t abs_int, double *eng) {
    return guard([&] {
        check_ctx(c);
        if (n > 0 && !eng) throw Error(SPHGPU_ERR_ARG, "null output array");
        eos_io(c, n, material, dens, ent);
        for (int64_t i = 0; i < n; ++i)
            if (c->mats[material[i]].kind != 0 || !c->mats[material[i]].sorted)
                throw Error(SPHGPU_ERR_EOS, "InternalEnergy needs a table (not an analytic material) whose density column "
                                             "is non-decreasing");
        launch_eos_internal_energy(n, c->e_mat.get(), c->e_in1.get(), c->e_in2.get(), c->d_mats.get(),
                                   SPHGPU_MAX_MATERIALS, abs_int ? 1 : 0, c->e_out1.get(), c->stream);
        SPH_HIP(hipMemcpyAsync(eng, c->e_out1.get(), sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
        SPH_HIP(hipStreamSynchronize(c->stream));
    });
}

// ---------------------------------------------------------------- simulation
void sphgpu_sim_default_params(sphgpu_sim_params *p) {
    if (!p) return;
    p->n_leaf = 16;
    p->n_group = 64;
    p->theta = 0.5;
    p->damping = 1.0;
    p->ext_force_until = 5000.0;
    p->nsys = 1;
    p->list_cap_sph = 4096;
    p->list_cap_epj = 8192;
    p->eos_density_fp64 = 1;
    p->list_cap_spj = 8192;
    p->list_cap_front = 4096;
    p->list_front_lds = kFrontLds;
    p->walk_waves = 0;
    p->persistent = 0;
    p->graph = 0;
    p->vsig_split = 1;
    p->hydro_compact = 1;
}

int32_t sphgpu_sim_init(sphgpu_ctx *c, const sphgpu_sim_params *prm, int64_t n, const double *pos, const double *vel,
                        const double *mass, const double *dens, const double *eng, const int32_t *material,
                        const int32_t *system, const int64_t *id) {
    return guard([&] {
        check_ctx(c);
        if (!prm) throw Error(SPHGPU_ERR_ARG, "null params");
        if (n <= 0 || n > INT32_MAX / 2) throw Error(SPHGPU_ERR_ARG, "particle count out of range");
        if (!pos || !vel || !mass || !dens || !eng || !material) throw Error(SPHGPU_ERR_ARG, "null particle array");
        if (prm->n_leaf < 1 || prm->n_leaf > kWave)
            throw Error(SPHGPU_ERR_ARG, "n_leaf must be in [1, 64]");
        if (prm->n_group < 1 || prm->n_group > SPHGPU_MAX_GROUP)
            throw Error(SPHGPU_ERR_ARG, "n_group must be in [1, SPHGPU_MAX_GROUP]");
        if (!(prm->theta >= 0)) throw Error(SPHGPU_ERR_ARG, "theta must be >= 0");
        if (prm->vsig_split < 0 || prm->vsig_split > 3) throw Error(SPHGPU_ERR_ARG, "vsig_split must be in [0, 3]");
        if (prm->hydro_compact < 0 || prm->hydro_compact > 1)
            throw Error(SPHGPU_ERR_ARG, "hydro_compact must be 0 or 1");
        const int nsys = prm->nsys > 0 ? prm->nsys : 1;
        if (nsys > 1 && !system) throw Error(SPHGPU_ERR_ARG, "nsys > 1 needs a system array");
        for (int64_t i = 0; i < n; ++i) {
            const int m = material[i];
            if (!mat_loaded(c, m))
                throw Error(SPHGPU_ERR_EOS, "material id " + std::to_string(m) + " has no EoS table");
        }
        std::unique_ptr<SimState> S(new SimState());
        S->prm = *prm;
        S->n = n;
        S->nsys = nsys;
        S->sys_start_h.assign(nsys + 1, 0);
        std::vector<int32_t> sysv(n, 0);
        if (system) {
            for (int64_t i = 0; i < n; ++i) {
                if (system[i] < 0 || system[i] >= nsys) throw Error(SPHGPU_ERR_ARG, "system id out of range");
                if (i > 0 && system[i] < system[i - 1]) throw Error(SPHGPU_ERR_ARG, "system ids must be non-decreasing");
                sysv[i] = system[i];
            }
        }
        std::vector<int64_t> cnt(nsys, 0);
        for (int64_t i = 0; i < n; ++i) cnt[sysv[i]]++;
        int64_t maxsys = 0;
        for (int q = 0; q < nsys; ++q) {
            S->sys_start_h[q + 1] = S->sys_start_h[q] + (int32_t) cnt[q];
            maxsys = std::max(maxsys, cnt[q]);
        }
        int sb = 0;
        while ((1 << sb) < nsys) ++sb;
        TreeParams &tp = S->tp;
        tp.n_leaf = prm->n_leaf;
        tp.n_group = prm->n_group;
        S->tree.max_group = prm->n_group;
        tp.nsys = nsys;
        tp.sysbits = sb;
        tp.bits = std::min(21, (63 - sb) / 3);
        tp.theta = prm->theta;
        int nl = 1;
        for (int64_t T = prm->n_leaf; T < maxsys; T *= 8) ++nl;
        if (nl > kMaxLevels) throw Error(SPHGPU_ERR_ARG, "tree too deep");
        tp.nlevels = nl;
        S->lsph.cap = prm->list_cap_sph > 0 ? prm->list_cap_sph : 4096;
        S->lgrav.cap = prm->list_cap_epj > 0 ? prm->list_cap_epj : 8192;
        S->lgrav.scap = prm->list_cap_spj > 0 ? prm->list_cap_spj : 8192;
        for (Lists *L : {&S->lsph, &S->lgrav}) {
            L->fcap = prm->list_cap_front > 0 ? prm->list_cap_front : 4096;
            L->flds = (prm->list_front_lds > 0 && prm->list_front_lds < kFrontLds) ? prm->list_front_lds : kFrontLds;
            L->waves = prm->walk_waves > 0 ? prm->walk_waves : 0;
        }
        sim_bind(*S);
        hipStream_t s = c->stream;
        PArrays &P = S->P;
        // upload
        std::vector<double> tmp(n);
        auto up = [&](double *d, const double *src, int stride, int comp) {
            for (int64_t i = 0; i < n; ++i) tmp[i] = src[i * stride + comp];
            SPH_HIP(hipMemcpy(d, tmp.data(), sizeof(double) * n, hipMemcpyHostToDevice));
        };
        up(P.x, pos, 3, 0); up(P.y, pos, 3, 1); up(P.z, pos, 3, 2);
        up(P.vx, vel, 3, 0); up(P.vy, vel, 3, 1); up(P.vz, vel, 3, 2);
        up(P.m, mass, 1, 0); up(P.rho, dens, 1, 0); up(P.u, eng, 1, 0);
        SPH_HIP(hipMemcpy(P.mat, material, sizeof(int32_t) * n, hipMemcpyHostToDevice));
        SPH_HIP(hipMemcpy(P.sys, sysv.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice));
        std::vector<int64_t> ids(n);
        for (int64_t i = 0; i < n; ++i) ids[i] = id ? id[i] : i;
        SPH_HIP(hipMemcpy(P.id, ids.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice));
        for (double *z : {P.ax, P.ay, P.az, P.gx, P.gy, P.gz, P.pot, P.du, P.dti, P.vhx, P.vhy, P.vhz, P.uh, P.p, P.cs,
                          P.divv, P.rotx, P.roty, P.rotz})
            SPH_HIP(hipMemsetAsync(z, 0, sizeof(double) * n, s));
        SPH_HIP(hipMemcpy(S->tree.sys_start.ensure(nsys + 1), S->sys_start_h.data(), sizeof(int32_t) * (nsys + 1),
                          hipMemcpyHostToDevice));
        S->tree.sys_start_h = S->sys_start_h;
        S->dt_sys.ensure(nsys);
        S->time_sys.ensure(nsys);
        SPH_HIP(hipMemsetAsync(S->time_sys.get(), 0, sizeof(double) * nsys, s));
        const int nchunk = (int) std::max<int64_t>(1, (maxsys + 4095) / 4096);
        S->dt_part.ensure((size_t) nsys * nchunk);
        S->stash.flags = hipHostMallocCoherent;
        S->stash.ensure(kStWords);
        for (auto &e : S->ev) SPH_HIP(hipEventCreate(&e));
        for (auto &e : S->xev) SPH_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        // src/main.cpp:69-104: initialize, forces, dt, external force
        launch_init_h(n, P, s);
        sim_forces(c, *S, false);
        sim_dt(c, *S);
        launch_ext_force(n, S->P, S->time_sys.get(), prm->ext_force_until, s);
        SPH_HIP(hipStreamSynchronize(s));
        S->ready = true;
        c->sim = std::move(S);
    });
}

int32_t sphgpu_sim_step(sphgpu_ctx *c, int32_t nsteps) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (nsteps < 0) throw Error(SPHGPU_ERR_ARG, "negative step count");
        hipStream_t s = c->stream;
        bool direct = false;
        for (int k = 0; k < nsteps; ++k) {
            if (step_graph(c, S)) continue;
            step_enqueue(c, S, true);
            SPH_HIP(hipStreamSynchronize(s));  // the one wait of the evaluation
            if (forces_check(S)) forces_stats(c, S);
            else sim_forces(c, S, true, true);
            step_finish(c, S, true);
            direct = true;
        }
        SPH_HIP(hipStreamSynchronize(s));
        if (direct)
            for (int k = 0; k < kStages; ++k) {
                float ms = 0;
                SPH_HIP(hipEventElapsedTime(&ms, S.ev[k], S.ev[k + 1]));
                S.stage_ms[k] = ms;
            }
    });
}

int32_t sphgpu_sim_size(sphgpu_ctx *c, int64_t *n) {
    return guard([&] {
        check_ctx(c);
        if (!n) throw Error(SPHGPU_ERR_ARG, "null pointer");
        *n = sim_of(c).n;
    });
}

int32_t sphgpu_sim_get(sphgpu_ctx *c, int32_t field, void *out) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!out) throw Error(SPHGPU_ERR_ARG, "null output");
        const int64_t n = S.n;
        PArrays &P = S.P;
        SPH_HIP(hipStreamSynchronize(c->stream));
        auto vec3 = [&](const double *a, const double *b, const double *d) {
            std::vector<double> t(3 * n);
            SPH_HIP(hipMemcpy(t.data(), a, sizeof(double) * n, hipMemcpyDeviceToHost));
            SPH_HIP(hipMemcpy(t.data() + n, b, sizeof(double) * n, hipMemcpyDeviceToHost));
            SPH_HIP(hipMemcpy(t.data() + 2 * n, d, sizeof(double) * n, hipMemcpyDeviceToHost));
            double *o = (double *) out;
            for (int64_t i = 0; i < n; ++i) {
                o[3 * i] = t[i];
                o[3 * i + 1] = t[n + i];
                o[3 * i + 2] = t[2 * n + i];
            }
        };
        auto sc = [&](const void *a, size_t el) { SPH_HIP(hipMemcpy(out, a, el * n, hipMemcpyDeviceToHost)); };
        switch (field) {
            case SPHGPU_F_POS: vec3(P.x, P.y, P.z); break;
            case SPHGPU_F_VEL: vec3(P.vx, P.vy, P.vz); break;
            case SPHGPU_F_ACC: vec3(P.ax, P.ay, P.az); break;
            case SPHGPU_F_GRAV: vec3(P.gx, P.gy, P.gz); break;
            case SPHGPU_F_ROTV: vec3(P.rotx, P.roty, P.rotz); break;
            case SPHGPU_F_VELHALF: vec3(P.vhx, P.vhy, P.vhz); break;
            case SPHGPU_F_POT: sc(P.pot, 8); break;
            case SPHGPU_F_MASS: sc(P.m, 8); break;
            case SPHGPU_F_SMTH: sc(P.h, 8); break;
            case SPHGPU_F_DENS: sc(P.rho, 8); break;
            case SPHGPU_F_ENG: sc(P.u, 8); break;
            case SPHGPU_F_PRES: sc(P.p, 8); break;
            case SPHGPU_F_SNDS: sc(P.cs, 8); break;
            case SPHGPU_F_DIVV: sc(P.divv, 8); break;
            case SPHGPU_F_GRADH: sc(P.gs, 8); break;
            case SPHGPU_F_BAL: sc(P.bal, 8); break;
            case SPHGPU_F_ENGDOT: sc(P.du, 8); break;
            case SPHGPU_F_DT: sc(P.dti, 8); break;
            case SPHGPU_F_ENGHALF: sc(P.uh, 8); break;
            case SPHGPU_F_NGB: sc(P.ngb, 4); break;
            case SPHGPU_F_MAT: sc(P.mat, 4); break;
            case SPHGPU_F_SYS: sc(P.sys, 4); break;
            case SPHGPU_F_ID: sc(P.id, 8); break;
            default: throw Error(SPHGPU_ERR_ARG, "unknown field");
        }
    });
}

int32_t sphgpu_sim_get_time(sphgpu_ctx *c, double *dt_sys, double *time_sys) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        SPH_HIP(hipStreamSynchronize(c->stream));
        if (dt_sys) SPH_HIP(hipMemcpy(dt_sys, S.dt_sys.get(), sizeof(double) * S.nsys, hipMemcpyDeviceToHost));
        if (time_sys) SPH_HIP(hipMemcpy(time_sys, S.time_sys.get(), sizeof(double) * S.nsys, hipMemcpyDeviceToHost));
    });
}

int32_t sphgpu_sim_stage_ms(sphgpu_ctx *c, double *ms, int32_t n) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!ms || n < 0) throw Error(SPHGPU_ERR_ARG, "bad output");
        for (int k = 0; k < std::min(n, kStages); ++k) ms[k] = S.stage_ms[k];
    });
}

int32_t sphgpu_sim_list_builds(sphgpu_ctx *c, int64_t *gather_builds, int64_t *force_evals) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!gather_builds || !force_evals) throw Error(SPHGPU_ERR_ARG, "null output");
        *gather_builds = S.gather_builds;
        *force_evals = S.force_evals;
    });
}

int32_t sphgpu_sim_set_capacities(sphgpu_ctx *c, int32_t list_cap_sph, int32_t list_cap_epj,
                                  int32_t list_cap_spj) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (list_cap_sph < 1 || list_cap_epj < 1 || list_cap_spj < 1)
            throw Error(SPHGPU_ERR_ARG, "list capacities must be >= 1");
        SPH_HIP(hipStreamSynchronize(c->stream));
        const bool shrink = list_cap_sph <= S.lsph.cap && list_cap_epj <= S.lgrav.cap && list_cap_spj <= S.lgrav.scap;
        S.lsph.cap = list_cap_sph;
        S.lgrav.cap = list_cap_epj;
        S.lgrav.scap = list_cap_spj;
        // smaller capacities fit the buffers allocated: a step may be recorded at once
        if (shrink) S.direct_sig = step_sig(c, S);
    });
}

int32_t sphgpu_sim_graph_stats(sphgpu_ctx *c, int64_t *replayed_steps, int64_t *captures) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!replayed_steps || !captures) throw Error(SPHGPU_ERR_ARG, "null output");
        *replayed_steps = S.graph_steps;
        *captures = S.graph_captures;
    });
}

int32_t sphgpu_sim_set_stats(sphgpu_ctx *c, int32_t on) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        SPH_HIP(hipStreamSynchronize(c->stream));
        S.stats = on != 0;
        for (auto &p : S.kev)
            for (auto &e : p)
                if (!e) SPH_HIP(hipEventCreate(&e));
        for (int k = 0; k < kStatKinds; ++k) {
            S.kms[k] = 0;
            S.klaunch[k] = 0;
        }
        SPH_HIP(hipMemset(S.kcount.ensure(4 * kStatKinds), 0, sizeof(long long) * 4 * kStatKinds));
    });
}

int32_t sphgpu_sim_kernel_stats(sphgpu_ctx *c, sphgpu_kernel_stat *out, int32_t n) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!out || n < 0) throw Error(SPHGPU_ERR_ARG, "bad output");
        SPH_HIP(hipStreamSynchronize(c->stream));
        long long cnt[4 * kStatKinds] = {0};
        if (S.kcount.get()) SPH_HIP(hipMemcpy(cnt, S.kcount.get(), sizeof(cnt), hipMemcpyDeviceToHost));
        for (int k = 0; k < std::min(n, kStatKinds); ++k) {
            out[k].ms = S.kms[k];
            out[k].launches = S.klaunch[k];
            out[k].n_i = cnt[4 * k + 0];
            out[k].n_j = cnt[4 * k + 1];
            out[k].n_s = cnt[4 * k + 2];
            out[k].pairs = cnt[4 * k + 3];
        }
    });
}

int32_t sphgpu_sim_lists(sphgpu_ctx *c, int32_t kind, int64_t *n_groups, int64_t *n_j, int64_t *n_s,
                         int32_t *group_start, int64_t *j_off, int32_t *j_idx, int64_t *s_off, int32_t *s_idx) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (kind < 0 || kind > 4) throw Error(SPHGPU_ERR_ARG, "kind must be in [0, 4]");
        if (kind == 4 && S.tree.max_group > kWave)
            throw Error(SPHGPU_ERR_ARG, "kind 4 needs groups of at most 64 particles");
        if (!n_groups || !n_j || !n_s) throw Error(SPHGPU_ERR_ARG, "null size output");
        hipStream_t s = c->stream;
        // The particles are in key order after every force evaluation (init and
        // each step sort before evaluating and do not move them afterwards), so
        // the tree is rebuilt in place: a re-sort would permute only the fields
        // the step carries and leave the others in the alternate buffers.
        tree_build_only(S.tree, S.tp, S.n, S.P, s);
        tree_sync_counts(S.tree, s);
        Lists &L = (kind == 2) ? S.lgrav : S.lsph;
        if (kind == 2) build_lists_grav(S.tree, S.tp, S.P, L, s);
        else if (kind == 3) {
            build_lists_sph(S.tree, S.tp, S.P, 0, L, s, kGatherMargin);
            enqueue_gather_compact(S.tree, S.P, L, kGatherMargin, s);
        } else {
            if (kind >= 1) tree_update_maxH(S.tree, S.P.h, s);
            build_lists_sph(S.tree, S.tp, S.P, kind == 4 ? 1 : kind, L, s);
        }
        const int32_t *idx_d = L.idx.get(), *cnt_d = L.cnt.get();
        if (kind == 4) {
            const size_t G = (size_t) std::max<int64_t>(S.tree.gcap, 1);
            int32_t *hidx = S.hidx.ensure(G * (size_t) L.cap), *hcnt = S.hcnt.ensure(G);
            enqueue_sym_compact(S.tree, S.P, L, nullptr, hidx, hcnt, s);
            idx_d = hidx;
            cnt_d = hcnt;
        }
        SPH_HIP(hipStreamSynchronize(s));
        const int64_t G = S.tree.ngroups;
        std::vector<int32_t> cnt(G), scnt(G);
        SPH_HIP(hipMemcpy(cnt.data(), cnt_d, sizeof(int32_t) * G, hipMemcpyDeviceToHost));
        if (kind == 2) SPH_HIP(hipMemcpy(scnt.data(), L.scnt.get(), sizeof(int32_t) * G, hipMemcpyDeviceToHost));
        int64_t tj = 0, ts = 0;
        for (int64_t g = 0; g < G; ++g) {
            tj += cnt[g];
            ts += (kind == 2) ? scnt[g] : 0;
        }
        *n_groups = G;
        *n_j = tj;
        *n_s = ts;
        if (!group_start && !j_off && !j_idx && !s_off && !s_idx) return;
        if (group_start) SPH_HIP(hipMemcpy(group_start, S.tree.gstart.get(), sizeof(int32_t) * (G + 1), hipMemcpyDeviceToHost));
        std::vector<int32_t> all((size_t) G * L.cap);
        SPH_HIP(hipMemcpy(all.data(), idx_d, sizeof(int32_t) * all.size(), hipMemcpyDeviceToHost));
        int64_t o = 0;
        for (int64_t g = 0; g < G; ++g) {
            if (j_off) j_off[g] = o;
            if (j_idx) memcpy(j_idx + o, all.data() + g * L.cap, sizeof(int32_t) * cnt[g]);
            o += cnt[g];
        }
        if (j_off) j_off[G] = o;
        if (kind == 2) {
            std::vector<int32_t> sall((size_t) G * L.scap);
            SPH_HIP(hipMemcpy(sall.data(), L.sidx.get(), sizeof(int32_t) * sall.size(), hipMemcpyDeviceToHost));
            int64_t so = 0;
            for (int64_t g = 0; g < G; ++g) {
                if (s_off) s_off[g] = so;
                if (s_idx) memcpy(s_idx + so, sall.data() + g * L.scap, sizeof(int32_t) * scnt[g]);
                so += scnt[g];
            }
            if (s_off) s_off[G] = so;
        }
    });
}

int32_t sphgpu_sim_nodes(sphgpu_ctx *c, int64_t *n_nodes, double *moments) {
    return guard([&] {
        check_ctx(c);
        SimState &S = sim_of(c);
        if (!n_nodes) throw Error(SPHGPU_ERR_ARG, "null pointer");
        const int64_t nn = S.tree.nnodes;
        *n_nodes = nn;
        if (!moments) return;
        SPH_HIP(hipStreamSynchronize(c->stream));
        const Tree &t = S.tree;
        const double *src[10] = {t.mass.get(), t.comx.get(), t.comy.get(), t.comz.get(), t.lox.get(),
                                 t.loy.get(), t.loz.get(), t.hix.get(), t.hiy.get(), t.hiz.get()};
        std::vector<double> tmp(nn);
        for (int f = 0; f < 10; ++f) {
            SPH_HIP(hipMemcpy(tmp.data(), src[f], sizeof(double) * nn, hipMemcpyDeviceToHost));
            for (int64_t k = 0; k < nn; ++k) moments[k * 10 + f] = tmp[k];
        }
    });
}

}  // extern "C"
