// kernels_fp64.hip -- FP64 interaction kernels (the "plain" path).
//
// One 64-lane wavefront per i-group: lane l owns i-particle gstart[g]+l, and
// the group's j-list is streamed through LDS in tiles of 64 (each lane
// gathers one j, then every lane reads the tile back as an LDS broadcast).
// The pair terms are those of src/force.h / src/kernel.h, evaluated in FP64
// with the algebra rearranged for instruction count: the kernel normalisation
// and constant factors come out of the sums, both gradients of a hydro pair are
// multiples of dr, |dr| comes from a refined reciprocal square root, and
// polynomials and dot products are fused multiply-adds (written out with
// fma(): the file is compiled with -ffp-contract=off, so nothing else is
// fused).  Every change moves results by a few units in the last place of the
// terms.  Each lane sums its pairs sequentially in tile order (see Tiles).
// Pairs outside the support are skipped: there W and gradW are exactly zero,
// so the skip changes no sum.
#include <map>
#include <mutex>

#include "internal.hpp"

namespace sphgpu {
namespace {

__device__ __forceinline__ double pow8_(double a) { const double a2 = a * a; const double a4 = a2 * a2; return a4 * a4; }

// WendlandC6 normalisation (1365/64)/(H^3 pi) (src/kernel.h:16)
__device__ __forceinline__ double wnorm_(double H) { return (1365. / 64.) / (H * H * H * kPi); }

// WendlandC6::W(r) with s = r/H (src/kernel.h:8-19) is norm * p(s) (1-s)_+^8,
// p(s) = 1 + 8s + 25s^2 + 32s^3.
__device__ __forceinline__ double wpoly_(double s) { return fma(fma(fma(32.0, s, 25.0), s, 8.0), s, 1.0); }

// WendlandC6::gradW (src/kernel.h:22-34) is dr * norm * dW/ds / (r H + 1e-6 h)
// with dW/ds = (1-s)^7 [(1-s)(8 + 50s + 96s^2) - 8 p(s)]; the bracket expands
// to -22 s (1 + 7s + 16s^2) (its constant terms cancel exactly).  Returned
// without the factor -22 norm, from o7 = (1-s)_+^7.
__device__ __forceinline__ double dwds_(double s, double o7) { return (s * fma(fma(16.0, s, 7.0), s, 1.0)) * o7; }
__device__ __forceinline__ double pow7f_(double o) { const double o2 = o * o, o4 = o2 * o2; return (o4 * o2) * o; }

// The polynomials' constants held in registers for the pair loops: v_fma_f64
// takes no literal operand (and at most one scalar one), so an fma of two
// constants otherwise rebuilds one of them with two moves in every pair.  Set
// up once per group; the empty asm statements emit nothing, they only hide
// the values from constant folding.  The arithmetic is unchanged.
struct PolyK {
    double k32 = 32.0, k25 = 25.0, k16 = 16.0, k7 = 7.0;
    __device__ __forceinline__ PolyK() {
        asm volatile("" : "+s"(k32));
        asm volatile("" : "+v"(k25));
        asm volatile("" : "+s"(k16));
        asm volatile("" : "+v"(k7));
    }
};
__device__ __forceinline__ double wpoly_(double s, const PolyK &K) {
    return fma(fma(fma(K.k32, s, K.k25), s, 8.0), s, 1.0);
}
__device__ __forceinline__ double dwds_(double s, double o7, const PolyK &K) {
    return (s * fma(fma(K.k16, s, K.k7), s, 1.0)) * o7;
}

// Reciprocal of a positive, finite x: the hardware estimate refined by two
// Newton steps (five instructions against ten for the IEEE divide; the
// reference's -ffast-math build also divides through reciprocals).
__device__ __forceinline__ double rcp_(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

// 1/sqrt(x) of a positive, finite x: the hardware estimate y0 refined by one
// step of the series (1 - e)^(-1/2) = 1 + e/2 + 3e^2/8 + O(e^3), e = 1 - x y0^2
// (Householder's method of order 2): six instructions against eighteen for
// the IEEE square root, which is followed by a divide besides.
__device__ __forceinline__ double rsqrt_(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(0.375, e, 0.5), y);
}
// |dr| and 1/|dr| from r2 = |dr|^2 >= 0.  The 1e-300, far below the rounding
// of any r2 of a physical distance, keeps r2 = 0 finite: |dr| = 0 exactly.
struct Dist {
    double r, rinv;
};
__device__ __forceinline__ Dist dist_(double r2) {
    const double rinv = rsqrt_(r2 + 1.0e-300);
    return {r2 * rinv, rinv};
}
__device__ __forceinline__ double norm2_(double x, double y, double z) { return fma(z, z, fma(y, y, x * x)); }

// group-level reductions (once per group) through LDS permutes; the per-tile
// ones take the VALU's lane paths (dpp_wave_*, xor32_f)
__device__ __forceinline__ double wave_min_(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmin(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ double wave_max_(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = dmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// The support test runs in FP32 on offsets o = (float)(x - c) from the centre c
// of the group's box, in dot form (|o_j|^2 - 2 o_i.o_j against a threshold,
// see dot_mask and dot_cut_); the FP64 pair code then evaluates the pairs it
// admits.  The thresholds make the FP32 test admit a superset of the
// in-support pairs: each offset component rounds by at most 2^-24 |o|, which
// the radius H (1 + 1e-4) + 2^-22 (|o_i|_1 + |o_j|_1) covers, and the FMA
// chain's rounding, a few ulp of |o_i|^2 + |o_j|^2, is covered by
// 2^-19 (|o_i|_1^2 + |o_j|_1^2).  Pairs admitted beyond the support give
// W = gradW = 0 exactly, so every sum is the one over the in-support pairs.
struct Frame {
    double cx, cy, cz;
};
// The hydro pass's j record in LDS with the entry's FP32 offset record (the
// support mask's operand) beside it: 144 bytes.  In the pair loop each lane
// reads the record of its own j; at a stride of 128 bytes (the span of the
// banks) the lanes' 16-byte reads all start on the same 4 banks and
// serialise, at 144 bytes they start on 8 positions, which lets the banks
// serve 8 lanes per cycle.
struct alignas(16) HydroSlot {
    HydroRec r;
    float4 o;  // FP32 offset from the frame centre, |offset|^2
};
static_assert(sizeof(HydroSlot) == 144, "LDS record stride");
// The derivative pass's j entry in LDS: position and mass, velocity and the
// FP32 offset record in an 80-byte slot (stride on 8 start banks, as above).
struct alignas(16) DrvtSlot {
    double x, y, z, m, vx, vy, vz, pad;
    float4 o;
};
static_assert(sizeof(DrvtSlot) == 80, "LDS record stride");
// The density pass's: position and mass, and the FP32 offset record: 48 bytes.
struct alignas(16) DensSlot {
    double x, y, z, m;
    float4 o;
};
static_assert(sizeof(DensSlot) == 48, "LDS record stride");
__device__ __forceinline__ Frame group_frame(double x, double y, double z, bool act) {
    const double lx = wave_min_(act ? x : INFINITY), hx = wave_max_(act ? x : -INFINITY);
    const double ly = wave_min_(act ? y : INFINITY), hy = wave_max_(act ? y : -INFINITY);
    const double lz = wave_min_(act ? z : INFINITY), hz = wave_max_(act ? z : -INFINITY);
    return {0.5 * (lx + hx), 0.5 * (ly + hy), 0.5 * (lz + hz)};
}
__device__ __forceinline__ float wave_maxf_(float v) { return dpp_wave_maxf(v); }

// ---------------------------------------------------------------- density
// CalcDensity::operator() (src/force.h:9-23)
// At least 5 waves per SIMD: the register cap this sets costs no spills and
// measured faster than the compiler's own choice (4 waves); the kernel now fits
// 71 registers (7 waves).  A cap for 8 waves (64 registers, two spilled)
// measured slower: 0.508 -> 0.516 ms at 1M.
__device__ __forceinline__ void dens_group(const DensArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const PolyK PK;  // polynomial constants in registers
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xi = a.xi[ip], yi = a.yi[ip], zi = a.zi[ip], hi = a.hi[ip], mi = a.mi[ip];
    const double H = kSupport * hi;
    const double invH = 1.0 / H;
    const double norm = wnorm_(H);
    const Frame fr = group_frame(xi, yi, zi, act);
    const float oxi = (float) (xi - fr.cx), oyi = (float) (yi - fr.cy), ozi = (float) (zi - fr.cz);
    const float Hc = (float) H * 1.0001f;
    const float wi = fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)), n1i = norm1_(oxi, oyi, ozi);
    const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    double dens = 0.0;
    int ngb = 0;
    const int64_t lb = list_begin(a.L, g);
    const Tiles tl(a.L.cnt[g]);
    __shared__ DensSlot sj[kWave];
    // Software-pipelined tiles: the j record of tile t+1 and the list index of
    // tile t+2 are in flight while tile t is evaluated (a lane without an entry
    // loads particle 0, which exists whenever the list is not empty).
    auto load = [&](int q) { return a.jpos ? a.jpos[q] : make_double4(a.xj[q], a.yj[q], a.zj[q], a.mj[q]); };
    double4 pq = make_double4(0.0, 0.0, 0.0, 0.0);
    int qn = 0;
    if (tl.T > 0) {
        const int k0 = tl.entry(0, lane), k1 = tl.entry(1, lane);
        pq = load(k0 < tl.nj ? a.L.idx[lb + k0] : 0);
        qn = (tl.T > 1 && k1 < tl.nj) ? a.L.idx[lb + k1] : 0;
    }
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        float nj1 = 0.f;
        if (k < tl.nj) {
            const double x = pq.x, y = pq.y, z = pq.z;
            const float ox = (float) (x - fr.cx), oy = (float) (y - fr.cy), oz = (float) (z - fr.cz);
            sj[lane] = DensSlot{x, y, z, pq.w, make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)))};
            nj1 = norm1_(ox, oy, oz);
        }
        if (t + 1 < tl.T) {
            pq = load(qn);
            const int k2 = tl.entry(t + 2, lane);
            qn = (t + 2 < tl.T && k2 < tl.nj) ? a.L.idx[lb + k2] : 0;
        }
        const float nj1max = wave_maxf_(nj1);  // every lane takes part: outside the condition
        const float negc = act ? -dot_cut_(Hc, wi, n1i, nj1max) : INFINITY;
        const float pnegc = xor32_f(negc);
        __syncthreads();
        const uint64_t mask = mfma_dot_mask_f([&](int e) { return sj[e].o; }, tl.count(t), tx, ty, tz, negc, ptx, pty,
                                              ptz, pnegc);
        for (uint64_t m = mask; m; m &= m - 1) {
            const DensSlot &J = sj[__builtin_ctzll(m)];
            const double dx = J.x - xi, dy = J.y - yi, dz = J.z - zi;
            const double s = dist_(norm2_(dx, dy, dz)).r * invH;  // r / H with 1/H hoisted out of the j loop
            const double o = fmax(1.0 - s, 0.0);
            dens = fma(J.m, wpoly_(s, PK) * pow8_(o), dens);  // sum of m_j W_ij / norm
            ngb += (o > 0.0) ? 1 : 0;
        }
        __syncthreads();
    }
    double hn = 0.0;  // lanes without a particle take no part in the check
    if (act) {
        dens = dmax(5.0, dens * norm);  // FLAG_GI floor (src/force.h:18-20)
        a.rho_out[ip] = dens;
        hn = kSMTH * pow(mi / dens, 1.0 / 3.0);
        a.h_out[ip] = hn;
        if (a.ngb_out) a.ngb_out[ip] = ngb;
    }
    if (a.flag) gather_cover_check(a.sbox, a.crad, a.flag, g, act, xi, yi, zi, hn);
}
__global__ __launch_bounds__(kWave, 5) void k_dens(DensArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) dens_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave, 5) void k_dens_persistent(DensArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        dens_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ---------------------------------------------------------------- derivative
// CalcDerivative::operator() (src/force.h:89-106) + Balsara switch of
// RealPtcl::copyFromForce (src/class.h:132-142) when csi is given.
__device__ __forceinline__ void drvt_group(const DrvtArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const PolyK PK;  // polynomial constants in registers
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xi = a.xi[ip], yi = a.yi[ip], zi = a.zi[ip];
    const double vxi = a.vxi[ip], vyi = a.vyi[ip], vzi = a.vzi[ip];
    const double hi = a.hi[ip], rhoi = a.rhoi[ip];
    const double H = kSupport * hi;
    const double invH = 1.0 / H;
    const double norm = wnorm_(H);
    const Frame fr = group_frame(xi, yi, zi, act);
    const float oxi = (float) (xi - fr.cx), oyi = (float) (yi - fr.cy), ozi = (float) (zi - fr.cz);
    const float Hc = (float) H * 1.0001f;
    const float wi = fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)), n1i = norm1_(oxi, oyi, ozi);
    const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    const double reg = 1.0e-6 * hi;
    // the sums leave out the factor K = -22 norm of gradW
    double divv = 0.0, rx = 0.0, ry = 0.0, rz = 0.0, gsw = 0.0, gsf = 0.0;
    const int64_t lb = list_begin(a.L, g);
    const Tiles tl(a.L.cnt[g]);
    __shared__ DrvtSlot sj[kWave];
    // software-pipelined tiles, as in the density pass
    auto load = [&](int q, double4 &pq, double4 &vq) {
        if (a.jpos) {
            pq = a.jpos[q];
            vq = a.jvel[q];
        } else {
            pq = make_double4(a.xj[q], a.yj[q], a.zj[q], a.mj[q]);
            vq = make_double4(a.vxj[q], a.vyj[q], a.vzj[q], 0.0);
        }
    };
    double4 pq = make_double4(0.0, 0.0, 0.0, 0.0), vq = pq;
    int qn = 0;
    if (tl.T > 0) {
        const int k0 = tl.entry(0, lane), k1 = tl.entry(1, lane);
        load(k0 < tl.nj ? a.L.idx[lb + k0] : 0, pq, vq);
        qn = (tl.T > 1 && k1 < tl.nj) ? a.L.idx[lb + k1] : 0;
    }
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        float nj1 = 0.f;
        if (k < tl.nj) {
            const double x = pq.x, y = pq.y, z = pq.z;
            const float ox = (float) (x - fr.cx), oy = (float) (y - fr.cy), oz = (float) (z - fr.cz);
            sj[lane] = DrvtSlot{x, y, z, pq.w, vq.x, vq.y, vq.z, 0.0,
                                make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)))};
            nj1 = norm1_(ox, oy, oz);
        }
        if (t + 1 < tl.T) {
            load(qn, pq, vq);
            const int k2 = tl.entry(t + 2, lane);
            qn = (t + 2 < tl.T && k2 < tl.nj) ? a.L.idx[lb + k2] : 0;
        }
        const float nj1max = wave_maxf_(nj1);  // every lane takes part: outside the condition
        const float negc = act ? -dot_cut_(Hc, wi, n1i, nj1max) : INFINITY;
        const float pnegc = xor32_f(negc);
        __syncthreads();
        const uint64_t mask = mfma_dot_mask_f([&](int e) { return sj[e].o; }, tl.count(t), tx, ty, tz, negc, ptx, pty,
                                              ptz, pnegc);
        for (uint64_t msk = mask; msk; msk &= msk - 1) {
            const int jj = __builtin_ctzll(msk);
            const DrvtSlot &J = sj[jj];
            const double drx = xi - J.x, dry = yi - J.y, drz = zi - J.z;
            const double dvx = vxi - J.vx, dvy = vyi - J.vy, dvz = vzi - J.vz;
            const double r2 = norm2_(drx, dry, drz);
            const double r = dist_(r2).r;
            const double s = r * invH;
            const double o = fmax(1.0 - s, 0.0);
            const double o7 = pow7f_(o);
            const double mf = J.m * (dwds_(s, o7, PK) * rcp_(fma(r, H, reg)));  // m_j gradW / (K dr)
            const double dvdr = fma(dvz, drz, fma(dvy, dry, dvx * drx));
            divv = fma(mf, dvdr, divv);                     // -divv / K: sum of m_j dv . gradW / K
            rx = fma(mf, fma(dvy, drz, -(dvz * dry)), rx);  // -rot / K: sum of m_j dv x gradW / K
            ry = fma(mf, fma(dvz, drx, -(dvx * drz)), ry);
            rz = fma(mf, fma(dvx, dry, -(dvy * drx)), rz);
            gsw = fma(J.m, wpoly_(s, PK) * (o7 * o), gsw);  // sum of m_j W / norm
            gsf = fma(mf, r2, gsf);                       // sum of m_j dr . gradW / K
        }
        __syncthreads();
    }
    if (act) {
        // divv = -sum m_j dv.gradW, rot = -sum m_j dv x gradW,
        // gs = -sum m_j / h (3 W + dr.gradW)  (src/force.h:95-101)
        const double K = -22.0 * norm;
        divv *= -K;
        rx *= -K;
        ry *= -K;
        rz *= -K;
        double gs = -(3.0 * norm * gsw + K * gsf) / hi;
        gs = 1.0 / (1.0 + hi * gs / (3.0 * rhoi));
        divv /= rhoi;
        const double inv = 1.0 / rhoi;
        rx *= inv;
        ry *= inv;
        rz *= inv;
        a.divv[ip] = divv;
        a.rotx[ip] = rx;
        a.roty[ip] = ry;
        a.rotz[ip] = rz;
        a.gs[ip] = gs;
        if (a.csi)
            a.bal[ip] = fabs(divv) / (fabs(divv) + sqrt(rx * rx + ry * ry + rz * rz) + 1.0e-4 * a.csi[ip] / hi);
    }
}
__global__ __launch_bounds__(kWave) void k_drvt(DrvtArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) drvt_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave) void k_drvt_persistent(DrvtArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        drvt_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ---------------------------------------------------------------- hydro force
// CalcHydroForce::operator() (src/force.h:112-147), `#if 1` branch.  v_sig_max
// runs over every j of the list; the force terms only inside either support.
using HydroJ = HydroRec;  // per-j record of a tile (LDS)

__device__ __forceinline__ HydroRec make_hydro_rec(double x, double y, double z, double vx, double vy, double vz,
                                                   double m, double rho, double cs, double bal, double gs, double p,
                                                   double h) {
    HydroRec r;
    r.x = x; r.y = y; r.z = z;
    r.vx = vx; r.vy = vy; r.vz = vz;
    r.m = m;
    r.rho = rho;
    r.cs = cs;
    r.bal = bal;
    r.gs = gs;
    r.A = gs * p / (rho * rho);
    r.reg = 1.0e-6 * h;
    r.H = kSupport * h;
    r.norm = wnorm_(r.H);
    r.invH = 1.0 / r.H;
    return r;
}

__global__ void k_pack_kin(int64_t n, const double *x, const double *y, const double *z, const double *m,
                           const double *vx, const double *vy, const double *vz, double4 *pos, double4 *vel,
                           const double *h, double *hcopy) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) {
        pos[k] = make_double4(x[k], y[k], z[k], m[k]);
        vel[k] = make_double4(vx[k], vy[k], vz[k], 0.0);
        if (hcopy) hcopy[k] = h[k];
    }
}

__global__ void k_pack_hydro(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                             const double *vy, const double *vz, const double *m, const double *rho,
                             const double *cs, const double *bal, const double *gs, const double *p, const double *h,
                             HydroRec *out, VsigRec *vout) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    out[k] = make_hydro_rec(x[k], y[k], z[k], vx[k], vy[k], vz[k], m[k], rho[k], cs[k], bal[k], gs[k], p[k], h[k]);
    if (vout) vout[k] = VsigRec{x[k], y[k], z[k], vx[k], vy[k], vz[k], cs[k], 0.0};
}
__global__ void k_pack_hydro_f(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                               const double *vy, const double *vz, const double *m, const double *rho,
                               const double *cs, const double *bal, const double *gs, const double *p,
                               const double *h, HydroRecF *out, VsigRec *vout) {
    const int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (vout) vout[k] = VsigRec{x[k], y[k], z[k], vx[k], vy[k], vz[k], cs[k], 0.0};
    const HydroRec r = make_hydro_rec(x[k], y[k], z[k], vx[k], vy[k], vz[k], m[k], rho[k], cs[k], bal[k], gs[k], p[k],
                                      h[k]);
    HydroRecF f;
    f.x = r.x; f.y = r.y; f.z = r.z;
    f.vx = r.vx; f.vy = r.vy; f.vz = r.vz;
    f.k = make_float4((float) r.rho, (float) r.bal, (float) r.invH, (float) r.H);
    f.w = make_float4((float) r.norm, (float) r.reg, (float) r.gs, (float) r.m);
    f.c = make_float4((float) r.cs, (float) r.A, 0.f, 0.f);
    out[k] = f;
}
// 4 waves per SIMD (128 registers, seven spilled outside the pair loops):
// measured 2.09 ms at 1M against 2.14 ms at the compiler's 3 waves.  P1: the
// kernel also forms v_sig_max over the list and writes dt (else k_vsig does;
// without pass 1 the kernel needs 126 registers and spills none).
template <bool P1>
__device__ __forceinline__ void hydro_group(const HydroArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const PolyK PK;  // polynomial constants in registers
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xi = a.xi[ip], yi = a.yi[ip], zi = a.zi[ip];
    const double vxi = a.vxi[ip], vyi = a.vyi[ip], vzi = a.vzi[ip];
    const double hi = a.hi[ip], rhoi = a.rhoi[ip], pi = a.pi[ip], csi = a.csi[ip], gsi = a.gsi[ip],
                 bali = a.bali[ip];
    const double Hi = kSupport * hi;
    const double invHi = 1.0 / Hi;
    const double normi = wnorm_(Hi);
    const double regi = 1.0e-6 * hi;
    const double Ai = gsi * pi / (rhoi * rhoi);
    const int64_t lb = list_begin(a.L, g);
    const Tiles tl(a.L.cnt[g]);
    const Frame fr = group_frame(xi, yi, zi, act);
    const float oxi = (float) (xi - fr.cx), oyi = (float) (yi - fr.cy), ozi = (float) (zi - fr.cz);
    // dot-form support tests (see mfma_dot_mask_sym): the lane's own threshold per
    // tile from dot_cut_, j's threshold Q_j with the group's largest |o_i|_1
    const float Hci = (float) Hi * 1.0001f;
    const float wi = fmaf(oxi, oxi, fmaf(oyi, oyi, ozi * ozi)), n1i = norm1_(oxi, oyi, ozi);
    const float n1max = wave_maxf_(act ? n1i : 0.f);
    const float tx = -2.f * oxi, ty = -2.f * oyi, tz = -2.f * ozi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    // v_sig bound of pass 1: with v0 the centre of the group's velocity box,
    // c_i + c_j - 3w <= c_i + c_j + 3|v_i - v_j| <= (c_i + 3|v_i - v0|) + (c_j + 3|v_j - v0|)
    const double cvx = 0.5 * (wave_min_(act ? vxi : INFINITY) + wave_max_(act ? vxi : -INFINITY));
    const double cvy = 0.5 * (wave_min_(act ? vyi : INFINITY) + wave_max_(act ? vyi : -INFINITY));
    const double cvz = 0.5 * (wave_min_(act ? vzi : INFINITY) + wave_max_(act ? vzi : -INFINITY));
    const double bi = csi + 3.0 * sqrt((vxi - cvx) * (vxi - cvx) + (vyi - cvy) * (vyi - cvy) + (vzi - cvz) * (vzi - cvz));
    double ax = 0.0, ay = 0.0, az = 0.0, du = 0.0, vmax = 0.0;
    __shared__ HydroSlot sj[kWave];
    __shared__ float sQ[kWave];  // threshold of j's support
    // the list index of tile t+1 is in flight while tile t is evaluated (the
    // j record itself would hold 32 more registers through the pair loops)
    int qn = (tl.T > 0 && tl.entry(0, lane) < tl.nj) ? a.L.idx[lb + tl.entry(0, lane)] : 0;
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        float nj1 = 0.f;
        double bj = 0.0;
        const int q = qn;
        if (t + 1 < tl.T && tl.entry(t + 1, lane) < tl.nj) qn = a.L.idx[lb + tl.entry(t + 1, lane)];
        if (k < tl.nj) {
            const HydroJ r = a.jrec ? a.jrec[q]
                                    : make_hydro_rec(a.xj[q], a.yj[q], a.zj[q], a.vxj[q], a.vyj[q], a.vzj[q], a.mj[q],
                                                     a.rhoj[q], a.csj[q], a.balj[q], a.gsj[q], a.pj[q], a.hj[q]);
            sj[lane].r = r;
            if (P1) {
                const double ux = r.vx - cvx, uy = r.vy - cvy, uz = r.vz - cvz;
                bj = r.cs + 3.0 * sqrt(ux * ux + uy * uy + uz * uz);
            }
            const float ox = (float) (r.x - fr.cx), oy = (float) (r.y - fr.cy), oz = (float) (r.z - fr.cz);
            sj[lane].o = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            nj1 = norm1_(ox, oy, oz);
            sQ[lane] = dot_cut_j_((float) r.H * 1.0001f, nj1, n1max);
        }
        const float nj1max = wave_maxf_(nj1);  // every lane takes part: outside the condition
        const float negc = act ? -dot_cut_(Hci, wi, n1i, nj1max) : INFINITY;
        const float pnegc = xor32_f(negc);
        const float dsym = act ? wi - negc : 0.f, pdsym = xor32_f(dsym);
        // the entries that can raise some lane's maximum: b_j >= T, T the least
        // over the lanes of vmax_i (1 - 1e-10) - b_i.  For every other entry
        // b_i + b_j stays below each lane's maximum by a margin far above the
        // rounding of either side, so pass 1 visits only these (none: the tile
        // is skipped)
        const double T = P1 ? dpp_wave_min(act ? vmax * (1.0 - 1.0e-10) - bi : INFINITY) : 0.0;
        const uint64_t cand = P1 ? __ballot(k < tl.nj && !(bj < T)) : 0ULL;
        __syncthreads();
        const int nt = tl.count(t);
        // pass 1, every entry that may matter: v_sig_max, with w = dv.dr / |dr|
        // when negative (else 0) formed as in pass 2.  Evaluated for every lane
        // without branches: the earlier form, which skipped the square root
        // where a bound on |w| kept the maximum, spent more on its divergent
        // branches than it saved (FP64 hydro 1.96 -> 1.73 ms at 1M).  Two
        // entries per pass, two chains in flight (an odd last one is taken
        // twice, which the maximum absorbs): 1.72 -> 1.69 ms.
        auto vsig = [&](const HydroJ &J) {
            const double drx = xi - J.x, dry = yi - J.y, drz = zi - J.z;
            const double dvx = vxi - J.vx, dvy = vyi - J.vy, dvz = vzi - J.vz;
            const double dvdr = fma(dvz, drz, fma(dvy, dry, dvx * drx));
            const double w = fmin(dvdr, 0.0) * dist_(norm2_(drx, dry, drz)).rinv;
            return fma(-3.0, w, csi + J.cs);
        };
        for (uint64_t cm = cand; cm;) {
            const int j1 = __builtin_ctzll(cm);
            cm &= cm - 1;
            const int j2 = cm ? __builtin_ctzll(cm) : j1;
            if (cm) cm &= cm - 1;
            const double v1 = vsig(sj[j1].r), v2 = vsig(sj[j2].r);
            vmax = dmax(vmax, dmax(v1, v2));
        }
        // pass 2: the force terms of the pairs inside either support.  Both
        // gradients are multiples of dr: gradW_i = K F_i dr with K = -22 and
        // F_i = norm_i s(1 + 7s + 16s^2)(1-s)^7 / (r H_i + 1e-6 h_i); the two
        // reciprocals come from one, 1/(a_i a_j).  AVh is half the viscosity
        // AV of the reference, and the sums leave out K.
        const uint64_t mask = mfma_dot_mask_sym([&](int e) { return sj[e].o; }, [&](int e) { return sQ[e]; }, nt, tx,
                                                ty, tz, negc, dsym, ptx, pty, ptz, pnegc, pdsym);
        // one pair's terms: C (the acceleration term is C dr) and U (the du/dt term is U dv.dr)
        auto pair = [&](const HydroJ &J, double &drx, double &dry, double &drz, double &dvdr, double &C, double &U) {
            drx = xi - J.x;
            dry = yi - J.y;
            drz = zi - J.z;
            const double dvx = vxi - J.vx, dvy = vyi - J.vy, dvz = vzi - J.vz;
            dvdr = fma(dvz, drz, fma(dvy, dry, dvx * drx));
            const Dist d = dist_(norm2_(drx, dry, drz));
            const double w = fmin(dvdr, 0.0) * d.rinv;  // dv.dr / r when negative, else 0
            const double v_sig = fma(-3.0, w, csi + J.cs);
            const double AVh = ((-0.25 * kAV) * (v_sig * w)) * ((bali + J.bal) * rcp_(rhoi + J.rho));
            const double si = d.r * invHi, sj = d.r * J.invH;
            const double ai = fma(d.r, Hi, regi), aj = fma(d.r, J.H, J.reg);
            const double rab = rcp_(ai * aj);
            const double Fi = (normi * dwds_(si, pow7f_(fmax(1.0 - si, 0.0)), PK)) * (aj * rab);
            const double Fj = (J.norm * dwds_(sj, pow7f_(fmax(1.0 - sj, 0.0)), PK)) * (ai * rab);
            const double Fm = fma(Fj, J.gs, Fi * gsi);  // 2 (gradW_i gs_i + gradW_j gs_j) / 2 / (K dr)
            const double mj = J.m;
            C = mj * fma(AVh, Fm, fma(J.A, Fj, Ai * Fi));
            U = (mj * (Ai + AVh)) * Fm;
        };
        // two pairs per pass (two independent chains in flight; measured
        // 2.00 -> 1.97 ms at 1M), summed in mask order: a lane whose count is
        // odd evaluates its last pair twice and adds it once
        for (uint64_t msk = mask; msk;) {
            const int j1 = __builtin_ctzll(msk);
            msk &= msk - 1;
            const bool two = msk != 0;
            const int j2 = two ? __builtin_ctzll(msk) : j1;
            if (two) msk &= msk - 1;
            double x1, y1, z1, p1, C1, U1, x2, y2, z2, p2, C2, U2;
            pair(sj[j1].r, x1, y1, z1, p1, C1, U1);
            pair(sj[j2].r, x2, y2, z2, p2, C2, U2);
            ax = fma(C1, x1, ax);
            ay = fma(C1, y1, ay);
            az = fma(C1, z1, az);
            du = fma(U1, p1, du);
            if (two) {
                ax = fma(C2, x2, ax);
                ay = fma(C2, y2, ay);
                az = fma(C2, z2, az);
                du = fma(U2, p2, du);
            }
        }
        __syncthreads();
    }
    if (act) {
        // acc = -sum m_j (A_i gradW_i + A_j gradW_j + AV gradW_m) = 22 sum C dr,
        // du = sum m_j (A_i + AV/2) dv . gradW_m = -11 sum ...  (src/force.h:126-145)
        a.ax[ip] = 22.0 * ax;
        a.ay[ip] = 22.0 * ay;
        a.az[ip] = 22.0 * az;
        a.du[ip] = -11.0 * du;
        if (P1) a.dt[ip] = kCFL * 2.0 * hi / vmax;
    }
}
template <bool P1>
__global__ __launch_bounds__(kWave, 4) void k_hydro(HydroArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) hydro_group<P1>(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
template <bool P1>
__global__ __launch_bounds__(kWave, 4) void k_hydro_persistent(HydroArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        hydro_group<P1>(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ------------------------------------------------------------ v_sig maximum
// The dt of CalcHydroForce (src/force.h:122-124, 145): v_sig_max over every j
// of the list, i.e. pass 1 of the hydro kernel, as a kernel of its own.  It
// needs only positions, velocities and sound speeds, so the force evaluation
// runs it on its second stream beside the gravity kernel, and the hydro kernel
// (pass 2 alone) keeps its registers for the force terms.  Each tile is tested
// with the same bound, each candidate's v_sig is formed with the same
// arithmetic, and the maximum does not depend on the order: dt is the number
// the one-kernel form gives.
using VsigSlot = VsigRec;
// the entry's position, velocity and sound speed: four 16-byte loads of its
// dt record (none for q < 0)
__device__ __forceinline__ VsigSlot vsig_load(const HydroArgs &a, int q) {
    VsigSlot r{};
    if (q < 0) return r;
    const double2 *p = reinterpret_cast<const double2 *>(a.vrec + q);
    const double2 u = p[0], v = p[1], w = p[2], c = p[3];
    r.x = u.x; r.y = u.y; r.z = v.x; r.vx = v.y; r.vy = w.x; r.vz = w.y; r.cs = c.x;
    return r;
}
__device__ __forceinline__ void vsig_group(const HydroArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xi = a.xi[ip], yi = a.yi[ip], zi = a.zi[ip];
    const double vxi = a.vxi[ip], vyi = a.vyi[ip], vzi = a.vzi[ip];
    const double csi = a.csi[ip];
    // the bound of the hydro kernel's pass 1: v_sig <= b_i + b_j, b = c + 3|v - v0|
    const double cvx = 0.5 * (wave_min_(act ? vxi : INFINITY) + wave_max_(act ? vxi : -INFINITY));
    const double cvy = 0.5 * (wave_min_(act ? vyi : INFINITY) + wave_max_(act ? vyi : -INFINITY));
    const double cvz = 0.5 * (wave_min_(act ? vzi : INFINITY) + wave_max_(act ? vzi : -INFINITY));
    const double bi = csi + 3.0 * sqrt((vxi - cvx) * (vxi - cvx) + (vyi - cvy) * (vyi - cvy) + (vzi - cvz) * (vzi - cvz));
    const int64_t lb = list_begin(a.L, g);
    const Tiles tl(a.L.cnt[g]);  // interleaved as in the hydro kernel: 20 % fewer candidates than list order
    __shared__ VsigSlot sv[kWave];
    auto index_at = [&](int t) {
        const int k = tl.entry(t, lane);
        return (t < tl.T && k < tl.nj) ? a.L.idx[lb + k] : -1;
    };
    // list indices two tiles ahead, records one tile ahead
    VsigSlot rn = vsig_load(a, index_at(0));
    int qn = index_at(1);
    double vmax = 0.0;
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        const VsigSlot r = rn;
        const int q2 = index_at(t + 2);
        rn = vsig_load(a, qn);
        qn = q2;
        double cj = 0.0, u2 = 0.0;
        if (k < tl.nj) {
            sv[lane] = r;
            const double ux = r.vx - cvx, uy = r.vy - cvy, uz = r.vz - cvz;
            u2 = ux * ux + uy * uy + uz * uz;
            cj = r.cs;
        }
        // the entry is a candidate unless c_j + 3|u_j| < T, tested without
        // the square root as e = T - c_j > 0 and 9|u_j|^2 < e^2 (the 1e-10
        // margin of T covers the rounding of either form)
        const double T = dpp_wave_min(act ? vmax * (1.0 - 1.0e-10) - bi : INFINITY);
        const double e = T - cj;
        const uint64_t cand = __ballot(k < tl.nj && !(e > 0.0 && 9.0 * u2 < e * e));
        __syncthreads();
        auto vsig = [&](const VsigSlot &J) {
            const double drx = xi - J.x, dry = yi - J.y, drz = zi - J.z;
            const double dvx = vxi - J.vx, dvy = vyi - J.vy, dvz = vzi - J.vz;
            const double dvdr = fma(dvz, drz, fma(dvy, dry, dvx * drx));
            const double w = fmin(dvdr, 0.0) * dist_(norm2_(drx, dry, drz)).rinv;
            return fma(-3.0, w, csi + J.cs);
        };
        // four candidates per pass (a short tail repeats the first, which the
        // maximum absorbs); fmax is std::max here (v_sig >= +0, vmax never
        // NaN).  Measured at 1M beside the gravity kernel, FP64 step: two per
        // pass 7.54 ms, four 7.47 ms (96 registers: fewer waves of it share
        // the compute units with gravity; capping them further through LDS
        // made it slower, 7.56 and 7.97 ms)
        for (uint64_t cm = cand; cm;) {
            const int j1 = __builtin_ctzll(cm);
            cm &= cm - 1;
            const int j2 = cm ? __builtin_ctzll(cm) : j1;
            if (cm) cm &= cm - 1;
            const int j3 = cm ? __builtin_ctzll(cm) : j1;
            if (cm) cm &= cm - 1;
            const int j4 = cm ? __builtin_ctzll(cm) : j1;
            if (cm) cm &= cm - 1;
            const double v1 = vsig(sv[j1]), v2 = vsig(sv[j2]), v3 = vsig(sv[j3]), v4 = vsig(sv[j4]);
            vmax = fmax(fmax(vmax, fmax(v1, v2)), fmax(v3, v4));
        }
        __syncthreads();
    }
    if (act) a.dt[ip] = kCFL * 2.0 * a.hi[ip] / vmax;
}
__global__ __launch_bounds__(kWave) void k_vsig(HydroArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) vsig_group(a, xcd_remap((int) blockIdx.x, G));
}

// ---------------------------------------------------------------- gravity
// CalcGravityForce<EPJ::Grav> then CalcGravityForce<SPJMonopole>
// (src/force.h:154-168): the particle list first, then the monopoles.  The
// reference's 1.0 / sqrt(r^2 + eps^2) is rsqrt_ (within an ulp or so of the
// quotient of the correctly rounded square root), r^2 + eps^2 and the
// accumulations are fused multiply-adds, and G m_j is formed once per j when
// the tile is staged.  Each changes results at the rounding level only.
__device__ __forceinline__ void grav_tile_(const double4 *sp, int nt, double xi, double yi, double zi, double eps2,
                                           double &ax, double &ay, double &az, double &pot) {
    // eight interactions in flight: their chains (sum of squares, reciprocal
    // square root and its refinement, force) are long, and the ILP measured
    // faster than occupancy (114 registers at 4 waves per SIMD: 2.00 ms at 1M,
    // against 2.09 ms unrolled by 2 at 8 waves and 2.08 ms by 16 at 8 waves)
    // The force needs x^(-3/2), x = r^2 + eps^2, and the potential x^(-1/2) =
    // x^(-3/2) x: from the hardware estimate y of x^(-1/2), x^(-3/2) =
    // y^3 (1 - e)^(-3/2) with e = 1 - x y^2, whose series 1 + 3e/2 + 15e^2/8
    // leaves O(e^3) ~ 1e-21 (|e| <= 2^-23): 18 instructions per interaction
    // against 19 through a refined 1/r.
    // The series' constants stay in registers: v_fma_f64 takes no literal
    // operand, and the compiler otherwise rebuilds 1.5 with two moves in every
    // interaction (20 instructions instead of 18).  The empty asm statements
    // emit nothing; they only keep the values opaque to constant folding.
    double c15 = 1.5, c1875 = 1.875;
    asm volatile("" : "+v"(c15));
    asm volatile("" : "+s"(c1875));
#pragma unroll 8
    for (int jj = 0; jj < nt; ++jj) {
        const double4 pj = sp[jj];  // x, y, z, G m
        const double drx = xi - pj.x, dry = yi - pj.y, drz = zi - pj.z;
        const double x2 = fma(drz, drz, fma(dry, dry, fma(drx, drx, eps2)));
        const double y = __builtin_amdgcn_rsq(x2);
        const double y2 = y * y;
        const double e = fma(-x2, y2, 1.0);
        const double y3 = y2 * y;
        const double g3 = pj.w * fma(y3, e * fma(c1875, e, c15), y3);  // G m_j / (r^2 + eps^2)^(3/2)
        ax = fma(-g3, drx, ax);
        ay = fma(-g3, dry, ay);
        az = fma(-g3, drz, az);
        pot = fma(-g3, x2, pot);  // G m_j / (r^2 + eps^2)^(1/2)
    }
}

// j position and mass of list entry q of one pass, from the packed record
// when there is one
__device__ __forceinline__ double4 grav_load_(const double4 *rec, const double *X, const double *Y, const double *Z,
                                              const double *M, int q) {
    return rec ? rec[q] : make_double4(X[q], Y[q], Z[q], M[q]);
}

__device__ __forceinline__ void grav_group(const GravArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xi = a.xi[ip], yi = a.yi[ip], zi = a.zi[ip];
    double eps2;
    if (a.eps2i) eps2 = a.eps2i[ip];
    else { const double h = a.hi[ip]; eps2 = h * h; }
    double ax = 0.0, ay = 0.0, az = 0.0, pot = 0.0;
    __shared__ double4 sp[kWave];
    for (int pass = 0; pass < 2; ++pass) {
        if (pass && !a.ms && !a.srec) break;
        const ListRef &L = pass ? a.Ls : a.Lj;
        const double4 *R = pass ? a.srec : a.jrec;
        const double *X = pass ? a.xs : a.xj, *Y = pass ? a.ys : a.yj, *Z = pass ? a.zs : a.zj,
                     *Mm = pass ? a.ms : a.mj;
        const int64_t lb = list_begin(L, g);
        const int nj = L.cnt[g];
        // Software-pipelined tiles: the record of tile t+1 and the list index
        // of tile t+2 are in flight while tile t is evaluated (a lane without
        // an entry loads entry 0's record, which exists whenever nj > 0).
        double4 r = make_double4(0.0, 0.0, 0.0, 0.0);
        int qn = 0;
        if (nj > 0) {
            r = grav_load_(R, X, Y, Z, Mm, lane < nj ? L.idx[lb + lane] : L.idx[lb]);
            qn = (kWave + lane < nj) ? L.idx[lb + kWave + lane] : L.idx[lb];
        }
        for (int t = 0; t < nj; t += kWave) {
            if (t + lane < nj) sp[lane] = make_double4(r.x, r.y, r.z, kG * r.w);
            if (t + kWave < nj) {
                r = grav_load_(R, X, Y, Z, Mm, qn);
                qn = (t + 2 * kWave + lane < nj) ? L.idx[lb + t + 2 * kWave + lane] : L.idx[lb];
            }
            __syncthreads();
            grav_tile_(sp, min(kWave, nj - t), xi, yi, zi, eps2, ax, ay, az, pot);
            __syncthreads();
        }
    }
    if (act) {
        a.ax[ip] = ax;
        a.ay[ip] = ay;
        a.az[ip] = az;
        a.pot[ip] = pot;
    }
}
__global__ __launch_bounds__(kWave) void k_grav(GravArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) grav_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave) void k_grav_persistent(GravArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        grav_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

}  // namespace

void launch_pack_kin(int64_t n, const double *x, const double *y, const double *z, const double *m, const double *vx,
                     const double *vy, const double *vz, double4 *pos, double4 *vel, hipStream_t s, const double *h,
                     double *hcopy) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_kin, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, n, x, y, z, m, vx, vy, vz, pos,
                       vel, h, hcopy);
    SPH_LAUNCH_CHECK();
}
void launch_pack_hydro(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                       const double *vy, const double *vz, const double *m, const double *rho, const double *cs,
                       const double *bal, const double *gs, const double *p, const double *h, HydroRec *out,
                       hipStream_t s, VsigRec *vout) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_hydro, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, n, x, y, z, vx, vy, vz, m,
                       rho, cs, bal, gs, p, h, out, vout);
    SPH_LAUNCH_CHECK();
}
void launch_pack_hydro_f(int64_t n, const double *x, const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz, const double *m, const double *rho, const double *cs,
                         const double *bal, const double *gs, const double *p, const double *h, HydroRecF *out,
                         hipStream_t s, VsigRec *vout) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_pack_hydro_f, dim3((unsigned) ((n + 255) / 256)), dim3(256), 0, s, n, x, y, z, vx, vy, vz,
                       m, rho, cs, bal, gs, p, h, out, vout);
    SPH_LAUNCH_CHECK();
}
namespace {
__global__ void k_sched_reset(int32_t *sched) {
    if (threadIdx.x < 8) sched[threadIdx.x] = 0;
}
}  // namespace

void launch_sched_reset(int32_t *sched, hipStream_t s) {
    hipLaunchKernelGGL(k_sched_reset, dim3(1), dim3(kWave), 0, s, sched);
    SPH_LAUNCH_CHECK();
}

int persistent_grid(const void *kernel, int ngroups) {
    // resident workgroups per (device, kernel), looked up once
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, int> cache;
    int dev = 0;
    SPH_HIP(hipGetDevice(&dev));
    int grid;
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find({dev, kernel});
        if (it == cache.end()) {
            int ncu = 0, per_cu = 0;
            SPH_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
            SPH_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kWave, 0));
            it = cache.emplace(std::make_pair(dev, kernel), std::max(8, (ncu * std::max(per_cu, 1)) & ~7)).first;
        }
        grid = it->second;
    }
    return std::min(grid, (ngroups + 7) & ~7);
}

template <class A>
static void launch_groups(void (*k)(A), void (*kp)(A), const A &a, hipStream_t s) {
    if (a.ngroups <= 0) return;
    if (a.sched) {
        launch_sched_reset(a.sched, s);
        hipLaunchKernelGGL(kp, dim3(persistent_grid((const void *) kp, a.ngroups)), dim3(kWave), 0, s, a);
    } else {
        hipLaunchKernelGGL(k, dim3(a.ngroups), dim3(kWave), 0, s, a);
    }
    SPH_LAUNCH_CHECK();
}

void launch_dens_fp64(const DensArgs &a, hipStream_t s) { launch_groups(k_dens, k_dens_persistent, a, s); }
void launch_drvt_fp64(const DrvtArgs &a, hipStream_t s) { launch_groups(k_drvt, k_drvt_persistent, a, s); }
void launch_hydro_fp64(const HydroArgs &a, hipStream_t s) {
    if (a.pass1) launch_groups(k_hydro<true>, k_hydro_persistent<true>, a, s);
    else launch_groups(k_hydro<false>, k_hydro_persistent<false>, a, s);
}
// one workgroup per group (no persistent form: the persistent counters belong
// to the kernels of the main stream)
void launch_vsig(const HydroArgs &a, hipStream_t s) {
    if (a.ngroups <= 0) return;
    if (!a.vrec) throw Error(-1, "launch_vsig needs the dt records (vrec)");  // -1: SPHGPU_ERR_ARG
    hipLaunchKernelGGL(k_vsig, dim3(a.ngroups), dim3(kWave), 0, s, a);
    SPH_LAUNCH_CHECK();
}
void launch_grav_fp64(const GravArgs &a, hipStream_t s) { launch_groups(k_grav, k_grav_persistent, a, s); }

}  // namespace sphgpu
