// kernels_mixed.hip -- mixed-precision interaction kernels (the precision form).
//
// Positions and velocities stay FP64 in HBM.  For each i-group the wave takes
// the centre of the group's position box (and of its velocity box) in FP64;
// every i and j position / velocity enters the pair loop as an FP32 offset
// from that centre, formed in FP64 and rounded once, so dr and dv carry
// relative errors of order 2^-24 of the group's extent rather than of the
// absolute coordinates.  Pair arithmetic is FP32 with the hardware rsqrt and
// reciprocal; the per-i kernel constants (1/H, norm) are formed in FP64 and
// rounded once.  The error floor of this form is the FP32 rounding of the
// pair offsets themselves: half an ulp of s = r/H, amplified by
// |d ln W / d ln s| (2 to 13 over the support), gives about 1e-7 rms relative
// error in the density sums.  The kernel polynomials use the algebraically equal forms
//   W'(s) = -22 s (1 + 7 s + 16 s^2) (1 - s)^7 * norm
// that avoid the cancellation of the reference's (src/kernel.h:28-29) form.
// Sums: each 64-j tile is accumulated in FP32 and the tile sums are combined
// with Kahan compensation.  Finalisation (density floor, smoothing length,
// grad-h factor, Balsara switch, dt) is FP64.
#include "internal.hpp"

namespace sphgpu {
namespace {

struct Kahan {
    float s = 0.f, c = 0.f;
    __device__ __forceinline__ void add(float x) {
        const float y = x - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double) s - (double) c; }
};

__device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }

// group-level reductions (once per group) through LDS permutes; the per-tile
// ones take the VALU's lane paths (dpp_wave_*, xor32_f)
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
__device__ __forceinline__ float wave_maxf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, kWave));
    return v;
}
__device__ __forceinline__ float wave_minf(float v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, kWave));
    return v;
}
// centre of the box spanned by the active lanes' values
__device__ __forceinline__ double wave_centre(double v, bool act) {
    const double lo = wave_min(act ? v : INFINITY);
    const double hi = wave_max(act ? v : -INFINITY);
    return 0.5 * (lo + hi);
}

constexpr double kNormD = 1365.0 / 64.0 / kPi;

// Per-particle kernel constants, formed in FP64 and rounded once: every pair
// term of particle i is scaled by norm, so an error there would bias the
// whole sum rather than average out over the pairs.
struct KernelConsts {
    float H, invH, norm;
};
__device__ __forceinline__ KernelConsts kernel_consts(double h) {
    const double Hd = kSupport * h;
    return {(float) Hd, (float) (1.0 / Hd), (float) (kNormD / (Hd * Hd * Hd))};
}

// j position and mass, from the packed record when there is one
__device__ __forceinline__ void load_xyzm(const DensArgs &a, int q, double &x, double &y, double &z, double &m) {
    if (a.jpos) {
        const double4 r = a.jpos[q];
        x = r.x; y = r.y; z = r.z; m = r.w;
    } else {
        x = a.xj[q]; y = a.yj[q]; z = a.zj[q]; m = a.mj[q];
    }
}

// ---------------------------------------------------------------- density
// At least 5 waves per SIMD (no spills; measured faster than 4).
__device__ __forceinline__ void dens_mixed_group(const DensArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xd = a.xi[ip], yd = a.yi[ip], zd = a.zi[ip];
    const double cx = wave_centre(xd, act), cy = wave_centre(yd, act), cz = wave_centre(zd, act);
    const float xi = (float) (xd - cx), yi = (float) (yd - cy), zi = (float) (zd - cz);
    const double hid = a.hi[ip];
    const KernelConsts kc = kernel_consts(hid);
    const float H = kc.H, invH = kc.invH, norm = kc.norm;
    // mask constants (see mfma_dot_mask): inside when |o_j - o_i|^2 < H^2 (1 + 1e-4)
    const float negc = act ? fmaf(xi, xi, fmaf(yi, yi, zi * zi)) - H * H * 1.0001f : INFINITY;
    const float tx = -2.f * xi, ty = -2.f * yi, tz = -2.f * zi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    const float pnegc = __shfl_xor(negc, 32, kWave);
    Kahan dens;
    int ngb = 0;
    const int64_t lb = list_begin(a.L, g);
    const int nj = a.L.cnt[g];
    // Software-pipelined tiles: while tile t is evaluated from LDS buffer t&1,
    // the particle loads of tile t+1 and the list index of tile t+2 are in
    // flight (their first use comes after the evaluation).
    __shared__ float4 so[2][kWave];  // FP32 offset from the centre, |offset|^2
    __shared__ float sm[2][kWave];
    const Tiles tl(nj);
    int qn = 0;  // list entry of tile t+1 for this lane (0 if none)
    double px = 0, py = 0, pz = 0, pm = 0;  // particle data of tile t (lane's entry)
    if (tl.T > 0) {
        const int k0 = tl.entry(0, lane);
        const int q0 = (k0 < nj) ? a.L.idx[lb + k0] : 0;
        load_xyzm(a, q0, px, py, pz, pm);
        const int k1 = tl.entry(1, lane);
        qn = (tl.T > 1 && k1 < nj) ? a.L.idx[lb + k1] : 0;
    }
    for (int t = 0; t < tl.T; ++t) {
        const int b = t & 1;
        if (tl.entry(t, lane) < nj) {
            sm[b][lane] = (float) pm;
            const float ox = (float) (px - cx), oy = (float) (py - cy), oz = (float) (pz - cz);
            so[b][lane] = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
        }
        __syncthreads();
        if (t + 1 < tl.T) {  // issue tile t+1's particle loads and tile t+2's index load
            load_xyzm(a, qn, px, py, pz, pm);
            const int k2 = tl.entry(t + 2, lane);
            qn = (t + 2 < tl.T && k2 < nj) ? a.L.idx[lb + k2] : 0;
        }
        const int nt = tl.count(t);
        float part = 0.f;
        const uint64_t mask = mfma_dot_mask(so[b], nt, tx, ty, tz, negc, ptx, pty, ptz, pnegc);
        for (uint64_t msk = mask; msk; msk &= msk - 1) {
            const int jj = __builtin_ctzll(msk);
            const float4 oj = so[b][jj];
            const float dx = oj.x - xi, dy = oj.y - yi, dz = oj.z - zi;
            const float r2 = dx * dx + dy * dy + dz * dz;
            {
                const float r = r2 * rsq(r2 + 1e-30f);
                const float s = r * invH;
                const float om = fmaxf(1.f - s, 0.f);
                const float om2 = om * om, om4 = om2 * om2;
                const float w = (1.f + s * (8.f + s * (25.f + s * 32.f))) * (om4 * om4);
                part += sm[b][jj] * w;
                ngb += (om > 0.f) ? 1 : 0;
            }
        }
        dens.add(part * norm);
    }
    double hn = 0.0;  // lanes without a particle take no part in the check
    if (act) {
        const double d = dmax(5.0, dens.value());
        a.rho_out[ip] = d;
        hn = kSMTH * pow(a.mi[ip] / d, 1.0 / 3.0);
        a.h_out[ip] = hn;
        if (a.ngb_out) a.ngb_out[ip] = ngb;
    }
    // the positions again from memory (cached): keeping the FP64 ones live
    // through the tile loop cost a wave of occupancy
    if (a.flag) gather_cover_check(a.sbox, a.crad, a.flag, g, act, a.xi[ip], a.yi[ip], a.zi[ip], hn);
}
__global__ __launch_bounds__(kWave, 5) void k_dens_mixed(DensArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) dens_mixed_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave, 5) void k_dens_mixed_persistent(DensArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        dens_mixed_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ---------------------------------------------------------------- derivative
__device__ __forceinline__ void drvt_mixed_group(const DrvtArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xd = a.xi[ip], yd = a.yi[ip], zd = a.zi[ip];
    const double vxd = a.vxi[ip], vyd = a.vyi[ip], vzd = a.vzi[ip];
    const double cx = wave_centre(xd, act), cy = wave_centre(yd, act), cz = wave_centre(zd, act);
    const double cvx = wave_centre(vxd, act), cvy = wave_centre(vyd, act), cvz = wave_centre(vzd, act);
    const float xi = (float) (xd - cx), yi = (float) (yd - cy), zi = (float) (zd - cz);
    const float vxi = (float) (vxd - cvx), vyi = (float) (vyd - cvy), vzi = (float) (vzd - cvz);
    const double hid = a.hi[ip], rhoi = a.rhoi[ip];
    const KernelConsts kc = kernel_consts(hid);
    const float H = kc.H, invH = kc.invH, norm = kc.norm;
    const float reg = (float) (1.0e-6 * hid);
    const float negc = act ? fmaf(xi, xi, fmaf(yi, yi, zi * zi)) - H * H * 1.0001f : INFINITY;
    const float tx = -2.f * xi, ty = -2.f * yi, tz = -2.f * zi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    const float pnegc = __shfl_xor(negc, 32, kWave);
    Kahan divv, rx, ry, rz, gsum;
    const int64_t lb = list_begin(a.L, g);
    const int nj = a.L.cnt[g];
    __shared__ float4 so[kWave];  // FP32 offset from the centre, |offset|^2
    __shared__ float sm[kWave], svx[kWave], svy[kWave], svz[kWave];
    const Tiles tl(nj);
    // software-pipelined tiles: the j records of tile t+1 and the list index of
    // tile t+2 are in flight while tile t is evaluated (a lane without an entry
    // loads particle 0, which exists whenever the list is not empty)
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
        load(k0 < nj ? a.L.idx[lb + k0] : 0, pq, vq);
        qn = (tl.T > 1 && k1 < nj) ? a.L.idx[lb + k1] : 0;
    }
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        if (k < nj) {
            sm[lane] = (float) pq.w;
            const float ox = (float) (pq.x - cx), oy = (float) (pq.y - cy), oz = (float) (pq.z - cz);
            so[lane] = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            svx[lane] = (float) (vq.x - cvx);
            svy[lane] = (float) (vq.y - cvy);
            svz[lane] = (float) (vq.z - cvz);
        }
        if (t + 1 < tl.T) {
            load(qn, pq, vq);
            const int k2 = tl.entry(t + 2, lane);
            qn = (t + 2 < tl.T && k2 < nj) ? a.L.idx[lb + k2] : 0;
        }
        __syncthreads();
        const int nt = tl.count(t);
        float pd = 0.f, px = 0.f, py = 0.f, pz = 0.f, pg = 0.f;
        const uint64_t mask = mfma_dot_mask(so, nt, tx, ty, tz, negc, ptx, pty, ptz, pnegc);
        for (uint64_t msk = mask; msk; msk &= msk - 1) {
            const int jj = __builtin_ctzll(msk);
            const float4 oj = so[jj];
            const float drx = xi - oj.x, dry = yi - oj.y, drz = zi - oj.z;
            const float r2 = drx * drx + dry * dry + drz * drz;
            {
                const float dvx = vxi - svx[jj], dvy = vyi - svy[jj], dvz = vzi - svz[jj];
                const float r = r2 * rsq(r2 + 1e-30f);
                const float s = r * invH;
                const float om = fmaxf(1.f - s, 0.f);
                const float om2 = om * om, om4 = om2 * om2, om7 = om4 * om2 * om;
                const float w = (1.f + s * (8.f + s * (25.f + s * 32.f))) * (om4 * om4);
                // gradW = dr * W'(s) / (r H + 1e-6 h), W'(s) = -22 s (1+7s+16s^2) (1-s)^7 (norm folded in later)
                const float f = (-22.f * s * (1.f + s * (7.f + 16.f * s)) * om7) * rcp(r * H + reg);
                const float m = sm[jj];
                const float dvdr = dvx * drx + dvy * dry + dvz * drz;
                pd -= m * f * dvdr;
                // (-m dv) x (dr f)
                px -= m * f * (dvy * drz - dvz * dry);
                py -= m * f * (dvz * drx - dvx * drz);
                pz -= m * f * (dvx * dry - dvy * drx);
                pg -= m * (3.f * w + f * r2);
            }
        }
        divv.add(pd * norm);
        rx.add(px * norm);
        ry.add(py * norm);
        rz.add(pz * norm);
        gsum.add(pg * norm);
        __syncthreads();
    }
    if (act) {
        const double gs = 1.0 / (1.0 + hid * (gsum.value() / hid) / (3.0 * rhoi));
        const double dv = divv.value() / rhoi;
        const double inv = 1.0 / rhoi;
        const double ox = rx.value() * inv, oy = ry.value() * inv, oz = rz.value() * inv;
        a.divv[ip] = dv;
        a.rotx[ip] = ox;
        a.roty[ip] = oy;
        a.rotz[ip] = oz;
        a.gs[ip] = gs;
        if (a.csi) a.bal[ip] = fabs(dv) / (fabs(dv) + sqrt(ox * ox + oy * oy + oz * oz) + 1.0e-4 * a.csi[ip] / hid);
    }
}
__global__ __launch_bounds__(kWave) void k_drvt_mixed(DrvtArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) drvt_mixed_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave) void k_drvt_mixed_persistent(DrvtArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        drvt_mixed_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ---------------------------------------------------------------- hydro force
// One entry of a hydro tile in LDS, read by the pair loops as five 16-byte
// loads.  Its 80-byte stride puts the lanes' reads (each of its own j) on 8
// starting banks, so that the banks serve 8 lanes per cycle.
struct alignas(16) HydroSlotF {
    float4 o;  // FP32 offset from the centre, |offset|^2
    float4 v;  // velocity offset from the velocity centre, sound speed
    float4 k;  // rho, Balsara switch, 1/H, H
    float4 w;  // kernel normalisation, gradient regularisation, grad-h term, mass
    float4 a;  // A = f P / rho^2, -, -, -
};
static_assert(sizeof(HydroSlotF) == 80, "LDS record stride");

// 4 waves per SIMD: at 5 waves (96 registers) 27 spill, and the kernel took
// 1.33 ms at 1M against 1.20 ms at 4 waves (124 registers).  P1: the kernel
// also forms v_sig_max (in FP32) and writes dt; else the FP64 k_vsig does.
template <bool P1>
__device__ __forceinline__ void hydro_mixed_group(const HydroArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xd = a.xi[ip], yd = a.yi[ip], zd = a.zi[ip];
    const double vxd = a.vxi[ip], vyd = a.vyi[ip], vzd = a.vzi[ip];
    const double cx = wave_centre(xd, act), cy = wave_centre(yd, act), cz = wave_centre(zd, act);
    const double cvx = wave_centre(vxd, act), cvy = wave_centre(vyd, act), cvz = wave_centre(vzd, act);
    const float xi = (float) (xd - cx), yi = (float) (yd - cy), zi = (float) (zd - cz);
    const float vxi = (float) (vxd - cvx), vyi = (float) (vyd - cvy), vzi = (float) (vzd - cvz);
    const double hid = a.hi[ip], rhoid = a.rhoi[ip];
    const KernelConsts kc = kernel_consts(hid);
    const float Hi = kc.H, invHi = kc.invH, normi = kc.norm;
    const float regi = (float) (1.0e-6 * hid);
    // dot-form support tests from the matrix cores (mfma_dot_mask_sym), with
    // the thresholds that make them a superset of the in-support pairs
    const float Hci = (float) (kSupport * hid) * 1.0001f;
    const float wi = fmaf(xi, xi, fmaf(yi, yi, zi * zi)), n1i = norm1_(xi, yi, zi);
    const float n1max = wave_maxf(act ? n1i : 0.f);
    const float tx = -2.f * xi, ty = -2.f * yi, tz = -2.f * zi;
    const float ptx = __shfl_xor(tx, 32, kWave), pty = __shfl_xor(ty, 32, kWave), ptz = __shfl_xor(tz, 32, kWave);
    const float rhoi = (float) rhoid, csi = (float) a.csi[ip], bali = (float) a.bali[ip];
    const float fAi = (float) (a.gsi[ip] * a.pi[ip] / (rhoid * rhoid));  // f_i P_i / rho_i^2
    const float gsi = (float) a.gsi[ip];
    // v_sig bound: c_i + c_j - 3w <= (c_i + 3|v_i - v0|) + (c_j + 3|v_j - v0|), v0 the velocity centre
    const float bi = csi + 3.f * sqrtf(vxi * vxi + vyi * vyi + vzi * vzi);
    Kahan ax, ay, az, du;
    float vmax = 0.f;
    const int64_t lb = list_begin(a.L, g);
    const int nj = a.L.cnt[g];
    __shared__ HydroSlotF slot[kWave];
    __shared__ float sQ[kWave];
    const Tiles tl(nj);
    // the list index of tile t+1 is in flight while tile t is evaluated
    int qn = (tl.T > 0 && tl.entry(0, lane) < nj) ? a.L.idx[lb + tl.entry(0, lane)] : 0;
    for (int t = 0; t < tl.T; ++t) {
        const int k = tl.entry(t, lane);
        float bnd = 0.f, nj1 = 0.f;
        const int q = qn;
        if (t + 1 < tl.T && tl.entry(t + 1, lane) < nj) qn = a.L.idx[lb + tl.entry(t + 1, lane)];
        if (k < nj) {
            double x, y, z, vx, vy, vz;
            float4 ek, ew;  // (rho, Balsara, 1/H, H), (kernel normalisation, 1e-6 h, grad-h term, mass)
            float cjf, Ajf;
            if (a.jrecf) {  // one 96-byte record per entry, its FP32 fields converted once per particle
                const HydroRecF r = a.jrecf[q];
                x = r.x; y = r.y; z = r.z; vx = r.vx; vy = r.vy; vz = r.vz;
                ek = r.k;
                ew = r.w;
                cjf = r.c.x;
                Ajf = r.c.y;
            } else {
                double m, rj, cj, bj, gj, Aj, regj, Hj, nj3;
                if (a.jrec) {  // one 128-byte record per entry
                    const HydroRec r = a.jrec[q];
                    x = r.x; y = r.y; z = r.z; vx = r.vx; vy = r.vy; vz = r.vz; m = r.m;
                    rj = r.rho; cj = r.cs; bj = r.bal; gj = r.gs; Aj = r.A; regj = r.reg; Hj = r.H; nj3 = r.norm;
                } else {
                    x = a.xj[q]; y = a.yj[q]; z = a.zj[q]; vx = a.vxj[q]; vy = a.vyj[q]; vz = a.vzj[q]; m = a.mj[q];
                    rj = a.rhoj[q]; cj = a.csj[q]; bj = a.balj[q]; gj = a.gsj[q];
                    const double hj = a.hj[q];
                    regj = 1.0e-6 * hj;
                    Aj = gj * a.pj[q] / (rj * rj);
                    Hj = kSupport * hj;
                    nj3 = kNormD / (Hj * Hj * Hj);
                }
                ek = make_float4((float) rj, (float) bj, (float) (1.0 / Hj), (float) Hj);
                ew = make_float4((float) nj3, (float) regj, (float) gj, (float) m);
                cjf = (float) cj;
                Ajf = (float) Aj;
            }
            const float ox = (float) (x - cx), oy = (float) (y - cy), oz = (float) (z - cz);
            nj1 = norm1_(ox, oy, oz);
            sQ[lane] = dot_cut_j_(ek.w * 1.0001f, nj1, n1max);
            const float ux = (float) (vx - cvx), uy = (float) (vy - cvy), uz = (float) (vz - cvz);
            if (P1) bnd = cjf + 3.f * sqrtf(ux * ux + uy * uy + uz * uz);
            HydroSlotF e;
            e.o = make_float4(ox, oy, oz, fmaf(ox, ox, fmaf(oy, oy, oz * oz)));
            e.v = make_float4(ux, uy, uz, cjf);
            e.k = ek;
            e.w = ew;
            e.a = make_float4(Ajf, 0.f, 0.f, 0.f);
            slot[lane] = e;
        }
        const float nj1max = dpp_wave_maxf(nj1);  // every lane takes part: outside the condition
        const float negc = act ? -dot_cut_(Hci, wi, n1i, nj1max) : INFINITY;
        const float pnegc = xor32_f(negc);
        const float dsym = act ? wi - negc : 0.f, pdsym = xor32_f(dsym);
        // the entries that can raise some lane's v_sig maximum: bound_j >= T,
        // T the least over the lanes of vmax_i (1 - 1e-5) - b_i (a margin far
        // above the FP32 rounding); no other entry can change any maximum
        const float T = P1 ? wave_minf(act ? vmax * (1.f - 1.0e-5f) - bi : INFINITY) : 0.f;
        const uint64_t cand = P1 ? __ballot(k < nj && !(bnd < T)) : 0ULL;
        __syncthreads();
        const int nt = tl.count(t);
        float px = 0.f, py = 0.f, pz = 0.f, pu = 0.f;
        // pass 1: v_sig_max over the candidates, two per pass (two chains in
        // flight; an odd last one is taken twice): 1.11 -> 1.03 ms at 1M
        auto vsig = [&](int jj) {
            const float4 oj = slot[jj].o, vj = slot[jj].v;
            const float drx = xi - oj.x, dry = yi - oj.y, drz = zi - oj.z;
            const float dvx = vxi - vj.x, dvy = vyi - vj.y, dvz = vzi - vj.z;
            const float dvdr = dvx * drx + dvy * dry + dvz * drz;
            const float r2 = drx * drx + dry * dry + drz * drz;
            const float w = (dvdr < 0.f) ? dvdr * rsq(r2 + 1e-30f) : 0.f;
            return csi + vj.w - 3.f * w;
        };
        for (uint64_t cm = cand; cm;) {
            const int j1 = __builtin_ctzll(cm);
            cm &= cm - 1;
            const int j2 = cm ? __builtin_ctzll(cm) : j1;
            if (cm) cm &= cm - 1;
            const float v1 = vsig(j1), v2 = vsig(j2);
            vmax = fmaxf(vmax, fmaxf(v1, v2));
        }
        // the pairs inside either support (a superset: pairs beyond both give zero terms)
        const uint64_t mask = mfma_dot_mask_sym([&](int e) { return slot[e].o; }, [&](int e) { return sQ[e]; }, nt, tx,
                                                ty, tz, negc, dsym, ptx, pty, ptz, pnegc, pdsym);
        // pass 2: force terms of this lane's in-support pairs
        for (uint64_t msk = mask; msk; msk &= msk - 1) {
            const int jj = __builtin_ctzll(msk);
            const HydroSlotF &J = slot[jj];
            const float4 oj = J.o, vj = J.v, kj = J.k, wj = J.w;
            const float drx = xi - oj.x, dry = yi - oj.y, drz = zi - oj.z;
            const float dvx = vxi - vj.x, dvy = vyi - vj.y, dvz = vzi - vj.z;
            const float dvdr = dvx * drx + dvy * dry + dvz * drz;
            const float r2 = drx * drx + dry * dry + drz * drz;
            const float rinv = rsq(r2 + 1e-30f);
            const float w = (dvdr < 0.f) ? dvdr * rinv : 0.f;
            const float v_sig = csi + vj.w - 3.f * w;
            {
                const float r = r2 * rinv;
                const float AV = (-0.5f * v_sig * w) * rcp(rhoi + kj.x) * (bali + kj.y);
                // W'(s) factors for h_i and h_j
                const float si = r * invHi;
                const float oi = fmaxf(1.f - si, 0.f);
                const float oi2 = oi * oi, oi7 = oi2 * oi2 * oi2 * oi;
                const float fi = normi * (-22.f * si * (1.f + si * (7.f + 16.f * si)) * oi7) * rcp(r * Hi + regi);
                const float Hj = kj.w;
                const float sj = r * kj.z;
                const float oj = fmaxf(1.f - sj, 0.f);
                const float oj2 = oj * oj, oj7 = oj2 * oj2 * oj2 * oj;
                const float fj = wj.x * (-22.f * sj * (1.f + sj * (7.f + 16.f * sj)) * oj7) * rcp(r * Hj + wj.y);
                // gradW_i = dr fi, gradW_j = dr fj, mean = 0.5 (fi gs_i + fj gs_j) dr
                const float fm = 0.5f * (fi * gsi + fj * wj.z);
                const float m = wj.w;
                const float coef = m * (fAi * fi + J.a.x * fj + AV * fm);
                px -= coef * drx;
                py -= coef * dry;
                pz -= coef * drz;
                pu += m * (fAi + 0.5f * AV) * fm * dvdr;
            }
        }
        ax.add(px);
        ay.add(py);
        az.add(pz);
        du.add(pu);
        __syncthreads();
    }
    if (act) {
        a.ax[ip] = ax.value();
        a.ay[ip] = ay.value();
        a.az[ip] = az.value();
        a.du[ip] = du.value();
        if (P1) a.dt[ip] = kCFL * 2.0 * hid / (double) vmax;
    }
}
template <bool P1>
__global__ __launch_bounds__(kWave, 4) void k_hydro_mixed(HydroArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) hydro_mixed_group<P1>(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
template <bool P1>
__global__ __launch_bounds__(kWave, 4) void k_hydro_mixed_persistent(HydroArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        hydro_mixed_group<P1>(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

// ---------------------------------------------------------------- gravity
__device__ __forceinline__ void grav_tile_mixed(const float *sx, const float *sy, const float *sz, const float *sgm,
                                                int nt, float xi, float yi, float zi, float eps2, float &px, float &py,
                                                float &pz, float &pp) {
#pragma unroll 4
    for (int jj = 0; jj < nt; ++jj) {
        const float drx = xi - sx[jj], dry = yi - sy[jj], drz = zi - sz[jj];
        const float r2 = fmaf(drz, drz, fmaf(dry, dry, fmaf(drx, drx, eps2)));  // r^2 + eps^2 in three FMAs
        const float inv = rsq(r2);
        const float gm_inv = sgm[jj] * inv;
        const float gm_inv3 = gm_inv * inv * inv;
        px -= gm_inv3 * drx;
        py -= gm_inv3 * dry;
        pz -= gm_inv3 * drz;
        pp -= gm_inv;
    }
}

__device__ __forceinline__ void grav_mixed_group(const GravArgs &a, const int item) {
    const int lane = threadIdx.x;
    const int g = group_of(a, item);
    const int i0 = a.gstart[g] + kWave * (item - g * a.chunks), ni = min(kWave, a.gstart[g + 1] - i0);
    if (ni <= 0) return;  // a chunk beyond the group's particles
    const bool act = lane < ni;
    const int ip = i0 + (act ? lane : 0);
    const double xd = a.xi[ip], yd = a.yi[ip], zd = a.zi[ip];
    const double cx = wave_centre(xd, act), cy = wave_centre(yd, act), cz = wave_centre(zd, act);
    const float xi = (float) (xd - cx), yi = (float) (yd - cy), zi = (float) (zd - cz);
    double e2;
    if (a.eps2i) e2 = a.eps2i[ip];
    else { const double h = a.hi[ip]; e2 = h * h; }
    const float eps2 = (float) e2;
    Kahan ax, ay, az, pot;
    __shared__ float sx[kWave], sy[kWave], sz[kWave], sgm[kWave];
    for (int pass = 0; pass < 2; ++pass) {
        const ListRef &L = pass ? a.Ls : a.Lj;
        if (pass && !a.ms && !a.srec) break;
        const double4 *R = pass ? a.srec : a.jrec;
        const double *X = pass ? a.xs : a.xj, *Y = pass ? a.ys : a.yj, *Z = pass ? a.zs : a.zj,
                     *M = pass ? a.ms : a.mj;
        auto load = [&](int q) { return R ? R[q] : make_double4(X[q], Y[q], Z[q], M[q]); };
        const int64_t lb = list_begin(L, g);
        const int nj = L.cnt[g];
        // (no prefetch of the next tile: the kernel is bound by its VALU issue,
        // and the registers a prefetch holds cost a wave of occupancy)
        for (int t = 0; t < nj; t += kWave) {
            if (t + lane < nj) {
                const double4 r = load(L.idx[lb + t + lane]);
                sx[lane] = (float) (r.x - cx);
                sy[lane] = (float) (r.y - cy);
                sz[lane] = (float) (r.z - cz);
                sgm[lane] = (float) (kG * r.w);
            }
            __syncthreads();
            float px = 0.f, py = 0.f, pz = 0.f, pp = 0.f;
            grav_tile_mixed(sx, sy, sz, sgm, min(kWave, nj - t), xi, yi, zi, eps2, px, py, pz, pp);
            ax.add(px);
            ay.add(py);
            az.add(pz);
            pot.add(pp);
            __syncthreads();
        }
    }
    if (act) {
        a.ax[ip] = ax.value();
        a.ay[ip] = ay.value();
        a.az[ip] = az.value();
        a.pot[ip] = pot.value();
    }
}
__global__ __launch_bounds__(kWave) void k_grav_mixed(GravArgs a) {
    const int G = group_count(a);
    if ((int) blockIdx.x < G) grav_mixed_group(a, xcd_remap((int) blockIdx.x, G));
}
// persistent form: a fixed grid taking groups from the per-XCD counters
__global__ __launch_bounds__(kWave) void k_grav_mixed_persistent(GravArgs a) {
    GroupIter it(a.sched, group_count(a));
    for (int g = it.first(); g >= 0; g = it.next()) {
        grav_mixed_group(a, g);
        __syncthreads();  // the next group reuses the LDS tiles
    }
}

}  // namespace

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

void launch_dens_mixed(const DensArgs &a, hipStream_t s) {
    launch_groups(k_dens_mixed, k_dens_mixed_persistent, a, s);
}
void launch_drvt_mixed(const DrvtArgs &a, hipStream_t s) {
    launch_groups(k_drvt_mixed, k_drvt_mixed_persistent, a, s);
}
void launch_hydro_mixed(const HydroArgs &a, hipStream_t s) {
    if (a.pass1) launch_groups(k_hydro_mixed<true>, k_hydro_mixed_persistent<true>, a, s);
    else launch_groups(k_hydro_mixed<false>, k_hydro_mixed_persistent<false>, a, s);
}
void launch_grav_mixed(const GravArgs &a, hipStream_t s) {
    launch_groups(k_grav_mixed, k_grav_mixed_persistent, a, s);
}

}  // namespace sphgpu
