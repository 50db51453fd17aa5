// integrate.hip -- FP64 step glue on the GPU: RealPtcl::initialize /
// initialKick / fullDrift / predict / finalKick / dampMotion
// (src/class.h:196-229), addExternalForce (src/GI.h:458-465) and
// getTimeStepGlobal (src/integral.h:3-21) as a per-system min-reduction.
// Compiled with -ffp-contract=off (the reference's operand order).
#include <algorithm>

#include "internal.hpp"

namespace sphgpu {
namespace {

constexpr int kChunk = 4096;

inline int nblk(int64_t n, int b) { return (int) ((n + b - 1) / b); }

__global__ void k_init_h(int64_t n, PArrays P) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    P.h[p] = kSMTH * pow(P.m[p] / P.rho[p], 1.0 / 3.0);
    P.gs[p] = 1.0;
    P.bal[p] = 1.0;
}

__global__ void k_kick_drift(int64_t n, PArrays P, const double *dt_sys, int nsys, double *time_sys) {
    if (blockIdx.x == 0)  // and each system's time advances by its dt
        for (int q = threadIdx.x; q < nsys; q += blockDim.x) time_sys[q] += dt_sys[q];
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double dt = dt_sys[P.sys[p]];
    const double ht = 0.5 * dt;
    const double axp = P.ax[p] + P.gx[p], ayp = P.ay[p] + P.gy[p], azp = P.az[p] + P.gz[p];
    const double vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
    const double vhx = vx + ht * axp, vhy = vy + ht * ayp, vhz = vz + ht * azp;
    P.vhx[p] = vhx;
    P.vhy[p] = vhy;
    P.vhz[p] = vhz;
    P.uh[p] = P.u[p] + ht * P.du[p];
    P.x[p] += dt * vhx;
    P.y[p] += dt * vhy;
    P.z[p] += dt * vhz;
    P.vx[p] = vx + dt * axp;
    P.vy[p] = vy + dt * ayp;
    P.vz[p] = vz + dt * azp;
    P.u[p] += dt * P.du[p];
}


// addExternalForce: acc += -vel * 0.05 / dt while time < until
__device__ __forceinline__ void ext_force(PArrays &P, int64_t p, const double *time_sys, double until) {
    if (time_sys[P.sys[p]] >= until) return;
    const double inv = 1.0 / P.dti[p];
    P.ax[p] += ((-P.vx[p]) * 0.05) * inv;
    P.ay[p] += ((-P.vy[p]) * 0.05) * inv;
    P.az[p] += ((-P.vz[p]) * 0.05) * inv;
}

__global__ void k_ext_force(int64_t n, PArrays P, const double *time_sys, double until) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    ext_force(P, p, time_sys, until);
}

__global__ void k_finalize(int64_t n, PArrays P, const double *dt_sys, const double *time_sys, double until,
                           double damping, StepGuard guard) {
    const int64_t p = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || guard.failed()) return;
    ext_force(P, p, time_sys, until);
    const double ht = 0.5 * dt_sys[P.sys[p]];
    double vx = P.vhx[p] + ht * (P.ax[p] + P.gx[p]);
    double vy = P.vhy[p] + ht * (P.ay[p] + P.gy[p]);
    double vz = P.vhz[p] + ht * (P.az[p] + P.gz[p]);
    P.u[p] = P.uh[p] + ht * P.du[p];
    vx *= damping;
    vy *= damping;
    vz *= damping;
    P.vx[p] = vx;
    P.vy[p] = vy;
    P.vz[p] = vz;
}

__global__ void k_dt_partial(const int32_t *sys_start, int nchunk, const double *dti, double *part) {
    const int s = blockIdx.y, c = blockIdx.x;
    const int b = sys_start[s] + c * kChunk;
    const int e = min(sys_start[s + 1], b + kChunk);
    double v = 1.0e+30;
    for (int p = b + threadIdx.x; p < e; p += blockDim.x) v = dmin(v, dmin(dti[p], 1.0e+30));
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) sh[threadIdx.x] = dmin(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t) s * nchunk + c] = sh[0];
}

__global__ void k_dt_final(int nchunk, const double *part, double *dt_sys, StepGuard guard) {
    if (guard.failed()) return;
    const int s = blockIdx.x;
    double v = 1.0e+30;
    for (int c = threadIdx.x; c < nchunk; c += blockDim.x) v = dmin(v, part[(int64_t) s * nchunk + c]);
    __shared__ double sh[256];
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int) threadIdx.x < o) sh[threadIdx.x] = dmin(sh[threadIdx.x], sh[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) dt_sys[s] = sh[0];
}

}  // namespace

void launch_init_h(int64_t n, PArrays P, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_init_h, dim3(nblk(n, 256)), dim3(256), 0, s, n, P);
    SPH_LAUNCH_CHECK();
}

void launch_kick_drift(int64_t n, PArrays P, const double *dt_sys, int nsys, double *time_sys, hipStream_t s) {
    hipLaunchKernelGGL(k_kick_drift, dim3(std::max(1, nblk(n, 256))), dim3(256), 0, s, n, P, dt_sys, nsys, time_sys);
    SPH_LAUNCH_CHECK();
}

void launch_ext_force(int64_t n, PArrays P, const double *time_sys, double until, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_ext_force, dim3(nblk(n, 256)), dim3(256), 0, s, n, P, time_sys, until);
    SPH_LAUNCH_CHECK();
}

void launch_finalize(int64_t n, PArrays P, const double *dt_sys, const double *time_sys, double until,
                     double damping, StepGuard guard, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_finalize, dim3(nblk(n, 256)), dim3(256), 0, s, n, P, dt_sys, time_sys, until, damping,
                       guard);
    SPH_LAUNCH_CHECK();
}

void launch_dt_reduce(int nsys, const int32_t *sys_start_host, const int32_t *sys_start, int64_t n,
                      const double *dti, double *part, double *dt_sys, StepGuard guard, hipStream_t s) {
    (void) n;
    int maxn = 0;
    for (int q = 0; q < nsys; ++q) maxn = std::max(maxn, sys_start_host[q + 1] - sys_start_host[q]);
    const int nchunk = std::max(1, (maxn + kChunk - 1) / kChunk);
    hipLaunchKernelGGL(k_dt_partial, dim3(nchunk, nsys), dim3(256), 0, s, sys_start, nchunk, dti, part);
    SPH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_dt_final, dim3(nsys), dim3(256), 0, s, nchunk, part, dt_sys, guard);
    SPH_LAUNCH_CHECK();
}

}  // namespace sphgpu
