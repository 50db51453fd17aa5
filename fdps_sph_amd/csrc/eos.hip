// eos.hip -- EoS table evaluation on the GPU (src/EoS.h, src/interpolation.h).
//
// Tables live in HBM as one (rho axis, u axis, 6 property planes) set per
// material; a particle's material id selects the set.  The lookup restates
// BilinearInterpolation::interpolate (src/interpolation.h:60-102): lower_bound
// on both axes clamped to n-1, four-corner blend only when both indices are
// strictly inside (0, n-1), otherwise the clamped corner value.  Compiled with
// -ffp-contract=off: the blend is evaluated in the reference's operand order.
#include "internal.hpp"

namespace sphgpu {
namespace {

// std::lower_bound (libstdc++ bisection)
__device__ __forceinline__ int lower_bound_d(const double *a, int n, double v) {
    int first = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (a[first + half] < v) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}
// std::upper_bound
__device__ __forceinline__ int upper_bound_d(const double *a, int n, double v) {
    int first = 0, len = n;
    while (len > 0) {
        const int half = len >> 1;
        if (!(v < a[first + half])) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

__device__ __forceinline__ double bilinear(const MatDev &t, double v1, double v2, int prop) {
    int li = lower_bound_d(t.rho, t.nr, v1);
    int ci = lower_bound_d(t.u, t.nu, v2);
    if (li > t.nr - 1) li = t.nr - 1;
    if (ci > t.nu - 1) ci = t.nu - 1;
    const double *T = t.plane + (int64_t) prop * t.nr * t.nu;
    if (li < t.nr - 1 && ci < t.nu - 1 && li > 0 && ci > 0) {
        const double xi = (v1 - t.rho[li - 1]) / (t.rho[li] - t.rho[li - 1]);
        const double eta = (v2 - t.u[ci - 1]) / (t.u[ci] - t.u[ci - 1]);
        const double *r0 = T + (int64_t) (li - 1) * t.nu, *r1 = T + (int64_t) li * t.nu;
        return (1 - xi) * (1 - eta) * r0[ci - 1] + xi * (1 - eta) * r1[ci - 1] + (1 - xi) * eta * r0[ci] +
               xi * eta * r1[ci];
    }
    return T[(int64_t) li * t.nu + ci];
}

__device__ __forceinline__ bool mat_ok(const MatDev &t) { return t.kind == 1 || t.nr > 0; }

// Analytic Tillotson EoS: the compressed/cold branch P_co and the expanded
// branch P_ex of src/EoS.h:91-103 (eta = rho/rho0, mu = eta - 1,
// w = u/(u0 eta^2) + 1, z = 1/eta - 1), each with its exact partial
// derivatives in rho and u.
__device__ void till_co(const double *q, double rho, double u, double &P, double &dr, double &du) {
    const double rho0 = q[0], u0 = q[1], A = q[4], B = q[5], a = q[6], b = q[7];
    const double eta = rho / rho0, mu = eta - 1.0;
    const double w = u / u0 / eta / eta + 1.0;
    const double dw_dr = -2.0 * u / (u0 * eta * eta * rho), dw_du = 1.0 / (u0 * eta * eta);
    const double k = a + b / w;
    P = k * rho * u + A * mu + B * mu * mu;
    dr = k * u - b * rho * u / (w * w) * dw_dr + A / rho0 + 2.0 * B * mu / rho0;
    du = k * rho - b * rho * u / (w * w) * dw_du;
}

__device__ void till_ex(const double *q, double rho, double u, double &P, double &dr, double &du) {
    const double rho0 = q[0], u0 = q[1], A = q[4], a = q[6], b = q[7], alpha = q[8], beta = q[9];
    const double eta = rho / rho0, mu = eta - 1.0;
    const double w = u / u0 / eta / eta + 1.0;
    const double dw_dr = -2.0 * u / (u0 * eta * eta * rho), dw_du = 1.0 / (u0 * eta * eta);
    const double z = 1.0 / eta - 1.0, dz_dr = -1.0 / (eta * rho);
    const double ea = exp(-alpha * z), eb = exp(-beta * z * z);
    const double F = b * rho * u / w + A * mu * ea;
    const double dF_dr = b * u / w - b * rho * u / (w * w) * dw_dr + A / rho0 * ea - alpha * A * mu * ea * dz_dr;
    const double dF_du = b * rho / w - b * rho * u / (w * w) * dw_du;
    P = a * rho * u + F * eb;
    dr = a * u + dF_dr * eb + F * eb * (-2.0 * beta * z * dz_dr);
    du = a * rho + dF_du * eb;
}

// Tillotson's regimes: compressed (rho >= rho0) or cold (u <= u_iv) -> P_co;
// expanded and hot (u >= u_cv) -> P_ex; linear blend in u in between.
// c^2 = dP/drho|_u + P/rho^2 dP/du|_rho; c_s at least 1e-12 m/s, the floor the
// ANEOS tables carry (at c_s = 0 the Balsara switch of a particle at rest
// would be 0/0).
__device__ void tillotson(const double *q, double rho, double u, double &P, double &cs) {
    const double rho0 = q[0], uiv = q[2], ucv = q[3];
    double dr, du;
    if (rho >= rho0 || u <= uiv) {
        till_co(q, rho, u, P, dr, du);
    } else if (u >= ucv) {
        till_ex(q, rho, u, P, dr, du);
    } else {
        double Pc, drc, duc, Pe, dre, due;
        till_co(q, rho, u, Pc, drc, duc);
        till_ex(q, rho, u, Pe, dre, due);
        const double d = ucv - uiv;
        P = ((u - uiv) * Pe + (ucv - u) * Pc) / d;
        dr = ((u - uiv) * dre + (ucv - u) * drc) / d;
        du = (Pe + (u - uiv) * due - Pc + (ucv - u) * duc) / d;
    }
    const double c2 = dr + P / (rho * rho) * du;
    cs = sqrt(c2 > 1e-24 ? c2 : 1e-24);
}

__global__ void k_eos_pressure(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                               int nmat, double *p, double *cs) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = mat[i];
    if (m < 0 || m >= nmat || !mat_ok(mats[m])) {
        p[i] = NAN;
        cs[i] = NAN;
        return;
    }
    const double r = rho[i], e = u[i];
    if (mats[m].kind == 1) {
        double P, c;
        tillotson(mats[m].till, r, e, P, c);
        p[i] = P;
        cs[i] = c;
        return;
    }
    const MatDev &t = mats[m];
    p[i] = bilinear(t, r, e, 3);
    cs[i] = bilinear(t, r, e, 4);
}

__global__ void k_eos_property(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                               int nmat, int prop, double *out) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = mat[i];
    if (m < 0 || m >= nmat || !mat_ok(mats[m])) {
        out[i] = NAN;
        return;
    }
    if (mats[m].kind == 1) {  // an analytic material has P and c_s only
        double P, c;
        tillotson(mats[m].till, rho[i], u[i], P, c);
        out[i] = prop == 3 ? P : (prop == 4 ? c : NAN);
        return;
    }
    out[i] = bilinear(mats[m], rho[i], u[i], prop);
}

// RestrictedBilinearInterpolation::get_var2_neighbors (src/interpolation.h:189-236)
// on the inclusive slice v2[b1..b2]; false where the reference indexes out of range.
__device__ bool restricted_pair(const double *v2, int64_t N, uint32_t b1, uint32_t b2, double S, uint32_t gl,
                                int abs_int, int64_t &first, int64_t &second) {
    if (b2 < b1 || (int64_t) b2 >= N + 1) return false;
    const int64_t L = (int64_t) b2 - b1 + 1;
    bool init = true;
    double min_d = 0;
    uint32_t idx = 0;
    for (int64_t a = 0; a < L; ++a) {
        if ((int64_t) b1 + a >= N) return false;
        const double d = S - v2[b1 + a];
        if (init) {
            min_d = d;
            idx = (uint32_t) a;
            init = false;
        } else {
            bool lt;
            if (abs_int) {
                const int di = (int) d, mi = (int) min_d;
                lt = abs(di) < abs(mi);
            } else {
                lt = fabs(d) < fabs(min_d);
            }
            if (lt) {
                min_d = d;
                idx = (uint32_t) a;
            }
        }
    }
    int32_t f, s;
    if (min_d < 0) {
        if (idx <= 0) { f = (int32_t) idx; s = (int32_t) (idx + 1); }
        else if (idx + 1 == gl) { f = (int32_t) (idx - 2); s = (int32_t) (idx - 1); }
        else { f = (int32_t) (idx - 1); s = (int32_t) idx; }
    } else if (min_d > 0) {
        if (idx + 1 == gl) { f = (int32_t) (idx - 2); s = (int32_t) (idx - 1); }
        else { f = (int32_t) idx; s = (int32_t) (idx + 1); }
    } else {
        if (idx >= gl - 1) { f = (int32_t) (idx - 2); s = (int32_t) (idx - 1); }
        else { f = (int32_t) (idx - 1); s = (int32_t) (idx + 1); }
    }
    if (f < 0 || s < 0 || f >= L || s >= L) return false;
    first = f;
    second = s;
    return true;
}

__device__ __forceinline__ double lin(double x1, double x2, double x, double q1, double q2) {
    return (((x2 - x) / (x2 - x1)) * q1) + (((x - x1) / (x2 - x1)) * q2);
}

// RestrictedBilinearInterpolation::interpolate (src/interpolation.h:277-334)
// with grid_length 120 (src/EoS.h:404).  The reference's linear scans for the
// first entry >= rho and the last entry <= rho are binary searches here,
// which the host only allows for a non-decreasing density column.
__global__ void k_eos_internal_energy(int64_t n, const int32_t *mat, const double *rho, const double *S,
                                      const MatDev *mats, int nmat, int abs_int, double *out) {
    const int64_t i = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int m = mat[i];
    if (m < 0 || m >= nmat || mats[m].kind != 0 || mats[m].nr <= 0 || !mats[m].sorted_rho) {
        out[i] = NAN;
        return;
    }
    const MatDev t = mats[m];
    const uint32_t gl = 120;
    const int64_t N = (int64_t) t.nr * t.nu;
    const int32_t v_size = (int32_t) N;
    const double val1 = rho[i], val2 = S[i];
    const double *v1 = t.full_rho;
    uint32_t lb1 = 0, lb2 = 0, ub1 = 0, ub2 = 0;
    // "lower": last entry <= val1
    if (val1 == val1) {
        const int ub = upper_bound_d(v1, (int) N, val1);
        if (ub > 0) {
            const uint32_t a = (uint32_t) (N - 1 - (ub - 1));  // position in the reversed vector
            lb1 = (uint32_t) (v_size - 1) - a - (gl - 1);
            lb2 = lb1 + gl;
            if (lb2 == 0) lb2 = gl;
        }
    }
    if (lb1 == 0 && lb2 == 0) { lb1 = 0; lb2 = gl; }
    // "upper": first entry >= val1
    {
        const int lb = (val1 == val1) ? lower_bound_d(v1, (int) N, val1) : (int) N;
        if (lb < N) {
            ub1 = (uint32_t) lb;
            ub2 = (uint32_t) lb + gl;
            if (ub2 > (uint32_t) (N - 1)) ub2 = (uint32_t) (N - 1);
        }
    }
    if (ub1 == 0 && ub2 == 0) { ub1 = 0; ub2 = gl; }
    int64_t lf, ls, uf, us;
    if (!restricted_pair(t.full_S, N, lb1, lb2, val2, gl, abs_int, lf, ls) ||
        !restricted_pair(t.full_S, N, ub1, ub2, val2, gl, abs_int, uf, us)) {
        out[i] = NAN;
        return;
    }
    const double *v2 = t.full_S, *v3 = t.full_u;
    const double p1 = v1[lb1], p2 = v1[ub1];
    const double u1 = lin(v2[lb1 + lf], v2[lb1 + ls], val2, v3[lb1 + lf], v3[lb1 + ls]);
    const double u2 = lin(v2[ub1 + uf], v2[ub1 + us], val2, v3[ub1 + uf], v3[ub1 + us]);
    out[i] = lin(p1, p2, val1, u1, u2);
}

inline int nblk(int64_t n, int b) { return (int) ((n + b - 1) / b); }

}  // namespace

void launch_eos_pressure(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                         int nmat, double *p, double *cs, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eos_pressure, dim3(nblk(n, 256)), dim3(256), 0, s, n, mat, rho, u, mats, nmat, p, cs);
    SPH_LAUNCH_CHECK();
}

void launch_eos_property(int64_t n, const int32_t *mat, const double *rho, const double *u, const MatDev *mats,
                         int nmat, int prop, double *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eos_property, dim3(nblk(n, 256)), dim3(256), 0, s, n, mat, rho, u, mats, nmat, prop, out);
    SPH_LAUNCH_CHECK();
}

void launch_eos_internal_energy(int64_t n, const int32_t *mat, const double *rho, const double *S,
                                const MatDev *mats, int nmat, int abs_int, double *out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_eos_internal_energy, dim3(nblk(n, 128)), dim3(128), 0, s, n, mat, rho, S, mats, nmat,
                       abs_int, out);
    SPH_LAUNCH_CHECK();
}

}  // namespace sphgpu
