/*
 * sph_oracle.c -- CPU oracle for the FDPS_SPH per-step interaction path.
 *
 * TEST INFRASTRUCTURE ONLY (see sph_oracle.h).  Built twice by oracle/Makefile:
 *   liboracle_strict.so  -O2 -ffp-contract=off, no fast-math, serial:
 *                        the parity checker (sequential j-order, IEEE ops);
 *   liboracle_fast.so    the reference's flags -O3 -ffast-math -funroll-loops
 *                        -fopenmp (CMakeLists.txt:26-27): the CPU baseline.
 *
 * Arithmetic follows the reference expression by expression (operand order,
 * left-to-right association).  FDPS's Vector3 is not available here; its
 * operators are taken as dot = x*x'+y*y'+z*z' (left to right), cross =
 * (y z'-z y', z x'-x z', x y'-y x'), and vector/scalar = vector*(1.0/scalar)
 * (the form -ffast-math's reciprocal-math produces in any case).
 */
#define _POSIX_C_SOURCE 200809L
#include "sph_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* math::pi = atan(1.0)*4.0 (src/mathfunc.h:4); equals the double nearest pi. */
#define ORC_PI 3.14159265358979323846
/* CalcGravityForce<T>::G (src/force.h:171-172) */
#define ORC_G 6.67e-11
/* PARAM (src/param.h:9-17) */
#define ORC_SMTH 1.2
#define ORC_C_CFL 0.3
#define ORC_AV_STRENGTH 1.0
/* kernel_t::supportRadius() (src/kernel.h:36-38) */
#define ORC_SUPPORT 2.5

/* math::plus, pow7, pow8 (src/mathfunc.h:9-12,46-57) */
static inline double plus_(double a) { return (a > 0) ? a : 0; }
static inline double pow7_(double a) { const double a3 = a * a * a; return a3 * a3 * a; }
static inline double pow8_(double a) { const double a2 = a * a; const double a4 = a2 * a2; return a4 * a4; }
static inline double dot3_(const double a[3], const double b[3]) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
/* std::max(a, b) == (a < b) ? b : a */
static inline double stdmax_(double a, double b) { return (a < b) ? b : a; }
static inline double stdmin_(double a, double b) { return (b < a) ? b : a; }

/* WendlandC6::W (src/kernel.h:8-19), 3-D normalisation. */
double orc_W(const double dr[3], double h)
{
    const double H = ORC_SUPPORT * h;
    const double s = sqrt(dot3_(dr, dr)) / H;
    double r = (1.0 + s * (8.0 + s * (25.0 + s * (32.0)))) * pow8_(plus_(1.0 - s));
    r *= (1365. / 64.) / (H * H * H * ORC_PI);
    return r;
}

/* WendlandC6::gradW (src/kernel.h:22-34), incl. the 1e-6*h regulariser. */
void orc_gradW(const double dr[3], double h, double out[3])
{
    const double H = ORC_SUPPORT * h;
    const double s = sqrt(dot3_(dr, dr)) / H;
    double r = pow7_(plus_(1.0 - s)) *
               (plus_(1.0 - s) * (8.0 + s * (50.0 + s * (96.0))) - 8.0 * (1.0 + s * (8.0 + s * (25.0 + s * (32.0)))));
    r *= (1365. / 64.) / (H * H * H * ORC_PI);
    const double inv = 1.0 / (sqrt(dot3_(dr, dr)) * H + 1.0e-6 * h);
    for (int k = 0; k < 3; ++k) out[k] = (dr[k] * r) * inv;
}

/* RESULT::*::clear (src/class.h:41-43,53-57,67-71,81-85) */
void orc_clear_dens(orc_res_dens *r, int n) { for (int i = 0; i < n; ++i) { r[i].dens = 0; r[i].smth = 0; } }
void orc_clear_drvt(orc_res_drvt *r, int n)
{
    for (int i = 0; i < n; ++i) { r[i].div_v = 0; r[i].grad_smth = 0; r[i].rot_v[0] = r[i].rot_v[1] = r[i].rot_v[2] = 0; }
}
void orc_clear_hydro(orc_res_hydro *r, int n)
{
    for (int i = 0; i < n; ++i) { r[i].acc[0] = r[i].acc[1] = r[i].acc[2] = 0; r[i].eng_dot = 0; r[i].dt = 1.0e+30; }
}
void orc_clear_grav(orc_res_grav *r, int n)
{
    for (int i = 0; i < n; ++i) { r[i].acc[0] = r[i].acc[1] = r[i].acc[2] = 0; r[i].pot = 0; r[i].dt = 1.0e+30; }
}

/* CalcDensity::operator() (src/force.h:9-23), FLAG_GI floor (src/GI.h:5). */
void orc_calc_density(const orc_epi_dens *epi, int nip, const orc_epj_dens *epj, int njp, orc_res_dens *res)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_dens *ith = &epi[i];
        for (int j = 0; j < njp; ++j) {
            const orc_epj_dens *jth = &epj[j];
            double dr[3];
            for (int k = 0; k < 3; ++k) dr[k] = jth->pos[k] - ith->pos[k];
            res[i].dens += jth->mass * orc_W(dr, ith->smth);
        }
        res[i].dens = stdmax_(5.0, res[i].dens);
        res[i].smth = ORC_SMTH * pow(ith->mass / res[i].dens, 1.0 / 3.0);
    }
}

void orc_count_ngb(const orc_epi_dens *epi, int nip, const orc_epj_dens *epj, int njp, int32_t *ngb)
{
    for (int i = 0; i < nip; ++i) {
        const double H = ORC_SUPPORT * epi[i].smth;
        int32_t c = 0;
        for (int j = 0; j < njp; ++j) {
            double dr[3];
            for (int k = 0; k < 3; ++k) dr[k] = epj[j].pos[k] - epi[i].pos[k];
            const double s = sqrt(dot3_(dr, dr)) / H;
            if (plus_(1.0 - s) > 0) ++c;
        }
        ngb[i] = c;
    }
}

/* CalcDerivative::operator() (src/force.h:89-106). */
void orc_calc_derivative(const orc_epi_drvt *epi, int nip, const orc_epj_drvt *epj, int njp, orc_res_drvt *res)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_drvt *ith = &epi[i];
        for (int j = 0; j < njp; ++j) {
            const orc_epj_drvt *jth = &epj[j];
            double dr[3], dv[3], gw[3], a[3];
            for (int k = 0; k < 3; ++k) { dr[k] = ith->pos[k] - jth->pos[k]; dv[k] = ith->vel[k] - jth->vel[k]; }
            orc_gradW(dr, ith->smth, gw);
            const double w = orc_W(dr, ith->smth);
            const double mneg = -jth->mass;
            for (int k = 0; k < 3; ++k) a[k] = mneg * dv[k];
            res[i].div_v += dot3_(a, gw);
            res[i].rot_v[0] += a[1] * gw[2] - a[2] * gw[1];
            res[i].rot_v[1] += a[2] * gw[0] - a[0] * gw[2];
            res[i].rot_v[2] += a[0] * gw[1] - a[1] * gw[0];
            res[i].grad_smth -= jth->mass / ith->smth * (3.0 * w + dot3_(dr, gw));
        }
        res[i].grad_smth = 1.0 / (1.0 + ith->smth * res[i].grad_smth / (3.0 * ith->dens));
        res[i].div_v /= ith->dens;
        const double inv = 1.0 / ith->dens;
        for (int k = 0; k < 3; ++k) res[i].rot_v[k] *= inv;
    }
}

/* CalcHydroForce::operator() (src/force.h:112-147), live `#if 1` branch,
 * FLAG_R00 == false so AVa is not read. */
void orc_calc_hydro(const orc_epi_hydro *epi, int nip, const orc_epj_hydro *epj, int njp, orc_res_hydro *res)
{
    for (int i = 0; i < nip; ++i) {
        double v_sig_max = 0.0;
        const orc_epi_hydro *ith = &epi[i];
        const double Ai = ith->grad_smth * ith->pres / (ith->dens * ith->dens);
        for (int j = 0; j < njp; ++j) {
            const orc_epj_hydro *jth = &epj[j];
            double dr[3], dv[3], gwi[3], gwj[3], gw[3];
            for (int k = 0; k < 3; ++k) { dr[k] = ith->pos[k] - jth->pos[k]; dv[k] = ith->vel[k] - jth->vel[k]; }
            const double dvdr = dot3_(dv, dr);
            const double w_ij = (dvdr < 0) ? dvdr / sqrt(dot3_(dr, dr)) : 0;
            const double v_sig = ith->snds + jth->snds - 3.0 * w_ij;
            v_sig_max = stdmax_(v_sig_max, v_sig);
            const double AV = -ORC_AV_STRENGTH * 0.5 * v_sig * w_ij / (0.5 * (ith->dens + jth->dens)) * 0.5 *
                              (ith->Bal + jth->Bal);
            orc_gradW(dr, ith->smth, gwi);
            orc_gradW(dr, jth->smth, gwj);
            for (int k = 0; k < 3; ++k) gw[k] = 0.5 * (gwi[k] * ith->grad_smth + gwj[k] * jth->grad_smth);
            const double Aj = jth->grad_smth * jth->pres / (jth->dens * jth->dens);
            for (int k = 0; k < 3; ++k) res[i].acc[k] -= jth->mass * (Ai * gwi[k] + Aj * gwj[k] + AV * gw[k]);
            const double sc = jth->mass * (Ai + 0.5 * AV);
            double sdv[3];
            for (int k = 0; k < 3; ++k) sdv[k] = sc * dv[k];
            res[i].eng_dot += dot3_(sdv, gw);
        }
        res[i].dt = ORC_C_CFL * 2.0 * ith->smth / v_sig_max;
    }
}

/* CalcGravityForce<TPtclJ>::operator() (src/force.h:154-168), P2P form. */
void orc_calc_grav_ep(const orc_epi_grav *epi, int nip, const orc_epj_grav *epj, int njp, orc_res_grav *res)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_grav *ith = &epi[i];
        for (int j = 0; j < njp; ++j) {
            const orc_epj_grav *jth = &epj[j];
            double dr[3];
            for (int k = 0; k < 3; ++k) dr[k] = ith->pos[k] - jth->pos[k];
            const double dr2 = dot3_(dr, dr);
            const double dr_inv = 1.0 / sqrt(dr2 + ith->eps2);
            const double m_dr3_inv = jth->mass * (dr_inv * dr_inv * dr_inv);
            for (int k = 0; k < 3; ++k) res[i].acc[k] -= ORC_G * m_dr3_inv * dr[k];
            res[i].pot -= ORC_G * jth->mass * dr_inv;
        }
    }
}

/* Same functor instantiated for PS::SPJMonopole (P2M). */
void orc_calc_grav_sp(const orc_epi_grav *epi, int nip, const orc_spj_mono *spj, int njp, orc_res_grav *res)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_grav *ith = &epi[i];
        for (int j = 0; j < njp; ++j) {
            const orc_spj_mono *jth = &spj[j];
            double dr[3];
            for (int k = 0; k < 3; ++k) dr[k] = ith->pos[k] - jth->pos[k];
            const double dr2 = dot3_(dr, dr);
            const double dr_inv = 1.0 / sqrt(dr2 + ith->eps2);
            const double m_dr3_inv = jth->mass * (dr_inv * dr_inv * dr_inv);
            for (int k = 0; k < 3; ++k) res[i].acc[k] -= ORC_G * m_dr3_inv * dr[k];
            res[i].pot -= ORC_G * jth->mass * dr_inv;
        }
    }
}

static inline double absmax3_(const double v[3])
{
    double m = fabs(v[0]);
    if (fabs(v[1]) > m) m = fabs(v[1]);
    if (fabs(v[2]) > m) m = fabs(v[2]);
    return m;
}

/* |d gradW / d ln h| per component (central difference), the response of a
 * kernel gradient to a relative change of the smoothing length. */
static void gradW_dlnh_(const double dr[3], double h, double out[3])
{
    const double d = 1e-5;
    double a[3], b[3];
    orc_gradW(dr, h * (1.0 + d), a);
    orc_gradW(dr, h * (1.0 - d), b);
    for (int k = 0; k < 3; ++k) out[k] = fabs(a[k] - b[k]) / (2.0 * d);
}

void orc_hydro_term_scale(const orc_epi_hydro *epi, int nip, const orc_epj_hydro *epj, int njp, double *acc_scale,
                          double *du_scale)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_hydro *ith = &epi[i];
        const double Ai = ith->grad_smth * ith->pres / (ith->dens * ith->dens);
        /* |A| plus its response to a relative change of rho through P: f c^2 / rho */
        const double Si = fabs(Ai) + ith->grad_smth * ith->snds * ith->snds / ith->dens;
        double sa = 0, su = 0;
        for (int j = 0; j < njp; ++j) {
            const orc_epj_hydro *jth = &epj[j];
            double dr[3], dv[3], gwi[3], gwj[3], gw[3], di[3], dj[3], t[3];
            for (int k = 0; k < 3; ++k) { dr[k] = ith->pos[k] - jth->pos[k]; dv[k] = ith->vel[k] - jth->vel[k]; }
            const double dvdr = dot3_(dv, dr);
            const double w_ij = (dvdr < 0) ? dvdr / sqrt(dot3_(dr, dr)) : 0;
            const double v_sig = ith->snds + jth->snds - 3.0 * w_ij;
            const double AV = -0.5 * v_sig * w_ij / (0.5 * (ith->dens + jth->dens)) * 0.5 * (ith->Bal + jth->Bal);
            orc_gradW(dr, ith->smth, gwi);
            orc_gradW(dr, jth->smth, gwj);
            gradW_dlnh_(dr, ith->smth, di);
            gradW_dlnh_(dr, jth->smth, dj);
            for (int k = 0; k < 3; ++k) gw[k] = 0.5 * (gwi[k] * ith->grad_smth + gwj[k] * jth->grad_smth);
            const double Aj = jth->grad_smth * jth->pres / (jth->dens * jth->dens);
            const double Sj = fabs(Aj) + jth->grad_smth * jth->snds * jth->snds / jth->dens;
            double su_k = 0;
            for (int k = 0; k < 3; ++k) {
                const double gbar = fabs(gw[k]) + 0.5 * (ith->grad_smth * di[k] + jth->grad_smth * dj[k]);
                t[k] = jth->mass * (Si * (fabs(gwi[k]) + di[k]) + Sj * (fabs(gwj[k]) + dj[k]) + fabs(AV) * gbar);
                su_k += fabs(dv[k]) * gbar;
            }
            sa += absmax3_(t);
            su += jth->mass * (Si + 0.5 * fabs(AV)) * su_k;
        }
        acc_scale[i] = sa;
        du_scale[i] = su;
    }
}

void orc_drvt_term_scale(const orc_epi_drvt *epi, int nip, const orc_epj_drvt *epj, int njp, double *div_scale,
                         double *rot_scale)
{
    for (int i = 0; i < nip; ++i) {
        const orc_epi_drvt *ith = &epi[i];
        double sd = 0, sr = 0;
        for (int j = 0; j < njp; ++j) {
            const orc_epj_drvt *jth = &epj[j];
            double dr[3], dv[3], gw[3], dg[3], g[3], c[3];
            for (int k = 0; k < 3; ++k) { dr[k] = ith->pos[k] - jth->pos[k]; dv[k] = ith->vel[k] - jth->vel[k]; }
            orc_gradW(dr, ith->smth, gw);
            gradW_dlnh_(dr, ith->smth, dg);
            for (int k = 0; k < 3; ++k) g[k] = fabs(gw[k]) + dg[k];
            sd += jth->mass * (fabs(dv[0]) * g[0] + fabs(dv[1]) * g[1] + fabs(dv[2]) * g[2]);
            c[0] = fabs(dv[1]) * g[2] + fabs(dv[2]) * g[1];
            c[1] = fabs(dv[2]) * g[0] + fabs(dv[0]) * g[2];
            c[2] = fabs(dv[0]) * g[1] + fabs(dv[1]) * g[0];
            sr += jth->mass * absmax3_(c);
        }
        div_scale[i] = sd / ith->dens;
        rot_scale[i] = sr / ith->dens;
    }
}

/* RealPtcl::copyFromForce(const RESULT::Drvt&) Balsara switch (src/class.h:136-137). */
double orc_balsara(double div_v, const double rot_v[3], double snds, double smth)
{
    return fabs(div_v) / (fabs(div_v) + sqrt(dot3_(rot_v, rot_v)) + 1.0e-4 * snds / smth);
}

/* ------------------------------------------------------------------------ */
/* EoS tables                                                               */
/* ------------------------------------------------------------------------ */
struct orc_table {
    int kind;         /* 0: rho-u table, 1: analytic Tillotson (till[10]) */
    double till[10];
    int nr, nu;
    double *cells;    /* [nr][nu][6] */
    double *rho_axis; /* column 0, field 0 of each row */
    double *u_axis;   /* row 0, field 1 */
    double *full_rho, *full_u, *full_S;
    int64_t nfull;
};

static orc_table *table_from_cells_(int nr, int nu, const double *cells)
{
    orc_table *t = (orc_table *) calloc(1, sizeof(orc_table));
    const int64_t nc = (int64_t) nr * nu;
    t->nr = nr; t->nu = nu; t->nfull = nc;
    t->cells = (double *) malloc(sizeof(double) * 6 * nc);
    memcpy(t->cells, cells, sizeof(double) * 6 * nc);
    t->rho_axis = (double *) malloc(sizeof(double) * nr);
    t->u_axis = (double *) malloc(sizeof(double) * nu);
    t->full_rho = (double *) malloc(sizeof(double) * nc);
    t->full_u = (double *) malloc(sizeof(double) * nc);
    t->full_S = (double *) malloc(sizeof(double) * nc);
    for (int r = 0; r < nr; ++r) t->rho_axis[r] = cells[((int64_t) r * nu) * 6 + 0];
    for (int c = 0; c < nu; ++c) t->u_axis[c] = cells[(int64_t) c * 6 + 1];
    for (int64_t k = 0; k < nc; ++k) {
        t->full_rho[k] = cells[k * 6 + 0];
        t->full_u[k] = cells[k * 6 + 1];
        t->full_S[k] = cells[k * 6 + 5];
    }
    return t;
}

orc_table *orc_table_from_cells(int nrho, int nu, const double *cells)
{
    if (nrho <= 0 || nu <= 0 || !cells) return NULL;
    return table_from_cells_(nrho, nu, cells);
}

/* Table reader of src/EoS.h:290-385: header line "# nrow ncol", further '#'
 * lines skipped, one cell (up to 6 numbers) per line, a new row every ncol
 * cells.  Blank lines are skipped (the reference would push an
 * uninitialised cell for them). */
orc_table *orc_table_load(const char *path)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    char line[4096];
    int nr = 0, nc = 0, got_header = 0;
    int64_t ncell = 0, cap = 0;
    double *cells = NULL;
    while (fgets(line, sizeof line, fp)) {
        if (!got_header) {
            char tag[64];
            if (sscanf(line, "%63s %d %d", tag, &nr, &nc) != 3 || nr <= 0 || nc <= 0) { fclose(fp); free(cells); return NULL; }
            got_header = 1;
            cap = (int64_t) nr * nc;
            cells = (double *) calloc((size_t) cap * 6, sizeof(double));
            continue;
        }
        if (line[0] == '#') continue;
        char *p = line;
        while (*p == ' ' || *p == '\t') ++p;
        if (*p == '\n' || *p == '\r' || *p == '\0') continue;
        if (ncell >= cap) break;
        double *c = cells + ncell * 6;
        for (int f = 0; f < 6; ++f) {
            char *end;
            const double v = strtod(p, &end);
            if (end == p) break;
            c[f] = v;
            p = end;
        }
        ++ncell;
    }
    fclose(fp);
    if (!got_header || ncell != cap) { free(cells); return NULL; }
    orc_table *t = table_from_cells_(nr, nc, cells);
    free(cells);
    return t;
}

void orc_table_free(orc_table *t)
{
    if (!t) return;
    free(t->cells); free(t->rho_axis); free(t->u_axis); free(t->full_rho); free(t->full_u); free(t->full_S);
    free(t);
}

int orc_table_dims(const orc_table *t, int *nrho, int *nu) { *nrho = t->nr; *nu = t->nu; return 0; }
const double *orc_table_cells(const orc_table *t) { return t->cells; }

/* std::lower_bound (libstdc++ bisection) on an ascending axis. */
static int64_t lower_bound_(const double *a, int64_t n, double v)
{
    int64_t first = 0, len = n;
    while (len > 0) {
        const int64_t half = len >> 1;
        if (a[first + half] < v) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

/* Analytic Tillotson EoS.  The branches restate Tillotson::P_co / P_ex of
 * src/EoS.h:91-103 (eta = rho/rho0, mu = eta - 1, w = u/(u0 eta^2) + 1):
 *   P_co = (a + b/w) rho u + A mu + B mu^2
 *   P_ex = a rho u + (b rho u / w + A mu exp(-alpha z)) exp(-beta z^2),  z = 1/eta - 1.
 * The reference defines them but never calls them (its Tillotson::Pressure is
 * the table lookup).  The regime selection is Tillotson's (compressed or cold:
 * P_co; expanded and hot beyond u_cv: P_ex; linear blend in u between u_iv and
 * u_cv), and c^2 = dP/drho|_u + P/rho^2 dP/du|_rho from the exact derivatives
 * of those expressions, with c_s at least 1e-12 m/s: the floor the ANEOS
 * tables carry, which keeps the Balsara switch |div v| / (|div v| + |rot v| +
 * 1e-4 c_s/h) defined for a particle at rest in tension.  prm = (rho0, u0, u_iv, u_cv, A, B, a, b, alpha,
 * beta), the order of the Tillotson constructor (src/EoS.h:446). */
static void till_co_(const double *q, double rho, double u, double *P, double *dr, double *du)
{
    const double rho0 = q[0], u0 = q[1], A = q[4], B = q[5], a = q[6], b = q[7];
    const double eta = rho / rho0, mu = eta - 1.0;
    const double w = u / u0 / eta / eta + 1.0;
    const double dw_dr = -2.0 * u / (u0 * eta * eta * rho), dw_du = 1.0 / (u0 * eta * eta);
    const double k = a + b / w;
    *P = k * rho * u + A * mu + B * mu * mu;
    *dr = k * u - b * rho * u / (w * w) * dw_dr + A / rho0 + 2.0 * B * mu / rho0;
    *du = k * rho - b * rho * u / (w * w) * dw_du;
}

static void till_ex_(const double *q, double rho, double u, double *P, double *dr, double *du)
{
    const double rho0 = q[0], u0 = q[1], A = q[4], a = q[6], b = q[7], alpha = q[8], beta = q[9];
    const double eta = rho / rho0, mu = eta - 1.0;
    const double w = u / u0 / eta / eta + 1.0;
    const double dw_dr = -2.0 * u / (u0 * eta * eta * rho), dw_du = 1.0 / (u0 * eta * eta);
    const double z = 1.0 / eta - 1.0, dz_dr = -1.0 / (eta * rho);
    const double ea = exp(-alpha * z), eb = exp(-beta * z * z);
    const double F = b * rho * u / w + A * mu * ea;
    const double dF_dr = b * u / w - b * rho * u / (w * w) * dw_dr + A / rho0 * ea - alpha * A * mu * ea * dz_dr;
    const double dF_du = b * rho / w - b * rho * u / (w * w) * dw_du;
    *P = a * rho * u + F * eb;
    *dr = a * u + dF_dr * eb + F * eb * (-2.0 * beta * z * dz_dr);
    *du = a * rho + dF_du * eb;
}

void orc_tillotson(const double prm[10], double rho, double u, double *pres, double *snds)
{
    const double rho0 = prm[0], uiv = prm[2], ucv = prm[3];
    double P, dr, du;
    if (rho >= rho0 || u <= uiv) {
        till_co_(prm, rho, u, &P, &dr, &du);
    } else if (u >= ucv) {
        till_ex_(prm, rho, u, &P, &dr, &du);
    } else {
        double Pc, drc, duc, Pe, dre, due;
        till_co_(prm, rho, u, &Pc, &drc, &duc);
        till_ex_(prm, rho, u, &Pe, &dre, &due);
        const double d = ucv - uiv;
        P = ((u - uiv) * Pe + (ucv - u) * Pc) / d;
        dr = ((u - uiv) * dre + (ucv - u) * drc) / d;
        du = (Pe + (u - uiv) * due - Pc + (ucv - u) * duc) / d;
    }
    const double c2 = dr + P / (rho * rho) * du;
    *pres = P;
    *snds = sqrt(c2 > 1e-24 ? c2 : 1e-24);
}

orc_table *orc_table_tillotson(const double prm[10])
{
    if (!prm) return NULL;
    orc_table *t = (orc_table *) calloc(1, sizeof(orc_table));
    t->kind = 1;
    memcpy(t->till, prm, sizeof(double) * 10);
    return t;
}

/* BilinearInterpolation::interpolate (src/interpolation.h:33-102); for an
 * analytic Tillotson material P (prop 3) and c_s (prop 4), NaN otherwise. */
double orc_eos_bilinear(const orc_table *t, double v1, double v2, int prop)
{
    if (t->kind == 1) {
        double P, c;
        orc_tillotson(t->till, v1, v2, &P, &c);
        return prop == 3 ? P : (prop == 4 ? c : NAN);
    }
    int64_t li = lower_bound_(t->rho_axis, t->nr, v1);
    int64_t ci = lower_bound_(t->u_axis, t->nu, v2);
    if (li > t->nr - 1) li = t->nr - 1;
    if (ci > t->nu - 1) ci = t->nu - 1;
#define CELL_(r, c) t->cells[(((int64_t) (r)) * t->nu + (c)) * 6 + prop]
    if (li < t->nr - 1 && ci < t->nu - 1 && li > 0 && ci > 0) {
        const double xi = (v1 - t->rho_axis[li - 1]) / (t->rho_axis[li] - t->rho_axis[li - 1]);
        const double eta = (v2 - t->u_axis[ci - 1]) / (t->u_axis[ci] - t->u_axis[ci - 1]);
        return (1 - xi) * (1 - eta) * CELL_(li - 1, ci - 1) + xi * (1 - eta) * CELL_(li, ci - 1) +
               (1 - xi) * eta * CELL_(li - 1, ci) + xi * eta * CELL_(li, ci);
    }
    return CELL_(li, ci);
#undef CELL_
}

/* RestrictedBilinearInterpolation (src/interpolation.h:107-335), grid_length
 * 120 as hard-coded at src/EoS.h:251,404.  Unsigned index arithmetic of the
 * reference is kept (uint32); reads the reference makes out of range return
 * NaN here. */
static int restricted_pair_(const double *v2, int64_t nfull, uint32_t b1, uint32_t b2, double S, uint32_t gl,
                            int abs_int, int64_t *first, int64_t *second)
{
    if (b2 < b1 || (int64_t) b2 >= nfull + 1) return -1;
    const int64_t L = (int64_t) b2 - b1 + 1; /* inclusive slice */
    int init = 1;
    double min_d = 0;
    uint32_t idx = 0;
    for (int64_t a = 0; a < L; ++a) {
        if ((int64_t) b1 + a >= nfull) return -1; /* slice reads past the vector */
        const double d = S - v2[b1 + a];
        if (init) { min_d = d; idx = (uint32_t) a; init = 0; }
        else {
            int lt;
            if (abs_int) { const int di = (int) d, mi = (int) min_d; lt = abs(di) < abs(mi); }
            else lt = fabs(d) < fabs(min_d);
            if (lt) { min_d = d; idx = (uint32_t) a; }
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
    if (f < 0 || s < 0 || f >= L || s >= L) return -1;
    *first = f; *second = s;
    return 0;
}

static double lin_(double x1, double x2, double x, double q1, double q2)
{
    return (((x2 - x) / (x2 - x1)) * q1) + (((x - x1) / (x2 - x1)) * q2);
}

double orc_eos_internal_energy(const orc_table *t, double val1, double val2, int abs_int)
{
    if (t->kind != 0) return NAN;
    const uint32_t gl = 120;
    const double *v1 = t->full_rho, *v2 = t->full_S, *v3 = t->full_u;
    const int64_t N = t->nfull;
    const int32_t v_size = (int32_t) N;
    uint32_t lb1 = 0, lb2 = 0, ub1 = 0, ub2 = 0;
    /* "lower" bound: last entry <= value, searched on the reversed vector */
    for (uint32_t a = 0; a < (uint32_t) N; ++a) {
        const double val = v1[N - 1 - a];
        if (val <= val1) {
            lb1 = (uint32_t) (v_size - 1) - a - (gl - 1);
            lb2 = lb1 + gl;
            if (lb2 == 0) lb2 = gl;
            break;
        }
    }
    if (lb1 == 0 && lb2 == 0) { lb1 = 0; lb2 = gl; }
    /* "upper" bound: first entry >= value */
    for (uint32_t a = 0; a < (uint32_t) N; ++a) {
        if (v1[a] >= val1) {
            ub1 = a;
            ub2 = a + gl;
            if (ub2 > (uint32_t) (N - 1)) ub2 = (uint32_t) (N - 1);
            break;
        }
    }
    if (ub1 == 0 && ub2 == 0) { ub1 = 0; ub2 = gl; }
    int64_t lf, ls, uf, us;
    if (restricted_pair_(v2, N, lb1, lb2, val2, gl, abs_int, &lf, &ls)) return NAN;
    if (restricted_pair_(v2, N, ub1, ub2, val2, gl, abs_int, &uf, &us)) return NAN;
    const double p1 = v1[lb1], p2 = v1[ub1];
    const double s11 = v2[lb1 + lf], s12 = v2[lb1 + ls], s21 = v2[ub1 + uf], s22 = v2[ub1 + us];
    const double u11 = v3[lb1 + lf], u12 = v3[lb1 + ls], u21 = v3[ub1 + uf], u22 = v3[ub1 + us];
    const double u1 = lin_(s11, s12, val2, u11, u12);
    const double u2 = lin_(s21, s22, val2, u21, u22);
    return lin_(p1, p2, val1, u1, u2);
}

/* ------------------------------------------------------------------------ */
/* Morton keys, sorting, tree, lists (restates the GPU library's design)    */
/* ------------------------------------------------------------------------ */
static uint64_t expand_bits_(uint64_t v)
{
    v &= 0x1fffffULL;
    v = (v | (v << 32)) & 0x1f00000000ffffULL;
    v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
    v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

static inline uint64_t quantise_(double x, double lo, double scale, uint64_t qmax)
{
    const double f = (x - lo) * scale;
    uint64_t q = (f > 0) ? (uint64_t) f : 0;
    return q > qmax ? qmax : q;
}

uint64_t orc_morton(double x, double y, double z, const double lo[3], double side, int bits)
{
    const double scale = (double) (1ULL << bits) / side;
    const uint64_t qmax = (1ULL << bits) - 1;
    const uint64_t ix = quantise_(x, lo[0], scale, qmax);
    const uint64_t iy = quantise_(y, lo[1], scale, qmax);
    const uint64_t iz = quantise_(z, lo[2], scale, qmax);
    return (expand_bits_(ix) << 2) | (expand_bits_(iy) << 1) | expand_bits_(iz);
}

static int sys_bits_(int nsys)
{
    int b = 0;
    while ((1 << b) < nsys) ++b;
    return b;
}
static int morton_bits_(int sysbits) { int b = (63 - sysbits) / 3; return b > 21 ? 21 : b; }

typedef struct { uint64_t key; int64_t idx; } keyidx_;
static int cmp_keyidx_(const void *a, const void *b)
{
    const keyidx_ *x = (const keyidx_ *) a, *y = (const keyidx_ *) b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->idx < y->idx) ? -1 : (x->idx > y->idx);
}

#define PERMUTE_D_(f)                                                       \
    do { if (s->f) { for (int64_t k = 0; k < n; ++k) tmpd[k] = s->f[ord[k].idx]; memcpy(s->f, tmpd, sizeof(double) * n); } } while (0)
#define PERMUTE_I_(f)                                                       \
    do { if (s->f) { for (int64_t k = 0; k < n; ++k) tmpi[k] = s->f[ord[k].idx]; memcpy(s->f, tmpi, sizeof(int32_t) * n); } } while (0)

/* Per-system bounding cube (min corner, max side) used for the Morton keys. */
static void sys_boxes_(const orc_state *s, int nsys, double *lo, double *side)
{
    double *hi = (double *) malloc(sizeof(double) * 3 * nsys);
    for (int q = 0; q < nsys; ++q)
        for (int k = 0; k < 3; ++k) { lo[q * 3 + k] = INFINITY; hi[q * 3 + k] = -INFINITY; }
    for (int64_t p = 0; p < s->n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        const double c[3] = {s->x[p], s->y[p], s->z[p]};
        for (int k = 0; k < 3; ++k) {
            lo[q * 3 + k] = stdmin_(lo[q * 3 + k], c[k]);
            hi[q * 3 + k] = stdmax_(hi[q * 3 + k], c[k]);
        }
    }
    for (int q = 0; q < nsys; ++q) {
        double sd = 0;
        for (int k = 0; k < 3; ++k) sd = stdmax_(sd, hi[q * 3 + k] - lo[q * 3 + k]);
        side[q] = (sd > 0) ? sd : 1.0;
    }
    free(hi);
}

int orc_sort_state(orc_state *s, int nsys)
{
    const int64_t n = s->n;
    if (n <= 0) return 0;
    const int sb = sys_bits_(nsys), bits = morton_bits_(sb);
    double *lo = (double *) malloc(sizeof(double) * 3 * nsys), *side = (double *) malloc(sizeof(double) * nsys);
    sys_boxes_(s, nsys, lo, side);
    keyidx_ *ord = (keyidx_ *) malloc(sizeof(keyidx_) * n);
    for (int64_t p = 0; p < n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        ord[p].key = ((uint64_t) q << (3 * bits)) | orc_morton(s->x[p], s->y[p], s->z[p], lo + 3 * q, side[q], bits);
        ord[p].idx = p;
    }
    qsort(ord, (size_t) n, sizeof(keyidx_), cmp_keyidx_);
    double *tmpd = (double *) malloc(sizeof(double) * n);
    int32_t *tmpi = (int32_t *) malloc(sizeof(int32_t) * n);
    PERMUTE_D_(x); PERMUTE_D_(y); PERMUTE_D_(z); PERMUTE_D_(vx); PERMUTE_D_(vy); PERMUTE_D_(vz);
    PERMUTE_D_(ax); PERMUTE_D_(ay); PERMUTE_D_(az); PERMUTE_D_(gx); PERMUTE_D_(gy); PERMUTE_D_(gz);
    PERMUTE_D_(pot); PERMUTE_D_(m); PERMUTE_D_(h); PERMUTE_D_(rho); PERMUTE_D_(u); PERMUTE_D_(p);
    PERMUTE_D_(cs); PERMUTE_D_(divv); PERMUTE_D_(rotx); PERMUTE_D_(roty); PERMUTE_D_(rotz); PERMUTE_D_(gs);
    PERMUTE_D_(bal); PERMUTE_D_(du); PERMUTE_D_(dti); PERMUTE_D_(vhx); PERMUTE_D_(vhy); PERMUTE_D_(vhz);
    PERMUTE_D_(uh);
    PERMUTE_I_(mat); PERMUTE_I_(sys); PERMUTE_I_(ngb);
    if (s->id) {
        int64_t *tmpl = (int64_t *) malloc(sizeof(int64_t) * n);
        for (int64_t k = 0; k < n; ++k) tmpl[k] = s->id[ord[k].idx];
        memcpy(s->id, tmpl, sizeof(int64_t) * n);
        free(tmpl);
    }
    free(tmpd); free(tmpi); free(ord); free(lo); free(side);
    return 0;
}

/* Tree over the sorted particles.  A boundary between particles k-1 and k
 * exists at threshold T iff the smallest octree cell holding both contains
 * more than T particles; leaves use T = n_leaf, level l uses n_leaf*8^l, the
 * top level holds one node per system, i-groups use T = n_group. */
struct orc_tree {
    int64_t n;
    int nlevels;                /* levels 0..nlevels-1, top = nlevels-1 */
    int64_t *level_off;         /* nlevels+1 */
    int64_t nnodes;
    int64_t *nstart, *nend, *c0, *c1;
    double *mass, *mx, *com, *lo, *hi, *maxH; /* com/lo/hi: 3 per node */
    int64_t ngroups;
    int64_t *gstart;            /* ngroups+1 */
    int nsys;
    int64_t *sys_root;          /* node id of each system's root, -1 if empty */
};

static int64_t upper_bound_u64_(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t first = 0, len = n;
    while (len > 0) {
        const int64_t half = len >> 1;
        if (!(v < a[first + half])) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}
static int64_t lower_bound_u64_(const uint64_t *a, int64_t n, uint64_t v)
{
    int64_t first = 0, len = n;
    while (len > 0) {
        const int64_t half = len >> 1;
        if (a[first + half] < v) { first += half + 1; len = len - half - 1; }
        else len = half;
    }
    return first;
}

static int64_t *starts_from_counts_(const int64_t *cnt, int64_t n, int64_t T, int64_t *nout)
{
    int64_t c = 1;
    for (int64_t k = 1; k < n; ++k) c += (cnt[k] > T);
    int64_t *st = (int64_t *) malloc(sizeof(int64_t) * (c + 1));
    int64_t q = 0;
    st[q++] = 0;
    for (int64_t k = 1; k < n; ++k) if (cnt[k] > T) st[q++] = k;
    st[c] = n;
    *nout = c;
    return st;
}

orc_tree *orc_tree_build(const orc_state *s, const orc_params *prm)
{
    if (prm->n_group < 1 || prm->n_group > ORC_MAX_GROUP) {
        fprintf(stderr, "oracle: n_group %d outside [1, %d]\n", prm->n_group, ORC_MAX_GROUP);
        abort();
    }
    const int64_t n = s->n;
    const int nsys = prm->nsys > 0 ? prm->nsys : 1;
    const int sb = sys_bits_(nsys), bits = morton_bits_(sb);
    const int unused = 63 - sb - 3 * bits;
    orc_tree *t = (orc_tree *) calloc(1, sizeof(orc_tree));
    t->n = n; t->nsys = nsys;
    t->sys_root = (int64_t *) malloc(sizeof(int64_t) * nsys);
    for (int q = 0; q < nsys; ++q) t->sys_root[q] = -1;
    if (n <= 0) { t->nlevels = 0; t->level_off = (int64_t *) calloc(1, sizeof(int64_t)); t->gstart = (int64_t *) calloc(1, sizeof(int64_t)); return t; }

    /* keys of the (already sorted) particles */
    double *blo = (double *) malloc(sizeof(double) * 3 * nsys), *side = (double *) malloc(sizeof(double) * nsys);
    sys_boxes_(s, nsys, blo, side);
    uint64_t *key = (uint64_t *) malloc(sizeof(uint64_t) * n);
    for (int64_t p = 0; p < n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        key[p] = ((uint64_t) q << (3 * bits)) | orc_morton(s->x[p], s->y[p], s->z[p], blo + 3 * q, side[q], bits);
    }
    free(blo); free(side);
    /* LCA cell counts */
    int64_t *cnt = (int64_t *) malloc(sizeof(int64_t) * n);
    cnt[0] = INT64_MAX;
    int64_t max_sys = 0;
    {
        int64_t run = 1;
        for (int64_t k = 1; k < n; ++k) {
            const uint64_t a = key[k - 1], b = key[k];
            if ((a >> (3 * bits)) != (b >> (3 * bits))) { cnt[k] = INT64_MAX; if (run > max_sys) max_sys = run; run = 1; continue; }
            ++run;
            const uint64_t xo = a ^ b;
            const int p = (xo == 0) ? 63 : (__builtin_clzll(xo) - 1);
            int lev = (p - unused - sb) / 3;
            if (lev > bits) lev = bits;
            const int shift = 3 * (bits - lev);
            const uint64_t mask = (shift >= 64) ? ~0ULL : ((1ULL << shift) - 1);
            const uint64_t clo = a & ~mask, chi = clo | mask;
            cnt[k] = upper_bound_u64_(key, n, chi) - lower_bound_u64_(key, n, clo);
        }
        if (run > max_sys) max_sys = run;
    }
    /* levels */
    int nl = 1;
    {
        int64_t T = prm->n_leaf;
        while (T < max_sys) { T *= 8; ++nl; }
    }
    t->nlevels = nl;
    t->level_off = (int64_t *) malloc(sizeof(int64_t) * (nl + 1));
    int64_t **lst = (int64_t **) malloc(sizeof(int64_t *) * nl);
    int64_t *lcnt = (int64_t *) malloc(sizeof(int64_t) * nl);
    {
        int64_t T = prm->n_leaf;
        for (int l = 0; l < nl; ++l) {
            if (l == nl - 1) T = INT64_MAX - 1; /* top: systems only */
            lst[l] = starts_from_counts_(cnt, n, T, &lcnt[l]);
            T *= 8;
        }
    }
    t->level_off[0] = 0;
    for (int l = 0; l < nl; ++l) t->level_off[l + 1] = t->level_off[l] + lcnt[l];
    const int64_t nn = t->level_off[nl];
    t->nnodes = nn;
    t->nstart = (int64_t *) malloc(sizeof(int64_t) * nn);
    t->nend = (int64_t *) malloc(sizeof(int64_t) * nn);
    t->c0 = (int64_t *) malloc(sizeof(int64_t) * nn);
    t->c1 = (int64_t *) malloc(sizeof(int64_t) * nn);
    t->mass = (double *) malloc(sizeof(double) * nn);
    t->mx = (double *) malloc(sizeof(double) * 3 * nn);
    t->com = (double *) malloc(sizeof(double) * 3 * nn);
    t->lo = (double *) malloc(sizeof(double) * 3 * nn);
    t->hi = (double *) malloc(sizeof(double) * 3 * nn);
    t->maxH = (double *) malloc(sizeof(double) * nn);
    for (int l = 0; l < nl; ++l) {
        for (int64_t q = 0; q < lcnt[l]; ++q) {
            const int64_t id = t->level_off[l] + q;
            t->nstart[id] = lst[l][q];
            t->nend[id] = lst[l][q + 1];
            if (l == 0) { t->c0[id] = t->c1[id] = -1; }
            else {
                /* children: level l-1 nodes starting inside [start, end) */
                const int64_t a = lower_bound_u64_((const uint64_t *) lst[l - 1], lcnt[l - 1], (uint64_t) lst[l][q]);
                const int64_t b = lower_bound_u64_((const uint64_t *) lst[l - 1], lcnt[l - 1], (uint64_t) lst[l][q + 1]);
                t->c0[id] = t->level_off[l - 1] + a;
                t->c1[id] = t->level_off[l - 1] + b;
            }
        }
    }
    /* moments: leaves from particles, then each level from its children, all
     * summed sequentially in index order */
    for (int64_t id = 0; id < nn; ++id) {
        double M = 0, mx[3] = {0, 0, 0}, lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, mh = 0;
        if (t->c0[id] < 0) {
            for (int64_t p = t->nstart[id]; p < t->nend[id]; ++p) {
                const double c[3] = {s->x[p], s->y[p], s->z[p]};
                M += s->m[p];
                for (int k = 0; k < 3; ++k) { mx[k] += s->m[p] * c[k]; lo[k] = stdmin_(lo[k], c[k]); hi[k] = stdmax_(hi[k], c[k]); }
                mh = stdmax_(mh, ORC_SUPPORT * s->h[p]);
            }
        } else {
            for (int64_t c = t->c0[id]; c < t->c1[id]; ++c) {
                M += t->mass[c];
                for (int k = 0; k < 3; ++k) {
                    mx[k] += t->mx[3 * c + k];
                    lo[k] = stdmin_(lo[k], t->lo[3 * c + k]);
                    hi[k] = stdmax_(hi[k], t->hi[3 * c + k]);
                }
                mh = stdmax_(mh, t->maxH[c]);
            }
        }
        t->mass[id] = M; t->maxH[id] = mh;
        for (int k = 0; k < 3; ++k) {
            t->mx[3 * id + k] = mx[k];
            t->com[3 * id + k] = mx[k] / M;
            t->lo[3 * id + k] = lo[k];
            t->hi[3 * id + k] = hi[k];
        }
    }
    for (int64_t q = 0; q < lcnt[nl - 1]; ++q) {
        const int64_t id = t->level_off[nl - 1] + q;
        const int sy = s->sys ? s->sys[t->nstart[id]] : 0;
        t->sys_root[sy] = id;
    }
    /* i-groups: inside each level-2 node (the top level when the tree is
     * shallower), consecutive leaves are packed greedily while the group stays
     * within n_group particles */
    {
        int64_t *gs = (int64_t *) malloc(sizeof(int64_t) * (n + 1));
        int64_t ng = 0;
        const int pl = (nl - 1 < 2) ? nl - 1 : 2;
        /* leaf range of every level-pl node: leaves whose start lies inside it */
        for (int64_t q = t->level_off[pl]; q < t->level_off[pl + 1]; ++q) {
            const int64_t a = lower_bound_u64_((const uint64_t *) t->nstart, t->level_off[1], (uint64_t) t->nstart[q]);
            const int64_t b = lower_bound_u64_((const uint64_t *) t->nstart, t->level_off[1], (uint64_t) t->nend[q]);
            int64_t run = 0;
            for (int64_t c = a; c < b; ++c) {
                const int64_t sz = t->nend[c] - t->nstart[c];
                if (c == a || run + sz > prm->n_group) { gs[ng++] = t->nstart[c]; run = 0; }
                run += sz;
            }
        }
        gs[ng] = n;
        t->gstart = gs;
        t->ngroups = ng;
    }
    for (int l = 0; l < nl; ++l) free(lst[l]);
    free(lst); free(lcnt); free(cnt); free(key);
    return t;
}

void orc_tree_free(orc_tree *t)
{
    if (!t) return;
    free(t->level_off); free(t->nstart); free(t->nend); free(t->c0); free(t->c1); free(t->mass); free(t->mx);
    free(t->com); free(t->lo); free(t->hi); free(t->maxH); free(t->gstart); free(t->sys_root);
    free(t);
}

int64_t orc_tree_ngroups(const orc_tree *t) { return t->ngroups; }
const int64_t *orc_tree_group_start(const orc_tree *t) { return t->gstart; }
int64_t orc_tree_nnodes(const orc_tree *t) { return t->nnodes; }
int orc_tree_nlevels(const orc_tree *t) { return t->nlevels; }
const int64_t *orc_tree_level_offset(const orc_tree *t) { return t->level_off; }

void orc_tree_node(const orc_tree *t, int64_t k, int64_t *range2, int64_t *child2, double *mom)
{
    range2[0] = t->nstart[k]; range2[1] = t->nend[k];
    child2[0] = t->c0[k]; child2[1] = t->c1[k];
    mom[0] = t->mass[k];
    for (int d = 0; d < 3; ++d) { mom[1 + d] = t->com[3 * k + d]; mom[4 + d] = t->lo[3 * k + d]; mom[7 + d] = t->hi[3 * k + d]; }
    mom[10] = t->maxH[k];
}

void orc_tree_update_maxH(orc_tree *t, const orc_state *s)
{
    for (int64_t id = 0; id < t->nnodes; ++id) {
        double mh = 0;
        if (t->c0[id] < 0) for (int64_t p = t->nstart[id]; p < t->nend[id]; ++p) mh = stdmax_(mh, ORC_SUPPORT * s->h[p]);
        else for (int64_t c = t->c0[id]; c < t->c1[id]; ++c) mh = stdmax_(mh, t->maxH[c]);
        t->maxH[id] = mh;
    }
}

typedef struct { double slo[3], shi[3], glo[3], ghi[3], r2; } gbox_;

/* group box (positions), search box (positions +- 2.5 h) and, for the gather
 * lists, the squared reach beyond the group box: (max 2.5 h)^2 (1 + 1e-12) */
static void group_boxes_(const orc_state *s, int64_t i0, int64_t i1, gbox_ *b)
{
    double R = 0.0;
    for (int k = 0; k < 3; ++k) { b->slo[k] = b->glo[k] = INFINITY; b->shi[k] = b->ghi[k] = -INFINITY; }
    for (int64_t p = i0; p < i1; ++p) {
        const double c[3] = {s->x[p], s->y[p], s->z[p]};
        const double H = ORC_SUPPORT * s->h[p];
        for (int k = 0; k < 3; ++k) {
            b->glo[k] = stdmin_(b->glo[k], c[k]);
            b->ghi[k] = stdmax_(b->ghi[k], c[k]);
            b->slo[k] = stdmin_(b->slo[k], c[k] - H);
            b->shi[k] = stdmax_(b->shi[k], c[k] + H);
        }
        R = stdmax_(R, H);
    }
    b->r2 = (R * R) * (1.0 + 1.0e-12);
}

/* a gather list holds the particles inside the search box that are also
 * within the largest search radius of the group box (the others lie beyond
 * every member's support) */
static inline double box_dist2_(const gbox_ *b, const double c[3])
{
    double d2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        const double d = (c[k] < b->glo[k]) ? (b->glo[k] - c[k]) : ((c[k] > b->ghi[k]) ? (c[k] - b->ghi[k]) : 0.0);
        d2 = d2 + d * d;
    }
    return d2;
}

static inline int ptest_(const gbox_ *b, int kind, double x, double y, double z, double H)
{
    const double c[3] = {x, y, z};
    int in = 1;
    for (int k = 0; k < 3; ++k) in &= (b->slo[k] <= c[k]) & (c[k] <= b->shi[k]);
    if (kind == 0) return in && box_dist2_(b, c) <= b->r2;
    if (in) return in;
    int ov = 1;
    for (int k = 0; k < 3; ++k) ov &= (c[k] - H <= b->ghi[k]) & (c[k] + H >= b->glo[k]);
    return ov;
}

static inline int ntest_(const orc_tree *t, int64_t id, const gbox_ *b, int kind)
{
    const double *lo = t->lo + 3 * id, *hi = t->hi + 3 * id;
    int ov = 1;
    for (int k = 0; k < 3; ++k) ov &= (lo[k] <= b->shi[k]) & (hi[k] >= b->slo[k]);
    if (kind == 0) {
        double d2 = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = (lo[k] > b->ghi[k]) ? (lo[k] - b->ghi[k]) : ((hi[k] < b->glo[k]) ? (b->glo[k] - hi[k]) : 0.0);
            d2 = d2 + d * d;
        }
        return ov && d2 <= b->r2;
    }
    if (ov) return ov;
    const double mh = t->maxH[id];
    int ov2 = 1;
    for (int k = 0; k < 3; ++k) ov2 &= (lo[k] - mh <= b->ghi[k]) & (hi[k] + mh >= b->glo[k]);
    return ov2;
}

int64_t orc_list_sph(const orc_tree *t, const orc_state *s, int64_t g, int kind, int brute, int32_t *out, int64_t cap)
{
    const int64_t i0 = t->gstart[g], i1 = t->gstart[g + 1];
    gbox_ b;
    group_boxes_(s, i0, i1, &b);
    const int sy = s->sys ? s->sys[i0] : 0;
    const int64_t root = t->sys_root[sy];
    int64_t nout = 0;
    if (brute) {
        for (int64_t p = t->nstart[root]; p < t->nend[root]; ++p) {
            if (ptest_(&b, kind, s->x[p], s->y[p], s->z[p], ORC_SUPPORT * s->h[p])) {
                if (nout >= cap) return -1;
                out[nout++] = (int32_t) p;
            }
        }
        return nout;
    }
    /* level-synchronous walk from the system root */
    int64_t fcap = 1024, nf = 1;
    int64_t *cur = (int64_t *) malloc(sizeof(int64_t) * fcap), *nxt = (int64_t *) malloc(sizeof(int64_t) * fcap);
    cur[0] = root;
    for (int l = t->nlevels - 1; l >= 1; --l) {
        int64_t nn = 0;
        for (int64_t q = 0; q < nf; ++q) {
            const int64_t id = cur[q];
            if (!ntest_(t, id, &b, kind)) continue;
            for (int64_t c = t->c0[id]; c < t->c1[id]; ++c) {
                if (nn >= fcap) { fcap *= 2; nxt = (int64_t *) realloc(nxt, sizeof(int64_t) * fcap); cur = (int64_t *) realloc(cur, sizeof(int64_t) * fcap); }
                nxt[nn++] = c;
            }
        }
        int64_t *tmp = cur; cur = nxt; nxt = tmp; nf = nn;
    }
    for (int64_t q = 0; q < nf; ++q) {
        const int64_t id = cur[q];
        if (!ntest_(t, id, &b, kind)) continue;
        for (int64_t p = t->nstart[id]; p < t->nend[id]; ++p) {
            if (ptest_(&b, kind, s->x[p], s->y[p], s->z[p], ORC_SUPPORT * s->h[p])) {
                if (nout >= cap) { free(cur); free(nxt); return -1; }
                out[nout++] = (int32_t) p;
            }
        }
    }
    free(cur); free(nxt);
    return nout;
}

/* Monopole acceptance: the node's box must not touch the group's box, and
 * (node box max side)^2 < theta^2 * (distance from the group box to the
 * node's centre of mass)^2. */
static inline int accept_(const orc_tree *t, int64_t id, const gbox_ *b, double theta2)
{
    const double *lo = t->lo + 3 * id, *hi = t->hi + 3 * id, *c = t->com + 3 * id;
    int ov = 1;
    for (int k = 0; k < 3; ++k) ov &= (lo[k] <= b->ghi[k]) & (hi[k] >= b->glo[k]);
    if (ov) return 0;
    double d[3];
    for (int k = 0; k < 3; ++k) d[k] = (c[k] < b->glo[k]) ? (b->glo[k] - c[k]) : ((c[k] > b->ghi[k]) ? (c[k] - b->ghi[k]) : 0.0);
    const double d2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    double sz = hi[0] - lo[0];
    sz = stdmax_(sz, hi[1] - lo[1]);
    sz = stdmax_(sz, hi[2] - lo[2]);
    return (sz * sz) < (theta2 * d2);
}

int orc_list_grav(const orc_tree *t, const orc_state *s, int64_t g, double theta,
                  int32_t *epj, int64_t cap_j, int64_t *nj, int32_t *spj, int64_t cap_s, int64_t *ns)
{
    const int64_t i0 = t->gstart[g], i1 = t->gstart[g + 1];
    gbox_ b;
    group_boxes_(s, i0, i1, &b);
    const double theta2 = theta * theta;
    const int sy = s->sys ? s->sys[i0] : 0;
    int64_t fcap = 1024, nf = 1, cj = 0, cs = 0;
    int64_t *cur = (int64_t *) malloc(sizeof(int64_t) * fcap), *nxt = (int64_t *) malloc(sizeof(int64_t) * fcap);
    cur[0] = t->sys_root[sy];
    int rc = 0;
    for (int l = t->nlevels - 1; l >= 0 && rc == 0; --l) {
        int64_t nn = 0;
        for (int64_t q = 0; q < nf && rc == 0; ++q) {
            const int64_t id = cur[q];
            if (accept_(t, id, &b, theta2)) {
                if (cs >= cap_s) { rc = -1; break; }
                spj[cs++] = (int32_t) id;
            } else if (l == 0) {
                for (int64_t p = t->nstart[id]; p < t->nend[id]; ++p) {
                    if (cj >= cap_j) { rc = -1; break; }
                    epj[cj++] = (int32_t) p;
                }
            } else {
                for (int64_t c = t->c0[id]; c < t->c1[id]; ++c) {
                    if (nn >= fcap) { fcap *= 2; nxt = (int64_t *) realloc(nxt, sizeof(int64_t) * fcap); cur = (int64_t *) realloc(cur, sizeof(int64_t) * fcap); }
                    nxt[nn++] = c;
                }
            }
        }
        int64_t *tmp = cur; cur = nxt; nxt = tmp; nf = nn;
    }
    free(cur); free(nxt);
    *nj = cj; *ns = cs;
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Step driver (src/main.cpp:69-104, 125-173)                               */
/* ------------------------------------------------------------------------ */
#define LIST_CAP_ 262144

static void list_fail_(const char *what)
{
    fprintf(stderr, "oracle: %s list exceeds %d entries\n", what, LIST_CAP_);
    abort();
}

static double eos_eval_(const orc_eos_set *eos, int mat, double rho, double u, int prop)
{
    if (mat < 0 || mat >= eos->nmat || !eos->tab[mat]) return NAN;
    return orc_eos_bilinear(eos->tab[mat], rho, u, prop);
}

/* one density pass: lists from the current h, functor, write-back after all groups */
static void pass_density_(orc_state *s, orc_tree *t, const orc_params *prm)
{
    const int64_t G = t->ngroups;
    double *nrho = (double *) malloc(sizeof(double) * s->n), *nh = (double *) malloc(sizeof(double) * s->n);
    int32_t *nngb = (int32_t *) malloc(sizeof(int32_t) * s->n);
#pragma omp parallel
    {
        int32_t *lst = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
        orc_epj_dens *epj = (orc_epj_dens *) malloc(sizeof(orc_epj_dens) * LIST_CAP_);
        orc_epi_dens epi[ORC_MAX_GROUP];
        orc_res_dens res[ORC_MAX_GROUP];
        int32_t ngb[ORC_MAX_GROUP];
#pragma omp for schedule(dynamic, 4)
        for (int64_t g = 0; g < G; ++g) {
            const int64_t i0 = t->gstart[g], ni = t->gstart[g + 1] - i0;
            const int64_t nj = orc_list_sph(t, s, g, 0, !prm->fast_lists, lst, LIST_CAP_);
            if (nj < 0) list_fail_("density");
            for (int64_t q = 0; q < nj; ++q) {
                const int64_t p = lst[q];
                epj[q].mass = s->m[p]; epj[q].pos[0] = s->x[p]; epj[q].pos[1] = s->y[p]; epj[q].pos[2] = s->z[p]; epj[q].smth = s->h[p];
            }
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                epi[a].pos[0] = s->x[p]; epi[a].pos[1] = s->y[p]; epi[a].pos[2] = s->z[p];
                epi[a].mass = s->m[p]; epi[a].smth = s->h[p]; epi[a].id = s->id ? s->id[p] : p;
            }
            orc_clear_dens(res, (int) ni);
            orc_calc_density(epi, (int) ni, epj, (int) nj, res);
            orc_count_ngb(epi, (int) ni, epj, (int) nj, ngb);
            for (int64_t a = 0; a < ni; ++a) { nrho[i0 + a] = res[a].dens; nh[i0 + a] = res[a].smth; nngb[i0 + a] = ngb[a]; }
        }
        free(lst); free(epj);
    }
    memcpy(s->rho, nrho, sizeof(double) * s->n);
    memcpy(s->h, nh, sizeof(double) * s->n);
    if (s->ngb) memcpy(s->ngb, nngb, sizeof(int32_t) * s->n);
    free(nrho); free(nh); free(nngb);
}

static void pass_eos_(orc_state *s, const orc_eos_set *eos)
{
#pragma omp parallel for
    for (int64_t p = 0; p < s->n; ++p) {
        const int mt = s->mat ? s->mat[p] : 0;
        s->p[p] = eos_eval_(eos, mt, s->rho[p], s->u[p], 3);
        s->cs[p] = eos_eval_(eos, mt, s->rho[p], s->u[p], 4);
    }
}

static void pass_derivative_(orc_state *s, orc_tree *t, const orc_params *prm)
{
    const int64_t G = t->ngroups;
#pragma omp parallel
    {
        int32_t *lst = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
        orc_epj_drvt *epj = (orc_epj_drvt *) malloc(sizeof(orc_epj_drvt) * LIST_CAP_);
        orc_epi_drvt epi[ORC_MAX_GROUP];
        orc_res_drvt res[ORC_MAX_GROUP];
#pragma omp for schedule(dynamic, 4)
        for (int64_t g = 0; g < G; ++g) {
            const int64_t i0 = t->gstart[g], ni = t->gstart[g + 1] - i0;
            const int64_t nj = orc_list_sph(t, s, g, 0, !prm->fast_lists, lst, LIST_CAP_);
            if (nj < 0) list_fail_("derivative");
            for (int64_t q = 0; q < nj; ++q) {
                const int64_t p = lst[q];
                epj[q].mass = s->m[p]; epj[q].pos[0] = s->x[p]; epj[q].pos[1] = s->y[p]; epj[q].pos[2] = s->z[p];
                epj[q].vel[0] = s->vx[p]; epj[q].vel[1] = s->vy[p]; epj[q].vel[2] = s->vz[p]; epj[q].smth = s->h[p];
            }
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                epi[a].pos[0] = s->x[p]; epi[a].pos[1] = s->y[p]; epi[a].pos[2] = s->z[p];
                epi[a].vel[0] = s->vx[p]; epi[a].vel[1] = s->vy[p]; epi[a].vel[2] = s->vz[p];
                epi[a].smth = s->h[p]; epi[a].dens = s->rho[p];
            }
            orc_clear_drvt(res, (int) ni);
            orc_calc_derivative(epi, (int) ni, epj, (int) nj, res);
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                s->divv[p] = res[a].div_v;
                s->rotx[p] = res[a].rot_v[0]; s->roty[p] = res[a].rot_v[1]; s->rotz[p] = res[a].rot_v[2];
                s->bal[p] = orc_balsara(res[a].div_v, res[a].rot_v, s->cs[p], s->h[p]);
                s->gs[p] = res[a].grad_smth;
            }
        }
        free(lst); free(epj);
    }
}

static void pass_hydro_(orc_state *s, orc_tree *t, const orc_params *prm)
{
    const int64_t G = t->ngroups;
    orc_tree_update_maxH(t, s);
    double *nax = (double *) malloc(sizeof(double) * 5 * s->n);
#pragma omp parallel
    {
        int32_t *lst = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
        orc_epj_hydro *epj = (orc_epj_hydro *) malloc(sizeof(orc_epj_hydro) * LIST_CAP_);
        orc_epi_hydro epi[ORC_MAX_GROUP];
        orc_res_hydro res[ORC_MAX_GROUP];
#pragma omp for schedule(dynamic, 4)
        for (int64_t g = 0; g < G; ++g) {
            const int64_t i0 = t->gstart[g], ni = t->gstart[g + 1] - i0;
            const int64_t nj = orc_list_sph(t, s, g, 1, !prm->fast_lists, lst, LIST_CAP_);
            if (nj < 0) list_fail_("hydro");
            for (int64_t q = 0; q < nj; ++q) {
                const int64_t p = lst[q];
                orc_epj_hydro *e = &epj[q];
                e->pos[0] = s->x[p]; e->pos[1] = s->y[p]; e->pos[2] = s->z[p];
                e->vel[0] = s->vx[p]; e->vel[1] = s->vy[p]; e->vel[2] = s->vz[p];
                e->dens = s->rho[p]; e->mass = s->m[p]; e->smth = s->h[p]; e->pres = s->p[p];
                e->grad_smth = s->gs[p]; e->snds = s->cs[p]; e->Bal = s->bal[p]; e->AVa = 1.0; e->id = p;
            }
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                orc_epi_hydro *e = &epi[a];
                e->pos[0] = s->x[p]; e->pos[1] = s->y[p]; e->pos[2] = s->z[p];
                e->vel[0] = s->vx[p]; e->vel[1] = s->vy[p]; e->vel[2] = s->vz[p];
                e->smth = s->h[p]; e->dens = s->rho[p]; e->pres = s->p[p]; e->snds = s->cs[p];
                e->grad_smth = s->gs[p]; e->Bal = s->bal[p]; e->AVa = 1.0; e->id = p;
            }
            orc_clear_hydro(res, (int) ni);
            orc_calc_hydro(epi, (int) ni, epj, (int) nj, res);
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                nax[5 * p + 0] = res[a].acc[0]; nax[5 * p + 1] = res[a].acc[1]; nax[5 * p + 2] = res[a].acc[2];
                nax[5 * p + 3] = res[a].eng_dot; nax[5 * p + 4] = res[a].dt;
            }
        }
        free(lst); free(epj);
    }
    for (int64_t p = 0; p < s->n; ++p) {
        s->ax[p] = nax[5 * p + 0]; s->ay[p] = nax[5 * p + 1]; s->az[p] = nax[5 * p + 2];
        s->du[p] = nax[5 * p + 3]; s->dti[p] = nax[5 * p + 4];
    }
    free(nax);
}

static void pass_gravity_(orc_state *s, orc_tree *t, const orc_params *prm)
{
    const int64_t G = t->ngroups;
#pragma omp parallel
    {
        int32_t *lj = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_), *ls = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
        orc_epj_grav *epj = (orc_epj_grav *) malloc(sizeof(orc_epj_grav) * LIST_CAP_);
        orc_spj_mono *spj = (orc_spj_mono *) malloc(sizeof(orc_spj_mono) * LIST_CAP_);
        orc_epi_grav epi[ORC_MAX_GROUP];
        orc_res_grav res[ORC_MAX_GROUP];
#pragma omp for schedule(dynamic, 4)
        for (int64_t g = 0; g < G; ++g) {
            const int64_t i0 = t->gstart[g], ni = t->gstart[g + 1] - i0;
            int64_t nj = 0, ns = 0;
            if (orc_list_grav(t, s, g, prm->theta, lj, LIST_CAP_, &nj, ls, LIST_CAP_, &ns)) list_fail_("gravity");
            for (int64_t q = 0; q < nj; ++q) {
                const int64_t p = lj[q];
                epj[q].pos[0] = s->x[p]; epj[q].pos[1] = s->y[p]; epj[q].pos[2] = s->z[p]; epj[q].mass = s->m[p]; epj[q].id = p;
            }
            for (int64_t q = 0; q < ns; ++q) {
                const int64_t id = ls[q];
                spj[q].mass = t->mass[id];
                for (int k = 0; k < 3; ++k) spj[q].pos[k] = t->com[3 * id + k];
            }
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                epi[a].pos[0] = s->x[p]; epi[a].pos[1] = s->y[p]; epi[a].pos[2] = s->z[p];
                epi[a].eps2 = s->h[p] * s->h[p]; epi[a].id = p;
            }
            orc_clear_grav(res, (int) ni);
            orc_calc_grav_ep(epi, (int) ni, epj, (int) nj, res);
            orc_calc_grav_sp(epi, (int) ni, spj, (int) ns, res);
            for (int64_t a = 0; a < ni; ++a) {
                const int64_t p = i0 + a;
                s->gx[p] = res[a].acc[0]; s->gy[p] = res[a].acc[1]; s->gz[p] = res[a].acc[2]; s->pot[p] = res[a].pot;
            }
        }
        free(lj); free(ls); free(epj); free(spj);
    }
}

/* addExternalForce (src/GI.h:458-465): acc += -vel * 0.05 / dt */
static void ext_force_(orc_state *s, const orc_params *prm, const double *time_sys)
{
    for (int64_t p = 0; p < s->n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        if (time_sys[q] >= prm->ext_force_until) continue;
        const double inv = 1.0 / s->dti[p];
        s->ax[p] += ((-s->vx[p]) * 0.05) * inv;
        s->ay[p] += ((-s->vy[p]) * 0.05) * inv;
        s->az[p] += ((-s->vz[p]) * 0.05) * inv;
    }
}

/* getTimeStepGlobal (src/integral.h:3-21), live branch, per system. */
static void timestep_(const orc_state *s, int nsys, double *dt_sys)
{
    for (int q = 0; q < nsys; ++q) dt_sys[q] = 1.0e+30;
    for (int64_t p = 0; p < s->n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        dt_sys[q] = stdmin_(dt_sys[q], stdmin_(s->dti[p], 1.0e+30));
    }
}

static void forces_(orc_state *s, const orc_eos_set *eos, const orc_params *prm)
{
    orc_tree *t = orc_tree_build(s, prm);
    for (int loop = 0; loop <= 3; ++loop) pass_density_(s, t, prm);
    pass_eos_(s, eos);
    pass_derivative_(s, t, prm);
    pass_hydro_(s, t, prm);
    pass_gravity_(s, t, prm);
    orc_tree_free(t);
}

int orc_sim_init(orc_state *s, const orc_eos_set *eos, const orc_params *prm, double *dt_sys, double *time_sys)
{
    const int nsys = prm->nsys > 0 ? prm->nsys : 1;
    /* RealPtcl::initialize (src/class.h:196-199) and its constructor (:120-124) */
    for (int64_t p = 0; p < s->n; ++p) {
        s->h[p] = ORC_SMTH * pow(s->m[p] / s->rho[p], 1.0 / 3.0);
        s->gs[p] = 1.0;
        s->bal[p] = 1.0;
    }
    orc_sort_state(s, nsys);
    forces_(s, eos, prm);
    timestep_(s, nsys, dt_sys);
    ext_force_(s, prm, time_sys);
    return 0;
}

int orc_sim_step(orc_state *s, const orc_eos_set *eos, const orc_params *prm, double *dt_sys, double *time_sys)
{
    const int nsys = prm->nsys > 0 ? prm->nsys : 1;
    for (int64_t p = 0; p < s->n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        const double dt = dt_sys[q];
        const double ht = 0.5 * dt;
        /* initialKick (src/class.h:201-205) */
        s->vhx[p] = s->vx[p] + ht * (s->ax[p] + s->gx[p]);
        s->vhy[p] = s->vy[p] + ht * (s->ay[p] + s->gy[p]);
        s->vhz[p] = s->vz[p] + ht * (s->az[p] + s->gz[p]);
        s->uh[p] = s->u[p] + ht * s->du[p];
        /* fullDrift (:207-213) */
        s->x[p] += dt * s->vhx[p];
        s->y[p] += dt * s->vhy[p];
        s->z[p] += dt * s->vhz[p];
        /* predict (:215-219) */
        s->vx[p] += dt * (s->ax[p] + s->gx[p]);
        s->vy[p] += dt * (s->ay[p] + s->gy[p]);
        s->vz[p] += dt * (s->az[p] + s->gz[p]);
        s->u[p] += dt * s->du[p];
    }
    for (int q = 0; q < nsys; ++q) time_sys[q] += dt_sys[q];
    orc_sort_state(s, nsys);
    forces_(s, eos, prm);
    ext_force_(s, prm, time_sys);
    for (int64_t p = 0; p < s->n; ++p) {
        const int q = s->sys ? s->sys[p] : 0;
        const double ht = 0.5 * dt_sys[q];
        /* finalKick (:221-225) and dampMotion (:227-229) */
        s->vx[p] = s->vhx[p] + ht * (s->ax[p] + s->gx[p]);
        s->vy[p] = s->vhy[p] + ht * (s->ay[p] + s->gy[p]);
        s->vz[p] = s->vhz[p] + ht * (s->az[p] + s->gz[p]);
        s->u[p] = s->uh[p] + ht * s->du[p];
        s->vx[p] *= prm->damping; s->vy[p] *= prm->damping; s->vz[p] *= prm->damping;
    }
    timestep_(s, nsys, dt_sys);
    return 0;
}

int orc_state_term_scales(const orc_state *s, const orc_params *prm, double *acc_scale, double *du_scale,
                          double *div_scale, double *rot_scale, double *grav_scale)
{
    orc_tree *t = orc_tree_build(s, prm);
    orc_tree_update_maxH(t, s);
    int32_t *lst = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
    int32_t *ls = (int32_t *) malloc(sizeof(int32_t) * LIST_CAP_);
    orc_epj_hydro *eh = (orc_epj_hydro *) malloc(sizeof(orc_epj_hydro) * LIST_CAP_);
    orc_epj_drvt *ed = (orc_epj_drvt *) malloc(sizeof(orc_epj_drvt) * LIST_CAP_);
    for (int64_t g = 0; g < t->ngroups; ++g) {
        const int64_t i0 = t->gstart[g], ni = t->gstart[g + 1] - i0;
        orc_epi_hydro ih[ORC_MAX_GROUP];
        orc_epi_drvt id[ORC_MAX_GROUP];
        for (int64_t a = 0; a < ni; ++a) {
            const int64_t p = i0 + a;
            ih[a].pos[0] = s->x[p]; ih[a].pos[1] = s->y[p]; ih[a].pos[2] = s->z[p];
            ih[a].vel[0] = s->vx[p]; ih[a].vel[1] = s->vy[p]; ih[a].vel[2] = s->vz[p];
            ih[a].smth = s->h[p]; ih[a].dens = s->rho[p]; ih[a].pres = s->p[p]; ih[a].snds = s->cs[p];
            ih[a].grad_smth = s->gs[p]; ih[a].Bal = s->bal[p]; ih[a].AVa = 1.0; ih[a].id = p;
            memcpy(id[a].pos, ih[a].pos, sizeof id[a].pos);
            memcpy(id[a].vel, ih[a].vel, sizeof id[a].vel);
            id[a].smth = s->h[p]; id[a].dens = s->rho[p];
        }
        int64_t nj = orc_list_sph(t, s, g, 1, 0, lst, LIST_CAP_);
        if (nj < 0) list_fail_("scale");
        for (int64_t q = 0; q < nj; ++q) {
            const int64_t p = lst[q];
            orc_epj_hydro *e = &eh[q];
            e->pos[0] = s->x[p]; e->pos[1] = s->y[p]; e->pos[2] = s->z[p];
            e->vel[0] = s->vx[p]; e->vel[1] = s->vy[p]; e->vel[2] = s->vz[p];
            e->dens = s->rho[p]; e->mass = s->m[p]; e->smth = s->h[p]; e->pres = s->p[p];
            e->grad_smth = s->gs[p]; e->snds = s->cs[p]; e->Bal = s->bal[p]; e->AVa = 1.0; e->id = p;
        }
        orc_hydro_term_scale(ih, (int) ni, eh, (int) nj, acc_scale + i0, du_scale + i0);
        nj = orc_list_sph(t, s, g, 0, 0, lst, LIST_CAP_);
        if (nj < 0) list_fail_("scale");
        for (int64_t q = 0; q < nj; ++q) {
            const int64_t p = lst[q];
            ed[q].mass = s->m[p]; ed[q].pos[0] = s->x[p]; ed[q].pos[1] = s->y[p]; ed[q].pos[2] = s->z[p];
            ed[q].vel[0] = s->vx[p]; ed[q].vel[1] = s->vy[p]; ed[q].vel[2] = s->vz[p]; ed[q].smth = s->h[p];
        }
        orc_drvt_term_scale(id, (int) ni, ed, (int) nj, div_scale + i0, rot_scale + i0);
        int64_t ngj = 0, ngs = 0;
        if (orc_list_grav(t, s, g, prm->theta, lst, LIST_CAP_, &ngj, ls, LIST_CAP_, &ngs)) list_fail_("scale");
        for (int64_t a = 0; a < ni; ++a) {
            const int64_t p = i0 + a;
            const double e2 = s->h[p] * s->h[p];
            double sg = 0;
            for (int64_t q = 0; q < ngj; ++q) {
                const int64_t j = lst[q];
                const double d[3] = {s->x[p] - s->x[j], s->y[p] - s->y[j], s->z[p] - s->z[j]};
                sg += ORC_G * s->m[j] / (dot3_(d, d) + e2);
            }
            for (int64_t q = 0; q < ngs; ++q) {
                const int64_t k = ls[q];
                const double d[3] = {s->x[p] - t->com[3 * k], s->y[p] - t->com[3 * k + 1], s->z[p] - t->com[3 * k + 2]};
                sg += ORC_G * t->mass[k] / (dot3_(d, d) + e2);
            }
            grav_scale[p] = sg;
        }
    }
    free(lst); free(ls); free(eh); free(ed);
    orc_tree_free(t);
    return 0;
}

double orc_wtime(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
