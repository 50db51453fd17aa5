/*
 * sph_oracle.h -- CPU oracle for the FDPS_SPH per-step interaction path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; it is used by tests/, by __graft_entry__.smoke() and by the
 * cpu_baseline leg of bench.py, always as the checker, never as the product.
 *
 * It restates, in plain C with a fixed sequential j-order, the arithmetic of
 *   src/kernel.h      (Wendland C6 W / gradW)
 *   src/force.h       (CalcDensity, CalcDerivative, CalcHydroForce,
 *                      CalcGravityForce<EPJ::Grav>, CalcGravityForce<SPJMonopole>)
 *   src/class.h       (EPI/EPJ/RESULT layouts, copyFromForce incl. Balsara,
 *                      initialize / initialKick / fullDrift / predict / finalKick)
 *   src/EoS.h, src/interpolation.h (table reader, BilinearInterpolation,
 *                      RestrictedBilinearInterpolation)
 *   src/integral.h    (getTimeStepGlobal)
 *   src/GI.h:458-465  (addExternalForce)
 *   src/main.cpp:92-104,125-173 (step sequence)
 * plus a restatement of the GPU library's own tree / group / interaction-list
 * construction (which replaces FDPS, absent here), so that the oracle and the
 * GPU path can be fed -- or can build -- identical lists.
 */
#ifndef SPH_ORACLE_H
#define SPH_ORACLE_H

#define ORC_MAX_GROUP 1024 /* the most particles an i-group may hold */
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- boundary structs: byte layouts of src/class.h (all members 8 B) ---- */
typedef struct { double pos[3]; double mass; double smth; int64_t id; } orc_epi_dens;      /* 48 */
typedef struct { double mass; double pos[3]; double smth; } orc_epj_dens;                 /* 40 */
typedef struct { double dens; double smth; } orc_res_dens;                                /* 16 */
typedef struct { double pos[3]; double vel[3]; double smth; double dens; } orc_epi_drvt;  /* 64 */
typedef struct { double mass; double pos[3]; double vel[3]; double smth; } orc_epj_drvt;  /* 64 */
typedef struct { double div_v; double rot_v[3]; double grad_smth; } orc_res_drvt;         /* 40 */
typedef struct { double pos[3]; double vel[3]; double smth, dens, pres, snds, grad_smth, Bal, AVa; int64_t id; } orc_epi_hydro; /* 112 */
typedef struct { double pos[3]; double vel[3]; double dens, mass, smth, pres, grad_smth, snds, Bal, AVa; int64_t id; } orc_epj_hydro; /* 120 */
typedef struct { double acc[3]; double eng_dot; double dt; } orc_res_hydro;               /* 40 */
typedef struct { double pos[3]; double eps2; int64_t id; } orc_epi_grav;                  /* 40 */
typedef struct { double pos[3]; double mass; int64_t id; } orc_epj_grav;                  /* 40 */
typedef struct { double mass; double pos[3]; } orc_spj_mono;                              /* 32 */
typedef struct { double acc[3]; double pot; double dt; } orc_res_grav;                    /* 40 */

/* ---- kernel (src/kernel.h) ---- */
double orc_W(const double dr[3], double h);
void   orc_gradW(const double dr[3], double h, double out[3]);

/* ---- functors (src/force.h); results are accumulated into res[], which the
 *      caller clears with the matching orc_clear_* (RESULT::*::clear) ---- */
void orc_clear_dens(orc_res_dens *r, int n);
void orc_clear_drvt(orc_res_drvt *r, int n);
void orc_clear_hydro(orc_res_hydro *r, int n);
void orc_clear_grav(orc_res_grav *r, int n);
void orc_calc_density(const orc_epi_dens *epi, int nip, const orc_epj_dens *epj, int njp, orc_res_dens *res);
void orc_calc_derivative(const orc_epi_drvt *epi, int nip, const orc_epj_drvt *epj, int njp, orc_res_drvt *res);
void orc_calc_hydro(const orc_epi_hydro *epi, int nip, const orc_epj_hydro *epj, int njp, orc_res_hydro *res);
void orc_calc_grav_ep(const orc_epi_grav *epi, int nip, const orc_epj_grav *epj, int njp, orc_res_grav *res);
void orc_calc_grav_sp(const orc_epi_grav *epi, int nip, const orc_spj_mono *spj, int njp, orc_res_grav *res);
/* in-support neighbour count of each i (j with W(dr,h_i) support 1-s > 0) */
void orc_count_ngb(const orc_epi_dens *epi, int nip, const orc_epj_dens *epj, int njp, int32_t *ngb);
/* Balsara switch of RealPtcl::copyFromForce(RESULT::Drvt) (src/class.h:132-142) */
double orc_balsara(double div_v, const double rot_v[3], double snds, double smth);

/* ---- EoS tables (src/EoS.h, src/interpolation.h) ---- */
typedef struct orc_table orc_table;
orc_table *orc_table_load(const char *path);                          /* NULL on failure */
orc_table *orc_table_from_cells(int nrho, int nu, const double *cells); /* cells[nrho][nu][6] */
/* Analytic Tillotson material, prm = (rho0, u0, u_iv, u_cv, A, B, a, b,
 * alpha, beta) (src/EoS.h:91-113, 446); evaluated by orc_eos_bilinear for
 * props 3 (P) and 4 (c_s), NaN for the others; no cells (dims 0 x 0). */
orc_table *orc_table_tillotson(const double prm[10]);
void   orc_tillotson(const double prm[10], double rho, double u, double *pres, double *snds);
void   orc_table_free(orc_table *t);
int    orc_table_dims(const orc_table *t, int *nrho, int *nu);
const double *orc_table_cells(const orc_table *t);                    /* [nrho][nu][6] */
double orc_eos_bilinear(const orc_table *t, double rho, double u, int prop);
/* RestrictedBilinearInterpolation::interpolate (u from rho, S); abs_int selects
 * the `int abs(int)` overload resolution for the nearest-S search (1) or the
 * double one (0).  Index paths the reference leaves undefined return NaN. */
double orc_eos_internal_energy(const orc_table *t, double rho, double S, int abs_int);

/* ---- particle state (SoA, caller-owned arrays of length n) ---- */
typedef struct {
    int64_t n;
    double *x, *y, *z, *vx, *vy, *vz;
    double *ax, *ay, *az;        /* hydro acc (+ external force) */
    double *gx, *gy, *gz, *pot;  /* gravity */
    double *m, *h, *rho, *u, *p, *cs;
    double *divv, *rotx, *roty, *rotz, *gs, *bal, *du, *dti;
    double *vhx, *vhy, *vhz, *uh;
    int32_t *mat, *sys, *ngb;
    int64_t *id;
} orc_state;

typedef struct {
    int n_leaf, n_group;     /* tree leaf (<= 64) / i-group (<= ORC_MAX_GROUP) particle limits */
    double theta;            /* gravity opening angle */
    double damping;          /* dampMotion factor (src/GI.h:32) */
    double ext_force_until;  /* addExternalForce window (5000 s, src/GI.h:459) */
    int nsys;                /* independent systems (1 for a single impact) */
    int fast_lists;          /* 1: tree-walk lists (same output as brute force) */
} orc_params;

/* One EoS table per material id. */
typedef struct { int nmat; const orc_table *const *tab; } orc_eos_set;

/* Sort particles by (system, Morton key) and permute every state array. */
int orc_sort_state(orc_state *s, int nsys);
/* Initial force evaluation (src/main.cpp:69-104); dt_sys[nsys] out. */
int orc_sim_init(orc_state *s, const orc_eos_set *eos, const orc_params *prm, double *dt_sys, double *time_sys);
/* One time step (src/main.cpp:125-173); dt_sys/time_sys in/out. */
int orc_sim_step(orc_state *s, const orc_eos_set *eos, const orc_params *prm, double *dt_sys, double *time_sys);

/* ---- tree / list restatement (valid after orc_sort_state) ---- */
typedef struct orc_tree orc_tree;
orc_tree *orc_tree_build(const orc_state *s, const orc_params *prm);
void   orc_tree_free(orc_tree *t);
int64_t orc_tree_ngroups(const orc_tree *t);
const int64_t *orc_tree_group_start(const orc_tree *t);   /* ngroups+1 */
int64_t orc_tree_nnodes(const orc_tree *t);
int    orc_tree_nlevels(const orc_tree *t);
const int64_t *orc_tree_level_offset(const orc_tree *t);   /* nlevels+1 */
void   orc_tree_node(const orc_tree *t, int64_t k, int64_t *range2, int64_t *child2, double *mom /*[mass, com3, lo3, hi3, maxH]*/);
void   orc_tree_update_maxH(orc_tree *t, const orc_state *s);
/* kind: 0 gather, 1 symmetric.  brute=1 uses an O(N) scan per group.
 * Returns the list length, -1 if cap exceeded. */
int64_t orc_list_sph(const orc_tree *t, const orc_state *s, int64_t g, int kind, int brute, int32_t *out, int64_t cap);
/* Gravity list: epj particle indices, spj node ids; returns 0 or -1 (cap). */
int orc_list_grav(const orc_tree *t, const orc_state *s, int64_t g, double theta,
                  int32_t *epj, int64_t cap_j, int64_t *nj, int32_t *spj, int64_t cap_s, int64_t *ns);

/* Scale on which sums whose terms cancel are compared (test diagnostics),
 * per i: the size of the summed terms plus their first-order response to a
 * relative change of the inputs that two correct evaluations compute only to
 * rounding -- the smoothing lengths (|d gradW / d ln h|, large for pairs near
 * the edge of the support, where gradW ~ (1 - s)^7) and, through the
 * pressure, the density (f c^2 / rho beside |A| = |f P / rho^2|).  hydro:
 * sum_j max_k m_j (S_i (|gWi_k| + |dgWi_k|) + S_j (...) + |AV| (|gW_k| +
 * |dgW_k|)) and sum_j m_j (S_i + |AV|/2) sum_k |dv_k| (|gW_k| + |dgW_k|);
 * derivative: sum_j m_j sum_k |dv_k| (|gW_k| + |dgW_k|) / rho_i and the same
 * with the terms of dv x gW; gravity: sum_j G m_j / (r^2 + eps_i^2).
 * Positions and velocities are taken as shared bit for bit. */
void orc_hydro_term_scale(const orc_epi_hydro *epi, int nip, const orc_epj_hydro *epj, int njp, double *acc_scale,
                          double *du_scale);
void orc_drvt_term_scale(const orc_epi_drvt *epi, int nip, const orc_epj_drvt *epj, int njp, double *div_scale,
                         double *rot_scale);
/* The same scales for every particle of a sorted state, over the lists the
 * step would build (symmetric for hydro, gather for the derivative, the
 * monopole tree for gravity). */
int orc_state_term_scales(const orc_state *s, const orc_params *prm, double *acc_scale, double *du_scale,
                          double *div_scale, double *rot_scale, double *grav_scale);

/* Morton key of one position inside a cube [lo, lo+side) with `bits` per axis. */
uint64_t orc_morton(double x, double y, double z, const double lo[3], double side, int bits);

/* Wall-clock timing helper for the baseline. */
double orc_wtime(void);
int    orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
