/*
 * fdps_sph_gpu.h -- C-ABI of the MI355X (gfx950) implementation of the
 * FDPS_SPH per-step interaction path.
 *
 * Three families of entry points:
 *
 *  1. FDPS multiwalk dispatch/retrieve pairs, one per interaction functor.
 *     They replace the CPU functors of src/force.h, which FDPS calls as
 *        f(const EPI* epi, S32 n_ip, const EPJ* epj, S32 n_jp, RESULT* res)
 *     from TreeForForce*::calcForceAllAndWriteBack (src/main.cpp:139-146).
 *     FDPS's calcForceAllAndWriteBackMultiWalk(dispatch, retrieve, tag_max,
 *     psys, dinfo, n_walk_limit) hands a batch of n_walk (i-group, j-list)
 *     pairs to `dispatch` and collects results with `retrieve`.  The structs
 *     below are byte-compatible with the reference's EPI/EPJ/RESULT classes
 *     (src/class.h:36-86, 233-453) and FDPS's PS::SPJMonopole, so FDPS's
 *     arrays can be passed unchanged.
 *
 *  2. EoS evaluation (src/EoS.h, src/interpolation.h): table loading in the
 *     reference's rho-u format and per-particle P, c_s, T, S and u(rho,S).
 *
 *  3. A device-resident simulation that keeps particles in HBM and runs the
 *     whole step of src/main.cpp:125-173 on the GPU (tree, interaction lists,
 *     the four functors, EoS, integrator, dt reduction).
 *
 * All functions return SPHGPU_OK (0) or a negative error code; the message
 * of the last error on the calling thread is available from
 * sphgpu_last_error().  A context is used by one host thread at a time.
 */
#ifndef FDPS_SPH_GPU_H
#define FDPS_SPH_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define SPHGPU_OK 0
#define SPHGPU_ERR_ARG (-1)      /* null pointer, negative size, bad enum */
#define SPHGPU_ERR_CAPACITY (-2) /* a list or batch exceeds its capacity */
#define SPHGPU_ERR_TAG (-3)      /* retrieve of a tag with no dispatched batch */
#define SPHGPU_ERR_HIP (-4)      /* HIP runtime failure */
#define SPHGPU_ERR_EOS (-5)      /* material id without a loaded table, bad table */
#define SPHGPU_ERR_STATE (-6)    /* call out of order (e.g. step before init) */

/* ---- precision of a context ---- */
#define SPHGPU_PREC_FP64 0  /* the reference's FP64 arithmetic, sequential j-order */
#define SPHGPU_PREC_MIXED 1 /* FP32 interactions relative to each group's centre,
                               Kahan-compensated FP32 sums; FP64 positions, EoS,
                               integrator */

/* ---- boundary structs (byte layouts of src/class.h; all members 8 B) ---- */
typedef struct { double pos[3]; double mass; double smth; int64_t id; } sphgpu_epi_dens;      /* EPI::Dens   src/class.h:233-258 */
typedef struct { double mass; double pos[3]; double smth; } sphgpu_epj_dens;                 /* EPJ::Dens   src/class.h:342-365 */
typedef struct { double dens; double smth; } sphgpu_res_dens;                                /* RESULT::Dens src/class.h:36-45 */
typedef struct { double pos[3]; double vel[3]; double smth; double dens; } sphgpu_epi_drvt;  /* EPI::Drvt   src/class.h:260-281 */
typedef struct { double mass; double pos[3]; double vel[3]; double smth; } sphgpu_epj_drvt;  /* EPJ::Drvt   src/class.h:367-392 */
typedef struct { double div_v; double rot_v[3]; double grad_smth; } sphgpu_res_drvt;         /* RESULT::Drvt src/class.h:47-58 */
typedef struct {
    double pos[3]; double vel[3]; double smth, dens, pres, snds, grad_smth, Bal, AVa; int64_t id;
} sphgpu_epi_hydro;                                                                          /* EPI::Hydro  src/class.h:283-315 */
typedef struct {
    double pos[3]; double vel[3]; double dens, mass, smth, pres, grad_smth, snds, Bal, AVa; int64_t id;
} sphgpu_epj_hydro;                                                                          /* EPJ::Hydro  src/class.h:394-432 */
typedef struct { double acc[3]; double eng_dot; double dt; } sphgpu_res_hydro;               /* RESULT::Hydro src/class.h:61-72 */
typedef struct { double pos[3]; double eps2; int64_t id; } sphgpu_epi_grav;                  /* EPI::Grav   src/class.h:317-338 */
typedef struct { double pos[3]; double mass; int64_t id; } sphgpu_epj_grav;                  /* EPJ::Grav   src/class.h:434-453 */
typedef struct { double mass; double pos[3]; } sphgpu_spj_monopole;                          /* PS::SPJMonopole (FDPS) */
typedef struct { double acc[3]; double pot; double dt; } sphgpu_res_grav;                    /* RESULT::Grav src/class.h:75-86 */

typedef struct sphgpu_ctx sphgpu_ctx;

/* ---- context ---- */
int32_t sphgpu_create(sphgpu_ctx **ctx, int32_t device, int32_t precision);
int32_t sphgpu_destroy(sphgpu_ctx *ctx);
/* Switches a context between SPHGPU_PREC_FP64 and SPHGPU_PREC_MIXED for the
 * work it launches from then on (a simulation's state is the same in both).
 * A process is given a few hardware queues (4 by default) and a context runs
 * on two streams, so a second context may share a queue with the first and
 * lose the overlap of its two streams: one context switched between the
 * precision forms keeps both forms on the same two queues. */
int32_t sphgpu_set_precision(sphgpu_ctx *ctx, int32_t precision);
const char *sphgpu_last_error(void);
int32_t sphgpu_version(void);

/* ---- multiwalk functors ----
 * dispatch copies the inputs (they may be reused by the caller as soon as it
 * returns) and launches the kernels on the context's stream; retrieve(tag)
 * waits for that batch and ASSIGNS the final RESULT values (including the
 * in-functor finalisation: smth for dens, grad_smth and /dens for drvt, dt for
 * hydro) to res[w][0..n_epi[w]).  tag in [0, 64). */

/* CalcDensity (src/force.h:6-24) */
int32_t sphgpu_dens_dispatch(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk,
                             const sphgpu_epi_dens *const epi[], const int32_t n_epi[],
                             const sphgpu_epj_dens *const epj[], const int32_t n_epj[]);
int32_t sphgpu_dens_retrieve(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_dens *const res[]);
/* CalcDerivative (src/force.h:86-107) */
int32_t sphgpu_drvt_dispatch(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk,
                             const sphgpu_epi_drvt *const epi[], const int32_t n_epi[],
                             const sphgpu_epj_drvt *const epj[], const int32_t n_epj[]);
int32_t sphgpu_drvt_retrieve(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_drvt *const res[]);
/* CalcHydroForce (src/force.h:109-148) */
int32_t sphgpu_hydro_dispatch(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk,
                              const sphgpu_epi_hydro *const epi[], const int32_t n_epi[],
                              const sphgpu_epj_hydro *const epj[], const int32_t n_epj[]);
int32_t sphgpu_hydro_retrieve(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                              sphgpu_res_hydro *const res[]);
/* CalcGravityForce<EPJ::Grav> + CalcGravityForce<PS::SPJMonopole>
 * (src/force.h:150-172); spj/n_spj may be NULL for a P2P-only batch.  The
 * EPJ list of a walk is summed before its SPJ list, as FDPS calls the two
 * functors. */
int32_t sphgpu_grav_dispatch(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk,
                             const sphgpu_epi_grav *const epi[], const int32_t n_epi[],
                             const sphgpu_epj_grav *const epj[], const int32_t n_epj[],
                             const sphgpu_spj_monopole *const spj[], const int32_t n_spj[]);
int32_t sphgpu_grav_retrieve(sphgpu_ctx *ctx, int32_t tag, int32_t n_walk, const int32_t n_epi[],
                             sphgpu_res_grav *const res[]);

/* ---- EoS (src/EoS.h, src/interpolation.h) ---- */
#define SPHGPU_EOS_TEMPERATURE 2
#define SPHGPU_EOS_PRESSURE 3
#define SPHGPU_EOS_SOUNDSPEED 4
#define SPHGPU_EOS_ENTROPY 5
#define SPHGPU_MAX_MATERIALS 16
/* Reads a table in the reference format (src/EoS.h:290-385) into material id. */
int32_t sphgpu_eos_load_table(sphgpu_ctx *ctx, int32_t material, const char *path);
/* Sets a table from cells[nrho][nu][6] = (rho, u, T, P, c_s, S). */
int32_t sphgpu_eos_set_table(sphgpu_ctx *ctx, int32_t material, int32_t nrho, int32_t nu, const double *cells);
/* Analytic Tillotson material (src/EoS.h:91-113): the compressed/cold
 * branch P_co and the expanded branch P_ex of the reference's Tillotson
 * class, which the reference defines but never calls, in Tillotson's regimes
 * (P_co for rho >= rho0 or u <= u_iv, P_ex for u >= u_cv, linear in u between),
 * c_s^2 = dP/drho + P/rho^2 dP/du from their exact derivatives, c_s at least
 * 1e-12 m/s (the ANEOS tables' floor).  params =
 * {rho0, u0, u_iv, u_cv, A, B, a, b, alpha, beta}, the constructor order of
 * src/EoS.h:446 (granite: 2680, 16e6, 3.5e6, 18e6, 18e9, 18e9, 0.5, 1.3, 5, 5).
 * Such a material gives P and c_s (properties 3, 4) and NaN for T and S;
 * sphgpu_eos_internal_energy refuses it. */
int32_t sphgpu_eos_set_tillotson(sphgpu_ctx *ctx, int32_t material, const double params[10]);
/* CalcPressure (src/force.h:26-32): pres/snds of n particles (host arrays). */
int32_t sphgpu_eos_pressure(sphgpu_ctx *ctx, int64_t n, const int32_t *material, const double *dens,
                            const double *eng, double *pres, double *snds);
/* CalcPressure with no allocation.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as sphgpu_eos_pressure, bit for bit.  Inputs and
 * outputs are staged through one fixed device array of SPHGPU_NOALLOC_MAX_N
 * particles (36 MiB), which each GPU the library runs on holds until the
 * process ends; n above SPHGPU_NOALLOC_MAX_N returns SPHGPU_ERR_CAPACITY.  All
 * calls on one GPU share that array, so calls that overlap on one GPU (from
 * other host threads or other contexts on the same device) race.  The EoS
 * tables must be loaded beforehand (loading a table allocates). */
#define SPHGPU_NOALLOC_MAX_N 1048576
int32_t sphgpu_eos_pressure_noalloc(sphgpu_ctx *ctx, int64_t n, const int32_t *material, const double *dens,
                                    const double *eng, double *pres, double *snds);
/* EoS_t::{Temperature,Pressure,SoundSpeed,Entropy} as one property (2..5). */
int32_t sphgpu_eos_property(sphgpu_ctx *ctx, int64_t n, const int32_t *material, const double *dens,
                            const double *eng, int32_t prop, double *out);
/* EoS_t::InternalEnergy (src/EoS.h:402-405): u(rho, S) by the restricted
 * bilinear scheme of src/interpolation.h:107-335 with grid_length 120.
 * abs_int selects the `int abs(int)` overload for the nearest-S search (1)
 * or the double one (0).  Index paths that the reference leaves undefined
 * give NaN. */
int32_t sphgpu_eos_internal_energy(sphgpu_ctx *ctx, int64_t n, const int32_t *material, const double *dens,
                                   const double *ent, int32_t abs_int, double *eng);

/* ---- device-resident simulation (src/main.cpp:69-173) ---- */
#define SPHGPU_MAX_GROUP 1024
typedef struct {
    int32_t n_leaf;          /* tree leaf limit (1..64), default 16 */
    int32_t n_group;         /* i-group limit (1..SPHGPU_MAX_GROUP), default 64; a group of more than
                                64 particles runs as one work item per wave of 64, sharing its list */
    double theta;            /* gravity opening angle, default 0.5 */
    double damping;          /* dampMotion factor (src/GI.h:32), default 1 */
    double ext_force_until;  /* addExternalForce window (src/GI.h:459), default 5000 s */
    int32_t nsys;            /* independent systems (impacts), default 1 */
    int32_t list_cap_sph;    /* per-group j-list capacity, default 4096 */
    int32_t list_cap_epj;    /* per-group gravity particle-list capacity, default 8192 */
    int32_t list_cap_spj;    /* per-group gravity node-list capacity, default 8192 */
    int32_t eos_density_fp64; /* precision form only: run the last density pass, whose
                                 density feeds the EoS, in FP64 (1, default) or FP32 (0).
                                 The EoS tables amplify density errors by
                                 d ln P / d ln rho (20 and more near the cold curve). */
    int32_t list_cap_front;  /* per-workgroup tree-walk frontier capacity in global memory,
                                default 4096 (grown on overflow like the list capacities) */
    int32_t list_front_lds;  /* frontier entries a tree walk keeps in LDS before it spills
                                to global memory (1..512), default 512 */
    int32_t persistent;      /* bitmask of interaction kernels run as one persistent launch of
                                the workgroups the GPU holds at once, taking the groups of
                                all systems from per-XCD counters: 1 density, 2 derivative,
                                4 hydro, 8 gravity; 0 (default): one workgroup per group.
                                Results are identical either way. */
    int32_t graph;           /* 1: each step runs as the replay of a recorded HIP graph (one per
                                parity of the particle buffer sets, recorded after a direct step
                                with the same capacities, recorded again whenever a capacity,
                                the precision or an allocation changes); 0 (default): launched
                                step by step.  A replayed step whose force evaluation overflows
                                a capacity leaves the state untouched and is finished directly.
                                Results are identical either way.  Ignored while kernel
                                statistics are on. */
    int32_t walk_waves;      /* persistent waves of each tree walk that builds interaction lists
                                (rounded up to a multiple of 8, at most 8192); 0 (default): 4096.
                                Walks beside force kernels share the GPU with them; fewer
                                waves leave the kernels more of it.  Results are identical. */
    int32_t vsig_split;      /* bitmask of the precision forms (1 FP64, 2 the precision form)
                                whose hydro pass forms dt -- v_sig_max over each list -- in a
                                kernel of its own (FP64 in both forms), run on the second
                                stream beside the gravity kernel, the hydro kernel forming the
                                force terms only; default 1.  The FP64 form's results are
                                identical either way; the precision form's dt becomes the
                                FP64 dt of its state. */
    int32_t hydro_compact;   /* 1 (default): where dt runs in its own kernel (vsig_split) and
                                groups hold at most 64 particles, the hydro force terms read
                                each symmetric list reduced, in order, to the entries within
                                max(max_i 2.5 h_i, 2.5 h_j) of some particle of the group (every
                                pair inside either support; dt still runs over the whole list,
                                after the compaction on the second stream, beside the gravity
                                and hydro kernels); 0: the whole lists.  The force sums take
                                the same terms in another order (results move in the last
                                bits). */
} sphgpu_sim_params;

void sphgpu_sim_default_params(sphgpu_sim_params *p);

/* Uploads n particles and runs the initial force evaluation of
 * src/main.cpp:69-104.  pos/vel are n x 3 row-major.  system may be NULL
 * (one system); otherwise it must be non-decreasing.  id may be NULL. */
int32_t sphgpu_sim_init(sphgpu_ctx *ctx, const sphgpu_sim_params *prm, int64_t n, const double *pos,
                        const double *vel, const double *mass, const double *dens, const double *eng,
                        const int32_t *material, const int32_t *system, const int64_t *id);
/* Advances nsteps steps (src/main.cpp:125-173). */
int32_t sphgpu_sim_step(sphgpu_ctx *ctx, int32_t nsteps);
int32_t sphgpu_sim_size(sphgpu_ctx *ctx, int64_t *n);

/* Fields for sphgpu_sim_get (particles in the library's current order, which
 * is the (system, Morton key) order of the last tree build; SPHGPU_F_ID maps
 * them back). */
enum {
    SPHGPU_F_POS = 0, SPHGPU_F_VEL, SPHGPU_F_ACC, SPHGPU_F_GRAV, SPHGPU_F_POT, SPHGPU_F_MASS,
    SPHGPU_F_SMTH, SPHGPU_F_DENS, SPHGPU_F_ENG, SPHGPU_F_PRES, SPHGPU_F_SNDS, SPHGPU_F_DIVV,
    SPHGPU_F_ROTV, SPHGPU_F_GRADH, SPHGPU_F_BAL, SPHGPU_F_ENGDOT, SPHGPU_F_DT, SPHGPU_F_NGB,
    SPHGPU_F_MAT, SPHGPU_F_SYS, SPHGPU_F_ID, SPHGPU_F_VELHALF, SPHGPU_F_ENGHALF, SPHGPU_F_COUNT
};
/* Copies one field to host memory: doubles (x3 for vector fields), int32
 * for NGB/MAT/SYS, int64 for ID. */
int32_t sphgpu_sim_get(sphgpu_ctx *ctx, int32_t field, void *out);
/* dt and simulated time of each system. */
int32_t sphgpu_sim_get_time(sphgpu_ctx *ctx, double *dt_sys, double *time_sys);
/* Wall time in ms of each stage of the last step (hipEvents on the main stream):
 * [0] kick/drift, [1] sort+tree, [2] density x4 (gather lists + kernels; the
 * gravity lists are built beside them on a second stream), [3] EoS,
 * [4] derivative, [5] gravity (the hydro pass's lists are built beside it),
 * [6] hydro, [7] final kick + dt; n <= 8.  With kernel statistics on, the
 * second stream's work runs in line: the gravity lists in [2], the hydro
 * lists in [5]. */
int32_t sphgpu_sim_stage_ms(sphgpu_ctx *ctx, double *ms, int32_t n);

/* Gather-list constructions and force evaluations since sphgpu_sim_init.  A
 * force evaluation runs 4 density passes and one derivative pass; they share
 * one gather list (built with an 8 % margin on h) for as long as it covers every
 * particle's support, so gather_builds / force_evals is 1 when h settles. */
int32_t sphgpu_sim_list_builds(sphgpu_ctx *ctx, int64_t *gather_builds, int64_t *force_evals);

/* Sets the per-group interaction-list capacities of the next force
 * evaluations (all >= 1): gather/symmetric j-list, gravity particle list,
 * gravity monopole list, as sphgpu_sim_params.list_cap_*.  For callers that
 * bound memory: an evaluation that needs more runs again with the capacity
 * doubled (a recorded step leaves the state untouched first), so results do
 * not depend on the capacities. */
int32_t sphgpu_sim_set_capacities(sphgpu_ctx *ctx, int32_t list_cap_sph, int32_t list_cap_epj,
                                  int32_t list_cap_spj);

/* Steps run as replays of a recorded graph and graphs recorded since
 * sphgpu_sim_init (sphgpu_sim_params.graph). */
int32_t sphgpu_sim_graph_stats(sphgpu_ctx *ctx, int64_t *replayed_steps, int64_t *captures);

/* Per-kernel statistics of the interaction kernels over the steps run since
 * statistics were switched on: summed kernel time (hipEvents on the
 * context's stream, one pair per launch), launches, and the interaction-list
 * sizes those launches processed: n_i = i-particles, n_j = j-list entries
 * (particles), n_s = monopole entries, pairs = sum over groups of
 * n_i * (n_j + n_s).  Index 0 density, 1 derivative, 2 hydro (force
 * terms), 3 gravity, 4 the density passes the precision form runs in FP64
 * (eos_density_fp64), 5 the hydro pass's dt (v_sig_max over each list, a
 * kernel of its own in a simulation step); n may be 4 to 6. */
typedef struct {
    double ms;
    int64_t launches, n_i, n_j, n_s, pairs;
} sphgpu_kernel_stat;
int32_t sphgpu_sim_set_stats(sphgpu_ctx *ctx, int32_t on);
int32_t sphgpu_sim_kernel_stats(sphgpu_ctx *ctx, sphgpu_kernel_stat *out, int32_t n);

/* Interaction-list export for the current particle state (sorts and builds
 * the tree first).  kind: 0 gather (density/derivative), 1 symmetric
 * (hydro), 2 gravity, 3 the gather lists as a force evaluation uses them:
 * built with the 8 % margin on h and compacted to the entries within R_g of
 * some particle of the group, R_g = 1.08 max_i 2.5 h_i, 4 the symmetric
 * lists as the hydro force terms read them with hydro_compact: reduced, in
 * order, to the entries within max(max_i 2.5 h_i, 2.5 h_j) of some particle
 * of the group (groups of at most 64 particles).  First call with NULL
 * arrays to get the sizes. */
int32_t sphgpu_sim_lists(sphgpu_ctx *ctx, int32_t kind, int64_t *n_groups, int64_t *n_j, int64_t *n_s,
                         int32_t *group_start, int64_t *j_off, int32_t *j_idx, int64_t *s_off, int32_t *s_idx);
/* Tree-node moments after the last tree build: per node mass, com[3],
 * box lo[3], box hi[3] (10 doubles) and the node count. */
int32_t sphgpu_sim_nodes(sphgpu_ctx *ctx, int64_t *n_nodes, double *moments);

#ifdef __cplusplus
}
#endif
#endif
