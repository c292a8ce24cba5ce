/* joemi -- MI355X-native steady-solve hot path of the Joe compressible
 * finite-volume solver, behind a plain C ABI.
 *
 * The reference (btracey/myjoe, "MUM/Joe") exposes this path as C++
 * inheritance: a case class derived from JoeWithModels overrides hook virtuals
 * and calls run(), which drives runBackwardEuler / calcRhs /
 * solveCoupledLinSysNS on the Ugp mesh. Each entry point below names the
 * reference interface it stands in for. All pointers are host pointers unless
 * the name says "device"; arrays use the reference's own layouts and
 * numbering (Ugp/UgpWithCv members): rhou[ncv][3], A[nbocv_s][5][5], etc.
 *
 * Errors never cross the ABI as exceptions: every call returns a jm_status
 * and jm_last_error() describes the last failure of the calling thread.
 */
#ifndef JOEMI_H
#define JOEMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  JM_OK = 0,
  JM_ERR_ARG = 1,          /* bad argument / size */
  JM_ERR_MESH = 2,         /* mesh cannot be built (connectivity, geometry) */
  JM_ERR_PARAM = 3,        /* Joe.in parameter missing or not understood */
  JM_ERR_HIP = 4,          /* HIP runtime failure, including no GPU */
  JM_ERR_UNSUPPORTED = 5,  /* option of the reference not provided here */
  JM_ERR_STATE = 6         /* call out of order (e.g. no state set) */
} jm_status;

enum { JM_FP64 = 0, JM_MIXED = 1 };            /* Jacobian/Krylov precision */
enum { JM_RHS_IMPLICIT = 1 };                   /* jm_calc_rhs flags */

typedef struct jm_mesh jm_mesh;
typedef struct jm_ctx jm_ctx;

/* text of the last error on this thread ("" if none) */
const char *jm_last_error(void);
const char *jm_version(void);

/* ---------------------------------------------------------------------------
 * Mesh. Replaces UgpWithCv::initializeFromRestartFile for .msh input
 * (UgpWithCv.h:316-347): MshFilter::initUgpMin (MshFilter.cpp:44-90),
 * Ugp::init -> redistReconnectReorder (Ugp.cpp:1341-4277, serial numbering),
 * UgpWithCv::init (buildNbocv, calcGeometry; UgpWithCv.cpp:768-1095) and
 * calcLsgCoeff (UgpWithCv.cpp:1783-1913). Host-side; the result is uploaded by
 * jm_ctx_create.
 * ------------------------------------------------------------------------- */

/* Faces in mesh-file order, oriented as after the reader's flip rule
 * (MshFilter.cpp:453-472): node order gives a normal from face_cells[2f] to
 * face_cells[2f+1]; boundary faces have face_cells[2f+1] = -1. */
jm_status jm_mesh_from_faces(int nno, const double *x_no, int nfa, const int64_t *face_ptr, const int *face_nodes,
                             const int *face_cells, const int *face_zone, int ncv, int nzone, const int *zone_ids,
                             const char *const *zone_names, jm_mesh **out);
/* Fluent ASCII .msh, as MshFilter reads it */
jm_status jm_mesh_read_msh(const char *path, jm_mesh **out);

/* A Ugp restart file (restart.out; replaces UgpWithCv::initializeFromRestartFile
 * for restart input: Ugp::readRestart, Ugp.cpp:12473-13240, then the same
 * init as .msh input). Serial files only (one partition, no periodic zones).
 * The file's STEP, TIME and cell fields (RHO, RHOU, RHOE, nuSA, ...) are kept
 * with the mesh, in its numbering: jm_mesh_restart_field copies one (count =
 * ncv for a scalar, 3 ncv for a vector, [ncv][3]; out may be NULL to ask for
 * the count) or fails with JM_ERR_ARG if the file had none of that name. */
jm_status jm_mesh_read_restart(const char *path, jm_mesh **out);
jm_status jm_mesh_restart_field(const jm_mesh *m, const char *name, double *out, int64_t *count);
jm_status jm_mesh_restart_step(const jm_mesh *m, int *step, double *time);
/* Ugp::writeRestart (Ugp.cpp:9902-10660) of a mesh and cell fields in its
 * numbering (dims[k] = 1: ncv values, 3: [ncv][3]), in the serial layout the
 * reference writes and reads: the reference's initializeFromRestartFile takes
 * the file as its own. */
jm_status jm_write_restart(const jm_mesh *m, const char *path, int nfield, const char *const *names,
                           const double *const *data, const int *dims, int step, double time);

/* A host that already holds the Ugp arrays (the reference's own objects)
 * hands them over without renumbering: UgpWithCv members cvofa, faocv_i/v,
 * nbocv_i/v, x_fa, fa_normal, x_cv, cv_volume, nbocv_lsg_coeff,
 * boundary_fa_lsg_coeff (UgpWithCv.h:201-207) and faZoneList (Ugp.h:132-310)
 * in list order. */
typedef struct {
  int ncv, nfa, nfa_b;
  const int *cvofa, *faocv_i, *faocv_v, *nbocv_i, *nbocv_v;
  const double *x_fa, *fa_normal, *x_cv, *cv_volume, *nbocv_lsg, *bfa_lsg;
  int nzone;
  const char *const *zone_names;
  const int *zone_kind; /* 1 boundary, 6 internal (Ugp.h:41-55) */
  const int *zone_ifa_f, *zone_ifa_l;
} jm_ugp_view;
jm_status jm_mesh_from_ugp(const jm_ugp_view *v, jm_mesh **out);

void jm_mesh_destroy(jm_mesh *m);
/* counts[0..7] = nno, nfa, nfa_b, ncv, nnz (nbocv_s), nzone, faocv_s, noofa_s */
jm_status jm_mesh_counts(const jm_mesh *m, int64_t counts[8]);
/* copy a mesh array out by its reference name: X_NO NOOFA_I NOOFA_V CVOFA
 * FAOCV_I FAOCV_V NBOCV_I NBOCV_V (int32) X_FA FA_NORMAL X_CV CV_VOLUME
 * NBOCV_LSG BFA_LSG (float64). out may be NULL to query *count. */
jm_status jm_mesh_array(const jm_mesh *m, const char *name, void *out, int64_t *count);
/* zone i in faZoneList order */
jm_status jm_mesh_zone(const jm_mesh *m, int i, char *name, int name_len, int *kind, int *ifa_f, int *ifa_l);

/* ---------------------------------------------------------------------------
 * Solve context: the per-solve state of UgpWithCvCompFlow / JoeWithModels,
 * resident on the current HIP device. Re-entrant: no process-wide statics
 * (the reference keeps function statics in calcDt, setBC and the BiCGStab,
 * UgpWithCv.cpp:2664-2673). joe_in is the text of a Joe.in file
 * (UgpWithCvCompFlow::init, UgpWithCvCompFlow.h:57-502). stream may be NULL
 * (the default stream) or a hipStream_t.
 * ------------------------------------------------------------------------- */
jm_status jm_ctx_create(const jm_mesh *m, const char *joe_in, int precision, void *stream, jm_ctx **out);
void jm_ctx_destroy(jm_ctx *c);

/* conservative state in the reference layout (rho[ncv], rhou[ncv][3],
 * rhoE[ncv]); stands in for writing the registered RHO/RHOU/RHOE arrays.
 * Setting the state starts a new solve: the step counter and CFL restart. */
jm_status jm_set_state(jm_ctx *c, const double *rho, const double *rhou, const double *rhoE);
jm_status jm_get_state(jm_ctx *c, double *rho, double *rhou, double *rhoE);
/* After jm_set_state: the step count a solve resumed from a restart file
 * continues from (the STEP value read with it; runBackwardEuler keeps it,
 * JoeWithModels.cpp:977, unless RESET_STEP, :12-17). The CFL still starts from
 * Joe.in's, as in the reference. */
jm_status jm_set_step(jm_ctx *c, int step);
/* JoeWithModels::initialHook uniform initial condition (JoeWithModels.cpp:156-186);
 * with TURB_MODEL = SA also nuSA = INITIAL_CONDITION_TURB (:219-240) */
jm_status jm_init_uniform(jm_ctx *c);
/* a transported scalar's field in reference numbering, [ncv]: stands in for
 * writing ScalarTranspEq::phi (UgpWithCv.h:50-113) of "nuSA", the one scalar
 * of the Spalart-Allmaras mix-in (RansTurbSA, TurbModel_SA.h:31), or of
 * "kine" / "omega", the two of the SST mix-in (RansTurbKOmSST,
 * TurbModel_KOMSST.h:36-52). The context must have been created with
 * TURB_MODEL = SA or SST. */
jm_status jm_set_scalar(jm_ctx *c, const char *name, const double *phi);

/* calcStateVariables + setBC + calcDt + calcRhs (JoeWithModels.cpp:943-1005,
 * 2564-3589; calcRhsCoupled, :4596-4611, with TIME_INTEGRATION =
 * BACKWARD_EULER_COUPLED) on the resident state; with JM_RHS_IMPLICIT the
 * block Jacobian A[nbocv_s][b][b] is assembled too (b = 5; 6 in the coupled
 * loop with TURB_MODEL = SA, the scalar's row and column in every block).
 * Asynchronous on the context's stream. */
jm_status jm_calc_rhs(jm_ctx *c, int flags);

/* copy a result out (synchronises the stream): RHS_RHO RHS_RHOU RHS_RHOE A
 * LOCAL_DT GRAD_RHO GRAD_P GRAD_U ALPHA_RHO VEL PRESS RHO_BFA VEL_BFA P_BFA
 * T_BFA H_BFA DQ RESID; with MU_REF > 0 GRAD_ENTHALPY; with TURB_MODEL = SA
 * NUSA NUSA_BFA GRAD_NUSA WALL_DIST STR_MAG VORT_MAG RHS_NUSA ASCAL
 * ([6][nbocv_s], runBackwardEuler's AScal[0]) DSCAL, and RESID has 6 entries
 * (in the coupled loop A is [nbocv_s][6][6] and DQ [ncv][6], and ASCAL and
 * DSCAL do not exist); with TURB_MODEL = SST (TurbModel_KOMSST.h) KINE OMEGA
 * KINE_BFA OMEGA_BFA GRAD_KINE GRAD_OMEGA WALL_DIST STR_MAG VORT_MAG DIVERG
 * CROSS_DIFF F1 F2 RHS_KINE RHS_OMEGA ASCAL ([2][6][nbocv_s]) DSCAL
 * ([2][ncv]), and RESID has 7 entries. Evaluated from the current state when asked (each
 * runs the residual once more): FLUX_FA, each face's inviscid flux [nfa][5]
 * (the terms calcRhs accumulates, JoeWithModels.cpp:2786-3105); MUL_FA and
 * MUT_FA, each face's laminar and eddy viscosity [nfa] (mul_fa of
 * calcMaterialProperties, UgpWithCvCompFlow.h:1701-1753; mut_fa of
 * calcRansTurbViscMuet, TurbModel_SA.h:74-128, TurbModel_KOMSST.h:100-158), zero when the case is
 * inviscid or has no turbulence model. out may be NULL to query *count. */
jm_status jm_get(jm_ctx *c, const char *name, double *out, int64_t *count);

/* One pass of the runBackwardEuler loop (JoeWithModels.cpp:975-1158): CFL
 * ramp, local time step, implicit residual, under-relaxed block system with
 * V/dt on the diagonal, block-Jacobi BiCGStab (UgpWithCv.cpp:2660-2876),
 * update, state and BC refresh, L1 residual. */
typedef struct {
  int step;            /* step counter after this call */
  int lin_iters;       /* BiCGStab iterations */
  int lin_status;      /* 1 converged, 2 hit MAX_ITER (reference "done" codes) */
  double cfl;          /* CFL used */
  double dt_min;       /* minimum local time step */
  double resid[5];     /* L1 residuals of rho, rhou(3), rhoE */
  double resid_scal;   /* L1 residual of nuSA (TURB_MODEL = SA; else 0): over UNDER_RELAXATION
                          in the segregated loop (JoeWithModels.cpp:1132-1134), as it is in the
                          coupled loop (:4443-4453) */
  int scal_iters;      /* its Jacobi BiCGStab iterations (solveCvScalarBcgstab; 0 in the coupled
                          loop, where nuSA is part of the block system) */
  double resid_l2[5];  /* L2 norms (square root of the sum of squares) of the same residuals;
                          the reference reports L1 only */
  double resid_scal2;  /* with TURB_MODEL = SST omega's L1 residual (resid_scal is kine's,
                          scal_iters the two solves' iterations); else 0 */
} jm_step_info;
jm_status jm_step(jm_ctx *c, jm_step_info *info);
/* The same step driven by the device: CFL ramp and step counter on the
 * device, the step captured on first use into three hipGraphs (residual,
 * Jacobian and the first BiCGStab iterations; further iterations in chunks;
 * update and refresh) that are replayed, the host choosing chunk or tail
 * from a flag the device writes to mapped memory. Results are bit-identical
 * to jm_step's. */
jm_status jm_step_graph(jm_ctx *c, jm_step_info *info);

/* The same step split where the reference calls its hooks, for a host whose
 * JoeWithModels subclass overrides sourceHook or boundaryHook:
 *   jm_step_begin  CFL ramp, calcDt, calcRhs with the Jacobian
 *                  (JoeWithModels.cpp:975-1005); with JM_HOST_SOURCE the
 *                  default sourceHook (gravity, JoeWithModels.h:111-129) is
 *                  not applied on the device because the host's replaces it;
 *   jm_add_source  the host sourceHook's terms, reference numbering:
 *                  rhs_rho[ncv], rhs_rhou[ncv][3], rhs_rhoE[ncv] added to the
 *                  residual and a_diag[ncv][5][5] (the A[nbocv_i[icv]] blocks)
 *                  added to the diagonal blocks; any pointer may be NULL;
 *   jm_step_end    A_diag += V/dt, relaxation, BiCGStab, update, state and
 *                  setBC refresh, L1 residual (:1011-1136).
 * After new HOOK boundary values (jm_set_boundary_state), jm_update_bc
 * re-runs calcStateVariables + setBC (:3857-4300) on the resident state. */
enum { JM_HOST_SOURCE = 1 };                    /* jm_step_begin flags */
jm_status jm_step_begin(jm_ctx *c, int flags);
jm_status jm_add_source(jm_ctx *c, const double *rhs_rho, const double *rhs_rhou, const double *rhs_rhoE,
                        const double *a_diag);
jm_status jm_step_end(jm_ctx *c, jm_step_info *info);
jm_status jm_update_bc(jm_ctx *c);
/* sourceHookCoupled (JoeWithModels.h:134-152) between jm_step_begin and
 * jm_step_end of a TIME_INTEGRATION = BACKWARD_EULER_COUPLED context: rhs
 * [ncv][b] added to the residual rows (rho, rhou(3), rhoE, then the
 * scalars) and a_diag [ncv][b][b] to the diagonal blocks A[nbocv_i[icv]],
 * reference numbering; b = 5 + nScal must be the system's block size (5, or
 * 6 with TURB_MODEL = SA); either pointer may be NULL. The host's terms are
 * added after the model's source (sourceHookRansTurbCoupled), where the
 * reference adds them before it (JoeWithModels.cpp:4607-4609): a scalar row
 * term may differ in its last bit. */
jm_status jm_add_source_coupled(jm_ctx *c, int b, const double *rhs, const double *a_diag);

/* JoeWithModels::run's steady loop until doneSolver (JoeWithModels.h:42-95):
 * NSTEPS, RESID_ENERG_TH or NaN. resid_hist may be NULL, else [max_steps][5]. */
jm_status jm_run(jm_ctx *c, int max_steps, double *resid_hist, int *steps_done);

/* boundaryHook bridge (JoeWithModels.h:105): values for a zone whose Joe.in
 * keyword is HOOK, per face of the zone (T[n], vel[n][3], p[n]). */
jm_status jm_set_boundary_state(jm_ctx *c, int zone, const double *T, const double *vel, const double *p);
/* the read side of the same bridge: the boundary values setBC last computed
 * on the zone's faces (T_bfa[n], vel_bfa[n][3], p_bfa[n]; JoeWithModels.cpp
 * :3857-4300), which the reference hands to boundaryHook (:3872-3881) before
 * the hook overwrites them. Any pointer may be NULL. Needs a state. */
jm_status jm_get_boundary_state(jm_ctx *c, int zone, double *T, double *vel, double *p);

/* writeWallValues of the reference's case class (joe.cpp:43-130) on boundary
 * zone `zone` at the resident state: per face 10 columns as the reference
 * writes them (x_fa[3], rho, press, temp of the face's cell, tau_wall =
 * mulam |u_t| / walld sign(u_x), qdot with the reference's 300 K wall, yplus,
 * mulam of the face), then 6 more: the face's pressure force on the wall
 * p_bfa n A[3] and its shear force mulam u_t / walld A[3] (their sums over the
 * zone give the wall's force, hence lift and drag). out [n][16] may be NULL
 * to query *count = n. Needs MU_REF > 0. */
jm_status jm_wall_values(jm_ctx *c, int zone, double *out, int64_t *count);

/* device views for callers that keep data on the GPU: the conservative state
 * as SoA double[5][ncv] (rho, rhou_x, rhou_y, rhou_z, rhoE) and the residual
 * in the same layout. */
jm_status jm_device_state(jm_ctx *c, double **q_dev);
jm_status jm_device_rhs(jm_ctx *c, double **rhs_dev);
/* wait for all work queued on the context's stream */
jm_status jm_sync(jm_ctx *c);

/* solveCvVectorR5Bcgstab (UgpWithCv.cpp:2660-2876: block-Jacobi BiCGStab,
 * unpivoted LU of each diagonal block, stop when max |LU^-1 r| <= zero_abs,
 * or <= zero_rel times its first value, or after max_iter iterations) on a
 * 5x5 block system the caller holds in DEVICE memory in the reference's
 * layout: nbocv_i[ncv+1], nbocv_v[nnz] (int32, the diagonal first in each
 * row), A[nnz][5][5], rhs[ncv][5], phi[ncv][5] (the initial guess in, the
 * solution out). The reference's callers (JoeWithModels::runBackwardEuler via
 * solveCoupledLinSysNS, JoeWithModels.cpp:1026) pass dq = 0 as the guess.
 * Runs on stream (NULL: the default stream) and returns when done; the work
 * space (block LU, eight vectors, partials) is allocated for the call.
 * iters / status (1 converged, 2 hit max_iter) may be NULL. */
jm_status jm_block_bcgstab(int ncv, const int *nbocv_i, const int *nbocv_v, const double *A, const double *rhs,
                           double *phi, double zero_abs, double zero_rel, int max_iter, void *stream, int *iters,
                           int *status);

/* No counterpart in the reference: jm_block_bcgstab for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, kernels and results; the work space is a
 * static array of 64 MiB that the library holds on each device it runs on
 * until the process ends, which bounds ncv to
 * jm_block_bcgstab_noalloc_max_cells() (about 129 000 cells); a larger ncv
 * fails with JM_ERR_ARG before any device call. No hipMalloc or hipFree is
 * made. Every call on a device shares that device's array, so calls that
 * overlap on one device (two host threads, or two streams) race. */
jm_status jm_block_bcgstab_noalloc(int ncv, const int *nbocv_i, const int *nbocv_v, const double *A,
                                   const double *rhs, double *phi, double zero_abs, double zero_rel, int max_iter,
                                   void *stream, int *iters, int *status);
int64_t jm_block_bcgstab_noalloc_max_cells(void);

/* Measurement hook, no counterpart in the reference: launches each kernel
 * of one residual evaluation (state, bc, grad, flux, dt) reps times on the
 * context's stream and returns the mean duration of each in ms, measured
 * with HIP events on that stream. ms_out has room for 8 values; names gets
 * a comma-separated list of the kernels timed. */
jm_status jm_time_residual_kernels(jm_ctx *c, int reps, double *ms_out, char *names, int names_len);

/* Measurement hook, no counterpart in the reference: the parts of one
 * implicit step on the context's stream, reps times each, mean ms into
 * ms_out[0..3]: the residual (calcStateVariables, setBC, calcDt, calcRhs or
 * calcRhsCoupled), the Jacobian assembly with V/dt and the relaxed
 * right-hand side, the preconditioner set-up (block LU, first residual) and
 * one BiCGStab iteration, timed with its stop tests off. Leaves the
 * context's Jacobian and Krylov vectors overwritten (the state is kept). */
jm_status jm_time_step_kernels(jm_ctx *c, int reps, double *ms_out, char *names, int names_len);

/* Mesh quality of the reference's geometry checks (UgpWithCv.cpp:1000-1091):
 * minimum over internal faces of s.n/(|s||n|) and the number of sub-tets
 * with non-positive volume. */
jm_status jm_mesh_quality(const jm_mesh *m, double *min_sdotn, int *neg_subtets);

/* ---------------------------------------------------------------------------
 * Serving: many steady solves on one GPU. No counterpart in the reference
 * (one JoeWithModels::run per process); each solve follows the same path as
 * jm_run. Requests are admitted in submission order while their device
 * footprint fits mem_budget_bytes and fewer than max_active run; each active
 * solve keeps its state and Jacobian resident on its own stream and advances
 * by graph-captured steps (jm_step_graph), overlapping the others. A solve
 * stops by NSTEPS, RESID_ENERG_TH, NaN or RESID_ENERG_REL (energy L1 below
 * this fraction of the first step's; this build's own key). The mesh passed
 * to jm_server_submit must stay alive until the ticket is released. Results
 * are bit-identical to a solve run alone.
 * ------------------------------------------------------------------------- */
/* device bytes a context for this mesh holds once it has stepped (the
 * server's admission estimate) */
jm_status jm_ctx_footprint(const jm_mesh *m, int precision, int64_t *bytes);

typedef struct jm_server jm_server;
typedef struct {
  int status;         /* 0 queued, 1 running, 2 done, 3 failed */
  int steps;          /* pseudo-time steps taken */
  double resid[5];    /* L1 residuals of the last step */
  double resid0[5];   /* L1 residuals of the first step */
  double queued_s;    /* submission to admission, seconds */
  double run_s;       /* admission (context build included) to retirement; a retried
                         solve's covers both attempts */
  int fp64_retries;   /* 1 if the solve was mixed precision, went non-finite and
                         was run again in FP64 (steps and resid are then the FP64 run's) */
  int steps_abandoned; /* steps of that abandoned mixed-precision attempt */
} jm_solve_report;
/* flags: JM_SERVER_BATCHED steps the active solves in lockstep (2 groups, each
 * on its own stream; JOEMI_SERVER_GROUPS = 1..4 in the environment changes it),
 * each kernel of the step and of the BiCGStab launched once for all of a
 * group's solves; without it each solve steps on its own stream through
 * jm_step_graph's captured step. Either way a solve's result is the same bits
 * as its own jm_step loop's. Batched serving runs the per-entry flux kernel
 * (GPU_FLUX_KERNEL = ENTRY) and no SA scalar; other requests fail with
 * JM_ERR_UNSUPPORTED.
 * A mixed-precision solve whose residual turns non-finite (its FP32 Krylov
 * solve stopped at the iteration limit on a far-off iterate) is queued again
 * once, at the front, in FP64 from its initial state, when the FP64 footprint
 * fits the budget; the report's fp64_retries says so. JM_SERVER_NO_FP64_RETRY
 * retires it with its non-finite state instead. */
#define JM_SERVER_BATCHED 1
#define JM_SERVER_NO_FP64_RETRY 2
jm_status jm_server_create(int64_t mem_budget_bytes, int max_active, int flags, jm_server **out);
void jm_server_destroy(jm_server *s);
/* initial state may be NULL (uniform, as jm_init_uniform) */
jm_status jm_server_submit(jm_server *s, const jm_mesh *m, const char *joe_in, int precision, const double *rho,
                           const double *rhou, const double *rhoE, int *ticket);
jm_status jm_server_status(jm_server *s, int ticket, int *status);
/* pseudo-time steps a submitted solve has taken so far (0 while queued) */
jm_status jm_server_progress(jm_server *s, int ticket, int *steps);
jm_status jm_server_wait(jm_server *s, int ticket);
/* final state (any pointer may be NULL) and report of a finished solve */
jm_status jm_server_result(jm_server *s, int ticket, double *rho, double *rhou, double *rhoE, jm_solve_report *rep);
jm_status jm_server_release(jm_server *s, int ticket);

/* ---------------------------------------------------------------------------
 * FPVA chemistry table. Replaces ChemtableCartesianLinear_single
 * (mixing_combustion/ChemtableCartesianLinear_single.h:44-199): a structured
 * Cartesian table of FP32 values over (Zmean, Zvar, progress variable),
 * looked up by linear interpolation along each axis with FP64 weights. The
 * reference looks up one point per call (Lookup, LookupCoeff,
 * LookupSelectedCoeff, :374-550); here one call looks up a whole field of
 * points on the device, each value equal to the reference's bit for bit.
 * ------------------------------------------------------------------------- */
typedef struct jm_chemtable jm_chemtable;

/* ChemtableCartesianLinear_single::Load (:19-246): the reference's binary
 * table file (version 2.0), checked, then uploaded to the current device.
 * A file that is not such a table fails with JM_ERR_PARAM, a table with an
 * axis of fewer than two strictly increasing values with JM_ERR_ARG, both
 * before any device call. */
jm_status jm_chemtable_load(const char *path, jm_chemtable **out);
/* The same table from arrays: x1[n1], x2[n2], x3[n3], data[n1][n2][n3][nvar]
 * (the variable fastest, the file's order); names[nvar] may be NULL (then
 * "0", "1", ...). */
jm_status jm_chemtable_create(int n1, int n2, int n3, int nvar, const double *x1, const double *x2, const double *x3,
                              const float *data, const char *const *names, jm_chemtable **out);
void jm_chemtable_destroy(jm_chemtable *t);
/* GetChemtableDimension1..4 into dims[4]; GetReferencePressure, GetToxi,
 * GetTfuel into ref[3]; either may be NULL */
jm_status jm_chemtable_info(const jm_chemtable *t, int *dims, double *ref);
/* the index of a variable by name, as Lookup's tag (:394-400); an unknown
 * name fails with JM_ERR_ARG */
jm_status jm_chemtable_var(const jm_chemtable *t, const char *name, int *index);
/* Lookup / LookupCoeff / LookupSelectedCoeff over n points: device arrays
 * zmean_dev, zvar_dev, prog_dev [n] (clipped to the table as the reference
 * clips them); out_dev[k][n] (device) receives variable vars[k] (host array,
 * 1 <= nout <= 32; LookupCoeff is vars 0..9, LookupSelectedCoeff 0..4).
 * Launched on stream (NULL: the null stream) and not waited for. */
jm_status jm_chemtable_lookup(const jm_chemtable *t, int64_t n, const double *zmean_dev, const double *zvar_dev,
                              const double *prog_dev, int nout, const int *vars, double *out_dev, void *stream);
/* jm_chemtable_lookup on host arrays (zmean, zvar, prog [n]; out [nout][n]):
 * copied to the device and back, the call returns when out holds the values.
 * Serves the reference's one-point Lookup / LookupCoeff calls (n = 1) and
 * batches of cells kept on the host. */
jm_status jm_chemtable_lookup_host(const jm_chemtable *t, int64_t n, const double *zmean, const double *zvar,
                                   const double *prog, int nout, const int *vars, double *out);
/* GetChemtableCoordinate1..3 (:89-94): axis d (1, 2, 3) into x[dims[d-1]] */
jm_status jm_chemtable_axis(const jm_chemtable *t, int d, double *x);
/* GetCombustionModel (which = -2), GetChemtableType (-1), GetChemtableVariables(i)
 * (which = i >= 0): the text into buf (len bytes, NUL-terminated) */
jm_status jm_chemtable_text(const jm_chemtable *t, int which, char *buf, int len);
/* Measurement hook, no counterpart in the reference: jm_chemtable_lookup
 * launched reps times on stream after one untimed launch; *ms = the mean
 * duration of one, from HIP events recorded on that stream. */
jm_status jm_chemtable_time_lookup(const jm_chemtable *t, int64_t n, const double *zmean_dev, const double *zvar_dev,
                                   const double *prog_dev, int nout, const int *vars, double *out_dev, void *stream,
                                   int reps, double *ms);

#ifdef __cplusplus
}
#endif
#endif /* JOEMI_H */
