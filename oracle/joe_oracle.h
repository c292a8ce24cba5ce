/* CPU restatement of Joe's steady-solve hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load this library, and only as the checker or the timed CPU baseline. The
 * product (libjoemi) never links or calls it.
 *
 * Layout and numbering are the reference's own (Ugp/UgpWithCv members, AoS):
 * cvofa[nfa][2], faocv/nbocv CSR, x_fa/fa_normal[nfa][3], x_cv[ncv][3],
 * nbocv_lsg[nnz][3], bfa_lsg[nfa_b][3]; rhou[ncv][3]; A[nnz][5][5].
 *
 * Every routine follows the reference file:line given at its definition in
 * joe_oracle.cpp and is pinned against the reference build (oracle/_ref) and
 * the committed golden vectors under tests/golden/.
 */
#ifndef JOE_ORACLE_H
#define JOE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* boundary-condition keywords of Joe.in (JoeWithModels.cpp:3857-4300) */
enum {
  OR_BC_CBC = 1,
  OR_BC_CBC_SUBSONIC_INLET = 2,
  OR_BC_CBC_SUBSONIC_OUTLET = 3,
  OR_BC_SYMMETRY = 4,
  OR_BC_NEUMANN = 5,
  OR_BC_WALL = 6,
  OR_BC_FARFIELD = 7,
  OR_BC_PROFILES = 8
};
enum { OR_HLLC = 0, OR_ROE = 1, OR_AUSM = 2 };
enum { OR_GRAD_GG = 0, OR_GRAD_LS = 1 };
enum { OR_SYMM_ORIGINAL_NUMERIC = 0, OR_SYMM_NEW_NUMERIC = 1, OR_SYMM_ANALYTICAL = 2 };

typedef struct {
  int ncv, nfa, nfa_b, nnz;
  const int *cvofa, *faocv_i, *faocv_v, *nbocv_i, *nbocv_v;
  const double *x_fa, *fa_normal, *x_cv, *cv_volume, *nbocv_lsg, *bfa_lsg;
  int nzone;                 /* boundary zones, in faZoneList order */
  const int *zone_bc;        /* OR_BC_* */
  const int *zone_ifa;       /* [nzone][2] first, last face */
  const double *zone_par;    /* [nzone][8] keyword values (see or_ctx_create) */
  const double *zone_scal;   /* [nzone][2] first scalar's boundary (nuSA, or kine with SST): kind (0 none,
                              1 value, 2 HOOK, 3 PROFILES), value; NULL without a turbulence model */
  const double *zone_scal2;  /* [nzone][2] the same for omega (SST); NULL otherwise */
  /* the boundary-profile file of PROFILES zones (JoeWithModels.cpp:262-290):
   * nprof rows of nval_prof values (y, rho, u, v, p, scalars); NULL without */
  const double *profiles;
  int nprof, nval_prof;
} or_mesh;

typedef struct {
  double gamma, R_gas, mu_ref;
  double gravity[3];
  int scheme, second_order, grad_scheme, limiter, symm_bc;
  double eps_limiter;        /* GRAD_RECONSTRUCTION EPS */
  double cfl;
  int time_step_min;         /* 1: CFL_MIN (global dt) */
  double under_relax;
  int ls_maxiter;
  double ls_abs, ls_rel;
  int cfl_after, cfl_interval;
  double cfl_factor, cfl_max;
  int nsteps;
  double resid_energ_th;
  /* viscous terms (MU_REF > 0): MU_MODE 0 POWERLAW / 1 SUTHERLAND,
   * MU_POWER_LAW, SL_TREF, SL_SREF, T_REF, PR, PRTURB, VISC_LIMITER */
  int visc_mode;
  double mu_power_law, sl_tref, sl_sref, T_ref, pr, pr_turb;
  int visc_limiter;
  /* Spalart-Allmaras (RansTurbSA, turbModels/TurbModel_SA.h) as one
   * segregated scalar nuSA: initial value (INITIAL_CONDITION_TURB) and
   * UNDER_RELAXATION_SCALARS */
  int sa;
  double sa_init, relax_scal;
  /* TIME_STEP: 0 MAX_BASED (default), 1 SUM_BASED */
  int time_step_sum;
  /* TIME_INTEGRATION = BACKWARD_EULER_COUPLED: calcRhsCoupled's fluxes and
   * Jacobian (nScal = 0, MU_REF = 0, HLLC); the loop is runBackwardEuler's
   * with the block-Jacobi BiCGStab in place of PETSc GMRES */
  int coupled;
  /* Menter SST k-omega (RansTurbKOmSST, turbModels/TurbModel_KOMSST.h) as two
   * segregated scalars kine, omega (registration order): initial values
   * (INITIAL_CONDITION_TURB), SST_LIMIT_PK, SST_LIMITER_SHEARSTRESS; the model
   * constants are the reference's defaults (:16-23) */
  int sst;
  double sst_init[2];
  int sst_limit_pk, sst_limiter_shear;
} or_params;

typedef struct or_ctx or_ctx;

or_ctx *or_ctx_create(const or_mesh *m, const or_params *p);
void or_ctx_destroy(or_ctx *c);
/* state in reference layout; rhou is [ncv][3] */
void or_set_state(or_ctx *c, const double *rho, const double *rhou, const double *rhoE);
void or_get_state(const or_ctx *c, double *rho, double *rhou, double *rhoE);
/* calcStateVariables + setBC + calcDt + calcRhs (implicit: also A). */
void or_residual(or_ctx *c, int implicit);
/* the transported scalars, reference numbering: nuSA [ncv] (SA) or kine, omega [2][ncv] (SST) */
void or_set_scalar(or_ctx *c, const double *phi);
/* one backward-Euler step (JoeWithModels.cpp:975-1158); returns BiCGStab iterations */
int or_step(or_ctx *c);
/* the step count of a solve resumed from a restart file (its STEP) */
void or_set_step(or_ctx *c, int step);
/* runBackwardEuler from the current state; returns steps taken. */
int or_run(or_ctx *c, int max_steps, double *resid_hist /* [max_steps][5 + nScal] or NULL */);
/* named arrays: RHS_RHO RHS_RHOU RHS_RHOE A LOCAL_DT GRAD_RHO GRAD_P GRAD_U
 * ALPHA_RHO VEL PRESS TEMP ENTHALPY SOS RHO_BFA VEL_BFA P_BFA T_BFA H_BFA DQ
 * RESID (5 + nScal) MUL_FA GRAD_ENTHALPY; with SA also NUSA NUSA_BFA
 * GRAD_NUSA MUT_FA WALL_DIST STR_MAG VORT_MAG DIFF_NUSA RHS_NUSA ASCAL
 * ([6][nnz]) DSCAL; with SST KINE OMEGA KINE_BFA OMEGA_BFA GRAD_KINE GRAD_OMEGA
 * DIFF_KINE DIFF_OMEGA RHS_KINE RHS_OMEGA MUT_FA WALL_DIST STR_MAG VORT_MAG DIVERG
 * CROSS_DIFF F1 F2 ASCAL ([2][6][nnz]) DSCAL ([2][ncv]); returns element count or -1 */
long long or_get(const or_ctx *c, const char *name, double *out);
/* current CFL and step counter */
double or_cfl(const or_ctx *c);
int or_nsteps_done(const or_ctx *c);

/* writeWallValues (joe.cpp:43-130) on a boundary zone: out [n][16] = x_fa (3),
 * rho, press, temp, tau_wall, qdot, yplus, mulam, pressure force (3), shear
 * force (3); returns n (out may be NULL) or -1 */
int or_wall_values(or_ctx *c, int zone, double *out);

#ifdef __cplusplus
}
#endif
#endif
