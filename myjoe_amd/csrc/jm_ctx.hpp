// Per-solve device context of the steady-solve hot path (host-side view).
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <vector>

#include "jm_mesh.hpp"

namespace jm {

enum BcKind : int {
  kBcCbc = 1,
  kBcCbcSubsonicInlet = 2,
  kBcCbcSubsonicOutlet = 3,
  kBcSymmetry = 4,
  kBcNeumann = 5,
  kBcWall = 6,
  kBcFarfield = 7,
  kBcHook = 8,
  kBcProfiles = 9
};
// kHllcC: the coupled HLLC of runBackwardEulerCoupled (energy from the side
// enthalpies, its own Jacobian and boundary treatments)
enum Scheme : int { kHllc = 0, kRoe = 1, kAusm = 2, kHllcC = 3 };
enum SymmBc : int { kSymmOriginal = 0, kSymmNew = 1, kSymmAnalytical = 2 };
// device cell numbering: the reference's own, face-count groups inside
// 1024-cell tiles, Morton order of centroids, or both
enum CellOrder : int { kOrderNative = 0, kOrderGrouped = 1, kOrderMorton = 2, kOrderMortonGrouped = 3 };
// residual kernels: one lane per cell-face entry, one thread per cell, or
// (inviscid flux only) one thread per internal face plus an entry gather
enum ResidKernels : int { kResidEntry = 0, kResidCell = 1, kResidFace = 2 };

// Parameters of UgpWithCvCompFlow::init / runBackwardEuler read from Joe.in.
struct SolverParams {
  double gamma = 1.4, R_gas = 1.0, p_ref = 1.0, rho_ref = 1.0, T_ref = 1.0;
  double gravity[3] = {0, 0, 0};
  int scheme = kHllc;
  bool second_order = false;
  bool grad_ls = false;  // false: Green-Gauss (the reference default)
  bool limiter = true;
  double eps_limiter = 0.001;
  int symm_bc = kSymmOriginal;
  int dt_mode = 0;  // 0 CFL (local), 1 CFL_MIN, 2 DT
  bool dt_sum = false;  // TIME_STEP = SUM_BASED (MAX_BASED otherwise)
  // TIME_INTEGRATION = BACKWARD_EULER_COUPLED: calcRhsCoupled's fluxes and
  // Jacobian (scheme kHllcC; with TURB_MODEL = SA the scalar in 6x6 blocks) in
  // the same backward-Euler loop, the block-Jacobi BiCGStab standing in for
  // the reference's PETSc GMRES
  bool coupled = false;
  double cfl = 0.0, dt_fixed = 0.0;
  double under_relax = 0.3;
  int ls_maxiter = 30;
  double ls_abs = 1e-8, ls_rel = 1e-2;
  int cfl_after = 100, cfl_interval = 10;
  double cfl_factor = 1.0, cfl_max = 100.0;
  int nsteps = 100;
  double resid_energ_th = 0.0, resid_energ_rel = 0.0;
  double rho_init = 1.0, p_init = 1.0, u_init[3] = {0, 0, 0};
  int cell_order = kOrderMorton;
  double mu_ref = 0.0, mu_power_law = 0.76, sl_tref = 273.15, sl_sref = 110.4, pr = 0.72, pr_turb = 0.9;
  int visc_mode = 0;
  bool visc_limiter = true;
  int grad_kernel = kResidCell, flux_kernel = kResidEntry;
  int jac_kernel = kResidEntry;  // GPU_JAC_KERNEL: ENTRY (one lane per cell-face entry) or CELL (one thread per row)
  int scal_kernel = kResidEntry;  // GPU_SCAL_KERNEL: the SST scalar rows per entry or per cell
  // preconditioner of the step's BiCGStab (GPU_PRECONDITIONER): 0 = block
  // Jacobi (the reference's), else ILU0: block ILU(0) on subdomains of this
  // many consecutive cells of the device order (GPU_ILU_SUBDOMAIN)
  int ilu_sub = 0;
  // the per-cell gradient's lanes take a tile's cells by face count (GPU_GRAD_THREAD_ORDER)
  bool grad_thread_sorted = false;
  // the FP32 BiCGStab's residual by the recurrence s - omega t (GPU_KRYLOV_RESIDUAL = RECURRENCE)
  bool krylov_recurrence = true;
  // Spalart-Allmaras as one segregated scalar nuSA (TURB_MODEL = SA): initial
  // value (INITIAL_CONDITION_TURB) and UNDER_RELAXATION_SCALARS
  bool sa = false;
  double sa_init = 0.0, relax_scal = 0.2;
  // Menter SST k-omega (TURB_MODEL = SST): kine, omega initial values,
  // SST_LIMIT_PK and SST_LIMITER_SHEARSTRESS
  bool sst = false;
  double sst_init[2] = {0.0, 0.0};
  int sst_limit_pk = 1, sst_limiter_shear = 0;
  // the boundary-profile file of PROFILES zones (the last such zone's)
  std::string profiles_file;
};

struct ZoneBc {
  int bc = 0;
  double par[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int skind = 0;  // 1: "<zone>.nuSA = sval" (Dirichlet on inflow faces); with SST kine's
  double sval = 0.0;
  int skind2 = 0;  // SST omega: 1 "<zone>.omega = sval2", 2 "= HOOK" (the model's wall value)
  double sval2 = 0.0;
};

void parse_params(const std::string &joe_in, const Mesh &m, SolverParams &p, std::vector<ZoneBc> &zbc);

// One solve's block system for the batched workgroup BiCGStab (device
// pointers): A[nnz][5][5], its diagonal-block LU [ncv][25], ten work vectors
// [ncv][5] (rhs, dq, res, res0, p, v, phat, s, shat, t) in the solve's
// precision, and its solver scalars.
struct KrylovJob {
  int ncv = 0, maxiter = 0;
  const int *nbocv_i = nullptr, *nbocv_v = nullptr;
  const void *A = nullptr;
  void *LU = nullptr;
  void *v[10] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void *S = nullptr;
  // first block and block count of this solve in each launch shape of the
  // batch (5 lanes per cell, a thread per cell, a thread per unknown);
  // filled by setup_krylov_batch
  int b5 = 0, nb5 = 0, b1 = 0, nb1 = 0, bx = 0, nbx = 0;
};

// device buffers of the batched BiCGStab, grown on demand
struct KrylovBatchBuf {
  KrylovJob *jobs_d = nullptr;
  int *map_d = nullptr;
  double *partials_d = nullptr;
  int *flag_h = nullptr, *flag_d = nullptr;
  size_t cap_jobs = 0, cap_map = 0, cap_partials = 0;
  int njobs = 0, g5 = 0, g1 = 0, gx = 0, max_it = 0;  // the batch set up last
  int last_launched = 0;  // iterations the last run launched before every solve stopped
  KrylovBatchBuf() = default;
  KrylovBatchBuf(const KrylovBatchBuf &) = delete;
  KrylovBatchBuf &operator=(const KrylovBatchBuf &) = delete;
  ~KrylovBatchBuf();
};

// the block-Jacobi BiCGStab of njobs solves of one precision, each kernel
// launched once for all of them: setup_krylov_batch (synchronous, on stream)
// builds the batch's tables, run_krylov_batch runs the batch set up last on
// stream, synchronising between chunks of iterations. jobs_h must be pinned
// host memory.
void setup_krylov_batch(KrylovJob *jobs_h, int njobs, void *stream, KrylovBatchBuf &buf);
void run_krylov_batch(int precision, void *stream, KrylovBatchBuf &buf);

class Ctx;

// A 5x5 block system on the device in the reference's layout (UgpWithCv
// members nbocv_i[ncv+1], nbocv_v[nnz] with the diagonal first, A[nnz][5][5],
// rhs[ncv][5], phi[ncv][5] the initial guess and solution) and the stop rule
// of solveCvVectorR5Bcgstab.
struct BlockSolve {
  int ncv = 0;
  const int *nbocv_i = nullptr, *nbocv_v = nullptr;
  const double *A = nullptr, *rhs = nullptr;
  double *phi = nullptr;
  double zero_abs = 1e-8, zero_rel = 1e-2;
  int max_iter = 30;
  void *stream = nullptr;
};
// solveCvVectorR5Bcgstab on it with a work space allocated for the call
void block_bcgstab(const BlockSolve &bs, int &iters, int &status);
// the same with the library's per-device static work space (no allocation):
// kNoAllocScratchDoubles doubles (64 MiB), enough for
// block_bcgstab_noalloc_max_cells() cells; larger systems are rejected
// before any device call. Calls that overlap on one device race.
constexpr size_t kNoAllocScratchDoubles = size_t(8) << 20;
void block_bcgstab_noalloc(const BlockSolve &bs, int &iters, int &status);
size_t block_bcgstab_noalloc_max_cells();

// Solves of one kernel configuration (equal Ctx::batch_key) stepping in
// lockstep for batched serving: each kernel of the step before the linear
// solve (CFL ramp, gradients, residual and local time step, Jacobian, solver
// set-up: pre) and after it (update, state and boundary refresh, L1, report:
// post) launched once for all of them, each solve keeping its own block
// partition, partials and scalars (so the same bits as its own step).
class StepBatch {
 public:
  StepBatch();
  ~StepBatch();
  StepBatch(const StepBatch &) = delete;
  StepBatch &operator=(const StepBatch &) = delete;
  // the batch's tables for these solves (each batch_prepare'd); synchronous
  void set(Ctx *const *ctxs, int n, void *stream);
  void pre(void *stream) const;
  void post(void *stream) const;
  int size() const;

  struct Impl;

 private:
  std::unique_ptr<Impl> d_;
};

struct StepInfo {
  int step = 0, lin_iters = 0, lin_status = 0;
  double cfl = 0, dt_min = 0;
  double resid[5] = {0, 0, 0, 0, 0};
  double resid_l2[5] = {0, 0, 0, 0, 0};  // L2 norms of the same residual
  double resid_s = 0;   // L1 of the nuSA residual (Spalart-Allmaras), or SST's kine
  double resid_s2 = 0;  // SST: omega's
  int scal_iters = 0;   // the scalars' Jacobi BiCGStab iterations
};

class Ctx {
 public:
  Ctx(const Mesh &m, const std::string &joe_in, int precision, void *stream);
  ~Ctx();
  Ctx(const Ctx &) = delete;
  Ctx &operator=(const Ctx &) = delete;

  void set_state(const double *rho, const double *rhou, const double *rhoE);
  // the step count a solve resumed from a restart file continues from
  void set_step(int step);
  // the transported scalar nuSA (Spalart-Allmaras), reference numbering
  // the transported scalar s (0: nuSA or kine, 1: omega) in reference numbering
  void set_scalar(const double *phi, int s = 0);
  void get_state(double *rho, double *rhou, double *rhoE);
  void init_uniform();
  void calc_rhs(bool implicit);
  // returns element count; copies when out != nullptr
  int64_t get(const std::string &name, double *out);
  StepInfo step();
  // the same step in two halves, for hosts whose sourceHook / boundaryHook
  // run between them: step_begin does the CFL ramp, local time step,
  // residual and Jacobian (without the V/dt term; the device's own gravity
  // source is left out when host_source is set); add_source adds the host's
  // per-cell source terms (reference numbering) to the residual and to the
  // diagonal blocks; step_end adds V/dt, relaxes, solves, updates, refreshes
  // state and boundary values and returns the step's residuals.
  void step_begin(bool host_source);
  void add_source(const double *rhs_rho, const double *rhs_rhou, const double *rhs_rhoE, const double *a_diag);
  // sourceHookCoupled's terms in the system's own blocks: rhs [ncv][b],
  // a_diag [ncv][b][b] (reference numbering; b the system's block size)
  void add_source_coupled(const double *rhs, const double *a_diag, int b);
  StepInfo step_end();
  // calcStateVariables + setBC on the resident state (after new HOOK values)
  void update_bc();
  // the same step with no host synchronisation inside: captured once into a
  // hipGraph and replayed on the context's stream (the CFL ramp and step
  // counter live on the device); _async launches it, _ready polls it,
  // _result waits and returns its report
  void device_step_async();
  bool device_step_ready();
  StepInfo device_step_result();
  StepInfo device_step();
  // batched serving: a step split around its linear solve, the halves run
  // by a StepBatch and the solve by run_krylov_batch, for many solves at
  // once. batch_prepare makes the buffers, the first refresh and the device
  // step counter (and rejects what the batched step does not run);
  // batch_key names the kernel configuration (solves of equal keys share a
  // StepBatch); batch_job gives the solve's linear system; batch_result reads
  // the step's report after the caller synchronised its stream.
  void batch_prepare();
  uint32_t batch_key() const;
  void batch_job(KrylovJob &job);
  StepInfo batch_result();
  int precision() const;
  // device bytes a context for this mesh holds once it has stepped
  static size_t footprint(const Mesh &m, int precision);
  static size_t footprint(const Mesh &m, int precision, bool sa);
  int run(int max_steps, double *hist, bool graph = false);
  // doneSolver (JoeWithModels.h:42-95) after a step, plus RESID_ENERG_REL
  bool done_solver(const StepInfo &s);
  void set_boundary_state(int ifa_f, int nface, const double *T, const double *vel, const double *p);
  // writeWallValues (joe.cpp:43-130) on a boundary face range at the current
  // state: out [nface][16] (x_fa, rho, press, temp, tau_wall, qdot, yplus,
  // mulam, pressure and shear force of the face); returns nface
  int wall_values(int ifa_f, int nface, double *out);
  // the boundary values setBC last computed on a face range (T_bfa, vel_bfa, p_bfa)
  void get_boundary_state(int ifa_f, int nface, double *T, double *vel, double *p);
  double *device_state();
  double *device_rhs();
  void sync();

  // timing hooks for the bench: launch the residual kernels only
  void launch_residual();
  // mean ms per kernel of one residual evaluation; returns the kernel names
  std::string time_kernels(int reps, double *ms_out);
  // mean ms of the residual, the Jacobian assembly, the Krylov set-up and one
  // BiCGStab iteration of an implicit step; returns their names
  std::string time_step_kernels(int reps, double *ms_out);
  int ncv() const;

  struct Impl;

 private:
  friend class StepBatch;
  std::unique_ptr<Impl> d_;
};

}  // namespace jm
