// Host C++ mirror of JoeWithModels over the joemi C ABI; see JoeWithModels.hpp.
#include "JoeWithModels.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

JoeWithModels::JoeWithModels(const char *inputFile) : UgpWithCvCompFlow(inputFile) {
  init(inputFile);  // when the most-derived class constructed the base without it
  for (const FaZone &z : faZoneList)
    if (z.kind_ == FA_ZONE_BOUNDARY) {
      auto it = params_.find(z.name_);
      if (it != params_.end() && it->second[0].size() > 1 && it->second[0][1] == "HOOK") hook_zones_.push_back(z.index_);
    }
  const std::string prec = getStringParam("IMPLICIT_PRECISION", "FP64");
  if (prec != "FP64" && prec != "MIXED") throw std::runtime_error("IMPLICIT_PRECISION must be FP64 or MIXED");
  if (from_restart_) {  // the registered fields the file held (Ugp.cpp:12570-12700)
    auto field = [&](const char *name, double *out, size_t n) {
      int64_t cnt = 0;
      if (jm_mesh_restart_field(mesh_, name, nullptr, &cnt) != JM_OK) return false;
      if ((size_t)cnt != n) throw std::runtime_error(std::string("restart field ") + name + " has the wrong size");
      require(jm_mesh_restart_field(mesh_, name, out, &cnt), name);
      return true;
    };
    read_rho_ = field("RHO", rho, ncv);
    read_rhou_ = field("RHOU", &rhou[0][0], 3 * (size_t)ncv);
    read_rhoE_ = field("RHOE", rhoE, ncv);
    require(jm_mesh_restart_step(mesh_, &restart_step_, nullptr), "jm_mesh_restart_step");
    step = restart_step_;
  }
}

bool JoeWithModels::checkDataFlag(const void *field) const {
  if (field == nullptr) return false;
  if (field == rho) return read_rho_;
  if (field == rhou) return read_rhou_;
  if (field == rhoE) return read_rhoE_;
  const RansTurbSA *sa = dynamic_cast<const RansTurbSA *>(this);
  if (sa && field == sa->nuSA) return read_nuSA_;
  const RansTurbKOmSST *sst = dynamic_cast<const RansTurbKOmSST *>(this);
  if (sst && field == kine) return read_kine_;
  return sst && field == sst->omega && read_omega_;
}

void JoeWithModels::writeRestart(const std::string &filename) {
  std::vector<const char *> names = {"RHO", "RHOU", "RHOE"};
  std::vector<const double *> data = {rho, &rhou[0][0], rhoE};
  std::vector<int> dims = {1, 3, 1};
  const RansTurbSA *sa = dynamic_cast<const RansTurbSA *>(this);
  if (sa && sa->nuSA) {
    names.push_back("nuSA");
    data.push_back(sa->nuSA);
    dims.push_back(1);
  }
  const RansTurbKOmSST *sst = dynamic_cast<const RansTurbKOmSST *>(this);
  if (sst && kine && sst->omega) {
    names.push_back("kine");
    data.push_back(kine);
    dims.push_back(1);
    names.push_back("omega");
    data.push_back(sst->omega);
    dims.push_back(1);
  }
  require(jm_write_restart(mesh_, filename.c_str(), (int)names.size(), names.data(), data.data(), dims.data(), step,
                           0.0),
          "jm_write_restart");
}

// The device context, once the object is complete: IMPLICIT_PRECISION = FP64
// (default) | MIXED (Jacobian, preconditioner and Krylov in FP32 around the
// FP64 residual); a RansTurbSA mix-in adds the nuSA equation.
void JoeWithModels::ensureContext() {
  if (ctx_) return;
  std::string text = joe_in_;
  if (dynamic_cast<RansTurbSA *>(this) && !checkParam("TURB_MODEL")) text += "\nTURB_MODEL = SA\n";
  if (dynamic_cast<RansTurbKOmSST *>(this) && !checkParam("TURB_MODEL")) text += "\nTURB_MODEL = SST\n";
  const std::string prec = getStringParam("IMPLICIT_PRECISION", "FP64");
  require(jm_ctx_create(mesh_, text.c_str(), prec == "MIXED" ? JM_MIXED : JM_FP64, nullptr, &ctx_), "jm_ctx_create");
}

JoeWithModels::~JoeWithModels() {
  if (ctx_) jm_ctx_destroy(ctx_);
}

// uniform branch of JoeWithModels::initialHook (JoeWithModels.cpp:180-211)
void JoeWithModels::initialHook() {
  const double rho_init = checkParam("RHO_INITIAL") ? getDoubleParam("RHO_INITIAL") : rho_ref;
  const double p_init = checkParam("P_INITIAL") ? getDoubleParam("P_INITIAL") : p_ref;
  double u_init[3] = {0.0, 0.0, 0.0};
  if (checkParam("U_INITIAL")) {
    const std::vector<std::string> &t = params_.at("U_INITIAL")[0];
    if (t.size() < 4) throw std::runtime_error("U_INITIAL needs three values");
    for (int i = 0; i < 3; i++) u_init[i] = std::stod(t[1 + i]);
  }
  double uu = 0.0;
  for (int i = 0; i < 3; i++) uu += u_init[i] * u_init[i];
  // fields read from a restart file keep their values (JoeWithModels.cpp:200-211)
  for (int icv = 0; icv < ncv; icv++) {
    if (!checkDataFlag(rho)) rho[icv] = rho_init;
    if (!checkDataFlag(rhou))
      for (int i = 0; i < 3; i++) rhou[icv][i] = rho_init * u_init[i];
    if (!checkDataFlag(rhoE)) rhoE[icv] = p_init / (gamma[icv] - 1.0) + 0.5 * rho_init * uu;
  }
}

// default sourceHook: gravity (JoeWithModels.h:111-129). When a case does not
// override sourceHook the device applies the same terms itself.
void JoeWithModels::sourceHook(double *rhs_rho, double (*rhs_rhou)[3], double *rhs_rhoE, double (*A)[5][5]) {
  for (int icv = 0; icv < ncv; icv++) {
    rhs_rhou[icv][0] += cv_volume[icv] * rho[icv] * gravity[0];
    rhs_rhou[icv][1] += cv_volume[icv] * rho[icv] * gravity[1];
    rhs_rhou[icv][2] += cv_volume[icv] * rho[icv] * gravity[2];
    rhs_rhoE[icv] += cv_volume[icv] * (gravity[0] * rhou[icv][0] + gravity[1] * rhou[icv][1] + gravity[2] * rhou[icv][2]);
    if (A != nullptr) {
      A[nbocv_i[icv]][1][0] -= cv_volume[icv] * gravity[0];
      A[nbocv_i[icv]][2][0] -= cv_volume[icv] * gravity[1];
      A[nbocv_i[icv]][3][0] -= cv_volume[icv] * gravity[2];
      A[nbocv_i[icv]][4][1] -= cv_volume[icv] * gravity[0];
      A[nbocv_i[icv]][4][2] -= cv_volume[icv] * gravity[1];
      A[nbocv_i[icv]][4][3] -= cv_volume[icv] * gravity[2];
    }
  }
}

// True when the dynamic type replaces sourceHook: compares the function each
// object would call (GNU bound-member-function extraction).
bool JoeWithModels::sourceHookOverridden() {
  typedef void (*Fn)(JoeWithModels *, double *, double (*)[3], double *, double (*)[5][5]);
  JoeWithModels probe{ProbeTag{}};
  Fn mine = (Fn)(this->*(&JoeWithModels::sourceHook));
  Fn base = (Fn)(probe.*(&JoeWithModels::sourceHook));
  return mine != base;
}

bool JoeWithModels::temporalHookOverridden() {
  typedef void (*Fn)(JoeWithModels *);
  JoeWithModels probe{ProbeTag{}};
  Fn mine = (Fn)(this->*(&JoeWithModels::temporalHook));
  Fn base = (Fn)(probe.*(&JoeWithModels::temporalHook));
  return mine != base;
}

void JoeWithModels::syncStateFromDevice() {
  require(jm_get_state(ctx_, rho, &rhou[0][0], rhoE), "jm_get_state");
  // calcStateVariables (UgpWithCvCompFlow.h:1662-1694), ideal gas, no
  // kinetic energy; like the reference, a non-positive pressure keeps the
  // previous value
  for (int icv = 0; icv < ncv; icv++) {
    for (int i = 0; i < 3; i++) vel[icv][i] = rhou[icv][i] / rho[icv];
    const double pr = (gamma[icv] - 1.0) * (rhoE[icv] - 0.5 * vecDotVec3d(rhou[icv], rhou[icv]) / rho[icv]);
    if (pr > 0.0) press[icv] = pr;
    temp[icv] = press[icv] / (rho[icv] * RoM[icv]);
    enthalpy[icv] = gamma[icv] * RoM[icv] / (gamma[icv] - 1.0) * temp[icv];
    sos[icv] = std::sqrt(gamma[icv] * press[icv] / rho[icv]);
  }
}

void JoeWithModels::syncFieldsFromDevice() {
  syncStateFromDevice();
  int64_t n = nfa;
  require(jm_get(ctx_, "MUL_FA", mul_fa, &n), "jm_get MUL_FA");
  n = 9 * (int64_t)ncv;
  require(jm_get(ctx_, "GRAD_U", &grad_u[0][0][0], &n), "jm_get GRAD_U");
  if (RansTurbSA *sa = dynamic_cast<RansTurbSA *>(this)) {
    n = ncv;
    require(jm_get(ctx_, "NUSA", sa->nuSA, &n), "jm_get NUSA");
    require(jm_get(ctx_, "WALL_DIST", sa->wallDist, &n), "jm_get WALL_DIST");
    n = nfa;
    require(jm_get(ctx_, "MUT_FA", sa->mut_fa_v_.data(), &n), "jm_get MUT_FA");
    for (int icv = 0; icv < ncv; icv++)  // calcRansTurbViscMuet's muT (TurbModel_SA.h:126)
      sa->muT[icv] = InterpolateAtCellCenterFromFaceValues(sa->mut_fa_v_.data(), icv);
  }
  if (RansTurbKOmSST *sst = dynamic_cast<RansTurbKOmSST *>(this)) {
    n = ncv;
    require(jm_get(ctx_, "KINE", kine, &n), "jm_get KINE");
    require(jm_get(ctx_, "OMEGA", sst->omega, &n), "jm_get OMEGA");
    require(jm_get(ctx_, "WALL_DIST", sst->wallDist, &n), "jm_get WALL_DIST");
    require(jm_get(ctx_, "CROSS_DIFF", sst->crossDiff, &n), "jm_get CROSS_DIFF");
    require(jm_get(ctx_, "F1", sst->blendFuncF1, &n), "jm_get F1");
    require(jm_get(ctx_, "F2", sst->blendFuncF2, &n), "jm_get F2");
    n = nfa;
    require(jm_get(ctx_, "MUT_FA", sst->mut_fa_v_.data(), &n), "jm_get MUT_FA");
    for (int icv = 0; icv < ncv; icv++)  // calcRansTurbViscMuet's muT (TurbModel_KOMSST.h:152)
      sst->muT[icv] = InterpolateAtCellCenterFromFaceValues(sst->mut_fa_v_.data(), icv);
  }
}

// setBC's HOOK branch (JoeWithModels.cpp:3870-3881) for every HOOK zone, then
// the device's state and boundary refresh with the new values
void JoeWithModels::applyBoundaryHooks() {
  if (hook_zones_.empty()) return;
  syncStateFromDevice();
  require(jm_update_bc(ctx_), "jm_update_bc");
  t_bfa_.assign(nfa_b, 0.0);
  vel_bfa_.assign(3 * (size_t)nfa_b, 0.0);
  p_bfa_.assign(nfa_b, 0.0);
  int64_t n = nfa_b;
  require(jm_get(ctx_, "T_BFA", t_bfa_.data(), &n), "jm_get T_BFA");
  n = 3 * (int64_t)nfa_b;
  require(jm_get(ctx_, "VEL_BFA", vel_bfa_.data(), &n), "jm_get VEL_BFA");
  n = nfa_b;
  require(jm_get(ctx_, "P_BFA", p_bfa_.data(), &n), "jm_get P_BFA");
  for (FaZone &z : faZoneList) {
    bool hooked = false;
    for (int k : hook_zones_) hooked |= (k == z.index_);
    if (!hooked) continue;
    boundaryHook(t_bfa_.data(), reinterpret_cast<double(*)[3]>(vel_bfa_.data()), p_bfa_.data(), &z);
    require(jm_set_boundary_state(ctx_, z.index_, t_bfa_.data() + z.ifa_f, vel_bfa_.data() + 3 * (size_t)z.ifa_f,
                                  p_bfa_.data() + z.ifa_f),
            "jm_set_boundary_state");
  }
  require(jm_update_bc(ctx_), "jm_update_bc");
}

// doneSolver (JoeWithModels.h:42-95); RUNTIME is not read (the reference's
// serial timers read zero, so it never stops a serial run either)
int JoeWithModels::doneSolver(double residEnergy) {
  int done = 0;
  if ((nsteps >= 0) && (step >= nsteps)) done = 1;
  if (residEnergy < resid_energ_th) done = 1;
  if (residEnergy != residEnergy) done = 1;
  if (std::FILE *fp = std::fopen("killjoe", "r")) {
    std::fclose(fp);
    std::remove("killjoe");
    done = 1;
  }
  return done;
}

// JoeWithModels::run -> runBackwardEuler (JoeWithModels.cpp:881-1198)
void JoeWithModels::run() {
  // the device mesh is the file's as read: transformMeshHook, which the
  // reference calls while reading it (Ugp.h:3045), is not called here
  if (checkParam("TRANSFORM_MESH"))
    throw std::runtime_error("TRANSFORM_MESH: mesh transforms are not supported (transformMeshHook is not called)");
  ensureContext();
  RansTurbSA *sa = dynamic_cast<RansTurbSA *>(this);
  if (sa) {  // initialHook's uniform nuSA unless the restart file held it (JoeWithModels.cpp:219-240)
    sa->nuSA_v_.assign(ncv, std::stod(params_.count("INITIAL_CONDITION_TURB")
                                          ? getStringParam("INITIAL_CONDITION_TURB")
                                          : std::string("0.0")));
    int64_t cnt = 0;
    if (jm_mesh_restart_field(mesh_, "nuSA", nullptr, &cnt) == JM_OK && cnt == ncv) {
      require(jm_mesh_restart_field(mesh_, "nuSA", sa->nuSA_v_.data(), &cnt), "nuSA");
      read_nuSA_ = true;
    }
    sa->nuSA = sa->nuSA_v_.data();
    sa->muT_v_.assign(ncv, 0.0);
    sa->wallDist_v_.assign(ncv, 0.0);
    sa->mut_fa_v_.assign(nfa, 0.0);
    sa->muT = sa->muT_v_.data();
    sa->wallDist = sa->wallDist_v_.data();
    // the model's registered fields and its transported scalar (TurbModel_SA.h:30-50)
    registerScalar(sa->nuSA, "nuSA");
    registerScalar(sa->muT, "muT");
    registerScalar(sa->wallDist, "wallDist");
    if (!getScalarTransportData("nuSA")) {
      scalarTranspEqVector.emplace_back("nuSA");
      scalarTranspEqVector.back().phi = sa->nuSA;
    }
  }
  RansTurbKOmSST *sst = dynamic_cast<RansTurbKOmSST *>(this);
  if (sst) {  // initialHook's uniform kine and omega unless the restart file held them (JoeWithModels.cpp:219-240)
    Param *ic = nullptr;
    if (!getParam(ic, "INITIAL_CONDITION_TURB") || ic->getSize() < 3)
      throw std::runtime_error("INITIAL_CONDITION_TURB needs kine and omega");
    sst->kine_v_.assign(ncv, ic->getDouble(1));
    sst->omega_v_.assign(ncv, ic->getDouble(2));
    int64_t cnt = 0;
    if (jm_mesh_restart_field(mesh_, "kine", nullptr, &cnt) == JM_OK && cnt == ncv) {
      require(jm_mesh_restart_field(mesh_, "kine", sst->kine_v_.data(), &cnt), "kine");
      read_kine_ = true;
    }
    cnt = 0;
    if (jm_mesh_restart_field(mesh_, "omega", nullptr, &cnt) == JM_OK && cnt == ncv) {
      require(jm_mesh_restart_field(mesh_, "omega", sst->omega_v_.data(), &cnt), "omega");
      read_omega_ = true;
    }
    kine = sst->kine_v_.data();
    sst->omega = sst->omega_v_.data();
    for (auto *v : {&sst->muT_v_, &sst->wallDist_v_, &sst->crossDiff_v_, &sst->f1_v_, &sst->f2_v_}) v->assign(ncv, 0.0);
    sst->mut_fa_v_.assign(nfa, 0.0);
    sst->muT = sst->muT_v_.data();
    sst->wallDist = sst->wallDist_v_.data();
    sst->crossDiff = sst->crossDiff_v_.data();
    sst->blendFuncF1 = sst->f1_v_.data();
    sst->blendFuncF2 = sst->f2_v_.data();
    // the model's registered fields and its transported scalars (TurbModel_KOMSST.h:36-68)
    registerScalar(kine, "kine");
    registerScalar(sst->omega, "omega");
    registerScalar(sst->muT, "muT");
    registerScalar(sst->wallDist, "wallDist");
    registerScalar(sst->crossDiff, "crossDiff");
    registerScalar(sst->blendFuncF1, "blendFuncF1");
    registerScalar(sst->blendFuncF2, "blendFuncF2");
    if (!getScalarTransportData("kine")) {
      scalarTranspEqVector.emplace_back("kine");
      scalarTranspEqVector.back().phi = kine;
    }
    if (!getScalarTransportData("omega")) {
      scalarTranspEqVector.emplace_back("omega");
      scalarTranspEqVector.back().phi = sst->omega;
    }
  }
  initialHook();
  require(jm_set_state(ctx_, rho, &rhou[0][0], rhoE), "jm_set_state");
  if (sa) require(jm_set_scalar(ctx_, "nuSA", sa->nuSA), "jm_set_scalar");
  if (sst) {
    require(jm_set_scalar(ctx_, "kine", kine), "jm_set_scalar kine");
    require(jm_set_scalar(ctx_, "omega", sst->omega), "jm_set_scalar omega");
  }
  // the STEP of a restart file carries on (runBackwardEuler does not reset
  // it) unless RESET_STEP (JoeWithModels.cpp:12-17)
  step = checkParam("RESET_STEP") ? 0 : restart_step_;
  if (step > 0) require(jm_set_step(ctx_, step), "jm_set_step");
  const int write_restart = getIntParam("WRITE_RESTART", "-1");
  // host copies for an overridden temporalHook, every HOST_SYNC_INTERVAL
  // steps (this build's key; 0 = never)
  const int sync_every = temporalHookOverridden() ? getIntParam("HOST_SYNC_INTERVAL", "1") : 0;
  applyBoundaryHooks();
  // runBackwardEulerCoupled calls sourceHookCoupled (default: gravity, which
  // the device applies) instead of sourceHook (JoeWithModels.cpp:4596-4611),
  // so a sourceHook override takes no part in a coupled run
  const bool coupled = getStringParam("TIME_INTEGRATION", "BACKWARD_EULER") == "BACKWARD_EULER_COUPLED";
  const bool host_source = !coupled && sourceHookOverridden();
  const bool split = host_source || !hook_zones_.empty();
  if (host_source) {
    src_rho_.assign(ncv, 0.0);
    src_rhou_.assign(3 * (size_t)ncv, 0.0);
    src_rhoE_.assign(ncv, 0.0);
    src_A_.assign(25 * (size_t)nbocv_s, 0.0);
    a_diag_.assign(25 * (size_t)ncv, 0.0);
  }
  int done = doneSolver(1.0e20);
  while (!done) {
    jm_step_info info;
    if (!split) {
      require(jm_step(ctx_, &info), "jm_step");
    } else {
      require(jm_step_begin(ctx_, host_source ? JM_HOST_SOURCE : 0), "jm_step_begin");
      if (host_source) {
        syncStateFromDevice();
        std::fill(src_rho_.begin(), src_rho_.end(), 0.0);
        std::fill(src_rhou_.begin(), src_rhou_.end(), 0.0);
        std::fill(src_rhoE_.begin(), src_rhoE_.end(), 0.0);
        for (int icv = 0; icv < ncv; icv++)
          for (int k = 0; k < 25; k++) src_A_[25 * (size_t)nbocv_i[icv] + k] = 0.0;
        sourceHook(src_rho_.data(), reinterpret_cast<double(*)[3]>(src_rhou_.data()), src_rhoE_.data(),
                   reinterpret_cast<double(*)[5][5]>(src_A_.data()));
        for (int icv = 0; icv < ncv; icv++)
          for (int k = 0; k < 25; k++) a_diag_[25 * (size_t)icv + k] = src_A_[25 * (size_t)nbocv_i[icv] + k];
        require(jm_add_source(ctx_, src_rho_.data(), src_rhou_.data(), src_rhoE_.data(), a_diag_.data()),
                "jm_add_source");
      }
      require(jm_step_end(ctx_, &info), "jm_step_end");
    }
    step = info.step;
    cfl = info.cfl;
    lastLinearIterations = info.lin_iters;
    for (int k = 0; k < 5; k++) Residual[k] = info.resid[k];
    applyBoundaryHooks();
    if (sync_every > 0 && step % sync_every == 0) syncFieldsFromDevice();
    temporalHook();
    done = doneSolver(Residual[4]);
    if ((write_restart > 0) && (step % write_restart == 0)) {  // JoeWithModels.cpp:1155-1156
      syncStateFromDevice();
      if (sa) {
        int64_t n = ncv;
        require(jm_get(ctx_, "NUSA", sa->nuSA_v_.data(), &n), "jm_get NUSA");
      }
      if (sst) {
        int64_t n = ncv;
        require(jm_get(ctx_, "KINE", kine, &n), "jm_get KINE");
        require(jm_get(ctx_, "OMEGA", sst->omega, &n), "jm_get OMEGA");
      }
      char name[32];
      std::snprintf(name, sizeof name, "restart.%06d.out", step);
      writeRestart(name);
    }
  }
  syncFieldsFromDevice();
  finalHook();
  writeRestart();  // JoeWithModels.cpp:1175-1176
}
