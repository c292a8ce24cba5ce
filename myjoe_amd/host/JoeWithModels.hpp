// Host C++ mirror of the reference's case-class plug-in surface, running the
// steady solve on the GPU through the C ABI in include/joemi.h.
//
// A reference case is a class derived from JoeWithModels that overrides hook
// virtuals and calls run() (JoeWithModels.h:11-160, joe.cpp:963-987). This
// header keeps those names, signatures and call points, and JoeWithModels
// derives virtually from UgpWithCvCompFlow (UgpWithCvCompFlow.hpp) like the
// reference's, so such a class compiles unchanged, its constructor's
// UgpWithCvCompFlow(name) included:
//
//   initialHook()   before the loop; default = uniform RHO/P/U_INITIAL
//                   (JoeWithModels.cpp:156-211). Writes rho, rhou, rhoE.
//   boundaryHook()  inside setBC for zones whose Joe.in value is HOOK
//                   (JoeWithModels.cpp:3870-3881): fill T_bfa, vel_bfa,
//                   p_bfa over [zone->ifa_f, zone->ifa_l]. The first argument
//                   receives temperatures although the reference names it
//                   rho (JoeWithModels.h:105 vs JoeWithModels.cpp:3877).
//   sourceHook()    after the fluxes in calcRhs (JoeWithModels.h:111-129);
//                   default = gravity from GX/GY/GZ, run on the device.
//   temporalHook()  after every step; finalHook() after the loop.
//   doneSolver()    NSTEPS, RESID_ENERG_TH, NaN, a file named killjoe.
//
// RESTART names a Fluent .msh or a restart file of the reference's format
// (Ugp::readRestart): then the mesh, RHO/RHOU/RHOE, nuSA and STEP come from
// it (checkDataFlag tells initialHook which fields were read), and the step
// count carries on unless RESET_STEP. Like runBackwardEuler, run() writes
// restart.<step>.out every WRITE_RESTART steps and restart.out at the end
// (writeRestart).
//
// A case class that also derives from RansTurbSA (as joe.cpp's MyJoeSA,
// joe.cpp:286-296) solves the Spalart-Allmaras scalar nuSA with the flow.
//
// Host copies. The state members (rho, rhou, rhoE and vel, press, temp,
// enthalpy, sos), mul_fa, grad_u and the registered model fields (nuSA,
// muT, wallDist) live on the device during the loop. They are copied to the
// host before initialHook's successors read them: before each sourceHook and
// boundaryHook call (the state), before temporalHook when a case overrides it
// (every HOST_SYNC_INTERVAL steps, default 1; this build's Joe.in key, since
// each copy costs a device-to-host transfer), and before finalHook. A hook
// may call syncStateFromDevice() itself. The arrays handed to sourceHook
// start at zero (the hook's terms are added to the device residual and to
// the diagonal blocks A[nbocv_i[icv]]; off-diagonal blocks are not read back).
#pragma once

#include <string>
#include <vector>

#include "UgpWithCvCompFlow.hpp"

// Spalart-Allmaras mix-in (RansTurbSA, TurbModel_SA.h:9-444). Joe.in needs
// INITIAL_CONDITION_TURB and MU_REF > 0 and may give "<zone>.nuSA = value"
// boundary values, as for the reference; the model constants are the
// reference's defaults. nuSA holds the initial field when initialHook runs
// (it may change it) and the current field when hooks run; muT (the cell
// average of the face eddy viscosity, TurbModel_SA.h:126) and wallDist are
// refreshed with it.
class RansTurbSA : virtual public UgpWithCvCompFlow {
 public:
  RansTurbSA() {}
  virtual ~RansTurbSA() {}
  double *nuSA = nullptr;
  double *muT = nullptr;
  double *wallDist = nullptr;

 private:
  friend class JoeWithModels;
  std::vector<double> nuSA_v_, muT_v_, wallDist_v_, mut_fa_v_;
};

// the Menter SST k-omega mix-in (TurbModel_KOMSST.h:8-75): a case class
// deriving from it (joe.cpp's MyJoeSST) solves kine (the base's kine) and
// omega; the model's fields are host copies like RansTurbSA's
class RansTurbKOmSST : virtual public UgpWithCvCompFlow {
 public:
  RansTurbKOmSST() {}
  virtual ~RansTurbKOmSST() {}
  double *omega = nullptr;
  double *muT = nullptr;
  double *wallDist = nullptr;
  double *crossDiff = nullptr, *blendFuncF1 = nullptr, *blendFuncF2 = nullptr;

 private:
  friend class JoeWithModels;
  std::vector<double> kine_v_, omega_v_, muT_v_, wallDist_v_, crossDiff_v_, f1_v_, f2_v_, mut_fa_v_;
};

class JoeWithModels : virtual public UgpWithCvCompFlow {
 public:
  // inputFile: a Joe.in (JoeWithModels.h:18-21); read here unless the
  // most-derived class already constructed UgpWithCvCompFlow(inputFile)
  explicit JoeWithModels(const char *inputFile);
  virtual ~JoeWithModels();
  JoeWithModels(const JoeWithModels &) = delete;
  JoeWithModels &operator=(const JoeWithModels &) = delete;

  virtual void initialHook();
  virtual void boundaryHook(double *T_bfa, double (*vel_bfa)[3], double *p_bfa, FaZone *zone) {}
  virtual void sourceHook(double *rhs_rho, double (*rhs_rhou)[3], double *rhs_rhoE, double (*A)[5][5]);
  virtual void temporalHook() {}
  virtual void finalHook() {}

  // steady solve: initialHook, then runBackwardEuler until doneSolver
  virtual void run();
  virtual int doneSolver(double residEnergy);

  // whether a field (rho, rhou, rhoE, or a RansTurbSA's nuSA) was read from
  // the restart file (Ugp::checkDataFlag)
  bool checkDataFlag(const void *field) const;
  // Ugp::writeRestart of the current host state (restart.out by default)
  void writeRestart(const std::string &filename = "restart.out");

  // copy the device state into rho/rhou/rhoE and the derived vel, press,
  // temp, enthalpy, sos (calcStateVariables)
  void syncStateFromDevice();
  // also mul_fa, grad_u and the model fields (nuSA, muT, wallDist)
  void syncFieldsFromDevice();
  // the device solve, created on first use (run() creates it): the model
  // mix-ins of the most-derived class choose its transported scalars
  jm_ctx *deviceContext() {
    ensureContext();
    return ctx_;
  }
  const jm_mesh *deviceMesh() const { return mesh_; }

 private:
  struct ProbeTag {};
  explicit JoeWithModels(ProbeTag) {}
  void ensureContext();
  bool sourceHookOverridden();
  bool temporalHookOverridden();
  void applyBoundaryHooks();

  jm_ctx *ctx_ = nullptr;
  std::vector<int> hook_zones_;  // faZoneList indices (mesh zone order) with HOOK
  std::vector<double> src_rho_, src_rhou_, src_rhoE_, src_A_, a_diag_;
  std::vector<double> t_bfa_, vel_bfa_, p_bfa_;
  // read from a restart file: the fields it held, its STEP
  bool read_rho_ = false, read_rhou_ = false, read_rhoE_ = false, read_nuSA_ = false;
  bool read_kine_ = false, read_omega_ = false;
  int restart_step_ = 0;
};
