// Host C++ mirror of the reference's case-class plug-in surface, running the
// steady solve on the GPU through the C ABI in include/joemi.h.
//
// A reference case is a class derived from JoeWithModels that overrides hook
// virtuals and calls run() (JoeWithModels.h:11-160, joe.cpp:963-987). This
// header keeps those names, signatures and call points so such a class
// compiles against it unchanged for the inviscid steady path:
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
// Differences a case author sees: the arrays handed to sourceHook start at
// zero (the hook's terms are added to the device residual and to the
// diagonal blocks A[nbocv_i[icv]]; off-diagonal blocks are not read back),
// and the state members rho/rhou/rhoE/vel/press/temp are host copies that
// are refreshed before each hook that needs them (syncStateFromDevice()).
#pragma once

#include <list>
#include <map>
#include <string>
#include <vector>

#include "../../include/joemi.h"

#define FA_ZONE_BOUNDARY 1  // Ugp.h:43
#define FA_ZONE_INTERNAL 6  // Ugp.h:50

class FaZone {
 public:
  int ifa_f = 0, ifa_l = -1;
  const char *getName() const { return name_.c_str(); }
  std::string getNameString() const { return name_; }
  int getKind() const { return kind_; }
  int getIndex() const { return index_; }

 private:
  friend class JoeWithModels;
  std::string name_;
  int kind_ = 0, index_ = 0;
};

// Spalart-Allmaras mix-in (RansTurbSA, TurbModel_SA.h:9-444). Joe.in needs
// INITIAL_CONDITION_TURB and MU_REF > 0 and may give "<zone>.nuSA = value"
// boundary values, as for the reference; the model constants are the
// reference's defaults. nuSA holds the initial field when initialHook runs
// (it may change it) and the final field after run(); wallDist the wall
// distance after run().
class RansTurbSA {
 public:
  RansTurbSA() {}
  virtual ~RansTurbSA() {}
  double *nuSA = nullptr;
  double *wallDist = nullptr;

 private:
  friend class JoeWithModels;
  std::vector<double> nuSA_v_, wallDist_v_;
};

class JoeWithModels {
 public:
  // inputFile: a Joe.in (JoeWithModels.h:18-21)
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
  void run();
  virtual int doneSolver(double residEnergy);

  // whether a field (rho, rhou, rhoE, or a RansTurbSA's nuSA) was read from
  // the restart file (Ugp::checkDataFlag)
  bool checkDataFlag(const void *field) const;
  // Ugp::writeRestart of the current host state (restart.out by default)
  void writeRestart(const std::string &filename = "restart.out");

  // Joe.in access (Param.h): the first entry of a key, as the reference reads it
  bool checkParam(const std::string &name) const;
  double getDoubleParam(const std::string &name, const std::string &def = "") const;
  int getIntParam(const std::string &name, const std::string &def = "") const;
  std::string getStringParam(const std::string &name, const std::string &def = "") const;

  // copy the device state into rho/rhou/rhoE and the derived vel/press/temp
  void syncStateFromDevice();
  // the device solve, created on first use (run() creates it): the model
  // mix-ins of the most-derived class choose its transported scalars
  jm_ctx *deviceContext() {
    ensureContext();
    return ctx_;
  }
  const jm_mesh *deviceMesh() const { return mesh_; }

  // mesh in the reference's numbering (Ugp/UgpWithCv members)
  int nno = 0, nfa = 0, nfa_b = 0, ncv = 0, ncv_g = 0, nbocv_s = 0;
  int (*cvofa)[2] = nullptr;
  int *nbocv_i = nullptr, *nbocv_v = nullptr;
  double (*x_cv)[3] = nullptr, (*x_fa)[3] = nullptr, (*fa_normal)[3] = nullptr;
  double *cv_volume = nullptr;
  std::list<FaZone> faZoneList;

  // flow state (host copies)
  double *rho = nullptr, (*rhou)[3] = nullptr, *rhoE = nullptr;
  double (*vel)[3] = nullptr, *press = nullptr, *temp = nullptr;
  double *gamma = nullptr, *RoM = nullptr;  // per cell, constant for an ideal gas
  double GAMMA = 1.4, R_gas = 1.0, rho_ref = 1.0, p_ref = 1.0, T_ref = 1.0;
  double gravity[3] = {0.0, 0.0, 0.0};
  int step = 0, nsteps = 0;
  double cfl = 0.0, resid_energ_th = 0.0;
  double Residual[5] = {0, 0, 0, 0, 0};
  int lastLinearIterations = 0;
  int mpi_rank = 0, mpi_size = 1;

 private:
  struct ProbeTag {};
  explicit JoeWithModels(ProbeTag) {}
  void require(jm_status st, const char *what) const;
  void ensureContext();
  bool sourceHookOverridden();
  void applyBoundaryHooks();
  void allocHost();

  std::string joe_in_;
  std::map<std::string, std::vector<std::vector<std::string>>> params_;
  jm_mesh *mesh_ = nullptr;
  jm_ctx *ctx_ = nullptr;
  std::vector<int> hook_zones_;  // faZoneList indices (mesh zone order) with HOOK
  std::vector<double> src_rho_, src_rhou_, src_rhoE_, src_A_, a_diag_;
  std::vector<double> t_bfa_, vel_bfa_, p_bfa_;
  std::vector<int> cvofa_v_, nbocv_i_v_, nbocv_v_v_;
  std::vector<double> x_cv_v_, x_fa_v_, fa_normal_v_, cv_volume_v_;
  std::vector<double> rho_v_, rhou_v_, rhoE_v_, vel_v_, press_v_, temp_v_, gamma_v_, rom_v_;
  // read from a restart file: the fields it held, its STEP
  bool read_rho_ = false, read_rhou_ = false, read_rhoE_ = false, read_nuSA_ = false;
  int restart_step_ = 0;
};
