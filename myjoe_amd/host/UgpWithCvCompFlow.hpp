// Host C++ mirror of the virtual base a reference case class reads its data
// from: UgpWithCvCompFlow (JoeWithModels.h:11, UgpWithCvCompFlow.h) with the
// Ugp/UgpWithCv members it inherits, the Joe.in access of Param.h, the
// registered-field lookups of Ugp.h, the vector helpers of tc_vec3d.h and the
// MPI calls a case makes, served for one rank. With it a case class written
// for the reference, such as joe.cpp's MyJoe and MyJoeSA, compiles unchanged
// against JoeWithModels.hpp; the steady solve itself runs on the GPU.
//
// The mesh arrays are in the reference's numbering and fixed for the run.
// The flow arrays (rho, rhou, rhoE and what calcStateVariables derives from
// them), mul_fa, grad_u and the registered model fields are host copies of
// the device state: JoeWithModels refreshes them before the hooks that read
// them (see JoeWithModels.hpp).
#pragma once

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <list>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/joemi.h"

// The reference's headers bring std into scope, and case code uses string,
// list, cout and the like unqualified.
using namespace std;

// ---------------------------------------------------------------------------
// MPI as a case class calls it, for one rank (the steady solve here runs on
// one GPU; several GPUs are not built). Reductions copy their input;
// point-to-point calls have no partner and throw.
// ---------------------------------------------------------------------------
typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Op;
struct MPI_Status {
  int MPI_SOURCE = 0, MPI_TAG = 0, MPI_ERROR = 0;
};
constexpr MPI_Comm MPI_COMM_WORLD = 0;
constexpr MPI_Datatype MPI_CHAR = 1, MPI_INT = 2, MPI_LONG = 3, MPI_FLOAT = 4, MPI_DOUBLE = 5;
constexpr MPI_Op MPI_SUM = 1, MPI_MAX = 2, MPI_MIN = 3;
inline MPI_Comm mpi_comm = MPI_COMM_WORLD;

inline size_t jm_mpi_size_of(MPI_Datatype t) {
  switch (t) {
    case MPI_CHAR: return 1;
    case MPI_INT: return sizeof(int);
    case MPI_LONG: return sizeof(long);
    case MPI_FLOAT: return sizeof(float);
    case MPI_DOUBLE: return sizeof(double);
  }
  throw std::invalid_argument("unknown MPI datatype");
}
inline int MPI_Allreduce(const void *in, void *out, int n, MPI_Datatype t, MPI_Op, MPI_Comm) {
  if (in != out) std::memmove(out, in, (size_t)n * jm_mpi_size_of(t));
  return 0;
}
inline int MPI_Reduce(const void *in, void *out, int n, MPI_Datatype t, MPI_Op op, int, MPI_Comm c) {
  return MPI_Allreduce(in, out, n, t, op, c);
}
inline int MPI_Bcast(void *, int, MPI_Datatype, int, MPI_Comm) { return 0; }
inline int MPI_Barrier(MPI_Comm) { return 0; }
inline int MPI_Send(const void *, int, MPI_Datatype, int, int, MPI_Comm) {
  throw std::logic_error("MPI_Send: one rank, no partner");
}
inline int MPI_Recv(void *, int, MPI_Datatype, int, int, MPI_Comm, MPI_Status *) {
  throw std::logic_error("MPI_Recv: one rank, no partner");
}
inline double MPI_Wtime() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ---------------------------------------------------------------------------
// tc_vec3d.h helpers and UgpWithCv.h's sign, same arithmetic order
// ---------------------------------------------------------------------------
inline double vecDotVec3d(const double *a, const double *b) {
  double s = 0.0;
  for (int i = 0; i < 3; i++) s += a[i] * b[i];
  return s;
}
inline void vecMinVec3d(double *out, const double *a, const double *b) {
  for (int i = 0; i < 3; i++) out[i] = a[i] - b[i];
}
inline double normVec3d(double *out, const double *v) {
  const double mag = std::sqrt(vecDotVec3d(v, v));
  for (int i = 0; i < 3; i++) out[i] = v[i] / mag;
  return mag;
}
inline double normVec3d(double *v) { return normVec3d(v, v); }
inline double sign(double a) { return a > 0 ? 1.0 : (a < 0.0 ? -1.0 : 0.0); }

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
  friend class UgpWithCvCompFlow;
  friend class JoeWithModels;
  std::string name_;
  int kind_ = 0, index_ = 0;
};

// one Joe.in line: its tokens, name first (Param.h:32-260)
class Param {
 public:
  explicit Param(std::vector<std::string> tokens) : t_(std::move(tokens)) {}
  const std::string &getName() const { return t_[0]; }
  int getSize() const { return (int)t_.size(); }
  std::string getString(size_t pos = 1) const { return at(pos); }
  int getInt(size_t pos = 1) const { return std::stoi(at(pos)); }
  double getDouble(size_t pos = 1) const { return std::stod(at(pos)); }
  // the value after a token, as "CFL_RAMP AFTER_ITER=20 ..." is read
  std::string getString(const std::string &after) const { return at(pos_after(after)); }
  int getInt(const std::string &after) const { return std::stoi(at(pos_after(after))); }
  double getDouble(const std::string &after) const { return std::stod(at(pos_after(after))); }

 private:
  const std::string &at(size_t pos) const {
    if (pos >= t_.size()) throw std::out_of_range("parameter " + t_[0] + " has no value at this position");
    return t_[pos];
  }
  size_t pos_after(const std::string &name) const {
    for (size_t i = 1; i + 1 < t_.size(); i++)
      if (t_[i] == name) return i + 1;
    throw std::out_of_range("parameter " + t_[0] + " has no " + name);
  }
  std::vector<std::string> t_;
};

// a registered cell field (Ugp.h:418-480): the flag says whether it holds
// data read from a file or computed already
class DoubleScalar {
 public:
  DoubleScalar(std::string name, double **ptr) : name_(std::move(name)), ptr_(ptr) {}
  const char *getName() const { return name_.c_str(); }
  double *getData() const { return *ptr_; }
  void setFlag() { flag_ = 1; }
  void clearFlag() { flag_ = 0; }
  int checkFlag() const { return flag_; }

 private:
  std::string name_;
  double **ptr_;
  int flag_ = 0;
};

// a transported scalar's arrays (UgpWithCv.h:50-113): the field and its
// gradient, cell data in the reference's numbering
class ScalarTranspEq {
 public:
  explicit ScalarTranspEq(std::string name) : name_(std::move(name)) {}
  const char *getName() const { return name_.c_str(); }
  double *phi = nullptr;
  double (*grad_phi)[3] = nullptr;
  double *phi_bfa = nullptr;

 private:
  std::string name_;
};
typedef std::vector<ScalarTranspEq>::iterator ScalarTranspEqIterator;

class UgpWithCvCompFlow {
 public:
  // the most-derived class constructs this virtual base; either form works
  // (a case that omits it gets the Joe.in read by JoeWithModels' constructor)
  UgpWithCvCompFlow() {}
  explicit UgpWithCvCompFlow(const char *inputFile) { init(inputFile); }
  virtual ~UgpWithCvCompFlow();
  UgpWithCvCompFlow(const UgpWithCvCompFlow &) = delete;
  UgpWithCvCompFlow &operator=(const UgpWithCvCompFlow &) = delete;

  // Joe.in access (Param.h): the first line of a key, as the reference reads it
  bool checkParam(const std::string &name) const;
  bool getParam(Param *&p, const std::string &name);
  double getDoubleParam(const std::string &name, const std::string &def = "") const;
  double getDoubleParam(const std::string &name, double def) const;
  int getIntParam(const std::string &name, const std::string &def = "") const;
  int getIntParam(const std::string &name, int def) const;
  std::string getStringParam(const std::string &name, const std::string &def = "") const;

  // registered cell fields (Ugp.h:2147-2180): RHO, RHOE, press, temp,
  // and the model's (nuSA, muT, wallDist); nullptr when not registered
  double *getR1(const char *name);
  double *getR1(const std::string &name) { return getR1(name.c_str()); }
  DoubleScalar *getScalarData(const char *name);
  DoubleScalar *getScalarData(const std::string &name) { return getScalarData(name.c_str()); }
  ScalarTranspEq *getScalarTransportData(const std::string &name);
  int getScalarTransportIndex(const std::string &name);

  // average of a face field over a cell's faces in faocv order
  // (UgpWithCvCompFlow.h:1538-1551)
  double InterpolateAtCellCenterFromFaceValues(double *phi, int icv) const;

  int getNno() const { return nno; }
  int getNcv() const { return ncv; }
  int getNfa() const { return nfa; }

  // mesh in the reference's numbering (Ugp/UgpWithCv members)
  int nno = 0, nfa = 0, nfa_b = 0, ncv = 0, ncv_g = 0, nbocv_s = 0;
  int (*cvofa)[2] = nullptr;
  int *faocv_i = nullptr, *faocv_v = nullptr;
  int *nbocv_i = nullptr, *nbocv_v = nullptr;
  double (*x_no)[3] = nullptr, (*x_cv)[3] = nullptr, (*x_fa)[3] = nullptr, (*fa_normal)[3] = nullptr;
  double *cv_volume = nullptr;
  std::list<FaZone> faZoneList;

  // flow state and derived fields (host copies)
  double *rho = nullptr, (*rhou)[3] = nullptr, *rhoE = nullptr;
  double (*vel)[3] = nullptr, *press = nullptr, *temp = nullptr, *enthalpy = nullptr, *sos = nullptr;
  double *gamma = nullptr, *RoM = nullptr;        // per cell, constant for an ideal gas
  double *mul_fa = nullptr;                       // laminar viscosity per face (zero when inviscid)
  double (*grad_u)[3][3] = nullptr;               // velocity gradient of the last residual
  double (*rij_diag)[3] = nullptr, (*rij_offdiag)[3] = nullptr;  // Reynolds stresses (no model here: zero)
  double *kine = nullptr;  // turbulent kinetic energy of a model with one (SST; UgpWithCvCompFlow.h:656)
  std::vector<ScalarTranspEq> scalarTranspEqVector;

  double GAMMA = 1.4, R_gas = 1.0, rho_ref = 1.0, p_ref = 1.0, T_ref = 1.0, Pr = 0.72, PrTurb = 0.9;
  double gravity[3] = {0.0, 0.0, 0.0};
  int step = 0, nsteps = 0;
  double cfl = 0.0, resid_energ_th = 0.0;
  double Residual[5] = {0, 0, 0, 0, 0};
  int lastLinearIterations = 0;
  int mpi_rank = 0, mpi_size = 1;

 protected:
  void init(const char *inputFile);
  bool initialised() const { return mesh_ != nullptr; }
  void require(jm_status st, const char *what) const;
  void allocHost();
  // registers a model's cell field under its reference name
  void registerScalar(double *&field, const char *name);

  std::string joe_in_;
  std::map<std::string, std::vector<std::vector<std::string>>> params_;
  std::map<std::string, Param> param_objs_;
  jm_mesh *mesh_ = nullptr;
  bool from_restart_ = false;
  std::vector<int> cvofa_v_, faocv_i_v_, faocv_v_v_, nbocv_i_v_, nbocv_v_v_;
  std::vector<double> x_no_v_, x_cv_v_, x_fa_v_, fa_normal_v_, cv_volume_v_;
  std::vector<double> rho_v_, rhou_v_, rhoE_v_, vel_v_, press_v_, temp_v_, enthalpy_v_, sos_v_, gamma_v_, rom_v_;
  std::vector<double> mul_fa_v_, grad_u_v_, rij_diag_v_, rij_offdiag_v_;
  std::list<DoubleScalar> registered_;  // a list: getScalarData's pointers stay valid
};
