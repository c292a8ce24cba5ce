// ChemtableCartesianLinear_single over the joemi C ABI: the reference's class
// (mixing_combustion/ChemtableCartesianLinear_single.h:44-199) with its names,
// so a combustion model written against it compiles unchanged, and each
// lookup runs on the device with the reference's bits. A one-point call
// (Lookup, LookupCoeff, LookupSelectedCoeff) is one device round trip; the
// batched forms below look up a whole field of cells in one.
//
// Header-only; link libjoemi.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/joemi.h"

class ChemtableCartesianLinear_single {
 public:
  ChemtableCartesianLinear_single() = default;
  ~ChemtableCartesianLinear_single() { Unload(); }
  ChemtableCartesianLinear_single(const ChemtableCartesianLinear_single &) = delete;
  ChemtableCartesianLinear_single &operator=(const ChemtableCartesianLinear_single &) = delete;

  // Load (ChemtableCartesianLinear_single.cpp:19-246); the reference prints
  // what it read and throws -1 on a bad file, this throws std::runtime_error
  void Load(std::string tablename) {
    Unload();
    check(jm_chemtable_load(tablename.c_str(), &t_), "Load");
    int dims[4];
    double ref[3];
    check(jm_chemtable_info(t_, dims, ref), "Load");
    n1_ = dims[0], n2_ = dims[1], n3_ = dims[2], nvar_ = dims[3];
    pref_ = ref[0], toxi_ = ref[1], tfuel_ = ref[2];
    x_[0].resize(n1_), x_[1].resize(n2_), x_[2].resize(n3_);
    for (int d = 0; d < 3; d++) check(jm_chemtable_axis(t_, d + 1, x_[d].data()), "Load");
    model_ = text(-2), type_ = text(-1);
    names_.clear();
    for (int l = 0; l < nvar_; l++) names_.push_back(text(l));
  }
  void Unload() {
    if (t_) jm_chemtable_destroy(t_);
    t_ = nullptr;
  }

  int GetChemtableDimension1() { return n1_; }
  int GetChemtableDimension2() { return n2_; }
  int GetChemtableDimension3() { return n3_; }
  int GetChemtableDimension4() { return nvar_; }
  double GetReferencePressure() { return pref_; }
  double GetToxi() { return toxi_; }
  double GetTfuel() { return tfuel_; }
  std::string GetCombustionModel() { return model_; }
  std::string GetChemtableType() { return type_; }
  std::string GetChemtableVariables(int i) { return names_.at(i); }
  // the reference's axes are 1-based with padding at 0 and n+1; these take
  // the node index i1 = 1 .. n1 as the reference's callers do
  double GetChemtableCoordinate1(int i1) { return x_[0].at(i1 - 1); }
  double GetChemtableCoordinate2(int i2) { return x_[1].at(i2 - 1); }
  double GetChemtableCoordinate3(int i3) { return x_[2].at(i3 - 1); }

  // Lookup (:374-429): one variable by name at one point
  double Lookup(double A1, double A2, double A3, std::string tag) {
    const int v = var(tag);
    double out = 0.0;
    check(jm_chemtable_lookup_host(t_, 1, &A1, &A2, &A3, 1, &v, &out), "Lookup");
    return out;
  }
  // LookupCoeff (:438-492): the ten coefficients at one point
  void LookupCoeff(double &RoM, double &T0, double &E0, double &Gam0, double &a_Gam, double &mu0, double &a_mu,
                   double &lamOcp0, double &a_lamOcp, double &src_prog, double A1, double A2, double A3) {
    double c[10];
    lookup_n(1, &A1, &A2, &A3, 10, c);
    RoM = c[0], T0 = c[1], E0 = c[2], Gam0 = c[3], a_Gam = c[4], mu0 = c[5], a_mu = c[6], lamOcp0 = c[7],
    a_lamOcp = c[8], src_prog = c[9];
  }
  // LookupSelectedCoeff (:501-550): the first five at one point
  void LookupSelectedCoeff(double &RoM, double &T0, double &E0, double &Gam0, double &a_Gam, double A1, double A2,
                           double A3) {
    double c[5];
    lookup_n(1, &A1, &A2, &A3, 5, c);
    RoM = c[0], T0 = c[1], E0 = c[2], Gam0 = c[3], a_Gam = c[4];
  }

  // Batched forms (no counterpart in the reference): n points of host arrays
  // in one device call. coeff[k][n] receives LookupCoeff's k-th output (ncoeff
  // = 10) or LookupSelectedCoeff's (ncoeff = 5); Lookup of one variable by
  // name into out[n].
  void LookupCoeff(int n, const double *A1, const double *A2, const double *A3, double *coeff, int ncoeff = 10) {
    lookup_n(n, A1, A2, A3, ncoeff, coeff);
  }
  void Lookup(int n, const double *A1, const double *A2, const double *A3, const std::string &tag, double *out) {
    const int v = var(tag);
    check(jm_chemtable_lookup_host(t_, n, A1, A2, A3, 1, &v, out), "Lookup");
  }
  // the device table, for jm_chemtable_lookup on device arrays
  jm_chemtable *handle() { return t_; }

 private:
  static void check(jm_status s, const char *what) {
    if (s != JM_OK) throw std::runtime_error(std::string("ChemtableCartesianLinear_single::") + what + ": " +
                                             jm_last_error());
  }
  std::string text(int which) {
    char buf[128];
    check(jm_chemtable_text(t_, which, buf, sizeof buf), "Load");
    return buf;
  }
  int var(const std::string &tag) {  // an unknown tag is an error, as in the reference (:396-400)
    int v = -1;
    check(jm_chemtable_var(t_, tag.c_str(), &v), "Lookup");
    return v;
  }
  void lookup_n(int n, const double *A1, const double *A2, const double *A3, int nc, double *out) {
    if (nc < 1 || nc > 10) throw std::runtime_error("ChemtableCartesianLinear_single: 1 to 10 coefficients");
    if (nc > nvar_) throw std::runtime_error("ChemtableCartesianLinear_single: the table has fewer variables");
    int vars[10];
    for (int k = 0; k < nc; k++) vars[k] = k;
    check(jm_chemtable_lookup_host(t_, n, A1, A2, A3, nc, vars, out), "LookupCoeff");
  }

  jm_chemtable *t_ = nullptr;
  int n1_ = 0, n2_ = 0, n3_ = 0, nvar_ = 0;
  double pref_ = 0.0, toxi_ = 0.0, tfuel_ = 0.0;
  std::vector<double> x_[3];
  std::string model_, type_;
  std::vector<std::string> names_;
};
