// The FPVA chemistry table on the device: a structured Cartesian table of FP32
// values over (Zmean, Zvar, progress variable), looked up by trilinear
// interpolation with FP64 weights for a whole field of points per call
// (ChemtableCartesianLinear_single.h:44-199, ChemtableCartesianLinear_single.cpp).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace jm {

// What the table file holds (ChemtableCartesianLinear_single.cpp:19-246):
// data is [n1][n2][n3][nvar], the variable fastest.
struct ChemtableHost {
  std::string model, type;
  double pref = 0.0, toxi = 0.0, tfuel = 0.0;
  int n1 = 0, n2 = 0, n3 = 0, nvar = 0;
  std::vector<double> x1, x2, x3;
  std::vector<std::string> names;
  std::vector<float> data;
};

// Reads the reference's binary table file (version 2.0); throws
// std::invalid_argument for a file that is not one.
ChemtableHost read_chemtable(const std::string &path);

class Chemtable {
 public:
  // checks the table (each axis at least two strictly increasing finite
  // values, one name per variable, no name twice) before any device call,
  // then uploads it
  explicit Chemtable(ChemtableHost h);
  ~Chemtable();
  Chemtable(const Chemtable &) = delete;
  Chemtable &operator=(const Chemtable &) = delete;

  const ChemtableHost &meta() const { return h_; }  // everything but the data
  int var_index(const std::string &name) const;     // -1 if the table has none of that name
  // out[k][n] = variable vars[k] at the points (zm, zv, pc)[n]; device
  // pointers; launched on stream (0: the null stream), not waited for
  void lookup(int64_t n, const double *zm, const double *zv, const double *pc, int nout, const int *vars,
              double *out, void *stream) const;
  // the same on host arrays: copied in, looked up, copied back, waited for
  void lookup_host(int64_t n, const double *zm, const double *zv, const double *pc, int nout, const int *vars,
                   double *out) const;
  // the lookup kernel launched reps times on stream, mean ms from HIP events
  double time_lookup(int64_t n, const double *zm, const double *zv, const double *pc, int nout, const int *vars,
                     double *out, void *stream, int reps) const;

  static constexpr int kMaxOut = 32;

 private:
  ChemtableHost h_;
  int nvp_ = 0;              // floats per table node on the device (nvar rounded up to 4)
  float *data_d_ = nullptr;  // [n1][n2][n3][nvp]
  double *ax_d_ = nullptr;   // x1 | x2 | x3
};

}  // namespace jm
