/* CPU restatement of the FPVA chemistry-table lookup -- TEST INFRASTRUCTURE
 * ONLY. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it, as the checker; the product (libjoemi) never links or calls it.
 *
 * ChemtableCartesianLinear_single (mixing_combustion/ChemtableCartesianLinear_single.cpp):
 * the axes are kept 1-based as the reference keeps them (x[1..n]; x[0] and
 * x[n+1] are padding that a linear lookup never reads), the table as FP32
 * data[n1][n2][n3][nvar]. Pinned against the reference's own class, built
 * from its sources (oracle/_ref/chem_ref_dump), through the fixtures
 * tests/golden/chem_*.npz.
 */
#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

/* BinarySearch (CombustionGeneralDefinitions.cpp:211-245): i with
 * X[i] <= x < X[i+1] for i1 <= i < i2, x clipped to [X[i1], X[i2]] */
void binary_search(int &i, const double *X, int i1, int i2, double &x) {
  if (x <= X[i1]) {
    i = i1;
    x = X[i1];
    return;
  }
  if (x >= X[i2]) {
    i = i2 - 1;
    x = X[i2];
    return;
  }
  while (i1 < i2 - 1) {
    int mid = (i1 + i2) / 2;
    if (x > X[mid])
      i1 = mid;
    else if (x < X[mid])
      i2 = mid;
    else {
      i = mid;
      return;
    }
  }
  i = i1;
}

/* Compute_Weights (ChemtableCartesianLinear_single.cpp:359-371) */
void weights(double x, double x0, double xp1, double *w) {
  double d = xp1 - x0;
  w[0] = (xp1 - x) / d;
  w[1] = (x - x0) / d;
}

double cmin(double a, double b) { return (b < a) ? b : a; } /* std::min */
double cmax(double a, double b) { return (a < b) ? b : a; } /* std::max */

}  // namespace

extern "C" {

/* Lookup (:374-429) of variables vars[0..nout-1] at n points; out[k][n] */
void or_chemtable_lookup(int n1, int n2, int n3, int nvar, const double *x1_, const double *x2_, const double *x3_,
                         const float *data, int64_t n, const double *A1s, const double *A2s, const double *A3s,
                         int nout, const int *vars, double *out) {
  std::vector<double> x1(n1 + 2, 0.0), x2(n2 + 2, 0.0), x3(n3 + 2, 0.0);
  for (int i = 0; i < n1; i++) x1[i + 1] = x1_[i];
  for (int i = 0; i < n2; i++) x2[i + 1] = x2_[i];
  for (int i = 0; i < n3; i++) x3[i + 1] = x3_[i];
  auto D = [&](int i, int j, int k, int l) {  /* Data[i][j][k][l], 1-based node indices */
    return data[(((size_t)(i - 1) * n2 + (j - 1)) * n3 + (k - 1)) * nvar + l];
  };
  for (int64_t p = 0; p < n; p++) {
    double A1 = cmin(cmax(A1s[p], x1[1]), x1[n1]);  /* :390-392 */
    double A2 = cmin(cmax(A2s[p], x2[1]), x2[n2]);
    double A3 = cmin(cmax(A3s[p], x3[1]), x3[n3]);
    int idx[3];
    double w[3][2];
    /* SetCoordinates (:333-356) */
    binary_search(idx[0], x1.data(), 1, n1, A1);
    weights(A1, x1[idx[0]], x1[idx[0] + 1], w[0]);
    binary_search(idx[1], x2.data(), 1, n2, A2);
    weights(A2, x2[idx[1]], x2[idx[1] + 1], w[1]);
    binary_search(idx[2], x3.data(), 1, n3, A3);
    weights(A3, x3[idx[2]], x3[idx[2] + 1], w[2]);
    for (int q = 0; q < nout; q++) {
      int ll = vars[q];
      double val = 0.0, g[2], f[2];
      for (int k = 0; k < 2; k++) {  /* :414-428 */
        g[k] = 0.0;
        for (int j = 0; j < 2; j++) {
          f[j] = 0.0;
          for (int i = 0; i < 2; i++) f[j] += w[0][i] * (double)D(idx[0] + i, idx[1] + j, idx[2] + k, ll);
          g[k] += w[1][j] * f[j];
        }
        val += w[2][k] * g[k];
      }
      out[(size_t)q * n + p] = val;
    }
  }
}

}  // extern "C"
