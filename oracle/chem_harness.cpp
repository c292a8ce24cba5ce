// Test harness around the reference's own ChemtableCartesianLinear_single,
// built from its sources into oracle/_ref/ (never shipped in the product).
// Writes the reference's lookups for the fixtures under tests/golden/:
//
//   chem_ref_dump <table file> <points file> <out file>
//
// points file: int64 n, then Zmean[n], Zvar[n], C[n] (doubles). out file,
// doubles: Lookup(tag) of every variable [n][nvar], then LookupCoeff [n][10]
// when nvar >= 10, then LookupSelectedCoeff [n][5] when nvar >= 5.
#include <cstdio>
#include <cstdint>
#include <vector>

#include "ChemtableCartesianLinear_single.h"

int main(int argc, char **argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: chem_ref_dump <table> <points> <out>\n");
    return 2;
  }
  ChemtableCartesianLinear_single tab;
  tab.Load(argv[1]);
  FILE *fp = std::fopen(argv[2], "rb");
  if (!fp) return 3;
  int64_t n = 0;
  if (std::fread(&n, sizeof(n), 1, fp) != 1) return 3;
  std::vector<double> a(3 * n);
  if (std::fread(a.data(), sizeof(double), 3 * n, fp) != (size_t)(3 * n)) return 3;
  std::fclose(fp);
  const int nvar = tab.GetChemtableDimension4();
  std::vector<double> look((size_t)n * nvar), coeff, sel;
  for (int64_t p = 0; p < n; p++)
    for (int l = 0; l < nvar; l++) look[p * nvar + l] = tab.Lookup(a[p], a[n + p], a[2 * n + p], tab.GetChemtableVariables(l));
  if (nvar >= 10) {
    coeff.resize((size_t)n * 10);
    for (int64_t p = 0; p < n; p++) {
      double *c = &coeff[p * 10];
      tab.LookupCoeff(c[0], c[1], c[2], c[3], c[4], c[5], c[6], c[7], c[8], c[9], a[p], a[n + p], a[2 * n + p]);
    }
  }
  if (nvar >= 5) {
    sel.resize((size_t)n * 5);
    for (int64_t p = 0; p < n; p++) {
      double *c = &sel[p * 5];
      tab.LookupSelectedCoeff(c[0], c[1], c[2], c[3], c[4], a[p], a[n + p], a[2 * n + p]);
    }
  }
  fp = std::fopen(argv[3], "wb");
  if (!fp) return 4;
  std::fwrite(look.data(), sizeof(double), look.size(), fp);
  std::fwrite(coeff.data(), sizeof(double), coeff.size(), fp);
  std::fwrite(sel.data(), sizeof(double), sel.size(), fp);
  std::fclose(fp);
  tab.Unload();
  return 0;
}
