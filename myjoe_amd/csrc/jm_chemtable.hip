// The FPVA chemistry-table lookup on the device (ChemtableCartesianLinear_single,
// mixing_combustion/ChemtableCartesianLinear_single.cpp): each point's
// (Zmean, Zvar, progress variable) is clipped to the table, located on each
// axis by binary descent, and the requested variables are interpolated trilinearly
// from the FP32 nodes with FP64 weights, in the reference's order of
// operations (no FMA contraction), so each value equals the reference's bit
// for bit.
//
// Layout in HBM: the nodes [n1][n2][n3][nvp] as FP32, the variable fastest and
// padded to nvp = nvar rounded up to 4, so a node's variables are whole
// 16-byte vectors and a lookup of a contiguous run of variables (the
// reference's LookupCoeff reads 0-9, LookupSelectedCoeff 0-4) loads, for each
// four variables, one dwordx4 from each of the eight corners. The three axes
// (x1 | x2 | x3, FP64) are staged in LDS per block where they fit in 32 KB.
// The points and the results are separate FP64 arrays (out[k][n]), read and
// written once and coalesced: per point 24 bytes in and 8 bytes per variable
// out; the table (a few MB) is read from the caches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>

#include "jm_chemtable.hpp"
#include "jm_param.hpp"

#define CTCHECK(x)                                                                                      \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                                   " at " + #x);                                       \
  } while (0)

namespace jm {
namespace {

constexpr int kAxisLds = 4096;  // axis values staged in LDS when n1 + n2 + n3 is at most this

struct TabDev {
  const double *ax;    // x1 | x2 | x3
  const float *data;   // [n1][n2][n3][nvp]
  int n1, n2, n3, nvp;
  size_t s1, s2;       // floats between nodes along x1 and along x2
};

struct Sel {
  int nout;
  int first;  // >= 0: the variables are first, first + 1, ..., first + nout - 1
  int v[Chemtable::kMaxOut];
};

// std::min / std::max (the reference's clip)
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }

// BinarySearch (CombustionGeneralDefinitions.cpp:211-245) returns, for x in
// [X[1], X[n]] (1-based), the i with X[i] <= x < X[i+1], and n-1 for x = X[n]:
// its bisection keeps X[i1] < x < X[i2] and returns at once where x equals a
// node, so the result is the largest i <= n-1 with X[i] <= x whatever path
// the bisection takes. Here that index is found by a branchless descent over
// X[0 .. n-2] (0-based), the three axes interleaved so their LDS reads overlap.
// Only for NaN do the two differ (the reference stops at its first mid, this
// at 0); the weights are NaN there and so is every interpolated value.
__device__ __forceinline__ int top_pow2(int m) {  // the largest power of two <= m (m >= 1)
  return 1 << (31 - __clz(m));
}

// Lookup's clip (ChemtableCartesianLinear_single.cpp:390-392), the interval
// on each axis, and Compute_Weights (:359-371)
__device__ __forceinline__ void locate(const double *ax, int n1, int n2, int n3, double a1, double a2, double a3,
                                       int idx[3], double w[3][2]) {
  const double *X[3] = {ax, ax + n1, ax + n1 + n2};
  const int n[3] = {n1, n2, n3};
  double a[3] = {a1, a2, a3};
  int step[3];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a[d] = dmin(dmax(a[d], X[d][0]), X[d][n[d] - 1]);
    idx[d] = 0;
    step[d] = top_pow2(n[d] - 1);
  }
  int s = step[0] > step[1] ? step[0] : step[1];
  s = s > step[2] ? s : step[2];
  for (; s > 0; s >>= 1) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int c = idx[d] + s;
      if (s <= step[d] && c <= n[d] - 2 && X[d][c] <= a[d]) idx[d] = c;
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const double x0 = X[d][idx[d]], xp1 = X[d][idx[d] + 1];
    const double xp1x0 = xp1 - x0;
    w[d][0] = (xp1 - a[d]) / xp1x0;
    w[d][1] = (a[d] - x0) / xp1x0;
  }
}

// One thread per point. A contiguous run of variables (Sel::first >= 0) is
// interpolated one 16-byte vector of four variables at a time: eight float4
// gathers per vector, few registers, many waves; any other list one variable
// at a time. The sums follow Lookup / LookupCoeff (:414-428, 469-483):
// f = 0 + w0[0] d(i) + w0[1] d(i+1) for each (j, k), g = 0 + w1[0] f(j=0) +
// w1[1] f(j=1) for each k, value = 0 + w2[0] g(k=0) + w2[1] g(k=1).
template <bool LDS>
__global__ void __launch_bounds__(256) k_chem_lookup(TabDev T, int64_t n, const double *__restrict__ zm,
                                                     const double *__restrict__ zv, const double *__restrict__ pc,
                                                     Sel S, double *__restrict__ out) {
  extern __shared__ double ax_lds[];
  const double *ax = T.ax;
  if (LDS) {
    const int na = T.n1 + T.n2 + T.n3;
    for (int t = threadIdx.x; t < na; t += blockDim.x) ax_lds[t] = T.ax[t];
    __syncthreads();
    ax = ax_lds;
  }
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  int idx[3];
  double w[3][2];
  // the points and the results stream through once: non-temporal, so they do
  // not push the table's lines out of the caches
  locate(ax, T.n1, T.n2, T.n3, __builtin_nontemporal_load(zm + p), __builtin_nontemporal_load(zv + p),
         __builtin_nontemporal_load(pc + p), idx, w);
  const float *node = T.data + (size_t)idx[0] * T.s1 + (size_t)idx[1] * T.s2 + (size_t)idx[2] * T.nvp;
  if (S.first >= 0) {
    const int c0 = S.first >> 2, c1 = (S.first + S.nout - 1) >> 2;
#pragma unroll 1
    for (int c = c0; c <= c1; c++) {
      double val[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int kk = 0; kk < 2; kk++) {
        double g[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int jj = 0; jj < 2; jj++) {
          const float4 x = reinterpret_cast<const float4 *>(node + jj * T.s2 + kk * T.nvp)[c];
          const float4 y = reinterpret_cast<const float4 *>(node + T.s1 + jj * T.s2 + kk * T.nvp)[c];
          const float da[4] = {x.x, x.y, x.z, x.w}, db[4] = {y.x, y.y, y.z, y.w};
#pragma unroll
          for (int e = 0; e < 4; e++) {
            double f = 0.0;
            f += w[0][0] * (double)da[e];
            f += w[0][1] * (double)db[e];
            g[e] += w[1][jj] * f;
          }
        }
#pragma unroll
        for (int e = 0; e < 4; e++) val[e] += w[2][kk] * g[e];
      }
#pragma unroll
      for (int e = 0; e < 4; e++) {
        const int q = 4 * c + e - S.first;
        if (q >= 0 && q < S.nout) __builtin_nontemporal_store(val[e], out + (size_t)q * n + p);
      }
    }
  } else {
    for (int q = 0; q < S.nout; q++) {
      const int l = S.v[q];
      double val = 0.0;
      for (int kk = 0; kk < 2; kk++) {
        double g = 0.0;
        for (int jj = 0; jj < 2; jj++) {
          const float *a = node + jj * T.s2 + kk * T.nvp + l;
          double f = 0.0;
          f += w[0][0] * (double)a[0];
          f += w[0][1] * (double)a[T.s1];
          g += w[1][jj] * f;
        }
        val += w[2][kk] * g;
      }
      __builtin_nontemporal_store(val, out + (size_t)q * n + p);
    }
  }
}

std::string chem_str(const char *b) {
  char s[64];
  std::memcpy(s, b, 64);
  s[63] = '\0';  // the reference keeps what precedes the first NUL of the 63 (Load, :73-91)
  return std::string(s);
}

}  // namespace

// Load (ChemtableCartesianLinear_single.cpp:19-246): version 2.0 (double),
// combustion model and table type (64 characters each), reference pressure,
// oxidizer and fuel temperatures (doubles), n1, n2, n3, nvar (ints), the three
// axes (doubles), nvar names (64 characters each), then the nodes as floats
// in i, j, k, variable order.
ChemtableHost read_chemtable(const std::string &path) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) throw std::invalid_argument("cannot open the chemistry table " + path);
  std::unique_ptr<FILE, int (*)(FILE *)> hold(fp, std::fclose);
  auto rd = [&](void *p, size_t size, size_t count, const char *what) {
    if (std::fread(p, size, count, fp) != count)
      throw std::invalid_argument("chemistry table " + path + ": file ends in " + what);
  };
  ChemtableHost h;
  double version = 0.0;
  rd(&version, sizeof(double), 1, "the version");
  if (version != 2.0)
    throw std::invalid_argument("chemistry table " + path + ": version " + std::to_string(version) +
                                ", the lookup reads version 2.0");
  char b[64];
  rd(b, 1, 64, "the combustion model");
  h.model = chem_str(b);
  rd(b, 1, 64, "the table type");
  h.type = chem_str(b);
  rd(&h.pref, sizeof(double), 1, "Preference");
  rd(&h.toxi, sizeof(double), 1, "Toxi");
  rd(&h.tfuel, sizeof(double), 1, "Tfuel");
  int dims[4];
  rd(dims, sizeof(int), 4, "the table dimensions");
  for (int d = 0; d < 4; d++)
    if (dims[d] < 1 || dims[d] > (1 << 24))
      throw std::invalid_argument("chemistry table " + path + ": dimension " + std::to_string(dims[d]));
  h.n1 = dims[0], h.n2 = dims[1], h.n3 = dims[2], h.nvar = dims[3];
  // the file must hold what the header announces before anything that size is allocated
  const long here = std::ftell(fp);
  std::fseek(fp, 0, SEEK_END);
  const long end = std::ftell(fp);
  std::fseek(fp, here, SEEK_SET);
  const double need = 8.0 * (h.n1 + h.n2 + h.n3) + 64.0 * h.nvar + 4.0 * h.n1 * h.n2 * (double)h.n3 * h.nvar;
  if ((double)(end - here) < need)
    throw std::invalid_argument("chemistry table " + path + ": the file is shorter than its " + std::to_string(h.n1) +
                                " x " + std::to_string(h.n2) + " x " + std::to_string(h.n3) + " x " +
                                std::to_string(h.nvar) + " table");
  h.x1.resize(h.n1), h.x2.resize(h.n2), h.x3.resize(h.n3);
  rd(h.x1.data(), sizeof(double), h.n1, "x1");
  rd(h.x2.data(), sizeof(double), h.n2, "x2");
  rd(h.x3.data(), sizeof(double), h.n3, "x3");
  for (int l = 0; l < h.nvar; l++) {
    rd(b, 1, 64, "the variable names");
    h.names.push_back(chem_str(b));
  }
  h.data.resize((size_t)h.n1 * h.n2 * h.n3 * h.nvar);
  rd(h.data.data(), sizeof(float), h.data.size(), "the table data");
  return h;
}

Chemtable::Chemtable(ChemtableHost h) : h_(std::move(h)) {
  const int dims[3] = {h_.n1, h_.n2, h_.n3};
  const std::vector<double> *axes[3] = {&h_.x1, &h_.x2, &h_.x3};
  for (int d = 0; d < 3; d++) {
    if (dims[d] < 2) throw ArgError("chemistry table: each axis needs at least two values (x" + std::to_string(d + 1) + ")");
    if ((int)axes[d]->size() != dims[d]) throw ArgError("chemistry table: axis size");
    for (int q = 0; q < dims[d]; q++) {
      const double x = (*axes[d])[q];
      if (!std::isfinite(x) || (q > 0 && !(x > (*axes[d])[q - 1])))
        throw ArgError("chemistry table: axis x" + std::to_string(d + 1) +
                       " must be finite and strictly increasing (value " + std::to_string(q) + ")");
    }
  }
  if (h_.nvar < 1) throw ArgError("chemistry table: no variables");
  if ((int)h_.names.size() != h_.nvar) throw ArgError("chemistry table: one name per variable");
  std::set<std::string> seen;
  for (const auto &s : h_.names)  // Load refuses a name twice (:204-209)
    if (!seen.insert(s).second) throw ArgError("chemistry table: variable " + s + " appears twice");
  const size_t nodes = (size_t)h_.n1 * h_.n2 * h_.n3;
  if (h_.data.size() != nodes * h_.nvar) throw ArgError("chemistry table: data size");
  nvp_ = (h_.nvar + 3) & ~3;
  std::vector<float> dev(nodes * nvp_, 0.0f);
  for (size_t q = 0; q < nodes; q++)
    std::memcpy(&dev[q * nvp_], &h_.data[q * h_.nvar], sizeof(float) * h_.nvar);
  std::vector<double> ax;
  ax.insert(ax.end(), h_.x1.begin(), h_.x1.end());
  ax.insert(ax.end(), h_.x2.begin(), h_.x2.end());
  ax.insert(ax.end(), h_.x3.begin(), h_.x3.end());
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::runtime_error("HIP device: none found");
  try {
    CTCHECK(hipMalloc((void **)&data_d_, sizeof(float) * dev.size()));
    CTCHECK(hipMalloc((void **)&ax_d_, sizeof(double) * ax.size()));
    CTCHECK(hipMemcpy(data_d_, dev.data(), sizeof(float) * dev.size(), hipMemcpyHostToDevice));
    CTCHECK(hipMemcpy(ax_d_, ax.data(), sizeof(double) * ax.size(), hipMemcpyHostToDevice));
  } catch (...) {
    if (data_d_) (void)hipFree(data_d_);
    if (ax_d_) (void)hipFree(ax_d_);
    throw;
  }
  h_.data.clear();
  h_.data.shrink_to_fit();
}

Chemtable::~Chemtable() {
  if (data_d_) (void)hipFree(data_d_);
  if (ax_d_) (void)hipFree(ax_d_);
}

int Chemtable::var_index(const std::string &name) const {
  for (int l = 0; l < h_.nvar; l++)
    if (h_.names[l] == name) return l;
  return -1;
}

void Chemtable::lookup(int64_t n, const double *zm, const double *zv, const double *pc, int nout, const int *vars,
                       double *out, void *stream) const {
  if (n < 0) throw ArgError("chemistry lookup: n < 0");
  if (nout < 1 || nout > kMaxOut) throw ArgError("chemistry lookup: nout must be 1.." + std::to_string(kMaxOut));
  if (!vars) throw ArgError("chemistry lookup: no variables");
  Sel S{};
  S.nout = nout;
  bool contig = true;
  for (int q = 0; q < nout; q++) {
    if (vars[q] < 0 || vars[q] >= h_.nvar)
      throw ArgError("chemistry lookup: variable " + std::to_string(vars[q]) + " is not in the table (nvar " +
                     std::to_string(h_.nvar) + ")");
    S.v[q] = vars[q];
    if (vars[q] != vars[0] + q) contig = false;
  }
  if (n == 0) return;
  if (!zm || !zv || !pc || !out) throw ArgError("chemistry lookup: null device pointer");
  S.first = contig ? vars[0] : -1;
  TabDev T{ax_d_, data_d_, h_.n1, h_.n2, h_.n3, nvp_, (size_t)h_.n2 * h_.n3 * nvp_, (size_t)h_.n3 * nvp_};
  const int64_t nblk = (n + 255) / 256;
  if (nblk > 0x7fffffff) throw ArgError("chemistry lookup: too many points");
  const int na = h_.n1 + h_.n2 + h_.n3;
  const hipStream_t st = static_cast<hipStream_t>(stream);
  if (na <= kAxisLds)
    hipLaunchKernelGGL(k_chem_lookup<true>, dim3((unsigned)nblk), dim3(256), sizeof(double) * na, st, T, n, zm, zv,
                       pc, S, out);
  else
    hipLaunchKernelGGL(k_chem_lookup<false>, dim3((unsigned)nblk), dim3(256), 0, st, T, n, zm, zv, pc, S, out);
  CTCHECK(hipGetLastError());
}

void Chemtable::lookup_host(int64_t n, const double *zm, const double *zv, const double *pc, int nout,
                            const int *vars, double *out) const {
  if (n < 0) throw ArgError("chemistry lookup: n < 0");
  if (n == 0) {
    lookup(0, nullptr, nullptr, nullptr, nout, vars, nullptr, nullptr);  // checks the selection
    return;
  }
  double *d = nullptr;
  const size_t bytes = sizeof(double) * (size_t)n;
  CTCHECK(hipMalloc((void **)&d, bytes * (3 + (size_t)std::max(nout, 0))));
  try {
    CTCHECK(hipMemcpy(d, zm, bytes, hipMemcpyHostToDevice));
    CTCHECK(hipMemcpy(d + n, zv, bytes, hipMemcpyHostToDevice));
    CTCHECK(hipMemcpy(d + 2 * n, pc, bytes, hipMemcpyHostToDevice));
    lookup(n, d, d + n, d + 2 * n, nout, vars, d + 3 * n, nullptr);
    CTCHECK(hipMemcpy(out, d + 3 * n, bytes * nout, hipMemcpyDeviceToHost));
  } catch (...) {
    (void)hipFree(d);
    throw;
  }
  CTCHECK(hipFree(d));
}

double Chemtable::time_lookup(int64_t n, const double *zm, const double *zv, const double *pc, int nout,
                              const int *vars, double *out, void *stream, int reps) const {
  if (reps < 1) throw ArgError("chemistry lookup timing: reps < 1");
  const hipStream_t st = static_cast<hipStream_t>(stream);
  hipEvent_t e0, e1;
  CTCHECK(hipEventCreate(&e0));
  CTCHECK(hipEventCreate(&e1));
  float ms = 0.0f;
  try {
    lookup(n, zm, zv, pc, nout, vars, out, stream);  // untimed first launch
    CTCHECK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; r++) lookup(n, zm, zv, pc, nout, vars, out, stream);
    CTCHECK(hipEventRecord(e1, st));
    CTCHECK(hipEventSynchronize(e1));
    CTCHECK(hipEventElapsedTime(&ms, e0, e1));
  } catch (...) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return (double)ms / reps;
}

}  // namespace jm
