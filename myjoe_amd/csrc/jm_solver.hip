// Device side of the steady-solve hot path: data layout in HBM, kernels and
// the per-solve context. Compiled for gfx950 with -ffp-contract=off so the
// FP64 residual path reproduces the reference's arithmetic bit for bit.
//
// HBM layout (all per-solve, resident for the life of the context):
//   q      [5][ncv]   conservative state, SoA (streams)
//   prim   [ncv][8]   rho u v w p | x_cv (gathered by neighbours: 64 B rows)
//   grad   [ncv][16]  limited grad rho(3), grad p(3), grad u(9), alpha_rho
//   bfa    [nfa_b][8] boundary-face state rho u v w p T h
//   fgeo   [nfa][8]   unit normal, area, face centroid (64 B rows)
//   cfl    per cell, the cell's faces in the reference's accumulation order
//          (internal faces ascending, then boundary faces ascending) as
//          {face | side << 31, neighbour cell or -(zone + 1)}
//   A      [nnz][5][5] block Jacobian in nbocv order (FP64 or FP32)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "jm_ctx.hpp"
#include "jm_param.hpp"
#include "jm_physics.hpp"

#define HIPCHECK(x)                                                                                     \
  do {                                                                                                  \
    hipError_t e_ = (x);                                                                                \
    if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + \
                                                   " at " + #x);                                       \
  } while (0)

namespace jm {

struct ZoneDev {
  int bc;
  int skind;   // first scalar's boundary (nuSA, or SST's kine): 0 zero gradient, 1 "<zone>.<name> = sval",
               // 3 nuSA = PROFILES (the face's value from the profile file)
  int skind2;  // SST omega: 0 zero gradient, 1 "<zone>.omega = sval2", 2 HOOK (wall value)
  int pad;
  double par[8];
  double sval, sval2;
};

struct DMesh {
  int ncv, nfa, nfa_b, nnz;
  const double *fgeo;     // [nfa][8]
  const double *fnorm;    // [nfa][3] area-weighted normal (Green-Gauss or Spalart-Allmaras)
  const int *cfl_i;       // [ncv+1]
  const int2 *cfl;        // [entries]
  const int *cfl_noc;     // [entries] off-diagonal block; bit 30 = accumulate
  const int *nbocv_i;     // [ncv+1]
  const int *nbocv_v;     // [nnz]
  const double *lsg;      // [nnz][3]
  const double *bfa_lsg;  // [nfa_b][3]
  const int *bfa_cell;    // [nfa_b]
  const int *bfa_zone;    // [nfa_b]
  const double *vol;      // [ncv]
  const ZoneDev *zones;
  const int2 *ecp;        // [entries] cell of each entry and its position in the cell's list
  const double *cfl_r0;   // [entries][3] x_fa - x_cv
  const int *wave_e0;     // [nwave+1] first entry of each 64-lane wavefront
  int nwave;
  const int *foc;         // [entries] entry offset of the cell's k-th face in faocv order (Spalart-Allmaras)
  const int *cell_perm;   // [ncv] cell of each thread of the per-cell gradient, or null (thread t -> cell t)
};

struct DParams {
  double gamma, R, eps, g[3];
  int scheme, second_order, grad_ls, limiter, symm_bc, gravity_on;
  // viscous terms (MU_REF > 0): viscosity law, conductivity, gradient limiting
  int visc, visc_mode, visc_limiter, pad;
  double mu_ref, mu_power_law, sl_tref, sl_sref, T_ref, pr, pr_turb;
  // Spalart-Allmaras: nuSA [ncv] and per-cell model fields [ncv][8] =
  // grad nuSA (3), face-averaged mul, strain and vorticity magnitudes, wall distance
  int sa, pad2;
  const double *sa_phi, *sa_aux;
  const double *sa_prof;  // [nfa_b] nuSA of "<zone>.nuSA = PROFILES" faces, from the profile file
  // Menter SST k-omega: kine, omega [2][ncv], per-cell model fields
  // [ncv][kSstN], boundary values [nfa_b][2]; SST_LIMIT_PK, SST_LIMITER_SHEARSTRESS
  int sst, sst_limit_pk, sst_lim_shear, pad5;
  const double *sst_phi, *sst_aux, *sst_bfa;
  double sst_alfa[2];  // alfa_1, alfa_2 (TurbModel_KOMSST.h:24-25)
  // parity readout: when set, the per-entry flux kernel writes each face's
  // inviscid flux [nfa][5] (device face order) once, from the face's first cell
  double *fdump;
  // when set, the viscous per-entry flux kernel writes each face's laminar
  // and eddy viscosity [nfa][2] (device face order), as it used them
  double *vdump;
};

// per-cell Spalart-Allmaras fields in DParams::sa_aux
enum SaAux { kAuxGrad = 0, kAuxMue = 3, kAuxStr = 4, kAuxVort = 5, kAuxWall = 6, kAuxN = 8 };
// per-cell SST fields in DParams::sst_aux: limited kine and omega gradients,
// face-averaged mul, strain and vorticity magnitudes, divergence, cross
// diffusion, blending functions F1 and F2, wall distance
enum SstAux {
  kSGradK = 0, kSGradW = 3, kSMue = 6, kSStr = 7, kSVort = 8, kSDiv = 9, kSCd = 10, kSF1 = 11, kSF2 = 12,
  kSWall = 13, kSstN = 16
};

// per-solve scalars of the BiCGStab and of reductions (device resident)
struct SolverScalars {
  double rho, rho_prime, alpha, omega, beta, gamma;
  double res0_max, res_max;
  double dt_min;
  double resid[5];
  double resid_l2[5];  // L2 norms of the same residual
  double resid_s;  // L1 of the transported scalar's residual (Spalart-Allmaras; SST's kine)
  double resid_s2;  // SST: omega's
  int iter, done, maxiter, pad;
  double zero_abs, zero_rel;
  // pseudo-time step counter and CFL of the device-driven (graph) step
  int step_no, stop;  // stop: the host-polled loop's launch budget is spent
  double cfl_cur;
  int max_launch, pad3;
};

// what a device-driven step reports back, written to host-mapped memory
struct StepMail {
  int step, lin_iters, lin_status, solver_stop;
  double cfl, dt_min, resid[5], resid_l2[5];
};

namespace {

constexpr int kBlock = 256;
constexpr int kRedSlots = 8;  // partial-sum slots per block for fused reductions
inline int nblk(long long n, int b = kBlock) { return (int)((n + b - 1) / b); }

// A block's index and the grid size within one solve's launch: the hardware
// values for an ordinary launch, the solve's own share of a batched one.
struct Blk {
  int id, n;
};
__device__ __forceinline__ Blk hw_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

__device__ __forceinline__ double2 ld2(const double *p) { return *reinterpret_cast<const double2 *>(p); }

// Workgroups are dealt round-robin to the 8 XCDs (each with its own L2);
// renumber them so every XCD walks one contiguous range of the schedule and
// neighbouring cells share an L2.
__device__ __forceinline__ int xcd_block(int b, int nb) {
  const int X = 8;
  const int per = nb / X, rem = nb % X, x = b % X;
  return x * per + (x < rem ? x : rem) + b / X;
}

// ---------------------------------------------------------------------------
// calcStateVariables, UgpWithCvCompFlow.h:1662-1694 (gamma, R constant)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void d_state(const Blk &B, int ncv, const double *__restrict__ q, double *__restrict__ prim,
                                                  double gam, int *__restrict__ negp) {
  int c = B.id * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  double rho = q[c], r0 = q[ncv + c], r1 = q[2 * ncv + c], r2 = q[3 * ncv + c], rE = q[4 * ncv + c];
  double *P = prim + 8 * (size_t)c;
  double d = 0.0;
  d += r0 * r0;
  d += r1 * r1;
  d += r2 * r2;
  double pr = (gam - 1.0) * (rE - 0.5 * d / rho);
  double p = P[4];
  if (pr > 0.0)
    p = pr;
  else
    atomicAdd(negp, 1);
  P[0] = rho;
  P[1] = r0 / rho;
  P[2] = r1 / rho;
  P[3] = r2 / rho;
  P[4] = p;
}

__global__ void __launch_bounds__(kBlock) k_state(int ncv, const double *__restrict__ q, double *__restrict__ prim,
                                                  double gam, int *__restrict__ negp) {
  d_state(hw_blk(), ncv, q, prim, gam, negp);
}

// calcStateVariables with SST's kinetic energy in the pressure
// (UgpWithCvCompFlow.h:1672-1676)
__global__ void __launch_bounds__(kBlock) k_state_sst(int ncv, const double *__restrict__ q, double *__restrict__ prim,
                                                      double gam, int *__restrict__ negp,
                                                      const double *__restrict__ kine) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  double rho = q[c], r0 = q[ncv + c], r1 = q[2 * ncv + c], r2 = q[3 * ncv + c], rE = q[4 * ncv + c];
  double *P = prim + 8 * (size_t)c;
  double d = 0.0;
  d += r0 * r0;
  d += r1 * r1;
  d += r2 * r2;
  double pr = (gam - 1.0) * (rE - 0.5 * d / rho - rho * kine[c]);
  double p = P[4];
  if (pr > 0.0)
    p = pr;
  else
    atomicAdd(negp, 1);
  P[0] = rho;
  P[1] = r0 / rho;
  P[2] = r1 / rho;
  P[3] = r2 / rho;
  P[4] = p;
}

// ---------------------------------------------------------------------------
// setBC, JoeWithModels.cpp:3857-4300 with ComputeBCProperties_T/_H
// (UgpWithCvCompFlow.h:1783-1839) and the EOS density (:4292-4293)
// ---------------------------------------------------------------------------
// the characteristic far-field state of one boundary face (FARFIELD,
// JoeWithModels.cpp:4100-4183; PROFILES, :4190-4270)
__device__ __forceinline__ void farfield_state(const double u_bc[3], double p_bc, double rho_bc, double gamma_bc,
                                               const double nVec[3], const double vel[3], double rho, double p,
                                               double GAMMA, double R_gas, double vb[3], double &p_b, double &T_b) {
  double gm1 = gamma_bc - 1.0;
  double ovgm1 = 1.0 / gm1;
  double c_bc = sqrt(gamma_bc * p_bc / rho_bc);
  double s_bc = pow(rho_bc, gamma_bc) / p_bc;
  double qn_bc = dot3(nVec, u_bc);
  double vn_bc = qn_bc;
  double qne = dot3(nVec, vel);
  double ce = sqrt(GAMMA * p / rho);
  double ac1 = vn_bc > -c_bc ? qne + 2.0 * ovgm1 * ce : qn_bc + 2.0 * ovgm1 * c_bc;
  double ac2 = vn_bc > c_bc ? qne - 2.0 * ovgm1 * ce : qn_bc - 2.0 * ovgm1 * c_bc;
  double qnf = 0.5 * (ac1 + ac2);
  double cf = 0.25 * (ac1 - ac2) * gm1;
  double sf;
  if (vn_bc > 0) {
    for (int i = 0; i < 3; i++) vb[i] = vel[i] + (qnf - qne) * nVec[i];
    sf = pow(rho, GAMMA) / p;
  } else {
    for (int i = 0; i < 3; i++) vb[i] = u_bc[i] + (qnf - qn_bc) * nVec[i];
    sf = s_bc;
  }
  double rb = pow((sf * cf * cf / GAMMA), ovgm1);
  p_b = rb * cf * cf / GAMMA;
  T_b = p_b / (R_gas * rb);
}

__device__ __forceinline__ void d_bc(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                               const double *__restrict__ hook, double *__restrict__ bfa) {
  int f = B.id * blockDim.x + threadIdx.x;
  if (f >= m.nfa_b) return;
  const ZoneDev &z = m.zones[m.bfa_zone[f]];
  const int c = m.bfa_cell[f];
  const double *pc = prim + 8 * (size_t)c;
  const double rho = pc[0], p = pc[4];
  const double vel[3] = {pc[1], pc[2], pc[3]};
  const double GAMMA = P.gamma, R_gas = P.R;
  const double temp = p / (rho * R_gas);
  const double *geo = m.fgeo + 8 * (size_t)f;
  const double nVec[3] = {geo[0], geo[1], geo[2]};
  double T_b = 0.0, p_b = 0.0, h_b, vb[3] = {0.0, 0.0, 0.0};
  bool h_set = false;
  switch (z.bc) {
    case kBcCbc: {
      const double u_bc[3] = {z.par[0], z.par[1], z.par[2]};
      if (dot3(u_bc, nVec) > 0.0) {
        double velMagN = dot3(vel, nVec);
        double sos = sqrt(GAMMA * p / rho);
        double mach = fabs(velMagN) / sos;
        T_b = temp;
        vb[0] = vel[0]; vb[1] = vel[1]; vb[2] = vel[2];
        p_b = mach >= 1.0 ? p : z.par[4];
      } else {
        T_b = z.par[3];
        vb[0] = u_bc[0]; vb[1] = u_bc[1]; vb[2] = u_bc[2];
        p_b = z.par[4];
      }
      break;
    }
    case kBcCbcSubsonicInlet: {
      double wPow2 = dot3(vel, vel);
      double v = sqrt(wPow2);
      vb[0] = z.par[0] * v; vb[1] = z.par[1] * v; vb[2] = z.par[2] * v;
      double Ttot = z.par[3], ptot = z.par[4];
      h_b = GAMMA * R_gas / (GAMMA - 1.0) * Ttot;
      double w2 = dot3(vb, vb);
      h_b -= 0.5 * w2;
      T_b = h_b * (GAMMA - 1.0) / (GAMMA * R_gas);
      p_b = ptot * pow(T_b / Ttot, GAMMA / (GAMMA - 1.0));
      h_set = true;
      break;
    }
    case kBcCbcSubsonicOutlet:
      p_b = z.par[0];
      vb[0] = vel[0]; vb[1] = vel[1]; vb[2] = vel[2];
      T_b = temp;
      break;
    case kBcSymmetry: {
      double un = dot3(nVec, vel);
      for (int i = 0; i < 3; i++) vb[i] = vel[i] - 1.0 * un * nVec[i];
      T_b = temp;
      p_b = p;
      break;
    }
    case kBcNeumann:
      vb[0] = vel[0]; vb[1] = vel[1]; vb[2] = vel[2];
      T_b = temp;
      p_b = p;
      break;
    case kBcWall:
      T_b = z.par[0] > 0.0 ? z.par[0] : temp;
      p_b = p;
      break;
    case kBcFarfield: {
      const double u_bc[3] = {z.par[0], z.par[1], z.par[2]};
      farfield_state(u_bc, z.par[3], z.par[4], z.par[5], nVec, vel, rho, p, GAMMA, R_gas, vb, p_b, T_b);
      break;
    }
    case kBcProfiles: {  // JoeWithModels.cpp:4190-4270: the profile's rho, u, p and gamma_bc = 1.4
      const double *h = hook + 5 * (size_t)f;
      const double u_bc[3] = {h[1], h[2], h[3]};
      farfield_state(u_bc, h[4], h[0], 1.4, nVec, vel, rho, p, GAMMA, R_gas, vb, p_b, T_b);
      break;
    }
    case kBcHook: {
      const double *h = hook + 5 * (size_t)f;
      T_b = h[0];
      vb[0] = h[1]; vb[1] = h[2]; vb[2] = h[3];
      p_b = h[4];
      break;
    }
    default:
      break;
  }
  if (!h_set) h_b = GAMMA * R_gas / (GAMMA - 1.0) * T_b;
  // setScalarBC (Scalars.cpp:289-360): a "<zone>.nuSA" value, or with
  // PROFILES the face's profile value, on inflow faces (vel_bfa . fa_normal
  // <= 1e-8), the cell value otherwise
  double phi_b = 0.0;
  if (P.sa) {
    const double *fa_n = m.fnorm + 3 * (size_t)f;
    phi_b = (z.skind && !(dot3(vb, fa_n) > 1.0e-8)) ? (z.skind == 3 ? P.sa_prof[f] : z.sval) : P.sa_phi[c];
  }
  double *o = bfa + 8 * (size_t)f;
  o[0] = p_b / (R_gas * T_b);
  o[1] = vb[0];
  o[2] = vb[1];
  o[3] = vb[2];
  o[4] = p_b;
  o[5] = T_b;
  o[6] = h_b;
  o[7] = phi_b;
}

__global__ void __launch_bounds__(kBlock) k_bc(DMesh m, DParams P, const double *__restrict__ prim,
                                               const double *__restrict__ hook, double *__restrict__ bfa) {
  d_bc(hw_blk(), m, P, prim, hook, bfa);
}

// smooth Barth-Jespersen factor of one face (UgpWithCv.h:913-925, BJ undefined)
__device__ __forceinline__ double bj_face(double alfa, double phi, double phifa, double phiMax, double phiMin,
                                          double eps2) {
  double dm = phifa - phi;
  double dp = (phifa - phi) >= 0.0 ? phiMax - phi : phiMin - phi;
  return Dev::mn(alfa, (dp * dp + 2.0 * dm * dp + eps2) / (dp * dp + 2 * dm * dm + dm * dp + eps2));
}

// ---------------------------------------------------------------------------
// Gradients of u (vector), rho and p with boundary values, and their smooth
// Barth-Jespersen limiters: calcCvVectorGrad / calcCvScalarGrad
// (UgpWithCv.h:803-872), least squares (:1009-1083) or Green-Gauss
// (:1091-1188), limiters (:881-1000). One thread per cell; the cell's sums
// follow the reference's order (diagonal, nbocv neighbours, boundary faces).
// ---------------------------------------------------------------------------
// enthalpy of a cell state (calcStateVariables: T = p / (rho R),
// h = gamma R / (gamma - 1) T; UgpWithCvCompFlow.h:1680-1690)
// entries of a cell whose loads the per-cell gradient issues together
constexpr int kGradChunk = 4;

__device__ __forceinline__ double enthalpy_of(const DParams &P, double rho, double p) {
  const double T = p / (rho * P.R);
  return P.gamma * P.R / (P.gamma - 1.0) * T;
}

// With VISC the kernel also forms the enthalpy gradient and the velocity and
// enthalpy gradients the viscous flux uses (calcViscousFluxNS,
// JoeWithModels.cpp:3616-3630): limited when VISC_LIMITER and LIMITER are on,
// raw otherwise; written to gradv[ncv][12] (grad u 9, grad h 3).
template <bool LS, bool VISC>
__device__ __forceinline__ void d_grad(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                                 const double *__restrict__ bfa, double *__restrict__ grad,
                                                 double *__restrict__ gradv) {
  const int t = xcd_block(B.id, B.n) * blockDim.x + threadIdx.x;
  if (t >= m.ncv) return;
  const int c = m.cell_perm ? m.cell_perm[t] : t;
  const double *pc = prim + 8 * (size_t)c;
  const double rho = pc[0], u[3] = {pc[1], pc[2], pc[3]}, p = pc[4];
  const double xcv[3] = {pc[5], pc[6], pc[7]};
  double gr[3], gp[3], gu[3][3], gh[3] = {0.0, 0.0, 0.0};
  double rmax = rho, rmin = rho, pmax = p, pmin = p;
  double umax[3] = {u[0], u[1], u[2]}, umin[3] = {u[0], u[1], u[2]};
  const double h = VISC ? enthalpy_of(P, rho, p) : 0.0;
  double hmax = h, hmin = h;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  int e_int = e0;  // end of the internal entries (they come first)
  if (LS) {
    const int noc_f = m.nbocv_i[c];
    const double *l0 = m.lsg + 3 * (size_t)noc_f;
    for (int j = 0; j < 3; j++) {
      gr[j] = l0[j] * rho;
      gp[j] = l0[j] * p;
      if (VISC) gh[j] = l0[j] * h;
    }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) gu[i][j] = l0[j] * u[i];
    // internal entries (they come first) in chunks of kGradChunk: a chunk's
    // entries, then its neighbour rows and coefficients, are loaded together
    // (slots past the internal ones read the cell's own row, always valid)
    // and added in entry order
    const int noc_l = m.nbocv_i[c + 1] - 1;
    bool open = true;
    for (int eb = e0; open && eb < e1; eb += kGradChunk) {
      int2 en[kGradChunk];
#pragma unroll
      for (int q = 0; q < kGradChunk; q++) en[q] = eb + q < e1 ? m.cfl[eb + q] : make_int2(0, -1);
      double2 a[kGradChunk], b[kGradChunk];
      double pv[kGradChunk], l[kGradChunk][3];
#pragma unroll
      for (int q = 0; q < kGradChunk; q++) {
        const double *pn = prim + 8 * (size_t)(en[q].y >= 0 ? en[q].y : c);
        a[q] = ld2(pn);
        b[q] = ld2(pn + 2);
        pv[q] = pn[4];
        const double *lq = m.lsg + 3 * (size_t)min(noc_f + 1 + (eb - e0) + q, noc_l);
        l[q][0] = lq[0];
        l[q][1] = lq[1];
        l[q][2] = lq[2];
      }
#pragma unroll
      for (int q = 0; q < kGradChunk; q++) {
        open = open && en[q].y >= 0;
        if (!open) break;
        const double rn = a[q].x, un[3] = {a[q].y, b[q].x, b[q].y}, pnv = pv[q];
        const double l_0 = l[q][0], l_1 = l[q][1], l_2 = l[q][2];
        gr[0] += l_0 * rn; gr[1] += l_1 * rn; gr[2] += l_2 * rn;
        gp[0] += l_0 * pnv; gp[1] += l_1 * pnv; gp[2] += l_2 * pnv;
        if (VISC) {
          const double hn = enthalpy_of(P, rn, pnv);
          gh[0] += l_0 * hn; gh[1] += l_1 * hn; gh[2] += l_2 * hn;
          hmax = Dev::mx(hmax, hn); hmin = Dev::mn(hmin, hn);
        }
        for (int i = 0; i < 3; i++) {
          gu[i][0] += l_0 * un[i];
          gu[i][1] += l_1 * un[i];
          gu[i][2] += l_2 * un[i];
          umax[i] = Dev::mx(umax[i], un[i]);
          umin[i] = Dev::mn(umin[i], un[i]);
        }
        rmax = Dev::mx(rmax, rn); rmin = Dev::mn(rmin, rn);
        pmax = Dev::mx(pmax, pnv); pmin = Dev::mn(pmin, pnv);
        e_int = eb + q + 1;
      }
    }
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      if (en.y >= 0) continue;
      int f = en.x & 0x7fffffff;
      const double *bl = m.bfa_lsg + 3 * (size_t)f;
      const double *bb = bfa + 8 * (size_t)f;
      double b0 = bl[0], b1 = bl[1], b2 = bl[2];
      double rb = bb[0], ub[3] = {bb[1], bb[2], bb[3]}, pb = bb[4];
      for (int i = 0; i < 3; i++) {
        gu[i][0] += b0 * ub[i];
        gu[i][1] += b1 * ub[i];
        gu[i][2] += b2 * ub[i];
      }
      gr[0] += b0 * rb; gr[1] += b1 * rb; gr[2] += b2 * rb;
      gp[0] += b0 * pb; gp[1] += b1 * pb; gp[2] += b2 * pb;
      if (VISC) {
        const double hb = bb[6];
        gh[0] += b0 * hb; gh[1] += b1 * hb; gh[2] += b2 * hb;
      }
    }
  } else {
    for (int j = 0; j < 3; j++) {
      gr[j] = 0.0;
      gp[j] = 0.0;
      gu[0][j] = gu[1][j] = gu[2][j] = 0.0;
    }
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      int f = en.x & 0x7fffffff;
      const double *fn = m.fnorm + 3 * (size_t)f;
      double n[3] = {fn[0], fn[1], fn[2]};
      if (en.y >= 0) {
        const double *pn = prim + 8 * (size_t)en.y;
        double rn = pn[0], un[3] = {pn[1], pn[2], pn[3]}, pnv = pn[4];
        // distance weights (GG_WITH_DIST, UgpWithCv.h:37, :1104-1110)
        const double *xf = m.fgeo + 8 * (size_t)f + 4;
        double ds[3], dn[3];
        for (int j = 0; j < 3; j++) {
          ds[j] = xf[j] - xcv[j];
          dn[j] = xf[j] - pn[5 + j];
        }
        const bool s0 = (en.x >= 0);  // side 0: this cell is icv0
        const double wself = sqrt(dot3(ds, ds)), wnb = sqrt(dot3(dn, dn));
        const double w0 = s0 ? wself : wnb, w1 = s0 ? wnb : wself;
        double fr = s0 ? (w1 * rho + w0 * rn) / (w0 + w1) : (w1 * rn + w0 * rho) / (w0 + w1);
        double fp = s0 ? (w1 * p + w0 * pnv) / (w0 + w1) : (w1 * pnv + w0 * p) / (w0 + w1);
        for (int j = 0; j < 3; j++) {
          if (s0) {
            gr[j] += fr * n[j];
            gp[j] += fp * n[j];
          } else {
            gr[j] -= fr * n[j];
            gp[j] -= fp * n[j];
          }
        }
        if (VISC) {
          const double hn = enthalpy_of(P, rn, pnv);
          const double fh = s0 ? (w1 * h + w0 * hn) / (w0 + w1) : (w1 * hn + w0 * h) / (w0 + w1);
          for (int j = 0; j < 3; j++) {
            if (s0)
              gh[j] += fh * n[j];
            else
              gh[j] -= fh * n[j];
          }
          hmax = Dev::mx(hmax, hn); hmin = Dev::mn(hmin, hn);
        }
        for (int i = 0; i < 3; i++) {
          double fu = s0 ? (w1 * u[i] + w0 * un[i]) / (w0 + w1) : (w1 * un[i] + w0 * u[i]) / (w0 + w1);
          for (int j = 0; j < 3; j++) {
            if (s0)
              gu[i][j] += fu * n[j];
            else
              gu[i][j] -= fu * n[j];
          }
          umax[i] = Dev::mx(umax[i], un[i]);
          umin[i] = Dev::mn(umin[i], un[i]);
        }
        rmax = Dev::mx(rmax, rn); rmin = Dev::mn(rmin, rn);
        pmax = Dev::mx(pmax, pnv); pmin = Dev::mn(pmin, pnv);
        e_int = e + 1;
      } else {
        const double *bb = bfa + 8 * (size_t)f;
        double rb = bb[0], ub[3] = {bb[1], bb[2], bb[3]}, pb = bb[4];
        for (int j = 0; j < 3; j++) {
          gr[j] += rb * n[j];
          gp[j] += pb * n[j];
          if (VISC) gh[j] += bb[6] * n[j];
        }
        for (int i = 0; i < 3; i++)
          for (int j = 0; j < 3; j++) gu[i][j] += ub[i] * n[j];
      }
    }
    const double vol = m.vol[c];
    for (int j = 0; j < 3; j++) {
      gr[j] /= vol;
      gp[j] /= vol;
      if (VISC) gh[j] /= vol;
    }
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) gu[i][j] /= vol;
  }

  double au[3] = {1.0, 1.0, 1.0}, ar = 1.0, ap = 1.0, ah = 1.0;
  if (P.limiter) {
    const double sos = sqrt(P.gamma * p / rho);
    double e2u = P.eps * sos;
    e2u = e2u * e2u;
    double e2r = Dev::mx(P.eps * fabs(rho), 1.0e-8);
    e2r = e2r * e2r;
    double e2p = Dev::mx(P.eps * fabs(p), 1.0e-8);
    e2p = e2p * e2p;
    bool lu[3] = {(umax[0] - umin[0]) > 1.0e-8, (umax[1] - umin[1]) > 1.0e-8, (umax[2] - umin[2]) > 1.0e-8};
    bool lr = (rmax - rmin) > 1.0e-12, lp = (pmax - pmin) > 1.0e-12;
    // face offsets in chunks loaded together (slots past e_int repeat the
    // last entry and are not applied)
    for (int eb = e0; eb < e_int; eb += kGradChunk) {
      double r0[kGradChunk][3];
#pragma unroll
      for (int q = 0; q < kGradChunk; q++) {
        const double *rr = m.cfl_r0 + 3 * (size_t)min(eb + q, e_int - 1);
        r0[q][0] = rr[0];
        r0[q][1] = rr[1];
        r0[q][2] = rr[2];
      }
#pragma unroll
      for (int q = 0; q < kGradChunk; q++) {
        if (eb + q >= e_int) break;
        for (int i = 0; i < 3; i++)
          if (lu[i]) au[i] = bj_face(au[i], u[i], u[i] + dot3(r0[q], gu[i]), umax[i], umin[i], e2u);
        if (lr) ar = bj_face(ar, rho, rho + dot3(r0[q], gr), rmax, rmin, e2r);
        if (lp) ap = bj_face(ap, p, p + dot3(r0[q], gp), pmax, pmin, e2p);
      }
    }
    for (int i = 0; i < 3; i++)
      if (!lu[i]) au[i] = 0.0;
    if (!lr) ar = 0.0;
    if (!lp) ap = 0.0;
    if (VISC && P.visc_limiter) {  // enthalpy, reference value h (JoeWithModels.cpp:3619)
      double e2h = Dev::mx(P.eps * fabs(h), 1.0e-8);
      e2h = e2h * e2h;
      if ((hmax - hmin) > 1.0e-12) {
        for (int e = e0; e < e_int; e++) {
          const double *rr = m.cfl_r0 + 3 * (size_t)e;
          double r0[3] = {rr[0], rr[1], rr[2]};
          ah = bj_face(ah, h, h + dot3(r0, gh), hmax, hmin, e2h);
        }
      } else {
        ah = 0.0;
      }
    }
  }
  if (VISC) {
    const bool lim = P.limiter && P.visc_limiter;
    double *v = gradv + 12 * (size_t)c;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) v[3 * i + j] = lim ? gu[i][j] * au[i] : gu[i][j];
    for (int j = 0; j < 3; j++) v[9 + j] = lim ? gh[j] * ah : gh[j];
  }
  double *o = grad + 16 * (size_t)c;
  o[0] = gr[0] * ar; o[1] = gr[1] * ar; o[2] = gr[2] * ar;
  o[3] = gp[0] * ap; o[4] = gp[1] * ap; o[5] = gp[2] * ap;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[6 + 3 * i + j] = gu[i][j] * au[i];
  o[15] = ar;
}

template <bool LS, bool VISC>
__global__ void __launch_bounds__(kBlock) k_grad(DMesh m, DParams P, const double *__restrict__ prim,
                                                 const double *__restrict__ bfa, double *__restrict__ grad,
                                                 double *__restrict__ gradv) {
  d_grad<LS, VISC>(hw_blk(), m, P, prim, bfa, grad, gradv);
}

// Entry-parallel form of k_grad: one lane per cell-face entry, each 64-lane
// wavefront covering whole cells. Every lane computes its entry's terms;
// the cell's first lane (its owner) adds them from LDS in entry order, which
// is the reference's order, so the sums are the same bits as k_grad's.
// Phase 1: gradient terms. Phase 2: neighbour values and face offsets for
// the limiter, reduced by the owner, which writes the limited row.
constexpr int kWaves = kBlock / 64;

struct EntryLane {
  int e, c, pos;  // entry (-1: idle lane), its cell, position in the cell's list
};

template <int W = kWaves>
__device__ __forceinline__ EntryLane entry_lane(const DMesh &m, const Blk &B) {
  EntryLane L{-1, -1, 0};
  const int w = xcd_block(B.id, B.n) * W + (threadIdx.x >> 6);
  if (w < m.nwave) {
    const int e = m.wave_e0[w] + (threadIdx.x & 63);
    if (e < m.wave_e0[w + 1]) {
      const int2 cp = m.ecp[e];
      L.e = e;
      L.c = cp.x;
      L.pos = cp.y;
    }
  }
  return L;
}

__device__ __forceinline__ EntryLane entry_lane(const DMesh &m) { return entry_lane<>(m, hw_blk()); }

template <bool LS, int W = kWaves>
__device__ __forceinline__ void d_grad_e(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                                   const double *__restrict__ bfa, double *__restrict__ grad) {
  __shared__ double lds[64 * W * 15];
  double *slot = lds + 15 * threadIdx.x;
  const EntryLane L = entry_lane<W>(m, B);
  const bool owner = L.e >= 0 && L.pos == 0;
  double rho = 0.0, u[3] = {0.0, 0.0, 0.0}, p = 0.0, xcv[3] = {0.0, 0.0, 0.0};
  double nv[5] = {0.0, 0.0, 0.0, 0.0, 0.0};  // neighbour rho, u, v, w, p
  double xfa[3] = {0.0, 0.0, 0.0};
  bool internal = false;
  if (L.e >= 0) {
    const double *pc = prim + 8 * (size_t)L.c;
    double2 a = ld2(pc), b = ld2(pc + 2), cc = ld2(pc + 4), dd = ld2(pc + 6);
    rho = a.x; u[0] = a.y; u[1] = b.x; u[2] = b.y; p = cc.x;
    xcv[0] = cc.y; xcv[1] = dd.x; xcv[2] = dd.y;
    const int2 en = m.cfl[L.e];
    const int f = en.x & 0x7fffffff;
    internal = en.y >= 0;
    double t[15];
    if (internal) {
      const double *pn = prim + 8 * (size_t)en.y;
      double2 na = ld2(pn), nb = ld2(pn + 2), nc = ld2(pn + 4), nd = ld2(pn + 6);
      nv[0] = na.x; nv[1] = na.y; nv[2] = nb.x; nv[3] = nb.y; nv[4] = nc.x;
      const double *geo = m.fgeo + 8 * (size_t)f;
      xfa[0] = geo[4]; xfa[1] = geo[5]; xfa[2] = geo[6];
      if (LS) {
        const double *l = m.lsg + 3 * (size_t)(m.nbocv_i[L.c] + L.pos + 1);
        const double l0 = l[0], l1 = l[1], l2 = l[2];
        t[0] = l0 * nv[0]; t[1] = l1 * nv[0]; t[2] = l2 * nv[0];
        t[3] = l0 * nv[4]; t[4] = l1 * nv[4]; t[5] = l2 * nv[4];
        for (int i = 0; i < 3; i++) {
          t[6 + 3 * i] = l0 * nv[1 + i];
          t[7 + 3 * i] = l1 * nv[1 + i];
          t[8 + 3 * i] = l2 * nv[1 + i];
        }
      } else {
        // distance weights (GG_WITH_DIST, UgpWithCv.h:37, :1104-1110)
        const double xn[3] = {nc.y, nd.x, nd.y};
        double ds[3], dn[3];
        for (int j = 0; j < 3; j++) {
          ds[j] = xfa[j] - xcv[j];
          dn[j] = xfa[j] - xn[j];
        }
        const bool s0 = (en.x >= 0);
        const double wself = sqrt(dot3(ds, ds)), wnb = sqrt(dot3(dn, dn));
        const double w0 = s0 ? wself : wnb, w1 = s0 ? wnb : wself;
        const double *fn = m.fnorm + 3 * (size_t)f;
        const double n[3] = {fn[0], fn[1], fn[2]};
        const double fr = s0 ? (w1 * rho + w0 * nv[0]) / (w0 + w1) : (w1 * nv[0] + w0 * rho) / (w0 + w1);
        const double fp = s0 ? (w1 * p + w0 * nv[4]) / (w0 + w1) : (w1 * nv[4] + w0 * p) / (w0 + w1);
        for (int j = 0; j < 3; j++) {
          t[j] = s0 ? fr * n[j] : -(fr * n[j]);
          t[3 + j] = s0 ? fp * n[j] : -(fp * n[j]);
        }
        for (int i = 0; i < 3; i++) {
          const double fu = s0 ? (w1 * u[i] + w0 * nv[1 + i]) / (w0 + w1) : (w1 * nv[1 + i] + w0 * u[i]) / (w0 + w1);
          for (int j = 0; j < 3; j++) t[6 + 3 * i + j] = s0 ? fu * n[j] : -(fu * n[j]);
        }
      }
    } else {
      const double *bb = bfa + 8 * (size_t)f;
      const double rb = bb[0], ub[3] = {bb[1], bb[2], bb[3]}, pb = bb[4];
      double w[3];
      if (LS) {
        const double *bl = m.bfa_lsg + 3 * (size_t)f;
        w[0] = bl[0]; w[1] = bl[1]; w[2] = bl[2];
      } else {
        const double *fn = m.fnorm + 3 * (size_t)f;
        w[0] = fn[0]; w[1] = fn[1]; w[2] = fn[2];
      }
      for (int j = 0; j < 3; j++) {
        t[j] = w[j] * rb;
        t[3 + j] = w[j] * pb;
      }
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) t[6 + 3 * i + j] = w[j] * ub[i];
    }
    for (int k = 0; k < 15; k++) slot[k] = t[k];
  }
  __syncthreads();
  // Ordered sums, spread over the cell's lanes: the lane at position pos of a
  // cell with n entries (n >= 4) sums fields pos, pos + n, ... in entry order,
  // starting from the diagonal term (LS) or zero (GG).
  const double *base = lds + 15 * (threadIdx.x - L.pos);
  const int n = L.e >= 0 ? m.cfl_i[L.c + 1] - m.cfl_i[L.c] : 0;
  double gs[4] = {0.0, 0.0, 0.0, 0.0};
  if (L.e >= 0) {
    double l0[3] = {0.0, 0.0, 0.0};
    if (LS) {
      const double *l = m.lsg + 3 * (size_t)m.nbocv_i[L.c];
      l0[0] = l[0]; l0[1] = l[1]; l0[2] = l[2];
    }
    const double vol = LS ? 1.0 : m.vol[L.c];
    for (int r = 0; r < 4; r++) {
      const int k = L.pos + r * n;
      if (k >= 15) break;
      const double val = k < 3 ? rho : (k < 6 ? p : u[(k - 6) / 3]);
      double a = LS ? val * l0[k % 3] : 0.0;
      for (int j = 0; j < n; j++) a += base[15 * j + k];
      gs[r] = LS ? a : a / vol;
    }
  }
  __syncthreads();
  // neighbour values (internal faces) and the partial sums; field k is then
  // at base[15 * (k % n) + 6 + k / n]
  if (L.e >= 0) {
    slot[0] = internal ? 1.0 : 0.0;
    for (int k = 0; k < 5; k++) slot[1 + k] = nv[k];
    for (int r = 0; r < 4; r++) slot[6 + r] = gs[r];
  }
  __syncthreads();
  // bounds over the cell and its neighbours, spread the same way: quantity
  // q < 5 is the max of rho, u, v, w, p (q), q >= 5 the min (q - 5); bound q
  // is then at base[15 * (q % n) + 10 + q / n]
  if (L.e >= 0 && P.limiter) {
    for (int r = 0; r < 3; r++) {
      const int q = L.pos + r * n;
      if (q >= 10) break;
      const int fld = q % 5;
      const double own = fld == 0 ? rho : (fld == 4 ? p : u[fld - 1]);
      double a = own;
      for (int j = 0; j < n; j++) {
        const double *sj = base + 15 * j;
        if (sj[0] == 0.0) continue;
        a = q < 5 ? Dev::mx(a, sj[1 + fld]) : Dev::mn(a, sj[1 + fld]);
      }
      slot[10 + r] = a;
    }
  }
  __syncthreads();
  // gather the cell's gradient and bounds (no division: walk lanes, wrap columns)
  double g[15], bnd[10];
  if (L.e >= 0) {
    int li = 0, col = 0;
    for (int k = 0; k < 15; k++) {
      g[k] = base[15 * li + 6 + col];
      if (++li == n) {
        li = 0;
        col++;
      }
    }
    if (P.limiter) {
      li = 0;
      col = 0;
      for (int q = 0; q < 10; q++) {
        bnd[q] = base[15 * li + 10 + col];
        if (++li == n) {
          li = 0;
          col++;
        }
      }
    }
  }
  // every internal-face lane: its face's smooth BJ factors (UgpWithCv.h:913-925)
  double fac[5] = {1.0, 1.0, 1.0, 1.0, 1.0};  // rho, p, u, v, w
  if (L.e >= 0 && internal && P.limiter) {
    const double sos = sqrt(P.gamma * p / rho);
    double e2u = P.eps * sos;
    e2u = e2u * e2u;
    double e2r = Dev::mx(P.eps * fabs(rho), 1.0e-8);
    e2r = e2r * e2r;
    double e2p = Dev::mx(P.eps * fabs(p), 1.0e-8);
    e2p = e2p * e2p;
    const double r0[3] = {xfa[0] - xcv[0], xfa[1] - xcv[1], xfa[2] - xcv[2]};
    fac[0] = bj_face(1.0, rho, rho + dot3(r0, g), bnd[0], bnd[5], e2r);
    fac[1] = bj_face(1.0, p, p + dot3(r0, g + 3), bnd[4], bnd[9], e2p);
    for (int i = 0; i < 3; i++)
      fac[2 + i] = bj_face(1.0, u[i], u[i] + dot3(r0, g + 6 + 3 * i), bnd[1 + i], bnd[6 + i], e2u);
  }
  // slots 0..4 are free now (the reads above use 6..12)
  if (L.e >= 0)
    for (int k = 0; k < 5; k++) slot[k] = fac[k];
  __syncthreads();
  if (!owner) return;
  double au[3] = {1.0, 1.0, 1.0}, ar = 1.0, ap = 1.0;
  if (P.limiter) {
    // min over faces is exact in any order
    for (int j = 0; j < n; j++) {
      const double *sj = base + 15 * j;
      ar = Dev::mn(ar, sj[0]);
      ap = Dev::mn(ap, sj[1]);
      for (int i = 0; i < 3; i++) au[i] = Dev::mn(au[i], sj[2 + i]);
    }
    bool lu[3] = {(bnd[1] - bnd[6]) > 1.0e-8, (bnd[2] - bnd[7]) > 1.0e-8, (bnd[3] - bnd[8]) > 1.0e-8};
    bool lr = (bnd[0] - bnd[5]) > 1.0e-12, lp = (bnd[4] - bnd[9]) > 1.0e-12;
    for (int i = 0; i < 3; i++)
      if (!lu[i]) au[i] = 0.0;
    if (!lr) ar = 0.0;
    if (!lp) ap = 0.0;
  }
  double *o = grad + 16 * (size_t)L.c;
  o[0] = g[0] * ar; o[1] = g[1] * ar; o[2] = g[2] * ar;
  o[3] = g[3] * ap; o[4] = g[4] * ap; o[5] = g[5] * ap;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) o[6 + 3 * i + j] = g[6 + 3 * i + j] * au[i];
  o[15] = ar;
}

// one wavefront per block, as the flux kernel
constexpr int kGradWaves = 1;

template <bool LS>
__global__ void __launch_bounds__(64 * kGradWaves) k_grad_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                            const double *__restrict__ bfa,
                                                            double *__restrict__ grad) {
  d_grad_e<LS, kGradWaves>(hw_blk(), m, P, prim, bfa, grad);
}

// Reconstructed face state of one side (JoeWithModels.cpp:2832-2921): rho and
// p extrapolated first; both revert to the cell value when either is not
// positive, and the velocity is then left first order.
struct Side {
  double rho, u[3], p;
};

__device__ __forceinline__ bool reconstruct(Side &s, const double *pc, const double *gc, const double *xfa) {
  double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
  s.rho = pc[0] + dot3(r, gc);
  s.p = pc[4] + dot3(r, gc + 3);
  s.u[0] = pc[1];
  s.u[1] = pc[2];
  s.u[2] = pc[3];
  if ((s.p <= 0.0) || (s.rho <= 0.0)) {
    s.p = pc[4];
    s.rho = pc[0];
    return false;
  }
  s.u[0] += dot3(r, gc + 6);
  s.u[1] += dot3(r, gc + 9);
  s.u[2] += dot3(r, gc + 12);
  return true;
}

// kL, kR: the sides' turbulent kinetic energy (SST, HLLC only; else 0)
__device__ __forceinline__ void face_flux(int scheme, double F[5], const Side &L, const Side &R, double gam,
                                          double area, const double *n, double kL = 0.0, double kR = 0.0) {
  if (scheme == kRoe)
    flux_roe(F, L.rho, L.u, L.p, R.rho, R.u, R.p, gam, area, n);
  else if (scheme == kAusm)
    flux_ausm(F, L.rho, L.u, L.p, R.rho, R.u, R.p, gam, area, n);
  else
    flux_hllc(F, L.rho, L.u, L.p, R.rho, R.u, R.p, gam, area, n, 0.0, 0.0, nullptr, kL, kR);
}

// ---------------------------------------------------------------------------
// Residual: calcEulerFlux face loops (JoeWithModels.cpp:2786-3569) and the
// default sourceHook (JoeWithModels.h:111-129), gathered per cell in the
// reference's accumulation order so every cell sum matches it exactly.
// ---------------------------------------------------------------------------
// Local time step arguments (calcDt, UgpWithCvCompFlow.cpp:5-110): mode 0
// CFL (local), 1 CFL_MIN (global minimum filled afterwards), 2 fixed DT.
struct DtArgs {
  double cfl, dt_fixed;
  int mode;
  int sum;                // TIME_STEP = SUM_BASED (faocv-ordered sums; needs DMesh::foc)
  const double *cfl_dev;  // when set, the CFL number is read from device memory (graph-captured steps)
};

// One cell's residual; also its MAX_BASED local time step (UgpWithCvCompFlow.cpp:27-60,
// inviscid: the viscous bound 4(mu+mut)A^2/(rho V) is zero), which needs the
// same face normals and areas.
template <int SCHEME, bool SND>
__device__ __forceinline__ double flux_cell(const DMesh &m, const DParams &P, const double *__restrict__ prim,
                                            const double *__restrict__ grad, const double *__restrict__ bfa,
                                            const double *__restrict__ q, double *__restrict__ rhs,
                                            int *__restrict__ nfirst, const DtArgs &dta, double *__restrict__ dt,
                                            const int c) {
  const double gam = P.gamma;
  double pc[8], gc[15];
  {
    const double *s = prim + 8 * (size_t)c;
    for (int k = 0; k < 8; k += 2) {
      double2 v = ld2(s + k);
      pc[k] = v.x;
      pc[k + 1] = v.y;
    }
    if (SND) {
      const double *g = grad + 16 * (size_t)c;
      for (int k = 0; k < 14; k += 2) {
        double2 v = ld2(g + k);
        gc[k] = v.x;
        gc[k + 1] = v.y;
      }
      gc[14] = g[14];
    }
  }
  Side own_cc;
  own_cc.rho = pc[0];
  own_cc.u[0] = pc[1]; own_cc.u[1] = pc[2]; own_cc.u[2] = pc[3];
  own_cc.p = pc[4];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int nlow = 0;
  const double ru[3] = {q[m.ncv + c], q[2 * m.ncv + c], q[3 * m.ncv + c]};
  const double cs = sqrt(gam * pc[4] / pc[0]);
  double lambdaMax = 0.0;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  for (int e = e0; e < e1; e++) {
    int2 en = m.cfl[e];
    int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    double n[3], area, xfa[3];
    {
      double2 a = ld2(geo), b = ld2(geo + 2), cc = ld2(geo + 4);
      n[0] = a.x; n[1] = a.y; n[2] = b.x; area = b.y;
      xfa[0] = cc.x; xfa[1] = cc.y; xfa[2] = geo[6];
    }
    {
      const double Uk = dot3(ru, n) / pc[0];
      lambdaMax = Dev::mx(lambdaMax, (fabs(Uk) + cs) * area);
    }
    double F[5];
    if (en.y >= 0) {
      const bool s0 = (en.x >= 0);
      double pn[8], gn[15];
      const double *s = prim + 8 * (size_t)en.y;
      for (int k = 0; k < 8; k += 2) {
        double2 v = ld2(s + k);
        pn[k] = v.x;
        pn[k + 1] = v.y;
      }
      Side so, sn;
      if (SND) {
        const double *g = grad + 16 * (size_t)en.y;
        for (int k = 0; k < 14; k += 2) {
          double2 v = ld2(g + k);
          gn[k] = v.x;
          gn[k + 1] = v.y;
        }
        gn[14] = g[14];
        if (!reconstruct(so, pc, gc, xfa)) nlow++;
        reconstruct(sn, pn, gn, xfa);
      } else {
        so = own_cc;
        sn.rho = pn[0]; sn.u[0] = pn[1]; sn.u[1] = pn[2]; sn.u[2] = pn[3]; sn.p = pn[4];
      }
      // one flux on the face's (icv0, icv1) sides; acc - F == acc + (-F)
      face_flux(SCHEME, F, s0 ? so : sn, s0 ? sn : so, gam, area, n);
      for (int k = 0; k < 5; k++) acc[k] += s0 ? -F[k] : F[k];
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      Side sb;
      sb.rho = bb[0]; sb.u[0] = bb[1]; sb.u[1] = bb[2]; sb.u[2] = bb[3]; sb.p = bb[4];
      if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        face_flux(SCHEME, F, sb, sb, gam, area, n);
        if (z.bc == kBcSymmetry && P.symm_bc == kSymmAnalytical) {
          acc[0] -= 0.0;
          for (int i = 0; i < 3; i++) acc[1 + i] -= sb.p * n[i] * area;
          acc[4] -= 0.0;
        } else {
          for (int k = 0; k < 5; k++) acc[k] -= F[k];
        }
      } else {
        Side so = own_cc;
        if (SND && !reconstruct(so, pc, gc, xfa)) nlow++;
        face_flux(SCHEME, F, so, sb, gam, area, n);
        for (int k = 0; k < 5; k++) acc[k] -= F[k];
      }
    }
  }
  // sourceHook: gravity (JoeWithModels.h:111-129)
  const double vol = m.vol[c];
  if (P.gravity_on) {
    acc[1] += vol * pc[0] * P.g[0];
    acc[2] += vol * pc[0] * P.g[1];
    acc[3] += vol * pc[0] * P.g[2];
    acc[4] += vol * (P.g[0] * ru[0] + P.g[1] * ru[1] + P.g[2] * ru[2]);
  } else {
    // the same terms with g = 0: +0 for rhou, and -0 for rhoE only when
    // all three 0*rhou products are -0
    acc[1] += 0.0;
    acc[2] += 0.0;
    acc[3] += 0.0;
    bool allneg = signbit(pc[1]) && signbit(pc[2]) && signbit(pc[3]);
    acc[4] += allneg ? -0.0 : 0.0;
  }
  for (int k = 0; k < 5; k++) rhs[(size_t)k * m.ncv + c] = acc[k];
  if (nlow) atomicAdd(nfirst, nlow);
  double dtc;
  if (dta.mode == 2) {
    dtc = dta.dt_fixed;
  } else if (dta.sum) {  // SUM_BASED (UgpWithCvCompFlow.cpp:62-99), inviscid: dt_visc = 1e20
    double lc = 0.0;
    for (int k = 0; k < e1 - e0; k++) {
      const int f = m.cfl[e0 + m.foc[e0 + k]].x & 0x7fffffff;
      const double *geo = m.fgeo + 8 * (size_t)f;
      const double n[3] = {geo[0], geo[1], geo[2]};
      lc += (fabs(dot3(ru, n) / pc[0]) + cs) * geo[3];
    }
    dtc = Dev::mn((dta.cfl_dev ? *dta.cfl_dev : dta.cfl) * vol / lc, 1.0e20);
  } else {
    dtc = (dta.cfl_dev ? *dta.cfl_dev : dta.cfl) * vol / lambdaMax;
  }
  dt[c] = dtc;
  return dtc;
}

// ---------------------------------------------------------------------------
// Laminar viscous face flux, addViscFlux (UgpWithCvCompFlow.cpp:1521-1763)
// with the eddy viscosity mut of the Spalart-Allmaras model (0 without a
// turbulence model), nonLinear = rij = kine = 0 and SIMPLEX_VISCOUS = NO;
// zero turbulent terms are kept so that signed zeros come out as in the
// reference.
// ---------------------------------------------------------------------------
struct VSide {
  double rho, u[3], h, T;
  const double *gu, *gh;  // grad u [3][3], grad h [3]
  double k;               // turbulent kinetic energy (SST; else 0)
};

__device__ __forceinline__ double mu_lam(const DParams &P, double T) {
  if (P.visc_mode == 1) return P.mu_ref * pow(T / P.sl_tref, 1.5) * (P.sl_tref + P.sl_sref) / (T + P.sl_sref);
  return P.mu_ref * pow(T / P.T_ref, P.mu_power_law);
}

struct VFace {
  double tau[3];  // tau_ij n_j
  double keff, alpha;
};

// With SST the kinetic-energy terms: -2/3 rho k on the normal stress and the
// triple correlation (mul + mut) dk/ds in the energy flux (:1614-1615, 1657-1660)
template <bool SST = false>
__device__ __forceinline__ void visc_face(const DParams &P, const VSide &s0, const VSide &s1, double mul, double mut,
                                       const double *u_fa, double area, const double *nVec, double smag,
                                       const double *sVec, double Frhou[3], double &FrhoE, VFace &vf,
                                       double kine_fa = 0.0) {
  const double nonL = 0.0;
  const double alpha = dot3(nVec, sVec);
  double guf[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) guf[i][j] = 0.5 * (s0.gu[3 * i + j] + s1.gu[3 * i + j]);
  const double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  double tauLam[3], tauTurb[3], tmp;
  for (int i = 0; i < 3; i++)
    tauLam[i] = mul * (alpha * (s1.u[i] - s0.u[i]) / smag + guf[i][0] * fCorr[0] + guf[0][i] * nVec[0] +
                       guf[i][1] * fCorr[1] + guf[1][i] * nVec[1] + guf[i][2] * fCorr[2] + guf[2][i] * nVec[2]);
  tmp = 2.0 / 3.0 * mul * (guf[0][0] + guf[1][1] + guf[2][2]);
  for (int i = 0; i < 3; i++) tauLam[i] -= tmp * nVec[i];
  for (int i = 0; i < 3; i++)
    tauTurb[i] = (1.0 - nonL) * mut *
                 (alpha * (s1.u[i] - s0.u[i]) / smag + guf[i][0] * fCorr[0] + guf[0][i] * nVec[0] +
                  guf[i][1] * fCorr[1] + guf[1][i] * nVec[1] + guf[i][2] * fCorr[2] + guf[2][i] * nVec[2]);
  tmp = (1.0 - nonL) * 2.0 / 3.0 * mut * (guf[0][0] + guf[1][1] + guf[2][2]);
  if (SST) tmp += (1.0 - nonL) * 1.0 / 3.0 * (s0.rho + s1.rho) * kine_fa;
  for (int i = 0; i < 3; i++) tauTurb[i] -= tmp * nVec[i];
  tauTurb[0] += nonL * (0.0 * nVec[0] + 0.0 * nVec[1] + 0.0 * nVec[2]);
  tauTurb[1] += nonL * (0.0 * nVec[0] + 0.0 * nVec[1] + 0.0 * nVec[2]);
  tauTurb[2] += nonL * (0.0 * nVec[0] + 0.0 * nVec[1] + 0.0 * nVec[2]);
  for (int i = 0; i < 3; i++) vf.tau[i] = tauLam[i] + tauTurb[i];
  for (int i = 0; i < 3; i++) Frhou[i] = -area * vf.tau[i];
  const double keff = mul / P.pr + mut / P.pr_turb;
  const double enth = keff * (alpha * (s1.h - s0.h) / smag + 0.5 * ((s0.gh[0] + s1.gh[0]) * fCorr[0] +
                                                                    (s0.gh[1] + s1.gh[1]) * fCorr[1] +
                                                                    (s0.gh[2] + s1.gh[2]) * fCorr[2]));
  const double psi = SST ? (mul + mut) * (alpha * (s1.k - s0.k) / smag) : 0.0;
  FrhoE = -area * (enth + vf.tau[0] * u_fa[0] + vf.tau[1] * u_fa[1] + vf.tau[2] * u_fa[2] + psi);
  vf.keff = keff;
  vf.alpha = alpha;
}

// one half of the viscous Jacobian (A0: sgn -1 with side 0's state, A1: +1
// with side 1's), dFdU x dUdQ (:1662-1762, :712-745); rows 1-4 only
template <bool SST = false>
__device__ __forceinline__ void visc_jac(const DParams &P, double J[5][5], double sgn, const VSide &s, double mul,
                                      double mut, const VFace &vf, const double *u_fa, double area, double smag) {
  const double kine = SST ? s.k : 0.0;
  const double gam = P.gamma, R = P.R;
  double f11 = sgn * (mul + mut) * vf.alpha / smag * area;
  double f22 = sgn * (mul + mut) * vf.alpha / smag * area;
  double f33 = sgn * (mul + mut) * vf.alpha / smag * area;
  double f41 = area * vf.tau[0] * 0.5 + u_fa[0] * f11;
  double f42 = area * vf.tau[1] * 0.5 + u_fa[1] * f22;
  double f43 = area * vf.tau[2] * 0.5 + u_fa[2] * f33;
  double f44 = sgn * area * vf.keff * vf.alpha / smag;
  const double f12 = 0.0, f13 = 0.0, f21 = 0.0, f23 = 0.0, f31 = 0.0, f32 = 0.0;
  const double invRho = 1.0 / s.rho;
  const double q10 = -s.u[0] * invRho, q20 = -s.u[1] * invRho, q30 = -s.u[2] * invRho;
  const double q40 = gam * invRho * (-s.h - kine + R * s.T + 0.5 * dot3(s.u, s.u));
  const double q11 = invRho, q41 = -gam * s.u[0] * invRho;
  const double q22 = invRho, q42 = -gam * s.u[1] * invRho;
  const double q33 = invRho, q43 = -gam * s.u[2] * invRho;
  const double q44 = gam * invRho;
  const double q14 = 0.0, q24 = 0.0, q34 = 0.0;
  J[1][0] = f11 * q10 + f12 * q20 + f13 * q30;
  J[2][0] = f21 * q10 + f22 * q20 + f23 * q30;
  J[3][0] = f31 * q10 + f32 * q20 + f33 * q30;
  J[4][0] = f41 * q10 + f42 * q20 + f43 * q30 + f44 * q40;
  J[1][1] = f11 * q11;
  J[2][1] = f21 * q11;
  J[3][1] = f31 * q11;
  J[4][1] = f41 * q11 + f44 * q41;
  J[1][2] = f12 * q22;
  J[2][2] = f22 * q22;
  J[3][2] = f32 * q22;
  J[4][2] = f42 * q22 + f44 * q42;
  J[1][3] = f13 * q33;
  J[2][3] = f23 * q33;
  J[3][3] = f33 * q33;
  J[4][3] = f43 * q33 + f44 * q43;
  J[4][4] = f41 * q14 + f42 * q24 + f43 * q34 + f44 * q44;
  J[1][4] = J[2][4] = J[3][4] = 0.0;  // A0/A1 keep these zero (only A[4][4] is set in column 4)
}

// The viscous part of one cell-face entry: the face's viscous flux signed for
// this cell (flux(icv0 -> icv1) taken from icv0, added to icv1), and the
// face's mul for the viscous time-step bound. Symmetry faces contribute
// nothing (signed -0.0 terms leave the sums unchanged).
struct VEntry {
  double c[4];  // rhou(3), rhoE
  double mul;
};

__device__ __forceinline__ VSide vside_cell(const DParams &P, const double *pc, const double *gv) {
  VSide s;
  s.rho = pc[0];
  s.u[0] = pc[1]; s.u[1] = pc[2]; s.u[2] = pc[3];
  s.T = pc[4] / (pc[0] * P.R);
  s.h = P.gamma * P.R / (P.gamma - 1.0) * s.T;
  s.gu = gv;
  s.gh = gv + 9;
  s.k = 0.0;
  return s;
}

__device__ __forceinline__ VSide vside_bfa(const double *bb, const double *gv) {
  VSide s;
  s.rho = bb[0];
  s.u[0] = bb[1]; s.u[1] = bb[2]; s.u[2] = bb[3];
  s.T = bb[5];
  s.h = bb[6];
  s.gu = gv;
  s.gh = gv + 9;
  s.k = 0.0;
  return s;
}

// geometry of an internal face for the viscous terms: icv0 -> icv1 unit
// vector and distance, and the distance-weighted face temperature's mul
// (calcMaterialProperties, UgpWithCvCompFlow.h:1700-1750)
__device__ __forceinline__ void visc_internal_geom(const DParams &P, const double *x0, const double *x1,
                                                   const double *xfa, double T0, double T1, double sVec[3],
                                                   double &smag, double &mul, double &w0, double &w1) {
  for (int i = 0; i < 3; i++) sVec[i] = x1[i] - x0[i];
  smag = sqrt(dot3(sVec, sVec));
  for (int i = 0; i < 3; i++) sVec[i] /= smag;
  double dx0[3], dx1[3];
  for (int i = 0; i < 3; i++) {
    dx0[i] = xfa[i] - x0[i];
    dx1[i] = xfa[i] - x1[i];
  }
  w0 = sqrt(dot3(dx0, dx0));
  w1 = sqrt(dot3(dx1, dx1));
  mul = mu_lam(P, (w1 * T0 + w0 * T1) / (w0 + w1));
}

// one of two sides, selected field by field (a conditional between two
// local structs bound to a reference would put both in scratch memory)
__device__ __forceinline__ VSide vsel(bool c, const VSide &x, const VSide &y) {
  VSide r;
  r.rho = c ? x.rho : y.rho;
  for (int i = 0; i < 3; i++) r.u[i] = c ? x.u[i] : y.u[i];
  r.h = c ? x.h : y.h;
  r.T = c ? x.T : y.T;
  r.gu = c ? x.gu : y.gu;
  r.gh = c ? x.gh : y.gh;
  r.k = c ? x.k : y.k;
  return r;
}

// The coupled loop's viscous flux, calcViscousFluxCoupled
// (UgpWithCvCompFlow.cpp:3040-3124): F[5] (F[0] = 0), and tau_ij n_j and keff
// for the blocks below. With a turbulence model (turb) the kinetic-energy
// terms are formed with kine = 0 (:3084-3085, 3108-3110).
__device__ __forceinline__ void visc_coupled(const DParams &P, const VSide &s0, const VSide &s1, double mul,
                                             double mut, const double *u_fa, double area, const double *nVec,
                                             double smag, const double *sVec, double alpha, double F[5],
                                             double tau[3], double &keff, bool turb = false) {
  double g[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g[i][j] = 0.5 * (s0.gu[3 * i + j] + s1.gu[3 * i + j]);
  double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  const double mu = mul + mut;
  F[0] = 0.0;
  for (int i = 0; i < 3; i++)
    tau[i] = mu * (alpha * (s1.u[i] - s0.u[i]) / smag + g[i][0] * fCorr[0] + g[0][i] * nVec[0] + g[i][1] * fCorr[1] +
                   g[1][i] * nVec[1] + g[i][2] * fCorr[2] + g[2][i] * nVec[2]);
  double tmp = 2.0 / 3.0 * mu * (g[0][0] + g[1][1] + g[2][2]);
  if (turb) tmp += 1.0 / 3.0 * (s0.rho + s1.rho) * 0.0;
  for (int i = 0; i < 3; i++) tau[i] -= tmp * nVec[i];
  for (int i = 0; i < 3; i++) F[1 + i] = -area * tau[i];
  keff = mul / P.pr + mut / P.pr_turb;
  double enth = keff * (alpha * (s1.h - s0.h) / smag +
                        0.5 * ((s0.gh[0] + s1.gh[0]) * fCorr[0] + (s0.gh[1] + s1.gh[1]) * fCorr[1] +
                               (s0.gh[2] + s1.gh[2]) * fCorr[2]));
  double psi = 0.0;
  if (turb) psi = mu * (alpha * (0.0 - 0.0) / smag);
  F[4] = -area * (enth + tau[0] * u_fa[0] + tau[1] * u_fa[1] + tau[2] * u_fa[2] + psi);
}

// its blocks dF/dU dU/dQ (:3140-3268): which 0 is A0 (side 0's state, + sign),
// 1 is A1 (side 1's, - sign); entries the reference does not write are zero
__device__ __forceinline__ void visc_coupled_jac(const DParams &P, double J[5][5], int which, const VSide &s,
                                                 double mu, double keff, const double tau[3], const double *u_fa,
                                                 double area, double alpha, double smag) {
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) J[i][j] = 0.0;
  double aas = area * alpha / smag;
  double d11 = which == 0 ? mu * aas : -mu * aas;
  double d44 = which == 0 ? keff * aas : -keff * aas;
  double d4[3];
  for (int i = 0; i < 3; i++) d4[i] = -area * tau[i] * 0.5 + u_fa[i] * d11;
  double invRho = 1.0 / s.rho;
  double q0[3], q4[3];
  for (int i = 0; i < 3; i++) q0[i] = -invRho * s.u[i];
  double q40 = invRho * P.gamma * (-s.h - 0.0 + P.R * s.T + 0.5 * dot3(s.u, s.u));
  for (int i = 0; i < 3; i++) q4[i] = -invRho * s.u[i] * P.gamma;
  double q44 = invRho * P.gamma;
  for (int i = 0; i < 3; i++) {
    J[1 + i][0] = d11 * q0[i];
    J[1 + i][1 + i] = d11 * invRho;
  }
  J[4][0] = d4[0] * q0[0] + d4[1] * q0[1] + d4[2] * q0[2] + d44 * q40;
  for (int i = 0; i < 3; i++) J[4][1 + i] = d4[i] * invRho + d44 * q4[i];
  J[4][4] = d44 * q44;
}

// the coupled loop's boundary-face viscous geometry (JoeWithModels.cpp
// :5373-5404, 5548-5575, 5744-5775): along the normal with alpha = 1 on
// NEUMANN and WALL faces, the normalised cell-to-face vector otherwise
__device__ __forceinline__ void visc_coupled_bgeom(int bc, const double *xfa, const double *xc, const double *n,
                                                   double sArg[3], double &smag, double &alpha) {
  double sV[3] = {xfa[0] - xc[0], xfa[1] - xc[1], xfa[2] - xc[2]};
  if (bc == kBcNeumann || bc == kBcWall) {
    smag = fabs(dot3(sV, n));
    for (int i = 0; i < 3; i++) sArg[i] = n[i];
    alpha = 1.0;
  } else {
    smag = sqrt(dot3(sV, sV));
    for (int i = 0; i < 3; i++) sArg[i] = sV[i] / smag;
    alpha = dot3(n, sArg);
  }
}

// ---------------------------------------------------------------------------
// Menter SST k-omega (RansTurbKOmSST, turbModels/TurbModel_KOMSST.h) as two
// segregated scalars, kine and omega, in the reference's registration order
// ---------------------------------------------------------------------------
// model constants, getDoubleParam defaults (TurbModel_KOMSST.h:16-23); alfa_1
// and alfa_2 come from the host (DParams::sst_alfa), formed there as the
// reference forms them
constexpr double kSstSigmaK1 = 0.85, kSstSigmaK2 = 1.0, kSstSigmaW1 = 0.5, kSstSigmaW2 = 0.856, kSstBeta1 = 0.075,
                 kSstBeta2 = 0.0828, kSstBetaStar = 0.09, kSstA1 = 0.31;
// kine and omega transport settings (TurbModel_KOMSST.h:36-52)
constexpr double kSstPhiZero = 1.0e-8, kSstPhiZeroRel = 1.0e-2;
constexpr int kSstMaxIter = 1000;
constexpr double kSstLower[2] = {1.0e-10, 1.0e-4}, kSstUpper[2] = {1.0e10, 1.0e15};

// the shear-stress limiter's field: vorticity, or with SST_LIMITER_SHEARSTRESS = 1 strain (:90-93)
__device__ __forceinline__ double sst_lim(const DParams &P, const double *a) {
  return P.sst_lim_shear == 1 ? a[kSStr] : a[kSVort];
}

// mut_fa of calcRansTurbViscMuet (TurbModel_KOMSST.h:108-131) at an internal
// face: distance-weighted rho, kine, omega, limiter and F2 (icv0, icv1 order)
__device__ __forceinline__ double sst_mut_face(const DParams &P, int ncv, double w0, double w1, double rho0,
                                               double rho1, int c0, int c1) {
  const double *a0 = P.sst_aux + kSstN * (size_t)c0, *a1 = P.sst_aux + kSstN * (size_t)c1;
  const double rho_fa = (w1 * rho0 + w0 * rho1) / (w0 + w1);
  const double kine_fa = (w1 * P.sst_phi[c0] + w0 * P.sst_phi[c1]) / (w0 + w1);
  const double om_fa = (w1 * P.sst_phi[ncv + c0] + w0 * P.sst_phi[ncv + c1]) / (w0 + w1);
  const double limitFunc_fa = (w1 * sst_lim(P, a0) + w0 * sst_lim(P, a1)) / (w0 + w1);
  const double blendFuncF2_fa = (w1 * a0[kSF2] + w0 * a1[kSF2]) / (w0 + w1);
  const double zeta = Dev::mn(1.0 / om_fa, kSstA1 / (limitFunc_fa * blendFuncF2_fa));
  return Dev::mn(Dev::mx(rho_fa * kine_fa * zeta, 0.0), 1.0);
}

// the same on a non-wall boundary face, from its cell (:143-149)
__device__ __forceinline__ double sst_mut_cell(const DParams &P, int ncv, double rho, int c) {
  const double *a = P.sst_aux + kSstN * (size_t)c;
  const double zeta = Dev::mn(1.0 / P.sst_phi[ncv + c], kSstA1 / (sst_lim(P, a) * a[kSF2]));
  return Dev::mn(Dev::mx(rho * P.sst_phi[c] * zeta, 0.0), 1.0);
}

// ---------------------------------------------------------------------------
// Spalart-Allmaras (RansTurbSA, TurbModel_SA.h) as one segregated scalar nuSA
// ---------------------------------------------------------------------------
// model constants, getDoubleParam defaults (TurbModel_SA.h:20-27)
constexpr double kSaCv1_3 = 357.911, kSaCw1 = 3.239067817, kSaCw2 = 0.3, kSaCw3_6 = 64.0, kSaSigma = 0.6666666,
                 kSaCb1 = 0.1355, kSaCb2 = 0.622, kSaKarman2 = 0.1681;
// nuSA transport settings (TurbModel_SA.h:31-37)
constexpr double kSaPhiZero = 1.0e-9, kSaPhiZeroRel = 1.0e-2, kSaLower = 1.0e-10, kSaUpper = 1.0;
constexpr int kSaMaxIter = 1000;

// eddy viscosity from rho nuSA and the laminar viscosity (calcRansTurbViscMuet,
// TurbModel_SA.h:82-99 internal faces, :112-120 non-wall boundary faces)
__device__ __forceinline__ double sa_mut(double muSA, double mul) {
  double chi_3 = pow(muSA / mul, 3.0);
  double fv1 = chi_3 / (chi_3 + kSaCv1_3);
  return Dev::mx(muSA * fv1, 0.0);
}

// the same at an internal face from distance-weighted rho and nuSA (icv0, icv1 order)
__device__ __forceinline__ double sa_mut_face(double w0, double w1, double rho0, double rho1, double nu0, double nu1,
                                              double mul) {
  double rho_fa = (w1 * rho0 + w0 * rho1) / (w0 + w1);
  double nuSA_fa = (w1 * nu0 + w0 * nu1) / (w0 + w1);
  return sa_mut(rho_fa * nuSA_fa, mul);
}

// ---------------------------------------------------------------------------
// Block Jacobian: the implicit parts of the same face loops
// (JoeWithModels.cpp:3029-3101, 3213-3277, 3359-3385, 3517-3563) with
// cell-centre states, gathered per row; optionally the backward-Euler
// diagonal V/dt and the under-relaxed right-hand side (:1011-1024).
// ---------------------------------------------------------------------------
// the flow rows of a block of row stride NB (5, or 6 in the coupled loop with
// one scalar, whose flow rows have a zero scalar column)
template <typename T, int NB = 5>
__device__ __forceinline__ void store_block(T *dst, const double J[5][5], bool negate, bool accumulate) {
  for (int i = 0; i < 5; i++) {
    for (int j = 0; j < 5; j++) {
      double v = negate ? (accumulate ? (double)dst[NB * i + j] - J[i][j] : 0.0 - J[i][j])
                        : (accumulate ? (double)dst[NB * i + j] + J[i][j] : 0.0 + J[i][j]);
      dst[NB * i + j] = (T)v;
    }
    if (NB == 6 && !accumulate) dst[NB * i + 5] = (T)0.0;
  }
}

__device__ __forceinline__ void face_jac(int scheme, double J[5][5], int which, const Side &L, const Side &R,
                                         double gam, double area, const double *n, double kL = 0.0, double kR = 0.0) {
  if (scheme == kRoe)
    jac_roe(J, which, L.rho, L.u, L.p, R.rho, R.u, R.p, gam, area, n);
  else
    jac_hllc(J, which, L.rho, L.u, L.p, R.rho, R.u, R.p, gam, area, n, kL, kR);
}

template <typename T, int SCHEME, bool VISC, int NB = 5>
__device__ __forceinline__ void d_jac(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                                const double *__restrict__ bfa, T *__restrict__ A, int step_mode,
                                                const double *__restrict__ dt, const double *__restrict__ rhs,
                                                double relax, T *__restrict__ rhs_ls, T *__restrict__ phi,
                                                const double *__restrict__ gradv) {
  const int t = xcd_block(B.id, B.n) * blockDim.x + threadIdx.x;
  if (t >= m.ncv) return;
  const int c = t;
  const double gam = P.gamma;
  const double *pc = prim + 8 * (size_t)c;
  Side own;
  own.rho = pc[0]; own.u[0] = pc[1]; own.u[1] = pc[2]; own.u[2] = pc[3]; own.p = pc[4];
  double D[5][5];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) D[i][j] = 0.0;
  double J[5][5];
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  const int row = m.nbocv_i[c];
  for (int e = e0; e < e1; e++) {
    int2 en = m.cfl[e];
    int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double area = geo[3];
    if (en.y >= 0) {
      const double *pn = prim + 8 * (size_t)en.y;
      Side nb;
      nb.rho = pn[0]; nb.u[0] = pn[1]; nb.u[1] = pn[2]; nb.u[2] = pn[3]; nb.p = pn[4];
      const int noc = m.cfl_noc[e];
      const bool accum = (noc >> 30) & 1;
      T *off = A + NB * NB * (size_t)(noc & 0x3fffffff);
      // this cell is icv0: A[noc00] += Apl, A[noc01] += Ami; icv1:
      // A[noc11] -= Ami, A[noc10] -= Apl. Both blocks are evaluated once on
      // the lane's own (icv0, icv1) sides, so lanes of either kind share the
      // arithmetic; the diagonal and the off-diagonal block are different
      // memory, so the order of their updates does not matter.
      const bool s0 = en.x >= 0;
      const Side l = s0 ? own : nb, r = s0 ? nb : own;
      double hl = 0.0, hr = 0.0;
      if (SCHEME == kHllcC) {  // the cells' enthalpies (JoeWithModels.cpp:5035-5038)
        hl = enthalpy_of(P, l.rho, l.p);
        hr = enthalpy_of(P, r.rho, r.p);
        jac_hllc_coupled(J, 0, l.rho, l.u, l.p, hl, r.rho, r.u, r.p, hr, gam, area, n);
      } else {
        face_jac(SCHEME, J, 0, l, r, gam, area, n);  // Apl
      }
      if (s0) {
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
      } else {
        store_block<T, NB>(off, J, true, accum);
      }
      if (SCHEME == kHllcC)
        jac_hllc_coupled(J, 1, l.rho, l.u, l.p, hl, r.rho, r.u, r.p, hr, gam, area, n);
      else
        face_jac(SCHEME, J, 1, l, r, gam, area, n);  // Ami
      if (s0) {
        store_block<T, NB>(off, J, false, accum);
      } else {
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) D[i][j] -= J[i][j];
      }
      if (VISC && SCHEME == kHllcC) {  // the same face's viscous blocks next (JoeWithModels.cpp:5096-5104)
        const VSide vo = vside_cell(P, pc, gradv + 12 * (size_t)c);
        const VSide vn = vside_cell(P, pn, gradv + 12 * (size_t)en.y);
        const VSide a = vsel(s0, vo, vn), b = vsel(s0, vn, vo);
        const double xa[3] = {s0 ? pc[5] : pn[5], s0 ? pc[6] : pn[6], s0 ? pc[7] : pn[7]};
        const double xb[3] = {s0 ? pn[5] : pc[5], s0 ? pn[6] : pc[6], s0 ? pn[7] : pc[7]};
        const double xfa[3] = {geo[4], geo[5], geo[6]};
        double sV[3], smag, mul, w0, w1;
        visc_internal_geom(P, xa, xb, xfa, a.T, b.T, sV, smag, mul, w0, w1);
        const double alpha = dot3(n, sV);
        double mut = 0.0;
        if (P.sa) {
          const double nuo = P.sa_phi[c], nun = P.sa_phi[en.y];
          mut = sa_mut_face(w0, w1, a.rho, b.rho, s0 ? nuo : nun, s0 ? nun : nuo, mul);
        }
        const double ufa[3] = {0.5 * (a.u[0] + b.u[0]), 0.5 * (a.u[1] + b.u[1]), 0.5 * (a.u[2] + b.u[2])};
        double VF[5], tau[3], keff;
        visc_coupled(P, a, b, mul, mut, ufa, area, n, smag, sV, alpha, VF, tau, keff, P.sa != 0);
        // icv0: A[noc00] += A0, A[noc01] += A1; icv1: A[noc11] -= A1, A[noc10] -= A0
        visc_coupled_jac(P, J, s0 ? 0 : 1, vsel(s0, a, b), mul + mut, keff, tau, ufa, area, alpha, smag);
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++)
            if (s0) D[i][j] += J[i][j];
            else D[i][j] -= J[i][j];
        visc_coupled_jac(P, J, s0 ? 1 : 0, vsel(s0, b, a), mul + mut, keff, tau, ufa, area, alpha, smag);
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++)
            off[NB * i + j] = (T)(s0 ? (double)off[NB * i + j] + J[i][j] : (double)off[NB * i + j] - J[i][j]);
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      Side sb;
      sb.rho = bb[0]; sb.u[0] = bb[1]; sb.u[1] = bb[2]; sb.u[2] = bb[3]; sb.p = bb[4];
      if (SCHEME == kHllcC) {  // calcFluxCoupled boundary Jacobians (JoeWithModels.cpp:5150-5740)
        const double hb = bb[6];
        if (z.bc == kBcSymmetry || z.bc == kBcWall || z.bc == kBcNeumann) {
          if (z.bc != kBcNeumann && P.symm_bc == kSymmAnalytical) {
            jac_only_pressure(J, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int i = 0; i < 5; i++)
              for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
          } else if (P.symm_bc == kSymmNew) {
            double K[5][5];
            jac_hllc_coupled(J, 0, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            jac_hllc_coupled(K, 1, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int i = 0; i < 5; i++)
              for (int j = 0; j < 5; j++) D[i][j] += J[i][j] + K[i][j];
          } else {
            jac_hllc_coupled(J, 0, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int i = 0; i < 5; i++)
              for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
          }
        } else {  // cell-centre state against the boundary state
          jac_hllc_coupled(J, 0, own.rho, own.u, own.p, enthalpy_of(P, own.rho, own.p), sb.rho, sb.u, sb.p, hb, gam,
                           area, n);
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
        }
        if (VISC && z.bc != kBcSymmetry) {  // the face's viscous block A0 (no implicit term on symmetry faces)
          const double *gvo = gradv + 12 * (size_t)c;
          const VSide vo = vside_cell(P, pc, gvo);
          const VSide vb = vside_bfa(bb, gvo);
          const double xfa[3] = {geo[4], geo[5], geo[6]};
          double sArg[3], smag, alpha;
          visc_coupled_bgeom(z.bc, xfa, pc + 5, n, sArg, smag, alpha);
          const double mul = mu_lam(P, bb[5]);
          const double mut = (P.sa && z.bc != kBcWall)
                                 ? sa_mut(pc[0] * P.sa_phi[c], P.sa_aux[kAuxN * (size_t)c + kAuxMue])
                                 : 0.0;
          const VSide l = vsel(z.bc == kBcNeumann, vb, vo);
          double VF[5], tau[3], keff;
          visc_coupled(P, l, vb, mul, mut, vb.u, area, n, smag, sArg, alpha, VF, tau, keff, P.sa != 0);
          visc_coupled_jac(P, J, 0, l, mul + mut, keff, tau, vb.u, area, alpha, smag);
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
        }
      } else if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        if (z.bc == kBcWall || P.symm_bc == kSymmOriginal) {
          face_jac(SCHEME, J, 0, sb, sb, gam, area, n);
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
        } else if (P.symm_bc == kSymmNew) {
          double K[5][5];
          face_jac(SCHEME, J, 0, sb, sb, gam, area, n);
          face_jac(SCHEME, K, 1, sb, sb, gam, area, n);
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] += J[i][j] + K[i][j];
        } else {  // ANALYTICAL, JoeWithModels.cpp:3243-3269
          double a2 = gam - 1.0;
          double phi2 = 0.5 * a2 * (sb.u[0] * sb.u[0] + sb.u[1] * sb.u[1] + sb.u[2] * sb.u[2]);
          for (int i = 0; i < 3; i++) {
            D[i + 1][0] += phi2 * n[i] * area;
            for (int j = 0; j < 3; j++) D[i + 1][j + 1] += -a2 * sb.u[j] * n[i] * area;
            D[i + 1][4] += a2 * n[i] * area;
          }
          for (int j = 0; j < 5; j++) {
            D[0][j] += 0.0;
            D[4][j] += 0.0;
          }
        }
      } else {
        face_jac(SCHEME, J, 0, own, sb, gam, area, n);
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
      }
    }
  }
  if (VISC && SCHEME != kHllcC) {  // viscous Jacobian blocks, after the Euler ones (JoeWithModels.cpp:3680-3697, 3789-3796, 3835-3842)
    const VSide vo = vside_cell(P, pc, gradv + 12 * (size_t)c);
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      int f = en.x & 0x7fffffff;
      const double *geo = m.fgeo + 8 * (size_t)f;
      const double n[3] = {geo[0], geo[1], geo[2]};
      const double area = geo[3];
      const double xfa[3] = {geo[4], geo[5], geo[6]};
      double Fu[3], FE;
      VFace vf;
      if (en.y >= 0) {
        const double *pn = prim + 8 * (size_t)en.y;
        const VSide vn = vside_cell(P, pn, gradv + 12 * (size_t)en.y);
        const bool s0 = en.x >= 0;
        const VSide &a = s0 ? vo : vn, &b = s0 ? vn : vo;
        double sV[3], smag, mul, w0, w1;
        visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, a.T, b.T, sV, smag, mul, w0, w1);
        double mut = 0.0;
        if (P.sa) {
          const double nuo = P.sa_phi[c], nun = P.sa_phi[en.y];
          mut = sa_mut_face(w0, w1, a.rho, b.rho, s0 ? nuo : nun, s0 ? nun : nuo, mul);
        }
        const double ufa[3] = {0.5 * (a.u[0] + b.u[0]), 0.5 * (a.u[1] + b.u[1]), 0.5 * (a.u[2] + b.u[2])};
        visc_face(P, a, b, mul, mut, ufa, area, n, smag, sV, Fu, FE, vf);
        T *off = A + 25 * (size_t)(m.cfl_noc[e] & 0x3fffffff);
        // icv0: A[noc00] -= A0, A[noc01] -= A1; icv1: A[noc11] += A1,
        // A[noc10] += A0 (each block evaluated once, routed by the lane's side)
        visc_jac(P, J, -1.0, a, mul, mut, vf, ufa, area, smag);  // A0
        if (s0) {
          for (int i = 1; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] -= J[i][j];
        } else {
          for (int i = 1; i < 5; i++)
            for (int j = 0; j < 5; j++) off[5 * i + j] = (T)((double)off[5 * i + j] + J[i][j]);
        }
        visc_jac(P, J, 1.0, b, mul, mut, vf, ufa, area, smag);  // A1
        if (s0) {
          for (int i = 1; i < 5; i++)
            for (int j = 0; j < 5; j++) off[5 * i + j] = (T)((double)off[5 * i + j] - J[i][j]);
        } else {
          for (int i = 1; i < 5; i++)
            for (int j = 0; j < 5; j++) D[i][j] += J[i][j];
        }
      } else {
        const ZoneDev &z = m.zones[-en.y - 1];
        if (z.bc == kBcSymmetry) continue;
        const double *bb = bfa + 8 * (size_t)f;
        const double mul = mu_lam(P, bb[5]);
        const VSide vb = vside_bfa(bb, gradv + 12 * (size_t)c);
        double sV[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
        double smag;
        double sArg[3];  // a copy, not a pointer to n or sV, keeps both in registers
        if (z.bc == kBcWall) {
          smag = fabs(dot3(sV, n));
          for (int i = 0; i < 3; i++) sArg[i] = n[i];
        } else {
          smag = sqrt(dot3(sV, sV));
          for (int i = 0; i < 3; i++) sArg[i] = sV[i] / smag;
        }
        const double mut = (P.sa && z.bc != kBcWall) ? sa_mut(pc[0] * P.sa_phi[c], P.sa_aux[kAuxN * (size_t)c + kAuxMue])
                                                      : 0.0;
        visc_face(P, vo, vb, mul, mut, vb.u, area, n, smag, sArg, Fu, FE, vf);
        visc_jac(P, J, -1.0, vo, mul, mut, vf, vb.u, area, smag);
        for (int i = 1; i < 5; i++)
          for (int j = 0; j < 5; j++) D[i][j] -= J[i][j];
      }
    }
  }
  const double vol = m.vol[c];
  if (P.gravity_on) {
    D[1][0] -= vol * P.g[0];
    D[2][0] -= vol * P.g[1];
    D[3][0] -= vol * P.g[2];
    D[4][1] -= vol * P.g[0];
    D[4][2] -= vol * P.g[1];
    D[4][3] -= vol * P.g[2];
  }
  if (step_mode) {
    double tmp = vol / (dt[c]);
    for (int i = 0; i < 5; i++) D[i][i] += tmp;
    for (int k = 0; k < 5; k++) {
      rhs_ls[NB * (size_t)c + k] = (T)(relax * rhs[(size_t)k * m.ncv + c]);
      phi[NB * (size_t)c + k] = (T)0.0;
    }
  }
  T *d = A + NB * NB * (size_t)row;
  for (int i = 0; i < 5; i++) {
    for (int j = 0; j < 5; j++) d[NB * i + j] = (T)D[i][j];
    if (NB == 6) d[NB * i + 5] = (T)0.0;
  }
}

template <typename T, int SCHEME, bool VISC, int NB = 5>
__global__ void __launch_bounds__(kBlock) k_jac(DMesh m, DParams P, const double *__restrict__ prim,
                                                const double *__restrict__ bfa, T *__restrict__ A, int step_mode,
                                                const double *__restrict__ dt, const double *__restrict__ rhs,
                                                double relax, T *__restrict__ rhs_ls, T *__restrict__ phi,
                                                const double *__restrict__ gradv) {
  d_jac<T, SCHEME, VISC, NB>(hw_blk(), m, P, prim, bfa, A, step_mode, dt, rhs, relax, rhs_ls, phi, gradv);
}

// ---------------------------------------------------------------------------
// The same Jacobian assembly with one lane per cell-face entry (the flux
// kernel's wave table: whole cells per wavefront). Each lane evaluates its
// face's two blocks once: the off-diagonal one goes straight to A, the
// diagonal contribution (signed: D -= J is D + (-J) exactly) to the lane's LDS
// slot. The lanes of a cell then sum the diagonal block component-wise in the
// cell's entry order (lane pos owns components pos, pos + nent, ...), which is
// the reference's accumulation order, so the blocks are the same bits as
// k_jac's. The segregated viscous blocks are a second pass over the same
// slots (the reference adds them after every Euler block); the coupled loop
// adds each face's viscous block right after its Euler block, so there the
// slot also keeps the viscous block's eleven possibly nonzero entries.
// Duplicate neighbour pairs (two faces between one pair of cells) need their
// off-diagonal block accumulated in entry order and use k_jac.
// ---------------------------------------------------------------------------
// the entries of a coupled viscous block (visc_coupled_jac) that may be
// nonzero: [1+i][0], [1+i][1+i] for i < 3 and row 4; -1 for the others
__device__ __forceinline__ int vc_index(int k) {
  switch (k) {
    case 5: return 0;
    case 6: return 1;
    case 10: return 2;
    case 12: return 3;
    case 15: return 4;
    case 18: return 5;
    case 20: return 6;
    case 21: return 7;
    case 22: return 8;
    case 23: return 9;
    case 24: return 10;
    default: return -1;
  }
}
constexpr int kVcN = 11;
constexpr int kMaxOwned = 7;  // components of the diagonal block per lane (cells have at least 4 faces)

// SST: the cells' (or the boundary's) turbulent kinetic energy in the HLLC
// Jacobian (kineCV / kineFA, JoeWithModels.cpp:3034-3037, 3202-3206, 3490-3495)
// and in the viscous blocks' dU/dQ; the SST eddy viscosity
template <typename T, int SCHEME, bool VISC, int NB = 5, int W = kWaves, bool SST = false>
__device__ __forceinline__ void d_jac_e(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                        const double *__restrict__ bfa, T *__restrict__ A, int step_mode,
                                        const double *__restrict__ dt, const double *__restrict__ rhs, double relax,
                                        T *__restrict__ rhs_ls, T *__restrict__ phi,
                                        const double *__restrict__ gradv) {
  constexpr bool CPL = SCHEME == kHllcC;
  constexpr bool CV = VISC && CPL;
  constexpr int NS = CV ? 25 + kVcN : 25;
  __shared__ double lds[64 * W * NS];
  __shared__ signed char vsg[64 * W];  // the entry's viscous block: 0 none, +1 added, -1 subtracted
  double *slot = lds + NS * threadIdx.x;
  const EntryLane L = entry_lane<W>(m, B);
  const double gam = P.gamma;
  const int c = L.c, pos = L.pos;
  const int nent = L.e >= 0 ? m.cfl_i[c + 1] - m.cfl_i[c] : 0;
  const double *pc = prim + 8 * (size_t)(L.e >= 0 ? c : 0);
  Side own;
  own.rho = pc[0]; own.u[0] = pc[1]; own.u[1] = pc[2]; own.u[2] = pc[3]; own.p = pc[4];
  double J[5][5];
  if (L.e >= 0) {
    const int2 en = m.cfl[L.e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double area = geo[3];
    signed char vs = 0;
    if (en.y >= 0) {
      const double *pn = prim + 8 * (size_t)en.y;
      Side nb;
      nb.rho = pn[0]; nb.u[0] = pn[1]; nb.u[1] = pn[2]; nb.u[2] = pn[3]; nb.p = pn[4];
      const bool s0 = en.x >= 0;
      const Side l = s0 ? own : nb, r = s0 ? nb : own;
      const double kcl = SST ? P.sst_phi[s0 ? c : en.y] : 0.0, kcr = SST ? P.sst_phi[s0 ? en.y : c] : 0.0;
      double hl = 0.0, hr = 0.0;
      if (CPL) {
        hl = enthalpy_of(P, l.rho, l.p);
        hr = enthalpy_of(P, r.rho, r.p);
      }
      // t = 0: the diagonal's block (icv0 adds Apl, which 0; icv1 subtracts
      // Ami, which 1); t = 1: the off-diagonal block (icv0's A[noc01] = +Ami,
      // icv1's A[noc10] = -Apl). One loop, so one copy of the Jacobian code.
      T *off = A + NB * NB * (size_t)(m.cfl_noc[L.e] & 0x3fffffff);
#pragma unroll 1
      for (int t = 0; t < 2; t++) {
        const int which = (t == 0) == s0 ? 0 : 1;
        if (CPL) jac_hllc_coupled(J, which, l.rho, l.u, l.p, hl, r.rho, r.u, r.p, hr, gam, area, n);
        else face_jac(SCHEME, J, which, l, r, gam, area, n, kcl, kcr);
        if (t == 0) {
          for (int i = 0; i < 5; i++)
            for (int j = 0; j < 5; j++) slot[5 * i + j] = s0 ? J[i][j] : -J[i][j];
        } else {
          store_block<T, NB>(off, J, !s0, false);
        }
      }
      if (CV) {  // the same face's viscous blocks next (JoeWithModels.cpp:5096-5104)
        const VSide vo = vside_cell(P, pc, gradv + 12 * (size_t)c);
        const VSide vn = vside_cell(P, pn, gradv + 12 * (size_t)en.y);
        const VSide a = vsel(s0, vo, vn), b = vsel(s0, vn, vo);
        const double xa[3] = {s0 ? pc[5] : pn[5], s0 ? pc[6] : pn[6], s0 ? pc[7] : pn[7]};
        const double xb[3] = {s0 ? pn[5] : pc[5], s0 ? pn[6] : pc[6], s0 ? pn[7] : pc[7]};
        const double xfa[3] = {geo[4], geo[5], geo[6]};
        double sV[3], smag, mul, w0, w1;
        visc_internal_geom(P, xa, xb, xfa, a.T, b.T, sV, smag, mul, w0, w1);
        const double alpha = dot3(n, sV);
        double mut = 0.0;
        if (P.sa) {
          const double nuo = P.sa_phi[c], nun = P.sa_phi[en.y];
          mut = sa_mut_face(w0, w1, a.rho, b.rho, s0 ? nuo : nun, s0 ? nun : nuo, mul);
        }
        const double ufa[3] = {0.5 * (a.u[0] + b.u[0]), 0.5 * (a.u[1] + b.u[1]), 0.5 * (a.u[2] + b.u[2])};
        double VF[5], tau[3], keff;
        visc_coupled(P, a, b, mul, mut, ufa, area, n, smag, sV, alpha, VF, tau, keff, P.sa != 0);
#pragma unroll 1
        for (int t = 0; t < 2; t++) {
          // t = 0: the diagonal's block from this cell's side; t = 1: the other side's, off the diagonal
          visc_coupled_jac(P, J, (t == 0) == s0 ? 0 : 1, vsel((t == 0) == s0, a, b), mul + mut, keff, tau, ufa, area,
                           alpha, smag);
          if (t == 0) {
            for (int k = 0; k < 25; k++) {
              const int vi = vc_index(k);
              if (vi >= 0) slot[25 + vi] = s0 ? J[k / 5][k % 5] : -J[k / 5][k % 5];
            }
          } else {
            for (int i = 0; i < 5; i++)
              for (int j = 0; j < 5; j++)
                off[NB * i + j] =
                    (T)(s0 ? (double)off[NB * i + j] + J[i][j] : (double)off[NB * i + j] - J[i][j]);
          }
        }
        vs = s0 ? 1 : -1;
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      Side sb;
      sb.rho = bb[0]; sb.u[0] = bb[1]; sb.u[1] = bb[2]; sb.u[2] = bb[3]; sb.p = bb[4];
      if (CPL) {  // calcFluxCoupled boundary Jacobians (JoeWithModels.cpp:5150-5740)
        const double hb = bb[6];
        if (z.bc == kBcSymmetry || z.bc == kBcWall || z.bc == kBcNeumann) {
          if (z.bc != kBcNeumann && P.symm_bc == kSymmAnalytical) {
            jac_only_pressure(J, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5];
          } else if (P.symm_bc == kSymmNew) {
            double K[5][5];
            jac_hllc_coupled(J, 0, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            jac_hllc_coupled(K, 1, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5] + K[k / 5][k % 5];
          } else {
            jac_hllc_coupled(J, 0, sb.rho, sb.u, sb.p, hb, sb.rho, sb.u, sb.p, hb, gam, area, n);
            for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5];
          }
        } else {  // cell-centre state against the boundary state
          jac_hllc_coupled(J, 0, own.rho, own.u, own.p, enthalpy_of(P, own.rho, own.p), sb.rho, sb.u, sb.p, hb, gam,
                           area, n);
          for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5];
        }
        if (CV && z.bc != kBcSymmetry) {  // the face's viscous block (none on symmetry faces)
          const double *gvo = gradv + 12 * (size_t)c;
          const VSide vo = vside_cell(P, pc, gvo);
          const VSide vb = vside_bfa(bb, gvo);
          const double xfa[3] = {geo[4], geo[5], geo[6]};
          double sArg[3], smag, alpha;
          visc_coupled_bgeom(z.bc, xfa, pc + 5, n, sArg, smag, alpha);
          const double mul = mu_lam(P, bb[5]);
          const double mut = (P.sa && z.bc != kBcWall)
                                 ? sa_mut(pc[0] * P.sa_phi[c], P.sa_aux[kAuxN * (size_t)c + kAuxMue])
                                 : 0.0;
          const VSide lv = vsel(z.bc == kBcNeumann, vb, vo);
          double VF[5], tau[3], keff;
          visc_coupled(P, lv, vb, mul, mut, vb.u, area, n, smag, sArg, alpha, VF, tau, keff, P.sa != 0);
          visc_coupled_jac(P, J, 0, lv, mul + mut, keff, tau, vb.u, area, alpha, smag);
          for (int k = 0; k < 25; k++) {
            const int vi = vc_index(k);
            if (vi >= 0) slot[25 + vi] = J[k / 5][k % 5];
          }
          vs = 1;
        }
      } else if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        const double kb = SST ? P.sst_bfa[2 * (size_t)f] : 0.0;
        if (z.bc == kBcWall || P.symm_bc == kSymmOriginal) {
          face_jac(SCHEME, J, 0, sb, sb, gam, area, n, kb, kb);
          for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5];
        } else if (P.symm_bc == kSymmNew) {
          double K[5][5];
          face_jac(SCHEME, J, 0, sb, sb, gam, area, n, kb, kb);
          face_jac(SCHEME, K, 1, sb, sb, gam, area, n, kb, kb);
          for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5] + K[k / 5][k % 5];
        } else {  // ANALYTICAL, JoeWithModels.cpp:3243-3269
          const double a2 = gam - 1.0;
          const double phi2 = 0.5 * a2 * (sb.u[0] * sb.u[0] + sb.u[1] * sb.u[1] + sb.u[2] * sb.u[2]);
          for (int j = 0; j < 5; j++) {
            slot[j] = 0.0;
            slot[20 + j] = 0.0;
          }
          for (int i = 0; i < 3; i++) {
            slot[5 * (i + 1)] = phi2 * n[i] * area;
            for (int j = 0; j < 3; j++) slot[5 * (i + 1) + j + 1] = -a2 * sb.u[j] * n[i] * area;
            slot[5 * (i + 1) + 4] = a2 * n[i] * area;
          }
        }
      } else {
        if (SST) face_jac(SCHEME, J, 0, own, sb, gam, area, n, P.sst_phi[c], P.sst_bfa[2 * (size_t)f]);
        else face_jac(SCHEME, J, 0, own, sb, gam, area, n);
        for (int k = 0; k < 25; k++) slot[k] = J[k / 5][k % 5];
      }
    }
    vsg[threadIdx.x] = vs;
  }
  __syncthreads();
  // the cell's diagonal block, summed in entry order by the cell's lanes
  double Dk[kMaxOwned];
  const int first = threadIdx.x - pos;
  if (L.e >= 0) {
    int q = 0;
    for (int k = pos; k < 25; k += nent, q++) {
      double d = 0.0;
      for (int j = 0; j < nent; j++) {
        const double *sj = lds + NS * (first + j);
        d += sj[k];
        if (CV) {
          const int sg = vsg[first + j];
          if (sg != 0) {
            const int vi = vc_index(k);
            d += vi >= 0 ? sj[25 + vi] : (sg > 0 ? 0.0 : -0.0);
          }
        }
      }
      Dk[q] = d;
    }
  }
  if (VISC && !CPL) {  // viscous blocks after every Euler block (JoeWithModels.cpp:3680-3697, 3789-3796, 3835-3842)
    __syncthreads();
    if (L.e >= 0) {
      const int2 en = m.cfl[L.e];
      const int f = en.x & 0x7fffffff;
      const double *geo = m.fgeo + 8 * (size_t)f;
      const double n[3] = {geo[0], geo[1], geo[2]};
      const double area = geo[3];
      const double xfa[3] = {geo[4], geo[5], geo[6]};
      VSide vo = vside_cell(P, pc, gradv + 12 * (size_t)c);
      if (SST) vo.k = P.sst_phi[c];
      double Fu[3], FE;
      VFace vf;
      signed char vs = 0;
      if (en.y >= 0) {
        const double *pn = prim + 8 * (size_t)en.y;
        VSide vn = vside_cell(P, pn, gradv + 12 * (size_t)en.y);
        if (SST) vn.k = P.sst_phi[en.y];
        const bool s0 = en.x >= 0;
        const VSide &a = s0 ? vo : vn, &b = s0 ? vn : vo;
        double sV[3], smag, mul, w0, w1;
        visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, a.T, b.T, sV, smag, mul, w0, w1);
        double mut = 0.0, kine_fa = 0.0;
        if (SST) {
          mut = sst_mut_face(P, m.ncv, w0, w1, a.rho, b.rho, s0 ? c : en.y, s0 ? en.y : c);
          kine_fa = (w1 * a.k + w0 * b.k) / (w0 + w1);
        } else if (P.sa) {
          const double nuo = P.sa_phi[c], nun = P.sa_phi[en.y];
          mut = sa_mut_face(w0, w1, a.rho, b.rho, s0 ? nuo : nun, s0 ? nun : nuo, mul);
        }
        const double ufa[3] = {0.5 * (a.u[0] + b.u[0]), 0.5 * (a.u[1] + b.u[1]), 0.5 * (a.u[2] + b.u[2])};
        visc_face<SST>(P, a, b, mul, mut, ufa, area, n, smag, sV, Fu, FE, vf, kine_fa);
        T *off = A + NB * NB * (size_t)(m.cfl_noc[L.e] & 0x3fffffff);
        // icv0: A[noc00] -= A0, A[noc01] -= A1; icv1: A[noc11] += A1, A[noc10] += A0
        if (s0) visc_jac<SST>(P, J, 1.0, b, mul, mut, vf, ufa, area, smag);  // A1 -> off
        else visc_jac<SST>(P, J, -1.0, a, mul, mut, vf, ufa, area, smag);    // A0 -> off
        for (int i = 1; i < 5; i++)
          for (int j = 0; j < 5; j++)
            off[NB * i + j] = (T)(s0 ? (double)off[NB * i + j] - J[i][j] : (double)off[NB * i + j] + J[i][j]);
        if (s0) visc_jac<SST>(P, J, -1.0, a, mul, mut, vf, ufa, area, smag);  // A0 -> diagonal, subtracted
        else visc_jac<SST>(P, J, 1.0, b, mul, mut, vf, ufa, area, smag);     // A1 -> diagonal, added
        for (int i = 1; i < 5; i++)
          for (int j = 0; j < 5; j++) slot[5 * i + j] = s0 ? -J[i][j] : J[i][j];
        vs = 1;
      } else {
        const ZoneDev &z = m.zones[-en.y - 1];
        if (z.bc != kBcSymmetry) {
          const double *bb = bfa + 8 * (size_t)f;
          const double mul = mu_lam(P, bb[5]);
          VSide vb = vside_bfa(bb, gradv + 12 * (size_t)c);
          double kine_fa = 0.0;
          if (SST) {  // walls: kine1 = kine_fa = 0; other zones phi_bfa
            vb.k = z.bc == kBcWall ? 0.0 : P.sst_bfa[2 * (size_t)f];
            kine_fa = vb.k;
          }
          double sV[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          double smag;
          double sArg[3];
          if (z.bc == kBcWall) {
            smag = fabs(dot3(sV, n));
            for (int i = 0; i < 3; i++) sArg[i] = n[i];
          } else {
            smag = sqrt(dot3(sV, sV));
            for (int i = 0; i < 3; i++) sArg[i] = sV[i] / smag;
          }
          const double mut = z.bc == kBcWall ? 0.0
                             : SST       ? sst_mut_cell(P, m.ncv, pc[0], c)
                             : P.sa      ? sa_mut(pc[0] * P.sa_phi[c], P.sa_aux[kAuxN * (size_t)c + kAuxMue])
                                         : 0.0;
          visc_face<SST>(P, vo, vb, mul, mut, vb.u, area, n, smag, sArg, Fu, FE, vf, kine_fa);
          visc_jac<SST>(P, J, -1.0, vo, mul, mut, vf, vb.u, area, smag);
          for (int i = 1; i < 5; i++)
            for (int j = 0; j < 5; j++) slot[5 * i + j] = -J[i][j];
          vs = 1;
        }
      }
      vsg[threadIdx.x] = vs;
    }
    __syncthreads();
    if (L.e >= 0) {
      int q = 0;
      for (int k = pos; k < 25; k += nent, q++) {
        if (k < 5) continue;  // rows 1-4 only
        double d = Dk[q];
        for (int j = 0; j < nent; j++)
          if (vsg[first + j]) d += lds[NS * (first + j) + k];
        Dk[q] = d;
      }
    }
  }
  if (L.e < 0) return;
  // sourceHook (gravity), the time term, and the row's diagonal block
  const double vol = m.vol[c];
  const double tmp = step_mode ? vol / (dt[c]) : 0.0;
  T *dgl = A + NB * NB * (size_t)m.nbocv_i[c];
  int q = 0;
  for (int k = pos; k < 25; k += nent, q++) {
    const int i = k / 5, j = k % 5;
    double d = Dk[q];
    if (P.gravity_on) {
      if (k == 5) d -= vol * P.g[0];
      else if (k == 10) d -= vol * P.g[1];
      else if (k == 15) d -= vol * P.g[2];
      else if (k == 21) d -= vol * P.g[0];
      else if (k == 22) d -= vol * P.g[1];
      else if (k == 23) d -= vol * P.g[2];
    }
    if (step_mode && i == j) d += tmp;
    dgl[NB * i + j] = (T)d;
    if (NB == 6 && j == 4) dgl[NB * i + 5] = (T)0.0;
  }
  if (step_mode)
    for (int k = pos; k < 5; k += nent) {
      rhs_ls[NB * (size_t)c + k] = (T)(relax * rhs[(size_t)k * m.ncv + c]);
      phi[NB * (size_t)c + k] = (T)0.0;
    }
}

// one wavefront per block, as the flux kernel (the diagonal sums wait only
// for their own wavefront)
constexpr int kJacWaves = 1;

template <typename T, int SCHEME, bool VISC, int NB = 5, bool SST = false>
__global__ void __launch_bounds__(64 * kJacWaves) k_jac_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                  const double *__restrict__ bfa, T *__restrict__ A, int step_mode,
                                                  const double *__restrict__ dt, const double *__restrict__ rhs,
                                                  double relax, T *__restrict__ rhs_ls, T *__restrict__ phi,
                                                  const double *__restrict__ gradv) {
  d_jac_e<T, SCHEME, VISC, NB, kJacWaves, SST>(hw_blk(), m, P, prim, bfa, A, step_mode, dt, rhs, relax, rhs_ls,
                                              phi, gradv);
}

// ---------------------------------------------------------------------------
// Spalart-Allmaras per-cell fields after k_grad: the limited nuSA gradient
// (calcCvScalarGrad with limiterNavierS and reference nuSA; JoeWithModels.cpp
// :2736-2741, Scalars.cpp:134), the face-averaged laminar viscosity in faocv
// order (InterpolateAtCellCenterFromFaceValues, UgpWithCvCompFlow.h:1538-1551)
// and the strain and vorticity magnitudes of the limited velocity gradient
// (calcGradVel, calcStrainRate, calcVorticity, :1305-1392).
// ---------------------------------------------------------------------------
template <bool LS>
__global__ void __launch_bounds__(kBlock) k_sa_cell(DMesh m, DParams P, const double *__restrict__ prim,
                                                    const double *__restrict__ bfa, const double *__restrict__ grad,
                                                    double *__restrict__ aux) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= m.ncv) return;
  const int c = t;
  const double *pc = prim + 8 * (size_t)c;
  const double *phi = P.sa_phi;
  const double ph = phi[c];
  const double xcv[3] = {pc[5], pc[6], pc[7]};
  double g[3];
  double pmax = ph, pmin = ph;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  int e_int = e0;
  if (LS) {
    const int noc_f = m.nbocv_i[c];
    const double *l0 = m.lsg + 3 * (size_t)noc_f;
    for (int j = 0; j < 3; j++) g[j] = l0[j] * ph;
    int k = 1;
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      if (en.y < 0) break;
      const double pn = phi[en.y];
      const double *l = m.lsg + 3 * (size_t)(noc_f + k);
      for (int j = 0; j < 3; j++) g[j] += l[j] * pn;
      pmax = Dev::mx(pmax, pn);
      pmin = Dev::mn(pmin, pn);
      k++;
      e_int = e + 1;
    }
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      if (en.y >= 0) continue;
      int f = en.x & 0x7fffffff;
      const double *bl = m.bfa_lsg + 3 * (size_t)f;
      const double pb = bfa[8 * (size_t)f + 7];
      for (int j = 0; j < 3; j++) g[j] += bl[j] * pb;
    }
  } else {
    for (int j = 0; j < 3; j++) g[j] = 0.0;
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      int f = en.x & 0x7fffffff;
      const double *fn = m.fnorm + 3 * (size_t)f;
      if (en.y >= 0) {
        const double *pn = prim + 8 * (size_t)en.y;
        const double phn = phi[en.y];
        const double *xf = m.fgeo + 8 * (size_t)f + 4;
        double ds[3], dn[3];
        for (int j = 0; j < 3; j++) {
          ds[j] = xf[j] - xcv[j];
          dn[j] = xf[j] - pn[5 + j];
        }
        const bool s0 = (en.x >= 0);
        const double wself = sqrt(dot3(ds, ds)), wnb = sqrt(dot3(dn, dn));
        const double w0 = s0 ? wself : wnb, w1 = s0 ? wnb : wself;
        const double fv = s0 ? (w1 * ph + w0 * phn) / (w0 + w1) : (w1 * phn + w0 * ph) / (w0 + w1);
        for (int j = 0; j < 3; j++) {
          if (s0)
            g[j] += fv * fn[j];
          else
            g[j] -= fv * fn[j];
        }
        pmax = Dev::mx(pmax, phn);
        pmin = Dev::mn(pmin, phn);
        e_int = e + 1;
      } else {
        const double pb = bfa[8 * (size_t)f + 7];
        for (int j = 0; j < 3; j++) g[j] += pb * fn[j];
      }
    }
    const double vol = m.vol[c];
    for (int j = 0; j < 3; j++) g[j] /= vol;
  }
  double al = 1.0;
  if (P.limiter) {  // limitCvScalarGradBJ (UgpWithCv.h:881-937), reference value nuSA
    double e2 = Dev::mx(P.eps * fabs(ph), 1.0e-8);
    e2 = e2 * e2;
    if ((pmax - pmin) > 1.0e-12) {
      for (int e = e0; e < e_int; e++) {
        const double *rr = m.cfl_r0 + 3 * (size_t)e;
        double r0[3] = {rr[0], rr[1], rr[2]};
        al = bj_face(al, ph, ph + dot3(r0, g), pmax, pmin, e2);
      }
    } else {
      al = 0.0;
    }
  }
  // face-averaged laminar viscosity over the faocv list
  const double Tc = pc[4] / (pc[0] * P.R);
  const int nf = e1 - e0;
  double mue = 0.0;
  for (int k = 0; k < nf; k++) {
    const int e = e0 + m.foc[e0 + k];
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    double mul;
    if (en.y >= 0) {
      const double *pn = prim + 8 * (size_t)en.y;
      const double Tn = pn[4] / (pn[0] * P.R);
      const bool s0 = (en.x >= 0);
      double sV[3], smag, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, m.fgeo + 8 * (size_t)f + 4, s0 ? Tc : Tn,
                         s0 ? Tn : Tc, sV, smag, mul, w0, w1);
    } else {
      mul = mu_lam(P, bfa[8 * (size_t)f + 5]);
    }
    mue += mul;
  }
  mue /= (double)nf;
  // strain and vorticity magnitudes (pow(x, 2.0) is x * x)
  const double *gu = grad + 16 * (size_t)c + 6;
  double str = 0.0, vort = 0.0;
  const double diverg = gu[0] + gu[4] + gu[8];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = i == j ? 0.5 * (gu[3 * i + j] + gu[3 * j + i]) - 1. / 3. * diverg : 0.5 * (gu[3 * i + j] + gu[3 * j + i]);
      str += v * v;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0.5 * (gu[3 * i + j] - gu[3 * j + i]);
      vort += v * v;
    }
  double *o = aux + kAuxN * (size_t)c;
  o[kAuxGrad] = g[0] * al;
  o[kAuxGrad + 1] = g[1] * al;
  o[kAuxGrad + 2] = g[2] * al;
  o[kAuxMue] = mue;
  o[kAuxStr] = sqrt(2.0 * str);
  o[kAuxVort] = sqrt(2.0 * vort);
}

// ---------------------------------------------------------------------------
// SST boundary values of kine and omega: setScalarBC (Scalars.cpp:289-360) with
// the omega wall hook 60 mu / (rho beta_1 d^2) (boundaryHookScalarRansTurb,
// TurbModel_KOMSST.h:342-358). After k_bc, whose boundary velocity decides
// inflow (vel_bfa . fa_normal <= 1e-8).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_sst_bc(DMesh m, DParams P, const double *__restrict__ prim,
                                                   const double *__restrict__ bfa, double *__restrict__ sbfa) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= m.nfa_b) return;
  const int c = m.bfa_cell[f];
  const ZoneDev &z = m.zones[m.bfa_zone[f]];
  const double *bb = bfa + 8 * (size_t)f;
  const double vb[3] = {bb[1], bb[2], bb[3]};
  const bool inflow = !(dot3(vb, m.fnorm + 3 * (size_t)f) > 1.0e-8);
  sbfa[2 * (size_t)f] = (z.skind && inflow) ? z.sval : P.sst_phi[c];
  double w;
  if (z.skind2 == 2) {
    const double *pc = prim + 8 * (size_t)c;
    const double muLamCV = mu_lam(P, pc[4] / (pc[0] * P.R));
    const double d = P.sst_aux[kSstN * (size_t)c + kSWall];
    w = 60.0 * muLamCV / (pc[0] * kSstBeta1 * d * d);
  } else {
    w = (z.skind2 == 1 && inflow) ? z.sval2 : P.sst_phi[m.ncv + c];
  }
  sbfa[2 * (size_t)f + 1] = w;
}

// the limited gradient of one SST scalar with its boundary values
// (calcCvScalarGrad with limiterNavierS and the scalar as its own reference:
// UgpWithCv.h:881-937, 1009-1188), as k_sa_cell forms nuSA's
template <bool LS>
__device__ __forceinline__ void sst_grad(const DMesh &m, const DParams &P, const double *__restrict__ prim,
                                         const double *__restrict__ phi, int s, int c, double gout[3]) {
  const double *pc = prim + 8 * (size_t)c;
  const double ph = phi[c];
  const double xcv[3] = {pc[5], pc[6], pc[7]};
  double g[3];
  double pmax = ph, pmin = ph;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  int e_int = e0;
  if (LS) {
    const int noc_f = m.nbocv_i[c];
    const double *l0 = m.lsg + 3 * (size_t)noc_f;
    for (int j = 0; j < 3; j++) g[j] = l0[j] * ph;
    int k = 1;
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      if (en.y < 0) break;
      const double pn = phi[en.y];
      const double *l = m.lsg + 3 * (size_t)(noc_f + k);
      for (int j = 0; j < 3; j++) g[j] += l[j] * pn;
      pmax = Dev::mx(pmax, pn);
      pmin = Dev::mn(pmin, pn);
      k++;
      e_int = e + 1;
    }
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      if (en.y >= 0) continue;
      int f = en.x & 0x7fffffff;
      const double *bl = m.bfa_lsg + 3 * (size_t)f;
      const double pb = P.sst_bfa[2 * (size_t)f + s];
      for (int j = 0; j < 3; j++) g[j] += bl[j] * pb;
    }
  } else {
    for (int j = 0; j < 3; j++) g[j] = 0.0;
    for (int e = e0; e < e1; e++) {
      int2 en = m.cfl[e];
      int f = en.x & 0x7fffffff;
      const double *fn = m.fnorm + 3 * (size_t)f;
      if (en.y >= 0) {
        const double *pn = prim + 8 * (size_t)en.y;
        const double phn = phi[en.y];
        const double *xf = m.fgeo + 8 * (size_t)f + 4;
        double ds[3], dn[3];
        for (int j = 0; j < 3; j++) {
          ds[j] = xf[j] - xcv[j];
          dn[j] = xf[j] - pn[5 + j];
        }
        const bool s0 = (en.x >= 0);
        const double wself = sqrt(dot3(ds, ds)), wnb = sqrt(dot3(dn, dn));
        const double w0 = s0 ? wself : wnb, w1 = s0 ? wnb : wself;
        const double fv = s0 ? (w1 * ph + w0 * phn) / (w0 + w1) : (w1 * phn + w0 * ph) / (w0 + w1);
        for (int j = 0; j < 3; j++) {
          if (s0)
            g[j] += fv * fn[j];
          else
            g[j] -= fv * fn[j];
        }
        pmax = Dev::mx(pmax, phn);
        pmin = Dev::mn(pmin, phn);
        e_int = e + 1;
      } else {
        const double pb = P.sst_bfa[2 * (size_t)f + s];
        for (int j = 0; j < 3; j++) g[j] += pb * fn[j];
      }
    }
    const double vol = m.vol[c];
    for (int j = 0; j < 3; j++) g[j] /= vol;
  }
  double al = 1.0;
  if (P.limiter) {
    double e2 = Dev::mx(P.eps * fabs(ph), 1.0e-8);
    e2 = e2 * e2;
    if ((pmax - pmin) > 1.0e-12) {
      for (int e = e0; e < e_int; e++) {
        const double *rr = m.cfl_r0 + 3 * (size_t)e;
        double r0[3] = {rr[0], rr[1], rr[2]};
        al = bj_face(al, ph, ph + dot3(r0, g), pmax, pmin, e2);
      }
    } else {
      al = 0.0;
    }
  }
  for (int j = 0; j < 3; j++) gout[j] = g[j] * al;
}

// ---------------------------------------------------------------------------
// SST per-cell fields after k_grad and k_sst_bc: the limited kine and omega
// gradients, the face-averaged laminar viscosity, strain and vorticity
// magnitudes and the divergence of the limited velocity gradient (calcGradVel,
// calcStrainRateAndDivergence, calcVorticity, UgpWithCvCompFlow.h:1300-1392),
// the cross diffusion and the blending functions F1, F2
// (calcMenterBlendingFunctions, TurbModel_KOMSST.h:360-392)
// ---------------------------------------------------------------------------
template <bool LS>
__global__ void __launch_bounds__(kBlock) k_sst_cell(DMesh m, DParams P, const double *__restrict__ prim,
                                                     const double *__restrict__ bfa, const double *__restrict__ grad,
                                                     double *__restrict__ aux) {
  const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.ncv) return;
  const double *pc = prim + 8 * (size_t)c;
  double gk[3], gw[3];
  sst_grad<LS>(m, P, prim, P.sst_phi, 0, c, gk);
  sst_grad<LS>(m, P, prim, P.sst_phi + m.ncv, 1, c, gw);
  // face-averaged laminar viscosity over the faocv list
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  const double Tc = pc[4] / (pc[0] * P.R);
  const int nf = e1 - e0;
  double mue = 0.0;
  for (int k = 0; k < nf; k++) {
    const int e = e0 + m.foc[e0 + k];
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    double mul;
    if (en.y >= 0) {
      const double *pn = prim + 8 * (size_t)en.y;
      const double Tn = pn[4] / (pn[0] * P.R);
      const bool s0 = (en.x >= 0);
      double sV[3], smag, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, m.fgeo + 8 * (size_t)f + 4, s0 ? Tc : Tn,
                         s0 ? Tn : Tc, sV, smag, mul, w0, w1);
    } else {
      mul = mu_lam(P, bfa[8 * (size_t)f + 5]);
    }
    mue += mul;
  }
  mue /= (double)nf;
  const double *gu = grad + 16 * (size_t)c + 6;
  double str = 0.0, vort = 0.0;
  const double diverg = gu[0] + gu[4] + gu[8];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = i == j ? 0.5 * (gu[3 * i + j] + gu[3 * j + i]) - 1. / 3. * diverg : 0.5 * (gu[3 * i + j] + gu[3 * j + i]);
      str += v * v;
    }
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      double v = 0.5 * (gu[3 * i + j] - gu[3 * j + i]);
      vort += v * v;
    }
  double *o = aux + kSstN * (size_t)c;
  const double rho = pc[0], kine = P.sst_phi[c], omega = P.sst_phi[m.ncv + c], d = o[kSWall];
  const double crossDiff = Dev::mx(2.0 * rho * kSstSigmaW2 / omega * dot3(gk, gw), 1.0e-20);
  const double gamma1 = 500.0 * mue / (d * d * rho * omega);
  const double gamma2 = 4.0 * kSstSigmaW2 * rho * kine / (d * d * crossDiff);
  const double gamma3 = sqrt(kine) / (kSstBetaStar * omega * d);
  double gamma = Dev::mn(Dev::mx(gamma1, gamma3), gamma2);
  const double f1 = tanh(pow(gamma, 4.0));
  gamma = Dev::mx(2.0 * gamma3, gamma1);
  const double f2 = tanh(gamma * gamma);
  for (int j = 0; j < 3; j++) {
    o[kSGradK + j] = gk[j];
    o[kSGradW + j] = gw[j];
  }
  o[kSMue] = mue;
  o[kSStr] = sqrt(2.0 * str);
  o[kSVort] = sqrt(2.0 * vort);
  o[kSDiv] = diverg;
  o[kSCd] = crossDiff;
  o[kSF1] = f1;
  o[kSF2] = f2;
}

// ---------------------------------------------------------------------------
// The nuSA equation of one cell, gathered in the reference's order: the HLLC
// convective flux (calcEulerFlux scalar terms, JoeWithModels.cpp:2824-2869,
// 3014-3026, 3134-3198, 3291-3342, 3413-3454, 3490-3495) with its Jacobian
// rows AScal[0..5] at cell-centre states (:3088-3101, 3270-3274, 3385,
// 3559-3561); diffusion with the SA diffusivity (calcViscousFluxScalar_new,
// Scalars.cpp:117-247; diffusivityHookScalarRansTurb, TurbModel_SA.h:130-172);
// the SA source and its diagonal (sourceHookScalarRansTurb_new,
// TurbModel_SA.h:282-357). AS is [nnz][6] in device row order.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_scal(DMesh m, DParams P, const double *__restrict__ prim,
                                                 const double *__restrict__ grad, const double *__restrict__ bfa,
                                                 double *__restrict__ rhs_s, double *__restrict__ AS) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= m.ncv) return;
  const int c = t;
  const double gam = P.gamma;
  const bool snd = P.second_order != 0;
  const double *pc = prim + 8 * (size_t)c;
  const double *gc = grad + 16 * (size_t)c;
  const double *phi = P.sa_phi;
  const double *auxc = P.sa_aux + kAuxN * (size_t)c;
  const double ph = phi[c];
  Side own;
  own.rho = pc[0]; own.u[0] = pc[1]; own.u[1] = pc[2]; own.u[2] = pc[3]; own.p = pc[4];
  double acc = 0.0;
  double D[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double S[6], F[5], Fs;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  for (int e = e0; e < e1; e++) {
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double area = geo[3];
    const double xfa[3] = {geo[4], geo[5], geo[6]};
    if (en.y >= 0) {
      const int nb = en.y;
      const double *pn = prim + 8 * (size_t)nb;
      const double phn = phi[nb];
      Side nbs;
      nbs.rho = pn[0]; nbs.u[0] = pn[1]; nbs.u[1] = pn[2]; nbs.u[2] = pn[3]; nbs.p = pn[4];
      Side so = own, sn = nbs;
      double sco = ph, scn = phn;
      if (snd) {
        if (reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          sco = sco + dot3(r, auxc + kAuxGrad);
        }
        if (reconstruct(sn, pn, grad + 16 * (size_t)nb, xfa)) {
          const double r[3] = {xfa[0] - pn[5], xfa[1] - pn[6], xfa[2] - pn[7]};
          scn = scn + dot3(r, P.sa_aux + kAuxN * (size_t)nb + kAuxGrad);
        }
      }
      const bool s0 = en.x >= 0;
      const int noc = m.cfl_noc[e];
      const bool accum = (noc >> 30) & 1;
      double *off = AS + 6 * (size_t)(noc & 0x3fffffff);
      // the flux and both AScal rows once, on the face's (icv0, icv1) sides,
      // routed by this cell's side (acc - x == acc + (-x))
      const Side fl = s0 ? so : sn, fr = s0 ? sn : so;
      const Side jl = s0 ? own : nbs, jr = s0 ? nbs : own;
      const double phl = s0 ? sco : scn, phr = s0 ? scn : sco, pjl = s0 ? ph : phn, pjr = s0 ? phn : ph;
      flux_hllc(F, fl.rho, fl.u, fl.p, fr.rho, fr.u, fr.p, gam, area, n, phl, phr, &Fs);
      acc += s0 ? -Fs : Fs;
      jac_hllc_scal(S, 0, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, pjl, pjr);
      if (s0) {
        for (int k = 0; k < 6; k++) D[k] += S[k];
      } else {
        for (int k = 0; k < 6; k++) off[k] = accum ? off[k] - S[k] : 0.0 - S[k];
      }
      jac_hllc_scal(S, 1, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, pjl, pjr);
      if (s0) {
        for (int k = 0; k < 6; k++) off[k] = accum ? off[k] + S[k] : 0.0 + S[k];
      } else {
        for (int k = 0; k < 6; k++) D[k] -= S[k];
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      const double phb = bb[7];
      const double ub[3] = {bb[1], bb[2], bb[3]};
      if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        flux_hllc(F, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, phb, phb, &Fs);
        acc += -Fs;
        jac_hllc_scal(S, 0, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, phb, phb);
      } else {
        Side so = own;
        double sco = ph;
        if (snd && reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          sco = sco + dot3(r, auxc + kAuxGrad);
        }
        flux_hllc(F, so.rho, so.u, so.p, bb[0], ub, bb[4], gam, area, n, sco, phb, &Fs);
        acc += -Fs;
        jac_hllc_scal(S, 0, own.rho, own.u, own.p, bb[0], ub, bb[4], gam, area, n, ph, phb);
      }
      for (int k = 0; k < 6; k++) D[k] += S[k];
    }
  }
  // diffusion, after the convective pass (JoeWithModels.cpp:2595-2610)
  const double invSigma = 1.0 / kSaSigma;
  const double Tc = pc[4] / (pc[0] * P.R);
  for (int e = e0; e < e1; e++) {
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double nmag = geo[3];
    const double xfa[3] = {geo[4], geo[5], geo[6]};
    if (en.y >= 0) {
      const int nb = en.y;
      const double *pn = prim + 8 * (size_t)nb;
      const double Tn = pn[4] / (pn[0] * P.R);
      const bool s0 = en.x >= 0;
      double sV[3], smag, mul, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, s0 ? Tc : Tn, s0 ? Tn : Tc, sV, smag,
                         mul, w0, w1);
      const double rho0 = s0 ? pc[0] : pn[0], rho1 = s0 ? pn[0] : pc[0];
      const double nu0 = s0 ? ph : phi[nb], nu1 = s0 ? phi[nb] : ph;
      const double *g0 = s0 ? auxc + kAuxGrad : P.sa_aux + kAuxN * (size_t)nb + kAuxGrad;
      const double *g1 = s0 ? P.sa_aux + kAuxN * (size_t)nb + kAuxGrad : auxc + kAuxGrad;
      const double rho_fa = (w1 * rho0 + w0 * rho1) / (w0 + w1);
      const double nuSA_fa = (w1 * nu0 + w0 * nu1) / (w0 + w1);
      const double diff = invSigma * (mul + rho_fa * nuSA_fa);
      const double alpha = dot3(n, sV);
      const double viscFlux = diff * nmag *
                              (alpha * (nu1 - nu0) / smag + 0.5 * (g0[0] + g1[0]) * (n[0] - alpha * sV[0]) +
                               0.5 * (g0[1] + g1[1]) * (n[1] - alpha * sV[1]) +
                               0.5 * (g0[2] + g1[2]) * (n[2] - alpha * sV[2]));
      const double a = diff * nmag * alpha / smag;
      double *off = AS + 6 * (size_t)(m.cfl_noc[e] & 0x3fffffff);
      if (s0) {
        acc += viscFlux;
        D[5] += a / rho0;
        off[5] = off[5] + -(a / rho1);
      } else {
        acc += -viscFlux;
        D[5] += a / rho1;
        off[5] = off[5] + -(a / rho0);
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      if (!z.skind) continue;  // no diffusive flux without a nuSA value (Scalars.cpp:239-245)
      const double *bb = bfa + 8 * (size_t)f;
      const double mul = mu_lam(P, bb[5]);
      const double diff = z.bc == kBcWall ? invSigma * mul : invSigma * (mul + bb[0] * bb[7]);
      const double s_half[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
      const double smag_half = fabs(dot3(s_half, n));
      acc += diff * nmag * (bb[7] - ph) / smag_half;
      D[5] += diff * nmag / smag_half / pc[0];
    }
  }
  // source (TurbModel_SA.h:289-354)
  {
    const double rho = pc[0], vol = m.vol[c];
    const double mue = auxc[kAuxMue];
    const double Om = Dev::mx(auxc[kAuxVort], 1.e-8);
    const double Strainrate = Dev::mx(auxc[kAuxStr], 1.e-8);
    const double d = auxc[kAuxWall];
    const double chi = ph * rho / mue;
    const double chi_3 = pow(chi, 3.0);
    const double inv_KarmanConst2_d2 = 1.0 / (kSaKarman2 * d * d);
    const double fv1 = chi_3 / (chi_3 + kSaCv1_3);
    const double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    const double BlendTurbulentProduction = 0.0;
    const double Sv = Dev::mx(Om + Dev::mn(0.0, Strainrate - Om) * BlendTurbulentProduction, 0.0);
    const double Shat = Sv + ph * fv2 * inv_KarmanConst2_d2;
    const double inv_Shat = 1.0 / Dev::mx(Shat, 1.0e-8);
    const double r = Dev::mn(ph * inv_Shat * inv_KarmanConst2_d2, 10.0);
    const double gg = r + kSaCw2 * (pow(r, 6.0) - r);
    const double g_6 = pow(gg, 6.0);
    const double fw = gg * pow((1.0 + kSaCw3_6) / (g_6 + kSaCw3_6), 1.0 / 6.0);
    const double H1 = kSaCb1 * Shat * ph;
    const double H2 = -kSaCw1 * fw * ph * ph / pow(d, 2.0);
    const double H3 = kSaCb2 / kSaSigma * dot3(auxc + kAuxGrad, auxc + kAuxGrad);
    acc += rho * vol * (H1 + H2 + H3);
    const double d_H1 = 0.0;
    const double d_H2 = -2.0 * kSaCw1 * fw * ph / pow(d, 2.0);
    D[5] -= (d_H1 + d_H2) * vol;
  }
  rhs_s[c] = acc;
  double *dg = AS + 6 * (size_t)m.nbocv_i[c];
  for (int k = 0; k < 6; k++) dg[k] = D[k];
}

// ---------------------------------------------------------------------------
// Both SST scalars' equations of one cell (kine, omega), gathered in the
// reference's order as k_scal gathers nuSA's: the HLLC convective flux with
// the sides' kinetic energy in the enthalpy (calcEulerFlux scalar terms,
// JoeWithModels.cpp:2824-2930, 3014-3026, 3134-3198, 3408-3495) and the AScal
// rows at cell-centre states (:3088-3101, 3270-3274, 3559-3561), one face
// evaluation serving both scalars; diffusion with the blended SST
// diffusivities (calcViscousFluxScalar_new, Scalars.cpp:117-247;
// diffusivityHookScalarRansTurb, TurbModel_KOMSST.h:160-254); the sources and
// their diagonals (sourceHookScalarRansTurb_new, :256-304). rhs_s is
// [2][ncv], AS [2][nnz][6].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_sst_scal(DMesh m, DParams P, const double *__restrict__ prim,
                                                     const double *__restrict__ grad, const double *__restrict__ bfa,
                                                     double *__restrict__ rhs_s, double *__restrict__ AS) {
  const int c = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (c >= m.ncv) return;
  const int ncv = m.ncv;
  const size_t nnz = m.nnz;
  const double gam = P.gamma;
  const bool snd = P.second_order != 0;
  const double *pc = prim + 8 * (size_t)c;
  const double *gc = grad + 16 * (size_t)c;
  const double *kin = P.sst_phi, *omg = P.sst_phi + ncv;
  const double *auxc = P.sst_aux + kSstN * (size_t)c;
  const double kc = kin[c], wc = omg[c];
  double *AS1 = AS + 6 * nnz;
  Side own;
  own.rho = pc[0]; own.u[0] = pc[1]; own.u[1] = pc[2]; own.u[2] = pc[3]; own.p = pc[4];
  double acc0 = 0.0, acc1 = 0.0;
  double D0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, D1[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double S0[6], S1[6], F[5], Fs0, Fs1;
  const int e0 = m.cfl_i[c], e1 = m.cfl_i[c + 1];
  for (int e = e0; e < e1; e++) {
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double area = geo[3];
    const double xfa[3] = {geo[4], geo[5], geo[6]};
    if (en.y >= 0) {
      const int nb = en.y;
      const double *pn = prim + 8 * (size_t)nb;
      const double *auxn = P.sst_aux + kSstN * (size_t)nb;
      const double kn = kin[nb], wn = omg[nb];
      Side nbs;
      nbs.rho = pn[0]; nbs.u[0] = pn[1]; nbs.u[1] = pn[2]; nbs.u[2] = pn[3]; nbs.p = pn[4];
      Side so = own, sn = nbs;
      double kfo = kc, kfn = kn, wfo = wc, wfn = wn;
      if (snd) {
        if (reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          kfo = kfo + dot3(r, auxc + kSGradK);
          wfo = wfo + dot3(r, auxc + kSGradW);
        }
        if (reconstruct(sn, pn, grad + 16 * (size_t)nb, xfa)) {
          const double r[3] = {xfa[0] - pn[5], xfa[1] - pn[6], xfa[2] - pn[7]};
          kfn = kfn + dot3(r, auxn + kSGradK);
          wfn = wfn + dot3(r, auxn + kSGradW);
        }
      }
      const bool s0 = en.x >= 0;
      const int noc = m.cfl_noc[e];
      const bool accum = (noc >> 30) & 1;
      double *off0 = AS + 6 * (size_t)(noc & 0x3fffffff), *off1 = AS1 + 6 * (size_t)(noc & 0x3fffffff);
      const Side fl = s0 ? so : sn, fr = s0 ? sn : so;
      const Side jl = s0 ? own : nbs, jr = s0 ? nbs : own;
      const double kl = s0 ? kfo : kfn, kr = s0 ? kfn : kfo, kjl = s0 ? kc : kn, kjr = s0 ? kn : kc;
      const double wl = s0 ? wfo : wfn, wr = s0 ? wfn : wfo, wjl = s0 ? wc : wn, wjr = s0 ? wn : wc;
      flux_hllc(F, fl.rho, fl.u, fl.p, fr.rho, fr.u, fr.p, gam, area, n, kl, kr, &Fs0, kl, kr, &Fs1, wl, wr);
      acc0 += s0 ? -Fs0 : Fs0;
      acc1 += s0 ? -Fs1 : Fs1;
      jac_hllc_scal2(S0, S1, 0, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, kjl, kjr, wjl, wjr, kjl, kjr);
      if (s0) {
        for (int k = 0; k < 6; k++) {
          D0[k] += S0[k];
          D1[k] += S1[k];
        }
      } else {
        for (int k = 0; k < 6; k++) {
          off0[k] = accum ? off0[k] - S0[k] : 0.0 - S0[k];
          off1[k] = accum ? off1[k] - S1[k] : 0.0 - S1[k];
        }
      }
      jac_hllc_scal2(S0, S1, 1, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, kjl, kjr, wjl, wjr, kjl, kjr);
      if (s0) {
        for (int k = 0; k < 6; k++) {
          off0[k] = accum ? off0[k] + S0[k] : 0.0 + S0[k];
          off1[k] = accum ? off1[k] + S1[k] : 0.0 + S1[k];
        }
      } else {
        for (int k = 0; k < 6; k++) {
          D0[k] -= S0[k];
          D1[k] -= S1[k];
        }
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      const double kb = P.sst_bfa[2 * (size_t)f], wb = P.sst_bfa[2 * (size_t)f + 1];
      const double ub[3] = {bb[1], bb[2], bb[3]};
      if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        flux_hllc(F, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, kb, kb, &Fs0, kb, kb, &Fs1, wb, wb);
        acc0 += -Fs0;
        acc1 += -Fs1;
        jac_hllc_scal2(S0, S1, 0, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, kb, kb, wb, wb, kb, kb);
      } else {
        Side so = own;
        double kfo = kc, wfo = wc;
        if (snd && reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          kfo = kfo + dot3(r, auxc + kSGradK);
          wfo = wfo + dot3(r, auxc + kSGradW);
        }
        flux_hllc(F, so.rho, so.u, so.p, bb[0], ub, bb[4], gam, area, n, kfo, kb, &Fs0, kfo, kb, &Fs1, wfo, wb);
        acc0 += -Fs0;
        acc1 += -Fs1;
        jac_hllc_scal2(S0, S1, 0, own.rho, own.u, own.p, bb[0], ub, bb[4], gam, area, n, kc, kb, wc, wb, kc, kb);
      }
      for (int k = 0; k < 6; k++) {
        D0[k] += S0[k];
        D1[k] += S1[k];
      }
    }
  }
  // diffusion, after the convective pass (JoeWithModels.cpp:2595-2610)
  const double Tc = pc[4] / (pc[0] * P.R);
  for (int e = e0; e < e1; e++) {
    const int2 en = m.cfl[e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double nmag = geo[3];
    const double xfa[3] = {geo[4], geo[5], geo[6]};
    if (en.y >= 0) {
      const int nb = en.y;
      const double *pn = prim + 8 * (size_t)nb;
      const double Tn = pn[4] / (pn[0] * P.R);
      const bool s0 = en.x >= 0;
      double sV[3], smag, mul, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, s0 ? Tc : Tn, s0 ? Tn : Tc, sV, smag,
                         mul, w0, w1);
      const int c0 = s0 ? c : nb, c1 = s0 ? nb : c;
      const double rho0 = s0 ? pc[0] : pn[0], rho1 = s0 ? pn[0] : pc[0];
      const double *a0 = P.sst_aux + kSstN * (size_t)c0, *a1 = P.sst_aux + kSstN * (size_t)c1;
      const double mut = sst_mut_face(P, ncv, w0, w1, rho0, rho1, c0, c1);
      const double f1_fa = (w1 * a0[kSF1] + w0 * a1[kSF1]) / (w0 + w1);
      const double alpha = dot3(n, sV);
      const int noc = m.cfl_noc[e] & 0x3fffffff;
      // one scalar's face term, called with that scalar's own accumulators
      auto face = [&](double sg1, double sg2, const double *phi, int gS, double &acc, double &d5, double *off) {
        const double coeff = sg1 * f1_fa + (1. - f1_fa) * sg2;
        const double diff = mul + coeff * mut;
        const double phi0 = phi[c0], phi1 = phi[c1];
        const double *g0 = a0 + gS, *g1 = a1 + gS;
        const double viscFlux = diff * nmag *
                                (alpha * (phi1 - phi0) / smag + 0.5 * (g0[0] + g1[0]) * (n[0] - alpha * sV[0]) +
                                 0.5 * (g0[1] + g1[1]) * (n[1] - alpha * sV[1]) +
                                 0.5 * (g0[2] + g1[2]) * (n[2] - alpha * sV[2]));
        const double a = diff * nmag * alpha / smag;
        if (s0) {
          acc += viscFlux;
          d5 += a / rho0;
          off[5] = off[5] + -(a / rho1);
        } else {
          acc += -viscFlux;
          d5 += a / rho1;
          off[5] = off[5] + -(a / rho0);
        }
      };
      face(kSstSigmaK1, kSstSigmaK2, kin, kSGradK, acc0, D0[5], AS + 6 * (size_t)noc);
      face(kSstSigmaW1, kSstSigmaW2, omg, kSGradW, acc1, D1[5], AS1 + 6 * (size_t)noc);
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      if (!z.skind && !z.skind2) continue;  // no diffusive flux without a value or hook (Scalars.cpp:239-245)
      const double *bb = bfa + 8 * (size_t)f;
      const double mul = mu_lam(P, bb[5]);
      const double mutb = z.bc != kBcWall ? sst_mut_cell(P, ncv, pc[0], c) : 0.0;
      const double s_half[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
      const double smag_half = fabs(dot3(s_half, n));
      auto bface = [&](double sg1, double sg2, double phb, double ph, double &acc, double &d5) {
        double diff = mul;
        if (z.bc != kBcWall) {
          const double coeff = sg1 * auxc[kSF1] + (1. - auxc[kSF1]) * sg2;
          diff = mul + coeff * mutb;
        }
        acc += diff * nmag * (phb - ph) / smag_half;
        d5 += diff * nmag / smag_half / pc[0];
      };
      if (z.skind) bface(kSstSigmaK1, kSstSigmaK2, P.sst_bfa[2 * (size_t)f], kc, acc0, D0[5]);
      if (z.skind2) bface(kSstSigmaW1, kSstSigmaW2, P.sst_bfa[2 * (size_t)f + 1], wc, acc1, D1[5]);
    }
  }
  // sources (TurbModel_KOMSST.h:256-304)
  {
    const double rho = pc[0], vol = m.vol[c];
    const double kine = kc, omega = wc;
    const double lim = sst_lim(P, auxc), F2 = auxc[kSF2], str = auxc[kSStr], div = auxc[kSDiv];
    {
      const double zeta = Dev::mn(1.0 / omega, kSstA1 / (lim * F2));
      const double mut = Dev::mn(Dev::mx(rho * kine * zeta, 0.0), 1.0e5);
      double Pk = mut * str * str - 2. / 3. * rho * kine * div;
      if (P.sst_limit_pk == 1) Pk = Dev::mn(Pk, 20.0 * kSstBetaStar * rho * kine * omega);
      Pk = Dev::mx(Pk, 0.0);
      const double src = Pk - kSstBetaStar * rho * omega * kine;
      acc0 += src * vol;
      const double dsrcdphi = -kSstBetaStar * omega;
      D0[5] -= dsrcdphi * vol;
    }
    {
      const double F1 = auxc[kSF1];
      const double alfa = F1 * P.sst_alfa[0] + (1.0 - F1) * P.sst_alfa[1];
      const double beta = F1 * kSstBeta1 + (1.0 - F1) * kSstBeta2;
      const double zeta = Dev::mx(omega, lim * F2 / kSstA1);
      double Pk = str * str - 2. / 3. * zeta * div;
      Pk = Dev::mx(Pk, 0.0);
      const double src = alfa * rho * Pk + (1.0 - F1) * auxc[kSCd] - beta * rho * omega * omega;
      acc1 += src * vol;
      const double dsrcdphi = -2.0 * beta * omega;
      D1[5] -= dsrcdphi * vol;
    }
  }
  rhs_s[c] = acc0;
  rhs_s[ncv + c] = acc1;
  const size_t dgi = 6 * (size_t)m.nbocv_i[c];
  for (int k = 0; k < 6; k++) {
    AS[dgi + k] = D0[k];
    AS1[dgi + k] = D1[k];
  }
}

// The same rows per cell-face entry: a lane evaluates its entry's face for
// both scalars, writes the off-diagonal AScal entries it owns, and leaves its
// contributions to the cell's sums in LDS: the convective acc (2) and
// diagonal entries (2 x 6, signed for the cell: d - s == d + (-s)), then the
// diffusive acc (2) and diagonal [5] (2), -0.0 where it adds nothing. The
// owner lane sums the convective terms of the cell's entries in entry order,
// then the diffusive ones, then adds the sources, as k_sst_scal does; so the
// bits are k_sst_scal's. Each off-diagonal block has one writer (the SST path
// requires a mesh without duplicate neighbour pairs).
constexpr int kSstScalWaves = 1;

__global__ void __launch_bounds__(64 * kSstScalWaves) k_sst_scal_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                                  const double *__restrict__ grad,
                                                                  const double *__restrict__ bfa,
                                                                  double *__restrict__ rhs_s, double *__restrict__ AS) {
  constexpr int NS = 18;
  __shared__ double lds[64 * kSstScalWaves * NS];
  double *slot = lds + NS * threadIdx.x;
  const EntryLane L = entry_lane<kSstScalWaves>(m, hw_blk());
  const int ncv = m.ncv;
  const size_t nnz = m.nnz;
  double *AS1 = AS + 6 * nnz;
  const double *kin = P.sst_phi, *omg = P.sst_phi + ncv;
  const int c = L.e >= 0 ? L.c : 0;
  const double *pc = prim + 8 * (size_t)c;
  const double *auxc = P.sst_aux + kSstN * (size_t)c;
  const double kc = kin[c], wc = omg[c];
  if (L.e >= 0) {
    const double gam = P.gamma;
    const bool snd = P.second_order != 0;
    const double *gc = grad + 16 * (size_t)c;
    Side own;
    own.rho = pc[0]; own.u[0] = pc[1]; own.u[1] = pc[2]; own.u[2] = pc[3]; own.p = pc[4];
    double S0[6], S1[6], F[5], Fs0, Fs1;
    const int2 en = m.cfl[L.e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    const double n[3] = {geo[0], geo[1], geo[2]};
    const double area = geo[3];
    const double xfa[3] = {geo[4], geo[5], geo[6]};
    double dacc0 = -0.0, dacc1 = -0.0, dd0 = -0.0, dd1 = -0.0;  // diffusive terms
    if (en.y >= 0) {
      const int nb = en.y;
      const double *pn = prim + 8 * (size_t)nb;
      const double *auxn = P.sst_aux + kSstN * (size_t)nb;
      const double kn = kin[nb], wn = omg[nb];
      Side nbs;
      nbs.rho = pn[0]; nbs.u[0] = pn[1]; nbs.u[1] = pn[2]; nbs.u[2] = pn[3]; nbs.p = pn[4];
      Side so = own, sn = nbs;
      double kfo = kc, kfn = kn, wfo = wc, wfn = wn;
      if (snd) {
        if (reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          kfo = kfo + dot3(r, auxc + kSGradK);
          wfo = wfo + dot3(r, auxc + kSGradW);
        }
        if (reconstruct(sn, pn, grad + 16 * (size_t)nb, xfa)) {
          const double r[3] = {xfa[0] - pn[5], xfa[1] - pn[6], xfa[2] - pn[7]};
          kfn = kfn + dot3(r, auxn + kSGradK);
          wfn = wfn + dot3(r, auxn + kSGradW);
        }
      }
      const bool s0 = en.x >= 0;
      const int noc = m.cfl_noc[L.e] & 0x3fffffff;
      double *off0 = AS + 6 * (size_t)noc, *off1 = AS1 + 6 * (size_t)noc;
      const Side fl = s0 ? so : sn, fr = s0 ? sn : so;
      const Side jl = s0 ? own : nbs, jr = s0 ? nbs : own;
      const double kl = s0 ? kfo : kfn, kr = s0 ? kfn : kfo, kjl = s0 ? kc : kn, kjr = s0 ? kn : kc;
      const double wl = s0 ? wfo : wfn, wr = s0 ? wfn : wfo, wjl = s0 ? wc : wn, wjr = s0 ? wn : wc;
      flux_hllc(F, fl.rho, fl.u, fl.p, fr.rho, fr.u, fr.p, gam, area, n, kl, kr, &Fs0, kl, kr, &Fs1, wl, wr);
      slot[0] = s0 ? -Fs0 : Fs0;
      slot[1] = s0 ? -Fs1 : Fs1;
      // the diagonal's side (icv0 adds A_L's rows, icv1 subtracts A_R's)
      jac_hllc_scal2(S0, S1, s0 ? 0 : 1, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, kjl, kjr, wjl, wjr, kjl,
                     kjr);
      for (int k = 0; k < 6; k++) {
        slot[2 + k] = s0 ? S0[k] : -S0[k];
        slot[8 + k] = s0 ? S1[k] : -S1[k];
      }
      // the off-diagonal block (icv0's = +A_R's rows, icv1's = -A_L's)
      jac_hllc_scal2(S0, S1, s0 ? 1 : 0, jl.rho, jl.u, jl.p, jr.rho, jr.u, jr.p, gam, area, n, kjl, kjr, wjl, wjr, kjl,
                     kjr);
      for (int k = 0; k < 6; k++) {
        off0[k] = s0 ? 0.0 + S0[k] : 0.0 - S0[k];
        off1[k] = s0 ? 0.0 + S1[k] : 0.0 - S1[k];
      }
      // diffusion (the cell's second pass)
      const double Tc = pc[4] / (pc[0] * P.R), Tn = pn[4] / (pn[0] * P.R);
      double sV[3], smag, mul, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, s0 ? Tc : Tn, s0 ? Tn : Tc, sV, smag,
                         mul, w0, w1);
      const int c0 = s0 ? c : nb, c1 = s0 ? nb : c;
      const double rho0 = s0 ? pc[0] : pn[0], rho1 = s0 ? pn[0] : pc[0];
      const double *a0 = P.sst_aux + kSstN * (size_t)c0, *a1 = P.sst_aux + kSstN * (size_t)c1;
      const double mut = sst_mut_face(P, ncv, w0, w1, rho0, rho1, c0, c1);
      const double f1_fa = (w1 * a0[kSF1] + w0 * a1[kSF1]) / (w0 + w1);
      const double alpha = dot3(n, sV);
      const double nmag = area;
      auto face = [&](double sg1, double sg2, const double *phi, int gS, double &dacc, double &dd, double *off) {
        const double coeff = sg1 * f1_fa + (1. - f1_fa) * sg2;
        const double diff = mul + coeff * mut;
        const double phi0 = phi[c0], phi1 = phi[c1];
        const double *g0 = a0 + gS, *g1 = a1 + gS;
        const double viscFlux = diff * nmag *
                                (alpha * (phi1 - phi0) / smag + 0.5 * (g0[0] + g1[0]) * (n[0] - alpha * sV[0]) +
                                 0.5 * (g0[1] + g1[1]) * (n[1] - alpha * sV[1]) +
                                 0.5 * (g0[2] + g1[2]) * (n[2] - alpha * sV[2]));
        const double a = diff * nmag * alpha / smag;
        dacc = s0 ? viscFlux : -viscFlux;
        dd = s0 ? a / rho0 : a / rho1;
        off[5] = off[5] + -(s0 ? a / rho1 : a / rho0);
      };
      face(kSstSigmaK1, kSstSigmaK2, kin, kSGradK, dacc0, dd0, off0);
      face(kSstSigmaW1, kSstSigmaW2, omg, kSGradW, dacc1, dd1, off1);
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      const double kb = P.sst_bfa[2 * (size_t)f], wb = P.sst_bfa[2 * (size_t)f + 1];
      const double ub[3] = {bb[1], bb[2], bb[3]};
      if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        flux_hllc(F, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, kb, kb, &Fs0, kb, kb, &Fs1, wb, wb);
        jac_hllc_scal2(S0, S1, 0, bb[0], ub, bb[4], bb[0], ub, bb[4], gam, area, n, kb, kb, wb, wb, kb, kb);
      } else {
        Side so = own;
        double kfo = kc, wfo = wc;
        if (snd && reconstruct(so, pc, gc, xfa)) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          kfo = kfo + dot3(r, auxc + kSGradK);
          wfo = wfo + dot3(r, auxc + kSGradW);
        }
        flux_hllc(F, so.rho, so.u, so.p, bb[0], ub, bb[4], gam, area, n, kfo, kb, &Fs0, kfo, kb, &Fs1, wfo, wb);
        jac_hllc_scal2(S0, S1, 0, own.rho, own.u, own.p, bb[0], ub, bb[4], gam, area, n, kc, kb, wc, wb, kc, kb);
      }
      slot[0] = -Fs0;
      slot[1] = -Fs1;
      for (int k = 0; k < 6; k++) {
        slot[2 + k] = S0[k];
        slot[8 + k] = S1[k];
      }
      if (z.skind || z.skind2) {
        const double mul = mu_lam(P, bb[5]);
        const double mutb = z.bc != kBcWall ? sst_mut_cell(P, ncv, pc[0], c) : 0.0;
        const double s_half[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
        const double smag_half = fabs(dot3(s_half, n));
        const double nmag = area;
        auto bface = [&](double sg1, double sg2, double phb, double ph, double &dacc, double &dd) {
          double diff = mul;
          if (z.bc != kBcWall) {
            const double coeff = sg1 * auxc[kSF1] + (1. - auxc[kSF1]) * sg2;
            diff = mul + coeff * mutb;
          }
          dacc = diff * nmag * (phb - ph) / smag_half;
          dd = diff * nmag / smag_half / pc[0];
        };
        if (z.skind) bface(kSstSigmaK1, kSstSigmaK2, kb, kc, dacc0, dd0);
        if (z.skind2) bface(kSstSigmaW1, kSstSigmaW2, wb, wc, dacc1, dd1);
      }
    }
    slot[14] = dacc0;
    slot[15] = dacc1;
    slot[16] = dd0;
    slot[17] = dd1;
  }
  __syncthreads();
  if (L.e < 0 || L.pos != 0) return;
  const int nent = m.cfl_i[c + 1] - m.cfl_i[c];
  double acc0 = 0.0, acc1 = 0.0;
  double D0[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, D1[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < nent; j++) {
    const double *s = slot + NS * j;
    acc0 += s[0];
    acc1 += s[1];
    for (int k = 0; k < 6; k++) {
      D0[k] += s[2 + k];
      D1[k] += s[8 + k];
    }
  }
  for (int j = 0; j < nent; j++) {
    const double *s = slot + NS * j;
    acc0 += s[14];
    acc1 += s[15];
    D0[5] += s[16];
    D1[5] += s[17];
  }
  // sources (TurbModel_KOMSST.h:256-304)
  {
    const double rho = pc[0], vol = m.vol[c];
    const double kine = kc, omega = wc;
    const double lim = sst_lim(P, auxc), F2 = auxc[kSF2], str = auxc[kSStr], div = auxc[kSDiv];
    {
      const double zeta = Dev::mn(1.0 / omega, kSstA1 / (lim * F2));
      const double mut = Dev::mn(Dev::mx(rho * kine * zeta, 0.0), 1.0e5);
      double Pk = mut * str * str - 2. / 3. * rho * kine * div;
      if (P.sst_limit_pk == 1) Pk = Dev::mn(Pk, 20.0 * kSstBetaStar * rho * kine * omega);
      Pk = Dev::mx(Pk, 0.0);
      const double src = Pk - kSstBetaStar * rho * omega * kine;
      acc0 += src * vol;
      const double dsrcdphi = -kSstBetaStar * omega;
      D0[5] -= dsrcdphi * vol;
    }
    {
      const double F1 = auxc[kSF1];
      const double alfa = F1 * P.sst_alfa[0] + (1.0 - F1) * P.sst_alfa[1];
      const double beta = F1 * kSstBeta1 + (1.0 - F1) * kSstBeta2;
      const double zeta = Dev::mx(omega, lim * F2 / kSstA1);
      double Pk = str * str - 2. / 3. * zeta * div;
      Pk = Dev::mx(Pk, 0.0);
      const double src = alfa * rho * Pk + (1.0 - F1) * auxc[kSCd] - beta * rho * omega * omega;
      acc1 += src * vol;
      const double dsrcdphi = -2.0 * beta * omega;
      D1[5] -= dsrcdphi * vol;
    }
  }
  rhs_s[c] = acc0;
  rhs_s[ncv + c] = acc1;
  const size_t dgi = 6 * (size_t)m.nbocv_i[c];
  for (int k = 0; k < 6; k++) {
    AS[dgi + k] = D0[k];
    AS1[dgi + k] = D1[k];
  }
}

// the segregated scalar's update and clip with its own bounds
// (JoeWithModels.cpp:1084-1088)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sc_update_b(int ncv, const double *__restrict__ q,
                                                        const T *__restrict__ dq, const double *__restrict__ x,
                                                        double *__restrict__ phi, double lo, double hi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double rho = q[c];
  const double v = (phi[c] * (rho - (double)dq[5 * (size_t)c]) + x[c]) / rho;
  phi[c] = Dev::mn(Dev::mx(v, lo), hi);
}

// ---------------------------------------------------------------------------
// The transported scalar's row of the coupled system (runBackwardEulerCoupled
// with Spalart-Allmaras, 6x6 blocks), in the reference's order: per face of
// calcFluxCoupled (JoeWithModels.cpp :4797-5796) the HLLC scalar flux
// (calcEulerFluxCoupled_HLLC, scalars upwinded by sM, side enthalpies after
// reconstruction) and its Jacobian row at cell-centre states (:5035-5048,
// 5150-5740), then the same face's scalar diffusion with the SA diffusivity
// and its row entries [5][0], [5][5] (calcViscousFluxCoupled, :5076-5105,
// 5373-5404, 5548-5575, 5744-5775; diffusivityHookScalarRansTurb,
// TurbModel_SA.h:130-172); then the source (sourceHookRansTurbCoupled,
// TurbModel_SA.h:359-434).
// ---------------------------------------------------------------------------
// The cell's own values every entry of the scalar row reads.
struct ScalOwn {
  const double *pc, *gc, *gpc;
  double ph, hown, Tc;
  Side own;
};

__device__ __forceinline__ ScalOwn scal_own(const DParams &P, const double *__restrict__ prim,
                                            const double *__restrict__ grad, int c) {
  ScalOwn o;
  o.pc = prim + 8 * (size_t)c;
  o.gc = grad + 16 * (size_t)c;
  o.gpc = P.sa_aux + kAuxN * (size_t)c + kAuxGrad;
  o.ph = P.sa_phi[c];
  o.own.rho = o.pc[0]; o.own.u[0] = o.pc[1]; o.own.u[1] = o.pc[2]; o.own.u[2] = o.pc[3]; o.own.p = o.pc[4];
  o.hown = enthalpy_of(P, o.own.rho, o.own.p);
  o.Tc = o.pc[4] / (o.pc[0] * P.R);
  return o;
}

// One entry (face) of the scalar row: what it adds to the cell's residual,
// the Euler term then the diffusive one (out[0], out[1]), and to row 5 of
// the cell's diagonal block, from the Euler terms (out[2..7]) then from the
// diffusion (out[8] to [5][0], out[9] to [5][5]); -0.0 where the entry adds
// nothing (x + -0.0 is x, signed zeros included), and subtractions stored
// negated (a - b is a + (-b)). With A, row 5 of the entry's off-diagonal
// block is written (internal faces).
template <typename T, bool SND>
__device__ __forceinline__ void scal_c_entry(const DMesh &m, const DParams &P, const double *__restrict__ prim,
                                             const double *__restrict__ grad, const double *__restrict__ bfa,
                                             T *__restrict__ A, const ScalOwn &o, int e, double out[10]) {
  const double gam = P.gamma;
  const double invSigma = 1.0 / kSaSigma;
  const double *pc = o.pc, *gc = o.gc, *gpc = o.gpc;
  const double ph = o.ph;
  const Side &own = o.own;
  const double *phi = P.sa_phi;
  double S[6], R6[6], F[5], Fs, a0[2], a1[2];
  for (int k = 0; k < 10; k++) out[k] = -0.0;
  const int2 en = m.cfl[e];
  const int f = en.x & 0x7fffffff;
  const double *geo = m.fgeo + 8 * (size_t)f;
  const double n[3] = {geo[0], geo[1], geo[2]};
  const double area = geo[3];
  const double xfa[3] = {geo[4], geo[5], geo[6]};
  if (en.y >= 0) {
    const int nb = en.y;
    const double *pn = prim + 8 * (size_t)nb;
    const double phn = phi[nb];
    const double *gpn = P.sa_aux + kAuxN * (size_t)nb + kAuxGrad;
    Side nbs;
    nbs.rho = pn[0]; nbs.u[0] = pn[1]; nbs.u[1] = pn[2]; nbs.u[2] = pn[3]; nbs.p = pn[4];
    Side so = own, sn = nbs;
    double sco = ph, scn = phn;
    if (SND) {
      if (reconstruct(so, pc, gc, xfa)) {
        const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
        sco = sco + dot3(r, gpc);
      }
      if (reconstruct(sn, pn, grad + 16 * (size_t)nb, xfa)) {
        const double r[3] = {xfa[0] - pn[5], xfa[1] - pn[6], xfa[2] - pn[7]};
        scn = scn + dot3(r, gpn);
      }
    }
    const bool s0 = en.x >= 0;
    const int noc = m.cfl_noc[e];
    const bool accum = (noc >> 30) & 1;
    T *off = A ? A + 36 * (size_t)(noc & 0x3fffffff) + 30 : nullptr;
    T ov[6];  // row 5 of the off-diagonal block, kept until its one store
    // the face's (icv0, icv1) sides; the flux and both rows evaluated once,
    // routed by this cell's side
    const Side fl = s0 ? so : sn, fr = s0 ? sn : so;
    const double phl = s0 ? sco : scn, phr = s0 ? scn : sco;
    flux_hllc_coupled(F, fl.rho, fl.u, fl.p, enthalpy_of(P, fl.rho, fl.p), fr.rho, fr.u, fr.p,
                      enthalpy_of(P, fr.rho, fr.p), gam, area, n, phl, phr, &Fs);
    out[0] = s0 ? -Fs : Fs;
    if (A) {
      const Side jl = s0 ? own : nbs, jr = s0 ? nbs : own;
      const double hjl = enthalpy_of(P, jl.rho, jl.p), hjr = enthalpy_of(P, jr.rho, jr.p);
      const double pjl = s0 ? ph : phn, pjr = s0 ? phn : ph;
      jac_hllc_coupled_srow(S, 0, jl.rho, jl.u, jl.p, hjl, pjl, jr.rho, jr.u, jr.p, hjr, pjr, gam, area, n);
      jac_hllc_coupled_srow(R6, 1, jl.rho, jl.u, jl.p, hjl, pjl, jr.rho, jr.u, jr.p, hjr, pjr, gam, area, n);
      // icv0: A[noc00] += SL, A[noc01] += SR; icv1: A[noc11] -= SR, A[noc10] -= SL
      for (int k = 0; k < 6; k++) {
        if (s0) {
          out[2 + k] = S[k];
          ov[k] = (T)(accum ? (double)off[k] + R6[k] : 0.0 + R6[k]);
        } else {
          out[2 + k] = -R6[k];
          ov[k] = (T)(accum ? (double)off[k] - S[k] : 0.0 - S[k]);
        }
      }
    }
    {  // the face's scalar diffusion (:5076-5105)
      const double Tn = pn[4] / (pn[0] * P.R);
      double sV[3], smag, mul, w0, w1;
      visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, s0 ? o.Tc : Tn, s0 ? Tn : o.Tc, sV,
                         smag, mul, w0, w1);
      const double rho0 = s0 ? pc[0] : pn[0], rho1 = s0 ? pn[0] : pc[0];
      const double nu0 = s0 ? ph : phn, nu1 = s0 ? phn : ph;
      const double *g0 = s0 ? gpc : gpn, *g1 = s0 ? gpn : gpc;
      const double rho_fa = (w1 * rho0 + w0 * rho1) / (w0 + w1);
      const double nuSA_fa = (w1 * nu0 + w0 * nu1) / (w0 + w1);
      const double diff = invSigma * (mul + rho_fa * nuSA_fa);
      const double VFs =
          visc_coupled_srow(a0, a1, rho0, nu0, g0, rho1, nu1, g1, diff, 1.0, area, n, smag, sV, dot3(n, sV));
      out[1] = s0 ? -VFs : VFs;
      if (A) {  // icv0: [noc00] += a0, [noc01] += a1; icv1: [noc11] -= a1, [noc10] -= a0
        if (s0) {
          out[8] = a0[0];
          out[9] = a0[1];
          ov[0] = (T)((double)ov[0] + a1[0]);
          ov[5] = (T)((double)ov[5] + a1[1]);
        } else {
          out[8] = -a1[0];
          out[9] = -a1[1];
          ov[0] = (T)((double)ov[0] - a0[0]);
          ov[5] = (T)((double)ov[5] - a0[1]);
        }
        for (int k = 0; k < 6; k++) off[k] = ov[k];
      }
    }
  } else {
    const ZoneDev &z = m.zones[-en.y - 1];
    const double *bb = bfa + 8 * (size_t)f;
    const double rb = bb[0], pb = bb[4], hb = bb[6], phb = bb[7];
    const double ub[3] = {bb[1], bb[2], bb[3]};
    // a zone without a nuSA value carries no diffusive scalar flux
    // (DiffTerm = 0, :5453-5461, 5599-5607); NEUMANN zones keep it (:5293-5296)
    const double DiffTerm = (z.bc == kBcNeumann || z.skind) ? 1.0 : 0.0;
    const double mul = mu_lam(P, bb[5]);
    const double diff = z.bc == kBcWall ? invSigma * mul : invSigma * (mul + rb * phb);
    if (z.bc == kBcSymmetry || z.bc == kBcWall || z.bc == kBcNeumann) {
      const bool op = z.bc != kBcNeumann && P.symm_bc == kSymmAnalytical;  // ONLY_PRESSURE (:1910-1917)
      if (op) {
        Fs = 0.0 * area;
      } else {
        flux_hllc_coupled(F, rb, ub, pb, hb, rb, ub, pb, hb, gam, area, n, phb, phb, &Fs);
      }
      out[0] = -Fs;
      if (A) {
        if (op) {  // a zero row through dU/dQ
          for (int k = 0; k < 6; k++) S[k] = 0.0 * 1.0;
          S[0] -= phb / rb * S[5];
          S[5] /= rb;
          for (int k = 0; k < 6; k++) out[2 + k] = S[k];
        } else {
          jac_hllc_coupled_srow(S, 0, rb, ub, pb, hb, phb, rb, ub, pb, hb, phb, gam, area, n);
          if (P.symm_bc == kSymmNew) {
            jac_hllc_coupled_srow(R6, 1, rb, ub, pb, hb, phb, rb, ub, pb, hb, phb, gam, area, n);
            for (int k = 0; k < 6; k++) out[2 + k] = S[k] + R6[k];
          } else {
            for (int k = 0; k < 6; k++) out[2 + k] = S[k];
          }
        }
      }
      if (z.bc != kBcSymmetry) {  // wall (:5548-5575) and NEUMANN (:5373-5404) diffusion along the normal
        const double sV[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
        const double smag = fabs(dot3(sV, n));
        const bool neu = z.bc == kBcNeumann;
        // the gradient by value (zero on NEUMANN faces): a pointer chosen
        // between a local array and global memory puts the array in scratch
        const double gv[3] = {neu ? 0.0 : gpc[0], neu ? 0.0 : gpc[1], neu ? 0.0 : gpc[2]};
        const double VFs = visc_coupled_srow(a0, a1, neu ? rb : pc[0], neu ? phb : ph, gv, rb, phb, gv, diff,
                                             DiffTerm, area, n, smag, n, 1.0);
        out[1] = -VFs;
        if (A) {
          out[8] = a0[0];
          out[9] = a0[1];
        }
      }
    } else {
      Side so = own;
      double sco = ph;
      if (SND && reconstruct(so, pc, gc, xfa)) {
        const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
        sco = sco + dot3(r, gpc);
      }
      flux_hllc_coupled(F, so.rho, so.u, so.p, enthalpy_of(P, so.rho, so.p), rb, ub, pb, hb, gam, area, n, sco, phb,
                        &Fs);
      out[0] = -Fs;
      if (A) {  // cell-centre state against the boundary state (:5731-5740)
        jac_hllc_coupled_srow(S, 0, own.rho, own.u, own.p, o.hown, ph, rb, ub, pb, hb, phb, gam, area, n);
        for (int k = 0; k < 6; k++) out[2 + k] = S[k];
      }
      {  // the cell against the boundary state; its side-1 scalar is the
         // reconstructed face value (:5744-5775)
        double sArg[3], smag, alpha;
        visc_coupled_bgeom(z.bc, xfa, pc + 5, n, sArg, smag, alpha);
        const double VFs =
            visc_coupled_srow(a0, a1, pc[0], ph, gpc, rb, sco, gpc, diff, DiffTerm, area, n, smag, sArg, alpha);
        out[1] = -VFs;
        if (A) {
          out[8] = a0[0];
          out[9] = a0[1];
        }
      }
    }
  }
}

// The cell's source (sourceHookRansTurbCoupled, TurbModel_SA.h:359-434) after
// its faces' terms acc and D, then the step-mode terms and the stores.
template <typename T>
__device__ __forceinline__ void scal_c_finish(const DMesh &m, const DParams &P, const ScalOwn &o, int c, double acc,
                                              double (&D)[6], double *__restrict__ rhs_s, T *__restrict__ A,
                                              int step_mode, const double *__restrict__ dt, double relax,
                                              T *__restrict__ rhs_ls, T *__restrict__ x0) {
  const double *pc = o.pc, *gpc = o.gpc;
  const double ph = o.ph;
  const double *auxc = P.sa_aux + kAuxN * (size_t)c;
  // source and its diagonal entry (TurbModel_SA.h:359-434)
  {
    const double rho = pc[0], vol = m.vol[c];
    const double mue = auxc[kAuxMue];
    const double Om = Dev::mx(auxc[kAuxVort], 1.e-8);
    const double Strainrate = Dev::mx(auxc[kAuxStr], 1.e-8);
    const double d = auxc[kAuxWall];
    const double chi = ph * rho / mue;
    const double chi_3 = pow(chi, 3.0);
    const double inv_KarmanConst2_d2 = 1.0 / (kSaKarman2 * d * d);
    const double fv1 = chi_3 / (chi_3 + kSaCv1_3);
    const double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    const double BlendTurbulentProduction = 0.0;
    const double Sv = Dev::mx(Om + Dev::mn(0.0, Strainrate - Om) * BlendTurbulentProduction, 0.0);
    const double Shat = Sv + ph * fv2 * inv_KarmanConst2_d2;
    const double inv_Shat = 1.0 / Dev::mx(Shat, 1.0e-8);
    const double r = Dev::mn(ph * inv_Shat * inv_KarmanConst2_d2, 10.0);
    const double gg = r + kSaCw2 * (pow(r, 6.0) - r);
    const double g_6 = pow(gg, 6.0);
    const double fw = gg * pow((1.0 + kSaCw3_6) / (g_6 + kSaCw3_6), 1.0 / 6.0);
    const double H1 = kSaCb1 * Shat * ph;
    const double H2 = -kSaCw1 * fw * ph * ph / pow(d, 2.0);
    const double H3 = kSaCb2 / kSaSigma * dot3(gpc, gpc);
    acc += rho * vol * (H1 + H2 + H3);
    const double d_H1 = 0.0;
    const double d_H2 = -2.0 * kSaCw1 * fw * ph / pow(d, 2.0);
    D[5] -= (d_H1 + d_H2) * vol;
    if (step_mode) {
      D[5] += vol / (dt[c]);
      rhs_ls[6 * (size_t)c + 5] = (T)(relax * acc);
      x0[6 * (size_t)c + 5] = (T)0.0;
    }
  }
  rhs_s[c] = acc;
  if (A) {
    T *dg = A + 36 * (size_t)m.nbocv_i[c] + 30;
    for (int k = 0; k < 6; k++) dg[k] = (T)D[k];
  }
}

// ---------------------------------------------------------------------------
// The nuSA row of the coupled system (runBackwardEulerCoupled with
// Spalart-Allmaras, 6x6 blocks), gathered in the reference's order (see
// scal_c_entry for the face terms, scal_c_finish for the source). Writes
// rhs_s and, with A, row 5 of every block of the cell's row (A [nnz][6][6]
// in precision T); in step mode also V/dt on [5][5], the relaxed right-hand
// side and a zero initial guess of entry 5 (:4455-4473). k_scal_c: one
// thread per cell; k_scal_e: one lane per cell-face entry, the cell's first
// lane summing its entries' terms from LDS in entry order (the same sums).
// ---------------------------------------------------------------------------
template <typename T, bool SND>
__global__ void __launch_bounds__(kBlock) k_scal_c(DMesh m, DParams P, const double *__restrict__ prim,
                                                   const double *__restrict__ grad, const double *__restrict__ bfa,
                                                   const double *__restrict__ gradv, double *__restrict__ rhs_s,
                                                   T *__restrict__ A, int step_mode, const double *__restrict__ dt,
                                                   double relax, T *__restrict__ rhs_ls, T *__restrict__ x0) {
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (t >= m.ncv) return;
  const int c = t;
  const ScalOwn o = scal_own(P, prim, grad, c);
  double acc = 0.0;
  double D[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int e = m.cfl_i[c]; e < m.cfl_i[c + 1]; e++) {
    double t10[10];
    scal_c_entry<T, SND>(m, P, prim, grad, bfa, A, o, e, t10);
    acc += t10[0];
    acc += t10[1];
    for (int k = 0; k < 6; k++) D[k] += t10[2 + k];
    D[0] += t10[8];
    D[5] += t10[9];
  }
  scal_c_finish<T>(m, P, o, c, acc, D, rhs_s, A, step_mode, dt, relax, rhs_ls, x0);
}

constexpr int kScalWaves = 1;

template <typename T, bool SND>
__global__ void __launch_bounds__(64 * kScalWaves) k_scal_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                            const double *__restrict__ grad,
                                                            const double *__restrict__ bfa,
                                                            const double *__restrict__ gradv,
                                                            double *__restrict__ rhs_s, T *__restrict__ A,
                                                            int step_mode, const double *__restrict__ dt,
                                                            double relax, T *__restrict__ rhs_ls,
                                                            T *__restrict__ x0) {
  constexpr int NS = 10;
  __shared__ double lds[64 * kScalWaves * NS];
  double *slot = lds + NS * threadIdx.x;
  const EntryLane L = entry_lane<kScalWaves>(m, hw_blk());
  ScalOwn o;
  if (L.e >= 0) {
    o = scal_own(P, prim, grad, L.c);
    double t10[NS];
    scal_c_entry<T, SND>(m, P, prim, grad, bfa, A, o, L.e, t10);
    for (int k = 0; k < NS; k++) slot[k] = t10[k];
  }
  __syncthreads();
  if (L.e < 0 || L.pos != 0) return;
  const int c = L.c;
  const int nent = m.cfl_i[c + 1] - m.cfl_i[c];
  double acc = 0.0;
  double D[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int j = 0; j < nent; j++) {
    const double *s = slot + NS * j;
    acc += s[0];
    acc += s[1];
    for (int k = 0; k < 6; k++) D[k] += s[2 + k];
    D[0] += s[8];
    D[5] += s[9];
  }
  scal_c_finish<T>(m, P, o, c, acc, D, rhs_s, A, step_mode, dt, relax, rhs_ls, x0);
}

// ---------------------------------------------------------------------------
// writeWallValues of the reference's case class (joe.cpp:43-130), one thread
// per face of a boundary face range [f0, f0 + n): x_fa, the cell's rho, p, T,
// tau_wall = mulam |u_t| / walld sign(u_x), qdot with the reference's wall
// temperature 300, y+, mulam (the face's laminar viscosity from T_bfa), then
// the face's force on the wall: pressure p_bfa n A and shear mulam u_t / walld A
// (n the unit normal, A the area). out [n][16].
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_wall_values(DMesh m, DParams P, const double *__restrict__ q,
                                                        const double *__restrict__ prim,
                                                        const double *__restrict__ bfa, int f0, int n,
                                                        double *__restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int f = f0 + k;
  const int c = m.bfa_cell[f];
  const double *geo = m.fgeo + 8 * (size_t)f;
  const double *pc = prim + 8 * (size_t)c;
  const double *bb = bfa + 8 * (size_t)f;
  const double nv[3] = {geo[0], geo[1], geo[2]};
  const double area = geo[3];
  const double rho = q[c];
  double s_half[3], vel[3], velTang[3];
  for (int i = 0; i < 3; i++) s_half[i] = geo[4 + i] - pc[5 + i];
  for (int i = 0; i < 3; i++) vel[i] = q[(size_t)(1 + i) * m.ncv + c] / rho;
  const double un = dot3(nv, vel);
  for (int i = 0; i < 3; i++) velTang[i] = vel[i] - un * nv[i];
  const double velMag = sqrt(dot3(velTang, velTang));
  const double walld = fabs(dot3(s_half, nv));
  const double wallTemp = 300.0;
  const double mulam = mu_lam(P, bb[5]);
  const double sgn = vel[0] > 0 ? 1.0 : (vel[0] < 0.0 ? -1.0 : 0.0);
  const double tau = mulam * velMag / walld * sgn;
  const double utau = sqrt(fabs(tau) / rho);
  const double yplus = walld * utau / (mulam / rho);
  const double temp = pc[4] / (pc[0] * P.R);
  const double kTotal = P.R * P.gamma / (P.gamma - 1.0) * mulam / P.pr;
  const double qDot = kTotal * (temp - wallTemp) / walld;
  double *o = out + 16 * (size_t)k;
  for (int i = 0; i < 3; i++) o[i] = geo[4 + i];
  o[3] = rho;
  o[4] = pc[4];
  o[5] = temp;
  o[6] = tau;
  o[7] = qDot;
  o[8] = yplus;
  o[9] = mulam;
  for (int i = 0; i < 3; i++) o[10 + i] = bb[4] * nv[i] * area;
  for (int i = 0; i < 3; i++) o[13 + i] = mulam * velTang[i] / walld * area;
}

// ---------------------------------------------------------------------------
// Deterministic two-stage reductions: every block writes its partial and a
// one-block finalize kernel combines them in a fixed order. No device-scope
// fences or tickets: on a multi-XCD part a device-scope release writes back
// the XCD's L2, which made a one-ticket-per-block tail cost milliseconds.
// ---------------------------------------------------------------------------
enum RedOp { kSum = 0, kMax = 1, kMin = 2 };

__device__ __forceinline__ double rcomb(double a, double b, int op) {
  return op == kSum ? a + b : (op == kMax ? Dev::mx(a, b) : Dev::mn(a, b));
}

// result valid in thread 0; sh holds NV * 16 doubles
template <int NV>
__device__ __forceinline__ void block_reduce(double (&v)[NV], double *sh, int op) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int k = 0; k < NV; k++) {
    double x = v[k];
    for (int off = 32; off > 0; off >>= 1) x = rcomb(x, __shfl_down(x, off, 64), op);
    if (lane == 0) sh[k * 16 + wid] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 0; k < NV; k++) {
      double x = sh[k * 16];
      for (int w = 1; w < nw; w++) x = rcomb(x, sh[k * 16 + w], op);
      v[k] = x;
    }
  }
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials, double *sh, int op, const Blk &B) {
  block_reduce<NV>(v, sh, op);
  if (threadIdx.x == 0)
    for (int k = 0; k < NV; k++) partials[(size_t)k * B.n + B.id] = v[k];
}

template <int NV>
__device__ __forceinline__ void block_partials(double (&v)[NV], double *partials, double *sh, int op) {
  block_partials<NV>(v, partials, sh, op, hw_blk());
}

// fixed-order combination of nb partials; valid in thread 0
__device__ double fin_slot(const double *__restrict__ p, int nb, int op, double *sh) {
  const double id = op == kSum ? 0.0 : (op == kMax ? -__builtin_huge_val() : __builtin_huge_val());
  // four independent chains per thread keep four loads in flight
  double x0 = id, x1 = id, x2 = id, x3 = id;
  const int st = blockDim.x;
  int b = threadIdx.x;
  for (; b + 3 * st < nb; b += 4 * st) {
    x0 = rcomb(x0, p[b], op);
    x1 = rcomb(x1, p[b + st], op);
    x2 = rcomb(x2, p[b + 2 * st], op);
    x3 = rcomb(x3, p[b + 3 * st], op);
  }
  for (; b < nb; b += st) x0 = rcomb(x0, p[b], op);
  double v[1] = {rcomb(rcomb(x0, x1, op), rcomb(x2, x3, op), op)};
  __syncthreads();
  block_reduce<1>(v, sh, op);
  return v[0];
}

enum FinOp {
  kFinDt = 0, kFinL1 = 1, kFinAlpha = 2, kFinOmega = 3, kFinResid0 = 4, kFinResid = 5,
  kFinResid0J = 6,  // kFinResid0 of the Jacobi (scalar) solver, which stops at once if already converged
  kFinL1S = 7,      // L1 of the scalar residual
  kFinL1S2 = 8      // L1 of the second scalar's residual (SST omega)
};

// one block: finishes a reduction and applies the scalar update of its owner
constexpr int kFinBlock = 1024;
// slot k of the partials at partials + k * stride (nb values each)
__device__ void fin_body(int nb, int op, const double *__restrict__ partials, size_t stride, SolverScalars *S,
                         double *sh) {
  if ((op == kFinAlpha || op == kFinOmega || op == kFinResid) && (S->done || S->stop)) return;
  const bool t0 = threadIdx.x == 0;
  if (op == kFinDt) {
    double m = fin_slot(partials, nb, kMin, sh);
    if (t0) S->dt_min = m;
  } else if (op == kFinL1) {
    for (int k = 0; k < 5; k++) {
      double v = fin_slot(partials + (size_t)k * stride, nb, kSum, sh);
      if (t0) S->resid[k] = v;
    }
    for (int k = 0; k < 5; k++) {
      double v = fin_slot(partials + (size_t)(5 + k) * stride, nb, kSum, sh);
      if (t0) S->resid_l2[k] = sqrt(v);
    }
  } else if (op == kFinL1S) {
    double v = fin_slot(partials, nb, kSum, sh);
    if (t0) S->resid_s = v;
  } else if (op == kFinL1S2) {
    double v = fin_slot(partials, nb, kSum, sh);
    if (t0) S->resid_s2 = v;
  } else if (op == kFinAlpha) {  // UgpWithCv.cpp:2775-2782
    double gamma = fin_slot(partials, nb, kSum, sh);
    if (t0) {
      if (fabs(gamma) < 1.0E-20) gamma = 1.0E-20;
      S->gamma = gamma;
      S->alpha = S->rho / gamma;
    }
  } else if (op == kFinOmega) {  // :2808-2810
    double st = fin_slot(partials, nb, kSum, sh);
    double tt = fin_slot(partials + stride, nb, kSum, sh);
    if (t0) S->omega = st / (tt + 1.0E-20);
  } else {  // residual: stop test (:2847-2856) and the next iteration's rho, beta (:2736-2749)
    double rho_next = fin_slot(partials, nb, kSum, sh);
    double mxv = fin_slot(partials + stride, nb, kMax, sh);
    if (t0) {
      if (op == kFinResid0 || op == kFinResid0J) {
        S->res0_max = mxv;
        S->rho = 1.0;
        S->alpha = 1.0;
        S->omega = 1.0;
        S->iter = 0;
        S->done = (op == kFinResid0J && mxv <= S->zero_abs) ? 1 : 0;  // UgpWithCv.cpp:2356-2361
      } else {
        S->res_max = mxv;
        if ((mxv <= S->zero_abs) || (mxv / (S->res0_max + 1.0E-12) <= S->zero_rel))
          S->done = 1;
        else if (S->iter > S->maxiter)
          S->done = 2;
      }
      // the host-polled loop launches at most max_launch iterations; a
      // device-driven step stops at the same point (iter counts them)
      if (op == kFinResid && !S->done && S->iter + 1 >= S->max_launch) S->stop = 1;
      if (!S->done) {
        S->iter += 1;
        double rho_prime = -S->omega * S->rho;
        S->rho = rho_next;
        if (fabs(rho_prime) < 1.0E-20) rho_prime = 1.0E-20;
        S->rho_prime = rho_prime;
        S->beta = S->alpha * S->rho / rho_prime;
      }
    }
  }
}

__global__ void __launch_bounds__(kFinBlock) k_fin(int nb, int op, const double *__restrict__ partials,
                                                SolverScalars *S) {
  __shared__ double sh[16];
  fin_body(nb, op, partials, (size_t)nb, S, sh);
}

// first stage of a minimum over many partials: block b reduces
// in[b * kMinPre, (b + 1) * kMinPre) to out[b]. A minimum is exact in any
// order, so two stages give the one-stage result's bits.
constexpr int kMinPre = 2048;
__global__ void __launch_bounds__(256) k_min_pre(const double *__restrict__ in, int n, double *__restrict__ out) {
  __shared__ double sh[16];
  const int b0 = blockIdx.x * kMinPre;
  double m = __builtin_huge_val();
  for (int k = b0 + (int)threadIdx.x; k < min(n, b0 + kMinPre); k += blockDim.x) m = Dev::mn(m, in[k]);
  double v[1] = {m};
  block_reduce<1>(v, sh, kMin);
  if (threadIdx.x == 0) out[blockIdx.x] = v[0];
}

// fin_slot of a kFinBlock-thread block computed by a smaller block (a divisor
// of kFinBlock, whole wavefronts): thread t plays the full block's threads
// t, t + blockDim.x, ..., each wave's lanes in the same shuffle tree, and the
// sixteen wave results are combined in the same order, so the value is the
// same bits as k_fin's. Valid in every thread; sh holds 16 doubles.
__device__ double fin_slot_v(const double *__restrict__ p, int nb, int op, double *sh) {
  const double id = op == kSum ? 0.0 : (op == kMax ? -__builtin_huge_val() : __builtin_huge_val());
  const int st = kFinBlock, lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();  // sh may still be read by a previous call
  for (int r = 0; r < kFinBlock / (int)blockDim.x; r++) {
    double x0 = id, x1 = id, x2 = id, x3 = id;
    int b = threadIdx.x + r * blockDim.x;
    for (; b + 3 * st < nb; b += 4 * st) {
      x0 = rcomb(x0, p[b], op);
      x1 = rcomb(x1, p[b + st], op);
      x2 = rcomb(x2, p[b + 2 * st], op);
      x3 = rcomb(x3, p[b + 3 * st], op);
    }
    for (; b < nb; b += st) x0 = rcomb(x0, p[b], op);
    double x = rcomb(rcomb(x0, x1, op), rcomb(x2, x3, op), op);
    for (int off = 32; off > 0; off >>= 1) x = rcomb(x, __shfl_down(x, off, 64), op);
    if (lane == 0) sh[wid + r * nw] = x;
  }
  __syncthreads();
  double v = sh[0];
  for (int w = 1; w < kFinBlock / 64; w++) v = rcomb(v, sh[w], op);
  return v;
}

// Residual kernel: one cell per thread (flux_cell), block minima of the local
// time step to partials (k_fin kFinDt); CFL_MIN mode: k_dt_fill afterwards.
template <int SCHEME, bool SND>
__global__ void __launch_bounds__(kBlock) k_flux(DMesh m, DParams P, const double *__restrict__ prim,
                                                 const double *__restrict__ grad, const double *__restrict__ bfa,
                                                 const double *__restrict__ q, double *__restrict__ rhs,
                                                 int *__restrict__ nfirst, DtArgs dta, double *__restrict__ dt,
                                                 double *__restrict__ partials) {
  __shared__ double sh[16];
  const int t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  double dtc = 1.0e20;
  if (t < m.ncv) dtc = flux_cell<SCHEME, SND>(m, P, prim, grad, bfa, q, rhs, nfirst, dta, dt, t);
  double v[1] = {dtc};
  block_partials<1>(v, partials, sh, kMin);
}

// Entry-parallel form of k_flux: one lane per cell-face entry computes that
// face's flux (signed for this cell) and the face's wave-speed bound; the
// cell's owner lane adds the terms from LDS in entry order (the reference's
// order), applies the gravity source, and writes the residual and local dt.
// SST: the sides' turbulent kinetic energy in the HLLC flux (kineFA, JoeWithModels.cpp:2924-2930,
// 3140-3147, 3466-3471), the kinetic-energy viscous terms and the SST eddy viscosity
template <int SCHEME, bool SND, bool VISC, int W = kWaves, bool SST = false>
__device__ __forceinline__ void d_flux_e(const Blk &B, DMesh m, DParams P, const double *__restrict__ prim,
                                                   const double *__restrict__ grad, const double *__restrict__ bfa,
                                                   const double *__restrict__ q, double *__restrict__ rhs,
                                                   int *__restrict__ nfirst, DtArgs dta, double *__restrict__ dt,
                                                   double *__restrict__ partials, const double *__restrict__ gradv) {
  // inviscid 5, wave speed, viscous rhou/rhoE 4 (coupled: all 5), SUM_BASED diffusive term
  constexpr bool CPL = SCHEME == kHllcC;
  constexpr int NS = VISC ? (CPL ? 12 : 11) : 6;
  constexpr int kSumSlot = CPL ? 11 : 10;
  __shared__ double lds[64 * W * NS];
  __shared__ double sh[16];
  double *slot = lds + NS * threadIdx.x;
  const EntryLane L = entry_lane<W>(m, B);
  const double gam = P.gamma;
  double pc[8], gc[15];
  if (L.e >= 0) {
    const double *s = prim + 8 * (size_t)L.c;
    for (int k = 0; k < 8; k += 2) {
      double2 v = ld2(s + k);
      pc[k] = v.x;
      pc[k + 1] = v.y;
    }
    if (SND) {
      const double *g = grad + 16 * (size_t)L.c;
      for (int k = 0; k < 14; k += 2) {
        double2 v = ld2(g + k);
        gc[k] = v.x;
        gc[k + 1] = v.y;
      }
      gc[14] = g[14];
    }
    const int2 en = m.cfl[L.e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    double n[3], area, xfa[3];
    {
      double2 a = ld2(geo), b = ld2(geo + 2), cc = ld2(geo + 4);
      n[0] = a.x; n[1] = a.y; n[2] = b.x; area = b.y;
      xfa[0] = cc.x; xfa[1] = cc.y; xfa[2] = geo[6];
    }
    double lambda;
    {
      const double ru[3] = {q[m.ncv + L.c], q[2 * m.ncv + L.c], q[3 * m.ncv + L.c]};
      const double cs = sqrt(gam * pc[4] / pc[0]);
      const double Uk = dot3(ru, n) / pc[0];
      lambda = (fabs(Uk) + cs) * area;
    }
    Side own_cc;
    own_cc.rho = pc[0];
    own_cc.u[0] = pc[1]; own_cc.u[1] = pc[2]; own_cc.u[2] = pc[3];
    own_cc.p = pc[4];
    double F[5];
    double sgn = -1.0;
    bool low = false;
    double vc[4] = {-0.0, -0.0, -0.0, -0.0};
    double vcc[5] = {-0.0, -0.0, -0.0, -0.0, -0.0};  // the coupled viscous terms (-0.0: nothing added)
    double mul = 0.0, mut = 0.0;
    // each lane's two sides are chosen first (this cell reconstructed at the
    // face, the neighbour, or the boundary state) and the flux evaluated once
    // after the branches, so a wavefront whose lanes face different ways, or
    // mixes internal and boundary faces, runs the flux arithmetic once
    Side fl, fr;
    double kl = 0.0, kr = 0.0;  // SST: kineFA of the flux's two sides
    bool symm_analytic = false;
    double hl = 0.0, hr = 0.0;  // side enthalpies of the coupled flux
    if (en.y >= 0) {
      const bool s0 = (en.x >= 0);
      double pn[8], gn[15];
      const double *sp = prim + 8 * (size_t)en.y;
      for (int k = 0; k < 8; k += 2) {
        double2 v = ld2(sp + k);
        pn[k] = v.x;
        pn[k + 1] = v.y;
      }
      Side so, sn;
      bool okn = true;
      if (SND) {
        const double *g = grad + 16 * (size_t)en.y;
        for (int k = 0; k < 14; k += 2) {
          double2 v = ld2(g + k);
          gn[k] = v.x;
          gn[k + 1] = v.y;
        }
        gn[14] = g[14];
        low = !reconstruct(so, pc, gc, xfa);
        okn = reconstruct(sn, pn, gn, xfa);
      } else {
        so = own_cc;
        sn.rho = pn[0]; sn.u[0] = pn[1]; sn.u[1] = pn[2]; sn.u[2] = pn[3]; sn.p = pn[4];
      }
      fl = s0 ? so : sn;
      fr = s0 ? sn : so;
      if (!s0) sgn = 1.0;
      double ko = 0.0, kn = 0.0;  // SST: kine of the cell and the neighbour
      if (SST) {
        ko = P.sst_phi[L.c];
        kn = P.sst_phi[en.y];
        double kfo = ko, kfn = kn;
        if (SND && !low) {
          const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          kfo = kfo + dot3(r, P.sst_aux + kSstN * (size_t)L.c + kSGradK);
        }
        if (SND && okn) {
          const double r[3] = {xfa[0] - pn[5], xfa[1] - pn[6], xfa[2] - pn[7]};
          kfn = kfn + dot3(r, P.sst_aux + kSstN * (size_t)en.y + kSGradK);
        }
        kl = s0 ? kfo : kfn;
        kr = s0 ? kfn : kfo;
      }
      if (SCHEME == kHllcC) {  // calcThermoProp_p after reconstruction, or the cell's (JoeWithModels.cpp:4920-4928)
        hl = enthalpy_of(P, fl.rho, fl.p);
        hr = enthalpy_of(P, fr.rho, fr.p);
      }
      if (VISC) {  // calcViscousFluxNS internal faces (JoeWithModels.cpp:3635-3710)
        VSide vo = vside_cell(P, pc, gradv + 12 * (size_t)L.c);
        VSide vn = vside_cell(P, pn, gradv + 12 * (size_t)en.y);
        if (SST) {
          vo.k = ko;
          vn.k = kn;
        }
        const VSide &a = s0 ? vo : vn, &b = s0 ? vn : vo;
        double sV[3], smag, w0, w1;
        visc_internal_geom(P, s0 ? pc + 5 : pn + 5, s0 ? pn + 5 : pc + 5, xfa, a.T, b.T, sV, smag, mul, w0, w1);
        double kine_fa = 0.0;
        if (SST) {
          mut = sst_mut_face(P, m.ncv, w0, w1, a.rho, b.rho, s0 ? L.c : en.y, s0 ? en.y : L.c);
          kine_fa = (w1 * a.k + w0 * b.k) / (w0 + w1);
        } else if (P.sa) {
          const double nuo = P.sa_phi[L.c], nun = P.sa_phi[en.y];
          mut = sa_mut_face(w0, w1, a.rho, b.rho, s0 ? nuo : nun, s0 ? nun : nuo, mul);
        }
        const double ufa[3] = {0.5 * (a.u[0] + b.u[0]), 0.5 * (a.u[1] + b.u[1]), 0.5 * (a.u[2] + b.u[2])};
        if (CPL) {  // calcFluxCoupled internal faces (JoeWithModels.cpp:5076-5094)
          double VF[5], tau[3], keff;
          visc_coupled(P, a, b, mul, mut, ufa, area, n, smag, sV, dot3(n, sV), VF, tau, keff, P.sa != 0);
          for (int k = 0; k < 5; k++) vcc[k] = s0 ? -VF[k] : VF[k];
        } else {
          double Fu[3], FE;
          VFace vf;
          visc_face<SST>(P, a, b, mul, mut, ufa, area, n, smag, sV, Fu, FE, vf, kine_fa);
          for (int k = 0; k < 3; k++) vc[k] = s0 ? -Fu[k] : Fu[k];
          vc[3] = s0 ? -FE : FE;
        }
      }
    } else {
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      Side sb;
      sb.rho = bb[0]; sb.u[0] = bb[1]; sb.u[1] = bb[2]; sb.u[2] = bb[3]; sb.p = bb[4];
      hl = bb[6];
      hr = bb[6];
      const double kb = SST ? P.sst_bfa[2 * (size_t)f] : 0.0;
      if (z.bc == kBcSymmetry || z.bc == kBcWall || (SCHEME == kHllcC && z.bc == kBcNeumann)) {
        fl = sb;
        fr = sb;
        kl = kb;
        kr = kb;
        // the coupled loop treats walls as symmetry faces and NEUMANN faces
        // with the boundary state on both sides (JoeWithModels.cpp:5150-5470)
        symm_analytic = SCHEME == kHllcC ? (z.bc != kBcNeumann && P.symm_bc == kSymmAnalytical)
                                         : (z.bc == kBcSymmetry && P.symm_bc == kSymmAnalytical);
      } else {
        Side so = own_cc;
        if (SND) low = !reconstruct(so, pc, gc, xfa);
        fl = so;
        fr = sb;
        if (SST) {
          kl = P.sst_phi[L.c];
          if (SND && !low) {
            const double r[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
            kl = kl + dot3(r, P.sst_aux + kSstN * (size_t)L.c + kSGradK);
          }
          kr = kb;
        }
        if (SCHEME == kHllcC) hl = enthalpy_of(P, so.rho, so.p);
      }
      if (VISC && CPL) {  // calcFluxCoupled boundary faces; mul from T_bfa
        mul = mu_lam(P, bb[5]);
        if (P.sa && z.bc != kBcWall)  // mut_fa of calcRansTurbViscMuet (TurbModel_SA.h:112-120)
          mut = sa_mut(pc[0] * P.sa_phi[L.c], P.sa_aux[kAuxN * (size_t)L.c + kAuxMue]);
        const double *gvo = gradv + 12 * (size_t)L.c;
        if (z.bc == kBcSymmetry) {  // 2/3 (mu + mut) tr(grad u) on the momentum only (:5272-5284)
          double tmp = 2.0 / 3.0 * ((mul + mut) * (gvo[0] + gvo[4] + gvo[8]) + bb[0] * 0.0);
          for (int i = 0; i < 3; i++) vcc[1 + i] = -(area * tmp * n[i]);
        } else {
          const VSide vo = vside_cell(P, pc, gvo);
          const VSide vb = vside_bfa(bb, gvo);
          double sArg[3], smag, alpha;
          visc_coupled_bgeom(z.bc, xfa, pc + 5, n, sArg, smag, alpha);
          double VF[5], tau[3], keff;
          visc_coupled(P, vsel(z.bc == kBcNeumann, vb, vo), vb, mul, z.bc == kBcWall ? 0.0 : mut, vb.u, area, n,
                       smag, sArg, alpha, VF, tau, keff, P.sa != 0);
          for (int k = 0; k < 5; k++) vcc[k] = -VF[k];
        }
      } else if (VISC) {  // boundary faces (:3714-3851); boundary mul from T_bfa (ComputeBCProperties_T)
        mul = mu_lam(P, bb[5]);
        if (SST && z.bc != kBcWall)  // mut of every non-wall boundary face, symmetry included (calcDt reads it)
          mut = sst_mut_cell(P, m.ncv, pc[0], L.c);
        else if (P.sa && z.bc != kBcWall)
          mut = sa_mut(pc[0] * P.sa_phi[L.c], P.sa_aux[kAuxN * (size_t)L.c + kAuxMue]);
        if (SST && z.bc == kBcSymmetry) {  // -1/3 (rho + rho_bfa) kine_fa on the momentum (:3727-3745)
          const double tmp = 1.0 / 3.0 * (pc[0] + bb[0]) * kb;
          const double *fan = m.fnorm + 3 * (size_t)f;
          for (int k = 0; k < 3; k++) vc[k] = -(tmp * fan[k]);
        }
        if (z.bc != kBcSymmetry) {
          const double *gvo = gradv + 12 * (size_t)L.c;
          VSide vo = vside_cell(P, pc, gvo);
          VSide vb = vside_bfa(bb, gvo);
          double kine_fa = 0.0;
          if (SST) {  // walls: kine1 = kine_fa = 0 (:3767-3776); other zones phi_bfa (:3817-3826)
            vo.k = P.sst_phi[L.c];
            vb.k = z.bc == kBcWall ? 0.0 : kb;
            kine_fa = vb.k;
          }
          double sV[3] = {xfa[0] - pc[5], xfa[1] - pc[6], xfa[2] - pc[7]};
          double smag;
          const double *sArg;
          if (z.bc == kBcWall) {
            smag = fabs(dot3(sV, n));
            sArg = n;
          } else {
            smag = sqrt(dot3(sV, sV));
            for (int i = 0; i < 3; i++) sV[i] /= smag;
            sArg = sV;
          }
          double Fu[3], FE;
          VFace vf;
          visc_face<SST>(P, vo, vb, mul, mut, vb.u, area, n, smag, sArg, Fu, FE, vf, kine_fa);
          for (int k = 0; k < 3; k++) vc[k] = -Fu[k];
          vc[3] = -FE;
        }
      }
    }
    if (SCHEME == kHllcC)
      flux_hllc_coupled(F, fl.rho, fl.u, fl.p, hl, fr.rho, fr.u, fr.p, hr, gam, area, n);
    else
      face_flux(SCHEME, F, fl, fr, gam, area, n, kl, kr);
    if (symm_analytic) {  // the ONLY_PRESSURE flux
      F[0] = 0.0;
      for (int i = 0; i < 3; i++) F[1 + i] = fr.p * n[i] * area;
      F[4] = 0.0;
    }
    if (P.fdump && en.x >= 0)
      for (int k = 0; k < 5; k++) P.fdump[5 * (size_t)f + k] = F[k];
    if (VISC && P.vdump && en.x >= 0) {
      P.vdump[2 * (size_t)f] = mul;
      P.vdump[2 * (size_t)f + 1] = mut;
    }
    if (VISC) {
      if (CPL)
        for (int k = 0; k < 5; k++) slot[6 + k] = vcc[k];
      else
        for (int k = 0; k < 4; k++) slot[6 + k] = vc[k];
      if (dta.sum) {  // SUM_BASED diffusive term (UgpWithCvCompFlow.cpp:75-88)
        const double volume = en.y >= 0 ? 0.5 * (m.vol[L.c] + m.vol[en.y]) : 0.5 * m.vol[L.c];
        const double KinVisc = (4.0 / 3.0) * (mul + mut) / pc[0];
        slot[kSumSlot] = area * area * KinVisc / volume;
      } else {  // MAX_BASED viscous bound (:50-51)
        const double lamdiff = 4.0 * (mul + mut) / pc[0] * area * area / m.vol[L.c];
        lambda = Dev::mx(lambda, lamdiff);
      }
    }
    slot[5] = lambda;
    // acc - F == acc + (-F) exactly, so the owner only adds
    for (int k = 0; k < 5; k++) slot[k] = sgn < 0.0 ? -F[k] : F[k];
    if (low) atomicAdd(nfirst, 1);
  }
  __syncthreads();
  double dtc = 1.0e20;
  if (L.e >= 0 && L.pos == 0) {
    const int c = L.c;
    const int nent = m.cfl_i[c + 1] - m.cfl_i[c];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lambdaMax = 0.0;
    for (int j = 0; j < nent; j++) {
      const double *s = slot + NS * j;
      for (int k = 0; k < 5; k++) acc[k] += s[k];
      if (VISC && CPL)  // the coupled loop adds each face's viscous flux after its Euler flux (:5024-5094)
        for (int k = 0; k < 5; k++) acc[k] += s[6 + k];
      lambdaMax = Dev::mx(lambdaMax, s[5]);
    }
    if (VISC && !CPL)  // the viscous pass follows the Euler pass (JoeWithModels.cpp:2585-2589)
      for (int j = 0; j < nent; j++) {
        const double *s = slot + NS * j + 6;
        for (int k = 0; k < 4; k++) acc[1 + k] += s[k];
      }
    // sourceHook: gravity (JoeWithModels.h:111-129)
    const double vol = m.vol[c];
    if (P.gravity_on) {
      const double ru[3] = {q[m.ncv + c], q[2 * m.ncv + c], q[3 * m.ncv + c]};
      acc[1] += vol * pc[0] * P.g[0];
      acc[2] += vol * pc[0] * P.g[1];
      acc[3] += vol * pc[0] * P.g[2];
      acc[4] += vol * (P.g[0] * ru[0] + P.g[1] * ru[1] + P.g[2] * ru[2]);
    } else {
      acc[1] += 0.0;
      acc[2] += 0.0;
      acc[3] += 0.0;
      bool allneg = signbit(pc[1]) && signbit(pc[2]) && signbit(pc[3]);
      acc[4] += allneg ? -0.0 : 0.0;
    }
    for (int k = 0; k < 5; k++) rhs[(size_t)k * m.ncv + c] = acc[k];
    const double cflv = dta.cfl_dev ? *dta.cfl_dev : dta.cfl;
    if (dta.mode == 2) {
      dtc = dta.dt_fixed;
    } else if (dta.sum) {  // SUM_BASED (UgpWithCvCompFlow.cpp:62-99): sums in faocv order
      double lc = 0.0, ld = 0.0;
      for (int k = 0; k < nent; k++) {
        const double *s = slot + NS * m.foc[m.cfl_i[c] + k];
        lc += s[5];
        if (VISC) ld += s[kSumSlot];
      }
      dtc = Dev::mn(cflv * vol / lc, VISC ? 0.25 * cflv * vol / ld : 1.0e20);
    } else {
      dtc = cflv * vol / lambdaMax;
    }
    dt[c] = dtc;
  }
  double v[1] = {dtc};
  block_partials<1>(v, partials, sh, kMin, B);
}

// one wavefront per block: the owner lanes' sum waits only for its own
// wavefront's lanes, not for the slowest of a larger block
constexpr int kFluxWaves = 1;

template <int SCHEME, bool SND, bool VISC, bool SST = false>
__global__ void __launch_bounds__(64 * kFluxWaves) k_flux_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                            const double *__restrict__ grad,
                                                            const double *__restrict__ bfa,
                                                            const double *__restrict__ q, double *__restrict__ rhs,
                                                            int *__restrict__ nfirst, DtArgs dta,
                                                            double *__restrict__ dt, double *__restrict__ partials,
                                                            const double *__restrict__ gradv) {
  d_flux_e<SCHEME, SND, VISC, kFluxWaves, SST>(hw_blk(), m, P, prim, grad, bfa, q, rhs, nfirst, dta, dt, partials,
                                               gradv);
}

// Face form of the inviscid flux pass: k_face_flux evaluates each internal
// face once (one thread per face, device face order, icv0 -> icv1
// orientation) into fflux[nfa_i][5]; k_gather_e adds the faces' fluxes to
// their cells in entry order (the reference's), computes the boundary fluxes
// and the local time step. The flux of a face does not depend on which cell
// asks for it, so the sums are the same bits as k_flux_e's, with half the
// flux arithmetic and neighbour gathers, for 2 x 40 B per internal face.
template <int SCHEME, bool SND>
__global__ void __launch_bounds__(kBlock) k_face_flux(DMesh m, DParams P, const double *__restrict__ prim,
                                                      const double *__restrict__ grad, const int2 *__restrict__ fcell,
                                                      double *__restrict__ fflux, int *__restrict__ nfirst) {
  const int i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int nfi = m.nfa - m.nfa_b;
  if (i >= nfi) return;
  const int f = m.nfa_b + i;
  const int2 cc = fcell[i];
  const double *geo = m.fgeo + 8 * (size_t)f;
  double n[3], area, xfa[3];
  {
    double2 a = ld2(geo), b = ld2(geo + 2), c = ld2(geo + 4);
    n[0] = a.x; n[1] = a.y; n[2] = b.x; area = b.y;
    xfa[0] = c.x; xfa[1] = c.y; xfa[2] = geo[6];
  }
  double p0[8], p1[8];
  {
    const double *s0 = prim + 8 * (size_t)cc.x, *s1 = prim + 8 * (size_t)cc.y;
    for (int k = 0; k < 8; k += 2) {
      double2 v = ld2(s0 + k), w = ld2(s1 + k);
      p0[k] = v.x; p0[k + 1] = v.y;
      p1[k] = w.x; p1[k + 1] = w.y;
    }
  }
  Side L, R;
  int low = 0;
  if (SND) {
    double g[15];
    const double *g0 = grad + 16 * (size_t)cc.x;
    for (int k = 0; k < 14; k += 2) {
      double2 v = ld2(g0 + k);
      g[k] = v.x; g[k + 1] = v.y;
    }
    g[14] = g0[14];
    low += !reconstruct(L, p0, g, xfa);
    const double *g1 = grad + 16 * (size_t)cc.y;
    for (int k = 0; k < 14; k += 2) {
      double2 v = ld2(g1 + k);
      g[k] = v.x; g[k + 1] = v.y;
    }
    g[14] = g1[14];
    low += !reconstruct(R, p1, g, xfa);
  } else {
    L.rho = p0[0]; L.u[0] = p0[1]; L.u[1] = p0[2]; L.u[2] = p0[3]; L.p = p0[4];
    R.rho = p1[0]; R.u[0] = p1[1]; R.u[1] = p1[2]; R.u[2] = p1[3]; R.p = p1[4];
  }
  double F[5];
  face_flux(SCHEME, F, L, R, P.gamma, area, n);
  double *o = fflux + 5 * (size_t)i;
  for (int k = 0; k < 5; k++) o[k] = F[k];
  if (low) atomicAdd(nfirst, low);
}

template <int SCHEME, bool SND>
__global__ void __launch_bounds__(kBlock) k_gather_e(DMesh m, DParams P, const double *__restrict__ prim,
                                                     const double *__restrict__ grad, const double *__restrict__ bfa,
                                                     const double *__restrict__ q, const double *__restrict__ fflux,
                                                     double *__restrict__ rhs, int *__restrict__ nfirst, DtArgs dta,
                                                     double *__restrict__ dt, double *__restrict__ partials) {
  constexpr int NS = 6;
  __shared__ double lds[kBlock * NS];
  __shared__ double sh[16];
  double *slot = lds + NS * threadIdx.x;
  const EntryLane L = entry_lane(m);
  const double gam = P.gamma;
  double pc[8];
  if (L.e >= 0) {
    const double *s = prim + 8 * (size_t)L.c;
    for (int k = 0; k < 8; k += 2) {
      double2 v = ld2(s + k);
      pc[k] = v.x;
      pc[k + 1] = v.y;
    }
    const int2 en = m.cfl[L.e];
    const int f = en.x & 0x7fffffff;
    const double *geo = m.fgeo + 8 * (size_t)f;
    double n[3], area;
    {
      double2 a = ld2(geo), b = ld2(geo + 2);
      n[0] = a.x; n[1] = a.y; n[2] = b.x; area = b.y;
    }
    double lambda;
    {
      const double ru[3] = {q[m.ncv + L.c], q[2 * m.ncv + L.c], q[3 * m.ncv + L.c]};
      const double cs = sqrt(gam * pc[4] / pc[0]);
      const double Uk = dot3(ru, n) / pc[0];
      lambda = (fabs(Uk) + cs) * area;
    }
    double F[5];
    double sgn = -1.0;
    if (en.y >= 0) {
      const double *ff = fflux + 5 * (size_t)(f - m.nfa_b);
      for (int k = 0; k < 5; k++) F[k] = ff[k];
      if (en.x < 0) sgn = 1.0;
    } else {
      const double xfa[3] = {geo[4], geo[5], geo[6]};
      const ZoneDev &z = m.zones[-en.y - 1];
      const double *bb = bfa + 8 * (size_t)f;
      Side sb;
      sb.rho = bb[0]; sb.u[0] = bb[1]; sb.u[1] = bb[2]; sb.u[2] = bb[3]; sb.p = bb[4];
      if (z.bc == kBcSymmetry || z.bc == kBcWall) {
        face_flux(SCHEME, F, sb, sb, gam, area, n);
        if (z.bc == kBcSymmetry && P.symm_bc == kSymmAnalytical) {
          F[0] = 0.0;
          for (int i = 0; i < 3; i++) F[1 + i] = sb.p * n[i] * area;
          F[4] = 0.0;
        }
      } else {
        Side so;
        so.rho = pc[0]; so.u[0] = pc[1]; so.u[1] = pc[2]; so.u[2] = pc[3]; so.p = pc[4];
        if (SND) {
          double gc[15];
          const double *g = grad + 16 * (size_t)L.c;
          for (int k = 0; k < 15; k++) gc[k] = g[k];
          if (!reconstruct(so, pc, gc, xfa)) atomicAdd(nfirst, 1);
        }
        face_flux(SCHEME, F, so, sb, gam, area, n);
      }
    }
    slot[5] = lambda;
    for (int k = 0; k < 5; k++) slot[k] = sgn < 0.0 ? -F[k] : F[k];
  }
  __syncthreads();
  double dtc = 1.0e20;
  if (L.e >= 0 && L.pos == 0) {
    const int c = L.c;
    const int nent = m.cfl_i[c + 1] - m.cfl_i[c];
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lambdaMax = 0.0;
    for (int j = 0; j < nent; j++) {
      const double *s = slot + NS * j;
      for (int k = 0; k < 5; k++) acc[k] += s[k];
      lambdaMax = Dev::mx(lambdaMax, s[5]);
    }
    const double vol = m.vol[c];
    if (P.gravity_on) {
      const double ru[3] = {q[m.ncv + c], q[2 * m.ncv + c], q[3 * m.ncv + c]};
      acc[1] += vol * pc[0] * P.g[0];
      acc[2] += vol * pc[0] * P.g[1];
      acc[3] += vol * pc[0] * P.g[2];
      acc[4] += vol * (P.g[0] * ru[0] + P.g[1] * ru[1] + P.g[2] * ru[2]);
    } else {
      acc[1] += 0.0;
      acc[2] += 0.0;
      acc[3] += 0.0;
      bool allneg = signbit(pc[1]) && signbit(pc[2]) && signbit(pc[3]);
      acc[4] += allneg ? -0.0 : 0.0;
    }
    for (int k = 0; k < 5; k++) rhs[(size_t)k * m.ncv + c] = acc[k];
    const double cflv = dta.cfl_dev ? *dta.cfl_dev : dta.cfl;
    if (dta.mode == 2) {
      dtc = dta.dt_fixed;
    } else if (dta.sum) {  // SUM_BASED, inviscid (UgpWithCvCompFlow.cpp:62-99): sums in faocv order
      double lc = 0.0;
      for (int k = 0; k < nent; k++) lc += slot[NS * m.foc[m.cfl_i[c] + k] + 5];
      dtc = Dev::mn(cflv * vol / lc, 1.0e20);
    } else {
      dtc = cflv * vol / lambdaMax;
    }
    dt[c] = dtc;
  }
  double v[1] = {dtc};
  block_partials<1>(v, partials, sh, kMin);
}

__device__ __forceinline__ void d_dt_fill(const Blk &B, int ncv, double *dt, const SolverScalars *S) {
  int c = B.id * blockDim.x + threadIdx.x;
  if (c < ncv) dt[c] = S->dt_min;
}

__global__ void k_dt_fill(int ncv, double *dt, const SolverScalars *S) {
  d_dt_fill(hw_blk(), ncv, dt, S);
}

// CFL ramp of the pseudo-time loop (JoeWithModels.cpp:975-978) on the device
__device__ __forceinline__ void d_ramp(const Blk &B, SolverScalars *S, int after, int interval, double factor, double cfl_max) {
  if (threadIdx.x != 0) return;
  const int st = S->step_no + 1;
  S->step_no = st;
  if ((st >= after) && (interval > 0) && (st % interval == 0) && (S->cfl_cur < cfl_max)) S->cfl_cur *= factor;
}

__global__ void k_ramp(SolverScalars *S, int after, int interval, double factor, double cfl_max) {
  d_ramp(hw_blk(), S, after, interval, factor, cfl_max);
}

__global__ void k_set_step(SolverScalars *S, int step, double cfl) {
  if (threadIdx.x != 0) return;
  S->step_no = step;
  S->cfl_cur = cfl;
}

// BiCGStab scalars at the start of a solve (leaves dt_min, residuals, step)
__device__ __forceinline__ void d_solver_init(const Blk &B, SolverScalars *S, int maxiter, double zero_abs, double zero_rel) {
  if (threadIdx.x != 0) return;
  S->rho = S->rho_prime = S->alpha = S->omega = S->beta = S->gamma = 0.0;
  S->res0_max = S->res_max = 0.0;
  S->iter = 0;
  S->done = 0;
  S->stop = 0;
  S->max_launch = maxiter + 1;
  S->maxiter = maxiter;
  S->zero_abs = zero_abs;
  S->zero_rel = zero_rel;
}

__global__ void k_solver_init(SolverScalars *S, int maxiter, double zero_abs, double zero_rel) {
  d_solver_init(hw_blk(), S, maxiter, zero_abs, zero_rel);
}

template <typename T>
__device__ __forceinline__ void d_fill(const Blk &B, size_t n, T *__restrict__ p, T v) {
  const size_t i = (size_t)B.id * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

template <typename T>
__global__ void k_fill(size_t n, T *__restrict__ p, T v) {
  d_fill<T>(hw_blk(), n, p, v);
}

__device__ __forceinline__ void d_publish(const Blk &B, const SolverScalars *S, StepMail *out) {
  if (threadIdx.x != 0) return;
  out->solver_stop = S->done || S->stop;
  out->step = S->step_no;
  out->lin_iters = S->iter;
  out->lin_status = S->done;
  out->cfl = S->cfl_cur;
  out->dt_min = S->dt_min;
  for (int k = 0; k < 5; k++) out->resid[k] = S->resid[k];
  for (int k = 0; k < 5; k++) out->resid_l2[k] = S->resid_l2[k];
  __threadfence_system();
}

__global__ void k_publish(const SolverScalars *S, StepMail *out) {
  d_publish(hw_blk(), S, out);
}

// ---------------------------------------------------------------------------
// block-Jacobi BiCGStab, UgpWithCv.cpp:2660-2876, on NB x NB blocks (NB = 5,
// or 5 + nScal in the coupled loop). Vectors are [ncv][NB]; the
// matrix-vector products keep the reference's per-row order (blocks in
// nbocv order, columns ascending) using NB lanes per cell.
// ---------------------------------------------------------------------------
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_lu(int ncv, const int *__restrict__ nbocv_i, const T *__restrict__ A,
                                               T *__restrict__ LU) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  T a[NB][NB];
  const T *s = A + NB * NB * (size_t)nbocv_i[c];
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) a[i][j] = s[NB * i + j];
  ludeco_n<T, NB>(a);
  T *d = LU + NB * NB * (size_t)c;
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) d[NB * i + j] = a[i][j];
}

// y = A x for row component (c, i) = (tid / NB, tid % NB)
template <typename T, int NB>
__device__ __forceinline__ T spmv_row(const int *__restrict__ nbocv_i, const int *__restrict__ nbocv_v,
                                      const T *__restrict__ A, const T *__restrict__ x, int c, int i) {
  T acc = (T)0.0;
  for (int noc = nbocv_i[c]; noc < nbocv_i[c + 1]; noc++) {
    const T *a = A + NB * NB * (size_t)noc + NB * i;
    const T *xv = x + NB * (size_t)nbocv_v[noc];
    for (int j = 0; j < NB; j++) acc += a[j] * xv[j];
  }
  return acc;
}

// cells per block of the NB-lanes-per-cell kernels (5: 51 cells, 255 lanes)
template <int NB>
constexpr int cells_per_block() {
  return kBlock / NB;
}
constexpr int kCellsPerBlock = cells_per_block<5>();

// res = rhs - A phi; init: res0 = res. Partials: res.res0 (slot 0) and
// max |LU^-1 res| (slot 1), finished by k_fin kFinResid0 / kFinResid.
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_resid(int ncv, const int *__restrict__ nbocv_i,
                                                  const int *__restrict__ nbocv_v, const T *__restrict__ A,
                                                  const T *__restrict__ LU, const T *__restrict__ phi,
                                                  const T *__restrict__ rhs, T *__restrict__ res, T *__restrict__ res0,
                                                  int init, double *partials, const SolverScalars *S) {
  constexpr int CPB = cells_per_block<NB>();
  __shared__ double sh[32];
  __shared__ T cellv[CPB][NB];
  if (!init && (S->done || S->stop)) return;
  const int lc = threadIdx.x / NB, i = threadIdx.x % NB;
  const int c = blockIdx.x * CPB + lc;
  const bool active = (lc < CPB) && (c < ncv);
  T r = (T)0.0;
  if (active) {
    T ax = spmv_row<T, NB>(nbocv_i, nbocv_v, A, phi, c, i);
    r = rhs[NB * (size_t)c + i] - ax;
    res[NB * (size_t)c + i] = r;
    if (init) res0[NB * (size_t)c + i] = r;
    cellv[lc][i] = r;
  }
  __syncthreads();
  double vs[1] = {0.0}, vm[1] = {0.0};
  if (active) {
    T r0v = init ? r : res0[NB * (size_t)c + i];
    vs[0] = (double)r * (double)r0v;
    T tmp[NB];
    for (int k = 0; k < NB; k++) tmp[k] = cellv[lc][k];
    lusolv_n<T, NB>(LU + NB * NB * (size_t)c, tmp);
    vm[0] = fabs((double)tmp[i]);
  }
  block_reduce<1>(vs, sh, kSum);
  __syncthreads();
  block_reduce<1>(vm, sh + 16, kMax);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = vs[0];
    partials[gridDim.x + blockIdx.x] = vm[0];
  }
}

// res = s - omega t: the residual by the BiCGStab recurrence instead of
// rhs - A phi (no product with A), for the FP32 solve when asked
// (GPU_KRYLOV_RESIDUAL = RECURRENCE); partials as k_resid's
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_resid_rec(int ncv, const T *__restrict__ LU, const T *__restrict__ s,
                                                      const T *__restrict__ t, T *__restrict__ res,
                                                      const T *__restrict__ res0, double *partials,
                                                      const SolverScalars *S) {
  constexpr int CPB = cells_per_block<NB>();
  __shared__ double sh[32];
  __shared__ T cellv[CPB][NB];
  if (S->done || S->stop) return;
  const int lc = threadIdx.x / NB, i = threadIdx.x % NB;
  const int c = blockIdx.x * CPB + lc;
  const bool active = (lc < CPB) && (c < ncv);
  T r = (T)0.0;
  if (active) {
    const size_t o = NB * (size_t)c + i;
    r = s[o] - (T)S->omega * t[o];
    res[o] = r;
    cellv[lc][i] = r;
  }
  __syncthreads();
  double vs[1] = {0.0}, vm[1] = {0.0};
  if (active) {
    vs[0] = (double)r * (double)res0[NB * (size_t)c + i];
    T tmp[NB];
    for (int k = 0; k < NB; k++) tmp[k] = cellv[lc][k];
    lusolv_n<T, NB>(LU + NB * NB * (size_t)c, tmp);
    vm[0] = fabs((double)tmp[i]);
  }
  block_reduce<1>(vs, sh, kSum);
  __syncthreads();
  block_reduce<1>(vm, sh + 16, kMax);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = vs[0];
    partials[gridDim.x + blockIdx.x] = vm[0];
  }
}

// p = res - beta (p - omega v); phat = LU^-1 p   (one thread per cell). With
// BJ = false only p: the ILU solve forms phat after it.
template <typename T, int NB, bool BJ = true>
__global__ void __launch_bounds__(kBlock) k_p(int ncv, const T *__restrict__ LU, const T *__restrict__ res,
                                              const T *__restrict__ v, T *__restrict__ p, T *__restrict__ phat,
                                              const SolverScalars *S) {
  if (S->done || S->stop) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const T beta = (T)S->beta, omega = (T)S->omega;
  T t[NB];
  for (int k = 0; k < NB; k++) {
    size_t o = NB * (size_t)c + k;
    T pv = res[o] - beta * (p[o] - omega * v[o]);
    p[o] = pv;
    t[k] = pv;
  }
  if (!BJ) return;
  lusolv_n<T, NB>(LU + NB * NB * (size_t)c, t);
  for (int k = 0; k < NB; k++) phat[NB * (size_t)c + k] = t[k];
}

// y = A x with partial dots (y.other, y.y): mode 0 gives gamma = v.res0
// (k_fin kFinAlpha), mode 1 gives (s.t, t.t) (k_fin kFinOmega)
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_spmv_dot(int ncv, const int *__restrict__ nbocv_i,
                                                     const int *__restrict__ nbocv_v, const T *__restrict__ A,
                                                     const T *__restrict__ x, T *__restrict__ y,
                                                     const T *__restrict__ other, double *partials,
                                                     const SolverScalars *S) {
  constexpr int CPB = cells_per_block<NB>();
  __shared__ double sh[32];
  if (S->done || S->stop) return;
  const int lc = threadIdx.x / NB, i = threadIdx.x % NB;
  const int c = blockIdx.x * CPB + lc;
  double v[2] = {0.0, 0.0};
  if (lc < CPB && c < ncv) {
    T yv = spmv_row<T, NB>(nbocv_i, nbocv_v, A, x, c, i);
    y[NB * (size_t)c + i] = yv;
    double o = (double)other[NB * (size_t)c + i];
    v[0] = (double)yv * o;
    v[1] = (double)yv * (double)yv;
  }
  block_partials<2>(v, partials, sh, kSum);
}

// s = res - alpha v; shat = LU^-1 s (with BJ = false only s)
template <typename T, int NB, bool BJ = true>
__global__ void __launch_bounds__(kBlock) k_s(int ncv, const T *__restrict__ LU, const T *__restrict__ res,
                                              const T *__restrict__ v, T *__restrict__ s, T *__restrict__ shat,
                                              const SolverScalars *S) {
  if (S->done || S->stop) return;
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const T alpha = (T)S->alpha;
  T t[NB];
  for (int k = 0; k < NB; k++) {
    size_t o = NB * (size_t)c + k;
    T sv = res[o] - alpha * v[o];
    s[o] = sv;
    t[k] = sv;
  }
  if (!BJ) return;
  lusolv_n<T, NB>(LU + NB * NB * (size_t)c, t);
  for (int k = 0; k < NB; k++) shat[NB * (size_t)c + k] = t[k];
}

// phi += alpha phat + omega shat
template <typename T>
__global__ void __launch_bounds__(kBlock) k_x(int n5, T *__restrict__ phi, const T *__restrict__ phat,
                                              const T *__restrict__ shat, const SolverScalars *S) {
  if (S->done || S->stop) return;
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n5) return;
  phi[k] += (T)S->alpha * phat[k] + (T)S->omega * shat[k];
}

// ---------------------------------------------------------------------------
// Block ILU(0) on subdomains: this build's alternative to the block-Jacobi
// preconditioner (GPU_PRECONDITIONER = ILU0). The reference's BCGSTAB has
// block Jacobi only; its coupled loop hands the system to PETSc GMRES with an
// additive-Schwarz preconditioner whose subdomain solve is ILU
// (PetscSolver.h:373-387). Here a subdomain is `sub` consecutive cells of the
// device order (Morton tiles, so spatially compact); couplings that leave it
// are dropped (no overlap), and one thread factors and solves each subdomain,
// sequentially over its rows, so subdomains run in parallel.
//
// Factors, in place of a copy F of A ([nnz][NB][NB], the blocks of A's
// pattern): for row c in order and each in-subdomain lower neighbour k in
// ascending order, F_ck <- F_ck D_k^-1, then F_cj -= F_ck F_kj for every
// upper entry (k, j) of row k whose column j is also in row c; D_c is the
// final diagonal block, kept inverted in Dinv. So L (unit lower) and U
// (D on the diagonal) satisfy (L U)_cj = A_cj on the subdomain's pattern.
// ---------------------------------------------------------------------------
template <typename T, int NB>
__device__ __forceinline__ void ilu_invert(const T *__restrict__ D, T *__restrict__ Dinv) {
  T a[NB][NB];
  for (int i = 0; i < NB; i++)
    for (int j = 0; j < NB; j++) a[i][j] = D[NB * i + j];
  ludeco_n<T, NB>(a);
  for (int j = 0; j < NB; j++) {
    T e[NB];
    for (int i = 0; i < NB; i++) e[i] = i == j ? (T)1 : (T)0;
    lusolv_n<T, NB>(&a[0][0], e);
    for (int i = 0; i < NB; i++) Dinv[NB * i + j] = e[i];
  }
}

template <typename T, int NB>
__global__ void __launch_bounds__(64) k_ilu_factor(int ncv, int sub, const int *__restrict__ nbocv_i,
                                                   const int *__restrict__ nbocv_v, const T *__restrict__ A,
                                                   T *__restrict__ F, T *__restrict__ Dinv) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = d * sub;
  if (c0 >= ncv) return;
  const int c1 = min(ncv, c0 + sub);
  constexpr int BB = NB * NB;
  for (int c = c0; c < c1; c++) {
    const int e0 = nbocv_i[c], e1 = nbocv_i[c + 1];
    for (int e = e0; e < e1; e++) {
      const int k = nbocv_v[e];
      if (k < c0 || k >= c1) continue;  // the coupling leaves the subdomain: dropped
      for (int q = 0; q < BB; q++) F[BB * (size_t)e + q] = A[BB * (size_t)e + q];
    }
    int last = -1;
    for (;;) {  // lower neighbours in ascending column order
      int ek = -1, k = c;
      for (int e = e0 + 1; e < e1; e++) {
        const int col = nbocv_v[e];
        if (col >= c0 && col > last && col < k) {
          k = col;
          ek = e;
        }
      }
      if (ek < 0) break;
      last = k;
      T *L = F + BB * (size_t)ek;
      const T *Di = Dinv + BB * (size_t)k;
      for (int i = 0; i < NB; i++) {  // L_ck = F_ck D_k^-1, one row at a time
        T row[NB];
        for (int j = 0; j < NB; j++) {
          T acc = (T)0;
          for (int m = 0; m < NB; m++) acc += L[NB * i + m] * Di[NB * m + j];
          row[j] = acc;
        }
        for (int j = 0; j < NB; j++) L[NB * i + j] = row[j];
      }
      for (int f = nbocv_i[k] + 1; f < nbocv_i[k + 1]; f++) {  // upper entries (k, j) of row k
        const int j = nbocv_v[f];
        if (j <= k || j >= c1) continue;
        int ej = -1;
        for (int e = e0; e < e1; e++)
          if (nbocv_v[e] == j) {
            ej = e;
            break;
          }
        if (ej < 0) continue;  // not in row c's pattern: ILU(0) drops the fill
        const T *U = F + BB * (size_t)f;
        T *X = F + BB * (size_t)ej;
        for (int i = 0; i < NB; i++)
          for (int jj = 0; jj < NB; jj++) {
            T acc = X[NB * i + jj];
            for (int m = 0; m < NB; m++) acc -= L[NB * i + m] * U[NB * m + jj];
            X[NB * i + jj] = acc;
          }
      }
    }
    ilu_invert<T, NB>(F + BB * (size_t)e0, Dinv + BB * (size_t)c);
  }
}

// z = (L U)^-1 r on each subdomain: forward y_c = r_c - sum_{k<c} L_ck y_k,
// backward z_c = D_c^-1 (y_c - sum_{j>c} U_cj z_j); y is kept in z
template <typename T, int NB>
__global__ void __launch_bounds__(64) k_ilu_solve(int ncv, int sub, const int *__restrict__ nbocv_i,
                                                  const int *__restrict__ nbocv_v, const T *__restrict__ F,
                                                  const T *__restrict__ Dinv, const T *__restrict__ r,
                                                  T *__restrict__ z, const SolverScalars *S) {
  if (S->done || S->stop) return;
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  const int c0 = d * sub;
  if (c0 >= ncv) return;
  const int c1 = min(ncv, c0 + sub);
  constexpr int BB = NB * NB;
  for (int c = c0; c < c1; c++) {
    T y[NB];
    for (int i = 0; i < NB; i++) y[i] = r[NB * (size_t)c + i];
    for (int e = nbocv_i[c] + 1; e < nbocv_i[c + 1]; e++) {
      const int k = nbocv_v[e];
      if (k < c0 || k >= c) continue;
      const T *L = F + BB * (size_t)e;
      const T *yk = z + NB * (size_t)k;
      T v[NB];
      for (int m = 0; m < NB; m++) v[m] = yk[m];
      for (int i = 0; i < NB; i++)
        for (int m = 0; m < NB; m++) y[i] -= L[NB * i + m] * v[m];
    }
    for (int i = 0; i < NB; i++) z[NB * (size_t)c + i] = y[i];
  }
  for (int c = c1 - 1; c >= c0; c--) {
    T t[NB];
    for (int i = 0; i < NB; i++) t[i] = z[NB * (size_t)c + i];
    for (int e = nbocv_i[c] + 1; e < nbocv_i[c + 1]; e++) {
      const int j = nbocv_v[e];
      if (j <= c || j >= c1) continue;
      const T *U = F + BB * (size_t)e;
      const T *zj = z + NB * (size_t)j;
      T v[NB];
      for (int m = 0; m < NB; m++) v[m] = zj[m];
      for (int i = 0; i < NB; i++)
        for (int m = 0; m < NB; m++) t[i] -= U[NB * i + m] * v[m];
    }
    const T *Di = Dinv + BB * (size_t)c;
    for (int i = 0; i < NB; i++) {
      T acc = (T)0;
      for (int m = 0; m < NB; m++) acc += Di[NB * i + m] * t[m];
      z[NB * (size_t)c + i] = acc;
    }
  }
}

// One block system and its work space for the kernels above: A[nnz][NB][NB]
// in nbocv order, LU[ncv][NB][NB], ten vectors [ncv][NB] (v[0] rhs, v[1] phi
// (the solution), v[2] res, v[3] res0, v[4] p, v[5] v, v[6] phat, v[7] s,
// v[8] shat, v[9] t), reduction partials and the solver scalars.
struct BcgSys {
  int ncv = 0, maxiter = 30;
  double zero_abs = 1e-8, zero_rel = 1e-2;
  const int *nbocv_i = nullptr, *nbocv_v = nullptr;
  const void *A = nullptr;
  void *LU = nullptr;
  void *v[10] = {};
  double *partials = nullptr;
  SolverScalars *S = nullptr;
  hipStream_t stream = nullptr;
  // block ILU(0) on subdomains of ilu_sub cells instead of block Jacobi for
  // phat and shat (the stop test keeps the block-Jacobi residual): F
  // [nnz][NB][NB] and Dinv [ncv][NB][NB]; 0 = block Jacobi
  int ilu_sub = 0;
  void *F = nullptr, *Dinv = nullptr;
  // the residual of each iteration by the recurrence s - omega t instead of
  // rhs - A phi (FP32 solves only)
  bool rec_resid = false;
};

// start of a solve (UgpWithCv.cpp:2660-2735): p = v = 0, block LU of the
// diagonal, the first residual and its stop test
template <typename T, int NB>
void bcg_prologue(const BcgSys &y) {
  const int ncv = y.ncv, n = NB * ncv;
  const T *A = (const T *)y.A;
  T *LU = (T *)y.LU;
  T *rhs_ = (T *)y.v[0], *phi = (T *)y.v[1], *res = (T *)y.v[2], *res0 = (T *)y.v[3], *p = (T *)y.v[4],
    *vv = (T *)y.v[5];
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(kBlock), 0, y.stream, (size_t)n, p, (T)0);
  hipLaunchKernelGGL(k_fill<T>, dim3(nblk(n)), dim3(kBlock), 0, y.stream, (size_t)n, vv, (T)0);
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, y.stream, y.S, y.maxiter, y.zero_abs, y.zero_rel);
  hipLaunchKernelGGL((k_lu<T, NB>), dim3(nblk(ncv)), dim3(kBlock), 0, y.stream, ncv, y.nbocv_i, A, LU);
  if (y.ilu_sub > 0) {
    const int nsub = (ncv + y.ilu_sub - 1) / y.ilu_sub;
    hipLaunchKernelGGL((k_ilu_factor<T, NB>), dim3((nsub + 63) / 64), dim3(64), 0, y.stream, ncv, y.ilu_sub,
                       y.nbocv_i, y.nbocv_v, A, (T *)y.F, (T *)y.Dinv);
  }
  const int gb = (ncv + cells_per_block<NB>() - 1) / cells_per_block<NB>();
  hipLaunchKernelGGL((k_resid<T, NB>), dim3(gb), dim3(kBlock), 0, y.stream, ncv, y.nbocv_i, y.nbocv_v, A, LU, phi,
                     rhs_, res, res0, 1, y.partials, y.S);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, y.stream, gb, (int)kFinResid0, y.partials, y.S);
}

// count BiCGStab iterations (van der Vorst, UgpWithCv.cpp:2725-2856); each
// returns at once once the device has set the convergence or stop flag
template <typename T, int NB>
void bcg_iterations(const BcgSys &y, int count) {
  const int ncv = y.ncv, n = NB * ncv;
  const T *A = (const T *)y.A, *LU = (const T *)y.LU;
  T *rhs_ = (T *)y.v[0], *phi = (T *)y.v[1], *res = (T *)y.v[2], *res0 = (T *)y.v[3], *p = (T *)y.v[4],
    *vv = (T *)y.v[5], *phat = (T *)y.v[6], *s = (T *)y.v[7], *shat = (T *)y.v[8], *t = (T *)y.v[9];
  const int gb = (ncv + cells_per_block<NB>() - 1) / cells_per_block<NB>();
  hipStream_t st = y.stream;
  const bool ilu = y.ilu_sub > 0;
  const int nsub = ilu ? (ncv + y.ilu_sub - 1) / y.ilu_sub : 0;
  auto ilu_solve = [&](const T *r, T *z) {
    hipLaunchKernelGGL((k_ilu_solve<T, NB>), dim3((nsub + 63) / 64), dim3(64), 0, st, ncv, y.ilu_sub, y.nbocv_i,
                       y.nbocv_v, (const T *)y.F, (const T *)y.Dinv, r, z, y.S);
  };
  for (int k = 0; k < count; k++) {
    if (ilu) {
      hipLaunchKernelGGL((k_p<T, NB, false>), dim3(nblk(ncv)), dim3(kBlock), 0, st, ncv, LU, res, vv, p, phat, y.S);
      ilu_solve(p, phat);
    } else {
      hipLaunchKernelGGL((k_p<T, NB>), dim3(nblk(ncv)), dim3(kBlock), 0, st, ncv, LU, res, vv, p, phat, y.S);
    }
    hipLaunchKernelGGL((k_spmv_dot<T, NB>), dim3(gb), dim3(kBlock), 0, st, ncv, y.nbocv_i, y.nbocv_v, A, phat, vv,
                       res0, y.partials, y.S);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, st, gb, (int)kFinAlpha, y.partials, y.S);
    if (ilu) {
      hipLaunchKernelGGL((k_s<T, NB, false>), dim3(nblk(ncv)), dim3(kBlock), 0, st, ncv, LU, res, vv, s, shat, y.S);
      ilu_solve(s, shat);
    } else {
      hipLaunchKernelGGL((k_s<T, NB>), dim3(nblk(ncv)), dim3(kBlock), 0, st, ncv, LU, res, vv, s, shat, y.S);
    }
    hipLaunchKernelGGL((k_spmv_dot<T, NB>), dim3(gb), dim3(kBlock), 0, st, ncv, y.nbocv_i, y.nbocv_v, A, shat, t, s,
                       y.partials, y.S);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, st, gb, (int)kFinOmega, y.partials, y.S);
    hipLaunchKernelGGL(k_x<T>, dim3(nblk(n)), dim3(kBlock), 0, st, n, phi, phat, shat, y.S);
    if (y.rec_resid)
      hipLaunchKernelGGL((k_resid_rec<T, NB>), dim3(gb), dim3(kBlock), 0, st, ncv, LU, s, t, res, res0, y.partials,
                         y.S);
    else
      hipLaunchKernelGGL((k_resid<T, NB>), dim3(gb), dim3(kBlock), 0, st, ncv, y.nbocv_i, y.nbocv_v, A, LU, phi,
                         rhs_, res, res0, 0, y.partials, y.S);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, st, gb, (int)kFinResid, y.partials, y.S);
  }
}

// a whole solve, host-polled: the flag is read between chunks (2, then 4
// iterations) into hs (any host memory: the copy is waited for); iterations
// past convergence are no-ops, so the chunking does not change the result
template <typename T, int NB>
void bcg_solve(const BcgSys &y, SolverScalars *hs, int &iters, int &status) {
  bcg_prologue<T, NB>(y);
  const int max_it = y.maxiter + 1;
  int launched = 0;
  while (launched < max_it) {
    int chunk = std::min(launched == 0 ? 2 : 4, max_it - launched);
    bcg_iterations<T, NB>(y, chunk);
    launched += chunk;
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(hs, y.S, sizeof(SolverScalars), hipMemcpyDeviceToHost, y.stream));
    HIPCHECK(hipStreamSynchronize(y.stream));
    if (hs->done) break;
  }
  iters = hs->iter;
  status = hs->done;
}

// ---------------------------------------------------------------------------
// Batched block-Jacobi BiCGStab for serving: the kernels above, each launched
// once for every active solve (a block -> solve map per grid shape, each
// solve keeping its own blocks, partials and solver scalars), so a batch
// costs nine launches per iteration whatever the number of solves. Every
// solve is cut into blocks exactly as its own solve would be and its
// partials are finished by one block in the same order, so its result is the
// same bits as the per-solve solver's.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_lu(const KrylovJob *__restrict__ jobs, const int *__restrict__ map1) {
  const KrylovJob &J = jobs[map1[blockIdx.x]];
  const int c = (blockIdx.x - J.b1) * blockDim.x + threadIdx.x;
  if (c >= J.ncv) return;
  T a[5][5];
  const T *src = (const T *)J.A + 25 * (size_t)J.nbocv_i[c];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) a[i][j] = src[5 * i + j];
  ludeco5<T>(a);
  T *d = (T *)J.LU + 25 * (size_t)c;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) d[5 * i + j] = a[i][j];
  for (int k = 0; k < 5; k++) {  // p = v = 0
    ((T *)J.v[4])[5 * (size_t)c + k] = (T)0;
    ((T *)J.v[5])[5 * (size_t)c + k] = (T)0;
  }
}

// REC: the residual by the recurrence s - omega t, as k_resid_rec (the FP32
// solves; omega is stored by kb_x)
template <typename T, bool REC = false>
__global__ void __launch_bounds__(kBlock) kb_resid(const KrylovJob *__restrict__ jobs, const int *__restrict__ map5,
                                                   int init, double *partials) {
  __shared__ double sh[32];
  __shared__ T cellv[kCellsPerBlock][5];
  const KrylovJob &J = jobs[map5[blockIdx.x]];
  const SolverScalars *S = (const SolverScalars *)J.S;
  if (!init && (S->done || S->stop)) return;
  const T *A = (const T *)J.A, *LU = (const T *)J.LU, *phi = (const T *)J.v[1], *rhs = (const T *)J.v[0];
  T *res = (T *)J.v[2], *res0 = (T *)J.v[3];
  const int lc = threadIdx.x / 5, i = threadIdx.x % 5;
  const int c = (blockIdx.x - J.b5) * kCellsPerBlock + lc;
  const bool active = (lc < kCellsPerBlock) && (c < J.ncv);
  T r = (T)0.0;
  if (active && REC && !init) {
    const size_t o = 5 * (size_t)c + i;
    r = ((const T *)J.v[7])[o] - (T)S->omega * ((const T *)J.v[9])[o];
    res[o] = r;
    cellv[lc][i] = r;
  } else if (active) {
    T ax = spmv_row<T, 5>(J.nbocv_i, J.nbocv_v, A, phi, c, i);
    r = rhs[5 * (size_t)c + i] - ax;
    res[5 * (size_t)c + i] = r;
    if (init) res0[5 * (size_t)c + i] = r;
    cellv[lc][i] = r;
  }
  __syncthreads();
  double vs[1] = {0.0}, vm[1] = {0.0};
  if (active) {
    T r0v = init ? r : res0[5 * (size_t)c + i];
    vs[0] = (double)r * (double)r0v;
    T tmp[5] = {cellv[lc][0], cellv[lc][1], cellv[lc][2], cellv[lc][3], cellv[lc][4]};
    lusolv5<T>(LU + 25 * (size_t)c, tmp);
    vm[0] = fabs((double)tmp[i]);
  }
  block_reduce<1>(vs, sh, kSum);
  __syncthreads();
  block_reduce<1>(vm, sh + 16, kMax);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = vs[0];
    partials[gridDim.x + blockIdx.x] = vm[0];
  }
}

// which = 0: p = res - beta (p - omega v), phat = LU^-1 p; 1: s = res - alpha v, shat = LU^-1 s
// (which = 1 also finishes alpha = rho / (v.res0) from kb_spmv_dot's
// partials, as k_fin kFinAlpha does, in every block; the solve's first block
// stores it)
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_ps(const KrylovJob *__restrict__ jobs, const int *__restrict__ map1,
                                                int which, const double *__restrict__ partials, int grid5) {
  __shared__ double sh[16];
  const KrylovJob &J = jobs[map1[blockIdx.x]];
  SolverScalars *S = (SolverScalars *)J.S;
  if (S->done || S->stop) return;
  double alpha = 0.0;
  if (which == 1) {  // UgpWithCv.cpp:2775-2782
    double gamma = fin_slot_v(partials + J.b5, J.nb5, kSum, sh);
    if (fabs(gamma) < 1.0E-20) gamma = 1.0E-20;
    alpha = S->rho / gamma;
    if ((int)blockIdx.x == J.b1 && threadIdx.x == 0) {
      S->gamma = gamma;
      S->alpha = alpha;
    }
  }
  const int c = (blockIdx.x - J.b1) * blockDim.x + threadIdx.x;
  if (c >= J.ncv) return;
  const T *res = (const T *)J.v[2], *vv = (const T *)J.v[5], *LU = (const T *)J.LU;
  T *out = (T *)J.v[which == 0 ? 4 : 7], *hat = (T *)J.v[which == 0 ? 6 : 8];
  T t[5];
  if (which == 0) {
    const T beta = (T)S->beta, omega = (T)S->omega;
    for (int k = 0; k < 5; k++) {
      size_t o = 5 * (size_t)c + k;
      T pv = res[o] - beta * (out[o] - omega * vv[o]);
      out[o] = pv;
      t[k] = pv;
    }
  } else {
    const T alpha_t = (T)alpha;
    for (int k = 0; k < 5; k++) {
      size_t o = 5 * (size_t)c + k;
      T sv = res[o] - alpha_t * vv[o];
      out[o] = sv;
      t[k] = sv;
    }
  }
  lusolv5<T>(LU + 25 * (size_t)c, t);
  for (int k = 0; k < 5; k++) hat[5 * (size_t)c + k] = t[k];
}

// which = 0: v = A phat with v.res0; 1: t = A shat with s.t, t.t
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_spmv_dot(const KrylovJob *__restrict__ jobs,
                                                      const int *__restrict__ map5, int which, double *partials) {
  __shared__ double sh[32];
  const KrylovJob &J = jobs[map5[blockIdx.x]];
  const SolverScalars *S = (const SolverScalars *)J.S;
  if (S->done || S->stop) return;
  const T *x = (const T *)J.v[which == 0 ? 6 : 8], *other = (const T *)J.v[which == 0 ? 3 : 7];
  T *y = (T *)J.v[which == 0 ? 5 : 9];
  const int lc = threadIdx.x / 5, i = threadIdx.x % 5;
  const int c = (blockIdx.x - J.b5) * kCellsPerBlock + lc;
  double v[2] = {0.0, 0.0};
  if (lc < kCellsPerBlock && c < J.ncv) {
    T yv = spmv_row<T, 5>(J.nbocv_i, J.nbocv_v, (const T *)J.A, x, c, i);
    y[5 * (size_t)c + i] = yv;
    double o = (double)other[5 * (size_t)c + i];
    v[0] = (double)yv * o;
    v[1] = (double)yv * (double)yv;
  }
  block_partials<2>(v, partials, sh, kSum);
}

// phi += alpha phat + omega shat, with omega = (s.t) / (t.t + 1e-20) finished
// from kb_spmv_dot's partials in every block as k_fin kFinOmega does (the
// solve's first block stores it)
template <typename T>
__global__ void __launch_bounds__(kBlock) kb_x(const KrylovJob *__restrict__ jobs, const int *__restrict__ mapx,
                                               const double *__restrict__ partials, int grid5) {
  __shared__ double sh[16];
  const KrylovJob &J = jobs[mapx[blockIdx.x]];
  SolverScalars *S = (SolverScalars *)J.S;
  if (S->done || S->stop) return;
  const double st = fin_slot_v(partials + J.b5, J.nb5, kSum, sh);  // UgpWithCv.cpp:2808-2810
  const double tt = fin_slot_v(partials + J.b5 + grid5, J.nb5, kSum, sh);
  const double omega = st / (tt + 1.0E-20);
  if ((int)blockIdx.x == J.bx && threadIdx.x == 0) S->omega = omega;
  const int k = (blockIdx.x - J.bx) * blockDim.x + threadIdx.x;
  if (k >= 5 * J.ncv) return;
  T *phi = (T *)J.v[1];
  phi[k] += (T)S->alpha * ((const T *)J.v[6])[k] + (T)omega * ((const T *)J.v[8])[k];
}

// one block per solve: finish its partials (at its first block, stride = the whole grid)
__global__ void __launch_bounds__(kFinBlock) kb_fin(const KrylovJob *__restrict__ jobs, int op,
                                                    const double *__restrict__ partials, int grid5) {
  __shared__ double sh[16];
  const KrylovJob &J = jobs[blockIdx.x];
  fin_body(J.nb5, op, partials + J.b5, (size_t)grid5, (SolverScalars *)J.S, sh);
}

// *flag = 1 when every solve of the batch has converged or spent its launches
__global__ void kb_all_done(const KrylovJob *__restrict__ jobs, int njobs, int *flag) {
  __shared__ int open;
  if (threadIdx.x == 0) open = 0;
  __syncthreads();
  for (int j = threadIdx.x; j < njobs; j += blockDim.x) {
    const SolverScalars *S = (const SolverScalars *)jobs[j].S;
    if (!(S->done || S->stop)) atomicOr(&open, 1);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    *flag = open ? 0 : 1;
    __threadfence_system();
  }
}

// Q += dQ (JoeWithModels.cpp:1028-1035)
template <typename T>
__device__ __forceinline__ void d_update(const Blk &B, int ncv, double *__restrict__ q, const T *__restrict__ dq) {
  int c = B.id * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  for (int k = 0; k < 5; k++) q[(size_t)k * ncv + c] += (double)dq[5 * (size_t)c + k];
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_update(int ncv, double *__restrict__ q, const T *__restrict__ dq) {
  d_update<T>(hw_blk(), ncv, q, dq);
}

// the coupled loop's update with the scalar (JoeWithModels.cpp:4497-4516):
// Q += dQ[0..4], then nuSA = (nuSA (rho - dQ0) + dQ5) / rho with the updated
// rho, clipped to the model's bounds
template <typename T>
__global__ void __launch_bounds__(kBlock) k_update6(int ncv, double *__restrict__ q, const T *__restrict__ dq,
                                                    double *__restrict__ phi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const T *d = dq + 6 * (size_t)c;
  for (int k = 0; k < 5; k++) q[(size_t)k * ncv + c] += (double)d[k];
  const double rho = q[c];
  const double v = (phi[c] * (rho - (double)d[0]) + (double)d[5]) / rho;
  phi[c] = Dev::mn(Dev::mx(v, kSaLower), kSaUpper);
}

// sourceHook / sourceHookCoupled terms computed by the host, in device
// order: src[c] holds NB residual terms then the NB x NB terms of the
// diagonal block (JoeWithModels.h:111, 134); with NB = 6 the sixth residual
// term is the scalar's (rhs_s).
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_add_source(int ncv, const int *__restrict__ nbocv_i,
                                                       const double *__restrict__ src, int with_a,
                                                       double *__restrict__ rhs, double *__restrict__ rhs_s,
                                                       T *__restrict__ A) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double *s = src + (NB + NB * NB) * (size_t)c;
  for (int k = 0; k < 5; k++) rhs[(size_t)k * ncv + c] += s[k];
  if (NB == 6) rhs_s[c] += s[5];
  if (with_a) {
    T *d = A + NB * NB * (size_t)nbocv_i[c];
    for (int k = 0; k < NB * NB; k++) d[k] = (T)((double)d[k] + s[NB + k]);
  }
}

// backward-Euler time term and relaxed right-hand side (JoeWithModels.cpp
// :1011-1024, 4455-4473) for a Jacobian assembled without them; with NB = 6
// the scalar's row too
template <typename T, int NB>
__global__ void __launch_bounds__(kBlock) k_time_term(int ncv, const int *__restrict__ nbocv_i,
                                                      const double *__restrict__ vol, const double *__restrict__ dt,
                                                      double relax, const double *__restrict__ rhs,
                                                      const double *__restrict__ rhs_s, T *__restrict__ A,
                                                      T *__restrict__ rhs_ls, T *__restrict__ phi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double tmp = vol[c] / (dt[c]);
  T *d = A + NB * NB * (size_t)nbocv_i[c];
  for (int i = 0; i < NB; i++) d[(NB + 1) * i] = (T)((double)d[(NB + 1) * i] + tmp);
  for (int k = 0; k < NB; k++) {
    rhs_ls[NB * (size_t)c + k] = (T)(relax * (k < 5 ? rhs[(size_t)k * ncv + c] : rhs_s[c]));
    phi[NB * (size_t)c + k] = (T)0.0;
  }
}

// L1 residual sums (JoeWithModels.cpp:1119-1130) and the sums of squares of
// the L2 norms beside them, finished by k_fin kFinL1
__device__ __forceinline__ void d_l1(const Blk &B, int ncv, const double *__restrict__ rhs, double *partials) {
  __shared__ double sh[16 * 10];
  int c = B.id * blockDim.x + threadIdx.x;
  double v[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (c < ncv)
    for (int k = 0; k < 5; k++) {
      const double r = rhs[(size_t)k * ncv + c];
      v[k] = fabs(r);
      v[5 + k] = r * r;
    }
  block_partials<10>(v, partials, sh, kSum, B);
}

__global__ void __launch_bounds__(kBlock) k_l1(int ncv, const double *__restrict__ rhs, double *partials) {
  d_l1(hw_blk(), ncv, rhs, partials);
}

// ---------------------------------------------------------------------------
// Batched step for serving: the kernels of a step before and after its
// linear solve, each launched once for several solves of one kernel
// configuration (a block -> solve map per launch shape). Each solve runs the
// same bodies on its own block partition (Blk), partials and scalars as its
// own launches would, so its results are the same bits.
// ---------------------------------------------------------------------------
enum StepShape { kShCell = 0, kShGrad, kShFlux, kShBc, kShFill, kShDtFill, kShapes };

struct StepJob {
  DMesh dm;
  DParams dp;
  double *prim, *bfa, *grad, *gradv, *q, *rhs, *dt, *partials;
  const double *hook;
  int *counters;
  void *A, *rhs_ls, *phi;  // in the solve's precision
  SolverScalars *scal;
  StepMail *mail;
  DtArgs dta;
  size_t nA;
  double gamma, relax, cfl_factor, cfl_max, ls_abs, ls_rel;
  int ncv, cfl_after, cfl_interval, ls_maxiter;
  int b[kShapes], n[kShapes];  // first block and block count in each launch shape
};

__device__ __forceinline__ Blk job_blk(const StepJob &J, int shape) {
  return Blk{(int)blockIdx.x - J.b[shape], J.n[shape]};
}

// one block per solve
__global__ void ks_ramp(const StepJob *__restrict__ jobs) {
  const StepJob &J = jobs[blockIdx.x];
  d_ramp(Blk{0, 1}, J.scal, J.cfl_after, J.cfl_interval, J.cfl_factor, J.cfl_max);
}

__global__ void ks_solver_init(const StepJob *__restrict__ jobs) {
  const StepJob &J = jobs[blockIdx.x];
  d_solver_init(Blk{0, 1}, J.scal, J.ls_maxiter, J.ls_abs, J.ls_rel);
}

__global__ void ks_publish(const StepJob *__restrict__ jobs) {
  const StepJob &J = jobs[blockIdx.x];
  d_publish(Blk{0, 1}, J.scal, J.mail);
}

// k_fin of each solve's partials of one launch shape
__global__ void __launch_bounds__(kFinBlock) ks_fin(const StepJob *__restrict__ jobs, int op, int shape) {
  __shared__ double sh[16];
  const StepJob &J = jobs[blockIdx.x];
  fin_body(J.n[shape], op, J.partials, (size_t)J.n[shape], J.scal, sh);
}

// block -> solve map of the launch shape at map
template <bool LS, bool VISC, bool ENTRY>
__global__ void __launch_bounds__(kBlock) ks_grad(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  if constexpr (ENTRY)
    d_grad_e<LS>(job_blk(J, kShGrad), J.dm, J.dp, J.prim, J.bfa, J.grad);
  else
    d_grad<LS, VISC>(job_blk(J, kShGrad), J.dm, J.dp, J.prim, J.bfa, J.grad, J.gradv);
}

template <int SCHEME, bool SND, bool VISC>
__global__ void __launch_bounds__(kBlock) ks_flux(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_flux_e<SCHEME, SND, VISC>(job_blk(J, kShFlux), J.dm, J.dp, J.prim, J.grad, J.bfa, J.q, J.rhs, J.counters + 1,
                              J.dta, J.dt, J.partials, J.gradv);
}

__global__ void __launch_bounds__(kBlock) ks_dt_fill(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_dt_fill(job_blk(J, kShDtFill), J.ncv, J.dt, J.scal);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) ks_fill(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_fill<T>(job_blk(J, kShFill), J.nA, (T *)J.A, (T)0);
}

template <typename T, int SCHEME, bool VISC>
__global__ void __launch_bounds__(kBlock) ks_jac(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_jac<T, SCHEME, VISC>(job_blk(J, kShCell), J.dm, J.dp, J.prim, J.bfa, (T *)J.A, 1, J.dt, J.rhs, J.relax,
                         (T *)J.rhs_ls, (T *)J.phi, J.gradv);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) ks_update(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_update<T>(job_blk(J, kShCell), J.ncv, J.q, (const T *)J.phi);
}

__global__ void __launch_bounds__(kBlock) ks_state(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_state(job_blk(J, kShCell), J.ncv, J.q, J.prim, J.gamma, J.counters);
}

__global__ void __launch_bounds__(kBlock) ks_bc(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_bc(job_blk(J, kShBc), J.dm, J.dp, J.prim, J.hook, J.bfa);
}

__global__ void __launch_bounds__(kBlock) ks_l1(const StepJob *__restrict__ jobs, const int *__restrict__ map) {
  const StepJob &J = jobs[map[blockIdx.x]];
  d_l1(job_blk(J, kShCell), J.ncv, J.rhs, J.partials);
}

__global__ void k_to_double(int n, const float *__restrict__ s, double *__restrict__ d) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) d[k] = (double)s[k];
}

// ---------------------------------------------------------------------------
// The segregated nuSA solve (JoeWithModels.cpp:1040-1091) in FP64: right-hand
// side and V/dt, Jacobi-preconditioned BiCGStab (solveCvScalarBcgstab,
// UgpWithCv.cpp:2298-2479) on the diagonal-first rows of AS[.][5], update and clip.
// ---------------------------------------------------------------------------
// b = relax rhs - sum_noc AScal[0..4][noc] dq[nbocv_v[noc]]; AScal[5][diag] += V/dt
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sc_prep(int ncv, const int *__restrict__ nbocv_i,
                                                    const int *__restrict__ nbocv_v, const double *__restrict__ vol,
                                                    const double *__restrict__ dt, double relax,
                                                    double *__restrict__ rhs_s, double *__restrict__ AS,
                                                    const T *__restrict__ dq, double *__restrict__ b,
                                                    double *__restrict__ x) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  double r = rhs_s[c] * relax;
  const int noc_f = nbocv_i[c], noc_l = nbocv_i[c + 1] - 1;
  const double tmp = vol[c] / (dt[c]);
  AS[6 * (size_t)noc_f + 5] += tmp;
  for (int noc = noc_f; noc <= noc_l; noc++) {
    const double *a = AS + 6 * (size_t)noc;
    const T *d = dq + 5 * (size_t)nbocv_v[noc];
    r = r - a[0] * (double)d[0] - a[1] * (double)d[1] - a[2] * (double)d[2] - a[3] * (double)d[3] -
        a[4] * (double)d[4];
  }
  rhs_s[c] = r;
  b[c] = r;
  x[c] = 0.0;
}

__device__ __forceinline__ double sc_row(const int *__restrict__ nbocv_i, const int *__restrict__ nbocv_v,
                                         const double *__restrict__ AS, const double *__restrict__ x, int c) {
  const int noc_f = nbocv_i[c], noc_l = nbocv_i[c + 1] - 1;
  double y = AS[6 * (size_t)noc_f + 5] * x[c];
  for (int noc = noc_f + 1; noc <= noc_l; noc++) y += AS[6 * (size_t)noc + 5] * x[nbocv_v[noc]];
  return y;
}

// res = b - A x (init: res0 = res); partials res.res0 and max |res / a_diag|
__global__ void __launch_bounds__(kBlock) k_sc_resid(int ncv, const int *__restrict__ nbocv_i,
                                                     const int *__restrict__ nbocv_v, const double *__restrict__ AS,
                                                     const double *__restrict__ x, const double *__restrict__ b,
                                                     double *__restrict__ res, double *__restrict__ res0, int init,
                                                     double *partials, const SolverScalars *S) {
  __shared__ double sh[32];
  if (!init && (S->done || S->stop)) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double vs[1] = {0.0}, vm[1] = {0.0};
  if (c < ncv) {
    double r = b[c] - sc_row(nbocv_i, nbocv_v, AS, x, c);
    res[c] = r;
    if (init) res0[c] = r;
    vs[0] = r * (init ? r : res0[c]);
    vm[0] = fabs(r / AS[6 * (size_t)nbocv_i[c] + 5]);
  }
  block_reduce<1>(vs, sh, kSum);
  __syncthreads();
  block_reduce<1>(vm, sh + 16, kMax);
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = vs[0];
    partials[gridDim.x + blockIdx.x] = vm[0];
  }
}

// p = res - beta (p - omega v); phat = p / a_diag
__global__ void __launch_bounds__(kBlock) k_sc_p(int ncv, const int *__restrict__ nbocv_i,
                                                 const double *__restrict__ AS, const double *__restrict__ res,
                                                 const double *__restrict__ v, double *__restrict__ p,
                                                 double *__restrict__ phat, const SolverScalars *S) {
  if (S->done || S->stop) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double pv = res[c] - S->beta * (p[c] - S->omega * v[c]);
  p[c] = pv;
  phat[c] = pv / AS[6 * (size_t)nbocv_i[c] + 5];
}

// y = A x with partial dots (y.other, y.y)
__global__ void __launch_bounds__(kBlock) k_sc_spmv_dot(int ncv, const int *__restrict__ nbocv_i,
                                                        const int *__restrict__ nbocv_v, const double *__restrict__ AS,
                                                        const double *__restrict__ x, double *__restrict__ y,
                                                        const double *__restrict__ other, double *partials,
                                                        const SolverScalars *S) {
  __shared__ double sh[32];
  if (S->done || S->stop) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double v[2] = {0.0, 0.0};
  if (c < ncv) {
    const double yv = sc_row(nbocv_i, nbocv_v, AS, x, c);
    y[c] = yv;
    v[0] = yv * other[c];
    v[1] = yv * yv;
  }
  block_partials<2>(v, partials, sh, kSum);
}

// s = res - alpha v; shat = s / a_diag
__global__ void __launch_bounds__(kBlock) k_sc_s(int ncv, const int *__restrict__ nbocv_i,
                                                 const double *__restrict__ AS, const double *__restrict__ res,
                                                 const double *__restrict__ v, double *__restrict__ sv,
                                                 double *__restrict__ shat, const SolverScalars *S) {
  if (S->done || S->stop) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double s = res[c] - S->alpha * v[c];
  sv[c] = s;
  shat[c] = s / AS[6 * (size_t)nbocv_i[c] + 5];
}

__global__ void __launch_bounds__(kBlock) k_sc_x(int ncv, double *__restrict__ x, const double *__restrict__ phat,
                                                 const double *__restrict__ shat, const SolverScalars *S) {
  if (S->done || S->stop) return;
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  x[c] += S->alpha * phat[c] + S->omega * shat[c];
}

// phi = min(max((phi (rho - dq_rho) + dphi) / rho, lower), upper) with the updated rho
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sc_update(int ncv, const double *__restrict__ q,
                                                      const T *__restrict__ dq, const double *__restrict__ x,
                                                      double *__restrict__ phi) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double rho = q[c];
  const double v = (phi[c] * (rho - (double)dq[5 * (size_t)c]) + x[c]) / rho;
  phi[c] = Dev::mn(Dev::mx(v, kSaLower), kSaUpper);
}

// L1 of the scalar residual divided by the mean-flow relaxation (JoeWithModels.cpp:1132-1134)
__global__ void __launch_bounds__(kBlock) k_l1s(int ncv, const double *__restrict__ rhs_s, double relax,
                                                double *partials) {
  __shared__ double sh[16];
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  double v[1] = {c < ncv ? fabs(rhs_s[c] / relax) : 0.0};
  block_partials<1>(v, partials, sh, kSum);
}

// calcWallDistance (UgpWithCvCompFlow.h:1395-1535): nearest wall-face centroid
// (first in zone order on ties), then the normal distance where the cell
// centre projects within the face's "spectral radius". wx/wn: [nw][3] face
// centroids and area-weighted normals of the wall zones, in zone order.
__global__ void __launch_bounds__(kBlock) k_wall_dist(int ncv, const double *__restrict__ prim, int nw,
                                                      const double *__restrict__ wx, const double *__restrict__ wn,
                                                      double *__restrict__ aux, int stride, int slot) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncv) return;
  const double x[3] = {prim[8 * (size_t)c + 5], prim[8 * (size_t)c + 6], prim[8 * (size_t)c + 7]};
  double minDist = 1.0e20;
  int save = 0;
  for (int k = 0; k < nw; k++) {
    const double dx = x[0] - wx[3 * k], dy = x[1] - wx[3 * k + 1], dz = x[2] - wx[3 * k + 2];
    const double dist = sqrt(dx * dx + dy * dy + dz * dz);
    if (dist < minDist) {
      save = k;
      minDist = dist;
    }
  }
  const double *fn = wn + 3 * save;
  const double nmag = sqrt(dot3(fn, fn));
  const double n[3] = {fn[0] / nmag, fn[1] / nmag, fn[2] / nmag};
  const double s_half[3] = {wx[3 * save] - x[0], wx[3 * save + 1] - x[1], wx[3 * save + 2] - x[2]};
  const double sMag2 = dot3(s_half, s_half);
  const double dNorm2 = dot3(s_half, n) * dot3(s_half, n);
  const double dTang2 = sMag2 - dNorm2;
  const double specRadius2 = 0.5 * nmag;
  aux[stride * (size_t)c + slot] = dTang2 > specRadius2 ? minDist : fabs(dot3(s_half, n));
}


template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    free();
    n = count;
    if (count) HIPCHECK(hipMalloc((void **)&p, sizeof(T) * count));
  }
  void upload(const T *h, size_t count) {
    alloc(count);
    if (count) HIPCHECK(hipMemcpy(p, h, sizeof(T) * count, hipMemcpyHostToDevice));
  }
  void free() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { free(); }
};

template <typename T>
void ensure(DBuf<T> &b, size_t n) {
  if (b.n != n) b.alloc(n);
}

}  // namespace

// ---------------------------------------------------------------------------
// Joe.in -> parameters (UgpWithCvCompFlow.h:57-502, JoeWithModels.cpp:881-929,
// boundary keywords JoeWithModels.cpp:3857-4300)
// ---------------------------------------------------------------------------
// The boundary-profile file of PROFILES zones as the reference reads it
// ("n=<rows>\td=<values>", then rows of y, rho, u, v, p and one value per
// transported scalar; JoeWithModels.cpp:262-290) and its values at each
// PROFILES face: linear in the face's y, clamped at both ends, v dropped
// (:296-326). Returns [nfa_b][5] = rho, u, v, w, p (zero on other faces);
// with scal, the first scalar's column there too, interpolated alike (:331-357).
std::vector<double> profile_faces(const std::string &file, const Mesh &m, const std::vector<ZoneBc> &zbc, int nscal,
                                  std::vector<double> *scal = nullptr) {
  FILE *fp = std::fopen(file.c_str(), "rt");
  if (!fp) throw std::invalid_argument("could not open the boundary-profile file " + file);
  int nn = 0, nval = 0;
  const bool head = std::fscanf(fp, "n=%d\td=%d", &nn, &nval) == 2;
  std::vector<double> pr(head && nn > 0 && nval > 0 ? (size_t)nn * nval : 0);
  size_t got = 0;
  for (; got < pr.size(); got++)
    if (std::fscanf(fp, "%lf", &pr[got]) != 1) break;
  std::fclose(fp);
  if (!head || nn < 2 || got != pr.size())
    throw std::invalid_argument("boundary-profile file " + file + ": needs n=<rows> d=<values> and the rows");
  if (nval - 5 != nscal)
    throw std::invalid_argument("boundary-profile file " + file + " does not have the required number of values");
  std::vector<double> h(5 * (size_t)std::max(1, m.nfa_b), 0.0);
  if (scal) scal->assign((size_t)std::max(1, m.nfa_b), 0.0);
  for (size_t iz = 0; iz < m.zones.size(); iz++) {
    if (m.zones[iz].kind != kZoneBoundary || zbc[iz].bc != kBcProfiles) continue;
    for (int ifa = m.zones[iz].ifa_f; ifa <= m.zones[iz].ifa_l; ifa++) {
      const double y = m.x_fa[3 * (size_t)ifa + 1];
      int pos = 1;
      while ((pr[(size_t)pos * nval] < y) && (pos < nn - 1)) pos++;
      const double *r1 = &pr[(size_t)(pos - 1) * nval], *r2 = &pr[(size_t)pos * nval];
      double f;
      if (y > r2[0]) f = 1.0;
      else if (y < r1[0]) f = 0.0;
      else f = (y - r1[0]) / (r2[0] - r1[0]);
      double *o = &h[5 * (size_t)ifa];
      o[0] = r1[1] + f * (r2[1] - r1[1]);
      o[1] = r1[2] + f * (r2[2] - r1[2]);
      o[2] = 0.0;
      o[3] = 0.0;
      o[4] = r1[4] + f * (r2[4] - r1[4]);
      if (scal && nscal > 0) (*scal)[ifa] = r1[5] + f * (r2[5] - r1[5]);
    }
  }
  return h;
}

void parse_params(const std::string &joe_in, const Mesh &m, SolverParams &p, std::vector<ZoneBc> &zbc) {
  ParamMap pm(joe_in);
  p.gamma = pm.dbl("GAMMA", 1.4);
  if (pm.has("ALTITUDE")) throw Unsupported("ALTITUDE is not supported");
  p.T_ref = pm.dbl("T_REF");
  p.p_ref = pm.dbl("P_REF");
  p.rho_ref = pm.dbl("RHO_REF");
  p.R_gas = p.p_ref / p.rho_ref / p.T_ref;
  // viscous terms (UgpWithCvCompFlow.h:155-186)
  p.mu_ref = pm.dbl("MU_REF", 0.0);
  if (p.mu_ref > 0.0) {
    const std::string vm = pm.str("MU_MODE", "POWERLAW");
    if (vm == "POWERLAW") p.visc_mode = 0;
    else if (vm == "SUTHERLAND") p.visc_mode = 1;
    else throw std::invalid_argument("MU_MODE = " + vm + " (POWERLAW, SUTHERLAND)");
    p.mu_power_law = pm.dbl("MU_POWER_LAW", 0.76);
    p.sl_tref = pm.dbl("SL_TREF", 273.15);
    p.sl_sref = pm.dbl("SL_SREF", 110.4);
    p.pr = pm.dbl("PR", 0.72);
    p.pr_turb = pm.dbl("PRTURB", 0.9);
    p.visc_limiter = pm.str("VISC_LIMITER", "YES") == "YES";
    if (pm.str("SIMPLEX_VISCOUS", "NO") != "NO") throw Unsupported("SIMPLEX_VISCOUS = YES is not supported");
  }
  // device cell numbering (this build's own option; results do not depend on it)
  {
    const std::string o = pm.str("GPU_CELL_ORDER", "MORTON");
    if (o == "NATIVE") p.cell_order = kOrderNative;
    else if (o == "GROUPED") p.cell_order = kOrderGrouped;
    else if (o == "MORTON") p.cell_order = kOrderMorton;
    else if (o == "MORTON_GROUPED") p.cell_order = kOrderMortonGrouped;
    else throw std::invalid_argument("GPU_CELL_ORDER = " + o + " (NATIVE, GROUPED, MORTON, MORTON_GROUPED)");
  }
  // residual kernel forms (this build's own options; results do not depend on them)
  auto kform = [&](const char *key, const char *def) {
    const std::string o = pm.str(key, def);
    if (o == "ENTRY") return (int)kResidEntry;
    if (o == "CELL") return (int)kResidCell;
    if (o == "FACE" && std::string(key) == "GPU_FLUX_KERNEL") return (int)kResidFace;
    throw std::invalid_argument(std::string(key) + " = " + o + " (ENTRY, CELL; FACE for the flux)");
  };
  p.grad_kernel = kform("GPU_GRAD_KERNEL", "CELL");
  p.flux_kernel = kform("GPU_FLUX_KERNEL", "ENTRY");
  p.jac_kernel = kform("GPU_JAC_KERNEL", "ENTRY");
  p.scal_kernel = kform("GPU_SCAL_KERNEL", "ENTRY");
  {  // the per-cell gradient's thread order (this build's option; results do not depend on it)
    const std::string o = pm.str("GPU_GRAD_THREAD_ORDER", "NATIVE");
    if (o != "SORTED" && o != "NATIVE")
      throw std::invalid_argument("GPU_GRAD_THREAD_ORDER = " + o + " (SORTED, NATIVE)");
    p.grad_thread_sorted = o == "SORTED";
  }
  {  // the FP32 BiCGStab's residual: RECURRENCE (s - omega t) or TRUE (rhs - A phi, as the reference);
     // FP64 solves always take the reference's
    const std::string o = pm.str("GPU_KRYLOV_RESIDUAL", "RECURRENCE");
    if (o != "TRUE" && o != "RECURRENCE") throw std::invalid_argument("GPU_KRYLOV_RESIDUAL = " + o + " (TRUE, RECURRENCE)");
    p.krylov_recurrence = o == "RECURRENCE";
  }
  {  // this build's preconditioner option (the reference's BCGSTAB is block Jacobi)
    const std::string pc = pm.str("GPU_PRECONDITIONER", "BLOCK_JACOBI");
    if (pc == "ILU0") {
      p.ilu_sub = pm.integer("GPU_ILU_SUBDOMAIN", 64);
      if (p.ilu_sub < 1 || p.ilu_sub > 4096) throw std::invalid_argument("GPU_ILU_SUBDOMAIN must be 1..4096");
    } else if (pc != "BLOCK_JACOBI") {
      throw std::invalid_argument("GPU_PRECONDITIONER = " + pc + " (BLOCK_JACOBI, ILU0)");
    }
  }
  p.gravity[0] = pm.dbl("GX", 0.0);
  p.gravity[1] = pm.dbl("GY", 0.0);
  p.gravity[2] = pm.dbl("GZ", 0.0);
  p.second_order = pm.has("SPATIAL_SECOND_ORDER");
  // turbulence model: the reference chooses its class by run number (joe.cpp
  // :963-976); here TURB_MODEL = SA selects the RansTurbSA mix-in
  {
    const std::string tm = pm.str("TURB_MODEL", "NONE");
    if (tm == "SA") {
      p.sa = true;
      if (!(p.mu_ref > 0.0)) throw std::invalid_argument("TURB_MODEL = SA needs MU_REF > 0");
      const auto *t = pm.get("INITIAL_CONDITION_TURB");
      if (!t || t->size() < 2) throw std::invalid_argument("INITIAL_CONDITION_TURB missing (JoeWithModels.cpp:223-232)");
      p.sa_init = ParamMap::to_d((*t)[1], "INITIAL_CONDITION_TURB");
      p.relax_scal = pm.dbl("UNDER_RELAXATION_SCALARS", 0.2);
      if (pm.has("SCALAR_FIRST_ORDER")) throw Unsupported("SCALAR_FIRST_ORDER is not supported");
    } else if (tm == "SST") {  // RansTurbKOmSST (TurbModel_KOMSST.h), kine and omega segregated
      p.sst = true;
      if (!(p.mu_ref > 0.0)) throw std::invalid_argument("TURB_MODEL = SST needs MU_REF > 0");
      const auto *t = pm.get("INITIAL_CONDITION_TURB");
      if (!t || t->size() < 3)
        throw std::invalid_argument("INITIAL_CONDITION_TURB needs kine and omega (JoeWithModels.cpp:223-232)");
      p.sst_init[0] = ParamMap::to_d((*t)[1], "INITIAL_CONDITION_TURB");
      p.sst_init[1] = ParamMap::to_d((*t)[2], "INITIAL_CONDITION_TURB");
      p.relax_scal = pm.dbl("UNDER_RELAXATION_SCALARS", 0.2);
      p.sst_limit_pk = pm.integer("SST_LIMIT_PK", 1);
      p.sst_limiter_shear = pm.integer("SST_LIMITER_SHEARSTRESS", 0);
      if (pm.has("SCALAR_FIRST_ORDER")) throw Unsupported("SCALAR_FIRST_ORDER is not supported");
      if (pm.str("SCALAR_TURB_RECONSTRUCTION", "STANDARD") != "STANDARD")
        throw Unsupported("SCALAR_TURB_RECONSTRUCTION other than STANDARD is not supported");
      // the model constants and solver settings are the reference's defaults (TurbModel_KOMSST.h:16-52)
      for (const char *k : {"sigma_k1", "sigma_k2", "sigma_om1", "sigma_om2", "beta_1", "beta_2", "betaStar", "a1",
                            "ZERO_kine", "ZERO_REL_kine", "ZERO_omega", "ZERO_REL_omega"})
        if (pm.has(k)) throw Unsupported(std::string(k) + ": only the SST defaults are supported");
    } else if (tm != "NONE") {
      throw Unsupported("TURB_MODEL = " + tm + " is not supported (NONE, SA, SST)");
    }
  }
  std::string si = pm.str("SPACE_INTEGRATION", "HLLC");
  if (si == "HLLC") p.scheme = kHllc;
  else if (si == "ROE") p.scheme = kRoe;
  else if (si == "AUSM") p.scheme = kAusm;
  else throw Unsupported("SPACE_INTEGRATION = " + si + " is not supported (HLLC, ROE, AUSM)");
  if (p.sa && p.scheme != kHllc) throw Unsupported("TURB_MODEL = SA is supported with HLLC only");
  if (p.sst && p.scheme != kHllc) throw Unsupported("TURB_MODEL = SST is supported with HLLC only");
  if (p.sst && p.jac_kernel != kResidEntry)
    throw Unsupported("TURB_MODEL = SST needs the per-entry Jacobian (GPU_JAC_KERNEL = ENTRY)");
  if (pm.has("GRAD_RECONSTRUCTION")) {
    std::string s = pm.sub("GRAD_RECONSTRUCTION", "SCHEME");
    if (s == "LEASTSQUARES") p.grad_ls = true;
    else if (s == "GREENGAUSS") p.grad_ls = false;
    else throw Unsupported("GRAD_RECONSTRUCTION SCHEME=" + s + " is not supported");
    std::string l = pm.sub("GRAD_RECONSTRUCTION", "LIMITER");
    if (l == "NO" || l == "NOLIMITER") p.limiter = false;
    else if (l == "YES" || l == "BARTH_JESPERSEN_MOD") p.limiter = true;
    else throw std::invalid_argument("GRAD_RECONSTRUCTION LIMITER=" + l + " not recognised");
    p.eps_limiter = pm.sub_d("GRAD_RECONSTRUCTION", "EPS");
  } else {
    p.grad_ls = false;
    p.limiter = true;
    p.eps_limiter = 0.001;
  }
  std::string sb = pm.str("SYMM_BC", "ORIGINAL_NUMERIC");
  if (sb == "ORIGINAL_NUMERIC") p.symm_bc = kSymmOriginal;
  else if (sb == "NEW_NUMERIC") p.symm_bc = kSymmNew;
  else if (sb == "ANALYTICAL") p.symm_bc = kSymmAnalytical;
  else throw std::invalid_argument("SYMM_BC = " + sb + " not recognised");
  if (pm.has("CFL")) {
    p.dt_mode = 0;
    p.cfl = pm.dbl("CFL");
  } else if (pm.has("CFL_MIN")) {
    p.dt_mode = 1;
    p.cfl = pm.dbl("CFL_MIN");
  } else if (pm.has("DT")) {
    p.dt_mode = 2;
    p.dt_fixed = pm.dbl("DT");
  } else {
    throw std::invalid_argument("DT, CFL or CFL_MIN not specified");
  }
  const std::string tstep = pm.str("TIME_STEP", "MAX_BASED");
  if (tstep != "MAX_BASED" && tstep != "SUM_BASED")
    throw std::invalid_argument("TIME_STEP = " + tstep + " (MAX_BASED or SUM_BASED)");
  p.dt_sum = tstep == "SUM_BASED";
  std::string ti = pm.str("TIME_INTEGRATION", "BACKWARD_EULER");
  if (ti == "BACKWARD_EULER_COUPLED") p.coupled = true;  // JoeWithModels.cpp:39, 4326-4594
  else if (ti != "BACKWARD_EULER")
    throw Unsupported("TIME_INTEGRATION = " + ti + " is not supported (BACKWARD_EULER, BACKWARD_EULER_COUPLED)");
  if (p.coupled && p.sst) throw Unsupported("TURB_MODEL = SST in the coupled loop is not supported");
  std::string ls = pm.str("LINEAR_SOLVER_NS", "BCGSTAB");
  if (p.coupled) {
    // the reference's coupled loop solves with PETSc GMRES only
    // (UgpWithCvCompFlow.h:1000-1037); this build runs its block-Jacobi
    // BiCGStab with LINEAR_SOLVER_NS_TRESHOLDS for either keyword
    if (ls != "PETSC_GMRES" && ls != "BCGSTAB")
      throw Unsupported("LINEAR_SOLVER_NS = " + ls + " is not supported (PETSC_GMRES, BCGSTAB)");
    if (p.scheme != kHllc) throw Unsupported("BACKWARD_EULER_COUPLED is supported with HLLC only");
    // the coupled face loop limits grad u and grad h with the limiter of
    // GRAD_RECONSTRUCTION alone (JoeWithModels.cpp:4691-4716)
    p.visc_limiter = true;
    if (p.flux_kernel != kResidEntry) throw Unsupported("BACKWARD_EULER_COUPLED needs GPU_FLUX_KERNEL = ENTRY");
    p.scheme = kHllcC;
  } else if (ls != "BCGSTAB") {
    throw Unsupported("LINEAR_SOLVER_NS = " + ls + " is not supported (BCGSTAB)");
  }
  p.under_relax = pm.dbl("UNDER_RELAXATION", 0.3);
  if (pm.has("LINEAR_SOLVER_NS_TRESHOLDS")) {
    p.ls_maxiter = pm.sub_i("LINEAR_SOLVER_NS_TRESHOLDS", "MAX_ITER");
    p.ls_abs = pm.sub_d("LINEAR_SOLVER_NS_TRESHOLDS", "ABS_RESID");
    p.ls_rel = pm.sub_d("LINEAR_SOLVER_NS_TRESHOLDS", "REL_RESID");
  }
  if (pm.has("CFL_RAMP")) {
    p.cfl_after = pm.sub_i("CFL_RAMP", "AFTER_ITER");
    p.cfl_interval = pm.sub_i("CFL_RAMP", "INTERVAL_ITER");
    p.cfl_factor = pm.sub_d("CFL_RAMP", "FACTOR_CFL");
    p.cfl_max = pm.sub_d("CFL_RAMP", "MAX_CFL");
  }
  p.nsteps = pm.integer("NSTEPS", 100);
  p.resid_energ_th = pm.dbl("RESID_ENERG_TH", 0.0);
  // this build's own stop rule for served solves: energy L1 residual below
  // RESID_ENERG_REL times that of the first step (0: off)
  p.resid_energ_rel = pm.dbl("RESID_ENERG_REL", 0.0);
  p.rho_init = pm.has("RHO_INITIAL") ? pm.dbl("RHO_INITIAL") : p.rho_ref;
  p.p_init = pm.has("P_INITIAL") ? pm.dbl("P_INITIAL") : p.p_ref;
  if (pm.has("U_INITIAL")) {
    const auto *t = pm.get("U_INITIAL");
    if (t->size() < 4) throw std::invalid_argument("U_INITIAL needs 3 values");
    for (int i = 0; i < 3; i++) p.u_init[i] = ParamMap::to_d((*t)[1 + i], "U_INITIAL");
  }
  zbc.assign(m.zones.size(), ZoneBc());
  for (size_t iz = 0; iz < m.zones.size(); iz++) {
    const Zone &z = m.zones[iz];
    if (z.kind != kZoneBoundary) continue;
    const auto *t = pm.get(z.name);
    if (!t || t->size() < 2) throw std::invalid_argument("no bc set for zone " + z.name);
    const std::string &kw = (*t)[1];
    auto val = [&](size_t pos) {
      if (pos >= t->size()) throw std::invalid_argument("zone " + z.name + ": missing value for " + kw);
      return ParamMap::to_d((*t)[pos], z.name);
    };
    ZoneBc &b = zbc[iz];
    if (kw == "CBC" || kw == "CBC_SUBSONIC_INLET") {
      b.bc = kw == "CBC" ? kBcCbc : kBcCbcSubsonicInlet;
      for (int k = 0; k < 5; k++) b.par[k] = val(2 + k);
    } else if (kw == "CBC_SUBSONIC_OUTLET") {
      b.bc = kBcCbcSubsonicOutlet;
      b.par[0] = val(2);
    } else if (kw == "SYMMETRY") {
      b.bc = kBcSymmetry;
    } else if (kw == "NEUMANN") {
      b.bc = kBcNeumann;
    } else if (kw == "WALL") {
      b.bc = kBcWall;
      for (size_t k = 0; k < t->size(); k++)
        if ((*t)[k] == "TEMP" && k > 0) b.par[0] = val(k + 1);
    } else if (kw == "FARFIELD") {
      b.bc = kBcFarfield;
      for (int k = 0; k < 6; k++) b.par[k] = val(2 + k);
    } else if (kw == "HOOK") {
      b.bc = kBcHook;
    } else if (kw == "PROFILES") {  // JoeWithModels.cpp:249-256: one file, the last PROFILES zone's
      b.bc = kBcProfiles;
      if (t->size() < 3) throw std::invalid_argument("zone " + z.name + ": PROFILES needs a file name");
      p.profiles_file = (*t)[2];
    } else {
      throw std::invalid_argument("cannot recognize bc for zone " + z.name + ": " + kw);
    }
    // nuSA boundary value (scalarZoneIsHook / Profiles / Dirichlet / Flux,
    // UgpWithCvCompFlow.h:1881-1958)
    if (p.sa) {
      if (pm.has(z.name + ".nuSA.flux")) throw Unsupported("zone " + z.name + ": scalar flux boundaries");
      if (const auto *st = pm.get(z.name + ".nuSA")) {
        if (st->size() < 2) throw std::invalid_argument("zone " + z.name + ".nuSA: missing value");
        const std::string &v = (*st)[1];
        if (v == "HOOK" || v == "Hook" || v == "hook")
          throw Unsupported("zone " + z.name + ".nuSA = " + v + " is not supported");
        if (v == "PROFILES") {
          // the reference fills the scalar's profile values on the zones whose
          // flow boundary is PROFILES only (JoeWithModels.cpp:296-357)
          if (b.bc != kBcProfiles)
            throw Unsupported("zone " + z.name + ".nuSA = PROFILES needs the zone's own PROFILES boundary");
          b.skind = 3;
        } else {
          b.skind = 1;
          b.sval = ParamMap::to_d(v, z.name + ".nuSA");
        }
      }
    }
    // kine and omega boundary values (the same rules; omega = HOOK on a WALL
    // zone takes boundaryHookScalarRansTurb's wall value, TurbModel_KOMSST.h:342-358)
    if (p.sst) {
      for (int s = 0; s < 2; s++) {
        const std::string nm = s == 0 ? "kine" : "omega";
        if (pm.has(z.name + "." + nm + ".flux")) throw Unsupported("zone " + z.name + ": scalar flux boundaries");
        const auto *st = pm.get(z.name + "." + nm);
        if (!st) continue;
        if (st->size() < 2) throw std::invalid_argument("zone " + z.name + "." + nm + ": missing value");
        const std::string &v = (*st)[1];
        int kind = 1;
        double val = 0.0;
        if (v == "HOOK" || v == "Hook" || v == "hook") {
          if (s == 0 || b.bc != kBcWall)
            throw Unsupported("zone " + z.name + "." + nm + " = HOOK: the SST hook sets omega on WALL zones only");
          kind = 2;
        } else if (v == "PROFILES") {
          throw Unsupported("zone " + z.name + "." + nm + " = PROFILES is not supported");
        } else {
          val = ParamMap::to_d(v, z.name + "." + nm);
        }
        if (s == 0) {
          b.skind = kind;
          b.sval = val;
        } else {
          b.skind2 = kind;
          b.sval2 = val;
        }
      }
    }
  }
  if (p.sa || p.sst) {
    bool wall = false;
    for (const ZoneBc &b : zbc) wall = wall || b.bc == kBcWall;
    if (!wall) throw std::invalid_argument("the turbulence model needs a WALL zone for the wall distance");
  }
  if (p.sst && !p.profiles_file.empty()) throw Unsupported("PROFILES boundaries with SST are not supported");
  if (!p.profiles_file.empty()) (void)profile_faces(p.profiles_file, m, zbc, p.sa ? 1 : 0);  // checked here, read again at set-up
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
struct Ctx::Impl {
  int ncv = 0, nfa = 0, nfa_b = 0, nnz = 0, nzone = 0;
  int precision = 0;
  hipStream_t stream = nullptr;
  SolverParams par;
  std::vector<ZoneBc> zbc;
  std::vector<int> zone_slot;  // mesh zone index -> device zone slot
  DParams dp;
  DMesh dm;
  double cfl = 0.0;
  int step_no = 0;
  int step0 = 0;  // the step count the current solve started from (set_state: 0, set_step)
  bool have_state = false;
  // mesh buffers
  DBuf<double> fgeo, fnorm, lsg, bfa_lsg, vol;
  DBuf<int> cfl_i, cfl_noc, nbocv_i, nbocv_v, bfa_cell, bfa_zone;
  DBuf<int2> cfl_e;
  DBuf<int2> ecp;
  DBuf<int> wave_e0;
  DBuf<double> cfl_r0;
  int nwave = 0;
  DBuf<ZoneDev> zones;
  std::vector<int> perm;              // device cell -> reference cell
  std::vector<int> fnew;              // reference face -> device face
  std::vector<int> h_dev_nbocv_i, h_ref_nbocv_i;
  bool dup_blocks = false;
  int min_faces = 1 << 30;  // fewest faces of a cell
  // the per-entry Jacobian kernel: needs no duplicate neighbour pairs (their
  // shared block is accumulated in entry order by k_jac) and at least four
  // faces per cell (its lanes own at most seven components of a block)
  bool jac_entry() const { return par.jac_kernel == kResidEntry && !dup_blocks && min_faces >= 4; }
  // state
  DBuf<double> q, prim, grad, bfa, rhs, dt, hook;
  DBuf<double> gradv;  // [ncv][12] velocity and enthalpy gradients of the viscous flux
  DBuf<int2> fcell;    // [nfa_i] device cells (icv0, icv1) of each internal face (face form of the flux)
  DBuf<double> fflux;  // [nfa_i][5] face fluxes of the face form
  bool face_form() const { return par.flux_kernel == kResidFace && !dp.visc; }
  DBuf<int> counters;  // [0] negative pressure, [1] first-order fallbacks
  // solver (T = double or float, both allocated lazily)
  DBuf<double> A64, LU64, v64[10];
  DBuf<float> A32, LU32, v32[10];
  DBuf<double> F64, Dinv64;  // block ILU(0) factors (GPU_PRECONDITIONER = ILU0)
  DBuf<float> F32, Dinv32;
  DBuf<double> partials;
  DBuf<SolverScalars> scal;
  SolverScalars *h_scal = nullptr;  // pinned
  // Spalart-Allmaras: nuSA, per-cell fields [ncv][kAuxN], residual, AScal
  // rows [nnz][6], Jacobi BiCGStab vectors (b, x, res, res0, p, v, phat, s,
  // shat, t), its scalars, the faocv entry map
  DBuf<double> sa_phi, sa_aux, sa_prof, rhs_s, AS, sw[10];
  // Menter SST: kine, omega [2][ncv], per-cell fields [ncv][kSstN], boundary
  // values [nfa_b][2]; residuals rhs_s [2][ncv] and AScal rows AS [2][nnz][6]
  DBuf<double> sst_phi, sst_aux, sst_bfa;
  DBuf<double> sst_dscal;  // the last step's kine and omega corrections [2][ncv] (DSCAL)
  int scal_iters2 = 0;
  DBuf<SolverScalars> scal_s;
  SolverScalars *h_scal_s = nullptr;  // pinned
  DBuf<int> foc;
  DBuf<int> cell_perm;  // the per-cell gradient's thread -> cell map (GPU_GRAD_THREAD_ORDER = SORTED)
  int scal_iters = 0;
  // block size of the linear system: 5, or 6 in the coupled loop with the
  // Spalart-Allmaras scalar (its row and column in every block)
  int nbs = 5;
  bool coupled_sa() const { return nbs == 6; }
  // the Jacobian, preconditioner and Krylov vectors of the block system in
  // precision T (nvec work vectors)
  template <typename T>
  void ensure_solver(int nvec, bool with_lu);
  template <typename T>
  void launch_scal_c(T *A, bool step_mode);
  template <typename T>
  void update();
  template <typename T>
  void time_term();
  void push_source(bool with_a);
  template <typename T>
  void scalar_step(const T *dq);
  template <typename T>
  void sst_scalar_step(const T *dq);
  void finish_step(StepInfo &info, int it, int st);
  // split step (host hooks): host source terms in device order, phase flag
  DBuf<double> src;
  std::vector<double> h_src;
  bool begun = false;

  template <typename T>
  void solve(int &iters, int &status);
  template <typename T>
  void solve_prologue();
  template <typename T>
  void solve_iterations(int count);
  template <typename T>
  DBuf<T> *krylov();
  template <typename T>
  BcgSys bcg_sys();
  template <typename T>
  DBuf<T> *lu_();
  // device-driven step in three captured graphs: head (CFL ramp, residual,
  // Jacobian, Krylov start, kDevChunk iterations), chunk (kDevChunk more),
  // tail (update, refresh, L1); each ends with a report to host-mapped
  // memory and the host picks chunk or tail from the solver flag
  static constexpr int kDevChunk = 4;
  template <typename T>
  void enqueue_step_head();
  template <typename T>
  void enqueue_step_chunk();
  template <typename T>
  void enqueue_step_tail();
  hipGraphExec_t capture(void (Ctx::Impl::*enq)());
  hipStream_t cap_stream = nullptr;
  hipGraphExec_t g_head = nullptr, g_chunk = nullptr, g_tail = nullptr;
  int dev_phase = 0;  // 1: head or chunk in flight, 2: tail in flight
  hipEvent_t ev_step = nullptr;
  StepMail *mail_h = nullptr, *mail_d = nullptr;
  bool dev_step_dirty = true;
  bool graph_inflight = false;
  double resid0_energy = 0.0;
  template <typename T>
  T *A_();
  void launch_state_bc();
  void launch_grad(hipStream_t s);
  void launch_flux(hipStream_t s, const double *cfl_dev = nullptr);
  int flux_blocks() const;
  void launch_dt_fin();
  bool flux_entry() const;
  // jac_follows: the Jacobian pass comes next (it forms the coupled
  // scalar row together with its blocks)
  void launch_grad_flux(bool jac_follows = false);
  template <typename T>
  void launch_jac(bool step_mode);
};

template <>
double *Ctx::Impl::A_<double>() {
  return A64.p;
}
template <>
DBuf<double> *Ctx::Impl::krylov<double>() {
  return v64;
}
template <>
DBuf<float> *Ctx::Impl::krylov<float>() {
  return v32;
}
template <>
DBuf<double> *Ctx::Impl::lu_<double>() {
  return &LU64;
}
template <>
DBuf<float> *Ctx::Impl::lu_<float>() {
  return &LU32;
}
template <>
float *Ctx::Impl::A_<float>() {
  return A32.p;
}

// Morton key of a point in the unit cube [0,1)^3 (21 bits per axis)
static uint64_t morton3(double x, double y, double z) {
  auto spread = [](uint64_t v) {
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
  };
  auto q = [](double t) {
    double u = t < 0.0 ? 0.0 : (t >= 1.0 ? 0.99999999 : t);
    return (uint64_t)(u * (double)(1 << 21));
  };
  return spread(q(x)) | spread(q(y)) << 1 | spread(q(z)) << 2;
}

Ctx::Ctx(const Mesh &m, const std::string &joe_in, int precision, void *stream) : d_(new Impl) {
  Impl &d = *d_;
  d.precision = precision;
  d.stream = (hipStream_t)stream;
  parse_params(joe_in, m, d.par, d.zbc);
  d.ncv = m.ncv;
  d.nfa = m.nfa;
  d.nfa_b = m.nfa_b;
  d.nnz = m.nbocv_s();
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw std::runtime_error("no HIP device available");
  const int ncv = m.ncv, nfa = m.nfa, nfa_b = m.nfa_b;

  // zones: device slots for boundary zones
  std::vector<ZoneDev> zd;
  d.zone_slot.assign(m.zones.size(), -1);
  for (size_t iz = 0; iz < m.zones.size(); iz++) {
    if (m.zones[iz].kind != kZoneBoundary) continue;
    ZoneDev z;
    std::memset(&z, 0, sizeof(z));
    z.bc = d.zbc[iz].bc;
    for (int k = 0; k < 8; k++) z.par[k] = d.zbc[iz].par[k];
    z.skind = d.zbc[iz].skind;
    z.sval = d.zbc[iz].sval;
    z.skind2 = d.zbc[iz].skind2;
    z.sval2 = d.zbc[iz].sval2;
    d.zone_slot[iz] = (int)zd.size();
    zd.push_back(z);
  }
  d.nzone = (int)zd.size();
  if (zd.empty()) zd.push_back(ZoneDev());
  d.zones.upload(zd.data(), zd.size());
  std::vector<int> bz(std::max(1, nfa_b), 0);
  for (size_t iz = 0; iz < m.zones.size(); iz++) {
    const Zone &z = m.zones[iz];
    if (z.kind != kZoneBoundary) continue;
    for (int f = z.ifa_f; f <= z.ifa_l; f++) bz[f] = d.zone_slot[iz];
  }

  // per-cell face lists in the reference's accumulation order (reference ids)
  std::vector<int> ci(ncv + 1, 0), slot_of;  // slot_of: position in the reference nbocv row
  std::vector<int2> cf;
  cf.reserve(m.faocv_v.size());
  slot_of.reserve(m.faocv_v.size());
  std::vector<int> fi, fb, used;
  for (int c = 0; c < ncv; c++) {
    fi.clear();
    fb.clear();
    used.clear();
    for (int k = m.faocv_i[c]; k < m.faocv_i[c + 1]; k++) {
      int f = m.faocv_v[k];
      (f >= nfa_b ? fi : fb).push_back(f);
    }
    std::sort(fi.begin(), fi.end());
    std::sort(fb.begin(), fb.end());
    ci[c] = (int)cf.size();
    for (int f : fi) {
      int c0 = m.cvofa[2 * (size_t)f], c1 = m.cvofa[2 * (size_t)f + 1];
      int side = (c0 == c) ? 0 : 1;
      int nb = side == 0 ? c1 : c0;
      // getImplDependencyIndex (UgpWithCv.h:745-755): first match in the row
      int slot = m.nbocv_i[c];
      while (m.nbocv_v[++slot] != nb) {
      }
      bool acc = std::find(used.begin(), used.end(), slot) != used.end();
      if (acc) d.dup_blocks = true;
      used.push_back(slot);
      cf.push_back(make_int2(f | (side << 31), nb));
      slot_of.push_back((slot - m.nbocv_i[c]) | (acc ? (1 << 30) : 0));
    }
    for (int f : fb) {
      cf.push_back(make_int2(f, -(bz[f] + 1)));
      slot_of.push_back(0);
    }
    d.min_faces = std::min(d.min_faces, (int)cf.size() - ci[c]);
  }
  ci[ncv] = (int)cf.size();
  // SST runs the per-entry Jacobian only
  if (d.par.sst && !d.jac_entry())
    throw Unsupported("TURB_MODEL = SST needs the per-entry Jacobian (GPU_JAC_KERNEL = ENTRY, cells of four or more "
                      "faces, no duplicate neighbour pairs)");

  // Device numbering (internal only; the ABI speaks reference numbering),
  // chosen by GPU_CELL_ORDER: the reference's own order (NATIVE), or Morton
  // order of the centroids (MORTON), optionally with the cells of every tile
  // of kTile cells grouped by face count so a wavefront's lanes loop equally
  // long (GROUPED, MORTON_GROUPED). Boundary faces keep their index; internal
  // faces are numbered in order of first visit by the device-ordered cells.
  // Per-cell sums keep the reference order, so results do not depend on it.
  const int kTile = 1024;
  const int order = d.par.cell_order;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int c = 0; c < ncv; c++)
    for (int i = 0; i < 3; i++) {
      lo[i] = std::min(lo[i], m.x_cv[3 * (size_t)c + i]);
      hi[i] = std::max(hi[i], m.x_cv[3 * (size_t)c + i]);
    }
  double ext = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  if (!(ext > 0.0)) ext = 1.0;
  std::vector<uint64_t> key(ncv);
  for (int c = 0; c < ncv; c++)
    key[c] = morton3((m.x_cv[3 * (size_t)c] - lo[0]) / ext, (m.x_cv[3 * (size_t)c + 1] - lo[1]) / ext,
                     (m.x_cv[3 * (size_t)c + 2] - lo[2]) / ext);
  std::vector<int> &perm = d.perm;
  perm.resize(ncv);
  for (int c = 0; c < ncv; c++) perm[c] = c;
  if (order == kOrderMorton || order == kOrderMortonGrouped)
    std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return key[a] < key[b]; });
  if (order == kOrderMortonGrouped || order == kOrderGrouped)
    for (int t0 = 0; t0 < ncv; t0 += kTile) {
      int t1 = std::min(ncv, t0 + kTile);
      std::stable_sort(perm.begin() + t0, perm.begin() + t1,
                       [&](int a, int b) { return (ci[a + 1] - ci[a]) < (ci[b + 1] - ci[b]); });
    }
  std::vector<int> inv(ncv);
  for (int cd = 0; cd < ncv; cd++) inv[perm[cd]] = cd;
  std::vector<int> fnew(nfa, -1);
  for (int f = 0; f < nfa_b; f++) fnew[f] = f;
  int nf = nfa_b;
  for (int cd = 0; cd < ncv; cd++) {
    int c = perm[cd];
    for (int e = ci[c]; e < ci[c + 1]; e++) {
      int f = cf[e].x & 0x7fffffff;
      if (cf[e].y >= 0 && fnew[f] < 0) fnew[f] = nf++;
    }
  }
  if (nf != nfa) throw std::runtime_error("internal face not reached by any cell");
  d.fnew = fnew;
  if (d.par.flux_kernel == kResidFace && !(d.par.mu_ref > 0.0) && nfa > nfa_b) {
    std::vector<int2> fc(nfa - nfa_b);
    for (int f = nfa_b; f < nfa; f++)
      fc[fnew[f] - nfa_b] = make_int2(inv[m.cvofa[2 * (size_t)f]], inv[m.cvofa[2 * (size_t)f + 1]]);
    d.fcell.upload(fc.data(), fc.size());
    d.fflux.alloc(5 * (size_t)(nfa - nfa_b));
  }

  // face records in device order: unit normal and area as normVec3d, x_fa
  std::vector<double> fg(8 * (size_t)nfa);
  std::vector<double> fn;
  if (!d.par.grad_ls || d.par.sa || d.par.sst) fn.resize(3 * (size_t)nfa);
  for (int f = 0; f < nfa; f++) {
    const double *n = &m.fa_normal[3 * (size_t)f];
    double mag = 0.0;
    mag += n[0] * n[0];
    mag += n[1] * n[1];
    mag += n[2] * n[2];
    mag = std::sqrt(mag);
    double *g = &fg[8 * (size_t)fnew[f]];
    g[0] = n[0] / mag;
    g[1] = n[1] / mag;
    g[2] = n[2] / mag;
    g[3] = mag;
    g[4] = m.x_fa[3 * (size_t)f];
    g[5] = m.x_fa[3 * (size_t)f + 1];
    g[6] = m.x_fa[3 * (size_t)f + 2];
    g[7] = 0.0;
    if (!fn.empty())
      for (int i = 0; i < 3; i++) fn[3 * (size_t)fnew[f] + i] = n[i];
  }
  d.fgeo.upload(fg.data(), fg.size());
  std::vector<double>().swap(fg);
  if (!fn.empty()) d.fnorm.upload(fn.data(), fn.size());

  // rows of nbocv / LS coefficients / Jacobian in device cell order
  std::vector<int> &dni = d.h_dev_nbocv_i;
  dni.assign(ncv + 1, 0);
  for (int cd = 0; cd < ncv; cd++) {
    int c = perm[cd];
    dni[cd + 1] = dni[cd] + (m.nbocv_i[c + 1] - m.nbocv_i[c]);
  }
  {
    std::vector<int> nv(d.nnz);
    std::vector<double> lg(3 * (size_t)d.nnz);
    for (int cd = 0; cd < ncv; cd++) {
      int c = perm[cd];
      for (int j = 0; j < m.nbocv_i[c + 1] - m.nbocv_i[c]; j++) {
        nv[dni[cd] + j] = inv[m.nbocv_v[m.nbocv_i[c] + j]];
        for (int i = 0; i < 3; i++) lg[3 * (size_t)(dni[cd] + j) + i] = m.nbocv_lsg[3 * (size_t)(m.nbocv_i[c] + j) + i];
      }
    }
    d.nbocv_i.upload(dni.data(), dni.size());
    d.nbocv_v.upload(nv.data(), nv.size());
    d.lsg.upload(lg.data(), lg.size());
  }
  {
    std::vector<int> dci(ncv + 1, 0), noc;
    std::vector<int2> dcf;
    dcf.reserve(cf.size());
    noc.reserve(cf.size());
    for (int cd = 0; cd < ncv; cd++) {
      int c = perm[cd];
      dci[cd] = (int)dcf.size();
      for (int e = ci[c]; e < ci[c + 1]; e++) {
        int f = cf[e].x & 0x7fffffff;
        int side_bit = cf[e].x & (int)0x80000000;
        if (cf[e].y >= 0) {
          dcf.push_back(make_int2(fnew[f] | side_bit, inv[cf[e].y]));
          noc.push_back((dni[cd] + (slot_of[e] & 0x3fffffff)) | (slot_of[e] & (1 << 30)));
        } else {
          dcf.push_back(make_int2(f, cf[e].y));
          noc.push_back(dni[cd]);
        }
      }
    }
    dci[ncv] = (int)dcf.size();
    d.cfl_i.upload(dci.data(), dci.size());
    d.cfl_e.upload(dcf.data(), dcf.size());
    // the per-cell gradient's thread -> cell map: within each tile of kBlock
    // device cells (one block), the cells by ascending face count, so the
    // lanes of a wavefront loop over similar numbers of faces; the cells and
    // the memory layout stay where they are (GPU_GRAD_THREAD_ORDER = SORTED;
    // measured slower at 160^3, 2.76 against 2.61 ms: the lanes' rows spread
    // over the tile, so NATIVE is the default)
    if (d.par.grad_thread_sorted) {
      std::vector<int> tp(ncv);
      for (int t0 = 0; t0 < ncv; t0 += kBlock) {
        const int t1 = std::min(ncv, t0 + kBlock);
        for (int t = t0; t < t1; t++) tp[t] = t;
        std::stable_sort(tp.begin() + t0, tp.begin() + t1,
                         [&](int a, int b) { return dci[a + 1] - dci[a] < dci[b + 1] - dci[b]; });
      }
      d.cell_perm.upload(tp.data(), tp.size());
    }
    // face-centroid offset x_fa - x_cv of every entry (the limiter's r0), so
    // the per-cell limiter loop reads it without gathering face rows
    {
      std::vector<double> r0(3 * std::max<size_t>(1, dcf.size()), 0.0);
      size_t k = 0;
      for (int cd = 0; cd < ncv; cd++) {
        const int c = perm[cd];
        for (int e = ci[c]; e < ci[c + 1]; e++, k++) {
          const int f = cf[e].x & 0x7fffffff;
          for (int i = 0; i < 3; i++) r0[3 * k + i] = m.x_fa[3 * (size_t)f + i] - m.x_cv[3 * (size_t)c + i];
        }
      }
      d.cfl_r0.upload(r0.data(), r0.size());
    }
    // Spalart-Allmaras and SUM_BASED time steps: for each cell, the entry
    // offsets of its faces in the reference's faocv order (the order of
    // InterpolateAtCellCenterFromFaceValues and of calcDt's face loop)
    if (d.par.sa || d.par.sst || d.par.dt_sum) {
      std::vector<int> fo(std::max<size_t>(1, dcf.size()), 0);
      size_t k = 0;
      for (int cd = 0; cd < ncv; cd++) {
        const int c = perm[cd];
        const int n = ci[c + 1] - ci[c];
        if (m.faocv_i[c + 1] - m.faocv_i[c] != n) throw std::runtime_error("faocv and the cell's entries disagree");
        for (int j = 0; j < n; j++, k++) {
          const int f = m.faocv_v[m.faocv_i[c] + j];
          int pos = -1;
          for (int e = ci[c]; e < ci[c + 1]; e++)
            if ((cf[e].x & 0x7fffffff) == f) pos = e - ci[c];
          if (pos < 0) throw std::runtime_error("face of faocv not among the cell's entries");
          fo[k] = pos;
        }
      }
      d.foc.upload(fo.data(), fo.size());
    }
    // entry-parallel kernels: the cell of every entry and its position in
    // the cell's list, and wavefronts of 64 lanes that each cover whole cells
    // (consecutive cells while their entries fit), by first entry
    std::vector<int2> ecp(std::max<size_t>(1, dcf.size()));
    for (int cd = 0; cd < ncv; cd++)
      for (int e = dci[cd]; e < dci[cd + 1]; e++) ecp[e] = make_int2(cd, e - dci[cd]);
    std::vector<int> we;
    int c = 0;
    while (c < ncv) {
      we.push_back(dci[c]);
      const int eb = dci[c];
      if (dci[c + 1] - eb > 64) throw std::runtime_error("a cell with more than 64 faces");
      while (c < ncv && dci[c + 1] - eb <= 64) c++;
    }
    we.push_back(dci[ncv]);
    d.nwave = (int)we.size() - 1;
    d.ecp.upload(ecp.data(), ecp.size());
    d.wave_e0.upload(we.data(), we.size());
    d.cfl_noc.upload(noc.data(), noc.size());
  }
  std::vector<int> bcell(std::max(1, nfa_b), 0);
  for (int f = 0; f < nfa_b; f++) bcell[f] = inv[m.cvofa[2 * (size_t)f]];
  d.bfa_zone.upload(bz.data(), bz.size());
  d.bfa_cell.upload(bcell.data(), bcell.size());
  if (nfa_b > 0)
    d.bfa_lsg.upload(m.bfa_lsg.data(), m.bfa_lsg.size());
  else
    d.bfa_lsg.alloc(3);
  {
    std::vector<double> v(ncv);
    for (int cd = 0; cd < ncv; cd++) v[cd] = m.cv_volume[perm[cd]];
    d.vol.upload(v.data(), v.size());
  }
  d.h_ref_nbocv_i = m.nbocv_i;

  // state buffers; prim carries x_cv in its last three slots
  const size_t n = ncv;
  d.q.alloc(5 * n);
  std::vector<double> pr(8 * n, 0.0);
  for (size_t cd = 0; cd < n; cd++)
    for (int i = 0; i < 3; i++) pr[8 * cd + 5 + i] = m.x_cv[3 * (size_t)perm[cd] + i];
  d.prim.upload(pr.data(), pr.size());
  d.grad.alloc(16 * n);
  HIPCHECK(hipMemset(d.grad.p, 0, sizeof(double) * 16 * n));
  d.gradv.alloc(d.par.mu_ref > 0.0 ? 12 * n : 12);
  HIPCHECK(hipMemset(d.gradv.p, 0, sizeof(double) * d.gradv.n));
  d.bfa.alloc(8 * (size_t)std::max(1, nfa_b));
  d.rhs.alloc(5 * n);
  d.dt.alloc(n);
  d.hook.alloc(5 * (size_t)std::max(1, nfa_b));
  HIPCHECK(hipMemset(d.hook.p, 0, d.hook.n * sizeof(double)));
  if (!d.par.profiles_file.empty()) {
    // PROFILES zones: the face values rho, u, 0, 0, p interpolated in y from
    // the profile file (JoeWithModels.cpp:262-326), kept in the per-face
    // boundary buffer that HOOK zones fill from the host
    std::vector<double> sp;
    std::vector<double> h = profile_faces(d.par.profiles_file, m, d.zbc, d.par.sa ? 1 : 0, &sp);
    HIPCHECK(hipMemcpy(d.hook.p, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    if (d.par.sa) d.sa_prof.upload(sp.data(), sp.size());
  }
  if (!d.sa_prof.p) {
    const double zero = 0.0;
    d.sa_prof.upload(&zero, 1);
  }
  d.counters.alloc(4);
  HIPCHECK(hipMemset(d.counters.p, 0, 4 * sizeof(int)));
  // reduction partials: the largest grid that writes them (five per block
  // of the NB-lanes-per-cell kernels, or one per flux wavefront)
  d.partials.alloc(std::max<size_t>(8 * (size_t)(nblk(5 * n) + 64), (size_t)d.nwave + d.nwave / kMinPre + 64));
  d.scal.alloc(1);
  HIPCHECK(hipMemset(d.scal.p, 0, sizeof(SolverScalars)));
  HIPCHECK(hipHostMalloc((void **)&d.h_scal, sizeof(SolverScalars)));
  if (d.par.sa) {
    std::vector<double> ph(n, d.par.sa_init);  // initialHook (JoeWithModels.cpp:219-240)
    d.sa_phi.upload(ph.data(), n);
    d.sa_aux.alloc(kAuxN * n);
    HIPCHECK(hipMemset(d.sa_aux.p, 0, sizeof(double) * kAuxN * n));
    d.rhs_s.alloc(n);
    if (!d.par.coupled) {  // the segregated scalar's AScal rows and Jacobi BiCGStab
      d.AS.alloc(6 * (size_t)d.nnz);
      HIPCHECK(hipMemset(d.AS.p, 0, sizeof(double) * 6 * (size_t)d.nnz));
      for (int k = 0; k < 10; k++) d.sw[k].alloc(n);
    }
    d.scal_s.alloc(1);
    HIPCHECK(hipMemset(d.scal_s.p, 0, sizeof(SolverScalars)));
    HIPCHECK(hipHostMalloc((void **)&d.h_scal_s, sizeof(SolverScalars)));
    // wall faces in zone order (calcWallDistance, initialHookScalarRansTurbModel)
    std::vector<double> wx, wn;
    for (size_t iz = 0; iz < m.zones.size(); iz++) {
      const Zone &z = m.zones[iz];
      if (z.kind != kZoneBoundary || d.zbc[iz].bc != kBcWall) continue;
      for (int f = z.ifa_f; f <= z.ifa_l; f++)
        for (int i = 0; i < 3; i++) {
          wx.push_back(m.x_fa[3 * (size_t)f + i]);
          wn.push_back(m.fa_normal[3 * (size_t)f + i]);
        }
    }
    if (wx.empty()) throw std::invalid_argument("TURB_MODEL = SA: the WALL zones have no faces");
    DBuf<double> dwx, dwn;
    dwx.upload(wx.data(), wx.size());
    dwn.upload(wn.data(), wn.size());
    hipLaunchKernelGGL(k_wall_dist, dim3(nblk(ncv)), dim3(kBlock), 0, d.stream, ncv, d.prim.p, (int)(wx.size() / 3),
                       dwx.p, dwn.p, d.sa_aux.p, (int)kAuxN, (int)kAuxWall);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(d.stream));
  }
  if (d.par.sst) {
    std::vector<double> ph(2 * (size_t)n);  // initialHook (JoeWithModels.cpp:219-240)
    for (size_t c = 0; c < n; c++) {
      ph[c] = d.par.sst_init[0];
      ph[n + c] = d.par.sst_init[1];
    }
    d.sst_phi.upload(ph.data(), ph.size());
    d.sst_aux.alloc(kSstN * n);
    HIPCHECK(hipMemset(d.sst_aux.p, 0, sizeof(double) * kSstN * n));
    d.sst_bfa.alloc(2 * (size_t)std::max(nfa_b, 1));
    HIPCHECK(hipMemset(d.sst_bfa.p, 0, sizeof(double) * 2 * (size_t)std::max(nfa_b, 1)));
    d.rhs_s.alloc(2 * n);
    d.AS.alloc(12 * (size_t)d.nnz);
    HIPCHECK(hipMemset(d.AS.p, 0, sizeof(double) * 12 * (size_t)d.nnz));
    for (int k = 0; k < 10; k++) d.sw[k].alloc(n);
    d.sst_dscal.alloc(2 * n);
    HIPCHECK(hipMemset(d.sst_dscal.p, 0, sizeof(double) * 2 * n));
    d.scal_s.alloc(1);
    HIPCHECK(hipMemset(d.scal_s.p, 0, sizeof(SolverScalars)));
    HIPCHECK(hipHostMalloc((void **)&d.h_scal_s, sizeof(SolverScalars)));
    // wall distance (calcWallDistance, initialHookScalarRansTurbModel, TurbModel_KOMSST.h:82-98)
    std::vector<double> wx, wn;
    for (size_t iz = 0; iz < m.zones.size(); iz++) {
      const Zone &z = m.zones[iz];
      if (z.kind != kZoneBoundary || d.zbc[iz].bc != kBcWall) continue;
      for (int f = z.ifa_f; f <= z.ifa_l; f++)
        for (int i = 0; i < 3; i++) {
          wx.push_back(m.x_fa[3 * (size_t)f + i]);
          wn.push_back(m.fa_normal[3 * (size_t)f + i]);
        }
    }
    if (wx.empty()) throw std::invalid_argument("TURB_MODEL = SST: the WALL zones have no faces");
    DBuf<double> dwx, dwn;
    dwx.upload(wx.data(), wx.size());
    dwn.upload(wn.data(), wn.size());
    hipLaunchKernelGGL(k_wall_dist, dim3(nblk(ncv)), dim3(kBlock), 0, d.stream, ncv, d.prim.p, (int)(wx.size() / 3),
                       dwx.p, dwn.p, d.sst_aux.p, (int)kSstN, (int)kSWall);
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(d.stream));
  }

  d.dm.ncv = ncv;
  d.dm.nfa = nfa;
  d.dm.nfa_b = nfa_b;
  d.dm.nnz = d.nnz;
  d.dm.fgeo = d.fgeo.p;
  d.dm.fnorm = d.fnorm.p;
  d.dm.cfl_i = d.cfl_i.p;
  d.dm.cfl = d.cfl_e.p;
  d.dm.cfl_noc = d.cfl_noc.p;
  d.dm.nbocv_i = d.nbocv_i.p;
  d.dm.nbocv_v = d.nbocv_v.p;
  d.dm.lsg = d.lsg.p;
  d.dm.bfa_lsg = d.bfa_lsg.p;
  d.dm.bfa_cell = d.bfa_cell.p;
  d.dm.bfa_zone = d.bfa_zone.p;
  d.dm.vol = d.vol.p;
  d.dm.zones = d.zones.p;
  d.dm.ecp = d.ecp.p;
  d.dm.cfl_r0 = d.cfl_r0.p;
  d.dm.wave_e0 = d.wave_e0.p;
  d.dm.nwave = d.nwave;
  d.dm.foc = d.foc.p;
  d.dm.cell_perm = d.cell_perm.p;

  d.dp.gamma = d.par.gamma;
  d.dp.R = d.par.R_gas;
  d.dp.eps = d.par.eps_limiter;
  for (int i = 0; i < 3; i++) d.dp.g[i] = d.par.gravity[i];
  d.dp.scheme = d.par.scheme;
  d.dp.second_order = d.par.second_order;
  d.dp.grad_ls = d.par.grad_ls;
  d.dp.limiter = d.par.limiter;
  d.dp.symm_bc = d.par.symm_bc;
  d.dp.gravity_on = (d.par.gravity[0] != 0.0 || d.par.gravity[1] != 0.0 || d.par.gravity[2] != 0.0);
  d.dp.visc = d.par.mu_ref > 0.0;
  d.dp.visc_mode = d.par.visc_mode;
  d.dp.visc_limiter = d.par.visc_limiter;
  d.dp.mu_ref = d.par.mu_ref;
  d.dp.mu_power_law = d.par.mu_power_law;
  d.dp.sl_tref = d.par.sl_tref;
  d.dp.sl_sref = d.par.sl_sref;
  d.dp.T_ref = d.par.T_ref;
  d.dp.pr = d.par.pr;
  d.dp.pr_turb = d.par.pr_turb;
  d.dp.sa = d.par.sa ? 1 : 0;
  d.nbs = (d.par.coupled && d.par.sa) ? 6 : 5;
  d.dp.sa_phi = d.sa_phi.p;
  d.dp.sa_prof = d.sa_prof.p;
  d.dp.fdump = nullptr;
  d.dp.vdump = nullptr;
  d.dp.sa_aux = d.sa_aux.p;
  d.dp.sst = d.par.sst ? 1 : 0;
  d.dp.sst_limit_pk = d.par.sst_limit_pk;
  d.dp.sst_lim_shear = d.par.sst_limiter_shear;
  d.dp.sst_phi = d.sst_phi.p;
  d.dp.sst_aux = d.sst_aux.p;
  d.dp.sst_bfa = d.sst_bfa.p;
  d.dp.sst_alfa[0] = 0.075 / 0.09 - 0.5 * std::pow(0.41, 2.0) / std::sqrt(0.09);
  d.dp.sst_alfa[1] = 0.0828 / 0.09 - 0.856 * std::pow(0.41, 2.0) / std::sqrt(0.09);
  d.cfl = d.par.cfl;
  HIPCHECK(hipDeviceSynchronize());
}

Ctx::~Ctx() {
  if (!d_) return;
  Impl &d = *d_;
  if (d.graph_inflight && d.ev_step) (void)hipEventSynchronize(d.ev_step);
  (void)hipStreamSynchronize(d.stream);
  for (hipGraphExec_t g : {d.g_head, d.g_chunk, d.g_tail})
    if (g) (void)hipGraphExecDestroy(g);
  if (d.cap_stream) (void)hipStreamDestroy(d.cap_stream);
  if (d.ev_step) (void)hipEventDestroy(d.ev_step);
  if (d.mail_h) (void)hipHostFree(d.mail_h);
  if (d.h_scal) (void)hipHostFree(d.h_scal);
  if (d.h_scal_s) (void)hipHostFree(d.h_scal_s);
}

// Device bytes a context holds once it has stepped (mesh, state, Jacobian,
// preconditioner, Krylov vectors, reduction partials).
size_t Ctx::footprint(const Mesh &m, int precision) {
  const size_t ncv = m.ncv, nfa = m.nfa, nfa_b = m.nfa_b, nnz = m.nbocv_s();
  const size_t ent = 2 * (nfa - nfa_b) + nfa_b;
  const size_t sJ = precision == 0 ? 8 : 4;
  size_t b = 0;
  b += nfa * (8 + 3) * 8;                          // fgeo, fnorm
  b += nnz * (3 * 8 + 4) + nfa_b * (3 * 8 + 4 + 4);  // lsg, nbocv_v, bfa_lsg, bfa_cell, bfa_zone
  b += (ncv + 1) * 4 * 2 + ncv * 4;                // cfl_i, nbocv_i, wave table
  b += ent * (8 + 4 + 8 + 24);                     // entries, their blocks, their cells and positions, limiter offsets
  b += ncv * (5 + 8 + 16 + 5 + 1 + 1) * 8;         // q, prim, grad, rhs, dt, vol
  b += nfa_b * (8 + 5) * 8;                        // bfa, hook
  b += (nnz + ncv) * 25 * sJ + 10 * ncv * 5 * sJ;  // A, LU, Krylov vectors
  b += 8 * ((size_t)nblk(5 * (long long)ncv) + 64) * 8;
  return b + (1 << 20);
}

size_t Ctx::footprint(const Mesh &m, int precision, bool sa) {
  size_t b = footprint(m, precision);
  if (sa) {
    const size_t ncv = m.ncv, nnz = m.nbocv_s(), ent = 2 * (size_t)(m.nfa - m.nfa_b) + m.nfa_b;
    b += ncv * (1 + kAuxN + 1 + 10) * 8 + nnz * 6 * 8 + ent * 4;  // nuSA, fields, rhs, Krylov, AScal, faocv map
  }
  return b;
}

// One pseudo-time step driven entirely by the device, captured once into a
// hipGraph and replayed; results are identical to step()'s.
void Ctx::device_step_async() {
  if (d_->par.sa) throw Unsupported("the graph-captured step does not carry the Spalart-Allmaras scalar");
  if (d_->par.sst) throw Unsupported("the graph-captured step does not carry the SST scalars");
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  if (d.begun) throw std::logic_error("device step between step_begin and step_end");
  if (d.graph_inflight) throw std::logic_error("a device step is already in flight");
  if (d.precision == 0) {
    ensure(d.A64, 25 * (size_t)d.nnz);
    ensure(d.LU64, 25 * (size_t)d.ncv);
    for (int k = 0; k < 10; k++) ensure(d.v64[k], 5 * (size_t)d.ncv);
  } else {
    ensure(d.A32, 25 * (size_t)d.nnz);
    ensure(d.LU32, 25 * (size_t)d.ncv);
    for (int k = 0; k < 10; k++) ensure(d.v32[k], 5 * (size_t)d.ncv);
  }
  if (!d.mail_h) {
    HIPCHECK(hipHostMalloc((void **)&d.mail_h, sizeof(StepMail), hipHostMallocMapped));
    std::memset(d.mail_h, 0, sizeof(StepMail));
    HIPCHECK(hipHostGetDevicePointer((void **)&d.mail_d, d.mail_h, 0));
    HIPCHECK(hipEventCreateWithFlags(&d.ev_step, hipEventDisableTiming));
  }
  if (d.step_no == d.step0) d.launch_state_bc();
  if (d.dev_step_dirty) {
    hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, d.stream, d.scal.p, d.step_no, d.cfl);
    d.dev_step_dirty = false;
  }
  if (!d.g_head) {
    if (d.precision == 0) {
      d.g_head = d.capture(&Ctx::Impl::enqueue_step_head<double>);
      d.g_chunk = d.capture(&Ctx::Impl::enqueue_step_chunk<double>);
      d.g_tail = d.capture(&Ctx::Impl::enqueue_step_tail<double>);
    } else {
      d.g_head = d.capture(&Ctx::Impl::enqueue_step_head<float>);
      d.g_chunk = d.capture(&Ctx::Impl::enqueue_step_chunk<float>);
      d.g_tail = d.capture(&Ctx::Impl::enqueue_step_tail<float>);
    }
  }
  HIPCHECK(hipGraphLaunch(d.g_head, d.stream));
  HIPCHECK(hipEventRecord(d.ev_step, d.stream));
  d.dev_phase = 1;
  d.graph_inflight = true;
}

// Advances an in-flight device step: after the head or a chunk, launches
// another chunk or the tail depending on the solver flag in the report;
// true once the tail has finished.
bool Ctx::device_step_ready() {
  Impl &d = *d_;
  if (!d.graph_inflight) return true;
  hipError_t e = hipEventQuery(d.ev_step);
  if (e == hipErrorNotReady) return false;
  HIPCHECK(e);
  if (d.dev_phase == 2) return true;
  const volatile StepMail *m = d.mail_h;
  if (m->solver_stop) {
    HIPCHECK(hipGraphLaunch(d.g_tail, d.stream));
    d.dev_phase = 2;
  } else {
    HIPCHECK(hipGraphLaunch(d.g_chunk, d.stream));
  }
  HIPCHECK(hipEventRecord(d.ev_step, d.stream));
  return false;
}

StepInfo Ctx::device_step_result() {
  Impl &d = *d_;
  if (!d.graph_inflight) throw std::logic_error("no device step in flight");
  while (!device_step_ready()) HIPCHECK(hipEventSynchronize(d.ev_step));
  d.graph_inflight = false;
  d.dev_phase = 0;
  const volatile StepMail *m = d.mail_h;
  StepInfo info;
  info.step = m->step;
  info.lin_iters = m->lin_iters;
  info.lin_status = m->lin_status;
  info.cfl = m->cfl;
  info.dt_min = m->dt_min;
  for (int k = 0; k < 5; k++) info.resid[k] = m->resid[k];
  for (int k = 0; k < 5; k++) info.resid_l2[k] = m->resid_l2[k];
  d.step_no = info.step;
  d.cfl = info.cfl;
  return info;
}

StepInfo Ctx::device_step() {
  device_step_async();
  return device_step_result();
}

KrylovBatchBuf::~KrylovBatchBuf() {
  if (jobs_d) (void)hipFree(jobs_d);
  if (map_d) (void)hipFree(map_d);
  if (partials_d) (void)hipFree(partials_d);
  if (flag_h) (void)hipHostFree(flag_h);
}

// The batched BiCGStab of njobs solves of one precision. Setup (not
// capturable): block offsets of each solve in the three launch shapes, the
// block -> solve maps, and the job table on the device.
void setup_krylov_batch(KrylovJob *jobs_h, int njobs, void *stream, KrylovBatchBuf &buf) {
  hipStream_t st = (hipStream_t)stream;
  int g5 = 0, g1 = 0, gx = 0, max_it = 0;
  for (int j = 0; j < njobs; j++) {
    KrylovJob &J = jobs_h[j];
    J.b5 = g5;
    J.nb5 = (J.ncv + kCellsPerBlock - 1) / kCellsPerBlock;
    J.b1 = g1;
    J.nb1 = nblk(J.ncv);
    J.bx = gx;
    J.nbx = nblk(5 * (long long)J.ncv);
    g5 += J.nb5;
    g1 += J.nb1;
    gx += J.nbx;
    max_it = std::max(max_it, J.maxiter + 1);
  }
  buf.njobs = njobs;
  buf.g5 = g5;
  buf.g1 = g1;
  buf.gx = gx;
  buf.max_it = max_it;
  if (njobs <= 0) return;
  std::vector<int> map((size_t)g5 + g1 + gx);
  for (int j = 0; j < njobs; j++) {
    const KrylovJob &J = jobs_h[j];
    for (int b = 0; b < J.nb5; b++) map[J.b5 + b] = j;
    for (int b = 0; b < J.nb1; b++) map[(size_t)g5 + J.b1 + b] = j;
    for (int b = 0; b < J.nbx; b++) map[(size_t)g5 + g1 + J.bx + b] = j;
  }
  if (buf.cap_jobs < (size_t)njobs) {
    if (buf.jobs_d) HIPCHECK(hipFree(buf.jobs_d));
    buf.cap_jobs = std::max<size_t>(njobs, 2 * buf.cap_jobs);
    HIPCHECK(hipMalloc((void **)&buf.jobs_d, sizeof(KrylovJob) * buf.cap_jobs));
  }
  if (buf.cap_map < map.size()) {
    if (buf.map_d) HIPCHECK(hipFree(buf.map_d));
    buf.cap_map = std::max(map.size(), 2 * buf.cap_map);
    HIPCHECK(hipMalloc((void **)&buf.map_d, sizeof(int) * buf.cap_map));
  }
  if (buf.cap_partials < 2 * (size_t)g5 + 8) {
    if (buf.partials_d) HIPCHECK(hipFree(buf.partials_d));
    buf.cap_partials = std::max<size_t>(2 * (size_t)g5 + 8, 2 * buf.cap_partials);
    HIPCHECK(hipMalloc((void **)&buf.partials_d, sizeof(double) * buf.cap_partials));
  }
  if (!buf.flag_h) {
    HIPCHECK(hipHostMalloc((void **)&buf.flag_h, sizeof(int), hipHostMallocMapped));
    HIPCHECK(hipHostGetDevicePointer((void **)&buf.flag_d, buf.flag_h, 0));
  }
  HIPCHECK(hipMemcpyAsync(buf.jobs_d, jobs_h, sizeof(KrylovJob) * njobs, hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(buf.map_d, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
}

template <typename T>
static void krylov_batch_prologue(hipStream_t st, const KrylovBatchBuf &buf) {
  const int *m5 = buf.map_d, *m1 = buf.map_d + buf.g5;
  hipLaunchKernelGGL(kb_lu<T>, dim3(buf.g1), dim3(kBlock), 0, st, buf.jobs_d, m1);
  hipLaunchKernelGGL(kb_resid<T>, dim3(buf.g5), dim3(kBlock), 0, st, buf.jobs_d, m5, 1, buf.partials_d);
  hipLaunchKernelGGL(kb_fin, dim3(buf.njobs), dim3(kFinBlock), 0, st, buf.jobs_d, (int)kFinResid0, buf.partials_d,
                     buf.g5);
}

template <typename T>
static void krylov_batch_iterations(hipStream_t st, const KrylovBatchBuf &buf, int count) {
  const int *m5 = buf.map_d, *m1 = buf.map_d + buf.g5, *mx = buf.map_d + buf.g5 + buf.g1;
  const KrylovJob *jd = buf.jobs_d;
  double *pt = buf.partials_d;
  const int g5 = buf.g5, g1 = buf.g1, gx = buf.gx, nj = buf.njobs;
  for (int k = 0; k < count; k++) {  // seven launches: alpha and omega are finished in kb_ps and kb_x
    hipLaunchKernelGGL(kb_ps<T>, dim3(g1), dim3(kBlock), 0, st, jd, m1, 0, pt, g5);
    hipLaunchKernelGGL(kb_spmv_dot<T>, dim3(g5), dim3(kBlock), 0, st, jd, m5, 0, pt);
    hipLaunchKernelGGL(kb_ps<T>, dim3(g1), dim3(kBlock), 0, st, jd, m1, 1, pt, g5);
    hipLaunchKernelGGL(kb_spmv_dot<T>, dim3(g5), dim3(kBlock), 0, st, jd, m5, 1, pt);
    hipLaunchKernelGGL(kb_x<T>, dim3(gx), dim3(kBlock), 0, st, jd, mx, pt, g5);
    // FP32 solves take the residual by the recurrence (GPU_KRYLOV_RESIDUAL = RECURRENCE, the server's rule)
    if (std::is_same<T, float>::value)
      hipLaunchKernelGGL((kb_resid<T, true>), dim3(g5), dim3(kBlock), 0, st, jd, m5, 0, pt);
    else
      hipLaunchKernelGGL(kb_resid<T>, dim3(g5), dim3(kBlock), 0, st, jd, m5, 0, pt);
    hipLaunchKernelGGL(kb_fin, dim3(nj), dim3(kFinBlock), 0, st, jd, (int)kFinResid, pt, g5);
  }
}

// The batch set up last, with host polling between chunks of iterations: the
// first chunk is the previous call's count less two (consecutive steps of the
// same solves take similar counts), then chunks of two. Launches after a
// solve's stop return at once, so the chunking changes no result.
void run_krylov_batch(int precision, void *stream, KrylovBatchBuf &buf) {
  if (buf.njobs <= 0) return;
  hipStream_t st = (hipStream_t)stream;
  if (precision == 0) krylov_batch_prologue<double>(st, buf);
  else krylov_batch_prologue<float>(st, buf);
  int launched = 0, chunk = std::max(2, buf.last_launched - 2);
  while (launched < buf.max_it) {
    chunk = std::min(chunk, buf.max_it - launched);
    if (precision == 0) krylov_batch_iterations<double>(st, buf, chunk);
    else krylov_batch_iterations<float>(st, buf, chunk);
    launched += chunk;
    hipLaunchKernelGGL(kb_all_done, dim3(1), dim3(256), 0, st, buf.jobs_d, buf.njobs, buf.flag_d);
    HIPCHECK(hipStreamSynchronize(st));
    if (*(volatile int *)buf.flag_h) break;
    chunk = 2;
  }
  buf.last_launched = launched;
  HIPCHECK(hipGetLastError());
}

int Ctx::precision() const { return d_->precision; }

uint32_t Ctx::batch_key() const {
  const Impl &d = *d_;
  const bool grad = d.par.second_order || d.dp.visc;
  const bool ent = grad && d.par.grad_kernel == kResidEntry && !d.dp.visc;
  return (uint32_t)d.precision | (uint32_t)d.par.scheme << 1 | (uint32_t)(d.par.second_order ? 1 : 0) << 3 |
         (uint32_t)(d.par.grad_ls ? 1 : 0) << 4 | (uint32_t)(d.dp.visc ? 1 : 0) << 5 | (uint32_t)(ent ? 1 : 0) << 6;
}

struct StepBatch::Impl {
  StepJob *jobs_d = nullptr;
  int *map_d = nullptr;
  size_t cap_jobs = 0, cap_map = 0;
  int njobs = 0;
  int tot[kShapes] = {};
  size_t off[kShapes] = {};
  int precision = 0, scheme = 0;
  bool snd = false, ls = false, visc = false, entry = false;
  ~Impl() {
    if (jobs_d) (void)hipFree(jobs_d);
    if (map_d) (void)hipFree(map_d);
  }
};

StepBatch::StepBatch() : d_(new Impl) {}
StepBatch::~StepBatch() = default;
int StepBatch::size() const { return d_->njobs; }

void StepBatch::set(Ctx *const *ctxs, int n, void *stream) {
  Impl &B = *d_;
  hipStream_t st = (hipStream_t)stream;
  B.njobs = 0;
  if (n <= 0) return;
  const uint32_t key = ctxs[0]->batch_key();
  std::vector<StepJob> jobs(n);
  int tot[kShapes] = {};
  for (int j = 0; j < n; j++) {
    if (ctxs[j]->batch_key() != key) throw std::logic_error("a step batch mixes kernel configurations");
    Ctx::Impl &d = *ctxs[j]->d_;
    StepJob &J = jobs[j];
    std::memset(&J, 0, sizeof J);
    J.dm = d.dm;
    J.dp = d.dp;
    J.prim = d.prim.p;
    J.bfa = d.bfa.p;
    J.grad = d.grad.p;
    J.gradv = d.gradv.p;
    J.q = d.q.p;
    J.rhs = d.rhs.p;
    J.dt = d.dt.p;
    J.partials = d.partials.p;
    J.hook = d.hook.p;
    J.counters = d.counters.p;
    if (d.precision == 0) {
      J.A = d.A64.p;
      J.rhs_ls = d.v64[0].p;
      J.phi = d.v64[1].p;
    } else {
      J.A = d.A32.p;
      J.rhs_ls = d.v32[0].p;
      J.phi = d.v32[1].p;
    }
    J.scal = d.scal.p;
    J.mail = d.mail_d;
    J.dta.cfl = d.cfl;
    J.dta.dt_fixed = d.par.dt_fixed;
    J.dta.mode = d.par.dt_mode;
    J.dta.sum = d.par.dt_sum ? 1 : 0;
    J.dta.cfl_dev = &d.scal.p->cfl_cur;
    J.nA = 25 * (size_t)d.nnz;
    J.gamma = d.par.gamma;
    J.relax = d.par.under_relax;
    J.cfl_factor = d.par.cfl_factor;
    J.cfl_max = d.par.cfl_max;
    J.ls_abs = d.par.ls_abs;
    J.ls_rel = d.par.ls_rel;
    J.ncv = d.ncv;
    J.cfl_after = d.par.cfl_after;
    J.cfl_interval = d.par.cfl_interval;
    J.ls_maxiter = d.par.ls_maxiter;
    // the grids of the solve's own launches (launch_grad, launch_flux,
    // launch_jac, launch_state_bc)
    const bool grad = d.par.second_order || d.dp.visc;
    const bool ent = d.par.grad_kernel == kResidEntry && !d.dp.visc;
    J.n[kShCell] = nblk(d.ncv);
    J.n[kShGrad] = !grad ? 0 : ent ? (d.nwave + kWaves - 1) / kWaves : nblk(d.ncv);
    // the batched flux kernel runs blocks of kWaves wavefronts
    J.n[kShFlux] = d.flux_entry() ? (d.nwave + kWaves - 1) / kWaves : nblk(d.ncv);
    J.n[kShBc] = d.nfa_b > 0 ? nblk(d.nfa_b) : 0;
    J.n[kShFill] = d.dup_blocks ? (int)((J.nA + kBlock - 1) / kBlock) : 0;
    J.n[kShDtFill] = d.par.dt_mode == 1 ? nblk(d.ncv) : 0;
    for (int k = 0; k < kShapes; k++) {
      J.b[k] = tot[k];
      tot[k] += J.n[k];
    }
    if (j == 0) {
      B.precision = d.precision;
      B.scheme = d.par.scheme;
      B.snd = d.par.second_order;
      B.ls = d.par.grad_ls;
      B.visc = d.dp.visc != 0;
      B.entry = grad && ent;
    }
  }
  size_t nmap = 0;
  for (int k = 0; k < kShapes; k++) {
    B.off[k] = nmap;
    B.tot[k] = tot[k];
    nmap += tot[k];
  }
  std::vector<int> map(nmap);
  for (int j = 0; j < n; j++)
    for (int k = 0; k < kShapes; k++)
      for (int b = 0; b < jobs[j].n[k]; b++) map[B.off[k] + jobs[j].b[k] + b] = j;
  if (B.cap_jobs < (size_t)n) {
    if (B.jobs_d) HIPCHECK(hipFree(B.jobs_d));
    B.jobs_d = nullptr;
    B.cap_jobs = std::max<size_t>(n, 2 * B.cap_jobs);
    HIPCHECK(hipMalloc((void **)&B.jobs_d, sizeof(StepJob) * B.cap_jobs));
  }
  if (B.cap_map < nmap) {
    if (B.map_d) HIPCHECK(hipFree(B.map_d));
    B.map_d = nullptr;
    B.cap_map = std::max(nmap, 2 * B.cap_map);
    HIPCHECK(hipMalloc((void **)&B.map_d, sizeof(int) * B.cap_map));
  }
  HIPCHECK(hipMemcpyAsync(B.jobs_d, jobs.data(), sizeof(StepJob) * n, hipMemcpyHostToDevice, st));
  HIPCHECK(hipMemcpyAsync(B.map_d, map.data(), sizeof(int) * nmap, hipMemcpyHostToDevice, st));
  HIPCHECK(hipStreamSynchronize(st));
  B.njobs = n;
}

template <typename T>
static void step_batch_jac(const StepBatch::Impl &B, hipStream_t st) {
  const StepJob *jd = B.jobs_d;
  if (B.tot[kShFill])
    hipLaunchKernelGGL(ks_fill<T>, dim3(B.tot[kShFill]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShFill]);
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(B.tot[kShCell]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShCell]);
  };
  if (B.visc) {
    if (B.scheme == kRoe) go(ks_jac<T, kRoe, true>);
    else if (B.scheme == kHllcC) go(ks_jac<T, kHllcC, true>);
    else go(ks_jac<T, kHllc, true>);
  } else {
    if (B.scheme == kRoe) go(ks_jac<T, kRoe, false>);
    else if (B.scheme == kHllcC) go(ks_jac<T, kHllcC, false>);
    else go(ks_jac<T, kHllc, false>);
  }
}

void StepBatch::pre(void *stream) const {
  const Impl &B = *d_;
  if (B.njobs == 0) return;
  hipStream_t st = (hipStream_t)stream;
  const StepJob *jd = B.jobs_d;
  hipLaunchKernelGGL(ks_ramp, dim3(B.njobs), dim3(64), 0, st, jd);
  if (B.tot[kShGrad]) {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(B.tot[kShGrad]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShGrad]);
    };
    if (B.entry) {
      if (B.ls) go(ks_grad<true, false, true>);
      else go(ks_grad<false, false, true>);
    } else if (B.visc) {
      if (B.ls) go(ks_grad<true, true, false>);
      else go(ks_grad<false, true, false>);
    } else {
      if (B.ls) go(ks_grad<true, false, false>);
      else go(ks_grad<false, false, false>);
    }
  }
  {
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(B.tot[kShFlux]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShFlux]);
    };
    if (B.visc) {
      if (B.snd) {
        if (B.scheme == kRoe) go(ks_flux<kRoe, true, true>);
        else if (B.scheme == kAusm) go(ks_flux<kAusm, true, true>);
        else if (B.scheme == kHllcC) go(ks_flux<kHllcC, true, true>);
        else go(ks_flux<kHllc, true, true>);
      } else {
        if (B.scheme == kRoe) go(ks_flux<kRoe, false, true>);
        else if (B.scheme == kAusm) go(ks_flux<kAusm, false, true>);
        else if (B.scheme == kHllcC) go(ks_flux<kHllcC, false, true>);
        else go(ks_flux<kHllc, false, true>);
      }
    } else {
      if (B.snd) {
        if (B.scheme == kRoe) go(ks_flux<kRoe, true, false>);
        else if (B.scheme == kAusm) go(ks_flux<kAusm, true, false>);
        else if (B.scheme == kHllcC) go(ks_flux<kHllcC, true, false>);
        else go(ks_flux<kHllc, true, false>);
      } else {
        if (B.scheme == kRoe) go(ks_flux<kRoe, false, false>);
        else if (B.scheme == kAusm) go(ks_flux<kAusm, false, false>);
        else if (B.scheme == kHllcC) go(ks_flux<kHllcC, false, false>);
        else go(ks_flux<kHllc, false, false>);
      }
    }
  }
  hipLaunchKernelGGL(ks_fin, dim3(B.njobs), dim3(kFinBlock), 0, st, jd, (int)kFinDt, (int)kShFlux);
  if (B.tot[kShDtFill])
    hipLaunchKernelGGL(ks_dt_fill, dim3(B.tot[kShDtFill]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShDtFill]);
  if (B.precision == 0) step_batch_jac<double>(B, st);
  else step_batch_jac<float>(B, st);
  hipLaunchKernelGGL(ks_solver_init, dim3(B.njobs), dim3(64), 0, st, jd);
  HIPCHECK(hipGetLastError());
}

void StepBatch::post(void *stream) const {
  const Impl &B = *d_;
  if (B.njobs == 0) return;
  hipStream_t st = (hipStream_t)stream;
  const StepJob *jd = B.jobs_d;
  const int *mc = B.map_d + B.off[kShCell];
  const dim3 gc(B.tot[kShCell]);
  if (B.precision == 0) hipLaunchKernelGGL(ks_update<double>, gc, dim3(kBlock), 0, st, jd, mc);
  else hipLaunchKernelGGL(ks_update<float>, gc, dim3(kBlock), 0, st, jd, mc);
  hipLaunchKernelGGL(ks_state, gc, dim3(kBlock), 0, st, jd, mc);
  if (B.tot[kShBc])
    hipLaunchKernelGGL(ks_bc, dim3(B.tot[kShBc]), dim3(kBlock), 0, st, jd, B.map_d + B.off[kShBc]);
  hipLaunchKernelGGL(ks_l1, gc, dim3(kBlock), 0, st, jd, mc);
  hipLaunchKernelGGL(ks_fin, dim3(B.njobs), dim3(kFinBlock), 0, st, jd, (int)kFinL1, (int)kShCell);
  hipLaunchKernelGGL(ks_publish, dim3(B.njobs), dim3(64), 0, st, jd);
  HIPCHECK(hipGetLastError());
}

void Ctx::batch_prepare() {
  Impl &d = *d_;
  if (d.par.sa) throw Unsupported("batched serving does not carry the Spalart-Allmaras scalar");
  if (d.par.sst) throw Unsupported("batched serving does not carry the SST scalars");
  if (!d.flux_entry() || d.face_form())
    throw Unsupported("batched serving runs the per-entry flux kernel (GPU_FLUX_KERNEL = ENTRY)");
  if (!d.have_state) throw std::logic_error("no state set");
  if (d.precision == 0) {
    ensure(d.A64, 25 * (size_t)d.nnz);
    ensure(d.LU64, 25 * (size_t)d.ncv);
    for (int k = 0; k < 10; k++) ensure(d.v64[k], 5 * (size_t)d.ncv);
  } else {
    ensure(d.A32, 25 * (size_t)d.nnz);
    ensure(d.LU32, 25 * (size_t)d.ncv);
    for (int k = 0; k < 10; k++) ensure(d.v32[k], 5 * (size_t)d.ncv);
  }
  if (!d.mail_h) {
    HIPCHECK(hipHostMalloc((void **)&d.mail_h, sizeof(StepMail), hipHostMallocMapped));
    std::memset(d.mail_h, 0, sizeof(StepMail));
    HIPCHECK(hipHostGetDevicePointer((void **)&d.mail_d, d.mail_h, 0));
  }
  if (d.step_no == d.step0) d.launch_state_bc();
  hipLaunchKernelGGL(k_set_step, dim3(1), dim3(64), 0, d.stream, d.scal.p, d.step_no, d.cfl);
  d.dev_step_dirty = false;
  HIPCHECK(hipStreamSynchronize(d.stream));
}

void Ctx::batch_job(KrylovJob &job) {
  Impl &d = *d_;
  job.ncv = d.ncv;
  job.maxiter = d.par.ls_maxiter;
  job.nbocv_i = d.nbocv_i.p;
  job.nbocv_v = d.nbocv_v.p;
  job.S = d.scal.p;
  if (d.precision == 0) {
    job.A = d.A64.p;
    job.LU = d.LU64.p;
    for (int k = 0; k < 10; k++) job.v[k] = d.v64[k].p;
  } else {
    job.A = d.A32.p;
    job.LU = d.LU32.p;
    for (int k = 0; k < 10; k++) job.v[k] = d.v32[k].p;
  }
}

StepInfo Ctx::batch_result() {
  Impl &d = *d_;
  const volatile StepMail *m = d.mail_h;
  StepInfo info;
  info.step = m->step;
  info.lin_iters = m->lin_iters;
  info.lin_status = m->lin_status;
  info.cfl = m->cfl;
  info.dt_min = m->dt_min;
  for (int k = 0; k < 5; k++) info.resid[k] = m->resid[k];
  for (int k = 0; k < 5; k++) info.resid_l2[k] = m->resid_l2[k];
  d.step_no = info.step;
  d.cfl = info.cfl;
  return info;
}

int Ctx::ncv() const { return d_->ncv; }

void Ctx::set_state(const double *rho, const double *rhou, const double *rhoE) {
  Impl &d = *d_;
  const size_t n = d.ncv;
  std::vector<double> h(5 * n);
  for (size_t cd = 0; cd < n; cd++) {
    const size_t c = d.perm[cd];
    h[cd] = rho[c];
    h[n + cd] = rhou[3 * c];
    h[2 * n + cd] = rhou[3 * c + 1];
    h[3 * n + cd] = rhou[3 * c + 2];
    h[4 * n + cd] = rhoE[c];
  }
  HIPCHECK(hipMemcpyAsync(d.q.p, h.data(), sizeof(double) * 5 * n, hipMemcpyHostToDevice, d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
  d.have_state = true;
  d.step_no = d.step0 = 0;
  d.cfl = d.par.cfl;
  d.dev_step_dirty = true;
}

// a solve resumed from a restart file continues its step count (the reference
// keeps the STEP it read: runBackwardEuler does not reset it); the CFL starts
// again from Joe.in's
void Ctx::set_step(int step) {
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  if (step < 0) throw std::invalid_argument("step must be non-negative");
  d.step_no = d.step0 = step;
  d.dev_step_dirty = true;
}

void Ctx::get_state(double *rho, double *rhou, double *rhoE) {
  Impl &d = *d_;
  const size_t n = d.ncv;
  std::vector<double> h(5 * n);
  HIPCHECK(hipMemcpyAsync(h.data(), d.q.p, sizeof(double) * 5 * n, hipMemcpyDeviceToHost, d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
  for (size_t cd = 0; cd < n; cd++) {
    const size_t c = d.perm[cd];
    rho[c] = h[cd];
    rhou[3 * c] = h[n + cd];
    rhou[3 * c + 1] = h[2 * n + cd];
    rhou[3 * c + 2] = h[3 * n + cd];
    rhoE[c] = h[4 * n + cd];
  }
}

// initialHook, uniform branch (JoeWithModels.cpp:156-186)
void Ctx::init_uniform() {
  Impl &d = *d_;
  const size_t n = d.ncv;
  std::vector<double> rho(n), rhou(3 * n), rhoE(n);
  const double *u = d.par.u_init;
  double uu = 0.0;
  uu += u[0] * u[0];
  uu += u[1] * u[1];
  uu += u[2] * u[2];
  for (size_t c = 0; c < n; c++) {
    rho[c] = d.par.rho_init;
    for (int i = 0; i < 3; i++) rhou[3 * c + i] = d.par.rho_init * u[i];
    rhoE[c] = d.par.p_init / (d.par.gamma - 1.0) + 0.5 * d.par.rho_init * uu;
  }
  set_state(rho.data(), rhou.data(), rhoE.data());
  if (d.par.sa) {  // uniform nuSA (JoeWithModels.cpp:219-240)
    std::vector<double> ph(n, d.par.sa_init);
    set_scalar(ph.data());
  }
  if (d.par.sst)  // uniform kine and omega
    for (int s = 0; s < 2; s++) {
      std::vector<double> ph(n, d.par.sst_init[s]);
      set_scalar(ph.data(), s);
    }
}

void Ctx::set_scalar(const double *phi, int s) {
  Impl &d = *d_;
  if (!d.par.sa && !d.par.sst) throw std::logic_error("no transported scalar (TURB_MODEL = SA or SST)");
  if (s < 0 || s > (d.par.sst ? 1 : 0)) throw std::invalid_argument("no such transported scalar");
  const size_t n = d.ncv;
  std::vector<double> h(n);
  for (size_t cd = 0; cd < n; cd++) h[cd] = phi[d.perm[cd]];
  double *dst = d.par.sst ? d.sst_phi.p + (size_t)s * n : d.sa_phi.p;
  HIPCHECK(hipMemcpyAsync(dst, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
  d.dev_step_dirty = true;
}

void Ctx::Impl::launch_state_bc() {
  if (par.sst)
    hipLaunchKernelGGL(k_state_sst, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, q.p, prim.p, par.gamma, counters.p,
                       sst_phi.p);
  else
    hipLaunchKernelGGL(k_state, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, q.p, prim.p, par.gamma, counters.p);
  if (nfa_b > 0) {
    hipLaunchKernelGGL(k_bc, dim3(nblk(nfa_b)), dim3(kBlock), 0, stream, dm, dp, prim.p, hook.p, bfa.p);
    if (par.sst)  // kine and omega on the boundary faces (setScalarBC)
      hipLaunchKernelGGL(k_sst_bc, dim3(nblk(nfa_b)), dim3(kBlock), 0, stream, dm, dp, prim.p, bfa.p, sst_bfa.p);
  }
}

void Ctx::Impl::launch_grad(hipStream_t s) {
  if (!par.second_order && !dp.visc) return;
  // the per-entry gradient kernel has no viscous variant
  const bool ent = par.grad_kernel == kResidEntry && !dp.visc;
  const dim3 g(ent ? (nwave + kGradWaves - 1) / kGradWaves : nblk(ncv));
  if (ent) {
    const dim3 b(64 * kGradWaves);
    if (par.grad_ls) hipLaunchKernelGGL(k_grad_e<true>, g, b, 0, s, dm, dp, prim.p, bfa.p, grad.p);
    else hipLaunchKernelGGL(k_grad_e<false>, g, b, 0, s, dm, dp, prim.p, bfa.p, grad.p);
  } else {
    if (dp.visc) {
      if (par.grad_ls) hipLaunchKernelGGL((k_grad<true, true>), g, dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, gradv.p);
      else hipLaunchKernelGGL((k_grad<false, true>), g, dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, gradv.p);
    } else {
      if (par.grad_ls) hipLaunchKernelGGL((k_grad<true, false>), g, dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, gradv.p);
      else hipLaunchKernelGGL((k_grad<false, false>), g, dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, gradv.p);
    }
  }
  if (par.sa) {  // nuSA gradient, face-averaged mul, strain and vorticity
    if (par.grad_ls)
      hipLaunchKernelGGL(k_sa_cell<true>, dim3(nblk(ncv)), dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, sa_aux.p);
    else
      hipLaunchKernelGGL(k_sa_cell<false>, dim3(nblk(ncv)), dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p, sa_aux.p);
  }
  if (par.sst) {  // kine and omega gradients, strain, vorticity, divergence, cross diffusion, F1, F2
    if (par.grad_ls)
      hipLaunchKernelGGL(k_sst_cell<true>, dim3(nblk(ncv)), dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p,
                         sst_aux.p);
    else
      hipLaunchKernelGGL(k_sst_cell<false>, dim3(nblk(ncv)), dim3(kBlock), 0, s, dm, dp, prim.p, bfa.p, grad.p,
                         sst_aux.p);
  }
}

// the per-cell flux kernel has no viscous variant
bool Ctx::Impl::flux_entry() const { return par.flux_kernel != kResidCell || dp.visc; }

// the local time step's minimum over the flux blocks' partials into
// SolverScalars::dt_min; many partials (one per flux wavefront) are first
// reduced in parallel
void Ctx::Impl::launch_dt_fin() {
  const int nb = flux_blocks();
  if (nb > 4 * kMinPre) {
    const int nb2 = (nb + kMinPre - 1) / kMinPre;
    hipLaunchKernelGGL(k_min_pre, dim3(nb2), dim3(256), 0, stream, partials.p, nb, partials.p + nb);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb2, (int)kFinDt, partials.p + nb, scal.p);
  } else {
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinDt, partials.p, scal.p);
  }
}

int Ctx::Impl::flux_blocks() const {
  if (face_form()) return (nwave + kWaves - 1) / kWaves;  // k_gather_e
  return flux_entry() ? (nwave + kFluxWaves - 1) / kFluxWaves : nblk(ncv);
}

void Ctx::Impl::launch_flux(hipStream_t s, const double *cfl_dev) {
  DtArgs dta;
  dta.cfl = cfl;
  dta.dt_fixed = par.dt_fixed;
  dta.mode = par.dt_mode;
  dta.sum = par.dt_sum ? 1 : 0;
  dta.cfl_dev = cfl_dev;
  const dim3 g(flux_blocks());
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, dim3(kBlock), 0, s, dm, dp, prim.p, grad.p, bfa.p, q.p, rhs.p, counters.p + 1, dta,
                       dt.p, partials.p);
  };
  auto goe = [&](auto kern) {
    hipLaunchKernelGGL(kern, g, dim3(64 * kFluxWaves), 0, s, dm, dp, prim.p, grad.p, bfa.p, q.p, rhs.p,
                       counters.p + 1, dta, dt.p, partials.p, gradv.p);
  };
  if (face_form()) {
    const dim3 gf(nblk(nfa - nfa_b));
    auto gof = [&](auto fk, auto gk) {
      if (nfa > nfa_b) hipLaunchKernelGGL(fk, gf, dim3(kBlock), 0, s, dm, dp, prim.p, grad.p, fcell.p, fflux.p, counters.p + 1);
      hipLaunchKernelGGL(gk, g, dim3(kBlock), 0, s, dm, dp, prim.p, grad.p, bfa.p, q.p, fflux.p, rhs.p, counters.p + 1,
                         dta, dt.p, partials.p);
    };
    if (par.second_order) {
      if (par.scheme == kRoe) gof(k_face_flux<kRoe, true>, k_gather_e<kRoe, true>);
      else if (par.scheme == kAusm) gof(k_face_flux<kAusm, true>, k_gather_e<kAusm, true>);
      else gof(k_face_flux<kHllc, true>, k_gather_e<kHllc, true>);
    } else {
      if (par.scheme == kRoe) gof(k_face_flux<kRoe, false>, k_gather_e<kRoe, false>);
      else if (par.scheme == kAusm) gof(k_face_flux<kAusm, false>, k_gather_e<kAusm, false>);
      else gof(k_face_flux<kHllc, false>, k_gather_e<kHllc, false>);
    }
    return;
  }
  if (flux_entry()) {
    if (par.sst) {  // HLLC, viscous (the parser requires both)
      if (par.second_order) goe(k_flux_e<kHllc, true, true, true>);
      else goe(k_flux_e<kHllc, false, true, true>);
      return;
    }
    if (dp.visc) {
      if (par.second_order) {
        if (par.scheme == kRoe) goe(k_flux_e<kRoe, true, true>);
        else if (par.scheme == kAusm) goe(k_flux_e<kAusm, true, true>);
        else if (par.scheme == kHllcC) goe(k_flux_e<kHllcC, true, true>);
        else goe(k_flux_e<kHllc, true, true>);
      } else {
        if (par.scheme == kRoe) goe(k_flux_e<kRoe, false, true>);
        else if (par.scheme == kAusm) goe(k_flux_e<kAusm, false, true>);
        else if (par.scheme == kHllcC) goe(k_flux_e<kHllcC, false, true>);
        else goe(k_flux_e<kHllc, false, true>);
      }
    } else {
      if (par.second_order) {
        if (par.scheme == kRoe) goe(k_flux_e<kRoe, true, false>);
        else if (par.scheme == kAusm) goe(k_flux_e<kAusm, true, false>);
        else if (par.scheme == kHllcC) goe(k_flux_e<kHllcC, true, false>);
        else goe(k_flux_e<kHllc, true, false>);
      } else {
        if (par.scheme == kRoe) goe(k_flux_e<kRoe, false, false>);
        else if (par.scheme == kAusm) goe(k_flux_e<kAusm, false, false>);
        else if (par.scheme == kHllcC) goe(k_flux_e<kHllcC, false, false>);
        else goe(k_flux_e<kHllc, false, false>);
      }
    }
  } else {
    if (par.second_order) {
      if (par.scheme == kRoe) go(k_flux<kRoe, true>);
      else if (par.scheme == kAusm) go(k_flux<kAusm, true>);
      else go(k_flux<kHllc, true>);
    } else {
      if (par.scheme == kRoe) go(k_flux<kRoe, false>);
      else if (par.scheme == kAusm) go(k_flux<kAusm, false>);
      else go(k_flux<kHllc, false>);
    }
  }
}

void Ctx::Impl::launch_grad_flux(bool jac_follows) {
  launch_grad(stream);
  launch_flux(stream);
  // dt_min of calcDt; CFL_MIN sets every local dt to it
  launch_dt_fin();
  if (par.dt_mode == 1) hipLaunchKernelGGL(k_dt_fill, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, dt.p, scal.p);
  if (coupled_sa()) {
    // the scalar row of the coupled system: its residual here, or with its
    // Jacobian row in launch_jac
    if (!jac_follows) launch_scal_c<double>(nullptr, false);
  } else if (par.sa) {
    // the nuSA equation: residual and AScal rows (calcRhs scalar terms)
    hipLaunchKernelGGL(k_scal, dim3(nblk(ncv)), dim3(kBlock), 0, stream, dm, dp, prim.p, grad.p, bfa.p, rhs_s.p, AS.p);
  } else if (par.sst) {
    // kine's and omega's equations: residuals and AScal rows, per entry or per cell
    if (par.scal_kernel == kResidEntry && nwave > 0)
      hipLaunchKernelGGL(k_sst_scal_e, dim3((nwave + kSstScalWaves - 1) / kSstScalWaves), dim3(64 * kSstScalWaves), 0,
                         stream, dm, dp, prim.p, grad.p, bfa.p, rhs_s.p, AS.p);
    else
      hipLaunchKernelGGL(k_sst_scal, dim3(nblk(ncv)), dim3(kBlock), 0, stream, dm, dp, prim.p, grad.p, bfa.p, rhs_s.p,
                         AS.p);
  }
}

template <typename T>
void Ctx::Impl::launch_scal_c(T *A, bool step_mode) {
  DBuf<T> *v = krylov<T>();
  T *b = step_mode ? v[0].p : nullptr, *x0 = step_mode ? v[1].p : nullptr;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblk(ncv)), dim3(kBlock), 0, stream, dm, dp, prim.p, grad.p, bfa.p, gradv.p,
                       rhs_s.p, A, step_mode ? 1 : 0, dt.p, par.under_relax, b, x0);
  };
  // one lane per entry, as the entry Jacobian (each off-diagonal block written
  // by one lane, so not with duplicate neighbour pairs); GPU_JAC_KERNEL = CELL
  // takes the per-cell form
  auto goe = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((nwave + kScalWaves - 1) / kScalWaves), dim3(64 * kScalWaves), 0, stream, dm, dp,
                       prim.p, grad.p, bfa.p, gradv.p, rhs_s.p, A, step_mode ? 1 : 0, dt.p, par.under_relax, b, x0);
  };
  const bool ent = par.jac_kernel == kResidEntry && !dup_blocks && nwave > 0;
  if (ent) {
    if (par.second_order) goe(k_scal_e<T, true>);
    else goe(k_scal_e<T, false>);
  } else {
    if (par.second_order) go(k_scal_c<T, true>);
    else go(k_scal_c<T, false>);
  }
}

template <typename T>
void Ctx::Impl::ensure_solver(int nvec, bool with_lu) {
  const size_t b = nbs;
  const bool ilu = with_lu && par.ilu_sub > 0;
  if constexpr (std::is_same<T, double>::value) {
    ensure(A64, b * b * (size_t)nnz);
    if (with_lu) ensure(LU64, b * b * (size_t)ncv);
    if (ilu) {
      ensure(F64, b * b * (size_t)nnz);
      ensure(Dinv64, b * b * (size_t)ncv);
    }
    for (int k = 0; k < nvec; k++) ensure(v64[k], b * (size_t)ncv);
  } else {
    ensure(A32, b * b * (size_t)nnz);
    if (with_lu) ensure(LU32, b * b * (size_t)ncv);
    if (ilu) {
      ensure(F32, b * b * (size_t)nnz);
      ensure(Dinv32, b * b * (size_t)ncv);
    }
    for (int k = 0; k < nvec; k++) ensure(v32[k], b * (size_t)ncv);
  }
}

// The segregated nuSA half of a step (JoeWithModels.cpp:1040-1091) after the
// mean flow took dq: right-hand side and V/dt, Jacobi BiCGStab polled by the
// host in chunks (2, then 4 iterations; iterations after convergence are
// no-ops), update and clip. Work vectors sw: b, x, res, res0, p, v, phat, s, shat, t.
template <typename T>
void Ctx::Impl::scalar_step(const T *dq) {
  double *b = sw[0].p, *x = sw[1].p, *res = sw[2].p, *res0 = sw[3].p, *p = sw[4].p, *vv = sw[5].p, *phat = sw[6].p,
         *sv = sw[7].p, *shat = sw[8].p, *tv = sw[9].p;
  const dim3 g(nblk(ncv)), bl(kBlock);
  hipLaunchKernelGGL(k_sc_prep<T>, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, vol.p, dt.p, par.relax_scal, rhs_s.p,
                     AS.p, dq, b, x);
  hipLaunchKernelGGL(k_fill<double>, g, bl, 0, stream, (size_t)ncv, p, 0.0);
  hipLaunchKernelGGL(k_fill<double>, g, bl, 0, stream, (size_t)ncv, vv, 0.0);
  hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, stream, scal_s.p, kSaMaxIter, kSaPhiZero, kSaPhiZeroRel);
  const int nb = nblk(ncv);
  hipLaunchKernelGGL(k_sc_resid, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, AS.p, x, b, res, res0, 1, partials.p,
                     scal_s.p);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinResid0J, partials.p, scal_s.p);
  const int max_it = kSaMaxIter + 1;
  int launched = 0;
  while (launched < max_it) {
    const int chunk = std::min(launched == 0 ? 2 : 4, max_it - launched);
    for (int k = 0; k < chunk; k++) {
      hipLaunchKernelGGL(k_sc_p, g, bl, 0, stream, ncv, nbocv_i.p, AS.p, res, vv, p, phat, scal_s.p);
      hipLaunchKernelGGL(k_sc_spmv_dot, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, AS.p, phat, vv, res0, partials.p,
                         scal_s.p);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinAlpha, partials.p, scal_s.p);
      hipLaunchKernelGGL(k_sc_s, g, bl, 0, stream, ncv, nbocv_i.p, AS.p, res, vv, sv, shat, scal_s.p);
      hipLaunchKernelGGL(k_sc_spmv_dot, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, AS.p, shat, tv, sv, partials.p,
                         scal_s.p);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinOmega, partials.p, scal_s.p);
      hipLaunchKernelGGL(k_sc_x, g, bl, 0, stream, ncv, x, phat, shat, scal_s.p);
      hipLaunchKernelGGL(k_sc_resid, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, AS.p, x, b, res, res0, 0,
                         partials.p, scal_s.p);
      hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinResid, partials.p, scal_s.p);
    }
    launched += chunk;
    HIPCHECK(hipMemcpyAsync(h_scal_s, scal_s.p, sizeof(SolverScalars), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    if (h_scal_s->done) break;
  }
  scal_iters = h_scal_s->iter;
  hipLaunchKernelGGL(k_sc_update<T>, g, bl, 0, stream, ncv, q.p, dq, x, sa_phi.p);
}

// SST's half of a segregated step: kine, then omega, each as scalar_step
// solves nuSA (JoeWithModels.cpp:1050-1090: relaxation, V/dt, the coupling
// columns moved right with dq, Jacobi BiCGStab, update and clip with the
// scalar's bounds); omega's system was formed before kine's update, so the
// order only sets the order of the two updates
template <typename T>
void Ctx::Impl::sst_scalar_step(const T *dq) {
  double *b = sw[0].p, *x = sw[1].p, *res = sw[2].p, *res0 = sw[3].p, *p = sw[4].p, *vv = sw[5].p, *phat = sw[6].p,
         *sv = sw[7].p, *shat = sw[8].p, *tv = sw[9].p;
  const dim3 g(nblk(ncv)), bl(kBlock);
  const int nb = nblk(ncv);
  int iters[2] = {0, 0};
  for (int s = 0; s < 2; s++) {
    double *rs = rhs_s.p + (size_t)s * ncv, *as = AS.p + 6 * (size_t)nnz * s;
    hipLaunchKernelGGL(k_sc_prep<T>, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, vol.p, dt.p, par.relax_scal, rs, as,
                       dq, b, x);
    hipLaunchKernelGGL(k_fill<double>, g, bl, 0, stream, (size_t)ncv, p, 0.0);
    hipLaunchKernelGGL(k_fill<double>, g, bl, 0, stream, (size_t)ncv, vv, 0.0);
    hipLaunchKernelGGL(k_solver_init, dim3(1), dim3(64), 0, stream, scal_s.p, kSstMaxIter, kSstPhiZero,
                       kSstPhiZeroRel);
    hipLaunchKernelGGL(k_sc_resid, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, as, x, b, res, res0, 1, partials.p,
                       scal_s.p);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinResid0J, partials.p, scal_s.p);
    const int max_it = kSstMaxIter + 1;
    int launched = 0;
    while (launched < max_it) {
      const int chunk = std::min(launched == 0 ? 2 : 4, max_it - launched);
      for (int k = 0; k < chunk; k++) {
        hipLaunchKernelGGL(k_sc_p, g, bl, 0, stream, ncv, nbocv_i.p, as, res, vv, p, phat, scal_s.p);
        hipLaunchKernelGGL(k_sc_spmv_dot, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, as, phat, vv, res0, partials.p,
                           scal_s.p);
        hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinAlpha, partials.p, scal_s.p);
        hipLaunchKernelGGL(k_sc_s, g, bl, 0, stream, ncv, nbocv_i.p, as, res, vv, sv, shat, scal_s.p);
        hipLaunchKernelGGL(k_sc_spmv_dot, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, as, shat, tv, sv, partials.p,
                           scal_s.p);
        hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinOmega, partials.p, scal_s.p);
        hipLaunchKernelGGL(k_sc_x, g, bl, 0, stream, ncv, x, phat, shat, scal_s.p);
        hipLaunchKernelGGL(k_sc_resid, g, bl, 0, stream, ncv, nbocv_i.p, nbocv_v.p, as, x, b, res, res0, 0,
                           partials.p, scal_s.p);
        hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nb, (int)kFinResid, partials.p, scal_s.p);
      }
      launched += chunk;
      HIPCHECK(hipMemcpyAsync(h_scal_s, scal_s.p, sizeof(SolverScalars), hipMemcpyDeviceToHost, stream));
      HIPCHECK(hipStreamSynchronize(stream));
      if (h_scal_s->done) break;
    }
    iters[s] = h_scal_s->iter;
    HIPCHECK(hipMemcpyAsync(sst_dscal.p + (size_t)s * ncv, x, sizeof(double) * ncv, hipMemcpyDeviceToDevice, stream));
    hipLaunchKernelGGL(k_sc_update_b<T>, g, bl, 0, stream, ncv, q.p, dq, x, sst_phi.p + (size_t)s * ncv,
                       kSstLower[s], kSstUpper[s]);
  }
  scal_iters = iters[0];
  scal_iters2 = iters[1];
}

template <typename T>
void Ctx::Impl::launch_jac(bool step_mode) {
  DBuf<T> *v = nullptr;
  if constexpr (std::is_same<T, double>::value) v = v64; else v = v32;
  T *A = A_<T>();
  if (dup_blocks) {
    const size_t nA = (size_t)nbs * nbs * (size_t)nnz;
    hipLaunchKernelGGL(k_fill<T>, dim3((unsigned)((nA + kBlock - 1) / kBlock)), dim3(kBlock), 0, stream, nA, A, (T)0);
  }
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(nblk(ncv)), dim3(kBlock), 0, stream, dm, dp, prim.p, bfa.p, A, step_mode ? 1 : 0,
                       dt.p, rhs.p, par.under_relax, v[0].p, v[1].p, gradv.p);
  };
  const bool ent = jac_entry();
  auto goe = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((nwave + kJacWaves - 1) / kJacWaves), dim3(64 * kJacWaves), 0, stream, dm, dp,
                       prim.p, bfa.p, A, step_mode ? 1 : 0, dt.p, rhs.p, par.under_relax, v[0].p, v[1].p, gradv.p);
  };
  if (coupled_sa()) {  // 6x6 blocks: the flow rows, then the scalar row (SA needs MU_REF > 0)
    if (ent) goe(k_jac_e<T, kHllcC, true, 6>);
    else go(k_jac<T, kHllcC, true, 6>);
    launch_scal_c<T>(A, step_mode);
    return;
  }
  if (par.sst) {  // HLLC, viscous, per entry (the constructor checks the mesh allows it)
    goe(k_jac_e<T, kHllc, true, 5, true>);
    return;
  }
  if (ent) {
    if (dp.visc) {
      if (par.scheme == kRoe) goe(k_jac_e<T, kRoe, true>);
      else if (par.scheme == kHllcC) goe(k_jac_e<T, kHllcC, true>);
      else goe(k_jac_e<T, kHllc, true>);
    } else {
      if (par.scheme == kRoe) goe(k_jac_e<T, kRoe, false>);
      else if (par.scheme == kHllcC) goe(k_jac_e<T, kHllcC, false>);
      else goe(k_jac_e<T, kHllc, false>);
    }
    return;
  }
  if (dp.visc) {
    if (par.scheme == kRoe) go(k_jac<T, kRoe, true>);
    else if (par.scheme == kHllcC) go(k_jac<T, kHllcC, true>);
    else go(k_jac<T, kHllc, true>);
  } else {
    if (par.scheme == kRoe) go(k_jac<T, kRoe, false>);
    else if (par.scheme == kHllcC) go(k_jac<T, kHllcC, false>);
    else go(k_jac<T, kHllc, false>);
  }
}

// Work vectors: v[0] rhs, v[1] phi (= dq), v[2] res, v[3] res0, v[4] p,
// v[5] v, v[6] phat, v[7] s, v[8] shat, v[9] t.
// the context's block system as the BiCGStab kernels take it
template <typename T>
BcgSys Ctx::Impl::bcg_sys() {
  BcgSys y;
  y.ncv = ncv;
  y.maxiter = par.ls_maxiter;
  y.zero_abs = par.ls_abs;
  y.zero_rel = par.ls_rel;
  y.nbocv_i = nbocv_i.p;
  y.nbocv_v = nbocv_v.p;
  y.A = A_<T>();
  y.LU = lu_<T>()->p;
  DBuf<T> *v = krylov<T>();
  for (int k = 0; k < 10; k++) y.v[k] = v[k].p;
  y.partials = partials.p;
  y.S = scal.p;
  y.stream = stream;
  y.rec_resid = par.krylov_recurrence && !std::is_same<T, double>::value;
  if (par.ilu_sub > 0) {
    y.ilu_sub = par.ilu_sub;
    if constexpr (std::is_same<T, double>::value) {
      y.F = F64.p;
      y.Dinv = Dinv64.p;
    } else {
      y.F = F32.p;
      y.Dinv = Dinv32.p;
    }
    if (!y.F || !y.Dinv) throw std::logic_error("ILU factors not allocated");
  }
  return y;
}

template <typename T>
void Ctx::Impl::solve_prologue() {
  if (coupled_sa()) bcg_prologue<T, 6>(bcg_sys<T>());
  else bcg_prologue<T, 5>(bcg_sys<T>());
}

template <typename T>
void Ctx::Impl::solve_iterations(int count) {
  if (coupled_sa()) bcg_iterations<T, 6>(bcg_sys<T>(), count);
  else bcg_iterations<T, 5>(bcg_sys<T>(), count);
}

template <typename T>
void Ctx::Impl::solve(int &iters, int &status) {
  if (coupled_sa()) bcg_solve<T, 6>(bcg_sys<T>(), h_scal, iters, status);
  else bcg_solve<T, 5>(bcg_sys<T>(), h_scal, iters, status);
}

template <typename T>
void Ctx::Impl::enqueue_step_head() {
  hipLaunchKernelGGL(k_ramp, dim3(1), dim3(64), 0, stream, scal.p, par.cfl_after, par.cfl_interval, par.cfl_factor,
                     par.cfl_max);
  launch_grad(stream);
  launch_flux(stream, &scal.p->cfl_cur);
  launch_dt_fin();
  if (par.dt_mode == 1) hipLaunchKernelGGL(k_dt_fill, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, dt.p, scal.p);
  launch_jac<T>(true);
  solve_prologue<T>();
  solve_iterations<T>(kDevChunk);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, scal.p, mail_d);
}

template <typename T>
void Ctx::Impl::enqueue_step_chunk() {
  solve_iterations<T>(kDevChunk);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, scal.p, mail_d);
}

template <typename T>
void Ctx::Impl::enqueue_step_tail() {
  DBuf<T> *v = krylov<T>();
  hipLaunchKernelGGL(k_update<T>, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, q.p, v[1].p);
  launch_state_bc();
  hipLaunchKernelGGL(k_l1, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, rhs.p, partials.p);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nblk(ncv), (int)kFinL1, partials.p, scal.p);
  hipLaunchKernelGGL(k_publish, dim3(1), dim3(64), 0, stream, scal.p, mail_d);
}

hipGraphExec_t Ctx::Impl::capture(void (Ctx::Impl::*enq)()) {
  if (!cap_stream) HIPCHECK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
  hipStream_t keep = stream;
  stream = cap_stream;
  HIPCHECK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeThreadLocal));
  hipGraph_t g = nullptr;
  try {
    (this->*enq)();
  } catch (...) {
    (void)hipStreamEndCapture(cap_stream, &g);
    if (g) (void)hipGraphDestroy(g);
    stream = keep;
    throw;
  }
  stream = keep;
  HIPCHECK(hipStreamEndCapture(cap_stream, &g));
  hipGraphExec_t e = nullptr;
  hipError_t st = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHECK(st);
  return e;
}
void Ctx::launch_residual() {
  Impl &d = *d_;
  d.launch_state_bc();
  d.launch_grad_flux();
}

std::string Ctx::time_kernels(int reps, double *ms) {
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto timed = [&](auto &&launch) {
    launch();  // warm
    HIPCHECK(hipEventRecord(e0, d.stream));
    for (int r = 0; r < reps; r++) launch();
    HIPCHECK(hipEventRecord(e1, d.stream));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    return (double)t / reps;
  };
  ms[0] = timed([&] {
    hipLaunchKernelGGL(k_state, dim3(nblk(d.ncv)), dim3(kBlock), 0, d.stream, d.ncv, d.q.p, d.prim.p, d.par.gamma,
                       d.counters.p);
  });
  ms[1] = d.nfa_b > 0 ? timed([&] {
    hipLaunchKernelGGL(k_bc, dim3(nblk(d.nfa_b)), dim3(kBlock), 0, d.stream, d.dm, d.dp, d.prim.p, d.hook.p, d.bfa.p);
  }) : 0.0;
  ms[2] = d.par.second_order ? timed([&] { d.launch_grad(d.stream); }) : 0.0;
  ms[3] = timed([&] { d.launch_flux(d.stream); });
  // (the local time step is computed inside k_flux)
  ms[4] = timed([&] {
    d.launch_state_bc();
    d.launch_grad_flux();
  });
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
  return "k_state,k_bc,k_grad,k_flux,residual";
}

// Mean ms of the parts of an implicit step: the residual (state, BC,
// gradients, fluxes and dt; with the coupled scalar its row), the Jacobian
// assembly of a step (blocks, V/dt, relaxed right-hand side), the block LU
// of the preconditioner, and one BiCGStab iteration (its nine kernels) with
// the stop tests switched off so that every timed iteration runs.
std::string Ctx::time_step_kernels(int reps, double *ms) {
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  auto timed = [&](auto &&launch, int n) {
    launch();  // warm
    HIPCHECK(hipEventRecord(e0, d.stream));
    for (int r = 0; r < n; r++) launch();
    HIPCHECK(hipEventRecord(e1, d.stream));
    HIPCHECK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHECK(hipEventElapsedTime(&t, e0, e1));
    return (double)t / n;
  };
  auto parts = [&](auto tag) {
    using T = decltype(tag);
    d.ensure_solver<T>(10, true);
    ms[0] = timed([&] {
      d.launch_state_bc();
      d.launch_grad_flux(false);
    }, reps);
    ms[1] = timed([&] { d.launch_jac<T>(true); }, reps);
    BcgSys y = d.bcg_sys<T>();
    y.maxiter = 1 << 30;
    y.zero_abs = 0.0;
    y.zero_rel = 0.0;
    const int nb = d.nbs;
    ms[2] = timed([&] {
      if (nb == 6) bcg_prologue<T, 6>(y);
      else bcg_prologue<T, 5>(y);
    }, reps);
    ms[3] = timed([&] {
      if (nb == 6) bcg_iterations<T, 6>(y, 1);
      else bcg_iterations<T, 5>(y, 1);
    }, reps);
  };
  if (d.precision == 0) parts(double());
  else parts(float());
  HIPCHECK(hipEventDestroy(e0));
  HIPCHECK(hipEventDestroy(e1));
  return "residual,jacobian,krylov_setup,krylov_iteration";
}

void Ctx::calc_rhs(bool implicit) {
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  d.launch_state_bc();
  d.launch_grad_flux(implicit);
  if (implicit) {
    if (d.precision == 0) {
      d.ensure_solver<double>(2, false);
      d.launch_jac<double>(false);
    } else {
      d.ensure_solver<float>(2, false);
      d.launch_jac<float>(false);
    }
  }
  HIPCHECK(hipGetLastError());
}

StepInfo Ctx::step() {
  Impl &d = *d_;
  if (d.graph_inflight) throw std::logic_error("a device step is in flight");
  d.dev_step_dirty = true;
  if (!d.have_state) throw std::logic_error("no state set");
  if (d.begun) throw std::logic_error("step called between step_begin and step_end");
  StepInfo info;
  if (d.step_no == d.step0) d.launch_state_bc();
  d.step_no++;
  // CFL ramp (JoeWithModels.cpp:978)
  if ((d.step_no >= d.par.cfl_after) && (d.step_no % d.par.cfl_interval == 0) && (d.cfl < d.par.cfl_max))
    d.cfl *= d.par.cfl_factor;
  d.launch_grad_flux(true);
  int it = 0, st = 0;
  if (d.precision == 0) {
    d.ensure_solver<double>(10, true);
    d.launch_jac<double>(true);
    d.solve<double>(it, st);
    d.update<double>();
  } else {
    d.ensure_solver<float>(10, true);
    d.launch_jac<float>(true);
    d.solve<float>(it, st);
    d.update<float>();
  }
  d.launch_state_bc();
  d.finish_step(info, it, st);
  return info;
}

// Q += dQ and the scalar's update: segregated (its own solve) or coupled
template <typename T>
void Ctx::Impl::update() {
  const T *dq = krylov<T>()[1].p;
  if (coupled_sa()) {
    hipLaunchKernelGGL(k_update6<T>, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, q.p, dq, sa_phi.p);
    return;
  }
  hipLaunchKernelGGL(k_update<T>, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, q.p, dq);
  if (par.sa) scalar_step<T>(dq);
  if (par.sst) sst_scalar_step<T>(dq);
}

// residual L1 sums of the step just taken (JoeWithModels.cpp:1119-1136) and its report
void Ctx::Impl::finish_step(StepInfo &info, int it, int st) {
  hipLaunchKernelGGL(k_l1, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, rhs.p, partials.p);
  hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nblk(ncv), (int)kFinL1, partials.p, scal.p);
  if (par.sa) {
    // the segregated loop reports the scalar residual over the mean-flow
    // relaxation (JoeWithModels.cpp:1132-1134), the coupled loop as it is (:4443-4453)
    hipLaunchKernelGGL(k_l1s, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, rhs_s.p,
                       coupled_sa() ? 1.0 : par.under_relax, partials.p);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nblk(ncv), (int)kFinL1S, partials.p, scal.p);
  }
  if (par.sst) {  // kine's and omega's residuals over the mean-flow relaxation (:1132-1134)
    hipLaunchKernelGGL(k_l1s, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, rhs_s.p, par.under_relax, partials.p);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nblk(ncv), (int)kFinL1S, partials.p, scal.p);
    hipLaunchKernelGGL(k_l1s, dim3(nblk(ncv)), dim3(kBlock), 0, stream, ncv, rhs_s.p + ncv, par.under_relax,
                       partials.p);
    hipLaunchKernelGGL(k_fin, dim3(1), dim3(kFinBlock), 0, stream, nblk(ncv), (int)kFinL1S2, partials.p, scal.p);
  }
  HIPCHECK(hipMemcpyAsync(h_scal, scal.p, sizeof(SolverScalars), hipMemcpyDeviceToHost, stream));
  HIPCHECK(hipStreamSynchronize(stream));
  info.step = step_no;
  info.lin_iters = it;
  info.lin_status = st;
  info.cfl = cfl;
  info.dt_min = h_scal->dt_min;
  for (int k = 0; k < 5; k++) info.resid[k] = h_scal->resid[k];
  for (int k = 0; k < 5; k++) info.resid_l2[k] = h_scal->resid_l2[k];
  info.resid_s = (par.sa || par.sst) ? h_scal->resid_s : 0.0;
  info.resid_s2 = par.sst ? h_scal->resid_s2 : 0.0;
  info.scal_iters = (par.sa && !coupled_sa()) ? scal_iters : (par.sst ? scal_iters + scal_iters2 : 0);
}

void Ctx::step_begin(bool host_source) {
  Impl &d = *d_;
  if (d.graph_inflight) throw std::logic_error("a device step is in flight");
  d.dev_step_dirty = true;
  if (!d.have_state) throw std::logic_error("no state set");
  if (d.begun) throw std::logic_error("step_begin called twice without step_end");
  if (d.step_no == d.step0) d.launch_state_bc();
  d.step_no++;
  if ((d.step_no >= d.par.cfl_after) && (d.step_no % d.par.cfl_interval == 0) && (d.cfl < d.par.cfl_max))
    d.cfl *= d.par.cfl_factor;
  // a host sourceHook replaces the default one (gravity)
  const DParams keep = d.dp;
  if (host_source) d.dp.gravity_on = 0;
  d.launch_grad_flux(true);
  if (d.precision == 0) {
    d.ensure_solver<double>(10, true);
    d.launch_jac<double>(false);
  } else {
    d.ensure_solver<float>(10, true);
    d.launch_jac<float>(false);
  }
  d.dp = keep;
  HIPCHECK(hipGetLastError());
  d.begun = true;
}

void Ctx::add_source(const double *rhs_rho, const double *rhs_rhou, const double *rhs_rhoE, const double *a_diag) {
  Impl &d = *d_;
  if (!d.begun) throw std::logic_error("add_source outside step_begin/step_end");
  const size_t n = d.ncv, b = d.nbs, rec = b + b * b;
  d.h_src.assign(rec * n, 0.0);
  for (size_t cd = 0; cd < n; cd++) {
    const size_t c = d.perm[cd];
    double *o = &d.h_src[rec * cd];
    o[0] = rhs_rho ? rhs_rho[c] : 0.0;
    for (int i = 0; i < 3; i++) o[1 + i] = rhs_rhou ? rhs_rhou[3 * c + i] : 0.0;
    o[4] = rhs_rhoE ? rhs_rhoE[c] : 0.0;
    if (a_diag)  // the flow part of the system's diagonal block
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) o[b + b * i + j] = a_diag[25 * c + 5 * i + j];
  }
  d.push_source(a_diag != nullptr);
}

void Ctx::add_source_coupled(const double *rhs, const double *a_diag, int nb) {
  Impl &d = *d_;
  if (!d.begun) throw std::logic_error("add_source outside step_begin/step_end");
  if (nb != d.nbs) throw std::invalid_argument("block size " + std::to_string(nb) + " != the system's " + std::to_string(d.nbs));
  const size_t n = d.ncv, b = d.nbs, rec = b + b * b;
  d.h_src.assign(rec * n, 0.0);
  for (size_t cd = 0; cd < n; cd++) {
    const size_t c = d.perm[cd];
    double *o = &d.h_src[rec * cd];
    if (rhs)
      for (size_t k = 0; k < b; k++) o[k] = rhs[b * c + k];
    if (a_diag)
      for (size_t k = 0; k < b * b; k++) o[b + k] = a_diag[b * b * c + k];
  }
  d.push_source(a_diag != nullptr);
}

// the host's source terms (h_src, device order) into the residual and the
// diagonal blocks
void Ctx::Impl::push_source(bool with_a) {
  const size_t n = ncv, rec = nbs + nbs * nbs;
  ensure(src, rec * n);
  HIPCHECK(hipMemcpyAsync(src.p, h_src.data(), rec * n * sizeof(double), hipMemcpyHostToDevice, stream));
  const dim3 g(nblk(ncv)), bl(kBlock);
  const int wa = with_a ? 1 : 0;
  if (precision == 0) {
    if (coupled_sa()) hipLaunchKernelGGL((k_add_source<double, 6>), g, bl, 0, stream, ncv, nbocv_i.p, src.p, wa, rhs.p, rhs_s.p, A64.p);
    else hipLaunchKernelGGL((k_add_source<double, 5>), g, bl, 0, stream, ncv, nbocv_i.p, src.p, wa, rhs.p, rhs_s.p, A64.p);
  } else {
    if (coupled_sa()) hipLaunchKernelGGL((k_add_source<float, 6>), g, bl, 0, stream, ncv, nbocv_i.p, src.p, wa, rhs.p, rhs_s.p, A32.p);
    else hipLaunchKernelGGL((k_add_source<float, 5>), g, bl, 0, stream, ncv, nbocv_i.p, src.p, wa, rhs.p, rhs_s.p, A32.p);
  }
  // the host buffer must outlive the copy
  HIPCHECK(hipStreamSynchronize(stream));
}

template <typename T>
void Ctx::Impl::time_term() {
  DBuf<T> *v = krylov<T>();
  const dim3 g(nblk(ncv)), bl(kBlock);
  if (coupled_sa())
    hipLaunchKernelGGL((k_time_term<T, 6>), g, bl, 0, stream, ncv, nbocv_i.p, vol.p, dt.p, par.under_relax, rhs.p,
                       rhs_s.p, A_<T>(), v[0].p, v[1].p);
  else
    hipLaunchKernelGGL((k_time_term<T, 5>), g, bl, 0, stream, ncv, nbocv_i.p, vol.p, dt.p, par.under_relax, rhs.p,
                       rhs_s.p, A_<T>(), v[0].p, v[1].p);
}

StepInfo Ctx::step_end() {
  Impl &d = *d_;
  if (!d.begun) throw std::logic_error("step_end without step_begin");
  d.begun = false;
  int it = 0, st = 0;
  if (d.precision == 0) {
    d.time_term<double>();
    d.solve<double>(it, st);
    d.update<double>();
  } else {
    d.time_term<float>();
    d.solve<float>(it, st);
    d.update<float>();
  }
  d.launch_state_bc();
  StepInfo info;
  d.finish_step(info, it, st);
  return info;
}

void Ctx::update_bc() {
  Impl &d = *d_;
  if (!d.have_state) throw std::logic_error("no state set");
  d.launch_state_bc();
  HIPCHECK(hipGetLastError());
}

// doneSolver (JoeWithModels.h:42-95): NSTEPS, RESID_ENERG_TH, NaN
bool Ctx::done_solver(const StepInfo &s) {
  Impl &d = *d_;
  if (s.step == d.step0 + 1) d.resid0_energy = s.resid[4];
  const double r = s.resid[4];
  if ((d.par.nsteps >= 0) && (d.step_no >= d.par.nsteps)) return true;
  if (r < d.par.resid_energ_th) return true;
  if (r != r) return true;
  if (d.par.resid_energ_rel > 0.0 && d.step_no > d.step0 && r < d.par.resid_energ_rel * d.resid0_energy) return true;
  return false;
}

int Ctx::run(int max_steps, double *hist, bool graph) {
  Impl &d = *d_;
  {
    StepInfo s0;
    s0.step = d.step_no;
    s0.resid[4] = 1.0e20;
    if (d.step_no == d.step0 && done_solver(s0)) return 0;
    if (d.step_no > d.step0 && ((d.par.nsteps >= 0) && (d.step_no >= d.par.nsteps))) return 0;
  }
  int n = 0;
  while (n < max_steps) {
    StepInfo s = graph ? device_step() : step();
    if (hist)
      for (int k = 0; k < 5; k++) hist[5 * n + k] = s.resid[k];
    n++;
    if (done_solver(s)) break;
  }
  return n;
}

void Ctx::set_boundary_state(int ifa_f, int nface, const double *T, const double *vel, const double *p) {
  Impl &d = *d_;
  if (ifa_f < 0 || nface < 0 || ifa_f + nface > d.nfa_b) throw std::out_of_range("boundary face range");
  std::vector<double> h(5 * (size_t)nface);
  for (int k = 0; k < nface; k++) {
    h[5 * k] = T[k];
    h[5 * k + 1] = vel[3 * k];
    h[5 * k + 2] = vel[3 * k + 1];
    h[5 * k + 3] = vel[3 * k + 2];
    h[5 * k + 4] = p[k];
  }
  HIPCHECK(hipMemcpyAsync(d.hook.p + 5 * (size_t)ifa_f, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice,
                          d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
}

int Ctx::wall_values(int ifa_f, int nface, double *out) {
  Impl &d = *d_;
  if (ifa_f < 0 || nface < 0 || ifa_f + nface > d.nfa_b) throw std::out_of_range("boundary face range");
  if (!d.have_state) throw std::logic_error("no state set");
  if (!out || nface == 0) return nface;
  if (!d.dp.visc) throw std::invalid_argument("wall values need MU_REF > 0");
  DBuf<double> o;
  o.alloc(16 * (size_t)nface);
  hipLaunchKernelGGL(k_wall_values, dim3(nblk(nface)), dim3(kBlock), 0, d.stream, d.dm, d.dp, d.q.p, d.prim.p, d.bfa.p,
                     ifa_f, nface, o.p);
  HIPCHECK(hipGetLastError());
  HIPCHECK(hipMemcpyAsync(out, o.p, 16 * (size_t)nface * sizeof(double), hipMemcpyDeviceToHost, d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
  return nface;
}

void Ctx::get_boundary_state(int ifa_f, int nface, double *T, double *vel, double *p) {
  Impl &d = *d_;
  if (ifa_f < 0 || nface < 0 || ifa_f + nface > d.nfa_b) throw std::out_of_range("boundary face range");
  if (!d.have_state) throw std::logic_error("no state set");
  std::vector<double> h(8 * (size_t)nface);
  if (nface)
    HIPCHECK(hipMemcpyAsync(h.data(), d.bfa.p + 8 * (size_t)ifa_f, h.size() * sizeof(double), hipMemcpyDeviceToHost,
                            d.stream));
  HIPCHECK(hipStreamSynchronize(d.stream));
  for (int k = 0; k < nface; k++) {  // bfa rows: rho u v w p T h phi
    const double *b = &h[8 * (size_t)k];
    if (T) T[k] = b[5];
    if (vel)
      for (int i = 0; i < 3; i++) vel[3 * k + i] = b[1 + i];
    if (p) p[k] = b[4];
  }
}

double *Ctx::device_state() { return d_->q.p; }
double *Ctx::device_rhs() { return d_->rhs.p; }
void Ctx::sync() { HIPCHECK(hipStreamSynchronize(d_->stream)); }

int64_t Ctx::get(const std::string &name, double *out) {
  Impl &d = *d_;
  HIPCHECK(hipStreamSynchronize(d.stream));
  const size_t n = d.ncv, nb = d.nfa_b;
  const std::vector<int> &perm = d.perm;
  std::vector<double> h;
  auto copy = [&](const void *src, size_t bytes) {
    h.resize(bytes / sizeof(double));
    HIPCHECK(hipMemcpy(h.data(), src, bytes, hipMemcpyDeviceToHost));
  };
  // SoA [5][n] device-ordered component(s) -> reference-ordered AoS
  auto soa = [&](int comp0, int count) -> int64_t {
    if (out) {
      copy(d.rhs.p, 5 * n * sizeof(double));
      for (size_t cd = 0; cd < n; cd++)
        for (int i = 0; i < count; i++) out[(size_t)perm[cd] * count + i] = h[(comp0 + i) * n + cd];
    }
    return (int64_t)(n * count);
  };
  // per-cell records in device order -> reference order
  auto cellrec = [&](const double *src, int stride, int first, int count) -> int64_t {
    if (out) {
      copy(src, n * stride * sizeof(double));
      for (size_t cd = 0; cd < n; cd++)
        for (int k = 0; k < count; k++) out[(size_t)perm[cd] * count + k] = h[cd * stride + first + k];
    }
    return (int64_t)(n * count);
  };
  // boundary-face records keep the reference face index
  auto bfarec = [&](int first, int count) -> int64_t {
    if (out) {
      copy(d.bfa.p, nb * 8 * sizeof(double));
      for (size_t f = 0; f < nb; f++)
        for (int k = 0; k < count; k++) out[f * count + k] = h[f * 8 + first + k];
    }
    return (int64_t)(nb * count);
  };
  if (name == "RHS_RHO") return soa(0, 1);
  if (name == "RHS_RHOU") return soa(1, 3);
  if (name == "RHS_RHOE") return soa(4, 1);
  if (name == "LOCAL_DT") return cellrec(d.dt.p, 1, 0, 1);
  if (name == "PRESS") return cellrec(d.prim.p, 8, 4, 1);
  if (name == "VEL") return cellrec(d.prim.p, 8, 1, 3);
  if (name == "GRAD_RHO") return cellrec(d.grad.p, 16, 0, 3);
  if (name == "GRAD_P") return cellrec(d.grad.p, 16, 3, 3);
  // with viscous terms the reference's grad_u is the one calcViscousFluxNS
  // last formed (JoeWithModels.cpp:3616-3630)
  if (name == "GRAD_U") return d.dp.visc ? cellrec(d.gradv.p, 12, 0, 9) : cellrec(d.grad.p, 16, 6, 9);
  if (name == "GRAD_ENTHALPY") {
    if (!d.dp.visc) throw std::invalid_argument("GRAD_ENTHALPY exists only with MU_REF > 0");
    return cellrec(d.gradv.p, 12, 9, 3);
  }
  if (name == "ALPHA_RHO") return cellrec(d.grad.p, 16, 15, 1);
  if (name == "RHO_BFA") return bfarec(0, 1);
  if (name == "VEL_BFA") return bfarec(1, 3);
  if (name == "P_BFA") return bfarec(4, 1);
  if (name == "T_BFA") return bfarec(5, 1);
  if (name == "H_BFA") return bfarec(6, 1);
  if (name == "RESID") {  // L1 residuals of the last step; with SA nuSA's sixth, with SST kine's and omega's
    const int nr = d.par.sa ? 6 : (d.par.sst ? 7 : 5);
    if (out) {
      HIPCHECK(hipMemcpy(d.h_scal, d.scal.p, sizeof(SolverScalars), hipMemcpyDeviceToHost));
      for (int k = 0; k < 5; k++) out[k] = d.h_scal->resid[k];
      if (d.par.sa || d.par.sst) out[5] = d.h_scal->resid_s;
      if (d.par.sst) out[6] = d.h_scal->resid_s2;
    }
    return nr;
  }
  if (d.par.sst) {
    auto sbfa = [&](int s) -> int64_t {  // boundary values keep the reference face index
      if (out) {
        copy(d.sst_bfa.p, 2 * nb * sizeof(double));
        for (size_t f = 0; f < nb; f++) out[f] = h[2 * f + s];
      }
      return (int64_t)nb;
    };
    if (name == "KINE") return cellrec(d.sst_phi.p, 1, 0, 1);
    if (name == "OMEGA") return cellrec(d.sst_phi.p + n, 1, 0, 1);
    if (name == "KINE_BFA") return sbfa(0);
    if (name == "OMEGA_BFA") return sbfa(1);
    if (name == "GRAD_KINE") return cellrec(d.sst_aux.p, kSstN, kSGradK, 3);
    if (name == "GRAD_OMEGA") return cellrec(d.sst_aux.p, kSstN, kSGradW, 3);
    if (name == "WALL_DIST") return cellrec(d.sst_aux.p, kSstN, kSWall, 1);
    if (name == "STR_MAG") return cellrec(d.sst_aux.p, kSstN, kSStr, 1);
    if (name == "VORT_MAG") return cellrec(d.sst_aux.p, kSstN, kSVort, 1);
    if (name == "DIVERG") return cellrec(d.sst_aux.p, kSstN, kSDiv, 1);
    if (name == "CROSS_DIFF") return cellrec(d.sst_aux.p, kSstN, kSCd, 1);
    if (name == "F1") return cellrec(d.sst_aux.p, kSstN, kSF1, 1);
    if (name == "F2") return cellrec(d.sst_aux.p, kSstN, kSF2, 1);
    if (name == "RHS_KINE") return cellrec(d.rhs_s.p, 1, 0, 1);
    if (name == "RHS_OMEGA") return cellrec(d.rhs_s.p + n, 1, 0, 1);
    if (name == "DSCAL") {  // [2][ncv]: the last step's kine and omega corrections
      if (out) {
        copy(d.sst_dscal.p, 2 * n * sizeof(double));
        for (int s = 0; s < 2; s++)
          for (size_t cd = 0; cd < n; cd++) out[(size_t)s * n + perm[cd]] = h[(size_t)s * n + cd];
      }
      return (int64_t)(2 * n);
    }
    if (name == "ASCAL") {  // [2][6][nnz] in the reference's nbocv order, runBackwardEuler's AScal
      const size_t nnz = d.nnz;
      if (out) {
        copy(d.AS.p, 12 * nnz * sizeof(double));
        for (int s = 0; s < 2; s++)
          for (size_t cd = 0; cd < n; cd++) {
            const size_t c = perm[cd];
            const size_t len = d.h_dev_nbocv_i[cd + 1] - d.h_dev_nbocv_i[cd];
            for (size_t j = 0; j < len; j++)
              for (int k = 0; k < 6; k++)
                out[(6 * (size_t)s + k) * nnz + d.h_ref_nbocv_i[c] + j] =
                    h[6 * nnz * s + 6 * (d.h_dev_nbocv_i[cd] + j) + k];
          }
      }
      return (int64_t)(12 * nnz);
    }
  }
  if (d.par.sa) {
    if (name == "NUSA") return cellrec(d.sa_phi.p, 1, 0, 1);
    if (name == "NUSA_BFA") return bfarec(7, 1);
    if (name == "GRAD_NUSA") return cellrec(d.sa_aux.p, kAuxN, kAuxGrad, 3);
    if (name == "WALL_DIST") return cellrec(d.sa_aux.p, kAuxN, kAuxWall, 1);
    if (name == "STR_MAG") return cellrec(d.sa_aux.p, kAuxN, kAuxStr, 1);
    if (name == "VORT_MAG") return cellrec(d.sa_aux.p, kAuxN, kAuxVort, 1);
    if (name == "RHS_NUSA") return cellrec(d.rhs_s.p, 1, 0, 1);
    if ((name == "DSCAL" || name == "ASCAL") && d.coupled_sa())
      throw std::invalid_argument(name + ": the coupled loop keeps the scalar in the 6x6 system (A, DQ)");
    if (name == "DSCAL") return cellrec(d.sw[1].p, 1, 0, 1);
    if (name == "ASCAL") {  // [6][nnz] in the reference's nbocv order, like runBackwardEuler's AScal[iScal]
      const size_t nnz = d.nnz;
      if (out) {
        copy(d.AS.p, 6 * nnz * sizeof(double));
        for (size_t cd = 0; cd < n; cd++) {
          const size_t c = perm[cd];
          const size_t len = d.h_dev_nbocv_i[cd + 1] - d.h_dev_nbocv_i[cd];
          for (size_t j = 0; j < len; j++)
            for (int k = 0; k < 6; k++)
              out[k * nnz + d.h_ref_nbocv_i[c] + j] = h[6 * (d.h_dev_nbocv_i[cd] + j) + k];
        }
      }
      return (int64_t)(6 * nnz);
    }
  }
  if (name == "FLUX_FA") {  // each face's inviscid flux [nfa][5], reference numbering
    const size_t nf = d.nfa;
    if (!out) return (int64_t)(5 * nf);
    if (!d.have_state) throw std::logic_error("no state set");
    if (!d.flux_entry() || d.face_form()) throw Unsupported("FLUX_FA is written by the per-entry flux kernel");
    DBuf<double> fd;
    fd.alloc(5 * nf);
    d.dp.fdump = fd.p;
    try {
      d.launch_state_bc();
      d.launch_grad_flux(false);
    } catch (...) {
      d.dp.fdump = nullptr;
      throw;
    }
    d.dp.fdump = nullptr;
    copy(fd.p, 5 * nf * sizeof(double));
    for (size_t f = 0; f < nf; f++)
      for (int k = 0; k < 5; k++) out[5 * f + k] = h[5 * (size_t)d.fnew[f] + k];
    return (int64_t)(5 * nf);
  }
  if (name == "MUL_FA" || name == "MUT_FA") {
    // each face's laminar (mul_fa, calcMaterialProperties,
    // UgpWithCvCompFlow.h:1701-1753) or eddy viscosity (mut_fa,
    // calcRansTurbViscMuet, TurbModel_SA.h:74-128) [nfa], reference
    // numbering, as the viscous flux kernel forms them from the current state.
    // Zero when the case is inviscid (the reference leaves mul_fa unset then)
    // or has no turbulence model
    const size_t nf = d.nfa;
    if (!out) return (int64_t)nf;
    if (!d.have_state) throw std::logic_error("no state set");
    if (!d.dp.visc || (name == "MUT_FA" && !d.dp.sa && !d.dp.sst)) {
      for (size_t f = 0; f < nf; f++) out[f] = 0.0;
      return (int64_t)nf;
    }
    DBuf<double> vd;
    vd.alloc(2 * nf);
    d.dp.vdump = vd.p;
    try {
      d.launch_state_bc();
      d.launch_grad_flux(false);
    } catch (...) {
      d.dp.vdump = nullptr;
      throw;
    }
    d.dp.vdump = nullptr;
    copy(vd.p, 2 * nf * sizeof(double));
    const int k = name == "MUL_FA" ? 0 : 1;
    for (size_t f = 0; f < nf; f++) out[f] = h[2 * (size_t)d.fnew[f] + k];
    return (int64_t)nf;
  }
  if (name == "COUNTERS") {  // negative pressures, first-order fallbacks
    if (out) {
      int c[4];
      HIPCHECK(hipMemcpy(c, d.counters.p, sizeof(c), hipMemcpyDeviceToHost));
      for (int k = 0; k < 4; k++) out[k] = c[k];
    }
    return 4;
  }
  if (name == "A" || name == "DQ") {  // [nnz][b][b] and [ncv][b], b = 5, or 6 in the coupled loop with SA
    const bool isA = name == "A";
    const size_t b = d.nbs, bb = b * b;
    const size_t cnt = isA ? bb * (size_t)d.nnz : b * n;
    if (!out) return (int64_t)cnt;
    std::vector<double> dev(cnt);
    if (d.precision == 0) {
      const double *src = isA ? d.A64.p : d.v64[1].p;
      if (!src) throw std::logic_error(name + " not computed yet");
      HIPCHECK(hipMemcpy(dev.data(), src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      const float *src = isA ? d.A32.p : d.v32[1].p;
      if (!src) throw std::logic_error(name + " not computed yet");
      std::vector<float> f(cnt);
      HIPCHECK(hipMemcpy(f.data(), src, cnt * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < cnt; k++) dev[k] = f[k];
    }
    if (isA) {  // device rows -> reference nbocv rows (same order inside a row)
      for (size_t cd = 0; cd < n; cd++) {
        const size_t c = perm[cd];
        const size_t len = d.h_dev_nbocv_i[cd + 1] - d.h_dev_nbocv_i[cd];
        std::memcpy(out + bb * (size_t)d.h_ref_nbocv_i[c], dev.data() + bb * (size_t)d.h_dev_nbocv_i[cd],
                    bb * len * sizeof(double));
      }
    } else {
      for (size_t cd = 0; cd < n; cd++)
        for (size_t k = 0; k < b; k++) out[b * (size_t)perm[cd] + k] = dev[b * cd + k];
    }
    return (int64_t)cnt;
  }
  if (name == "ILU_F" || name == "ILU_DINV") {
    // the block ILU(0) factors of the last solve (GPU_PRECONDITIONER = ILU0):
    // F [nnz][b][b] in the reference's nbocv order like A (unit-lower L
    // below the diagonal, U above it, within each subdomain; other blocks
    // unused), and the inverted pivot blocks [ncv][b][b]
    if (d.par.ilu_sub <= 0) throw Unsupported(name + " needs GPU_PRECONDITIONER = ILU0");
    const bool isF = name == "ILU_F";
    const size_t bb = (size_t)d.nbs * d.nbs;
    const size_t cnt = bb * (isF ? (size_t)d.nnz : n);
    if (!out) return (int64_t)cnt;
    std::vector<double> dev(cnt);
    if (d.precision == 0) {
      const double *src = isF ? d.F64.p : d.Dinv64.p;
      if (!src) throw std::logic_error(name + " not computed yet");
      HIPCHECK(hipMemcpy(dev.data(), src, cnt * sizeof(double), hipMemcpyDeviceToHost));
    } else {
      const float *src = isF ? d.F32.p : d.Dinv32.p;
      if (!src) throw std::logic_error(name + " not computed yet");
      std::vector<float> f(cnt);
      HIPCHECK(hipMemcpy(f.data(), src, cnt * sizeof(float), hipMemcpyDeviceToHost));
      for (size_t k = 0; k < cnt; k++) dev[k] = f[k];
    }
    for (size_t cd = 0; cd < n; cd++) {
      const size_t c = perm[cd];
      if (isF) {
        const size_t len = d.h_dev_nbocv_i[cd + 1] - d.h_dev_nbocv_i[cd];
        std::memcpy(out + bb * (size_t)d.h_ref_nbocv_i[c], dev.data() + bb * (size_t)d.h_dev_nbocv_i[cd],
                    bb * len * sizeof(double));
      } else {
        std::memcpy(out + bb * c, dev.data() + bb * cd, bb * sizeof(double));
      }
    }
    return (int64_t)cnt;
  }
  if (name == "CELL_ORDER") {  // device position of each reference cell
    if (out)
      for (size_t cd = 0; cd < n; cd++) out[perm[cd]] = (double)cd;
    return (int64_t)n;
  }
  throw std::invalid_argument("unknown array " + name);
}

// ---------------------------------------------------------------------------
// solveCvVectorR5Bcgstab (UgpWithCv.cpp:2660-2876) on a 5x5 block system the
// caller holds on the device, with the same kernels and launch sequence as
// the context's own linear solve. The work space (block LU, eight vectors,
// partials, solver scalars) is carved from one device buffer.
// ---------------------------------------------------------------------------
namespace {
// doubles of work space a solve of ncv cells takes
size_t bcg_workspace_doubles(int ncv) {
  const size_t n = (size_t)ncv;
  const size_t gb = (n + kCellsPerBlock - 1) / kCellsPerBlock;
  const size_t scal = (sizeof(SolverScalars) + 7) / 8;
  return 25 * n + 8 * 5 * n + 2 * (gb + 64) + scal + 32;
}

BcgSys bcg_on_workspace(const BlockSolve &bs, double *ws, hipStream_t stream) {
  const size_t n = (size_t)bs.ncv;
  const size_t gb = (n + kCellsPerBlock - 1) / kCellsPerBlock;
  BcgSys y;
  y.ncv = bs.ncv;
  y.maxiter = bs.max_iter;
  y.zero_abs = bs.zero_abs;
  y.zero_rel = bs.zero_rel;
  y.nbocv_i = bs.nbocv_i;
  y.nbocv_v = bs.nbocv_v;
  y.A = bs.A;
  y.v[0] = (void *)bs.rhs;
  y.v[1] = bs.phi;
  size_t o = 0;
  y.S = (SolverScalars *)(ws + o);  // first, so it keeps the buffer's alignment
  o += (sizeof(SolverScalars) + 7) / 8 + 32;
  y.LU = ws + o;
  o += 25 * n;
  for (int k = 2; k < 10; k++) {
    y.v[k] = ws + o;
    o += 5 * n;
  }
  y.partials = ws + o;
  o += 2 * (gb + 64);
  y.stream = stream;
  return y;
}

void bcg_check(const BlockSolve &bs) {
  if (bs.ncv <= 0 || !bs.nbocv_i || !bs.nbocv_v || !bs.A || !bs.rhs || !bs.phi || bs.max_iter < 0)
    throw std::invalid_argument("block BiCGStab: ncv > 0, device arrays and max_iter >= 0 needed");
}
}  // namespace

void block_bcgstab(const BlockSolve &bs, int &iters, int &status) {
  bcg_check(bs);
  hipStream_t st = (hipStream_t)bs.stream;
  DBuf<double> ws;
  ws.alloc(bcg_workspace_doubles(bs.ncv));
  SolverScalars hs;
  bcg_solve<double, 5>(bcg_on_workspace(bs, ws.p, st), &hs, iters, status);
}

// ---------------------------------------------------------------------------
// The same solve for callers that cannot allocate device memory: the work
// space is this file-scope array, one copy per device, held until the process
// ends. Host code never passes the array's name to a kernel (on the host it
// is a stand-in, not a device address): each device's address is looked up
// once with hipGetSymbolAddress, with that device current, and kept. Every
// call on a device shares its array and leaves the last call's contents in
// it; bcg_prologue clears what must start at zero (p, v) with fill kernels
// and writes the solver scalars, LU and residuals before they are read, so
// nothing carries over. Calls that overlap on one device race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[kNoAllocScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_dev[kMaxDevices] = {};

// the current device's address of scratch
double *scratch_on_current_device() {
  int dev = -1;
  HIPCHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device ordinal out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      throw std::runtime_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    g_scratch_dev[dev] = (double *)ptr;
  }
  return g_scratch_dev[dev];
}
}  // namespace

size_t block_bcgstab_noalloc_max_cells() {
  size_t lo = 0, hi = kNoAllocScratchDoubles;  // largest ncv whose work space fits
  while (lo < hi) {
    const size_t mid = (lo + hi + 1) / 2;
    if (bcg_workspace_doubles((int)mid) <= kNoAllocScratchDoubles) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

void block_bcgstab_noalloc(const BlockSolve &bs, int &iters, int &status) {
  // the size is checked first, before any device call
  if (bs.ncv > 0 && bcg_workspace_doubles(bs.ncv) > kNoAllocScratchDoubles)
    throw ArgError("block BiCGStab without allocation: ncv = " + std::to_string(bs.ncv) +
                                " needs more than the " + std::to_string(kNoAllocScratchDoubles * 8 >> 20) +
                                " MiB work space (at most " + std::to_string(block_bcgstab_noalloc_max_cells()) +
                                " cells)");
  bcg_check(bs);
  double *ws = scratch_on_current_device();
  SolverScalars hs;
  bcg_solve<double, 5>(bcg_on_workspace(bs, ws, (hipStream_t)bs.stream), &hs, iters, status);
}

}  // namespace jm
