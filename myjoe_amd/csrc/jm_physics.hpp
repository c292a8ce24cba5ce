// Pointwise device physics of the steady-solve hot path: HLLC / Roe / AUSM
// face fluxes, their Jacobians, and the 5x5 block LU of the block-Jacobi
// preconditioner. Every expression keeps the operand order of the reference
// routine cited above it, so that compiled without FMA contraction
// (-ffp-contract=off) the FP64 results equal the reference's bit for bit.
#pragma once

#include <hip/hip_runtime.h>

namespace jm {

struct Dev {
  // std::min / std::max semantics (return the first argument on ties / NaN)
  static __device__ __forceinline__ double mn(double a, double b) { return (b < a) ? b : a; }
  static __device__ __forceinline__ double mx(double a, double b) { return (a < b) ? b : a; }
};

// vecDotVec3d (tc_vec3d.h:14-22): sum starts at 0.0
__device__ __forceinline__ double dot3(const double *a, const double *b) {
  double d = 0.0;
  d += a[0] * b[0];
  d += a[1] * b[1];
  d += a[2] * b[2];
  return d;
}

// HLLC flux, UgpWithCvCompFlow.cpp:120-235 (surfVeloc = 0; kL, kR the sides'
// turbulent kinetic energy, 0 without SST). With Fs, one transported scalar:
// FrhoScal = Frho * (upwind side's value) * area; with Fs2 a second one.
__device__ __forceinline__ void flux_hllc(double F[5], double rhoL, const double *uL, double pL, double rhoR,
                                          const double *uR, double pR, double gam, double area,
                                          const double *nVec, double phiL = 0.0, double phiR = 0.0,
                                          double *Fs = nullptr, double kL = 0.0, double kR = 0.0,
                                          double *Fs2 = nullptr, double phiL2 = 0.0, double phiR2 = 0.0) {
  const double gammaL = gam, gammaR = gam, surfv = 0.0;
  double unL = dot3(uL, nVec) - surfv;
  double unLr = unL - surfv;
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double hL = gammaL / (gammaL - 1.0) * pL / rhoL + 0.5 * uLuL + kL;
  double eL = hL * rhoL - pL;
  double unR = dot3(uR, nVec) - surfv;
  double unRr = unR - surfv;
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double hR = gammaR / (gammaR - 1.0) * pR / rhoR + 0.5 * uRuR + kR;
  double eR = hR * rhoR - pR;
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
#pragma unroll
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unLr - cL);
  double sR = Dev::mx(uRoe + cRoe, unRr + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double f0, f1, f2, f3, f4;
  if (sM >= 0.0) {
    if (sL > 0.0) {
      f0 = rhoL * unLr;
      f1 = rhoL * uL[0] * unLr + pL * nVec[0];
      f2 = rhoL * uL[1] * unLr + pL * nVec[1];
      f3 = rhoL * uL[2] * unLr + pL * nVec[2];
      f4 = eL * unLr + pL * unL;
    } else {
      double invSLmSs = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * invSLmSs;
      double r0 = (rhoL * uL[0] * sLmuL + (pStar - pL) * nVec[0]) * invSLmSs;
      double r1 = (rhoL * uL[1] * sLmuL + (pStar - pL) * nVec[1]) * invSLmSs;
      double r2 = (rhoL * uL[2] * sLmuL + (pStar - pL) * nVec[2]) * invSLmSs;
      double eSL = (sLmuL * eL - pL * unL + pStar * sM) * invSLmSs;
      f0 = rhoSL * sM;
      f1 = r0 * sM + pStar * nVec[0];
      f2 = r1 * sM + pStar * nVec[1];
      f3 = r2 * sM + pStar * nVec[2];
      f4 = (eSL + pStar) * sM;
    }
  } else {
    if (sR >= 0.0) {
      double invSRmSs = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * invSRmSs;
      double r0 = (rhoR * uR[0] * sRmuR + (pStar - pR) * nVec[0]) * invSRmSs;
      double r1 = (rhoR * uR[1] * sRmuR + (pStar - pR) * nVec[1]) * invSRmSs;
      double r2 = (rhoR * uR[2] * sRmuR + (pStar - pR) * nVec[2]) * invSRmSs;
      double eSR = (sRmuR * eR - pR * unR + pStar * sM) * invSRmSs;
      f0 = rhoSR * sM;
      f1 = r0 * sM + pStar * nVec[0];
      f2 = r1 * sM + pStar * nVec[1];
      f3 = r2 * sM + pStar * nVec[2];
      f4 = (eSR + pStar) * sM;
    } else {
      f0 = rhoR * unRr;
      f1 = rhoR * uR[0] * unRr + pR * nVec[0];
      f2 = rhoR * uR[1] * unRr + pR * nVec[1];
      f3 = rhoR * uR[2] * unRr + pR * nVec[2];
      f4 = eR * unRr + pR * unR;
    }
  }
  F[0] = f0 * area;
  F[1] = f1 * area;
  F[2] = f2 * area;
  F[3] = f3 * area;
  F[4] = f4 * area;
  if (Fs) *Fs = f0 * (sM >= 0.0 ? phiL : phiR) * area;
  if (Fs2) *Fs2 = f0 * (sM >= 0.0 ? phiL2 : phiR2) * area;
}

// Roe flux without entropy fix, UgpWithCvCompFlow.cpp:3281-3380; returns the
// central part alone when the Roe-averaged a^2 < 1e-12 (:3318-3319).
__device__ __forceinline__ void flux_roe(double F[5], double rho0, const double *u0, double p0, double rho1,
                                         const double *u1, double p1, double gam, double area, const double *nVec) {
  double gamma = 0.5 * (gam + gam);
  double gm1 = gamma - 1.0;
  double rhoun0 = rho0 * dot3(u0, nVec);
  double rhoun1 = rho1 * dot3(u1, nVec);
  double rhoE0 = p0 / gm1 + 0.5 * rho0 * dot3(u0, u0);
  double rhoE1 = p1 / gm1 + 0.5 * rho1 * dot3(u1, u1);
  double H0 = (rhoE0 + p0) / rho0;
  double H1 = (rhoE1 + p1) / rho1;
  F[0] = 0.5 * area * (rhoun0 + rhoun1);
  F[1] = 0.5 * area * (rhoun0 * u0[0] + rhoun1 * u1[0] + (p0 + p1) * nVec[0]);
  F[2] = 0.5 * area * (rhoun0 * u0[1] + rhoun1 * u1[1] + (p0 + p1) * nVec[1]);
  F[3] = 0.5 * area * (rhoun0 * u0[2] + rhoun1 * u1[2] + (p0 + p1) * nVec[2]);
  F[4] = 0.5 * area * (rhoun0 * H0 + rhoun1 * H1);
  double sqrt_rho0 = sqrt(rho0), sqrt_rho1 = sqrt(rho1);
  double tmp = 1.0 / (sqrt_rho0 + sqrt_rho1);
  double uAvg = tmp * (u0[0] * sqrt_rho0 + u1[0] * sqrt_rho1);
  double vAvg = tmp * (u0[1] * sqrt_rho0 + u1[1] * sqrt_rho1);
  double wAvg = tmp * (u0[2] * sqrt_rho0 + u1[2] * sqrt_rho1);
  double hAvg = tmp * (H0 * sqrt_rho0 + H1 * sqrt_rho1);
  double alphaAvg = 0.5 * (uAvg * uAvg + vAvg * vAvg + wAvg * wAvg);
  double a2Avg = gm1 * (hAvg - alphaAvg);
  if (a2Avg < 1.0E-12) return;
  double aAvg = sqrt(a2Avg);
  double unAvg = uAvg * nVec[0] + vAvg * nVec[1] + wAvg * nVec[2];
  double dr = rho1 - rho0;
  double dru = rho1 * u1[0] - rho0 * u0[0];
  double drv = rho1 * u1[1] - rho0 * u0[1];
  double drw = rho1 * u1[2] - rho0 * u0[2];
  double drE = rhoE1 - rhoE0;
  double lam1 = fabs(unAvg);
  double lam2 = fabs(unAvg + aAvg);
  double lam3 = fabs(unAvg - aAvg);
  double abv1 = 0.5 * (lam2 + lam3);
  double abv2 = 0.5 * (lam2 - lam3);
  double abv3 = abv1 - lam1;
  double abv4 = gm1 * (alphaAvg * dr - uAvg * dru - vAvg * drv - wAvg * drw + drE);
  double abv5 = unAvg * dr - nVec[0] * dru - nVec[1] * drv - nVec[2] * drw;
  double abv6 = abv3 * abv4 / a2Avg - abv2 * abv5 / aAvg;
  double abv7 = abv3 * abv5 - abv2 * abv4 / aAvg;
  F[0] -= 0.5 * area * (lam1 * dr + abv6);
  F[1] -= 0.5 * area * (lam1 * dru + uAvg * abv6 - nVec[0] * abv7);
  F[2] -= 0.5 * area * (lam1 * drv + vAvg * abv6 - nVec[1] * abv7);
  F[3] -= 0.5 * area * (lam1 * drw + wAvg * abv6 - nVec[2] * abv7);
  F[4] -= 0.5 * area * (lam1 * drE + hAvg * abv6 - unAvg * abv7);
}

// AUSM (Liou-Steffen), UgpWithCvCompFlow.cpp:3566-3641
__device__ __forceinline__ void flux_ausm(double F[5], double rhoL, const double *uL, double pL, double rhoR,
                                          const double *uR, double pR, double gam, double area,
                                          const double *nVec) {
  const double gammaL = gam, gammaR = gam, kL = 0.0, kR = 0.0;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double hL = (gammaL / (gammaL - 1.0)) * (pL / rhoL) + 0.5 * uLuL + kL;
  double mL = unL / cL;
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double hR = (gammaR / (gammaR - 1.0)) * (pR / rhoR) + 0.5 * uRuR + kR;
  double mR = unR / cR;
  double mLP = fabs(mL) <= 1.0 ? 0.25 * (mL + 1.0) * (mL + 1.0) : 0.5 * (mL + fabs(mL));
  double mRM = fabs(mR) <= 1.0 ? -0.25 * (mR - 1.0) * (mR - 1.0) : 0.5 * (mR - fabs(mR));
  double mF = mLP + mRM;
  double pLP = fabs(mL) <= 1.0 ? 0.25 * pL * (mL + 1.0) * (mL + 1.0) * (2.0 - mL) : 0.5 * pL * (mL + fabs(mL)) / mL;
  double pRM = fabs(mR) <= 1.0 ? 0.25 * pR * (mR - 1.0) * (mR - 1.0) * (2.0 + mR) : 0.5 * pR * (mR - fabs(mR)) / mR;
  double pF = pLP + pRM;
  double Phi = fabs(mF);
  double rcL = rhoL * cL, rcR = rhoR * cR;
  F[0] = 0.5 * (mF * ((rcL) + (rcR)) - Phi * ((rcR) - (rcL)));
#pragma unroll
  for (int i = 0; i < 3; i++)
    F[1 + i] = 0.5 * (mF * ((rcL * uL[i]) + (rcR * uR[i])) - Phi * ((rcR * uR[i]) - (rcL * uL[i]))) + nVec[i] * pF;
  F[4] = 0.5 * (mF * ((rcL * hL) + (rcR * hR)) - Phi * ((rcR * hR) - (rcL * hL)));
#pragma unroll
  for (int i = 0; i < 5; i++) F[i] *= area;
}

// calcJacobianA, UgpWithCvCompFlow.cpp:666-709 (nV area-weighted)
__device__ __forceinline__ void jac_a(double A[5][5], const double *vel, double pp, double rrho, const double *nV,
                                      double gamma) {
  const double surfVeloc = 0.0;
  double kapm1 = (gamma - 1.0);
  double nVel0 = vel[0] * nV[0], nVel1 = vel[1] * nV[1], nVel2 = vel[2] * nV[2];
  double U_k = nVel0 + nVel1 + nVel2;
  double vSquHlf = 0.5 * dot3(vel, vel);
  double c = sqrt(gamma * pp / rrho);
  double inv_kap_m1 = 1.0 / kapm1;
  A[0][0] = -surfVeloc;
  A[0][1] = nV[0];
  A[0][2] = nV[1];
  A[0][3] = nV[2];
  A[0][4] = 0.0;
  A[1][0] = -vel[0] * (nVel1 + nVel2) + nV[0] * (kapm1 * vSquHlf - vel[0] * vel[0]);
  A[1][1] = (2. - gamma) * nVel0 + U_k - surfVeloc;
  A[1][2] = vel[0] * nV[1] - kapm1 * vel[1] * nV[0];
  A[1][3] = vel[0] * nV[2] - kapm1 * vel[2] * nV[0];
  A[1][4] = kapm1 * nV[0];
  A[2][0] = -vel[1] * (nVel0 + nVel2) + nV[1] * (kapm1 * vSquHlf - vel[1] * vel[1]);
  A[2][1] = -kapm1 * vel[0] * nV[1] + vel[1] * nV[0];
  A[2][2] = (2. - gamma) * nVel1 + U_k - surfVeloc;
  A[2][3] = vel[1] * nV[2] - kapm1 * vel[2] * nV[1];
  A[2][4] = kapm1 * nV[1];
  A[3][0] = -vel[2] * (nVel0 + nVel1) + nV[2] * (kapm1 * vSquHlf - vel[2] * vel[2]);
  A[3][1] = -kapm1 * vel[0] * nV[2] + vel[2] * nV[0];
  A[3][2] = -kapm1 * vel[1] * nV[2] + vel[2] * nV[1];
  A[3][3] = (2. - gamma) * nVel2 + U_k - surfVeloc;
  A[3][4] = kapm1 * nV[2];
  A[4][0] = U_k * ((gamma - 2.) * vSquHlf - c * c * inv_kap_m1);
  A[4][1] = c * c * inv_kap_m1 * nV[0] - kapm1 * vel[0] * (nVel1 + nVel2) - (kapm1 * vel[0] * vel[0] - vSquHlf) * nV[0];
  A[4][2] = c * c * inv_kap_m1 * nV[1] - kapm1 * vel[1] * (nVel0 + nVel2) - (kapm1 * vel[1] * vel[1] - vSquHlf) * nV[1];
  A[4][3] = c * c * inv_kap_m1 * nV[2] - kapm1 * vel[2] * (nVel0 + nVel1) - (kapm1 * vel[2] * vel[2] - vSquHlf) * nV[2];
  A[4][4] = gamma * U_k - surfVeloc;
}

// One side of the HLLC Jacobian pair, UgpWithCvCompFlow.cpp:242-659.
// which = 0 computes A_L (d F / d U_L), which = 1 computes A_R. The reference
// computes the two independently (calcSubSonicJacobeanHLLC uses only the
// matching dSMdU/dpsdU), so evaluating one at a time gives the same bits.
__device__ __forceinline__ void jac_hllc(double J[5][5], int which, double rhoL, const double *uL, double pL, double rhoR,
                                      const double *uR, double pR, double gam, double area, const double *nVec,
                                      double kL = 0.0, double kR = 0.0) {
  const double gammaL = gam, gammaR = gam;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double hL = gammaL / (gammaL - 1.0) * pL / rhoL + 0.5 * uLuL + kL;
  double eL = hL * rhoL - pL;
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double hR = gammaR / (gammaR - 1.0) * pR / rhoR + 0.5 * uRuR + kR;
  double eR = hR * rhoR - pR;
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
#pragma unroll
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double gAvg = 0.5 * (gammaL + gammaR);

  // fully supersonic to one side: F = F(U_L) or F(U_R)
  bool left_sup = (sM >= 0.0) && (sL > 0.0);
  bool right_sup = (sM < 0.0) && !(sR >= 0.0);
  if (left_sup || right_sup) {
    bool active = left_sup ? (which == 0) : (which == 1);
    if (active) {
      double nVecArea[3] = {nVec[0] * area, nVec[1] * area, nVec[2] * area};
      if (which == 0)
        jac_a(J, uL, pL, rhoL, nVecArea, gAvg);
      else
        jac_a(J, uR, pR, rhoR, nVecArea, gAvg);
    } else {
#pragma unroll
      for (int i = 0; i < 5; i++)
#pragma unroll
        for (int j = 0; j < 5; j++) J[i][j] = 0.0;
    }
    return;
  }
  // star region on the side of the contact: K = L (sM >= 0) or R (sM < 0)
  const bool K_is_L = (sM >= 0.0);
  double rhoK = K_is_L ? rhoL : rhoR;
  const double *uK = K_is_L ? uL : uR;
  double pK = K_is_L ? pL : pR, eK = K_is_L ? eL : eR, unK = K_is_L ? unL : unR;
  double uKuK = K_is_L ? uLuL : uRuR, sK = K_is_L ? sL : sR;
  double invSKmSs = 1.0 / (sK - sM);
  double sKmuK = sK - unK;
  double rhoSK = rhoK * sKmuK * invSKmSs;
  double rhouSK[3];
#pragma unroll
  for (int i = 0; i < 3; i++) rhouSK[i] = (rhoK * uK[i] * sKmuK + (pStar - pK) * nVec[i]) * invSKmSs;
  double eSK = (sKmuK * eK - pK * unK + pStar * sM) * invSKmSs;
  double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
  // derivative of sM and p* with respect to the requested side
  double dSM[5], dps[5];
  if (which == 0) {
    double gammaLM1 = (gammaL - 1.0);
    dSM[0] = -unL * unL + uLuL * gammaLM1 / 2.0 + sM * sL;
    dSM[1] = nVec[0] * (2.0 * unL - sL - sM) - gammaLM1 * uL[0];
    dSM[2] = nVec[1] * (2.0 * unL - sL - sM) - gammaLM1 * uL[1];
    dSM[3] = nVec[2] * (2.0 * unL - sL - sM) - gammaLM1 * uL[2];
    dSM[4] = gammaLM1;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      dSM[i] *= invrhotld;
      dps[i] = rhoR * (sR - unR) * dSM[i];
    }
  } else {
    double gammaRM1 = (gammaR - 1.0);
    dSM[0] = unR * unR - uRuR * gammaRM1 / 2.0 - sM * sR;
    dSM[1] = -nVec[0] * (2.0 * unR - sR - sM) + gammaRM1 * uR[0];
    dSM[2] = -nVec[1] * (2.0 * unR - sR - sM) + gammaRM1 * uR[1];
    dSM[3] = -nVec[2] * (2.0 * unR - sR - sM) + gammaRM1 * uR[2];
    dSM[4] = -gammaRM1;
#pragma unroll
    for (int i = 0; i < 5; i++) {
      dSM[i] *= invrhotld;
      dps[i] = rhoL * (sL - unL) * dSM[i];
    }
  }
  // calcSubSonicJacobeanHLLC (UgpWithCvCompFlow.cpp:605-659) called with the
  // star side K first: the requested matrix is its "AL" when it is side K
  // (own-state terms), its "AR" otherwise.
  const double gamma = gAvg, gammaMinus1 = (gAvg - 1.0);
  const double *nV = nVec;
  double omL = 1.0 / (sK - sM);
  const bool own = (which == 0) == K_is_L;
  if (own) {
    const double qL = unK, psiL = uKuK, SL = sK, SM = sM, pS = pStar;
    const double *dSMdUL = dSM, *dpsdUL = dps;
    J[0][0] = SL + rhoSK * dSMdUL[0];
    J[0][1] = -nV[0] + rhoSK * dSMdUL[1];
    J[0][2] = -nV[1] + rhoSK * dSMdUL[2];
    J[0][3] = -nV[2] + rhoSK * dSMdUL[3];
    J[0][4] = +rhoSK * dSMdUL[4];
    J[1][0] = qL * uK[0] - nV[0] * psiL * gammaMinus1 / 2.0 + nV[0] * dpsdUL[0] + rhouSK[0] * dSMdUL[0];
    J[1][1] = SL - qL + nV[0] * (gamma - 2.0) * uK[0] + nV[0] * dpsdUL[1] + rhouSK[0] * dSMdUL[1];
    J[1][2] = -uK[0] * nV[1] + nV[0] * gammaMinus1 * uK[1] + nV[0] * dpsdUL[2] + rhouSK[0] * dSMdUL[2];
    J[1][3] = -uK[0] * nV[2] + nV[0] * gammaMinus1 * uK[2] + nV[0] * dpsdUL[3] + rhouSK[0] * dSMdUL[3];
    J[1][4] = -gammaMinus1 * nV[0] + nV[0] * dpsdUL[4] + rhouSK[0] * dSMdUL[4];
    J[2][0] = qL * uK[1] - nV[1] * psiL * gammaMinus1 / 2.0 + nV[1] * dpsdUL[0] + rhouSK[1] * dSMdUL[0];
    J[2][1] = -uK[1] * nV[0] + nV[1] * gammaMinus1 * uK[0] + nV[1] * dpsdUL[1] + rhouSK[1] * dSMdUL[1];
    J[2][2] = SL - qL + nV[1] * (gamma - 2.0) * uK[1] + nV[1] * dpsdUL[2] + rhouSK[1] * dSMdUL[2];
    J[2][3] = -uK[1] * nV[2] + nV[1] * gammaMinus1 * uK[2] + nV[1] * dpsdUL[3] + rhouSK[1] * dSMdUL[3];
    J[2][4] = -gammaMinus1 * nV[1] + nV[1] * dpsdUL[4] + rhouSK[1] * dSMdUL[4];
    J[3][0] = qL * uK[2] - nV[2] * psiL * gammaMinus1 / 2.0 + nV[2] * dpsdUL[0] + rhouSK[2] * dSMdUL[0];
    J[3][1] = -uK[2] * nV[0] + nV[2] * gammaMinus1 * uK[0] + nV[2] * dpsdUL[1] + rhouSK[2] * dSMdUL[1];
    J[3][2] = -uK[2] * nV[1] + nV[2] * gammaMinus1 * uK[1] + nV[2] * dpsdUL[2] + rhouSK[2] * dSMdUL[2];
    J[3][3] = SL - qL + nV[2] * (gamma - 2.0) * uK[2] + nV[2] * dpsdUL[3] + rhouSK[2] * dSMdUL[3];
    J[3][4] = -gammaMinus1 * nV[2] + nV[2] * dpsdUL[4] + rhouSK[2] * dSMdUL[4];
    J[4][0] = qL * (eK + pK) / rhoK - qL * psiL * (gamma - 1.0) / 2.0 + SM * dpsdUL[0] + (pS + eSK) * dSMdUL[0];
    J[4][1] = -nV[0] * (eK + pK) / rhoK + gammaMinus1 * uK[0] * qL + SM * dpsdUL[1] + (pS + eSK) * dSMdUL[1];
    J[4][2] = -nV[1] * (eK + pK) / rhoK + gammaMinus1 * uK[1] * qL + SM * dpsdUL[2] + (pS + eSK) * dSMdUL[2];
    J[4][3] = -nV[2] * (eK + pK) / rhoK + gammaMinus1 * uK[2] * qL + SM * dpsdUL[3] + (pS + eSK) * dSMdUL[3];
    J[4][4] = SL - qL * gamma + SM * dpsdUL[4] + (pS + eSK) * dSMdUL[4];
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) J[i][j] *= omL;
  } else {
#pragma unroll
    for (int i = 0; i < 5; i++) J[0][i] = omL * rhoSK * dSM[i];
#pragma unroll
    for (int i = 0; i < 5; i++) J[1][i] = omL * (nV[0] * dps[i] + rhouSK[0] * dSM[i]);
#pragma unroll
    for (int i = 0; i < 5; i++) J[2][i] = omL * (nV[1] * dps[i] + rhouSK[1] * dSM[i]);
#pragma unroll
    for (int i = 0; i < 5; i++) J[3][i] = omL * (nV[2] * dps[i] + rhouSK[2] * dSM[i]);
#pragma unroll
    for (int i = 0; i < 5; i++) J[4][i] = omL * (dps[i] * sM + (pStar + eSK) * dSM[i]);
  }
  // back in calcEulerFluxMatrices_HLLC (:361-383 / :492-514 and :415-437 / :523-545)
#pragma unroll
  for (int i = 0; i < 5; i++) J[0][i] = J[0][i] * sM + dSM[i] * rhoSK;
#pragma unroll
  for (int i = 0; i < 5; i++) J[1][i] = J[1][i] * sM + dSM[i] * rhouSK[0] + dps[i] * nVec[0];
#pragma unroll
  for (int i = 0; i < 5; i++) J[2][i] = J[2][i] * sM + dSM[i] * rhouSK[1] + dps[i] * nVec[1];
#pragma unroll
  for (int i = 0; i < 5; i++) J[3][i] = J[3][i] * sM + dSM[i] * rhouSK[2] + dps[i] * nVec[2];
#pragma unroll
  for (int i = 0; i < 5; i++) J[4][i] = (J[4][i] + dps[i]) * sM + (eSK + pStar) * dSM[i];
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) J[i][j] *= area;
}

// Scalar row of one side of the HLLC Jacobian pair (A_L_Scal / A_R_Scal,
// UgpWithCvCompFlow.cpp:296-311, 320-324, 389-405, 420-435, 504-519,
// 534-550, 561-577, 586-590): d(rho phi un)/dQ is row 0 of the 5x5 block
// times the upwind scalar, less the upwind star mass flux per unit density
// ("FrhostarOverrho") in column 0, which is also d/d(rho phi); the
// downwind side has no d/d(rho phi), and a side outside the fan is zero.
__device__ __forceinline__ void jac_hllc_scal(double S[6], int which, double rhoL, const double *uL, double pL,
                                           double rhoR, const double *uR, double pR, double gam, double area,
                                           const double *nVec, double phiL, double phiR, double kL = 0.0,
                                           double kR = 0.0) {
  double J[5][5];
  jac_hllc(J, which, rhoL, uL, pL, rhoR, uR, pR, gam, area, nVec, kL, kR);
  // the wave speeds again, with the same operations as jac_hllc (they do not
  // depend on the kinetic energy)
  const double gammaL = gam, gammaR = gam;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  int k;
  double fso;
  if (sM >= 0.0) {
    if (sL > 0.0) {
      k = 0;
      fso = unL * area;
    } else {
      k = 1;
      fso = 1.0 / (sL - sM) * (sL - unL) * sM * area;
    }
  } else {
    if (sR >= 0.0) {
      k = 2;
      fso = 1.0 / (sR - sM) * (sR - unR) * sM * area;
    } else {
      k = 3;
      fso = unR * area;
    }
  }
  const bool upL = k <= 1;
  const double phi = upL ? phiL : phiR;
  const bool zero = which == 0 ? (k == 3) : (k == 0);
  const bool upwind = which == 0 ? upL : !upL;
  if (zero) {
    for (int i = 0; i <= 5; i++) S[i] = 0.0;
  } else if (upwind) {
    S[0] = (J[0][0] - fso) * phi;
    for (int i = 1; i < 5; i++) S[i] = J[0][i] * phi;
    S[5] = fso;
  } else {
    for (int i = 0; i < 5; i++) S[i] = J[0][i] * phi;
    S[5] = 0.0;
  }
}

// the same for two transported scalars at once (one Jacobian, two rows)
__device__ __forceinline__ void jac_hllc_scal2(double S[6], double S2[6], int which, double rhoL,
                                            const double *uL, double pL, double rhoR, const double *uR,
                                            double pR, double gam, double area, const double *nVec,
                                            double phiL, double phiR, double phiL2, double phiR2,
                                            double kL, double kR) {
  double J[5][5];
  jac_hllc(J, which, rhoL, uL, pL, rhoR, uR, pR, gam, area, nVec, kL, kR);
  // the wave speeds again, with the same operations as jac_hllc (they do not
  // depend on the kinetic energy)
  const double gammaL = gam, gammaR = gam;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  int k;
  double fso;
  if (sM >= 0.0) {
    if (sL > 0.0) {
      k = 0;
      fso = unL * area;
    } else {
      k = 1;
      fso = 1.0 / (sL - sM) * (sL - unL) * sM * area;
    }
  } else {
    if (sR >= 0.0) {
      k = 2;
      fso = 1.0 / (sR - sM) * (sR - unR) * sM * area;
    } else {
      k = 3;
      fso = unR * area;
    }
  }
  const bool upL = k <= 1;
  const double phi = upL ? phiL : phiR, phi2 = upL ? phiL2 : phiR2;
  const bool zero = which == 0 ? (k == 3) : (k == 0);
  const bool upwind = which == 0 ? upL : !upL;
  // the three cases as jac_hllc_scal's, each row written once from selected
  // values (both rows by branch would keep one of them in scratch memory)
  const double j00 = upwind ? J[0][0] - fso : J[0][0];
#pragma unroll
  for (int i = 0; i < 5; i++) {
    const double ji = i == 0 ? j00 : J[0][i];
    S[i] = zero ? 0.0 : ji * phi;
    S2[i] = zero ? 0.0 : ji * phi2;
  }
  const double d = (zero || !upwind) ? 0.0 : fso;
  S[5] = d;
  S2[5] = d;
}

// One side of the Roe Jacobian pair, UgpWithCvCompFlow.cpp:3382-3559
// (A_L = A(U_L)/2 + P|L|P^-1 area/2, A_R = A(U_R)/2 - P|L|P^-1 area/2).
__device__ __forceinline__ void jac_roe(double J[5][5], int which, double rhoL, const double *uL, double pL, double rhoR,
                                     const double *uR, double pR, double gam, double area, const double *nVec) {
  const double gammaL = gam, gammaR = gam;
  double uLuL = dot3(uL, uL);
  double uRuR = dot3(uR, uR);
  double Rrho = sqrt(rhoR / rhoL);
  double rhoRoe = Rrho * rhoL;
  double tmp = 1.0 / (1.0 + Rrho);
  double vel[3];
#pragma unroll
  for (int i = 0; i < 3; i++) vel[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(vel, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double c = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(vel, vel));
  double nVecArea[3] = {nVec[0] * area, nVec[1] * area, nVec[2] * area};
  double Lam[5] = {uRoe, uRoe, uRoe, uRoe + c, uRoe - c};
  const double gamma = 0.5 * (gammaL + gammaR);
  const double *nV = nVec;
  const double rrho = rhoRoe;
  double kapm1 = (gamma - 1.0);
  double P[5][5], invP[5][5];
  {  // calcPMatrix, :3480-3521
    double rhooc = rrho / c;
    double rhoxc = rrho * c;
    double vSqu = dot3(vel, vel);
    P[0][0] = nV[0];
    P[0][1] = nV[1];
    P[0][2] = nV[2];
    P[0][3] = 0.5 * rhooc;
    P[0][4] = 0.5 * rhooc;
    P[1][0] = vel[0] * nV[0];
    P[1][1] = vel[0] * nV[1] - rrho * nV[2];
    P[1][2] = vel[0] * nV[2] + rrho * nV[1];
    P[1][3] = 0.5 * (vel[0] * rhooc + rrho * nV[0]);
    P[1][4] = 0.5 * (vel[0] * rhooc - rrho * nV[0]);
    P[2][0] = vel[1] * nV[0] + rrho * nV[2];
    P[2][1] = vel[1] * nV[1];
    P[2][2] = vel[1] * nV[2] - rrho * nV[0];
    P[2][3] = 0.5 * (vel[1] * rhooc + rrho * nV[1]);
    P[2][4] = 0.5 * (vel[1] * rhooc - rrho * nV[1]);
    P[3][0] = vel[2] * nV[0] - rrho * nV[1];
    P[3][1] = vel[2] * nV[1] + rrho * nV[0];
    P[3][2] = vel[2] * nV[2];
    P[3][3] = 0.5 * (vel[2] * rhooc + rrho * nV[2]);
    P[3][4] = 0.5 * (vel[2] * rhooc - rrho * nV[2]);
    P[4][0] = 0.5 * vSqu * nV[0] + rrho * vel[1] * nV[2] - rrho * vel[2] * nV[1];
    P[4][1] = 0.5 * vSqu * nV[1] - rrho * vel[0] * nV[2] + rrho * vel[2] * nV[0];
    P[4][2] = 0.5 * vSqu * nV[2] + rrho * vel[0] * nV[1] - rrho * vel[1] * nV[0];
    P[4][3] = 0.5 * (0.5 * vSqu * rhooc + rrho * (vel[0] * nV[0] + vel[1] * nV[1] + vel[2] * nV[2]) + rhoxc / kapm1);
    P[4][4] = 0.5 * (0.5 * vSqu * rhooc - rrho * (vel[0] * nV[0] + vel[1] * nV[1] + vel[2] * nV[2]) + rhoxc / kapm1);
  }
  {  // calcInvPMatrix, :3523-3559
    double rhoxc = rrho * c;
    double c2 = c * c;
    double vSqu = dot3(vel, vel);
    invP[0][0] = nV[0] - nV[2] * vel[1] / rrho + nV[1] * vel[2] / rrho - nV[0] * 0.5 * kapm1 * vSqu / c2;
    invP[0][1] = nV[0] * kapm1 * vel[0] / c2;
    invP[0][2] = nV[2] / rrho + nV[0] * kapm1 * vel[1] / c2;
    invP[0][3] = -nV[1] / rrho + nV[0] * kapm1 * vel[2] / c2;
    invP[0][4] = -nV[0] * kapm1 / c2;
    invP[1][0] = nV[1] + nV[2] * vel[0] / rrho - nV[0] * vel[2] / rrho - nV[1] * 0.5 * kapm1 * vSqu / c2;
    invP[1][1] = -nV[2] / rrho + nV[1] * kapm1 * vel[0] / c2;
    invP[1][2] = nV[1] * kapm1 * vel[1] / c2;
    invP[1][3] = nV[0] / rrho + nV[1] * kapm1 * vel[2] / c2;
    invP[1][4] = -nV[1] * kapm1 / c2;
    invP[2][0] = nV[2] - nV[1] * vel[0] / rrho + nV[0] * vel[1] / rrho - nV[2] * 0.5 * kapm1 * vSqu / c2;
    invP[2][1] = nV[1] / rrho + nV[2] * kapm1 * vel[0] / c2;
    invP[2][2] = -nV[0] / rrho + nV[2] * kapm1 * vel[1] / c2;
    invP[2][3] = nV[2] * kapm1 * vel[2] / c2;
    invP[2][4] = -nV[2] * kapm1 / c2;
    invP[3][0] = -(nV[0] * vel[0] + nV[1] * vel[1] + nV[2] * vel[2]) / rrho + 0.5 * kapm1 * vSqu / rhoxc;
    invP[3][1] = nV[0] / rrho - kapm1 * vel[0] / rhoxc;
    invP[3][2] = nV[1] / rrho - kapm1 * vel[1] / rhoxc;
    invP[3][3] = nV[2] / rrho - kapm1 * vel[2] / rhoxc;
    invP[3][4] = kapm1 / rhoxc;
    invP[4][0] = (nV[0] * vel[0] + nV[1] * vel[1] + nV[2] * vel[2]) / rrho + 0.5 * kapm1 * vSqu / rhoxc;
    invP[4][1] = -nV[0] / rrho - kapm1 * vel[0] / rhoxc;
    invP[4][2] = -nV[1] / rrho - kapm1 * vel[1] / rhoxc;
    invP[4][3] = -nV[2] / rrho - kapm1 * vel[2] / rhoxc;
    invP[4][4] = kapm1 / rhoxc;
  }
  if (which == 0)
    jac_a(J, uL, pL, rhoL, nVecArea, gamma);
  else
    jac_a(J, uR, pR, rhoR, nVecArea, gamma);
#pragma unroll
  for (int i = 0; i < 5; i++)
#pragma unroll
    for (int j = 0; j < 5; j++) {
      double pm = 0.0;
#pragma unroll
      for (int k = 0; k < 5; k++) pm += P[i][k] * fabs(Lam[k]) * invP[k][j];
      J[i][j] = which == 0 ? 0.5 * J[i][j] + 0.5 * pm * area : 0.5 * J[i][j] - 0.5 * pm * area;
    }
}

// The coupled HLLC of runBackwardEulerCoupled: calcEulerFluxCoupled_HLLC
// (UgpWithCvCompFlow.cpp:1808-1927) with kine = 0. The energy is built from
// the given side enthalpies h0, h1 (calcThermoProp_p's, or the boundary
// state's), not from gamma/(gamma-1) p/rho as flux_hllc does. With Fs, one
// transported scalar (ConvTerm = 1) gets the upwind side's value (scL or scR
// by the sign of sM) times the unscaled mass flux, then times the area
// (:1851-1896, 1926).
__device__ __forceinline__ void flux_hllc_coupled(double F[5], double rhoL, const double *uL, double pL, double h0,
                                                  double rhoR, const double *uR, double pR, double h1, double gam,
                                                  double area, const double *nVec, double scL = 0.0,
                                                  double scR = 0.0, double *Fs = nullptr) {
  const double gammaL = gam, gammaR = gam, kL = 0.0, kR = 0.0;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double rhoeL = (h0 + 0.5 * uLuL + kL) * rhoL - pL;
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double rhoeR = (h1 + 0.5 * uRuR + kR) * rhoR - pR;
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
#pragma unroll
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  // the upwind side and whether the star state is used, chosen per lane
  // before one evaluation
  const bool left = sM >= 0.0;
  const bool star = left ? !(sL > 0.0) : (sR >= 0.0);
  const double rho = left ? rhoL : rhoR, p = left ? pL : pR, un = left ? unL : unR, rhoe = left ? rhoeL : rhoeR;
  const double S = left ? sL : sR;
  // values, not a pointer, are selected: a pointer chosen between the two
  // sides' arrays would put both in scratch memory
  const double u[3] = {left ? uL[0] : uR[0], left ? uL[1] : uR[1], left ? uL[2] : uR[2]};
  if (!star) {
    F[0] = rho * un;
#pragma unroll
    for (int i = 0; i < 3; i++) F[1 + i] = rho * u[i] * un + p * nVec[i];
    F[4] = (rhoe + p) * un;
  } else {
    double inv = 1.0 / (S - sM);
    double smu = S - un;
    double rhoS = rho * smu * inv;
    double rhouS[3];
#pragma unroll
    for (int i = 0; i < 3; i++) rhouS[i] = (rho * u[i] * smu + (pStar - p) * nVec[i]) * inv;
    double rhoeS = (smu * rhoe - p * un + pStar * sM) * inv;
    F[0] = rhoS * sM;
#pragma unroll
    for (int i = 0; i < 3; i++) F[1 + i] = rhouS[i] * sM + pStar * nVec[i];
    F[4] = (rhoeS + pStar) * sM;
  }
  if (Fs) *Fs = (1.0 * F[0] * (left ? scL : scR)) * area;
#pragma unroll
  for (int i = 0; i < 5; i++) F[i] *= area;
}

// calcJacobianCoupled_HLLC, UgpWithCvCompFlow.cpp:2180-2241 (nV not normalised)
__device__ __forceinline__ void jac_coupled_super(double A[5][5], const double *vel, double pp, double rrho, double hh,
                                                  const double *nV, double gamma) {
  double kapm1 = (gamma - 1.0);
  double nVel[3] = {vel[0] * nV[0], vel[1] * nV[1], vel[2] * nV[2]};
  double U_k = nVel[0] + nVel[1] + nVel[2];
  double vSquHlf = 0.5 * dot3(vel, vel);
  A[0][0] = 0.0;
  A[0][1] = nV[0];
  A[0][2] = nV[1];
  A[0][3] = nV[2];
  A[0][4] = 0.0;
#pragma unroll
  for (int r = 0; r < 3; r++) {
    A[1 + r][0] = -vel[r] * U_k + nV[r] * (kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh);
#pragma unroll
    for (int c = 0; c < 3; c++)
      A[1 + r][1 + c] = r == c ? (2.0 - gamma) * nVel[r] + U_k : vel[r] * nV[c] - kapm1 * vel[c] * nV[r];
    A[1 + r][4] = kapm1 * nV[r];
  }
  A[4][0] = U_k * ((gamma - 2.0) * vSquHlf - gamma * (hh - pp / rrho));
#pragma unroll
  for (int c = 0; c < 3; c++) A[4][1 + c] = (hh + vSquHlf) * nV[c] - kapm1 * vel[c] * U_k;
  A[4][4] = gamma * U_k;
}

// dF/dQ of one side of a face for the coupled HLLC (which 0: A_L, 1: A_R),
// calcEulerFluxMatricesCoupled_HLLC (UgpWithCvCompFlow.cpp:1934-2174) with
// calcdSMdU_HLLC and calcSubSonicJacobianCoupled_HLLC (:2246-2287), nScal = 0,
// kine = 0. In every subsonic case the side's own dSM/dU is used with the
// other side's rho (S - u_n), the star state is the upwind side's, and the
// supersonic part and diagonal term belong to the upwind side only.
__device__ __forceinline__ void jac_hllc_coupled(double J[5][5], int which, double rhoL, const double *uL, double pL,
                                              double h0, double rhoR, const double *uR, double pR, double h1,
                                              double gam, double area, const double *nVec) {
  const double gammaL = gam, gammaR = gam;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double hL = h0 + 0.0;
  double rhoeL = (hL + 0.5 * uLuL) * rhoL - pL;
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double hR = h1 + 0.0;
  double rhoeR = (hR + 0.5 * uRuR) * rhoR - pR;
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
#pragma unroll
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double nVecArea[3];
#pragma unroll
  for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
  const bool left = sM >= 0.0;  // upwind side
  const bool sup = left ? (sL > 0.0) : !(sR >= 0.0);
  const bool own = (which == 0) == left;  // this side is the upwind one
  // this side's state (values selected, not a pointer: see flux_hllc_coupled)
  const double u[3] = {which == 0 ? uL[0] : uR[0], which == 0 ? uL[1] : uR[1], which == 0 ? uL[2] : uR[2]};
  const double p = which == 0 ? pL : pR, rho = which == 0 ? rhoL : rhoR, h = which == 0 ? hL : hR;
  if (own) {
    jac_coupled_super(J, u, p, rho, h, nVecArea, gam);
  } else {
#pragma unroll
    for (int i = 0; i < 5; i++)
#pragma unroll
      for (int j = 0; j < 5; j++) J[i][j] = 0.0;
  }
  if (sup) return;
  // the upwind side's star state
  const double uu[3] = {left ? uL[0] : uR[0], left ? uL[1] : uR[1], left ? uL[2] : uR[2]};
  const double Su = left ? sL : sR, unu = left ? unL : unR, rhou_ = left ? rhoL : rhoR, pu = left ? pL : pR;
  const double rhoeu = left ? rhoeL : rhoeR;
  double Omega = 1.0 / (Su - sM);
  double smu = Su - unu;
  double rhoS = rhou_ * smu * Omega;
  double rhouS[3];
#pragma unroll
  for (int i = 0; i < 3; i++) rhouS[i] = (rhou_ * uu[i] * smu + (pStar - pu) * nVec[i]) * Omega;
  double rhoeS = (smu * rhoeu - pu * unu + pStar * sM) * Omega;
  // this side's dSM/dU (calcdSMdU_HLLC with nV the unit normal)
  const double S = which == 0 ? sL : sR, Factor = which == 0 ? 1.0 : -1.0;
  double d[5];
  {
    double kapm1 = (gam - 1.0);
    double nVel[3] = {u[0] * nVec[0], u[1] * nVec[1], u[2] * nVec[2]};
    double U_k = nVel[0] + nVel[1] + nVel[2];
    double vSquHlf = 0.5 * dot3(u, u);
    d[0] = Factor * invrhotld * (-U_k * U_k + kapm1 * vSquHlf + gam * p / rho - kapm1 * h + sM * S);
#pragma unroll
    for (int i = 0; i < 3; i++) d[1 + i] = Factor * invrhotld * ((2 * U_k - S - sM) * nVec[i] - kapm1 * u[i]);
    d[4] = Factor * invrhotld * kapm1;
  }
  // the other side's rho (S - u_n)
  const double rhoSmq = which == 0 ? rhoR * (sR - unR) : rhoL * (sL - unL);
  double OmSM = Omega * sM;
  double OmSMp1 = OmSM + 1.0;
#pragma unroll
  for (int i = 0; i < 5; i++) {
    J[0][i] = -OmSM * J[0][i] + area * OmSMp1 * d[i] * rhoS;
    J[1][i] = -OmSM * J[1][i] + area * OmSMp1 * d[i] * (rhouS[0] + rhoSmq * nVec[0]);
    J[2][i] = -OmSM * J[2][i] + area * OmSMp1 * d[i] * (rhouS[1] + rhoSmq * nVec[1]);
    J[3][i] = -OmSM * J[3][i] + area * OmSMp1 * d[i] * (rhouS[2] + rhoSmq * nVec[2]);
    J[4][i] = -OmSM * J[4][i] + area * OmSMp1 * d[i] * (rhoeS + pStar + rhoSmq * sM);
  }
  if (own)
#pragma unroll
    for (int i = 0; i < 5; i++) J[i][i] += area * Omega * sM * S;
}

// The transported scalar's row of the coupled HLLC Jacobian of one face side
// (which 0: A_L[5][0..5], 1: A_R[5][0..5]): calcEulerFluxMatricesCoupled_HLLC
// (UgpWithCvCompFlow.cpp:1934-2174) with one scalar and dpress_dscal = 0
// (RansTurbSA leaves pressureDerivativeHookScalarRansTurb empty,
// UgpWithCvCompFlow.h:1590): the supersonic row of calcJacobianCoupled_HLLC
// (:2233-2240), the subsonic transformation (:2281-2284) with the star-state
// terms (:2022-2026, 2074-2078), then dU/dQ (:2150-2174). With dpress_dscal = 0
// the flow rows keep a zero scalar column and their column 0 unchanged, so
// only this row is new. The same side selection as jac_hllc_coupled.
__device__ __forceinline__ void jac_hllc_coupled_srow(double S[6], int which, double rhoL, const double *uL,
                                                      double pL, double h0, double scL, double rhoR,
                                                      const double *uR, double pR, double h1, double scR,
                                                      double gam, double area, const double *nVec) {
  const double gammaL = gam, gammaR = gam;
  double unL = dot3(uL, nVec);
  double uLuL = dot3(uL, uL);
  double cL = sqrt(gammaL * pL / rhoL);
  double hL = h0 + 0.0;
  double unR = dot3(uR, nVec);
  double uRuR = dot3(uR, uR);
  double cR = sqrt(gammaR * pR / rhoR);
  double hR = h1 + 0.0;
  double Rrho = sqrt(rhoR / rhoL);
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
#pragma unroll
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = Dev::mn(uRoe - cRoe, unL - cL);
  double sR = Dev::mx(uRoe + cRoe, unR + cR);
  double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double nVecArea[3];
#pragma unroll
  for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
  const bool left = sM >= 0.0;  // upwind side
  const bool sup = left ? (sL > 0.0) : !(sR >= 0.0);
  const bool own = (which == 0) == left;
  const double u[3] = {which == 0 ? uL[0] : uR[0], which == 0 ? uL[1] : uR[1], which == 0 ? uL[2] : uR[2]};
  const double p = which == 0 ? pL : pR, rho = which == 0 ? rhoL : rhoR, h = which == 0 ? hL : hR;
  const double sc = which == 0 ? scL : scR;
  if (own) {  // calcJacobianCoupled_HLLC's scalar row, nV = nVecArea
    double nVel[3] = {u[0] * nVecArea[0], u[1] * nVecArea[1], u[2] * nVecArea[2]};
    double U_k = nVel[0] + nVel[1] + nVel[2];
    S[0] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; i++) S[1 + i] = sc * nVecArea[i];
    S[4] = 0.0;
    S[5] = rho * U_k;
  } else {
#pragma unroll
    for (int j = 0; j < 6; j++) S[j] = 0.0;
  }
  if (!sup) {
    const double Su = left ? sL : sR, unu = left ? unL : unR, rhou_ = left ? rhoL : rhoR, scu = left ? scL : scR;
    double Omega = 1.0 / (Su - sM);
    double rhoS = rhou_ * (Su - unu) * Omega;
    // this side's dSM/dU (calcdSMdU_HLLC with nV the unit normal; the
    // scalar entry is Factor * invrhotld * dpress_dscal = 0)
    const double Sd = which == 0 ? sL : sR, Factor = which == 0 ? 1.0 : -1.0;
    double d[6];
    {
      double kapm1 = (gam - 1.0);
      double nVel[3] = {u[0] * nVec[0], u[1] * nVec[1], u[2] * nVec[2]};
      double U_k = nVel[0] + nVel[1] + nVel[2];
      double vSquHlf = 0.5 * dot3(u, u);
      d[0] = Factor * invrhotld * (-U_k * U_k + kapm1 * vSquHlf + gam * p / rho - kapm1 * h + sM * Sd);
#pragma unroll
      for (int i = 0; i < 3; i++) d[1 + i] = Factor * invrhotld * ((2 * U_k - Sd - sM) * nVec[i] - kapm1 * u[i]);
      d[4] = Factor * invrhotld * kapm1;
      d[5] = Factor * invrhotld * 0.0;
    }
    double OmSM = Omega * sM;
    double OmSMp1 = OmSM + 1.0;
#pragma unroll
    for (int i = 0; i < 6; i++) S[i] = -OmSM * S[i] + area * OmSMp1 * d[i] * rhoS * scu;
    if (own) {
      S[0] += area * Omega * sM * Su * sc;
      S[5] += area * Omega * sM * Su * rho;
    }
  }
  // dU/dQ of this side (ConvTerm = 1)
#pragma unroll
  for (int j = 0; j < 6; j++) S[j] *= 1.0;
  S[0] -= sc / rho * S[5];
  S[5] /= rho;
}

// The transported scalar's row of calcViscousFluxCoupled (UgpWithCvCompFlow.cpp
// :3113-3120 flux, :3158-3167 and :3217-3226 dF/dU, :3185-3190 and :3244-3249
// dU/dQ): Fs and the row's entries [5][0], [5][5] of side 0 (a0) and side 1 (a1)
__device__ __forceinline__ double visc_coupled_srow(double a0[2], double a1[2], double rho0, double scal0,
                                                    const double *gs0, double rho1, double scal1, const double *gs1,
                                                    double diff, double DiffTerm, double area, const double *nVec,
                                                    double smag, const double *sVec, double alpha) {
  double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  double Fs = -area * DiffTerm * diff *
              (alpha * (scal1 - scal0) / smag +
               0.5 * ((gs0[0] + gs1[0]) * fCorr[0] + (gs0[1] + gs1[1]) * fCorr[1] + (gs0[2] + gs1[2]) * fCorr[2]));
  double area_alpha_smag = area * alpha / smag;
  {
    double dFdU = DiffTerm * diff * area_alpha_smag;
    double invRho = 1.0 / rho0;
    a0[0] = dFdU * (-invRho * scal0);
    a0[1] = dFdU * invRho;
  }
  {
    double dFdU = -DiffTerm * diff * area_alpha_smag;
    double invRho = 1.0 / rho1;
    a1[0] = dFdU * (-invRho * scal1);
    a1[1] = dFdU * invRho;
  }
  return Fs;
}

// The ONLY_PRESSURE Jacobian of a symmetry or wall face (A_L), coupled form
// (UgpWithCvCompFlow.cpp:2091-2137)
__device__ __forceinline__ void jac_only_pressure(double J[5][5], double rhoL, const double *uL, double pL,
                                                  double h0, double gam, double area, const double *nVec) {
  double uLuL = dot3(uL, uL);
  double hL = h0 + 0.0;
#pragma unroll
  for (int j = 0; j < 5; j++) J[0][j] = 0.0;
  double dpdrho = 0.5 * (gam - 1.0) * uLuL - (gam - 1.0) * hL + gam * pL / rhoL;
#pragma unroll
  for (int r = 1; r <= 3; r++) {
    J[r][0] = dpdrho * nVec[r - 1] * area;
    J[r][1] = -(gam - 1.0) * uL[0] * nVec[r - 1] * area;
    J[r][2] = -(gam - 1.0) * uL[1] * nVec[r - 1] * area;
    J[r][3] = -(gam - 1.0) * uL[2] * nVec[r - 1] * area;
    J[r][4] = (gam - 1.0) * nVec[r - 1] * area;
  }
#pragma unroll
  for (int j = 0; j < 5; j++) J[4][j] = 0.0;
}

// ludeco / lusolv on an NB x NB block (UgpWithCv.cpp:2489-2540; NB = 5 for
// the flow equations, 5 + nScal in the coupled loop; T = double or float),
// in the reference's loop order
template <typename T, int NB>
__device__ __forceinline__ void ludeco_n(T A[NB][NB]) {
#pragma unroll
  for (int jc = 1; jc < NB; jc++) A[0][jc] /= A[0][0];
#pragma unroll
  for (int jrjc = 1; jrjc < NB; jrjc++) {
#pragma unroll
    for (int jr = jrjc; jr < NB; jr++) {
      T sum = A[jr][jrjc];
#pragma unroll
      for (int jm = 0; jm < jrjc; jm++) sum -= A[jr][jm] * A[jm][jrjc];
      A[jr][jrjc] = sum;
    }
    if (jrjc == NB - 1) break;
#pragma unroll
    for (int jc = jrjc + 1; jc < NB; jc++) {
      T sum = A[jrjc][jc];
#pragma unroll
      for (int jm = 0; jm < jrjc; jm++) sum -= A[jrjc][jm] * A[jm][jc];
      A[jrjc][jc] = sum / A[jrjc][jrjc];
    }
  }
}

template <typename T, int NB>
__device__ __forceinline__ void lusolv_n(const T *A /* [NB * NB] */, T c[NB]) {
  c[0] = c[0] / A[0];
#pragma unroll
  for (int jr = 1; jr < NB; jr++) {
    T sum = c[jr];
#pragma unroll
    for (int jm = 0; jm < jr; jm++) sum -= A[NB * jr + jm] * c[jm];
    c[jr] = sum / A[NB * jr + jr];
  }
#pragma unroll
  for (int jrjr = 1; jrjr < NB; jrjr++) {
    int jr = (NB - 1) - jrjr;
    T sum = c[jr];
#pragma unroll
    for (int jmjm = jr + 1; jmjm < NB; jmjm++) {
      int jm = (NB - 1) - jmjm + jr + 1;
      sum -= A[NB * jr + jm] * c[jm];
    }
    c[jr] = sum;
  }
}

template <typename T>
__device__ __forceinline__ void ludeco5(T A[5][5]) {
  ludeco_n<T, 5>(A);
}

template <typename T>
__device__ __forceinline__ void lusolv5(const T *A /* [25] */, T c[5]) {
  lusolv_n<T, 5>(A, c);
}

}  // namespace jm
