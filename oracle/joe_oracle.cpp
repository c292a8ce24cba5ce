// CPU restatement of Joe's steady-solve hot path -- TEST INFRASTRUCTURE ONLY.
// See joe_oracle.h for who may load it. Each routine restates the reference
// routine named at its definition, in the reference's loop structure (face
// loops scattering into cells, sequential sums), so that compiled like the
// reference (x86-64 SSE2, no contraction) it reproduces the reference's bits.
#include "joe_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

using std::max;
using std::min;

namespace {

inline double dot3(const double *a, const double *b) {  // tc_vec3d.h:14-22
  double d = 0.0;
  for (int i = 0; i < 3; i++) d += a[i] * b[i];
  return d;
}
inline double norm3(double *n, const double *v) {  // tc_vec3d.h:36-46
  double mag = sqrt(dot3(v, v));
  for (int i = 0; i < 3; i++) n[i] = v[i] / mag;
  return mag;
}

// ---------------------------------------------------------------------------
// pointwise flux functions
// ---------------------------------------------------------------------------

// HLLC, UgpWithCvCompFlow.cpp:120-235 (surfVeloc = 0). With Fs, nsc
// transported scalars: FrhoScal[s] = Frho * (upwind side value) * area.
void flux_hllc(double F[5], double rhoL, const double *uL, double pL, double gammaL, double kL, double rhoR,
               const double *uR, double pR, double gammaR, double kR, double area, const double *nVec,
               const double *phiL = nullptr, const double *phiR = nullptr, double *Fs = nullptr, int nsc = 0) {
  double surfv = 0.0;
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
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = min(uRoe - cRoe, unLr - cL);
  double sR = max(uRoe + cRoe, unRr + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double Frho, Frhou[3], FrhoE;
  bool left;
  if (sM >= 0.0) {
    left = true;
    if (sL > 0.0) {
      Frho = rhoL * unLr;
      for (int i = 0; i < 3; i++) Frhou[i] = rhoL * uL[i] * unLr + pL * nVec[i];
      FrhoE = eL * unLr + pL * unL;
    } else {
      double invSLmSs = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * invSLmSs;
      double rhouSL[3];
      for (int i = 0; i < 3; i++) rhouSL[i] = (rhoL * uL[i] * sLmuL + (pStar - pL) * nVec[i]) * invSLmSs;
      double eSL = (sLmuL * eL - pL * unL + pStar * sM) * invSLmSs;
      Frho = rhoSL * sM;
      for (int i = 0; i < 3; i++) Frhou[i] = rhouSL[i] * sM + pStar * nVec[i];
      FrhoE = (eSL + pStar) * sM;
    }
  } else {
    left = false;
    if (sR >= 0.0) {
      double invSRmSs = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * invSRmSs;
      double rhouSR[3];
      for (int i = 0; i < 3; i++) rhouSR[i] = (rhoR * uR[i] * sRmuR + (pStar - pR) * nVec[i]) * invSRmSs;
      double eSR = (sRmuR * eR - pR * unR + pStar * sM) * invSRmSs;
      Frho = rhoSR * sM;
      for (int i = 0; i < 3; i++) Frhou[i] = rhouSR[i] * sM + pStar * nVec[i];
      FrhoE = (eSR + pStar) * sM;
    } else {
      Frho = rhoR * unRr;
      for (int i = 0; i < 3; i++) Frhou[i] = rhoR * uR[i] * unRr + pR * nVec[i];
      FrhoE = eR * unRr + pR * unR;
    }
  }
  F[0] = Frho * area;
  for (int i = 0; i < 3; i++) F[1 + i] = Frhou[i] * area;
  F[4] = FrhoE * area;
  if (Fs)
    for (int s = 0; s < nsc; s++) {
      double FrhoScal = Frho * (left ? phiL[s] : phiR[s]);
      Fs[s] = FrhoScal * area;
    }
}

// Roe, UgpWithCvCompFlow.cpp:3281-3380. When the Roe-averaged a^2 < 1e-12 the
// reference returns after the central part (:3318-3319); so does this.
void flux_roe(double F[5], double rhoL, const double *uL, double pL, double gammaL, double rhoR, const double *uR,
              double pR, double gammaR, double area, const double *nVec) {
  double rho0 = rhoL, rho1 = rhoR;
  const double *u0 = uL, *u1 = uR;
  double p0 = pL, p1 = pR;
  double gamma = 0.5 * (gammaL + gammaR);
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
void flux_ausm(double F[5], double rhoL, const double *uL, double pL, double gammaL, double kL, double rhoR,
               const double *uR, double pR, double gammaR, double kR, double area, const double *nVec) {
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
  F[0] = 0.5 * (mF * ((rhoL * cL) + (rhoR * cR)) - Phi * ((rhoR * cR) - (rhoL * cL)));
  for (int i = 0; i < 3; i++)
    F[1 + i] = 0.5 * (mF * ((rhoL * cL * uL[i]) + (rhoR * cR * uR[i])) - Phi * ((rhoR * cR * uR[i]) - (rhoL * cL * uL[i]))) +
               nVec[i] * pF;
  F[4] = 0.5 * (mF * ((rhoL * cL * hL) + (rhoR * cR * hR)) - Phi * ((rhoR * cR * hR) - (rhoL * cL * hL)));
  for (int i = 0; i < 5; i++) F[i] *= area;
}

// calcJacobianA, UgpWithCvCompFlow.cpp:666-709 (nV area-weighted)
void jac_a(double A[5][5], const double *vel, double pp, double rrho, const double *nV, double gamma) {
  double surfVeloc = 0.0;
  double kapm1 = (gamma - 1.0);
  double nVel[3] = {vel[0] * nV[0], vel[1] * nV[1], vel[2] * nV[2]};
  double U_k = nVel[0] + nVel[1] + nVel[2];
  double vSquHlf = 0.5 * dot3(vel, vel);
  double c = sqrt(gamma * pp / rrho);
  double inv_kap_m1 = 1.0 / kapm1;
  A[0][0] = -surfVeloc;
  A[0][1] = nV[0];
  A[0][2] = nV[1];
  A[0][3] = nV[2];
  A[0][4] = 0.0;
  A[1][0] = -vel[0] * (nVel[1] + nVel[2]) + nV[0] * (kapm1 * vSquHlf - vel[0] * vel[0]);
  A[1][1] = (2. - gamma) * nVel[0] + U_k - surfVeloc;
  A[1][2] = vel[0] * nV[1] - kapm1 * vel[1] * nV[0];
  A[1][3] = vel[0] * nV[2] - kapm1 * vel[2] * nV[0];
  A[1][4] = kapm1 * nV[0];
  A[2][0] = -vel[1] * (nVel[0] + nVel[2]) + nV[1] * (kapm1 * vSquHlf - vel[1] * vel[1]);
  A[2][1] = -kapm1 * vel[0] * nV[1] + vel[1] * nV[0];
  A[2][2] = (2. - gamma) * nVel[1] + U_k - surfVeloc;
  A[2][3] = vel[1] * nV[2] - kapm1 * vel[2] * nV[1];
  A[2][4] = kapm1 * nV[1];
  A[3][0] = -vel[2] * (nVel[0] + nVel[1]) + nV[2] * (kapm1 * vSquHlf - vel[2] * vel[2]);
  A[3][1] = -kapm1 * vel[0] * nV[2] + vel[2] * nV[0];
  A[3][2] = -kapm1 * vel[1] * nV[2] + vel[2] * nV[1];
  A[3][3] = (2. - gamma) * nVel[2] + U_k - surfVeloc;
  A[3][4] = kapm1 * nV[2];
  A[4][0] = U_k * ((gamma - 2.) * vSquHlf - c * c * inv_kap_m1);
  A[4][1] = c * c * inv_kap_m1 * nV[0] - kapm1 * vel[0] * (nVel[1] + nVel[2]) - (kapm1 * vel[0] * vel[0] - vSquHlf) * nV[0];
  A[4][2] = c * c * inv_kap_m1 * nV[1] - kapm1 * vel[1] * (nVel[0] + nVel[2]) - (kapm1 * vel[1] * vel[1] - vSquHlf) * nV[1];
  A[4][3] = c * c * inv_kap_m1 * nV[2] - kapm1 * vel[2] * (nVel[0] + nVel[1]) - (kapm1 * vel[2] * vel[2] - vSquHlf) * nV[2];
  A[4][4] = gamma * U_k - surfVeloc;
}

// calcSubSonicJacobeanHLLC, UgpWithCvCompFlow.cpp:605-659
void jac_hllc_subsonic(double (*AL)[5], double (*AR)[5], double rhoL, const double *uL, double pL, double eL,
                       double qL, double psiL, double SL, double rhoSL, const double *rhouSL, double eSL,
                       const double *dSMdUL, const double *dSMdUR, const double *dpsdUL, const double *dpsdUR,
                       double SM, double pS, double gamma, const double *nV) {
  double gammaMinus1 = (gamma - 1.0);
  double omL = 1.0 / (SL - SM);
  if (AL != NULL) {
    AL[0][0] = SL + rhoSL * dSMdUL[0];
    AL[0][1] = -nV[0] + rhoSL * dSMdUL[1];
    AL[0][2] = -nV[1] + rhoSL * dSMdUL[2];
    AL[0][3] = -nV[2] + rhoSL * dSMdUL[3];
    AL[0][4] = +rhoSL * dSMdUL[4];
    AL[1][0] = qL * uL[0] - nV[0] * psiL * gammaMinus1 / 2.0 + nV[0] * dpsdUL[0] + rhouSL[0] * dSMdUL[0];
    AL[1][1] = SL - qL + nV[0] * (gamma - 2.0) * uL[0] + nV[0] * dpsdUL[1] + rhouSL[0] * dSMdUL[1];
    AL[1][2] = -uL[0] * nV[1] + nV[0] * gammaMinus1 * uL[1] + nV[0] * dpsdUL[2] + rhouSL[0] * dSMdUL[2];
    AL[1][3] = -uL[0] * nV[2] + nV[0] * gammaMinus1 * uL[2] + nV[0] * dpsdUL[3] + rhouSL[0] * dSMdUL[3];
    AL[1][4] = -gammaMinus1 * nV[0] + nV[0] * dpsdUL[4] + rhouSL[0] * dSMdUL[4];
    AL[2][0] = qL * uL[1] - nV[1] * psiL * gammaMinus1 / 2.0 + nV[1] * dpsdUL[0] + rhouSL[1] * dSMdUL[0];
    AL[2][1] = -uL[1] * nV[0] + nV[1] * gammaMinus1 * uL[0] + nV[1] * dpsdUL[1] + rhouSL[1] * dSMdUL[1];
    AL[2][2] = SL - qL + nV[1] * (gamma - 2.0) * uL[1] + nV[1] * dpsdUL[2] + rhouSL[1] * dSMdUL[2];
    AL[2][3] = -uL[1] * nV[2] + nV[1] * gammaMinus1 * uL[2] + nV[1] * dpsdUL[3] + rhouSL[1] * dSMdUL[3];
    AL[2][4] = -gammaMinus1 * nV[1] + nV[1] * dpsdUL[4] + rhouSL[1] * dSMdUL[4];
    AL[3][0] = qL * uL[2] - nV[2] * psiL * gammaMinus1 / 2.0 + nV[2] * dpsdUL[0] + rhouSL[2] * dSMdUL[0];
    AL[3][1] = -uL[2] * nV[0] + nV[2] * gammaMinus1 * uL[0] + nV[2] * dpsdUL[1] + rhouSL[2] * dSMdUL[1];
    AL[3][2] = -uL[2] * nV[1] + nV[2] * gammaMinus1 * uL[1] + nV[2] * dpsdUL[2] + rhouSL[2] * dSMdUL[2];
    AL[3][3] = SL - qL + nV[2] * (gamma - 2.0) * uL[2] + nV[2] * dpsdUL[3] + rhouSL[2] * dSMdUL[3];
    AL[3][4] = -gammaMinus1 * nV[2] + nV[2] * dpsdUL[4] + rhouSL[2] * dSMdUL[4];
    AL[4][0] = qL * (eL + pL) / rhoL - qL * psiL * (gamma - 1.0) / 2.0 + SM * dpsdUL[0] + (pS + eSL) * dSMdUL[0];
    AL[4][1] = -nV[0] * (eL + pL) / rhoL + gammaMinus1 * uL[0] * qL + SM * dpsdUL[1] + (pS + eSL) * dSMdUL[1];
    AL[4][2] = -nV[1] * (eL + pL) / rhoL + gammaMinus1 * uL[1] * qL + SM * dpsdUL[2] + (pS + eSL) * dSMdUL[2];
    AL[4][3] = -nV[2] * (eL + pL) / rhoL + gammaMinus1 * uL[2] * qL + SM * dpsdUL[3] + (pS + eSL) * dSMdUL[3];
    AL[4][4] = SL - qL * gamma + SM * dpsdUL[4] + (pS + eSL) * dSMdUL[4];
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) AL[i][j] *= omL;
  }
  if (AR != NULL) {
    for (int i = 0; i < 5; i++) AR[0][i] = omL * rhoSL * dSMdUR[i];
    for (int i = 0; i < 5; i++) AR[1][i] = omL * (nV[0] * dpsdUR[i] + rhouSL[0] * dSMdUR[i]);
    for (int i = 0; i < 5; i++) AR[2][i] = omL * (nV[1] * dpsdUR[i] + rhouSL[1] * dSMdUR[i]);
    for (int i = 0; i < 5; i++) AR[3][i] = omL * (nV[2] * dpsdUR[i] + rhouSL[2] * dSMdUR[i]);
    for (int i = 0; i < 5; i++) AR[4][i] = omL * (dpsdUR[i] * SM + (pS + eSL) * dSMdUR[i]);
  }
}

// calcEulerFluxMatrices_HLLC, UgpWithCvCompFlow.cpp:242-595 (nScal = 0)
// Which HLLC wave configuration the Jacobian was taken in, and the mass flux
// per unit density of the upwind star (or unstarred) state, "FrhostarOverrho"
// of the scalar rows (UgpWithCvCompFlow.cpp:296-311, 389-405, 534-550, 561-577).
struct HllcBranch {
  int k;       // 0: sL > 0; 1: sM >= 0 >= sL; 2: sR >= 0 > sM; 3: sR < 0
  double fso;  // FrhostarOverrho
};

void jac_hllc(double (*A_L)[5], double (*A_R)[5], double rhoL, const double *uL, double pL, double gammaL, double kL,
              double rhoR, const double *uR, double pR, double gammaR, double kR, double area, const double *nVec,
              HllcBranch *br = nullptr) {
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
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = min(uRoe - cRoe, unL - cL);
  double sR = max(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double gAvg = 0.5 * (gammaL + gammaR);
  if (sM >= 0.0) {
    if (sL > 0.0) {
      if (br) *br = {0, unL * area};
      if (A_L != NULL) {
        double nVecArea[3];
        for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
        jac_a(A_L, uL, pL, rhoL, nVecArea, gAvg);
      }
      if (A_R != NULL)
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_R[i][j] = 0.0;
    } else {
      double invSLmSs = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * invSLmSs;
      double rhouSL[3];
      for (int i = 0; i < 3; i++) rhouSL[i] = (rhoL * uL[i] * sLmuL + (pStar - pL) * nVec[i]) * invSLmSs;
      double eSL = (sLmuL * eL - pL * unL + pStar * sM) * invSLmSs;
      if (br) *br = {1, invSLmSs * sLmuL * sM * area};
      double gammaLM1 = (gammaL - 1.0), gammaRM1 = (gammaR - 1.0);
      double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
      double dSMdUL[5], dSMdUR[5], dpsdUL[5], dpsdUR[5];
      if (A_L != NULL) {
        dSMdUL[0] = -unL * unL + uLuL * gammaLM1 / 2.0 + sM * sL;
        dSMdUL[1] = nVec[0] * (2.0 * unL - sL - sM) - gammaLM1 * uL[0];
        dSMdUL[2] = nVec[1] * (2.0 * unL - sL - sM) - gammaLM1 * uL[1];
        dSMdUL[3] = nVec[2] * (2.0 * unL - sL - sM) - gammaLM1 * uL[2];
        dSMdUL[4] = gammaLM1;
        for (int i = 0; i < 5; i++) {
          dSMdUL[i] *= invrhotld;
          dpsdUL[i] = rhoR * (sR - unR) * dSMdUL[i];
        }
      }
      if (A_R != NULL) {
        dSMdUR[0] = unR * unR - uRuR * gammaRM1 / 2.0 - sM * sR;
        dSMdUR[1] = -nVec[0] * (2.0 * unR - sR - sM) + gammaRM1 * uR[0];
        dSMdUR[2] = -nVec[1] * (2.0 * unR - sR - sM) + gammaRM1 * uR[1];
        dSMdUR[3] = -nVec[2] * (2.0 * unR - sR - sM) + gammaRM1 * uR[2];
        dSMdUR[4] = -gammaRM1;
        for (int i = 0; i < 5; i++) {
          dSMdUR[i] *= invrhotld;
          dpsdUR[i] = rhoL * (sL - unL) * dSMdUR[i];
        }
      }
      jac_hllc_subsonic(A_L, A_R, rhoL, uL, pL, eL, unL, uLuL, sL, rhoSL, rhouSL, eSL, dSMdUL, dSMdUR, dpsdUL, dpsdUR,
                        sM, pStar, gAvg, nVec);
      if (A_L != NULL) {
        for (int i = 0; i < 5; i++) A_L[0][i] = A_L[0][i] * sM + dSMdUL[i] * rhoSL;
        for (int i = 0; i < 5; i++) A_L[1][i] = A_L[1][i] * sM + dSMdUL[i] * rhouSL[0] + dpsdUL[i] * nVec[0];
        for (int i = 0; i < 5; i++) A_L[2][i] = A_L[2][i] * sM + dSMdUL[i] * rhouSL[1] + dpsdUL[i] * nVec[1];
        for (int i = 0; i < 5; i++) A_L[3][i] = A_L[3][i] * sM + dSMdUL[i] * rhouSL[2] + dpsdUL[i] * nVec[2];
        for (int i = 0; i < 5; i++) A_L[4][i] = (A_L[4][i] + dpsdUL[i]) * sM + (eSL + pStar) * dSMdUL[i];
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_L[i][j] *= area;
      }
      if (A_R != NULL) {
        for (int i = 0; i < 5; i++) A_R[0][i] = A_R[0][i] * sM + dSMdUR[i] * rhoSL;
        for (int i = 0; i < 5; i++) A_R[1][i] = A_R[1][i] * sM + dSMdUR[i] * rhouSL[0] + dpsdUR[i] * nVec[0];
        for (int i = 0; i < 5; i++) A_R[2][i] = A_R[2][i] * sM + dSMdUR[i] * rhouSL[1] + dpsdUR[i] * nVec[1];
        for (int i = 0; i < 5; i++) A_R[3][i] = A_R[3][i] * sM + dSMdUR[i] * rhouSL[2] + dpsdUR[i] * nVec[2];
        for (int i = 0; i < 5; i++) A_R[4][i] = (A_R[4][i] + dpsdUR[i]) * sM + (eSL + pStar) * dSMdUR[i];
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_R[i][j] *= area;
      }
    }
  } else {
    if (sR >= 0.0) {
      double invSRmSs = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * invSRmSs;
      double rhouSR[3];
      for (int i = 0; i < 3; i++) rhouSR[i] = (rhoR * uR[i] * sRmuR + (pStar - pR) * nVec[i]) * invSRmSs;
      double eSR = (sRmuR * eR - pR * unR + pStar * sM) * invSRmSs;
      if (br) *br = {2, invSRmSs * sRmuR * sM * area};
      double gammaLM1 = (gammaL - 1.0), gammaRM1 = (gammaR - 1.0);
      double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
      double dSMdUL[5], dSMdUR[5], dpsdUL[5], dpsdUR[5];
      if (A_L != NULL) {
        dSMdUL[0] = -unL * unL + uLuL * gammaLM1 / 2.0 + sM * sL;
        dSMdUL[1] = nVec[0] * (2.0 * unL - sL - sM) - gammaLM1 * uL[0];
        dSMdUL[2] = nVec[1] * (2.0 * unL - sL - sM) - gammaLM1 * uL[1];
        dSMdUL[3] = nVec[2] * (2.0 * unL - sL - sM) - gammaLM1 * uL[2];
        dSMdUL[4] = gammaLM1;
        for (int i = 0; i < 5; i++) {
          dSMdUL[i] *= invrhotld;
          dpsdUL[i] = rhoR * (sR - unR) * dSMdUL[i];
        }
      }
      if (A_R != NULL) {
        dSMdUR[0] = unR * unR - uRuR * gammaRM1 / 2.0 - sM * sR;
        dSMdUR[1] = -nVec[0] * (2.0 * unR - sR - sM) + gammaRM1 * uR[0];
        dSMdUR[2] = -nVec[1] * (2.0 * unR - sR - sM) + gammaRM1 * uR[1];
        dSMdUR[3] = -nVec[2] * (2.0 * unR - sR - sM) + gammaRM1 * uR[2];
        dSMdUR[4] = -gammaRM1;
        for (int i = 0; i < 5; i++) {
          dSMdUR[i] *= invrhotld;
          dpsdUR[i] = rhoL * (sL - unL) * dSMdUR[i];
        }
      }
      jac_hllc_subsonic(A_R, A_L, rhoR, uR, pR, eR, unR, uRuR, sR, rhoSR, rhouSR, eSR, dSMdUR, dSMdUL, dpsdUR, dpsdUL,
                        sM, pStar, gAvg, nVec);
      if (A_L != NULL) {
        for (int i = 0; i < 5; i++) A_L[0][i] = A_L[0][i] * sM + dSMdUL[i] * rhoSR;
        for (int i = 0; i < 5; i++) A_L[1][i] = A_L[1][i] * sM + dSMdUL[i] * rhouSR[0] + dpsdUL[i] * nVec[0];
        for (int i = 0; i < 5; i++) A_L[2][i] = A_L[2][i] * sM + dSMdUL[i] * rhouSR[1] + dpsdUL[i] * nVec[1];
        for (int i = 0; i < 5; i++) A_L[3][i] = A_L[3][i] * sM + dSMdUL[i] * rhouSR[2] + dpsdUL[i] * nVec[2];
        for (int i = 0; i < 5; i++) A_L[4][i] = (A_L[4][i] + dpsdUL[i]) * sM + (eSR + pStar) * dSMdUL[i];
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_L[i][j] *= area;
      }
      if (A_R != NULL) {
        for (int i = 0; i < 5; i++) A_R[0][i] = A_R[0][i] * sM + dSMdUR[i] * rhoSR;
        for (int i = 0; i < 5; i++) A_R[1][i] = A_R[1][i] * sM + dSMdUR[i] * rhouSR[0] + dpsdUR[i] * nVec[0];
        for (int i = 0; i < 5; i++) A_R[2][i] = A_R[2][i] * sM + dSMdUR[i] * rhouSR[1] + dpsdUR[i] * nVec[1];
        for (int i = 0; i < 5; i++) A_R[3][i] = A_R[3][i] * sM + dSMdUR[i] * rhouSR[2] + dpsdUR[i] * nVec[2];
        for (int i = 0; i < 5; i++) A_R[4][i] = (A_R[4][i] + dpsdUR[i]) * sM + (eSR + pStar) * dSMdUR[i];
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_R[i][j] *= area;
      }
    } else {
      if (br) *br = {3, unR * area};
      if (A_R != NULL) {
        double nVecArea[3];
        for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
        jac_a(A_R, uR, pR, rhoR, nVecArea, gAvg);
      }
      if (A_L != NULL)
        for (int i = 0; i < 5; i++)
          for (int j = 0; j < 5; j++) A_L[i][j] = 0.0;
    }
  }
}

// Scalar rows of the HLLC Jacobian (A_L_Scal / A_R_Scal, the same lines as
// HllcBranch): d(rho phi un)/dQ from row 0 of the 5x5 block times the upwind
// scalar, and d/d(rho phi) on the upwind side. A side whose block was not
// computed (NULL) is left untouched.
void jac_hllc_scal(double *SL, double *SR, const double (*A_L)[5], const double (*A_R)[5], const HllcBranch &br,
                   double phiL, double phiR) {
  const bool upL = br.k <= 1;
  const double phi = upL ? phiL : phiR;
  if (SL && A_L) {
    if (br.k == 3) {
      for (int i = 0; i <= 5; i++) SL[i] = 0.0;
    } else if (upL) {
      SL[0] = (A_L[0][0] - br.fso) * phi;
      for (int i = 1; i < 5; i++) SL[i] = A_L[0][i] * phi;
      SL[5] = br.fso;
    } else {
      for (int i = 0; i < 5; i++) SL[i] = A_L[0][i] * phi;
      SL[5] = 0.0;
    }
  }
  if (SR && A_R) {
    if (br.k == 0) {
      for (int i = 0; i <= 5; i++) SR[i] = 0.0;
    } else if (!upL) {
      SR[0] = (A_R[0][0] - br.fso) * phi;
      for (int i = 1; i < 5; i++) SR[i] = A_R[0][i] * phi;
      SR[5] = br.fso;
    } else {
      for (int i = 0; i < 5; i++) SR[i] = A_R[0][i] * phi;
      SR[5] = 0.0;
    }
  }
}

// calcPMatrix / calcInvPMatrix, UgpWithCvCompFlow.cpp:3480-3559
void p_matrix(double P[5][5], const double *vel, double c, double rrho, const double *nV, double gamma) {
  double kapm1 = (gamma - 1.0);
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

void inv_p_matrix(double invP[5][5], const double *vel, double c, double rrho, const double *nV, double gamma) {
  double kapm1 = (gamma - 1.0);
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

// calcEulerFluxMatrices_Roe, UgpWithCvCompFlow.cpp:3382-3476 (nScal = 0)
void jac_roe(double (*A_L)[5], double (*A_R)[5], double rhoL, const double *uL, double pL, double gammaL,
             double rhoR, const double *uR, double pR, double gammaR, double area, const double *nVec) {
  double uLuL = dot3(uL, uL);
  double uRuR = dot3(uR, uR);
  double Rrho = sqrt(rhoR / rhoL);
  double rhoRoe = Rrho * rhoL;
  double tmp = 1.0 / (1.0 + Rrho);
  double velRoe[3];
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double nVecArea[3];
  for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
  double Lambda[5];
  for (int i = 0; i < 3; i++) Lambda[i] = uRoe;
  Lambda[3] = uRoe + cRoe;
  Lambda[4] = uRoe - cRoe;
  double P[5][5], InvP[5][5];
  p_matrix(P, velRoe, cRoe, rhoRoe, nVec, 0.5 * (gammaL + gammaR));
  inv_p_matrix(InvP, velRoe, cRoe, rhoRoe, nVec, 0.5 * (gammaL + gammaR));
  if (A_L != NULL) jac_a(A_L, uL, pL, rhoL, nVecArea, 0.5 * (gammaL + gammaR));
  if (A_R != NULL) jac_a(A_R, uR, pR, rhoR, nVecArea, 0.5 * (gammaL + gammaR));
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) {
      double Proj_ModJac = 0.0;
      for (int k = 0; k < 5; k++) Proj_ModJac += P[i][k] * fabs(Lambda[k]) * InvP[k][j];
      if (A_L != NULL) A_L[i][j] = 0.5 * A_L[i][j] + 0.5 * Proj_ModJac * area;
      if (A_R != NULL) A_R[i][j] = 0.5 * A_R[i][j] - 0.5 * Proj_ModJac * area;
    }
}

// ludeco / lusolv, UgpWithCv.cpp:2489-2540
// calcEulerFluxCoupled_HLLC, UgpWithCvCompFlow.cpp:1808-1927, nScal = 0,
// kine = 0: the energy is built from the given enthalpy h (calcThermoProp_p's
// or the cell's), not from gamma/(gamma-1) p/rho as the segregated flux does.
// only_pressure: the "ONLY_PRESSURE" treatment of symmetry and wall faces.
// With scal (one transported scalar, ConvTerm = 1; scal[0] left, scal[1]
// right), *Fs gets its flux: the upwind side's value times the unscaled mass
// flux, then times the area (:1851-1896, 1926).
void flux_hllc_coupled(double F[5], double rhoL, const double *uL, double pL, double h0, double gammaL, double rhoR,
                       const double *uR, double pR, double h1, double gammaR, double area, const double *nVec,
                       bool only_pressure, const double *scal = NULL, double *Fs = NULL) {
  if (only_pressure) {
    F[0] = 0.0;
    for (int i = 1; i < 4; i++) F[i] = pL * nVec[i - 1];
    F[4] = 0.0;
    for (int i = 0; i < 5; i++) F[i] *= area;
    if (Fs) *Fs = 0.0 * area;
    return;
  }
  const double kL = 0.0, kR = 0.0;
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
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = min(uRoe - cRoe, unL - cL);
  double sR = max(uRoe + cRoe, unR + cR);
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  if (sM >= 0.0) {
    if (sL > 0.0) {
      F[0] = rhoL * unL;
      for (int i = 0; i < 3; i++) F[1 + i] = rhoL * uL[i] * unL + pL * nVec[i];
      F[4] = (rhoeL + pL) * unL;
    } else {
      double invSLmSs = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * invSLmSs;
      double rhouSL[3];
      for (int i = 0; i < 3; i++) rhouSL[i] = (rhoL * uL[i] * sLmuL + (pStar - pL) * nVec[i]) * invSLmSs;
      double rhoeSL = (sLmuL * rhoeL - pL * unL + pStar * sM) * invSLmSs;
      F[0] = rhoSL * sM;
      for (int i = 0; i < 3; i++) F[1 + i] = rhouSL[i] * sM + pStar * nVec[i];
      F[4] = (rhoeSL + pStar) * sM;
    }
  } else {
    if (sR >= 0.0) {
      double invSRmSs = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * invSRmSs;
      double rhouSR[3];
      for (int i = 0; i < 3; i++) rhouSR[i] = (rhoR * uR[i] * sRmuR + (pStar - pR) * nVec[i]) * invSRmSs;
      double rhoeSR = (sRmuR * rhoeR - pR * unR + pStar * sM) * invSRmSs;
      F[0] = rhoSR * sM;
      for (int i = 0; i < 3; i++) F[1 + i] = rhouSR[i] * sM + pStar * nVec[i];
      F[4] = (rhoeSR + pStar) * sM;
    } else {
      F[0] = rhoR * unR;
      for (int i = 0; i < 3; i++) F[1 + i] = rhoR * uR[i] * unR + pR * nVec[i];
      F[4] = (rhoeR + pR) * unR;
    }
  }
  if (Fs) *Fs = (1.0 * F[0] * (sM >= 0.0 ? scal[0] : scal[1])) * area;
  for (int i = 0; i < 5; i++) F[i] *= area;
}

// The scalar row of calcEulerFluxMatricesCoupled_HLLC (UgpWithCvCompFlow.cpp
// :1934-2174) for one transported scalar with no pressure derivative
// (dpress_dscal = 0: RansTurbSA leaves pressureDerivativeHookScalarRansTurb
// empty, UgpWithCvCompFlow.h:1590): the supersonic rows of
// calcJacobianCoupled_HLLC (:2233-2240), the subsonic transformation
// (:2281-2284) and star-state terms (:2022-2026, 2074-2078), then dU/dQ
// (:2150-2174). With dpress_dscal = 0 the flow rows' scalar column stays zero
// and their column 0 unchanged, so only this row is formed: SL = A_L[5][0..5],
// SR = A_R[5][0..5] (either may be NULL).
void jac_hllc_coupled_srow(double *SL, double *SR, double rhoL, const double *uL, double pL, double h0, double scalL,
                           double gammaL, double rhoR, const double *uR, double pR, double h1, double scalR,
                           double gammaR, double area, const double *nVec, bool only_pressure) {
  auto zero = [](double *r) {
    for (int j = 0; j < 6; j++) r[j] = 0.0;
  };
  auto finish = [](double *r, double scal, double rho) {
    for (int j = 0; j < 6; j++) r[j] *= 1.0;  // ConvTerm
    r[0] -= scal / rho * r[5];
    r[5] /= rho;
  };
  if (only_pressure) {
    if (SL) {
      zero(SL);
      finish(SL, scalL, rhoL);
    }
    if (SR) {
      zero(SR);
      finish(SR, scalR, rhoR);
    }
    return;
  }
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
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = min(uRoe - cRoe, unL - cL);
  double sR = max(uRoe + cRoe, unR + cR);
  double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double nVecArea[3];
  for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
  // calcJacobianCoupled_HLLC's scalar row (nV = nVecArea)
  auto super = [&](double *r, const double *vel, double rrho, double scal) {
    double nVel[3] = {vel[0] * nVecArea[0], vel[1] * nVecArea[1], vel[2] * nVecArea[2]};
    double U_k = nVel[0] + nVel[1] + nVel[2];
    r[0] = 0.0;
    r[1] = scal * nVecArea[0];
    r[2] = scal * nVecArea[1];
    r[3] = scal * nVecArea[2];
    r[4] = 0.0;
    r[5] = 0.0;
    r[5] = rrho * U_k;
  };
  // calcdSMdU_HLLC (unit normal), the scalar entry dpress_dscal = 0
  auto dsm = [&](double d[6], const double *vel, double pp, double rrho, double hh, double gamma, double S,
                 double Factor) {
    double kapm1 = (gamma - 1.0);
    double nVel[3] = {vel[0] * nVec[0], vel[1] * nVec[1], vel[2] * nVec[2]};
    double U_k = nVel[0] + nVel[1] + nVel[2];
    double vSquHlf = 0.5 * dot3(vel, vel);
    d[0] = Factor * invrhotld * (-U_k * U_k + kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh + sM * S);
    for (int i = 0; i < 3; i++) d[1 + i] = Factor * invrhotld * ((2 * U_k - S - sM) * nVec[i] - kapm1 * vel[i]);
    d[4] = Factor * invrhotld * kapm1;
    d[5] = Factor * invrhotld * 0.0;
  };
  auto sub = [&](double *r, const double *d, double rhoS, double scal, double Omega) {
    double OmSM = Omega * sM;
    double OmSMp1 = OmSM + 1.0;
    for (int i = 0; i < 6; i++) r[i] = -OmSM * r[i] + area * OmSMp1 * d[i] * rhoS * scal;
  };
  double d[6];
  if (sM >= 0.0) {
    if (sL > 0.0) {
      if (SL) super(SL, uL, rhoL, scalL);
      if (SR) zero(SR);
    } else {
      double OmegaL = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * OmegaL;
      if (SL) {
        dsm(d, uL, pL, rhoL, hL, gammaL, sL, 1.0);
        super(SL, uL, rhoL, scalL);
        sub(SL, d, rhoSL, scalL, OmegaL);
        SL[0] += area * OmegaL * sM * sL * scalL;
        SL[5] += area * OmegaL * sM * sL * rhoL;
      }
      if (SR) {
        dsm(d, uR, pR, rhoR, hR, gammaR, sR, -1.0);
        zero(SR);
        sub(SR, d, rhoSL, scalL, OmegaL);
      }
    }
  } else {
    if (sR >= 0.0) {
      double OmegaR = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * OmegaR;
      if (SR) {
        dsm(d, uR, pR, rhoR, hR, gammaR, sR, -1.0);
        super(SR, uR, rhoR, scalR);
        sub(SR, d, rhoSR, scalR, OmegaR);
        SR[0] += area * OmegaR * sM * sR * scalR;
        SR[5] += area * OmegaR * sM * sR * rhoR;
      }
      if (SL) {
        dsm(d, uL, pL, rhoL, hL, gammaL, sL, 1.0);
        zero(SL);
        sub(SL, d, rhoSR, scalR, OmegaR);
      }
    } else {
      if (SR) super(SR, uR, rhoR, scalR);
      if (SL) zero(SL);
    }
  }
  if (SL) finish(SL, scalL, rhoL);
  if (SR) finish(SR, scalR, rhoR);
}

// calcJacobianCoupled_HLLC, UgpWithCvCompFlow.cpp:2180-2241 (nV not normalised)
static void jac_coupled_super(double (*A)[5], const double *vel, double pp, double rrho, double hh, const double *nV,
                              double gamma) {
  double kapm1 = (gamma - 1.0);
  double nVel[3] = {vel[0] * nV[0], vel[1] * nV[1], vel[2] * nV[2]};
  double U_k = nVel[0] + nVel[1] + nVel[2];
  double vSquHlf = 0.5 * dot3(vel, vel);
  A[0][0] = 0.0;
  A[0][1] = nV[0];
  A[0][2] = nV[1];
  A[0][3] = nV[2];
  A[0][4] = 0.0;
  A[1][0] = -vel[0] * U_k + nV[0] * (kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh);
  A[1][1] = (2.0 - gamma) * nVel[0] + U_k;
  A[1][2] = vel[0] * nV[1] - kapm1 * vel[1] * nV[0];
  A[1][3] = vel[0] * nV[2] - kapm1 * vel[2] * nV[0];
  A[1][4] = kapm1 * nV[0];
  A[2][0] = -vel[1] * U_k + nV[1] * (kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh);
  A[2][1] = vel[1] * nV[0] - kapm1 * vel[0] * nV[1];
  A[2][2] = (2.0 - gamma) * nVel[1] + U_k;
  A[2][3] = vel[1] * nV[2] - kapm1 * vel[2] * nV[1];
  A[2][4] = kapm1 * nV[1];
  A[3][0] = -vel[2] * U_k + nV[2] * (kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh);
  A[3][1] = vel[2] * nV[0] - kapm1 * vel[0] * nV[2];
  A[3][2] = vel[2] * nV[1] - kapm1 * vel[1] * nV[2];
  A[3][3] = (2.0 - gamma) * nVel[2] + U_k;
  A[3][4] = kapm1 * nV[2];
  A[4][0] = U_k * ((gamma - 2.0) * vSquHlf - gamma * (hh - pp / rrho));
  A[4][1] = (hh + vSquHlf) * nV[0] - kapm1 * vel[0] * U_k;
  A[4][2] = (hh + vSquHlf) * nV[1] - kapm1 * vel[1] * U_k;
  A[4][3] = (hh + vSquHlf) * nV[2] - kapm1 * vel[2] * U_k;
  A[4][4] = gamma * U_k;
}

// calcdSMdU_HLLC, UgpWithCvCompFlow.cpp:2246-2265 (nV the unit normal)
static void dsm_du(double d[5], const double *vel, double pp, double rrho, double hh, const double *nV, double gamma,
                   double SM, double S, double invertrho, double Factor) {
  double kapm1 = (gamma - 1.0);
  double nVel[3] = {vel[0] * nV[0], vel[1] * nV[1], vel[2] * nV[2]};
  double U_k = nVel[0] + nVel[1] + nVel[2];
  double vSquHlf = 0.5 * dot3(vel, vel);
  d[0] = Factor * invertrho * (-U_k * U_k + kapm1 * vSquHlf + gamma * pp / rrho - kapm1 * hh + SM * S);
  for (int i = 0; i < 3; i++) d[1 + i] = Factor * invertrho * ((2 * U_k - S - SM) * nV[i] - kapm1 * vel[i]);
  d[4] = Factor * invertrho * kapm1;
}

// calcSubSonicJacobianCoupled_HLLC, UgpWithCvCompFlow.cpp:2270-2287
static void jac_coupled_sub(double (*A)[5], const double *dSMdU, double rhoS, const double *rhouS, double rhoeS,
                            double pS, double Omega, double sM, double rhoSmq, const double *nV, double area) {
  double OmSM = Omega * sM;
  double OmSMp1 = OmSM + 1.0;
  for (int i = 0; i < 5; i++) {
    A[0][i] = -OmSM * A[0][i] + area * OmSMp1 * dSMdU[i] * rhoS;
    A[1][i] = -OmSM * A[1][i] + area * OmSMp1 * dSMdU[i] * (rhouS[0] + rhoSmq * nV[0]);
    A[2][i] = -OmSM * A[2][i] + area * OmSMp1 * dSMdU[i] * (rhouS[1] + rhoSmq * nV[1]);
    A[3][i] = -OmSM * A[3][i] + area * OmSMp1 * dSMdU[i] * (rhouS[2] + rhoSmq * nV[2]);
    A[4][i] = -OmSM * A[4][i] + area * OmSMp1 * dSMdU[i] * (rhoeS + pS + rhoSmq * sM);
  }
}

// calcEulerFluxMatricesCoupled_HLLC, UgpWithCvCompFlow.cpp:1934-2174, nScal = 0,
// kine = 0. A_L or A_R may be NULL.
void jac_hllc_coupled(double (*A_L)[5], double (*A_R)[5], double rhoL, const double *uL, double pL, double h0,
                      double gammaL, double rhoR, const double *uR, double pR, double h1, double gammaR, double area,
                      const double *nVec, bool only_pressure) {
  if (only_pressure) {
    if (A_L) {
      double uLuL = dot3(uL, uL);
      double hL = h0 + 0.0;
      for (int j = 0; j < 5; j++) A_L[0][j] = 0.0;
      double dpdrho = 0.5 * (gammaL - 1.0) * uLuL - (gammaL - 1.0) * hL + gammaL * pL / rhoL;
      for (int r = 1; r <= 3; r++) {
        A_L[r][0] = dpdrho * nVec[r - 1] * area;
        A_L[r][1] = -(gammaL - 1.0) * uL[0] * nVec[r - 1] * area;
        A_L[r][2] = -(gammaL - 1.0) * uL[1] * nVec[r - 1] * area;
        A_L[r][3] = -(gammaL - 1.0) * uL[2] * nVec[r - 1] * area;
        A_L[r][4] = (gammaL - 1.0) * nVec[r - 1] * area;
      }
      for (int j = 0; j < 5; j++) A_L[4][j] = 0.0;
    }
    if (A_R)
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) A_R[i][j] = 0.0;
    return;
  }
  double dSMdUL[5], dSMdUR[5];
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
  for (int i = 0; i < 3; i++) velRoe[i] = tmp * (uL[i] + uR[i] * Rrho);
  double uRoe = dot3(velRoe, nVec);
  double gamPdivRho = tmp * ((gammaL * pL / rhoL + 0.5 * (gammaL - 1.0) * uLuL) +
                             (gammaR * pR / rhoR + 0.5 * (gammaR - 1.0) * uRuR) * Rrho);
  double cRoe = sqrt(gamPdivRho - ((gammaL + gammaR) * 0.5 - 1.0) * 0.5 * dot3(velRoe, velRoe));
  double sL = min(uRoe - cRoe, unL - cL);
  double sR = max(uRoe + cRoe, unR + cR);
  double invrhotld = 1.0 / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double sM = (pL - pR - rhoL * unL * (sL - unL) + rhoR * unR * (sR - unR)) / (rhoR * (sR - unR) - rhoL * (sL - unL));
  double pStar = rhoR * (unR - sR) * (unR - sM) + pR;
  double nVecArea[3];
  for (int i = 0; i < 3; i++) nVecArea[i] = nVec[i] * area;
  auto zero = [](double (*A)[5]) {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) A[i][j] = 0.0;
  };
  if (sM >= 0.0) {
    if (sL > 0.0) {
      if (A_L) jac_coupled_super(A_L, uL, pL, rhoL, hL, nVecArea, gammaL);
      if (A_R) zero(A_R);
    } else {
      double OmegaL = 1.0 / (sL - sM);
      double sLmuL = sL - unL;
      double rhoSL = rhoL * sLmuL * OmegaL;
      double rhouSL[3];
      for (int i = 0; i < 3; i++) rhouSL[i] = (rhoL * uL[i] * sLmuL + (pStar - pL) * nVec[i]) * OmegaL;
      double rhoeSL = (sLmuL * rhoeL - pL * unL + pStar * sM) * OmegaL;
      if (A_L) {
        dsm_du(dSMdUL, uL, pL, rhoL, hL, nVec, gammaL, sM, sL, invrhotld, 1.0);
        jac_coupled_super(A_L, uL, pL, rhoL, hL, nVecArea, gammaL);
        double rhoSmqR = rhoR * (sR - unR);
        jac_coupled_sub(A_L, dSMdUL, rhoSL, rhouSL, rhoeSL, pStar, OmegaL, sM, rhoSmqR, nVec, area);
        for (int i = 0; i < 5; i++) A_L[i][i] += area * OmegaL * sM * sL;
      }
      if (A_R) {
        dsm_du(dSMdUR, uR, pR, rhoR, hR, nVec, gammaR, sM, sR, invrhotld, -1.0);
        zero(A_R);
        double rhoSmqL = rhoL * (sL - unL);
        jac_coupled_sub(A_R, dSMdUR, rhoSL, rhouSL, rhoeSL, pStar, OmegaL, sM, rhoSmqL, nVec, area);
      }
    }
  } else {
    if (sR >= 0.0) {
      double OmegaR = 1.0 / (sR - sM);
      double sRmuR = sR - unR;
      double rhoSR = rhoR * sRmuR * OmegaR;
      double rhouSR[3];
      for (int i = 0; i < 3; i++) rhouSR[i] = (rhoR * uR[i] * sRmuR + (pStar - pR) * nVec[i]) * OmegaR;
      double rhoeSR = (sRmuR * rhoeR - pR * unR + pStar * sM) * OmegaR;
      if (A_R) {
        dsm_du(dSMdUR, uR, pR, rhoR, hR, nVec, gammaR, sM, sR, invrhotld, -1.0);
        jac_coupled_super(A_R, uR, pR, rhoR, hR, nVecArea, gammaR);
        double rhoSmqL = rhoL * (sL - unL);
        jac_coupled_sub(A_R, dSMdUR, rhoSR, rhouSR, rhoeSR, pStar, OmegaR, sM, rhoSmqL, nVec, area);
        for (int i = 0; i < 5; i++) A_R[i][i] += area * OmegaR * sM * sR;
      }
      if (A_L) {
        dsm_du(dSMdUL, uL, pL, rhoL, hL, nVec, gammaL, sM, sL, invrhotld, 1.0);
        zero(A_L);
        double rhoSmqR = rhoR * (sR - unR);
        jac_coupled_sub(A_L, dSMdUL, rhoSR, rhouSR, rhoeSR, pStar, OmegaR, sM, rhoSmqR, nVec, area);
      }
    } else {
      if (A_R) jac_coupled_super(A_R, uR, pR, rhoR, hR, nVecArea, gammaR);
      if (A_L) zero(A_L);
    }
  }
}

void ludeco(double (*A)[5], int order) {
  for (int jc = 1; jc < order; jc++) A[0][jc] /= A[0][0];
  int jrjc = 0;
  for (;;) {
    jrjc++;
    int jrjcm1 = jrjc - 1;
    int jrjcp1 = jrjc + 1;
    for (int jr = jrjc; jr < order; jr++) {
      double sum = A[jr][jrjc];
      for (int jm = 0; jm <= jrjcm1; jm++) sum -= A[jr][jm] * A[jm][jrjc];
      A[jr][jrjc] = sum;
    }
    if (jrjc == (order - 1)) return;
    for (int jc = jrjcp1; jc < order; jc++) {
      double sum = A[jrjc][jc];
      for (int jm = 0; jm <= jrjcm1; jm++) sum -= A[jrjc][jm] * A[jm][jc];
      A[jrjc][jc] = sum / A[jrjc][jrjc];
    }
  }
}

void lusolv(double (*A)[5], double *c, int order) {
  c[0] = c[0] / A[0][0];
  for (int jr = 1; jr < order; jr++) {
    int jrm1 = jr - 1;
    double sum = c[jr];
    for (int jm = 0; jm <= jrm1; jm++) sum -= A[jr][jm] * c[jm];
    c[jr] = sum / A[jr][jr];
  }
  for (int jrjr = 1; jrjr < order; jrjr++) {
    int jr = (order - 1) - jrjr;
    int jrp1 = jr + 1;
    double sum = c[jr];
    for (int jmjm = jrp1; jmjm < order; jmjm++) {
      int jm = (order - 1) - jmjm + jrp1;
      sum -= A[jr][jm] * c[jm];
    }
    c[jr] = sum;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// context: the member arrays of UgpWithCvCompFlow/JoeWithModels the path uses
// ---------------------------------------------------------------------------
struct or_ctx {
  or_mesh m;
  or_params p;
  int ncv, nfa, nfa_b, nnz;
  std::vector<double> rho, rhou, rhoE, vel, press, temp, enthalpy, sos;
  std::vector<double> rho_bfa, T_bfa, vel_bfa, p_bfa, h_bfa;
  std::vector<double> rho_bpr, vel_bpr, p_bpr;  // PROFILES face values (UgpWithCvCompFlow.h:679)
  std::vector<double> grad_rho, grad_p, grad_u, alpha_rho, local_dt;
  std::vector<double> mul_fa, lamOcp_fa, grad_h;
  std::vector<double> rhs_rho, rhs_rhou, rhs_rhoE, A, dq, resid;
  // Spalart-Allmaras: the transported scalar nuSA (ScalarTranspEq, UgpWithCv.h:50-113),
  // its matrix rows AScal[6][nnz] (JoeWithModels.cpp:892), mut_fa and the model fields
  bool sa = false;
  std::vector<double> phi, phi_bfa, phi_bpr, grad_phi, diff, mut_fa, wall_dist, str_mag, vort_mag;
  std::vector<double> rhs_scal, AScal, dscal;
  // Menter SST (RansTurbKOmSST): kine and omega, registered in that order
  // (TurbModel_KOMSST.h:36-52), in the same arrays scalar by scalar (phi
  // [nsc][ncv], phi_bfa [nsc][nfa_b], grad_phi [nsc][ncv][3], diff [nsc][nfa],
  // rhs_scal and dscal [nsc][ncv], AScal [nsc][6][nnz]); its model fields
  bool sst = false;
  int nsc = 0;
  std::vector<double> diverg, cross_diff, f1, f2;
  // the coupled loop with the scalar: its 6x6 blocks and correction [ncv][6]
  std::vector<double> A6, dq6;
  // each face's inviscid flux [nfa][5] as the face loops last formed it
  // (area included, icv0 -> icv1 orientation; boundary faces out of the domain)
  std::vector<double> flux_fa;
  double cfl;
  int step;
  bool have_press;
  bool resume_refresh = false;  // or_set_step: the first step refreshes as at step 0
};

namespace {

// laminar viscosity of a temperature (UgpWithCvCompFlow.h:1712, 1734, 1795-1798)
double mu_lam(const or_ctx &c, double T) {
  if (c.p.visc_mode == 1)
    return c.p.mu_ref * pow(T / c.p.sl_tref, 1.5) * (c.p.sl_tref + c.p.sl_sref) / (T + c.p.sl_sref);
  return c.p.mu_ref * pow(T / c.p.T_ref, c.p.mu_power_law);
}

// distance weights of a face between two cells (GG_WITH_DIST, UgpWithCv.h:37,
// :1104-1110; the same in calcMaterialProperties)
static void gg_weights(const or_mesh &m, int ifa, int icv0, int icv1, double &w0, double &w1) {
  double dx0[3], dx1[3];
  for (int i = 0; i < 3; i++) {
    dx0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
    dx1[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv1 + i];
  }
  double s0 = 0.0, s1 = 0.0;
  for (int i = 0; i < 3; i++) s0 += dx0[i] * dx0[i];
  for (int i = 0; i < 3; i++) s1 += dx1[i] * dx1[i];
  w0 = sqrt(s0);
  w1 = sqrt(s1);
}

// calcStateVariables, UgpWithCvCompFlow.h:1662-1694 (kine = NULL but with SST;
// gamma, RoM constant: initialHookScalarRansCombModel sets them to GAMMA, R_gas)
void calc_state(or_ctx &c) {
  const double gam = c.p.gamma, R = c.p.R_gas;
  for (int icv = 0; icv < c.ncv; icv++) {
    for (int i = 0; i < 3; i++) c.vel[3 * icv + i] = c.rhou[3 * icv + i] / c.rho[icv];
    double kinecv = 0.0;
    if (c.sst) kinecv = c.phi[icv];
    double pr = (gam - 1.0) * (c.rhoE[icv] - 0.5 * dot3(&c.rhou[3 * icv], &c.rhou[3 * icv]) / c.rho[icv] - c.rho[icv] * kinecv);
    if (pr > 0.0) c.press[icv] = pr;
    c.temp[icv] = c.press[icv] / (c.rho[icv] * R);
    c.enthalpy[icv] = gam * R / (gam - 1.0) * c.temp[icv];
    c.sos[icv] = sqrt(gam * c.press[icv] / c.rho[icv]);
  }
}

// setBC, JoeWithModels.cpp:3857-4300, with ComputeBCProperties_T
// (UgpWithCvCompFlow.h:1783-1811) and rho_bfa from the EOS (:4292-4293).
void set_bc(or_ctx &c) {
  const double GAMMA = c.p.gamma, R_gas = c.p.R_gas;
  const or_mesh &m = c.m;
  for (int iz = 0; iz < m.nzone; iz++) {
    const double *par = &m.zone_par[8 * iz];
    const int f0 = m.zone_ifa[2 * iz], f1 = m.zone_ifa[2 * iz + 1];
    switch (m.zone_bc[iz]) {
      case OR_BC_CBC: {
        double u_bc[3] = {par[0], par[1], par[2]}, T_bc = par[3], p_bc = par[4];
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          double nVec[3];
          norm3(nVec, &m.fa_normal[3 * ifa]);
          if (dot3(u_bc, nVec) > 0.0) {
            double velMagN = dot3(&c.vel[3 * icv0], nVec);
            double mach = fabs(velMagN) / c.sos[icv0];
            c.T_bfa[ifa] = c.temp[icv0];
            for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = c.vel[3 * icv0 + i];
            c.p_bfa[ifa] = mach >= 1.0 ? c.press[icv0] : p_bc;
          } else {
            c.T_bfa[ifa] = T_bc;
            for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = u_bc[i];
            c.p_bfa[ifa] = p_bc;
          }
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_CBC_SUBSONIC_INLET: {
        double angleU[3] = {par[0], par[1], par[2]}, Ttot = par[3], ptot = par[4];
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          double u[3] = {c.rhou[3 * icv0] / c.rho[icv0], c.rhou[3 * icv0 + 1] / c.rho[icv0], c.rhou[3 * icv0 + 2] / c.rho[icv0]};
          double wPow2 = dot3(u, u);
          double vel = sqrt(wPow2);
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = angleU[i] * vel;
          c.T_bfa[ifa] = Ttot;
          c.p_bfa[ifa] = ptot;
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        for (int ifa = f0; ifa <= f1; ifa++) {
          double wPow2 = dot3(&c.vel_bfa[3 * ifa], &c.vel_bfa[3 * ifa]);
          c.h_bfa[ifa] -= 0.5 * wPow2;
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.T_bfa[ifa] = c.h_bfa[ifa] * (GAMMA - 1.0) / (GAMMA * R_gas);
        for (int ifa = f0; ifa <= f1; ifa++) c.p_bfa[ifa] = ptot * pow(c.T_bfa[ifa] / Ttot, GAMMA / (GAMMA - 1.0));
        break;
      }
      case OR_BC_CBC_SUBSONIC_OUTLET: {
        double p_bc = par[0];
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          c.p_bfa[ifa] = p_bc;
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = c.rhou[3 * icv0 + i] / c.rho[icv0];
          c.T_bfa[ifa] = c.temp[icv0];
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_SYMMETRY: {
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          double nVec[3];
          norm3(nVec, &m.fa_normal[3 * ifa]);
          double u0[3] = {c.vel[3 * icv0], c.vel[3 * icv0 + 1], c.vel[3 * icv0 + 2]};
          double un = dot3(nVec, u0);
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = u0[i] - 1.0 * un * nVec[i];
          c.T_bfa[ifa] = c.temp[icv0];
          c.p_bfa[ifa] = c.press[icv0];
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_NEUMANN: {
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = c.rhou[3 * icv0 + i] / c.rho[icv0];
          c.T_bfa[ifa] = c.temp[icv0];
          c.p_bfa[ifa] = c.press[icv0];
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_WALL: {
        double T_bc = par[0];
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = 0.0;
          c.T_bfa[ifa] = T_bc > 0.0 ? T_bc : c.temp[icv0];
          c.p_bfa[ifa] = c.press[icv0];
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_FARFIELD: {
        double u_bc[3] = {par[0], par[1], par[2]};
        double p_bc = par[3], rho_bc = par[4], gamma_bc = par[5];
        double gm1 = gamma_bc - 1.0;
        double ovgm1 = 1.0 / gm1;
        double c_bc = sqrt(gamma_bc * p_bc / rho_bc);
        double s_bc = pow(rho_bc, gamma_bc) / p_bc;
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          double nVec[3];
          norm3(nVec, &m.fa_normal[3 * ifa]);
          double qn_bc = dot3(nVec, u_bc);
          double vn_bc = qn_bc;
          double qne = dot3(nVec, &c.vel[3 * icv0]);
          double ce = sqrt(GAMMA * c.press[icv0] / c.rho[icv0]);
          double ac1 = vn_bc > -c_bc ? qne + 2.0 * ovgm1 * ce : qn_bc + 2.0 * ovgm1 * c_bc;
          double ac2 = vn_bc > c_bc ? qne - 2.0 * ovgm1 * ce : qn_bc - 2.0 * ovgm1 * c_bc;
          double qnf = 0.5 * (ac1 + ac2);
          double cf = 0.25 * (ac1 - ac2) * gm1;
          double velf[3], sf;
          if (vn_bc > 0) {
            for (int i = 0; i < 3; i++) velf[i] = c.vel[3 * icv0 + i] + (qnf - qne) * nVec[i];
            sf = pow(c.rho[icv0], GAMMA) / c.press[icv0];
          } else {
            for (int i = 0; i < 3; i++) velf[i] = u_bc[i] + (qnf - qn_bc) * nVec[i];
            sf = s_bc;
          }
          c.rho_bfa[ifa] = pow((sf * cf * cf / GAMMA), ovgm1);
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = velf[i];
          c.p_bfa[ifa] = c.rho_bfa[ifa] * cf * cf / GAMMA;
          c.T_bfa[ifa] = c.p_bfa[ifa] / (R_gas * c.rho_bfa[ifa]);
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
      case OR_BC_PROFILES: {  // JoeWithModels.cpp:4190-4270: FARFIELD's characteristic state, per-face profile
        const double gamma_bc = 1.4;
        double gm1 = gamma_bc - 1.0;
        double ovgm1 = 1.0 / gm1;
        for (int ifa = f0; ifa <= f1; ifa++) {
          double rho_bc = c.rho_bpr[ifa];
          double u_bc[3] = {c.vel_bpr[3 * ifa], c.vel_bpr[3 * ifa + 1], c.vel_bpr[3 * ifa + 2]};
          double p_bc = c.p_bpr[ifa];
          double c_bc = sqrt(gamma_bc * p_bc / rho_bc);
          double s_bc = pow(rho_bc, gamma_bc) / p_bc;
          int icv0 = m.cvofa[2 * ifa];
          double nVec[3];
          norm3(nVec, &m.fa_normal[3 * ifa]);
          double qn_bc = dot3(nVec, u_bc);
          double vn_bc = qn_bc;
          double qne = dot3(nVec, &c.vel[3 * icv0]);
          double ce = sqrt(GAMMA * c.press[icv0] / c.rho[icv0]);
          double ac1 = vn_bc > -c_bc ? qne + 2.0 * ovgm1 * ce : qn_bc + 2.0 * ovgm1 * c_bc;
          double ac2 = vn_bc > c_bc ? qne - 2.0 * ovgm1 * ce : qn_bc - 2.0 * ovgm1 * c_bc;
          double qnf = 0.5 * (ac1 + ac2);
          double cf = 0.25 * (ac1 - ac2) * gm1;
          double velf[3], sf;
          if (vn_bc > 0) {
            for (int i = 0; i < 3; i++) velf[i] = c.vel[3 * icv0 + i] + (qnf - qne) * nVec[i];
            sf = pow(c.rho[icv0], GAMMA) / c.press[icv0];
          } else {
            for (int i = 0; i < 3; i++) velf[i] = u_bc[i] + (qnf - qn_bc) * nVec[i];
            sf = s_bc;
          }
          c.rho_bfa[ifa] = pow((sf * cf * cf / GAMMA), ovgm1);
          for (int i = 0; i < 3; i++) c.vel_bfa[3 * ifa + i] = velf[i];
          c.p_bfa[ifa] = c.rho_bfa[ifa] * cf * cf / GAMMA;
          c.T_bfa[ifa] = c.p_bfa[ifa] / (R_gas * c.rho_bfa[ifa]);
        }
        for (int ifa = f0; ifa <= f1; ifa++) c.h_bfa[ifa] = GAMMA * R_gas / (GAMMA - 1.0) * c.T_bfa[ifa];
        break;
      }
    }
  }
  // every boundary keyword ends with ComputeBCProperties_T (the subsonic
  // inlet with _H, from the static temperature): laminar viscosity and
  // conductivity of the boundary faces (UgpWithCvCompFlow.h:1783-1839)
  if (c.p.mu_ref > 0.0)
    for (int ifa = 0; ifa < c.nfa_b; ifa++) {
      c.mul_fa[ifa] = mu_lam(c, c.T_bfa[ifa]);
      c.lamOcp_fa[ifa] = c.mul_fa[ifa] / c.p.pr;
    }
  for (int ifa = 0; ifa < c.nfa_b; ifa++) c.rho_bfa[ifa] = c.p_bfa[ifa] / (R_gas * c.T_bfa[ifa]);
}

// calcMaterialProperties, internal faces (UgpWithCvCompFlow.h:1700-1750):
// distance-weighted face temperature, then Sutherland or power law
void calc_material(or_ctx &c) {
  if (!(c.p.mu_ref > 0.0)) return;
  const or_mesh &m = c.m;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double temperature = (w1 * c.temp[icv0] + w0 * c.temp[icv1]) / (w0 + w1);
    c.mul_fa[ifa] = mu_lam(c, temperature);
    c.lamOcp_fa[ifa] = c.mul_fa[ifa] / c.p.pr;
  }
}

// calcDt, UgpWithCvCompFlow.cpp:5-110: MAX_BASED (:30-60) or SUM_BASED
// (:62-99). mut_fa is zero without a turbulence model
// (UgpWithCvCompFlow.h:1568-1580).
double calc_dt(or_ctx &c) {
  const or_mesh &m = c.m;
  double dt_min = 1.0e20;
  if (c.p.time_step_sum) {
    for (int icv = 0; icv < c.ncv; icv++) {
      double cs = sqrt(c.p.gamma * c.press[icv] / c.rho[icv]);
      double lambda_conv = 0.0, lambda_diff = 0.0;
      for (int foc = m.faocv_i[icv]; foc <= m.faocv_i[icv + 1] - 1; foc++) {
        int ifa = m.faocv_v[foc];
        double nVec[3];
        double area = norm3(nVec, &m.fa_normal[3 * ifa]);
        double Uk = dot3(&c.rhou[3 * icv], nVec) / c.rho[icv];
        int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
        double volume = (icv0 >= 0 && icv1 >= 0) ? 0.5 * (m.cv_volume[icv0] + m.cv_volume[icv1])
                                                 : 0.5 * m.cv_volume[icv];
        lambda_conv += (fabs(Uk) + cs) * area;
        if (c.p.mu_ref > 0.0) {
          double KinVisc = (4.0 / 3.0) * (c.mul_fa[ifa] + c.mut_fa[ifa]) / c.rho[icv];
          lambda_diff += area * area * KinVisc / volume;
        } else {
          lambda_diff = 0.0;
        }
      }
      double dt_conv = c.cfl * m.cv_volume[icv] / lambda_conv;
      double dt_visc = c.p.mu_ref > 0.0 ? 0.25 * c.cfl * m.cv_volume[icv] / lambda_diff : 1.0e20;
      double dt_cv = min(dt_conv, dt_visc);
      dt_min = min(dt_min, dt_cv);
      c.local_dt[icv] = dt_cv;
    }
    if (c.p.time_step_min)
      for (int icv = 0; icv < c.ncv; icv++) c.local_dt[icv] = dt_min;
    return dt_min;
  }
  for (int icv = 0; icv < c.ncv; icv++) {
    double lambdaMax = 0.0;
    double cs = sqrt(c.p.gamma * c.press[icv] / c.rho[icv]);
    for (int foc = m.faocv_i[icv]; foc <= m.faocv_i[icv + 1] - 1; foc++) {
      int ifa = m.faocv_v[foc];
      double nVec[3];
      double area = norm3(nVec, &m.fa_normal[3 * ifa]);
      double Uk = dot3(&c.rhou[3 * icv], nVec) / c.rho[icv];
      double lambda = (fabs(Uk) + cs) * area;
      double lamdiff = 4.0 * (c.mul_fa[ifa] + c.mut_fa[ifa]) / c.rho[icv] * area * area / m.cv_volume[icv];
      lambda = max(lambda, lamdiff);
      lambdaMax = max(lambdaMax, lambda);
    }
    double dt_cv = c.cfl * m.cv_volume[icv] / lambdaMax;
    dt_min = min(dt_min, dt_cv);
    c.local_dt[icv] = dt_cv;
  }
  if (c.p.time_step_min)
    for (int icv = 0; icv < c.ncv; icv++) c.local_dt[icv] = dt_min;
  return dt_min;
}

// LS gradients with boundary values, UgpWithCv.h:1009-1083
void grad_scalar_ls(or_ctx &c, double *g, const double *phi, const double *phi_fa) {
  const or_mesh &m = c.m;
  for (int icv = 0; icv < c.ncv; icv++) {
    int noc_f = m.nbocv_i[icv], noc_l = m.nbocv_i[icv + 1] - 1;
    for (int i = 0; i < 3; i++) g[3 * icv + i] = m.nbocv_lsg[3 * noc_f + i] * phi[icv];
    for (int noc = noc_f + 1; noc <= noc_l; noc++) {
      int nb = m.nbocv_v[noc];
      for (int i = 0; i < 3; i++) g[3 * icv + i] += m.nbocv_lsg[3 * noc + i] * phi[nb];
    }
  }
  for (int ifa = 0; ifa < c.nfa_b; ifa++) {
    int icv = m.cvofa[2 * ifa];
    for (int i = 0; i < 3; i++) g[3 * icv + i] += m.bfa_lsg[3 * ifa + i] * phi_fa[ifa];
  }
}

void grad_vector_ls(or_ctx &c, double *g, const double *u, const double *u_fa) {
  const or_mesh &m = c.m;
  for (int icv = 0; icv < c.ncv; icv++) {
    int noc_f = m.nbocv_i[icv], noc_l = m.nbocv_i[icv + 1] - 1;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) g[9 * icv + 3 * i + j] = m.nbocv_lsg[3 * noc_f + j] * u[3 * icv + i];
    for (int noc = noc_f + 1; noc <= noc_l; noc++) {
      int nb = m.nbocv_v[noc];
      for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) g[9 * icv + 3 * i + j] += m.nbocv_lsg[3 * noc + j] * u[3 * nb + i];
    }
  }
  for (int ifa = 0; ifa < c.nfa_b; ifa++) {
    int icv = m.cvofa[2 * ifa];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) g[9 * icv + 3 * i + j] += m.bfa_lsg[3 * ifa + j] * u_fa[3 * ifa + i];
  }
}

// Green-Gauss gradients, UgpWithCv.h:1091-1188. The reference defines
// GG_WITH_DIST (UgpWithCv.h:37): face values are weighted by the distances
// from the face centroid to the two cell centroids (:1104-1110).

void grad_scalar_gg(or_ctx &c, double *g, const double *p, const double *p_fa) {
  const or_mesh &m = c.m;
  for (int k = 0; k < 3 * c.ncv; k++) g[k] = 0.0;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double faceValue = (w1 * p[icv0] + w0 * p[icv1]) / (w0 + w1);
    for (int i = 0; i < 3; i++) {
      g[3 * icv0 + i] += faceValue * m.fa_normal[3 * ifa + i];
      g[3 * icv1 + i] -= faceValue * m.fa_normal[3 * ifa + i];
    }
  }
  for (int ifa = 0; ifa < c.nfa_b; ifa++) {
    int icv = m.cvofa[2 * ifa];
    for (int i = 0; i < 3; i++) g[3 * icv + i] += p_fa[ifa] * m.fa_normal[3 * ifa + i];
  }
  for (int icv = 0; icv < c.ncv; icv++)
    for (int i = 0; i < 3; i++) g[3 * icv + i] /= m.cv_volume[icv];
}

void grad_vector_gg(or_ctx &c, double *g, const double *u, const double *u_fa) {
  const or_mesh &m = c.m;
  for (int k = 0; k < 9 * c.ncv; k++) g[k] = 0.0;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    for (int i = 0; i < 3; i++) {
      double faceValue = (w1 * u[3 * icv0 + i] + w0 * u[3 * icv1 + i]) / (w0 + w1);
      for (int j = 0; j < 3; j++) {
        g[9 * icv0 + 3 * i + j] += faceValue * m.fa_normal[3 * ifa + j];
        g[9 * icv1 + 3 * i + j] -= faceValue * m.fa_normal[3 * ifa + j];
      }
    }
  }
  for (int ifa = 0; ifa < c.nfa_b; ifa++) {
    int icv = m.cvofa[2 * ifa];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) g[9 * icv + 3 * i + j] += u_fa[3 * ifa + i] * m.fa_normal[3 * ifa + j];
  }
  for (int icv = 0; icv < c.ncv; icv++)
    for (int k = 0; k < 9; k++) g[9 * icv + k] /= m.cv_volume[icv];
}

// smooth Barth-Jespersen limiters, UgpWithCv.h:881-1000 (BJ undefined, :877)
void limit_scalar(or_ctx &c, double *g, const double *phi, const double *ref, double *alpha) {
  const or_mesh &m = c.m;
  const double eps = c.p.eps_limiter;
  for (int icv = 0; icv < c.ncv; icv++) {
    double phiMax = phi[icv], phiMin = phiMax;
    for (int noc = m.nbocv_i[icv] + 1; noc <= m.nbocv_i[icv + 1] - 1; noc++) {
      int nb = m.nbocv_v[noc];
      phiMax = max(phiMax, phi[nb]);
      phiMin = min(phiMin, phi[nb]);
    }
    double alfa = 1.0;
    if ((phiMax - phiMin) > 1.0e-12) {
      for (int foc = m.faocv_i[icv]; foc <= m.faocv_i[icv + 1] - 1; foc++) {
        int ifa = m.faocv_v[foc];
        if (ifa >= c.nfa_b) {
          double r0[3];
          for (int i = 0; i < 3; i++) r0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv + i];
          double phifa = phi[icv] + dot3(r0, &g[3 * icv]);
          double eps2 = max(eps * fabs(ref[icv]), 1.0e-8);  // SDWLSClippingScalar
          eps2 = pow(eps2, 2.0);
          double dm = phifa - phi[icv];
          double dp = (phifa - phi[icv]) >= 0.0 ? phiMax - phi[icv] : phiMin - phi[icv];
          alfa = min(alfa, (dp * dp + 2.0 * dm * dp + eps2) / (dp * dp + 2 * dm * dm + dm * dp + eps2));
        }
      }
    } else
      alfa = 0.0;
    for (int i = 0; i < 3; i++) g[3 * icv + i] *= alfa;
    if (alpha) alpha[icv] = alfa;
  }
}

void limit_vector(or_ctx &c, double *g, const double *phi, const double *ref) {
  const or_mesh &m = c.m;
  const double eps = c.p.eps_limiter;
  for (int icv = 0; icv < c.ncv; icv++) {
    for (int i = 0; i < 3; i++) {
      double phiMax = phi[3 * icv + i], phiMin = phiMax;
      for (int noc = m.nbocv_i[icv] + 1; noc <= m.nbocv_i[icv + 1] - 1; noc++) {
        int nb = m.nbocv_v[noc];
        phiMax = max(phiMax, phi[3 * nb + i]);
        phiMin = min(phiMin, phi[3 * nb + i]);
      }
      double alfa = 1.0;
      if ((phiMax - phiMin) > 1.0e-8) {
        for (int foc = m.faocv_i[icv]; foc <= m.faocv_i[icv + 1] - 1; foc++) {
          int ifa = m.faocv_v[foc];
          if (ifa >= c.nfa_b) {
            double r0[3];
            for (int k = 0; k < 3; k++) r0[k] = m.x_fa[3 * ifa + k] - m.x_cv[3 * icv + k];
            double phifa = phi[3 * icv + i] + dot3(r0, &g[9 * icv + 3 * i]);
            double eps2 = pow(eps * ref[icv], 2.0);
            double dm = phifa - phi[3 * icv + i];
            double dp = (phifa - phi[3 * icv + i]) >= 0.0 ? phiMax - phi[3 * icv + i] : phiMin - phi[3 * icv + i];
            alfa = min(alfa, (dp * dp + 2.0 * dm * dp + eps2) / (dp * dp + 2 * dm * dm + dm * dp + eps2));
          }
        }
      } else
        alfa = 0.0;
      for (int j = 0; j < 3; j++) g[9 * icv + 3 * i + j] *= alfa;
    }
  }
}

// Scalars are transported with HLLC only (the build rejects other schemes
// together with a turbulence model), so Fs / SL / SR are HLLC's.
// k0, k1: the turbulent kinetic energy of each side (kineFA / kineCV of the
// face loops, 0 without SST); phi0/phi1/Fs: the c.nsc transported scalars.
void face_flux(const or_ctx &c, double F[5], double rho0, const double *u0, double p0, double k0, double rho1,
               const double *u1, double p1, double k1, double area, const double *nVec,
               const double *phi0 = nullptr, const double *phi1 = nullptr, double *Fs = nullptr) {
  const double g = c.p.gamma;
  if (c.p.scheme == OR_ROE)
    flux_roe(F, rho0, u0, p0, g, rho1, u1, p1, g, area, nVec);
  else if (c.p.scheme == OR_AUSM)
    flux_ausm(F, rho0, u0, p0, g, k0, rho1, u1, p1, g, k1, area, nVec);
  else
    flux_hllc(F, rho0, u0, p0, g, k0, rho1, u1, p1, g, k1, area, nVec, phi0, phi1, Fs, Fs ? c.nsc : 0);
}

// SL/SR: [nsc][6], the scalar rows of each scalar (AplScal / AmiScal)
void face_jac(const or_ctx &c, double (*Apl)[5], double (*Ami)[5], double rho0, const double *u0, double p0, double k0,
              double rho1, const double *u1, double p1, double k1, double area, const double *nVec,
              const double *phi0 = nullptr, const double *phi1 = nullptr, double *SL = nullptr, double *SR = nullptr) {
  const double g = c.p.gamma;
  if (c.p.scheme == OR_ROE) {
    jac_roe(Apl, Ami, rho0, u0, p0, g, rho1, u1, p1, g, area, nVec);
  } else {  // HLLC, and AUSM which uses the HLLC Jacobian (JoeWithModels.cpp:3057-3061)
    HllcBranch br;
    jac_hllc(Apl, Ami, rho0, u0, p0, g, k0, rho1, u1, p1, g, k1, area, nVec, &br);
    if (SL || SR)
      for (int s = 0; s < c.nsc; s++)
        jac_hllc_scal(SL ? SL + 6 * s : nullptr, SR ? SR + 6 * s : nullptr, Apl, Ami, br, phi0[s], phi1[s]);
  }
}

// AScal[s][k][noc] (k = 0..5) += sgn * S[k]
inline void add_ascal(or_ctx &c, int noc, double sgn, const double *S, int s = 0) {
  const size_t nnz = c.nnz;
  double *AS = &c.AScal[6 * nnz * (size_t)s];
  for (int k = 0; k <= 5; k++) {
    if (sgn > 0.0) AS[k * nnz + noc] += S[k];
    else AS[k * nnz + noc] -= S[k];
  }
}

// calcEulerFlux, JoeWithModels.cpp:2621-3589 (nScal = 0, no shock fix)
// the gradients calcEulerFlux and calcFluxCoupled take before their face
// loops (JoeWithModels.cpp:2682-2741, 4691-4745)
static void euler_gradients(or_ctx &c) {
  {
    // calcCvVectorGrad / calcCvScalarGrad (JoeWithModels.cpp:2682-2694)
    if (c.p.grad_scheme == OR_GRAD_LS) {
      grad_vector_ls(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
    } else {
      grad_vector_gg(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
    }
    if (c.p.limiter) limit_vector(c, c.grad_u.data(), c.vel.data(), c.sos.data());
    for (int icv = 0; icv < c.ncv; icv++) c.alpha_rho[icv] = 1.0;
    if (c.p.grad_scheme == OR_GRAD_LS)
      grad_scalar_ls(c, c.grad_rho.data(), c.rho.data(), c.rho_bfa.data());
    else
      grad_scalar_gg(c, c.grad_rho.data(), c.rho.data(), c.rho_bfa.data());
    if (c.p.limiter) limit_scalar(c, c.grad_rho.data(), c.rho.data(), c.rho.data(), c.alpha_rho.data());
    if (c.p.grad_scheme == OR_GRAD_LS)
      grad_scalar_ls(c, c.grad_p.data(), c.press.data(), c.p_bfa.data());
    else
      grad_scalar_gg(c, c.grad_p.data(), c.press.data(), c.p_bfa.data());
    if (c.p.limiter) limit_scalar(c, c.grad_p.data(), c.press.data(), c.press.data(), NULL);
    // STANDARD reconstruction of the transported scalars (:2736-2741)
    for (int s = 0; s < c.nsc; s++) {
      double *g = &c.grad_phi[3 * (size_t)c.ncv * s], *ph = &c.phi[(size_t)c.ncv * s];
      const double *pb = &c.phi_bfa[(size_t)c.nfa_b * s];
      if (c.p.grad_scheme == OR_GRAD_LS)
        grad_scalar_ls(c, g, ph, pb);
      else
        grad_scalar_gg(c, g, ph, pb);
      if (c.p.limiter) limit_scalar(c, g, ph, ph, NULL);
    }
  }
}

// the face's inviscid flux as the face loop adds it (FLUX_FA)
inline void keep_flux(or_ctx &c, int ifa, const double *F) {
  for (int k = 0; k < 5; k++) c.flux_fa[5 * (size_t)ifa + k] = F[k];
}

void calc_euler_flux(or_ctx &c, double *A) {
  const or_mesh &m = c.m;
  const bool snd = c.p.second_order != 0;
  if (snd) euler_gradients(c);
  double F[5], Apl[5][5], Ami[5][5], Fs[2] = {0.0, 0.0}, SL[12], SR[12];
  const int nsc = c.nsc;
  const size_t n = c.ncv, nb = c.nfa_b;
  // scalar s of cell icv, boundary face ifa, its gradient; its residual
  auto ph = [&](int s, int icv) { return c.phi[n * s + icv]; };
  auto phb = [&](int s, int ifa) { return c.phi_bfa[nb * s + ifa]; };
  auto gph = [&](int s, int icv) { return &c.grad_phi[3 * (n * s + icv)]; };
  auto rs = [&](int s, int icv) -> double & { return c.rhs_scal[n * s + icv]; };
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    int noc00 = 0, noc01 = 0, noc11 = 0, noc10 = 0;
    if (A) {  // getImplDependencyIndex, UgpWithCv.h:745-755
      noc00 = m.nbocv_i[icv0];
      noc01 = noc00;
      while (m.nbocv_v[++noc01] != icv1) {}
      noc11 = m.nbocv_i[icv1];
      noc10 = noc11;
      while (m.nbocv_v[++noc10] != icv0) {}
    }
    double nVec[3];
    double area = norm3(nVec, &m.fa_normal[3 * ifa]);
    double rho0 = c.rho[icv0], p0 = c.press[icv0];
    double u0[3] = {c.vel[3 * icv0], c.vel[3 * icv0 + 1], c.vel[3 * icv0 + 2]};
    double rho1 = c.rho[icv1], p1 = c.press[icv1];
    double u1[3] = {c.vel[3 * icv1], c.vel[3 * icv1 + 1], c.vel[3 * icv1 + 2]};
    double sc0[2], sc1[2], cv0[2], cv1[2];  // Scalar0/1 (faces), ScalCV0/1 (cells)
    for (int s = 0; s < nsc; s++) {
      sc0[s] = cv0[s] = ph(s, icv0);
      sc1[s] = cv1[s] = ph(s, icv1);
    }
    if (snd) {
      double r0[3], r1[3];
      for (int i = 0; i < 3; i++) {
        r0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
        r1[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv1 + i];
      }
      rho0 += dot3(r0, &c.grad_rho[3 * icv0]);
      p0 += dot3(r0, &c.grad_p[3 * icv0]);
      if ((p0 <= 0.0) || (rho0 <= 0.0)) {
        p0 = c.press[icv0];
        rho0 = c.rho[icv0];
      } else {
        for (int i = 0; i < 3; i++) u0[i] += dot3(r0, &c.grad_u[9 * icv0 + 3 * i]);
        for (int s = 0; s < nsc; s++) sc0[s] = sc0[s] + dot3(r0, gph(s, icv0));  // :2862-2869
      }
      rho1 += dot3(r1, &c.grad_rho[3 * icv1]);
      p1 += dot3(r1, &c.grad_p[3 * icv1]);
      if ((p1 <= 0.0) || (rho1 <= 0.0)) {
        p1 = c.press[icv1];
        rho1 = c.rho[icv1];
      } else {
        for (int i = 0; i < 3; i++) u1[i] += dot3(r1, &c.grad_u[9 * icv1 + 3 * i]);
        for (int s = 0; s < nsc; s++) sc1[s] = sc1[s] + dot3(r1, gph(s, icv1));
      }
    }
    // kineFA0/1 and kineCV0/1 (:2924-2930)
    const double kfa0 = c.sst ? sc0[0] : 0.0, kfa1 = c.sst ? sc1[0] : 0.0;
    const double kcv0 = c.sst ? cv0[0] : 0.0, kcv1 = c.sst ? cv1[0] : 0.0;
    face_flux(c, F, rho0, u0, p0, kfa0, rho1, u1, p1, kfa1, area, nVec, sc0, sc1, nsc ? Fs : nullptr);
    keep_flux(c, ifa, F);
    c.rhs_rho[icv0] -= F[0];
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
    c.rhs_rhoE[icv0] -= F[4];
    for (int s = 0; s < nsc; s++) rs(s, icv0) -= 1.0 * Fs[s];  // ScalConvTerm = 1 (:3014-3015)
    c.rhs_rho[icv1] += F[0];
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv1 + i] += F[1 + i];
    c.rhs_rhoE[icv1] += F[4];
    for (int s = 0; s < nsc; s++) rs(s, icv1) += 1.0 * Fs[s];
    if (A) {
      face_jac(c, Apl, Ami, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], kcv0, c.rho[icv1], &c.vel[3 * icv1],
               c.press[icv1], kcv1, area, nVec, cv0, cv1, nsc ? SL : nullptr, nsc ? SR : nullptr);
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc00 + 5 * i + j] += Apl[i][j];
          A[25 * noc01 + 5 * i + j] += Ami[i][j];
        }
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc11 + 5 * i + j] -= Ami[i][j];
          A[25 * noc10 + 5 * i + j] -= Apl[i][j];
        }
      for (int s = 0; s < nsc; s++) {  // :3088-3101
        add_ascal(c, noc00, 1.0, SL + 6 * s, s);
        add_ascal(c, noc01, 1.0, SR + 6 * s, s);
        add_ascal(c, noc11, -1.0, SR + 6 * s, s);
        add_ascal(c, noc10, -1.0, SL + 6 * s, s);
      }
    }
  }
  // boundary zones in faZoneList order (JoeWithModels.cpp:3112-3569)
  for (int iz = 0; iz < m.nzone; iz++) {
    const int bc = m.zone_bc[iz];
    for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) {
      int icv0 = m.cvofa[2 * ifa];
      double nVec[3];
      double area = norm3(nVec, &m.fa_normal[3 * ifa]);
      const double *ub = &c.vel_bfa[3 * ifa];
      double sb[2];  // phi_bfa on both sides
      for (int s = 0; s < nsc; s++) sb[s] = phb(s, ifa);
      if (bc == OR_BC_SYMMETRY || bc == OR_BC_WALL) {
        // both sides phi_bfa, kineFA = kine's phi_bfa (:3134-3147, 3291-3294)
        const double kfa = c.sst ? sb[0] : 0.0;
        face_flux(c, F, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, area, nVec, sb,
                  sb, nsc ? Fs : nullptr);
        if (bc == OR_BC_SYMMETRY && c.p.symm_bc == OR_SYMM_ANALYTICAL) {
          const double Fp[5] = {0.0, c.p_bfa[ifa] * nVec[0] * area, c.p_bfa[ifa] * nVec[1] * area,
                                c.p_bfa[ifa] * nVec[2] * area, 0.0};
          keep_flux(c, ifa, Fp);
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= c.p_bfa[ifa] * nVec[i] * area;
          c.rhs_rho[icv0] -= 0.0;
          c.rhs_rhoE[icv0] -= 0.0;
        } else {
          keep_flux(c, ifa, F);
          c.rhs_rho[icv0] -= F[0];
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
          c.rhs_rhoE[icv0] -= F[4];
        }
        for (int s = 0; s < nsc; s++) rs(s, icv0) -= 1.0 * Fs[s];  // :3197-3198, 3341-3342
        if (A) {
          int noc00 = m.nbocv_i[icv0];
          if (nsc) {  // scalar rows from the bfa/bfa HLLC Jacobian (:3202-3206, 3346-3350, 3270-3274)
            double Al[5][5];
            face_jac(c, Al, NULL, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, area,
                     nVec, sb, sb, SL, nullptr);
            for (int s = 0; s < nsc; s++) add_ascal(c, noc00, 1.0, SL + 6 * s, s);
          }
          if (bc == OR_BC_WALL || c.p.symm_bc == OR_SYMM_ORIGINAL_NUMERIC) {
            face_jac(c, Apl, NULL, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, area,
                     nVec);
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += Apl[k / 5][k % 5];
          } else if (c.p.symm_bc == OR_SYMM_NEW_NUMERIC) {
            face_jac(c, Apl, Ami, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa, area,
                     nVec);
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += Apl[k / 5][k % 5] + Ami[k / 5][k % 5];
          } else {  // ANALYTICAL (JoeWithModels.cpp:3243-3269)
            double S[5][5];
            double a2 = c.p.gamma - 1.0;
            double phi = 0.5 * a2 * (ub[0] * ub[0] + ub[1] * ub[1] + ub[2] * ub[2]);
            for (int i = 0; i < 5; i++) {
              S[0][i] = 0.0;
              S[4][i] = 0.0;
            }
            for (int i = 0; i < 3; i++) {
              S[i + 1][0] = phi * nVec[i] * area;
              for (int j = 0; j < 3; j++) S[i + 1][j + 1] = -a2 * ub[j] * nVec[i] * area;
              S[i + 1][4] = a2 * nVec[i] * area;
            }
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += S[k / 5][k % 5];
          }
        }
      } else {
        double rho0 = c.rho[icv0], p0 = c.press[icv0];
        double u0[3] = {c.vel[3 * icv0], c.vel[3 * icv0 + 1], c.vel[3 * icv0 + 2]};
        double sc0[2], cv0[2];  // :3413-3419
        for (int s = 0; s < nsc; s++) sc0[s] = cv0[s] = ph(s, icv0);
        if (snd) {
          double r0[3];
          for (int i = 0; i < 3; i++) r0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
          rho0 += dot3(r0, &c.grad_rho[3 * icv0]);
          p0 += dot3(r0, &c.grad_p[3 * icv0]);
          if ((p0 <= 0.0) || (rho0 <= 0.0)) {
            p0 = c.press[icv0];
            rho0 = c.rho[icv0];
          } else {
            for (int i = 0; i < 3; i++) u0[i] += dot3(r0, &c.grad_u[9 * icv0 + 3 * i]);
            for (int s = 0; s < nsc; s++) sc0[s] = sc0[s] + dot3(r0, gph(s, icv0));  // :3445-3454
          }
        }
        // kineCV0, kineFA0 (reconstructed), kineFA1 = phi_bfa (:3466-3471)
        const double kcv0 = c.sst ? cv0[0] : 0.0, kfa0 = c.sst ? sc0[0] : 0.0, kfa1 = c.sst ? sb[0] : 0.0;
        face_flux(c, F, rho0, u0, p0, kfa0, c.rho_bfa[ifa], ub, c.p_bfa[ifa], kfa1, area, nVec, sc0, sb,
                  nsc ? Fs : nullptr);
        keep_flux(c, ifa, F);
        c.rhs_rho[icv0] -= F[0];
        for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
        c.rhs_rhoE[icv0] -= F[4];
        for (int s = 0; s < nsc; s++) rs(s, icv0) -= 1.0 * Fs[s];
        if (A) {
          face_jac(c, Apl, NULL, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], kcv0, c.rho_bfa[ifa], ub, c.p_bfa[ifa],
                   kfa1, area, nVec, cv0, sb, nsc ? SL : nullptr, nullptr);
          int noc00 = m.nbocv_i[icv0];
          for (int k = 0; k < 25; k++) A[25 * noc00 + k] += Apl[k / 5][k % 5];
          for (int s = 0; s < nsc; s++) add_ascal(c, noc00, 1.0, SL + 6 * s, s);  // :3559-3561
        }
      }
    }
  }
}

// calcViscousFluxCoupled, UgpWithCvCompFlow.cpp:3040-3274, nScal = 0, no
// turbulence model (kine = 0, psi = 0). gu0/gu1 are [3][3] velocity
// gradients, gh0/gh1 enthalpy gradients. A0 / A1 (may be NULL) get only the
// entries the reference writes; the caller zeroes them once.
static void visc_coupled(double F[5], double (*A0)[5], double (*A1)[5], double rho0, const double *u0,
                         const double *gu0, double h0, const double *gh0, double T0, double R0, double gam0,
                         double rho1, const double *u1, const double *gu1, double h1, const double *gh1, double T1,
                         double R1, double gam1, double mul, double mut, double lamOcp, double pr_turb,
                         const double *u_fa, double area, const double *nVec, double smag, const double *sVec,
                         double alpha, bool turb = false) {
  const double kine0 = 0.0, kine1 = 0.0;
  double g[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) g[i][j] = 0.5 * (gu0[3 * i + j] + gu1[3 * i + j]);
  double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  double mu = mul + mut;
  double tau[3];
  F[0] = 0.0;
  for (int i = 0; i < 3; i++)
    tau[i] = mu * (alpha * (u1[i] - u0[i]) / smag + g[i][0] * fCorr[0] + g[0][i] * nVec[0] + g[i][1] * fCorr[1] +
                   g[1][i] * nVec[1] + g[i][2] * fCorr[2] + g[2][i] * nVec[2]);
  double tmp = 2.0 / 3.0 * mu * (g[0][0] + g[1][1] + g[2][2]);
  if (turb) tmp += 1.0 / 3.0 * (rho0 + rho1) * 0.0;  // kine_fa = 0 (:3084-3085)
  for (int i = 0; i < 3; i++) tau[i] -= tmp * nVec[i];
  for (int i = 0; i < 3; i++) F[1 + i] = -area * tau[i];
  double keff = lamOcp + mut / pr_turb;
  double enth = keff * (alpha * (h1 - h0) / smag + 0.5 * ((gh0[0] + gh1[0]) * fCorr[0] + (gh0[1] + gh1[1]) * fCorr[1] +
                                                         (gh0[2] + gh1[2]) * fCorr[2]));
  double psi = 0.0;
  if (turb) psi = mu * (alpha * (kine1 - kine0) / smag);  // :3108-3110
  F[4] = -area * (enth + tau[0] * u_fa[0] + tau[1] * u_fa[1] + tau[2] * u_fa[2] + psi);
  auto side = [&](double (*A)[5], double sg, double rho, const double *u, double h, double T, double R, double gam,
                  double kine) {
    double aas = area * alpha / smag;
    double d11 = sg > 0.0 ? mu * aas : -mu * aas;
    double d44 = sg > 0.0 ? keff * aas : -keff * aas;
    double d4[3];
    for (int i = 0; i < 3; i++) d4[i] = -area * tau[i] * 0.5 + u_fa[i] * d11;
    double invRho = 1.0 / rho;
    double q0[3], q4[3];
    for (int i = 0; i < 3; i++) q0[i] = -invRho * u[i];
    double q40 = invRho * gam * (-h - kine + R * T + 0.5 * dot3(u, u));
    for (int i = 0; i < 3; i++) q4[i] = -invRho * u[i] * gam;
    double q44 = invRho * gam;
    for (int i = 0; i < 3; i++) {
      A[1 + i][0] = d11 * q0[i];
      A[1 + i][1 + i] = d11 * invRho;
    }
    A[4][0] = d4[0] * q0[0] + d4[1] * q0[1] + d4[2] * q0[2] + d44 * q40;
    for (int i = 0; i < 3; i++) A[4][1 + i] = d4[i] * invRho + d44 * q4[i];
    A[4][4] = d44 * q44;
  };
  if (A0) side(A0, 1.0, rho0, u0, h0, T0, R0, gam0, kine0);
  if (A1) side(A1, -1.0, rho1, u1, h1, T1, R1, gam1, kine1);
}

// The transported scalar's row of calcViscousFluxCoupled (UgpWithCvCompFlow.cpp
// :3113-3120 flux, :3158-3167 and :3217-3226 dF/dU, :3185-3190 and :3244-3249
// dU/dQ): *Fs and the row's entries A0[5][0], A0[5][5] (a0, may be NULL) and
// A1[5][0], A1[5][5] (a1, may be NULL).
static void visc_coupled_srow(double *Fs, double *a0, double *a1, double rho0, double scal0, const double *gs0,
                              double rho1, double scal1, const double *gs1, double diff, double DiffTerm, double area,
                              const double *nVec, double smag, const double *sVec, double alpha) {
  double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  *Fs = -area * DiffTerm * diff *
        (alpha * (scal1 - scal0) / smag +
         0.5 * ((gs0[0] + gs1[0]) * fCorr[0] + (gs0[1] + gs1[1]) * fCorr[1] + (gs0[2] + gs1[2]) * fCorr[2]));
  double area_alpha_smag = area * alpha / smag;
  if (a0) {
    double dFdU = DiffTerm * diff * area_alpha_smag;
    double invRho = 1.0 / rho0;
    a0[0] = dFdU * (-invRho * scal0);
    a0[1] = dFdU * invRho;
  }
  if (a1) {
    double dFdU = -DiffTerm * diff * area_alpha_smag;
    double invRho = 1.0 / rho1;
    a1[0] = dFdU * (-invRho * scal1);
    a1[1] = dFdU * invRho;
  }
}

// calcFluxCoupled, JoeWithModels.cpp:4613-5825, for nScal = 0, MU_REF = 0 and
// HLLC: the coupled HLLC flux and Jacobian, each face side's enthalpy from
// calcThermoProp_p after reconstruction (the cell's when first order), and
// the coupled boundary treatments: SYMMETRY and WALL by SYMM_BC (ANALYTICAL:
// the ONLY_PRESSURE flux and Jacobian), NEUMANN with the boundary state on
// both sides, the others reconstructed against the boundary state. The
// effective SYMM_BC default is ORIGINAL_NUMERIC here too: init registers it
// (UgpWithCvCompFlow.h:479) before calcFluxCoupled asks with ANALYTICAL.
void sa_diffusivity(or_ctx &c);

void calc_flux_coupled(or_ctx &c, double *A) {
  const or_mesh &m = c.m;
  const bool snd = c.p.second_order != 0;
  const bool visc = c.p.mu_ref > 0.0;
  if (snd) euler_gradients(c);
  const bool ls = c.p.grad_scheme == OR_GRAD_LS;
  if (visc) {  // grad_u also when first order, and grad_enthalpy, both limited by the limiter of GRAD_RECONSTRUCTION (:4691-4716)
    if (!snd) {
      if (ls) grad_vector_ls(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
      else grad_vector_gg(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
      if (c.p.limiter) limit_vector(c, c.grad_u.data(), c.vel.data(), c.sos.data());
    }
    if (ls) grad_scalar_ls(c, c.grad_h.data(), c.enthalpy.data(), c.h_bfa.data());
    else grad_scalar_gg(c, c.grad_h.data(), c.enthalpy.data(), c.h_bfa.data());
    if (c.p.limiter) limit_scalar(c, c.grad_h.data(), c.enthalpy.data(), c.enthalpy.data(), NULL);
  }
  // one transported scalar (Spalart-Allmaras) in the coupled system: its
  // limited gradient (STANDARD reconstruction or MU_REF > 0, :4718-4738; done
  // by euler_gradients when second order), the diffusivity (:4739-4743), and
  // its row of each 6x6 block, kept in AScal[k][nnz] = A[noc][5][k]
  const bool sa = c.sa;
  const size_t nnz = c.nnz;
  double *AS = (sa && A) ? c.AScal.data() : NULL;
  if (sa) {
    if (!snd) {
      if (ls) grad_scalar_ls(c, c.grad_phi.data(), c.phi.data(), c.phi_bfa.data());
      else grad_scalar_gg(c, c.grad_phi.data(), c.phi.data(), c.phi_bfa.data());
      if (c.p.limiter) limit_scalar(c, c.grad_phi.data(), c.phi.data(), c.phi.data(), NULL);
    }
    sa_diffusivity(c);
  }
  auto srow = [&](int noc, const double *r, double sg) {
    for (int k = 0; k < 6; k++) {
      if (sg > 0.0) AS[k * nnz + noc] += r[k];
      else AS[k * nnz + noc] -= r[k];
    }
  };
  auto vrow = [&](int noc, const double *a, double sg) {  // entries [5][0] and [5][5]
    if (sg > 0.0) {
      AS[0 * nnz + noc] += a[0];
      AS[5 * nnz + noc] += a[1];
    } else {
      AS[0 * nnz + noc] -= a[0];
      AS[5 * nnz + noc] -= a[1];
    }
  };
  const double zero3[3] = {0.0, 0.0, 0.0};
  double V0[5][5], V1[5][5], VF[5];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) V0[i][j] = V1[i][j] = 0.0;
  const double g = c.p.gamma, R = c.p.R_gas;
  auto thermo_h = [&](double rho, double p) {  // calcThermoProp_p, UgpWithCvCompFlow.h:1765-1771
    double T = p / (rho * R);
    return g * R / (g - 1.0) * T;
  };
  double F[5], Apl[5][5], Ami[5][5], Fs = 0.0, SL[6], SR[6], VFs = 0.0, a0[2], a1[2];
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {  // :4797-5105
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    int noc00 = 0, noc01 = 0, noc11 = 0, noc10 = 0;
    if (A) {
      noc00 = m.nbocv_i[icv0];
      noc01 = noc00;
      while (m.nbocv_v[++noc01] != icv1) {}
      noc11 = m.nbocv_i[icv1];
      noc10 = noc11;
      while (m.nbocv_v[++noc10] != icv0) {}
    }
    double nVec[3];
    double area = norm3(nVec, &m.fa_normal[3 * ifa]);
    double rho0 = c.rho[icv0], p0 = c.press[icv0], h0 = c.enthalpy[icv0];
    double u0[3] = {c.vel[3 * icv0], c.vel[3 * icv0 + 1], c.vel[3 * icv0 + 2]};
    double rho1 = c.rho[icv1], p1 = c.press[icv1], h1 = c.enthalpy[icv1];
    double u1[3] = {c.vel[3 * icv1], c.vel[3 * icv1 + 1], c.vel[3 * icv1 + 2]};
    double sc[2] = {sa ? c.phi[icv0] : 0.0, sa ? c.phi[icv1] : 0.0};
    if (snd) {
      double r0[3], r1[3];
      for (int i = 0; i < 3; i++) {
        r0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
        r1[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv1 + i];
      }
      rho0 += dot3(r0, &c.grad_rho[3 * icv0]);
      p0 += dot3(r0, &c.grad_p[3 * icv0]);
      if ((p0 <= 0.0) || (rho0 <= 0.0)) {
        p0 = c.press[icv0];
        rho0 = c.rho[icv0];
      } else {
        for (int i = 0; i < 3; i++) u0[i] += dot3(r0, &c.grad_u[9 * icv0 + 3 * i]);
        if (sa) sc[0] += dot3(r0, &c.grad_phi[3 * icv0]);
      }
      rho1 += dot3(r1, &c.grad_rho[3 * icv1]);
      p1 += dot3(r1, &c.grad_p[3 * icv1]);
      if ((p1 <= 0.0) || (rho1 <= 0.0)) {
        p1 = c.press[icv1];
        rho1 = c.rho[icv1];
      } else {
        for (int i = 0; i < 3; i++) u1[i] += dot3(r1, &c.grad_u[9 * icv1 + 3 * i]);
        if (sa) sc[1] += dot3(r1, &c.grad_phi[3 * icv1]);
      }
      h0 = thermo_h(rho0, p0);
      h1 = thermo_h(rho1, p1);
    }
    flux_hllc_coupled(F, rho0, u0, p0, h0, g, rho1, u1, p1, h1, g, area, nVec, false, sc, sa ? &Fs : NULL);
    keep_flux(c, ifa, F);
    c.rhs_rho[icv0] -= F[0];
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
    c.rhs_rhoE[icv0] -= F[4];
    if (sa) c.rhs_scal[icv0] -= Fs;
    c.rhs_rho[icv1] += F[0];
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv1 + i] += F[1 + i];
    c.rhs_rhoE[icv1] += F[4];
    if (sa) c.rhs_scal[icv1] += Fs;
    if (A) {  // cell-centre states (:5035-5038)
      jac_hllc_coupled(Apl, Ami, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], c.enthalpy[icv0], g, c.rho[icv1],
                       &c.vel[3 * icv1], c.press[icv1], c.enthalpy[icv1], g, area, nVec, false);
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc00 + 5 * i + j] += Apl[i][j];
          A[25 * noc01 + 5 * i + j] += Ami[i][j];
        }
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc11 + 5 * i + j] -= Ami[i][j];
          A[25 * noc10 + 5 * i + j] -= Apl[i][j];
        }
      if (sa) {
        jac_hllc_coupled_srow(SL, SR, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], c.enthalpy[icv0], c.phi[icv0], g,
                              c.rho[icv1], &c.vel[3 * icv1], c.press[icv1], c.enthalpy[icv1], c.phi[icv1], g, area,
                              nVec, false);
        srow(noc00, SL, 1.0);
        srow(noc01, SR, 1.0);
        srow(noc11, SR, -1.0);
        srow(noc10, SL, -1.0);
      }
    }
    if (visc) {  // the same face's viscous flux and blocks (:5076-5105)
      double sVec[3];
      for (int i = 0; i < 3; i++) sVec[i] = m.x_cv[3 * icv1 + i] - m.x_cv[3 * icv0 + i];
      double smag = sqrt(dot3(sVec, sVec));
      for (int i = 0; i < 3; i++) sVec[i] /= smag;
      double alpha = dot3(nVec, sVec);
      const double *v0 = &c.vel[3 * icv0], *v1 = &c.vel[3 * icv1];
      double uAux_fa[3] = {0.5 * (v0[0] + v1[0]), 0.5 * (v0[1] + v1[1]), 0.5 * (v0[2] + v1[2])};
      visc_coupled(VF, A ? V0 : NULL, A ? V1 : NULL, c.rho[icv0], v0, &c.grad_u[9 * icv0], c.enthalpy[icv0],
                   &c.grad_h[3 * icv0], c.temp[icv0], R, g, c.rho[icv1], v1, &c.grad_u[9 * icv1], c.enthalpy[icv1],
                   &c.grad_h[3 * icv1], c.temp[icv1], R, g, c.mul_fa[ifa], c.mut_fa[ifa], c.lamOcp_fa[ifa],
                   c.p.pr_turb, uAux_fa, area, nVec, smag, sVec, alpha, sa);
      if (sa)
        visc_coupled_srow(&VFs, A ? a0 : NULL, A ? a1 : NULL, c.rho[icv0], c.phi[icv0], &c.grad_phi[3 * icv0],
                          c.rho[icv1], c.phi[icv1], &c.grad_phi[3 * icv1], c.diff[ifa], 1.0, area, nVec, smag, sVec,
                          alpha);
      c.rhs_rho[icv0] -= VF[0];
      for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= VF[1 + i];
      c.rhs_rhoE[icv0] -= VF[4];
      if (sa) c.rhs_scal[icv0] -= VFs;
      c.rhs_rho[icv1] += VF[0];
      for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv1 + i] += VF[1 + i];
      c.rhs_rhoE[icv1] += VF[4];
      if (sa) c.rhs_scal[icv1] += VFs;
      if (A) {
        for (int k = 0; k < 25; k++) {
          A[25 * noc00 + k] += V0[k / 5][k % 5];
          A[25 * noc01 + k] += V1[k / 5][k % 5];
        }
        for (int k = 0; k < 25; k++) {
          A[25 * noc11 + k] -= V1[k / 5][k % 5];
          A[25 * noc10 + k] -= V0[k / 5][k % 5];
        }
        if (sa) {
          vrow(noc00, a0, 1.0);
          vrow(noc01, a1, 1.0);
          vrow(noc11, a1, -1.0);
          vrow(noc10, a0, -1.0);
        }
      }
    }
  }
  for (int iz = 0; iz < m.nzone; iz++) {  // :5113-5796
    const int bc = m.zone_bc[iz];
    // a zone without a nuSA value carries no viscous scalar flux (DiffTerm = 0,
    // :5453-5461, 5599-5607); NEUMANN zones keep it (:5293-5296)
    const double DiffTerm = (bc == OR_BC_NEUMANN || m.zone_scal[2 * iz] != 0.0) ? 1.0 : 0.0;
    for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) {
      int icv0 = m.cvofa[2 * ifa];
      int noc00 = m.nbocv_i[icv0];
      double nVec[3];
      double area = norm3(nVec, &m.fa_normal[3 * ifa]);
      const double *ub = &c.vel_bfa[3 * ifa];
      const double rb = c.rho_bfa[ifa], pb = c.p_bfa[ifa], hb = c.h_bfa[ifa];
      const double phib = sa ? c.phi_bfa[ifa] : 0.0;
      const double scb[2] = {phib, phib};
      if (bc == OR_BC_SYMMETRY || bc == OR_BC_WALL) {
        const bool op = c.p.symm_bc == OR_SYMM_ANALYTICAL;
        flux_hllc_coupled(F, rb, ub, pb, hb, g, rb, ub, pb, hb, g, area, nVec, op, scb, sa ? &Fs : NULL);
        keep_flux(c, ifa, F);
        c.rhs_rho[icv0] -= F[0];
        for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
        c.rhs_rhoE[icv0] -= F[4];
        if (sa) c.rhs_scal[icv0] -= Fs;
        if (A) {
          const bool nn = c.p.symm_bc == OR_SYMM_NEW_NUMERIC;
          jac_hllc_coupled(Apl, nn ? Ami : NULL, rb, ub, pb, hb, g, rb, ub, pb, hb, g, area, nVec, op);
          for (int k = 0; k < 25; k++)
            if (nn) A[25 * noc00 + k] += Apl[k / 5][k % 5] + Ami[k / 5][k % 5];
            else A[25 * noc00 + k] += Apl[k / 5][k % 5];
          if (sa) {
            jac_hllc_coupled_srow(SL, nn ? SR : NULL, rb, ub, pb, hb, phib, g, rb, ub, pb, hb, phib, g, area, nVec,
                                  op);
            if (nn)
              for (int k = 0; k < 6; k++) AS[k * nnz + noc00] += SL[k] + SR[k];
            else
              srow(noc00, SL, 1.0);
          }
        }
        if (visc && bc == OR_BC_SYMMETRY) {  // 2/3 (mu + mut) tr(grad u) only (:5272-5284)
          const double *gu = &c.grad_u[9 * icv0];
          double tmp = 2.0 / 3.0 * ((c.mul_fa[ifa] + c.mut_fa[ifa]) * (gu[0] + gu[4] + gu[8]) + rb * 0.0);
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= area * tmp * nVec[i];
        } else if (visc) {  // wall: the cell against the boundary state, mut = 0 (:5548-5575)
          double sVec[3];
          for (int i = 0; i < 3; i++) sVec[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
          double smag = fabs(dot3(sVec, nVec));
          const double *gu = &c.grad_u[9 * icv0], *gh = &c.grad_h[3 * icv0];
          visc_coupled(VF, A ? V0 : NULL, NULL, c.rho[icv0], &c.vel[3 * icv0], gu, c.enthalpy[icv0], gh, c.temp[icv0],
                       R, g, rb, ub, gu, hb, gh, c.T_bfa[ifa], R, g, c.mul_fa[ifa], 0.0, c.lamOcp_fa[ifa], c.p.pr_turb,
                       ub, area, nVec, smag, nVec, 1.0, sa);
          if (sa)
            visc_coupled_srow(&VFs, A ? a0 : NULL, NULL, c.rho[icv0], c.phi[icv0], &c.grad_phi[3 * icv0], rb, phib,
                              &c.grad_phi[3 * icv0], c.diff[ifa], DiffTerm, area, nVec, smag, nVec, 1.0);
          c.rhs_rho[icv0] -= VF[0];
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= VF[1 + i];
          c.rhs_rhoE[icv0] -= VF[4];
          if (sa) c.rhs_scal[icv0] -= VFs;
          if (A) {
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += V0[k / 5][k % 5];
            if (sa) vrow(noc00, a0, 1.0);
          }
        }
      } else if (bc == OR_BC_NEUMANN) {
        flux_hllc_coupled(F, rb, ub, pb, hb, g, rb, ub, pb, hb, g, area, nVec, false, scb, sa ? &Fs : NULL);
        keep_flux(c, ifa, F);
        c.rhs_rho[icv0] -= F[0];
        for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
        c.rhs_rhoE[icv0] -= F[4];
        if (sa) c.rhs_scal[icv0] -= Fs;
        if (A) {
          jac_hllc_coupled(Apl, Ami, rb, ub, pb, hb, g, rb, ub, pb, hb, g, area, nVec, false);
          for (int k = 0; k < 25; k++)
            if (c.p.symm_bc == OR_SYMM_NEW_NUMERIC) A[25 * noc00 + k] += Apl[k / 5][k % 5] + Ami[k / 5][k % 5];
            else A[25 * noc00 + k] += Apl[k / 5][k % 5];
          if (sa) {
            jac_hllc_coupled_srow(SL, SR, rb, ub, pb, hb, phib, g, rb, ub, pb, hb, phib, g, area, nVec, false);
            if (c.p.symm_bc == OR_SYMM_NEW_NUMERIC)
              for (int k = 0; k < 6; k++) AS[k * nnz + noc00] += SL[k] + SR[k];
            else
              srow(noc00, SL, 1.0);
          }
        }
        if (visc) {  // the boundary state on both sides, s along the normal (:5373-5404)
          double sVec[3];
          for (int i = 0; i < 3; i++) sVec[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
          double smag = fabs(dot3(sVec, nVec));
          const double *gu = &c.grad_u[9 * icv0], *gh = &c.grad_h[3 * icv0];
          visc_coupled(VF, A ? V0 : NULL, A ? V1 : NULL, rb, ub, gu, hb, gh, c.T_bfa[ifa], R, g, rb, ub, gu, hb, gh,
                       c.T_bfa[ifa], R, g, c.mul_fa[ifa], c.mut_fa[ifa], c.lamOcp_fa[ifa], c.p.pr_turb, ub, area,
                       nVec, smag, nVec, 1.0, sa);
          if (sa)  // the scalar gradients are zero here (:5381-5386)
            visc_coupled_srow(&VFs, A ? a0 : NULL, NULL, rb, phib, zero3, rb, phib, zero3, c.diff[ifa], DiffTerm, area,
                              nVec, smag, nVec, 1.0);
          c.rhs_rho[icv0] -= VF[0];
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= VF[1 + i];
          c.rhs_rhoE[icv0] -= VF[4];
          if (sa) c.rhs_scal[icv0] -= VFs;
          if (A) {
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += V0[k / 5][k % 5];
            if (sa) vrow(noc00, a0, 1.0);
          }
        }
      } else {
        double rho0 = c.rho[icv0], p0 = c.press[icv0], h0 = c.enthalpy[icv0];
        double u0[3] = {c.vel[3 * icv0], c.vel[3 * icv0 + 1], c.vel[3 * icv0 + 2]};
        double sc[2] = {sa ? c.phi[icv0] : 0.0, phib};
        if (snd) {
          double r0[3];
          for (int i = 0; i < 3; i++) r0[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
          rho0 += dot3(r0, &c.grad_rho[3 * icv0]);
          p0 += dot3(r0, &c.grad_p[3 * icv0]);
          if ((p0 <= 0.0) || (rho0 <= 0.0)) {
            p0 = c.press[icv0];
            rho0 = c.rho[icv0];
          } else {
            for (int i = 0; i < 3; i++) u0[i] += dot3(r0, &c.grad_u[9 * icv0 + 3 * i]);
            if (sa) sc[0] += dot3(r0, &c.grad_phi[3 * icv0]);
          }
          h0 = thermo_h(rho0, p0);
        }
        flux_hllc_coupled(F, rho0, u0, p0, h0, g, rb, ub, pb, hb, g, area, nVec, false, sc, sa ? &Fs : NULL);
        keep_flux(c, ifa, F);
        c.rhs_rho[icv0] -= F[0];
        for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= F[1 + i];
        c.rhs_rhoE[icv0] -= F[4];
        if (sa) c.rhs_scal[icv0] -= Fs;
        if (A) {  // cell-centre state against the boundary state (:5731-5735)
          jac_hllc_coupled(Apl, NULL, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], c.enthalpy[icv0], g, rb, ub, pb,
                           hb, g, area, nVec, false);
          for (int k = 0; k < 25; k++) A[25 * noc00 + k] += Apl[k / 5][k % 5];
          if (sa) {
            jac_hllc_coupled_srow(SL, NULL, c.rho[icv0], &c.vel[3 * icv0], c.press[icv0], c.enthalpy[icv0],
                                  c.phi[icv0], g, rb, ub, pb, hb, phib, g, area, nVec, false);
            srow(noc00, SL, 1.0);
          }
        }
        if (visc) {  // the cell against the boundary state (:5744-5775)
          double sVec[3];
          for (int i = 0; i < 3; i++) sVec[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
          double smag = sqrt(dot3(sVec, sVec));
          for (int i = 0; i < 3; i++) sVec[i] /= smag;
          double alpha = dot3(nVec, sVec);
          const double *gu = &c.grad_u[9 * icv0], *gh = &c.grad_h[3 * icv0];
          visc_coupled(VF, A ? V0 : NULL, NULL, c.rho[icv0], &c.vel[3 * icv0], gu, c.enthalpy[icv0], gh, c.temp[icv0],
                       R, g, rb, ub, gu, hb, gh, c.T_bfa[ifa], R, g, c.mul_fa[ifa], c.mut_fa[ifa], c.lamOcp_fa[ifa],
                       c.p.pr_turb, ub, area, nVec, smag, sVec, alpha, sa);
          // the boundary side's scalar is the cell's reconstructed face value
          // (Scalar0), as the reference passes it (:5764-5769)
          if (sa)
            visc_coupled_srow(&VFs, A ? a0 : NULL, NULL, c.rho[icv0], c.phi[icv0], &c.grad_phi[3 * icv0], rb, sc[0],
                              &c.grad_phi[3 * icv0], c.diff[ifa], DiffTerm, area, nVec, smag, sVec, alpha);
          c.rhs_rho[icv0] -= VF[0];
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= VF[1 + i];
          c.rhs_rhoE[icv0] -= VF[4];
          if (sa) c.rhs_scal[icv0] -= VFs;
          if (A) {
            for (int k = 0; k < 25; k++) A[25 * noc00 + k] += V0[k / 5][k % 5];
            if (sa) vrow(noc00, a0, 1.0);
          }
        }
      }
    }
  }
}

// calcRhs, JoeWithModels.cpp:2564-2594, with the default sourceHook
// (gravity, JoeWithModels.h:111-129).
// addViscFlux, UgpWithCvCompFlow.cpp:1521-1763, laminar (mut = 0, no
// turbulence model, nonLinear = rij = kine = 0 as allocated at
// UgpWithCvCompFlow.h:532-545), SIMPLEX_VISCOUS = NO. A0/A1 rows 1-4 are
// written by dFdU x dUdQ (:712-745); row 0 stays as the caller left it.
void add_visc_flux(const or_ctx &c, double *Frhou, double &FrhoE, double (*A0)[5], double (*A1)[5], double rho0,
                   const double *u0, const double *grad_u0, double h0, const double *grad_h0, double T0, double R0,
                   double gam0, double kine0, double rho1, const double *u1, const double *grad_u1, double h1,
                   const double *grad_h1, double T1, double R1, double gam1, double kine1, double mul, double mut,
                   double lambdaOverCp, const double *u_fa, double area, const double *nVec, double smag,
                   const double *sVec, double kine_fa = 0.0) {
  const double nonL = 0.0;
  const double rij_d[3] = {0.0, 0.0, 0.0}, rij_offd[3] = {0.0, 0.0, 0.0};
  double alpha = dot3(nVec, sVec);
  double grad_u_f[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) grad_u_f[i][j] = 0.5 * (grad_u0[3 * i + j] + grad_u1[3 * i + j]);
  double fCorr[3] = {nVec[0] - alpha * sVec[0], nVec[1] - alpha * sVec[1], nVec[2] - alpha * sVec[2]};
  double tauij_nj[3], tauLamij_nj[3], tauTurbij_nj[3], enth, tmp;
  for (int i = 0; i < 3; i++)
    tauLamij_nj[i] = mul * (alpha * (u1[i] - u0[i]) / smag + grad_u_f[i][0] * fCorr[0] + grad_u_f[0][i] * nVec[0] +
                            grad_u_f[i][1] * fCorr[1] + grad_u_f[1][i] * nVec[1] + grad_u_f[i][2] * fCorr[2] +
                            grad_u_f[2][i] * nVec[2]);
  tmp = 2.0 / 3.0 * mul * (grad_u_f[0][0] + grad_u_f[1][1] + grad_u_f[2][2]);
  tauLamij_nj[0] -= tmp * nVec[0];
  tauLamij_nj[1] -= tmp * nVec[1];
  tauLamij_nj[2] -= tmp * nVec[2];
  for (int i = 0; i < 3; i++)
    tauTurbij_nj[i] = (1.0 - nonL) * mut *
                      (alpha * (u1[i] - u0[i]) / smag + grad_u_f[i][0] * fCorr[0] + grad_u_f[0][i] * nVec[0] +
                       grad_u_f[i][1] * fCorr[1] + grad_u_f[1][i] * nVec[1] + grad_u_f[i][2] * fCorr[2] +
                       grad_u_f[2][i] * nVec[2]);
  tmp = (1.0 - nonL) * 2.0 / 3.0 * mut * (grad_u_f[0][0] + grad_u_f[1][1] + grad_u_f[2][2]);
  // -2/3 rho k (:1615; with SA kine_fa = 0 and the term is left out)
  if (c.sst) tmp += (1.0 - nonL) * 1.0 / 3.0 * (rho0 + rho1) * kine_fa;
  tauTurbij_nj[0] -= tmp * nVec[0];
  tauTurbij_nj[1] -= tmp * nVec[1];
  tauTurbij_nj[2] -= tmp * nVec[2];
  tauTurbij_nj[0] += nonL * (rij_d[0] * nVec[0] + rij_offd[0] * nVec[1] + rij_offd[1] * nVec[2]);
  tauTurbij_nj[1] += nonL * (rij_offd[0] * nVec[0] + rij_d[1] * nVec[1] + rij_offd[2] * nVec[2]);
  tauTurbij_nj[2] += nonL * (rij_offd[1] * nVec[0] + rij_offd[2] * nVec[1] + rij_d[2] * nVec[2]);
  for (int i = 0; i < 3; i++) tauij_nj[i] = tauLamij_nj[i] + tauTurbij_nj[i];
  for (int i = 0; i < 3; i++) Frhou[i] = -area * tauij_nj[i];
  double keff = lambdaOverCp + mut / c.p.pr_turb;
  enth = keff * (alpha * (h1 - h0) / smag + 0.5 * ((grad_h0[0] + grad_h1[0]) * fCorr[0] +
                                                   (grad_h0[1] + grad_h1[1]) * fCorr[1] +
                                                   (grad_h0[2] + grad_h1[2]) * fCorr[2]));
  double psi = 0.0;
  if (c.sst) psi = (mul + mut) * (alpha * (kine1 - kine0) / smag);  // triple correlation (:1657-1660)
  FrhoE = -area * (enth + tauij_nj[0] * u_fa[0] + tauij_nj[1] * u_fa[1] + tauij_nj[2] * u_fa[2] + psi);

  double dfdU[5][5] = {{0.0}}, dUdQ[5][5] = {{0.0}};
  auto mult = [&](double (*A)[5]) {  // dFdU_times_dUdQ, :717-745
    A[1][0] = dfdU[1][1] * dUdQ[1][0] + dfdU[1][2] * dUdQ[2][0] + dfdU[1][3] * dUdQ[3][0];
    A[2][0] = dfdU[2][1] * dUdQ[1][0] + dfdU[2][2] * dUdQ[2][0] + dfdU[2][3] * dUdQ[3][0];
    A[3][0] = dfdU[3][1] * dUdQ[1][0] + dfdU[3][2] * dUdQ[2][0] + dfdU[3][3] * dUdQ[3][0];
    A[4][0] = dfdU[4][1] * dUdQ[1][0] + dfdU[4][2] * dUdQ[2][0] + dfdU[4][3] * dUdQ[3][0] + dfdU[4][4] * dUdQ[4][0];
    A[1][1] = dfdU[1][1] * dUdQ[1][1];
    A[2][1] = dfdU[2][1] * dUdQ[1][1];
    A[3][1] = dfdU[3][1] * dUdQ[1][1];
    A[4][1] = dfdU[4][1] * dUdQ[1][1] + dfdU[4][4] * dUdQ[4][1];
    A[1][2] = dfdU[1][2] * dUdQ[2][2];
    A[2][2] = dfdU[2][2] * dUdQ[2][2];
    A[3][2] = dfdU[3][2] * dUdQ[2][2];
    A[4][2] = dfdU[4][2] * dUdQ[2][2] + dfdU[4][4] * dUdQ[4][2];
    A[1][3] = dfdU[1][3] * dUdQ[3][3];
    A[2][3] = dfdU[2][3] * dUdQ[3][3];
    A[3][3] = dfdU[3][3] * dUdQ[3][3];
    A[4][3] = dfdU[4][3] * dUdQ[3][3] + dfdU[4][4] * dUdQ[4][3];
    A[4][4] = dfdU[4][1] * dUdQ[1][4] + dfdU[4][2] * dUdQ[2][4] + dfdU[4][3] * dUdQ[3][4] + dfdU[4][4] * dUdQ[4][4];
  };
  auto side = [&](double sgn, double rho, const double *u, double h, double T, double R, double gam, double kine) {
    dfdU[1][1] = sgn * (mul + mut) * alpha / smag * area;
    dfdU[1][2] = 0.0;
    dfdU[1][3] = 0.0;
    dfdU[2][1] = 0.0;
    dfdU[2][2] = sgn * (mul + mut) * alpha / smag * area;
    dfdU[2][3] = 0.0;
    dfdU[3][1] = 0.0;
    dfdU[3][2] = 0.0;
    dfdU[3][3] = sgn * (mul + mut) * alpha / smag * area;
    dfdU[4][1] = area * tauij_nj[0] * 0.5 + u_fa[0] * dfdU[1][1];
    dfdU[4][2] = area * tauij_nj[1] * 0.5 + u_fa[1] * dfdU[2][2];
    dfdU[4][3] = area * tauij_nj[2] * 0.5 + u_fa[2] * dfdU[3][3];
    dfdU[4][4] = sgn * area * keff * alpha / smag;
    double invRho = 1.0 / rho;
    dUdQ[1][0] = -u[0] * invRho;
    dUdQ[2][0] = -u[1] * invRho;
    dUdQ[3][0] = -u[2] * invRho;
    dUdQ[4][0] = gam * invRho * (-h - kine + R * T + 0.5 * dot3(u, u));
    dUdQ[1][1] = invRho;
    dUdQ[4][1] = -gam * u[0] * invRho;
    dUdQ[2][2] = invRho;
    dUdQ[4][2] = -gam * u[1] * invRho;
    dUdQ[3][3] = invRho;
    dUdQ[4][3] = -gam * u[2] * invRho;
    dUdQ[4][4] = gam * invRho;
  };
  if (A0) {
    side(-1.0, rho0, u0, h0, T0, R0, gam0, kine0);
    mult(A0);
  }
  if (A1) {
    side(1.0, rho1, u1, h1, T1, R1, gam1, kine1);
    mult(A1);
  }
}

// calcViscousFluxNS, JoeWithModels.cpp:3591-3855 (kine from SST's first scalar)
void calc_viscous_flux(or_ctx &c, double *A) {
  const or_mesh &m = c.m;
  const double gam = c.p.gamma, R = c.p.R_gas;
  const bool ls = c.p.grad_scheme == OR_GRAD_LS;
  const bool lim = c.p.visc_limiter && c.p.limiter;
  // gradients with boundary values (:3616-3630)
  if (!c.p.visc_limiter || !c.p.second_order) {
    if (ls) grad_vector_ls(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
    else grad_vector_gg(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
    if (lim) limit_vector(c, c.grad_u.data(), c.vel.data(), c.sos.data());
  }
  if (ls) grad_scalar_ls(c, c.grad_h.data(), c.enthalpy.data(), c.h_bfa.data());
  else grad_scalar_gg(c, c.grad_h.data(), c.enthalpy.data(), c.h_bfa.data());
  if (lim) limit_scalar(c, c.grad_h.data(), c.enthalpy.data(), c.enthalpy.data(), NULL);

  double A0[5][5], A1[5][5];
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 5; j++) A0[i][j] = A1[i][j] = 0.0;
  double Frhou[3] = {0.0, 0.0, 0.0}, FrhoE = 0.0;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    int noc00 = 0, noc01 = 0, noc11 = 0, noc10 = 0;
    if (A) {
      noc00 = m.nbocv_i[icv0];
      noc01 = noc00;
      while (m.nbocv_v[++noc01] != icv1) {}
      noc11 = m.nbocv_i[icv1];
      noc10 = noc11;
      while (m.nbocv_v[++noc10] != icv0) {}
    }
    double nVec[3];
    double area = norm3(nVec, &m.fa_normal[3 * ifa]);
    double sVec[3];
    for (int i = 0; i < 3; i++) sVec[i] = m.x_cv[3 * icv1 + i] - m.x_cv[3 * icv0 + i];
    double smag = sqrt(dot3(sVec, sVec));
    for (int i = 0; i < 3; i++) sVec[i] /= smag;
    const double *v0 = &c.vel[3 * icv0], *v1 = &c.vel[3 * icv1];
    double uAux_fa[3] = {0.5 * (v0[0] + v1[0]), 0.5 * (v0[1] + v1[1]), 0.5 * (v0[2] + v1[2])};
    double kine0 = 0.0, kine1 = 0.0, kine_fa = 0.0;
    if (c.sst) {  // :3654-3669
      kine0 = c.phi[icv0];
      kine1 = c.phi[icv1];
      double w0, w1;
      gg_weights(m, ifa, icv0, icv1, w0, w1);
      kine_fa = (w1 * kine0 + w0 * kine1) / (w0 + w1);
    }
    add_visc_flux(c, Frhou, FrhoE, A ? A0 : NULL, A ? A1 : NULL, c.rho[icv0], v0, &c.grad_u[9 * icv0],
                  c.enthalpy[icv0], &c.grad_h[3 * icv0], c.temp[icv0], R, gam, kine0, c.rho[icv1], v1,
                  &c.grad_u[9 * icv1], c.enthalpy[icv1], &c.grad_h[3 * icv1], c.temp[icv1], R, gam, kine1,
                  c.mul_fa[ifa], c.mut_fa[ifa], c.lamOcp_fa[ifa], uAux_fa, area, nVec, smag, sVec, kine_fa);
    if (A) {
      for (int i = 1; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc00 + 5 * i + j] -= A0[i][j];
          A[25 * noc01 + 5 * i + j] -= A1[i][j];
        }
      for (int i = 1; i < 5; i++)
        for (int j = 0; j < 5; j++) {
          A[25 * noc11 + 5 * i + j] += A1[i][j];
          A[25 * noc10 + 5 * i + j] += A0[i][j];
        }
    }
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= Frhou[i];
    c.rhs_rhoE[icv0] -= FrhoE;
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv1 + i] += Frhou[i];
    c.rhs_rhoE[icv1] += FrhoE;
  }
  // boundary faces, zone by zone (:3714-3851)
  for (int iz = 0; iz < m.nzone; iz++) {
    const int bc = m.zone_bc[iz];
    const int f0 = m.zone_ifa[2 * iz], f1 = m.zone_ifa[2 * iz + 1];
    if (bc == OR_BC_SYMMETRY) {  // only the k term (:3725-3745)
      if (c.sst)
        for (int ifa = f0; ifa <= f1; ifa++) {
          int icv0 = m.cvofa[2 * ifa];
          double kine_fa = c.phi_bfa[ifa];
          double tmp = 1.0 / 3.0 * (c.rho[icv0] + c.rho_bfa[ifa]) * kine_fa;
          for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= tmp * m.fa_normal[3 * ifa + i];
        }
      continue;
    }
    for (int ifa = f0; ifa <= f1; ifa++) {
      int icv0 = m.cvofa[2 * ifa];
      double nVec[3];
      double area = norm3(nVec, &m.fa_normal[3 * ifa]);
      double sVec[3];
      for (int i = 0; i < 3; i++) sVec[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
      double smag;
      const double *sArg;
      if (bc == OR_BC_WALL) {
        smag = fabs(dot3(sVec, nVec));  // wall distance; the normal replaces s (:3777-3781)
        sArg = nVec;
      } else {
        smag = sqrt(dot3(sVec, sVec));
        for (int i = 0; i < 3; i++) sVec[i] /= smag;
        sArg = sVec;
      }
      const double *vb = &c.vel_bfa[3 * ifa];
      // walls: kine0 of the cell, kine1 = kine_fa = 0 (:3767-3776); other zones
      // kine1 = kine_fa = phi_bfa (:3817-3826)
      double kine0 = 0.0, kine1 = 0.0, kine_fa = 0.0;
      if (c.sst) {
        kine0 = c.phi[icv0];
        if (bc != OR_BC_WALL) kine1 = kine_fa = c.phi_bfa[ifa];
      }
      add_visc_flux(c, Frhou, FrhoE, A ? A0 : NULL, NULL, c.rho[icv0], &c.vel[3 * icv0], &c.grad_u[9 * icv0],
                    c.enthalpy[icv0], &c.grad_h[3 * icv0], c.temp[icv0], R, gam, kine0, c.rho_bfa[ifa], vb,
                    &c.grad_u[9 * icv0], c.h_bfa[ifa], &c.grad_h[3 * icv0], c.T_bfa[ifa], R, gam, kine1, c.mul_fa[ifa],
                    c.mut_fa[ifa], c.lamOcp_fa[ifa], vb, area, nVec, smag, sArg, kine_fa);
      if (A) {
        int noc00 = m.nbocv_i[icv0];
        for (int i = 1; i < 5; i++)
          for (int j = 0; j < 5; j++) A[25 * noc00 + 5 * i + j] -= A0[i][j];
      }
      for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv0 + i] -= Frhou[i];
      c.rhs_rhoE[icv0] -= FrhoE;
    }
  }
}

// ---------------------------------------------------------------------------
// Spalart-Allmaras as one segregated scalar (RansTurbSA, TurbModel_SA.h)
// ---------------------------------------------------------------------------
namespace sa {
// model constants, getDoubleParam defaults (TurbModel_SA.h:20-27)
const double cv1_3 = 357.911, cw1 = 3.239067817, cw2 = 0.3, cw3_6 = 64.0, sigma = 0.6666666, cb1 = 0.1355,
             cb2 = 0.622, KarmanConst_2 = 0.1681;
// ScalarTranspEq settings of nuSA (TurbModel_SA.h:31-37)
const double phiZero = 1.0e-9, phiZeroRel = 1.0e-2, lowerBound = 1.0e-10, upperBound = 1.0;
const int phiMaxiter = 1000;
}  // namespace sa

// calcWallDistance, UgpWithCvCompFlow.h:1395-1535 (one rank): nearest wall
// face centroid, then the normal distance when the cell centre lies within
// the face's "spectral radius"
void calc_wall_distance(or_ctx &c) {
  const or_mesh &m = c.m;
  std::vector<int> wf;
  for (int iz = 0; iz < m.nzone; iz++)
    if (m.zone_bc[iz] == OR_BC_WALL)
      for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) wf.push_back(ifa);
  for (int icv = 0; icv < c.ncv; icv++) {
    double minDist = 1.0e20;
    int save = 0;
    for (size_t k = 0; k < wf.size(); k++) {
      const double *xf = &m.x_fa[3 * wf[k]], *xc = &m.x_cv[3 * icv];
      double dist = sqrt(pow(xc[0] - xf[0], 2.0) + pow(xc[1] - xf[1], 2.0) + pow(xc[2] - xf[2], 2.0));
      if (dist < minDist) {
        save = (int)k;
        minDist = dist;
      }
    }
    const int f = wf[save];
    double n[3], s_half[3];
    double nmag = norm3(n, &m.fa_normal[3 * f]);
    for (int i = 0; i < 3; i++) s_half[i] = m.x_fa[3 * f + i] - m.x_cv[3 * icv + i];
    double sMag2 = dot3(s_half, s_half);
    double dNorm2 = dot3(s_half, n) * dot3(s_half, n);
    double dTang2 = sMag2 - dNorm2;
    double specRadius2 = 0.5 * nmag;
    c.wall_dist[icv] = dTang2 > specRadius2 ? minDist : fabs(dot3(s_half, n));
  }
}

// InterpolateAtCellCenterFromFaceValues, UgpWithCvCompFlow.h:1538-1551
double cell_avg(const or_ctx &c, const std::vector<double> &phi, int icv) {
  double phiC = 0.0;
  int foc_f = c.m.faocv_i[icv], foc_l = c.m.faocv_i[icv + 1] - 1;
  for (int foc = foc_f; foc <= foc_l; foc++) phiC += phi[c.m.faocv_v[foc]];
  phiC /= (double)(foc_l - foc_f + 1);
  return phiC;
}

// calcRansTurbViscMuet of RansTurbSA (TurbModel_SA.h:74-128) with calcGradVel,
// calcStrainRate, calcVorticity (UgpWithCvCompFlow.h:1305-1392)
void calc_turb_visc(or_ctx &c) {
  const or_mesh &m = c.m;
  if (c.p.grad_scheme == OR_GRAD_LS)
    grad_vector_ls(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
  else
    grad_vector_gg(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
  if (c.p.limiter) limit_vector(c, c.grad_u.data(), c.vel.data(), c.sos.data());
  for (int icv = 0; icv < c.ncv; icv++) {
    const double *g = &c.grad_u[9 * icv];
    double str = 0.0;
    double diverg = g[0] + g[4] + g[8];
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        if (i == j) str += pow(0.5 * (g[3 * i + j] + g[3 * j + i]) - 1. / 3. * diverg, 2.0);
        else str += pow(0.5 * (g[3 * i + j] + g[3 * j + i]), 2.0);
    c.str_mag[icv] = sqrt(2.0 * str);
    double vort = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) vort += pow(0.5 * (g[3 * i + j] - g[3 * j + i]), 2.0);
    c.vort_mag[icv] = sqrt(2.0 * vort);
  }
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double rho_fa = (w1 * c.rho[icv0] + w0 * c.rho[icv1]) / (w0 + w1);
    double nuSA_fa = (w1 * c.phi[icv0] + w0 * c.phi[icv1]) / (w0 + w1);
    double muSA_fa = rho_fa * nuSA_fa;
    double chi_3 = pow(muSA_fa / c.mul_fa[ifa], 3.0);
    double fv1 = chi_3 / (chi_3 + sa::cv1_3);
    c.mut_fa[ifa] = max(muSA_fa * fv1, 0.0);
  }
  for (int iz = 0; iz < m.nzone; iz++) {
    const int f0 = m.zone_ifa[2 * iz], f1 = m.zone_ifa[2 * iz + 1];
    if (m.zone_bc[iz] == OR_BC_WALL) {
      for (int ifa = f0; ifa <= f1; ifa++) c.mut_fa[ifa] = 0.0;
    } else {
      for (int ifa = f0; ifa <= f1; ifa++) {
        int icv0 = m.cvofa[2 * ifa];
        double muSA = c.rho[icv0] * c.phi[icv0];
        double mul_icv0 = cell_avg(c, c.mul_fa, icv0);
        double chi_3 = pow(muSA / mul_icv0, 3.0);
        double fv1 = chi_3 / (chi_3 + sa::cv1_3);
        c.mut_fa[ifa] = max(muSA * fv1, 0.0);
      }
    }
  }
}

namespace sst {
// model constants, getDoubleParam defaults (TurbModel_KOMSST.h:16-23)
const double sigma_k1 = 0.85, sigma_k2 = 1.0, sigma_om1 = 0.5, sigma_om2 = 0.856, beta_1 = 0.075, beta_2 = 0.0828,
             betaStar = 0.09, a1 = 0.31;
// ScalarTranspEq settings of kine and omega (:36-52)
const double phiZero = 1.0e-8, phiZeroRel = 1.0e-2;
const int phiMaxiter = 1000;
const double lowerBound[2] = {1.0e-10, 1.0e-4}, upperBound[2] = {1.0e10, 1.0e15};
}  // namespace sst

// the limited gradient of scalar s with its boundary values (calcCvScalarGrad
// as calcEulerFlux, calcViscousFluxScalar_new and calcMenterBlendingFunctions
// call it: gradreconstruction, limiterNavierS, the scalar as its own reference)
static void scalar_grad(or_ctx &c, int s) {
  double *g = &c.grad_phi[3 * (size_t)c.ncv * s], *ph = &c.phi[(size_t)c.ncv * s];
  const double *pb = &c.phi_bfa[(size_t)c.nfa_b * s];
  if (c.p.grad_scheme == OR_GRAD_LS)
    grad_scalar_ls(c, g, ph, pb);
  else
    grad_scalar_gg(c, g, ph, pb);
  if (c.p.limiter) limit_scalar(c, g, ph, ph, NULL);
}

// setScalarBC, Scalars.cpp:289-360, per scalar: a zone with "<zone>.<name> = v"
// takes v on inflow faces and the cell value on outflow faces ("= PROFILES":
// the face's value interpolated from the profile file); "= HOOK" calls
// boundaryHookScalarRansTurb, which for SST's omega on WALL zones sets
// 60 mu / (rho beta_1 d^2) (TurbModel_KOMSST.h:342-358) and otherwise leaves
// the face values alone; any other zone takes the cell value
void set_scalar_bc(or_ctx &c) {
  const or_mesh &m = c.m;
  for (int s = 0; s < c.nsc; s++) {
    const double *zs = s == 0 ? m.zone_scal : m.zone_scal2;
    double *pb = &c.phi_bfa[(size_t)c.nfa_b * s];
    const double *ph = &c.phi[(size_t)c.ncv * s];
    for (int iz = 0; iz < m.nzone; iz++) {
      const double kind = zs[2 * iz];
      const double val = zs[2 * iz + 1];
      for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) {
        int icv0 = m.cvofa[2 * ifa];
        if (kind == 2.0) {
          if (c.sst && s == 1 && m.zone_bc[iz] == OR_BC_WALL) {
            double muLamCV = mu_lam(c, c.temp[icv0]);
            pb[ifa] = 60.0 * muLamCV / (c.rho[icv0] * sst::beta_1 * c.wall_dist[icv0] * c.wall_dist[icv0]);
          }
        } else if (kind != 0.0 && !(dot3(&c.vel_bfa[3 * ifa], &m.fa_normal[3 * ifa]) > 1.0e-8)) {
          pb[ifa] = kind == 3.0 ? c.phi_bpr[(size_t)c.nfa_b * s + ifa] : val;  // PROFILES: the face's profile value
        } else {
          pb[ifa] = ph[icv0];
        }
      }
    }
  }
}

// calcRansTurbViscMuet of RansTurbKOmSST (TurbModel_KOMSST.h:100-158): calcGradVel,
// calcStrainRateAndDivergence, calcVorticity (UgpWithCvCompFlow.h:1300-1392),
// calcMenterBlendingFunctions (TurbModel_KOMSST.h:360-392), then mut_fa
void calc_turb_visc_sst(or_ctx &c) {
  const or_mesh &m = c.m;
  const size_t n = c.ncv;
  if (c.p.grad_scheme == OR_GRAD_LS)
    grad_vector_ls(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
  else
    grad_vector_gg(c, c.grad_u.data(), c.vel.data(), c.vel_bfa.data());
  if (c.p.limiter) limit_vector(c, c.grad_u.data(), c.vel.data(), c.sos.data());
  for (int icv = 0; icv < c.ncv; icv++) {
    const double *g = &c.grad_u[9 * icv];
    c.diverg[icv] = g[0] + g[4] + g[8];
    double str = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++)
        if (i == j) str += pow(0.5 * (g[3 * i + j] + g[3 * j + i]) - 1. / 3. * c.diverg[icv], 2.0);
        else str += pow(0.5 * (g[3 * i + j] + g[3 * j + i]), 2.0);
    c.str_mag[icv] = sqrt(2.0 * str);
    double vort = 0.0;
    for (int i = 0; i < 3; i++)
      for (int j = 0; j < 3; j++) vort += pow(0.5 * (g[3 * i + j] - g[3 * j + i]), 2.0);
    c.vort_mag[icv] = sqrt(2.0 * vort);
  }
  scalar_grad(c, 0);
  scalar_grad(c, 1);
  const double *kine = &c.phi[0], *omega = &c.phi[n];
  for (int icv = 0; icv < c.ncv; icv++) {
    const double d = c.wall_dist[icv], rho = c.rho[icv];
    const double mue = cell_avg(c, c.mul_fa, icv);
    c.cross_diff[icv] = max(2.0 * rho * sst::sigma_om2 / omega[icv] * dot3(&c.grad_phi[3 * icv], &c.grad_phi[3 * (n + icv)]),
                            1.0e-20);
    double gamma1 = 500.0 * mue / (pow(d, 2.0) * rho * omega[icv]);
    double gamma2 = 4.0 * sst::sigma_om2 * rho * kine[icv] / (d * d * c.cross_diff[icv]);
    double gamma3 = sqrt(kine[icv]) / (sst::betaStar * omega[icv] * d);
    double gamma = min(max(gamma1, gamma3), gamma2);
    c.f1[icv] = tanh(pow(gamma, 4.0));
    gamma = max(2.0 * gamma3, gamma1);
    c.f2[icv] = tanh(pow(gamma, 2.0));
  }
  const double *lim = c.p.sst_limiter_shear == 1 ? c.str_mag.data() : c.vort_mag.data();
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double rho_fa = (w1 * c.rho[icv0] + w0 * c.rho[icv1]) / (w0 + w1);
    double kine_fa = (w1 * kine[icv0] + w0 * kine[icv1]) / (w0 + w1);
    double om_fa = (w1 * omega[icv0] + w0 * omega[icv1]) / (w0 + w1);
    double limitFunc_fa = (w1 * lim[icv0] + w0 * lim[icv1]) / (w0 + w1);
    double blendFuncF2_fa = (w1 * c.f2[icv0] + w0 * c.f2[icv1]) / (w0 + w1);
    double zeta = min(1.0 / om_fa, sst::a1 / (limitFunc_fa * blendFuncF2_fa));
    c.mut_fa[ifa] = min(max(rho_fa * kine_fa * zeta, 0.0), 1.0);
  }
  for (int iz = 0; iz < m.nzone; iz++) {
    const int f0 = m.zone_ifa[2 * iz], f1 = m.zone_ifa[2 * iz + 1];
    for (int ifa = f0; ifa <= f1; ifa++) {
      if (m.zone_bc[iz] == OR_BC_WALL) {
        c.mut_fa[ifa] = 0.0;
        continue;
      }
      int icv0 = m.cvofa[2 * ifa];
      double zeta = min(1.0 / omega[icv0], sst::a1 / (lim[icv0] * c.f2[icv0]));
      c.mut_fa[ifa] = min(max(c.rho[icv0] * kine[icv0] * zeta, 0.0), 1.0);
    }
  }
}

// diffusivityHookScalarRansTurb of RansTurbKOmSST (TurbModel_KOMSST.h:160-254)
void sst_diffusivity(or_ctx &c, int s) {
  const or_mesh &m = c.m;
  const double sg1 = s == 0 ? sst::sigma_k1 : sst::sigma_om1, sg2 = s == 0 ? sst::sigma_k2 : sst::sigma_om2;
  double *diff = &c.diff[(size_t)c.nfa * s];
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double f1_fa = (w1 * c.f1[icv0] + w0 * c.f1[icv1]) / (w0 + w1);
    double coeff = sg1 * f1_fa + (1. - f1_fa) * sg2;
    diff[ifa] = c.mul_fa[ifa] + coeff * c.mut_fa[ifa];
  }
  for (int iz = 0; iz < m.nzone; iz++)
    for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) {
      if (m.zone_bc[iz] == OR_BC_WALL) {
        diff[ifa] = c.mul_fa[ifa];
      } else {
        int icv0 = m.cvofa[2 * ifa];
        double coeff = sg1 * c.f1[icv0] + (1. - c.f1[icv0]) * sg2;
        diff[ifa] = c.mul_fa[ifa] + coeff * c.mut_fa[ifa];
      }
    }
}

// sourceHookScalarRansTurb_new of RansTurbKOmSST (TurbModel_KOMSST.h:256-304)
void sst_source(or_ctx &c, int s, bool implicit) {
  const or_mesh &m = c.m;
  const size_t n = c.ncv, nnz = c.nnz;
  const double *kine = &c.phi[0], *omega = &c.phi[n];
  const double *lim = c.p.sst_limiter_shear == 1 ? c.str_mag.data() : c.vort_mag.data();
  double *rhs = &c.rhs_scal[n * s];
  double *A = &c.AScal[(6 * (size_t)s + 5) * nnz];
  const double alfa_1 = sst::beta_1 / sst::betaStar - sst::sigma_om1 * pow(0.41, 2.0) / sqrt(sst::betaStar);
  const double alfa_2 = sst::beta_2 / sst::betaStar - sst::sigma_om2 * pow(0.41, 2.0) / sqrt(sst::betaStar);
  for (int icv = 0; icv < c.ncv; icv++) {
    const double rho = c.rho[icv], vol = m.cv_volume[icv];
    if (s == 0) {
      double zeta = min(1.0 / omega[icv], sst::a1 / (lim[icv] * c.f2[icv]));
      double mut = min(max(rho * kine[icv] * zeta, 0.0), 1.0e5);
      double Pk = mut * c.str_mag[icv] * c.str_mag[icv] - 2. / 3. * rho * kine[icv] * c.diverg[icv];
      if (c.p.sst_limit_pk == 1) Pk = min(Pk, 20.0 * sst::betaStar * rho * kine[icv] * omega[icv]);
      Pk = max(Pk, 0.0);
      double src = Pk - sst::betaStar * rho * omega[icv] * kine[icv];
      rhs[icv] += src * vol;
      if (implicit) {
        double dsrcdphi = -sst::betaStar * omega[icv];
        A[m.nbocv_i[icv]] -= dsrcdphi * vol;
      }
    } else {
      double F1 = c.f1[icv];
      double alfa = F1 * alfa_1 + (1.0 - F1) * alfa_2;
      double beta = F1 * sst::beta_1 + (1.0 - F1) * sst::beta_2;
      double zeta = max(omega[icv], lim[icv] * c.f2[icv] / sst::a1);
      double Pk = c.str_mag[icv] * c.str_mag[icv] - 2. / 3. * zeta * c.diverg[icv];
      Pk = max(Pk, 0.0);
      double src = alfa * rho * Pk + (1.0 - F1) * c.cross_diff[icv] - beta * rho * omega[icv] * omega[icv];
      rhs[icv] += src * vol;
      if (implicit) {
        double dsrcdphi = -2.0 * beta * omega[icv];
        A[m.nbocv_i[icv]] -= dsrcdphi * vol;
      }
    }
  }
}

// diffusivityHookScalarRansTurb of RansTurbSA (TurbModel_SA.h:130-172)
void sa_diffusivity(or_ctx &c) {
  const or_mesh &m = c.m;
  const double invSigma = 1.0 / sa::sigma;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double w0, w1;
    gg_weights(m, ifa, icv0, icv1, w0, w1);
    double rho_fa = (w1 * c.rho[icv0] + w0 * c.rho[icv1]) / (w0 + w1);
    double nuSA_fa = (w1 * c.phi[icv0] + w0 * c.phi[icv1]) / (w0 + w1);
    c.diff[ifa] = invSigma * (c.mul_fa[ifa] + rho_fa * nuSA_fa);
  }
  for (int iz = 0; iz < m.nzone; iz++)
    for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++)
      c.diff[ifa] = m.zone_bc[iz] == OR_BC_WALL ? invSigma * c.mul_fa[ifa]
                                                : invSigma * (c.mul_fa[ifa] + c.rho_bfa[ifa] * c.phi_bfa[ifa]);
}

// calcViscousFluxScalar_new, Scalars.cpp:117-247, for scalar s with the model's
// diffusivity (diffusivityHookScalarRansTurb, TurbModel_SA.h:130-172,
// TurbModel_KOMSST.h:160-254)
void calc_viscous_flux_scalar(or_ctx &c, bool implicit, int s = 0) {
  const or_mesh &m = c.m;
  const size_t nnz = c.nnz, n = c.ncv;
  double *A5 = &c.AScal[(6 * (size_t)s + 5) * nnz];
  scalar_grad(c, s);
  if (c.sst) sst_diffusivity(c, s);
  else sa_diffusivity(c);
  const double *phi = &c.phi[n * s], *phi_bfa = &c.phi_bfa[(size_t)c.nfa_b * s], *grad_phi = &c.grad_phi[3 * n * s];
  const double *diff = &c.diff[(size_t)c.nfa * s];
  double *rhs_scal = &c.rhs_scal[n * s];
  const double *zs = s == 0 ? m.zone_scal : m.zone_scal2;
  for (int ifa = c.nfa_b; ifa < c.nfa; ifa++) {
    int icv0 = m.cvofa[2 * ifa], icv1 = m.cvofa[2 * ifa + 1];
    double n[3], sv[3];
    double nmag = norm3(n, &m.fa_normal[3 * ifa]);
    for (int i = 0; i < 3; i++) sv[i] = m.x_cv[3 * icv1 + i] - m.x_cv[3 * icv0 + i];
    double smag = sqrt(dot3(sv, sv));
    for (int i = 0; i < 3; i++) sv[i] /= smag;
    double alpha = dot3(n, sv);
    const double *g0 = &grad_phi[3 * icv0], *g1 = &grad_phi[3 * icv1];
    double viscFlux = diff[ifa] * nmag *
                      (alpha * (phi[icv1] - phi[icv0]) / smag + 0.5 * (g0[0] + g1[0]) * (n[0] - alpha * sv[0]) +
                       0.5 * (g0[1] + g1[1]) * (n[1] - alpha * sv[1]) + 0.5 * (g0[2] + g1[2]) * (n[2] - alpha * sv[2]));
    rhs_scal[icv0] += viscFlux;
    rhs_scal[icv1] -= viscFlux;
    if (implicit) {
      int noc00 = m.nbocv_i[icv0], noc01 = noc00;
      while (m.nbocv_v[++noc01] != icv1) {}
      int noc11 = m.nbocv_i[icv1], noc10 = noc11;
      while (m.nbocv_v[++noc10] != icv0) {}
      A5[noc00] += diff[ifa] * nmag * alpha / smag / c.rho[icv0];
      A5[noc01] += -diff[ifa] * nmag * alpha / smag / c.rho[icv1];
      A5[noc11] += diff[ifa] * nmag * alpha / smag / c.rho[icv1];
      A5[noc10] += -diff[ifa] * nmag * alpha / smag / c.rho[icv0];
    }
  }
  // boundary zones with a value or HOOK (:168-219); others carry no diffusive flux
  for (int iz = 0; iz < m.nzone; iz++) {
    if (zs[2 * iz] == 0.0) continue;
    for (int ifa = m.zone_ifa[2 * iz]; ifa <= m.zone_ifa[2 * iz + 1]; ifa++) {
      int icv0 = m.cvofa[2 * ifa];
      double n[3], s_half[3];
      double nmag = norm3(n, &m.fa_normal[3 * ifa]);
      for (int i = 0; i < 3; i++) s_half[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv0 + i];
      double smag_half = fabs(dot3(s_half, n));
      double viscFlux = diff[ifa] * nmag * (phi_bfa[ifa] - phi[icv0]) / smag_half;
      rhs_scal[icv0] += viscFlux;
      if (implicit) A5[m.nbocv_i[icv0]] += diff[ifa] * nmag / smag_half / c.rho[icv0];
    }
  }
}

// sourceHookScalarRansTurb_new, TurbModel_SA.h:282-357
void sa_source(or_ctx &c, bool implicit) {
  const or_mesh &m = c.m;
  double *A5 = &c.AScal[5 * (size_t)c.nnz];
  for (int icv = 0; icv < c.ncv; icv++) {
    const double nuSA = c.phi[icv], rho = c.rho[icv], vol = m.cv_volume[icv];
    double mue = cell_avg(c, c.mul_fa, icv);
    double Om = max(c.vort_mag[icv], 1.e-8);
    double Strainrate = max(c.str_mag[icv], 1.e-8);
    double d = c.wall_dist[icv];
    double chi = nuSA * rho / mue;
    double chi_3 = pow(chi, 3.0);
    double inv_KarmanConst2_d2 = 1.0 / (sa::KarmanConst_2 * d * d);
    double fv1 = chi_3 / (chi_3 + sa::cv1_3);
    double fv2 = 1.0 - chi / (1.0 + chi * fv1);
    double BlendTurbulentProduction = 0.0;
    double S = max(Om + min(0.0, Strainrate - Om) * BlendTurbulentProduction, 0.0);
    double Shat = S + nuSA * fv2 * inv_KarmanConst2_d2;
    double inv_Shat = 1.0 / max(Shat, 1.0e-8);
    double r = min(nuSA * inv_Shat * inv_KarmanConst2_d2, 10.0);
    double g = r + sa::cw2 * (pow(r, 6.0) - r);
    double g_6 = pow(g, 6.0);
    double fw = g * pow((1.0 + sa::cw3_6) / (g_6 + sa::cw3_6), 1.0 / 6.0);
    double H1 = sa::cb1 * Shat * nuSA;
    double H2 = -sa::cw1 * fw * nuSA * nuSA / pow(d, 2.0);
    double H3 = sa::cb2 / sa::sigma * dot3(&c.grad_phi[3 * icv], &c.grad_phi[3 * icv]);
    c.rhs_scal[icv] += rho * vol * (H1 + H2 + H3);
    if (implicit) {
      double d_H1 = 0.0;
      double d_H2 = -2.0 * sa::cw1 * fw * nuSA / pow(d, 2.0);
      A5[m.nbocv_i[icv]] -= (d_H1 + d_H2) * vol;
    }
  }
}

// solveCvScalarBcgstab, UgpWithCv.cpp:2298-2479 (Jacobi preconditioner; one rank)
int bcgstab_scalar(const or_ctx &c, double *phi, const double *Ap, const double *rhs, double zero, double zeroRel,
                   int maxiter) {
  const or_mesh &m = c.m;
  const int ncv = c.ncv;
  std::vector<double> res(ncv), res0(ncv), p(ncv, 0.0), v(ncv, 0.0), t(ncv), s(ncv), phat(ncv, 0.0), shat(ncv, 0.0);
  auto matvec = [&](double *out, const double *x) {
    for (int icv = 0; icv < ncv; icv++) {
      int noc_f = m.nbocv_i[icv], noc_l = m.nbocv_i[icv + 1] - 1;
      out[icv] = Ap[noc_f] * x[icv];
      for (int noc = noc_f + 1; noc <= noc_l; noc++) out[icv] += Ap[noc] * x[m.nbocv_v[noc]];
    }
  };
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  matvec(res.data(), phi);
  for (int icv = 0; icv < ncv; icv++) {
    res[icv] = rhs[icv] - res[icv];
    res0[icv] = res[icv];
  }
  double res0_max = 0.0;
  for (int icv = 0; icv < ncv; icv++) res0_max = max(res0_max, fabs(res[icv] / Ap[m.nbocv_i[icv]]));
  int iter = 0, done = 0;
  if (res0_max <= zero) done = 1;
  while (done == 0) {
    iter++;
    double rho_prime = -omega * rho;
    rho = 0.0;
    for (int icv = 0; icv < ncv; icv++) rho += res[icv] * res0[icv];
    if (fabs(rho_prime) < 1.0E-20) rho_prime = 1.0E-20;
    double beta = alpha * rho / rho_prime;
    for (int icv = 0; icv < ncv; icv++) p[icv] = res[icv] - beta * (p[icv] - omega * v[icv]);
    for (int icv = 0; icv < ncv; icv++) phat[icv] = p[icv] / Ap[m.nbocv_i[icv]];
    matvec(v.data(), phat.data());
    double gamma = 0.0;
    for (int icv = 0; icv < ncv; icv++) gamma += v[icv] * res0[icv];
    if (fabs(gamma) < 1.0E-20) gamma = 1.0E-20;
    alpha = rho / gamma;
    for (int icv = 0; icv < ncv; icv++) s[icv] = res[icv] - alpha * v[icv];
    for (int icv = 0; icv < ncv; icv++) shat[icv] = s[icv] / Ap[m.nbocv_i[icv]];
    matvec(t.data(), shat.data());
    double b0 = 0.0, b1 = 0.0;
    for (int icv = 0; icv < ncv; icv++) {
      b0 += s[icv] * t[icv];
      b1 += t[icv] * t[icv];
    }
    omega = b0 / (b1 + 1.0E-20);
    for (int icv = 0; icv < ncv; icv++) phi[icv] += alpha * phat[icv] + omega * shat[icv];
    matvec(res.data(), phi);
    double res_max = 0.0;
    for (int icv = 0; icv < ncv; icv++) {
      res[icv] = rhs[icv] - res[icv];
      res_max = max(res_max, fabs(res[icv] / Ap[m.nbocv_i[icv]]));
    }
    if ((res_max <= zero) || (res_max / (res0_max + 1.0E-12) <= zeroRel) || (iter > maxiter)) done = 1;
  }
  return iter;
}

// The scalar half of a backward-Euler step (JoeWithModels.cpp:1040-1091),
// after the mean flow took dq: relaxation, vol/dt, the coupling columns
// AScal[0..4] moved to the right with dq, Jacobi BiCGStab, update and clip.
void scalar_solve(or_ctx &c, bool update) {
  const or_mesh &m = c.m;
  const size_t nnz = c.nnz, n = c.ncv;
  std::fill(c.dscal.begin(), c.dscal.end(), 0.0);
  for (int s = 0; s < c.nsc; s++) {  // in scalarTranspEqVector order
    double *rhs_scal = &c.rhs_scal[n * s], *AS = &c.AScal[6 * nnz * s], *dscal = &c.dscal[n * s], *phi = &c.phi[n * s];
    for (int icv = 0; icv < c.ncv; icv++) {
      rhs_scal[icv] *= c.p.relax_scal;
      int noc_f = m.nbocv_i[icv], noc_l = m.nbocv_i[icv + 1] - 1;
      double tmp = m.cv_volume[icv] / (c.local_dt[icv]);
      AS[5 * nnz + noc_f] += tmp;
      for (int noc = noc_f; noc <= noc_l; noc++) {
        const double *dqn = &c.dq[5 * m.nbocv_v[noc]];
        rhs_scal[icv] = rhs_scal[icv] - AS[0 * nnz + noc] * dqn[0] - AS[1 * nnz + noc] * dqn[1] -
                        AS[2 * nnz + noc] * dqn[2] - AS[3 * nnz + noc] * dqn[3] - AS[4 * nnz + noc] * dqn[4];
      }
    }
    const double zero = c.sst ? sst::phiZero : sa::phiZero, zeroRel = c.sst ? sst::phiZeroRel : sa::phiZeroRel;
    const double lo = c.sst ? sst::lowerBound[s] : sa::lowerBound, hi = c.sst ? sst::upperBound[s] : sa::upperBound;
    bcgstab_scalar(c, dscal, &AS[5 * nnz], rhs_scal, zero, zeroRel, c.sst ? sst::phiMaxiter : sa::phiMaxiter);
    if (update)
      for (int icv = 0; icv < c.ncv; icv++)
        phi[icv] = min(max((phi[icv] * (c.rho[icv] - c.dq[5 * icv]) + dscal[icv]) / c.rho[icv], lo), hi);
  }
}

void calc_rhs(or_ctx &c, double *A) {
  for (int icv = 0; icv < c.ncv; icv++) {
    c.rhs_rho[icv] = 0.0;
    for (int i = 0; i < 3; i++) c.rhs_rhou[3 * icv + i] = 0.0;
    c.rhs_rhoE[icv] = 0.0;
  }
  if (c.nsc) std::fill(c.rhs_scal.begin(), c.rhs_scal.end(), 0.0);
  if (c.p.coupled) {  // calcRhsCoupled, JoeWithModels.cpp:4596-4611
    calc_flux_coupled(c, A);
  } else {
    calc_euler_flux(c, A);
    if (c.p.mu_ref > 0.0) calc_viscous_flux(c, A);
  }
  // sourceHook / sourceHookCoupled: gravity (JoeWithModels.h:111-152)
  const double *g = c.p.gravity;
  for (int icv = 0; icv < c.ncv; icv++) {
    double vol = c.m.cv_volume[icv];
    c.rhs_rhou[3 * icv + 0] += vol * c.rho[icv] * g[0];
    c.rhs_rhou[3 * icv + 1] += vol * c.rho[icv] * g[1];
    c.rhs_rhou[3 * icv + 2] += vol * c.rho[icv] * g[2];
    c.rhs_rhoE[icv] += vol * (g[0] * c.rhou[3 * icv] + g[1] * c.rhou[3 * icv + 1] + g[2] * c.rhou[3 * icv + 2]);
    if (A) {
      double *D = &A[25 * c.m.nbocv_i[icv]];
      D[5 * 1 + 0] -= vol * g[0];
      D[5 * 2 + 0] -= vol * g[1];
      D[5 * 3 + 0] -= vol * g[2];
      D[5 * 4 + 1] -= vol * g[0];
      D[5 * 4 + 2] -= vol * g[1];
      D[5 * 4 + 3] -= vol * g[2];
    }
  }
  // scalars: diffusion, then the model source (JoeWithModels.cpp:2595-2618)
  // (the coupled loop's scalar diffusion is part of calcFluxCoupled; its
  // source is sourceHookRansTurbCoupled, TurbModel_SA.h:359-434, the same
  // terms into the block's scalar row)
  if (c.sa) {
    if (c.p.mu_ref > 0.0 && !c.p.coupled) calc_viscous_flux_scalar(c, A != NULL);
    sa_source(c, A != NULL);
  }
  if (c.sst)
    for (int s = 0; s < 2; s++) {
      if (c.p.mu_ref > 0.0) calc_viscous_flux_scalar(c, A != NULL, s);
      sst_source(c, s, A != NULL);
    }
}

void mat_vec(const or_ctx &c, double *res, const double *A, const double *phi) {  // UgpWithCv.cpp:2580-2596
  const or_mesh &m = c.m;
  for (int icv = 0; icv < c.ncv; icv++) {
    for (int i = 0; i < 5; i++) res[5 * icv + i] = 0.0;
    for (int noc = m.nbocv_i[icv]; noc <= m.nbocv_i[icv + 1] - 1; noc++) {
      int nb = m.nbocv_v[noc];
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) res[5 * icv + i] += A[25 * noc + 5 * i + j] * phi[5 * nb + j];
    }
  }
}

// solveCvVectorR5Bcgstab, UgpWithCv.cpp:2660-2876 (one rank: halo updates are no-ops)
int bcgstab(const or_ctx &c, double *phi, const double *A, const double *rhs, double zeroAbs, double zeroRel,
            int maxiter) {
  const int ncv = c.ncv, n5 = 5 * ncv;
  std::vector<double> res(n5), res0(n5), p(n5, 0.0), v(n5, 0.0), t(n5), s(n5), phat(n5, 0.0), shat(n5, 0.0);
  std::vector<double> LU(25 * (size_t)ncv);
  for (int icv = 0; icv < ncv; icv++) {
    std::memcpy(&LU[25 * icv], &A[25 * c.m.nbocv_i[icv]], 25 * sizeof(double));
    ludeco((double(*)[5]) & LU[25 * icv], 5);
  }
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  mat_vec(c, res.data(), A, phi);
  for (int k = 0; k < n5; k++) {
    res[k] = (rhs[k] - res[k]);
    res0[k] = res[k];
  }
  double res0_max = 0.0;
  for (int icv = 0; icv < ncv; icv++) {
    double tmp[5];
    for (int i = 0; i < 5; i++) tmp[i] = res[5 * icv + i];
    lusolv((double(*)[5]) & LU[25 * icv], tmp, 5);
    for (int i = 0; i < 5; i++) res0_max = max(res0_max, fabs(tmp[i]));
  }
  int iter = 0, done = 0;
  while (done == 0) {
    iter++;
    double rho_prime = -omega * rho;
    rho = 0.0;
    for (int k = 0; k < n5; k++) rho += res[k] * res0[k];
    if (fabs(rho_prime) < 1.0E-20) rho_prime = 1.0E-20;
    double beta = alpha * rho / rho_prime;
    for (int k = 0; k < n5; k++) p[k] = res[k] - beta * (p[k] - omega * v[k]);
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < 5; i++) phat[5 * icv + i] = p[5 * icv + i];
      lusolv((double(*)[5]) & LU[25 * icv], &phat[5 * icv], 5);
    }
    mat_vec(c, v.data(), A, phat.data());
    double gamma = 0.0;
    for (int k = 0; k < n5; k++) gamma += v[k] * res0[k];
    if (fabs(gamma) < 1.0E-20) gamma = 1.0E-20;
    alpha = rho / gamma;
    for (int k = 0; k < n5; k++) s[k] = res[k] - alpha * v[k];
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < 5; i++) shat[5 * icv + i] = s[5 * icv + i];
      lusolv((double(*)[5]) & LU[25 * icv], &shat[5 * icv], 5);
    }
    mat_vec(c, t.data(), A, shat.data());
    double b0 = 0.0, b1 = 0.0;
    for (int k = 0; k < n5; k++) {
      b0 += s[k] * t[k];
      b1 += t[k] * t[k];
    }
    omega = b0 / (b1 + 1.0E-20);
    for (int k = 0; k < n5; k++) phi[k] += alpha * phat[k] + omega * shat[k];
    mat_vec(c, res.data(), A, phi);
    for (int k = 0; k < n5; k++) res[k] = rhs[k] - res[k];
    double res_max = 0.0;
    for (int icv = 0; icv < ncv; icv++) {
      double tmp[5];
      for (int i = 0; i < 5; i++) tmp[i] = res[5 * icv + i];
      lusolv((double(*)[5]) & LU[25 * icv], tmp, 5);
      for (int i = 0; i < 5; i++) res_max = max(res_max, fabs(tmp[i]));
    }
    if ((res_max <= zeroAbs) || (res_max / (res0_max + 1.0E-12) <= zeroRel))
      done = 1;
    else if (iter > maxiter)
      done = 2;
  }
  return iter;
}

// ludeco / lusolv (UgpWithCv.cpp:2489-2540) on a flat b x b block
void ludeco_b(double *A, int b) {
  for (int jc = 1; jc < b; jc++) A[jc] /= A[0];
  int jrjc = 0;
  for (;;) {
    jrjc++;
    int jrjcm1 = jrjc - 1;
    int jrjcp1 = jrjc + 1;
    for (int jr = jrjc; jr < b; jr++) {
      double sum = A[b * jr + jrjc];
      for (int jm = 0; jm <= jrjcm1; jm++) sum -= A[b * jr + jm] * A[b * jm + jrjc];
      A[b * jr + jrjc] = sum;
    }
    if (jrjc == (b - 1)) return;
    for (int jc = jrjcp1; jc < b; jc++) {
      double sum = A[b * jrjc + jc];
      for (int jm = 0; jm <= jrjcm1; jm++) sum -= A[b * jrjc + jm] * A[b * jm + jc];
      A[b * jrjc + jc] = sum / A[b * jrjc + jrjc];
    }
  }
}

void lusolv_b(const double *A, double *c, int b) {
  c[0] = c[0] / A[0];
  for (int jr = 1; jr < b; jr++) {
    int jrm1 = jr - 1;
    double sum = c[jr];
    for (int jm = 0; jm <= jrm1; jm++) sum -= A[b * jr + jm] * c[jm];
    c[jr] = sum / A[b * jr + jr];
  }
  for (int jrjr = 1; jrjr < b; jrjr++) {
    int jr = (b - 1) - jrjr;
    int jrp1 = jr + 1;
    double sum = c[jr];
    for (int jmjm = jrp1; jmjm < b; jmjm++) {
      int jm = (b - 1) - jmjm + jrp1;
      sum -= A[b * jr + jm] * c[jm];
    }
    c[jr] = sum;
  }
}

// The coupled system's block-Jacobi BiCGStab on b x b blocks (b = 5 + nScal):
// the recurrences and stop rule of solveCvVectorR5Bcgstab (UgpWithCv.cpp
// :2660-2876) on the wider blocks. The reference solves this system with
// PETSc GMRES (PetscSolver.h:278-310), which is not part of this path; the
// converged state is set by the residual, not by the Krylov method.
int bcgstab_b(const or_ctx &c, double *phi, const double *A, const double *rhs, double zeroAbs, double zeroRel,
              int maxiter, int b) {
  const or_mesh &m = c.m;
  const int ncv = c.ncv, n = b * ncv, bb = b * b;
  std::vector<double> res(n), res0(n), p(n, 0.0), v(n, 0.0), t(n), s(n), phat(n, 0.0), shat(n, 0.0);
  std::vector<double> LU((size_t)bb * ncv), tmp(b);
  for (int icv = 0; icv < ncv; icv++) {
    std::memcpy(&LU[(size_t)bb * icv], &A[(size_t)bb * m.nbocv_i[icv]], bb * sizeof(double));
    ludeco_b(&LU[(size_t)bb * icv], b);
  }
  auto mat_vec_b = [&](double *out, const double *x) {
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < b; i++) out[b * icv + i] = 0.0;
      for (int noc = m.nbocv_i[icv]; noc <= m.nbocv_i[icv + 1] - 1; noc++) {
        int nb = m.nbocv_v[noc];
        for (int i = 0; i < b; i++)
          for (int j = 0; j < b; j++) out[b * icv + i] += A[(size_t)bb * noc + b * i + j] * x[b * nb + j];
      }
    }
  };
  auto max_prec = [&](const std::vector<double> &r) {
    double mx = 0.0;
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < b; i++) tmp[i] = r[b * icv + i];
      lusolv_b(&LU[(size_t)bb * icv], tmp.data(), b);
      for (int i = 0; i < b; i++) mx = max(mx, fabs(tmp[i]));
    }
    return mx;
  };
  double rho = 1.0, alpha = 1.0, omega = 1.0;
  mat_vec_b(res.data(), phi);
  for (int k = 0; k < n; k++) {
    res[k] = (rhs[k] - res[k]);
    res0[k] = res[k];
  }
  double res0_max = max_prec(res);
  int iter = 0, done = 0;
  while (done == 0) {
    iter++;
    double rho_prime = -omega * rho;
    rho = 0.0;
    for (int k = 0; k < n; k++) rho += res[k] * res0[k];
    if (fabs(rho_prime) < 1.0E-20) rho_prime = 1.0E-20;
    double beta = alpha * rho / rho_prime;
    for (int k = 0; k < n; k++) p[k] = res[k] - beta * (p[k] - omega * v[k]);
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < b; i++) phat[b * icv + i] = p[b * icv + i];
      lusolv_b(&LU[(size_t)bb * icv], &phat[b * icv], b);
    }
    mat_vec_b(v.data(), phat.data());
    double gamma = 0.0;
    for (int k = 0; k < n; k++) gamma += v[k] * res0[k];
    if (fabs(gamma) < 1.0E-20) gamma = 1.0E-20;
    alpha = rho / gamma;
    for (int k = 0; k < n; k++) s[k] = res[k] - alpha * v[k];
    for (int icv = 0; icv < ncv; icv++) {
      for (int i = 0; i < b; i++) shat[b * icv + i] = s[b * icv + i];
      lusolv_b(&LU[(size_t)bb * icv], &shat[b * icv], b);
    }
    mat_vec_b(t.data(), shat.data());
    double b0 = 0.0, b1 = 0.0;
    for (int k = 0; k < n; k++) {
      b0 += s[k] * t[k];
      b1 += t[k] * t[k];
    }
    omega = b0 / (b1 + 1.0E-20);
    for (int k = 0; k < n; k++) phi[k] += alpha * phat[k] + omega * shat[k];
    mat_vec_b(res.data(), phi);
    for (int k = 0; k < n; k++) res[k] = rhs[k] - res[k];
    double res_max = max_prec(res);
    if ((res_max <= zeroAbs) || (res_max / (res0_max + 1.0E-12) <= zeroRel))
      done = 1;
    else if (iter > maxiter)
      done = 2;
  }
  return iter;
}

// the 6x6 blocks of the coupled system with one scalar: the flow rows of A
// (their scalar column is zero with no pressure derivative) and the scalar
// row from AScal
void assemble_blocks6(const or_ctx &c, std::vector<double> &A6) {
  const size_t nnz = c.nnz;
  A6.assign(36 * nnz, 0.0);
  for (size_t noc = 0; noc < nnz; noc++) {
    for (int i = 0; i < 5; i++)
      for (int j = 0; j < 5; j++) A6[36 * noc + 6 * i + j] = c.A[25 * noc + 5 * i + j];
    for (int k = 0; k < 6; k++) A6[36 * noc + 30 + k] = c.AScal[k * nnz + noc];
  }
}

// one coupled step's linear system and solve with the transported scalar
// (runBackwardEulerCoupled, JoeWithModels.cpp:4455-4473): every row of the
// residual under-relaxed by UNDER_RELAXATION, V/dt on all six diagonal
// entries; dq6 [ncv][6]
int coupled_solve6(or_ctx &c) {
  const int ncv = c.ncv;
  std::vector<double> rhs(6 * (size_t)ncv);
  for (int icv = 0; icv < ncv; icv++) {
    rhs[6 * icv + 0] = c.rhs_rho[icv] * c.p.under_relax;
    for (int i = 0; i < 3; i++) rhs[6 * icv + 1 + i] = c.rhs_rhou[3 * icv + i] * c.p.under_relax;
    rhs[6 * icv + 4] = c.rhs_rhoE[icv] * c.p.under_relax;
    rhs[6 * icv + 5] = c.rhs_scal[icv] * c.p.under_relax;
  }
  assemble_blocks6(c, c.A6);
  for (int icv = 0; icv < ncv; icv++) {
    double tmp = c.m.cv_volume[icv] / (c.local_dt[icv]);
    for (int i = 0; i < 6; i++) c.A6[36 * (size_t)c.m.nbocv_i[icv] + 7 * i] += tmp;
  }
  c.dq6.assign(6 * (size_t)ncv, 0.0);
  return bcgstab_b(c, c.dq6.data(), c.A6.data(), rhs.data(), c.p.ls_abs, c.p.ls_rel, c.p.ls_maxiter, 6);
}

// state + BC refresh done at the start of runBackwardEuler and after each update
void refresh(or_ctx &c) {
  calc_state(c);
  calc_material(c);
  set_bc(c);
  if (c.sa) {
    set_scalar_bc(c);
    calc_turb_visc(c);
  }
  if (c.sst) {
    set_scalar_bc(c);
    calc_turb_visc_sst(c);
  }
}

}  // namespace

extern "C" {

or_ctx *or_ctx_create(const or_mesh *m, const or_params *p) {
  or_ctx *c = new or_ctx();
  c->m = *m;
  c->p = *p;
  c->ncv = m->ncv;
  c->nfa = m->nfa;
  c->nfa_b = m->nfa_b;
  c->nnz = m->nnz;
  const size_t n = m->ncv, nb = m->nfa_b;
  c->rho.assign(n, 0.0);
  c->rhou.assign(3 * n, 0.0);
  c->rhoE.assign(n, 0.0);
  c->vel.assign(3 * n, 0.0);
  c->press.assign(n, 0.0);
  c->temp.assign(n, 0.0);
  c->enthalpy.assign(n, 0.0);
  c->sos.assign(n, 0.0);
  c->rho_bfa.assign(nb, 0.0);
  c->rho_bpr.assign(nb, 0.0);
  c->vel_bpr.assign(3 * nb, 0.0);
  c->p_bpr.assign(nb, 0.0);
  // boundary profiles of PROFILES zones, linear in y (JoeWithModels.cpp:296-326)
  if (m->profiles)
    for (int z = 0; z < m->nzone; z++) {
      if (m->zone_bc[z] != OR_BC_PROFILES) continue;
      const int nn = m->nprof, nv = m->nval_prof;
      const double *pr = m->profiles;
      for (int ifa = m->zone_ifa[2 * z]; ifa <= m->zone_ifa[2 * z + 1]; ifa++) {
        const double y = m->x_fa[3 * ifa + 1];
        int pos = 1;
        while ((pr[pos * nv] < y) && (pos < nn - 1)) pos++;
        double f;
        if (y > pr[pos * nv]) f = 1.0;
        else if (y < pr[(pos - 1) * nv]) f = 0.0;
        else f = (y - pr[(pos - 1) * nv]) / (pr[pos * nv] - pr[(pos - 1) * nv]);
        const double *r1 = pr + (pos - 1) * nv, *r2 = pr + pos * nv;
        c->rho_bpr[ifa] = r1[1] + f * (r2[1] - r1[1]);
        c->vel_bpr[3 * ifa] = r1[2] + f * (r2[2] - r1[2]);
        c->vel_bpr[3 * ifa + 1] = 0.0;  // the reference drops v (:318)
        c->vel_bpr[3 * ifa + 2] = 0.0;
        c->p_bpr[ifa] = r1[4] + f * (r2[4] - r1[4]);
      }
    }
  c->T_bfa.assign(nb, 0.0);
  c->vel_bfa.assign(3 * nb, 0.0);
  c->p_bfa.assign(nb, 0.0);
  c->h_bfa.assign(nb, 0.0);
  c->grad_rho.assign(3 * n, 0.0);
  c->grad_p.assign(3 * n, 0.0);
  c->grad_u.assign(9 * n, 0.0);
  c->alpha_rho.assign(n, 1.0);
  c->local_dt.assign(n, 0.0);
  c->rhs_rho.assign(n, 0.0);
  c->rhs_rhou.assign(3 * n, 0.0);
  c->rhs_rhoE.assign(n, 0.0);
  c->A.assign(25 * (size_t)m->nnz, 0.0);
  c->dq.assign(5 * n, 0.0);
  c->mul_fa.assign(m->nfa, 0.0);
  c->lamOcp_fa.assign(m->nfa, 0.0);
  c->grad_h.assign(3 * n, 0.0);
  c->sst = p->sst != 0;
  c->nsc = p->sa ? 1 : (p->sst ? 2 : 0);
  c->resid.assign(5 + c->nsc, 0.0);
  c->mut_fa.assign(m->nfa, 0.0);
  c->flux_fa.assign(5 * (size_t)m->nfa, 0.0);
  c->sa = p->sa != 0;
  if (c->nsc) {
    const size_t ns = c->nsc;
    c->phi.assign(ns * n, 0.0);  // initialHook, JoeWithModels.cpp:219-240
    for (size_t s = 0; s < ns; s++)
      std::fill(c->phi.begin() + s * n, c->phi.begin() + (s + 1) * n, c->sst ? p->sst_init[s] : p->sa_init);
    c->phi_bfa.assign(ns * nb, 0.0);
    // each scalar's boundary profile on PROFILES zones (JoeWithModels.cpp:331-357)
    c->phi_bpr.assign(ns * nb, 0.0);
    if (m->profiles)
      for (int z = 0; z < m->nzone; z++) {
        if (m->zone_bc[z] != OR_BC_PROFILES) continue;
        const int nn = m->nprof, nv = m->nval_prof;
        const double *pr = m->profiles;
        for (size_t s = 0; s < ns && (int)(5 + s) < nv; s++)
          for (int ifa = m->zone_ifa[2 * z]; ifa <= m->zone_ifa[2 * z + 1]; ifa++) {
            const double y = m->x_fa[3 * ifa + 1];
            int pos = 1;
            while (pr[pos * nv] < y && (pos < nn - 1)) pos++;
            double fi;
            if (y > pr[pos * nv]) fi = 1.0;
            else if (y < pr[(pos - 1) * nv]) fi = 0.0;
            else fi = (y - pr[(pos - 1) * nv]) / (pr[pos * nv] - pr[(pos - 1) * nv]);
            const double phi1 = pr[(pos - 1) * nv + 5 + s], phi2 = pr[pos * nv + 5 + s];
            c->phi_bpr[s * nb + ifa] = phi1 + fi * (phi2 - phi1);
          }
      }
    c->grad_phi.assign(3 * ns * n, 0.0);
    c->diff.assign(ns * m->nfa, 0.0);
    c->wall_dist.assign(n, 0.0);
    c->str_mag.assign(n, 0.0);
    c->vort_mag.assign(n, 0.0);
    c->rhs_scal.assign(ns * n, 0.0);
    c->AScal.assign(6 * ns * (size_t)m->nnz, 0.0);
    c->dscal.assign(ns * n, 0.0);
    if (c->sst) {
      c->diverg.assign(n, 0.0);
      c->cross_diff.assign(n, 0.0);
      c->f1.assign(n, 0.0);
      c->f2.assign(n, 0.0);
    }
    // initialHookScalarRansTurbModel, TurbModel_SA.h:62-72, TurbModel_KOMSST.h:82-98
    calc_wall_distance(*c);
  }
  c->cfl = p->cfl;
  c->step = 0;
  return c;
}

void or_set_scalar(or_ctx *c, const double *phi) {
  if (c->nsc) std::memcpy(c->phi.data(), phi, sizeof(double) * c->ncv * c->nsc);
}

void or_ctx_destroy(or_ctx *c) { delete c; }

void or_set_state(or_ctx *c, const double *rho, const double *rhou, const double *rhoE) {
  std::memcpy(c->rho.data(), rho, sizeof(double) * c->ncv);
  std::memcpy(c->rhou.data(), rhou, sizeof(double) * 3 * c->ncv);
  std::memcpy(c->rhoE.data(), rhoE, sizeof(double) * c->ncv);
}

void or_get_state(const or_ctx *c, double *rho, double *rhou, double *rhoE) {
  std::memcpy(rho, c->rho.data(), sizeof(double) * c->ncv);
  std::memcpy(rhou, c->rhou.data(), sizeof(double) * 3 * c->ncv);
  std::memcpy(rhoE, c->rhoE.data(), sizeof(double) * c->ncv);
}

void or_residual(or_ctx *c, int implicit) {
  refresh(*c);
  calc_dt(*c);
  if (implicit) std::fill(c->A.begin(), c->A.end(), 0.0);
  if (c->nsc) std::fill(c->AScal.begin(), c->AScal.end(), 0.0);
  calc_rhs(*c, implicit ? c->A.data() : NULL);
  if (implicit && c->sa && c->p.coupled) assemble_blocks6(*c, c->A6);
}

// one pass of the loop body of runBackwardEuler (JoeWithModels.cpp:975-1158),
// assuming state/BC are current (as after refresh).
static int step_once(or_ctx &c) {
  c.step++;
  if ((c.step >= c.p.cfl_after) && (c.step % c.p.cfl_interval == 0) && (c.cfl < c.p.cfl_max)) c.cfl *= c.p.cfl_factor;
  calc_dt(c);
  std::fill(c.A.begin(), c.A.end(), 0.0);
  std::fill(c.dq.begin(), c.dq.end(), 0.0);
  if (c.nsc) std::fill(c.AScal.begin(), c.AScal.end(), 0.0);
  calc_rhs(c, c.A.data());
  if (c.sa && c.p.coupled) {  // runBackwardEulerCoupled with the scalar (JoeWithModels.cpp:4437-4516)
    for (size_t i = 0; i < c.resid.size(); i++) c.resid[i] = 0.0;
    for (int icv = 0; icv < c.ncv; icv++) {
      c.resid[0] += fabs(c.rhs_rho[icv]);
      for (int i = 0; i < 3; i++) c.resid[i + 1] += fabs(c.rhs_rhou[3 * icv + i]);
      c.resid[4] += fabs(c.rhs_rhoE[icv]);
      c.resid[5] += fabs(c.rhs_scal[icv]);
    }
    int it = coupled_solve6(c);
    for (int icv = 0; icv < c.ncv; icv++) {
      const double *d = &c.dq6[6 * (size_t)icv];
      c.rho[icv] += d[0];
      c.rhou[3 * icv + 0] += d[1];
      c.rhou[3 * icv + 1] += d[2];
      c.rhou[3 * icv + 2] += d[3];
      c.rhoE[icv] += d[4];
      for (int k = 0; k < 5; k++) c.dq[5 * icv + k] = d[k];
    }
    for (int icv = 0; icv < c.ncv; icv++)  // update + clip (:4510-4516)
      c.phi[icv] = min(max((c.phi[icv] * (c.rho[icv] - c.dq6[6 * (size_t)icv]) + c.dq6[6 * (size_t)icv + 5]) / c.rho[icv],
                           sa::lowerBound),
                       sa::upperBound);
    refresh(c);
    return it;
  }
  std::vector<double> rhs(5 * (size_t)c.ncv);
  for (int icv = 0; icv < c.ncv; icv++) {
    rhs[5 * icv + 0] = c.p.under_relax * c.rhs_rho[icv];
    rhs[5 * icv + 1] = c.p.under_relax * c.rhs_rhou[3 * icv + 0];
    rhs[5 * icv + 2] = c.p.under_relax * c.rhs_rhou[3 * icv + 1];
    rhs[5 * icv + 3] = c.p.under_relax * c.rhs_rhou[3 * icv + 2];
    rhs[5 * icv + 4] = c.p.under_relax * c.rhs_rhoE[icv];
    double tmp = c.m.cv_volume[icv] / (c.local_dt[icv]);
    for (int i = 0; i < 5; i++) c.A[25 * c.m.nbocv_i[icv] + 6 * i] += tmp;
  }
  int it = bcgstab(c, c.dq.data(), c.A.data(), rhs.data(), c.p.ls_abs, c.p.ls_rel, c.p.ls_maxiter);
  for (int icv = 0; icv < c.ncv; icv++) {
    c.rho[icv] += c.dq[5 * icv + 0];
    c.rhou[3 * icv + 0] += c.dq[5 * icv + 1];
    c.rhou[3 * icv + 1] += c.dq[5 * icv + 2];
    c.rhou[3 * icv + 2] += c.dq[5 * icv + 3];
    c.rhoE[icv] += c.dq[5 * icv + 4];
  }
  if (c.nsc) scalar_solve(c, true);
  refresh(c);
  for (size_t i = 0; i < c.resid.size(); i++) c.resid[i] = 0.0;
  for (int icv = 0; icv < c.ncv; icv++) {
    c.resid[0] += fabs(c.rhs_rho[icv]);
    for (int i = 0; i < 3; i++) c.resid[i + 1] += fabs(c.rhs_rhou[3 * icv + i]);
    c.resid[4] += fabs(c.rhs_rhoE[icv]);
  }
  for (int s = 0; s < c.nsc; s++)  // JoeWithModels.cpp:1132-1134
    for (int icv = 0; icv < c.ncv; icv++) c.resid[5 + s] += fabs(c.rhs_scal[(size_t)c.ncv * s + icv] / c.p.under_relax);
  return it;
}

int or_step(or_ctx *c) {
  if (c->step == 0 || c->resume_refresh) refresh(*c);
  c->resume_refresh = false;
  return step_once(*c);
}

// a solve resumed from a restart file keeps the step count it read
// (runBackwardEuler does not reset it, JoeWithModels.cpp:977); CFL from Joe.in
void or_set_step(or_ctx *c, int step) {
  c->step = step;
  c->cfl = c->p.cfl;
  c->resume_refresh = true;
}

int or_run(or_ctx *c, int max_steps, double *hist) {
  refresh(*c);
  // doneSolver(1e20) before the loop, then after every step (JoeWithModels.h:42-95)
  auto done = [&](double r) {
    if ((c->p.nsteps >= 0) && (c->step >= c->p.nsteps)) return true;
    if (r < c->p.resid_energ_th) return true;
    if (r != r) return true;
    return false;
  };
  int n = 0;
  if (done(1.0e20)) return 0;
  while (n < max_steps) {
    step_once(*c);
    const int nr = (int)c->resid.size();
    if (hist)
      for (int i = 0; i < nr; i++) hist[nr * n + i] = c->resid[i];
    n++;
    if (done(c->resid[4])) break;
  }
  return n;
}

long long or_get(const or_ctx *c, const char *name, double *out) {
  const std::vector<double> *v = nullptr;
  std::string n(name);
  if (n == "RHS_RHO") v = &c->rhs_rho;
  else if (n == "RHS_RHOU") v = &c->rhs_rhou;
  else if (n == "RHS_RHOE") v = &c->rhs_rhoE;
  else if (n == "A") v = (c->sa && c->p.coupled) ? &c->A6 : &c->A;
  else if (n == "LOCAL_DT") v = &c->local_dt;
  else if (n == "GRAD_RHO") v = &c->grad_rho;
  else if (n == "GRAD_P") v = &c->grad_p;
  else if (n == "GRAD_U") v = &c->grad_u;
  else if (n == "ALPHA_RHO") v = &c->alpha_rho;
  else if (n == "VEL") v = &c->vel;
  else if (n == "PRESS") v = &c->press;
  else if (n == "TEMP") v = &c->temp;
  else if (n == "ENTHALPY") v = &c->enthalpy;
  else if (n == "MUL_FA") v = &c->mul_fa;
  else if (n == "FLUX_FA") v = &c->flux_fa;
  else if (n == "GRAD_ENTHALPY") v = &c->grad_h;
  else if (n == "SOS") v = &c->sos;
  else if (n == "RHO_BFA") v = &c->rho_bfa;
  else if (n == "VEL_BFA") v = &c->vel_bfa;
  else if (n == "P_BFA") v = &c->p_bfa;
  else if (n == "T_BFA") v = &c->T_bfa;
  else if (n == "H_BFA") v = &c->h_bfa;
  else if (n == "DQ") v = (c->sa && c->p.coupled) ? &c->dq6 : &c->dq;
  else if (n == "RESID") v = &c->resid;
  else if (n == "RHO") v = &c->rho;
  else if (n == "RHOU") v = &c->rhou;
  else if (n == "RHOE") v = &c->rhoE;
  else if (n == "MUT_FA") v = &c->mut_fa;
  else if (c->sa) {
    if (n == "NUSA") v = &c->phi;
    else if (n == "NUSA_BFA") v = &c->phi_bfa;
    else if (n == "GRAD_NUSA") v = &c->grad_phi;
    else if (n == "WALL_DIST") v = &c->wall_dist;
    else if (n == "STR_MAG") v = &c->str_mag;
    else if (n == "VORT_MAG") v = &c->vort_mag;
    else if (n == "DIFF_NUSA") v = &c->diff;
    else if (n == "RHS_NUSA") v = &c->rhs_scal;
    else if (n == "ASCAL") v = &c->AScal;
    else if (n == "DSCAL") v = &c->dscal;
  } else if (c->sst) {
    static thread_local std::vector<double> part;
    const size_t nc = c->ncv, nb = c->nfa_b, nf = c->nfa;
    auto half = [&](const std::vector<double> &src, size_t len, int s) {
      part.assign(src.begin() + len * s, src.begin() + len * (s + 1));
      return &part;
    };
    if (n == "KINE") v = half(c->phi, nc, 0);
    else if (n == "OMEGA") v = half(c->phi, nc, 1);
    else if (n == "KINE_BFA") v = half(c->phi_bfa, nb, 0);
    else if (n == "OMEGA_BFA") v = half(c->phi_bfa, nb, 1);
    else if (n == "GRAD_KINE") v = half(c->grad_phi, 3 * nc, 0);
    else if (n == "GRAD_OMEGA") v = half(c->grad_phi, 3 * nc, 1);
    else if (n == "DIFF_KINE") v = half(c->diff, nf, 0);
    else if (n == "DIFF_OMEGA") v = half(c->diff, nf, 1);
    else if (n == "RHS_KINE") v = half(c->rhs_scal, nc, 0);
    else if (n == "RHS_OMEGA") v = half(c->rhs_scal, nc, 1);
    else if (n == "WALL_DIST") v = &c->wall_dist;
    else if (n == "STR_MAG") v = &c->str_mag;
    else if (n == "VORT_MAG") v = &c->vort_mag;
    else if (n == "DIVERG") v = &c->diverg;
    else if (n == "CROSS_DIFF") v = &c->cross_diff;
    else if (n == "F1") v = &c->f1;
    else if (n == "F2") v = &c->f2;
    else if (n == "ASCAL") v = &c->AScal;
    else if (n == "DSCAL") v = &c->dscal;
  }
  if (!v) return -1;
  if (out) std::memcpy(out, v->data(), sizeof(double) * v->size());
  return (long long)v->size();
}

double or_cfl(const or_ctx *c) { return c->cfl; }
int or_nsteps_done(const or_ctx *c) { return c->step; }

/* One linear solve on the current A/rhs of the last or_residual call with
 * the first-step scaling (JoeWithModels.cpp:1011-1026); result in DQ (and,
 * with SA, the scalar correction in DSCAL). */
int or_linear_solve(or_ctx *c) {
  if (c->sa && c->p.coupled) return coupled_solve6(*c);  // DQ [ncv][6]
  std::vector<double> rhs(5 * (size_t)c->ncv);
  for (int icv = 0; icv < c->ncv; icv++) {
    rhs[5 * icv + 0] = c->p.under_relax * c->rhs_rho[icv];
    for (int i = 0; i < 3; i++) rhs[5 * icv + 1 + i] = c->p.under_relax * c->rhs_rhou[3 * icv + i];
    rhs[5 * icv + 4] = c->p.under_relax * c->rhs_rhoE[icv];
    double tmp = c->m.cv_volume[icv] / (c->local_dt[icv]);
    for (int i = 0; i < 5; i++) c->A[25 * c->m.nbocv_i[icv] + 6 * i] += tmp;
  }
  std::fill(c->dq.begin(), c->dq.end(), 0.0);
  int it = bcgstab(*c, c->dq.data(), c->A.data(), rhs.data(), c->p.ls_abs, c->p.ls_rel, c->p.ls_maxiter);
  if (c->nsc) scalar_solve(*c, false);  // DSCAL: the scalar corrections of that first step
  return it;
}

/* writeWallValues of the reference's case class (joe.cpp:43-130) on boundary
 * zone `zone` (index among the boundary zones) at the current state and
 * boundary values: per face 10 columns x_fa (3), rho, press, temp of its
 * cell, tau_wall = mulam |u_t| / walld sign(u_x), qdot with the reference's
 * wall temperature 300, yplus and mulam, the face's laminar viscosity
 * (mul_fa of the boundary face, from T_bfa); then the face's force on the
 * wall in 6 more columns: pressure p_bfa n A (3) and shear mulam u_t / walld
 * A (3), n the unit normal and A the area. out [n][16] (may be NULL);
 * returns n or -1. */
int or_wall_values(or_ctx *c, int zone, double *out) {
  const or_mesh &m = c->m;
  if (zone < 0 || zone >= m.nzone) return -1;
  const int f0 = m.zone_ifa[2 * zone], f1 = m.zone_ifa[2 * zone + 1];
  const int n = f1 - f0 + 1;
  if (!out) return n;
  const double R = c->p.R_gas, gam = c->p.gamma;
  for (int ifa = f0; ifa <= f1; ifa++) {
    const int icv = m.cvofa[2 * ifa];
    double nv[3], s_half[3], vel[3], velTang[3];
    const double area = norm3(nv, &m.fa_normal[3 * ifa]);
    for (int i = 0; i < 3; i++) s_half[i] = m.x_fa[3 * ifa + i] - m.x_cv[3 * icv + i];
    for (int i = 0; i < 3; i++) vel[i] = c->rhou[3 * icv + i] / c->rho[icv];
    const double un = dot3(nv, vel);
    for (int i = 0; i < 3; i++) velTang[i] = vel[i] - un * nv[i];
    const double velMag = sqrt(dot3(velTang, velTang));
    const double walld = fabs(dot3(s_half, nv));
    const double wallTemp = 300.0;
    const double mulam = c->mul_fa[ifa];
    const double sgn = vel[0] > 0 ? 1.0 : (vel[0] < 0.0 ? -1.0 : 0.0);
    const double tau = mulam * velMag / walld * sgn;
    const double utau = sqrt(fabs(tau) / c->rho[icv]);
    const double yplus = walld * utau / (mulam / c->rho[icv]);
    const double kTotal = R * gam / (gam - 1.0) * mulam / c->p.pr;
    const double qDot = kTotal * (c->temp[icv] - wallTemp) / walld;
    double *o = out + 16 * (size_t)(ifa - f0);
    for (int i = 0; i < 3; i++) o[i] = m.x_fa[3 * ifa + i];
    o[3] = c->rho[icv];
    o[4] = c->press[icv];
    o[5] = c->temp[icv];
    o[6] = tau;
    o[7] = qDot;
    o[8] = yplus;
    o[9] = mulam;
    for (int i = 0; i < 3; i++) o[10 + i] = c->p_bfa[ifa] * nv[i] * area;
    for (int i = 0; i < 3; i++) o[13 + i] = mulam * velTang[i] / walld * area;
  }
  return n;
}

}  // extern "C"
