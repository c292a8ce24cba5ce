// Case classes written the way a reference joe.cpp writes them (a class per
// run number deriving from JoeWithModels, joe.cpp:24-40, 963-987), compiled
// against the host mirror. Used by the tests to exercise every hook.
//
//   joe_cases <case> <Joe.in> <out.bin>
//     plain     JoeWithModels with no overrides
//     inlethook boundaryHook sets the zone "inlet" to the constant state a
//               CBC inflow would impose (T, u, p from INLET_HOOK_STATE)
//     gravity   sourceHook overridden, calling the default (gravity) hook
//     perturb   initialHook adds a smooth density perturbation
//     sa        the Spalart-Allmaras mix-in (MyJoeSA, joe.cpp:286-296); nuSA
//               is appended to out.bin
//     sst       the Menter SST mix-in (MyJoeSST, joe.cpp:301-310); kine and
//               omega are appended to out.bin
//     walls     constructed as joe.cpp's MyJoe is (naming the virtual base
//               UgpWithCvCompFlow), it reads the host copies in finalHook to
//               form each WALL face's shear, y+, heat flux and viscosity
//               (writeWallValues, joe.cpp:43-130) and appends their count
//               and values to out.bin
//
// out.bin: step count, the five L1 residuals of every step, then rho[ncv],
// rhou[ncv][3], rhoE[ncv] (reference numbering), all as doubles.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iostream>
#include <vector>

#include "../JoeWithModels.hpp"

class Recorder : public JoeWithModels {
 public:
  explicit Recorder(const char *f) : JoeWithModels(f) {}
  std::vector<double> hist;
  int temporal_calls = 0, final_calls = 0;
  void temporalHook() override {
    temporal_calls++;
    for (int k = 0; k < 5; k++) hist.push_back(Residual[k]);
  }
  void finalHook() override { final_calls++; }
  void write(const char *path) {
    std::FILE *fp = std::fopen(path, "wb");
    if (!fp) throw std::runtime_error("cannot write output");
    double n = (double)(hist.size() / 5);  // steps this run took (a resumed run starts past 0)
    std::fwrite(&n, sizeof(double), 1, fp);
    std::fwrite(hist.data(), sizeof(double), hist.size(), fp);
    std::fwrite(rho, sizeof(double), ncv, fp);
    std::fwrite(&rhou[0][0], sizeof(double), 3 * (size_t)ncv, fp);
    std::fwrite(rhoE, sizeof(double), ncv, fp);
    std::fclose(fp);
  }
};

class InletHook : public Recorder {
 public:
  explicit InletHook(const char *f) : Recorder(f) {}
  int calls = 0;
  void boundaryHook(double *T_bfa, double (*vel_bfa)[3], double *p_bfa, FaZone *zone) override {
    if (zone->getNameString() != "inlet") return;
    calls++;
    const double T = getDoubleParam("INLET_HOOK_T"), u = getDoubleParam("INLET_HOOK_U"),
                 p = getDoubleParam("INLET_HOOK_P");
    for (int ifa = zone->ifa_f; ifa <= zone->ifa_l; ifa++) {
      T_bfa[ifa] = T;
      vel_bfa[ifa][0] = u;
      vel_bfa[ifa][1] = 0.0;
      vel_bfa[ifa][2] = 0.0;
      p_bfa[ifa] = p;
    }
  }
};

class GravityHook : public Recorder {
 public:
  explicit GravityHook(const char *f) : Recorder(f) {}
  int calls = 0;
  void sourceHook(double *rhs_rho, double (*rhs_rhou)[3], double *rhs_rhoE, double (*A)[5][5]) override {
    calls++;
    JoeWithModels::sourceHook(rhs_rho, rhs_rhou, rhs_rhoE, A);
  }
};

class Perturb : public Recorder {
 public:
  explicit Perturb(const char *f) : Recorder(f) {}
  void initialHook() override {
    JoeWithModels::initialHook();
    for (int icv = 0; icv < ncv; icv++) {
      double s = 1.0 + 0.01 * std::sin(3.0 * x_cv[icv][0]) * std::cos(2.0 * x_cv[icv][1]);
      rho[icv] *= s;
      for (int i = 0; i < 3; i++) rhou[icv][i] *= s;
      rhoE[icv] *= s;
    }
  }
};

// the host copies a reference case class reads: mul_fa, the state, the gas
// constants, faZoneList and getParam, as MyJoe::writeWallValues does
class WallCase : public Recorder {
 public:
  explicit WallCase(const char *f) : Recorder(f), UgpWithCvCompFlow(f) {}
  std::vector<double> walls;  // per wall face: tau_wall, y+, heat flux, mul
  int fresh = 0;              // temporalHook calls that saw a new host state
  double seen = -1.0;
  void temporalHook() override {
    Recorder::temporalHook();
    if (rhoE[ncv / 2] != seen) fresh++;
    seen = rhoE[ncv / 2];
  }
  void finalHook() override {
    Recorder::finalHook();
    for (list<FaZone>::iterator zone = faZoneList.begin(); zone != faZoneList.end(); zone++) {
      Param *param;
      if (zone->getKind() != FA_ZONE_BOUNDARY || !getParam(param, zone->getName()) || param->getString() != "WALL")
        continue;
      for (int ifa = zone->ifa_f; ifa <= zone->ifa_l; ifa++) {
        const int icv = cvofa[ifa][0];
        double n[3], d[3], v[3], vt[3];
        normVec3d(n, fa_normal[ifa]);
        vecMinVec3d(d, x_fa[ifa], x_cv[icv]);
        for (int i = 0; i < 3; i++) v[i] = rhou[icv][i] / rho[icv];
        const double un = vecDotVec3d(n, v);
        for (int i = 0; i < 3; i++) vt[i] = v[i] - un * n[i];
        const double dw = fabs(vecDotVec3d(d, n));
        const double tau = mul_fa[ifa] * sqrt(vecDotVec3d(vt, vt)) / dw * sign(v[0]);
        const double yplus = dw * sqrt(fabs(tau) / rho[icv]) / (mul_fa[ifa] / rho[icv]);
        const double kappa = R_gas * gamma[icv] / (gamma[icv] - 1.0) * mul_fa[ifa] / Pr;
        walls.insert(walls.end(), {tau, yplus, kappa * (temp[icv] - 300.0) / dw, mul_fa[ifa]});
      }
    }
  }
};

class SaCase : public Recorder, public RansTurbSA {
 public:
  explicit SaCase(const char *f) : Recorder(f) {}
};

class SstCase : public Recorder, public RansTurbKOmSST {
 public:
  explicit SstCase(const char *f) : Recorder(f) {}
};

int main(int argc, char **argv) {
  if (argc != 4) {
    std::cerr << "usage: joe_cases <plain|inlethook|gravity|perturb|sa|sst|walls> <Joe.in> <out.bin>\n";
    return 2;
  }
  try {
    const std::string c = argv[1];
    if (c == "plain") {
      Recorder j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::printf("plain steps=%d temporal=%d final=%d\n", j.step, j.temporal_calls, j.final_calls);
    } else if (c == "inlethook") {
      InletHook j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::printf("inlethook steps=%d hook_calls=%d\n", j.step, j.calls);
      if (j.calls != j.step + 1) return 3;  // once before the loop, once per step
    } else if (c == "gravity") {
      GravityHook j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::printf("gravity steps=%d hook_calls=%d\n", j.step, j.calls);
      if (j.calls != j.step) return 3;
    } else if (c == "perturb") {
      Perturb j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::printf("perturb steps=%d\n", j.step);
    } else if (c == "walls") {
      WallCase j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::FILE *fp = std::fopen(argv[3], "ab");
      if (!fp) throw std::runtime_error("cannot write output");
      const double nw = (double)(j.walls.size() / 4);
      std::fwrite(&nw, sizeof(double), 1, fp);
      std::fwrite(j.walls.data(), sizeof(double), j.walls.size(), fp);
      std::fclose(fp);
      std::printf("walls steps=%d fresh=%d faces=%d\n", j.step, j.fresh, (int)nw);
    } else if (c == "sa") {
      SaCase j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::FILE *fp = std::fopen(argv[3], "ab");
      if (!fp) throw std::runtime_error("cannot write output");
      std::fwrite(j.nuSA, sizeof(double), j.ncv, fp);
      std::fclose(fp);
      std::printf("sa steps=%d\n", j.step);
    } else if (c == "sst") {
      SstCase j(argv[2]);
      j.run();
      j.write(argv[3]);
      std::FILE *fp = std::fopen(argv[3], "ab");
      if (!fp) throw std::runtime_error("cannot write output");
      std::fwrite(j.kine, sizeof(double), j.ncv, fp);
      std::fwrite(j.omega, sizeof(double), j.ncv, fp);
      std::fclose(fp);
      std::printf("sst steps=%d\n", j.step);
    } else {
      std::cerr << "unknown case " << c << "\n";
      return 2;
    }
  } catch (std::exception &e) {
    std::cerr << "error: " << e.what() << "\n";
    return 1;
  }
  return 0;
}
