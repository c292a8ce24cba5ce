// Test-only harness around the reference Joe solver (btracey/myjoe), compiled
// from the reference's own sources by oracle/Makefile into oracle/_ref/.
// It never ships with the product: tests and the golden-vector script run it
// to pin the CPU restatement (oracle/joe_oracle.cpp) and the HIP path.
//
// Modes (argv[1]):
//   init  <Joe.in> <out.bin>              mesh arrays after Ugp/UgpWithCv init
//   rhs   <Joe.in> <state.bin> <out.bin>  one implicit calcRhs on a given state
//   solve <Joe.in> <out.bin> [nsnap]      JoeWithModels::run(); state history
//   time  <Joe.in> <state.bin> <reps>     seconds per residual evaluation (state,
//                                         BC, dt, explicit calcRhs), printed
//   restart-write <Joe.in> <state.bin> <restart>  Ugp::writeRestart of a given state
//   restart-read  <Joe.in> <out.bin>      mesh arrays and RHO, RHOU, RHOE (NUSA)
//                                         after initializeFromRestartFile(RESTART)
//   rhs-coupled <Joe.in> <state.bin> <out.bin>  one calcRhsCoupled (the
//                                         residual and (5+nScal)^2 blocks of
//                                         runBackwardEulerCoupled); with -sa
//                                         the nuSA row and the model fields
// rhs-sa / solve-sa run the same with the Spalart-Allmaras mix-in (RansTurbSA,
// as joe.cpp's MyJoeSA); the state file may then carry NUSA.
// rhs-sst / solve-sst run it with the Menter SST mix-in (RansTurbKOmSST, as
// joe.cpp's MyJoeSST); the state file then carries KINE and OMEGA, and the
// scalar records are named after each scalar (KINE_BFA, RHS_OMEGA, ...).
//
// The dump format is a sequence of records:
//   char name[32]; int32 kind (0 = f64, 1 = i32, 2 = bytes); int32 pad;
//   int64 count; payload.
#include "JoeWithModels.h"
#include "TurbModel_SA.h"
#include "TurbModel_KOMSST.h"
#include "myMem.h"

#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct Dump {
  FILE *fp = nullptr;
  explicit Dump(const char *path) {
    fp = fopen(path, "wb");
    if (!fp) { fprintf(stderr, "cannot open %s\n", path); throw(-1); }
    fwrite("JMDUMP01", 1, 8, fp);
  }
  ~Dump() { if (fp) fclose(fp); }
  void rec(const char *name, int kind, const void *p, long long count) {
    char nm[32];
    memset(nm, 0, sizeof(nm));
    strncpy(nm, name, 31);
    int hdr[2] = {kind, 0};
    fwrite(nm, 1, 32, fp);
    fwrite(hdr, sizeof(int), 2, fp);
    fwrite(&count, sizeof(long long), 1, fp);
    size_t sz = kind == 0 ? 8 : (kind == 1 ? 4 : 1);
    if (count > 0) fwrite(p, sz, (size_t)count, fp);
  }
  void f64(const char *n, const double *p, long long c) { rec(n, 0, p, c); }
  void i32(const char *n, const int *p, long long c) { rec(n, 1, p, c); }
};

// Reads the same record format (used for the state file of mode "rhs").
struct Load {
  std::vector<std::pair<std::string, std::vector<char>>> recs;
  explicit Load(const char *path) {
    FILE *fp = fopen(path, "rb");
    if (!fp) { fprintf(stderr, "cannot open %s\n", path); throw(-1); }
    char magic[8];
    if (fread(magic, 1, 8, fp) != 8 || memcmp(magic, "JMDUMP01", 8)) throw(-1);
    while (true) {
      char nm[33] = {0};
      int hdr[2];
      long long count;
      if (fread(nm, 1, 32, fp) != 32) break;
      if (fread(hdr, sizeof(int), 2, fp) != 2) throw(-1);
      if (fread(&count, sizeof(long long), 1, fp) != 1) throw(-1);
      size_t sz = hdr[0] == 0 ? 8 : (hdr[0] == 1 ? 4 : 1);
      std::vector<char> buf(sz * (size_t)count);
      if (count > 0 && fread(buf.data(), sz, (size_t)count, fp) != (size_t)count) throw(-1);
      recs.emplace_back(std::string(nm), std::move(buf));
    }
    fclose(fp);
  }
  const double *f64(const char *name, long long expect) const {
    for (auto &r : recs)
      if (r.first == name) {
        if ((long long)r.second.size() != expect * 8) {
          fprintf(stderr, "record %s has wrong size\n", name);
          throw(-1);
        }
        return (const double *)r.second.data();
      }
    fprintf(stderr, "record %s missing\n", name);
    throw(-1);
  }
};

class RefJoe : public JoeWithModels {
 public:
  explicit RefJoe(char *name) : JoeWithModels(name), UgpWithCvCompFlow(name) {}

  Dump *snap = nullptr;
  int nsnap = 0;

  // The reference leaves mul_fa/lamOcp_fa uninitialised when MU_REF = 0 but
  // calcDt reads mul_fa (UgpWithCvCompFlow.h:533, UgpWithCvCompFlow.cpp:50);
  // zero them so the reference's local time step is defined.
  void initialHook() {
    JoeWithModels::initialHook();
    for (int ifa = 0; ifa < nfa; ifa++) {
      mul_fa[ifa] = 0.0;
      lamOcp_fa[ifa] = 0.0;
    }
  }

  // fields of a turbulence model mix-in (none here)
  virtual void dumpModel(Dump &d) {}

  // the record name of transported scalar s: its registered name in capitals
  // (NUSA; KINE, OMEGA)
  std::string key(int s) {
    std::string k = scalarTranspEqVector[s].getName();
    for (char &ch : k) ch = (char)toupper((unsigned char)ch);
    return k;
  }

  // every transported scalar from the state file
  void loadScalars(const Load &st) {
    for (int s = 0; s < (int)scalarTranspEqVector.size(); s++) {
      const double *v = st.f64(key(s).c_str(), ncv);
      for (int icv = 0; icv < ncv; icv++) scalarTranspEqVector[s].phi[icv] = v[icv];
    }
  }

  void dumpState(const char *suffix) {
    char n[32];
    snprintf(n, 32, "RHO%s", suffix);  snap->f64(n, rho, ncv);
    snprintf(n, 32, "RHOU%s", suffix); snap->f64(n, &rhou[0][0], 3LL * ncv);
    snprintf(n, 32, "RHOE%s", suffix); snap->f64(n, rhoE, ncv);
    for (int s = 0; s < (int)scalarTranspEqVector.size(); s++) {
      snprintf(n, 32, "%s%s", key(s).c_str(), suffix);
      snap->f64(n, scalarTranspEqVector[s].phi, ncv);
    }
  }

  // the first nsnap steps of this run, numbered from 1 (a run resumed from a
  // restart file starts at the file's STEP + 1)
  int snap_first = -1;
  void temporalHook() {
    if (snap_first < 0) snap_first = step;
    const int k = step - snap_first + 1;
    if (snap && k <= nsnap) {
      char s[16];
      snprintf(s, 16, "@%d", k);
      dumpState(s);
      char n[32];
      snprintf(n, 32, "RESIDFIELD@%d", k);
      snap->f64(n, residField, ncv);
      snprintf(n, 32, "LOCALDT@%d", k);
      snap->f64(n, local_dt, ncv);
    }
  }

  void finalHook() {
    if (snap) {
      dumpState("");
      int st = step;
      snap->i32("STEP", &st, 1);
    }
  }

  void dumpMesh(Dump &d) {
    int counts[8] = {nno, nfa, nfa_b, ncv, ncv_g, nbocv_s, faocv_s, noofa_s};
    d.i32("COUNTS", counts, 8);
    d.f64("X_NO", &x_no[0][0], 3LL * nno);
    d.i32("NOOFA_I", noofa_i, nfa + 1);
    d.i32("NOOFA_V", noofa_v, noofa_i[nfa]);
    d.i32("CVOFA", &cvofa[0][0], 2LL * nfa);
    d.i32("FAOCV_I", faocv_i, ncv + 1);
    d.i32("FAOCV_V", faocv_v, faocv_i[ncv]);
    d.i32("NBOCV_I", nbocv_i, ncv + 1);
    d.i32("NBOCV_V", nbocv_v, nbocv_s);
    d.f64("X_FA", &x_fa[0][0], 3LL * nfa);
    d.f64("FA_NORMAL", &fa_normal[0][0], 3LL * nfa);
    d.f64("X_CV", &x_cv[0][0], 3LL * ncv_g);
    d.f64("CV_VOLUME", cv_volume, ncv_g);
    d.f64("NBOCV_LSG", &nbocv_lsg_coeff[0][0], 3LL * nbocv_s);
    d.f64("BFA_LSG", &boundary_fa_lsg_coeff[0][0], 3LL * nfa_b);
    // zone table: kind, ifa_f, ifa_l per zone, names joined by '\n'
    std::vector<int> zt;
    std::string names;
    for (list<FaZone>::iterator z = faZoneList.begin(); z != faZoneList.end(); z++) {
      zt.push_back(z->getKind());
      zt.push_back(z->ifa_f);
      zt.push_back(z->ifa_l);
      names += z->getName();
      names += "\n";
    }
    d.i32("ZONES", zt.data(), (long long)zt.size());
    d.rec("ZONE_NAMES", 2, names.data(), (long long)names.size());
  }

  void dumpPrims(Dump &d) {
    d.f64("VEL", &vel[0][0], 3LL * ncv);
    d.f64("PRESS", press, ncv);
    d.f64("TEMP", temp, ncv);
    d.f64("ENTHALPY", enthalpy, ncv);
    d.f64("SOS", sos, ncv);
    d.f64("RHO_BFA", rho_bfa, nfa_b);
    d.f64("VEL_BFA", &vel_bfa[0][0], 3LL * nfa_b);
    d.f64("P_BFA", p_bfa, nfa_b);
    d.f64("T_BFA", T_bfa, nfa_b);
    d.f64("H_BFA", h_bfa, nfa_b);
    if (mu_ref > 0.0) d.f64("MUL_FA", mul_fa, nfa);
  }

  // reps explicit residual evaluations on a given state, timed: the work of
  // the restatement's or_residual(c, 0) (calcStateVariables, setBC, calcDt,
  // calcRhs without A; JoeWithModels.cpp:943-1005)
  double timeRhs(const Load &st, int reps) {
    initializeFromRestartFile(getStringParam("RESTART"));
    initialHook();
    initialHookScalarRansTurbModel();
    initialHookScalarRansCombModel();
    const double *r = st.f64("RHO", ncv);
    const double *ru = st.f64("RHOU", 3LL * ncv);
    const double *re = st.f64("RHOE", ncv);
    double *RHSrho = new double[ncv];
    double (*RHSrhou)[3] = new double[ncv][3];
    double *RHSrhoE = new double[ncv];
    auto t0 = std::chrono::steady_clock::now();
    for (int k = 0; k < reps; k++) {
      for (int icv = 0; icv < ncv; icv++) {
        rho[icv] = r[icv];
        for (int i = 0; i < 3; i++) rhou[icv][i] = ru[3 * icv + i];
        rhoE[icv] = re[icv];
      }
      calcStateVariables();
      calcMaterialProperties();
      setBC();
      calcRansTurbViscMuet();
      calcDt(cfl);
      calcRhs(RHSrho, RHSrhou, RHSrhoE, NULL, NULL, NULL, false);
    }
    auto t1 = std::chrono::steady_clock::now();
    delete[] RHSrho;
    delete[] RHSrhou;
    delete[] RHSrhoE;
    return std::chrono::duration<double>(t1 - t0).count() / reps;
  }

  // Ugp::writeRestart (Ugp.cpp:9902-10660) of a given state on the mesh of RESTART
  void writeState(const Load &st, const char *path) {
    initializeFromRestartFile(getStringParam("RESTART"));
    initialHook();
    initialHookScalarRansTurbModel();
    initialHookScalarRansCombModel();
    const double *r = st.f64("RHO", ncv);
    const double *ru = st.f64("RHOU", 3LL * ncv);
    const double *re = st.f64("RHOE", ncv);
    for (int icv = 0; icv < ncv; icv++) {
      rho[icv] = r[icv];
      for (int i = 0; i < 3; i++) rhou[icv][i] = ru[3 * icv + i];
      rhoE[icv] = re[icv];
    }
    loadScalars(st);
    calcStateVariables();
    writeRestart(path);
  }

  // the mesh and conserved state initializeFromRestartFile (Ugp.cpp:12473-13240) gives
  void readState(Dump &d) {
    initializeFromRestartFile(getStringParam("RESTART"));
    dumpMesh(d);
    d.f64("RHO", rho, ncv);
    d.f64("RHOU", &rhou[0][0], 3LL * ncv);
    d.f64("RHOE", rhoE, ncv);
    for (int s = 0; s < (int)scalarTranspEqVector.size(); s++) d.f64(key(s).c_str(), scalarTranspEqVector[s].phi, ncv);
  }

  // One coupled residual evaluation in the order runBackwardEulerCoupled uses
  // (JoeWithModels.cpp:4380-4429): state, BC, local time step, then
  // calcRhsCoupled into getMem2D / getMem3D arrays as that loop allocates them.
  void rhsCoupledOnce(const Load &st, Dump &d) {
    initializeFromRestartFile(getStringParam("RESTART"));
    initialHook();
    initialHookScalarRansTurbModel();
    initialHookScalarRansCombModel();
    const double *r = st.f64("RHO", ncv);
    const double *ru = st.f64("RHOU", 3LL * ncv);
    const double *re = st.f64("RHOE", ncv);
    for (int icv = 0; icv < ncv; icv++) {
      rho[icv] = r[icv];
      for (int i = 0; i < 3; i++) rhou[icv][i] = ru[3 * icv + i];
      rhoE[icv] = re[icv];
    }
    const int nScal = scalarTranspEqVector.size();
    const int b = 5 + nScal;
    loadScalars(st);
    calcStateVariables();
    calcMaterialProperties();
    setBC();
    calcRansTurbViscMuet();
    double dtmin = calcDt(cfl);
    double ***A;
    getMem3D(&A, 0, nbocv_s - 1, 0, b - 1, 0, b - 1, "harness A", true);
    double **rhs;
    getMem2D(&rhs, 0, ncv - 1, 0, b - 1, "harness rhs", true);
    for (int noc = 0; noc < nbocv_s; noc++)
      for (int i = 0; i < b; i++)
        for (int j = 0; j < b; j++) A[noc][i][j] = 0.0;
    calcRhsCoupled(rhs, A, nScal, true);
    std::vector<double> v(ncv), vu(3LL * ncv), ab((size_t)b * b * nbocv_s);
    for (int icv = 0; icv < ncv; icv++) v[icv] = rhs[icv][0];
    d.f64("RHS_RHO", v.data(), ncv);
    for (int icv = 0; icv < ncv; icv++)
      for (int i = 0; i < 3; i++) vu[3 * icv + i] = rhs[icv][1 + i];
    d.f64("RHS_RHOU", vu.data(), 3LL * ncv);
    for (int icv = 0; icv < ncv; icv++) v[icv] = rhs[icv][4];
    d.f64("RHS_RHOE", v.data(), ncv);
    if (nScal > 0) {  // the transported scalars' rows of the coupled residual and their fields
      for (int s = 0; s < nScal; s++) {
        for (int icv = 0; icv < ncv; icv++) v[icv] = rhs[icv][5 + s];
        const std::string k = key(s);
        d.f64(("RHS_" + k).c_str(), v.data(), ncv);
        ScalarTranspEq &eq = scalarTranspEqVector[s];
        d.f64(k.c_str(), eq.phi, ncv);
        d.f64((k + "_BFA").c_str(), eq.phi_bfa, nfa_b);
        d.f64(("GRAD_" + k).c_str(), &eq.grad_phi[0][0], 3LL * ncv);
        d.f64(("DIFF_" + k).c_str(), eq.diff, nfa);
      }
      d.f64("MUT_FA", mut_fa, nfa);
      dumpModel(d);
    }
    for (int noc = 0; noc < nbocv_s; noc++)
      for (int i = 0; i < b; i++)
        for (int j = 0; j < b; j++) ab[((size_t)noc * b + i) * b + j] = A[noc][i][j];
    d.f64("A", ab.data(), (long long)ab.size());
    dumpPrims(d);
    d.f64("LOCAL_DT", local_dt, ncv);
    d.f64("DT_MIN", &dtmin, 1);
    if (sndOrder) {
      d.f64("GRAD_RHO", &grad_rho[0][0], 3LL * ncv);
      d.f64("GRAD_P", &grad_p[0][0], 3LL * ncv);
    }
    if (sndOrder || mu_ref > 0.0) d.f64("GRAD_U", &grad_u[0][0][0], 9LL * ncv);
    if (mu_ref > 0.0) {
      d.f64("GRAD_ENTHALPY", &grad_enthalpy[0][0], 3LL * ncv);
      d.f64("MUL_FA", mul_fa, nfa);
    }
    freeMem3D(A, 0, nbocv_s - 1, 0, b - 1, 0, b - 1);
    freeMem2D(rhs, 0, ncv - 1, 0, b - 1);
  }

  // One implicit residual evaluation in the order runBackwardEuler uses
  // (JoeWithModels.cpp:943-1005).
  void rhsOnce(const Load &st, Dump &d) {
    initializeFromRestartFile(getStringParam("RESTART"));
    initialHook();
    initialHookScalarRansTurbModel();
    initialHookScalarRansCombModel();
    const double *r = st.f64("RHO", ncv);
    const double *ru = st.f64("RHOU", 3LL * ncv);
    const double *re = st.f64("RHOE", ncv);
    for (int icv = 0; icv < ncv; icv++) {
      rho[icv] = r[icv];
      for (int i = 0; i < 3; i++) rhou[icv][i] = ru[3 * icv + i];
      rhoE[icv] = re[icv];
    }
    const int nScal = scalarTranspEqVector.size();
    loadScalars(st);
    calcStateVariables();
    calcMaterialProperties();
    setBC();
    calcRansTurbViscMuet();
    double dtmin = calcDt(cfl);

    double *RHSrho = new double[ncv];
    double (*RHSrhou)[3] = new double[ncv][3];
    double *RHSrhoE = new double[ncv];
    double (*A)[5][5] = new double[nbocv_s][5][5];
    for (int noc = 0; noc < nbocv_s; noc++)
      for (int i = 0; i < 5; i++)
        for (int j = 0; j < 5; j++) A[noc][i][j] = 0.0;
    // scalar right-hand sides and matrix rows, laid out as runBackwardEuler's
    // getMem2D / getMem3D arrays ([nScal][ncv], [nScal][6][nbocv_s])
    double **rhsScal = NULL, ***AScal = NULL;
    if (nScal > 0) {
      rhsScal = new double *[nScal];
      AScal = new double **[nScal];
      for (int s = 0; s < nScal; s++) {
        rhsScal[s] = new double[ncv];
        AScal[s] = new double *[6];
        for (int k = 0; k < 6; k++) {
          AScal[s][k] = new double[nbocv_s];
          for (int noc = 0; noc < nbocv_s; noc++) AScal[s][k][noc] = 0.0;
        }
      }
    }

    auto t0 = std::chrono::steady_clock::now();
    calcRhs(RHSrho, RHSrhou, RHSrhoE, rhsScal, A, AScal, true);
    auto t1 = std::chrono::steady_clock::now();
    double secs = std::chrono::duration<double>(t1 - t0).count();
    if (nScal > 0) {
      std::vector<double> as(6LL * nbocv_s * nScal);
      for (int s = 0; s < nScal; s++) {
        ScalarTranspEq &eq = scalarTranspEqVector[s];
        const std::string k = key(s);
        d.f64(k.c_str(), eq.phi, ncv);
        d.f64((k + "_BFA").c_str(), eq.phi_bfa, nfa_b);
        d.f64(("GRAD_" + k).c_str(), &eq.grad_phi[0][0], 3LL * ncv);
        d.f64(("DIFF_" + k).c_str(), eq.diff, nfa);
        d.f64(("RHS_" + k).c_str(), rhsScal[s], ncv);
        for (int kk = 0; kk < 6; kk++)
          for (int noc = 0; noc < nbocv_s; noc++) as[((size_t)s * 6 + kk) * nbocv_s + noc] = AScal[s][kk][noc];
      }
      d.f64("MUT_FA", mut_fa, nfa);
      d.f64("ASCAL", as.data(), (long long)as.size());
      dumpModel(d);
    }

    dumpPrims(d);
    d.f64("LOCAL_DT", local_dt, ncv);
    d.f64("DT_MIN", &dtmin, 1);
    d.f64("RHS_RHO", RHSrho, ncv);
    d.f64("RHS_RHOU", &RHSrhou[0][0], 3LL * ncv);
    d.f64("RHS_RHOE", RHSrhoE, ncv);
    d.f64("A", &A[0][0][0], 25LL * nbocv_s);
    if (mu_ref > 0.0) d.f64("GRAD_ENTHALPY", &grad_enthalpy[0][0], 3LL * ncv);
    if (mu_ref > 0.0 && !sndOrder) d.f64("GRAD_U", &grad_u[0][0][0], 9LL * ncv);
    if (sndOrder) {
      d.f64("GRAD_RHO", &grad_rho[0][0], 3LL * ncv);
      d.f64("GRAD_P", &grad_p[0][0], 3LL * ncv);
      d.f64("GRAD_U", &grad_u[0][0][0], 9LL * ncv);
      d.f64("ALPHA_RHO", alpha_rho, ncv);
    }
    d.f64("CALCRHS_SECONDS", &secs, 1);

    // One linear solve of the first backward-Euler step on this state
    // (JoeWithModels.cpp:1011-1026), so the Krylov path is pinned too.
    double underRelax = getDoubleParam("UNDER_RELAXATION", "0.3");
    double (*rhs)[5] = new double[ncv][5];
    double (*dq)[5] = new double[ncv_g][5];
    for (int icv = 0; icv < ncv; icv++) {
      rhs[icv][0] = underRelax * RHSrho[icv];
      for (int i = 0; i < 3; i++) rhs[icv][1 + i] = underRelax * RHSrhou[icv][i];
      rhs[icv][4] = underRelax * RHSrhoE[icv];
      double tmp = cv_volume[icv] / (local_dt[icv]);
      for (int i = 0; i < 5; i++) A[nbocv_i[icv]][i][i] += tmp;
    }
    for (int icv = 0; icv < ncv_g; icv++)
      for (int i = 0; i < 5; i++) dq[icv][i] = 0.0;
    if (!checkParam("LINEAR_SOLVER_NS_TRESHOLDS"))
      ParamMap::add("LINEAR_SOLVER_NS_TRESHOLDS  MAX_ITER=30  ABS_RESID=1.0e-8  REL_RESID=1.0e-2");
    int maxIterLS = getParam("LINEAR_SOLVER_NS_TRESHOLDS")->getInt("MAX_ITER");
    double zeroAbsLS = getParam("LINEAR_SOLVER_NS_TRESHOLDS")->getDouble("ABS_RESID");
    double zeroRelLS = getParam("LINEAR_SOLVER_NS_TRESHOLDS")->getDouble("REL_RESID");
    solveCoupledLinSysNS(dq, A, rhs, zeroAbsLS, zeroRelLS, maxIterLS);
    d.f64("DQ", &dq[0][0], 5LL * ncv);
    if (nScal > 0) {
      // the scalar half of that step (JoeWithModels.cpp:1054-1082), without the update
      double underRelaxScalars = getDoubleParam("UNDER_RELAXATION_SCALARS", "0.2");
      std::vector<double> ds((size_t)nScal * ncv);
      double *dScal = new double[ncv_g];
      for (int s = 0; s < nScal; s++) {
        ScalarTranspEq &eq = scalarTranspEqVector[s];
        for (int icv = 0; icv < ncv_g; icv++) dScal[icv] = 0.0;
        for (int icv = 0; icv < ncv; ++icv) {
          rhsScal[s][icv] *= underRelaxScalars;
          int noc_f = nbocv_i[icv];
          int noc_l = nbocv_i[icv + 1] - 1;
          double tmp = cv_volume[icv] / (local_dt[icv]);
          AScal[s][5][noc_f] += tmp;
          for (int noc = noc_f; noc <= noc_l; noc++)
            rhsScal[s][icv] = rhsScal[s][icv] - AScal[s][0][noc] * dq[nbocv_v[noc]][0] -
                              AScal[s][1][noc] * dq[nbocv_v[noc]][1] - AScal[s][2][noc] * dq[nbocv_v[noc]][2] -
                              AScal[s][3][noc] * dq[nbocv_v[noc]][3] - AScal[s][4][noc] * dq[nbocv_v[noc]][4];
        }
        solveLinSysScalar(dScal, AScal[s][5], rhsScal[s], eq.phiZero, eq.phiZeroRel, eq.phiMaxiter, eq.getName());
        for (int icv = 0; icv < ncv; icv++) ds[(size_t)s * ncv + icv] = dScal[icv];
      }
      d.f64("DSCAL", ds.data(), (long long)ds.size());
      delete[] dScal;
    }

    delete[] RHSrho; delete[] RHSrhou; delete[] RHSrhoE; delete[] A;
    delete[] rhs; delete[] dq;
  }
};

// the reference's own Spalart-Allmaras case class (joe.cpp:286-296)
class RefJoeSA : public RefJoe, public RansTurbSA {
 public:
  explicit RefJoeSA(char *name) : RefJoe(name), UgpWithCvCompFlow(name) {}
  void dumpModel(Dump &d) {
    d.f64("WALL_DIST", wallDist, ncv);
    d.f64("STR_MAG", strMag, ncv);
    d.f64("VORT_MAG", vortMag, ncv);
  }
};

// the reference's own SST case class (joe.cpp:301-310)
class RefJoeSST : public RefJoe, public RansTurbKOmSST {
 public:
  explicit RefJoeSST(char *name) : RefJoe(name), UgpWithCvCompFlow(name) {}
  void dumpModel(Dump &d) {
    d.f64("WALL_DIST", wallDist, ncv);
    d.f64("STR_MAG", strMag, ncv);
    d.f64("VORT_MAG", vortMag, ncv);
    d.f64("DIVERG", diverg, ncv);
    d.f64("CROSS_DIFF", crossDiff, ncv);
    d.f64("F1", blendFuncF1, ncv);
    d.f64("F2", blendFuncF2, ncv);
  }
};

}  // namespace

int main(int argc, char **argv) {
  MPI_Init(&argc, &argv);
  initMpiStuff();
  if (argc < 4) {
    fprintf(stderr, "usage: %s init|rhs|solve Joe.in ...\n", argv[0]);
    return 2;
  }
  std::string mode = argv[1];
  const bool with_sa = mode.size() > 3 && mode.compare(mode.size() - 3, 3, "-sa") == 0;
  if (with_sa) mode.resize(mode.size() - 3);
  const bool with_sst = mode.size() > 4 && mode.compare(mode.size() - 4, 4, "-sst") == 0;
  if (with_sst) mode.resize(mode.size() - 4);
  try {
    RefJoe *jp = with_sa ? static_cast<RefJoe *>(new RefJoeSA(argv[2]))
                         : (with_sst ? static_cast<RefJoe *>(new RefJoeSST(argv[2])) : new RefJoe(argv[2]));
    RefJoe &joe = *jp;
    if (mode == "init") {
      joe.initializeFromRestartFile(joe.getStringParam("RESTART"));
      Dump d(argv[3]);
      joe.dumpMesh(d);
    } else if (mode == "rhs") {
      if (argc < 5) return 2;
      Load st(argv[3]);
      Dump d(argv[4]);
      joe.rhsOnce(st, d);
      joe.dumpMesh(d);
    } else if (mode == "rhs-coupled") {
      if (argc < 5) return 2;
      Load st(argv[3]);
      Dump d(argv[4]);
      joe.rhsCoupledOnce(st, d);
      joe.dumpMesh(d);
    } else if (mode == "restart-write") {
      if (argc < 5) return 2;
      Load st(argv[3]);
      joe.writeState(st, argv[4]);
    } else if (mode == "restart-read") {
      Dump d(argv[3]);
      joe.readState(d);
    } else if (mode == "time") {
      Load st(argv[3]);
      printf("SECONDS_PER_RESIDUAL %.9e\n", joe.timeRhs(st, atoi(argv[4])));
    } else if (mode == "solve") {
      Dump d(argv[3]);
      joe.snap = &d;
      joe.nsnap = argc > 4 ? atoi(argv[4]) : 0;
      auto t0 = std::chrono::steady_clock::now();
      joe.run();
      auto t1 = std::chrono::steady_clock::now();
      double secs = std::chrono::duration<double>(t1 - t0).count();
      d.f64("RUN_SECONDS", &secs, 1);
      joe.snap = nullptr;
    } else {
      fprintf(stderr, "unknown mode %s\n", mode.c_str());
      return 2;
    }
    delete jp;
  } catch (int e) {
    fprintf(stderr, "reference threw %d\n", e);
    MPI_Finalize();
    return 1;
  }
  MPI_Finalize();
  return 0;
}
