// Host mirror of UgpWithCvCompFlow over the joemi C ABI; see
// UgpWithCvCompFlow.hpp.
#include "UgpWithCvCompFlow.hpp"

#include <fstream>
#include <sstream>

namespace {

// Param.cpp:247-298 tokenizer: separators ' ', '=', tab, CR; '#' starts a comment
std::vector<std::string> tokens(const std::string &line) {
  std::vector<std::string> out;
  std::string cur;
  for (char ch : line) {
    if (ch == '#') break;
    if (ch == ' ' || ch == '=' || ch == '\t' || ch == '\r') {
      if (!cur.empty()) out.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(ch);
    }
  }
  if (!cur.empty()) out.push_back(cur);
  return out;
}

bool ends_with(const std::string &s, const std::string &suf) {
  return s.size() >= suf.size() && s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

}  // namespace

void UgpWithCvCompFlow::require(jm_status st, const char *what) const {
  if (st != JM_OK) {
    std::ostringstream os;
    os << "joemi: " << what << " failed (status " << (int)st << "): " << jm_last_error();
    throw std::runtime_error(os.str());
  }
}

// Joe.in, the gas constants (UgpWithCvCompFlow.h:150-200) and the mesh:
// RESTART names a Fluent .msh (the MshFilter path of UgpWithCv::init) or a
// restart file (Ugp::readRestart), UgpWithCv.h:316-347
void UgpWithCvCompFlow::init(const char *inputFile) {
  if (initialised()) return;
  std::ifstream f(inputFile);
  if (!f) throw std::runtime_error(std::string("cannot open input file ") + inputFile);
  std::stringstream ss;
  ss << f.rdbuf();
  joe_in_ = ss.str();
  std::istringstream is(joe_in_);
  std::string line;
  while (std::getline(is, line)) {
    std::vector<std::string> t = tokens(line);
    if (t.empty()) continue;
    params_[t[0]].push_back(t);
    param_objs_.emplace(t[0], Param(t));  // the first line of a key
  }
  GAMMA = getDoubleParam("GAMMA", "1.4");
  p_ref = getDoubleParam("P_REF");
  rho_ref = getDoubleParam("RHO_REF");
  T_ref = getDoubleParam("T_REF");
  R_gas = p_ref / rho_ref / T_ref;
  Pr = getDoubleParam("PR", "0.72");
  PrTurb = getDoubleParam("PRTURB", "0.9");
  gravity[0] = getDoubleParam("GX", "0.0");
  gravity[1] = getDoubleParam("GY", "0.0");
  gravity[2] = getDoubleParam("GZ", "0.0");
  nsteps = getIntParam("NSTEPS", "100");
  resid_energ_th = getDoubleParam("RESID_ENERG_TH", "0.0");

  const std::string restart = getStringParam("RESTART", "restart.out");
  from_restart_ = !ends_with(restart, ".msh");
  if (from_restart_)
    require(jm_mesh_read_restart(restart.c_str(), &mesh_), "jm_mesh_read_restart");
  else
    require(jm_mesh_read_msh(restart.c_str(), &mesh_), "jm_mesh_read_msh");
  int64_t counts[8];
  require(jm_mesh_counts(mesh_, counts), "jm_mesh_counts");
  nno = (int)counts[0];
  nfa = (int)counts[1];
  nfa_b = (int)counts[2];
  ncv = ncv_g = (int)counts[3];
  nbocv_s = (int)counts[4];
  const int nzone = (int)counts[5];

  auto load_i = [&](const char *name, std::vector<int> &v) {
    int64_t n = 0;
    require(jm_mesh_array(mesh_, name, nullptr, &n), name);
    v.resize(n);
    require(jm_mesh_array(mesh_, name, v.data(), &n), name);
  };
  auto load_d = [&](const char *name, std::vector<double> &v) {
    int64_t n = 0;
    require(jm_mesh_array(mesh_, name, nullptr, &n), name);
    v.resize(n);
    require(jm_mesh_array(mesh_, name, v.data(), &n), name);
  };
  load_i("CVOFA", cvofa_v_);
  load_i("FAOCV_I", faocv_i_v_);
  load_i("FAOCV_V", faocv_v_v_);
  load_i("NBOCV_I", nbocv_i_v_);
  load_i("NBOCV_V", nbocv_v_v_);
  {  // node coordinates, where the mesh source carries them
    int64_t n = 0;
    if (jm_mesh_array(mesh_, "X_NO", nullptr, &n) == JM_OK && n > 0) {
      x_no_v_.resize(n);
      require(jm_mesh_array(mesh_, "X_NO", x_no_v_.data(), &n), "X_NO");
    }
  }
  load_d("X_CV", x_cv_v_);
  load_d("X_FA", x_fa_v_);
  load_d("FA_NORMAL", fa_normal_v_);
  load_d("CV_VOLUME", cv_volume_v_);
  cvofa = reinterpret_cast<int(*)[2]>(cvofa_v_.data());
  faocv_i = faocv_i_v_.data();
  faocv_v = faocv_v_v_.data();
  nbocv_i = nbocv_i_v_.data();
  nbocv_v = nbocv_v_v_.data();
  x_no = x_no_v_.empty() ? nullptr : reinterpret_cast<double(*)[3]>(x_no_v_.data());
  x_cv = reinterpret_cast<double(*)[3]>(x_cv_v_.data());
  x_fa = reinterpret_cast<double(*)[3]>(x_fa_v_.data());
  fa_normal = reinterpret_cast<double(*)[3]>(fa_normal_v_.data());
  cv_volume = cv_volume_v_.data();

  char name[256];
  for (int i = 0; i < nzone; i++) {
    FaZone z;
    require(jm_mesh_zone(mesh_, i, name, sizeof(name), &z.kind_, &z.ifa_f, &z.ifa_l), "jm_mesh_zone");
    z.name_ = name;
    z.index_ = i;
    faZoneList.push_back(z);
  }
  allocHost();
  registerScalar(rho, "RHO");
  registerScalar(rhoE, "RHOE");
  registerScalar(press, "press");
  registerScalar(temp, "temp");
}

UgpWithCvCompFlow::~UgpWithCvCompFlow() {
  if (mesh_) jm_mesh_destroy(mesh_);
}

void UgpWithCvCompFlow::allocHost() {
  const size_t n = ncv;
  rho_v_.assign(n, 0.0);
  rhou_v_.assign(3 * n, 0.0);
  rhoE_v_.assign(n, 0.0);
  vel_v_.assign(3 * n, 0.0);
  press_v_.assign(n, 0.0);
  temp_v_.assign(n, 0.0);
  enthalpy_v_.assign(n, 0.0);
  sos_v_.assign(n, 0.0);
  gamma_v_.assign(n, GAMMA);
  rom_v_.assign(n, R_gas);
  mul_fa_v_.assign(nfa, 0.0);
  grad_u_v_.assign(9 * n, 0.0);
  rij_diag_v_.assign(3 * n, 0.0);
  rij_offdiag_v_.assign(3 * n, 0.0);
  rho = rho_v_.data();
  rhou = reinterpret_cast<double(*)[3]>(rhou_v_.data());
  rhoE = rhoE_v_.data();
  vel = reinterpret_cast<double(*)[3]>(vel_v_.data());
  press = press_v_.data();
  temp = temp_v_.data();
  enthalpy = enthalpy_v_.data();
  sos = sos_v_.data();
  gamma = gamma_v_.data();
  RoM = rom_v_.data();
  mul_fa = mul_fa_v_.data();
  grad_u = reinterpret_cast<double(*)[3][3]>(grad_u_v_.data());
  rij_diag = reinterpret_cast<double(*)[3]>(rij_diag_v_.data());
  rij_offdiag = reinterpret_cast<double(*)[3]>(rij_offdiag_v_.data());
}

void UgpWithCvCompFlow::registerScalar(double *&field, const char *name) {
  for (DoubleScalar &d : registered_)
    if (std::string(d.getName()) == name) return;
  registered_.emplace_back(name, &field);
}

double *UgpWithCvCompFlow::getR1(const char *name) {
  DoubleScalar *d = getScalarData(name);
  return d ? d->getData() : nullptr;
}

DoubleScalar *UgpWithCvCompFlow::getScalarData(const char *name) {
  for (DoubleScalar &d : registered_)
    if (std::string(d.getName()) == name) return &d;
  return nullptr;
}

ScalarTranspEq *UgpWithCvCompFlow::getScalarTransportData(const std::string &name) {
  for (ScalarTranspEq &e : scalarTranspEqVector)
    if (name == e.getName()) return &e;
  return nullptr;
}

int UgpWithCvCompFlow::getScalarTransportIndex(const std::string &name) {
  for (size_t i = 0; i < scalarTranspEqVector.size(); i++)
    if (name == scalarTranspEqVector[i].getName()) return (int)i;
  return -1;
}

double UgpWithCvCompFlow::InterpolateAtCellCenterFromFaceValues(double *phi, int icv) const {
  double sum = 0.0;
  const int f0 = faocv_i[icv], f1 = faocv_i[icv + 1];
  for (int k = f0; k < f1; k++) sum += phi[faocv_v[k]];
  return sum / (double)(f1 - f0);
}

bool UgpWithCvCompFlow::checkParam(const std::string &name) const { return params_.count(name) != 0; }

bool UgpWithCvCompFlow::getParam(Param *&p, const std::string &name) {
  auto it = param_objs_.find(name);
  p = it == param_objs_.end() ? nullptr : &it->second;
  return p != nullptr;
}

std::string UgpWithCvCompFlow::getStringParam(const std::string &name, const std::string &def) const {
  auto it = params_.find(name);
  if (it == params_.end() || it->second[0].size() < 2) {
    if (def.empty()) throw std::runtime_error("missing parameter " + name);
    return def;
  }
  return it->second[0][1];
}

double UgpWithCvCompFlow::getDoubleParam(const std::string &name, const std::string &def) const {
  return std::stod(getStringParam(name, def));
}

double UgpWithCvCompFlow::getDoubleParam(const std::string &name, double def) const {
  return checkParam(name) ? getDoubleParam(name, std::string()) : def;
}

int UgpWithCvCompFlow::getIntParam(const std::string &name, const std::string &def) const {
  return std::stoi(getStringParam(name, def));
}

int UgpWithCvCompFlow::getIntParam(const std::string &name, int def) const {
  return checkParam(name) ? getIntParam(name, std::string()) : def;
}
