// C ABI of libjoemi (include/joemi.h): thin wrappers that turn exceptions
// into status codes and a per-thread error string.
#include <cstring>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/joemi.h"
#include "jm_chemtable.hpp"
#include "jm_ctx.hpp"
#include "jm_param.hpp"
#include "jm_mesh.hpp"
#include "jm_server.hpp"

struct jm_mesh {
  jm::Mesh m;
};
struct jm_server {
  std::unique_ptr<jm::Server> s;
};
struct jm_chemtable {
  std::unique_ptr<jm::Chemtable> t;
};
struct jm_ctx {
  std::unique_ptr<jm::Ctx> c;
  const jm::Mesh *mesh = nullptr;
  std::vector<std::vector<double>> hook;  // per zone: [n][5] T u v w p
};

namespace {
thread_local std::string g_err;

template <typename F>
jm_status guard(F &&f) {
  try {
    g_err.clear();
    return f();
  } catch (const jm::Unsupported &e) {
    g_err = e.what();
    return JM_ERR_UNSUPPORTED;
  } catch (const jm::ArgError &e) {
    g_err = e.what();
    return JM_ERR_ARG;
  } catch (const std::invalid_argument &e) {
    g_err = e.what();
    return JM_ERR_PARAM;
  } catch (const std::logic_error &e) {
    g_err = e.what();
    return JM_ERR_STATE;
  } catch (const std::runtime_error &e) {
    g_err = e.what();
    std::string s = e.what();
    if (s.rfind("HIP", 0) == 0 || s.find("HIP device") != std::string::npos) return JM_ERR_HIP;
    return JM_ERR_MESH;
  } catch (const std::exception &e) {
    g_err = e.what();
    return JM_ERR_ARG;
  } catch (...) {
    g_err = "unknown error";
    return JM_ERR_ARG;
  }
}
}  // namespace

extern "C" {

const char *jm_last_error(void) { return g_err.c_str(); }
const char *jm_version(void) { return "joemi 0.1 (gfx950)"; }

jm_status jm_mesh_from_faces(int nno, const double *x_no, int nfa, const int64_t *face_ptr, const int *face_nodes,
                             const int *face_cells, const int *face_zone, int ncv, int nzone, const int *zone_ids,
                             const char *const *zone_names, jm_mesh **out) {
  return guard([&] {
    if (!out || !x_no || !face_ptr || !face_nodes || !face_cells || !face_zone || (nzone > 0 && (!zone_ids || !zone_names)))
      throw std::exception();
    jm::FaceInput in;
    in.nno = nno;
    in.nfa = nfa;
    in.ncv = ncv;
    in.x_no = x_no;
    in.face_ptr = face_ptr;
    in.face_nodes = face_nodes;
    in.face_cells = face_cells;
    in.face_zone = face_zone;
    for (int i = 0; i < nzone; i++) in.zone_names.emplace_back(zone_ids[i], zone_names[i]);
    std::unique_ptr<jm_mesh> m(new jm_mesh);
    jm::build_mesh(in, m->m);
    *out = m.release();
    return JM_OK;
  });
}

jm_status jm_mesh_read_msh(const char *path, jm_mesh **out) {
  return guard([&] {
    if (!path || !out) throw std::exception();
    std::unique_ptr<jm_mesh> m(new jm_mesh);
    jm::read_msh(path, m->m);
    *out = m.release();
    return JM_OK;
  });
}

jm_status jm_mesh_read_restart(const char *path, jm_mesh **out) {
  return guard([&] {
    if (!path || !out) throw std::exception();
    std::unique_ptr<jm_mesh> m(new jm_mesh);
    jm::read_restart(path, m->m);
    *out = m.release();
    return JM_OK;
  });
}

jm_status jm_mesh_restart_field(const jm_mesh *mm, const char *name, double *out, int64_t *count) {
  return guard([&] {
    if (!mm || !name || !count) throw std::exception();
    for (const auto &f : mm->m.restart_cv)
      if (f.first == name) {
        *count = (int64_t)f.second.size();
        if (out) std::memcpy(out, f.second.data(), f.second.size() * sizeof(double));
        return JM_OK;
      }
    throw std::invalid_argument(std::string("no cell field ") + name + " was read from a restart file");
  });
}

jm_status jm_mesh_restart_step(const jm_mesh *mm, int *step, double *time) {
  return guard([&] {
    if (!mm) throw std::exception();
    if (step) *step = mm->m.restart_step;
    if (time) *time = mm->m.restart_time;
    return JM_OK;
  });
}

jm_status jm_write_restart(const jm_mesh *mm, const char *path, int nfield, const char *const *names,
                           const double *const *data, const int *dims, int step, double time) {
  return guard([&] {
    if (!mm || !path || nfield < 0 || (nfield > 0 && (!names || !data || !dims))) throw std::exception();
    std::vector<jm::RestartField> fields;
    for (int k = 0; k < nfield; k++) {
      if (!names[k] || !data[k]) throw std::exception();
      fields.push_back({names[k], data[k], dims[k]});
    }
    jm::write_restart(path, mm->m, fields, step, time);
    return JM_OK;
  });
}

jm_status jm_mesh_from_ugp(const jm_ugp_view *v, jm_mesh **out) {
  return guard([&] {
    if (!v || !out) throw std::exception();
    std::unique_ptr<jm_mesh> mm(new jm_mesh);
    jm::Mesh &m = mm->m;
    m.ncv = v->ncv;
    m.nfa = v->nfa;
    m.nfa_b = v->nfa_b;
    m.cvofa.assign(v->cvofa, v->cvofa + 2 * (size_t)v->nfa);
    m.faocv_i.assign(v->faocv_i, v->faocv_i + v->ncv + 1);
    m.faocv_v.assign(v->faocv_v, v->faocv_v + m.faocv_i[v->ncv]);
    m.nbocv_i.assign(v->nbocv_i, v->nbocv_i + v->ncv + 1);
    m.nbocv_v.assign(v->nbocv_v, v->nbocv_v + m.nbocv_i[v->ncv]);
    m.x_fa.assign(v->x_fa, v->x_fa + 3 * (size_t)v->nfa);
    m.fa_normal.assign(v->fa_normal, v->fa_normal + 3 * (size_t)v->nfa);
    m.x_cv.assign(v->x_cv, v->x_cv + 3 * (size_t)v->ncv);
    m.cv_volume.assign(v->cv_volume, v->cv_volume + v->ncv);
    m.nbocv_lsg.assign(v->nbocv_lsg, v->nbocv_lsg + 3 * (size_t)m.nbocv_i[v->ncv]);
    m.bfa_lsg.assign(v->bfa_lsg, v->bfa_lsg + 3 * (size_t)v->nfa_b);
    for (int i = 0; i < v->nzone; i++) {
      jm::Zone z;
      z.name = v->zone_names[i];
      z.kind = v->zone_kind[i];
      z.ifa_f = v->zone_ifa_f[i];
      z.ifa_l = v->zone_ifa_l[i];
      z.index = i;
      m.zones.push_back(z);
    }
    for (int f = 0; f < m.nfa; f++) {
      int c0 = m.cvofa[2 * (size_t)f], c1 = m.cvofa[2 * (size_t)f + 1];
      if (c0 < 0 || c0 >= m.ncv || c1 >= m.ncv || (f >= m.nfa_b && c1 < 0) || (f < m.nfa_b && c1 >= 0))
        throw std::runtime_error("cvofa inconsistent with nfa_b");
    }
    *out = mm.release();
    return JM_OK;
  });
}

void jm_mesh_destroy(jm_mesh *m) { delete m; }

jm_status jm_mesh_counts(const jm_mesh *mm, int64_t counts[8]) {
  return guard([&] {
    if (!mm || !counts) throw std::exception();
    const jm::Mesh &m = mm->m;
    counts[0] = m.nno;
    counts[1] = m.nfa;
    counts[2] = m.nfa_b;
    counts[3] = m.ncv;
    counts[4] = m.nbocv_s();
    counts[5] = (int64_t)m.zones.size();
    counts[6] = (int64_t)m.faocv_v.size();
    counts[7] = (int64_t)m.noofa_v.size();
    return JM_OK;
  });
}

jm_status jm_mesh_array(const jm_mesh *mm, const char *name, void *out, int64_t *count) {
  return guard([&] {
    if (!mm || !name || !count) throw std::exception();
    const jm::Mesh &m = mm->m;
    std::string n(name);
    const std::vector<int> *vi = nullptr;
    const std::vector<double> *vd = nullptr;
    if (n == "NOOFA_I") vi = &m.noofa_i;
    else if (n == "NOOFA_V") vi = &m.noofa_v;
    else if (n == "CVOFA") vi = &m.cvofa;
    else if (n == "FAOCV_I") vi = &m.faocv_i;
    else if (n == "FAOCV_V") vi = &m.faocv_v;
    else if (n == "NBOCV_I") vi = &m.nbocv_i;
    else if (n == "NBOCV_V") vi = &m.nbocv_v;
    else if (n == "X_NO") vd = &m.x_no;
    else if (n == "X_FA") vd = &m.x_fa;
    else if (n == "FA_NORMAL") vd = &m.fa_normal;
    else if (n == "X_CV") vd = &m.x_cv;
    else if (n == "CV_VOLUME") vd = &m.cv_volume;
    else if (n == "NBOCV_LSG") vd = &m.nbocv_lsg;
    else if (n == "BFA_LSG") vd = &m.bfa_lsg;
    else throw std::invalid_argument(std::string("unknown mesh array ") + name);
    if (vi) {
      *count = (int64_t)vi->size();
      if (out) std::memcpy(out, vi->data(), vi->size() * sizeof(int));
    } else {
      *count = (int64_t)vd->size();
      if (out) std::memcpy(out, vd->data(), vd->size() * sizeof(double));
    }
    return JM_OK;
  });
}

jm_status jm_mesh_zone(const jm_mesh *mm, int i, char *name, int name_len, int *kind, int *ifa_f, int *ifa_l) {
  return guard([&] {
    if (!mm || i < 0 || i >= (int)mm->m.zones.size()) throw std::exception();
    const jm::Zone &z = mm->m.zones[i];
    if (name && name_len > 0) {
      std::strncpy(name, z.name.c_str(), name_len - 1);
      name[name_len - 1] = 0;
    }
    if (kind) *kind = z.kind;
    if (ifa_f) *ifa_f = z.ifa_f;
    if (ifa_l) *ifa_l = z.ifa_l;
    return JM_OK;
  });
}

jm_status jm_ctx_create(const jm_mesh *m, const char *joe_in, int precision, void *stream, jm_ctx **out) {
  return guard([&] {
    if (!m || !joe_in || !out || (precision != JM_FP64 && precision != JM_MIXED)) throw std::exception();
    std::unique_ptr<jm_ctx> c(new jm_ctx);
    c->c.reset(new jm::Ctx(m->m, joe_in, precision, stream));
    c->mesh = &m->m;
    *out = c.release();
    return JM_OK;
  });
}

void jm_ctx_destroy(jm_ctx *c) { delete c; }

jm_status jm_set_state(jm_ctx *c, const double *rho, const double *rhou, const double *rhoE) {
  return guard([&] {
    if (!c || !rho || !rhou || !rhoE) throw std::exception();
    c->c->set_state(rho, rhou, rhoE);
    return JM_OK;
  });
}

jm_status jm_set_step(jm_ctx *c, int step) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->set_step(step);
    return JM_OK;
  });
}

jm_status jm_set_scalar(jm_ctx *c, const char *name, const double *phi) {
  return guard([&] {
    if (!c || !name || !phi) throw std::exception();
    const std::string n(name);
    if (n == "nuSA" || n == "kine") c->c->set_scalar(phi, 0);
    else if (n == "omega") c->c->set_scalar(phi, 1);
    else throw std::invalid_argument("no transported scalar " + n);
    return JM_OK;
  });
}

jm_status jm_get_state(jm_ctx *c, double *rho, double *rhou, double *rhoE) {
  return guard([&] {
    if (!c || !rho || !rhou || !rhoE) throw std::exception();
    c->c->get_state(rho, rhou, rhoE);
    return JM_OK;
  });
}

jm_status jm_init_uniform(jm_ctx *c) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->init_uniform();
    return JM_OK;
  });
}

jm_status jm_calc_rhs(jm_ctx *c, int flags) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->calc_rhs((flags & JM_RHS_IMPLICIT) != 0);
    return JM_OK;
  });
}

jm_status jm_get(jm_ctx *c, const char *name, double *out, int64_t *count) {
  return guard([&] {
    if (!c || !name || !count) throw std::exception();
    *count = c->c->get(name, out);
    return JM_OK;
  });
}

static void fill_info(const jm::StepInfo &s, jm_step_info *info) {
  if (!info) return;
  info->step = s.step;
  info->lin_iters = s.lin_iters;
  info->lin_status = s.lin_status;
  info->cfl = s.cfl;
  info->dt_min = s.dt_min;
  for (int k = 0; k < 5; k++) info->resid[k] = s.resid[k];
  for (int k = 0; k < 5; k++) info->resid_l2[k] = s.resid_l2[k];
  info->resid_scal = s.resid_s;
  info->resid_scal2 = s.resid_s2;
  info->scal_iters = s.scal_iters;
}

jm_status jm_step_begin(jm_ctx *c, int flags) {
  return guard([&] {
    if (!c || (flags & ~JM_HOST_SOURCE)) throw std::exception();
    c->c->step_begin((flags & JM_HOST_SOURCE) != 0);
    return JM_OK;
  });
}

jm_status jm_add_source(jm_ctx *c, const double *rhs_rho, const double *rhs_rhou, const double *rhs_rhoE,
                        const double *a_diag) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->add_source(rhs_rho, rhs_rhou, rhs_rhoE, a_diag);
    return JM_OK;
  });
}

jm_status jm_add_source_coupled(jm_ctx *c, int b, const double *rhs, const double *a_diag) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->add_source_coupled(rhs, a_diag, b);
    return JM_OK;
  });
}

jm_status jm_step_end(jm_ctx *c, jm_step_info *info) {
  return guard([&] {
    if (!c) throw std::exception();
    fill_info(c->c->step_end(), info);
    return JM_OK;
  });
}

jm_status jm_update_bc(jm_ctx *c) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->update_bc();
    return JM_OK;
  });
}

jm_status jm_step_graph(jm_ctx *c, jm_step_info *info) {
  return guard([&] {
    if (!c) throw std::exception();
    fill_info(c->c->device_step(), info);
    return JM_OK;
  });
}

jm_status jm_step(jm_ctx *c, jm_step_info *info) {
  return guard([&] {
    if (!c) throw std::exception();
    fill_info(c->c->step(), info);
    return JM_OK;
  });
}

jm_status jm_run(jm_ctx *c, int max_steps, double *resid_hist, int *steps_done) {
  return guard([&] {
    if (!c || max_steps < 0) throw std::exception();
    int n = c->c->run(max_steps, resid_hist);
    if (steps_done) *steps_done = n;
    return JM_OK;
  });
}

jm_status jm_set_boundary_state(jm_ctx *c, int zone, const double *T, const double *vel, const double *p) {
  return guard([&] {
    if (!c || !T || !vel || !p || zone < 0 || zone >= (int)c->mesh->zones.size()) throw std::exception();
    const jm::Zone &z = c->mesh->zones[zone];
    if (z.kind != jm::kZoneBoundary) throw std::invalid_argument("zone is not a boundary zone");
    c->c->set_boundary_state(z.ifa_f, z.ifa_l - z.ifa_f + 1, T, vel, p);
    return JM_OK;
  });
}

jm_status jm_block_bcgstab(int ncv, const int *nbocv_i, const int *nbocv_v, const double *A, const double *rhs,
                           double *phi, double zero_abs, double zero_rel, int max_iter, void *stream, int *iters,
                           int *status) {
  return guard([&] {
    jm::BlockSolve bs;
    bs.ncv = ncv;
    bs.nbocv_i = nbocv_i;
    bs.nbocv_v = nbocv_v;
    bs.A = A;
    bs.rhs = rhs;
    bs.phi = phi;
    bs.zero_abs = zero_abs;
    bs.zero_rel = zero_rel;
    bs.max_iter = max_iter;
    bs.stream = stream;
    int it = 0, st = 0;
    jm::block_bcgstab(bs, it, st);
    if (iters) *iters = it;
    if (status) *status = st;
    return JM_OK;
  });
}

jm_status jm_block_bcgstab_noalloc(int ncv, const int *nbocv_i, const int *nbocv_v, const double *A,
                                   const double *rhs, double *phi, double zero_abs, double zero_rel, int max_iter,
                                   void *stream, int *iters, int *status) {
  return guard([&] {
    jm::BlockSolve bs;
    bs.ncv = ncv;
    bs.nbocv_i = nbocv_i;
    bs.nbocv_v = nbocv_v;
    bs.A = A;
    bs.rhs = rhs;
    bs.phi = phi;
    bs.zero_abs = zero_abs;
    bs.zero_rel = zero_rel;
    bs.max_iter = max_iter;
    bs.stream = stream;
    int it = 0, st = 0;
    jm::block_bcgstab_noalloc(bs, it, st);
    if (iters) *iters = it;
    if (status) *status = st;
    return JM_OK;
  });
}

int64_t jm_block_bcgstab_noalloc_max_cells(void) { return (int64_t)jm::block_bcgstab_noalloc_max_cells(); }

jm_status jm_chemtable_load(const char *path, jm_chemtable **out) {
  return guard([&] {
    if (!path || !out) throw jm::ArgError("jm_chemtable_load: null argument");
    *out = nullptr;
    auto t = std::make_unique<jm_chemtable>();
    t->t = std::make_unique<jm::Chemtable>(jm::read_chemtable(path));
    *out = t.release();
    return JM_OK;
  });
}

jm_status jm_chemtable_create(int n1, int n2, int n3, int nvar, const double *x1, const double *x2, const double *x3,
                              const float *data, const char *const *names, jm_chemtable **out) {
  return guard([&] {
    if (!out || !x1 || !x2 || !x3 || !data) throw jm::ArgError("jm_chemtable_create: null argument");
    *out = nullptr;
    if (n1 < 1 || n2 < 1 || n3 < 1 || nvar < 1) throw jm::ArgError("jm_chemtable_create: dimensions must be >= 1");
    jm::ChemtableHost h;
    h.n1 = n1, h.n2 = n2, h.n3 = n3, h.nvar = nvar;
    h.x1.assign(x1, x1 + n1);
    h.x2.assign(x2, x2 + n2);
    h.x3.assign(x3, x3 + n3);
    for (int l = 0; l < nvar; l++) h.names.push_back(names ? std::string(names[l]) : std::to_string(l));
    h.data.assign(data, data + (size_t)n1 * n2 * n3 * nvar);
    auto t = std::make_unique<jm_chemtable>();
    t->t = std::make_unique<jm::Chemtable>(std::move(h));
    *out = t.release();
    return JM_OK;
  });
}

void jm_chemtable_destroy(jm_chemtable *t) { delete t; }

jm_status jm_chemtable_info(const jm_chemtable *t, int *dims, double *ref) {
  return guard([&] {
    if (!t) throw jm::ArgError("jm_chemtable_info: null table");
    const auto &h = t->t->meta();
    if (dims) dims[0] = h.n1, dims[1] = h.n2, dims[2] = h.n3, dims[3] = h.nvar;
    if (ref) ref[0] = h.pref, ref[1] = h.toxi, ref[2] = h.tfuel;
    return JM_OK;
  });
}

jm_status jm_chemtable_var(const jm_chemtable *t, const char *name, int *index) {
  return guard([&] {
    if (!t || !name || !index) throw jm::ArgError("jm_chemtable_var: null argument");
    const int l = t->t->var_index(name);
    if (l < 0) throw jm::ArgError(std::string("Unknown variable in chemistry table: ") + name);
    *index = l;
    return JM_OK;
  });
}

jm_status jm_chemtable_lookup(const jm_chemtable *t, int64_t n, const double *zmean_dev, const double *zvar_dev,
                              const double *prog_dev, int nout, const int *vars, double *out_dev, void *stream) {
  return guard([&] {
    if (!t) throw jm::ArgError("jm_chemtable_lookup: null table");
    t->t->lookup(n, zmean_dev, zvar_dev, prog_dev, nout, vars, out_dev, stream);
    return JM_OK;
  });
}

jm_status jm_chemtable_lookup_host(const jm_chemtable *t, int64_t n, const double *zmean, const double *zvar,
                                   const double *prog, int nout, const int *vars, double *out) {
  return guard([&] {
    if (!t) throw jm::ArgError("jm_chemtable_lookup_host: null table");
    if (n > 0 && (!zmean || !zvar || !prog || !out)) throw jm::ArgError("jm_chemtable_lookup_host: null array");
    t->t->lookup_host(n, zmean, zvar, prog, nout, vars, out);
    return JM_OK;
  });
}

jm_status jm_chemtable_axis(const jm_chemtable *t, int d, double *x) {
  return guard([&] {
    if (!t || !x || d < 1 || d > 3) throw jm::ArgError("jm_chemtable_axis: axis 1, 2 or 3");
    const auto &h = t->t->meta();
    const std::vector<double> &a = d == 1 ? h.x1 : (d == 2 ? h.x2 : h.x3);
    std::memcpy(x, a.data(), sizeof(double) * a.size());
    return JM_OK;
  });
}

jm_status jm_chemtable_text(const jm_chemtable *t, int which, char *buf, int len) {
  return guard([&] {
    if (!t || !buf || len < 1) throw jm::ArgError("jm_chemtable_text: null argument");
    const auto &h = t->t->meta();
    if (which < -2 || which >= h.nvar) throw jm::ArgError("jm_chemtable_text: no such entry");
    const std::string &s = which == -2 ? h.model : (which == -1 ? h.type : h.names[which]);
    if ((int)s.size() >= len) throw jm::ArgError("jm_chemtable_text: buffer too small");
    std::memcpy(buf, s.c_str(), s.size() + 1);
    return JM_OK;
  });
}

jm_status jm_chemtable_time_lookup(const jm_chemtable *t, int64_t n, const double *zmean_dev, const double *zvar_dev,
                                   const double *prog_dev, int nout, const int *vars, double *out_dev, void *stream,
                                   int reps, double *ms) {
  return guard([&] {
    if (!t || !ms) throw jm::ArgError("jm_chemtable_time_lookup: null argument");
    *ms = t->t->time_lookup(n, zmean_dev, zvar_dev, prog_dev, nout, vars, out_dev, stream, reps);
    return JM_OK;
  });
}

jm_status jm_wall_values(jm_ctx *c, int zone, double *out, int64_t *count) {
  return guard([&] {
    if (!c || zone < 0 || zone >= (int)c->mesh->zones.size()) throw std::exception();
    const jm::Zone &z = c->mesh->zones[zone];
    if (z.kind != jm::kZoneBoundary) throw std::invalid_argument("zone is not a boundary zone");
    const int n = c->c->wall_values(z.ifa_f, z.ifa_l - z.ifa_f + 1, out);
    if (count) *count = n;
    return JM_OK;
  });
}

jm_status jm_get_boundary_state(jm_ctx *c, int zone, double *T, double *vel, double *p) {
  return guard([&] {
    if (!c || zone < 0 || zone >= (int)c->mesh->zones.size()) throw std::exception();
    const jm::Zone &z = c->mesh->zones[zone];
    if (z.kind != jm::kZoneBoundary) throw std::invalid_argument("zone is not a boundary zone");
    c->c->get_boundary_state(z.ifa_f, z.ifa_l - z.ifa_f + 1, T, vel, p);
    return JM_OK;
  });
}

jm_status jm_device_state(jm_ctx *c, double **q) {
  return guard([&] {
    if (!c || !q) throw std::exception();
    *q = c->c->device_state();
    return JM_OK;
  });
}

jm_status jm_device_rhs(jm_ctx *c, double **r) {
  return guard([&] {
    if (!c || !r) throw std::exception();
    *r = c->c->device_rhs();
    return JM_OK;
  });
}

jm_status jm_sync(jm_ctx *c) {
  return guard([&] {
    if (!c) throw std::exception();
    c->c->sync();
    return JM_OK;
  });
}

jm_status jm_time_residual_kernels(jm_ctx *c, int reps, double *ms_out, char *names, int names_len) {
  return guard([&] {
    if (!c || !ms_out || reps < 1) throw std::exception();
    std::string n = c->c->time_kernels(reps, ms_out);
    if (names && names_len > 0) {
      std::strncpy(names, n.c_str(), names_len - 1);
      names[names_len - 1] = 0;
    }
    return JM_OK;
  });
}

jm_status jm_time_step_kernels(jm_ctx *c, int reps, double *ms_out, char *names, int names_len) {
  return guard([&] {
    if (!c || !ms_out || reps < 1) throw std::exception();
    std::string n = c->c->time_step_kernels(reps, ms_out);
    if (names && names_len > 0) {
      std::strncpy(names, n.c_str(), names_len - 1);
      names[names_len - 1] = 0;
    }
    return JM_OK;
  });
}

jm_status jm_ctx_footprint(const jm_mesh *m, int precision, int64_t *bytes) {
  return guard([&] {
    if (!m || !bytes || (precision != JM_FP64 && precision != JM_MIXED)) throw std::exception();
    *bytes = (int64_t)jm::Ctx::footprint(m->m, precision);
    return JM_OK;
  });
}

jm_status jm_server_create(int64_t mem_budget_bytes, int max_active, int flags, jm_server **out) {
  return guard([&] {
    if (!out || mem_budget_bytes <= 0 || max_active < 1 || (flags & ~(JM_SERVER_BATCHED | JM_SERVER_NO_FP64_RETRY)))
      throw std::exception();
    std::unique_ptr<jm_server> s(new jm_server);
    s->s.reset(new jm::Server((size_t)mem_budget_bytes, max_active, (flags & JM_SERVER_BATCHED) != 0,
                              (flags & JM_SERVER_NO_FP64_RETRY) == 0));
    *out = s.release();
    return JM_OK;
  });
}

void jm_server_destroy(jm_server *s) { delete s; }

jm_status jm_server_submit(jm_server *s, const jm_mesh *m, const char *joe_in, int precision, const double *rho,
                           const double *rhou, const double *rhoE, int *ticket) {
  return guard([&] {
    if (!s || !m || !joe_in || !ticket || (precision != JM_FP64 && precision != JM_MIXED)) throw std::exception();
    if ((rho || rhou || rhoE) && !(rho && rhou && rhoE)) throw std::exception();
    *ticket = s->s->submit(&m->m, joe_in, precision, rho, rhou, rhoE);
    return JM_OK;
  });
}

jm_status jm_server_status(jm_server *s, int ticket, int *status) {
  return guard([&] {
    if (!s || !status) throw std::exception();
    *status = s->s->status(ticket);
    return JM_OK;
  });
}

jm_status jm_server_progress(jm_server *s, int ticket, int *steps) {
  return guard([&] {
    if (!s || !steps) throw std::exception();
    *steps = s->s->progress(ticket);
    return JM_OK;
  });
}

jm_status jm_server_wait(jm_server *s, int ticket) {
  return guard([&] {
    if (!s) throw std::exception();
    s->s->wait(ticket);
    return JM_OK;
  });
}

jm_status jm_server_result(jm_server *s, int ticket, double *rho, double *rhou, double *rhoE, jm_solve_report *rep) {
  return guard([&] {
    if (!s) throw std::exception();
    auto r = s->s->finished(ticket);
    if (rep) {
      rep->status = r->status;
      rep->steps = r->steps.load();
      for (int k = 0; k < 5; k++) {
        rep->resid[k] = r->last.resid[k];
        rep->resid0[k] = r->resid0[k];
      }
      rep->queued_s = r->t_start > 0 ? r->t_start - r->t_submit : r->t_done - r->t_submit;
      rep->run_s = r->t_start > 0 ? r->t_done - r->t_start : 0.0;
      rep->fp64_retries = r->fp64_retries;
      rep->steps_abandoned = r->steps_abandoned;
    }
    if (r->status != jm::Server::kDone) throw std::runtime_error("solve failed: " + r->error);
    const size_t n = r->rho.size();
    if (rho) std::memcpy(rho, r->rho.data(), n * sizeof(double));
    if (rhou) std::memcpy(rhou, r->rhou.data(), 3 * n * sizeof(double));
    if (rhoE) std::memcpy(rhoE, r->rhoE.data(), n * sizeof(double));
    return JM_OK;
  });
}

jm_status jm_server_release(jm_server *s, int ticket) {
  return guard([&] {
    if (!s) throw std::exception();
    s->s->release(ticket);
    return JM_OK;
  });
}

jm_status jm_mesh_quality(const jm_mesh *m, double *min_sdotn, int *neg_subtets) {
  return guard([&] {
    if (!m) throw std::exception();
    if (min_sdotn) *min_sdotn = m->m.min_sdotn;
    if (neg_subtets) *neg_subtets = m->m.count_neg_subtet;
    return JM_OK;
  });
}

}  // extern "C"
