// Host mesh pipeline: Joe's serial numbering, geometry and least-squares
// coefficients, rebuilt from the face list of a mesh file.
#include "jm_mesh.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <stdexcept>

namespace jm {

namespace {

[[noreturn]] void fail(const std::string &msg) { throw std::runtime_error(msg); }

}  // namespace

// ---------------------------------------------------------------------------
// Numbering. Mirrors the serial behaviour of Ugp::redistReconnectReorder:
// nodes and cells take the order in which the face stream first touches them
// (the de-duplicated request lists, Ugp.cpp:1760-2120), faces are bucketed by
// sorted zone (boundary zones by name, then internal; Ugp.cpp:3410-3538) and
// cells by "touches a boundary face" (Ugp.cpp:3639-3709), both stable.
// ---------------------------------------------------------------------------
void build_mesh(const FaceInput &in, Mesh &m) {
  const int nfa = in.nfa;
  if (nfa <= 0 || in.ncv <= 0 || in.nno <= 0) fail("empty mesh");

  // zone kinds from icv1 (MshFilter.cpp:684-741)
  std::map<int, int> kind_of;  // zone id -> kind
  for (int ifa = 0; ifa < nfa; ifa++) {
    int z = in.face_zone[ifa];
    int k = in.face_cells[2 * ifa + 1] < 0 ? kZoneBoundary : kZoneInternal;
    auto it = kind_of.find(z);
    if (it == kind_of.end())
      kind_of[z] = k;
    else if (it->second != k)
      fail("zone " + std::to_string(z) + " mixes boundary and internal faces");
  }

  // first-touch renumbering of nodes and cells
  std::vector<int> no_new(in.nno, -1), cv_new(in.ncv, -1);
  int nno = 0, ncv = 0;
  for (int ifa = 0; ifa < nfa; ifa++) {
    for (int64_t k = in.face_ptr[ifa]; k < in.face_ptr[ifa + 1]; k++) {
      int ino = in.face_nodes[k];
      if (ino < 0 || ino >= in.nno) fail("face node out of range");
      if (no_new[ino] < 0) no_new[ino] = nno++;
    }
    int c0 = in.face_cells[2 * ifa], c1 = in.face_cells[2 * ifa + 1];
    if (c0 < 0 || c0 >= in.ncv || c1 >= in.ncv) fail("face cell out of range");
    if (cv_new[c0] < 0) cv_new[c0] = ncv++;
    if (c1 >= 0 && cv_new[c1] < 0) cv_new[c1] = ncv++;
  }
  if (ncv != in.ncv) fail("cells not referenced by any face");

  // zone table sorted by (kind, name) (Ugp.h:155-157; list::sort is stable)
  m.zones.clear();
  for (auto &z : in.zone_names) {
    auto it = kind_of.find(z.first);
    if (it == kind_of.end()) continue;  // zone without faces (e.g. the cell zone)
    Zone zz;
    zz.index = z.first;
    zz.name = z.second;
    zz.kind = it->second;
    m.zones.push_back(zz);
  }
  for (auto &kv : kind_of) {
    bool named = false;
    for (auto &z : m.zones) named |= (z.index == kv.first);
    if (!named) fail("face zone " + std::to_string(kv.first) + " has no name");
  }
  std::stable_sort(m.zones.begin(), m.zones.end(), [](const Zone &a, const Zone &b) {
    return (a.kind < b.kind) || (a.kind == b.kind && std::strcmp(a.name.c_str(), b.name.c_str()) < 0);
  });
  const int nzone = (int)m.zones.size();
  int n_internal = 0;
  std::map<int, int> order_of;
  for (int i = 0; i < nzone; i++) {
    order_of[m.zones[i].index] = i;
    if (m.zones[i].kind == kZoneInternal) n_internal++;
  }

  // face buckets
  std::vector<int> bucket(nfa);
  std::vector<int> fa_count(nzone + n_internal + 1, 0);
  for (int ifa = 0; ifa < nfa; ifa++) {
    int pos = order_of[in.face_zone[ifa]];
    bucket[ifa] = m.zones[pos].kind == kZoneInternal ? pos + n_internal : pos;
    fa_count[bucket[ifa] + 1]++;
  }
  for (int i = 0; i < nzone + n_internal; i++) fa_count[i + 1] += fa_count[i];
  m.nfa_b = -1;
  for (int i = 0; i < nzone; i++) {
    Zone &z = m.zones[i];
    int b = z.kind == kZoneInternal ? i + n_internal : i;
    if (z.kind == kZoneInternal) {
      if (m.nfa_b < 0) m.nfa_b = fa_count[i];
      z.ifa_f = fa_count[i];  // inter-processor bucket, empty on one rank
      z.ifa_l = fa_count[i + 1] - 1;
      z.ifa_i_f = fa_count[b];
      z.ifa_i_l = fa_count[b + 1] - 1;
    } else {
      z.ifa_f = fa_count[b];
      z.ifa_l = fa_count[b + 1] - 1;
    }
  }
  if (m.nfa_b < 0) m.nfa_b = nfa;
  std::vector<int> fa_new(nfa);
  {
    std::vector<int> next(fa_count.begin(), fa_count.end() - 1);
    for (int ifa = 0; ifa < nfa; ifa++) fa_new[ifa] = next[bucket[ifa]]++;
  }

  // connectivity in first-touch node/cell numbering, new face order
  m.nfa = nfa;
  m.nno = nno;
  m.x_no.assign(3 * (size_t)nno, 0.0);
  for (int ino = 0; ino < in.nno; ino++)
    if (no_new[ino] >= 0)
      for (int i = 0; i < 3; i++) m.x_no[3 * (size_t)no_new[ino] + i] = in.x_no[3 * (size_t)ino + i];
  m.noofa_i.assign(nfa + 1, 0);
  for (int ifa = 0; ifa < nfa; ifa++)
    m.noofa_i[fa_new[ifa] + 1] = (int)(in.face_ptr[ifa + 1] - in.face_ptr[ifa]);
  for (int ifa = 0; ifa < nfa; ifa++) m.noofa_i[ifa + 1] += m.noofa_i[ifa];
  m.noofa_v.resize(m.noofa_i[nfa]);
  std::vector<int> cvofa_tmp(2 * (size_t)nfa);
  for (int ifa = 0; ifa < nfa; ifa++) {
    int f = fa_new[ifa];
    int nof = m.noofa_i[f];
    for (int64_t k = in.face_ptr[ifa]; k < in.face_ptr[ifa + 1]; k++) m.noofa_v[nof++] = no_new[in.face_nodes[k]];
    int c0 = in.face_cells[2 * ifa], c1 = in.face_cells[2 * ifa + 1];
    cvofa_tmp[2 * f] = cv_new[c0];
    cvofa_tmp[2 * f + 1] = c1 >= 0 ? cv_new[c1] : -1;
  }

  // cell buckets: cells of boundary faces first (Ugp.cpp:3639-3709)
  std::vector<int> cv_flag(ncv, -1);
  for (int ifa = 0; ifa < m.nfa_b; ifa++) cv_flag[cvofa_tmp[2 * ifa]] = 0;
  for (int ifa = m.nfa_b; ifa < nfa; ifa++) {
    if (cv_flag[cvofa_tmp[2 * ifa]] == -1) cv_flag[cvofa_tmp[2 * ifa]] = 2;
    if (cv_flag[cvofa_tmp[2 * ifa + 1]] == -1) cv_flag[cvofa_tmp[2 * ifa + 1]] = 2;
  }
  int cv_counts[4] = {0, 0, 0, 0};
  for (int icv = 0; icv < ncv; icv++) {
    if (cv_flag[icv] < 0) fail("cell without faces");
    cv_counts[cv_flag[icv] + 1]++;
  }
  for (int i = 0; i < 3; i++) cv_counts[i + 1] += cv_counts[i];
  for (int icv = 0; icv < ncv; icv++) {
    int i = cv_flag[icv];
    cv_flag[icv] = cv_counts[i]++;
  }
  m.ncv = ncv;
  m.cv_of_input.resize(in.ncv);
  for (int c = 0; c < in.ncv; c++) m.cv_of_input[c] = cv_flag[cv_new[c]];
  m.cvofa.resize(2 * (size_t)nfa);
  for (int ifa = 0; ifa < nfa; ifa++) {
    m.cvofa[2 * ifa] = cv_flag[cvofa_tmp[2 * ifa]];
    int c1 = cvofa_tmp[2 * ifa + 1];
    m.cvofa[2 * ifa + 1] = c1 >= 0 ? cv_flag[c1] : -1;
  }

  // faocv: faces of each cell in ascending face index (Ugp.cpp:3541-3561)
  m.faocv_i.assign(ncv + 1, 0);
  for (int ifa = 0; ifa < nfa; ifa++) {
    m.faocv_i[m.cvofa[2 * ifa] + 1]++;
    if (m.cvofa[2 * ifa + 1] >= 0) m.faocv_i[m.cvofa[2 * ifa + 1] + 1]++;
  }
  for (int icv = 0; icv < ncv; icv++) m.faocv_i[icv + 1] += m.faocv_i[icv];
  m.faocv_v.resize(m.faocv_i[ncv]);
  {
    std::vector<int> next(m.faocv_i.begin(), m.faocv_i.end() - 1);
    for (int ifa = 0; ifa < nfa; ifa++) {  // ascending ifa => rows already sorted
      m.faocv_v[next[m.cvofa[2 * ifa]]++] = ifa;
      if (m.cvofa[2 * ifa + 1] >= 0) m.faocv_v[next[m.cvofa[2 * ifa + 1]]++] = ifa;
    }
  }

  // canonical face order of hex/prism/pyramid/tet cells (Ugp.cpp:5584-5953)
  {
    const int NNOC_MAX = 64, NFOC_MAX = 64;
    std::vector<int> no_flag(nno, -1);
    int nno_cv;
    int noofa_i_cv[NFOC_MAX + 1];
    std::vector<int> noofa_v_cv(8 * NFOC_MAX);
    int no_flag_cv[NNOC_MAX], fa_flag_cv[NFOC_MAX], ifa_of_ifa_cv[NFOC_MAX];
    auto quad_match = [&](int &a, int &b, int &c, int &d, int nf) -> int {
      for (int f = 0; f < nf; f++) {
        int nf0 = noofa_i_cv[f], nl = noofa_i_cv[f + 1] - 1;
        if (nl - nf0 + 1 != 4) continue;
        for (int n0 = nf0; n0 <= nl; n0++) {
          if (a != -1 && a != noofa_v_cv[n0]) continue;
          int n1 = n0 + 1 > nl ? nf0 : n0 + 1;
          if (b != -1 && b != noofa_v_cv[n1]) continue;
          int n2 = n1 + 1 > nl ? nf0 : n1 + 1;
          if (c != -1 && c != noofa_v_cv[n2]) continue;
          int n3 = n2 + 1 > nl ? nf0 : n2 + 1;
          if (d != -1 && d != noofa_v_cv[n3]) continue;
          a = noofa_v_cv[n0]; b = noofa_v_cv[n1]; c = noofa_v_cv[n2]; d = noofa_v_cv[n3];
          return f;
        }
      }
      return -1;
    };
    auto tri_match = [&](int &a, int &b, int &c, int nf) -> int {
      for (int f = 0; f < nf; f++) {
        int nf0 = noofa_i_cv[f], nl = noofa_i_cv[f + 1] - 1;
        if (nl - nf0 + 1 != 3) continue;
        for (int n0 = nf0; n0 <= nl; n0++) {
          if (a != -1 && a != noofa_v_cv[n0]) continue;
          int n1 = n0 + 1 > nl ? nf0 : n0 + 1;
          if (b != -1 && b != noofa_v_cv[n1]) continue;
          int n2 = n1 + 1 > nl ? nf0 : n1 + 1;
          if (c != -1 && c != noofa_v_cv[n2]) continue;
          a = noofa_v_cv[n0]; b = noofa_v_cv[n1]; c = noofa_v_cv[n2];
          return f;
        }
      }
      return -1;
    };
    auto setf = [&](int f, int slot) {
      if (f < 0 || fa_flag_cv[f] != -1) fail("element face matching failed");
      fa_flag_cv[f] = slot;
    };
    for (int icv = 0; icv < ncv; icv++) {
      const int foc_f = m.faocv_i[icv], foc_l = m.faocv_i[icv + 1] - 1;
      const int nfa_cv = foc_l - foc_f + 1;
      if (nfa_cv > NFOC_MAX) fail("too many faces on a cell");
      // noocv order: first appearance walking faocv and noofa (Ugp.cpp:5520-5582)
      nno_cv = 0;
      std::vector<int> ino_of;  // local -> global
      int nof_cv = 0;
      noofa_i_cv[0] = 0;
      for (int foc = foc_f; foc <= foc_l; foc++) {
        int ifa = m.faocv_v[foc];
        for (int nof = m.noofa_i[ifa]; nof < m.noofa_i[ifa + 1]; nof++) {
          int ino = m.noofa_v[nof];
          if (no_flag[ino] == -1) {
            no_flag[ino] = nno_cv++;
            ino_of.push_back(ino);
          }
        }
      }
      if (nno_cv > NNOC_MAX) fail("too many nodes on a cell");
      for (int foc = foc_f; foc <= foc_l; foc++) {
        int ifa = m.faocv_v[foc];
        int k = foc - foc_f;
        noofa_i_cv[k + 1] = noofa_i_cv[k] + m.noofa_i[ifa + 1] - m.noofa_i[ifa];
        if (m.cvofa[2 * ifa] == icv) {
          for (int nof = m.noofa_i[ifa]; nof < m.noofa_i[ifa + 1]; nof++) noofa_v_cv[nof_cv++] = no_flag[m.noofa_v[nof]];
        } else {
          for (int nof = m.noofa_i[ifa + 1] - 1; nof >= m.noofa_i[ifa]; nof--) noofa_v_cv[nof_cv++] = no_flag[m.noofa_v[nof]];
        }
        ifa_of_ifa_cv[k] = ifa;
      }
      for (int i = 0; i < nno_cv; i++) no_flag_cv[i] = -1;
      for (int i = 0; i < nfa_cv; i++) fa_flag_cv[i] = -1;
      int tri = 0, quad = 0;
      for (int f = 0; f < nfa_cv; f++) {
        int nn = noofa_i_cv[f + 1] - noofa_i_cv[f];
        tri += nn == 3;
        quad += nn == 4;
      }
      int *v = no_flag_cv;
      if (nfa_cv == 6 && nno_cv == 8 && quad == 6 && tri == 0) {
        setf(quad_match(v[0], v[3], v[2], v[1], nfa_cv), 4);
        setf(quad_match(v[0], v[1], v[5], v[4], nfa_cv), 2);
        setf(quad_match(v[1], v[2], v[6], v[5], nfa_cv), 1);
        setf(quad_match(v[2], v[3], v[7], v[6], nfa_cv), 3);
        setf(quad_match(v[3], v[0], v[4], v[7], nfa_cv), 0);
        setf(quad_match(v[4], v[5], v[6], v[7], nfa_cv), 5);
      } else if (nfa_cv == 5 && nno_cv == 6 && quad == 3 && tri == 2) {
        setf(tri_match(v[0], v[2], v[1], nfa_cv), 3);
        setf(quad_match(v[0], v[1], v[4], v[3], nfa_cv), 0);
        setf(quad_match(v[1], v[2], v[5], v[4], nfa_cv), 1);
        setf(quad_match(v[2], v[0], v[3], v[5], nfa_cv), 2);
        setf(tri_match(v[3], v[4], v[5], nfa_cv), 4);
      } else if (nfa_cv == 5 && nno_cv == 5 && quad == 1 && tri == 4) {
        setf(quad_match(v[0], v[3], v[2], v[1], nfa_cv), 4);
        setf(tri_match(v[0], v[1], v[4], nfa_cv), 0);
        setf(tri_match(v[1], v[2], v[4], nfa_cv), 1);
        setf(tri_match(v[2], v[3], v[4], nfa_cv), 2);
        setf(tri_match(v[3], v[0], v[4], nfa_cv), 3);
      } else if (nfa_cv == 4 && nno_cv == 4 && quad == 0 && tri == 4) {
        setf(tri_match(v[0], v[2], v[1], nfa_cv), 3);
        setf(tri_match(v[0], v[1], v[3], nfa_cv), 2);
        setf(tri_match(v[1], v[2], v[3], nfa_cv), 0);
        setf(tri_match(v[2], v[0], v[3], nfa_cv), 1);
      } else {
        for (int f = 0; f < nfa_cv; f++) fa_flag_cv[f] = f;
      }
      for (int k = 0; k < nfa_cv; k++) m.faocv_v[foc_f + fa_flag_cv[k]] = ifa_of_ifa_cv[k];
      for (int ino : ino_of) no_flag[ino] = -1;
    }
  }

  build_nbocv(m);
  calc_geometry(m);
  calc_lsg_coeff(m);
}

// nbocv: diagonal first, then neighbours in ascending internal face order
// (UgpWithCv.cpp:768-834).
void build_nbocv(Mesh &m) {
  const int ncv = m.ncv;
  m.nbocv_i.assign(ncv + 1, 0);
  for (int icv = 0; icv < ncv; icv++) m.nbocv_i[icv + 1] = 1;
  for (int ifa = m.nfa_b; ifa < m.nfa; ifa++) {
    int c0 = m.cvofa[2 * ifa], c1 = m.cvofa[2 * ifa + 1];
    if (c1 < 0) fail("internal face without second cell");
    m.nbocv_i[c0 + 1]++;
    m.nbocv_i[c1 + 1]++;
  }
  for (int icv = 0; icv < ncv; icv++) m.nbocv_i[icv + 1] += m.nbocv_i[icv];
  m.nbocv_v.resize(m.nbocv_i[ncv]);
  std::vector<int> next(ncv);
  for (int icv = 0; icv < ncv; icv++) {
    m.nbocv_v[m.nbocv_i[icv]] = icv;
    next[icv] = m.nbocv_i[icv] + 1;
  }
  for (int ifa = m.nfa_b; ifa < m.nfa; ifa++) {
    int c0 = m.cvofa[2 * ifa], c1 = m.cvofa[2 * ifa + 1];
    m.nbocv_v[next[c0]++] = c1;
    m.nbocv_v[next[c1]++] = c0;
  }
}

// Face centroids/normals by iterated sub-triangle centroid, cell centroids and
// volumes by sub-tets around the mean of face centroids (UgpWithCv.cpp:842-1035).
void calc_geometry(Mesh &m) {
  const int nfa = m.nfa, ncv = m.ncv;
  const double *x_no = m.x_no.data();
  m.x_fa.assign(3 * (size_t)nfa, 0.0);
  m.fa_normal.assign(3 * (size_t)nfa, 0.0);
  for (int ifa = 0; ifa < nfa; ifa++) {
    double *xf = &m.x_fa[3 * (size_t)ifa];
    double *fn = &m.fa_normal[3 * (size_t)ifa];
    const int nof_f = m.noofa_i[ifa], nof_e = m.noofa_i[ifa + 1];
    for (int i = 0; i < 3; i++) xf[i] = 0.0;
    for (int nof = nof_f; nof < nof_e; nof++)
      for (int i = 0; i < 3; i++) xf[i] += x_no[3 * (size_t)m.noofa_v[nof] + i];
    for (int i = 0; i < 3; i++) xf[i] /= (double)(nof_e - nof_f);
    for (int i = 0; i < 3; i++) fn[i] = 0.0;
    int ino2 = m.noofa_v[nof_e - 1];
    for (int nof = nof_f; nof < nof_e; nof++) {
      int ino1 = ino2;
      ino2 = m.noofa_v[nof];
      double v1[3], v2[3];
      for (int i = 0; i < 3; i++) {
        v1[i] = x_no[3 * (size_t)ino1 + i] - xf[i];
        v2[i] = x_no[3 * (size_t)ino2 + i] - xf[i];
      }
      fn[0] += 0.5 * (v1[1] * v2[2] - v1[2] * v2[1]);
      fn[1] += 0.5 * (v1[2] * v2[0] - v1[0] * v2[2]);
      fn[2] += 0.5 * (v1[0] * v2[1] - v1[1] * v2[0]);
    }
    double fa_area = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
    double un[3];
    for (int i = 0; i < 3; i++) un[i] = fn[i] / fa_area;
    bool done = false;
    int iter = 0;
    while (!done) {
      double dx_fa[3] = {0.0, 0.0, 0.0}, x_tmp[3] = {0.0, 0.0, 0.0};
      double area_sum = 0.0;
      int jn2 = m.noofa_v[nof_e - 1];
      for (int nof = nof_f; nof < nof_e; nof++) {
        int jn1 = jn2;
        jn2 = m.noofa_v[nof];
        double v1[3], v2[3];
        for (int i = 0; i < 3; i++) {
          v1[i] = x_no[3 * (size_t)jn1 + i] - xf[i];
          v2[i] = x_no[3 * (size_t)jn2 + i] - xf[i];
        }
        double tn[3];
        tn[0] = (v1[1] * v2[2] - v1[2] * v2[1]);
        tn[1] = (v1[2] * v2[0] - v1[0] * v2[2]);
        tn[2] = (v1[0] * v2[1] - v1[1] * v2[0]);
        double ta = tn[0] * un[0] + tn[1] * un[1] + tn[2] * un[2];
        for (int i = 0; i < 3; i++) x_tmp[i] += ta * (x_no[3 * (size_t)jn1 + i] + x_no[3 * (size_t)jn2 + i] + xf[i]);
        for (int i = 0; i < 3; i++) dx_fa[i] += ta * (v1[i] + v2[i]);
        area_sum += ta;
      }
      double d2 = 0.0;
      for (int i = 0; i < 3; i++) {
        x_tmp[i] /= 3.0 * area_sum;
        dx_fa[i] /= 2.0 * area_sum;
        xf[i] += dx_fa[i];
        d2 += dx_fa[i] * dx_fa[i];
      }
      for (int i = 0; i < 3; i++) xf[i] = x_tmp[i];
      if (++iter > 200) fail("face centroid did not converge");
      done = (d2 / area_sum < 1.0E-15);
    }
  }

  m.x_cv.assign(3 * (size_t)ncv, 0.0);
  m.cv_volume.assign(ncv, 0.0);
  m.count_neg_subtet = 0;
  for (int icv = 0; icv < ncv; icv++) {
    double xa[3] = {0.0, 0.0, 0.0};
    for (int foc = m.faocv_i[icv]; foc < m.faocv_i[icv + 1]; foc++)
      for (int i = 0; i < 3; i++) xa[i] += m.x_fa[3 * (size_t)m.faocv_v[foc] + i];
    for (int i = 0; i < 3; i++) xa[i] /= (double)(m.faocv_i[icv + 1] - m.faocv_i[icv]);
    double vol = 0.0, xc[3] = {0.0, 0.0, 0.0};
    for (int foc = m.faocv_i[icv]; foc < m.faocv_i[icv + 1]; foc++) {
      int ifa = m.faocv_v[foc];
      const double *xf = &m.x_fa[3 * (size_t)ifa];
      double v1[3];
      for (int i = 0; i < 3; i++) v1[i] = xf[i] - xa[i];
      const int nof_f = m.noofa_i[ifa], nof_e = m.noofa_i[ifa + 1];
      auto subtet = [&](int ino1, int ino2) {
        double v2[3], v3[3];
        for (int i = 0; i < 3; i++) {
          v2[i] = x_no[3 * (size_t)ino1 + i] - xa[i];
          v3[i] = x_no[3 * (size_t)ino2 + i] - xa[i];
        }
        double tv = v1[0] * (v2[1] * v3[2] - v2[2] * v3[1]) + v1[1] * (v2[2] * v3[0] - v2[0] * v3[2]) +
                    v1[2] * (v2[0] * v3[1] - v2[1] * v3[0]);
        if (tv <= 0.0) m.count_neg_subtet++;
        vol += tv;
        for (int i = 0; i < 3; i++) xc[i] += tv * (xa[i] + xf[i] + x_no[3 * (size_t)ino1 + i] + x_no[3 * (size_t)ino2 + i]);
      };
      if (m.cvofa[2 * ifa] == icv) {
        int ino2 = m.noofa_v[nof_e - 1];
        for (int nof = nof_f; nof < nof_e; nof++) {
          int ino1 = ino2;
          ino2 = m.noofa_v[nof];
          subtet(ino1, ino2);
        }
      } else {
        int ino2 = m.noofa_v[nof_f];
        for (int nof = nof_e - 1; nof >= nof_f; nof--) {
          int ino1 = ino2;
          ino2 = m.noofa_v[nof];
          subtet(ino1, ino2);
        }
      }
    }
    for (int i = 0; i < 3; i++) m.x_cv[3 * (size_t)icv + i] = xc[i] / (4.0 * vol);
    m.cv_volume[icv] = vol / 6.0;
  }

  // s.n check of the reference (UgpWithCv.cpp:1064-1085)
  m.min_sdotn = 1.0;
  for (int ifa = m.nfa_b; ifa < nfa; ifa++) {
    int c0 = m.cvofa[2 * ifa], c1 = m.cvofa[2 * ifa + 1];
    double s[3], smag = 0.0, nmag = 0.0;
    for (int i = 0; i < 3; i++) {
      s[i] = m.x_cv[3 * (size_t)c1 + i] - m.x_cv[3 * (size_t)c0 + i];
      smag += s[i] * s[i];
      nmag += m.fa_normal[3 * (size_t)ifa + i] * m.fa_normal[3 * (size_t)ifa + i];
    }
    smag = std::sqrt(smag);
    nmag = std::sqrt(nmag);
    double dp = 0.0;
    for (int i = 0; i < 3; i++) dp += (s[i] / smag) * (m.fa_normal[3 * (size_t)ifa + i] / nmag);
    m.min_sdotn = std::min(m.min_sdotn, dp);
  }
}

// Weighted (1/|dx|^2) least-squares gradient coefficients with Dirichlet
// boundary faces (UgpWithCv.cpp:1783-1913).
void calc_lsg_coeff(Mesh &m) {
  const int ncv = m.ncv;
  const double *x_cv = m.x_cv.data();
  m.nbocv_lsg.assign(3 * (size_t)m.nbocv_s(), 0.0);
  m.bfa_lsg.assign(3 * (size_t)m.nfa_b, 0.0);
  for (int icv = 0; icv < ncv; icv++) {
    double sxx = 0.0, syy = 0.0, szz = 0.0, sxy = 0.0, sxz = 0.0, syz = 0.0;
    const int noc_f = m.nbocv_i[icv], noc_l = m.nbocv_i[icv + 1] - 1;
    auto acc = [&](const double *dx) {
      double w = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      sxx += w * dx[0] * dx[0];
      syy += w * dx[1] * dx[1];
      szz += w * dx[2] * dx[2];
      sxy += w * dx[0] * dx[1];
      sxz += w * dx[0] * dx[2];
      syz += w * dx[1] * dx[2];
    };
    for (int noc = noc_f + 1; noc <= noc_l; noc++) {
      int nb = m.nbocv_v[noc];
      double dx[3];
      for (int i = 0; i < 3; i++) dx[i] = x_cv[3 * (size_t)nb + i] - x_cv[3 * (size_t)icv + i];
      acc(dx);
    }
    const int foc_f = m.faocv_i[icv], foc_l = m.faocv_i[icv + 1] - 1;
    for (int foc = foc_f; foc <= foc_l; foc++) {
      int ifa = m.faocv_v[foc];
      if (ifa < m.nfa_b) {
        double dx[3];
        for (int i = 0; i < 3; i++) dx[i] = m.x_fa[3 * (size_t)ifa + i] - x_cv[3 * (size_t)icv + i];
        acc(dx);
      }
    }
    double denom = 2.0 * sxy * sxz * syz + sxx * syy * szz - sxx * syz * syz - syy * sxz * sxz - szz * sxy * sxy;
    if (!(std::fabs(denom) > 1.0E-12 * m.cv_volume[icv] * m.cv_volume[icv]))
      fail("singular least-squares system at cell " + std::to_string(icv));
    auto coef = [&](const double *dx, double *c) {
      double w = 1.0 / (dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
      c[0] = w * ((syy * szz - syz * syz) * dx[0] + (sxz * syz - sxy * szz) * dx[1] + (sxy * syz - sxz * syy) * dx[2]) / denom;
      c[1] = w * ((sxz * syz - sxy * szz) * dx[0] + (sxx * szz - sxz * sxz) * dx[1] + (sxy * sxz - syz * sxx) * dx[2]) / denom;
      c[2] = w * ((sxy * syz - sxz * syy) * dx[0] + (sxy * sxz - syz * sxx) * dx[1] + (sxx * syy - sxy * sxy) * dx[2]) / denom;
    };
    double *diag = &m.nbocv_lsg[3 * (size_t)noc_f];
    for (int noc = noc_f + 1; noc <= noc_l; noc++) {
      int nb = m.nbocv_v[noc];
      double dx[3];
      for (int i = 0; i < 3; i++) dx[i] = x_cv[3 * (size_t)nb + i] - x_cv[3 * (size_t)icv + i];
      double *c = &m.nbocv_lsg[3 * (size_t)noc];
      coef(dx, c);
      for (int i = 0; i < 3; i++) diag[i] -= c[i];
    }
    for (int foc = foc_f; foc <= foc_l; foc++) {
      int ifa = m.faocv_v[foc];
      if (ifa < m.nfa_b) {
        double dx[3];
        for (int i = 0; i < 3; i++) dx[i] = m.x_fa[3 * (size_t)ifa + i] - x_cv[3 * (size_t)icv + i];
        double *c = &m.bfa_lsg[3 * (size_t)ifa];
        coef(dx, c);
        for (int i = 0; i < 3; i++) diag[i] -= c[i];
      }
    }
  }
}

}  // namespace jm

// ---------------------------------------------------------------------------
// Fluent ASCII reader with MshFilter's token rules (MshFilter.cpp:3-69):
// parentheses set the level, quotes group, indices are hexadecimal.
// Sections 2 (dimension), 10 (nodes), 12 (cells), 13 (faces), 39/45 (zone
// names); periodic sections are not supported.
// ---------------------------------------------------------------------------
namespace jm {

namespace {

struct MshTok {
  FILE *fp;
  int level = 0;
  std::vector<char> buf;
  size_t pos = 0, max_pos = 0;
  explicit MshTok(FILE *f) : fp(f), buf(1 << 16) {}
  int next(std::string &tok) {
    tok.clear();
    int tl = level;
    bool quote = false;
    while (true) {
      if (pos >= max_pos) {
        max_pos = fread(buf.data(), 1, buf.size(), fp);
        pos = 0;
        if (max_pos == 0) return tok.empty() ? -1 : tl;
      }
      char c = buf[pos++];
      if (c == '"') {
        if (quote) break;
        quote = true;
        continue;
      } else if (quote) {
        tok.push_back(c);
      } else if (c == ' ' || c == '\t' || c == '\n' || c == '\r') {
        if (!tok.empty()) break;
      } else if (c == '(') {
        level++;
        if (!tok.empty()) break;
        tl++;
      } else if (c == ')') {
        level--;
        if (!tok.empty()) break;
        tl--;
      } else {
        tok.push_back(c);
      }
    }
    return tl;
  }
};

long hexv(const std::string &s) { return std::strtol(s.c_str(), nullptr, 16); }

}  // namespace

void read_msh(const std::string &path, Mesh &m) {
  FILE *fp = std::fopen(path.c_str(), "rb");
  if (!fp) fail("cannot open msh file " + path);
  MshTok tk(fp);
  std::string t;
  int nno = 0, ncv = 0, nfa = 0;
  std::vector<double> x_no;
  std::vector<int64_t> fptr(1, 0);
  std::vector<int> fnodes, fcells, fzone;
  std::vector<std::pair<int, std::string>> names;
  int lvl;
  try {
    while ((lvl = tk.next(t)) != -1) {
      if (lvl != 1) continue;
      if (t == "0") {
        tk.next(t);
      } else if (t == "2") {
        tk.next(t);
        if (t != "3") fail("dimension != 3");
      } else if (t == "10") {
        tk.next(t);
        if (t == "0") {
          tk.next(t);
          tk.next(t);
          nno = (int)hexv(t);
          x_no.assign(3 * (size_t)nno, 0.0);
          tk.next(t);
          tk.next(t);
        } else {
          tk.next(t);
          long i0 = hexv(t) - 1;
          tk.next(t);
          long i1 = hexv(t) - 1;
          tk.next(t);
          tk.next(t);
          if (t != "3") fail("node section must be 3D");
          if (i1 >= nno) fail("node index out of range");
          for (long i = i0; i <= i1; i++)
            for (int k = 0; k < 3; k++) {
              tk.next(t);
              x_no[3 * i + k] = std::strtod(t.c_str(), nullptr);
            }
        }
      } else if (t == "12") {
        tk.next(t);
        if (t == "0") {
          tk.next(t);
          tk.next(t);
          ncv = (int)hexv(t);
          tk.next(t);
        } else {
          tk.next(t);
          long i0 = hexv(t);
          tk.next(t);
          long i1 = hexv(t);
          tk.next(t);
          tk.next(t);
          if (hexv(t) == 0)
            for (long i = i0; i <= i1; i++) tk.next(t);
        }
      } else if (t == "13") {
        tk.next(t);
        if (t == "0") {
          tk.next(t);
          tk.next(t);
          nfa = (int)hexv(t);
          tk.next(t);
        } else {
          int zone = (int)hexv(t);
          tk.next(t);
          long i0 = hexv(t) - 1;
          tk.next(t);
          long i1 = hexv(t) - 1;
          tk.next(t);
          tk.next(t);
          int nnf = (int)hexv(t);
          if ((long)fzone.size() != i0) fail("face sections must be contiguous and in order");
          for (long i = i0; i <= i1; i++) {
            int nn = nnf;
            if (nnf == 0 || nnf == 5) {
              tk.next(t);
              nn = (int)hexv(t);
            }
            if (nn < 3 || nn > 8) fail("face with unsupported node count");
            int nl[8];
            for (int k = 0; k < nn; k++) {
              tk.next(t);
              nl[k] = (int)hexv(t) - 1;
            }
            tk.next(t);
            int c0 = (int)hexv(t) - 1;
            tk.next(t);
            int c1 = (int)hexv(t) - 1;
            // MshFilter flip rule (MshFilter.cpp:453-472)
            if (c0 >= 0) {
              for (int k = 0; k < nn; k++) fnodes.push_back(nl[nn - k - 1]);
              fcells.push_back(c0);
              fcells.push_back(c1);
            } else {
              if (c1 < 0) fail("face without cells");
              for (int k = 0; k < nn; k++) fnodes.push_back(nl[k]);
              fcells.push_back(c1);
              fcells.push_back(c0);
            }
            fptr.push_back((int64_t)fnodes.size());
            fzone.push_back(zone);
          }
        }
      } else if (t == "18") {
        fail("periodic face sections are not supported");
      } else if (t == "39" || t == "45") {
        tk.next(t);
        int zone = std::atoi(t.c_str());
        tk.next(t);
        tk.next(t);
        names.emplace_back(zone, t);
      }
    }
  } catch (...) {
    std::fclose(fp);
    throw;
  }
  std::fclose(fp);
  if ((int)fzone.size() != nfa) fail("face count mismatch in msh file");
  FaceInput in;
  in.nno = nno;
  in.nfa = nfa;
  in.ncv = ncv;
  in.x_no = x_no.data();
  in.face_ptr = fptr.data();
  in.face_nodes = fnodes.data();
  in.face_cells = fcells.data();
  in.face_zone = fzone.data();
  in.zone_names = names;
  build_mesh(in, m);
}

// ---------------------------------------------------------------------------
// Ugp restart files (Ugp.cpp:9792-9821 record ids, 9902-10660 writeRestart,
// 12473-13240 readRestart): int32 magic 123581321 and version 1, then records
// of a 256-byte header (char name[52], int32 id, int64 skip to the next
// header, int32 idata[16], double rdata[16]) followed by their data, in the
// writer's byte order (the reader swaps when the magic comes out swapped).
// ---------------------------------------------------------------------------
namespace {

enum : int {
  kIoMagic = 123581321, kIoVersion = 1, kIoCounts = 10, kIoI0 = 11, kIoD0 = 12, kIoFaCheck = 20,
  kIoNoofa = 21, kIoCvofa = 22, kIoFaZone = 23, kIoFaD1 = 24, kIoFaD3 = 25, kIoNoCheck = 30, kIoXNo = 31,
  kIoNoD1 = 32, kIoNoD3 = 33, kIoCvCheck = 40, kIoCvPart = 41, kIoCvD1 = 42, kIoCvD3 = 43, kIoCvD33 = 44,
  kIoData = 50, kIoEof = 51
};
constexpr int kIoHeader = 256, kIoNameLen = 52;

struct IoHeader {
  char name[kIoNameLen];
  int32_t id;
  int64_t skip;
  int32_t idata[16];
  double rdata[16];
};
static_assert(sizeof(IoHeader) == kIoHeader, "Ugp restart header is 256 bytes");

template <typename T>
void swap_bytes(T *p, size_t n) {
  for (size_t i = 0; i < n; i++) {
    unsigned char *b = reinterpret_cast<unsigned char *>(p + i);
    std::reverse(b, b + sizeof(T));
  }
}

struct RestartFile {
  std::vector<unsigned char> bytes;
  bool swap = false;
  template <typename T>
  std::vector<T> read(int64_t at, size_t n) const {
    if (at < 0 || at + (int64_t)(n * sizeof(T)) > (int64_t)bytes.size()) fail("restart file is truncated");
    std::vector<T> v(n);
    std::memcpy(v.data(), bytes.data() + at, n * sizeof(T));
    if (swap) swap_bytes(v.data(), n);
    return v;
  }
  IoHeader header(int64_t at) const {
    if (at < 0 || at + kIoHeader > (int64_t)bytes.size()) fail("restart file is truncated");
    IoHeader h;
    std::memcpy(&h, bytes.data() + at, kIoHeader);
    if (swap) {
      swap_bytes(&h.id, 1);
      swap_bytes(&h.skip, 1);
      swap_bytes(h.idata, 16);
      swap_bytes(h.rdata, 16);
    }
    h.name[kIoNameLen - 1] = 0;
    return h;
  }
};

}  // namespace

void read_restart(const std::string &path, Mesh &m) {
  RestartFile f;
  {
    FILE *fp = std::fopen(path.c_str(), "rb");
    if (!fp) fail("cannot open restart file " + path);
    std::fseek(fp, 0, SEEK_END);
    long n = std::ftell(fp);
    std::fseek(fp, 0, SEEK_SET);
    f.bytes.resize(n > 0 ? (size_t)n : 0);
    const size_t got = f.bytes.empty() ? 0 : std::fread(f.bytes.data(), 1, f.bytes.size(), fp);
    std::fclose(fp);
    if (got != f.bytes.size()) fail("cannot read restart file " + path);
  }
  auto head = f.read<int32_t>(0, 2);
  if (head[0] != kIoMagic) {
    f.swap = true;
    head = f.read<int32_t>(0, 2);
    if (head[0] != kIoMagic) fail(path + " is not a Ugp restart file");
  }
  if (head[1] != kIoVersion) fail(path + ": restart version " + std::to_string(head[1]) + " (1 expected)");
  int nno = -1, nfa = -1, ncv = -1, nnoofa = -1;
  std::vector<int32_t> counts, noofa_v, cvofa, fzone;
  std::vector<double> x_no;
  std::vector<std::pair<int, std::string>> zone_names;
  std::vector<std::pair<std::string, std::vector<double>>> cvf;
  std::vector<int> cvf_dim;
  int step = 0;
  double time = 0.0;
  int64_t at = 8;
  for (bool done = false; !done;) {
    const IoHeader h = f.header(at);
    const int64_t data = at + kIoHeader;
    switch (h.id) {
      case kIoCounts:
        nno = h.idata[0];
        nfa = h.idata[1];
        ncv = h.idata[2];
        nnoofa = h.idata[3];
        if (nno <= 0 || nfa <= 0 || ncv <= 0 || nnoofa <= 0) fail("restart file: bad counts");
        fzone.assign(nfa, -1);
        break;
      case kIoNoofa:
        if (h.idata[0] != nfa || h.idata[1] != nnoofa) fail("restart file: NOOFA counts disagree");
        counts = f.read<int32_t>(data, nfa);
        noofa_v = f.read<int32_t>(data + 4 * (int64_t)nfa, nnoofa);
        break;
      case kIoCvofa:
        if (h.idata[0] != nfa || h.idata[1] != 2) fail("restart file: CVOFA counts disagree");
        cvofa = f.read<int32_t>(data, 2 * (size_t)nfa);
        break;
      case kIoFaZone: {
        const int kind = h.idata[0], f0 = h.idata[1], f1 = h.idata[2];
        if (kind != kZoneBoundary && kind != kZoneInternal)
          fail(std::string("restart file: zone ") + h.name + ": periodic zones are not supported");
        if (f0 < 0 || f1 >= nfa) fail(std::string("restart file: zone ") + h.name + ": face range");
        const int id = (int)zone_names.size();
        zone_names.emplace_back(id, h.name);
        for (int i = f0; i <= f1; i++) fzone[i] = id;
        break;
      }
      case kIoXNo:
        if (h.idata[0] != nno || h.idata[1] != 3) fail("restart file: X_NO counts disagree");
        x_no = f.read<double>(data, 3 * (size_t)nno);
        break;
      case kIoI0:
        if (std::strcmp(h.name, "STEP") == 0) step = h.idata[0];
        break;
      case kIoD0:
        if (std::strcmp(h.name, "TIME") == 0) time = h.rdata[0];
        break;
      case kIoCvD1:
      case kIoCvD3: {
        const int dim = h.id == kIoCvD1 ? 1 : 3;
        if (h.idata[0] != ncv) fail(std::string("restart file: field ") + h.name + " has the wrong count");
        cvf.emplace_back(h.name, f.read<double>(data, (size_t)dim * ncv));
        cvf_dim.push_back(dim);
        break;
      }
      case kIoEof:
        done = true;
        break;
      default:  // checks, partition, face/node/tensor fields, DATA marker
        break;
    }
    if (h.skip < kIoHeader) fail("restart file: bad record length");
    at += h.skip;
  }
  if (counts.empty() || cvofa.empty() || x_no.empty()) fail("restart file: mesh records missing");
  std::vector<int64_t> fptr(nfa + 1, 0);
  for (int i = 0; i < nfa; i++) {
    if (counts[i] < 3) fail("restart file: a face with fewer than 3 nodes");
    fptr[i + 1] = fptr[i] + counts[i];
  }
  if (fptr[nfa] != nnoofa) fail("restart file: NOOFA sizes disagree");
  for (int i = 0; i < nfa; i++) {
    if (fzone[i] < 0) fail("restart file: a face in no zone");
    if (cvofa[2 * i + 1] < -1) fail("restart file: periodic faces are not supported");
  }
  std::vector<int> fnodes(noofa_v.begin(), noofa_v.end()), fcells(cvofa.begin(), cvofa.end());
  FaceInput in;
  in.nno = nno;
  in.nfa = nfa;
  in.ncv = ncv;
  in.x_no = x_no.data();
  in.face_ptr = fptr.data();
  in.face_nodes = fnodes.data();
  in.face_cells = fcells.data();
  in.face_zone = fzone.data();
  in.zone_names = zone_names;
  build_mesh(in, m);
  // the file's cell fields in the final numbering
  m.restart_cv.clear();
  for (size_t k = 0; k < cvf.size(); k++) {
    const int dim = cvf_dim[k];
    std::vector<double> v(cvf[k].second.size());
    for (int c = 0; c < ncv; c++)
      for (int i = 0; i < dim; i++) v[(size_t)dim * m.cv_of_input[c] + i] = cvf[k].second[(size_t)dim * c + i];
    m.restart_cv.emplace_back(cvf[k].first, std::move(v));
  }
  m.restart_step = step;
  m.restart_time = time;
}

void write_restart(const std::string &path, const Mesh &m, const std::vector<RestartField> &fields, int step,
                   double time) {
  std::vector<unsigned char> out;
  auto put = [&](const void *p, size_t n) {
    const unsigned char *b = static_cast<const unsigned char *>(p);
    out.insert(out.end(), b, b + n);
  };
  // writeRestart reuses one header for all zone records, so a short zone
  // name keeps the tail of a longer one after its terminator; the same here
  char zone_name[kIoNameLen] = {0};
  auto header = [&](const char *name, int id, int64_t data_bytes, std::initializer_list<int> idata,
                    double r0 = 0.0) {
    IoHeader h;
    std::memset(&h, 0, sizeof h);
    if (id == kIoFaZone) {
      std::snprintf(zone_name, kIoNameLen, "%s", name);
      std::memcpy(h.name, zone_name, kIoNameLen);
    } else {
      std::snprintf(h.name, kIoNameLen, "%s", name);
    }
    h.id = id;
    h.skip = kIoHeader + data_bytes;
    int k = 0;
    for (int v : idata) h.idata[k++] = v;
    h.rdata[0] = r0;
    put(&h, sizeof h);
  };
  const int nno = m.nno, nfa = m.nfa, ncv = m.ncv, nnoofa = m.noofa_i[nfa];
  // nodes: those of boundary faces first, then the rest, each in mesh order;
  // faces by zone and cells in mesh order (writeRestart's serial layout)
  std::vector<int> bnd(nno, 1), no_file(nno);
  for (int i = m.noofa_i[0]; i < m.noofa_i[m.nfa_b]; i++) bnd[m.noofa_v[i]] = 0;
  int nb = 0;
  for (int n = 0; n < nno; n++) nb += bnd[n] == 0;
  {
    int k0 = 0, k1 = nb;
    for (int n = 0; n < nno; n++) no_file[n] = bnd[n] == 0 ? k0++ : k1++;
  }
  const int32_t head[2] = {kIoMagic, kIoVersion};
  put(head, sizeof head);
  header("NO_FA_CV_NOOFA_COUNTS", kIoCounts, 0, {nno, nfa, ncv, nnoofa});
  std::vector<int32_t> iv;
  header("NO_CHECK", kIoNoCheck, 4LL * nno, {nno});
  iv.resize(nno);
  for (int n = 0; n < nno; n++) iv[n] = n;
  put(iv.data(), 4 * iv.size());
  header("FA_CHECK", kIoFaCheck, 4LL * nfa, {nfa});
  iv.resize(nfa);
  for (int i = 0; i < nfa; i++) iv[i] = i;
  put(iv.data(), 4 * iv.size());
  header("CV_CHECK", kIoCvCheck, 4LL * ncv, {ncv});
  iv.resize(ncv);
  for (int c = 0; c < ncv; c++) iv[c] = c;
  put(iv.data(), 4 * iv.size());
  header("NOOFA_I_AND_V", kIoNoofa, 4LL * nfa + 4LL * nnoofa, {nfa, nnoofa});
  iv.resize(nfa);
  for (int i = 0; i < nfa; i++) iv[i] = m.noofa_i[i + 1] - m.noofa_i[i];
  put(iv.data(), 4 * iv.size());
  iv.resize(nnoofa);
  for (int i = 0; i < nnoofa; i++) iv[i] = no_file[m.noofa_v[i]];
  put(iv.data(), 4 * iv.size());
  header("CVOFA", kIoCvofa, 8LL * nfa, {nfa, 2});
  iv.assign(m.cvofa.begin(), m.cvofa.end());
  put(iv.data(), 4 * iv.size());
  for (const Zone &z : m.zones) {
    const int f0 = z.kind == kZoneInternal ? z.ifa_i_f : z.ifa_f, f1 = z.kind == kZoneInternal ? z.ifa_i_l : z.ifa_l;
    header(z.name.c_str(), kIoFaZone, 0, {z.kind, f0, f1});
  }
  header("CV_PART", kIoCvPart, 4LL * ncv, {ncv, 1});
  iv.assign(ncv, 0);
  put(iv.data(), 4 * iv.size());
  header("X_NO", kIoXNo, 24LL * nno, {nno, 3});
  {
    std::vector<double> x(3 * (size_t)nno);
    for (int n = 0; n < nno; n++)
      for (int i = 0; i < 3; i++) x[3 * (size_t)no_file[n] + i] = m.x_no[3 * (size_t)n + i];
    put(x.data(), 8 * x.size());
  }
  header("DATA", kIoData, 0, {});
  header("STEP", kIoI0, 0, {step});
  header("TIME", kIoD0, 0, {}, time);
  for (const RestartField &fd : fields) {
    if (fd.dim != 1 && fd.dim != 3) fail("restart field " + fd.name + ": dimension 1 or 3");
    header(fd.name.c_str(), fd.dim == 1 ? kIoCvD1 : kIoCvD3, 8LL * fd.dim * ncv, {ncv, fd.dim == 3 ? 3 : 0});
    put(fd.data, 8 * (size_t)fd.dim * ncv);
  }
  header("EOF", kIoEof, 0, {});
  FILE *fp = std::fopen(path.c_str(), "wb");
  if (!fp) fail("cannot write restart file " + path);
  const size_t wrote = std::fwrite(out.data(), 1, out.size(), fp);
  const bool closed = std::fclose(fp) == 0;
  if (wrote != out.size() || !closed) fail("cannot write restart file " + path);
}

}  // namespace jm
