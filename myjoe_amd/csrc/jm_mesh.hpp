// Host-side mesh in the reference's (Ugp/UgpWithCv) numbering and layout.
//
// The numbering contract follows the reference's serial path:
//   MshFilter flip rule        MshFilter.cpp:453-472
//   first-touch renumbering    Ugp.cpp:1760-2120 (redistReconnectReorder, 1 rank)
//   zone sort                  Ugp.cpp:2942 with FaZone::operator< (Ugp.h:155-157)
//   face / cell buckets        Ugp.cpp:3390-3786
//   faocv build + reorder      Ugp.cpp:5452-5503, 5584-5953
//   nbocv                      UgpWithCv.cpp:768-834
//   geometry                   UgpWithCv.cpp:842-1095
//   least-squares coefficients UgpWithCv.cpp:1783-1913
#pragma once

#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace jm {

enum ZoneKind : int { kZoneBoundary = 1, kZoneInternal = 6 };

struct Zone {
  std::string name;
  int index = 0;  // zone id of the mesh file
  int kind = kZoneBoundary;
  int ifa_f = 0, ifa_l = -1;      // boundary: its faces; internal: the (empty) inter-processor part
  int ifa_i_f = 0, ifa_i_l = -1;  // internal zones: the fully internal faces (Ugp.h:150-151)
};

struct Mesh {
  int nno = 0, nfa = 0, nfa_b = 0, ncv = 0;
  std::vector<double> x_no;              // [nno][3]
  std::vector<int> noofa_i, noofa_v;     // CSR
  std::vector<int> cvofa;                // [nfa][2], icv1 = -1 on boundary
  std::vector<int> faocv_i, faocv_v;     // CSR
  std::vector<int> nbocv_i, nbocv_v;     // CSR, diagonal first
  std::vector<Zone> zones;               // sorted: boundary zones by name, then internal
  std::vector<double> x_fa, fa_normal;   // [nfa][3]
  std::vector<double> x_cv;              // [ncv][3]
  std::vector<double> cv_volume;         // [ncv]
  std::vector<double> nbocv_lsg;         // [nnz][3]
  std::vector<double> bfa_lsg;           // [nfa_b][3]
  int count_neg_subtet = 0;              // sub-tets with non-positive volume
  double min_sdotn = 1.0;                // min over internal faces of s.n / |s||n|
  std::vector<int> cv_of_input;          // final index of each input (file) cell
  // cell fields of a restart file, in the final numbering (read_restart)
  std::vector<std::pair<std::string, std::vector<double>>> restart_cv;
  int restart_step = 0;
  double restart_time = 0.0;

  int nbocv_s() const { return ncv ? nbocv_i[ncv] : 0; }
};

// Faces in file order, already in the Joe orientation (normal from cells[0]
// to cells[1], boundary faces carry -1 second). Throws std::runtime_error.
struct FaceInput {
  int nno = 0, nfa = 0, ncv = 0;
  const double *x_no = nullptr;         // [nno][3]
  const int64_t *face_ptr = nullptr;    // [nfa+1]
  const int *face_nodes = nullptr;
  const int *face_cells = nullptr;      // [nfa][2]
  const int *face_zone = nullptr;       // [nfa] zone ids
  std::vector<std::pair<int, std::string>> zone_names;  // (id, name)
};

void build_mesh(const FaceInput &in, Mesh &m);

// Reads a Fluent ASCII .msh the way MshFilter does and builds the mesh.
void read_msh(const std::string &path, Mesh &m);

// Reads a Ugp restart file (serial: one partition, no periodic faces) the way
// Ugp::readRestart does and builds the mesh; its STEP, TIME and cell fields
// (scalars and 3-vectors, by name) are kept in m.restart_*.
void read_restart(const std::string &path, Mesh &m);

// Writes a Ugp restart file of the mesh (Ugp::writeRestart's serial layout)
// with STEP, TIME and the given cell fields (ncv values for a scalar, 3 ncv
// for a vector) in the mesh's numbering.
struct RestartField {
  std::string name;
  const double *data;
  int dim;  // 1 or 3
};
void write_restart(const std::string &path, const Mesh &m, const std::vector<RestartField> &fields, int step,
                   double time);

// Geometry, nbocv and LS coefficients from connectivity + x_no.
void build_nbocv(Mesh &m);
void calc_geometry(Mesh &m);
void calc_lsg_coeff(Mesh &m);

}  // namespace jm
