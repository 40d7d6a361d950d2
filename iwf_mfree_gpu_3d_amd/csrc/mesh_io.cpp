// mesh_io.cpp -- input files of the set-up (host): the surface of an LS-DYNA keyword solid mesh
// and MeshLab JSON triangle meshes (indexed triangle meshes for mfg_tool_3d_gpu_create), and the
// warm-start inputs: particle positions from a legacy VTK file and steady-state temperature
// samples.
//
//   VTK       vtk_read_init_pos (src/vtk_reader.cpp:29-69): the lines after the line holding
//             "POINTS", up to the first line holding "CELLS", that split into exactly three
//             space-separated fields (lines of a single space are skipped)
//   steady    the "x y z T" lines of the steady-state temperature file read by the single-grain
//             set-up (src/Benchmarks/benchmarks_single_grain.cu:281-301): every line that splits
//             into exactly four space-separated fields
//   Fields are split at every single space, as std::getline with a ' ' delimiter does: two
//   spaces in a row make an empty field, a trailing space none.
//
//   LS-DYNA   ldynak_read_triangles_from_tetmesh / _quadmesh (src/ldynak_reader.cpp:308-434):
//             *NODE lines "nid x y z [tc rc]", *ELEMENT_SOLID lines "eid pid n1 .. n8" (tets
//             repeat n4 in n5..n8); every element face is listed with the node order of the
//             reference's face tables; faces that belong to exactly one element form the surface
//             (hexahedra: each boundary quad is split into (1 2 3) and (3 4 1))
//   JSON      json_read_triangles (src/json_reader.cpp:108-117): the comma-separated numbers
//             after "vertices" up to the first ']' are x y z triples, those after
//             "connectivity" up to the next ']' are triangle index triples
//
// The reference collects boundary faces and renumbers nodes through std::unordered_map, whose
// iteration order is implementation-defined.  Here the order is fixed: vertices by ascending
// node id, boundary faces in element order (and face-table order within an element).
#include <algorithm>
#include <array>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/mfree_gpu.h"

namespace {

struct Mesh {
  std::vector<double> v;  // x y z
  std::vector<int> t;     // i1 i2 i3
};

std::vector<std::string> tokens_ws(const std::string &s) {
  std::istringstream in(s);
  std::vector<std::string> out;
  std::string w;
  while (in >> w) out.push_back(w);
  return out;
}

// faces of a hexahedron and of a tetrahedron, node order as the reference lists them
const int HEX_FACES[6][4] = {{0, 1, 2, 3}, {7, 6, 5, 4}, {4, 5, 1, 0},
                             {3, 2, 6, 7}, {0, 3, 7, 4}, {5, 6, 2, 1}};
const int TET_FACES[4][3] = {{0, 1, 3}, {2, 1, 0}, {1, 2, 3}, {3, 2, 0}};

int read_ldynak(const char *path, int quads, Mesh &m, std::string &err) {
  std::ifstream in(path);
  if (!in) {
    err = std::string("cannot open ") + path;
    return MFG_ERR_ARG;
  }
  std::map<unsigned, std::array<double, 3>> nodes;
  std::vector<std::array<unsigned, 8>> elems;
  enum { NONE, NODE, ELEM } sec = NONE;
  std::string line;
  while (std::getline(in, line)) {
    if (line.find('$') != std::string::npos) continue;  // comment lines
    if (line.find('*') != std::string::npos) {
      if (line.find("*NODE") != std::string::npos) sec = NODE;
      else if (line.find("*ELEMENT") != std::string::npos) sec = ELEM;
      else sec = NONE;
      continue;
    }
    const std::vector<std::string> tk = tokens_ws(line);
    if (tk.empty()) continue;
    if (sec == NODE) {
      if (tk.size() != 4 && tk.size() != 6) {
        err = "malformed *NODE line: " + line;
        return MFG_ERR_ARG;
      }
      nodes[(unsigned)std::strtoul(tk[0].c_str(), nullptr, 0)] = {
          std::strtod(tk[1].c_str(), nullptr), std::strtod(tk[2].c_str(), nullptr),
          std::strtod(tk[3].c_str(), nullptr)};
    } else if (sec == ELEM) {
      if (tk.size() != 10) {
        err = "malformed *ELEMENT line: " + line;
        return MFG_ERR_ARG;
      }
      std::array<unsigned, 8> e;
      for (int k = 0; k < 8; k++) e[k] = (unsigned)std::strtoul(tk[2 + k].c_str(), nullptr, 0);
      elems.push_back(e);
    }
  }
  if (nodes.empty() || elems.empty()) {
    err = std::string("no nodes or elements in ") + path;
    return MFG_ERR_ARG;
  }
  // every face keyed by its sorted node set; count, first oriented occurrence
  struct Face {
    std::array<unsigned, 4> nodes;
    int count;
  };
  std::map<std::array<unsigned, 4>, size_t> index;
  std::vector<Face> faces;
  const int nf = quads ? 6 : 4, nv = quads ? 4 : 3;
  for (const auto &e : elems)
    for (int f = 0; f < nf; f++) {
      std::array<unsigned, 4> fn = {0, 0, 0, 0};
      for (int k = 0; k < nv; k++) fn[k] = e[quads ? HEX_FACES[f][k] : TET_FACES[f][k]];
      std::array<unsigned, 4> key = fn;
      std::sort(key.begin(), key.begin() + nv);
      auto it = index.find(key);
      if (it == index.end()) {
        index[key] = faces.size();
        faces.push_back({fn, 1});
      } else {
        faces[it->second].count++;
      }
    }
  std::map<unsigned, int> lin;
  for (const auto &kv : nodes) {
    lin[kv.first] = (int)(m.v.size() / 3);
    m.v.insert(m.v.end(), kv.second.begin(), kv.second.end());
  }
  auto id = [&](unsigned n, int &out) {
    auto it = lin.find(n);
    if (it == lin.end()) return false;
    out = it->second;
    return true;
  };
  for (const auto &f : faces) {
    if (f.count != 1) continue;
    int a, b, c, d = 0;
    if (!id(f.nodes[0], a) || !id(f.nodes[1], b) || !id(f.nodes[2], c) ||
        (quads && !id(f.nodes[3], d))) {
      err = "element references an undefined node";
      return MFG_ERR_ARG;
    }
    m.t.insert(m.t.end(), {a, b, c});
    if (quads) m.t.insert(m.t.end(), {c, d, a});
  }
  return MFG_OK;
}

// the comma-separated numbers on the lines after the line holding `key`, up to the first ']'
bool json_numbers(std::ifstream &in, const char *key, std::vector<std::string> &out) {
  std::string line;
  bool sec = false;
  while (std::getline(in, line)) {
    if (line.find(key) != std::string::npos) sec = true;
    if (line.find(']') != std::string::npos) {
      if (sec) return true;
      continue;
    }
    if (!sec) continue;
    std::stringstream ss(line);
    std::vector<std::string> tk;
    std::string item;
    while (std::getline(ss, item, ',')) tk.push_back(item);
    if (tk.size() >= 3) out.insert(out.end(), tk.begin(), tk.end());
  }
  return sec;
}

int read_json(const char *path, Mesh &m, std::string &err) {
  std::ifstream in(path);
  if (!in) {
    err = std::string("cannot open ") + path;
    return MFG_ERR_ARG;
  }
  std::vector<std::string> vs, ts;
  if (!json_numbers(in, "vertices", vs) || !json_numbers(in, "connectivity", ts) ||
      vs.size() % 3 || ts.size() % 3 || vs.empty() || ts.empty()) {
    err = std::string("no vertices / connectivity in ") + path;
    return MFG_ERR_ARG;
  }
  for (const auto &s : vs) m.v.push_back(std::atof(s.c_str()));
  const int nv = (int)(m.v.size() / 3);
  for (const auto &s : ts) {
    const int k = std::atoi(s.c_str());
    if (k < 0 || k >= nv) {
      err = "triangle index out of range";
      return MFG_ERR_ARG;
    }
    m.t.push_back(k);
  }
  return MFG_OK;
}

int deliver(const Mesh &m, int *n_verts, int *n_tris, double *verts3, int *tris3) {
  if (n_verts) *n_verts = (int)(m.v.size() / 3);
  if (n_tris) *n_tris = (int)(m.t.size() / 3);
  if (verts3) std::memcpy(verts3, m.v.data(), sizeof(double) * m.v.size());
  if (tris3) std::memcpy(tris3, m.t.data(), sizeof(int) * m.t.size());
  return MFG_OK;
}

thread_local std::string g_mesh_err;

std::vector<std::string> split_spaces(const std::string &s) {
  std::vector<std::string> out;
  size_t a = 0;
  while (a < s.size()) {
    const size_t b = s.find(' ', a);
    if (b == std::string::npos) {
      out.push_back(s.substr(a));
      break;
    }
    out.push_back(s.substr(a, b - a));
    a = b + 1;
  }
  return out;
}

bool parse_double(const std::string &t, double &v) {
  const char *p = t.c_str();
  char *end = nullptr;
  v = std::strtod(p, &end);
  return end != p;
}

int read_fields(const char *path, bool vtk, int width, std::vector<double> &out,
                std::string &err) {
  std::ifstream in(path);
  if (!in) {
    err = std::string("cannot open ") + path;
    return MFG_ERR_ARG;
  }
  std::string line;
  bool sec = !vtk;
  while (std::getline(in, line)) {
    if (vtk) {
      if (line.find("POINTS") != std::string::npos) {
        sec = true;
        continue;
      }
      if (line == " ") continue;
      if (sec && line.find("CELLS") != std::string::npos) break;
    }
    if (!sec) continue;
    const std::vector<std::string> tk = split_spaces(line);
    if ((int)tk.size() != width) continue;
    for (int k = 0; k < width; k++) {
      double v;
      if (!parse_double(tk[k], v)) {
        err = "malformed number in line: " + line;
        return MFG_ERR_ARG;
      }
      out.push_back(v);
    }
  }
  return MFG_OK;
}

int deliver_fields(const std::vector<double> &v, int width, int *n, double *out) {
  if (n) *n = (int)(v.size() / width);
  if (out) std::memcpy(out, v.data(), sizeof(double) * v.size());
  return MFG_OK;
}

}  // namespace

extern "C" {

int mfg_mesh_read_ldynak(const char *path, int quads, int *n_verts, int *n_tris, double *verts3,
                         int *tris3) {
  if (!path) return MFG_ERR_ARG;
  Mesh m;
  const int rc = read_ldynak(path, quads, m, g_mesh_err);
  if (rc != MFG_OK) return rc;
  return deliver(m, n_verts, n_tris, verts3, tris3);
}

int mfg_mesh_read_json(const char *path, int *n_verts, int *n_tris, double *verts3, int *tris3) {
  if (!path) return MFG_ERR_ARG;
  Mesh m;
  const int rc = read_json(path, m, g_mesh_err);
  if (rc != MFG_OK) return rc;
  return deliver(m, n_verts, n_tris, verts3, tris3);
}

int mfg_vtk_read_positions(const char *path, int *n, double *xyz) {
  if (!path) return MFG_ERR_ARG;
  std::vector<double> v;
  const int rc = read_fields(path, true, 3, v, g_mesh_err);
  if (rc != MFG_OK) return rc;
  return deliver_fields(v, 3, n, xyz);
}

int mfg_read_steady_temperatures(const char *path, int *n, double *xyzt) {
  if (!path) return MFG_ERR_ARG;
  std::vector<double> v;
  const int rc = read_fields(path, false, 4, v, g_mesh_err);
  if (rc != MFG_OK) return rc;
  return deliver_fields(v, 4, n, xyzt);
}

const char *mfg_mesh_last_error(void) { return g_mesh_err.c_str(); }

}  // extern "C"
