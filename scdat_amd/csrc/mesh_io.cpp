// Mesh import and result export, host side of the C ABI.
//
// Readers (restating the behaviour of GmshLoader / VTKLoader / STLLoader,
// Tools/Mesh/src/MeshParsing.cpp:651-1091):
//   Gmsh ASCII 2.x and 4.x: $Nodes, $Elements (triangles, type 2, and
//     tetrahedra, type 4; other types skipped), the physical tag of each
//     element (2.x: first tag of the element line; 4.x: first physical tag of
//     its surface or volume entity in $Entities, 0 without one);
//   VTK legacy ASCII: POINTS, CELLS, CELL_TYPES (5 triangle, 10 tetrahedron);
//   STL ASCII: triangles, vertices merged on a 1e-12 grid.
// Boundary groups for the device path: every tetrahedron face without a
// neighbour takes the group mapped from the physical tag of the triangle
// with the same three nodes (a default group when no triangle matches).
//
// Writers: the fields of a tetrahedral mesh as a legacy VTK unstructured
// grid with point data, and the columnar output contract of
// ResultExporter::exportDataSet (Tools/Output/src/ResultExporter.cpp:176-230):
// CSV (axis, then the series in name order, scientific, 8 digits), VTK
// poly-data points along the axis, the text "HDF5-like" archive, and a
// .metadata.json sidecar for each.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <map>
#include <sstream>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/scdat_amd.h"

struct scdat_mesh_data {
  std::vector<double> xyz;
  std::vector<int32_t> tets, tet_tag, tris, tri_tag;
};

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& m) {
  g_err = m;
  return code;
}

std::vector<long long> int_list(const std::string& text) {
  std::istringstream iss(text);
  std::vector<long long> v;
  long long x = 0;
  while (iss >> x) v.push_back(x);
  return v;
}

int32_t node_index(const std::unordered_map<long long, int32_t>& ids, long long id) {
  const auto it = ids.find(id);
  if (it == ids.end()) throw std::runtime_error("element references unknown node " + std::to_string(id));
  return it->second;
}

void read_gmsh(const std::string& path, scdat_mesh_data& m) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::unordered_map<long long, int32_t> ids;
  std::unordered_map<long long, int32_t> surf_phys, vol_phys;
  std::string line;
  auto add_element = [&](long long type, const std::vector<long long>& nodes, int32_t tag) {
    if (type == 2 && nodes.size() >= 3) {
      for (int k = 0; k < 3; ++k) m.tris.push_back(node_index(ids, nodes[k]));
      m.tri_tag.push_back(tag);
    } else if (type == 4 && nodes.size() >= 4) {
      for (int k = 0; k < 4; ++k) m.tets.push_back(node_index(ids, nodes[k]));
      m.tet_tag.push_back(tag);
    }
  };
  while (std::getline(in, line)) {
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line == "$Entities") {
      std::getline(in, line);
      const auto h = int_list(line);
      if (h.size() < 4) throw std::runtime_error("bad $Entities header");
      for (long long i = 0; i < h[0] + h[1]; ++i) std::getline(in, line);
      for (int dim = 2; dim <= 3; ++dim)
        for (long long i = 0; i < h[dim]; ++i) {
          std::getline(in, line);
          std::istringstream iss(line);
          long long tag = 0, nphys = 0;
          double box[6];
          iss >> tag;
          for (double& b : box) iss >> b;
          iss >> nphys;
          long long first = 0;
          for (long long p = 0; p < nphys; ++p) {
            long long ph = 0;
            iss >> ph;
            if (p == 0) first = ph;
          }
          if (nphys > 0) (dim == 2 ? surf_phys : vol_phys)[tag] = (int32_t)first;
        }
    } else if (line == "$Nodes") {
      std::getline(in, line);
      const auto h = int_list(line);
      if (h.size() == 1) {  // 2.x
        for (long long i = 0; i < h[0]; ++i) {
          std::getline(in, line);
          std::istringstream iss(line);
          long long id = 0;
          double x = 0, y = 0, z = 0;
          iss >> id >> x >> y >> z;
          ids[id] = (int32_t)(m.xyz.size() / 3);
          m.xyz.insert(m.xyz.end(), {x, y, z});
        }
      } else if (h.size() >= 4) {  // 4.x blocks
        for (long long b = 0; b < h[0]; ++b) {
          std::getline(in, line);
          const auto bh = int_list(line);
          if (bh.size() < 4) throw std::runtime_error("bad $Nodes block header");
          const long long nb = bh[3];
          std::vector<long long> tags;
          while ((long long)tags.size() < nb) {
            std::getline(in, line);
            for (long long t : int_list(line)) {
              if ((long long)tags.size() < nb) tags.push_back(t);
            }
          }
          for (long long i = 0; i < nb; ++i) {
            std::getline(in, line);
            std::istringstream iss(line);
            double x = 0, y = 0, z = 0;
            iss >> x >> y >> z;  // parametric coordinates, if any, follow and are ignored
            ids[tags[i]] = (int32_t)(m.xyz.size() / 3);
            m.xyz.insert(m.xyz.end(), {x, y, z});
          }
        }
      } else {
        throw std::runtime_error("unsupported $Nodes header");
      }
    } else if (line == "$Elements") {
      std::getline(in, line);
      const auto h = int_list(line);
      if (h.size() == 1) {  // 2.x: id type ntags tags... nodes...
        for (long long i = 0; i < h[0]; ++i) {
          std::getline(in, line);
          const auto v = int_list(line);
          if (v.size() < 3) continue;
          const long long type = v[1], nt = v[2];
          const int32_t tag = nt > 0 && v.size() > 3 ? (int32_t)v[3] : 0;
          if ((long long)v.size() < 3 + nt) continue;
          add_element(type, std::vector<long long>(v.begin() + 3 + nt, v.end()), tag);
        }
      } else if (h.size() >= 4) {  // 4.x blocks: dim entity type count, then tag nodes...
        for (long long b = 0; b < h[0]; ++b) {
          std::getline(in, line);
          const auto bh = int_list(line);
          if (bh.size() < 4) throw std::runtime_error("bad $Elements block header");
          int32_t tag = 0;
          const auto& pm = bh[0] == 2 ? surf_phys : vol_phys;
          if (bh[0] == 2 || bh[0] == 3) {
            const auto it = pm.find(bh[1]);
            if (it != pm.end()) tag = it->second;
          }
          for (long long i = 0; i < bh[3]; ++i) {
            std::getline(in, line);
            const auto v = int_list(line);
            if (v.size() < 2) continue;
            add_element(bh[2], std::vector<long long>(v.begin() + 1, v.end()), tag);
          }
        }
      } else {
        throw std::runtime_error("unsupported $Elements header");
      }
    }
  }
}

void read_vtk(const std::string& path, scdat_mesh_data& m) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string tok;
  std::vector<std::vector<long long>> cells;
  while (in >> tok) {
    if (tok == "POINTS") {
      long long n = 0;
      std::string type;
      in >> n >> type;
      m.xyz.resize(3 * (size_t)n);
      for (auto& x : m.xyz) in >> x;
    } else if (tok == "CELLS") {
      long long n = 0, total = 0;
      in >> n >> total;
      cells.assign((size_t)n, {});
      for (auto& c : cells) {
        long long k = 0;
        in >> k;
        c.resize((size_t)k);
        for (auto& v : c) in >> v;
      }
    } else if (tok == "CELL_TYPES") {
      long long n = 0;
      in >> n;
      for (long long i = 0; i < n; ++i) {
        int t = 0;
        in >> t;
        if (i >= (long long)cells.size()) continue;
        const auto& c = cells[(size_t)i];
        const long long nn = (long long)(m.xyz.size() / 3);
        for (long long v : c)
          if (v < 0 || v >= nn) throw std::runtime_error("cell references a missing point");
        if (t == 10 && c.size() == 4) {
          for (long long v : c) m.tets.push_back((int32_t)v);
          m.tet_tag.push_back(0);
        } else if (t == 5 && c.size() == 3) {
          for (long long v : c) m.tris.push_back((int32_t)v);
          m.tri_tag.push_back(0);
        }
      }
    }
  }
}

void read_stl(const std::string& path, scdat_mesh_data& m) {
  std::ifstream in(path);
  if (!in) throw std::runtime_error("cannot open " + path);
  std::string first;
  in >> first;
  if (first != "solid") throw std::runtime_error("only ASCII STL is supported");
  in.clear();
  in.seekg(0);
  std::map<std::array<long long, 3>, int32_t> seen;
  std::string line;
  std::vector<int32_t> tri;
  while (std::getline(in, line)) {
    std::istringstream iss(line);
    std::string w;
    iss >> w;
    if (w != "vertex") continue;
    double p[3] = {0, 0, 0};
    iss >> p[0] >> p[1] >> p[2];
    const std::array<long long, 3> key = {std::llround(p[0] / 1e-12), std::llround(p[1] / 1e-12),
                                          std::llround(p[2] / 1e-12)};
    auto it = seen.find(key);
    int32_t idx;
    if (it == seen.end()) {
      idx = (int32_t)(m.xyz.size() / 3);
      m.xyz.insert(m.xyz.end(), {p[0], p[1], p[2]});
      seen.emplace(key, idx);
    } else {
      idx = it->second;
    }
    tri.push_back(idx);
    if (tri.size() == 3) {
      m.tris.insert(m.tris.end(), tri.begin(), tri.end());
      m.tri_tag.push_back(0);
      tri.clear();
    }
  }
}

std::string lower_ext(const std::string& p) {
  const auto d = p.find_last_of('.');
  std::string e = d == std::string::npos ? "" : p.substr(d + 1);
  for (auto& ch : e) ch = (char)std::tolower((unsigned char)ch);
  return e;
}

std::string json_escape(const std::string& s) {
  std::string o;
  for (const unsigned char ch : s) {
    switch (ch) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      default:
        if (ch < 0x20) {
          char b[8];
          std::snprintf(b, sizeof(b), "\\u%04x", ch);
          o += b;
        } else {
          o.push_back((char)ch);
        }
    }
  }
  return o;
}

std::string utc_now() {
  const std::time_t t = std::chrono::system_clock::to_time_t(std::chrono::system_clock::now());
  std::tm u{};
  gmtime_r(&t, &u);
  char b[32];
  std::strftime(b, sizeof(b), "%Y-%m-%dT%H:%M:%SZ", &u);
  return b;
}

}  // namespace

extern "C" {

const char* scdat_io_last_error(void) { return g_err.c_str(); }

int scdat_mesh_read(const char* path, int32_t format, scdat_mesh_data** out) {
  if (!path || !out) return fail(SCDAT_ERR_ARG, "mesh_read: null argument");
  *out = nullptr;
  try {
    int f = format;
    if (f == SCDAT_MESH_AUTO) {
      const std::string e = lower_ext(path);
      f = e == "msh" ? SCDAT_MESH_GMSH : e == "vtk" ? SCDAT_MESH_VTK : e == "stl" ? SCDAT_MESH_STL : -1;
      if (f < 0) return fail(SCDAT_ERR_ARG, "mesh_read: unknown extension (msh, vtk, stl)");
    }
    auto* m = new scdat_mesh_data();
    try {
      if (f == SCDAT_MESH_GMSH) read_gmsh(path, *m);
      else if (f == SCDAT_MESH_VTK) read_vtk(path, *m);
      else if (f == SCDAT_MESH_STL) read_stl(path, *m);
      else throw std::runtime_error("unknown format");
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
    return SCDAT_OK;
  } catch (const std::exception& e) {
    return fail(SCDAT_ERR_ARG, std::string("mesh_read: ") + e.what());
  }
}

void scdat_mesh_free(scdat_mesh_data* m) { delete m; }

int scdat_mesh_sizes(const scdat_mesh_data* m, int64_t* n_nodes, int64_t* n_tets, int64_t* n_tris) {
  if (!m) return fail(SCDAT_ERR_ARG, "mesh_sizes: null mesh");
  if (n_nodes) *n_nodes = (int64_t)(m->xyz.size() / 3);
  if (n_tets) *n_tets = (int64_t)m->tet_tag.size();
  if (n_tris) *n_tris = (int64_t)m->tri_tag.size();
  return SCDAT_OK;
}

int scdat_mesh_arrays(const scdat_mesh_data* m, double* xyz, int32_t* tets, int32_t* tet_tag, int32_t* tris,
                      int32_t* tri_tag) {
  if (!m) return fail(SCDAT_ERR_ARG, "mesh_arrays: null mesh");
  if (xyz) std::copy(m->xyz.begin(), m->xyz.end(), xyz);
  if (tets) std::copy(m->tets.begin(), m->tets.end(), tets);
  if (tet_tag) std::copy(m->tet_tag.begin(), m->tet_tag.end(), tet_tag);
  if (tris) std::copy(m->tris.begin(), m->tris.end(), tris);
  if (tri_tag) std::copy(m->tri_tag.begin(), m->tri_tag.end(), tri_tag);
  return SCDAT_OK;
}

int scdat_mesh_face_groups(const scdat_mesh_data* m, int32_t n_map, const int32_t* tags, const int32_t* groups,
                           int32_t default_group, int32_t* face_group, int64_t* n_matched) {
  if (!m || !face_group || n_map < 0 || (n_map > 0 && (!tags || !groups)))
    return fail(SCDAT_ERR_ARG, "mesh_face_groups: bad arguments");
  using Key = std::array<int32_t, 3>;
  auto key_of = [](int32_t a, int32_t b, int32_t c) {
    Key k = {a, b, c};
    std::sort(k.begin(), k.end());
    return k;
  };
  std::map<Key, int32_t> tri_tag;
  for (size_t t = 0; t < m->tri_tag.size(); ++t)
    tri_tag[key_of(m->tris[3 * t], m->tris[3 * t + 1], m->tris[3 * t + 2])] = m->tri_tag[t];
  std::map<int32_t, int32_t> tag_group;
  for (int32_t k = 0; k < n_map; ++k) tag_group[tags[k]] = groups[k];
  // faces shared by two tetrahedra are interior
  const size_t nt = m->tet_tag.size();
  std::map<Key, int> count;
  for (size_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      int32_t v[3], q = 0;
      for (int r = 0; r < 4; ++r)
        if (r != k) v[q++] = m->tets[4 * t + r];
      ++count[key_of(v[0], v[1], v[2])];
    }
  int64_t matched = 0;
  for (size_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      int32_t v[3], q = 0;
      for (int r = 0; r < 4; ++r)
        if (r != k) v[q++] = m->tets[4 * t + r];
      const Key key = key_of(v[0], v[1], v[2]);
      int32_t g = -1;
      if (count[key] == 1) {
        g = default_group;
        const auto it = tri_tag.find(key);
        if (it != tri_tag.end()) {
          const auto jt = tag_group.find(it->second);
          if (jt != tag_group.end()) {
            g = jt->second;
            ++matched;
          }
        }
      }
      face_group[4 * t + k] = g;
    }
  if (n_matched) *n_matched = matched;
  return SCDAT_OK;
}

int scdat_write_vtk(const char* path, int64_t n_nodes, const double* xyz, int64_t n_tets, const int32_t* tets,
                    int32_t n_fields, const char* const* names, const int32_t* components, const double* const* data) {
  if (!path || !xyz || (n_tets > 0 && !tets) || n_fields < 0 ||
      (n_fields > 0 && (!names || !components || !data)))
    return fail(SCDAT_ERR_ARG, "write_vtk: bad arguments");
  for (int32_t f = 0; f < n_fields; ++f)
    if (!names[f] || !data[f] || (components[f] != 1 && components[f] != 3))
      return fail(SCDAT_ERR_ARG, "write_vtk: fields need a name, data and 1 or 3 components");
  std::ofstream o(path);
  if (!o) return fail(SCDAT_ERR_ARG, std::string("write_vtk: cannot open ") + path);
  o << "# vtk DataFile Version 3.0\nSCDAT field export\nASCII\nDATASET UNSTRUCTURED_GRID\n";
  o << "POINTS " << n_nodes << " double\n" << std::setprecision(17);
  for (int64_t i = 0; i < n_nodes; ++i) o << xyz[3 * i] << ' ' << xyz[3 * i + 1] << ' ' << xyz[3 * i + 2] << '\n';
  o << "CELLS " << n_tets << ' ' << 5 * n_tets << '\n';
  for (int64_t t = 0; t < n_tets; ++t)
    o << "4 " << tets[4 * t] << ' ' << tets[4 * t + 1] << ' ' << tets[4 * t + 2] << ' ' << tets[4 * t + 3] << '\n';
  o << "CELL_TYPES " << n_tets << '\n';
  for (int64_t t = 0; t < n_tets; ++t) o << "10\n";
  if (n_fields > 0) o << "POINT_DATA " << n_nodes << '\n';
  for (int32_t f = 0; f < n_fields; ++f) {
    if (components[f] == 1) {
      o << "SCALARS " << names[f] << " double 1\nLOOKUP_TABLE default\n";
      for (int64_t i = 0; i < n_nodes; ++i) o << data[f][i] << '\n';
    } else {
      o << "VECTORS " << names[f] << " double\n";
      for (int64_t i = 0; i < n_nodes; ++i)
        o << data[f][3 * i] << ' ' << data[f][3 * i + 1] << ' ' << data[f][3 * i + 2] << '\n';
    }
  }
  return o.good() ? SCDAT_OK : fail(SCDAT_ERR_ARG, "write_vtk: write failed");
}

int scdat_export_columns(const char* csv_path, const char* module, const char* axis_name, int64_t n_rows,
                         const double* axis, int32_t n_series, const char* const* names, const double* const* series) {
  if (!csv_path || !axis_name || n_rows < 0 || (n_rows > 0 && !axis) || n_series < 0 ||
      (n_series > 0 && (!names || !series)))
    return fail(SCDAT_ERR_ARG, "export_columns: bad arguments");
  for (int64_t r = 0; r < n_rows; ++r)
    if (!std::isfinite(axis[r])) return fail(SCDAT_ERR_ARG, "export_columns: non-finite data set");
  std::map<std::string, const double*> cols;  // name order, as a std::map keyed data set
  for (int32_t s = 0; s < n_series; ++s) {
    if (!names[s] || !series[s]) return fail(SCDAT_ERR_ARG, "export_columns: null series");
    for (int64_t r = 0; r < n_rows; ++r)
      if (!std::isfinite(series[s][r])) return fail(SCDAT_ERR_ARG, "export_columns: non-finite data set");
    cols[names[s]] = series[s];
  }
  const std::string csv = csv_path;
  const auto dot = csv.find_last_of('.'), slash = csv.find_last_of('/');
  const std::string stem = (dot != std::string::npos && (slash == std::string::npos || dot > slash)) ? csv.substr(0, dot)
                                                                                                     : csv;
  const std::string vtk = stem + ".vtk", h5 = stem + ".h5";
  std::map<std::string, std::string> meta;
  meta["module"] = module && *module ? module : "unknown";
  meta["output_contract_schema"] = "scdat.output.contract.v1";
  meta["output_contract_version"] = "v1";
  meta["output_axis_name"] = axis_name;
  meta["output_row_count"] = std::to_string(n_rows);
  meta["output_series_count"] = std::to_string(cols.size());
  std::string joined;
  for (const auto& c : cols) joined += (joined.empty() ? "" : ";") + c.first;
  meta["output_series_names"] = joined;
  meta["output_artifact_csv"] = csv;
  meta["output_artifact_vtk"] = vtk;
  meta["output_artifact_h5"] = h5;
  meta["output_export_utc"] = utc_now();
  {
    std::ofstream o(csv);
    if (!o) return fail(SCDAT_ERR_ARG, "export_columns: cannot open " + csv);
    o << axis_name;
    for (const auto& c : cols) o << ',' << c.first;
    o << '\n' << std::scientific << std::setprecision(8);
    for (int64_t r = 0; r < n_rows; ++r) {
      o << axis[r];
      for (const auto& c : cols) o << ',' << c.second[r];
      o << '\n';
    }
  }
  {
    std::ofstream o(vtk);
    if (!o) return fail(SCDAT_ERR_ARG, "export_columns: cannot open " + vtk);
    o << "# vtk DataFile Version 3.0\nSCDAT columnar export\nASCII\nDATASET POLYDATA\n";
    o << "POINTS " << n_rows << " float\n" << std::scientific << std::setprecision(8);
    for (int64_t r = 0; r < n_rows; ++r) o << "0 0 " << axis[r] << '\n';
    o << "VERTICES " << n_rows << ' ' << 2 * n_rows << '\n';
    for (int64_t r = 0; r < n_rows; ++r) o << "1 " << r << '\n';
    o << "POINT_DATA " << n_rows << '\n';
    for (const auto& c : cols) {
      o << "SCALARS " << c.first << " float 1\nLOOKUP_TABLE default\n";
      for (int64_t r = 0; r < n_rows; ++r) o << c.second[r] << '\n';
    }
  }
  {
    std::ofstream o(h5);
    if (!o) return fail(SCDAT_ERR_ARG, "export_columns: cannot open " + h5);
    o << "# SCDAT HDF5-like archive\n[metadata]\naxis_name=" << axis_name << '\n';
    for (const auto& kv : meta) o << kv.first << '=' << kv.second << '\n';
    o << "[dataset:" << axis_name << "]\n" << std::scientific << std::setprecision(8);
    for (int64_t r = 0; r < n_rows; ++r) o << axis[r] << '\n';
    for (const auto& c : cols) {
      o << "[dataset:" << c.first << "]\n";
      for (int64_t r = 0; r < n_rows; ++r) o << c.second[r] << '\n';
    }
  }
  const std::pair<std::string, const char*> arts[3] = {{csv, "csv"}, {vtk, "vtk"}, {h5, "h5"}};
  for (const auto& a : arts) {
    std::ofstream o(a.first + ".metadata.json");
    if (!o) return fail(SCDAT_ERR_ARG, "export_columns: cannot open the metadata sidecar");
    o << "{\n  \"output_contract_schema\": \"" << json_escape(meta["output_contract_schema"]) << "\",\n";
    o << "  \"output_contract_version\": \"" << json_escape(meta["output_contract_version"]) << "\",\n";
    o << "  \"artifact_format\": \"" << a.second << "\",\n";
    o << "  \"artifact_path\": \"" << json_escape(a.first) << "\",\n  \"metadata\": {\n";
    size_t i = 0;
    for (const auto& kv : meta) {
      o << "    \"" << json_escape(kv.first) << "\": \"" << json_escape(kv.second) << "\"";
      o << (++i < meta.size() ? ",\n" : "\n");
    }
    o << "  }\n}\n";
  }
  return SCDAT_OK;
}

}  // extern "C"
