// Harness over the reference's mesh loaders (GmshLoader, VTKLoader,
// STLLoader, Tools/Mesh/src/MeshParsing.cpp:1096-1125) for golden vectors of
// the repository's readers.  Test infrastructure only.
#include <cstdint>
#include <exception>
#include <memory>
#include <string>

#include "Tools/Mesh/include/MeshParsing.h"

using namespace SCDAT::Mesh;

extern "C" {

// format 1 Gmsh, 2 VTK, 3 STL.  Fills up to the capacities; returns 0 ok,
// -1 on a loader error, -2 when a capacity is too small.  Elements: type
// (3 triangle, 4 tetrahedron), 4 node indices (-1 padded), material id.
int ref_mesh_read(const char* path, int format, int64_t cap_nodes, int64_t cap_el, int64_t* n_nodes,
                  int64_t* n_el, double* xyz, int32_t* etype, int32_t* enodes, int32_t* etag) {
  try {
    MeshLoadOptions opt;
    opt.validate_mesh = false;
    std::unique_ptr<VolMesh> m;
    if (format == 1) m = GmshLoader().loadMeshFile(path, opt);
    else if (format == 2) m = VTKLoader().loadMeshFile(path, opt);
    else m = STLLoader().loadMeshFile(path, opt);
    *n_nodes = (int64_t)m->getNodeCount();
    const auto& rec = m->getElementRecords();
    *n_el = (int64_t)rec.size();
    if (*n_nodes > cap_nodes || *n_el > cap_el) return -2;
    for (int64_t i = 0; i < *n_nodes; ++i) {
      const auto& p = m->getNodes()[i]->getPosition();
      xyz[3 * i] = p.x();
      xyz[3 * i + 1] = p.y();
      xyz[3 * i + 2] = p.z();
    }
    for (int64_t e = 0; e < *n_el; ++e) {
      etype[e] = (int32_t)rec[e].type;
      for (int k = 0; k < 4; ++k)
        enodes[4 * e + k] = k < (int)rec[e].node_indices.size() ? (int32_t)rec[e].node_indices[k] : -1;
      etag[e] = rec[e].material_id;
    }
    return 0;
  } catch (const std::exception&) {
    return -1;
  }
}

}  // extern "C"
