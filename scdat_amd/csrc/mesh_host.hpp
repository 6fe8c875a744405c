// Host-side mesh preprocessing for the PIC-MCC device path.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace scdat {

struct BoxSpec {
  double lo[3], hi[3];
  int32_t n[3];
  int32_t conductor;
  double c_lo[3], c_hi[3];
};

// Per-axis grid coordinates: the conductor box faces fall on grid lines.
bool box_axis(const BoxSpec& s, int ax, std::vector<double>& coord, std::string& err);

// Structured box mesh, 6-tet Kuhn split of every hex, conductor hexes removed.
// node_kind: 0 free, 1 Dirichlet outer wall, 2 conductor surface.
// face_group: 4 per tet (face k opposite local vertex k); -1 interior,
// 0..5 outer walls (xmin,xmax,ymin,ymax,zmin,zmax), 6..11 conductor faces.
bool box_mesh(const BoxSpec& s, std::vector<double>& xyz, std::vector<int32_t>& tets,
              std::vector<int32_t>& kind, std::vector<int32_t>& group, std::string& err);

// Face-neighbour table by sorted vertex triples (face k opposite vertex k).
void face_neighbors(int64_t nt, const int32_t* tets, std::vector<int32_t>& nb);

// Partner tets across two opposite boundary groups (periodic pair): faces of
// group ga matched with faces of group gb translated by shift.
int periodic_partners(int64_t nn, const double* xyz, int64_t nt, const int32_t* tets,
                      const int32_t* group, int ga, int gb, const double* shift,
                      std::vector<int32_t>& partner);

// Node adjacency including self, self first then ascending.
void node_adjacency(int64_t nn, int64_t nt, const int32_t* tets, std::vector<int64_t>& rp,
                    std::vector<int32_t>& ci);

}  // namespace scdat
