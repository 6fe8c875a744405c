// Host-side mesh preprocessing: structured box mesher with a conforming 6-tet
// split, face neighbours, periodic partners and node adjacency.
//
// The reference box mesher (Tools/Mesh/src/MeshParsing.cpp:1208-1261) numbers
// nodes i*(ny+1)*(nz+1) + j*(nz+1) + k and splits every hex into 5 tets with
// one fixed pattern, which does not conform across neighbouring hexes, and it
// never fills element neighbours (Tools/Mesh/include/MeshAlgorithms.h:193).
// This mesher keeps the same lexicographic node order (k fastest) but uses
// the Kuhn (Freudenthal) split into 6 tets around the v000-v111 diagonal,
// which conforms across hexes, so a face-neighbour walk can relocate
// particles in O(1) per crossing.
#include "mesh_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <tuple>

namespace scdat {

bool box_axis(const BoxSpec& s, int ax, std::vector<double>& c, std::string& err) {
  const int n = s.n[ax];
  const double lo = s.lo[ax], hi = s.hi[ax];
  if (n < 1 || !(hi > lo)) {
    err = "box: invalid axis extent or cell count";
    return false;
  }
  c.clear();
  if (!s.conductor) {
    for (int i = 0; i <= n; ++i) c.push_back(lo + (hi - lo) * (double)i / n);
    c.back() = hi;
    return true;
  }
  const double a = s.c_lo[ax], b = s.c_hi[ax];
  if (!(a > lo && b > a && hi > b)) {
    err = "box: conductor must lie strictly inside the domain";
    return false;
  }
  const double L = hi - lo;
  int na = (int)std::lround(n * (a - lo) / L);
  int nb = (int)std::lround(n * (b - a) / L);
  na = std::max(na, 1);
  nb = std::max(nb, 1);
  int nc = n - na - nb;
  if (nc < 1) {
    err = "box: too few cells to resolve the conductor";
    return false;
  }
  for (int i = 0; i < na; ++i) c.push_back(lo + (a - lo) * (double)i / na);
  for (int i = 0; i < nb; ++i) c.push_back(a + (b - a) * (double)i / nb);
  for (int i = 0; i <= nc; ++i) c.push_back(b + (hi - b) * (double)i / nc);
  c.back() = hi;
  return true;
}

namespace {
// Kuhn split: tet for axis permutation (p0,p1,p2) is v000, +e_p0, +e_p0+e_p1, v111.
const int kPerm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

double orient(const double* a, const double* b, const double* c, const double* d) {
  const double u[3] = {b[0] - a[0], b[1] - a[1], b[2] - a[2]};
  const double v[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
  const double w[3] = {d[0] - a[0], d[1] - a[1], d[2] - a[2]};
  return u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) +
         u[2] * (v[0] * w[1] - v[1] * w[0]);
}
}  // namespace

bool box_mesh(const BoxSpec& s, std::vector<double>& xyz, std::vector<int32_t>& tets,
              std::vector<int32_t>& kind, std::vector<int32_t>& group, std::string& err) {
  std::vector<double> cx, cy, cz;
  if (!box_axis(s, 0, cx, err) || !box_axis(s, 1, cy, err) || !box_axis(s, 2, cz, err))
    return false;
  const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
  // Conductor index ranges [ia0, ia1) in hex units.
  int ic[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
  if (s.conductor) {
    const std::vector<double>* cc[3] = {&cx, &cy, &cz};
    for (int ax = 0; ax < 3; ++ax) {
      const auto& c = *cc[ax];
      for (int i = 0; i < (int)c.size(); ++i) {
        if (std::abs(c[i] - s.c_lo[ax]) < 1e-12 * (1 + std::abs(c[i]))) ic[ax][0] = i;
        if (std::abs(c[i] - s.c_hi[ax]) < 1e-12 * (1 + std::abs(c[i]))) ic[ax][1] = i;
      }
      if (ic[ax][0] < 0 || ic[ax][1] <= ic[ax][0]) {
        err = "box: conductor faces not on grid lines";
        return false;
      }
    }
  }
  auto in_cond_hex = [&](int i, int j, int k) {
    return s.conductor && i >= ic[0][0] && i < ic[0][1] && j >= ic[1][0] && j < ic[1][1] &&
           k >= ic[2][0] && k < ic[2][1];
  };
  // Node strictly inside the conductor (not on its surface) is removed.
  auto node_removed = [&](int i, int j, int k) {
    return s.conductor && i > ic[0][0] && i < ic[0][1] && j > ic[1][0] && j < ic[1][1] &&
           k > ic[2][0] && k < ic[2][1];
  };
  auto node_on_cond = [&](int i, int j, int k) {
    return s.conductor && i >= ic[0][0] && i <= ic[0][1] && j >= ic[1][0] && j <= ic[1][1] &&
           k >= ic[2][0] && k <= ic[2][1] && !node_removed(i, j, k);
  };
  const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
  std::vector<int32_t> id((size_t)(NX * NY * NZ), -1);
  xyz.clear();
  kind.clear();
  int32_t next = 0;
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY; ++j)
      for (int k = 0; k < NZ; ++k) {
        if (node_removed(i, j, k)) continue;
        id[(size_t)((i * NY + j) * NZ + k)] = next++;
        xyz.push_back(cx[i]);
        xyz.push_back(cy[j]);
        xyz.push_back(cz[k]);
        const bool wall = i == 0 || i == nx || j == 0 || j == ny || k == 0 || k == nz;
        kind.push_back(wall ? 1 : (node_on_cond(i, j, k) ? 2 : 0));
      }
  tets.clear();
  group.clear();
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        if (in_cond_hex(i, j, k)) continue;
        for (int t = 0; t < 6; ++t) {
          int ijk[4][3];
          int cur[3] = {i, j, k};
          for (int r = 0; r < 3; ++r) ijk[0][r] = cur[r];
          for (int st = 0; st < 3; ++st) {
            cur[kPerm[t][st]] += 1;
            for (int r = 0; r < 3; ++r) ijk[st + 1][r] = cur[r];
          }
          int32_t v[4];
          for (int q = 0; q < 4; ++q)
            v[q] = id[(size_t)((ijk[q][0] * NY + ijk[q][1]) * NZ + ijk[q][2])];
          const double* p[4];
          for (int q = 0; q < 4; ++q) p[q] = &xyz[3 * (size_t)v[q]];
          if (orient(p[0], p[1], p[2], p[3]) < 0) {
            std::swap(v[2], v[3]);
            std::swap(ijk[2], ijk[3]);
          }
          for (int q = 0; q < 4; ++q) tets.push_back(v[q]);
          // Boundary groups per face (opposite vertex f).
          for (int f = 0; f < 4; ++f) {
            int g = -1;
            int fv[3], m = 0;
            for (int q = 0; q < 4; ++q)
              if (q != f) fv[m++] = q;
            for (int ax = 0; ax < 3 && g < 0; ++ax) {
              const int a0 = ijk[fv[0]][ax], a1 = ijk[fv[1]][ax], a2 = ijk[fv[2]][ax];
              if (a0 != a1 || a1 != a2) continue;
              const int nax = ax == 0 ? nx : (ax == 1 ? ny : nz);
              if (a0 == 0) g = 2 * ax;
              else if (a0 == nax) g = 2 * ax + 1;
              else if (s.conductor) {
                // Face on a conductor plane, inside the conductor's extent.
                int oth[2], mm = 0;
                for (int b = 0; b < 3; ++b)
                  if (b != ax) oth[mm++] = b;
                bool inside = true;
                for (int b2 = 0; b2 < 2; ++b2) {
                  const int bx = oth[b2];
                  for (int q = 0; q < 3; ++q) {
                    const int c = ijk[fv[q]][bx];
                    inside = inside && c >= ic[bx][0] && c <= ic[bx][1];
                  }
                }
                if (inside && (a0 == ic[ax][0] || a0 == ic[ax][1])) {
                  // The face is a conductor face only if the hex on the other
                  // side is a conductor hex: this hex is outside by construction.
                  g = 6 + 2 * ax + (a0 == ic[ax][1] ? 1 : 0);
                }
              }
            }
            group.push_back(g);
          }
        }
      }
  return true;
}

void face_neighbors(int64_t nt, const int32_t* tets, std::vector<int32_t>& nb) {
  struct F { int32_t a, b, c; int32_t t; int8_t k; };
  std::vector<F> f((size_t)nt * 4);
  for (int64_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      int32_t v[3];
      int m = 0;
      for (int j = 0; j < 4; ++j)
        if (j != k) v[m++] = tets[4 * t + j];
      std::sort(v, v + 3);
      f[4 * t + k] = {v[0], v[1], v[2], (int32_t)t, (int8_t)k};
    }
  std::sort(f.begin(), f.end(), [](const F& x, const F& y) {
    return std::tie(x.a, x.b, x.c) < std::tie(y.a, y.b, y.c);
  });
  nb.assign((size_t)nt * 4, -1);
  for (size_t i = 0; i + 1 < f.size(); ++i) {
    if (f[i].a == f[i + 1].a && f[i].b == f[i + 1].b && f[i].c == f[i + 1].c) {
      nb[4 * (size_t)f[i].t + f[i].k] = f[i + 1].t;
      nb[4 * (size_t)f[i + 1].t + f[i + 1].k] = f[i].t;
      ++i;
    }
  }
}

int periodic_partners(int64_t nn, const double* xyz, int64_t nt, const int32_t* tets,
                      const int32_t* group, int ga, int gb, const double* shift,
                      std::vector<int32_t>& partner) {
  (void)nn;
  if (partner.size() != (size_t)nt * 4) partner.assign((size_t)nt * 4, -1);
  // Key: face centroid quantised on a fine lattice after translation.
  using Key = std::tuple<int64_t, int64_t, int64_t>;
  std::map<Key, int64_t> ma;
  auto key = [&](int64_t t, int k, const double* sh) {
    double c[3] = {0, 0, 0};
    for (int j = 0; j < 4; ++j)
      if (j != k)
        for (int r = 0; r < 3; ++r) c[r] += xyz[3 * (size_t)tets[4 * t + j] + r] / 3.0;
    for (int r = 0; r < 3; ++r) c[r] += sh[r];
    return Key(std::llround(c[0] * 1e9), std::llround(c[1] * 1e9), std::llround(c[2] * 1e9));
  };
  const double zero[3] = {0, 0, 0};
  for (int64_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k)
      if (group[4 * t + k] == ga) ma[key(t, k, shift)] = 4 * t + k;
  int matched = 0;
  for (int64_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k)
      if (group[4 * t + k] == gb) {
        auto it = ma.find(key(t, k, zero));
        if (it == ma.end()) continue;
        const int64_t fa = it->second;
        partner[fa] = (int32_t)t;               // leaving through ga enters t
        partner[4 * t + k] = (int32_t)(fa / 4); // leaving through gb enters fa's tet
        ++matched;
      }
  return matched;
}

void node_adjacency(int64_t nn, int64_t nt, const int32_t* tets, std::vector<int64_t>& rp,
                    std::vector<int32_t>& ci) {
  std::vector<int64_t> cnt((size_t)nn + 1, 0);
  // Upper bound per node: 4 entries per incident tet.
  for (int64_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) cnt[tets[4 * t + k] + 1] += 4;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> tmp((size_t)cnt[nn]);
  std::vector<int64_t> fill(cnt.begin(), cnt.end() - 1);
  for (int64_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      const int32_t a = tets[4 * t + k];
      for (int j = 0; j < 4; ++j) tmp[fill[a]++] = tets[4 * t + j];
    }
  rp.assign((size_t)nn + 1, 0);
  ci.clear();
  ci.reserve((size_t)cnt[nn] / 3);
  for (int64_t i = 0; i < nn; ++i) {
    auto b = tmp.begin() + cnt[i], e = tmp.begin() + fill[i];
    std::sort(b, e);
    e = std::unique(b, e);
    ci.push_back((int32_t)i);
    for (auto it = b; it != e; ++it)
      if (*it != i) ci.push_back(*it);
    rp[i + 1] = (int64_t)ci.size();
  }
}

}  // namespace scdat
