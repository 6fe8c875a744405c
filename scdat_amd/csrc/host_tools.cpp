// Host-side tools of the C ABI: counter-based RNG, analytic locate on the
// structured box mesh, and the synthetic Maxwellian loader.
//
// The reference loads particles with a random_device-seeded thread_local
// mt19937 for positions (Tools/Mesh/src/MeshParsing.cpp:87-127) and
// std::normal_distribution(mt19937(42)) for velocities
// (Main/main_PICcore.cpp:189-229), which is not reproducible across runs.
// Here every particle's state is a pure function of (seed, stream, particle
// index) through Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random
// numbers: as easy as 1, 2, 3", SC'11), so the device path and the CPU
// checker receive identical inputs at any size.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/scdat_amd.h"
#include "mesh_host.hpp"

namespace scdat {

struct Philox {
  static void round(uint32_t c[4], const uint32_t k[2]) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
  }
  static void gen(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; ++r) {
      if (r) {
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
      }
      round(c, k);
    }
    std::memcpy(out, c, sizeof(c));
  }
};

// (0,1) uniform from 32 bits: (u + 0.5) * 2^-32.
static inline double u01(uint32_t u) { return ((double)u + 0.5) * 2.3283064365386963e-10; }

bool box_locate(const BoxSpec& s, int64_t n, const double* x, int32_t* cell, std::string& err) {
  std::vector<double> ax[3];
  for (int r = 0; r < 3; ++r)
    if (!box_axis(s, r, ax[r], err)) return false;
  const int nx = s.n[0], ny = s.n[1], nz = s.n[2];
  int ic[3][2] = {{-1, -1}, {-1, -1}, {-1, -1}};
  if (s.conductor)
    for (int r = 0; r < 3; ++r)
      for (int i = 0; i < (int)ax[r].size(); ++i) {
        if (std::abs(ax[r][i] - s.c_lo[r]) < 1e-12 * (1 + std::abs(ax[r][i]))) ic[r][0] = i;
        if (std::abs(ax[r][i] - s.c_hi[r]) < 1e-12 * (1 + std::abs(ax[r][i]))) ic[r][1] = i;
      }
  auto in_cond = [&](int i, int j, int k) {
    return s.conductor && i >= ic[0][0] && i < ic[0][1] && j >= ic[1][0] && j < ic[1][1] &&
           k >= ic[2][0] && k < ic[2][1];
  };
  // rank of each hex among the kept hexes, in mesher order (k fastest)
  std::vector<int32_t> rank((size_t)nx * ny * nz, -1);
  int32_t r = 0;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k)
        if (!in_cond(i, j, k)) rank[((size_t)i * ny + j) * nz + k] = r++;
  // Kuhn tet index for the axis order of the local coordinates (descending),
  // matching the permutation order used by the mesher.
  const int perm_index[3][3][3] = {
      {{-1, -1, -1}, {-1, -1, 0}, {-1, 1, -1}},
      {{-1, -1, 2}, {-1, -1, -1}, {3, -1, -1}},
      {{-1, 4, -1}, {5, -1, -1}, {-1, -1, -1}},
  };
  for (int64_t p = 0; p < n; ++p) {
    int h[3];
    double u[3];
    bool ok = true;
    for (int a = 0; a < 3; ++a) {
      const auto& c = ax[a];
      const double v = x[3 * p + a];
      if (!(v >= c.front() && v <= c.back())) {
        ok = false;
        break;
      }
      int i = (int)(std::upper_bound(c.begin(), c.end(), v) - c.begin()) - 1;
      i = std::min(std::max(i, 0), (int)c.size() - 2);
      h[a] = i;
      u[a] = (v - c[i]) / (c[i + 1] - c[i]);
    }
    if (!ok) {
      cell[p] = -1;
      continue;
    }
    const int32_t hr = rank[((size_t)h[0] * ny + h[1]) * nz + h[2]];
    if (hr < 0) {
      cell[p] = -1;
      continue;
    }
    int o[3] = {0, 1, 2};
    std::sort(o, o + 3, [&](int a, int b) { return u[a] > u[b] || (u[a] == u[b] && a < b); });
    cell[p] = 6 * hr + perm_index[o[0]][o[1]][o[2]];
  }
  return true;
}

}  // namespace scdat

using namespace scdat;

extern "C" {

void scdat_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  Philox::gen(ctr, key, out);
}

int scdat_box_locate(const scdat_box_spec* spec, int64_t n, const double* x, int32_t* cell) {
  if (!spec || n < 0 || (n > 0 && (!x || !cell))) return SCDAT_ERR_ARG;
  BoxSpec s;
  std::memcpy(&s, spec, sizeof(s));
  std::string err;
  try {
    return box_locate(s, n, x, cell, err) ? SCDAT_OK : SCDAT_ERR_ARG;
  } catch (...) {
    return SCDAT_ERR_NOMEM;
  }
}

int scdat_load_maxwellian(const scdat_box_spec* spec, int64_t n, uint64_t seed, uint32_t stream,
                          double vth, const double* drift, double* x, double* v, int32_t* cell) {
  if (!spec || n < 0 || !(vth >= 0) || (n > 0 && (!x || !v || !cell))) return SCDAT_ERR_ARG;
  const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  const double L[3] = {spec->hi[0] - spec->lo[0], spec->hi[1] - spec->lo[1], spec->hi[2] - spec->lo[2]};
  for (int64_t p = 0; p < n; ++p) {
    // positions: rejection against the conductor box, retries on counter word 2
    for (uint32_t attempt = 0;; ++attempt) {
      const uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), 2u * attempt, stream};
      uint32_t o[4];
      Philox::gen(ctr, key, o);
      double q[3];
      for (int r = 0; r < 3; ++r) q[r] = spec->lo[r] + L[r] * u01(o[r]);
      bool inside = spec->conductor != 0;
      for (int r = 0; r < 3 && inside; ++r) inside = q[r] > spec->c_lo[r] && q[r] < spec->c_hi[r];
      if (!inside || attempt > 1000) {
        for (int r = 0; r < 3; ++r) x[3 * p + r] = q[r];
        break;
      }
    }
    const uint32_t ctr[4] = {(uint32_t)p, (uint32_t)(p >> 32), 1u, stream};
    uint32_t o[4];
    Philox::gen(ctr, key, o);
    // Box-Muller: two pairs -> three normals used
    const double r1 = std::sqrt(-2.0 * std::log(u01(o[0]))), t1 = 6.283185307179586 * u01(o[1]);
    const double r2 = std::sqrt(-2.0 * std::log(u01(o[2]))), t2 = 6.283185307179586 * u01(o[3]);
    const double g[3] = {r1 * std::cos(t1), r1 * std::sin(t1), r2 * std::cos(t2)};
    for (int r = 0; r < 3; ++r) v[3 * p + r] = vth * g[r] + (drift ? drift[r] : 0.0);
  }
  return scdat_box_locate(spec, n, x, cell);
}

}  // extern "C"
