// C-ABI implementation: context, mesh setup, particle I/O and the step
// orchestration of the PIC-MCC cycle on one gfx950 GPU.
//
// Step order follows PICCycle::executeTimeStep
// (Tools/PICcore/src/PICCycle.cpp:311-416): push + gather + walk + deposit
// with the previous step's E, Poisson with the floating conductor, nodal E.
// Everything stays resident on the device; the host only polls the PCG
// convergence flag between batches of iterations.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <map>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/scdat_amd.h"
#include "device_common.hpp"
#include "field.hpp"
#include "mesh_host.hpp"
#include "push.hpp"
#include "emit.hpp"
#include "sort.hpp"
#include "adjoint.hpp"

// Scratch of scdat_get_stats_noalloc: the results of its one reduction
// kernel (8 header words, then the fixed-point charge sum of each scenario).
// One copy per GPU, held until the process ends; every call overwrites
// every word it reads back, so nothing needs clearing between calls, and
// calls that overlap on one GPU race on it.
constexpr int kStatsScratchHead = 8;
constexpr int kStatsMaxSegments = 4096;
static __device__ __align__(256) unsigned long long scratch[kStatsScratchHead + kStatsMaxSegments];

namespace scdat {

constexpr double kElem = 1.602176634e-19;  // Tools/Basic/include/Constants.h:46

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                     \
  do {                                                                                \
    hipError_t e_ = (x);                                                              \
    if (e_ != hipSuccess)                                                             \
      throw Error(SCDAT_ERR_DEVICE, std::string(#x) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 16;
    if (hipMalloc(&p, b) != hipSuccess) {
      p = nullptr;
      throw Error(SCDAT_ERR_NOMEM, "hipMalloc failed (" + std::to_string(b) + " bytes)");
    }
    bytes = b;
  }
  template <typename U>
  U* as() const { return static_cast<U*>(p); }
  template <typename U>
  void upload(const std::vector<U>& v, hipStream_t s) {
    alloc(sizeof(U) * v.size());
    if (!v.empty()) HIPCHK(hipMemcpyAsync(p, v.data(), sizeof(U) * v.size(), hipMemcpyHostToDevice, s));
  }
};

// Host-side copy of the mesh and everything derived from it.
struct HostMesh {
  int64_t nn = 0, nt = 0;
  std::vector<double> xyz, dval;
  std::vector<int32_t> tets, kind, group, nb, partner;
  std::vector<double> A, X0, vol;  // per tet, FP64
  int W = 0;
  std::vector<int32_t> col;        // slot-major W*nn
  std::vector<double> kfull, aval, gx, gy, gz, lift, b1, dinv, dscale;
  std::vector<int32_t> kc_idx;
  std::vector<double> kc_val;
  bool has_cond = false;
  // two-level preconditioner (nc = 0: none): row-block aggregates
  int nc = 0;
  std::vector<double> dco, aci;
  // aggregates each row block touches (its own first), CSR over the blocks
  std::vector<int32_t> cnb_off, cnb;
  int cnb_max = 0;
  // tiled order: tile (x + nx (y + ny z)) of each of the first tile_blocks
  // row blocks, the tile grid, and whether every tile fills its block slot
  // for slot (a tile's point (x, y, z) at row x + 8 y + 64 z of its block)
  std::vector<int32_t> tile_of_block;
  int32_t tdim[3] = {0, 0, 0};
  int64_t tile_blocks = 0;
  bool tiles_aligned = false;
};

// Coarse space of the two-level preconditioner: aggregate b = the free rows
// of row block b (256 consecutive nodes of the internal order, a compact box
// in Morton order), Ac = P^T A P on the eliminated operator, inverted once
// by Cholesky (blocks without free rows keep zero rows), and the diagonal D
// on free rows (0 on constrained rows) with which the iteration kernel forms
// A P z from the column-scaled operator.  Leaves nc = 0 for meshes too small
// to gain from it or too large for a dense coarse inverse.
static void build_coarse(HostMesh& h) {
  h.nc = 0;
  h.dco.clear();
  h.aci.clear();
  const int64_t nn = h.nn;
  int64_t nfree = 0;
  for (int64_t i = 0; i < nn; ++i) nfree += h.kind[i] == SCDAT_NODE_FREE;
  const int G = (int)((nn + kCoarseBlockRows - 1) / kCoarseBlockRows);
  if (nfree < 4096 || G > kMaxCoarse) return;
  const int W = h.W;
  std::vector<double> Ac((size_t)G * G, 0.0);
  std::vector<char> used((size_t)G, 0);
  for (int64_t i = 0; i < nn; ++i) {
    if (h.kind[i] != SCDAT_NODE_FREE) continue;
    const int bi = (int)(i / kCoarseBlockRows);
    used[bi] = 1;
    for (int sl = 0; sl < W; ++sl) {
      const size_t o = (size_t)sl * nn + i;
      const double v = h.aval[o];
      const int32_t j = h.col[o];
      if (v == 0.0 || h.kind[j] != SCDAT_NODE_FREE) continue;
      Ac[(size_t)bi * G + j / kCoarseBlockRows] += v;
    }
  }
  std::vector<int> idx;  // used blocks
  for (int b = 0; b < G; ++b)
    if (used[b]) idx.push_back(b);
  const int m = (int)idx.size();
  std::vector<double> L((size_t)m * m);
  for (int r = 0; r < m; ++r)
    for (int q = 0; q < m; ++q) L[(size_t)r * m + q] = Ac[(size_t)idx[r] * G + idx[q]];
  for (int j = 0; j < m; ++j) {  // Cholesky, lower triangle in place
    double d = L[(size_t)j * m + j];
    for (int k = 0; k < j; ++k) d -= L[(size_t)j * m + k] * L[(size_t)j * m + k];
    if (!(d > 0.0)) return;
    d = std::sqrt(d);
    L[(size_t)j * m + j] = d;
    for (int i = j + 1; i < m; ++i) {
      double t = L[(size_t)i * m + j];
      const double* li = &L[(size_t)i * m];
      const double* lj = &L[(size_t)j * m];
      for (int k = 0; k < j; ++k) t -= li[k] * lj[k];
      L[(size_t)i * m + j] = t / d;
    }
  }
  // inverse: Linv lower triangular, Ac^-1 = Linv^T Linv
  std::vector<double> Li((size_t)m * m, 0.0);
  for (int j = 0; j < m; ++j) {
    Li[(size_t)j * m + j] = 1.0 / L[(size_t)j * m + j];
    for (int i = j + 1; i < m; ++i) {
      double t = 0.0;
      for (int k = j; k < i; ++k) t -= L[(size_t)i * m + k] * Li[(size_t)k * m + j];
      Li[(size_t)i * m + j] = t / L[(size_t)i * m + i];
    }
  }
  h.aci.assign((size_t)G * G, 0.0);
  for (int r = 0; r < m; ++r)
    for (int q = 0; q <= r; ++q) {
      double t = 0.0;
      for (int k = r; k < m; ++k) t += Li[(size_t)k * m + r] * Li[(size_t)k * m + q];
      h.aci[(size_t)idx[r] * G + idx[q]] = t;
      h.aci[(size_t)idx[q] * G + idx[r]] = t;
    }
  h.dco.assign((size_t)nn, 0.0);
  for (int64_t i = 0; i < nn; ++i)
    if (h.kind[i] == SCDAT_NODE_FREE) h.dco[i] = 1.0 / h.dinv[i];
  h.cnb_off.assign((size_t)G + 1, 0);
  h.cnb.clear();
  std::vector<int32_t> seen((size_t)G, -1);
  for (int b = 0; b < G; ++b) {
    h.cnb.push_back(b);
    seen[b] = b;
    const int64_t i1 = std::min<int64_t>(nn, (int64_t)(b + 1) * kCoarseBlockRows);
    for (int64_t i = (int64_t)b * kCoarseBlockRows; i < i1; ++i) {
      if (h.kind[i] != SCDAT_NODE_FREE) continue;
      for (int sl = 0; sl < W; ++sl) {
        const size_t o = (size_t)sl * nn + i;
        const int32_t j = h.col[o];
        if (h.aval[o] == 0.0 || h.kind[j] != SCDAT_NODE_FREE) continue;
        const int c = j / kCoarseBlockRows;
        if (seen[c] != b) {
          seen[c] = b;
          h.cnb.push_back(c);
        }
      }
    }
    h.cnb_off[b + 1] = (int32_t)h.cnb.size();
    h.cnb_max = std::max(h.cnb_max, h.cnb_off[b + 1] - h.cnb_off[b]);
  }
  h.nc = G;
}

static bool solve3(double M[3][3], double r[3], double out[3]) {
  for (int p = 0; p < 3; ++p) {
    int best = p;
    for (int row = p + 1; row < 3; ++row)
      if (std::abs(M[row][p]) > std::abs(M[best][p])) best = row;
    if (std::abs(M[best][p]) < 1e-300) return false;
    if (best != p) {
      for (int c = 0; c < 3; ++c) std::swap(M[best][c], M[p][c]);
      std::swap(r[best], r[p]);
    }
    for (int row = p + 1; row < 3; ++row) {
      const double f = M[row][p] / M[p][p];
      for (int c = p; c < 3; ++c) M[row][c] -= f * M[p][c];
      r[row] -= f * r[p];
    }
  }
  for (int row = 2; row >= 0; --row) {
    double s = r[row];
    for (int c = row + 1; c < 3; ++c) s -= M[row][c] * out[c];
    out[row] = s / M[row][row];
  }
  return true;
}

// Builds tet maps, ELL structure, operator, elimination, conductor vectors
// and the least-squares gradient weights.
static void prepare_mesh(HostMesh& h, int op, double eps) {
  const int64_t nn = h.nn, nt = h.nt;
  face_neighbors(nt, h.tets.data(), h.nb);
  for (int64_t f = 0; f < 4 * nt; ++f)
    if (h.nb[f] >= 0) h.group[f] = -1;
    else if (h.group[f] < 0 || h.group[f] >= kMaxGroups) h.group[f] = 0;
  h.partner.assign(4 * nt, -1);
  h.A.assign(9 * nt, 0);
  h.X0.assign(3 * nt, 0);
  h.vol.assign(nt, 0);
  for (int64_t t = 0; t < nt; ++t) {
    const double* p[4];
    for (int k = 0; k < 4; ++k) p[k] = &h.xyz[3 * (size_t)h.tets[4 * t + k]];
    double J[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) J[r][c] = p[c + 1][r] - p[0][r];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    if (!(std::abs(det) > 0)) throw Error(SCDAT_ERR_ARG, "degenerate tetrahedron " + std::to_string(t));
    h.vol[t] = std::abs(det) / 6.0;
    double* a = &h.A[9 * t];
    const double id = 1.0 / det;
    a[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
    a[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
    a[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
    a[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
    a[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
    a[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
    a[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
    a[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
    a[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
    for (int r = 0; r < 3; ++r) h.X0[3 * t + r] = p[0][r];
  }
  // ELL structure from node adjacency (self first).
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  node_adjacency(nn, nt, h.tets.data(), rp, ci);
  int W = 0;
  for (int64_t i = 0; i < nn; ++i) W = std::max<int>(W, (int)(rp[i + 1] - rp[i]));
  h.W = W;
  h.col.assign((size_t)W * nn, 0);
  for (int64_t i = 0; i < nn; ++i)
    for (int s = 0; s < W; ++s) {
      const int64_t k = rp[i] + s;
      h.col[(size_t)s * nn + i] = k < rp[i + 1] ? ci[k] : (int32_t)i;
    }
  auto slot = [&](int64_t i, int32_t j) -> int64_t {
    const int64_t len = rp[i + 1] - rp[i];
    for (int s = 0; s < len; ++s)
      if (ci[rp[i] + s] == j) return (int64_t)s * nn + i;
    return -1;
  };
  h.kfull.assign((size_t)W * nn, 0.0);
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t* nd = &h.tets[4 * t];
    const double V = h.vol[t];
    if (op == 0) {
      // P1 stiffness, MultiBoundaryElectricFieldSolver.cpp:1704-1759,1870-1894.
      const double* a = &h.A[9 * t];
      double g[4][3];
      for (int r = 0; r < 3; ++r) {
        g[1][r] = a[r];
        g[2][r] = a[3 + r];
        g[3][r] = a[6 + r];
        g[0][r] = -(g[1][r] + g[2][r] + g[3][r]);
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          h.kfull[slot(nd[i], nd[j])] +=
              eps * V * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
    } else {
      // Graph Laplacian, PoissonSolver.cpp:276-327.
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
          const double* pi = &h.xyz[3 * (size_t)nd[i]];
          const double* pj = &h.xyz[3 * (size_t)nd[j]];
          double d2 = 0;
          for (int r = 0; r < 3; ++r) d2 += (pi[r] - pj[r]) * (pi[r] - pj[r]);
          const double d = std::max(std::sqrt(d2), 1e-12);
          const double w = eps * (V / 6.0) / (d * d);
          h.kfull[slot(nd[i], nd[i])] += w;
          h.kfull[slot(nd[j], nd[j])] += w;
          h.kfull[slot(nd[i], nd[j])] -= w;
          h.kfull[slot(nd[j], nd[i])] -= w;
        }
    }
  }
  // Nodes that no tetrahedron uses (padding between batched scenarios) get a
  // unit row and are held at 0 V.
  for (int64_t i = 0; i < nn; ++i)
    if (rp[i + 1] - rp[i] == 1) {
      h.kfull[i] = 1.0;
      h.kind[i] = SCDAT_NODE_DIRICHLET;
      h.dval[i] = 0.0;
    }
  // Symmetric elimination of Dirichlet and conductor nodes.
  h.aval = h.kfull;
  h.lift.assign(nn, 0.0);
  h.b1.assign(nn, 0.0);
  h.dinv.assign(nn, 1.0);
  h.dscale.assign(nn, 1.0);
  h.has_cond = false;
  for (int64_t i = 0; i < nn; ++i) {
    if (h.kind[i] != SCDAT_NODE_FREE) {
      // Constrained row: K_ii x_i = K_ii g_i keeps the row on the operator's
      // own scale, so a relative residual test weighs all rows alike.
      const double kii = h.kfull[i];
      if (!(kii > 0)) throw Error(SCDAT_ERR_ARG, "operator diagonal not positive at node " + std::to_string(i));
      for (int s = 0; s < W; ++s) h.aval[(size_t)s * nn + i] = 0.0;
      h.aval[i] = kii;  // slot 0 is self
      h.dscale[i] = kii;
      h.dinv[i] = 1.0 / kii;
      if (h.kind[i] == SCDAT_NODE_CONDUCTOR) {
        h.b1[i] = kii;
        h.has_cond = true;
      }
      continue;
    }
    const int64_t len = rp[i + 1] - rp[i];
    for (int s = 1; s < len; ++s) {
      const size_t o = (size_t)s * nn + i;
      const int32_t j = h.col[o];
      if (h.kind[j] == SCDAT_NODE_DIRICHLET) {
        h.lift[i] -= h.kfull[o] * h.dval[j];
        h.aval[o] = 0.0;
      } else if (h.kind[j] == SCDAT_NODE_CONDUCTOR) {
        h.b1[i] -= h.kfull[o];
        h.aval[o] = 0.0;
      }
    }
    const double d = h.aval[i];
    if (!(d > 0)) throw Error(SCDAT_ERR_ARG, "operator diagonal not positive at node " + std::to_string(i));
    h.dinv[i] = 1.0 / d;
  }
  // kc = sum of the unconstrained conductor rows of K.
  h.kc_idx.clear();
  h.kc_val.clear();
  if (h.has_cond) {
    std::vector<double> kc(nn, 0.0);
    std::vector<char> touched(nn, 0);
    for (int64_t i = 0; i < nn; ++i) {
      if (h.kind[i] != SCDAT_NODE_CONDUCTOR) continue;
      const int64_t len = rp[i + 1] - rp[i];
      for (int s = 0; s < len; ++s) {
        const size_t o = (size_t)s * nn + i;
        kc[h.col[o]] += h.kfull[o];
        touched[h.col[o]] = 1;
      }
    }
    for (int64_t j = 0; j < nn; ++j)
      if (touched[j]) {
        h.kc_idx.push_back((int32_t)j);
        h.kc_val.push_back(kc[j]);
      }
  }
  // Least-squares gradient weights, PoissonSolver.cpp:474-529,749-787.
  h.gx.assign((size_t)W * nn, 0.0);
  h.gy.assign((size_t)W * nn, 0.0);
  h.gz.assign((size_t)W * nn, 0.0);
  for (int64_t i = 0; i < nn; ++i) {
    const double* xi = &h.xyz[3 * (size_t)i];
    const int64_t len = rp[i + 1] - rp[i];
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int s = 1; s < len; ++s) {
      const double* xj = &h.xyz[3 * (size_t)h.col[(size_t)s * nn + i]];
      const double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] += d[r] * d[c];
    }
    for (int s = 1; s < len; ++s) {
      const size_t o = (size_t)s * nn + i;
      const double* xj = &h.xyz[3 * (size_t)h.col[o]];
      double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
      double Mc[3][3];
      std::memcpy(Mc, M, sizeof(M));
      double w[3] = {0, 0, 0};
      if (!solve3(Mc, d, w)) w[0] = w[1] = w[2] = 0;
      h.gx[o] = w[0];
      h.gy[o] = w[1];
      h.gz[o] = w[2];
    }
  }
}

// Host bucket grid for locating particles uploaded without a cell.
static void host_locate(const HostMesh& h, int64_t n, const double* x, std::vector<int32_t>& cell) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int64_t i = 0; i < h.nn; ++i)
    for (int r = 0; r < 3; ++r) {
      lo[r] = std::min(lo[r], h.xyz[3 * i + r]);
      hi[r] = std::max(hi[r], h.xyz[3 * i + r]);
    }
  const int64_t nb = std::max<int64_t>(1, (int64_t)std::cbrt((double)h.nt / 4.0));
  const int64_t B = std::min<int64_t>(nb, 256);
  auto bidx = [&](double v, int r) {
    int64_t b = (int64_t)((v - lo[r]) / (hi[r] - lo[r] + 1e-300) * B);
    return std::min<int64_t>(std::max<int64_t>(b, 0), B - 1);
  };
  std::vector<std::vector<int32_t>> bucket((size_t)(B * B * B));
  for (int64_t t = 0; t < h.nt; ++t) {
    int64_t bl[3] = {B, B, B}, bh[3] = {-1, -1, -1};
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) {
        const int64_t b = bidx(h.xyz[3 * (size_t)h.tets[4 * t + k] + r], r);
        bl[r] = std::min(bl[r], b);
        bh[r] = std::max(bh[r], b);
      }
    for (int64_t a = bl[0]; a <= bh[0]; ++a)
      for (int64_t b = bl[1]; b <= bh[1]; ++b)
        for (int64_t c = bl[2]; c <= bh[2]; ++c) bucket[(a * B + b) * B + c].push_back((int32_t)t);
  }
  cell.assign(n, -1);
  for (int64_t i = 0; i < n; ++i) {
    const double* p = &x[3 * i];
    const auto& bk = bucket[(bidx(p[0], 0) * B + bidx(p[1], 1)) * B + bidx(p[2], 2)];
    double best = -1e300;
    int32_t bt = -1;
    for (int32_t t : bk) {
      const double* a = &h.A[9 * (size_t)t];
      const double* o = &h.X0[3 * (size_t)t];
      const double d0 = p[0] - o[0], d1 = p[1] - o[1], d2 = p[2] - o[2];
      const double l1 = a[0] * d0 + a[1] * d1 + a[2] * d2;
      const double l2 = a[3] * d0 + a[4] * d1 + a[5] * d2;
      const double l3 = a[6] * d0 + a[7] * d1 + a[8] * d2;
      const double mn = std::min(std::min(1.0 - l1 - l2 - l3, l1), std::min(l2, l3));
      if (mn > best) {
        best = mn;
        bt = t;
      }
    }
    if (best >= -1e-9) cell[i] = bt;
  }
}

constexpr int kMaxPatches = SCDAT_MAX_PATCHES;

struct SpeciesState {
  double q = 0, m = 0, w = 0;
  int64_t cap = 0, n = 0;  // n = slots in use
  int64_t cs = 0;
  int cur = 0;
  DevBuf buf[2][8];        // x,y,z,vx,vy,vz,cell,id
  DevBuf ke_partial, alive_partial;
  int32_t next_id = 0;
  int64_t stat_blocks = 0;  // partials written by the last push
  double v2_mean = 0.0;     // mean |v|^2 of the uploaded particles (sort cadence)
  int sort_every = 0;       // steps between sorts of this species (0: cfg.sort_interval)
  DevBuf mcc_cand;          // collision candidates the push flags (capacity entries)
};

struct Engine {
  virtual ~Engine() = default;
};

// Address of this process's statistics scratch on device `dev`, looked up
// once per device with that device current (each GPU holds its own copy of
// the array at its own address).  Null when the lookup fails.
unsigned long long* stats_scratch(int dev) {
  static std::mutex mu;
  static std::array<unsigned long long*, 64> ptr{};
  if (dev < 0 || dev >= (int)ptr.size()) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    ptr[dev] = static_cast<unsigned long long*>(p);
  }
  return ptr[dev];
}

// Inputs of the statistics reduction (scdat_get_stats_noalloc).
struct StatsArgs {
  const double* ke[kMaxSpecies];        // per-block kinetic-energy partials of the last push
  const int32_t* alive[kMaxSpecies];    // per-block alive counts
  int64_t nb[kMaxSpecies];
  double q[kMaxSpecies], w[kMaxSpecies];
  int ns;
  const unsigned long long* rho;        // fixed-point charge of the last push
  int64_t nn;
  const int32_t* seg_b0;                // scenario row blocks (null: one scenario)
  int n_seg;
  const double* seg_wscale;
  const unsigned long long* led_cnt;    // [scenario][group][species]
  const unsigned long long* led_emi;
  uint32_t surf_mask, cond_mask;        // groups with a condition; conductor groups
  double q_init;
  const unsigned long long* lost;
  const unsigned long long* pushes;
  const PcgScalars* sc;
  unsigned long long* out;              // the scratch
};

// Block 0: the scalar statistics, in the order and with the roundings of
// the host sums of scdat_get_stats (one thread, no fused multiply-adds);
// block 1 + s: the exact integer charge of scenario s.
__global__ void __launch_bounds__(kBlock) k_stats_noalloc(StatsArgs a) {
  if (blockIdx.x == 0) {
    if (threadIdx.x != 0) return;
    double ke = 0.0;
    long long alive = 0;
    for (int s = 0; s < a.ns; ++s)
      for (int64_t b = 0; b < a.nb[s]; ++b) {
        ke = __dadd_rn(ke, a.ke[s][b]);
        alive += a.alive[s][b];
      }
    double surf = 0.0, qc = a.q_init;
    for (int sg = 0; sg < a.n_seg; ++sg) {
      const double scale = a.seg_wscale ? a.seg_wscale[sg] : 1.0;
      for (int g = 0; g < kMaxGroups; ++g) {
        if (!((a.surf_mask >> g) & 1u)) continue;
        for (int sp = 0; sp < a.ns; ++sp) {
          const size_t o = ((size_t)sg * kMaxGroups + g) * kMaxSpecies + sp;
          const double qw = __dmul_rn(a.q[sp], __dmul_rn(a.w[sp], scale));
          const double qq = __dmul_rn((double)a.led_cnt[o], qw);
          surf = __dadd_rn(surf, qq);
          if (sg == 0 && ((a.cond_mask >> g) & 1u))
            qc = __dadd_rn(qc, __dsub_rn(qq, __dmul_rn((double)a.led_emi[o], qw)));
        }
      }
    }
    a.out[0] = (unsigned long long)__double_as_longlong(ke);
    a.out[1] = (unsigned long long)alive;
    a.out[2] = (unsigned long long)__double_as_longlong(surf);
    a.out[3] = (unsigned long long)__double_as_longlong(qc);
    a.out[4] = *a.lost;
    a.out[5] = *a.pushes;
    a.out[6] = (unsigned long long)__double_as_longlong(a.sc->vfloat);
    a.out[7] = 0ull;
    return;
  }
  __shared__ long long lds[kBlock / 64];
  const int sg = (int)blockIdx.x - 1;
  const int64_t i0 = a.seg_b0 ? (int64_t)a.seg_b0[sg] * kPcgBlock : 0;
  const int64_t e1 = a.seg_b0 ? (int64_t)a.seg_b0[sg + 1] * kPcgBlock : a.nn;
  const int64_t i1 = e1 < a.nn ? e1 : a.nn;
  long long t = 0;
  for (int64_t i = i0 + threadIdx.x; i < i1; i += kBlock) t += (long long)a.rho[i];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) t += __shfl_xor(t, off, 64);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = t;
  __syncthreads();
  if (threadIdx.x == 0) {
    long long tot = 0;
    for (int w = 0; w < kBlock / 64; ++w) tot += lds[w];
    a.out[kStatsScratchHead + sg] = (unsigned long long)tot;
  }
}

// Host-side figures of the last solve (statistics and the next solve's batch
// sizes), saved and restored around runs that must leave the context as it
// was (the charging forward and adjoint).
struct SolverFigures {
  int last_iters = 0, last_outer = 0, last_conv = 0;
  int inner_hist[16] = {};
  int64_t iters_total = 0, failed_solves = 0;
  double last_rel = 0.0, last_init_rel = 0.0;
  template <typename C>
  static SolverFigures of(const C* c) {
    SolverFigures f;
    f.last_iters = c->last_iters;
    f.last_outer = c->last_outer;
    f.last_conv = c->last_conv;
    for (int k = 0; k < 16; ++k) f.inner_hist[k] = c->inner_hist[k];
    f.iters_total = c->iters_total;
    f.failed_solves = c->failed_solves;
    f.last_rel = c->last_rel;
    f.last_init_rel = c->last_init_rel;
    return f;
  }
  template <typename C>
  void restore(C* c) const {
    c->last_iters = last_iters;
    c->last_outer = last_outer;
    c->last_conv = last_conv;
    for (int k = 0; k < 16; ++k) c->inner_hist[k] = inner_hist[k];
    c->iters_total = iters_total;
    c->failed_solves = failed_solves;
    c->last_rel = last_rel;
    c->last_init_rel = last_init_rel;
  }
};

// Particle state of one step of the charging adjoint plus that step's nodal E.
struct AdjStep {
  DevBuf v[7];  // x, y, z, vx, vy, vz, cell
  DevBuf E;
};

}  // namespace scdat

using namespace scdat;

struct scdat_ctx {
  scdat_config cfg{};
  std::string err;
  hipStream_t stream = nullptr;
  bool has_mesh = false, initialized = false;
  DevBuf d_xprev;          // previous step's solution (initial-guess extrapolation)
  bool have_xprev = false;
  HostMesh h;
  std::array<int, kMaxGroups> gtype{};
  std::array<int, kMaxGroups> gpartner{};
  std::array<std::array<double, 3>, kMaxGroups> gshift{};
  // device mesh
  DevBuf d_tets, d_nb, d_groups, d_partner, d_gshift;
  DevBuf d_wmap;  // FP32 walk records: tet map + face neighbours
  // ELL operator
  DevBuf d_col, d_v32, d_v64, d_vh32, d_vh64, d_dinv32, d_dinv64, d_g[3];
  DevBuf d_pat, d_poff, d_pval, d_pdinv;  // row patterns of the FP32 inner operator
  DevBuf d_pap;                           // and their A P entries (two-level, tiled order)
  bool pat_ap = false;
  DevBuf d_cnb_off, d_cnb;                // aggregates each row block touches
  // tile-staged iteration (tiled order): per-row box pattern ids, the
  // table (neighbour directions, A D^-1, A P, D^-1) and each block's 27
  // neighbour blocks
  DevBuf d_bpat, d_bd, d_bv, d_bap, d_bdi, d_nbb;
  int nbpat = 0;
  bool use_box = true;
  int npat = 0;                           // 0: none (too many, or FP64 inner solve)
  bool pat_wide = false;                  // 4-byte pattern ids
  bool pat_force_wide = false;            // 4-byte ids whatever the count (tests)
  int pcg_block = 0;                      // PCG row block: 0 by size, 256 or 512
  bool use_patterns = true;
  DevBuf d_kind, d_lift, d_dval, d_dscale, d_phi1, d_kc_idx, d_kc_val, d_b64, d_x64, d_r64, d_partial, d_sc;
  DevBuf d_inner[8];
  // two-level preconditioner
  DevBuf d_dco, d_aci, d_acip, d_coarse[6];
  int precond = SCDAT_PRECOND_JACOBI;
  // fields
  DevBuf d_E, d_phiT, d_phi64, d_rho[2];
  int rho_target = 0, rho_last = 0;
  bool rho_clean[2] = {false, false};
  // diagnostics
  DevBuf d_ledger_count, d_ledger_energy, d_lost, d_pushes;
  // sort scratch
  DevBuf d_count, d_offset, d_cursor, d_tiles, d_total, d_perm, d_stage;
  // rank of every cell in the sorted particle order (SCDAT_SHAPE_CELL_ORDER)
  int cell_order = 0;
  bool cell_key_ready = false;
  DevBuf d_cell_key;
  std::deque<SpeciesState> sp;
  double unit = 0.0, w_ref = 0.0, q_init = 0.0, ccap = 0.0;
  int64_t steps = 0;
  double sim_time = 0.0;
  PcgScalars* h_sc = nullptr;  // pinned mirror
  // background collisions and surface emission
  std::vector<scdat_mcc_process> mcc;
  uint32_t key[2] = {0x5CDA7u, 0x2026u};
  std::array<scdat_surface_emission, kMaxGroups> emit{};
  DevBuf d_counts, d_append, d_emit_count, d_events, d_n_events, d_xyz64, d_tet4;
  // null-collision state per species: majorant controls, max |v|^2 of the
  // last push, candidate count; recomputed from the particles after a load
  DevBuf d_mcc_ctl, d_mcc_v2max, d_mcc_ncand;
  bool mcc_v2_stale = true;
  bool push_mcc = false;  // the next push flags collision candidates
  DevBuf d_face_tet, d_face_k, d_face_cum;
  std::array<int64_t, kMaxGroups + 1> face_off{};
  std::array<double, kMaxGroups> face_area{};
  int64_t events_dropped = 0;
  int64_t event_cap = 0;
  unsigned long long* h_append = nullptr;  // pinned, kMaxSpecies + 1 (last: impacts)
  bool append_pending = false;
  // conductor patches held at the potentials of a surface circuit
  int n_patch = 0;
  std::vector<int32_t> node_patch;
  scdat_circuit* circuit = nullptr;  // caller-owned
  std::array<int32_t, kMaxPatches> patch_cnode{};
  std::array<int32_t, kMaxGroups> group_patch{};
  std::array<double, kMaxPatches> patch_q{};  // collected charge so far
  std::array<double, kMaxPatches> patch_didv{};
  DevBuf d_phik, d_patch_v;
  double* h_patch_v = nullptr;               // pinned, kMaxPatches
  unsigned long long* h_ledger = nullptr;    // pinned, 2 kMaxGroups kMaxSpecies
  unsigned long long* h_stats = nullptr;     // pinned mirror of the statistics scratch
  unsigned long long* d_stats = nullptr;     // this device's statistics scratch (symbol address)
  int device = 0;
  bool ledger_pending = false;
  // batched scenarios (segments); empty seg_node0: one scenario
  std::vector<int64_t> seg_node0;  // n_seg + 1 node offsets
  std::vector<int32_t> tet_seg_h, blk_seg_h, seg_b0_h;
  std::vector<double> seg_wscale, seg_dt, seg_ccap;
  std::vector<uint32_t> seg_key;
  std::vector<int64_t> kc_off_h;
  std::vector<std::array<scdat_surface_emission, kMaxGroups>> seg_emit;
  DevBuf d_tet_seg, d_blk_seg, d_seg_b0, d_seg_unit, d_seg_dt, d_seg_wscale, d_seg_qinit, d_seg_ccap, d_kc_off,
      d_seg_emit, d_seg_key;
  bool seg_emit_dirty = false;
  // per-scenario background-collision target densities [segment][kMaxProc]
  // (NaN: the process's own density) and FP64 time steps; photoemission face
  // ranges [group][segment + 1] of the per-scenario face table and their
  // areas [group][segment]
  std::vector<double> seg_mcc_n;
  bool seg_mcc_dirty = true;
  DevBuf d_seg_mcc_n, d_seg_dt64, d_seg_face_off;
  std::vector<int64_t> seg_face_off_h;
  std::vector<double> seg_face_area_h;
  // field-emission cathodes (ion-driven, FieldEmissionBoundaryCondition)
  std::array<scdat_field_emission, kMaxGroups> fe{};
  uint32_t fe_mask = 0;
  std::array<double, kMaxGroups> fe_pending{}, fe_pending_sec{};
  DevBuf d_fe_x, d_fe_v, d_fe_cell, d_fe_id;
  int node_order = SCDAT_ORDER_GIVEN;
  int order_in_use = SCDAT_ORDER_GIVEN;  // the internal order of the current mesh
  std::vector<int32_t> perm, iperm;  // internal node -> caller's node, and back (empty: same order)
  int push_chunks = 4;  // particles per thread of the push kernel (scdat_autotune)
  // charging adjoint: working steps, checkpoints, adjoint vectors, snapshot
  std::deque<AdjStep> adj_st, adj_ck;
  std::deque<DevBuf> adj_ck_x64;
  DevBuf d_adj_sp, d_adj_rho, d_adj_part, d_adj_xb, d_adj_vb, d_adj_phibar, d_adj_y, d_adj_rhobar, d_adj_Ebar[2],
      d_adj_dotp, d_snap_E, d_snap_x64, d_snap_phi64, d_snap_phiT, d_snap_sc;
  int64_t adj_np = 0;
  SolverFigures snap_solver;
  DevBuf d_adj_events;  // recorded absorption events of a charging forward
  DevBuf d_adj_path;    // and every particle's cell after each of its steps
  int64_t adj_events_np = 0;
  int adj_events_steps = 0;
  int n_seg() const { return seg_node0.empty() ? 1 : (int)seg_node0.size() - 1; }
  bool segmented() const { return !seg_node0.empty(); }
  int last_iters = 0, last_outer = 0, last_conv = 0;
  // Poisson solves that missed the tolerance (the previous potential was
  // kept): over the context's life, and within the current scdat_step call
  int64_t failed_solves = 0;
  int batch_failed = 0;
  int64_t first_failed_step = -1;
  int inner_hist[16] = {};  // inner iterations per outer pass of the last solve
  int64_t iters_total = 0;  // inner iterations of all solves
  double last_rel = 0.0, last_init_rel = 0.0;
  double stage_ms[5] = {0, 0, 0, 0, 0};
  double push_ms_sum = 0.0;
  int64_t push_launches = 0;
  std::vector<hipEvent_t> ev_pool;
  // per-kernel timing (HIP events on the context stream), enabled on request:
  // [0] push, [1] PCG iteration; sums in ms over executed launches
  int kernel_timing = 0;  // 0 off, else time every kernel_timing-th PCG iteration launch
  int64_t kt_seq = 0;     // PCG iteration launches so far (the sampling count)
  std::vector<hipEvent_t> kt_pool;
  double kt_sum[2] = {0, 0};
  int64_t kt_n[2] = {0, 0};
  bool is_fp64() const { return cfg.precision == 1; }
  // Captured steps (SCDAT_SHAPE_STEP_GRAPH): while a scdat_step call runs,
  // the slot counts and the step number live on the device and the launches
  // around the Poisson iteration are replayed from HIP graphs, one per
  // launch configuration (density buffer parity, particle buffer parities,
  // sorted species) of that call.
  // lattice location tables of the FP32 push (lat_dev.tet_hex null: the
  // mesh is not a Kuhn lattice) and SCDAT_SHAPE_LATTICE_PUSH
  DevBuf d_lat_tet_hex, d_lat_hex_tet;
  LatticeDev lat_dev{};
  int lat_mode = 0;
  // deposit buckets (SCDAT_SHAPE_DEPOSIT_BUCKETS): entry counts and entries
  // per 512-row range of the charge
  DevBuf d_dep_count, d_dep_ent;
  int dep_cap = 0;
  int dep_mode = 1;
  int step_graph = 0;
  bool capturing = false;  // launches are being recorded: device counts, no host reads
  DevBuf d_nslots, d_step;
  std::map<uint64_t, hipGraphExec_t> gexec;
  int64_t graph_replays = 0, graph_builds = 0;
  void drop_graphs() {
    for (auto& kv : gexec) (void)hipGraphExecDestroy(kv.second);
    gexec.clear();
  }
};

namespace {

int fail(scdat_ctx* c, int code, const std::string& m) {
  if (c) c->err = m;
  return code;
}
// Without a temporary string: no allocation while the message fits the
// context's reserved buffer.
int fail(scdat_ctx* c, int code, const char* m) {
  if (c) c->err.assign(m);
  return code;
}

template <typename F>
int guarded(scdat_ctx* c, F&& fn) {
  if (!c) return SCDAT_ERR_ARG;
  try {
    fn();
    return SCDAT_OK;
  } catch (const Error& e) {
    c->err = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    c->err = "host allocation failed";
    return SCDAT_ERR_NOMEM;
  } catch (const std::exception& e) {
    c->err = e.what();
    return SCDAT_ERR_STATE;
  }
}

size_t tsize(const scdat_ctx* c) { return c->is_fp64() ? 8 : 4; }

// Row patterns of the FP32 inner operator A D^-1: rows whose column offsets
// j - i, FP32 values and FP32 D^-1 agree bit for bit share one table entry.
// A slot whose value is zero (a column eliminated by a constraint, or ELL
// padding) is stored as offset 0 with value 0: it adds 0 either way, and its
// column no longer splits rows that differ only in where their constrained
// neighbours are numbered.  Structured meshes collapse to a few hundred
// patterns (a 64^3 box with a conductor: about a thousand for 273k rows; a
// sweep of 256 boxes: about a thousand per box).  Used while the ids and
// the table take at most half the ELL's bytes (8 W per row); 2-byte ids up
// to 65536 patterns, 4-byte beyond.  In the tiled order with the two-level
// coarse space the key also holds the row's FP32 entries of A P (A_ij on
// free rows and free columns), which the fused two-level iteration reads.
void build_row_patterns(scdat_ctx* c, const std::vector<float>& vh, const std::vector<float>& di) {
  const HostMesh& h = c->h;
  const int64_t n = h.nn;
  const int W = h.W;
  c->npat = 0;
  c->pat_ap = false;
  if (W < 7 || W > 20 || n == 0) return;  // the pattern kernels exist for these widths
  const bool with_ap = h.nc > 0 && c->order_in_use == SCDAT_ORDER_TILED;
  const int per = with_ap ? 12 : 8;  // key bytes per slot
  // table entries allowed: (npat (per W + 4) + 4 n) <= 4 W n
  const int64_t max_pat = std::max<int64_t>(0, (4 * (int64_t)W * n - 4 * n) / (per * W + 4));
  std::unordered_map<std::string, int> id;
  id.reserve(1024);
  std::vector<uint32_t> pat((size_t)n);
  std::vector<int32_t> poff;
  std::vector<float> pval, pdinv, pap;
  std::string key((size_t)(per * W + 4), '\0');
  for (int64_t i = 0; i < n; ++i) {
    char* k = key.data();
    const bool free_i = h.kind[i] == SCDAT_NODE_FREE;
    for (int sl = 0; sl < W; ++sl) {
      const size_t o = (size_t)sl * n + i;
      const int32_t j = h.col[o];
      const float ap = with_ap && free_i && h.kind[j] == SCDAT_NODE_FREE ? (float)h.aval[o] : 0.0f;
      const int32_t off = (vh[o] == 0.0f && ap == 0.0f) ? 0 : j - (int32_t)i;
      std::memcpy(k + per * sl, &off, 4);
      std::memcpy(k + per * sl + 4, &vh[o], 4);
      if (with_ap) std::memcpy(k + per * sl + 8, &ap, 4);
    }
    std::memcpy(k + per * W, &di[i], 4);
    auto it = id.find(key);
    if (it == id.end()) {
      if ((int64_t)id.size() >= max_pat) return;  // unstructured: stay on the ELL
      it = id.emplace(key, (int)id.size()).first;
      for (int sl = 0; sl < W; ++sl) {
        int32_t off;
        float v, ap = 0.0f;
        std::memcpy(&off, k + per * sl, 4);
        std::memcpy(&v, k + per * sl + 4, 4);
        if (with_ap) std::memcpy(&ap, k + per * sl + 8, 4);
        poff.push_back(off);
        pval.push_back(v);
        pap.push_back(ap);
      }
      pdinv.push_back(di[i]);
    }
    pat[(size_t)i] = (uint32_t)it->second;
  }
  c->pat_wide = id.size() > 65536 || c->pat_force_wide;
  if (c->pat_wide) {
    c->d_pat.upload(pat, c->stream);
  } else {
    std::vector<uint16_t> p16(pat.begin(), pat.end());
    c->d_pat.upload(p16, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));  // the staging vectors are locals
  }
  c->d_poff.upload(poff, c->stream);
  c->d_pval.upload(pval, c->stream);
  c->d_pdinv.upload(pdinv, c->stream);
  if (with_ap) c->d_pap.upload(pap, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));  // the staging vectors are locals
  c->npat = (int)id.size();
  c->pat_ap = with_ap;
}

// Box patterns of the tile-staged iteration (tiled order, every tile in its
// block slot).  Row i of block b sits at point (x, y, z) of b's tile (slot
// x + 8 y + 64 z); every column it reads with a non-zero entry must be a
// lattice neighbour p + delta, delta in {-1, 0, 1}^3, so the iteration can
// take r_{i+1} of all of them from the block's 10^3 box of staged values.
// Its pattern holds, per ELL slot, the direction code (delta + 1 in two bits
// per axis, x lowest; 21 = the row itself, also for zero entries), the FP32
// A D^-1 entry, the
// FP32 A P entry (two-level coarse space; 0 otherwise) and D^-1: the same
// values in the same slot order as the row patterns, so both iterations give
// the same bits.  Rows of blocks past the tiles (constrained nodes of the
// index-0 layers) must have only their diagonal.  Leaves nbpat = 0 when a
// condition fails or the table would need more than 65536 entries.
void build_box_patterns(scdat_ctx* c, const std::vector<float>& vh, const std::vector<float>& di) {
  const HostMesh& h = c->h;
  c->nbpat = 0;
  const int64_t n = h.nn;
  const int W = h.W;
  if (c->order_in_use != SCDAT_ORDER_TILED || !h.tiles_aligned || h.tile_blocks == 0 || W > 16 || W < 7) return;
  const int64_t G = (n + kPcgBlock - 1) / kPcgBlock;
  const int64_t nx = h.tdim[0], ny = h.tdim[1], nz = h.tdim[2];
  std::vector<int32_t> block_of_tile((size_t)(nx * ny * nz), -1);
  for (int64_t b = 0; b < h.tile_blocks; ++b) block_of_tile[h.tile_of_block[b]] = (int32_t)b;
  std::vector<int32_t> nbb((size_t)G * 27, -1);
  for (int64_t b = 0; b < G; ++b) {
    nbb[b * 27 + 13] = (int32_t)b;
    if (b >= h.tile_blocks) continue;
    const int64_t t = h.tile_of_block[b];
    const int64_t tx = t % nx, ty = (t / nx) % ny, tz = t / (nx * ny);
    for (int dz = -1; dz <= 1; ++dz)
      for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
          const int64_t ux = tx + dx, uy = ty + dy, uz = tz + dz;
          if (ux < 0 || uy < 0 || uz < 0 || ux >= nx || uy >= ny || uz >= nz) continue;
          nbb[b * 27 + (dx + 1) + 3 * (dy + 1) + 9 * (dz + 1)] = block_of_tile[ux + nx * (uy + ny * uz)];
        }
  }
  auto tile_xyz = [&](int64_t b, int64_t* q) {
    const int64_t t = h.tile_of_block[b];
    q[0] = t % nx;
    q[1] = (t / nx) % ny;
    q[2] = t / (nx * ny);
  };
  const bool with_ap = h.nc > 0;
  std::unordered_map<std::string, int> id;
  id.reserve(1024);
  std::vector<uint16_t> pat((size_t)n);
  std::vector<uint8_t> pd;
  std::vector<float> pv, pap, pdi;
  const int per = 9;  // key bytes per slot: direction, A D^-1, A P
  std::string key((size_t)(per * W + 4), '\0');
  for (int64_t i = 0; i < n; ++i) {
    const int64_t b = i / kPcgBlock, sl_i = i % kPcgBlock;
    const bool tile_row = b < h.tile_blocks;
    int64_t ti[3] = {0, 0, 0};
    if (tile_row) tile_xyz(b, ti);
    const int64_t pi[3] = {8 * ti[0] + (sl_i & 7), 8 * ti[1] + ((sl_i >> 3) & 7), 8 * ti[2] + (sl_i >> 6)};
    const bool free_i = h.kind[i] == SCDAT_NODE_FREE;
    char* k = key.data();
    for (int sl = 0; sl < W; ++sl) {
      const size_t o = (size_t)sl * n + i;
      const int32_t j = h.col[o];
      const float ap = with_ap && free_i && h.kind[j] == SCDAT_NODE_FREE ? (float)h.aval[o] : 0.0f;
      uint8_t code = 21;
      if (!(vh[o] == 0.0f && ap == 0.0f) && j != i) {
        const int64_t bj = j / kPcgBlock, sj = j % kPcgBlock;
        if (!tile_row || bj >= h.tile_blocks) return;  // only the diagonal off the tiles
        int64_t tj[3];
        tile_xyz(bj, tj);
        const int64_t pj[3] = {8 * tj[0] + (sj & 7), 8 * tj[1] + ((sj >> 3) & 7), 8 * tj[2] + (sj >> 6)};
        int dcode = 0;
        for (int r = 0; r < 3; ++r) {
          const int64_t d = pj[r] - pi[r];
          if (d < -1 || d > 1) return;  // not a lattice neighbour
          dcode |= (int)(d + 1) << (2 * r);
        }
        code = (uint8_t)dcode;
      }
      const float v = vh[o];
      k[per * sl] = (char)code;
      std::memcpy(k + per * sl + 1, &v, 4);
      std::memcpy(k + per * sl + 5, &ap, 4);
    }
    std::memcpy(k + per * W, &di[i], 4);
    auto it = id.find(key);
    if (it == id.end()) {
      if (id.size() >= 65536) return;
      it = id.emplace(key, (int)id.size()).first;
      for (int sl = 0; sl < 16; ++sl) pd.push_back(sl < W ? (uint8_t)k[per * sl] : (uint8_t)21);
      for (int sl = 0; sl < W; ++sl) {
        float v, ap;
        std::memcpy(&v, k + per * sl + 1, 4);
        std::memcpy(&ap, k + per * sl + 5, 4);
        pv.push_back(v);
        pap.push_back(ap);
      }
      pdi.push_back(di[i]);
    }
    pat[(size_t)i] = (uint16_t)it->second;
  }
  hipStream_t s = c->stream;
  c->d_bpat.upload(pat, s);
  c->d_bd.upload(pd, s);
  c->d_bv.upload(pv, s);
  c->d_bap.upload(pap, s);
  c->d_bdi.upload(pdi, s);
  c->d_nbb.upload(nbb, s);
  HIPCHK(hipStreamSynchronize(s));  // the staging vectors are locals
  c->nbpat = (int)id.size();
}

// Boundary faces per group, and within a group per scenario, in (tet, face)
// order with cumulative areas restarting at each (group, scenario): the
// photoemission source table.  A scenario's range is the table a context of
// that scenario alone builds for the group, so its face choice is the same.
void build_face_table(scdat_ctx* c) {
  const HostMesh& h = c->h;
  const int ns = c->segmented() ? c->n_seg() : 1;
  struct F {
    int32_t g, sg, t, k;
  };
  std::vector<F> all;
  for (int64_t t = 0; t < h.nt; ++t)
    for (int k = 0; k < 4; ++k) {
      const int g = h.group[4 * t + k];
      if (g < 0 || g >= kMaxGroups || h.nb[4 * t + k] >= 0) continue;
      all.push_back({g, c->segmented() ? c->tet_seg_h[t] : 0, (int32_t)t, k});
    }
  std::stable_sort(all.begin(), all.end(), [](const F& a, const F& b) { return a.g != b.g ? a.g < b.g : a.sg < b.sg; });
  std::vector<int32_t> ft, fk;
  std::vector<double> fc;
  c->seg_face_off_h.assign((size_t)kMaxGroups * (ns + 1), 0);
  c->seg_face_area_h.assign((size_t)kMaxGroups * ns, 0.0);
  size_t q = 0;
  for (int g = 0; g < kMaxGroups; ++g) {
    c->face_off[g] = (int64_t)ft.size();
    c->face_area[g] = 0.0;
    for (int sg = 0; sg < ns; ++sg) {
      c->seg_face_off_h[(size_t)g * (ns + 1) + sg] = (int64_t)ft.size();
      double acc = 0.0;
      for (; q < all.size() && all[q].g == g && all[q].sg == sg; ++q) {
        const int64_t t = all[q].t;
        const int k = all[q].k;
        const double* p[3];
        int m = 0;
        for (int r = 0; r < 4; ++r)
          if (r != k) p[m++] = &h.xyz[3 * (size_t)h.tets[4 * t + r]];
        const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double w[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
        acc += 0.5 * std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        ft.push_back((int32_t)t);
        fk.push_back(k);
        fc.push_back(acc);
      }
      c->seg_face_area_h[(size_t)g * ns + sg] = acc;
      c->face_area[g] += acc;
    }
    c->seg_face_off_h[(size_t)g * (ns + 1) + ns] = (int64_t)ft.size();
  }
  c->face_off[kMaxGroups] = (int64_t)ft.size();
  if (ft.empty()) {
    ft.push_back(0);
    fk.push_back(0);
    fc.push_back(0.0);
  }
  hipStream_t s = c->stream;
  c->d_face_tet.upload(ft, s);
  c->d_face_k.upload(fk, s);
  c->d_face_cum.upload(fc, s);
  c->d_seg_face_off.upload(c->seg_face_off_h, s);
}

static bool lattice_indices(int64_t nn, const double* xyz, std::vector<int32_t>& ijk, int32_t dims[3],
                            std::vector<double>* lines_out);

// Lattice location tables of the FP32 push (LatticeDev).  Built when every
// tet is a Kuhn tet of a hexahedron of a rectilinear grid (nodes v000,
// v000 + e_p0, v000 + e_p0 + e_p1, v111) and no two tets take one
// (hexahedron, permutation) slot, with at most kLatAxis hexahedra per axis
// whose grid lines form at most kLatSeg runs of equal spacing (the device
// maps every line of a run to its index within 1e-5 of a cell); otherwise
// the push walks every sub-step.  A hexahedron is whole when its six tets
// exist and none of their faces is a boundary face; its clearance is the
// number (at most 7) of nested cubes of whole hexahedra around it: 0 (not
// whole), 1 (whole), 2 (it and its 26 neighbours whole), ...
void build_lattice(scdat_ctx* c) {
  c->lat_dev = LatticeDev{};
  const HostMesh& h = c->h;
  if (h.nt >= ((int64_t)1 << 29) - 1) return;
  std::vector<int32_t> ijk;
  int32_t dims[3];
  std::vector<double> L[3];
  if (!lattice_indices(h.nn, h.xyz.data(), ijk, dims, L)) return;
  int nh[3];
  for (int r = 0; r < 3; ++r) {
    nh[r] = dims[r] - 1;
    if (nh[r] < 1 || nh[r] > kLatAxis) return;
  }
  const int64_t nhex = (int64_t)nh[0] * nh[1] * nh[2];
  if (nhex > ((int64_t)1 << 28)) return;
  // runs of equal spacing along each axis, checked as the device evaluates them
  LatticeDev dv{};
  for (int r = 0; r < 3; ++r) {
    dv.nh[r] = nh[r];
    int ns = 0, i0 = 0;
    while (i0 < nh[r]) {
      const double h0 = L[r][i0 + 1] - L[r][i0];
      int i1 = i0 + 1;
      while (i1 < nh[r] && std::abs((L[r][i1 + 1] - L[r][i1]) - h0) <= 1e-9 * h0) ++i1;
      if (ns == kLatSeg) return;
      dv.sx[r][ns] = (float)L[r][i0];
      dv.sinv[r][ns] = (float)(1.0 / h0);
      dv.si[r][ns] = i0;
      for (int j = i0; j <= i1; ++j) {
        const float t = ((float)L[r][j] - dv.sx[r][ns]) * dv.sinv[r][ns];
        if (std::abs((double)t - (double)(j - i0)) > 1e-5) return;
      }
      ++ns;
      i0 = i1;
    }
    dv.nseg[r] = ns;
    dv.si[r][ns] = nh[r];
    dv.lo[r] = (float)L[r].front();
    dv.hi[r] = (float)L[r].back();
  }
  // permutation index of (p0, p1), the order of lat_locate
  static const int kIdx[3][3] = {{-1, 0, 1}, {2, -1, 3}, {4, 5, -1}};
  std::vector<int32_t> hex_tet((size_t)(6 * nhex), -1);
  std::vector<int64_t> hex_of((size_t)h.nt);
  for (int64_t t = 0; t < h.nt; ++t) {
    int mn[3] = {INT32_MAX, INT32_MAX, INT32_MAX};
    for (int q = 0; q < 4; ++q)
      for (int r = 0; r < 3; ++r) mn[r] = std::min(mn[r], ijk[3 * (size_t)h.tets[4 * t + q] + r]);
    int n0 = 0, n3 = 0, p0 = -1, p2 = -1;
    for (int q = 0; q < 4; ++q) {
      int off[3], sum = 0;
      for (int r = 0; r < 3; ++r) {
        off[r] = ijk[3 * (size_t)h.tets[4 * t + q] + r] - mn[r];
        if (off[r] < 0 || off[r] > 1) return;
        sum += off[r];
      }
      if (sum == 0) ++n0;
      else if (sum == 3) ++n3;
      else if (sum == 1) p0 = off[0] ? 0 : (off[1] ? 1 : 2);
      else p2 = !off[0] ? 0 : (!off[1] ? 1 : 2);
    }
    if (n0 != 1 || n3 != 1 || p0 < 0 || p2 < 0 || p0 == p2) return;
    const int p1 = 3 - p0 - p2;
    const int64_t hx = ((int64_t)mn[0] * nh[1] + mn[1]) * nh[2] + mn[2];
    int32_t& slot = hex_tet[(size_t)(6 * hx + kIdx[p0][p1])];
    if (slot >= 0) return;  // two tets on one slot (overlapping meshes)
    slot = (int32_t)t;
    hex_of[t] = hx;
  }
  // whole hexahedra, then erosions by the 3x3x3 cube
  std::vector<uint8_t> e((size_t)nhex, 0), d((size_t)nhex, 0);
  for (int64_t hx = 0; hx < nhex; ++hx) {
    bool whole = true;
    for (int k = 0; k < 6 && whole; ++k) {
      const int32_t t = hex_tet[(size_t)(6 * hx + k)];
      whole = t >= 0;
      for (int f = 0; f < 4 && whole; ++f) whole = h.nb[4 * (size_t)t + f] >= 0;
    }
    e[hx] = whole ? 1 : 0;
  }
  const int64_t stride[3] = {(int64_t)nh[1] * nh[2], nh[2], 1};
  std::vector<uint8_t> b((size_t)nhex);
  for (int level = 0; level < 7; ++level) {
    for (int64_t hx = 0; hx < nhex; ++hx) d[hx] = (uint8_t)(d[hx] + e[hx]);
    for (int r = 0; r < 3; ++r) {
      for (int64_t hx = 0; hx < nhex; ++hx) {
        const int64_t ir = (hx / stride[r]) % nh[r];
        b[hx] = e[hx] && ir > 0 && ir + 1 < nh[r] && e[hx - stride[r]] && e[hx + stride[r]];
      }
      e.swap(b);
    }
  }
  std::vector<uint32_t> packed((size_t)(6 * nhex));
  for (int64_t q = 0; q < 6 * nhex; ++q)
    packed[q] = hex_tet[q] < 0 ? 0u : ((uint32_t)(hex_tet[q] + 1) | (uint32_t)d[q / 6] << 29);
  std::vector<uint32_t> tet_hex((size_t)h.nt);
  for (int64_t t = 0; t < h.nt; ++t) {
    const int64_t hx = hex_of[t];
    const uint32_t i = (uint32_t)(hx / stride[0]), j = (uint32_t)((hx / stride[1]) % nh[1]),
                   k = (uint32_t)(hx % nh[2]);
    tet_hex[t] = i | j << 10 | k << 20;
  }
  hipStream_t s = c->stream;
  c->d_lat_tet_hex.upload(tet_hex, s);
  c->d_lat_hex_tet.upload(packed, s);
  HIPCHK(hipStreamSynchronize(s));  // the staging vectors are local
  dv.tet_hex = c->d_lat_tet_hex.as<uint32_t>();
  dv.hex_tet = c->d_lat_hex_tet.as<uint32_t>();
  c->lat_dev = dv;
}

template <typename T>
void upload_mesh_device(scdat_ctx* c) {
  HostMesh& h = c->h;
  hipStream_t s = c->stream;
  std::vector<TetMap<T>> tm((size_t)h.nt);
  for (int64_t t = 0; t < h.nt; ++t) {
    TetMap<T>& m = tm[t];
    std::memset(&m, 0, sizeof(m));
    for (int k = 0; k < 9; ++k) m.a[k] = (T)h.A[9 * t + k];
    for (int k = 0; k < 3; ++k) m.x0[k] = (T)h.X0[3 * t + k];
    for (int k = 0; k < 4; ++k) m.node[k] = h.tets[4 * t + k];
  }
  c->d_tets.upload(tm, s);
  c->d_xyz64.upload(h.xyz, s);
  c->d_tet4.upload(h.tets, s);
  build_face_table(c);
  c->d_nb.upload(h.nb, s);
  if constexpr (sizeof(T) == 4) {
    std::vector<WalkMap<float>> wm((size_t)h.nt);
    for (int64_t t = 0; t < h.nt; ++t) {
      WalkMap<float>& m = wm[t];
      for (int k = 0; k < 9; ++k) m.a[k] = (float)h.A[9 * t + k];
      for (int k = 0; k < 3; ++k) m.x0[k] = (float)h.X0[3 * t + k];
      for (int k = 0; k < 4; ++k) m.nb[k] = h.nb[4 * t + k];
    }
    c->d_wmap.upload(wm, s);
    HIPCHK(hipStreamSynchronize(s));  // the staging vector is local
    build_lattice(c);
  } else {
    c->lat_dev = LatticeDev{};
  }
  c->d_groups.upload(h.group, s);
  c->d_partner.upload(h.partner, s);
  c->d_col.upload(h.col, s);
  std::vector<float> f32(h.aval.size());
  for (size_t i = 0; i < f32.size(); ++i) f32[i] = (float)h.aval[i];
  c->d_v32.upload(f32, s);
  c->d_v64.upload(h.aval, s);
  // column-scaled operator A D^-1 of the inner solve (rounded once from FP64)
  std::vector<double> vh(h.aval.size());
  for (size_t o = 0; o < vh.size(); ++o) vh[o] = h.aval[o] * h.dinv[h.col[o]];
  std::vector<float> di32(h.dinv.size());
  for (size_t i = 0; i < di32.size(); ++i) di32[i] = (float)h.dinv[i];
  c->npat = 0;
  if (sizeof(T) == 4) {
    for (size_t o = 0; o < vh.size(); ++o) f32[o] = (float)vh[o];
    c->d_vh32.upload(f32, s);
    build_row_patterns(c, f32, di32);
    build_box_patterns(c, f32, di32);
  } else {
    c->d_vh64.upload(vh, s);
  }
  c->d_dinv32.upload(di32, s);
  c->d_dinv64.upload(h.dinv, s);
  const std::vector<double>* g[3] = {&h.gx, &h.gy, &h.gz};
  for (int r = 0; r < 3; ++r) {
    if (sizeof(T) == 4) {
      std::vector<float> v(g[r]->size());
      for (size_t i = 0; i < v.size(); ++i) v[i] = (float)(*g[r])[i];
      c->d_g[r].upload(v, s);
    } else {
      c->d_g[r].upload(*g[r], s);
    }
  }
  if (h.nc > 0) {
    if (sizeof(T) == 4) {
      std::vector<float> a(h.dco.begin(), h.dco.end()), b(h.aci.begin(), h.aci.end());
      c->d_dco.upload(a, s);
      c->d_aci.upload(b, s);
      // the fused iteration's copy: rows padded to a multiple of 64 columns
      const size_t G = (size_t)h.nc, gp = (G + 63) / 64 * 64;
      std::vector<float> bp(G * gp, 0.0f);
      for (size_t r = 0; r < G; ++r)
        for (size_t q = 0; q < G; ++q) bp[r * gp + q] = (float)h.aci[r * G + q];
      c->d_acip.upload(bp, s);
      HIPCHK(hipStreamSynchronize(s));  // the staging vectors are locals
    } else {
      c->d_dco.upload(h.dco, s);
      c->d_aci.upload(h.aci, s);
    }
    for (auto& b : c->d_coarse) {
      b.alloc(sizeof(T) * (size_t)h.nc);
      HIPCHK(hipMemsetAsync(b.p, 0, b.bytes, s));
    }
    c->d_cnb_off.upload(h.cnb_off, s);
    c->d_cnb.upload(h.cnb, s);
  }
  c->d_kind.upload(h.kind, s);
  c->d_lift.upload(h.lift, s);
  c->d_dval.upload(h.dval, s);
  c->d_dscale.upload(h.dscale, s);
  c->d_kc_idx.upload(h.kc_idx, s);
  c->d_kc_val.upload(h.kc_val, s);
  const size_t nn = (size_t)h.nn;
  c->d_b64.alloc(8 * nn);
  c->d_x64.alloc(8 * nn);
  c->d_xprev.alloc(8 * nn);
  c->have_xprev = false;
  c->d_r64.alloc(8 * nn);
  c->d_phi1.alloc(8 * nn);
  HIPCHK(hipMemsetAsync(c->d_x64.p, 0, 8 * nn, s));
  HIPCHK(hipMemsetAsync(c->d_phi1.p, 0, 8 * nn, s));
  c->d_partial.alloc(8 * partial_doubles(h.nn));
  c->d_sc.alloc(sizeof(PcgScalars));
  HIPCHK(hipMemsetAsync(c->d_sc.p, 0, sizeof(PcgScalars), s));
  for (auto& b : c->d_inner) b.alloc(sizeof(T) * nn);
  c->d_E.alloc(sizeof(Vec4<T>) * nn);
  HIPCHK(hipMemsetAsync(c->d_E.p, 0, sizeof(Vec4<T>) * nn, s));
  c->d_phiT.alloc(sizeof(T) * nn);
  c->d_phi64.alloc(8 * nn);
  std::vector<double> phi0(nn, 0.0);
  for (size_t i = 0; i < nn; ++i)
    if (h.kind[i] == SCDAT_NODE_DIRICHLET) phi0[i] = h.dval[i];
  HIPCHK(hipMemcpyAsync(c->d_phi64.p, phi0.data(), 8 * nn, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_x64.p, phi0.data(), 8 * nn, hipMemcpyHostToDevice, s));
  for (int k = 0; k < 2; ++k) {
    c->d_rho[k].alloc(8 * nn);
    HIPCHK(hipMemsetAsync(c->d_rho[k].p, 0, 8 * nn, s));
    c->rho_clean[k] = true;
  }
  c->d_ledger_count.alloc(8 * kMaxGroups * kMaxSpecies);
  c->d_ledger_energy.alloc(8 * kMaxGroups * kMaxSpecies);
  c->d_counts.alloc(8 * C_COUNT);
  c->d_append.alloc(8 * (kMaxSpecies + 1));
  c->d_emit_count.alloc(8 * kMaxGroups * kMaxSpecies);
  c->d_n_events.alloc(8);
  HIPCHK(hipMemsetAsync(c->d_counts.p, 0, 8 * C_COUNT, s));
  HIPCHK(hipMemsetAsync(c->d_append.p, 0, 8 * (kMaxSpecies + 1), s));
  HIPCHK(hipMemsetAsync(c->d_emit_count.p, 0, 8 * kMaxGroups * kMaxSpecies, s));
  HIPCHK(hipMemsetAsync(c->d_n_events.p, 0, 8, s));
  c->d_lost.alloc(8);
  c->d_pushes.alloc(8);
  HIPCHK(hipMemsetAsync(c->d_pushes.p, 0, 8, s));
  HIPCHK(hipMemsetAsync(c->d_ledger_count.p, 0, 8 * kMaxGroups * kMaxSpecies, s));
  HIPCHK(hipMemsetAsync(c->d_ledger_energy.p, 0, 8 * kMaxGroups * kMaxSpecies, s));
  HIPCHK(hipMemsetAsync(c->d_lost.p, 0, 8, s));
  c->d_count.alloc(4 * (size_t)h.nt);
  c->d_offset.alloc(4 * (size_t)h.nt);
  c->d_cursor.alloc(4 * (size_t)h.nt);
  c->d_tiles.alloc(4 * sort_tiles(h.nt) + 16);
  c->d_total.alloc(16);
  HIPCHK(hipStreamSynchronize(s));
}

EllDev ell(scdat_ctx* c) {
  EllDev A{};
  A.n = c->h.nn;
  A.w = c->h.W;
  A.col = c->d_col.as<int32_t>();
  A.v32 = c->d_v32.as<float>();
  A.v64 = c->d_v64.as<double>();
  A.vh32 = c->d_vh32.as<float>();
  A.vh64 = c->d_vh64.as<double>();
  A.dinv32 = c->d_dinv32.as<float>();
  A.dinv64 = c->d_dinv64.as<double>();
  A.pb = kPcgBlock;
  if (c->use_patterns && c->npat > 0) {
    A.npat = c->npat;
    if (c->pat_wide) A.pat32 = c->d_pat.as<uint32_t>();
    else A.pat = c->d_pat.as<uint16_t>();
    A.poff = c->d_poff.as<int32_t>();
    A.pval32 = c->d_pval.as<float>();
    A.pdinv32 = c->d_pdinv.as<float>();
  }
  if (c->is_fp64()) {
    A.gx64 = c->d_g[0].as<double>();
    A.gy64 = c->d_g[1].as<double>();
    A.gz64 = c->d_g[2].as<double>();
  } else {
    A.gx32 = c->d_g[0].as<float>();
    A.gy32 = c->d_g[1].as<float>();
    A.gz32 = c->d_g[2].as<float>();
  }
  if (c->use_box && c->nbpat > 0 && !c->segmented() && !c->is_fp64() && A.w >= 13 && A.w <= 16) {
    A.box = 1;
    A.bpat = c->d_bpat.as<uint16_t>();
    A.bd = c->d_bd.as<uint8_t>();
    A.bv = c->d_bv.as<float>();
    A.bap = c->d_bap.as<float>();
    A.bdi = c->d_bdi.as<float>();
    A.nbb = c->d_nbb.as<int32_t>();
  }
  // two-level preconditioner: one scenario only (the coarse space spans the mesh)
  if (c->precond == SCDAT_PRECOND_TWO_LEVEL && c->h.nc > 0 && !c->segmented()) {
    A.nc = c->h.nc;
    if (c->is_fp64()) {
      A.dco64 = c->d_dco.as<double>();
      A.aci64 = c->d_aci.as<double>();
    } else {
      A.dco32 = c->d_dco.as<float>();
      A.aci32 = c->d_aci.as<float>();
      // the coarse solve runs inside the iteration kernel when the row
      // patterns or the box patterns carry A P (tiled order) and the width
      // has a fused kernel
      const bool rows_ap = A.npat > 0 && c->pat_ap && A.w >= 13 && A.w <= 18;
      if ((rows_ap || A.box) && c->h.cnb_max <= kMaxTouched) {
        A.fused = 1;
        A.cnb_off = c->d_cnb_off.as<int32_t>();
        A.cnb = c->d_cnb.as<int32_t>();
        if (rows_ap) A.pap32 = c->d_pap.as<float>();
        A.ncp = (A.nc + 63) / 64 * 64;
        A.acip32 = c->d_acip.as<float>();
      }
    }
  }
  // 512-row blocks halve the blocks whose partials every block sums (64^3:
  // 12.0 -> 11.0 us per iteration) but leave a small mesh's grid too narrow
  // (32^3, 36k rows: the charging adjoint 1.09 -> 1.63 s)
  // the tile-staged iteration needs the tiles' 512-row blocks
  if (A.box && c->pcg_block == 256) A.box = 0;
  const bool small = A.n < (int64_t)kPcgBlock * 256;
  if (!c->segmented() && A.nc == 0 && !A.box && (c->pcg_block == 256 || (c->pcg_block == 0 && small))) A.pb = 256;
  return A;
}

FieldBufs fbufs(scdat_ctx* c) {
  FieldBufs f{};
  f.kind = c->d_kind.as<int32_t>();
  f.lift = c->d_lift.as<double>();
  f.dval = c->d_dval.as<double>();
  f.dscale = c->d_dscale.as<double>();
  f.phi1 = c->h.has_cond ? c->d_phi1.as<double>() : nullptr;
  f.kc_idx = c->d_kc_idx.as<int32_t>();
  f.kc_val = c->d_kc_val.as<double>();
  f.n_kc = (int64_t)c->h.kc_idx.size();
  f.b64 = c->d_b64.as<double>();
  f.x64 = c->d_x64.as<double>();
  f.r64 = c->d_r64.as<double>();
  f.partial = c->d_partial.as<double>();
  f.sc = c->d_sc.as<PcgScalars>();
  f.n_basis = c->n_patch;
  f.basis = c->d_phik.as<double>();
  f.coef = c->d_patch_v.as<double>();
  f.n_seg = c->n_seg();
  if (c->segmented()) {
    f.blk_seg = c->d_blk_seg.as<int32_t>();
    f.seg_b0 = c->d_seg_b0.as<int32_t>();
    f.seg_unit = c->d_seg_unit.as<double>();
    f.kc_off = c->d_kc_off.as<int64_t>();
  }
  return f;
}

// Scalars slots on the device: one per segment, plus the summary slot when
// there are segments.
size_t n_scalars(const scdat_ctx* c) { return c->segmented() ? (size_t)c->n_seg() + 1 : 1; }

// Timing-only events: no system-scope fence (cache writeback and invalidate)
// when they complete, so the timed kernel runs as its untimed neighbours do.
// They are read only after a stream synchronisation.
hipEvent_t take_kt_event(scdat_ctx* c, size_t idx) {
  while (c->kt_pool.size() <= idx) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    c->kt_pool.push_back(e);
  }
  return c->kt_pool[idx];
}

template <typename T>
InnerBufs<T> ibufs(scdat_ctx* c) {
  InnerBufs<T> b;
  for (int k = 0; k < 2; ++k) {
    b.r[k] = c->d_inner[k].as<T>();
    b.w[k] = c->d_inner[2 + k].as<T>();
    b.s[k] = c->d_inner[4 + k].as<T>();
  }
  b.p = c->d_inner[6].as<T>();
  b.x = c->d_inner[7].as<T>();
  for (int k = 0; k < 2; ++k) {
    b.zr[k] = c->d_coarse[k].as<T>();
    b.zs[k] = c->d_coarse[2 + k].as<T>();
    b.zw[k] = c->d_coarse[4 + k].as<T>();
  }
  return b;
}

// The host reads one PcgScalars: the solve's own, or the segments' summary.
void read_scalars(scdat_ctx* c) {
  const PcgScalars* src = c->d_sc.as<PcgScalars>();
  if (c->segmented()) {
    launch_seg_summary(fbufs(c), c->stream);
    src += c->n_seg();
  }
  HIPCHK(hipMemcpyAsync(c->h_sc, src, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

// PCG driver: b64 and the scalars' bnorm2 are already set (k_rhs or host).
// `hist` (optional, >= max_outer + 1 entries) holds the inner iteration count
// of each outer pass of the previous solve.  A pass it predicts (more than 16
// iterations) is launched as one batch of that count + 2 followed by its
// apply, with no host poll: the launches past convergence exit at once, and
// an inner solve cut short by the batch is an inexact inner solve that the
// FP64 outer loop corrects.  The host reads the flags when the outer loop
// runs past the predicted passes, then polls batch by batch as before; a
// typical solve costs one host poll.  With segments (batched scenarios) a
// predicted pass is launched without reading the residual first, but its
// inner loop is polled.
template <typename T>
void pcg_drive(scdat_ctx* c, const FieldBufs& f, double tol, double inner_tol, int max_outer,
               int* hist) {
  const EllDev A = ell(c);
  const InnerBufs<T> ib = ibufs<T>(c);
  hipStream_t s = c->stream;
  launch_set_controls(f, (int)n_scalars(c), tol, inner_tol, c->cfg.pcg_max_iter, max_outer, s);
  const bool chain = hist && !c->segmented();
  // kernel-timing samples: (pass, iteration, event pair); a sample counts if
  // its launch ran an iteration (below that pass's count)
  struct Sample {
    int pass, it;
  };
  std::vector<Sample> samples;
  size_t tallied = 0;
  auto tally = [&](size_t upto, int pass, int count) {
    for (; tallied < upto; ++tallied) {
      const Sample& q = samples[tallied];
      if (q.pass != pass || q.it >= count) continue;
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->kt_pool[2 * tallied], c->kt_pool[2 * tallied + 1]));
      c->kt_sum[1] += ms;
      c->kt_n[1] += 1;
    }
  };
  bool fresh = false;  // h_sc was read after the last launch
  int passes = 0;
  for (int outer = 0; outer <= max_outer; ++outer) {
    launch_resid64(A, f, s);
    fresh = false;
    const bool spec = hist && hist[outer] > 16;
    if (!spec) {
      read_scalars(c);
      fresh = true;
      if (chain && c->h_sc->stall) {
        // a predicted pass ran out of its batch before its inner solve
        // stopped (the launches after it did nothing): resume it where it
        // stands, as a polled pass would have gone on
        const int pass = c->h_sc->outer_it;
        int it = c->h_sc->inner_it, batch = 4;
        launch_clear_stall(f, s);
        for (;;) {
          for (int k = 0; k < batch; ++k, ++it) launch_inner_iter<T>(A, f, ib, it, s, nullptr);
          read_scalars(c);
          if (c->h_sc->inner_done) break;
          batch = std::min(batch * 2, 64);
        }
        launch_apply<T>(A, f, ib, s);
        fresh = false;
        outer = pass;  // the loop goes on with the next pass's residual
        continue;
      }
      if (outer == 0 && !chain) {
        const double bn0 = std::sqrt(c->h_sc->bnorm2);
        c->last_init_rel = bn0 > 0 ? c->h_sc->rnorm / bn0 : 0.0;
      }
      if (c->h_sc->outer_done) break;
    }
    launch_inner_init<T>(A, f, ib, s);
    fresh = false;
    int it = 0, batch = 8;
    // the previous solve's count + 2 (polled) or + a sixteenth, 3 to 8
    // (launched without a poll: a pass that runs out of its batch stalls the
    // solve until the host resumes it, and the next pass's launches run
    // idle): most passes end in this batch
    if (spec) batch = std::min(1024, hist[outer] + (chain ? std::max(3, std::min(8, hist[outer] / 16)) : 2));
    auto launch_batch = [&](int count) {
      for (int k = 0; k < count; ++k, ++it) {
        // every kernel_timing-th launch, away from batch starts (the first
        // launch after a host poll finds an idle GPU and is not representative)
        const bool sample = c->kernel_timing > 0 && it >= 5 && ++c->kt_seq % c->kernel_timing == 0;
        if (sample) {
          const size_t j = samples.size();
          hipEvent_t ev[2] = {take_kt_event(c, 2 * j), take_kt_event(c, 2 * j + 1)};
          launch_inner_iter<T>(A, f, ib, it, s, ev);
          samples.push_back({outer, it});
        } else {
          launch_inner_iter<T>(A, f, ib, it, s, nullptr);
        }
      }
    };
    if (chain && spec) {
      launch_batch(batch);
      launch_apply<T>(A, f, ib, s);
      continue;
    }
    for (;;) {
      launch_batch(batch);
      read_scalars(c);
      fresh = true;
      if (spec && it == batch) {
        if (outer == 0) {
          const double bn0 = std::sqrt(c->h_sc->bnorm2);
          c->last_init_rel = bn0 > 0 ? c->h_sc->rnorm / bn0 : 0.0;
        }
        if (c->h_sc->outer_done) break;  // converged before this pass: nothing ran
      }
      if (c->h_sc->inner_done) break;
      batch = batch > 64 ? 4 : std::min(batch * 2, 64);  // after a predicted batch: small steps
    }
    if (c->h_sc->outer_done) {
      if (hist) hist[outer] = 0;  // no speculation at this index next time
      break;
    }
    if (!chain) tally(samples.size(), outer, c->h_sc->inner_it);
    ++passes;
    if (hist) hist[outer] = c->h_sc->inner_it;
    launch_apply<T>(A, f, ib, s);
    fresh = false;
  }
  if (!fresh) read_scalars(c);
  const PcgScalars& q = *c->h_sc;
  if (chain) {
    // every pass's count from the device: the samples and the next prediction
    for (size_t j = 0; j < samples.size(); ++j) {
      const Sample& sm = samples[j];
      if (sm.pass >= 16 || sm.pass >= q.outer_it || sm.it >= q.pass_it[sm.pass]) continue;
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, c->kt_pool[2 * j], c->kt_pool[2 * j + 1]));
      c->kt_sum[1] += ms;
      c->kt_n[1] += 1;
    }
    for (int k = 0; k <= max_outer; ++k) hist[k] = k < q.outer_it && k < 16 ? q.pass_it[k] : 0;
    const double bn0 = std::sqrt(q.bnorm2);
    c->last_init_rel = bn0 > 0 ? q.init_rnorm / bn0 : 0.0;
  } else if (hist) {
    for (int k = passes + 1; k <= max_outer; ++k) hist[k] = 0;
  }
  c->last_iters = q.total_it;
  c->iters_total += q.total_it;
  c->last_outer = q.outer_it;
  const double bn = std::sqrt(q.bnorm2);
  c->last_rel = bn > 0 ? q.rnorm / bn : q.rnorm;
  c->last_conv = q.failed ? 0 : 1;
}

// Unit-potential solution: the nodes with in_set[i] at 1 V, every other
// constrained node at 0 V, no charge.  Written to out (FP64, n).
template <typename T>
void solve_unit(scdat_ctx* c, const std::vector<char>& in_set, double* out) {
  const HostMesh& h = c->h;
  const int64_t nn = h.nn;
  FieldBufs f = fbufs(c);
  f.x64 = out;
  // right-hand side: K_ii on the set's rows, minus the set's columns of K on
  // free rows (the lift of a unit Dirichlet value)
  std::vector<double> b((size_t)nn, 0.0), x0((size_t)nn, 0.0);
  for (int64_t i = 0; i < nn; ++i) {
    if (h.kind[i] != SCDAT_NODE_FREE) {
      if (in_set[i]) {
        b[i] = h.dscale[i];
        x0[i] = 1.0;
      }
      continue;
    }
    for (int sl = 1; sl < h.W; ++sl) {
      const size_t o = (size_t)sl * nn + i;
      const int32_t j = h.col[o];
      if (j != i && in_set[j] && h.kind[j] != SCDAT_NODE_FREE) b[i] -= h.kfull[o];
    }
  }
  HIPCHK(hipMemcpyAsync(c->d_b64.p, b.data(), 8 * (size_t)nn, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(out, x0.data(), 8 * (size_t)nn, hipMemcpyHostToDevice, c->stream));
  // |b|^2 per segment
  std::vector<PcgScalars> init(n_scalars(c));
  for (int sg = 0; sg < c->n_seg(); ++sg) {
    const int64_t i0 = c->segmented() ? c->seg_node0[sg] : 0, i1 = c->segmented() ? c->seg_node0[sg + 1] : nn;
    double bb = 0;
    for (int64_t i = i0; i < i1; ++i) bb += b[i] * b[i];
    init[sg].bnorm2 = bb;
  }
  HIPCHK(hipMemcpyAsync(f.sc, init.data(), sizeof(PcgScalars) * init.size(), hipMemcpyHostToDevice, c->stream));
  pcg_drive<T>(c, f, 1e-13, c->is_fp64() ? 0.0 : 1e-5, 12, nullptr);
  HIPCHK(hipStreamSynchronize(c->stream));
  if (!c->last_conv) throw Error(SCDAT_ERR_SOLVER, "unit-potential solve did not converge");
}

template <typename T>
void solve_phi1(scdat_ctx* c) {
  if (!c->h.has_cond) return;
  const int64_t nn = c->h.nn;
  std::vector<char> cond((size_t)nn, 0);
  for (int64_t i = 0; i < nn; ++i) cond[i] = c->h.kind[i] == SCDAT_NODE_CONDUCTOR;
  solve_unit<T>(c, cond, c->d_phi1.as<double>());
  std::vector<double> phi1((size_t)nn);
  HIPCHK(hipMemcpyAsync(phi1.data(), c->d_phi1.p, 8 * phi1.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->segmented()) {
    // capacitance per scenario over its own conductor rows
    c->seg_ccap.assign((size_t)c->n_seg(), 0.0);
    for (int sg = 0; sg < c->n_seg(); ++sg) {
      double cap = 0;
      for (int64_t k = c->kc_off_h[sg]; k < c->kc_off_h[sg + 1]; ++k) cap += c->h.kc_val[k] * phi1[c->h.kc_idx[k]];
      if (c->kc_off_h[sg + 1] > c->kc_off_h[sg] && !(cap > 0))
        throw Error(SCDAT_ERR_SOLVER, "conductor capacitance not positive in scenario " + std::to_string(sg));
      c->seg_ccap[sg] = cap;
    }
    c->d_seg_ccap.upload(c->seg_ccap, c->stream);
    c->ccap = c->seg_ccap[0];
  } else {
    double cap = 0;
    for (size_t k = 0; k < c->h.kc_idx.size(); ++k) cap += c->h.kc_val[k] * phi1[c->h.kc_idx[k]];
    c->ccap = cap;
    if (!(cap > 0)) throw Error(SCDAT_ERR_SOLVER, "conductor capacitance not positive");
  }
  // patch basis for a circuit-driven conductor
  if (c->n_patch > 0) {
    c->d_phik.alloc(8 * (size_t)nn * c->n_patch);
    for (int k = 0; k < c->n_patch; ++k) {
      std::vector<char> in((size_t)nn, 0);
      for (int64_t i = 0; i < nn; ++i) in[i] = c->node_patch[i] == k;
      solve_unit<T>(c, in, c->d_phik.as<double>() + (size_t)k * nn);
    }
  }
  std::vector<PcgScalars> z(n_scalars(c));
  HIPCHK(hipMemcpyAsync(c->d_sc.p, z.data(), sizeof(PcgScalars) * z.size(), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
}

FloatArgs float_args(scdat_ctx* c) {
  FloatArgs fa{};
  fa.q_init = c->q_init;
  fa.ccap = c->ccap;
  fa.ledger_count = c->d_ledger_count.as<unsigned long long>();
  fa.emit_count = c->d_emit_count.as<unsigned long long>();
  for (size_t s = 0; s < c->sp.size() && s < (size_t)kMaxSpecies; ++s) fa.qw[s] = c->sp[s].q * c->sp[s].w;
  fa.cond_mask = 0;
  for (int g = 0; g < kMaxGroups; ++g)
    if (c->gtype[g] == SCDAT_BC_CONDUCTOR) fa.cond_mask |= 1u << g;
  if (c->segmented()) {
    fa.seg_qinit = c->d_seg_qinit.as<double>();
    fa.seg_wscale = c->d_seg_wscale.as<double>();
    fa.seg_ccap = c->d_seg_ccap.as<double>();
  }
  return fa;
}

// Right-hand side and initial guess of a step's solve.
template <typename T>
void solve_setup(scdat_ctx* c, int rho_idx) {
  const EllDev A = ell(c);
  FieldBufs f = fbufs(c);
  launch_rhs(A, f, c->d_rho[rho_idx].as<unsigned long long>(), c->unit, c->cfg.load_sign, c->stream);
  // initial guess: 2 x_n - x_{n-1} on free rows (x_n alone from the second solve)
  launch_extrapolate(f, c->d_xprev.as<double>(), A.n, !c->have_xprev, c->stream);
  c->have_xprev = true;
}

// The PCG solve (host-driven) and the failed-solve policy; float: then the
// floating potential.
template <typename T>
void solve_run(scdat_ctx* c, bool flt = true) {
  const EllDev A = ell(c);
  FieldBufs f = fbufs(c);
  const bool fp64 = c->is_fp64();
  pcg_drive<T>(c, f, c->cfg.pcg_tol, fp64 ? 0.0 : c->cfg.inner_tol, fp64 ? 4 : 12, c->inner_hist);
  if (!c->last_conv) {
    // PICCycle::solvePoissonEquation keeps the previous potential when every
    // solver fails (PICCycle.cpp:731-767) and updateElectricField then
    // rebuilds the same E from it: the failed segments take back the
    // previous solution x_n (the extrapolation saved it), k_float keeps their
    // floating potential, so the field kernel reproduces the previous phi
    // and E bit for bit.
    launch_keep_failed(f, c->d_xprev.as<double>(), A.n, c->stream);
    c->failed_solves++;
    if (c->batch_failed++ == 0) c->first_failed_step = c->steps;
  }
  if (flt && c->h.has_cond && c->n_patch == 0) launch_float(f, float_args(c), c->stream);
}

template <typename T>
void do_solve(scdat_ctx* c, int rho_idx) {
  solve_setup<T>(c, rho_idx);
  solve_run<T>(c);
}

// Snapshot of the cumulative ledger and emission counts (read after the next
// stream synchronisation) for the patch currents.
void snapshot_ledger(scdat_ctx* c) {
  if (c->n_patch == 0) return;
  const size_t m = (size_t)kMaxGroups * kMaxSpecies;
  HIPCHK(hipMemcpyAsync(c->h_ledger, c->d_ledger_count.p, 8 * m, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(c->h_ledger + m, c->d_emit_count.p, 8 * m, hipMemcpyDeviceToHost, c->stream));
  c->ledger_pending = true;
}

// Collected charge per patch since the last step -> plasma current -> one
// backward-Euler circuit step over dt -> patch potentials for the field.
void advance_patches(scdat_ctx* c, bool advance) {
  if (c->n_patch == 0) return;
  if (c->ledger_pending) {
    const size_t m = (size_t)kMaxGroups * kMaxSpecies;
    std::array<double, kMaxPatches> q{};
    for (int g = 0; g < kMaxGroups; ++g) {
      const int k = c->group_patch[g];
      if (k < 0 || k >= c->n_patch) continue;
      for (size_t sp = 0; sp < c->sp.size(); ++sp) {
        const double qw = c->sp[sp].q * c->sp[sp].w;
        q[k] += ((double)c->h_ledger[g * kMaxSpecies + sp] - (double)c->h_ledger[m + g * kMaxSpecies + sp]) * qw;
      }
    }
    if (advance) {
      int32_t nc = 0;
      std::vector<double> v;
      // circuit size: the largest node index used
      for (int k = 0; k < c->n_patch; ++k) nc = std::max(nc, c->patch_cnode[k] + 1);
      std::vector<double> cur(nc, 0.0), didv(nc, 0.0), pot(nc, 0.0);
      for (int k = 0; k < c->n_patch; ++k) {
        cur[c->patch_cnode[k]] += (q[k] - c->patch_q[k]) / c->cfg.dt;
        didv[c->patch_cnode[k]] += c->patch_didv[k];
      }
      const int rc = scdat_circuit_advance(c->circuit, c->cfg.dt, cur.data(), didv.data(), nullptr, nullptr, nullptr,
                                           0, std::numeric_limits<double>::infinity(), nullptr, nullptr, nullptr,
                                           nullptr);
      if (rc != SCDAT_OK) throw Error(rc, std::string("surface circuit: ") + scdat_circuit_last_error(c->circuit));
    }
    c->patch_q = q;
    c->ledger_pending = false;
  }
  int32_t ncn = 0, nbr = 0;
  scdat_circuit_size(c->circuit, &ncn, &nbr);
  std::vector<double> all((size_t)std::max(1, ncn), 0.0);
  if (scdat_circuit_get_potentials(c->circuit, all.data()) != SCDAT_OK)
    throw Error(SCDAT_ERR_STATE, "surface circuit: cannot read potentials");
  for (int k = 0; k < c->n_patch; ++k) c->h_patch_v[k] = all[c->patch_cnode[k]];
  HIPCHK(hipMemcpyAsync(c->d_patch_v.p, c->h_patch_v, 8 * (size_t)c->n_patch, hipMemcpyHostToDevice, c->stream));
}

template <typename T>
void do_field(scdat_ctx* c, unsigned long long* rho_clear) {
  launch_phi_efield<T>(ell(c), fbufs(c), c->d_phiT.as<T>(), c->d_phi64.as<double>(),
                       c->d_E.as<Vec4<T>>(), rho_clear, c->stream);
}

bool species_has_mcc(const scdat_ctx* c, int si) {
  for (const auto& q : c->mcc)
    if (q.species == si) return true;
  return false;
}

// Device form of the processes whose incident species is si; returns their count.
int pack_procs(const scdat_ctx* c, int si, MccProcDev* out) {
  int np = 0;
  for (size_t k = 0; k < c->mcc.size(); ++k) {
    const scdat_mcc_process& q = c->mcc[k];
    if (q.species != si) continue;
    MccProcDev& d = out[np++];
    d.type = q.type;
    d.e_species = q.product_electron_species;
    d.i_species = q.product_ion_species;
    d.slot = (int)k;
    d.n = q.target_density;
    d.thr = q.threshold_ev;
    d.smax = q.sigma_max_m2;
    d.loss = q.energy_loss_ev;
    d.eref = q.reference_energy_ev;
    d.pidx = q.power_index;
    d.alpha = q.alpha;
    d.beta = q.beta;
    d.clog = q.coulomb_log;
  }
  return np;
}

// Before a step's push: the majorant of every colliding species from the
// largest |v|^2 of the previous push (from the particles themselves after a
// load), which also clears the push's accumulators.
void upload_seg_mcc(scdat_ctx* c);

template <typename T>
void mcc_prepare(scdat_ctx* c, bool majorants = true) {
  hipStream_t s = c->stream;
  const int ns = (int)c->sp.size();
  // controls and |v|^2 maxima per (species, scenario)
  const size_t nseg = c->segmented() ? (size_t)c->n_seg() : 1;
  if (c->d_mcc_ctl.bytes < 8 * (size_t)kMccCtl * kMaxSpecies * nseg) {
    c->d_mcc_ctl.alloc(8 * (size_t)kMccCtl * kMaxSpecies * nseg);
    c->d_mcc_v2max.alloc(8 * (size_t)kMaxSpecies * nseg);
    c->d_mcc_ncand.alloc(4 * (size_t)kMaxSpecies);
    c->mcc_v2_stale = true;
  }
  if (c->segmented() && c->seg_mcc_dirty) upload_seg_mcc(c);
  unsigned long long* v2 = c->d_mcc_v2max.as<unsigned long long>();
  if (c->mcc_v2_stale) {
    launch_fill_u64(v2, (int64_t)(kMaxSpecies * nseg), s);
    for (int si = 0; si < ns; ++si) {
      SpeciesState& S = c->sp[si];
      if (!species_has_mcc(c, si) || S.n == 0) continue;
      launch_v2max<T>(S.buf[S.cur][3].as<T>(), S.buf[S.cur][4].as<T>(), S.buf[S.cur][5].as<T>(),
                      S.buf[S.cur][6].as<int32_t>(), S.n, v2 + si * nseg,
                      c->segmented() ? c->d_tet_seg.as<int32_t>() : nullptr, s);
    }
    c->mcc_v2_stale = false;
  }
  for (int si = 0; si < ns; ++si) {
    if (!species_has_mcc(c, si)) continue;
    SpeciesState& S = c->sp[si];
    if (S.mcc_cand.bytes < 4 * (size_t)S.cap) S.mcc_cand.alloc(4 * (size_t)S.cap);
    if (!majorants) continue;  // allocations and maxima only (before a stream capture)
    MccMajorantArgs m{};
    m.np = pack_procs(c, si, m.proc);
    m.mass = S.m;
    m.dt = c->cfg.dt;
    m.v2max = v2 + si * nseg;
    m.n_cand = c->d_mcc_ncand.as<unsigned int>() + si;
    m.ctl = c->d_mcc_ctl.as<double>() + (size_t)kMccCtl * si * nseg;
    if (c->segmented()) {
      m.seg_n = c->d_seg_mcc_n.as<double>();
      m.seg_dt = c->d_seg_dt64.as<double>();
      m.n_seg = (int)nseg;
    }
    launch_mcc_majorant(m, s);
  }
  HIPCHK(hipGetLastError());
}

template <typename T>
PushArgs<T> push_args(scdat_ctx* c, int si) {
  SpeciesState& S = c->sp[si];
  PushArgs<T> a{};
  a.tets = c->d_tets.as<TetMap<T>>();
  if constexpr (sizeof(T) == 4) a.wmap = c->d_wmap.as<WalkMap<T>>();
  a.nb = c->d_nb.as<int4>();
  a.groups = c->d_groups.as<int4>();
  a.partner = c->d_partner.as<int4>();
  a.E = c->d_E.as<Vec4<T>>();
  a.rho = c->d_rho[c->rho_target].as<unsigned long long>();
  a.periodic_mask = 0;
  a.reflect_mask = 0;
  for (int g = 0; g < kMaxGroups; ++g) {
    if (c->gtype[g] == SCDAT_BC_PERIODIC) a.periodic_mask |= 1u << g;
    if (c->gtype[g] == SCDAT_BC_REFLECT) a.reflect_mask |= 1u << g;
  }
  a.gshift = c->d_gshift.as<T>();
  a.x = S.buf[S.cur][0].as<T>();
  a.y = S.buf[S.cur][1].as<T>();
  a.z = S.buf[S.cur][2].as<T>();
  a.vx = S.buf[S.cur][3].as<T>();
  a.vy = S.buf[S.cur][4].as<T>();
  a.vz = S.buf[S.cur][5].as<T>();
  a.cell = S.buf[S.cur][6].as<int32_t>();
  a.n = S.n;
  if (c->capturing) {
    a.n = S.cap;
    a.n_dev = c->d_nslots.as<int64_t>() + si;
    a.step_dev = c->d_step.as<uint32_t>();
    a.n_pushed = c->d_nslots.as<int64_t>() + kMaxSpecies + si;
  }
  a.species = si;
  a.qm = (T)(S.q / S.m);
  a.cs = S.cs;
  a.mass = S.m;
  a.weight = S.w;
  a.energy_weight = std::max(1.0, S.w);
  a.dt = (T)c->cfg.dt;
  a.adaptive = c->cfg.adaptive_substep;
  a.floor_ratio = (T)c->cfg.floor_ratio;
  a.max_sub = c->cfg.max_substeps;
  a.max_walk = 256;
  a.tol = sizeof(T) == 4 ? (T)2e-6 : (T)1e-10;
  for (int r = 0; r < 3; ++r) a.B[r] = (T)c->cfg.bfield[r];
  const double bm = std::sqrt(c->cfg.bfield[0] * c->cfg.bfield[0] + c->cfg.bfield[1] * c->cfg.bfield[1] +
                              c->cfg.bfield[2] * c->cfg.bfield[2]);
  a.bmag = (T)bm;
  a.has_b = bm > 1e-20 ? 1 : 0;
  a.ledger_count = c->d_ledger_count.as<unsigned long long>();
  a.ledger_energy = c->d_ledger_energy.as<double>();
  a.lost = c->d_lost.as<unsigned long long>();
  a.pushes = c->d_pushes.as<unsigned long long>();
  a.ke_partial = S.ke_partial.as<double>();
  a.alive_partial = S.alive_partial.as<int32_t>();
  a.id = S.buf[S.cur][7].as<int32_t>();
  a.see_mask = 0;
  if (S.q > 0 && c->event_cap > 0) a.see_mask = c->fe_mask;  // ion impacts drive field emission
  if (S.q < 0 && c->event_cap > 0)
    for (int g = 0; g < kMaxGroups; ++g) {
      bool on = c->emit[g].see_model != SCDAT_SEE_OFF;
      for (const auto& se : c->seg_emit) on = on || se[g].see_model != SCDAT_SEE_OFF;
      if (on && (c->gtype[g] == SCDAT_BC_CONDUCTOR || c->gtype[g] == SCDAT_BC_ABSORB)) a.see_mask |= 1u << g;
    }
  if (c->event_cap > 0)
    for (int g = 0; g < kMaxGroups; ++g)
      if (c->gtype[g] == SCDAT_BC_HOST) a.see_mask |= 1u << g;  // impacts for the caller
  a.events = c->d_events.as<SeeEvent>();
  a.n_events = c->d_n_events.as<unsigned long long>();
  a.event_cap = c->event_cap;
  if (c->segmented()) {
    a.tet_seg = c->d_tet_seg.as<int32_t>();
    a.seg_dt = c->d_seg_dt.as<T>();
    a.seg_wscale = c->d_seg_wscale.as<double>();
  }
  a.chunks = c->push_chunks;
  if (sizeof(T) == 4 && c->lat_mode) a.lat = c->lat_dev;
  if (c->dep_mode && c->dep_cap > 0) a.dep = DepBuckets{c->d_dep_count.as<unsigned int>(), c->d_dep_ent.as<DepEntry>(),
                                                        c->dep_cap};
  if (c->push_mcc && species_has_mcc(c, si)) {
    const size_t nseg = c->segmented() ? (size_t)c->n_seg() : 1;
    a.mcc_ctl = c->d_mcc_ctl.as<double>() + (size_t)kMccCtl * si * nseg;
    a.mcc_cand = S.mcc_cand.as<uint32_t>();
    a.mcc_ncand = c->d_mcc_ncand.as<unsigned int>() + si;
    a.mcc_v2max = c->d_mcc_v2max.as<unsigned long long>() + si * nseg;
    if (c->segmented()) a.mcc_seg_key = c->d_seg_key.as<uint32_t>();
    a.mcc_step = (uint32_t)c->steps;
    a.mcc_key0 = c->key[0];
    a.mcc_key1 = c->key[1];
  }
  return a;
}

// Launch-shape tuner: for 1, 2, 4 and 8 particles per thread (each with and
// without lattice location where the mesh allows it), pushes
// cell-sorted scratch copies of every species for one sort cycle of that
// species (its sort cadence, at most 20 pushes; 10 without sorting) in the
// current field and times every push: a species' push slows as its
// particles drift from cell order, so the cycle's average is what the
// stepping loop sees.  Keeps the shape with the least summed average.  The
// context's particles, charge, ledger and statistics are untouched.
template <typename T>
uint32_t sort_species_copy(scdat_ctx* c, int si);

template <typename T>
void autotune_push(scdat_ctx* c) {
  const int ns = (int)c->sp.size();
  hipStream_t s = c->stream;
  const size_t m = (size_t)kMaxGroups * kMaxSpecies * c->n_seg();
  DevBuf rho, led_c, led_e, cnt, ke, alive;
  size_t kb = 16, ab = 16;
  for (const auto& S : c->sp) {
    kb = std::max(kb, S.ke_partial.bytes);
    ab = std::max(ab, S.alive_partial.bytes);
  }
  rho.alloc(8 * (size_t)c->h.nn);
  led_c.alloc(8 * m);
  led_e.alloc(8 * m);
  cnt.alloc(32);
  ke.alloc(kb);
  alive.alloc(ab);
  HIPCHK(hipMemsetAsync(led_c.p, 0, led_c.bytes, s));
  HIPCHK(hipMemsetAsync(led_e.p, 0, led_e.bytes, s));
  HIPCHK(hipMemsetAsync(cnt.p, 0, 32, s));
  std::vector<int> P((size_t)ns);
  int Pmax = 1;
  for (int si = 0; si < ns; ++si) {
    const int every = c->sp[si].sort_every > 0 ? c->sp[si].sort_every : c->cfg.sort_interval;
    P[si] = every > 0 ? std::min(every, 20) : 10;
    Pmax = std::max(Pmax, P[si]);
  }
  const int cand[4] = {1, 2, 4, 8};
  double best = 1e300;
  int pick = c->push_chunks, pick_lat = 0;
  std::vector<int64_t> nsc((size_t)ns, 0);
  // lattice location is tried where the mesh has its tables (FP32)
  const int n_lat = sizeof(T) == 4 && c->lat_dev.tet_hex ? 2 : 1;
  for (int kk = 0; kk < 4 * n_lat; ++kk) {
    const int k = kk % 4;
    c->lat_mode = kk / 4;
    // scratch copies in cell order, as the stepping loop has them after a sort
    for (int si = 0; si < ns; ++si)
      if (c->sp[si].n) nsc[si] = sort_species_copy<T>(c, si);
    double tsum = 0.0;
    for (int p = 0; p < Pmax; ++p) {
      HIPCHK(hipMemsetAsync(rho.p, 0, rho.bytes, s));
      for (int si = 0; si < ns; ++si) {
        SpeciesState& S = c->sp[si];
        if (!nsc[si] || p >= P[si]) continue;
        const int dst = 1 - S.cur;
        PushArgs<T> a = push_args<T>(c, si);
        a.n = nsc[si];
        a.x = S.buf[dst][0].as<T>();
        a.y = S.buf[dst][1].as<T>();
        a.z = S.buf[dst][2].as<T>();
        a.vx = S.buf[dst][3].as<T>();
        a.vy = S.buf[dst][4].as<T>();
        a.vz = S.buf[dst][5].as<T>();
        a.cell = S.buf[dst][6].as<int32_t>();
        a.id = S.buf[dst][7].as<int32_t>();
        a.rho = rho.as<unsigned long long>();
        a.ledger_count = led_c.as<unsigned long long>();
        a.ledger_energy = led_e.as<double>();
        a.lost = cnt.as<unsigned long long>();
        a.pushes = cnt.as<unsigned long long>() + 1;
        a.n_events = cnt.as<unsigned long long>() + 2;
        a.see_mask = 0;
        a.ke_partial = ke.as<double>();
        a.alive_partial = alive.as<int32_t>();
        a.chunks = cand[k];
        hipEvent_t e0 = take_kt_event(c, 2 * si), e1 = take_kt_event(c, 2 * si + 1);
        launch_push<T>(a, s, e0, e1);
        HIPCHK(hipGetLastError());
      }
      if (c->dep_mode && c->dep_cap > 0)
        launch_dep_reduce(DepBuckets{c->d_dep_count.as<unsigned int>(), c->d_dep_ent.as<DepEntry>(), c->dep_cap},
                          rho.as<unsigned long long>(), c->h.nn, s);
      HIPCHK(hipStreamSynchronize(s));
      for (int si = 0; si < ns; ++si) {
        if (!nsc[si] || p >= P[si]) continue;
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, c->kt_pool[2 * si], c->kt_pool[2 * si + 1]));
        tsum += t / P[si];
      }
    }
    if (tsum < best) {
      best = tsum;
      pick = cand[k];
      pick_lat = c->lat_mode;
    }
  }
  c->push_chunks = pick;
  c->lat_mode = pick_lat;
}

template <typename T>
void zero_rho(scdat_ctx* c, int k) {
  if (!c->rho_clean[k]) {
    // a fill kernel while recording (no memset nodes in captured graphs)
    if (c->capturing) launch_fill_u64(c->d_rho[k].as<unsigned long long>(), c->h.nn, c->stream);
    else HIPCHK(hipMemsetAsync(c->d_rho[k].p, 0, 8 * (size_t)c->h.nn, c->stream));
    c->rho_clean[k] = true;
  }
}

hipEvent_t take_event(scdat_ctx* c, size_t idx) {
  while (c->ev_pool.size() <= idx) {
    hipEvent_t e;
    HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
    c->ev_pool.push_back(e);
  }
  return c->ev_pool[idx];
}

// One push of all species into rho_target; events bracket each launch.
template <typename T>
void do_push(scdat_ctx* c, std::vector<std::pair<size_t, size_t>>* ev, size_t* evn) {
  zero_rho<T>(c, c->rho_target);
  for (size_t si = 0; si < c->sp.size(); ++si) {
    if ((c->capturing ? c->sp[si].cap : c->sp[si].n) == 0) continue;
    const PushArgs<T> a = push_args<T>(c, (int)si);
    c->sp[si].stat_blocks = push_blocks(a.n, a.chunks);
    if (ev) {
      const size_t e0 = (*evn)++, e1 = (*evn)++;
      launch_push<T>(a, c->stream, take_event(c, e0), take_event(c, e1));
      ev->push_back({e0, e1});
    } else {
      launch_push<T>(a, c->stream);
    }
    HIPCHK(hipGetLastError());
  }
  if (c->dep_mode && c->dep_cap > 0)
    launch_dep_reduce(DepBuckets{c->d_dep_count.as<unsigned int>(), c->d_dep_ent.as<DepEntry>(), c->dep_cap},
                      c->d_rho[c->rho_target].as<unsigned long long>(), c->h.nn, c->stream);
  c->rho_clean[c->rho_target] = false;
  c->rho_last = c->rho_target;
  c->rho_target = 1 - c->rho_target;
}

// Deposit buckets sized for the species' capacities: room for half an entry
// per particle slot and range (a push makes about a quarter: a block's
// table holds the nodes of its particles' tets), at most 32768 per range.
void dep_setup(scdat_ctx* c) {
  int64_t cap = 0;
  for (const auto& S : c->sp) cap += S.cap;
  const int64_t ranges = (c->h.nn + kDepRangeRows - 1) / kDepRangeRows;
  if (ranges <= 0 || cap <= 0) {
    c->dep_cap = 0;
    return;
  }
  int64_t per = std::max<int64_t>(1024, cap / (2 * ranges));
  int64_t p2 = 1024;
  while (p2 < per && p2 < 32768) p2 <<= 1;
  if (c->dep_cap == (int)p2 && c->d_dep_count.bytes >= 4 * (size_t)ranges) return;
  c->d_dep_count.alloc(4 * (size_t)ranges);
  HIPCHK(hipMemsetAsync(c->d_dep_count.p, 0, c->d_dep_count.bytes, c->stream));
  c->d_dep_ent.alloc(sizeof(DepEntry) * (size_t)ranges * (size_t)p2);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->dep_cap = (int)p2;
}

static uint64_t morton3(const double* p, const double* lo, const double* hi);

// Sort rank of every cell: its centroid's position along a Morton curve over
// the bounding box of its scenario (scenario ranges kept, ties by index), so
// that a stretch of sorted particles covers a compact region, in every
// direction, and keeps it while the particles drift between sorts.
void build_cell_key(scdat_ctx* c) {
  const HostMesh& h = c->h;
  const int64_t nt = h.nt;
  const int ns = c->segmented() ? c->n_seg() : 1;
  auto seg_of = [&](int64_t t) { return c->segmented() ? c->tet_seg_h[t] : 0; };
  std::vector<double> cen(3 * (size_t)nt);
  std::vector<double> lo(3 * (size_t)ns, 1e300), hi(3 * (size_t)ns, -1e300);
  for (int64_t t = 0; t < nt; ++t) {
    const int sg = seg_of(t);
    for (int r = 0; r < 3; ++r) {
      double x = 0.0;
      for (int v = 0; v < 4; ++v) x += h.xyz[3 * (size_t)h.tets[4 * t + v] + r];
      cen[3 * t + r] = 0.25 * x;
      lo[3 * sg + r] = std::min(lo[3 * sg + r], cen[3 * t + r]);
      hi[3 * sg + r] = std::max(hi[3 * sg + r], cen[3 * t + r]);
    }
  }
  std::vector<std::pair<uint64_t, int32_t>> key((size_t)nt);
  for (int64_t t = 0; t < nt; ++t) {
    const int sg = seg_of(t);
    key[t] = {morton3(&cen[3 * t], &lo[3 * sg], &hi[3 * sg]), (int32_t)t};
  }
  // scenarios are consecutive tet ranges: sort inside each
  int64_t b = 0;
  while (b < nt) {
    int64_t e = b + 1;
    while (e < nt && seg_of(e) == seg_of(b)) ++e;
    std::sort(key.begin() + b, key.begin() + e);
    b = e;
  }
  std::vector<int32_t> rank((size_t)nt);
  for (int64_t k = 0; k < nt; ++k) rank[key[k].second] = (int32_t)k;
  c->d_cell_key.upload(rank, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->cell_key_ready = true;
}

// Sorts species si from its current buffer into the other one and returns
// the alive count; the caller decides whether the sorted copy becomes the
// species' state.
template <typename T>
uint32_t sort_species_copy(scdat_ctx* c, int si) {
  SpeciesState& S = c->sp[si];
  if (c->cell_order == 1 && !c->cell_key_ready) build_cell_key(c);
  SortArgs<T> a{};
  const int s0 = S.cur, s1 = 1 - S.cur;
  a.src = {S.buf[s0][0].as<T>(), S.buf[s0][1].as<T>(), S.buf[s0][2].as<T>(), S.buf[s0][3].as<T>(),
           S.buf[s0][4].as<T>(), S.buf[s0][5].as<T>(), S.buf[s0][6].as<int32_t>(),
           S.buf[s0][7].as<int32_t>()};
  a.dst = {S.buf[s1][0].as<T>(), S.buf[s1][1].as<T>(), S.buf[s1][2].as<T>(), S.buf[s1][3].as<T>(),
           S.buf[s1][4].as<T>(), S.buf[s1][5].as<T>(), S.buf[s1][6].as<int32_t>(),
           S.buf[s1][7].as<int32_t>()};
  a.n = S.n;
  a.ncell = c->h.nt;
  a.key = c->cell_order == 1 ? c->d_cell_key.as<int32_t>() : nullptr;
  a.count = c->d_count.as<uint32_t>();
  a.offset = c->d_offset.as<uint32_t>();
  a.cursor = c->d_cursor.as<uint32_t>();
  a.tiles = c->d_tiles.as<uint32_t>();
  a.total = c->d_total.as<uint32_t>();
  if (c->d_perm.bytes < 4 * (size_t)S.n) throw Error(SCDAT_ERR_STATE, "sort: permutation buffer too small");
  a.perm = c->d_perm.as<uint32_t>();
  if (c->d_stage.bytes < sizeof(SortRec<T>) * (size_t)S.n) throw Error(SCDAT_ERR_STATE, "sort: staging buffer too small");
  a.stage = c->d_stage.as<SortRec<T>>();
  if (c->capturing) {
    // the grids cover the capacity; the count stays on the device
    a.n = S.cap;
    a.n_dev = c->d_nslots.as<int64_t>() + si;
    launch_sort<T>(a, c->stream);
    HIPCHK(hipGetLastError());
    return 0;
  }
  launch_sort<T>(a, c->stream);
  HIPCHK(hipGetLastError());
  uint32_t tot = 0;
  HIPCHK(hipMemcpyAsync(&tot, a.total, 4, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return tot;
}

template <typename T>
void do_sort_species(scdat_ctx* c, int si) {
  SpeciesState& S = c->sp[si];
  if ((c->capturing ? S.cap : S.n) == 0) return;
  const uint32_t tot = sort_species_copy<T>(c, si);
  S.cur = 1 - S.cur;
  if (!c->capturing) S.n = tot;
}

template <typename T>
void do_sort(scdat_ctx* c) {
  for (size_t si = 0; si < c->sp.size(); ++si) do_sort_species<T>(c, (int)si);
}

SurfEmitDev emit_dev(const scdat_surface_emission& q) {
  SurfEmitDev d{};
  d.see_model = q.see_model;
  d.e_species = q.electron_species;
  d.dmax = q.see_delta_max;
  d.emax = q.see_energy_max_ev;
  d.t_see = q.see_temperature_ev;
  d.photo_j = q.photo_current_density;
  d.t_ph = q.photo_temperature_ev;
  d.v_ref = q.photo_reference_v;
  return d;
}

void upload_seg_emit(scdat_ctx* c) {
  std::vector<SurfEmitDev> v((size_t)c->n_seg() * kMaxGroups);
  for (int sg = 0; sg < c->n_seg(); ++sg)
    for (int g = 0; g < kMaxGroups; ++g) v[(size_t)sg * kMaxGroups + g] = emit_dev(c->seg_emit[sg][g]);
  c->d_seg_emit.upload(v, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->seg_emit_dirty = false;
}

// Per-scenario device tables (after the species are known: the charge unit
// of scenario s is that of the reference weight times its multiplier).
template <typename T>
void upload_segments(scdat_ctx* c) {
  if (!c->segmented()) return;
  if (c->n_patch > 0) throw Error(SCDAT_ERR_STATE, "conductor patches need a single scenario");
  const int ns = c->n_seg();
  hipStream_t s = c->stream;
  c->d_tet_seg.upload(c->tet_seg_h, s);
  c->d_blk_seg.upload(c->blk_seg_h, s);
  c->d_seg_b0.upload(c->seg_b0_h, s);
  std::vector<double> unit(ns), qinit(ns, c->q_init);
  std::vector<T> dt(ns);
  for (int sg = 0; sg < ns; ++sg) {
    unit[sg] = kElem * (c->w_ref * c->seg_wscale[sg]) / 4294967296.0;
    dt[sg] = (T)c->seg_dt[sg];
  }
  c->d_seg_unit.upload(unit, s);
  c->d_seg_dt.upload(dt, s);
  c->d_seg_wscale.upload(c->seg_wscale, s);
  c->d_seg_qinit.upload(qinit, s);
  c->d_seg_key.upload(c->seg_key, s);
  // conductor-row sums kc are sorted by node: split them by scenario
  c->kc_off_h.assign((size_t)ns + 1, 0);
  for (int sg = 0; sg <= ns; ++sg)
    c->kc_off_h[sg] = std::lower_bound(c->h.kc_idx.begin(), c->h.kc_idx.end(), (int32_t)c->seg_node0[sg]) -
                      c->h.kc_idx.begin();
  c->d_kc_off.upload(c->kc_off_h, s);
  c->seg_ccap.assign((size_t)ns, 0.0);
  c->d_seg_ccap.upload(c->seg_ccap, s);
  c->d_seg_dt64.upload(c->seg_dt, s);
  build_face_table(c);
  c->seg_mcc_dirty = true;
  upload_seg_emit(c);
}

// Per-scenario collision densities on the device: each scenario's override,
// else the process's own target density.
void upload_seg_mcc(scdat_ctx* c) {
  const int ns = c->n_seg();
  std::vector<double> d((size_t)ns * kMaxProc, 0.0);
  for (int sg = 0; sg < ns; ++sg)
    for (size_t k = 0; k < c->mcc.size() && k < (size_t)kMaxProc; ++k) {
      const double o = c->seg_mcc_n.size() == d.size() ? c->seg_mcc_n[(size_t)sg * kMaxProc + k] : NAN;
      d[(size_t)sg * kMaxProc + k] = std::isnan(o) ? c->mcc[k].target_density : o;
    }
  c->d_seg_mcc_n.upload(d, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
  c->seg_mcc_dirty = false;
}

// Sort cadence per species: a species is re-sorted by cell once its
// particles have moved about kSortCells cells since the last sort (rms speed
// of the uploaded particles x dt / cell size), at most every
// cfg.sort_interval steps.  At 64^3 with 1 eV plasma, electrons (1.6 cells
// per step) sort every 5 steps, protons every 20: the electron push spends
// 20 % less and the sort costs less than that.
constexpr double kSortCells = 8.0;
void set_sort_cadence(scdat_ctx* c) {
  const HostMesh& h = c->h;
  double vsum = 0.0;
  for (int64_t t = 0; t < h.nt; ++t) vsum += h.vol[t];
  const double hcell = h.nt > 0 ? std::cbrt(6.0 * vsum / (double)h.nt) : 0.0;
  double dt = c->cfg.dt;
  for (double d : c->seg_dt) dt = std::max(dt, d);
  for (auto& S : c->sp) {
    S.sort_every = 0;
    const double cells = hcell > 0.0 ? std::sqrt(S.v2_mean) * dt / hcell : 0.0;
    if (c->cfg.sort_interval <= 0 || !(cells > 0.0)) continue;
    const double k = std::round(kSortCells / cells);
    const double lo = std::min(2, c->cfg.sort_interval);
    S.sort_every = (int)std::max(lo, std::min((double)c->cfg.sort_interval, k));
  }
}

template <typename T>
void do_initialize(scdat_ctx* c) {
  // Periodic partners.
  HostMesh& h = c->h;
  h.partner.assign(4 * (size_t)h.nt, -1);
  for (int g = 0; g < kMaxGroups; ++g) {
    if (c->gtype[g] != SCDAT_BC_PERIODIC) continue;
    const int pg = c->gpartner[g];
    if (pg < 0 || pg >= kMaxGroups || c->gtype[pg] != SCDAT_BC_PERIODIC)
      throw Error(SCDAT_ERR_ARG, "periodic group " + std::to_string(g) + " lacks a periodic partner");
    if (pg < g) continue;
    const int m = periodic_partners(h.nn, h.xyz.data(), h.nt, h.tets.data(), h.group.data(), g, pg,
                                    c->gshift[g].data(), h.partner);
    if (m == 0) throw Error(SCDAT_ERR_ARG, "periodic groups do not match under the given shift");
  }
  c->d_partner.upload(h.partner, c->stream);
  {
    std::vector<T> sh(3 * kMaxGroups);
    for (int g = 0; g < kMaxGroups; ++g)
      for (int r = 0; r < 3; ++r) sh[3 * g + r] = (T)c->gshift[g][r];
    c->d_gshift.upload(sh, c->stream);
  }
  upload_segments<T>(c);
  dep_setup(c);
  bool host_groups = false;
  for (int g = 0; g < kMaxGroups; ++g) host_groups = host_groups || c->gtype[g] == SCDAT_BC_HOST;
  if (host_groups && c->event_cap == 0) {
    // impacts per step: a 1/64 of all species' capacity, at least 65536
    int64_t cap = 0;
    for (const auto& S : c->sp) cap += S.cap;
    c->event_cap = std::max<int64_t>(65536, cap / 64);
    c->d_events.alloc(sizeof(SeeEvent) * (size_t)c->event_cap);
  }
  solve_phi1<T>(c);
  if (c->n_patch > 0) {
    c->patch_q.fill(0.0);
    c->ledger_pending = false;
    advance_patches(c, false);
  }
  // Initial deposit -> Poisson -> E (PICCycle::initialize + first field).
  zero_rho<T>(c, 0);
  c->rho_target = 0;
  for (size_t si = 0; si < c->sp.size(); ++si) {
    if (c->sp[si].n == 0) continue;
    launch_deposit<T>(push_args<T>(c, (int)si), c->stream);
  }
  c->rho_clean[0] = false;
  c->rho_last = 0;
  c->rho_target = 1;
  do_solve<T>(c, 0);
  zero_rho<T>(c, 1);
  do_field<T>(c, nullptr);
  HIPCHK(hipStreamSynchronize(c->stream));
  set_sort_cadence(c);
  c->mcc_v2_stale = true;
  c->initialized = true;
}

bool has_post_push(const scdat_ctx* c) {
  if (!c->mcc.empty() || c->event_cap > 0 || c->fe_mask) return true;
  for (const auto& e : c->emit)
    if (e.photo_current_density > 0.0) return true;
  for (const auto& se : c->seg_emit)
    for (const auto& e : se)
      if (e.photo_current_density > 0.0) return true;
  return false;
}

template <typename T>
SpeciesOut<T> species_out(scdat_ctx* c, int si) {
  SpeciesState& S = c->sp[si];
  SpeciesOut<T> o{};
  o.x = S.buf[S.cur][0].as<T>();
  o.y = S.buf[S.cur][1].as<T>();
  o.z = S.buf[S.cur][2].as<T>();
  o.vx = S.buf[S.cur][3].as<T>();
  o.vy = S.buf[S.cur][4].as<T>();
  o.vz = S.buf[S.cur][5].as<T>();
  o.cell = S.buf[S.cur][6].as<int32_t>();
  o.id = S.buf[S.cur][7].as<int32_t>();
  o.base = S.n;
  if (c->capturing) o.base_dev = c->d_nslots.as<int64_t>() + si;
  o.cap = S.cap;
  o.append = c->d_append.as<unsigned long long>() + si;
  o.m = S.m;
  o.cs = S.cs;
  return o;
}

// Collisions, secondary emission from the push's impacts and photoemission,
// after the push and before the Poisson solve; new particles are deposited
// into the charge of this step.  Species slot counts are updated once the
// stream has synchronised (apply_appends).
// Field-emission cathodes (FieldEmissionBoundaryCondition::processParticle,
// Tools/Boundary/src/FieldEmissionBoundaryCondition.cpp:88-178): every ion
// impact on such a group is absorbed and, in impact order (here: particle id
// order), adds J A dt / e expected field electrons and yield * w_ion / w_sec
// expected secondaries to the group's running accumulators; the integer
// parts (at most max_*_per_event) are emitted along the inward normal at the
// configured speeds, placed 1e-10 m apart off the impact point.  Host side:
// the impacts of a step are few; the new electrons are appended and
// deposited on the device.
template <typename T>
void field_emission(scdat_ctx* c, unsigned long long* counts) {
  if (!c->fe_mask || c->event_cap <= 0) return;
  hipStream_t s = c->stream;
  unsigned long long ne = 0;
  HIPCHK(hipMemcpyAsync(&ne, c->d_n_events.p, 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const int64_t n = std::min<int64_t>((int64_t)ne, c->event_cap);
  if (n == 0) return;
  std::vector<SeeEvent> ev((size_t)n);
  HIPCHK(hipMemcpyAsync(ev.data(), c->d_events.p, sizeof(SeeEvent) * (size_t)n, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  std::vector<const SeeEvent*> hits;
  for (const auto& e : ev) {
    const int g = e.group_face & 0xFF;
    if (c->sp[e.species].q > 0 && ((c->fe_mask >> g) & 1u)) hits.push_back(&e);
  }
  std::sort(hits.begin(), hits.end(), [](const SeeEvent* a, const SeeEvent* b) {
    return a->species != b->species ? a->species < b->species : a->id < b->id;
  });
  const HostMesh& h = c->h;
  struct Out {
    std::vector<double> xs, vs;
    std::vector<int32_t> cells, ids;
  };
  std::array<Out, kMaxSpecies> out;
  std::vector<unsigned long long> per((size_t)kMaxGroups * kMaxSpecies, 0ull);
  for (const SeeEvent* e : hits) {
    const int g = e->group_face & 0xFF, face = e->group_face >> 8;
    const scdat_field_emission& p = c->fe[g];
    const int es = p.electron_species;
    std::vector<double>& xs = out[es].xs;
    std::vector<double>& vs = out[es].vs;
    std::vector<int32_t>& cells = out[es].cells;
    std::vector<int32_t>& ids = out[es].ids;
    // inward normal of the impact face: grad lambda_face of its tet
    double nrm[3];
    const double* A = &h.A[9 * (size_t)e->cell];
    for (int r = 0; r < 3; ++r)
      nrm[r] = face == 0 ? -(A[r] + A[3 + r] + A[6 + r]) : A[3 * (face - 1) + r];
    const double len = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for (int r = 0; r < 3; ++r) nrm[r] = len > 1e-12 ? nrm[r] / len : (r == 2 ? 1.0 : 0.0);
    // The impact point projected onto the face's plane in FP64, as the
    // reference projects the particle onto the boundary face
    // (PICCycle.cpp:1262-1264): an FP32 walk's crossing point can sit a
    // few 1e-11 m off the plane, comparable to the emission offsets.
    const double* P = &h.xyz[3 * (size_t)h.tets[4 * (size_t)e->cell + ((face + 1) & 3)]];
    const double dn = (e->x[0] - P[0]) * nrm[0] + (e->x[1] - P[1]) * nrm[1] + (e->x[2] - P[2]) * nrm[2];
    const double hp[3] = {e->x[0] - dn * nrm[0], e->x[1] - dn * nrm[1], e->x[2] - dn * nrm[2]};
    int made = 0;
    auto emit = [&](int64_t count, double speed) {
      for (int64_t k = 0; k < count; ++k) {
        const double off = 1.0e-10 * (double)(made + 1);
        for (int r = 0; r < 3; ++r) {
          xs.push_back(hp[r] + nrm[r] * off);
          vs.push_back(nrm[r] * speed);
        }
        cells.push_back(e->cell);
        ids.push_back(c->sp[es].next_id++);
        ++made;
      }
    };
    const double dt = c->cfg.dt;
    if (dt > 0.0 && p.emission_current_density_a_per_m2 > 0.0) {
      const double expect = p.emission_current_density_a_per_m2 * p.emission_area_m2 * dt / kElem;
      if (std::isfinite(expect) && expect > 0.0) {
        c->fe_pending[g] += expect;
        const int64_t cnt = std::min<int64_t>((int64_t)std::floor(c->fe_pending[g]), p.max_emitted_particles_per_event);
        c->fe_pending[g] = std::max(0.0, c->fe_pending[g] - (double)cnt);
        emit(cnt, p.emitted_electron_speed_m_per_s);
      }
    }
    if (p.enable_secondary_electron_emission && p.secondary_electron_yield_per_ion > 0.0) {
      const double expect = p.secondary_electron_yield_per_ion * std::max(0.0, c->sp[e->species].w) /
                            std::max(1e-12, p.secondary_electron_weight);
      if (std::isfinite(expect) && expect > 0.0) {
        c->fe_pending_sec[g] += expect;
        const int64_t cnt =
            std::min<int64_t>((int64_t)std::floor(c->fe_pending_sec[g]), p.max_secondary_particles_per_event);
        c->fe_pending_sec[g] = std::max(0.0, c->fe_pending_sec[g] - (double)cnt);
        emit(cnt, p.secondary_electron_speed_m_per_s);
      }
    }
    per[(size_t)g * kMaxSpecies + es] += (unsigned long long)made;
  }
  unsigned long long* rho = c->d_rho[c->rho_last].as<unsigned long long>();
  bool any = false;
  for (int es = 0; es < (int)c->sp.size(); ++es) {
    const Out& o = out[es];
    const int64_t m = (int64_t)o.cells.size();
    if (m == 0) continue;
    any = true;
    if (c->sp[es].n + m > c->sp[es].cap) throw Error(SCDAT_ERR_NOMEM, "field emission: electron species capacity exceeded");
    // staging buffers are reused per species: wait for the previous inject
    HIPCHK(hipStreamSynchronize(s));
    c->d_fe_x.upload(o.xs, s);
    c->d_fe_v.upload(o.vs, s);
    c->d_fe_cell.upload(o.cells, s);
    c->d_fe_id.upload(o.ids, s);
    launch_inject<T>(species_out<T>(c, es), m, c->d_fe_x.as<double>(), c->d_fe_v.as<double>(),
                     c->d_fe_cell.as<int32_t>(), c->d_fe_id.as<int32_t>(), c->d_tets.as<TetMap<T>>(), rho, counts, s);
    HIPCHK(hipGetLastError());
  }
  if (!any) return;
  // emitted counts per group and species (ledger): read, add, write back
  std::vector<unsigned long long> em((size_t)kMaxGroups * kMaxSpecies);
  HIPCHK(hipMemcpyAsync(em.data(), c->d_emit_count.p, 8 * em.size(), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (size_t k = 0; k < em.size(); ++k) em[k] += per[k];
  HIPCHK(hipMemcpyAsync(c->d_emit_count.p, em.data(), 8 * em.size(), hipMemcpyHostToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
}

template <typename T>
void post_push(scdat_ctx* c) {
  hipStream_t s = c->stream;
  unsigned long long* counts = c->d_counts.as<unsigned long long>();
  launch_fill_u64(c->d_append.as<unsigned long long>(), kMaxSpecies, s);
  field_emission<T>(c, counts);
  const uint32_t step = (uint32_t)c->steps;
  unsigned long long* rho = c->d_rho[c->rho_last].as<unsigned long long>();
  const int ns = (int)c->sp.size();
  for (int si = 0; si < ns; ++si) {
    MccArgs<T> a{};
    const int np = pack_procs(c, si, a.proc);
    SpeciesState& S = c->sp[si];
    if (np == 0 || (c->capturing ? S.cap : S.n) == 0) continue;
    if (c->capturing) a.step_dev = c->d_step.as<uint32_t>();
    a.cand = S.mcc_cand.as<uint32_t>();
    a.n_cand = c->d_mcc_ncand.as<unsigned int>() + si;
    a.ctl = c->d_mcc_ctl.as<double>() + (size_t)kMccCtl * si * (c->segmented() ? (size_t)c->n_seg() : 1);
    if (c->segmented()) {
      a.tet_seg = c->d_tet_seg.as<int32_t>();
      a.seg_n = c->d_seg_mcc_n.as<double>();
      a.seg_dt = c->d_seg_dt64.as<double>();
      a.seg_key = c->d_seg_key.as<uint32_t>();
    }
    a.x = S.buf[S.cur][0].as<T>();
    a.y = S.buf[S.cur][1].as<T>();
    a.z = S.buf[S.cur][2].as<T>();
    a.vx = S.buf[S.cur][3].as<T>();
    a.vy = S.buf[S.cur][4].as<T>();
    a.vz = S.buf[S.cur][5].as<T>();
    a.cell = S.buf[S.cur][6].as<int32_t>();
    a.id = S.buf[S.cur][7].as<int32_t>();
    a.n = c->capturing ? S.cap : S.n;
    a.species = si;
    a.np = np;
    a.mass = S.m;
    a.dt = c->cfg.dt;
    a.step = step;
    a.key0 = c->key[0];
    a.key1 = c->key[1];
    for (int o = 0; o < ns; ++o) a.out[o] = species_out<T>(c, o);
    a.tets = c->d_tets.as<TetMap<T>>();
    a.rho = rho;
    a.counts = counts;
    launch_mcc<T>(a, s);
    HIPCHK(hipGetLastError());
  }
  EmitArgs<T> e{};
  for (int g = 0; g < kMaxGroups; ++g) {
    const scdat_surface_emission& q = c->emit[g];
    e.g[g].see_model = q.see_model;
    e.g[g].e_species = q.electron_species;
    e.g[g].dmax = q.see_delta_max;
    e.g[g].emax = q.see_energy_max_ev;
    e.g[g].t_see = q.see_temperature_ev;
    e.g[g].photo_j = q.photo_current_density;
    e.g[g].t_ph = q.photo_temperature_ev;
    e.g[g].v_ref = q.photo_reference_v;
  }
  for (int o = 0; o < ns; ++o) {
    e.out[o] = species_out<T>(c, o);
    e.q_abs_w[o] = std::abs(c->sp[o].q) * c->sp[o].w;
    e.mass[o] = c->sp[o].m;
    if (c->sp[o].q < 0) e.electron_mask |= 1u << o;
  }
  e.tets = c->d_tets.as<TetMap<T>>();
  e.xyz = c->d_xyz64.as<double>();
  e.tet4 = c->d_tet4.as<int32_t>();
  e.rho = rho;
  e.counts = counts;
  e.emit_count = c->d_emit_count.as<unsigned long long>();
  e.step = step;
  if (c->capturing) e.step_dev = c->d_step.as<uint32_t>();
  e.key0 = c->key[0];
  e.key1 = c->key[1];
  e.dt = c->cfg.dt;
  e.events = c->d_events.as<SeeEvent>();
  e.n_events = c->d_n_events.as<unsigned long long>();
  e.event_cap = c->event_cap;
  e.face_tet = c->d_face_tet.as<int32_t>();
  e.face_k = c->d_face_k.as<int32_t>();
  e.face_cum = c->d_face_cum.as<double>();
  e.vfloat = &c->d_sc.as<PcgScalars>()->vfloat;
  if (c->segmented()) {
    if (c->seg_emit_dirty) upload_seg_emit(c);
    e.tet_seg = c->d_tet_seg.as<int32_t>();
    e.gseg = c->d_seg_emit.as<SurfEmitDev>();
    e.seg_key = c->d_seg_key.as<uint32_t>();
  }
  if (c->event_cap > 0) launch_see<T>(e, c->event_cap, s);
  for (int g = 0; g < kMaxGroups; ++g) {
    const int64_t f0 = c->face_off[g], f1 = c->face_off[g + 1];
    if (f1 <= f0) continue;
    EmitArgs<T> eg = e;
    eg.f0 = f0;
    eg.f1 = f1;
    eg.group = g;
    eg.is_conductor = c->gtype[g] == SCDAT_BC_CONDUCTOR && c->h.has_cond ? 1 : 0;
    double bound = -1.0;
    if (c->segmented()) {
      // one grid row per scenario, sized for the largest expected count
      const int ns = c->n_seg();
      for (int sg = 0; sg < ns; ++sg) {
        const scdat_surface_emission& q = c->seg_emit[sg][g];
        if (!(q.photo_current_density > 0.0)) continue;
        const SpeciesState& S = c->sp[q.electron_species];
        const double area = c->seg_face_area_h[(size_t)g * ns + sg];
        bound = std::max(bound, q.photo_current_density * area * c->seg_dt[sg] /
                                    (std::abs(S.q) * S.w * c->seg_wscale[sg]));
      }
      eg.seg_face_off = c->d_seg_face_off.as<int64_t>() + (size_t)g * (ns + 1);
      eg.seg_dt = c->d_seg_dt64.as<double>();
      eg.seg_wscale = c->d_seg_wscale.as<double>();
      eg.vfloat_stride = (int64_t)(sizeof(PcgScalars) / sizeof(double));
    } else {
      const scdat_surface_emission& q = c->emit[g];
      if (q.photo_current_density > 0.0) {
        const SpeciesState& S = c->sp[q.electron_species];
        bound = q.photo_current_density * c->face_area[g] * c->cfg.dt / (std::abs(S.q) * S.w);
      }
    }
    if (bound < 0.0) continue;
    if (bound > 4.0e9) throw Error(SCDAT_ERR_ARG, "photoemission: too many macro-electrons per step");
    launch_photo<T>(eg, (int64_t)std::floor(bound) + 1, c->segmented() ? c->n_seg() : 1, s);
  }
  HIPCHK(hipGetLastError());
  if (c->capturing) {
    // the slot counts advance on the device
    CountArgs ca{};
    ca.n_species = ns;
    for (int si = 0; si < ns; ++si) ca.cap[si] = c->sp[si].cap;
    launch_commit_counts(c->d_nslots.as<int64_t>(), c->d_append.as<unsigned long long>(), ca, s);
    return;
  }
  // appended counts and impacts, read once the stream has synchronised
  HIPCHK(hipMemcpyAsync(c->h_append, c->d_append.p, 8 * kMaxSpecies, hipMemcpyDeviceToHost, s));
  HIPCHK(hipMemcpyAsync(c->h_append + kMaxSpecies, c->d_n_events.p, 8, hipMemcpyDeviceToHost, s));
  c->append_pending = true;
}

// After a stream synchronisation: new slot counts (capped at the capacity,
// overflow is counted on the device) and dropped impacts.
void apply_appends(scdat_ctx* c) {
  if (!c->append_pending) return;
  for (size_t si = 0; si < c->sp.size(); ++si) {
    SpeciesState& S = c->sp[si];
    S.n = std::min<int64_t>(S.cap, S.n + (int64_t)c->h_append[si]);
  }
  const int64_t ne = (int64_t)c->h_append[kMaxSpecies];
  if (ne > c->event_cap) c->events_dropped += ne - c->event_cap;
  c->append_pending = false;
}

template <typename T>
void do_step(scdat_ctx* c, int n) {
  std::vector<std::pair<size_t, size_t>> pev;
  size_t evn = 0;
  struct StageEv { size_t a, b, cc, d, e; };
  std::vector<StageEv> sev;
  c->push_ms_sum = 0;
  c->push_launches = 0;
  c->batch_failed = 0;
  for (int k = 0; k < n; ++k) {
    StageEv se;
    se.a = evn++;
    HIPCHK(hipEventRecord(take_event(c, se.a), c->stream));
    const bool post = has_post_push(c);
    if (c->event_cap > 0) launch_fill_u64(c->d_n_events.as<unsigned long long>(), 1, c->stream);
    if (!c->mcc.empty()) {
      mcc_prepare<T>(c);
      c->push_mcc = true;
    }
    do_push<T>(c, &pev, &evn);
    c->push_mcc = false;
    if (post) post_push<T>(c);
    snapshot_ledger(c);
    se.b = evn++;
    HIPCHK(hipEventRecord(take_event(c, se.b), c->stream));
    do_solve<T>(c, c->rho_last);
    apply_appends(c);
    advance_patches(c, true);
    se.cc = evn++;
    HIPCHK(hipEventRecord(take_event(c, se.cc), c->stream));
    // E for the next push; clears the buffer the next push deposits into.
    do_field<T>(c, c->d_rho[c->rho_target].as<unsigned long long>());
    c->rho_clean[c->rho_target] = true;
    se.d = evn++;
    HIPCHK(hipEventRecord(take_event(c, se.d), c->stream));
    c->steps++;
    c->sim_time += c->segmented() ? c->seg_dt[0] : c->cfg.dt;
    if (c->cfg.sort_interval > 0)
      for (size_t si = 0; si < c->sp.size(); ++si) {
        const int every = c->sp[si].sort_every > 0 ? c->sp[si].sort_every : c->cfg.sort_interval;
        if (c->steps % every == 0) do_sort_species<T>(c, (int)si);
      }
    se.e = evn++;
    HIPCHK(hipEventRecord(take_event(c, se.e), c->stream));
    sev.push_back(se);
  }
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int k = 0; k < 5; ++k) c->stage_ms[k] = 0;
  for (auto& se : sev) {
    float t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.a], c->ev_pool[se.b]));
    c->stage_ms[0] += t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.b], c->ev_pool[se.cc]));
    c->stage_ms[2] += t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.cc], c->ev_pool[se.d]));
    c->stage_ms[3] += t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.d], c->ev_pool[se.e]));
    c->stage_ms[4] += t;
  }
  for (auto& pe : pev) {
    float t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[pe.first], c->ev_pool[pe.second]));
    c->push_ms_sum += t;
    c->push_launches++;
    if (c->kernel_timing) {
      c->kt_sum[0] += t;
      c->kt_n[0] += 1;
    }
  }
}

// Host bookkeeping the launches of a captured step change, saved before a
// recording and restored after it, so that a replay and a recording leave
// the same host state.
struct HostLaunchState {
  bool rho_clean[2];
  int rho_last, rho_target;
  int cur[kMaxSpecies];
  bool push_mcc, have_xprev;
  static HostLaunchState of(const scdat_ctx* c) {
    HostLaunchState h{};
    h.rho_clean[0] = c->rho_clean[0];
    h.rho_clean[1] = c->rho_clean[1];
    h.rho_last = c->rho_last;
    h.rho_target = c->rho_target;
    for (size_t k = 0; k < c->sp.size(); ++k) h.cur[k] = c->sp[k].cur;
    h.push_mcc = c->push_mcc;
    h.have_xprev = c->have_xprev;
    return h;
  }
  void restore(scdat_ctx* c) const {
    c->rho_clean[0] = rho_clean[0];
    c->rho_clean[1] = rho_clean[1];
    c->rho_last = rho_last;
    c->rho_target = rho_target;
    for (size_t k = 0; k < c->sp.size(); ++k) c->sp[k].cur = cur[k];
    c->push_mcc = push_mcc;
    c->have_xprev = have_xprev;
  }
};

// Replays the graph recorded for `key`, recording it first (fn issues the
// launches) when this configuration has not run in the current call.
template <typename F>
void replay(scdat_ctx* c, uint64_t key, F&& fn) {
  auto it = c->gexec.find(key);
  if (it == c->gexec.end()) {
    const HostLaunchState saved = HostLaunchState::of(c);
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    try {
      fn();
    } catch (...) {
      c->capturing = false;
      (void)hipStreamEndCapture(c->stream, &g);
      if (g) (void)hipGraphDestroy(g);
      saved.restore(c);
      throw;
    }
    c->capturing = false;
    HIPCHK(hipStreamEndCapture(c->stream, &g));
    hipGraphExec_t e = nullptr;
    const hipError_t rc = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(rc);
    saved.restore(c);
    it = c->gexec.emplace(key, e).first;
    c->graph_builds++;
  }
  HIPCHK(hipGraphLaunch(it->second, c->stream));
  c->graph_replays++;
}

// Steps with the launches around the Poisson iteration replayed from HIP
// graphs (SCDAT_SHAPE_STEP_GRAPH): the same kernels with the same arguments
// in the same order as do_step, so the results are the same bits; the slot
// counts and the step number (the random streams' counter) are read by the
// kernels from the device, and the PCG iteration stays host-driven (its
// inner counts vary from step to step).  Per step: graph 1 = push, impacts,
// collisions, emission, slot counts, right-hand side and initial guess;
// the solve; graph 2 = floating potential, phi and E, the sorts due.
template <typename T>
void do_step_graph(scdat_ctx* c, int n) {
  hipStream_t s = c->stream;
  const int ns = (int)c->sp.size();
  c->push_ms_sum = 0;
  c->push_launches = 0;
  c->batch_failed = 0;
  for (int k = 0; k < 5; ++k) c->stage_ms[k] = 0;
  // before recording: uploads, allocations and stale maxima, and the counts
  if (c->segmented() && c->seg_emit_dirty) upload_seg_emit(c);
  if (!c->mcc.empty()) mcc_prepare<T>(c, false);
  if (c->cell_order == 1 && !c->cell_key_ready) build_cell_key(c);
  // slot counts, then the counts the last push of each species covered
  if (c->d_nslots.bytes < 16 * (size_t)kMaxSpecies) c->d_nslots.alloc(16 * (size_t)kMaxSpecies);
  if (c->d_step.bytes < 4) c->d_step.alloc(16);
  {
    std::vector<int64_t> cnt(kMaxSpecies, 0);
    for (int si = 0; si < ns; ++si) cnt[si] = c->sp[si].n;
    cnt.resize(2 * kMaxSpecies, 0);
    const uint32_t step = (uint32_t)c->steps;
    HIPCHK(hipMemcpyAsync(c->d_nslots.p, cnt.data(), 16 * (size_t)kMaxSpecies, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(c->d_step.p, &step, 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
  }
  c->drop_graphs();
  const bool post = has_post_push(c);
  struct StageEv { size_t a, b, cc, e; };
  std::vector<StageEv> sev;
  size_t evn = 0;
  try {
    for (int k = 0; k < n; ++k) {
      StageEv se;
      se.a = evn++;
      HIPCHK(hipEventRecord(take_event(c, se.a), s));
      uint64_t cur = 0;
      for (int si = 0; si < ns; ++si) cur |= (uint64_t)c->sp[si].cur << si;
      const uint64_t kpre = 1ull | (uint64_t)c->rho_target << 1 | (uint64_t)c->rho_clean[c->rho_target] << 2 |
                            cur << 8;
      replay(c, kpre, [&] {
        if (c->event_cap > 0) launch_fill_u64(c->d_n_events.as<unsigned long long>(), 1, s);
        if (!c->mcc.empty()) {
          mcc_prepare<T>(c);
          c->push_mcc = true;
        }
        do_push<T>(c, nullptr, nullptr);
        c->push_mcc = false;
        if (post) post_push<T>(c);
        solve_setup<T>(c, c->rho_last);
      });
      // what the recorded launches do to the host state
      c->rho_clean[c->rho_target] = false;
      c->rho_last = c->rho_target;
      c->rho_target = 1 - c->rho_target;
      c->have_xprev = true;
      se.b = evn++;
      HIPCHK(hipEventRecord(take_event(c, se.b), s));
      solve_run<T>(c, false);
      se.cc = evn++;
      HIPCHK(hipEventRecord(take_event(c, se.cc), s));
      uint64_t sorted = 0;
      const int64_t next = c->steps + 1;
      if (c->cfg.sort_interval > 0)
        for (int si = 0; si < ns; ++si) {
          const int every = c->sp[si].sort_every > 0 ? c->sp[si].sort_every : c->cfg.sort_interval;
          if (next % every == 0 && c->sp[si].cap > 0) sorted |= 1ull << si;
        }
      const uint64_t kpost = 2ull | (uint64_t)c->rho_target << 2 | cur << 8 | sorted << 24;
      replay(c, kpost, [&] {
        FieldBufs f = fbufs(c);
        if (c->h.has_cond && c->n_patch == 0) launch_float(f, float_args(c), s);
        do_field<T>(c, c->d_rho[c->rho_target].as<unsigned long long>());
        for (int si = 0; si < ns; ++si)
          if ((sorted >> si) & 1ull) do_sort_species<T>(c, si);
        launch_step_advance(c->d_step.as<uint32_t>(), s);
      });
      c->rho_clean[c->rho_target] = true;
      for (int si = 0; si < ns; ++si)
        if ((sorted >> si) & 1ull) c->sp[si].cur = 1 - c->sp[si].cur;
      c->steps++;
      c->sim_time += c->segmented() ? c->seg_dt[0] : c->cfg.dt;
      se.e = evn++;
      HIPCHK(hipEventRecord(take_event(c, se.e), s));
      sev.push_back(se);
    }
  } catch (...) {
    (void)hipStreamSynchronize(s);
    c->drop_graphs();
    throw;
  }
  // the counts back on the host
  std::vector<int64_t> cnt(2 * kMaxSpecies, 0);
  HIPCHK(hipMemcpyAsync(cnt.data(), c->d_nslots.p, 16 * (size_t)kMaxSpecies, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  for (int si = 0; si < ns; ++si) {
    c->sp[si].n = cnt[si];
    // the statistics read the partials of the blocks the last push needed
    // (the grid covered the capacity; the others hold zeros)
    if (c->sp[si].cap > 0) c->sp[si].stat_blocks = push_blocks(cnt[kMaxSpecies + si], c->push_chunks);
  }
  for (auto& se : sev) {
    float t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.a], c->ev_pool[se.b]));
    c->stage_ms[0] += t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.b], c->ev_pool[se.cc]));
    c->stage_ms[2] += t;
    HIPCHK(hipEventElapsedTime(&t, c->ev_pool[se.cc], c->ev_pool[se.e]));
    c->stage_ms[3] += t;
  }
  c->drop_graphs();
}

// Whether this context's steps can be recorded: host-side stages (surface
// circuit, field-emission cathodes) need the eager loop.
bool graph_capable(const scdat_ctx* c) { return c->n_patch == 0 && c->fe_mask == 0; }

template <typename T>
void set_particles_t(scdat_ctx* c, int si, int64_t n, const double* x, const double* v,
                     const int32_t* cell) {
  SpeciesState& S = c->sp[si];
  if (n > S.cap) throw Error(SCDAT_ERR_NOMEM, "particle count exceeds species capacity");
  std::vector<int32_t> cl;
  if (cell) cl.assign(cell, cell + n);
  else host_locate(c->h, n, x, cl);
  for (int64_t i = 0; i < n; ++i)
    if (cl[i] < -1 || cl[i] >= c->h.nt) throw Error(SCDAT_ERR_ARG, "cell index out of range");
  double v2 = 0.0;
  for (int64_t i = 0; i < 3 * n; ++i) v2 += v[i] * v[i];
  S.v2_mean = n > 0 ? v2 / (double)n : 0.0;
  std::vector<T> f[6];
  for (int r = 0; r < 6; ++r) f[r].resize(n);
  for (int64_t i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r) {
      f[r][i] = (T)x[3 * i + r];
      f[3 + r][i] = (T)v[3 * i + r];
    }
  std::vector<int32_t> id(n);
  if (c->segmented()) {
    // ids count within each scenario, as in a context of that scenario alone
    std::vector<int32_t> next((size_t)c->n_seg(), 0);
    for (int64_t i = 0; i < n; ++i) id[i] = cl[i] >= 0 ? next[c->tet_seg_h[cl[i]]]++ : (int32_t)i;
  } else {
    for (int64_t i = 0; i < n; ++i) id[i] = (int32_t)i;
  }
  S.cur = 0;
  for (int r = 0; r < 6; ++r)
    if (n) HIPCHK(hipMemcpyAsync(S.buf[0][r].p, f[r].data(), sizeof(T) * n, hipMemcpyHostToDevice, c->stream));
  if (n) HIPCHK(hipMemcpyAsync(S.buf[0][6].p, cl.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  if (n) HIPCHK(hipMemcpyAsync(S.buf[0][7].p, id.data(), 4 * n, hipMemcpyHostToDevice, c->stream));
  S.n = n;
  S.next_id = (int32_t)n;
  do_sort_species<T>(c, si);
}

template <typename T>
void get_particles_t(scdat_ctx* c, int si, double* x, double* v, int32_t* cell, int32_t* id) {
  SpeciesState& S = c->sp[si];
  const int64_t n = S.n;
  if (n == 0) return;
  std::vector<T> f[6];
  for (int r = 0; r < 6; ++r) {
    f[r].resize(n);
    HIPCHK(hipMemcpyAsync(f[r].data(), S.buf[S.cur][r].p, sizeof(T) * n, hipMemcpyDeviceToHost, c->stream));
  }
  std::vector<int32_t> cl(n), ids(n);
  HIPCHK(hipMemcpyAsync(cl.data(), S.buf[S.cur][6].p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(ids.data(), S.buf[S.cur][7].p, 4 * n, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  for (int64_t i = 0; i < n; ++i)
    for (int r = 0; r < 3; ++r) {
      if (x) x[3 * i + r] = (double)f[r][i];
      if (v) v[3 * i + r] = (double)f[3 + r][i];
    }
  if (cell) std::memcpy(cell, cl.data(), 4 * n);
  if (id) std::memcpy(id, ids.data(), 4 * n);
}

// Device ledger, emission counts and energies of every scenario.
struct LedgerHost {
  std::vector<unsigned long long> cnt, emi;
  std::vector<double> en;
};
LedgerHost read_ledger(scdat_ctx* c) {
  const size_t tot = (size_t)kMaxGroups * kMaxSpecies * c->n_seg();
  LedgerHost L;
  L.cnt.resize(tot);
  L.emi.resize(tot);
  L.en.resize(tot);
  HIPCHK(hipMemcpyAsync(L.cnt.data(), c->d_ledger_count.p, 8 * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(L.emi.data(), c->d_emit_count.p, 8 * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(L.en.data(), c->d_ledger_energy.p, 8 * tot, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return L;
}
double seg_scale(const scdat_ctx* c, int sg) { return c->segmented() ? c->seg_wscale[sg] : 1.0; }
// Charge of one fixed-point unit at node i (its scenario's weight).
double node_unit(const scdat_ctx* c, int64_t i) {
  if (!c->segmented()) return c->unit;
  const int sg = (int)(std::upper_bound(c->seg_node0.begin(), c->seg_node0.end(), i) - c->seg_node0.begin()) - 1;
  return kElem * (c->w_ref * c->seg_wscale[sg]) / 4294967296.0;
}

// ---- internal node order ------------------------------------------------------

// 63-bit Morton key of a point in [lo, hi]^3 (21 bits per axis).
static uint64_t morton3(const double* p, const double* lo, const double* hi) {
  uint64_t key = 0;
  uint32_t q[3];
  for (int r = 0; r < 3; ++r) {
    const double t = (p[r] - lo[r]) / (hi[r] - lo[r] + 1e-300);
    q[r] = (uint32_t)std::min(2097151.0, std::max(0.0, t * 2097152.0));
  }
  for (int b = 20; b >= 0; --b)
    for (int r = 0; r < 3; ++r) key = (key << 1) | ((q[r] >> b) & 1u);
  return key;
}

// Lattice indices of the nodes when they lie on a rectilinear grid (every
// coordinate is one of a few grid-line values per axis, as on a structured
// box mesh whose lines are graded to put a conductor's faces on them):
// false when the grid lines would not be mostly occupied by nodes.
static bool lattice_indices(int64_t nn, const double* xyz, std::vector<int32_t>& ijk, int32_t dims[3],
                            std::vector<double>* lines_out) {
  if (nn < 8) return false;
  std::vector<double> line[3];
  for (int r = 0; r < 3; ++r) {
    std::vector<double> v((size_t)nn);
    for (int64_t i = 0; i < nn; ++i) v[i] = xyz[3 * i + r];
    std::sort(v.begin(), v.end());
    const double ext = v.back() - v.front();
    if (!(ext > 0.0)) return false;
    const double eps = 1e-9 * ext;
    for (double x : v)
      if (line[r].empty() || x - line[r].back() > eps) line[r].push_back(x);
    dims[r] = (int32_t)line[r].size();
  }
  if ((double)dims[0] * dims[1] * dims[2] > 2.0 * (double)nn) return false;
  ijk.resize(3 * (size_t)nn);
  for (int64_t i = 0; i < nn; ++i)
    for (int r = 0; r < 3; ++r) {
      const std::vector<double>& L = line[r];
      const double x = xyz[3 * i + r];
      // the grid line within eps below or at x
      size_t k = (size_t)(std::upper_bound(L.begin(), L.end(), x + 1e-9 * (L.back() - L.front())) - L.begin());
      ijk[3 * i + r] = (int32_t)(k - 1);
    }
  if (lines_out)
    for (int r = 0; r < 3; ++r) lines_out[r] = line[r];
  return true;
}

// Tiled order of a lattice (rectilinear grid) mesh: the lattice points with
// every index >= 1 in tiles of 8 x 8 x 8 points (tiles in x, y, z order,
// points inside a tile in x, y, z order), each tile filling one 512-row
// block; a tile's missing
// points (outside the mesh, or a removed box) are filled with nodes of the
// index-0 layers, the rest of which follow the tiles.  A row block is then a
// compact 8^3 box: the aggregate of the two-level coarse space.  perm[k] =
// caller's node of internal row k.  False when the nodes are not a lattice.
static bool tiled_perm(int64_t nn, const double* xyz, std::vector<int32_t>& perm, HostMesh& h) {
  std::vector<int32_t> ijk;
  int32_t dims[3];
  if (!lattice_indices(nn, xyz, ijk, dims, nullptr)) return false;
  int32_t nt[3];
  for (int r = 0; r < 3; ++r) nt[r] = (dims[r] - 1 + 7) / 8;
  const int64_t ntiles = (int64_t)nt[0] * nt[1] * nt[2];
  if (ntiles <= 0 || ntiles > (int64_t)1 << 22) return false;
  std::vector<int32_t> slot((size_t)ntiles * 512, -1);
  std::vector<int32_t> extra;
  for (int64_t u = 0; u < nn; ++u) {
    const int32_t* q = &ijk[3 * u];
    if (q[0] == 0 || q[1] == 0 || q[2] == 0) {
      extra.push_back((int32_t)u);
      continue;
    }
    const int32_t a = q[0] - 1, b = q[1] - 1, cc = q[2] - 1;
    const int64_t t = (a >> 3) + (int64_t)nt[0] * ((b >> 3) + (int64_t)nt[1] * (cc >> 3));
    const int64_t k = t * 512 + (a & 7) + 8 * (b & 7) + 64 * (cc & 7);
    if (slot[k] >= 0) extra.push_back((int32_t)u);  // two nodes at one point
    else slot[k] = (int32_t)u;
  }
  perm.clear();
  perm.reserve((size_t)nn);
  size_t e = 0;
  bool aligned = true;
  h.tile_of_block.clear();
  for (int64_t t = 0; t < ntiles; ++t) {
    bool any = false;
    for (int k = 0; k < 512 && !any; ++k) any = slot[t * 512 + k] >= 0;
    if (!any) continue;  // a tile with no node takes no block
    h.tile_of_block.push_back((int32_t)t);
    for (int k = 0; k < 512; ++k) {
      const int32_t u = slot[t * 512 + k];
      if (u >= 0) perm.push_back(u);
      else if (e < extra.size()) perm.push_back(extra[e++]);
      else aligned = false;  // the following tiles slide off their blocks
    }
  }
  while (e < extra.size()) perm.push_back(extra[e++]);
  for (int r = 0; r < 3; ++r) h.tdim[r] = nt[r];
  h.tile_blocks = (int64_t)h.tile_of_block.size();
  h.tiles_aligned = aligned;
  return (int64_t)perm.size() == nn;
}

// Mesh preparation with the nodes in the context's internal order: tiled
// (lattice meshes; 512-row blocks are 8^3 boxes), Morton order of the
// coordinates (consecutive row blocks are compact boxes on any mesh), or
// the caller's order.  perm[internal] = the caller's node index.  The tiled
// order falls back to Morton when the nodes are not a lattice.
void mesh_setup(scdat_ctx* c, int64_t nn, const double* xyz, int64_t nt, const int32_t* tets,
                const int32_t* kind, const double* dval, const int32_t* group, int order) {
  HostMesh& h = c->h;
  h = HostMesh();
  h.nn = nn;
  h.nt = nt;
  c->perm.clear();
  c->iperm.clear();
  if (order == SCDAT_ORDER_TILED && !tiled_perm(nn, xyz, c->perm, h)) {
    c->perm.clear();
    h.tile_of_block.clear();
    h.tile_blocks = 0;
    h.tiles_aligned = false;
    order = SCDAT_ORDER_MORTON;
  }
  if (order == SCDAT_ORDER_MORTON) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = 0; i < nn; ++i)
      for (int r = 0; r < 3; ++r) {
        lo[r] = std::min(lo[r], xyz[3 * i + r]);
        hi[r] = std::max(hi[r], xyz[3 * i + r]);
      }
    std::vector<std::pair<uint64_t, int32_t>> key((size_t)nn);
    for (int64_t i = 0; i < nn; ++i) key[i] = {morton3(&xyz[3 * i], lo, hi), (int32_t)i};
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<uint64_t, int32_t>& a, const std::pair<uint64_t, int32_t>& b) {
                       return a.first < b.first;
                     });
    c->perm.resize((size_t)nn);
    for (int64_t k = 0; k < nn; ++k) c->perm[k] = key[k].second;
  }
  if (!c->perm.empty()) {
    c->iperm.resize((size_t)nn);
    for (int64_t k = 0; k < nn; ++k) c->iperm[c->perm[k]] = (int32_t)k;
  }
  c->order_in_use = order;
  auto user = [&](int64_t k) { return c->perm.empty() ? k : (int64_t)c->perm[k]; };
  h.xyz.resize(3 * (size_t)nn);
  h.kind.resize((size_t)nn);
  h.dval.resize((size_t)nn);
  for (int64_t k = 0; k < nn; ++k) {
    const int64_t u = user(k);
    for (int r = 0; r < 3; ++r) h.xyz[3 * k + r] = xyz[3 * u + r];
    h.kind[k] = kind[u];
    h.dval[k] = dval[u];
  }
  h.tets.resize(4 * (size_t)nt);
  for (int64_t k = 0; k < 4 * nt; ++k) h.tets[k] = c->iperm.empty() ? tets[k] : c->iperm[tets[k]];
  h.group.assign(group, group + 4 * nt);
  prepare_mesh(h, c->cfg.poisson_operator, c->cfg.permittivity);
  build_coarse(h);
  if (c->is_fp64()) upload_mesh_device<double>(c);
  else upload_mesh_device<float>(c);
  c->has_mesh = true;
  c->initialized = false;
  c->cell_key_ready = false;
}

// Back to the caller's node order (set_scenarios: scenarios are ranges of
// the caller's node numbering).
void mesh_natural_order(scdat_ctx* c) {
  if (c->perm.empty()) return;
  const HostMesh h = c->h;
  const int64_t nn = h.nn;
  std::vector<double> xyz(3 * (size_t)nn), dv((size_t)nn);
  std::vector<int32_t> kind((size_t)nn), tets(h.tets.size());
  for (int64_t k = 0; k < nn; ++k) {
    const int32_t u = c->perm[k];
    for (int r = 0; r < 3; ++r) xyz[3 * (size_t)u + r] = h.xyz[3 * k + r];
    kind[u] = h.kind[k];
    dv[u] = h.dval[k];
  }
  for (size_t k = 0; k < tets.size(); ++k) tets[k] = c->perm[h.tets[k]];
  std::vector<int32_t> group(h.group);
  mesh_setup(c, nn, xyz.data(), h.nt, tets.data(), kind.data(), dv.data(), group.data(), SCDAT_ORDER_GIVEN);
}

// node index conversions at the API boundary
inline int64_t to_user(const scdat_ctx* c, int64_t k) { return c->perm.empty() ? k : c->perm[k]; }

// ---- discrete adjoint of the charging cycle ---------------------------------

struct ChargingRun {
  scdat_charging_params p;
  double a_lit = 0.0;  // sunlit conductor area
  uint32_t cond_mask = 0;
};

double photo_j(const ChargingRun& r, double V, double* dj_dv) {
  const double t = std::max(1e-3, r.p.photo_temperature_ev);
  const double over = V - r.p.photo_reference_v;
  const double e = std::exp(-std::max(0.0, over) / t);
  if (dj_dv) *dj_dv = over > 0.0 ? -r.p.photo_current_density * e / t : 0.0;
  return r.p.photo_current_density * e;
}

template <typename T>
void adj_alloc_step(scdat_ctx* c, AdjStep& st) {
  const int64_t np = c->adj_np;
  for (int r = 0; r < 6; ++r) st.v[r].alloc(sizeof(T) * (size_t)np);
  st.v[6].alloc(4 * (size_t)np);
  st.E.alloc(sizeof(Vec4<T>) * (size_t)c->h.nn);
}

template <typename T>
AdjState<T> adj_state(AdjStep& st) {
  return {st.v[0].as<T>(), st.v[1].as<T>(), st.v[2].as<T>(), st.v[3].as<T>(),
          st.v[4].as<T>(), st.v[5].as<T>(), st.v[6].as<int32_t>()};
}

template <typename T>
void adj_copy_step(scdat_ctx* c, AdjStep& dst, const AdjStep& src) {
  for (int r = 0; r < 7; ++r)
    HIPCHK(hipMemcpyAsync(dst.v[r].p, src.v[r].p, src.v[r].bytes, hipMemcpyDeviceToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(dst.E.p, src.E.p, src.E.bytes, hipMemcpyDeviceToDevice, c->stream));
}

// Checks, buffers for `steps_kept` working steps, the concatenated initial
// particles in adj_st[0] and a snapshot of the context's field state.
template <typename T>
ChargingRun adj_begin(scdat_ctx* c, const scdat_charging_params* p, int steps_kept) {
  if (!c->initialized) throw Error(SCDAT_ERR_STATE, "charging adjoint: call scdat_initialize first");
  if (c->steps != 0) throw Error(SCDAT_ERR_STATE, "charging adjoint: starts from the initialized state (no steps)");
  if (c->segmented() || c->n_patch > 0 || !c->h.has_cond)
    throw Error(SCDAT_ERR_STATE, "charging adjoint: needs one scenario with a floating conductor");
  for (int g = 0; g < kMaxGroups; ++g)
    if (c->gtype[g] == SCDAT_BC_PERIODIC || c->gtype[g] == SCDAT_BC_REFLECT)
      throw Error(SCDAT_ERR_STATE, "charging adjoint: periodic and reflecting groups are not differentiated");
  for (int r = 0; r < 3; ++r)
    if (c->cfg.bfield[r] != 0.0) throw Error(SCDAT_ERR_STATE, "charging adjoint: needs B = 0");
  if (!(p->see_energy_max_ev > 0.0) || p->see_delta_max < 0.0 || p->surface_conductance_s < 0.0)
    throw Error(SCDAT_ERR_ARG, "charging adjoint: bad parameters");
  ChargingRun run;
  run.p = *p;
  for (int k = 0; k < 6; ++k)
    if ((p->sunlit_faces >> k) & 1) run.a_lit += c->face_area[kAdjFaceGroup0 + k];
  for (int g = 0; g < kMaxGroups; ++g)
    if (c->gtype[g] == SCDAT_BC_CONDUCTOR) run.cond_mask |= 1u << g;
  int64_t np = 0;
  for (const auto& S : c->sp) np += S.n;
  const bool realloc = np != c->adj_np || (int)c->adj_st.size() < steps_kept;
  c->adj_np = np;
  hipStream_t s = c->stream;
  if (realloc) {
    c->adj_st.clear();
    c->adj_ck.clear();
    c->adj_ck_x64.clear();
  }
  while ((int)c->adj_st.size() < steps_kept) {
    c->adj_st.emplace_back();
    adj_alloc_step<T>(c, c->adj_st.back());
  }
  // species index and the concatenated initial state
  std::vector<int32_t> sp((size_t)np);
  int64_t off = 0;
  AdjStep& s0 = c->adj_st[0];
  for (size_t si = 0; si < c->sp.size(); ++si) {
    SpeciesState& S = c->sp[si];
    std::fill(sp.begin() + off, sp.begin() + off + S.n, (int32_t)si);
    for (int r = 0; r < 7; ++r) {
      const size_t es = r < 6 ? sizeof(T) : 4;
      if (S.n)
        HIPCHK(hipMemcpyAsync(static_cast<char*>(s0.v[r].p) + es * off, S.buf[S.cur][r].p, es * S.n,
                              hipMemcpyDeviceToDevice, s));
    }
    off += S.n;
  }
  c->d_adj_sp.upload(sp, s);
  HIPCHK(hipMemcpyAsync(s0.E.p, c->d_E.p, s0.E.bytes, hipMemcpyDeviceToDevice, s));
  const size_t nn = (size_t)c->h.nn;
  c->d_adj_rho.alloc(8 * nn);
  c->d_adj_part.alloc(8 * (size_t)kAdjPartials * (size_t)std::max(1, adj_blocks(std::max(np, c->h.nn))));
  c->d_adj_dotp.alloc(8 * (size_t)adj_blocks(c->h.nn));
  // snapshot of the context's field state (restored by adj_end)
  c->d_snap_E.alloc(c->d_E.bytes);
  c->d_snap_x64.alloc(c->d_x64.bytes);
  c->d_snap_phi64.alloc(c->d_phi64.bytes);
  c->d_snap_phiT.alloc(c->d_phiT.bytes);
  c->d_snap_sc.alloc(sizeof(PcgScalars));
  HIPCHK(hipMemcpyAsync(c->d_snap_E.p, c->d_E.p, c->d_E.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_snap_x64.p, c->d_x64.p, c->d_x64.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_snap_phi64.p, c->d_phi64.p, c->d_phi64.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_snap_phiT.p, c->d_phiT.p, c->d_phiT.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_snap_sc.p, c->d_sc.p, sizeof(PcgScalars), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  c->snap_solver = SolverFigures::of(c);
  return run;
}

void adj_end(scdat_ctx* c) {
  hipStream_t s = c->stream;
  HIPCHK(hipMemcpyAsync(c->d_E.p, c->d_snap_E.p, c->d_E.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_x64.p, c->d_snap_x64.p, c->d_x64.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_phi64.p, c->d_snap_phi64.p, c->d_phi64.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_phiT.p, c->d_snap_phiT.p, c->d_phiT.bytes, hipMemcpyDeviceToDevice, s));
  HIPCHK(hipMemcpyAsync(c->d_sc.p, c->d_snap_sc.p, sizeof(PcgScalars), hipMemcpyDeviceToDevice, s));
  HIPCHK(hipStreamSynchronize(s));
  c->snap_solver.restore(c);
}

template <typename T>
AdjArgs<T> adj_args(scdat_ctx* c, const ChargingRun& run) {
  AdjArgs<T> a{};
  a.mesh = push_args<T>(c, 0);
  a.sp = c->d_adj_sp.as<int32_t>();
  a.n = c->adj_np;
  for (size_t si = 0; si < c->sp.size(); ++si) {
    const SpeciesState& S = c->sp[si];
    a.qm[si] = (T)(S.q / S.m);
    a.qw[si] = S.q * S.w;
    a.mass[si] = S.m;
    a.electron[si] = S.q < 0 ? 1 : 0;
  }
  a.dt = (T)c->cfg.dt;
  a.cond_mask = run.cond_mask;
  for (int k = 0; k < 6; ++k) {
    const int g = kAdjFaceGroup0 + k;
    a.yield[g] = run.p.face_yield[k];
    a.dmax[g] = run.p.see_delta_max * run.p.face_yield[k];
  }
  a.delta_max = run.p.see_delta_max;
  a.emax = run.p.see_energy_max_ev;
  a.rho = c->d_adj_rho.as<double>();
  a.partial = c->d_adj_part.as<double>();
  return a;
}

// Block partials of the last adjoint kernel, summed in block order.
std::array<double, kAdjPartials> adj_partials(scdat_ctx* c, int blocks) {
  std::vector<double> h((size_t)blocks * kAdjPartials);
  HIPCHK(hipMemcpyAsync(h.data(), c->d_adj_part.p, 8 * h.size(), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::array<double, kAdjPartials> t{};
  for (int b = 0; b < blocks; ++b)
    for (int k = 0; k < kAdjPartials; ++k) t[k] += h[(size_t)b * kAdjPartials + k];
  return t;
}

// One forward step s0 -> s1: push + conductor charge (Q, V of step n in,
// of step n + 1 out) + Poisson + E of step n + 1 into s1.E.
template <typename T>
void adj_forward_step(scdat_ctx* c, const ChargingRun& run, AdjArgs<T> a, AdjStep& s0, AdjStep& s1, double& Q,
                      double& V, int step = 0) {
  hipStream_t s = c->stream;
  a.step = step;
  HIPCHK(hipMemsetAsync(c->d_adj_rho.p, 0, 8 * (size_t)c->h.nn, s));
  a.mesh.E = s0.E.as<Vec4<T>>();
  a.s0 = adj_state<T>(s0);
  a.s1 = adj_state<T>(s1);
  launch_adj_forward<T>(a, s);
  HIPCHK(hipGetLastError());
  const auto part = adj_partials(c, adj_blocks(a.n));
  double hit = 0.0;
  for (int k = 0; k < kAdjPartials; ++k) hit += part[k];
  const double dt = c->cfg.dt;
  Q += hit + dt * (photo_j(run, V, nullptr) * run.a_lit - run.p.surface_conductance_s * V);
  const EllDev A = ell(c);
  FieldBufs f = fbufs(c);
  launch_rhs_real(A, f, c->d_adj_rho.as<double>(), 1.0, c->cfg.load_sign, 0, s);
  const bool fp64 = c->is_fp64();
  pcg_drive<T>(c, f, c->cfg.pcg_tol, fp64 ? 0.0 : c->cfg.inner_tol, fp64 ? 4 : 12, c->inner_hist);
  if (!c->last_conv) throw Error(SCDAT_ERR_SOLVER, "charging adjoint: forward Poisson solve did not converge");
  FloatArgs fa = float_args(c);
  fa.q_init = Q;
  fa.cond_mask = 0;  // the charge is Q itself
  fa.emit_count = nullptr;
  launch_float(f, fa, s);
  launch_phi_efield<T>(A, f, c->d_phiT.as<T>(), c->d_phi64.as<double>(), s1.E.as<Vec4<T>>(), nullptr, s);
  HIPCHK(hipMemcpyAsync(&V, &f.sc->vfloat, sizeof(double), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
}

template <typename T>
double charging_forward(scdat_ctx* c, const scdat_charging_params* p, int steps, int events) {
  const ChargingRun run = adj_begin<T>(c, p, 2);
  AdjArgs<T> a = adj_args<T>(c, run);
  if (events == SCDAT_EVENTS_RECORD) {
    c->d_adj_events.alloc(4 * (size_t)std::max<int64_t>(1, c->adj_np));
    HIPCHK(hipMemsetAsync(c->d_adj_events.p, 0xFF, 4 * (size_t)c->adj_np, c->stream));  // -1: never absorbed
    // each particle's cell after every step (4 B per particle-step)
    c->d_adj_path.alloc(4 * (size_t)std::max<int64_t>(1, c->adj_np) * (size_t)std::max(1, steps));
    c->adj_events_np = c->adj_np;
    c->adj_events_steps = steps;
  } else if (events == SCDAT_EVENTS_REPLAY) {
    if (c->adj_events_np != c->adj_np || c->adj_events_steps < steps)
      throw Error(SCDAT_ERR_STATE, "charging forward: no recorded events for these particles and steps");
  }
  a.event_mode = events;
  a.events = c->d_adj_events.as<int32_t>();
  a.path = c->d_adj_path.as<int32_t>();
  double Q = c->q_init, V = c->h_sc->vfloat;
  {
    PcgScalars q;
    HIPCHK(hipMemcpy(&q, c->d_sc.p, sizeof(q), hipMemcpyDeviceToHost));
    V = q.vfloat;
  }
  for (int n = 0; n < steps; ++n)
    adj_forward_step<T>(c, run, a, c->adj_st[n & 1], c->adj_st[(n + 1) & 1], Q, V, n);
  adj_end(c);
  return V;
}

// Adjoint of one step n (state s0 -> s1, V_n = Vn): in, Ebar of E_{n+1}
// (ebar_in), Vbar of V_{n+1}, Qbar of Q_{n+2}; out, Ebar of E_n (ebar_out),
// Vbar of V_n, Qbar of Q_{n+1}, gradient contributions.
template <typename T>
void adj_backward_step(scdat_ctx* c, const ChargingRun& run, AdjArgs<T> a, AdjStep& s0, AdjStep& s1, double Vn,
                       double* ebar_in, double* ebar_out, double& Vbar, double& Qbar, double* grad) {
  hipStream_t s = c->stream;
  const int64_t nn = c->h.nn;
  const EllDev A = ell(c);
  double* phibar = c->d_adj_phibar.as<double>();
  HIPCHK(hipMemsetAsync(phibar, 0, 8 * (size_t)nn, s));
  launch_adj_field_t<T>(A, ebar_in, phibar, s);
  launch_adj_dot(nn, c->d_phi1.as<double>(), phibar, c->d_adj_dotp.as<double>(), s);
  {
    const int nb = adj_blocks(nn);
    std::vector<double> h((size_t)nb);
    HIPCHK(hipMemcpyAsync(h.data(), c->d_adj_dotp.p, 8 * h.size(), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    for (double v : h) Vbar += v;
  }
  // V = (Q + sum_C rho - kc.phi_0) / C
  const double vc = Vbar / c->ccap;
  Qbar += vc;
  launch_adj_scatter((int64_t)c->h.kc_idx.size(), c->d_kc_idx.as<int32_t>(), c->d_kc_val.as<double>(), -vc, phibar,
                     s);
  // phi_0 = A^-1 b(rho): rhobar = A^-1 phibar_0 on free rows (A symmetric)
  FieldBufs f = fbufs(c);
  f.x64 = c->d_adj_y.as<double>();
  launch_rhs_real(A, f, phibar, 1.0, c->cfg.load_sign, 1, s);
  const bool fp64 = c->is_fp64();
  pcg_drive<T>(c, f, c->cfg.pcg_tol, fp64 ? 0.0 : c->cfg.inner_tol, fp64 ? 4 : 12, nullptr);
  if (!c->last_conv) throw Error(SCDAT_ERR_SOLVER, "charging adjoint: adjoint Poisson solve did not converge");
  launch_adj_rhobar(nn, c->d_kind.as<int32_t>(), c->d_adj_y.as<double>(), vc * (double)c->cfg.load_sign,
                    c->d_adj_rhobar.as<double>(), s);
  // particles: deposit, secondary charge, drift and kick of step n
  HIPCHK(hipMemsetAsync(ebar_out, 0, 24 * (size_t)nn, s));
  a.mesh.E = s0.E.as<Vec4<T>>();
  a.s0 = adj_state<T>(s0);
  a.s1 = adj_state<T>(s1);
  a.xb = c->d_adj_xb.as<double>();
  a.vb = c->d_adj_vb.as<double>();
  a.rhobar = c->d_adj_rhobar.as<double>();
  a.Ebar = ebar_out;
  a.qbar = Qbar;
  launch_adj_backward<T>(a, s);
  HIPCHK(hipGetLastError());
  const auto part = adj_partials(c, adj_blocks(a.n));
  grad[SCDAT_GRAD_DELTA_MAX] += part[0];
  for (int k = 0; k < 6; ++k) grad[SCDAT_GRAD_FACE_YIELD + k] += part[1 + k];
  // Q_{n+1} = Q_n + ... + dt (J(V_n) A_lit - G V_n)
  const double dt = c->cfg.dt;
  double djdv = 0.0;
  const double j = photo_j(run, Vn, &djdv);
  if (run.p.photo_current_density != 0.0) grad[SCDAT_GRAD_PHOTO_J] += Qbar * dt * run.a_lit * j / run.p.photo_current_density;
  else {
    const double t = std::max(1e-3, run.p.photo_temperature_ev);
    grad[SCDAT_GRAD_PHOTO_J] += Qbar * dt * run.a_lit * std::exp(-std::max(0.0, Vn - run.p.photo_reference_v) / t);
  }
  grad[SCDAT_GRAD_CONDUCTANCE] -= Qbar * dt * Vn;
  Vbar = Qbar * dt * (run.a_lit * djdv - run.p.surface_conductance_s);
}

template <typename T>
double charging_adjoint(scdat_ctx* c, const scdat_charging_params* p, int steps, int interval, double* grad) {
  const int K = std::max(1, std::min(interval, steps));
  const ChargingRun run = adj_begin<T>(c, p, K + 1);
  AdjArgs<T> a = adj_args<T>(c, run);
  hipStream_t s = c->stream;
  const int64_t nn = c->h.nn, np = c->adj_np;
  const int nseg = (steps + K - 1) / K;
  double Q = c->q_init, V;
  {
    PcgScalars q;
    HIPCHK(hipMemcpy(&q, c->d_sc.p, sizeof(q), hipMemcpyDeviceToHost));
    V = q.vfloat;
  }
  while ((int)c->adj_ck.size() < nseg) {
    c->adj_ck.emplace_back();
    adj_alloc_step<T>(c, c->adj_ck.back());
    c->adj_ck_x64.emplace_back();
    c->adj_ck_x64.back().alloc(8 * (size_t)nn);
  }
  std::vector<double> ckQ(nseg), ckV(nseg);
  // forward sweep, checkpoint at the start of every segment
  for (int k = 0; k < nseg; ++k) {
    adj_copy_step<T>(c, c->adj_ck[k], c->adj_st[0]);
    HIPCHK(hipMemcpyAsync(c->adj_ck_x64[k].p, c->d_x64.p, 8 * (size_t)nn, hipMemcpyDeviceToDevice, s));
    ckQ[k] = Q;
    ckV[k] = V;
    const int L = std::min(K, steps - k * K);
    for (int j = 0; j < L; ++j) {
      AdjStep& from = c->adj_st[j & 1];
      AdjStep& to = c->adj_st[(j + 1) & 1];
      adj_forward_step<T>(c, run, a, from, to, Q, V, k * K + j);
    }
    if (L & 1) adj_copy_step<T>(c, c->adj_st[0], c->adj_st[1]);
  }
  const double v_final = V;
  // reverse sweep: recompute each segment from its checkpoint, then go back
  c->d_adj_xb.alloc(24 * (size_t)np);
  c->d_adj_vb.alloc(24 * (size_t)np);
  c->d_adj_phibar.alloc(8 * (size_t)nn);
  c->d_adj_rhobar.alloc(8 * (size_t)nn);
  c->d_adj_y.alloc(8 * (size_t)nn);
  for (int k = 0; k < 2; ++k) c->d_adj_Ebar[k].alloc(24 * (size_t)nn);
  HIPCHK(hipMemsetAsync(c->d_adj_xb.p, 0, 24 * (size_t)np, s));
  HIPCHK(hipMemsetAsync(c->d_adj_vb.p, 0, 24 * (size_t)np, s));
  HIPCHK(hipMemsetAsync(c->d_adj_y.p, 0, 8 * (size_t)nn, s));
  HIPCHK(hipMemsetAsync(c->d_adj_Ebar[0].p, 0, 24 * (size_t)nn, s));
  for (int q = 0; q < SCDAT_N_GRAD; ++q) grad[q] = 0.0;
  double Vbar = 1.0, Qbar = 0.0;  // objective: V at the last step
  int eb = 0;                      // d_adj_Ebar[eb]: adjoint of E_{n+1}
  std::vector<double> Vs((size_t)K + 1);
  for (int k = nseg - 1; k >= 0; --k) {
    const int L = std::min(K, steps - k * K);
    adj_copy_step<T>(c, c->adj_st[0], c->adj_ck[k]);
    HIPCHK(hipMemcpyAsync(c->d_x64.p, c->adj_ck_x64[k].p, 8 * (size_t)nn, hipMemcpyDeviceToDevice, s));
    Q = ckQ[k];
    V = ckV[k];
    Vs[0] = V;
    for (int j = 0; j < L; ++j) {
      adj_forward_step<T>(c, run, a, c->adj_st[j], c->adj_st[j + 1], Q, V, k * K + j);
      Vs[j + 1] = V;
    }
    for (int j = L - 1; j >= 0; --j) {
      adj_backward_step<T>(c, run, a, c->adj_st[j], c->adj_st[j + 1], Vs[j], c->d_adj_Ebar[eb].as<double>(),
                           c->d_adj_Ebar[1 - eb].as<double>(), Vbar, Qbar, grad);
      eb = 1 - eb;
    }
  }
  adj_end(c);
  return v_final;
}

// Ledger entries of scenarios [s0, s1) summed, PICCycle::getSurfaceCurrentLedger
// (PICCycle.cpp:1295-1329) fields.
void fill_ledger(scdat_ctx* c, const LedgerHost& L, int s0, int s1, scdat_ledger_entry* out, int n_groups) {
  for (int g = 0; g < n_groups; ++g) std::memset(&out[g], 0, sizeof(out[g]));
  for (int sg = s0; sg < s1; ++sg) {
    const size_t lo = (size_t)sg * kMaxGroups * kMaxSpecies;
    const double ws = seg_scale(c, sg);
    for (int g = 0; g < n_groups; ++g) {
      scdat_ledger_entry& E = out[g];
      for (size_t sp = 0; sp < c->sp.size(); ++sp) {
        const unsigned long long k = L.emi[lo + g * kMaxSpecies + sp];
        E.emitted_electron_particles += (int64_t)k;
        E.emitted_electron_charge_c += (double)k * c->sp[sp].q * (c->sp[sp].w * ws);
      }
      // escaped: not recorded; caller-side conditions record their own
      if (c->gtype[g] == SCDAT_BC_ESCAPE || c->gtype[g] == SCDAT_BC_HOST) continue;
      for (size_t sp = 0; sp < c->sp.size(); ++sp) {
        const unsigned long long k = L.cnt[lo + g * kMaxSpecies + sp];
        const double q = (double)k * c->sp[sp].q * (c->sp[sp].w * ws);
        E.incident_energy_j += L.en[lo + g * kMaxSpecies + sp];
        // electron-like = negative charge
        if (c->sp[sp].q < 0) {
          E.absorbed_electron_charge_c += q;
          E.absorbed_electron_particles += (int64_t)k;
        } else {
          E.absorbed_ion_charge_c += q;
          E.absorbed_ion_particles += (int64_t)k;
        }
      }
    }
  }
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void scdat_config_default(scdat_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->precision = 0;
  c->dt = 1e-12;
  c->adaptive_substep = 1;
  c->floor_ratio = 1e-3;
  c->max_substeps = 64;
  c->poisson_operator = 0;
  c->permittivity = 8.8541878128e-12;
  c->load_sign = 1;
  c->pcg_tol = 1e-8;
  c->pcg_max_iter = 20000;
  c->inner_tol = 1e-5;  // fewest iterations in total at 64^3 (1e-4..1e-5; 1e-6: +7 %)
  c->sort_interval = 20;
}

int scdat_box_mesh_size(const scdat_box_spec* spec, int64_t* n_nodes, int64_t* n_tets) {
  if (!spec || !n_nodes || !n_tets) return SCDAT_ERR_ARG;
  BoxSpec s;
  std::memcpy(&s, spec, sizeof(s));
  std::vector<double> xyz;
  std::vector<int32_t> tets, kind, group;
  std::string err;
  try {
    if (!box_mesh(s, xyz, tets, kind, group, err)) return SCDAT_ERR_ARG;
  } catch (...) {
    return SCDAT_ERR_NOMEM;
  }
  *n_nodes = (int64_t)kind.size();
  *n_tets = (int64_t)(tets.size() / 4);
  return SCDAT_OK;
}

int scdat_box_mesh_build(const scdat_box_spec* spec, double* xyz, int32_t* tets, int32_t* node_kind,
                         int32_t* face_group) {
  if (!spec || !xyz || !tets || !node_kind) return SCDAT_ERR_ARG;
  BoxSpec s;
  std::memcpy(&s, spec, sizeof(s));
  std::vector<double> vx;
  std::vector<int32_t> vt, vk, vg;
  std::string err;
  try {
    if (!box_mesh(s, vx, vt, vk, vg, err)) return SCDAT_ERR_ARG;
  } catch (...) {
    return SCDAT_ERR_NOMEM;
  }
  std::memcpy(xyz, vx.data(), sizeof(double) * vx.size());
  std::memcpy(tets, vt.data(), sizeof(int32_t) * vt.size());
  std::memcpy(node_kind, vk.data(), sizeof(int32_t) * vk.size());
  if (face_group) std::memcpy(face_group, vg.data(), sizeof(int32_t) * vg.size());
  return SCDAT_OK;
}

int scdat_device_info(char* name, int32_t name_len, int32_t* n_cu) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) return SCDAT_ERR_DEVICE;
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, 0) != hipSuccess) return SCDAT_ERR_DEVICE;
  if (name && name_len > 0) {
    std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
  }
  if (n_cu) *n_cu = p.multiProcessorCount;
  return SCDAT_OK;
}

int scdat_create(const scdat_config* cfg, scdat_ctx** out) {
  if (!out) return SCDAT_ERR_ARG;
  *out = nullptr;
  scdat_ctx* c = new (std::nothrow) scdat_ctx();
  if (!c) return SCDAT_ERR_NOMEM;
  if (cfg) c->cfg = *cfg;
  else scdat_config_default(&c->cfg);
  if (c->cfg.precision != 0 && c->cfg.precision != 1) {
    delete c;
    return SCDAT_ERR_ARG;
  }
  if (!(c->cfg.dt > 0) || c->cfg.max_substeps < 1 || c->cfg.pcg_max_iter < 1) {
    delete c;
    return SCDAT_ERR_ARG;
  }
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n < 1) {
    delete c;
    return SCDAT_ERR_DEVICE;
  }
  if (hipGetDevice(&c->device) != hipSuccess || !(c->d_stats = stats_scratch(c->device))) {
    delete c;
    return SCDAT_ERR_DEVICE;
  }
  c->err.reserve(512);  // messages fit without an allocation (the no-alloc entry point)
  if (hipStreamCreate(&c->stream) != hipSuccess) {
    delete c;
    return SCDAT_ERR_DEVICE;
  }
  if (hipHostMalloc((void**)&c->h_sc, sizeof(PcgScalars), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_append, 8 * (kMaxSpecies + 1), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_patch_v, 8 * kMaxPatches, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ledger, 16 * kMaxGroups * kMaxSpecies, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->h_stats, sizeof(scratch), hipHostMallocDefault) != hipSuccess) {
    if (c->h_ledger) (void)hipHostFree(c->h_ledger);
    if (c->h_append) (void)hipHostFree(c->h_append);
    if (c->h_patch_v) (void)hipHostFree(c->h_patch_v);
    if (c->h_sc) (void)hipHostFree(c->h_sc);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return SCDAT_ERR_DEVICE;
  }
  for (int g = 0; g < kMaxGroups; ++g) {
    c->gtype[g] = SCDAT_BC_ESCAPE;
    c->gpartner[g] = -1;
  }
  *out = c;
  return SCDAT_OK;
}

void scdat_destroy(scdat_ctx* c) {
  if (!c) return;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->drop_graphs();
  for (auto e : c->ev_pool) (void)hipEventDestroy(e);
  for (auto e : c->kt_pool) (void)hipEventDestroy(e);
  if (c->h_sc) (void)hipHostFree(c->h_sc);
  if (c->h_append) (void)hipHostFree(c->h_append);
  if (c->h_patch_v) (void)hipHostFree(c->h_patch_v);
  if (c->h_ledger) (void)hipHostFree(c->h_ledger);
  if (c->h_stats) (void)hipHostFree(c->h_stats);
  hipStream_t s = c->stream;
  delete c;  // DevBuf destructors free device memory
  if (s) (void)hipStreamDestroy(s);
}

const char* scdat_last_error(const scdat_ctx* c) { return c ? c->err.c_str() : "null context"; }

int scdat_set_mesh(scdat_ctx* c, int64_t n_nodes, const double* xyz, int64_t n_tets,
                   const int32_t* tets, const int32_t* node_kind, const double* dirichlet,
                   const int32_t* face_group) {
  if (!c) return SCDAT_ERR_ARG;
  if (n_nodes < 4 || n_tets < 1 || !xyz || !tets || !node_kind)
    return fail(c, SCDAT_ERR_ARG, "set_mesh: empty mesh or null array");
  if (n_nodes > INT32_MAX || n_tets > INT32_MAX) return fail(c, SCDAT_ERR_ARG, "set_mesh: mesh too large");
  return guarded(c, [&] {
    for (int64_t k = 0; k < 4 * n_tets; ++k)
      if (tets[k] < 0 || tets[k] >= n_nodes) throw Error(SCDAT_ERR_ARG, "set_mesh: tet references a missing node");
    for (int64_t i = 0; i < n_nodes; ++i)
      if (node_kind[i] < 0 || node_kind[i] > 2) throw Error(SCDAT_ERR_ARG, "set_mesh: node kind must be 0, 1 or 2");
    std::vector<double> dv(dirichlet ? std::vector<double>(dirichlet, dirichlet + n_nodes)
                                     : std::vector<double>((size_t)n_nodes, 0.0));
    std::vector<int32_t> fg(face_group ? std::vector<int32_t>(face_group, face_group + 4 * n_tets)
                                       : std::vector<int32_t>(4 * (size_t)n_tets, 0));
    c->seg_node0.clear();
    c->seg_emit.clear();
    c->fe_mask = 0;
    c->fe_pending.fill(0.0);
    c->fe_pending_sec.fill(0.0);
    mesh_setup(c, n_nodes, xyz, n_tets, tets, node_kind, dv.data(), fg.data(), c->node_order);
  });
}

int scdat_set_boundary(scdat_ctx* c, int32_t group, int32_t type, int32_t partner_group,
                       const double* shift) {
  if (!c) return SCDAT_ERR_ARG;
  if (group < 0 || group >= kMaxGroups || type < 0 || type > SCDAT_BC_HOST)
    return fail(c, SCDAT_ERR_ARG, "set_boundary: bad group or type");
  c->gtype[group] = type;
  c->gpartner[group] = partner_group;
  c->fe_mask &= ~(1u << group);  // a new condition replaces field emission
  for (int r = 0; r < 3; ++r) c->gshift[group][r] = shift ? shift[r] : 0.0;
  c->initialized = false;
  return SCDAT_OK;
}

int scdat_get_host_impacts(scdat_ctx* c, int64_t max, scdat_impact* out, int64_t* n) {
  if (!c || !n || max < 0 || (max > 0 && !out)) return SCDAT_ERR_ARG;
  if (!c->initialized) return fail(c, SCDAT_ERR_STATE, "get_host_impacts: not initialized");
  return guarded(c, [&] {
    *n = 0;
    if (c->event_cap <= 0) return;
    unsigned long long ne = 0;
    HIPCHK(hipMemcpyAsync(&ne, c->d_n_events.p, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const int64_t m = std::min<int64_t>((int64_t)ne, c->event_cap);
    std::vector<SeeEvent> ev((size_t)m);
    if (m) HIPCHK(hipMemcpyAsync(ev.data(), c->d_events.p, sizeof(SeeEvent) * (size_t)m, hipMemcpyDeviceToHost,
                                 c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    // in (species, id) order: the order does not depend on the push's schedule
    std::vector<const SeeEvent*> hits;
    for (const auto& e : ev)
      if (c->gtype[e.group_face & 0xFF] == SCDAT_BC_HOST) hits.push_back(&e);
    std::sort(hits.begin(), hits.end(), [](const SeeEvent* a, const SeeEvent* b) {
      return a->species != b->species ? a->species < b->species : a->id < b->id;
    });
    int64_t k = 0;
    for (const SeeEvent* e : hits) {
      if (k < max) {
        scdat_impact& o = out[k];
        o.species = e->species;
        o.id = e->id;
        o.group = e->group_face & 0xFF;
        o.cell = e->cell;
        for (int r = 0; r < 3; ++r) {
          o.x[r] = e->x[r];
          o.v[r] = e->v[r];
        }
      }
      ++k;
    }
    *n = k;
    if ((int64_t)ne > c->event_cap)
      throw Error(SCDAT_ERR_NOMEM, "get_host_impacts: more impacts in the step than the impact buffer holds");
  });
}

int scdat_get_escaped(scdat_ctx* c, int32_t species, int64_t* n) {
  if (!c || !n) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "bad species");
  return guarded(c, [&] {
    const LedgerHost L = read_ledger(c);
    int64_t e = 0;
    for (int sg = 0; sg < c->n_seg(); ++sg)
      for (int g = 0; g < kMaxGroups; ++g)
        if (c->gtype[g] == SCDAT_BC_ESCAPE) e += (int64_t)L.cnt[((size_t)sg * kMaxGroups + g) * kMaxSpecies + species];
    *n = e;
  });
}

int scdat_add_species(scdat_ctx* c, double q, double m, double w, int64_t capacity, int32_t* id) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "add_species: set the mesh first");
  if (!(m > 0) || !(w > 0) || capacity < 0 || c->sp.size() >= (size_t)kMaxSpecies)
    return fail(c, SCDAT_ERR_ARG, "add_species: bad mass/weight/capacity or too many species");
  if (capacity > INT32_MAX) return fail(c, SCDAT_ERR_ARG, "add_species: capacity above 2^31");
  return guarded(c, [&] {
    if (c->sp.empty()) {
      c->w_ref = w;
      // rho_fx unit: one elementary charge of reference weight is 2^32 units
      // (2^8 charge resolution x 2^24 barycentric resolution).
      c->unit = kElem * w / 4294967296.0;
    }
    c->sp.emplace_back();
    SpeciesState& S = c->sp.back();
    S.q = q;
    S.m = m;
    S.w = w;
    S.cap = capacity;
    // the sort's permutation and staging buffers cover the largest species
    const size_t cap1 = (size_t)std::max<int64_t>(1, capacity);
    if (c->d_perm.bytes < 4 * cap1) c->d_perm.alloc(4 * cap1);
    const size_t rec = c->is_fp64() ? sizeof(SortRec<double>) : sizeof(SortRec<float>);
    if (c->d_stage.bytes < rec * cap1) c->d_stage.alloc(rec * cap1);
    const double csd = q * w / (kElem * c->w_ref) * 256.0;
    S.cs = (int64_t)std::llround(csd);
    if (std::abs(csd - (double)S.cs) > 1e-6 * std::max(1.0, std::abs(csd)))
      throw Error(SCDAT_ERR_ARG, "add_species: q*w must be a multiple of e*w_ref/256");
    const size_t ts = tsize(c);
    for (int b = 0; b < 2; ++b) {
      for (int r = 0; r < 6; ++r) S.buf[b][r].alloc(ts * (size_t)capacity);
      S.buf[b][6].alloc(4 * (size_t)capacity);
      S.buf[b][7].alloc(4 * (size_t)capacity);
    }
    const size_t nb = (size_t)((capacity + kBlock - 1) / kBlock) + 1;
    S.ke_partial.alloc(8 * nb);
    S.alive_partial.alloc(4 * nb);
    HIPCHK(hipMemsetAsync(S.ke_partial.p, 0, 8 * nb, c->stream));
    HIPCHK(hipMemsetAsync(S.alive_partial.p, 0, 4 * nb, c->stream));
    if (id) *id = (int32_t)c->sp.size() - 1;
  });
}

int scdat_set_particles(scdat_ctx* c, int32_t species, int64_t n, const double* x, const double* v,
                        const int32_t* cell) {
  if (!c) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "set_particles: bad species");
  if (n < 0 || (n > 0 && (!x || !v))) return fail(c, SCDAT_ERR_ARG, "set_particles: bad arrays");
  c->mcc_v2_stale = true;
  return guarded(c, [&] {
    if (c->is_fp64()) set_particles_t<double>(c, species, n, x, v, cell);
    else set_particles_t<float>(c, species, n, x, v, cell);
  });
}

int scdat_get_particle_count(scdat_ctx* c, int32_t species, int64_t* n_alive, int64_t* n_slots) {
  if (!c) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "bad species");
  return guarded(c, [&] {
    SpeciesState& S = c->sp[species];
    if (n_slots) *n_slots = S.n;
    if (n_alive) {
      std::vector<int32_t> cl(S.n);
      if (S.n) {
        HIPCHK(hipMemcpyAsync(cl.data(), S.buf[S.cur][6].p, 4 * S.n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
      int64_t a = 0;
      for (int32_t v : cl) a += v >= 0;
      *n_alive = a;
    }
  });
}

int scdat_get_particles(scdat_ctx* c, int32_t species, double* x, double* v, int32_t* cell,
                        int32_t* id) {
  if (!c) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "bad species");
  return guarded(c, [&] {
    if (c->is_fp64()) get_particles_t<double>(c, species, x, v, cell, id);
    else get_particles_t<float>(c, species, x, v, cell, id);
  });
}

int scdat_set_time_step(scdat_ctx* c, double dt) {
  if (!c) return SCDAT_ERR_ARG;
  if (dt > 0.0 && std::isfinite(dt)) c->cfg.dt = dt;
  return SCDAT_OK;
}

int scdat_set_pcg_controls(scdat_ctx* c, double tol, int32_t max_iter) {
  if (!c) return SCDAT_ERR_ARG;
  if (!(tol > 0.0) || !std::isfinite(tol)) return fail(c, SCDAT_ERR_ARG, "set_pcg_controls: tolerance must be > 0");
  c->cfg.pcg_tol = tol;
  if (max_iter > 0) c->cfg.pcg_max_iter = max_iter;
  return SCDAT_OK;
}

int scdat_set_conductor_charge(scdat_ctx* c, double q) {
  if (!c) return SCDAT_ERR_ARG;
  c->q_init = q;
  return SCDAT_OK;
}

int scdat_initialize(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "initialize: no mesh");
  c->batch_failed = 0;
  const int rc = guarded(c, [&] {
    if (c->is_fp64()) do_initialize<double>(c);
    else do_initialize<float>(c);
  });
  if (rc == SCDAT_OK && c->batch_failed > 0) {
    c->err = "initial Poisson solve did not reach the tolerance; the initial potential was kept";
    return SCDAT_ERR_SOLVER;
  }
  return rc;
}

int scdat_step(scdat_ctx* c, int32_t n) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->initialized) return fail(c, SCDAT_ERR_STATE, "step: call scdat_initialize first");
  if (n < 0) return fail(c, SCDAT_ERR_ARG, "step: negative count");
  const int rc = guarded(c, [&] {
    const bool graph = c->step_graph && graph_capable(c) && n > 0;
    if (c->is_fp64()) {
      if (graph) do_step_graph<double>(c, n);
      else do_step<double>(c, n);
    } else {
      if (graph) do_step_graph<float>(c, n);
      else do_step<float>(c, n);
    }
  });
  if (rc == SCDAT_OK && c->batch_failed > 0) {
    c->err = std::to_string(c->batch_failed) + " of " + std::to_string(n) +
             " Poisson solves did not reach the tolerance (first in step " +
             std::to_string(c->first_failed_step + 1) +
             "); those steps kept the previous potential and field, and every step ran";
    return SCDAT_ERR_SOLVER;
  }
  return rc;
}

int scdat_push(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->initialized) return fail(c, SCDAT_ERR_STATE, "push: call scdat_initialize first");
  return guarded(c, [&] {
    if (c->is_fp64()) do_push<double>(c, nullptr, nullptr);
    else do_push<float>(c, nullptr, nullptr);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int scdat_deposit(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "deposit: no mesh");
  return guarded(c, [&] {
    const int k = c->rho_target;
    HIPCHK(hipMemsetAsync(c->d_rho[k].p, 0, 8 * (size_t)c->h.nn, c->stream));
    for (size_t si = 0; si < c->sp.size(); ++si) {
      if (c->sp[si].n == 0) continue;
      if (c->is_fp64()) launch_deposit<double>(push_args<double>(c, (int)si), c->stream);
      else launch_deposit<float>(push_args<float>(c, (int)si), c->stream);
    }
    c->rho_clean[k] = false;
    c->rho_last = k;
    c->rho_target = 1 - k;
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int scdat_solve(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "solve: no mesh");
  c->batch_failed = 0;
  const int rc = guarded(c, [&] {
    if (c->h.has_cond && !(c->ccap > 0)) {
      if (c->is_fp64()) solve_phi1<double>(c);
      else solve_phi1<float>(c);
    }
    if (c->is_fp64()) {
      do_solve<double>(c, c->rho_last);
      do_field<double>(c, nullptr);
    } else {
      do_solve<float>(c, c->rho_last);
      do_field<float>(c, nullptr);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
  if (rc == SCDAT_OK && c->batch_failed > 0) {
    c->err = "Poisson PCG did not reach the tolerance; the previous potential and field were kept";
    return SCDAT_ERR_SOLVER;
  }
  return rc;
}

int scdat_compute_efield(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "compute_efield: no mesh");
  return guarded(c, [&] {
    const size_t nn = (size_t)c->h.nn;
    std::vector<double> p(nn);
    HIPCHK(hipMemcpyAsync(p.data(), c->d_phi64.p, 8 * nn, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->is_fp64()) {
      HIPCHK(hipMemcpyAsync(c->d_phiT.p, p.data(), 8 * nn, hipMemcpyHostToDevice, c->stream));
      launch_efield_only<double>(ell(c), c->d_phiT.as<double>(), c->d_E.as<Vec4<double>>(), c->stream);
    } else {
      std::vector<float> pf(nn);
      for (size_t i = 0; i < nn; ++i) pf[i] = (float)p[i];
      HIPCHK(hipMemcpyAsync(c->d_phiT.p, pf.data(), 4 * nn, hipMemcpyHostToDevice, c->stream));
      launch_efield_only<float>(ell(c), c->d_phiT.as<float>(), c->d_E.as<Vec4<float>>(), c->stream);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int scdat_sort(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "sort: no mesh");
  return guarded(c, [&] {
    if (c->is_fp64()) do_sort<double>(c);
    else do_sort<float>(c);
  });
}

int scdat_set_efield(scdat_ctx* c, const double* E) {
  if (!c || !E) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_efield: no mesh");
  return guarded(c, [&] {
    const size_t nn = (size_t)c->h.nn;
    if (c->is_fp64()) {
      std::vector<Vec4<double>> e(nn);
      for (size_t i = 0; i < nn; ++i) {
        const int64_t u = to_user(c, (int64_t)i);
        e[i] = {E[3 * u], E[3 * u + 1], E[3 * u + 2], 0.0};
      }
      HIPCHK(hipMemcpyAsync(c->d_E.p, e.data(), sizeof(Vec4<double>) * nn, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    } else {
      std::vector<Vec4<float>> e(nn);
      for (size_t i = 0; i < nn; ++i) {
        const int64_t u = to_user(c, (int64_t)i);
        e[i] = {(float)E[3 * u], (float)E[3 * u + 1], (float)E[3 * u + 2], 0.f};
      }
      HIPCHK(hipMemcpyAsync(c->d_E.p, e.data(), sizeof(Vec4<float>) * nn, hipMemcpyHostToDevice, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
    }
  });
}

int scdat_set_potential(scdat_ctx* c, const double* phi) {
  if (!c || !phi) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_potential: no mesh");
  const int rc = guarded(c, [&] {
    const size_t nn = (size_t)c->h.nn;
    std::vector<double> p(nn);
    for (size_t i = 0; i < nn; ++i) p[i] = phi[to_user(c, (int64_t)i)];
    HIPCHK(hipMemcpyAsync(c->d_phi64.p, p.data(), 8 * nn, hipMemcpyHostToDevice, c->stream));
    // The next solve holds Dirichlet nodes at these values (PICCycle.cpp:
    // 674-688 takes potential_[i] as the value of a DIRICHLET node): update
    // the values and the lifting of the load on their free neighbours.
    HostMesh& h = c->h;
    bool changed = false;
    for (size_t i = 0; i < nn; ++i)
      if (h.kind[i] == SCDAT_NODE_DIRICHLET && h.dval[i] != p[i]) {
        h.dval[i] = p[i];
        changed = true;
      }
    if (changed) {
      const int W = h.W;
      for (size_t i = 0; i < nn; ++i) {
        h.lift[i] = 0.0;
        if (h.kind[i] != SCDAT_NODE_FREE) continue;
        for (int sl = 1; sl < W; ++sl) {
          const size_t o = (size_t)sl * nn + i;
          const int32_t j = h.col[o];
          if (h.kind[j] == SCDAT_NODE_DIRICHLET) h.lift[i] -= h.kfull[o] * h.dval[j];
        }
      }
      c->d_lift.upload(h.lift, c->stream);
      c->d_dval.upload(h.dval, c->stream);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
  });
  return rc != SCDAT_OK ? rc : scdat_compute_efield(c);
}

int scdat_set_rho(scdat_ctx* c, const double* rho) {
  if (!c || !rho) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_rho: no mesh");
  return guarded(c, [&] {
    const size_t nn = (size_t)c->h.nn;
    if (c->sp.empty() && !c->segmented()) {
      // no species to fix the unit: 2^-60 of the largest |rho| of this call,
      // a power of two (a field solve without particles)
      double mx = 0.0;
      for (size_t i = 0; i < nn; ++i) {
        if (!std::isfinite(rho[i])) throw Error(SCDAT_ERR_ARG, "set_rho: non-finite value");
        mx = std::max(mx, std::abs(rho[i]));
      }
      int e = 0;
      std::frexp(mx, &e);
      c->unit = mx > 0.0 ? std::ldexp(1.0, e - 60) : 1.0;
      c->w_ref = c->unit * 4294967296.0 / kElem;
    }
    std::vector<long long> fx(nn);
    for (size_t i = 0; i < nn; ++i) fx[i] = std::llround(rho[to_user(c, (int64_t)i)] / node_unit(c, (int64_t)i));
    const int k = c->rho_target;
    HIPCHK(hipMemcpyAsync(c->d_rho[k].p, fx.data(), 8 * nn, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->rho_clean[k] = false;
    c->rho_last = k;
    c->rho_target = 1 - k;
  });
}

int scdat_get_fields(scdat_ctx* c, double* rho, double* phi, double* E) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_fields: no mesh");
  return guarded(c, [&] {
    const size_t nn = (size_t)c->h.nn;
    if (rho) {
      std::vector<long long> fx(nn);
      HIPCHK(hipMemcpyAsync(fx.data(), c->d_rho[c->rho_last].p, 8 * nn, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < nn; ++i) rho[to_user(c, (int64_t)i)] = (double)fx[i] * node_unit(c, (int64_t)i);
    }
    if (phi) {
      std::vector<double> p(nn);
      HIPCHK(hipMemcpyAsync(p.data(), c->d_phi64.p, 8 * nn, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (size_t i = 0; i < nn; ++i) phi[to_user(c, (int64_t)i)] = p[i];
    }
    if (E) {
      if (c->is_fp64()) {
        std::vector<Vec4<double>> e(nn);
        HIPCHK(hipMemcpyAsync(e.data(), c->d_E.p, sizeof(Vec4<double>) * nn, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < nn; ++i) {
          const int64_t u = to_user(c, (int64_t)i);
          E[3 * u] = e[i].x;
          E[3 * u + 1] = e[i].y;
          E[3 * u + 2] = e[i].z;
        }
      } else {
        std::vector<Vec4<float>> e(nn);
        HIPCHK(hipMemcpyAsync(e.data(), c->d_E.p, sizeof(Vec4<float>) * nn, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t i = 0; i < nn; ++i) {
          const int64_t u = to_user(c, (int64_t)i);
          E[3 * u] = e[i].x;
          E[3 * u + 1] = e[i].y;
          E[3 * u + 2] = e[i].z;
        }
      }
    }
  });
}

int scdat_get_rho_fixed(scdat_ctx* c, int64_t* rho_fx, double* unit) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_rho_fixed: no mesh");
  return guarded(c, [&] {
    if (unit) *unit = c->unit;
    if (rho_fx) {
      std::vector<int64_t> fx((size_t)c->h.nn);
      HIPCHK(hipMemcpyAsync(fx.data(), c->d_rho[c->rho_last].p, 8 * fx.size(), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (int64_t i = 0; i < c->h.nn; ++i) rho_fx[to_user(c, i)] = fx[i];
    }
  });
}

int scdat_get_tet_maps(scdat_ctx* c, float* maps) {
  if (!c || !maps) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_tet_maps: no mesh");
  const HostMesh& h = c->h;
  for (int64_t t = 0; t < h.nt; ++t) {
    for (int k = 0; k < 9; ++k) maps[12 * t + k] = (float)h.A[9 * t + k];
    for (int k = 0; k < 3; ++k) maps[12 * t + 9 + k] = (float)h.X0[3 * t + k];
  }
  return SCDAT_OK;
}

int scdat_get_species_charge_fixed(scdat_ctx* c, int32_t species, int64_t* cs) {
  if (!c || !cs) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "bad species");
  *cs = c->sp[species].cs;
  return SCDAT_OK;
}

int scdat_get_stats(scdat_ctx* c, scdat_stats* st) {
  if (!c || !st) return SCDAT_ERR_ARG;
  return guarded(c, [&] {
    std::memset(st, 0, sizeof(*st));
    st->completed_steps = c->steps;
    st->simulation_time = c->sim_time;
    double ke = 0;
    int64_t alive = 0;
    for (auto& S : c->sp) {
      const size_t nb = (size_t)S.stat_blocks;
      if (!nb) continue;
      std::vector<double> k(nb);
      std::vector<int32_t> a(nb);
      HIPCHK(hipMemcpyAsync(k.data(), S.ke_partial.p, 8 * nb, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(a.data(), S.alive_partial.p, 4 * nb, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      for (size_t b = 0; b < nb; ++b) {
        ke += k[b];
        alive += a[b];
      }
    }
    st->total_energy = ke;
    st->active_particles = alive;
    if (c->has_mesh) {
      const size_t nn = (size_t)c->h.nn;
      std::vector<long long> fx(nn);
      HIPCHK(hipMemcpyAsync(fx.data(), c->d_rho[c->rho_last].p, 8 * nn, hipMemcpyDeviceToHost, c->stream));
      const LedgerHost L = read_ledger(c);
      unsigned long long lost = 0, pushes = 0;
      HIPCHK(hipMemcpyAsync(&lost, c->d_lost.p, 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(&pushes, c->d_pushes.p, 8, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipMemcpyAsync(c->h_sc, c->d_sc.p, sizeof(PcgScalars), hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      long double tq = 0;
      for (int sg = 0; sg < c->n_seg(); ++sg) {
        const int64_t i0 = c->segmented() ? c->seg_node0[sg] : 0;
        const int64_t i1 = c->segmented() ? c->seg_node0[sg + 1] : (int64_t)nn;
        long double t = 0;
        for (int64_t i = i0; i < i1; ++i) t += (long double)fx[i];
        tq += t * (long double)(c->unit * seg_scale(c, sg));
      }
      st->total_charge = (double)tq;
      // surface charge of every scenario; conductor charge of the first
      double surf = 0, qc = c->q_init;
      for (int sg = 0; sg < c->n_seg(); ++sg)
        for (int g = 0; g < kMaxGroups; ++g)
          for (size_t s = 0; s < c->sp.size(); ++s) {
            if (c->gtype[g] == SCDAT_BC_ESCAPE || c->gtype[g] == SCDAT_BC_HOST) continue;
            const size_t o = ((size_t)sg * kMaxGroups + g) * kMaxSpecies + s;
            const double qw = c->sp[s].q * (c->sp[s].w * seg_scale(c, sg));
            const double q = (double)L.cnt[o] * qw;
            surf += q;
            if (sg == 0 && c->gtype[g] == SCDAT_BC_CONDUCTOR) qc += q - (double)L.emi[o] * qw;
          }
      st->surface_deposited_charge = surf;
      st->conductor_charge = qc;
      st->floating_potential = c->h.has_cond ? c->h_sc->vfloat : 0.0;
      st->lost_particles = (int64_t)lost;
      st->particle_pushes = (int64_t)pushes;
    }
    st->pcg_iterations = c->last_iters;
    st->pcg_outer = c->last_outer;
    st->pcg_rel_residual = c->last_rel;
    st->pcg_converged = c->last_conv;
    st->pcg_initial_rel_residual = c->last_init_rel;
    st->pcg_iterations_total = c->iters_total;
    st->pcg_failed_solves = c->failed_solves;
  });
}

// The statistics of scdat_get_stats with no allocation: the device sums go
// to this GPU's static scratch (one kernel, every result word written),
// the host reads them back into the context's pinned mirror (allocated at
// scdat_create) and fills *st.  Same values bit for bit.  No hipMalloc,
// hipFree or host allocation: error messages fit the context's reserved
// buffer.  At most kStatsMaxSegments scenarios.  Calls that overlap on one
// GPU (two host threads, two contexts) race on the scratch.
int scdat_get_stats_noalloc(scdat_ctx* c, scdat_stats* st) {
  if (!c || !st) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_stats_noalloc: no mesh");
  if (c->n_seg() > kStatsMaxSegments) return fail(c, SCDAT_ERR_ARG, "get_stats_noalloc: more than 4096 scenarios");
  if (c->sp.size() > (size_t)kMaxSpecies) return fail(c, SCDAT_ERR_ARG, "get_stats_noalloc: too many species");
  StatsArgs a{};
  a.ns = (int)c->sp.size();
  for (int s = 0; s < a.ns; ++s) {
    const SpeciesState& S = c->sp[s];
    a.ke[s] = S.ke_partial.as<double>();
    a.alive[s] = S.alive_partial.as<int32_t>();
    a.nb[s] = S.stat_blocks;
    a.q[s] = S.q;
    a.w[s] = S.w;
  }
  a.rho = c->d_rho[c->rho_last].as<unsigned long long>();
  a.nn = c->h.nn;
  a.n_seg = c->n_seg();
  if (c->segmented()) {
    a.seg_b0 = c->d_seg_b0.as<int32_t>();
    a.seg_wscale = c->d_seg_wscale.as<double>();
  }
  a.led_cnt = c->d_ledger_count.as<unsigned long long>();
  a.led_emi = c->d_emit_count.as<unsigned long long>();
  for (int g = 0; g < kMaxGroups; ++g) {
    if (c->gtype[g] != SCDAT_BC_ESCAPE && c->gtype[g] != SCDAT_BC_HOST) a.surf_mask |= 1u << g;
    if (c->gtype[g] == SCDAT_BC_CONDUCTOR) a.cond_mask |= 1u << g;
  }
  a.q_init = c->q_init;
  a.lost = c->d_lost.as<unsigned long long>();
  a.pushes = c->d_pushes.as<unsigned long long>();
  a.sc = c->d_sc.as<PcgScalars>();
  a.out = c->d_stats;
  hipLaunchKernelGGL(k_stats_noalloc, dim3(1 + a.n_seg), dim3(kBlock), 0, c->stream, a);
  const size_t words = (size_t)kStatsScratchHead + (size_t)a.n_seg;
  if (hipGetLastError() != hipSuccess ||
      hipMemcpyAsync(c->h_stats, c->d_stats, 8 * words, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return fail(c, SCDAT_ERR_DEVICE, "get_stats_noalloc: device error");
  const unsigned long long* h = c->h_stats;
  auto as_double = [](unsigned long long u) {
    double d;
    std::memcpy(&d, &u, sizeof d);
    return d;
  };
  std::memset(st, 0, sizeof(*st));
  st->completed_steps = c->steps;
  st->simulation_time = c->sim_time;
  st->total_energy = as_double(h[0]);
  st->active_particles = (int64_t)h[1];
  long double tq = 0;
  for (int sg = 0; sg < a.n_seg; ++sg)
    tq += (long double)(long long)h[kStatsScratchHead + sg] * (long double)(c->unit * seg_scale(c, sg));
  st->total_charge = (double)tq;
  st->surface_deposited_charge = as_double(h[2]);
  st->conductor_charge = as_double(h[3]);
  st->floating_potential = c->h.has_cond ? as_double(h[6]) : 0.0;
  st->lost_particles = (int64_t)h[4];
  st->particle_pushes = (int64_t)h[5];
  st->pcg_iterations = c->last_iters;
  st->pcg_outer = c->last_outer;
  st->pcg_rel_residual = c->last_rel;
  st->pcg_converged = c->last_conv;
  st->pcg_initial_rel_residual = c->last_init_rel;
  st->pcg_iterations_total = c->iters_total;
  st->pcg_failed_solves = c->failed_solves;
  return SCDAT_OK;
}

int scdat_get_ledger(scdat_ctx* c, scdat_ledger_entry* out, int32_t n_groups) {
  if (!c || !out || n_groups < 0 || n_groups > kMaxGroups) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_ledger: no mesh");
  return guarded(c, [&] { fill_ledger(c, read_ledger(c), 0, c->n_seg(), out, n_groups); });
}

// ---- background collisions and surface emission -----------------------------

int scdat_set_seed(scdat_ctx* c, uint64_t seed) {
  if (!c) return SCDAT_ERR_ARG;
  c->key[0] = (uint32_t)seed;
  c->key[1] = (uint32_t)(seed >> 32);
  return SCDAT_OK;
}

int scdat_mcc_process_default(int32_t type, scdat_mcc_process* p) {
  if (!p || type < SCDAT_MCC_ELASTIC || type > SCDAT_MCC_COULOMB) return SCDAT_ERR_ARG;
  std::memset(p, 0, sizeof(*p));
  p->type = type;
  p->product_electron_species = -1;
  p->product_ion_species = -1;
  p->reference_energy_ev = 1.0;
  p->power_index = 0.5;
  p->alpha = 1.2;
  p->beta = 0.35;
  p->coulomb_log = std::log(1.0e4);  // lambda_D / b_min with b_min = 1e-4 lambda_D
  // MonteCarloCollisionHandler::initializeDefaultCollisionParameters,
  // CollisionAlgorithm.cpp:590-646
  switch (type) {
    case SCDAT_MCC_ELASTIC:
      p->sigma_max_m2 = 2.0e-19;
      p->reference_energy_ev = 10.0;
      p->power_index = 0.3;
      break;
    case SCDAT_MCC_INELASTIC:
      p->threshold_ev = 1.0;
      p->sigma_max_m2 = 8.0e-20;
      p->energy_loss_ev = 1.0;
      break;
    case SCDAT_MCC_IONIZATION:
      p->threshold_ev = 15.76;
      p->sigma_max_m2 = 3.0e-20;
      p->energy_loss_ev = 15.76;
      break;
    case SCDAT_MCC_EXCITATION:
      p->threshold_ev = 11.5;
      p->sigma_max_m2 = 2.5e-20;
      p->energy_loss_ev = 11.5;
      break;
    case SCDAT_MCC_CHARGE_EXCHANGE:
      p->sigma_max_m2 = 1.2e-19;
      p->energy_loss_ev = 0.2;
      break;
    case SCDAT_MCC_RECOMBINATION:
      p->sigma_max_m2 = 5.0e-19;
      p->power_index = 0.7;
      break;
    default:
      p->sigma_max_m2 = 1.0e-18;
      break;
  }
  return SCDAT_OK;
}

int scdat_mcc_configure(scdat_ctx* c, const scdat_mcc_process* procs, int32_t n) {
  if (!c) return SCDAT_ERR_ARG;
  if (n < 0 || n > SCDAT_MAX_MCC_PROCESSES || (n > 0 && !procs))
    return fail(c, SCDAT_ERR_ARG, "mcc_configure: bad process count");
  const int ns = (int)c->sp.size();
  for (int k = 0; k < n; ++k) {
    const scdat_mcc_process& q = procs[k];
    if (q.type < SCDAT_MCC_ELASTIC || q.type > SCDAT_MCC_COULOMB || q.species < 0 || q.species >= ns ||
        !(q.target_density >= 0.0) || !(q.sigma_max_m2 >= 0.0))
      return fail(c, SCDAT_ERR_ARG, "mcc_configure: bad process " + std::to_string(k));
    if (q.type == SCDAT_MCC_IONIZATION &&
        (q.product_electron_species < 0 || q.product_electron_species >= ns || q.product_ion_species < 0 ||
         q.product_ion_species >= ns))
      return fail(c, SCDAT_ERR_ARG, "mcc_configure: ionisation needs product species");
  }
  c->mcc.assign(procs, procs + n);
  c->mcc_v2_stale = true;
  std::fill(c->seg_mcc_n.begin(), c->seg_mcc_n.end(), NAN);
  c->seg_mcc_dirty = true;
  return SCDAT_OK;
}

int scdat_mcc_set_density(scdat_ctx* c, int32_t process, double density) {
  if (!c) return SCDAT_ERR_ARG;
  if (process < 0 || process >= (int)c->mcc.size() || !(density >= 0.0))
    return fail(c, SCDAT_ERR_ARG, "mcc_set_density: bad process or density");
  c->mcc[process].target_density = density;
  c->seg_mcc_dirty = true;
  return SCDAT_OK;
}

int scdat_set_scenario_mcc_density(scdat_ctx* c, int32_t scenario, int32_t process, double density) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->segmented() || scenario < 0 || scenario >= c->n_seg())
    return fail(c, SCDAT_ERR_ARG, "set_scenario_mcc_density: no such scenario");
  if (process < 0 || process >= (int)c->mcc.size() || !(density >= 0.0) || !std::isfinite(density))
    return fail(c, SCDAT_ERR_ARG, "set_scenario_mcc_density: bad process or density");
  c->seg_mcc_n[(size_t)scenario * kMaxProc + process] = density;
  c->seg_mcc_dirty = true;
  return SCDAT_OK;
}

int scdat_set_surface_emission(scdat_ctx* c, int32_t group, const scdat_surface_emission* e) {
  if (!c || !e) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_surface_emission: set the mesh first");
  if (group < 0 || group >= kMaxGroups || e->see_model < SCDAT_SEE_OFF || e->see_model > SCDAT_SEE_STERNGLASS ||
      e->electron_species < 0 || e->electron_species >= (int)c->sp.size() || c->sp[e->electron_species].q >= 0)
    return fail(c, SCDAT_ERR_ARG, "set_surface_emission: bad group, model or electron species");
  if (e->see_model == SCDAT_SEE_STERNGLASS && !(e->see_energy_max_ev > 0.0))
    return fail(c, SCDAT_ERR_ARG, "set_surface_emission: Sternglass needs energy_max > 0");
  return guarded(c, [&] {
    c->emit[group] = *e;
    for (auto& se : c->seg_emit) se[group] = *e;
    c->seg_emit_dirty = c->segmented();
    bool see = false;
    for (const auto& q : c->emit) see = see || q.see_model != SCDAT_SEE_OFF;
    if (see && c->event_cap == 0) {
      // impacts per step: a 1/64 of the electron capacity, at least 65536
      int64_t cap = 0;
      for (const auto& S : c->sp)
        if (S.q < 0) cap += S.cap;
      c->event_cap = std::max<int64_t>(65536, cap / 64);
      c->d_events.alloc(sizeof(SeeEvent) * (size_t)c->event_cap);
    }
  });
}

int scdat_get_event_counts(scdat_ctx* c, scdat_event_counts* out) {
  if (!c || !out) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_event_counts: no mesh");
  return guarded(c, [&] {
    unsigned long long v[C_COUNT];
    HIPCHK(hipMemcpyAsync(v, c->d_counts.p, 8 * C_COUNT, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::memset(out, 0, sizeof(*out));
    for (int k = 0; k < SCDAT_MAX_MCC_PROCESSES; ++k) out->collisions[k] = (int64_t)v[k];
    out->created = (int64_t)v[C_CREATED];
    out->removed = (int64_t)v[C_REMOVED];
    out->see_emitted = (int64_t)v[C_SEE];
    out->photo_emitted = (int64_t)v[C_PHOTO];
    out->overflow = (int64_t)v[C_OVERFLOW];
    out->impacts_dropped = c->events_dropped;
    out->mcc_candidates = (int64_t)v[C_CAND];
    out->mcc_majorant_exceeded = (int64_t)v[C_MAJOR];
  });
}

// ---- field-emission cathodes ----------------------------------------------------

void scdat_field_emission_default(scdat_field_emission* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->emission_area_m2 = 1.0e-8;
  p->emitted_electron_speed_m_per_s = 2.0e6;
  p->emitted_electron_weight = 1.0;
  p->max_emitted_particles_per_event = 64;
  p->secondary_electron_speed_m_per_s = 8.0e5;
  p->secondary_electron_weight = 1.0;
  p->max_secondary_particles_per_event = 64;
}

int scdat_set_field_emission(scdat_ctx* c, int32_t group, const scdat_field_emission* e) {
  if (!c || !e) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_field_emission: set the mesh first");
  if (c->segmented()) return fail(c, SCDAT_ERR_STATE, "set_field_emission: needs a single scenario");
  if (group < 0 || group >= kMaxGroups || e->electron_species < 0 || e->electron_species >= (int)c->sp.size() ||
      c->sp[e->electron_species].q >= 0)
    return fail(c, SCDAT_ERR_ARG, "set_field_emission: bad group or electron species");
  return guarded(c, [&] {
    scdat_field_emission p = *e;
    // the constructor's defaults for non-positive settings
    // (FieldEmissionBoundaryCondition.cpp:23-55)
    if (p.emission_area_m2 <= 0.0) p.emission_area_m2 = 1.0e-8;
    if (p.emitted_electron_speed_m_per_s <= 0.0) p.emitted_electron_speed_m_per_s = 2.0e6;
    if (p.emitted_electron_weight <= 0.0) p.emitted_electron_weight = 1.0;
    if (p.max_emitted_particles_per_event <= 0) p.max_emitted_particles_per_event = 1;
    if (p.secondary_electron_speed_m_per_s <= 0.0) p.secondary_electron_speed_m_per_s = 8.0e5;
    if (p.secondary_electron_weight <= 0.0) p.secondary_electron_weight = 1.0;
    if (p.max_secondary_particles_per_event <= 0) p.max_secondary_particles_per_event = 1;
    // a species carries one weight: the emitted electrons take the species'
    const double ws = c->sp[p.electron_species].w;
    auto same = [&](double w) { return std::fabs(w - ws) <= 1e-12 * std::max(1.0, std::fabs(ws)); };
    if (!same(p.emitted_electron_weight) || (p.enable_secondary_electron_emission && !same(p.secondary_electron_weight)))
      throw Error(SCDAT_ERR_ARG, "set_field_emission: emitted electron weights must equal the electron species' weight");
    p.emission_current_density_a_per_m2 = std::max(0.0, p.emission_current_density_a_per_m2);
    // the emission condition replaces the group's particle condition and
    // absorbs what hits it (FieldEmissionBoundaryCondition.cpp:94-95)
    if (c->gtype[group] != SCDAT_BC_ABSORB) {
      c->gtype[group] = SCDAT_BC_ABSORB;
      c->gpartner[group] = -1;
      c->initialized = false;
    }
    c->fe[group] = p;
    c->fe_mask |= 1u << group;
    c->fe_pending[group] = 0.0;
    c->fe_pending_sec[group] = 0.0;
    if (c->event_cap == 0) {
      int64_t cap = 0;
      for (const auto& S : c->sp) cap += S.cap;
      c->event_cap = std::max<int64_t>(65536, cap / 64);
      c->d_events.alloc(sizeof(SeeEvent) * (size_t)c->event_cap);
    }
  });
}

int scdat_set_emission_current_density(scdat_ctx* c, int32_t group, double j) {
  if (!c) return SCDAT_ERR_ARG;
  if (group < 0 || group >= kMaxGroups || !((c->fe_mask >> group) & 1u))
    return fail(c, SCDAT_ERR_ARG, "set_emission_current_density: group has no field emission");
  c->fe[group].emission_current_density_a_per_m2 = std::max(0.0, j);
  return SCDAT_OK;
}

int scdat_append_particles(scdat_ctx* c, int32_t species, int64_t n, const double* x, const double* v,
                           const int32_t* cell) {
  if (!c) return SCDAT_ERR_ARG;
  if (species < 0 || species >= (int)c->sp.size()) return fail(c, SCDAT_ERR_ARG, "append_particles: bad species");
  if (n < 0 || (n > 0 && (!x || !v))) return fail(c, SCDAT_ERR_ARG, "append_particles: bad arrays");
  if (n == 0) return SCDAT_OK;
  return guarded(c, [&] {
    SpeciesState& S = c->sp[species];
    if (S.n + n > S.cap) throw Error(SCDAT_ERR_NOMEM, "append_particles: species capacity exceeded");
    std::vector<int32_t> cl;
    if (cell) cl.assign(cell, cell + n);
    else host_locate(c->h, n, x, cl);
    for (int64_t i = 0; i < n; ++i)
      if (cl[i] < -1 || cl[i] >= c->h.nt) throw Error(SCDAT_ERR_ARG, "append_particles: cell index out of range");
    const size_t ts = tsize(c);
    std::vector<char> f[6];
    for (int r = 0; r < 6; ++r) f[r].resize(ts * (size_t)n);
    for (int64_t i = 0; i < n; ++i)
      for (int r = 0; r < 3; ++r) {
        if (ts == 4) {
          reinterpret_cast<float*>(f[r].data())[i] = (float)x[3 * i + r];
          reinterpret_cast<float*>(f[3 + r].data())[i] = (float)v[3 * i + r];
        } else {
          reinterpret_cast<double*>(f[r].data())[i] = x[3 * i + r];
          reinterpret_cast<double*>(f[3 + r].data())[i] = v[3 * i + r];
        }
      }
    std::vector<int32_t> id((size_t)n);
    for (int64_t i = 0; i < n; ++i) id[i] = S.next_id++;
    for (int r = 0; r < 6; ++r)
      HIPCHK(hipMemcpyAsync(static_cast<char*>(S.buf[S.cur][r].p) + ts * (size_t)S.n, f[r].data(), ts * (size_t)n,
                            hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(S.buf[S.cur][6].as<int32_t>() + S.n, cl.data(), 4 * (size_t)n, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipMemcpyAsync(S.buf[S.cur][7].as<int32_t>() + S.n, id.data(), 4 * (size_t)n, hipMemcpyHostToDevice,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    S.n += n;
  });
}

// ---- layer profile (result export) ----------------------------------------------

namespace {
std::map<double, std::vector<int64_t>> layer_nodes(const scdat_ctx* c, int axis) {
  std::map<double, std::vector<int64_t>> layers;
  for (int64_t i = 0; i < c->h.nn; ++i)
    layers[(double)std::llround(c->h.xyz[3 * i + axis] * 1.0e9) * 1.0e-9].push_back(i);
  return layers;
}
}  // namespace

int scdat_layer_count(scdat_ctx* c, int32_t axis, int64_t* n) {
  if (!c || !n || axis < 0 || axis > 2) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "layer_count: no mesh");
  return guarded(c, [&] { *n = (int64_t)layer_nodes(c, axis).size(); });
}

int scdat_layer_profile(scdat_ctx* c, int32_t axis, double length, int32_t n_cells, double area, double* position,
                        double* potential, double* field, double* electron_density, double* ion_density,
                        double* electron_temperature_ev) {
  if (!c || axis < 0 || axis > 2 || !(length > 0.0) || n_cells < 1 || !(area > 0.0)) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "layer_profile: no mesh");
  return guarded(c, [&] {
    const auto layers = layer_nodes(c, axis);
    const size_t nl = layers.size(), nn = (size_t)c->h.nn;
    std::vector<double> phi(nn), E(3 * nn);
    {
      std::vector<double> pu(nn), eu(3 * nn);
      const int rc = scdat_get_fields(c, nullptr, pu.data(), eu.data());
      if (rc != SCDAT_OK) throw Error(rc, c->err);
      for (size_t i = 0; i < nn; ++i) {  // back to the internal numbering of h
        const int64_t u = to_user(c, (int64_t)i);
        phi[i] = pu[u];
        for (int r = 0; r < 3; ++r) E[3 * i + r] = eu[3 * u + r];
      }
    }
    size_t k = 0;
    for (const auto& L : layers) {
      double ps = 0.0, fs = 0.0;
      for (int64_t i : L.second) {
        ps += phi[i];
        fs += E[3 * i + axis];
      }
      if (position) position[k] = L.first;
      if (potential) potential[k] = ps / (double)L.second.size();
      if (field) field[k] = fs / (double)L.second.size();
      ++k;
    }
    const double dz = length / (double)n_cells, vol = area * dz;
    std::vector<double> ne(nl, 0.0), ni(nl, 0.0), acc(nl, 0.0);
    for (size_t si = 0; si < c->sp.size(); ++si) {
      const SpeciesState& S = c->sp[si];
      if (S.n == 0 || S.q == 0.0) continue;
      std::vector<double> x(3 * (size_t)S.n), v(3 * (size_t)S.n);
      std::vector<int32_t> cell((size_t)S.n), id((size_t)S.n);
      if (c->is_fp64()) get_particles_t<double>(c, (int)si, x.data(), v.data(), cell.data(), id.data());
      else get_particles_t<float>(c, (int)si, x.data(), v.data(), cell.data(), id.data());
      for (int64_t p = 0; p < S.n; ++p) {
        if (cell[p] < 0) continue;
        const double w = S.w * (c->segmented() ? c->seg_wscale[c->tet_seg_h[cell[p]]] : 1.0);
        const double z = std::min(std::max(x[3 * p + axis], 0.0), length);
        const long long b = std::min<long long>(std::max<long long>(std::llround(z / dz), 0), (long long)nl - 1);
        if (S.q < 0.0) {
          ne[b] += w / vol;
          const double ke = 0.5 * S.m * (v[3 * p] * v[3 * p] + v[3 * p + 1] * v[3 * p + 1] + v[3 * p + 2] * v[3 * p + 2]);
          acc[b] += w * ke / kElem;
        } else {
          ni[b] += w / vol;
        }
      }
    }
    for (size_t b = 0; b < nl; ++b) {
      if (electron_density) electron_density[b] = ne[b];
      if (ion_density) ion_density[b] = ni[b];
      if (electron_temperature_ev)
        electron_temperature_ev[b] = ne[b] > 0.0 ? acc[b] / std::max(1.0, ne[b] * vol) / 1.5 : 0.0;
    }
  });
}

// ---- batched scenarios ----------------------------------------------------------

int scdat_set_scenarios(scdat_ctx* c, int32_t n, const int64_t* node_start, const double* weight_scale,
                        const double* time_step, const uint64_t* seed) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_scenarios: set the mesh first");
  if (n < 1 || !node_start) return fail(c, SCDAT_ERR_ARG, "set_scenarios: need n >= 1 and node_start");
  for (const auto& S : c->sp)
    if (S.n > 0) return fail(c, SCDAT_ERR_STATE, "set_scenarios: call before set_particles");
  if (c->n_patch > 0) return fail(c, SCDAT_ERR_STATE, "set_scenarios: conductor patches need a single scenario");
  if (c->fe_mask) return fail(c, SCDAT_ERR_STATE, "set_scenarios: field emission needs a single scenario");
  return guarded(c, [&] {
    mesh_natural_order(c);  // scenarios are ranges of the caller's node numbering
    const HostMesh& h = c->h;
    const int64_t align = SCDAT_SEGMENT_ALIGN;
    static_assert(SCDAT_SEGMENT_ALIGN == kPcgBlock, "scenario rows align with the PCG row blocks");
    if (node_start[0] != 0 || node_start[n] != h.nn)
      throw Error(SCDAT_ERR_ARG, "set_scenarios: node_start must run from 0 to the node count");
    for (int sg = 0; sg < n; ++sg) {
      if (node_start[sg + 1] <= node_start[sg]) throw Error(SCDAT_ERR_ARG, "set_scenarios: empty scenario");
      if (node_start[sg] % align)
        throw Error(SCDAT_ERR_ARG, "set_scenarios: scenario starts must be multiples of SCDAT_SEGMENT_ALIGN");
      if (weight_scale && !(weight_scale[sg] > 0.0)) throw Error(SCDAT_ERR_ARG, "set_scenarios: weight scale <= 0");
      if (time_step && !(time_step[sg] > 0.0 && std::isfinite(time_step[sg])))
        throw Error(SCDAT_ERR_ARG, "set_scenarios: bad time step");
    }
    std::vector<int32_t> tseg((size_t)h.nt);
    for (int64_t t = 0; t < h.nt; ++t) {
      const int32_t a = h.tets[4 * t];
      const int sg = (int)(std::upper_bound(node_start, node_start + n + 1, (int64_t)a) - node_start) - 1;
      for (int k = 1; k < 4; ++k) {
        const int64_t b = h.tets[4 * t + k];
        if (b < node_start[sg] || b >= node_start[sg + 1])
          throw Error(SCDAT_ERR_ARG, "set_scenarios: tetrahedron " + std::to_string(t) + " spans two scenarios");
      }
      tseg[t] = sg;
    }
    const int64_t G = (h.nn + align - 1) / align;
    std::vector<int32_t> bseg((size_t)G), b0((size_t)n + 1);
    for (int sg = 0; sg <= n; ++sg) b0[sg] = (int32_t)((node_start[sg] + align - 1) / align);
    for (int sg = 0; sg < n; ++sg)
      for (int32_t b = b0[sg]; b < b0[sg + 1]; ++b) bseg[b] = sg;
    c->seg_node0.assign(node_start, node_start + n + 1);
    c->tet_seg_h = std::move(tseg);
    c->blk_seg_h = std::move(bseg);
    c->seg_b0_h = std::move(b0);
    c->seg_wscale.assign((size_t)n, 1.0);
    c->seg_dt.assign((size_t)n, c->cfg.dt);
    c->seg_key.assign(2 * (size_t)n, 0);
    for (int sg = 0; sg < n; ++sg) {
      if (weight_scale) c->seg_wscale[sg] = weight_scale[sg];
      if (time_step) c->seg_dt[sg] = time_step[sg];
      c->seg_key[2 * sg] = seed ? (uint32_t)seed[sg] : c->key[0];
      c->seg_key[2 * sg + 1] = seed ? (uint32_t)(seed[sg] >> 32) : c->key[1];
    }
    c->seg_emit.assign((size_t)n, c->emit);
    c->seg_mcc_n.assign((size_t)n * kMaxProc, NAN);
    c->seg_mcc_dirty = true;
    const size_t m = (size_t)kMaxGroups * kMaxSpecies * n;
    for (DevBuf* b : {&c->d_ledger_count, &c->d_ledger_energy, &c->d_emit_count}) {
      b->alloc(8 * m);
      HIPCHK(hipMemsetAsync(b->p, 0, 8 * m, c->stream));
    }
    c->d_sc.alloc(sizeof(PcgScalars) * ((size_t)n + 1));
    HIPCHK(hipMemsetAsync(c->d_sc.p, 0, sizeof(PcgScalars) * ((size_t)n + 1), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->initialized = false;
  });
}

int scdat_set_scenario_emission(scdat_ctx* c, int32_t scenario, int32_t group, const scdat_surface_emission* e) {
  if (!c || !e) return SCDAT_ERR_ARG;
  if (!c->segmented() || scenario < 0 || scenario >= c->n_seg())
    return fail(c, SCDAT_ERR_ARG, "set_scenario_emission: no such scenario");
  if (group < 0 || group >= kMaxGroups || e->see_model < SCDAT_SEE_OFF || e->see_model > SCDAT_SEE_STERNGLASS ||
      e->electron_species < 0 || e->electron_species >= (int)c->sp.size() || c->sp[e->electron_species].q >= 0)
    return fail(c, SCDAT_ERR_ARG, "set_scenario_emission: bad group, model or electron species");
  if (e->see_model == SCDAT_SEE_STERNGLASS && !(e->see_energy_max_ev > 0.0))
    return fail(c, SCDAT_ERR_ARG, "set_scenario_emission: Sternglass needs energy_max > 0");
  if (!(e->photo_current_density >= 0.0))
    return fail(c, SCDAT_ERR_ARG, "set_scenario_emission: photoemission current density < 0");
  return guarded(c, [&] {
    c->seg_emit[scenario][group] = *e;
    c->seg_emit_dirty = true;
    if (e->see_model != SCDAT_SEE_OFF && c->event_cap == 0) {
      int64_t cap = 0;
      for (const auto& S : c->sp)
        if (S.q < 0) cap += S.cap;
      c->event_cap = std::max<int64_t>(65536, cap / 64);
      c->d_events.alloc(sizeof(SeeEvent) * (size_t)c->event_cap);
    }
  });
}

int scdat_get_scenario_stats(scdat_ctx* c, int32_t scenario, scdat_scenario_stats* out) {
  if (!c || !out) return SCDAT_ERR_ARG;
  if (scenario < 0 || scenario >= c->n_seg()) return fail(c, SCDAT_ERR_ARG, "get_scenario_stats: no such scenario");
  if (!c->initialized) return fail(c, SCDAT_ERR_STATE, "get_scenario_stats: not initialized");
  return guarded(c, [&] {
    PcgScalars q;
    HIPCHK(hipMemcpyAsync(&q, c->d_sc.as<PcgScalars>() + scenario, sizeof(q), hipMemcpyDeviceToHost, c->stream));
    const LedgerHost L = read_ledger(c);
    std::memset(out, 0, sizeof(*out));
    const double bn = std::sqrt(q.bnorm2);
    out->completed_steps = c->steps;
    out->simulation_time = (double)c->steps * (c->segmented() ? c->seg_dt[scenario] : c->cfg.dt);
    out->floating_potential = c->h.has_cond ? q.vfloat : 0.0;
    out->pcg_iterations = q.total_it;
    out->pcg_converged = q.failed ? 0 : 1;
    out->pcg_rel_residual = bn > 0 ? q.rnorm / bn : q.rnorm;
    double qc = c->q_init;
    const size_t lo = (size_t)scenario * kMaxGroups * kMaxSpecies;
    for (int g = 0; g < kMaxGroups; ++g) {
      if (c->gtype[g] != SCDAT_BC_CONDUCTOR) continue;
      for (size_t sp = 0; sp < c->sp.size(); ++sp) {
        const double qw = c->sp[sp].q * (c->sp[sp].w * seg_scale(c, scenario));
        qc += ((double)L.cnt[lo + g * kMaxSpecies + sp] - (double)L.emi[lo + g * kMaxSpecies + sp]) * qw;
      }
    }
    out->conductor_charge = qc;
    out->time_step = c->segmented() ? c->seg_dt[scenario] : c->cfg.dt;
    out->weight_scale = seg_scale(c, scenario);
    out->first_node = c->segmented() ? c->seg_node0[scenario] : 0;
    out->n_nodes = c->segmented() ? c->seg_node0[scenario + 1] - c->seg_node0[scenario] : c->h.nn;
  });
}

int scdat_get_scenario_ledger(scdat_ctx* c, int32_t scenario, scdat_ledger_entry* out, int32_t n_groups) {
  if (!c || !out || n_groups < 0 || n_groups > kMaxGroups) return SCDAT_ERR_ARG;
  if (!c->has_mesh || scenario < 0 || scenario >= c->n_seg())
    return fail(c, SCDAT_ERR_ARG, "get_scenario_ledger: no such scenario");
  return guarded(c, [&] { fill_ledger(c, read_ledger(c), scenario, scenario + 1, out, n_groups); });
}

// ---- launch shapes ---------------------------------------------------------------

int scdat_set_node_order(scdat_ctx* c, int32_t order) {
  if (!c) return SCDAT_ERR_ARG;
  if (order != SCDAT_ORDER_MORTON && order != SCDAT_ORDER_GIVEN && order != SCDAT_ORDER_TILED)
    return fail(c, SCDAT_ERR_ARG, "set_node_order: unknown order");
  if (c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_node_order: call before scdat_set_mesh");
  c->node_order = order;
  return SCDAT_OK;
}

int scdat_set_preconditioner(scdat_ctx* c, int32_t kind) {
  if (!c) return SCDAT_ERR_ARG;
  if (kind != SCDAT_PRECOND_JACOBI && kind != SCDAT_PRECOND_TWO_LEVEL)
    return fail(c, SCDAT_ERR_ARG, "set_preconditioner: unknown kind");
  c->precond = kind;
  return SCDAT_OK;
}

int scdat_get_solver_info(scdat_ctx* c, scdat_solver_info* out) {
  if (!c || !out) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "get_solver_info: no mesh");
  std::memset(out, 0, sizeof(*out));
  const EllDev A = ell(c);
  out->rows = A.n;
  out->ell_width = A.w;
  int64_t nnz = 0;
  for (int64_t i = 0; i < c->h.nn; ++i)
    for (int sl = 0; sl < c->h.W; ++sl) nnz += c->h.aval[(size_t)sl * c->h.nn + i] != 0.0;
  out->nonzeros = nnz;
  out->coarse_unknowns = A.nc;
  out->row_patterns = A.npat;
  out->pcg_block = A.nc > 0 ? kPcgBlock : A.pb;
  out->preconditioner = A.nc > 0 ? SCDAT_PRECOND_TWO_LEVEL : SCDAT_PRECOND_JACOBI;
  out->coarse_in_kernel = A.fused;
  out->tile_staged = A.box && (A.nc == 0 || A.fused);
  out->box_patterns = c->nbpat;
  return SCDAT_OK;
}

int scdat_set_launch_shape(scdat_ctx* c, int32_t what, int32_t value) {
  if (!c) return SCDAT_ERR_ARG;
  if (what == SCDAT_SHAPE_PCG_BLOCK) {
    if (value != 0 && value != 256 && value != 512)
      return fail(c, SCDAT_ERR_ARG, "set_launch_shape: PCG block must be 0, 256 or 512");
    c->pcg_block = value;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_ROW_PATTERNS) {
    if (value < 0 || value > 2) return fail(c, SCDAT_ERR_ARG, "set_launch_shape: row patterns must be 0, 1 or 2");
    c->use_patterns = value != 0;
    const bool wide = value == 2;
    if (wide != c->pat_force_wide) {
      c->pat_force_wide = wide;
      if (c->has_mesh && !c->is_fp64())
        return guarded(c, [&] {
          // the FP32 values of A D^-1 and D^-1 as the upload rounds them
          const HostMesh& h = c->h;
          std::vector<float> vh(h.aval.size()), di(h.dinv.size());
          for (size_t o = 0; o < vh.size(); ++o) vh[o] = (float)(h.aval[o] * h.dinv[h.col[o]]);
          for (size_t i = 0; i < di.size(); ++i) di[i] = (float)h.dinv[i];
          build_row_patterns(c, vh, di);
        });
    }
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_TILE_STAGING) {
    if (value != 0 && value != 1) return fail(c, SCDAT_ERR_ARG, "set_launch_shape: tile staging must be 0 or 1");
    c->use_box = value != 0;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_CELL_ORDER) {
    if (value != 0 && value != 1) return fail(c, SCDAT_ERR_ARG, "set_launch_shape: cell order must be 0 or 1");
    if (value != c->cell_order) c->cell_key_ready = false;
    c->cell_order = value;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_STEP_GRAPH) {
    if (value != 0 && value != 1) return fail(c, SCDAT_ERR_ARG, "set_launch_shape: step graph must be 0 or 1");
    c->step_graph = value;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_LATTICE_PUSH) {
    if (value != 0 && value != 1) return fail(c, SCDAT_ERR_ARG, "set_launch_shape: lattice push must be 0 or 1");
    c->lat_mode = value;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_DEPOSIT_BUCKETS) {
    if (value != 0 && value != 1)
      return fail(c, SCDAT_ERR_ARG, "set_launch_shape: deposit buckets must be 0 or 1");
    c->dep_mode = value;
    return SCDAT_OK;
  }
  if (what != SCDAT_SHAPE_PUSH_CHUNKS || !(value == 1 || value == 2 || value == 4 || value == 8))
    return fail(c, SCDAT_ERR_ARG, "set_launch_shape: push chunks must be 1, 2, 4 or 8");
  c->push_chunks = value;
  return SCDAT_OK;
}

int scdat_get_launch_shape(scdat_ctx* c, int32_t what, int32_t* value) {
  if (!c || !value) return SCDAT_ERR_ARG;
  if (what == SCDAT_SHAPE_PCG_BLOCK) {
    *value = c->pcg_block;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_ROW_PATTERNS) {
    *value = c->use_patterns ? (c->pat_force_wide ? 2 : 1) : 0;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_TILE_STAGING) {
    *value = c->use_box ? 1 : 0;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_CELL_ORDER) {
    *value = c->cell_order;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_STEP_GRAPH) {
    *value = c->step_graph;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_DEPOSIT_BUCKETS) {
    *value = c->dep_mode;
    return SCDAT_OK;
  }
  if (what == SCDAT_SHAPE_LATTICE_PUSH) {
    // 1 only when the push locates on the lattice (FP32, lattice mesh)
    *value = c->lat_mode && c->lat_dev.tet_hex && !c->is_fp64() ? 1 : 0;
    return SCDAT_OK;
  }
  if (what != SCDAT_SHAPE_PUSH_CHUNKS) return SCDAT_ERR_ARG;
  *value = c->push_chunks;
  return SCDAT_OK;
}

int scdat_autotune(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->initialized) return fail(c, SCDAT_ERR_STATE, "autotune: call scdat_initialize first");
  return guarded(c, [&] {
    if (c->is_fp64()) autotune_push<double>(c);
    else autotune_push<float>(c);
  });
}

// ---- discrete adjoint of the charging cycle ------------------------------------

void scdat_charging_params_default(scdat_charging_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->see_delta_max = 2.0;
  p->see_energy_max_ev = 400.0;
  for (int k = 0; k < 6; ++k) p->face_yield[k] = 1.0;
  p->photo_current_density = 2.0e-5;
  p->photo_temperature_ev = 2.0;
  p->photo_reference_v = 0.0;
  p->sunlit_faces = 1 << 1;  // XMax face of the conductor (group 7)
  p->surface_conductance_s = 0.0;
}

int scdat_charging_forward(scdat_ctx* c, const scdat_charging_params* p, int32_t steps, double* v_final) {
  return scdat_charging_forward_events(c, p, steps, SCDAT_EVENTS_FREE, v_final);
}

int scdat_charging_forward_events(scdat_ctx* c, const scdat_charging_params* p, int32_t steps, int32_t events,
                                  double* v_final) {
  if (!c || !p || !v_final || steps < 0 || events < SCDAT_EVENTS_FREE || events > SCDAT_EVENTS_REPLAY)
    return SCDAT_ERR_ARG;
  return guarded(c, [&] {
    *v_final = c->is_fp64() ? charging_forward<double>(c, p, steps, events)
                            : charging_forward<float>(c, p, steps, events);
  });
}

int scdat_charging_adjoint(scdat_ctx* c, const scdat_charging_params* p, int32_t steps, int32_t checkpoint_interval,
                           double* v_final, double* grad) {
  if (!c || !p || !v_final || !grad || steps < 1 || checkpoint_interval < 1) return SCDAT_ERR_ARG;
  return guarded(c, [&] {
    *v_final = c->is_fp64() ? charging_adjoint<double>(c, p, steps, checkpoint_interval, grad)
                            : charging_adjoint<float>(c, p, steps, checkpoint_interval, grad);
  });
}

// ---- conductor patches driven by a surface circuit ----------------------------

int scdat_set_conductor_circuit(scdat_ctx* c, int32_t n_patches, const int32_t* node_patch, scdat_circuit* circuit,
                                const int32_t* patch_circuit_node, const int32_t* group_patch) {
  if (!c) return SCDAT_ERR_ARG;
  if (!c->has_mesh) return fail(c, SCDAT_ERR_STATE, "set_conductor_circuit: set the mesh first");
  if (c->segmented() && n_patches != 0)
    return fail(c, SCDAT_ERR_STATE, "set_conductor_circuit: conductor patches need a single scenario");
  if (n_patches == 0) {
    c->n_patch = 0;
    c->circuit = nullptr;
    c->initialized = false;
    return SCDAT_OK;
  }
  if (n_patches < 0 || n_patches > kMaxPatches || !circuit || !patch_circuit_node)
    return fail(c, SCDAT_ERR_ARG, "set_conductor_circuit: bad patch count, circuit or node map");
  if (!c->h.has_cond) return fail(c, SCDAT_ERR_STATE, "set_conductor_circuit: the mesh has no conductor nodes");
  int32_t ncn = 0, nbr = 0;
  scdat_circuit_size(circuit, &ncn, &nbr);
  for (int k = 0; k < n_patches; ++k)
    if (patch_circuit_node[k] < 0 || patch_circuit_node[k] >= ncn)
      return fail(c, SCDAT_ERR_ARG, "set_conductor_circuit: patch mapped to a missing circuit node");
  return guarded(c, [&] {
    const HostMesh& h = c->h;
    std::vector<int32_t> np((size_t)h.nn, -1);
    if (node_patch) {
      for (int64_t i = 0; i < h.nn; ++i) np[i] = node_patch[to_user(c, i)];
    } else {
      // six conductor faces (groups 6..11 -> patches 0..5) and the body
      // (patch 6: nodes on more than one face)
      if (n_patches != 7) throw Error(SCDAT_ERR_ARG, "set_conductor_circuit: default patches need n_patches = 7");
      std::vector<uint32_t> bits((size_t)h.nn, 0u);
      for (int64_t t = 0; t < h.nt; ++t)
        for (int k = 0; k < 4; ++k) {
          const int g = h.group[4 * t + k];
          if (g < 6 || g > 11) continue;
          for (int q = 0; q < 4; ++q)
            if (q != k) bits[h.tets[4 * t + q]] |= 1u << (g - 6);
        }
      for (int64_t i = 0; i < h.nn; ++i) {
        if (h.kind[i] != SCDAT_NODE_CONDUCTOR || bits[i] == 0u) continue;
        np[i] = __builtin_popcount(bits[i]) == 1 ? __builtin_ctz(bits[i]) : 6;
      }
    }
    for (int64_t i = 0; i < h.nn; ++i) {
      if (h.kind[i] == SCDAT_NODE_CONDUCTOR && (np[i] < 0 || np[i] >= n_patches))
        throw Error(SCDAT_ERR_ARG, "set_conductor_circuit: conductor node " + std::to_string(i) + " has no patch");
      if (h.kind[i] != SCDAT_NODE_CONDUCTOR) np[i] = -1;
    }
    c->node_patch = np;
    c->n_patch = n_patches;
    c->circuit = circuit;
    for (int k = 0; k < n_patches; ++k) c->patch_cnode[k] = patch_circuit_node[k];
    for (int g = 0; g < kMaxGroups; ++g)
      c->group_patch[g] = group_patch ? group_patch[g] : (g >= 6 && g < 12 && g - 6 < n_patches ? g - 6 : -1);
    c->patch_didv.fill(0.0);
    c->d_patch_v.alloc(8 * kMaxPatches);
    c->initialized = false;
  });
}

int scdat_set_patch_didv(scdat_ctx* c, const double* didv) {
  if (!c || !didv) return SCDAT_ERR_ARG;
  for (int k = 0; k < c->n_patch; ++k) c->patch_didv[k] = didv[k];
  return SCDAT_OK;
}

int scdat_get_patch_potentials(scdat_ctx* c, double* v) {
  if (!c || !v) return SCDAT_ERR_ARG;
  for (int k = 0; k < c->n_patch; ++k) v[k] = c->h_patch_v[k];
  return SCDAT_OK;
}

int scdat_get_patch_basis(scdat_ctx* c, int32_t patch, double* phi) {
  if (!c || !phi) return SCDAT_ERR_ARG;
  if (patch < 0 || patch >= c->n_patch || !c->initialized)
    return fail(c, SCDAT_ERR_STATE, "get_patch_basis: no such patch or not initialised");
  return guarded(c, [&] {
    std::vector<double> b((size_t)c->h.nn);
    HIPCHK(hipMemcpyAsync(b.data(), c->d_phik.as<double>() + (size_t)patch * c->h.nn, 8 * (size_t)c->h.nn,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < c->h.nn; ++i) phi[to_user(c, i)] = b[i];
  });
}

int scdat_get_stage_times(scdat_ctx* c, double* ms5) {
  if (!c || !ms5) return SCDAT_ERR_ARG;
  for (int k = 0; k < 5; ++k) ms5[k] = c->stage_ms[k];
  return SCDAT_OK;
}

int scdat_get_push_kernel_time(scdat_ctx* c, double* avg_ms, int64_t* launches) {
  if (!c) return SCDAT_ERR_ARG;
  if (avg_ms) *avg_ms = c->push_launches ? c->push_ms_sum / (double)c->push_launches : 0.0;
  if (launches) *launches = c->push_launches;
  return SCDAT_OK;
}

int scdat_set_kernel_timing(scdat_ctx* c, int32_t on) {
  if (!c) return SCDAT_ERR_ARG;
  c->kernel_timing = on > 0 ? on : 0;
  for (int k = 0; k < 2; ++k) {
    c->kt_sum[k] = 0.0;
    c->kt_n[k] = 0;
  }
  return SCDAT_OK;
}

int scdat_get_kernel_time(scdat_ctx* c, int32_t which, double* avg_ms, int64_t* launches) {
  if (!c || which < 0 || which > 1) return SCDAT_ERR_ARG;
  if (avg_ms) *avg_ms = c->kt_n[which] ? c->kt_sum[which] / (double)c->kt_n[which] : 0.0;
  if (launches) *launches = c->kt_n[which];
  return SCDAT_OK;
}

int scdat_sync(scdat_ctx* c) {
  if (!c) return SCDAT_ERR_ARG;
  return guarded(c, [&] { HIPCHK(hipStreamSynchronize(c->stream)); });
}

}  // extern "C"
