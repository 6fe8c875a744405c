// SCDAT PIC-MCC oracle: FP64 CPU restatement of the reference hot path.
//
// TEST INFRASTRUCTURE ONLY.  This library is the checker for the GPU path and
// the timed CPU baseline.  Only the test-suite, the smoke check and the CPU
// baseline leg of the benchmark load it; the product library never links,
// loads or calls it.
//
// Every routine below restates one reference routine and cites it
// (paths relative to the SCDAT source tree).  Where the oracle deliberately
// departs from the reference (physical sign of the Poisson load, sparse CSR
// instead of a dense matrix, face-neighbour walk instead of an O(N) scan) the
// departure is stated at the routine.
//
// Parity pinning status per routine is listed in oracle/README.md.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr double kE = 1.602176634e-19;        // Tools/Basic/include/Constants.h:46
constexpr double kEps0 = 8.8541878128e-12;    // Tools/Basic/include/Constants.h:53

enum BcType { BC_ABSORB = 0, BC_PERIODIC = 1, BC_REFLECT = 2, BC_CONDUCTOR = 3, BC_ESCAPE = 4 };
constexpr int kMaxGroups = 16;

struct Ledger {
  double abs_e_charge = 0, abs_i_charge = 0, emit_e_charge = 0, energy = 0;
  int64_t n_abs_e = 0, n_abs_i = 0, n_emit_e = 0;
};

struct Species {
  double q = 0, m = 1, w = 1;
  std::vector<double> x, v;  // 3 per particle
  std::vector<int32_t> cell; // -1 = removed
  std::vector<int32_t> id;   // upload index, or a generated id (bit 31 set)
};

// Monte Carlo collision process of one incident species with a uniform
// background.  Types follow CollisionProcessFactory
// (Tools/Interactions/Collisions/src/CollisionAlgorithm.cpp:738-812).
enum MccType { MCC_ELASTIC = 0, MCC_INELASTIC, MCC_IONIZATION, MCC_EXCITATION, MCC_CHARGE_EXCHANGE,
               MCC_RECOMBINATION, MCC_COULOMB };
struct MccProc {
  int type = 0, species = 0, e_species = -1, i_species = -1;
  double n = 0, thr = 0, smax = 0, loss = 0, eref = 1, pidx = 0.5, alpha = 1.2, beta = 0.35, clog = 10;
};
// Surface emission of one boundary group.
struct SurfEmit {
  int see_model = 0;  // 0 off, 1 BasicSurfaceMaterialModel secondaryYield, 2 Sternglass
  double dmax = 0, emax = 400, t_see = 2;
  double photo_j = 0, t_ph = 2, v_ref = 0;  // photo_j = J0 * Y (A/m^2)
  int e_species = 0;
};
struct SeeEvent { int species, id, cell, group, face; double x[3], v[3]; };

struct Csr {
  int n = 0;
  std::vector<int64_t> rp;
  std::vector<int32_t> ci;
  std::vector<double> va;
  void spmv(const double* x, double* y) const {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      double s = 0;
      for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += va[k] * x[ci[k]];
      y[i] = s;
    }
  }
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0;
  const int n = (int)a.size();
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  return s;
}

// Jacobi-preconditioned conjugate gradient.
// Follows Tools/FieldSolver/src/MultiBoundaryElectricFieldSolver.cpp:1432-1513
// (r = b - A x0, z = D^-1 r, alpha = rz/pAp, beta = rz_new/rz_old, breakdown
// guard 1e-30) and the OpenMP CSR structure of
// Tools/Parallel/src/ParallelSolver.cpp:33-106.  Departure: the stopping test
// is relative, ||r|| <= tol * ||b||, instead of the reference's absolute 1e-12.
struct PcgResult { int iters; double rel_res; int converged; };
PcgResult pcg(const Csr& A, const std::vector<double>& b, std::vector<double>& x, double tol,
              int max_it) {
  const int n = A.n;
  std::vector<double> r(n), z(n), p(n), q(n), dinv(n, 1.0);
  for (int i = 0; i < n; ++i)
    for (int64_t k = A.rp[i]; k < A.rp[i + 1]; ++k)
      if (A.ci[k] == i && std::abs(A.va[k]) > 0) dinv[i] = 1.0 / A.va[k];
  A.spmv(x.data(), q.data());
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  const double bnorm = std::sqrt(dot(b, b));
  const double stop = tol * (bnorm > 0 ? bnorm : 1.0);
  double rn = std::sqrt(dot(r, r));
  if (rn <= stop) return {0, bnorm > 0 ? rn / bnorm : rn, 1};
  for (int i = 0; i < n; ++i) { z[i] = r[i] * dinv[i]; p[i] = z[i]; }
  double rz = dot(r, z);
  for (int it = 1; it <= max_it; ++it) {
    A.spmv(p.data(), q.data());
    const double pq = dot(p, q);
    if (std::abs(pq) < 1e-300) return {it, rn / (bnorm > 0 ? bnorm : 1), 0};
    const double alpha = rz / pq;
    double rr = 0;
#pragma omp parallel for reduction(+ : rr) schedule(static)
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      rr += r[i] * r[i];
    }
    rn = std::sqrt(rr);
    if (rn <= stop) return {it, bnorm > 0 ? rn / bnorm : rn, 1};
    double rz_new = 0;
#pragma omp parallel for reduction(+ : rz_new) schedule(static)
    for (int i = 0; i < n; ++i) { z[i] = r[i] * dinv[i]; rz_new += r[i] * z[i]; }
    const double beta = rz_new / rz;
    rz = rz_new;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
  }
  return {max_it, rn / (bnorm > 0 ? bnorm : 1), 0};
}

// Boris velocity update, Tools/Particle/src/PICParticleAlgorithms.cpp:803-851.
inline void boris(double* v, const double* E, const double* B, double qm, double dt) {
  const double h = 0.5 * dt;
  double vm[3] = {v[0] + E[0] * (qm * h), v[1] + E[1] * (qm * h), v[2] + E[2] * (qm * h)};
  double vp[3] = {vm[0], vm[1], vm[2]};
  const double bmag = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  if (bmag > 1e-20) {
    const double t[3] = {B[0] * (qm * h), B[1] * (qm * h), B[2] * (qm * h)};
    const double t2 = t[0] * t[0] + t[1] * t[1] + t[2] * t[2];
    if (t2 < 100.0) {
      const double sf = 2.0 / (1.0 + t2);
      const double s[3] = {t[0] * sf, t[1] * sf, t[2] * sf};
      const double vs[3] = {vm[0] + (vm[1] * t[2] - vm[2] * t[1]),
                            vm[1] + (vm[2] * t[0] - vm[0] * t[2]),
                            vm[2] + (vm[0] * t[1] - vm[1] * t[0])};
      vp[0] = vm[0] + (vs[1] * s[2] - vs[2] * s[1]);
      vp[1] = vm[1] + (vs[2] * s[0] - vs[0] * s[2]);
      vp[2] = vm[2] + (vs[0] * s[1] - vs[1] * s[0]);
    }
  }
  v[0] = vp[0] + E[0] * (qm * h);
  v[1] = vp[1] + E[1] * (qm * h);
  v[2] = vp[2] + E[2] * (qm * h);
}

// Adaptive sub-step, Tools/Particle/src/PICParticleAlgorithms.cpp:750-801:
// dt = min(dt_max, 0.1|v0|/|qE/m| when |E|>1e-10, |a|>1e-10, |v0|>1 m/s;
// 0.2*pi/omega_c when |B|>1e-10), clamped to [floor*dt_max, dt_max].  v0 is the
// velocity at the start of the whole step (the reference reads the particle,
// which is only updated after the loop).
inline double adaptive_dt(const double* v0, const double* E, const double* B, double qm,
                          double dt_max, double floor_ratio) {
  double dt = dt_max;
  const double emag = std::sqrt(E[0] * E[0] + E[1] * E[1] + E[2] * E[2]);
  if (emag > 1e-10) {
    const double acc = emag * std::abs(qm);
    const double vel = std::sqrt(v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2]);
    if (acc > 1e-10 && vel > 1.0) dt = std::min(dt, 0.1 * vel / acc);
  }
  const double bmag = std::sqrt(B[0] * B[0] + B[1] * B[1] + B[2] * B[2]);
  if (bmag > 1e-10) {
    const double wc = bmag * std::abs(qm);
    if (wc > 1e-10) dt = std::min(dt, 0.1 * 2.0 * M_PI / wc);
  }
  dt = std::max(dt, dt_max * floor_ratio);
  return std::min(dt, dt_max);
}

struct Sim {
  // ---- mesh -------------------------------------------------------------
  int nn = 0, nt = 0;
  std::vector<double> xyz;
  std::vector<int32_t> tet, nb, fgroup, partner;
  std::vector<double> A, X0, vol;  // barycentric affine map per tet (FP64)
  std::vector<int32_t> kind;       // 0 free, 1 Dirichlet, 2 floating conductor
  std::vector<double> dval;        // Dirichlet values
  std::array<int, kMaxGroups> gtype{};
  std::array<std::array<double, 3>, kMaxGroups> gshift{};
  // ---- fields -----------------------------------------------------------
  std::vector<double> rho, phi, E, Bn;
  double Qcond = 0.0, Vfloat = 0.0;
  int has_cond = 0;
  // ---- Poisson ----------------------------------------------------------
  int op = 0;  // 0 = P1 FEM, 1 = reference graph Laplacian
  double eps = kEps0;
  int load_sign = +1;  // +1 physical, -1 reference convention
  Csr Kfull;           // unconstrained operator
  Csr Ared;            // reduced (Dirichlet-eliminated, conductor collapsed)
  std::vector<int32_t> dof;  // node -> dof (-1 Dirichlet)
  int ndof = 0, anchor = -1;
  std::vector<double> xred;  // warm start
  double pcg_tol = 1e-10;
  int pcg_max = 20000;
  PcgResult last{0, 0, 0};
  // LSQ gradient operator per node: neighbour list + weights (3 per edge)
  std::vector<int64_t> nrp;
  std::vector<int32_t> nci;
  std::vector<double> nw;
  // ---- particles ----------------------------------------------------------
  std::vector<Species> sp;
  double dt = 1e-12;
  int adaptive = 1;
  double floor_ratio = 1e-3;
  int max_sub = 100000;
  int max_walk = 256;
  // 0 = once at the end position (physics default), 1 = LinearSegmentCloud per
  // sub-step (reference default), 2 = NearestPoint per sub-step (reference)
  int deposit_mode = 0;
  int nseg = 10;
  std::array<Ledger, kMaxGroups> ledger{};
  double surface_charge = 0.0;
  int64_t lost = 0;
  int64_t steps = 0;
  // ---- collisions and surface emission ------------------------------------
  std::vector<MccProc> mcc;
  uint32_t key[2] = {0, 0};
  int64_t mcc_count[16] = {};
  int64_t mcc_created = 0, mcc_removed = 0, see_emitted = 0, photo_emitted = 0;
  // null-collision state: largest |v|^2 of the last push per species (from
  // the particles themselves after a load), the majorant of the coming step
  std::vector<double> v2max;
  bool v2_have = false;
  std::vector<std::array<double, 3>> mcc_ctl;  // P_null, v^2 cap, nu_max
  bool ctl_ready = false;
  int64_t mcc_candidates = 0, mcc_major = 0;
  std::array<SurfEmit, kMaxGroups> emit{};
  std::vector<SeeEvent> events;
};

// Barycentric coordinates through the affine inverse-Jacobian map, the fast
// path of Tools/Particle/src/PICParticleAlgorithms.cpp:216-244 (lambda_1..3 =
// J^-1 (P - P0), lambda_0 = 1 - sum).  Equal to the signed-volume path at
// :246-296 up to rounding.
inline void bary(const Sim& s, int t, const double* p, double* lam) {
  const double* a = &s.A[9 * (size_t)t];
  const double* o = &s.X0[3 * (size_t)t];
  const double d0 = p[0] - o[0], d1 = p[1] - o[1], d2 = p[2] - o[2];
  lam[1] = a[0] * d0 + a[1] * d1 + a[2] * d2;
  lam[2] = a[3] * d0 + a[4] * d1 + a[5] * d2;
  lam[3] = a[6] * d0 + a[7] * d1 + a[8] * d2;
  lam[0] = 1.0 - lam[1] - lam[2] - lam[3];
}

void build_geometry(Sim& s) {
  s.A.assign(9 * (size_t)s.nt, 0);
  s.X0.assign(3 * (size_t)s.nt, 0);
  s.vol.assign(s.nt, 0);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < s.nt; ++t) {
    const double* p[4];
    for (int k = 0; k < 4; ++k) p[k] = &s.xyz[3 * (size_t)s.tet[4 * (size_t)t + k]];
    double J[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) J[r][c] = p[c + 1][r] - p[0][r];
    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
    s.vol[t] = std::abs(det) / 6.0;
    double* a = &s.A[9 * (size_t)t];
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
    for (int r = 0; r < 3; ++r) s.X0[3 * (size_t)t + r] = p[0][r];
  }
}

// Face-neighbour table.  The reference declares Element neighbours
// (Tools/Mesh/include/MeshAlgorithms.h:193) but never fills them, so its
// relocation falls back to an O(N_elem) scan
// (Tools/Particle/src/PICParticleAlgorithms.cpp:882-888).  The oracle builds
// the table by matching sorted face vertex triples; face k is the face
// opposite local vertex k.
void build_neighbors(Sim& s) {
  struct F { int64_t a, b, c; int32_t t, k; };
  std::vector<F> f((size_t)s.nt * 4);
  for (int t = 0; t < s.nt; ++t)
    for (int k = 0; k < 4; ++k) {
      int v[3], m = 0;
      for (int j = 0; j < 4; ++j)
        if (j != k) v[m++] = s.tet[4 * (size_t)t + j];
      std::sort(v, v + 3);
      f[4 * (size_t)t + k] = {v[0], v[1], v[2], t, k};
    }
  std::sort(f.begin(), f.end(), [](const F& x, const F& y) {
    if (x.a != y.a) return x.a < y.a;
    if (x.b != y.b) return x.b < y.b;
    return x.c < y.c;
  });
  s.nb.assign(4 * (size_t)s.nt, -1);
  for (size_t i = 0; i + 1 < f.size(); ++i) {
    if (f[i].a == f[i + 1].a && f[i].b == f[i + 1].b && f[i].c == f[i + 1].c) {
      s.nb[4 * (size_t)f[i].t + f[i].k] = f[i + 1].t;
      s.nb[4 * (size_t)f[i + 1].t + f[i + 1].k] = f[i].t;
      ++i;
    }
  }
}

// Node adjacency (nodes sharing an element), as buildNodeNeighbors,
// Tools/FieldSolver/src/PoissonSolver.cpp:52-91.
void build_adjacency(const Sim& s, std::vector<int64_t>& rp, std::vector<int32_t>& ci,
                     bool with_self) {
  std::vector<std::vector<int32_t>> adj(s.nn);
  for (int t = 0; t < s.nt; ++t)
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j)
        if (i != j || with_self) adj[s.tet[4 * (size_t)t + i]].push_back(s.tet[4 * (size_t)t + j]);
  rp.assign(s.nn + 1, 0);
  for (int i = 0; i < s.nn; ++i) {
    auto& a = adj[i];
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
    rp[i + 1] = rp[i] + (int64_t)a.size();
  }
  ci.resize(rp[s.nn]);
  for (int i = 0; i < s.nn; ++i) std::copy(adj[i].begin(), adj[i].end(), ci.begin() + rp[i]);
}

// Operator assembly into CSR over the node adjacency pattern.
//  op 0: P1 FEM stiffness K_ij = eps * grad N_i . grad N_j * V_e,
//        Tools/FieldSolver/src/MultiBoundaryElectricFieldSolver.cpp:1704-1759
//        with gradients from :1870-1894.
//  op 1: reference PoissonSolver graph Laplacian, w = eps * (V_e/6) / d_ij^2,
//        Tools/FieldSolver/src/PoissonSolver.cpp:276-327.
void assemble(Sim& s) {
  build_adjacency(s, s.Kfull.rp, s.Kfull.ci, true);
  s.Kfull.n = s.nn;
  s.Kfull.va.assign(s.Kfull.ci.size(), 0.0);
  auto add = [&](int i, int j, double v) {
    auto b = s.Kfull.ci.begin() + s.Kfull.rp[i], e = s.Kfull.ci.begin() + s.Kfull.rp[i + 1];
    auto it = std::lower_bound(b, e, j);
    s.Kfull.va[it - s.Kfull.ci.begin()] += v;
  };
  for (int t = 0; t < s.nt; ++t) {
    const int32_t* nd = &s.tet[4 * (size_t)t];
    const double V = s.vol[t];
    if (s.op == 0) {
      const double* a = &s.A[9 * (size_t)t];
      double g[4][3];
      for (int r = 0; r < 3; ++r) {
        g[1][r] = a[0 + r];
        g[2][r] = a[3 + r];
        g[3][r] = a[6 + r];
        g[0][r] = -(g[1][r] + g[2][r] + g[3][r]);
      }
      for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j)
          add(nd[i], nd[j], s.eps * V * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]));
    } else {
      const double scale = V / 6.0;
      for (int i = 0; i < 4; ++i)
        for (int j = i + 1; j < 4; ++j) {
          const double* pi = &s.xyz[3 * (size_t)nd[i]];
          const double* pj = &s.xyz[3 * (size_t)nd[j]];
          double d2 = 0;
          for (int r = 0; r < 3; ++r) d2 += (pi[r] - pj[r]) * (pi[r] - pj[r]);
          const double d = std::max(std::sqrt(d2), 1e-12);
          const double w = s.eps * scale / (d * d);
          add(nd[i], nd[i], w);
          add(nd[j], nd[j], w);
          add(nd[i], nd[j], -w);
          add(nd[j], nd[i], -w);
        }
    }
  }
  // Reduced system: Dirichlet rows/cols eliminated (LinearSystem.cpp:37-58,
  // MultiBoundary :1229-1243) and all conductor nodes collapsed onto one
  // anchor dof by row+column summation (MultiBoundary :1172-1217).
  s.dof.assign(s.nn, -1);
  s.ndof = 0;
  s.anchor = -1;
  s.has_cond = 0;
  for (int i = 0; i < s.nn; ++i)
    if (s.kind[i] == 0) s.dof[i] = s.ndof++;
  for (int i = 0; i < s.nn; ++i)
    if (s.kind[i] == 2) {
      if (s.anchor < 0) s.anchor = s.ndof++;
      s.dof[i] = s.anchor;
      s.has_cond = 1;
    }
  std::vector<std::vector<std::pair<int32_t, double>>> rows(s.ndof);
  for (int i = 0; i < s.nn; ++i) {
    const int di = s.dof[i];
    if (di < 0) continue;
    for (int64_t k = s.Kfull.rp[i]; k < s.Kfull.rp[i + 1]; ++k) {
      const int dj = s.dof[s.Kfull.ci[k]];
      if (dj >= 0) rows[di].push_back({dj, s.Kfull.va[k]});
    }
  }
  s.Ared.n = s.ndof;
  s.Ared.rp.assign(s.ndof + 1, 0);
  s.Ared.ci.clear();
  s.Ared.va.clear();
  for (int d = 0; d < s.ndof; ++d) {
    auto& r = rows[d];
    std::sort(r.begin(), r.end(), [](auto& x, auto& y) { return x.first < y.first; });
    for (size_t k = 0; k < r.size(); ++k) {
      if (!s.Ared.ci.empty() && (int64_t)s.Ared.ci.size() > s.Ared.rp[d] &&
          s.Ared.ci.back() == r[k].first)
        s.Ared.va.back() += r[k].second;
      else {
        s.Ared.ci.push_back(r[k].first);
        s.Ared.va.push_back(r[k].second);
      }
    }
    s.Ared.rp[d + 1] = (int64_t)s.Ared.ci.size();
  }
  s.xred.assign(s.ndof, 0.0);
}

// LSQ nodal gradient operator: g_i = M_i^-1 sum_j d_ij (phi_j - phi_i),
// M_i = sum_j d_ij d_ij^T over edge neighbours, E_i = -g_i.
// Tools/FieldSolver/src/PoissonSolver.cpp:474-529 (neighbours, E = -grad)
// and :749-787 (normal equations) with solve3x3 pivoting at :124-170.
// The weights w_ij = M_i^-1 d_ij are precomputed once: the gradient is linear
// in phi, so this is the same operator applied as a sparse product.
bool solve3(double M[3][3], double r[3], double out[3]) {
  for (int p = 0; p < 3; ++p) {
    int best = p;
    for (int row = p + 1; row < 3; ++row)
      if (std::abs(M[row][p]) > std::abs(M[best][p])) best = row;
    if (std::abs(M[best][p]) < 1e-18) return false;
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

void build_lsq(Sim& s) {
  build_adjacency(s, s.nrp, s.nci, false);
  s.nw.assign(3 * s.nci.size(), 0.0);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < s.nn; ++i) {
    double M[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    const double* xi = &s.xyz[3 * (size_t)i];
    for (int64_t k = s.nrp[i]; k < s.nrp[i + 1]; ++k) {
      const double* xj = &s.xyz[3 * (size_t)s.nci[k]];
      const double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
      for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c) M[r][c] += d[r] * d[c];
    }
    for (int64_t k = s.nrp[i]; k < s.nrp[i + 1]; ++k) {
      const double* xj = &s.xyz[3 * (size_t)s.nci[k]];
      double d[3] = {xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]};
      double Mc[3][3];
      std::memcpy(Mc, M, sizeof(M));
      double w[3] = {0, 0, 0};
      if (!solve3(Mc, d, w)) w[0] = w[1] = w[2] = 0;
      for (int r = 0; r < 3; ++r) s.nw[3 * (size_t)k + r] = w[r];
    }
  }
}

void compute_efield(Sim& s) {
  s.E.assign(3 * (size_t)s.nn, 0.0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s.nn; ++i) {
    double g[3] = {0, 0, 0};
    for (int64_t k = s.nrp[i]; k < s.nrp[i + 1]; ++k) {
      const double dphi = s.phi[s.nci[k]] - s.phi[i];
      for (int r = 0; r < 3; ++r) g[r] += s.nw[3 * (size_t)k + r] * dphi;
    }
    for (int r = 0; r < 3; ++r) s.E[3 * (size_t)i + r] = -g[r];
  }
}

// Poisson solve on the reduced system.  Load vector: the deposited nodal
// charge (C).  The reference forms b_i = -rho_i * V_e / 4 summed over elements
// (PoissonSolver.cpp:335-357, MultiBoundary :1756); deposited P1 charge is
// already the consistent load sum_p q w N_i(x_p), so the oracle uses it
// directly with sign load_sign (+1 = physical -div(eps grad phi) = rho).
// Conductor anchor row gets + Q (MultiBoundary :1217, b_anchor += value).
PcgResult solve_poisson(Sim& s) {
  std::vector<double> b(s.ndof, 0.0);
  for (int i = 0; i < s.nn; ++i) {
    const int d = s.dof[i];
    if (d >= 0) b[d] += s.load_sign * s.rho[i];
  }
  if (s.anchor >= 0) b[s.anchor] += s.Qcond;
  // Dirichlet lifting: b_i -= K_ij g_j for constrained j.
  for (int i = 0; i < s.nn; ++i) {
    const int di = s.dof[i];
    if (di < 0) continue;
    for (int64_t k = s.Kfull.rp[i]; k < s.Kfull.rp[i + 1]; ++k) {
      const int j = s.Kfull.ci[k];
      if (s.kind[j] == 1 && s.dval[j] != 0.0) b[di] -= s.Kfull.va[k] * s.dval[j];
    }
  }
  PcgResult r = pcg(s.Ared, b, s.xred, s.pcg_tol, s.pcg_max);
  s.last = r;
  for (int i = 0; i < s.nn; ++i) {
    const int d = s.dof[i];
    s.phi[i] = d >= 0 ? s.xred[d] : s.dval[i];
  }
  if (s.anchor >= 0) s.Vfloat = s.xred[s.anchor];
  return r;
}

struct WalkOut { int cell; int dead; int group; int face; double hit[3]; };

// Straight-segment tet walk from x0 (inside `cell`) to x1.  Replaces
// CrossTetrahedron::dichotomyBoundaryIntersection + locateParticle
// (PICParticleAlgorithms.cpp:853-939) and PICCycle boundary detection
// (PICCycle.cpp:1074-1133,1221-1293): the exit face is the face whose
// barycentric coordinate reaches zero first along the segment.
//   absorb/conductor -> dead, hit point recorded
//   periodic -> positions shifted by the group's period, walk continues in the partner tet
//   reflect -> specular reflection about the face plane, walk continues
void walk(const Sim& s, double* x0, double* x1, double* v, int cell, WalkOut& o, double tol) {
  o.dead = 0;
  o.group = -1;
  int cur = cell;
  for (int it = 0; it < s.max_walk; ++it) {
    double l1[4];
    bary(s, cur, x1, l1);
    if (l1[0] >= -tol && l1[1] >= -tol && l1[2] >= -tol && l1[3] >= -tol) {
      o.cell = cur;
      return;
    }
    double l0[4];
    bary(s, cur, x0, l0);
    int kb = -1;
    double tb = 2.0;
    for (int k = 0; k < 4; ++k) {
      if (l1[k] < -tol && l1[k] < l0[k]) {
        const double t = l0[k] / (l0[k] - l1[k]);
        if (t < tb) { tb = t; kb = k; }
      }
    }
    if (kb < 0) {  // degenerate: take most negative coordinate
      double mn = 0;
      for (int k = 0; k < 4; ++k)
        if (l1[k] < mn) { mn = l1[k]; kb = k; }
      tb = 0;
    }
    tb = std::min(std::max(tb, 0.0), 1.0);
    const int n = s.nb[4 * (size_t)cur + kb];
    if (n >= 0) {
      cur = n;
      continue;
    }
    const int g = s.fgroup[4 * (size_t)cur + kb];
    double hit[3];
    for (int r = 0; r < 3; ++r) hit[r] = x0[r] + tb * (x1[r] - x0[r]);
    const int ty = (g >= 0 && g < kMaxGroups) ? s.gtype[g] : BC_ESCAPE;
    if (ty == BC_PERIODIC) {
      const int pt = s.partner[4 * (size_t)cur + kb];
      if (pt >= 0) {
        for (int r = 0; r < 3; ++r) {
          x0[r] = hit[r] + s.gshift[g][r];
          x1[r] += s.gshift[g][r];
        }
        cur = pt;
        continue;
      }
    } else if (ty == BC_REFLECT) {
      const double* a = &s.A[9 * (size_t)cur];
      double nrm[3];
      if (kb == 0) for (int r = 0; r < 3; ++r) nrm[r] = -(a[r] + a[3 + r] + a[6 + r]);
      else for (int r = 0; r < 3; ++r) nrm[r] = a[3 * (kb - 1) + r];
      const double nl = std::sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
      for (int r = 0; r < 3; ++r) nrm[r] /= nl;  // inward normal (grad lambda_k)
      const double dd = (x1[0] - hit[0]) * nrm[0] + (x1[1] - hit[1]) * nrm[1] + (x1[2] - hit[2]) * nrm[2];
      const double vn = v[0] * nrm[0] + v[1] * nrm[1] + v[2] * nrm[2];
      for (int r = 0; r < 3; ++r) {
        x1[r] -= 2.0 * dd * nrm[r];
        v[r] -= 2.0 * vn * nrm[r];
        x0[r] = hit[r];
      }
      continue;
    }
    o.dead = 1;
    o.group = g;
    o.face = kb;
    o.cell = cur;
    for (int r = 0; r < 3; ++r) o.hit[r] = hit[r];
    return;
  }
  o.dead = 2;  // walk limit: lost
  o.cell = cur;
}

void push_species(Sim& s, Species& p, std::vector<double>& rho_acc) {
  const int64_t n = (int64_t)p.cell.size();
  const double qm = p.q / p.m;
  const double qw = p.q * p.w;
  const double tol = 1e-10;
  const int nthr =
#ifdef _OPENMP
      omp_get_max_threads();
#else
      1;
#endif
  std::vector<std::vector<double>> rho_t(nthr);
  std::vector<std::array<Ledger, kMaxGroups>> led_t(nthr);
  std::vector<double> qc_t(nthr, 0.0);
  std::vector<int64_t> lost_t(nthr, 0);
  std::vector<std::vector<SeeEvent>> ev_t(nthr);
  const int si = (int)(&p - s.sp.data());
  // Thread-local charge buffers reduced per node afterwards, as the OpenMP
  // branch of PICCycle::pushParticles, Tools/PICcore/src/PICCycle.cpp:483-563.
#pragma omp parallel
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    auto& rl = rho_t[tid];
    rl.assign(s.nn, 0.0);
    auto& led = led_t[tid];
#pragma omp for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      int cell = p.cell[i];
      if (cell < 0) continue;
      double* x = &p.x[3 * i];
      double* v = &p.v[3 * i];
      const double v0[3] = {v[0], v[1], v[2]};
      double remaining = s.dt;
      int sub = 0;
      bool dead = false;
      // Sub-step loop, PICParticleAlgorithms.cpp:593-701.
      while (remaining > 1e-20 && sub < s.max_sub) {
        ++sub;
        double lam[4];
        bary(s, cell, x, lam);
        const int32_t* nd = &s.tet[4 * (size_t)cell];
        double Ep[3] = {0, 0, 0}, Bp[3] = {0, 0, 0};
        for (int k = 0; k < 4; ++k)
          for (int r = 0; r < 3; ++r) {
            Ep[r] += lam[k] * s.E[3 * (size_t)nd[k] + r];
            Bp[r] += lam[k] * s.Bn[3 * (size_t)nd[k] + r];
          }
        const double h = s.adaptive ? adaptive_dt(v0, Ep, Bp, qm, remaining, s.floor_ratio) : remaining;
        double xs[3] = {x[0], x[1], x[2]};
        boris(v, Ep, Bp, qm, h);
        double xn[3] = {x[0] + v[0] * h, x[1] + v[1] * h, x[2] + v[2] * h};
        WalkOut w{};
        walk(s, xs, xn, v, cell, w, tol);
        if (w.dead) {
          dead = true;
          if (w.dead == 2) {
            lost_t[tid]++;
          } else {
            const int g = w.group;
            const double ke = 0.5 * p.m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            if (g >= 0 && g < kMaxGroups && s.gtype[g] != BC_ESCAPE) {
              // Ledger, PICCycle::recordSurfaceBoundaryInteraction :1295-1329.
              Ledger& L = led[g];
              L.energy += ke * std::max(1.0, p.w);
              if (p.q < 0) { L.abs_e_charge += qw; L.n_abs_e++; }
              else { L.abs_i_charge += qw; L.n_abs_i++; }
              if (s.gtype[g] == BC_CONDUCTOR) qc_t[tid] += qw;
              // secondary emission is evaluated after the push from the
              // recorded impact (electrons on absorbing or conductor faces)
              if (p.q < 0 && s.emit[g].see_model != 0 &&
                  (s.gtype[g] == BC_CONDUCTOR || s.gtype[g] == BC_ABSORB)) {
                SeeEvent e{si, p.id.empty() ? (int)i : p.id[i], w.cell, g, w.face, {}, {}};
                for (int r = 0; r < 3; ++r) { e.x[r] = w.hit[r]; e.v[r] = v[r]; }
                ev_t[tid].push_back(e);
              }
            }
            for (int r = 0; r < 3; ++r) x[r] = w.hit[r];
          }
          break;
        }
        cell = w.cell;
        if (s.deposit_mode == 2) {
          // NearestPoint: full charge at the sub-step's end position in the
          // final element, every sub-step (PICParticleAlgorithms.cpp:388-392,
          // called from the sub-step loop at :686-692).
          double l[4];
          bary(s, cell, xn, l);
          bool in = true;
          for (int k = 0; k < 4; ++k) in = in && l[k] >= -1e-8 && l[k] <= 1 + 1e-8;
          if (in)
            for (int k = 0; k < 4; ++k) rl[s.tet[4 * (size_t)cell + k]] += qw * l[k];
        }
        if (s.deposit_mode == 1) {
          // LinearSegmentCloud: charge/nseg at each segment midpoint, into the
          // final element only, samples outside dropped
          // (PICParticleAlgorithms.cpp:340-343,423-432).
          for (int sg = 0; sg < s.nseg; ++sg) {
            const double t = (sg + 0.5) / s.nseg;
            double pt[3] = {xs[0] + (xn[0] - xs[0]) * t, xs[1] + (xn[1] - xs[1]) * t,
                            xs[2] + (xn[2] - xs[2]) * t};
            double l[4];
            bary(s, cell, pt, l);
            bool in = true;
            for (int k = 0; k < 4; ++k) in = in && l[k] >= -1e-8 && l[k] <= 1 + 1e-8;
            if (!in) continue;
            for (int k = 0; k < 4; ++k) rl[s.tet[4 * (size_t)cell + k]] += (qw / s.nseg) * l[k];
          }
        }
        for (int r = 0; r < 3; ++r) x[r] = xn[r];
        remaining -= h;
      }
      if (dead) {
        p.cell[i] = -1;
        continue;
      }
      p.cell[i] = cell;
      if (s.deposit_mode == 0) {
        double lam[4];
        bary(s, cell, x, lam);
        for (int k = 0; k < 4; ++k) rl[s.tet[4 * (size_t)cell + k]] += qw * lam[k];
      }
    }
  }
  for (int t = 0; t < nthr; ++t) {
    if (rho_t[t].empty()) continue;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < s.nn; ++i) rho_acc[i] += rho_t[t][i];
    for (int g = 0; g < kMaxGroups; ++g) {
      Ledger& L = s.ledger[g];
      const Ledger& M = led_t[t][g];
      L.abs_e_charge += M.abs_e_charge; L.abs_i_charge += M.abs_i_charge;
      L.emit_e_charge += M.emit_e_charge; L.energy += M.energy;
      L.n_abs_e += M.n_abs_e; L.n_abs_i += M.n_abs_i; L.n_emit_e += M.n_emit_e;
      s.surface_charge += M.abs_e_charge + M.abs_i_charge;
    }
    s.Qcond += qc_t[t];
    s.lost += lost_t[t];
    s.events.insert(s.events.end(), ev_t[t].begin(), ev_t[t].end());
  }
}


// =========================================================================
// Monte Carlo collisions and surface emission
// =========================================================================
constexpr double kPi = 3.14159265358979323846;

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC'11): ten rounds of the two 32x32->64 multiply-xor
// bijection with the Weyl key schedule.
void philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c0 = n0; c1 = (uint32_t)p1; c2 = n2; c3 = (uint32_t)p0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}
inline double u01(uint32_t x) { return ((double)x + 0.5) * 2.3283064365386963e-10; }
inline void draw(const Sim& s, uint32_t a, uint32_t b, uint32_t c, uint32_t tag, uint32_t out[4]) {
  const uint32_t ctr[4] = {a, b, c, tag};
  philox(ctr, s.key, out);
}
// stream tags (fourth counter word)
constexpr uint32_t kTagMccA = 0x4D43, kTagMccB = 0x4D44, kTagMccId = 0x4D45, kTagSee = 0x5345,
                   kTagSeeK = 0x5346, kTagSeeId = 0x5347, kTagPhoto = 0x5048, kTagPhotoK = 0x5049,
                   kTagPhotoK2 = 0x504A;
constexpr int kMaxEmit = 8;
constexpr double kMe = 9.1093837015e-31;  // Tools/Basic/include/Constants.h:47

// Cross sections in m^2 of energy in eV,
// Tools/Interactions/Collisions/src/CollisonTypes.cpp:56-176.
double sigma_of(const MccProc& p, double E) {
  switch (p.type) {
    case MCC_ELASTIC:
    case MCC_INELASTIC:
    case MCC_CHARGE_EXCHANGE: {  // ElasticCollisionCrossSection :56-77
      if (E < p.thr) return 0.0;
      const double sg = p.smax * std::pow(std::max(1e-6, p.eref) / E, p.pidx);
      return std::max(0.0, std::min(sg, p.smax));
    }
    case MCC_IONIZATION: {  // :79-93
      if (E < p.thr || E <= p.thr) return 0.0;
      const double I = std::max(1e-6, p.thr);
      const double sg = p.smax * std::log(E / I) * (1.0 - I / E) / E;
      return std::max(0.0, std::min(sg, p.smax));
    }
    case MCC_EXCITATION: {  // :95-115
      if (E < p.thr || E <= p.thr) return 0.0;
      const double x = E / std::max(1e-6, p.thr);
      const double sg = p.smax * std::pow(std::max(0.0, 1.0 - 1.0 / x), p.alpha) * std::pow(1.0 / x, p.beta);
      return std::max(0.0, std::min(sg, p.smax));
    }
    case MCC_RECOMBINATION: {  // :117-132
      const double er = std::max(1e-6, p.eref);
      const double sg = p.smax * std::pow(er / std::max(er, E), p.pidx);
      return std::max(0.0, std::min(sg, p.smax));
    }
    default: {  // MCC_COULOMB :160-176
      if (E <= 0.0) return 0.0;
      const double Ej = E * kE;
      const double k = 1.0 / (4.0 * kPi * kEps0);
      const double ke2 = k * kE * kE;
      const double sg = kPi * p.clog * ke2 * ke2 / std::max(1e-30, 4.0 * Ej * Ej);
      return std::max(0.0, std::min(sg, p.smax));
    }
  }
}

inline void iso_dir(double u2, double u3, double* d) {
  const double ct = 2.0 * u2 - 1.0;
  const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
  const double ph = 2.0 * kPi * u3;
  d[0] = st * std::cos(ph);
  d[1] = st * std::sin(ph);
  d[2] = ct;
}

void append(Species& t, const double* x, const double* v, int cell, int32_t id) {
  for (int r = 0; r < 3; ++r) { t.x.push_back(x[r]); t.v.push_back(v[r]); }
  t.cell.push_back(cell);
  t.id.push_back(id);
}

void deposit_point(Sim& s, int cell, const double* x, double qw) {
  double lam[4];
  bary(s, cell, x, lam);
  for (int k = 0; k < 4; ++k) s.rho[s.tet[4 * (size_t)cell + k]] += qw * lam[k];
}

// Null-collision majorant (Vahedi & Surendra, Comput. Phys. Commun. 87
// (1995) 179) of species si's processes: nu_max >= sum_p n_p sigma_p(E)
// sqrt(2E/m) for E up to the energy of the cap v^2 = 4 v2max, sampled at
// 1024 log-spaced energies over [1e-8, 1] x cap, just above each threshold,
// and bounded below the lowest sample by sum n sigma_max v (sigma <= sigma_max
// in every form); times 1.05.  Returns {P_null, v^2 cap, nu_max}.
constexpr int kMajorantPoints = 1024;
double nu_total(const Sim& s, const std::vector<int>& pr, double E, double m) {
  double t = 0.0;
  for (int j : pr) t += s.mcc[j].n * sigma_of(s.mcc[j], E);
  return t * std::sqrt(2.0 * E * kE / m);
}
std::array<double, 3> majorant(const Sim& s, const std::vector<int>& pr, double m, double v2max) {
  const double v2cap = 4.0 * v2max;
  const double ecap = 0.5 * m * v2cap / kE;
  double f = 0.0;
  if (ecap > 0.0) {
    const double elo = 1e-8 * ecap, la = std::log(elo), lb = std::log(ecap);
    for (int k = 0; k < kMajorantPoints; ++k)
      f = std::max(f, nu_total(s, pr, std::exp(la + (lb - la) * (double)k / (double)(kMajorantPoints - 1)), m));
    for (int j : pr) {
      const double thr = s.mcc[j].thr;
      if (thr > elo && thr < ecap) f = std::max(f, nu_total(s, pr, thr * (1.0 + 1e-9), m));
    }
    double b = 0.0;
    for (int j : pr) b += s.mcc[j].n * s.mcc[j].smax;
    f = std::max(f, b * std::sqrt(2.0 * elo * kE / m));
  }
  const double numax = 1.05 * f;
  return {-std::expm1(-numax * s.dt), v2cap, numax};
}
std::vector<int> procs_of(const Sim& s, size_t si) {
  std::vector<int> pr;
  for (size_t k = 0; k < s.mcc.size(); ++k)
    if (s.mcc[k].species == (int)si) pr.push_back((int)k);
  return pr;
}
// largest |v|^2 of the alive particles of each species
void measure_v2max(Sim& s) {
  s.v2max.assign(s.sp.size(), 0.0);
  for (size_t si = 0; si < s.sp.size(); ++si) {
    const Species& P = s.sp[si];
    for (size_t i = 0; i < P.cell.size(); ++i) {
      if (P.cell[i] < 0) continue;
      const double* v = &P.v[3 * i];
      s.v2max[si] = std::max(s.v2max[si], v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
  }
  s.v2_have = true;
}
// The coming step's majorants (before its push).
void mcc_prepare(Sim& s) {
  if (!s.v2_have || s.v2max.size() != s.sp.size()) measure_v2max(s);
  s.mcc_ctl.assign(s.sp.size(), {0.0, 0.0, 0.0});
  for (size_t si = 0; si < s.sp.size(); ++si) {
    const std::vector<int> pr = procs_of(s, si);
    if (!pr.empty()) s.mcc_ctl[si] = majorant(s, pr, s.sp[si].m, s.v2max[si]);
  }
  s.ctl_ready = true;
}

// One background-collision step over the particles present before it
// (replaces MonteCarloCollisionHandler::processCollisions,
// CollisionAlgorithm.cpp:500-538, inside executeBackgroundMccStep,
// CollisionPicAdapter.cpp:338-434).  Per particle: nu_p = n_p sigma_p(E) |v|,
// one collision with probability 1 - exp(-sum nu dt), process chosen with
// probability nu_p / sum nu (null-collision thinning with the particle's own
// total frequency as the majorant).  The reference's second draw inside
// executeCollision (effective probability p^2) is not reproduced.  Process
// outcomes follow CollisionAlgorithm.cpp:189-460: isotropic scatter
// (elastic, Coulomb), energy loss (inelastic, excitation above threshold),
// ionisation with a uniform split of E - I and an ion at rest, removal
// (charge exchange turns the ion neutral, recombination the electron).
// Candidates by the null-collision method: a particle is a candidate when
// its first uniform is below P_null of the step's majorant, or when it is
// faster than the cap; a candidate collides when u2 P_c < 1 - exp(-nu dt)
// (P_c = P_null, or 1 above the cap), so the collision probability stays
// exactly 1 - exp(-nu dt), and R = (u2 P_c / p) nu picks the process.
void mcc_step(Sim& s) {
  if (s.mcc.empty()) return;
  if (!s.ctl_ready) {  // collisions without a push (checkers): majorant of the present particles
    measure_v2max(s);
    mcc_prepare(s);
  }
  s.ctl_ready = false;
  const uint32_t step = (uint32_t)s.steps;
  std::vector<size_t> n0(s.sp.size());
  for (size_t k = 0; k < s.sp.size(); ++k) n0[k] = s.sp[k].cell.size();
  for (size_t si = 0; si < s.sp.size(); ++si) {
    std::vector<int> pr;
    for (size_t k = 0; k < s.mcc.size(); ++k)
      if (s.mcc[k].species == (int)si) pr.push_back((int)k);
    if (pr.empty()) continue;
    const double m = s.sp[si].m;
    const std::array<double, 3> ctl = s.mcc_ctl[si];
    for (size_t i = 0; i < n0[si]; ++i) {
      Species& P = s.sp[si];
      if (P.cell[i] < 0) continue;
      double v[3] = {P.v[3 * i], P.v[3 * i + 1], P.v[3 * i + 2]};
      const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      const uint32_t id = (uint32_t)P.id[i];
      uint32_t a[4];
      draw(s, id, step, (uint32_t)si, kTagMccA, a);
      const bool forced = v2 > ctl[1];
      if (!forced && !(u01(a[0]) < ctl[0])) continue;
      s.mcc_candidates++;
      const double speed = std::sqrt(v2);
      const double E = 0.5 * m * v2 / kE;
      double nu[16], tot = 0.0;
      for (size_t j = 0; j < pr.size(); ++j) {
        nu[j] = s.mcc[pr[j]].n * sigma_of(s.mcc[pr[j]], E) * speed;
        tot += nu[j];
      }
      if (!(tot > 0.0)) continue;
      const double pc = forced ? 1.0 : ctl[0];
      const double p = -std::expm1(-tot * s.dt);
      if (p > pc) s.mcc_major++;
      const double u = u01(a[1]) * pc;
      if (!(u < p)) continue;
      const double R = (u / p) * tot;
      int j = -1;
      double cum = 0.0;
      for (size_t q = 0; q < pr.size(); ++q) {
        cum += nu[q];
        if (nu[q] > 0.0) j = (int)q;
        if (R < cum) break;
      }
      const MccProc& q = s.mcc[pr[j]];
      s.mcc_count[pr[j]]++;
      double d[3];
      switch (q.type) {
        case MCC_ELASTIC:
        case MCC_COULOMB:
          iso_dir(u01(a[2]), u01(a[3]), d);
          for (int r = 0; r < 3; ++r) P.v[3 * i + r] = speed * d[r];
          break;
        case MCC_INELASTIC:
        case MCC_EXCITATION: {
          if (q.type == MCC_EXCITATION && E <= q.thr) break;
          const double loss = q.type == MCC_INELASTIC ? std::max(0.0, q.loss) : (q.loss > 0.0 ? q.loss : q.thr);
          const double sp = std::sqrt(2.0 * std::max(1e-6, E - loss) * kE / m);
          iso_dir(u01(a[2]), u01(a[3]), d);
          for (int r = 0; r < 3; ++r) P.v[3 * i + r] = sp * d[r];
          break;
        }
        case MCC_IONIZATION: {
          if (E <= q.thr) break;
          uint32_t b[4], c[4];
          draw(s, id, step, (uint32_t)si, kTagMccB, b);
          draw(s, id, step, (uint32_t)si, kTagMccId, c);
          const double avail = E - q.thr;
          const double sec = u01(b[0]) * avail;
          const double prim = avail - sec;
          const double sp = std::sqrt(2.0 * prim * kE / m);
          iso_dir(u01(a[2]), u01(a[3]), d);
          for (int r = 0; r < 3; ++r) P.v[3 * i + r] = sp * d[r];
          const double xp[3] = {P.x[3 * i], P.x[3 * i + 1], P.x[3 * i + 2]};
          const int cell = P.cell[i];
          Species& Se = s.sp[q.e_species];
          const double se = std::sqrt(2.0 * sec * kE / Se.m);
          iso_dir(u01(b[1]), u01(b[2]), d);
          const double ve[3] = {se * d[0], se * d[1], se * d[2]};
          append(Se, xp, ve, cell, (int32_t)(0x80000000u | (c[0] & 0x7FFFFFFFu)));
          deposit_point(s, cell, xp, Se.q * Se.w);
          Species& Si = s.sp[q.i_species];
          const double vi[3] = {0.0, 0.0, 0.0};
          append(Si, xp, vi, cell, (int32_t)(0x80000000u | (c[1] & 0x7FFFFFFFu)));
          deposit_point(s, cell, xp, Si.q * Si.w);
          s.mcc_created += 2;
          break;
        }
        default:  // charge exchange, recombination
          P.cell[i] = -1;
          s.mcc_removed++;
          break;
      }
    }
  }
}

// Secondary-electron yield.  Model 1: secondaryYield of
// Tools/Material/src/SurfaceMaterialModel.cpp:163-170
// (delta_max x e^{1-x} / sqrt(max(0.1, cos theta)), x = E / max(1, E_max)).
// Model 2: Sternglass, delta_max (E/E_max) exp(2 - 2 sqrt(E/E_max)).
double see_yield(int model, double dmax, double emax, double E, double cos_inc) {
  if (model == 1) {
    const double x = std::max(1e-6, E / std::max(1.0, emax));
    return std::max(0.0, dmax * x * std::exp(1.0 - x) / std::sqrt(std::max(0.1, cos_inc)));
  }
  if (model == 2) {
    if (!(E > 0.0)) return 0.0;
    const double r = E / emax;
    return dmax * r * std::exp(2.0 - 2.0 * std::sqrt(r));
  }
  return 0.0;
}

// Photoelectron current density, SurfaceMaterialModel.cpp:199-208.
double photo_current(double j0y, double t_ph, double v_ref, double V) {
  return std::max(0.0, j0y) * std::exp(-std::max(0.0, V - v_ref) / std::max(1e-3, t_ph));
}

// Inward unit normal of face k of tet t (gradient of lambda_k).
void inward_normal(const Sim& s, int t, int k, double* n) {
  const double* a = &s.A[9 * (size_t)t];
  if (k == 0) for (int r = 0; r < 3; ++r) n[r] = -(a[r] + a[3 + r] + a[6 + r]);
  else for (int r = 0; r < 3; ++r) n[r] = a[3 * (k - 1) + r];
  const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
  for (int r = 0; r < 3; ++r) n[r] /= l;
}

// Emission point: the surface point moved 1e-6 of the way to the tet centroid.
void nudge(const Sim& s, int t, const double* p, double* x) {
  for (int r = 0; r < 3; ++r) {
    double c = 0.0;
    for (int k = 0; k < 4; ++k) c += s.xyz[3 * (size_t)s.tet[4 * (size_t)t + k] + r];
    x[r] = p[r] + 1e-6 * (0.25 * c - p[r]);
  }
}

// Emitted velocity: flux-weighted Maxwellian of temperature T (eV), energy
// -T ln(u_a u_b) (Gamma(2, T)), cos(theta) = sqrt(u_c) about the inward
// normal, azimuth 2 pi u_d in the tangent frame (t1 = n x a / |n x a|,
// a = x-axis unless |n_x| >= 0.9, then y-axis; t2 = n x t1).
void emit_velocity(const double* n, double T, double m, double ua, double ub, double uc, double ud,
                   double* v) {
  const double E = -T * (std::log(ua) + std::log(ub));
  const double sp = std::sqrt(2.0 * E * kE / m);
  const double ax[3] = {std::abs(n[0]) < 0.9 ? 1.0 : 0.0, std::abs(n[0]) < 0.9 ? 0.0 : 1.0, 0.0};
  double t1[3] = {n[1] * ax[2] - n[2] * ax[1], n[2] * ax[0] - n[0] * ax[2], n[0] * ax[1] - n[1] * ax[0]};
  const double l = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
  for (int r = 0; r < 3; ++r) t1[r] /= l;
  const double t2[3] = {n[1] * t1[2] - n[2] * t1[1], n[2] * t1[0] - n[0] * t1[2], n[0] * t1[1] - n[1] * t1[0]};
  const double ct = std::sqrt(uc), st = std::sqrt(1.0 - uc), ph = 2.0 * kPi * ud;
  const double cp = std::cos(ph), sn = std::sin(ph);
  for (int r = 0; r < 3; ++r) v[r] = sp * (st * cp * t1[r] + st * sn * t2[r] + ct * n[r]);
}

void record_emission(Sim& s, int g, int count, const Species& T) {
  Ledger& L = s.ledger[g];
  L.n_emit_e += count;
  L.emit_e_charge += count * T.q * T.w;
  if (s.gtype[g] == BC_CONDUCTOR) s.Qcond -= count * T.q * T.w;
}

// Secondary emission from the impacts recorded by the last push.
void see_emit(Sim& s) {
  const uint32_t step = (uint32_t)s.steps;
  for (const SeeEvent& e : s.events) {
    const SurfEmit& cf = s.emit[e.group];
    const Species& P = s.sp[e.species];
    double n[3];
    inward_normal(s, e.cell, e.face, n);
    const double v2 = e.v[0] * e.v[0] + e.v[1] * e.v[1] + e.v[2] * e.v[2];
    const double sp = std::sqrt(v2);
    const double E = 0.5 * P.m * v2 / kE;
    const double cosi = sp > 0.0 ? std::abs(e.v[0] * n[0] + e.v[1] * n[1] + e.v[2] * n[2]) / sp : 1.0;
    const double dl = see_yield(cf.see_model, cf.dmax, cf.emax, E, cosi);
    uint32_t a[4];
    draw(s, (uint32_t)e.id, step, (uint32_t)e.species, kTagSee, a);
    const double fl = std::floor(dl);
    int nem = (int)std::min(fl, (double)kMaxEmit) + (u01(a[0]) < dl - fl ? 1 : 0);
    nem = std::min(nem, kMaxEmit);
    if (nem <= 0) continue;
    double xe[3];
    nudge(s, e.cell, e.x, xe);
    Species& T = s.sp[cf.e_species];
    for (int k = 0; k < nem; ++k) {
      const uint32_t sk = ((uint32_t)e.species << 8) | (uint32_t)k;
      uint32_t b[4], c[4];
      draw(s, (uint32_t)e.id, step, sk, kTagSeeK, b);
      draw(s, (uint32_t)e.id, step, sk, kTagSeeId, c);
      double v[3];
      emit_velocity(n, cf.t_see, T.m, u01(b[0]), u01(b[1]), u01(b[2]), u01(b[3]), v);
      append(T, xe, v, e.cell, (int32_t)(0x80000000u | (c[0] & 0x7FFFFFFFu)));
      deposit_point(s, e.cell, xe, T.q * T.w);
    }
    record_emission(s, e.group, nem, T);
    s.see_emitted += nem;
  }
  s.events.clear();
}

// Boundary faces of group g in (tet, face) order with cumulative areas.
struct FaceList { std::vector<int32_t> tet, face; std::vector<double> cum; };
FaceList group_faces(const Sim& s, int g) {
  FaceList f;
  double acc = 0.0;
  for (int t = 0; t < s.nt; ++t)
    for (int k = 0; k < 4; ++k) {
      if (s.fgroup[4 * (size_t)t + k] != g || s.nb[4 * (size_t)t + k] >= 0) continue;
      const double* p[3];
      int m = 0;
      for (int q = 0; q < 4; ++q)
        if (q != k) p[m++] = &s.xyz[3 * (size_t)s.tet[4 * (size_t)t + q]];
      const double u[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
      const double w[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
      const double c[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
      acc += 0.5 * std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
      f.tet.push_back(t);
      f.face.push_back(k);
      f.cum.push_back(acc);
    }
  return f;
}

// Photoemission: per group and step, J A dt / (|q| w) macro-electrons
// (integer part plus one with the fractional probability), uniform over the
// group's faces, J from photo_current at the group's potential (the floating
// potential of the last solve on conductor faces, 0 V elsewhere).
void photo_emit(Sim& s) {
  const uint32_t step = (uint32_t)s.steps;
  for (int g = 0; g < kMaxGroups; ++g) {
    const SurfEmit& cf = s.emit[g];
    if (!(cf.photo_j > 0.0)) continue;
    const FaceList f = group_faces(s, g);
    if (f.cum.empty()) continue;
    const double area = f.cum.back();
    Species& T = s.sp[cf.e_species];
    const double V = s.gtype[g] == BC_CONDUCTOR ? s.Vfloat : 0.0;
    const double expect = photo_current(cf.photo_j, cf.t_ph, cf.v_ref, V) * area * s.dt /
                          (std::abs(T.q) * T.w);
    uint32_t a[4];
    draw(s, (uint32_t)g, step, 0u, kTagPhoto, a);
    const double fl = std::floor(expect);
    const int64_t nem = (int64_t)fl + (u01(a[0]) < expect - fl ? 1 : 0);
    for (int64_t k = 0; k < nem; ++k) {
      uint32_t b[4], c[4];
      draw(s, (uint32_t)k, step, (uint32_t)g, kTagPhotoK, b);
      draw(s, (uint32_t)k, step, (uint32_t)g, kTagPhotoK2, c);
      const double r = u01(b[0]) * area;
      size_t fi = (size_t)(std::upper_bound(f.cum.begin(), f.cum.end(), r) - f.cum.begin());
      if (fi >= f.cum.size()) fi = f.cum.size() - 1;
      const int t = f.tet[fi], kf = f.face[fi];
      const double* p[3];
      int m = 0;
      for (int q = 0; q < 4; ++q)
        if (q != kf) p[m++] = &s.xyz[3 * (size_t)s.tet[4 * (size_t)t + q]];
      const double sq = std::sqrt(u01(b[1])), u2 = u01(b[2]);
      double pt[3], xe[3], n[3], v[3];
      for (int q = 0; q < 3; ++q) pt[q] = (1.0 - sq) * p[0][q] + sq * (1.0 - u2) * p[1][q] + sq * u2 * p[2][q];
      nudge(s, t, pt, xe);
      inward_normal(s, t, kf, n);
      emit_velocity(n, cf.t_ph, T.m, u01(b[3]), u01(c[0]), u01(c[1]), u01(c[2]), v);
      append(T, xe, v, t, (int32_t)(0x80000000u | (c[3] & 0x7FFFFFFFu)));
      deposit_point(s, t, xe, T.q * T.w);
    }
    record_emission(s, g, (int)nem, T);
    s.photo_emitted += nem;
  }
}

void post_push(Sim& s) {
  mcc_step(s);
  see_emit(s);
  photo_emit(s);
}

}  // namespace

// =========================================================================
// C interface (ctypes).  All arrays are caller-owned; sizes in elements.
// =========================================================================
extern "C" {

void* orc_create(int nn, const double* xyz, int nt, const int32_t* tet, const int32_t* kind,
                 const double* dval, const int32_t* fgroup) {
  auto* s = new Sim();
  s->nn = nn;
  s->nt = nt;
  s->xyz.assign(xyz, xyz + 3 * (size_t)nn);
  s->tet.assign(tet, tet + 4 * (size_t)nt);
  s->kind.assign(kind, kind + nn);
  s->dval.assign(nn, 0.0);
  if (dval) s->dval.assign(dval, dval + nn);
  s->fgroup.assign(4 * (size_t)nt, -1);
  if (fgroup) s->fgroup.assign(fgroup, fgroup + 4 * (size_t)nt);
  s->partner.assign(4 * (size_t)nt, -1);
  build_geometry(*s);
  build_neighbors(*s);
  s->rho.assign(nn, 0.0);
  s->phi.assign(nn, 0.0);
  s->E.assign(3 * (size_t)nn, 0.0);
  s->Bn.assign(3 * (size_t)nn, 0.0);
  for (int i = 0; i < nn; ++i) s->phi[i] = s->kind[i] == 1 ? s->dval[i] : 0.0;
  // no condition set: escape (PICCycle.cpp:1254-1257)
  for (int g = 0; g < kMaxGroups; ++g) s->gtype[g] = BC_ESCAPE;
  return s;
}

void orc_destroy(void* h) { delete static_cast<Sim*>(h); }

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

// op: 0 FEM, 1 graph Laplacian; eps: permittivity (F/m); load_sign +1/-1
void orc_set_poisson(void* h, int op, double eps, int load_sign, double tol, int max_it) {
  auto* s = static_cast<Sim*>(h);
  s->op = op;
  s->eps = eps;
  s->load_sign = load_sign;
  s->pcg_tol = tol;
  s->pcg_max = max_it;
  assemble(*s);
  build_lsq(*s);
}

void orc_set_push(void* h, double dt, int adaptive, double floor_ratio, int max_sub,
                  int deposit_mode, int nseg) {
  auto* s = static_cast<Sim*>(h);
  s->dt = dt;
  s->adaptive = adaptive;
  s->floor_ratio = floor_ratio;
  s->max_sub = max_sub;
  s->deposit_mode = deposit_mode;
  s->nseg = nseg;
}

void orc_set_group(void* h, int g, int type, const double* shift) {
  auto* s = static_cast<Sim*>(h);
  if (g < 0 || g >= kMaxGroups) return;
  s->gtype[g] = type;
  for (int r = 0; r < 3; ++r) s->gshift[g][r] = shift ? shift[r] : 0.0;
}

// partner: 4 per tet, partner tet across a periodic face or -1
void orc_set_partner(void* h, const int32_t* partner) {
  auto* s = static_cast<Sim*>(h);
  s->partner.assign(partner, partner + 4 * (size_t)s->nt);
}

void orc_get_neighbors(void* h, int32_t* nb) {
  auto* s = static_cast<Sim*>(h);
  std::memcpy(nb, s->nb.data(), sizeof(int32_t) * 4 * (size_t)s->nt);
}

void orc_get_geometry(void* h, double* A, double* X0, double* vol) {
  auto* s = static_cast<Sim*>(h);
  if (A) std::memcpy(A, s->A.data(), sizeof(double) * 9 * (size_t)s->nt);
  if (X0) std::memcpy(X0, s->X0.data(), sizeof(double) * 3 * (size_t)s->nt);
  if (vol) std::memcpy(vol, s->vol.data(), sizeof(double) * (size_t)s->nt);
}

int64_t orc_kfull_nnz(void* h) { return (int64_t)static_cast<Sim*>(h)->Kfull.ci.size(); }
void orc_get_kfull(void* h, int64_t* rp, int32_t* ci, double* va) {
  auto* s = static_cast<Sim*>(h);
  std::memcpy(rp, s->Kfull.rp.data(), sizeof(int64_t) * (s->nn + 1));
  std::memcpy(ci, s->Kfull.ci.data(), sizeof(int32_t) * s->Kfull.ci.size());
  std::memcpy(va, s->Kfull.va.data(), sizeof(double) * s->Kfull.va.size());
}

int orc_add_species(void* h, double q, double m, double w) {
  auto* s = static_cast<Sim*>(h);
  Species p;
  p.q = q; p.m = m; p.w = w;
  s->sp.push_back(std::move(p));
  return (int)s->sp.size() - 1;
}

// cell may be -1 for particles to be located by scanning (setup only).
void orc_set_particles(void* h, int si, int64_t n, const double* x, const double* v,
                       const int32_t* cell) {
  auto* s = static_cast<Sim*>(h);
  Species& p = s->sp[si];
  p.x.assign(x, x + 3 * n);
  p.v.assign(v, v + 3 * n);
  p.cell.assign(cell, cell + n);
  p.id.resize(n);
  for (int64_t i = 0; i < n; ++i) p.id[i] = (int32_t)i;
  s->v2_have = false;
}

int64_t orc_get_particles(void* h, int si, double* x, double* v, int32_t* cell) {
  auto* s = static_cast<Sim*>(h);
  Species& p = s->sp[si];
  const int64_t n = (int64_t)p.cell.size();
  if (x) std::memcpy(x, p.x.data(), sizeof(double) * 3 * n);
  if (v) std::memcpy(v, p.v.data(), sizeof(double) * 3 * n);
  if (cell) std::memcpy(cell, p.cell.data(), sizeof(int32_t) * n);
  return n;
}

void orc_set_efield(void* h, const double* E) {
  auto* s = static_cast<Sim*>(h);
  s->E.assign(E, E + 3 * (size_t)s->nn);
}
void orc_set_bfield(void* h, const double* B) {
  auto* s = static_cast<Sim*>(h);
  s->Bn.assign(B, B + 3 * (size_t)s->nn);
}
void orc_set_rho(void* h, const double* rho) {
  auto* s = static_cast<Sim*>(h);
  s->rho.assign(rho, rho + s->nn);
}
void orc_set_phi(void* h, const double* phi) {
  auto* s = static_cast<Sim*>(h);
  s->phi.assign(phi, phi + s->nn);
}
void orc_set_conductor_charge(void* h, double q) { static_cast<Sim*>(h)->Qcond = q; }
double orc_conductor_charge(void* h) { return static_cast<Sim*>(h)->Qcond; }
double orc_floating_potential(void* h) { return static_cast<Sim*>(h)->Vfloat; }

void orc_get_fields(void* h, double* rho, double* phi, double* E) {
  auto* s = static_cast<Sim*>(h);
  if (rho) std::memcpy(rho, s->rho.data(), sizeof(double) * s->nn);
  if (phi) std::memcpy(phi, s->phi.data(), sizeof(double) * s->nn);
  if (E) std::memcpy(E, s->E.data(), sizeof(double) * 3 * (size_t)s->nn);
}

// Returns PCG iterations; rel residual and convergence flag via out params.
int orc_solve(void* h, double* rel_res, int* converged) {
  auto* s = static_cast<Sim*>(h);
  PcgResult r = solve_poisson(*s);
  if (rel_res) *rel_res = r.rel_res;
  if (converged) *converged = r.converged;
  return r.iters;
}

void orc_efield(void* h) { compute_efield(*static_cast<Sim*>(h)); }

// Push all species one step (deposit into rho, which is cleared first).
void orc_push(void* h) {
  auto* s = static_cast<Sim*>(h);
  s->events.clear();
  std::fill(s->rho.begin(), s->rho.end(), 0.0);
  for (auto& p : s->sp) push_species(*s, p, s->rho);
}

// Deposit current particle positions (no push) into rho (cleared first).
void orc_deposit(void* h) {
  auto* s = static_cast<Sim*>(h);
  std::fill(s->rho.begin(), s->rho.end(), 0.0);
  for (auto& p : s->sp) {
    const double qw = p.q * p.w;
    for (size_t i = 0; i < p.cell.size(); ++i) {
      const int c = p.cell[i];
      if (c < 0) continue;
      double lam[4];
      bary(*s, c, &p.x[3 * i], lam);
      for (int k = 0; k < 4; ++k) s->rho[s->tet[4 * (size_t)c + k]] += qw * lam[k];
    }
  }
}

// One full cycle step in the order of PICCycle::executeTimeStep
// (Tools/PICcore/src/PICCycle.cpp:311-416): clear rho, push + deposit with the
// previous E, then (when configured) background collisions, secondary and
// photo-emission with their new particles deposited, Poisson, nodal E.
// Returns PCG iterations.
int orc_step(void* h, int nsteps) {
  auto* s = static_cast<Sim*>(h);
  int it = 0;
  for (int k = 0; k < nsteps; ++k) {
    if (!s->mcc.empty()) mcc_prepare(*s);
    orc_push(h);
    if (!s->mcc.empty()) measure_v2max(*s);  // the next step's majorant
    post_push(*s);
    it = solve_poisson(*s).iters;
    compute_efield(*s);
    s->steps++;
  }
  return it;
}

// Differentiable charging model (the forward that the device's discrete
// adjoint differentiates), FP64: steps of push + deposit with the previous E
// (no sub-steps when the push is configured non-adaptive), conductor charge
//   Q += hits (q w) + sum over electron hits on conductor group 6 + k of
//        |q| w delta_max yield_k s(E),  s(E) = (E/E_max) exp(2 - 2 sqrt(E/E_max))
//      + dt (J0 exp(-max(0, V - V_ref) / T_ph) A_sunlit - G V),
// then the floating-conductor solve (anchor collapse) and nodal E.
// p = {delta_max, E_max, yield[6], J0, T_ph, V_ref, sunlit bits, G}.
// Returns the final floating potential; the particles and fields advance.
double orc_charging_forward(void* h, const double* p, int steps) {
  auto* s = static_cast<Sim*>(h);
  const double dmax = p[0], emax = p[1], j0 = p[8], tph = std::max(1e-3, p[9]), vref = p[10], G = p[12];
  const int sunlit = (int)p[11];
  double area = 0.0;
  for (int k = 0; k < 6; ++k)
    if ((sunlit >> k) & 1) {
      const FaceList f = group_faces(*s, 6 + k);
      if (!f.cum.empty()) area += f.cum.back();
    }
  // impacts on the conductor faces are recorded as emission events
  std::array<SurfEmit, kMaxGroups> saved = s->emit;
  for (int k = 0; k < 6; ++k) {
    s->emit[6 + k].see_model = 2;
    s->emit[6 + k].dmax = dmax * p[2 + k];
    s->emit[6 + k].emax = emax;
  }
  for (int n = 0; n < steps; ++n) {
    const double V = s->Vfloat;
    orc_push(h);  // Qcond += q w of every conductor hit
    double see = 0.0;
    for (const SeeEvent& e : s->events) {
      if (e.group < 6 || e.group >= 12 || s->gtype[e.group] != BC_CONDUCTOR) continue;
      const Species& sp = s->sp[e.species];
      const double E = 0.5 * sp.m * (e.v[0] * e.v[0] + e.v[1] * e.v[1] + e.v[2] * e.v[2]) / kE;
      if (!(E > 0.0)) continue;
      const double r = E / emax;
      see += std::abs(sp.q * sp.w) * dmax * p[2 + e.group - 6] * r * std::exp(2.0 - 2.0 * std::sqrt(r));
    }
    s->events.clear();
    const double jph = j0 * std::exp(-std::max(0.0, V - vref) / tph);
    s->Qcond += see + s->dt * (jph * area - G * V);
    solve_poisson(*s);
    compute_efield(*s);
    s->steps++;
  }
  s->emit = saved;
  return s->Vfloat;
}

// Ledger: 7 values per group (abs_e_charge, abs_i_charge, emit_e_charge,
// energy, n_abs_e, n_abs_i, n_emit_e).
void orc_get_ledger(void* h, double* out, int ngroups) {
  auto* s = static_cast<Sim*>(h);
  for (int g = 0; g < ngroups && g < kMaxGroups; ++g) {
    const Ledger& L = s->ledger[g];
    double* o = out + 7 * g;
    o[0] = L.abs_e_charge; o[1] = L.abs_i_charge; o[2] = L.emit_e_charge; o[3] = L.energy;
    o[4] = (double)L.n_abs_e; o[5] = (double)L.n_abs_i; o[6] = (double)L.n_emit_e;
  }
}
int64_t orc_lost(void* h) { return static_cast<Sim*>(h)->lost; }

double orc_kinetic_energy(void* h) {
  auto* s = static_cast<Sim*>(h);
  double e = 0;
  for (auto& p : s->sp)
    for (size_t i = 0; i < p.cell.size(); ++i)
      if (p.cell[i] >= 0) {
        const double* v = &p.v[3 * i];
        e += 0.5 * p.m * p.w * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
      }
  return e;
}

// Standalone kernels used to pin the oracle against closed forms and the
// compiled reference.
void orc_boris(double* v, const double* E, const double* B, double qm, double dt) {
  boris(v, E, B, qm, dt);
}
double orc_adaptive_dt(const double* v0, const double* E, const double* B, double qm,
                       double dt_max, double floor_ratio) {
  return adaptive_dt(v0, E, B, qm, dt_max, floor_ratio);
}
void orc_barycentric(void* h, int t, const double* p, double* lam) {
  bary(*static_cast<Sim*>(h), t, p, lam);
}

// Fixed-point deposition of FP32 particle positions with FP32 tet maps, the
// bit-exact checker for the GPU deposit.  Per particle in tet t:
//   d = x - x0                          (FP32)
//   l_k = fma(a_k0, d0, fma(a_k1, d1, a_k2 * d2)), k = 1..3 (FP32)
//   f_k = (int32) trunc(clamp(l_k, 0, 1) * 2^24), f_0 = 2^24 - f_1 - f_2 - f_3
//   rho_fx[node_k] += c_s * f_k         (int64)
// geo: 12 floats per tet (a00..a22 row-major, x0 xyz).  Restates
// ChargeDepositor::depositCharge (PICParticleAlgorithms.cpp:325-369,
// q*w*lambda_k into node k) in integer units of (q w)/(c_s 2^24).
void orc_deposit_fixed_f32(int nt, const float* geo, const int32_t* tet, int64_t n,
                           const float* x, const int32_t* cell, int64_t cs, int64_t* rho_fx) {
  (void)nt;
  for (int64_t i = 0; i < n; ++i) {
    const int c = cell[i];
    if (c < 0) continue;
    const float* g = &geo[12 * (size_t)c];
    const float d0 = x[3 * i + 0] - g[9], d1 = x[3 * i + 1] - g[10], d2 = x[3 * i + 2] - g[11];
    int32_t f[4];
    int32_t sum = 0;
    for (int k = 0; k < 3; ++k) {
      float l = std::fmaf(g[3 * k + 0], d0, std::fmaf(g[3 * k + 1], d1, g[3 * k + 2] * d2));
      l = std::fminf(std::fmaxf(l, 0.0f), 1.0f);
      f[k + 1] = (int32_t)(l * 16777216.0f);
      sum += f[k + 1];
    }
    f[0] = 16777216 - sum;
    for (int k = 0; k < 4; ++k) rho_fx[tet[4 * (size_t)c + k]] += cs * (int64_t)f[k];
  }
}


// ---- collisions and surface emission -------------------------------------
void orc_set_seed(void* h, uint32_t lo, uint32_t hi) {
  auto* s = static_cast<Sim*>(h);
  s->key[0] = lo;
  s->key[1] = hi;
}
void orc_set_step(void* h, int64_t step) { static_cast<Sim*>(h)->steps = step; }

static MccProc unpack_proc(const double* q) {
  MccProc p;
  p.type = (int)q[0]; p.species = (int)q[1]; p.e_species = (int)q[2]; p.i_species = (int)q[3];
  p.n = q[4]; p.thr = q[5]; p.smax = q[6]; p.loss = q[7]; p.eref = q[8]; p.pidx = q[9];
  p.alpha = q[10]; p.beta = q[11]; p.clog = q[12];
  return p;
}
// 16 doubles per process: type, species, e_species, i_species, density,
// threshold, sigma_max, energy_loss, reference_energy, power_index, alpha,
// beta, coulomb_log, (3 unused).
void orc_mcc_configure(void* h, int n, const double* prm) {
  auto* s = static_cast<Sim*>(h);
  s->mcc.clear();
  for (int k = 0; k < n && k < 16; ++k) s->mcc.push_back(unpack_proc(prm + 16 * k));
  s->v2_have = false;
}
// 8 doubles: see_model, delta_max, energy_max, see_temperature, photo_j,
// photo_temperature, photo_reference_v, electron_species.
void orc_set_emission(void* h, int g, const double* q) {
  auto* s = static_cast<Sim*>(h);
  SurfEmit& e = s->emit[g];
  e.see_model = (int)q[0]; e.dmax = q[1]; e.emax = q[2]; e.t_see = q[3];
  e.photo_j = q[4]; e.t_ph = q[5]; e.v_ref = q[6]; e.e_species = (int)q[7];
}
// Collisions, secondary and photo-emission after a push (deposits into rho).
void orc_post_push(void* h) { post_push(*static_cast<Sim*>(h)); }
// out: 16 per-process collision counts, created, removed, SEE emitted, photo emitted
void orc_mcc_counts(void* h, int64_t* out) {
  auto* s = static_cast<Sim*>(h);
  for (int k = 0; k < 16; ++k) out[k] = s->mcc_count[k];
  out[16] = s->mcc_created; out[17] = s->mcc_removed; out[18] = s->see_emitted; out[19] = s->photo_emitted;
  out[20] = s->mcc_candidates; out[21] = s->mcc_major;
}
void orc_get_ids(void* h, int si, int32_t* ids) {
  auto* s = static_cast<Sim*>(h);
  const Species& p = s->sp[si];
  std::memcpy(ids, p.id.data(), sizeof(int32_t) * p.id.size());
}
void orc_sigma(const double* prm, int64_t n, const double* e_ev, double* out) {
  const MccProc p = unpack_proc(prm);
  for (int64_t i = 0; i < n; ++i) out[i] = sigma_of(p, e_ev[i]);
}
// Coulomb logarithm of CoulombCollisionCrossSection (CollisonTypes.cpp:178-198):
// ln(lambda_D / max(1e-12, 1e-4 lambda_D)) clamped to [1, 30],
// lambda_D = sqrt(eps0 k T / (n e^2)) (1e-3 m for T or n <= 0).
double orc_coulomb_log(double temperature_k, double density_m3) {
  const double kB = 1.380649e-23;  // Tools/Basic/include/Constants.h:51
  const double ld = (temperature_k <= 0.0 || density_m3 <= 0.0)
                        ? 1e-3 : std::sqrt(kEps0 * kB * temperature_k / (density_m3 * kE * kE));
  const double bmin = std::max(1e-12, ld * 1e-4);
  return std::min(30.0, std::max(1.0, std::log(std::max(1.000001, ld / bmin))));
}
void orc_see_yield(int model, double dmax, double emax, int64_t n, const double* e_ev,
                   const double* cos_inc, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = see_yield(model, dmax, emax, e_ev[i], cos_inc[i]);
}
void orc_photo_current(double j0y, double t_ph, double v_ref, int64_t n, const double* V, double* out) {
  for (int64_t i = 0; i < n; ++i) out[i] = photo_current(j0y, t_ph, v_ref, V[i]);
}
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

}  // extern "C"
