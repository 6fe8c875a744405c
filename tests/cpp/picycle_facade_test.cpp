// The reference's PICCycle boundary tests (Tools/PICcore/test/
// PICCycleBoundary_test.cpp:121-221) re-expressed against the C++ PICCycle
// facade over the gfx950 library.  Prints one line per check and exits
// non-zero on the first failure.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../scdat_amd/facade/PICCycle.hpp"

using scdat_amd::PICcore::BoundaryKind;
using scdat_amd::PICcore::PICCycle;
using scdat_amd::PICcore::Vector3D;

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond);             \
      std::exit(1);                                                          \
    }                                                                        \
  } while (0)

namespace {

struct Mesh {
  std::vector<double> xyz;
  std::vector<int32_t> tets, kind, group;
};

// 1 x 1 x 1 m plate, 2 x 2 x 6 hexes, 6 tets each (conforming split).
Mesh plate() {
  scdat_box_spec s{};
  for (int r = 0; r < 3; ++r) {
    s.lo[r] = 0.0;
    s.hi[r] = 1.0;
  }
  s.n[0] = 2;
  s.n[1] = 2;
  s.n[2] = 6;
  int64_t nn = 0, nt = 0;
  CHECK(scdat_box_mesh_size(&s, &nn, &nt) == SCDAT_OK);
  Mesh m;
  m.xyz.resize(3 * nn);
  m.tets.resize(4 * nt);
  m.kind.resize(nn);
  m.group.resize(4 * nt);
  CHECK(scdat_box_mesh_build(&s, m.xyz.data(), m.tets.data(), m.kind.data(), m.group.data()) == SCDAT_OK);
  return m;
}

Vector3D centroid_of_extreme_tet(const Mesh& m, int axis) {
  Vector3D best;
  double bv = -1e300;
  for (size_t t = 0; t < m.tets.size() / 4; ++t) {
    double c[3] = {0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int r = 0; r < 3; ++r) c[r] += 0.25 * m.xyz[3 * (size_t)m.tets[4 * t + k] + r];
    if (c[axis] > bv) {
      bv = c[axis];
      best = Vector3D{c[0], c[1], c[2]};
    }
  }
  return best;
}

// createConfiguredCycle (PICCycleBoundary_test.cpp:68-99): dt 1e-9, periodic
// x and y, absorbing z, zero potential.
PICCycle configured(const Mesh& m) {
  PICCycle::Parameters p;
  p.time_step = 1.0e-9;
  p.max_iterations = 4;
  p.statistics_frequency = 1;
  p.precision = 1;
  p.sort_interval = 0;
  PICCycle cycle(p);
  cycle.setMesh(m.xyz, m.tets, m.kind, {}, m.group);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::XMin, BoundaryKind::Periodic);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::XMax, BoundaryKind::Periodic);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::YMin, BoundaryKind::Periodic);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::YMax, BoundaryKind::Periodic);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::ZMin, BoundaryKind::Absorbing);
  cycle.setBoundaryCondition(PICCycle::BoundaryFace::ZMax, BoundaryKind::Absorbing);
  return cycle;
}

void start(PICCycle& cycle, const Mesh& m, const Vector3D& x, const Vector3D& v) {
  cycle.addElectrons({x}, {v}, 1.0);
  CHECK(cycle.initialize());
  cycle.setPotential(std::vector<double>(m.kind.size(), 0.0));
}

void periodic_boundary_keeps_particle_active() {
  const Mesh m = plate();
  PICCycle cycle = configured(m);
  start(cycle, m, centroid_of_extreme_tet(m, 0), Vector3D{2.0e9, 0.0, 0.0});
  CHECK(cycle.executeTimeStep());
  std::vector<Vector3D> x, v;
  std::vector<int32_t> cell, id;
  cycle.getParticles(0, x, v, cell, id);
  CHECK(cell.size() == 1 && cell[0] >= 0);
  CHECK(x[0].x >= -1e-9 && x[0].x <= 1.0 + 1e-9);
  std::printf("ok periodic: x = %.12f\n", x[0].x);
}

void absorbing_boundary_records_charge() {
  const Mesh m = plate();
  PICCycle cycle = configured(m);
  start(cycle, m, centroid_of_extreme_tet(m, 2), Vector3D{0.0, 0.0, 2.0e9});
  CHECK(cycle.executeTimeStep());
  std::vector<Vector3D> x, v;
  std::vector<int32_t> cell, id;
  cycle.getParticles(0, x, v, cell, id);
  CHECK(cell[0] == -1);
  CHECK(cycle.getStatistics().surface_deposited_charge < 0.0);
  const auto& led = cycle.getSurfaceCurrentLedger();
  const auto& top = led[(size_t)PICCycle::BoundaryFace::ZMax];
  CHECK(top.absorbed_electron_charge_c < 0.0);
  CHECK(top.absorbed_electron_particles == 1u);
  const double q = cycle.getAndResetSurfaceCharge();
  CHECK(std::abs(q + SCDAT_ELEMENTARY_CHARGE) < 1e-30);
  CHECK(cycle.getStatistics().surface_deposited_charge == 0.0);
  std::printf("ok absorbing: ledger ZMax electrons = %zu\n", top.absorbed_electron_particles);
}

void manual_potential_reconstructs_field() {
  const Mesh m = plate();
  PICCycle cycle = configured(m);
  start(cycle, m, Vector3D{0.5, 0.5, 0.5}, Vector3D{0.0, 0.0, 0.0});
  std::vector<double> phi(m.kind.size());
  for (size_t i = 0; i < phi.size(); ++i) phi[i] = 12.0 * m.xyz[3 * i + 2];
  cycle.setPotential(phi);
  const auto& E = cycle.getElectricField();
  CHECK(E.size() == m.kind.size());
  double mean = 0.0;
  for (const auto& e : E) {
    CHECK(std::isfinite(e.x) && std::isfinite(e.y) && std::isfinite(e.z));
    CHECK(std::abs(e.z + 12.0) < 1e-9);  // least squares is exact for linear phi
    mean += e.z;
  }
  mean /= (double)E.size();
  CHECK(mean < -1.0);
  std::printf("ok potential: mean Ez = %.12f\n", mean);
}

void no_condition_escapes_without_charge() {
  const Mesh m = plate();
  PICCycle cycle = configured(m);
  cycle.clearBoundaryConditions();
  start(cycle, m, centroid_of_extreme_tet(m, 2), Vector3D{0.0, 0.0, 5.0e12});
  CHECK(cycle.executeTimeStep());
  std::vector<Vector3D> x, v;
  std::vector<int32_t> cell, id;
  cycle.getParticles(0, x, v, cell, id);
  CHECK(cell[0] == -1);
  CHECK(std::abs(cycle.getStatistics().surface_deposited_charge) < 1e-20);
  std::printf("ok escape\n");
}

}  // namespace

int main() {
  periodic_boundary_keeps_particle_active();
  absorbing_boundary_records_charge();
  manual_potential_reconstructs_field();
  no_condition_escapes_without_charge();
  std::printf("all facade checks passed\n");
  return 0;
}
