// Harness exposing reference routines to the test-suite through a C ABI.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's own
// sources (oracle/ref/Makefile) into oracle/_ref/libscdat_ref.so; used to
// produce golden vectors that pin the FP64 oracle.  Nothing in the product
// links or loads it.
//
// Reference classes used (paths relative to the SCDAT source tree):
//   Tools/Particle/src/PICParticleAlgorithms.cpp  FieldInterpolator,
//       ChargeDepositor, CrossTetrahedron
//   Tools/FieldSolver/src/MultiBoundaryElectricFieldSolver.cpp  P1 FEM,
//       FLOATING conductor, Dirichlet
//   Tools/FieldSolver/src/PoissonSolver.cpp  graph Laplacian, LSQ nodal E
//   Tools/Solver/src/LinearSystem.cpp  dense Gaussian elimination
//   Tools/Mesh/src/MeshAlgorithms.cpp, MeshParsing.cpp  nodes, tets, box mesher
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "Tools/FieldSolver/include/MultiBoundaryElectricFieldSolver.h"
#include "Tools/Mesh/include/MeshAlgorithms.h"
#include "Tools/Mesh/include/MeshParsing.h"
#include "Tools/Particle/include/PICParticleAlgorithms.h"

using SCDAT::Geometry::Point3D;
using SCDAT::Geometry::Vector3D;
namespace M = SCDAT::Mesh;
namespace P = SCDAT::Particle;
namespace F = SCDAT::FieldSolver;

namespace {

struct Built {
  std::vector<M::NodePtr> nodes;
  std::vector<M::ElementPtr> elems;
};

Built build(int nn, const double* xyz, int nt, const int32_t* tets) {
  Built b;
  b.nodes.reserve(nn);
  for (int i = 0; i < nn; ++i)
    b.nodes.push_back(std::make_shared<M::Node>(
        (M::NodeId)i, Point3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2])));
  for (int t = 0; t < nt; ++t) {
    std::vector<M::NodePtr> v = {b.nodes[tets[4 * t]], b.nodes[tets[4 * t + 1]],
                                 b.nodes[tets[4 * t + 2]], b.nodes[tets[4 * t + 3]]};
    b.elems.push_back(std::make_shared<M::TetrahedronElement>((M::ElementId)t, v));
  }
  return b;
}

}  // namespace

extern "C" {

// FieldInterpolator::computeBarycentricCoordinates (PICParticleAlgorithms.cpp:140-148,216-296).
int ref_barycentric(int nn, const double* xyz, int nt, const int32_t* tets, int npts,
                    const int32_t* tet_of_point, const double* pts, double* lam, int32_t* inside) {
  Built b = build(nn, xyz, nt, tets);
  P::FieldInterpolator fi;
  fi.setMesh(b.nodes, b.elems);
  std::vector<double> l(4);
  for (int k = 0; k < npts; ++k) {
    const bool in = fi.computeBarycentricCoordinates(
        Point3D(pts[3 * k], pts[3 * k + 1], pts[3 * k + 2]), (P::ElementId)tet_of_point[k], l);
    for (int j = 0; j < 4; ++j) lam[4 * k + j] = l[j];
    inside[k] = in ? 1 : 0;
  }
  return 0;
}

// ChargeDepositor::depositCharge (PICParticleAlgorithms.cpp:325-369) of point
// charges into a nodal array.
int ref_deposit(int nn, const double* xyz, int nt, const int32_t* tets, int np, const double* pos,
                const int32_t* cell, double charge, double weight, double* rho) {
  Built b = build(nn, xyz, nt, tets);
  P::ChargeDepositor cd;
  cd.setMesh(b.nodes, b.elems);
  std::vector<double> r(nn, 0.0);
  for (int k = 0; k < np; ++k)
    cd.depositCharge(Point3D(pos[3 * k], pos[3 * k + 1], pos[3 * k + 2]), (P::ElementId)cell[k],
                     charge, weight, r);
  std::memcpy(rho, r.data(), sizeof(double) * nn);
  return 0;
}

// CrossTetrahedron::crossTetrahedron (PICParticleAlgorithms.cpp:493-740) for a
// batch of independent particles, one step each, with nodal E gathered by
// FieldInterpolator (mode 0) or a uniform E/B (mode 1, crossTetraUniformEwithB).
// kernel: 0 NearestPoint, 1 LinearSegmentCloud (segments), 2 QuadratureGauss3.
// out_state: 6 per particle (x, v); out_elem; out_lost; rho accumulates.
int ref_cross(int nn, const double* xyz, int nt, const int32_t* tets, const double* nodal_E,
              int mode, const double* uniform_E, const double* uniform_B, int kernel, int segments,
              int np, const double* state, const int32_t* hint, double mass, double charge,
              double weight, double dt, double* out_state, int32_t* out_elem, int32_t* out_lost,
              double* rho) {
  Built b = build(nn, xyz, nt, tets);
  auto fi = std::make_shared<P::FieldInterpolator>();
  fi->setMesh(b.nodes, b.elems);
  std::vector<Vector3D> E(nn);
  if (nodal_E)
    for (int i = 0; i < nn; ++i) E[i] = Vector3D(nodal_E[3 * i], nodal_E[3 * i + 1], nodal_E[3 * i + 2]);
  fi->setElectricField(E);
  fi->setMagneticField(std::vector<Vector3D>(nn, Vector3D(0, 0, 0)));
  auto cd = std::make_shared<P::ChargeDepositor>();
  cd->setMesh(b.nodes, b.elems);
  cd->setTrajectoryDepositionScheme((P::TrajectoryChargeDepositionKernel)kernel, (size_t)segments);
  P::CrossTetrahedron ct;
  ct.setMesh(b.nodes, b.elems);
  ct.setFieldInterpolator(fi);
  ct.setChargeDepositor(cd);
  std::vector<double> r(nn, 0.0);
  for (int k = 0; k < np; ++k) {
    const double* s = &state[6 * k];
    P::Particle p((size_t)k, P::ParticleType::ELECTRON, Point3D(s[0], s[1], s[2]),
                  Vector3D(s[3], s[4], s[5]), mass, charge, weight);
    P::CrossTetrahedron::TrajectoryResult res;
    if (mode == 1)
      res = ct.crossTetraUniformEwithB(&p, dt, Vector3D(uniform_E[0], uniform_E[1], uniform_E[2]),
                                       Vector3D(uniform_B[0], uniform_B[1], uniform_B[2]), r);
    else
      res = ct.crossTetrahedron(&p, dt, r, true, (P::ElementId)(hint ? hint[k] : -1));
    out_state[6 * k + 0] = res.final_position.x();
    out_state[6 * k + 1] = res.final_position.y();
    out_state[6 * k + 2] = res.final_position.z();
    out_state[6 * k + 3] = res.final_velocity.x();
    out_state[6 * k + 4] = res.final_velocity.y();
    out_state[6 * k + 5] = res.final_velocity.z();
    out_elem[k] = (int32_t)res.final_element;
    out_lost[k] = res.lost_particle ? 1 : 0;
  }
  std::memcpy(rho, r.data(), sizeof(double) * nn);
  return 0;
}

// MultiBoundaryElectricFieldSolver (MultiBoundaryElectricFieldSolver.cpp):
// P1 stiffness (:1704-1759), Dirichlet (:1229-1243), FLOATING node collapse
// with charge q_float (:1172-1217).  kind: 0 free, 1 Dirichlet (dval), 2
// floating conductor.  density: nodal charge density handed to
// setChargeDensity.  permittivity: value for material 0 (resolved by the
// reference: <= 0 vacuum, < 1e-8 absolute, else relative).  solver: the
// reference SolverType enum (0 CG, 3 DIRECT_LU).
int ref_multiboundary(int nn, const double* xyz, int nt, const int32_t* tets, const int32_t* kind,
                      const double* dval, const double* density, double q_float,
                      double permittivity, int solver, double tol, double* phi) {
  auto mesh = std::make_shared<M::Mesh>();
  for (int i = 0; i < nn; ++i) mesh->addNode(Point3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
  std::vector<M::NodeId> ids(4);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 4; ++k) ids[k] = (M::NodeId)tets[4 * t + k];
    mesh->addElement(M::ElementType::TETRAHEDRON, ids);
  }
  F::MultiBoundaryElectricFieldSolver s(mesh);
  s.setMaterialPermittivity(0, permittivity);
  F::SolverConfiguration cfg;
  cfg.tolerance = tol;
  cfg.max_iterations = 100000;
  cfg.solver_type = (F::SolverType)solver;
  s.setSolverConfiguration(cfg);
  std::map<double, std::vector<M::NodeId>> dir;
  std::vector<M::NodeId> flo;
  for (int i = 0; i < nn; ++i) {
    if (kind[i] == 1) dir[dval ? dval[i] : 0.0].push_back((M::NodeId)i);
    if (kind[i] == 2) flo.push_back((M::NodeId)i);
    if (density) s.setChargeDensity((M::NodeId)i, density[i]);
  }
  int k = 0;
  for (auto& kv : dir)
    s.addBoundaryCondition("dirichlet_" + std::to_string(k++), F::BoundaryConditionType::DIRICHLET,
                           kv.first, kv.second);
  if (!flo.empty())
    s.addBoundaryCondition("floating", F::BoundaryConditionType::FLOATING, q_float, flo);
  const bool ok = s.solve();
  for (int i = 0; i < nn; ++i) phi[i] = s.getPotential((M::NodeId)i);
  return ok ? 0 : 1;
}

}  // extern "C"
