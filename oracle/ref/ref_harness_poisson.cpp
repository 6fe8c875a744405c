// Harness exposing the reference PoissonSolver to the test-suite (C ABI).
//
// TEST INFRASTRUCTURE ONLY; compiled with the reference's own sources into
// oracle/_ref/libscdat_ref.so (see ref_harness.cpp).  Separate translation
// unit because PoissonSolver.h and MultiBoundaryElectricFieldSolver.h declare
// conflicting types of the same name.
#include <cmath>
#include <cstdint>
#include <memory>
#include <vector>

#include "Tools/FieldSolver/include/BoundaryCondition.h"
#include "Tools/FieldSolver/include/PoissonSolver.h"
#include "Tools/Mesh/include/MeshParsing.h"

using SCDAT::Geometry::Point3D;
using SCDAT::Geometry::Vector3D;
namespace M = SCDAT::Mesh;
namespace F = SCDAT::FieldSolver;

extern "C" {

// PoissonSolver (PoissonSolver.cpp:251-529): graph-Laplacian operator,
// -rho V/4 load, Dirichlet nodes, DIRECT solve, LSQ nodal E.
int ref_poisson(int nn, const double* xyz, int nt, const int32_t* tets, const int32_t* kind,
                const double* dval, const double* density, double* phi, double* E) {
  auto mesh = std::make_shared<M::VolMesh>();
  for (int i = 0; i < nn; ++i) mesh->addNode(Point3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
  std::vector<std::size_t> ids(4);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 4; ++k) ids[k] = (std::size_t)tets[4 * t + k];
    mesh->addElement(M::ElementType::TETRAHEDRON, ids);
  }
  F::PoissonSolver s(mesh);
  for (int i = 0; i < nn; ++i) {
    if (kind[i] == 1)
      s.addBoundaryCondition((M::NodeId)i, std::make_shared<F::BoundaryCondition>(
                                               F::BoundaryConditionType::DIRICHLET, dval ? dval[i] : 0.0));
    if (density) s.setChargeDensity((M::NodeId)i, density[i]);
  }
  const bool ok = s.solve(SCDAT::Solver::SolverType::DIRECT);
  for (int i = 0; i < nn; ++i) {
    phi[i] = s.getPotential((M::NodeId)i);
    const Vector3D e = s.getElectricField((M::NodeId)i);
    E[3 * i] = e.x();
    E[3 * i + 1] = e.y();
    E[3 * i + 2] = e.z();
  }
  return ok ? 0 : 1;
}

// Two solves on one solver: the first as ref_poisson; then new charge
// densities (density2), new Dirichlet values (dval2, same nodes) and NEUMANN
// fluxes (flux2, nodes with kind 0 and flux2 != 0), assembleSystem() and a
// second solve, whose potentials and fields are returned.
int ref_poisson_twice(int nn, const double* xyz, int nt, const int32_t* tets, const int32_t* kind,
                      const double* dval, const double* density, const double* dval2, const double* density2,
                      const double* flux2, double* phi, double* E) {
  auto mesh = std::make_shared<M::VolMesh>();
  for (int i = 0; i < nn; ++i) mesh->addNode(Point3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
  std::vector<std::size_t> ids(4);
  for (int t = 0; t < nt; ++t) {
    for (int k = 0; k < 4; ++k) ids[k] = (std::size_t)tets[4 * t + k];
    mesh->addElement(M::ElementType::TETRAHEDRON, ids);
  }
  F::PoissonSolver s(mesh);
  for (int i = 0; i < nn; ++i) {
    if (kind[i] == 1)
      s.addBoundaryCondition((M::NodeId)i,
                             std::make_shared<F::BoundaryCondition>(F::BoundaryConditionType::DIRICHLET, dval[i]));
    s.setChargeDensity((M::NodeId)i, density[i]);
  }
  if (!s.solve(SCDAT::Solver::SolverType::DIRECT)) return 1;
  for (int i = 0; i < nn; ++i) {
    if (kind[i] == 1)
      s.addBoundaryCondition((M::NodeId)i,
                             std::make_shared<F::BoundaryCondition>(F::BoundaryConditionType::DIRICHLET, dval2[i]));
    else if (flux2[i] != 0.0)
      s.addBoundaryCondition((M::NodeId)i,
                             std::make_shared<F::BoundaryCondition>(F::BoundaryConditionType::NEUMANN, flux2[i]));
    s.setChargeDensity((M::NodeId)i, density2[i]);
  }
  s.assembleSystem();
  const bool ok = s.solve(SCDAT::Solver::SolverType::DIRECT);
  for (int i = 0; i < nn; ++i) {
    phi[i] = s.getPotential((M::NodeId)i);
    const Vector3D e = s.getElectricField((M::NodeId)i);
    E[3 * i] = e.x();
    E[3 * i + 1] = e.y();
    E[3 * i + 2] = e.z();
  }
  return ok ? 0 : 1;
}

// PoissonSolver_test.cpp:48-117 (SolvesFiniteDirichletProfileOnPlateMesh):
// the 0.01 x 0.01 x 0.02 m plate of 1 x 1 x 6 cells, tetrahedralised by the
// reference's generator, 10 V on z_min and 0 V on z_max, a DIRECT solve.
// Per node in mesh order (at most cap): z, phi, E; totals = total energy,
// largest and smallest potential.  Returns 0 when the solve succeeded.
int ref_poisson_plate(int cap, int* nn, double* z, double* phi, double* E, double* totals) {
  M::GeometryDimensions dims;
  dims.shape = M::GeometryShape::PLATE;
  dims.length = 0.01;
  dims.width = 0.01;
  dims.thickness = 0.02;
  M::MeshGenerationOptions opt;
  opt.nx = 1;
  opt.ny = 1;
  opt.nz = 6;
  opt.tetrahedralize = true;
  std::shared_ptr<M::VolMesh> mesh(M::GeometryMeshGenerator::generateFromDimensions(dims, opt));
  if (!mesh) return 2;
  const double z0 = mesh->getBoundingBoxMin().z(), z1 = mesh->getBoundingBoxMax().z();
  F::PoissonSolver s(mesh);
  for (const auto& node : mesh->getNodes()) {
    const double zn = node->getPosition().z();
    if (std::abs(zn - z0) < 1e-12 || std::abs(zn - z1) < 1e-12) {
      node->setBoundaryType(M::BoundaryType::DIRICHLET);
      s.addBoundaryCondition(node->getId(), std::make_shared<F::BoundaryCondition>(
                                                F::BoundaryConditionType::DIRICHLET,
                                                std::abs(zn - z0) < 1e-12 ? 10.0 : 0.0));
    } else {
      node->setBoundaryType(M::BoundaryType::INTERIOR);
    }
  }
  const bool ok = s.solve(SCDAT::Solver::SolverType::DIRECT);
  int k = 0;
  for (const auto& node : mesh->getNodes()) {
    if (k >= cap) return 3;
    z[k] = node->getPosition().z();
    phi[k] = node->getPotential();
    const Vector3D e = s.getElectricField(node->getId());
    E[3 * k] = e.x();
    E[3 * k + 1] = e.y();
    E[3 * k + 2] = e.z();
    ++k;
  }
  *nn = k;
  totals[0] = s.getTotalEnergy();
  totals[1] = s.getMaxPotential();
  totals[2] = s.getMinPotential();
  return ok ? 0 : 1;
}

}  // extern "C"
