// Harness exposing the reference PoissonSolver to the test-suite (C ABI).
//
// TEST INFRASTRUCTURE ONLY; compiled with the reference's own sources into
// oracle/_ref/libscdat_ref.so (see ref_harness.cpp).  Separate translation
// unit because PoissonSolver.h and MultiBoundaryElectricFieldSolver.h declare
// conflicting types of the same name.
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

}  // extern "C"
