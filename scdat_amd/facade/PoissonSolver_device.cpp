// Device implementation of SCDAT::FieldSolver::PoissonSolver, the class the
// reference declares in Tools/FieldSolver/include/PoissonSolver.h
// (unchanged): build this file against the reference's headers and link it,
// with libscdat_amd.so, in place of Tools/FieldSolver/src/PoissonSolver.cpp.
// Callers keep their VolMesh, their FieldSolver::BoundaryCondition objects
// and their charge densities; the graph-Laplacian system is solved on the
// GPU through the C ABI, and the nodal field comes from the device's
// least-squares gradient.
//
// Per member (reference behaviour cited where it is mirrored):
//  * the boundary-condition, material and charge-density maps are kept as
//    PoissonSolver.cpp:221-250 keeps them; the mesh's nodes are indexed in
//    the order of VolMesh::getNodes (buildNodeIndexMapping, :627-643);
//  * assembleSystem() forms what the reference assembles into its dense
//    system (:251-357): tetrahedra with edge weights eps V / (6 d^2), the
//    element permittivity resolved as :21-29, the load -rho_i V/4 summed
//    over the elements of node i, DIRICHLET values held, NEUMANN fluxes added
//    to the load (other condition types are ignored, as :341-357 ignores
//    them), and hands the mesh to a device context made for it (remade when
//    the mesh, the Dirichlet set or the permittivity changes);
//  * solve() assembles first if nothing is assembled yet (:360-373: later
//    charge changes need assembleSystem(), as in the reference), then solves
//    on the device in FP64 to a relative residual of 1e-12 whatever the
//    SolverType (the reference's DIRECT elimination of an N x N matrix does
//    not fit at PIC sizes), writes the potentials and fields onto the
//    mesh's nodes (updateNodePotentials, updateNodeElectricFields) and
//    returns false, with a message, when the solve missed its tolerance;
//  * computeElectricFieldGradient() sends the nodes' potentials to the device
//    and writes back E = -grad(phi) of the least-squares gradient over edge
//    neighbours (:395-446); a node the mesh no longer holds reads the
//    device's gradient of the last solve (:375-385 computes it on the host);
//  * the getters read what the reference's read: potentials from the
//    solution, fields from the nodes, the field at a position through the
//    mesh's spatial index as :387-402,655-686 do.
// Departures: element-wise assembly (assembleElementMatrix / RHS and the
// tetrahedron / triangle forms) marks the system for a whole assembly at the
// next solve instead of adding one element into a dense matrix; only
// tetrahedral elements and one permittivity across them are supported
// (solve() returns false with a message otherwise).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <limits>
#include <memory>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "FieldSolver/include/PoissonSolver.h"
#include "scdat_amd.h"

namespace SCDAT {
namespace FieldSolver {

namespace {

constexpr double kGeomTol = 1.0e-12;  // PoissonSolver.cpp:19

// Element permittivity: <= 0 vacuum, below 1e-8 absolute, else relative
// (PoissonSolver.cpp:21-29).
double resolve_permittivity(double value) {
  if (value <= 0.0) return SCDAT_VACUUM_PERMITTIVITY;
  return value < 1.0e-8 ? value : value * SCDAT_VACUUM_PERMITTIVITY;
}

double measure_of(const Mesh::ElementPtr& e) { return e ? std::max(std::abs(e->getVolume()), kGeomTol) : 0.0; }

Mesh::NodePtr node_by_id(const std::vector<Mesh::NodePtr>& nodes, Mesh::NodeId id) {
  if (id < nodes.size() && nodes[id] && nodes[id]->getId() == id) return nodes[id];
  for (const auto& n : nodes)
    if (n && n->getId() == id) return n;
  return nullptr;
}

// The device context of one solver and the system it holds.  It rides in
// system_ (the reference's dense LinearSystem, which this implementation
// does not build) as the deleter of an empty shared pointer, so the
// solver's defaulted destructor releases it.
struct DeviceState {
  scdat_ctx* ctx = nullptr;
  // what the context was made for
  std::vector<double> xyz;
  std::vector<int32_t> tets, kind;
  double eps = 0.0;
  // the assembled problem, per node index
  std::vector<double> dval;     // Dirichlet values (kind 1)
  std::vector<double> load;     // rho_i times the node's share of the volume
  std::vector<double> flux;     // NEUMANN values added to the right-hand side
  std::string unsupported;      // why the mesh cannot be solved here ("" if it can)
  bool assembled = false;
  // the last solution
  std::vector<double> phi, E;
  bool solved = false;
  ~DeviceState() {
    if (ctx) scdat_destroy(ctx);
  }
};

struct StateHolder {
  std::shared_ptr<DeviceState> st;
  void operator()(LinearSystem*) const {}
};

DeviceState* state_of(const LinearSystemPtr& p) {
  StateHolder* h = std::get_deleter<StateHolder>(p);
  return h ? h->st.get() : nullptr;
}

LinearSystemPtr make_state() { return LinearSystemPtr(nullptr, StateHolder{std::make_shared<DeviceState>()}); }

bool report(const char* what, scdat_ctx* ctx, int rc) {
  std::cerr << "PoissonSolver (device): " << what << " failed: " << (ctx ? scdat_last_error(ctx) : "no context")
            << " (status " << rc << ")" << std::endl;
  return false;
}

}  // namespace

// ---- FieldSolver::BoundaryCondition (declared in BoundaryCondition.h; the
// reference defines these in PoissonSolver.cpp:177-209) ----------------------

double BoundaryCondition::getValue(const Utils::Point3D& position, double time) const {
  return (is_time_dependent_ && time_function_) ? time_function_(position, time) : value_;
}

void BoundaryCondition::apply(Mesh::NodePtr node, double time) const {
  if (node && isApplicable(node) && type_ == BoundaryConditionType::DIRICHLET)
    node->setPotential(getValue(node->getPosition(), time));
}

bool BoundaryCondition::isApplicable(Mesh::NodePtr node) const {
  return node && node->getBoundaryType() != Mesh::BoundaryType::INTERIOR;
}

std::string BoundaryCondition::toString() const {
  std::ostringstream s;
  s << "BoundaryCondition(type=" << static_cast<int>(type_) << ", value=" << value_
    << ", time_dependent=" << (is_time_dependent_ ? "true" : "false") << ")";
  return s.str();
}

// ---- PoissonSolver ------------------------------------------------------------

void PoissonSolver::initialize() {
  buildNodeIndexMapping();
  system_ = make_state();
}

void PoissonSolver::addBoundaryCondition(Mesh::NodeId node_id, BoundaryConditionPtr bc) {
  boundary_conditions_[node_id] = std::move(bc);
}

void PoissonSolver::clearBoundaryConditions() { boundary_conditions_.clear(); }

void PoissonSolver::addMaterialProperty(Mesh::MaterialId material_id, MaterialPropertyPtr property) {
  material_properties_[material_id] = std::move(property);
}

void PoissonSolver::clearMaterialProperties() { material_properties_.clear(); }

void PoissonSolver::setChargeDensity(Mesh::NodeId node_id, double density) { charge_densities_[node_id] = density; }

void PoissonSolver::setChargeDensity(const std::unordered_map<Mesh::NodeId, double>& densities) {
  charge_densities_ = densities;
}

void PoissonSolver::clearChargeDensities() { charge_densities_.clear(); }

void PoissonSolver::assembleSystem() {
  if (!hasValidMesh()) return;
  buildNodeIndexMapping();
  if (!state_of(system_)) system_ = make_state();
  DeviceState& st = *state_of(system_);
  const std::size_t nn = node_to_index_.size();
  const auto nodes = getMeshNodes();
  st.unsupported.clear();
  std::vector<double> xyz(3 * nn, 0.0);
  for (const auto& n : nodes) {
    if (!n) continue;
    const std::size_t i = node_to_index_.at(n->getId());
    xyz[3 * i] = n->getPosition().x();
    xyz[3 * i + 1] = n->getPosition().y();
    xyz[3 * i + 2] = n->getPosition().z();
  }
  // tetrahedra, their permittivity, and each node's share of the volume
  std::vector<int32_t> tets;
  std::vector<double> share(nn, 0.0);
  double eps = -1.0;
  for (const auto& e : getMeshElements()) {
    if (!e) continue;
    const auto& en = e->getNodes();
    std::array<int32_t, 4> ids{};
    std::size_t k = 0;
    for (const auto& n : en) {
      if (!n) continue;
      const auto it = node_to_index_.find(n->getId());
      if (it == node_to_index_.end()) continue;
      if (k < 4) ids[k] = (int32_t)it->second;
      ++k;
    }
    if (k == 0) continue;
    if (e->getType() != Mesh::ElementType::TETRAHEDRON || en.size() != 4 || k != 4) {
      if (k >= 2) st.unsupported = "the device assembles tetrahedra only (the mesh has other elements)";
      continue;
    }
    const double pe = resolve_permittivity(getPermittivity(e));
    if (eps < 0.0) eps = pe;
    if (pe != eps) st.unsupported = "the device assembles one permittivity (the elements differ)";
    tets.insert(tets.end(), ids.begin(), ids.end());
    const double s = measure_of(e) / 4.0;
    for (int q = 0; q < 4; ++q) share[ids[q]] += s;
  }
  if (tets.empty() && st.unsupported.empty()) st.unsupported = "the mesh has no tetrahedra";
  // load: the reference's right-hand side is -rho_i V/4 per element; the
  // device takes rho_i times the share, with the reference's load sign
  st.load.assign(nn, 0.0);
  for (const auto& [id, density] : charge_densities_) {
    const auto it = node_to_index_.find(id);
    if (it != node_to_index_.end()) st.load[it->second] = density * share[it->second];
  }
  // make the context when the mesh, its Dirichlet set or eps changed
  std::vector<int32_t> kind(nn, SCDAT_NODE_FREE);
  std::vector<double> dval(nn, 0.0);
  for (const auto& [id, bc] : boundary_conditions_) {
    const auto it = node_to_index_.find(id);
    if (it == node_to_index_.end() || !bc) continue;
    if (bc->getType() == BoundaryConditionType::DIRICHLET) {
      kind[it->second] = SCDAT_NODE_DIRICHLET;
      dval[it->second] = bc->getValue();
    }
  }
  st.dval = dval;
  applyBoundaryConditions();
  if (st.unsupported.empty() && (!st.ctx || st.xyz != xyz || st.tets != tets || st.kind != kind || st.eps != eps)) {
    if (st.ctx) scdat_destroy(st.ctx);
    st.ctx = nullptr;
    st.solved = false;
    scdat_config cfg;
    scdat_config_default(&cfg);
    cfg.precision = 1;
    cfg.poisson_operator = 1;
    cfg.load_sign = -1;
    cfg.permittivity = eps;
    cfg.pcg_tol = 1e-12;
    cfg.pcg_max_iter = 1000000;
    cfg.sort_interval = 0;
    int rc = scdat_create(&cfg, &st.ctx);
    if (rc != SCDAT_OK) {
      st.unsupported = std::string("no device context: ") + (st.ctx ? scdat_last_error(st.ctx) : "create failed");
    } else {
      rc = scdat_set_mesh(st.ctx, (int64_t)nn, xyz.data(), (int64_t)(tets.size() / 4), tets.data(), kind.data(),
                          dval.data(), nullptr);
      if (rc != SCDAT_OK) st.unsupported = std::string("mesh upload: ") + scdat_last_error(st.ctx);
    }
    st.xyz = std::move(xyz);
    st.tets = std::move(tets);
    st.kind = std::move(kind);
    st.eps = eps;
  }
  st.assembled = true;
}

void PoissonSolver::assembleElementMatrix(Mesh::ElementPtr) {
  if (DeviceState* st = state_of(system_)) st->assembled = false;
}

void PoissonSolver::assembleElementRHS(Mesh::ElementPtr) {
  if (DeviceState* st = state_of(system_)) st->assembled = false;
}

void PoissonSolver::applyBoundaryConditions() {
  DeviceState* st = state_of(system_);
  if (!st) return;
  // NEUMANN values add to the right-hand side of their row (LinearSystem::
  // applyNeumannBC); DIRICHLET values are read by assembleSystem
  st->flux.assign(node_to_index_.size(), 0.0);
  for (const auto& [id, bc] : boundary_conditions_) {
    const auto it = node_to_index_.find(id);
    if (it == node_to_index_.end() || !bc) continue;
    if (bc->getType() == BoundaryConditionType::NEUMANN) st->flux[it->second] += bc->getValue();
  }
}

bool PoissonSolver::solve(SolverType) {
  DeviceState* st = state_of(system_);
  if (!st || !st->assembled) {
    assembleSystem();
    st = state_of(system_);
  }
  if (!st) {
    std::cerr << "PoissonSolver (device): no mesh to solve on" << std::endl;
    return false;
  }
  if (!st->unsupported.empty()) {
    std::cerr << "PoissonSolver (device): " << st->unsupported << std::endl;
    return false;
  }
  const std::size_t nn = st->load.size();
  // right-hand side -rho_i share_i + flux_i with the load sign -1: the
  // device's charge is rho_i share_i - flux_i
  std::vector<double> q(nn);
  for (std::size_t i = 0; i < nn; ++i) q[i] = st->load[i] - st->flux[i];
  // Dirichlet values of this assembly (the starting potential elsewhere)
  std::vector<double> phi0 = st->solved ? st->phi : std::vector<double>(nn, 0.0);
  for (std::size_t i = 0; i < nn; ++i)
    if (st->kind[i] == SCDAT_NODE_DIRICHLET) phi0[i] = st->dval[i];
  int rc = scdat_set_potential(st->ctx, phi0.data());
  if (rc != SCDAT_OK) return report("set_potential", st->ctx, rc);
  rc = scdat_set_rho(st->ctx, q.data());
  if (rc != SCDAT_OK) return report("set_rho", st->ctx, rc);
  rc = scdat_solve(st->ctx);
  if (rc != SCDAT_OK) return report("solve", st->ctx, rc);
  st->phi.assign(nn, 0.0);
  st->E.assign(3 * nn, 0.0);
  rc = scdat_get_fields(st->ctx, nullptr, st->phi.data(), st->E.data());
  if (rc != SCDAT_OK) return report("get_fields", st->ctx, rc);
  st->solved = true;
  updateNodePotentials();
  const auto nodes = getMeshNodes();
  for (const auto& [index, id] : index_to_node_)
    if (const auto n = node_by_id(nodes, id))
      n->setElectricField(Utils::Vector3D(st->E[3 * index], st->E[3 * index + 1], st->E[3 * index + 2]));
  return true;
}

double PoissonSolver::getPotential(Mesh::NodeId node_id) const {
  const auto it = node_to_index_.find(node_id);
  const DeviceState* st = state_of(system_);
  if (!st || !st->solved || it == node_to_index_.end() || it->second >= st->phi.size()) return 0.0;
  return st->phi[it->second];
}

Utils::Vector3D PoissonSolver::getElectricField(Mesh::NodeId node_id) const {
  const auto nodes = getMeshNodes();
  if (const auto n = node_by_id(nodes, node_id)) return n->getElectricField();
  const auto it = node_to_index_.find(node_id);
  if (!system_ || it == node_to_index_.end()) return Utils::Vector3D(0, 0, 0);
  return computeSimpleFiniteDifferenceGradient(it->second);
}

Utils::Vector3D PoissonSolver::getElectricField(const Utils::Point3D& position) const {
  initializeSpatialIndex();
  if (spatial_index_) {
    for (const auto& e : spatial_index_->findElementsContaining(position))
      if (e && e->contains(position)) return interpolateElectricFieldInElement(e, position);
  }
  return getElectricFieldNearestNeighbor(position);
}

Utils::Vector3D PoissonSolver::getElectricFieldAtPosition(const Utils::Point3D& position) const {
  return getElectricField(position);
}

void PoissonSolver::updateNodePotentials() {
  const DeviceState* st = state_of(system_);
  if (!st || !st->solved) return;
  const auto nodes = getMeshNodes();
  for (const auto& [index, id] : index_to_node_)
    if (const auto n = node_by_id(nodes, id)) n->setPotential(st->phi[index]);
}

void PoissonSolver::updateNodeElectricFields() { computeElectricFieldGradient(); }

void PoissonSolver::computeElectricFieldGradient() {
  const auto nodes = getMeshNodes();
  if (nodes.empty()) return;
  DeviceState* st = state_of(system_);
  if (!st || !st->ctx || !st->unsupported.empty() || st->kind.size() != node_to_index_.size()) {
    std::cerr << "PoissonSolver (device): no device system for this mesh (assemble it first"
              << (st && !st->unsupported.empty() ? "; " + st->unsupported : std::string()) << ")" << std::endl;
    return;
  }
  // the nodes' potentials to the device, E = -grad(phi) back
  const std::size_t nn = node_to_index_.size();
  std::vector<double> phi(nn, 0.0), E(3 * nn, 0.0);
  for (const auto& [index, id] : index_to_node_)
    if (const auto n = node_by_id(nodes, id)) phi[index] = n->getPotential();
  int rc = scdat_set_potential(st->ctx, phi.data());
  if (rc == SCDAT_OK) rc = scdat_get_fields(st->ctx, nullptr, nullptr, E.data());
  if (rc != SCDAT_OK) {
    report("field gradient", st->ctx, rc);
    return;
  }
  for (const auto& [index, id] : index_to_node_)
    if (const auto n = node_by_id(nodes, id)) n->setElectricField(Utils::Vector3D(E[3 * index], E[3 * index + 1], E[3 * index + 2]));
}

double PoissonSolver::getTotalEnergy() const {
  const auto nodes = getMeshNodes();
  if (nodes.empty()) return 0.0;
  std::unordered_map<Mesh::NodeId, double> cv;
  for (const auto& e : getMeshElements()) {
    if (!e || e->getNodes().empty()) continue;
    const double s = measure_of(e) / (double)e->getNodes().size();
    for (const auto& n : e->getNodes())
      if (n) cv[n->getId()] += s;
  }
  double w = 0.0;
  for (const auto& n : nodes) {
    if (!n) continue;
    const auto it = cv.find(n->getId());
    w += 0.5 * SCDAT_VACUUM_PERMITTIVITY * n->getElectricField().magnitudeSquared() * (it != cv.end() ? it->second : 0.0);
  }
  return w;
}

double PoissonSolver::getMaxPotential() const {
  const DeviceState* st = state_of(system_);
  if (!st || node_to_index_.empty()) return 0.0;
  if (!st->solved) return 0.0;
  return *std::max_element(st->phi.begin(), st->phi.end());
}

double PoissonSolver::getMinPotential() const {
  const DeviceState* st = state_of(system_);
  if (!st || node_to_index_.empty()) return 0.0;
  if (!st->solved) return 0.0;
  return *std::min_element(st->phi.begin(), st->phi.end());
}

Utils::Vector3D PoissonSolver::getMaxElectricField() const {
  Utils::Vector3D best(0, 0, 0);
  double best_mag = 0.0;
  for (const auto& [id, index] : node_to_index_) {
    (void)index;
    const Utils::Vector3D e = getElectricField(id);
    if (e.magnitude() > best_mag) {
      best_mag = e.magnitude();
      best = e;
    }
  }
  return best;
}

void PoissonSolver::printSolution() const {
  const DeviceState* st = state_of(system_);
  if (!st || !st->solved) return;
  for (std::size_t i = 0; i < st->phi.size(); ++i) std::cout << "phi[" << i << "] = " << st->phi[i] << "\n";
  std::cout.flush();
}

std::string PoissonSolver::toString() const {
  std::ostringstream s;
  s << "PoissonSolver(nodes=" << node_to_index_.size() << ", bcs=" << boundary_conditions_.size()
    << ", materials=" << material_properties_.size() << ", charges=" << charge_densities_.size() << ")";
  return s.str();
}

void PoissonSolver::buildNodeIndexMapping() {
  node_to_index_.clear();
  index_to_node_.clear();
  std::size_t next = 0;
  for (const auto& n : getMeshNodes()) {
    if (!n) continue;
    node_to_index_[n->getId()] = next;
    index_to_node_[next] = n->getId();
    ++next;
  }
}

MaterialPropertyPtr PoissonSolver::getMaterialProperty(Mesh::MaterialId material_id) const {
  const auto it = material_properties_.find(material_id);
  if (it != material_properties_.end()) return it->second;
  return std::make_shared<MaterialProperty>(material_id, Mesh::MaterialType::VACUUM, "default");
}

double PoissonSolver::getPermittivity(Mesh::ElementPtr element) const {
  return element ? getMaterialProperty(element->getMaterialId())->getPermittivity() : 1.0;
}

void PoissonSolver::assembleTetrahedronElement(Mesh::TetrahedronPtr tetrahedron) {
  assembleElementMatrix(tetrahedron);
  assembleElementRHS(tetrahedron);
}

void PoissonSolver::assembleTriangleElement(Mesh::TrianglePtr triangle) {
  assembleElementMatrix(triangle);
  assembleElementRHS(triangle);
}

void PoissonSolver::initializeSpatialIndex() const {
  if (spatial_index_ || !mesh_) return;
  spatial_index_ = std::make_shared<Mesh::SpatialIndex>();
  spatial_index_->initialize(mesh_->getElements());
}

// mean of the element's nodal fields (PoissonSolver.cpp:688-713)
Utils::Vector3D PoissonSolver::interpolateElectricFieldInElement(Mesh::ElementPtr element,
                                                                 const Utils::Point3D&) const {
  if (!element) return Utils::Vector3D(0, 0, 0);
  double sx = 0.0, sy = 0.0, sz = 0.0, c = 0.0;
  for (const auto& n : element->getNodes()) {
    if (!n) continue;
    const auto& e = n->getElectricField();
    sx += e.x();
    sy += e.y();
    sz += e.z();
    c += 1.0;
  }
  return c > 0.0 ? Utils::Vector3D(sx / c, sy / c, sz / c) : Utils::Vector3D(0, 0, 0);
}

Utils::Vector3D PoissonSolver::getElectricFieldNearestNeighbor(const Utils::Point3D& position) const {
  const auto nodes = getMeshNodes();
  if (nodes.empty()) return Utils::Vector3D(0, 0, 0);
  Mesh::NodeId nearest = nodes.front()->getId();
  double best = std::numeric_limits<double>::max();
  for (const auto& n : nodes) {
    if (!n) continue;
    const double d = n->getPosition().distanceTo(position);
    if (d < best) {
      best = d;
      nearest = n->getId();
    }
  }
  return getElectricField(nearest);
}

// gradient of the potential at node `index`: the device's least-squares
// gradient of the last solve (-E)
Utils::Vector3D PoissonSolver::computeSimpleFiniteDifferenceGradient(std::size_t index) const {
  const DeviceState* st = state_of(system_);
  if (!st || !st->solved || 3 * index + 2 >= st->E.size()) return Utils::Vector3D(0, 0, 0);
  return Utils::Vector3D(-st->E[3 * index], -st->E[3 * index + 1], -st->E[3 * index + 2]);
}

Utils::Vector3D PoissonSolver::computeSimpleGradientFallback(std::size_t index) const {
  return computeSimpleFiniteDifferenceGradient(index);
}

bool PoissonSolver::hasValidMesh() const { return mesh_ && !mesh_->getNodes().empty() && !mesh_->getElements().empty(); }

std::vector<Mesh::ElementPtr> PoissonSolver::getMeshElements() const {
  return mesh_ ? mesh_->getElements() : std::vector<Mesh::ElementPtr>{};
}

std::vector<Mesh::NodePtr> PoissonSolver::getMeshNodes() const {
  return mesh_ ? mesh_->getNodes() : std::vector<Mesh::NodePtr>{};
}

}  // namespace FieldSolver
}  // namespace SCDAT
