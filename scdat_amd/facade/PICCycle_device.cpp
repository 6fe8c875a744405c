// Device implementation of SCDAT::PICcore::PICCycle, the class the reference
// declares in Tools/PICcore/include/PICCycle.h (unchanged): build this file
// against the reference's headers and link it, with libscdat_amd.so, in
// place of Tools/PICcore/src/PICCycle.cpp.  Callers keep their code, their
// Mesh::Node / Mesh::Element objects, their ParticleManager and their
// BoundaryConditionBase objects; the cycle itself (push, gather, walk,
// boundaries, deposition, Poisson, nodal E) runs on the GPU through the C ABI.
//
// Per member (reference behaviour cited where it is mirrored):
//  * setMesh / setParticleManager / setPoissonSolver / setBoundaryCondition
//    store their arguments as PICCycle.cpp:110-200 does;
//  * initialize() checks the same prerequisites (PICCycle.cpp:255-309),
//    zeroes rho, phi and E, and builds the device context: nodes indexed by
//    their position in setMesh's vector, DIRICHLET nodes held at potential_,
//    tetrahedral elements, boundary faces grouped by the face of the
//    domain's bounding box they lie on (the reference's detectBoundaryFace,
//    PICCycle.cpp:1074-1126; faces off the box, such as a spacecraft
//    surface, form group 6, absorbing), the graph-Laplacian operator with
//    the reference's load sign (PoissonSolver.cpp:251-401);
//  * boundary conditions map to device kinds: AbsorbingBoundaryCondition
//    with probability 1, PeriodicBoundaryCondition (its period vector along
//    the face's axis), ReflectiveBoundaryCondition; no condition lets
//    particles escape (PICCycle.cpp:1254-1257); any other subclass runs on
//    the host (the slow path): the device hands back each step's impacts on
//    its face (SCDAT_BC_HOST) and applyRuntimeBoundaryCondition calls its
//    processParticle as PICCycle.cpp:1221-1293 does, records the ledger,
//    and the survivors and emitted particles return with the next upload;
//  * executeTimeStep() uploads the manager's active particles (the caller
//    may have changed them between steps, as PicMccSurfaceCurrentSampler
//    does with its host collisions), takes one device step with the E of
//    the previous step (E = 0 before the first solve, as the reference's
//    zero-initialised electric_field_), and writes positions, velocities
//    and statuses back into the manager, rho / phi / E into the members and
//    the shared Node objects (PICCycle.cpp:676-678,713-714), the ledger and
//    the statistics.  executeSimulation(T) runs its steps on the device with
//    one upload and one write-back;
//  * setPotential (inline in the header) calls updateElectricField, which
//    here sends potential_ to the device (Dirichlet values included, as
//    solvePoissonEquation reads them from potential_) and reads E back;
//    setNodePotential values reach the device before the next step.
// Departures: deposition is once per step at the end position (the
// reference's LinearSegmentCloud per sub-step is not used); executeTimeStep
// returns false when the Poisson solve missed its tolerance (the previous
// potential is kept, PICCycle.cpp:731-767); particles must carry at most 8
// distinct (charge, mass, weight) combinations.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <map>
#include <memory>
#include <optional>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "Boundary/include/BoundaryConditions.h"
#include "PICcore/include/PICCycle.h"
#include "Particle/include/ParticleManager.h"
#include "scdat_amd.h"

namespace SCDAT {
namespace PICcore {

namespace {

// The device context of one PICCycle.  It rides in cross_tetrahedron_ (the
// reference's host walker, which this implementation does not use) as the
// deleter of an empty shared pointer, so the cycle's defaulted destructor
// releases it.
struct DeviceState {
  scdat_ctx* ctx = nullptr;
  struct Species {
    double q, m, w;
    std::vector<std::size_t> members;  // container indices uploaded, in upload order
  };
  std::vector<Species> species;
  std::vector<int32_t> cell_of;        // last known cell per container index (-1 unknown)
  std::vector<double> sent_potential;  // potential_ as the device last received it
  std::array<int, 6> face_kind{};      // device kind per bounding-box face
  double surface_charge_seen = 0.0;    // device surface_deposited_charge already counted
  // what caller-side conditions (SCDAT_BC_HOST faces) recorded: added to the
  // device ledger, which does not count those faces
  std::array<PICCycle::SurfaceCurrentLedgerEntry, 6> host_ledger{};
  bool host_faces = false;
  bool ready = false;
  ~DeviceState() {
    if (ctx) scdat_destroy(ctx);
  }
};

struct StateHolder {
  std::shared_ptr<DeviceState> st;
  void operator()(CrossTetrahedron*) const {}
};

DeviceState* state_of(const CrossTetrahedronPtr& p) {
  StateHolder* h = std::get_deleter<StateHolder>(p);
  return h ? h->st.get() : nullptr;
}

CrossTetrahedronPtr make_state() {
  return CrossTetrahedronPtr(nullptr, StateHolder{std::make_shared<DeviceState>()});
}

bool report(const char* what, scdat_ctx* ctx, int rc) {
  std::cerr << "PICCycle (device): " << what << " failed: " << (ctx ? scdat_last_error(ctx) : "no context")
            << " (status " << rc << ")" << std::endl;
  return false;
}

// Boundary faces of the tetrahedra: face k (opposite vertex k) of tet t is
// on the boundary when no other tet shares its three nodes.
std::vector<int32_t> boundary_groups(const std::vector<int32_t>& tets, const std::vector<double>& xyz,
                                     const double lo[3], const double hi[3]) {
  const std::size_t nt = tets.size() / 4;
  std::map<std::array<int32_t, 3>, std::pair<int64_t, int>> seen;
  std::vector<int32_t> group(4 * nt, -1);
  std::vector<char> shared(4 * nt, 0);
  for (std::size_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      std::array<int32_t, 3> key;
      int m = 0;
      for (int q = 0; q < 4; ++q)
        if (q != k) key[m++] = tets[4 * t + q];
      std::sort(key.begin(), key.end());
      auto it = seen.find(key);
      if (it == seen.end()) {
        seen.emplace(key, std::make_pair((int64_t)t, k));
      } else {
        shared[4 * t + k] = 1;
        shared[4 * it->second.first + it->second.second] = 1;
      }
    }
  double span = 0.0;
  for (int r = 0; r < 3; ++r) span = std::max(span, hi[r] - lo[r]);
  const double tol = std::max(1e-10, 1e-8 * span);
  for (std::size_t t = 0; t < nt; ++t)
    for (int k = 0; k < 4; ++k) {
      if (shared[4 * t + k]) continue;
      int g = 6;  // off the bounding box
      for (int f = 0; f < 6 && g == 6; ++f) {
        const int axis = f / 2;
        const double plane = (f & 1) ? hi[axis] : lo[axis];
        bool on = true;
        for (int q = 0; q < 4 && on; ++q)
          if (q != k) on = std::abs(xyz[3 * (std::size_t)tets[4 * t + q] + axis] - plane) <= tol;
        if (on) g = f;
      }
      group[4 * t + k] = g;
    }
  return group;
}

bool upload_particles(DeviceState* st, ParticleManager& pm);

}  // namespace

PICCycle::PICCycle() : PICCycle(Parameters{}) {}

PICCycle::PICCycle(const Parameters& params) : params_(params) { cross_tetrahedron_ = make_state(); }

void PICCycle::setMesh(const std::vector<Mesh::NodePtr>& nodes, const std::vector<Mesh::ElementPtr>& elements) {
  nodes_ = nodes;
  elements_ = elements;
  if (DeviceState* st = state_of(cross_tetrahedron_)) st->ready = false;
}

void PICCycle::setParticleManager(std::shared_ptr<ParticleManager> particle_manager) {
  particle_manager_ = std::move(particle_manager);
}

void PICCycle::setPoissonSolver(std::shared_ptr<PoissonSolver> poisson_solver) {
  field_poisson_solver_ = std::move(poisson_solver);
}

void PICCycle::setBoundaryCondition(BoundaryFace face, BoundaryConditionPtr condition) {
  boundary_conditions_[static_cast<std::size_t>(face)] = std::move(condition);
  if (DeviceState* st = state_of(cross_tetrahedron_)) st->ready = false;
}

void PICCycle::clearBoundaryConditions() {
  boundary_conditions_.fill(nullptr);
  if (DeviceState* st = state_of(cross_tetrahedron_)) st->ready = false;
}

void PICCycle::addParticleLocation(ParticlePtr particle, ElementId element_id) {
  if (!particle) return;
  ParticleLocation loc;
  loc.particle = particle;
  loc.current_element = element_id;
  loc.is_valid = true;
  particle_locations_[particle->getId()] = loc;
}

void PICCycle::updateParticleLocation(ParticlePtr particle, ElementId new_element_id) {
  addParticleLocation(particle, new_element_id);
}

ElementId PICCycle::findParticleElement(ParticlePtr particle) const {
  if (!particle) return static_cast<ElementId>(-1);
  const auto it = particle_locations_.find(particle->getId());
  return it != particle_locations_.end() && it->second.is_valid ? it->second.current_element
                                                                 : static_cast<ElementId>(-1);
}

void PICCycle::updateDomainBounds() {
  domain_bounds_ = DomainBounds{};
  bool first = true;
  for (const auto& n : nodes_) {
    if (!n) continue;
    const auto& p = n->getPosition();
    if (first) {
      domain_bounds_.min_corner = p;
      domain_bounds_.max_corner = p;
      first = false;
      continue;
    }
    domain_bounds_.min_corner = Utils::Point3D(std::min(domain_bounds_.min_corner.x(), p.x()),
                                               std::min(domain_bounds_.min_corner.y(), p.y()),
                                               std::min(domain_bounds_.min_corner.z(), p.z()));
    domain_bounds_.max_corner = Utils::Point3D(std::max(domain_bounds_.max_corner.x(), p.x()),
                                               std::max(domain_bounds_.max_corner.y(), p.y()),
                                               std::max(domain_bounds_.max_corner.z(), p.z()));
  }
  domain_bounds_.valid = !first;
}

bool PICCycle::initialize() {
  if (nodes_.empty() || elements_.empty() || !particle_manager_ || !field_poisson_solver_) {
    std::cerr << "PICCycle initialization failed: missing required components" << std::endl;
    return false;
  }
  DeviceState* st = state_of(cross_tetrahedron_);
  if (!st) {
    cross_tetrahedron_ = make_state();
    st = state_of(cross_tetrahedron_);
  }
  const std::size_t nn = nodes_.size();
  charge_density_.assign(nn, 0.0);
  potential_.resize(nn, 0.0);
  electric_field_.assign(nn, Utils::Vector3D(0.0, 0.0, 0.0));
  magnetic_field_.assign(nn, Utils::Vector3D(0.0, 0.0, 0.0));
  cycle_stats_ = CycleStatistics();
  surface_current_ledger_.fill(SurfaceCurrentLedgerEntry{});
  for (std::size_t f = 0; f < surface_boundary_metadata_.size(); ++f) {
    if (surface_boundary_metadata_[f].surface_id < 0) surface_boundary_metadata_[f].surface_id = (int)f;
    surface_current_ledger_[f].metadata = surface_boundary_metadata_[f];
  }
  field_solver_solution_valid_ = false;
  updateDomainBounds();

  // mesh arrays (node index = position in nodes_)
  std::vector<double> xyz(3 * nn), dval(nn, 0.0);
  std::vector<int32_t> kind(nn, SCDAT_NODE_FREE);
  for (std::size_t i = 0; i < nn; ++i) {
    if (!nodes_[i]) {
      std::cerr << "PICCycle (device): null node " << i << std::endl;
      return false;
    }
    const auto& p = nodes_[i]->getPosition();
    xyz[3 * i] = p.x();
    xyz[3 * i + 1] = p.y();
    xyz[3 * i + 2] = p.z();
    if (nodes_[i]->getBoundaryType() == Mesh::BoundaryType::DIRICHLET) {
      kind[i] = SCDAT_NODE_DIRICHLET;
      dval[i] = potential_[i];
    }
  }
  std::unordered_map<Mesh::NodeId, int32_t> index_of;
  for (std::size_t i = 0; i < nn; ++i) index_of[nodes_[i]->getId()] = (int32_t)i;
  std::vector<int32_t> tets;
  tets.reserve(4 * elements_.size());
  for (const auto& e : elements_) {
    if (!e || e->getType() != Mesh::ElementType::TETRAHEDRON || e->getNodes().size() != 4) continue;
    for (const auto& n : e->getNodes()) {
      const auto it = index_of.find(n->getId());
      if (it == index_of.end()) {
        std::cerr << "PICCycle (device): element node outside setMesh's nodes" << std::endl;
        return false;
      }
      tets.push_back(it->second);
    }
  }
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  const std::vector<int32_t> group = boundary_groups(tets, xyz, lo, hi);

  // context
  if (st->ctx) scdat_destroy(st->ctx);
  st->ctx = nullptr;
  st->ready = false;
  scdat_config cfg;
  scdat_config_default(&cfg);
  cfg.precision = 1;  // the reference computes in FP64
  cfg.dt = params_.time_step;
  cfg.adaptive_substep = params_.adaptive_time_step ? 1 : 0;
  cfg.poisson_operator = 1;  // graph Laplacian of PoissonSolver
  cfg.load_sign = -1;        // and its load sign
  cfg.pcg_tol = 1e-10;
  cfg.sort_interval = 20;
  int rc = scdat_create(&cfg, &st->ctx);
  if (rc != SCDAT_OK) return report("scdat_create", nullptr, rc);
  scdat_ctx* ctx = st->ctx;
  rc = scdat_set_mesh(ctx, (int64_t)nn, xyz.data(), (int64_t)(tets.size() / 4), tets.data(), kind.data(),
                      dval.data(), group.data());
  if (rc != SCDAT_OK) return report("setMesh", ctx, rc);

  // boundary conditions of the six box faces; group 6 absorbs
  for (int f = 0; f < 6; ++f) {
    const BoundaryConditionPtr& bc = boundary_conditions_[f];
    int kind_f = SCDAT_BC_ESCAPE, partner = -1;
    double shift[3] = {0.0, 0.0, 0.0};
    if (!bc) {
      kind_f = SCDAT_BC_ESCAPE;
    } else if (auto* a = dynamic_cast<Particle::AbsorbingBoundaryCondition*>(bc.get())) {
      if (a->getAbsorptionProbability() < 1.0) {
        std::cerr << "PICCycle (device): absorption probability < 1 is not supported on the device" << std::endl;
        return false;
      }
      kind_f = SCDAT_BC_ABSORB;
    } else if (auto* p = dynamic_cast<Particle::PeriodicBoundaryCondition*>(bc.get())) {
      const int axis = f / 2;
      double len = 0.0;
      for (const auto& v : p->getPeriodVectors()) {
        const double c[3] = {v.x(), v.y(), v.z()};
        if (std::abs(c[axis]) > std::abs(len)) len = c[axis];
      }
      if (len == 0.0) {
        std::cerr << "PICCycle (device): periodic condition without a period along face " << f << std::endl;
        return false;
      }
      kind_f = SCDAT_BC_PERIODIC;
      partner = f ^ 1;
      shift[axis] = (f & 1) ? -std::abs(len) : std::abs(len);  // leaving the min face re-enters at max
    } else if (dynamic_cast<Particle::ReflectiveBoundaryCondition*>(bc.get())) {
      kind_f = SCDAT_BC_REFLECT;
    } else {
      // no device form: its processParticle runs on the host after each step
      kind_f = SCDAT_BC_HOST;
    }
    st->face_kind[f] = kind_f;
    st->host_faces = st->host_faces || kind_f == SCDAT_BC_HOST;
    rc = scdat_set_boundary(ctx, f, kind_f, partner, shift);
    if (rc != SCDAT_OK) return report("setBoundaryCondition", ctx, rc);
  }
  rc = scdat_set_boundary(ctx, 6, SCDAT_BC_ABSORB, -1, nullptr);
  if (rc != SCDAT_OK) return report("setBoundaryCondition", ctx, rc);

  // species: one per (charge, mass, weight)
  const auto& cont = particle_manager_->getContainer();
  std::map<std::tuple<double, double, double>, int> sid;
  st->species.clear();
  std::size_t npart = 0;
  for (const auto& p : cont) {
    ++npart;
    const auto key = std::make_tuple(p.getCharge(), p.getMass(), p.getWeight());
    if (!sid.count(key)) {
      if (sid.size() == 8) {
        std::cerr << "PICCycle (device): more than 8 (charge, mass, weight) combinations" << std::endl;
        return false;
      }
      const int s = (int)sid.size();
      sid[key] = s;
      st->species.push_back({p.getCharge(), p.getMass(), p.getWeight(), {}});
    }
  }
  for (auto& s : st->species) {
    int32_t id = -1;
    rc = scdat_add_species(ctx, s.q, s.m, s.w, (int64_t)std::max<std::size_t>(1024, 2 * npart), &id);
    if (rc != SCDAT_OK) return report("species", ctx, rc);
  }
  st->cell_of.assign(npart, -1);
  {
    std::size_t k = 0;
    for (const auto& p : cont) {
      const ElementId e = findParticleElement(const_cast<Particle::Particle*>(&p));
      if (e != static_cast<ElementId>(-1) && e < tets.size() / 4) st->cell_of[k] = (int32_t)e;
      ++k;
    }
  }
  // particles, the device's initial state, then the reference's initial
  // fields: E zero (electric_field_ as initialize leaves it), Dirichlet
  // values from potential_
  if (!upload_particles(st, *particle_manager_)) return false;
  rc = scdat_initialize(ctx);
  if (rc != SCDAT_OK && rc != SCDAT_ERR_SOLVER) return report("initialize", ctx, rc);
  const std::vector<double> zero(3 * nn, 0.0);
  rc = scdat_set_potential(ctx, potential_.data());
  if (rc == SCDAT_OK) rc = scdat_set_efield(ctx, zero.data());
  if (rc != SCDAT_OK) return report("initial field", ctx, rc);
  st->surface_charge_seen = 0.0;
  st->sent_potential = potential_;
  st->ready = true;
  if (params_.verbose)
    std::cout << "PIC cycle initialized on the device: " << nn << " nodes, " << tets.size() / 4 << " tetrahedra, "
              << npart << " particles" << std::endl;
  return true;
}

namespace {

// Active particles of the manager to the device, one species at a time.
bool upload_particles(DeviceState* st, ParticleManager& pm) {
  auto& cont = pm.getContainer();
  std::vector<std::vector<double>> x(st->species.size()), v(st->species.size());
  std::vector<std::vector<int32_t>> cell(st->species.size());
  for (auto& s : st->species) s.members.clear();
  if (st->cell_of.size() < cont.size()) st->cell_of.resize(cont.size(), -1);
  std::size_t k = 0;
  for (auto& p : cont) {
    const std::size_t idx = k++;
    if (!p.isActive()) continue;
    int s = -1;
    for (std::size_t q = 0; q < st->species.size(); ++q)
      if (st->species[q].q == p.getCharge() && st->species[q].m == p.getMass() && st->species[q].w == p.getWeight())
        s = (int)q;
    if (s < 0) return false;
    const auto pos = p.getPosition();
    const auto vel = p.getVelocity();
    x[s].insert(x[s].end(), {pos.x(), pos.y(), pos.z()});
    v[s].insert(v[s].end(), {vel.x(), vel.y(), vel.z()});
    cell[s].push_back(st->cell_of[idx]);
    st->species[s].members.push_back(idx);
  }
  for (std::size_t s = 0; s < st->species.size(); ++s) {
    const int64_t n = (int64_t)st->species[s].members.size();
    bool all_known = true;
    for (int32_t c : cell[s]) all_known = all_known && c >= 0;
    const int rc = scdat_set_particles(st->ctx, (int32_t)s, n, x[s].data(), v[s].data(),
                                       all_known ? cell[s].data() : nullptr);
    if (rc != SCDAT_OK) return report("particle upload", st->ctx, rc);
  }
  return true;
}

// Device particles back into the manager: moved particles get their new
// state, removed ones their terminal status; particles created on the device
// (none without collisions or emission) are appended.
void download_particles(DeviceState* st, ParticleManager& pm, const std::array<int, 6>& face_kind,
                        const double lo[3], const double hi[3],
                        const std::vector<std::vector<char>>* host_hit = nullptr) {
  auto& cont = pm.getContainer();
  std::vector<Particle::Particle*> by_index;
  by_index.reserve(cont.size());
  for (auto& p : cont) by_index.push_back(&p);
  for (std::size_t s = 0; s < st->species.size(); ++s) {
    int64_t alive = 0, slots = 0;
    if (scdat_get_particle_count(st->ctx, (int32_t)s, &alive, &slots) != SCDAT_OK) continue;
    std::vector<double> x(3 * slots), v(3 * slots);
    std::vector<int32_t> cell(slots), id(slots);
    if (scdat_get_particles(st->ctx, (int32_t)s, x.data(), v.data(), cell.data(), id.data()) != SCDAT_OK) continue;
    const auto& members = st->species[s].members;
    for (int64_t j = 0; j < slots; ++j) {
      const int32_t i = id[j];
      if (i < 0 || (std::size_t)i >= members.size()) {
        if (cell[j] >= 0) {  // created on the device
          const auto& sp = st->species[s];
          if (sp.q < 0.0)
            pm.createElectron(Utils::Point3D(x[3 * j], x[3 * j + 1], x[3 * j + 2]),
                              Utils::Vector3D(v[3 * j], v[3 * j + 1], v[3 * j + 2]), sp.w);
        }
        continue;
      }
      const std::size_t idx = members[(std::size_t)i];
      Particle::Particle* p = by_index[idx];
      p->setPosition(Utils::Point3D(x[3 * j], x[3 * j + 1], x[3 * j + 2]));
      p->setVelocity(Utils::Vector3D(v[3 * j], v[3 * j + 1], v[3 * j + 2]));
      st->cell_of[idx] = cell[j];
      if (cell[j] < 0) {
        // the face it left by: an escape face without a condition, else absorbed
        const double c[3] = {x[3 * j], x[3 * j + 1], x[3 * j + 2]};
        int f = -1;
        double best = 0.0;
        for (int q = 0; q < 6; ++q) {
          const double d = std::abs(c[q / 2] - ((q & 1) ? hi[q / 2] : lo[q / 2]));
          if (f < 0 || d < best) {
            f = q;
            best = d;
          }
        }
        // a particle stopped at a caller-side face: its condition decides
        if (host_hit && (std::size_t)i < (*host_hit)[s].size() && (*host_hit)[s][(std::size_t)i]) continue;
        if (f >= 0 && face_kind[f] == SCDAT_BC_ESCAPE) p->markAsEscaped();
        else p->markAsAbsorbed();
      }
    }
  }
}

}  // namespace

// Fields, ledger and statistics after device steps.
void PICCycle::updateStatistics() {
  DeviceState* st = state_of(cross_tetrahedron_);
  if (!st || !st->ctx) return;
  const std::size_t nn = nodes_.size();
  std::vector<double> E(3 * nn);
  if (scdat_get_fields(st->ctx, charge_density_.data(), potential_.data(), E.data()) == SCDAT_OK) {
    for (std::size_t i = 0; i < nn; ++i) {
      electric_field_[i] = Utils::Vector3D(E[3 * i], E[3 * i + 1], E[3 * i + 2]);
      nodes_[i]->setChargeDensity(charge_density_[i]);
      nodes_[i]->setPotential(potential_[i]);
      nodes_[i]->setElectricField(electric_field_[i]);
    }
    st->sent_potential = potential_;
    field_solver_solution_valid_ = true;
  }
  scdat_stats s;
  if (scdat_get_stats(st->ctx, &s) == SCDAT_OK) {
    cycle_stats_.completed_steps = (int)s.completed_steps;
    cycle_stats_.simulation_time = s.simulation_time;
    cycle_stats_.total_charge = s.total_charge;
    cycle_stats_.total_energy = s.total_energy;
    cycle_stats_.active_particles = (int)s.active_particles;
    cycle_stats_.surface_deposited_charge += s.surface_deposited_charge - st->surface_charge_seen;
    st->surface_charge_seen = s.surface_deposited_charge;
  }
  scdat_ledger_entry led[6];
  if (scdat_get_ledger(st->ctx, led, 6) == SCDAT_OK)
    for (int f = 0; f < 6; ++f) {
      SurfaceCurrentLedgerEntry& e = surface_current_ledger_[f];
      e.metadata = surface_boundary_metadata_[f];
      e.absorbed_electron_charge_c = led[f].absorbed_electron_charge_c;
      e.absorbed_ion_charge_c = led[f].absorbed_ion_charge_c;
      e.emitted_electron_charge_c = led[f].emitted_electron_charge_c;
      e.incident_energy_j = led[f].incident_energy_j;
      e.absorbed_electron_particles = (std::size_t)led[f].absorbed_electron_particles;
      e.absorbed_ion_particles = (std::size_t)led[f].absorbed_ion_particles;
      e.emitted_electron_particles = (std::size_t)led[f].emitted_electron_particles;
      const SurfaceCurrentLedgerEntry& h = st->host_ledger[f];
      e.absorbed_electron_charge_c += h.absorbed_electron_charge_c;
      e.absorbed_ion_charge_c += h.absorbed_ion_charge_c;
      e.emitted_electron_charge_c += h.emitted_electron_charge_c;
      e.incident_energy_j += h.incident_energy_j;
      e.absorbed_electron_particles += h.absorbed_electron_particles;
      e.absorbed_ion_particles += h.absorbed_ion_particles;
      e.emitted_electron_particles += h.emitted_electron_particles;
    }
}

// potential_ to the device: its Dirichlet nodes take these values in later
// solves (solvePoissonEquation reads them from potential_), and E is rebuilt
// from it (PoissonSolver.cpp:474-529 least squares).
void PICCycle::updateElectricField() {
  DeviceState* st = state_of(cross_tetrahedron_);
  if (!st || !st->ctx || !st->ready || potential_.size() != nodes_.size()) return;
  if (scdat_set_potential(st->ctx, potential_.data()) != SCDAT_OK) return;
  std::vector<double> E(3 * nodes_.size());
  if (scdat_get_fields(st->ctx, nullptr, nullptr, E.data()) != SCDAT_OK) return;
  for (std::size_t i = 0; i < nodes_.size(); ++i) {
    electric_field_[i] = Utils::Vector3D(E[3 * i], E[3 * i + 1], E[3 * i + 2]);
    if (nodes_[i]) nodes_[i]->setElectricField(electric_field_[i]);
  }
  st->sent_potential = potential_;
}

namespace {

// Steps on the device from the members' state: potential_ changes made by
// setNodePotential reach the Dirichlet values first (E is kept: the push
// uses the previous step's field), then the particles, then the steps.
int run_steps(DeviceState* st, std::vector<double>& potential,
              const std::vector<Utils::Vector3D>& efield, ParticleManager& pm, int n) {
  if (potential != st->sent_potential) {
    std::vector<double> E(3 * efield.size());
    for (std::size_t i = 0; i < efield.size(); ++i) {
      E[3 * i] = efield[i].x();
      E[3 * i + 1] = efield[i].y();
      E[3 * i + 2] = efield[i].z();
    }
    int rc = scdat_set_potential(st->ctx, potential.data());
    if (rc == SCDAT_OK) rc = scdat_set_efield(st->ctx, E.data());
    if (rc != SCDAT_OK) return rc;
    st->sent_potential = potential;
  }
  if (!upload_particles(st, pm)) return SCDAT_ERR_ARG;
  return scdat_step(st->ctx, n);
}

}  // namespace

// ---- caller-side boundary conditions (the slow path) -------------------------

namespace {

Utils::Vector3D inward_normal(PICCycle::BoundaryFace face) {
  const int f = static_cast<int>(face);
  double c[3] = {0.0, 0.0, 0.0};
  c[f / 2] = (f & 1) ? -1.0 : 1.0;
  return Utils::Vector3D(c[0], c[1], c[2]);
}

}  // namespace

// PICCycle.cpp:1074-1133: the face whose plane the position lies beyond by
// more than outside_tolerance (the farthest), or within near_tolerance of it.
std::optional<PICCycle::BoundaryFace> PICCycle::detectBoundaryFace(const Utils::Point3D& position,
                                                                   double outside_tolerance, bool allow_near_match,
                                                                   double near_tolerance) const {
  const double x[3] = {position.x(), position.y(), position.z()};
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  int best_f = -1;
  double best = outside_tolerance;
  for (int f = 0; f < 6; ++f) {
    const int a = f / 2;
    const double d = (f & 1) ? x[a] - hi[a] : lo[a] - x[a];
    if (d > best) {
      best = d;
      best_f = f;
    }
  }
  if (best_f < 0 && allow_near_match) {
    double nearest = near_tolerance;
    for (int f = 0; f < 6; ++f) {
      const int a = f / 2;
      const double d = std::abs((f & 1) ? hi[a] - x[a] : x[a] - lo[a]);
      if (d <= nearest) {
        nearest = d;
        best_f = f;
      }
    }
  }
  if (best_f < 0) return std::nullopt;
  return static_cast<BoundaryFace>(best_f);
}

// PICCycle.cpp:1135-1176: clamp into the box, onto the face's plane, then
// inward by the offset.
Utils::Point3D PICCycle::projectToBoundaryFace(BoundaryFace face, const Utils::Point3D& position,
                                               double inward_offset) const {
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  double x[3] = {position.x(), position.y(), position.z()};
  const int f = static_cast<int>(face);
  for (int r = 0; r < 3; ++r) x[r] = std::clamp(x[r], lo[r], hi[r]);
  x[f / 2] = (f & 1) ? hi[f / 2] : lo[f / 2];
  x[f / 2] += ((f & 1) ? -1.0 : 1.0) * inward_offset;
  for (int r = 0; r < 3; ++r) x[r] = std::clamp(x[r], lo[r], hi[r]);
  return Utils::Point3D(x[0], x[1], x[2]);
}

// PICCycle.cpp:1295-1329, into the host part of the ledger.
void PICCycle::recordSurfaceBoundaryInteraction(BoundaryFace face, const ParticleTypeDef& particle,
                                                const std::vector<ParticleTypeDef>& emitted_particles) {
  DeviceState* st = state_of(cross_tetrahedron_);
  if (!st) return;
  SurfaceCurrentLedgerEntry& e = st->host_ledger[static_cast<std::size_t>(face)];
  const double qw = particle.getCharge() * particle.getWeight();
  if (particle.getStatus() == ::SCDAT::Particle::ParticleStatus::ABSORBED) {
    e.incident_energy_j += particle.getKineticEnergy() * std::max(1.0, particle.getWeight());
    if (particle.getCharge() < 0.0) {
      e.absorbed_electron_charge_c += qw;
      e.absorbed_electron_particles += 1;
    } else {
      e.absorbed_ion_charge_c += qw;
      e.absorbed_ion_particles += 1;
    }
  }
  for (const auto& p : emitted_particles)
    if (p.getCharge() < 0.0) {
      e.emitted_electron_charge_c += p.getCharge() * p.getWeight();
      e.emitted_electron_particles += 1;
    }
}

// PICCycle.cpp:1221-1293 on a particle that stopped at a caller-side face:
// up to six passes of processParticle; an active particle's element is left
// unknown (the next upload locates it).
bool PICCycle::applyRuntimeBoundaryCondition(ParticlePtr particle, ElementId hint_element,
                                             std::vector<ParticleTypeDef>& emitted_particles,
                                             ElementId& resolved_element) {
  (void)hint_element;
  resolved_element = static_cast<ElementId>(-1);
  if (!particle || !domain_bounds_.valid) return false;
  const double span = std::max({domain_bounds_.max_corner.x() - domain_bounds_.min_corner.x(),
                                domain_bounds_.max_corner.y() - domain_bounds_.min_corner.y(),
                                domain_bounds_.max_corner.z() - domain_bounds_.min_corner.z(), 1.0e-12});
  const double outside_tol = std::max(1.0e-10, 1.0e-8 * span);
  const double near_tol = std::max(5.0 * outside_tol, 1.0e-9 * span);
  for (int pass = 0; pass < 6; ++pass) {
    if (!particle->isActive()) return true;
    const auto face = detectBoundaryFace(particle->getPosition(), outside_tol, pass == 0, near_tol);
    if (!face) return true;
    const auto& condition = boundary_conditions_[static_cast<std::size_t>(*face)];
    if (!condition) return false;
    auto generated = condition->processParticle(*particle, projectToBoundaryFace(*face, particle->getPosition()),
                                                inward_normal(*face), params_.time_step);
    emitted_particles.insert(emitted_particles.end(), generated.begin(), generated.end());
    recordSurfaceBoundaryInteraction(*face, *particle, generated);
    if (particle->getStatus() == ::SCDAT::Particle::ParticleStatus::ABSORBED) {
      cycle_stats_.surface_deposited_charge += particle->getCharge() * particle->getWeight();
      return true;
    }
    if (particle->getStatus() == ::SCDAT::Particle::ParticleStatus::ESCAPED) return true;
  }
  return true;
}

// PICCycle.cpp:1331-1379 over the particles that stopped at caller-side
// faces in the last step: each goes through applyRuntimeBoundaryCondition;
// one that stays active is moved a hair inside the box (its crossing point
// lies on the face) and located by the next upload; a failed condition
// marks it escaped.
void PICCycle::processRuntimeBoundaryEvents(const std::vector<RuntimeBoundaryEvent>& events,
                                            std::vector<ParticleTypeDef>& emitted_particles) {
  if (!particle_manager_ || events.empty()) return;
  auto& container = particle_manager_->getContainer();
  auto it_begin = container.begin();
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  const double hair = 1e-12 * std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]});
  for (const auto& event : events) {
    if (event.particle_index >= container.size()) continue;
    auto& particle = *(it_begin + event.particle_index);
    if (!particle.isActive()) continue;
    ElementId resolved = static_cast<ElementId>(-1);
    if (!applyRuntimeBoundaryCondition(&particle, event.hint_element, emitted_particles, resolved)) {
      particle.markAsEscaped();
      continue;
    }
    if (particle.isActive()) {
      const auto pos = particle.getPosition();
      double x[3] = {pos.x(), pos.y(), pos.z()};
      for (int r = 0; r < 3; ++r) x[r] = std::clamp(x[r], lo[r] + hair, hi[r] - hair);
      particle.setPosition(Utils::Point3D(x[0], x[1], x[2]));
    }
  }
}

bool PICCycle::executeTimeStep() {
  DeviceState* st = state_of(cross_tetrahedron_);
  if (!st || !st->ready) {
    if (!initialize()) return false;
    st = state_of(cross_tetrahedron_);
  }
  const int rc = run_steps(st, potential_, electric_field_, *particle_manager_, 1);
  if (rc != SCDAT_OK && rc != SCDAT_ERR_SOLVER) return report("executeTimeStep", st->ctx, rc);
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  if (!st->host_faces) {
    download_particles(st, *particle_manager_, st->face_kind, lo, hi);
  } else {
    // the step's impacts on caller-side faces, in (species, id) order
    int64_t n = 0;
    if (scdat_get_host_impacts(st->ctx, 0, nullptr, &n) != SCDAT_OK) return report("host impacts", st->ctx, 1);
    std::vector<scdat_impact> imp((std::size_t)n);
    if (n && scdat_get_host_impacts(st->ctx, n, imp.data(), &n) != SCDAT_OK)
      return report("host impacts", st->ctx, 1);
    std::vector<std::vector<char>> hit(st->species.size());
    for (std::size_t q = 0; q < st->species.size(); ++q) hit[q].assign(st->species[q].members.size(), 0);
    for (const scdat_impact& e : imp)
      if (e.species >= 0 && (std::size_t)e.species < hit.size() && e.id >= 0 &&
          (std::size_t)e.id < hit[(std::size_t)e.species].size())
        hit[(std::size_t)e.species][(std::size_t)e.id] = 1;
    download_particles(st, *particle_manager_, st->face_kind, lo, hi, &hit);
    std::vector<RuntimeBoundaryEvent> events;
    for (const scdat_impact& e : imp) {
      const auto& members = st->species[(std::size_t)e.species].members;
      if (e.id < 0 || (std::size_t)e.id >= members.size()) continue;
      RuntimeBoundaryEvent ev;
      ev.particle_index = members[(std::size_t)e.id];
      ev.hint_element = static_cast<ElementId>(e.cell);
      st->cell_of[ev.particle_index] = -1;
      events.push_back(ev);
    }
    std::vector<ParticleTypeDef> emitted;
    processRuntimeBoundaryEvents(events, emitted);
    // emitted particles join the manager; they need one of the cycle's
    // (charge, mass, weight) species to reach the device
    for (const auto& p : emitted) particle_manager_->getContainer().addParticle(p);
  }
  updateStatistics();
  return rc == SCDAT_OK;
}

bool PICCycle::executeSimulation(double total_time) {
  if (!initialize()) return false;
  DeviceState* st = state_of(cross_tetrahedron_);
  const int n = (int)std::llround(total_time / params_.time_step);
  if (n <= 0) return true;
  if (st->host_faces) {
    // caller-side conditions see every step's impacts
    bool ok = true;
    for (int k = 0; k < n; ++k) ok = executeTimeStep() && ok;
    return ok;
  }
  const int rc = run_steps(st, potential_, electric_field_, *particle_manager_, n);
  if (rc != SCDAT_OK && rc != SCDAT_ERR_SOLVER) return report("executeSimulation", st->ctx, rc);
  const double lo[3] = {domain_bounds_.min_corner.x(), domain_bounds_.min_corner.y(), domain_bounds_.min_corner.z()};
  const double hi[3] = {domain_bounds_.max_corner.x(), domain_bounds_.max_corner.y(), domain_bounds_.max_corner.z()};
  download_particles(st, *particle_manager_, st->face_kind, lo, hi);
  updateStatistics();
  return rc == SCDAT_OK;
}

}  // namespace PICcore
}  // namespace SCDAT
