// C++ mirror of SCDAT::PICcore::PICCycle (Tools/PICcore/include/PICCycle.h:
// 33-185) over the gfx950 C ABI.  Same public names, parameter meaning and
// call order; mesh and particles are passed as plain arrays instead of the
// reference's shared Node/Element/Particle objects, and errors throw
// std::runtime_error with the library's message where the reference returns
// false and prints to std::cerr.  Header-only; link with libscdat_amd.so.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/scdat_amd.h"

namespace scdat_amd {
namespace PICcore {

struct Vector3D {
  double x = 0.0, y = 0.0, z = 0.0;
};

// Particle boundary conditions available on the device
// (Tools/Boundary/include/BoundaryConditions.h: Periodic, Absorbing; plus
// specular reflection, floating conductor and free escape).
enum class BoundaryKind {
  Absorbing = SCDAT_BC_ABSORB,
  Periodic = SCDAT_BC_PERIODIC,
  Reflective = SCDAT_BC_REFLECT,
  Conductor = SCDAT_BC_CONDUCTOR,
  Escape = SCDAT_BC_ESCAPE,
};

class PICCycle {
 public:
  enum class BoundaryFace { XMin = 0, XMax = 1, YMin = 2, YMax = 3, ZMin = 4, ZMax = 5 };

  struct Parameters {
    double time_step = 1e-12;
    int max_iterations = 1000;
    double convergence_tolerance = 1e-6;
    bool adaptive_time_step = false;
    bool charge_conservation = true;  // always: fixed-point deposition
    int statistics_frequency = 10;
    bool verbose = false;
    // device controls
    int precision = 0;               // 0 mixed FP32/FP64, 1 FP64
    double pcg_tolerance = 1e-8;     // relative, ||r|| <= tol ||b||
    double permittivity = 8.8541878128e-12;
    int sort_interval = 20;
  };

  struct CycleStatistics {
    int completed_steps = 0;
    double simulation_time = 0.0;
    double total_charge = 0.0;
    double total_energy = 0.0;
    int active_particles = 0;
    double average_step_time = 0.0;
    double surface_deposited_charge = 0.0;
    // device additions
    double conductor_charge = 0.0;
    double floating_potential = 0.0;
    int pcg_iterations = 0;
    bool pcg_converged = true;
  };

  struct SurfaceBoundaryMetadata {
    int surface_id = -1;
    int material_id = -1;
    int circuit_node_id = -1;
  };

  struct SurfaceCurrentLedgerEntry {
    SurfaceBoundaryMetadata metadata;
    double absorbed_electron_charge_c = 0.0;
    double absorbed_ion_charge_c = 0.0;
    double emitted_electron_charge_c = 0.0;
    double incident_energy_j = 0.0;
    std::size_t absorbed_electron_particles = 0;
    std::size_t absorbed_ion_particles = 0;
    std::size_t emitted_electron_particles = 0;
  };

  PICCycle() : PICCycle(Parameters{}) {}
  explicit PICCycle(const Parameters& p) : params_(p) {
    scdat_config c;
    scdat_config_default(&c);
    c.precision = p.precision;
    c.dt = p.time_step;
    c.adaptive_substep = p.adaptive_time_step ? 1 : 0;
    c.permittivity = p.permittivity;
    c.pcg_tol = p.pcg_tolerance;
    c.sort_interval = p.sort_interval;
    const int rc = scdat_create(&c, &ctx_);
    if (rc != SCDAT_OK) throw std::runtime_error("scdat_create failed (" + std::to_string(rc) + ")");
  }
  ~PICCycle() {
    if (ctx_) scdat_destroy(ctx_);
  }
  PICCycle(const PICCycle&) = delete;
  PICCycle& operator=(const PICCycle&) = delete;
  PICCycle(PICCycle&& o) noexcept { *this = std::move(o); }
  PICCycle& operator=(PICCycle&& o) noexcept {
    std::swap(ctx_, o.ctx_);
    std::swap(params_, o.params_);
    std::swap(n_nodes_, o.n_nodes_);
    std::swap(lo_, o.lo_);
    std::swap(hi_, o.hi_);
    return *this;
  }

  // ---- setup (PICCycle.h:63-69) ---------------------------------------------
  // xyz[3 n], tets[4 t], node_kind[n] (0 free, 1 Dirichlet, 2 conductor),
  // dirichlet[n] (optional), face_group[4 t] (face k opposite vertex k; groups
  // 0..5 = XMin..ZMax walls, 6..11 = conductor faces).
  void setMesh(const std::vector<double>& xyz, const std::vector<int32_t>& tets,
               const std::vector<int32_t>& node_kind, const std::vector<double>& dirichlet = {},
               const std::vector<int32_t>& face_group = {}) {
    n_nodes_ = node_kind.size();
    check(scdat_set_mesh(ctx_, (int64_t)n_nodes_, xyz.data(), (int64_t)(tets.size() / 4), tets.data(),
                         node_kind.data(), dirichlet.empty() ? nullptr : dirichlet.data(),
                         face_group.empty() ? nullptr : face_group.data()),
          "setMesh");
    for (int r = 0; r < 3; ++r) {
      lo_[r] = 1e300;
      hi_[r] = -1e300;
    }
    for (std::size_t i = 0; i < n_nodes_; ++i)
      for (int r = 0; r < 3; ++r) {
        lo_[r] = std::min(lo_[r], xyz[3 * i + r]);
        hi_[r] = std::max(hi_[r], xyz[3 * i + r]);
      }
  }

  // A wall face's condition.  Periodic faces are paired with the opposite
  // face and shifted by the mesh's bounding-box extent, the periods the
  // reference's PeriodicBoundaryCondition is given (PICCycleBoundary_test.cpp:
  // 84-87).
  void setBoundaryCondition(BoundaryFace face, BoundaryKind kind) {
    const int g = (int)face, ax = g / 2;
    double shift[3] = {0.0, 0.0, 0.0};
    int partner = -1;
    if (kind == BoundaryKind::Periodic) {
      partner = g ^ 1;
      shift[ax] = (g & 1 ? -1.0 : 1.0) * (hi_[ax] - lo_[ax]);
    }
    setBoundaryGroup(g, kind, partner, shift);
  }
  void setBoundaryGroup(int group, BoundaryKind kind, int partner_group = -1,
                        const double* shift = nullptr) {
    check(scdat_set_boundary(ctx_, group, (int32_t)kind, partner_group, shift), "setBoundaryCondition");
  }
  // No condition on any wall: particles leaving are escaped, with no ledger
  // entry (PICCycle.cpp:1254-1257).
  void clearBoundaryConditions() {
    for (int g = 0; g < 6; ++g) setBoundaryGroup(g, BoundaryKind::Escape);
  }
  void setBoundaryMetadata(BoundaryFace face, const SurfaceBoundaryMetadata& m) {
    ledger_[(std::size_t)face].metadata = m;
  }

  // ---- particles (ParticleManager::createElectron / createIon over arrays,
  // ParticleManager.cpp:707-729).  Returns the species index.
  int addSpecies(double charge, double mass, double weight, int64_t capacity) {
    int32_t id = -1;
    check(scdat_add_species(ctx_, charge, mass, weight, capacity, &id), "addSpecies");
    return id;
  }
  void setParticles(int species, const std::vector<Vector3D>& x, const std::vector<Vector3D>& v) {
    check(scdat_set_particles(ctx_, species, (int64_t)x.size(), &x.data()->x, &v.data()->x, nullptr),
          "setParticles");
  }
  int addElectrons(const std::vector<Vector3D>& x, const std::vector<Vector3D>& v, double weight,
                   int64_t capacity = 0) {
    const int s = addSpecies(-SCDAT_ELEMENTARY_CHARGE, SCDAT_ELECTRON_MASS, weight,
                             capacity ? capacity : (int64_t)x.size());
    setParticles(s, x, v);
    return s;
  }
  // Ion mass A m_p, as ParticleFactory::createIon (ParticleDefinitions.cpp:386-388).
  int addIons(const std::vector<Vector3D>& x, const std::vector<Vector3D>& v, double weight,
              int mass_number = 1, int charge_number = 1, int64_t capacity = 0) {
    const int s = addSpecies(charge_number * SCDAT_ELEMENTARY_CHARGE, mass_number * SCDAT_PROTON_MASS, weight,
                             capacity ? capacity : (int64_t)x.size());
    setParticles(s, x, v);
    return s;
  }
  // Device order; cell -1 = removed (absorbed, escaped, recombined).
  void getParticles(int species, std::vector<Vector3D>& x, std::vector<Vector3D>& v,
                    std::vector<int32_t>& cell, std::vector<int32_t>& id) const {
    int64_t slots = 0;
    check(scdat_get_particle_count(ctx_, species, nullptr, &slots), "getParticles");
    x.resize(slots);
    v.resize(slots);
    cell.resize(slots);
    id.resize(slots);
    check(scdat_get_particles(ctx_, species, &x.data()->x, &v.data()->x, cell.data(), id.data()),
          "getParticles");
  }
  void setConductorCharge(double q) { check(scdat_set_conductor_charge(ctx_, q), "setConductorCharge"); }

  // ---- cycle (PICCycle.h:71-73) ---------------------------------------------
  bool initialize() {
    check(scdat_initialize(ctx_), "initialize");
    dirty_ = true;
    return true;
  }
  // false when the Poisson solve missed its tolerance (the field is still
  // updated from the best iterate, see the statistics)
  bool executeTimeStep() { return run(1); }
  bool executeSimulation(double total_time) {
    const long n = std::lround(total_time / params_.time_step);
    return n > 0 ? run((int)n) : true;
  }

  // ---- fields (PICCycle.h:73-117) -------------------------------------------
  const std::vector<double>& getChargeDensity() const {
    sync();
    return rho_;
  }
  const std::vector<double>& getPotential() const {
    sync();
    return phi_;
  }
  const std::vector<Vector3D>& getElectricField() const {
    sync();
    return E_;
  }
  void setTimeStep(double dt) {
    if (dt > 0.0) {
      params_.time_step = dt;
      check(scdat_set_time_step(ctx_, dt), "setTimeStep");
    }
  }
  double getTimeStep() const { return params_.time_step; }
  // Sets the potential and recomputes the nodal field (PICCycle.h:90-105).
  void setPotential(const std::vector<double>& phi) {
    check(scdat_set_potential(ctx_, phi.data()), "setPotential");
    dirty_ = true;
  }
  void setNodePotential(std::size_t node, double value) {
    std::vector<double> phi = getPotential();
    if (node < phi.size()) {
      phi[node] = value;
      setPotential(phi);
    }
  }

  // ---- statistics (PICCycle.h:135-185) --------------------------------------
  const CycleStatistics& getStatistics() const {
    scdat_stats s;
    check(scdat_get_stats(ctx_, &s), "getStatistics");
    stats_.completed_steps = (int)s.completed_steps;
    stats_.simulation_time = s.simulation_time;
    stats_.total_charge = s.total_charge;
    stats_.total_energy = s.total_energy;
    stats_.active_particles = (int)s.active_particles;
    stats_.surface_deposited_charge = s.surface_deposited_charge - surface_mark_;
    stats_.conductor_charge = s.conductor_charge;
    stats_.floating_potential = s.floating_potential;
    stats_.pcg_iterations = s.pcg_iterations;
    stats_.pcg_converged = s.pcg_converged != 0;
    return stats_;
  }
  const std::array<SurfaceCurrentLedgerEntry, 6>& getSurfaceCurrentLedger() const {
    scdat_ledger_entry e[6];
    check(scdat_get_ledger(ctx_, e, 6), "getSurfaceCurrentLedger");
    for (int g = 0; g < 6; ++g) {
      SurfaceCurrentLedgerEntry& L = ledger_[g];
      L.absorbed_electron_charge_c = e[g].absorbed_electron_charge_c;
      L.absorbed_ion_charge_c = e[g].absorbed_ion_charge_c;
      L.emitted_electron_charge_c = e[g].emitted_electron_charge_c;
      L.incident_energy_j = e[g].incident_energy_j;
      L.absorbed_electron_particles = (std::size_t)e[g].absorbed_electron_particles;
      L.absorbed_ion_particles = (std::size_t)e[g].absorbed_ion_particles;
      L.emitted_electron_particles = (std::size_t)e[g].emitted_electron_particles;
    }
    return ledger_;
  }
  double getAndResetSurfaceCharge() {
    const double q = getStatistics().surface_deposited_charge;
    surface_mark_ += q;
    stats_.surface_deposited_charge = 0.0;
    return q;
  }

  scdat_ctx* handle() const { return ctx_; }

 private:
  void check(int rc, const char* what) const {
    if (rc != SCDAT_OK)
      throw std::runtime_error(std::string(what) + ": " + scdat_last_error(ctx_) + " (" + std::to_string(rc) + ")");
  }
  bool run(int n) {
    const int rc = scdat_step(ctx_, n);
    dirty_ = true;
    if (rc == SCDAT_ERR_SOLVER) return false;
    check(rc, "executeTimeStep");
    return true;
  }
  void sync() const {
    if (!dirty_) return;
    rho_.resize(n_nodes_);
    phi_.resize(n_nodes_);
    E_.resize(n_nodes_);
    check(scdat_get_fields(ctx_, rho_.data(), phi_.data(), &E_.data()->x), "get fields");
    dirty_ = false;
  }

  scdat_ctx* ctx_ = nullptr;
  Parameters params_;
  std::size_t n_nodes_ = 0;
  double lo_[3] = {0, 0, 0}, hi_[3] = {0, 0, 0};
  double surface_mark_ = 0.0;
  mutable bool dirty_ = true;
  mutable std::vector<double> rho_, phi_;
  mutable std::vector<Vector3D> E_;
  mutable CycleStatistics stats_;
  mutable std::array<SurfaceCurrentLedgerEntry, 6> ledger_{};
};

static_assert(sizeof(Vector3D) == 3 * sizeof(double), "Vector3D must be three packed doubles");

}  // namespace PICcore
}  // namespace scdat_amd
