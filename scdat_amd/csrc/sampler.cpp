// Live-PIC surface current sampler: a short PIC-MCC window in a plate gap
// with the sampled surface at z = 0 held at the surface potential, the far
// plane at the plasma reference, periodic sides and absorbing z faces; the
// collected charge at z = 0 over the window gives the electron, ion and net
// current densities.  With a probe step the window runs at V, V + d and V - d
// on three device contexts at once (three host threads, three streams) and
// the central difference gives dJ/dV.
//
// Follows PicMccSurfaceCurrentSampler::sample / sampleWithDerivative
// ("Toolkit/Surface Charging/src/PicMccSurfaceCurrentSampler.cpp":489-670):
// plate of area A and gap h (createSamplingMesh :190-231), stable step
// h / (40 max speed) in [5e-12, 5e-9] s (computeStableTimeStep :176-188),
// Maxwellian electrons and ions with v_z directed at the surface
// (sampleVelocity :242-253), per-element macro-particle count, collisions of
// the SurfacePicV1 set (CollisionPicAdapter.cpp:129-149) when the neutral
// density exceeds 1e6 m^-3, and ledger currents at z = 0 (:602-617).
// Differences: positions come from the Philox loader (uniform in the gap, not
// per element), ions take the electrons' macro weight with their count
// scaled to the ion density (the device's fixed-point charge needs equal
// weights), and the mesh is the conforming 6-tet split.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/scdat_amd.h"

namespace {

constexpr double kAmu = 1.66053906660e-27;  // Tools/Basic/include/Constants.h:50

struct Window {
  scdat_sampler_config cfg;
  scdat_current_sample out{};
  int rc = SCDAT_OK;
  std::string err;
};

double stable_dt(const scdat_sampler_config& c) {
  const double gap = std::max(1e-6, c.gap_distance_m);
  const double ve = std::sqrt(std::max(1e-3, c.electron_temperature_ev) * SCDAT_ELEMENTARY_CHARGE /
                              SCDAT_ELECTRON_MASS);
  const double mi = std::max(1.0, c.ion_mass_amu) * kAmu;
  const double vi = std::sqrt(std::max(1e-3, c.ion_temperature_ev) * SCDAT_ELEMENTARY_CHARGE / mi);
  const double vmax = std::max(1.0, std::max(ve, vi));
  return std::min(std::max(gap / (40.0 * vmax), 5e-12), 5e-9);
}

#define WCHK(call)                                        \
  do {                                                    \
    const int rc_ = (call);                               \
    if (rc_ != SCDAT_OK) {                                \
      w.rc = rc_;                                         \
      w.err = std::string(#call) + ": " + (ctx ? scdat_last_error(ctx) : ""); \
      if (ctx) scdat_destroy(ctx);                        \
      return;                                             \
    }                                                     \
  } while (0)

void run_window(Window& w) {
  const scdat_sampler_config& c = w.cfg;
  scdat_ctx* ctx = nullptr;
  scdat_box_spec spec{};
  scdat_sampler_plate(&c, &spec);
  int64_t nn = 0, nt = 0;
  WCHK(scdat_box_mesh_size(&spec, &nn, &nt));
  std::vector<double> xyz(3 * nn);
  std::vector<int32_t> tets(4 * nt), kind(nn), group(4 * nt);
  WCHK(scdat_box_mesh_build(&spec, xyz.data(), tets.data(), kind.data(), group.data()));
  std::vector<double> dval(nn, 0.0);
  for (int64_t i = 0; i < nn; ++i) {
    const double z = xyz[3 * i + 2];
    kind[i] = SCDAT_NODE_FREE;
    if (std::abs(z - spec.lo[2]) < 1e-12 * std::max(1.0, spec.hi[2])) {
      kind[i] = SCDAT_NODE_DIRICHLET;
      dval[i] = c.surface_potential_v;
    } else if (std::abs(z - spec.hi[2]) < 1e-12 * std::max(1.0, spec.hi[2])) {
      kind[i] = SCDAT_NODE_DIRICHLET;
      dval[i] = c.plasma_reference_potential_v;
    }
  }
  int64_t ne = 0, ni = 0;
  double weight = 0.0;
  scdat_sampler_counts(&c, &ne, &ni, &weight);
  scdat_config cfg;
  scdat_config_default(&cfg);
  cfg.precision = c.precision;
  cfg.dt = c.time_step > 0.0 ? c.time_step : stable_dt(c);
  cfg.adaptive_substep = 1;
  cfg.sort_interval = 0;
  WCHK(scdat_create(&cfg, &ctx));
  WCHK(scdat_set_mesh(ctx, nn, xyz.data(), nt, tets.data(), kind.data(), dval.data(), group.data()));
  const double Lx = spec.hi[0] - spec.lo[0], Ly = spec.hi[1] - spec.lo[1];
  const double sx[3] = {Lx, 0, 0}, mx[3] = {-Lx, 0, 0}, sy[3] = {0, Ly, 0}, my[3] = {0, -Ly, 0};
  WCHK(scdat_set_boundary(ctx, 0, SCDAT_BC_PERIODIC, 1, sx));
  WCHK(scdat_set_boundary(ctx, 1, SCDAT_BC_PERIODIC, 0, mx));
  WCHK(scdat_set_boundary(ctx, 2, SCDAT_BC_PERIODIC, 3, sy));
  WCHK(scdat_set_boundary(ctx, 3, SCDAT_BC_PERIODIC, 2, my));
  WCHK(scdat_set_boundary(ctx, 4, SCDAT_BC_ABSORB, -1, nullptr));
  WCHK(scdat_set_boundary(ctx, 5, SCDAT_BC_ABSORB, -1, nullptr));
  const double mi = std::max(1.0, c.ion_mass_amu) * kAmu;
  const bool mcc = c.enable_mcc && c.neutral_density_m3 > 1e6;
  const int64_t cap_e = mcc ? 4 * ne + 64 : ne, cap_i = mcc ? 2 * ni + 64 : ni;
  int32_t se = -1, si = -1, sp = -1;
  WCHK(scdat_add_species(ctx, -SCDAT_ELEMENTARY_CHARGE, SCDAT_ELECTRON_MASS, weight, cap_e, &se));
  WCHK(scdat_add_species(ctx, SCDAT_ELEMENTARY_CHARGE, mi, weight, cap_i, &si));
  if (mcc) WCHK(scdat_add_species(ctx, SCDAT_ELEMENTARY_CHARGE, 40.0 * SCDAT_PROTON_MASS, weight, 2 * ne + 64, &sp));
  const double vth[2] = {std::sqrt(std::max(1e-3, c.electron_temperature_ev) * SCDAT_ELEMENTARY_CHARGE /
                                   SCDAT_ELECTRON_MASS),
                         std::sqrt(std::max(1e-3, c.ion_temperature_ev) * SCDAT_ELEMENTARY_CHARGE / mi)};
  const int64_t cnt[2] = {ne, ni};
  for (int s = 0; s < 2; ++s) {
    std::vector<double> x(3 * cnt[s]), v(3 * cnt[s]);
    std::vector<int32_t> cell(cnt[s]);
    WCHK(scdat_load_maxwellian(&spec, cnt[s], c.seed, (uint32_t)s, vth[s], nullptr, x.data(), v.data(),
                               cell.data()));
    scdat_sampler_direct(cnt[s], v.data());
    WCHK(scdat_set_particles(ctx, s == 0 ? se : si, cnt[s], x.data(), v.data(), cell.data()));
  }
  if (mcc) {
    // SurfacePicV1 (CollisionPicAdapter.cpp:129-149): elastic 2e-19 m^2,
    // excitation 11.5 eV / 2.5e-20, ionisation 15.76 eV / 3e-20 with Ar+.
    scdat_mcc_process p[3];
    scdat_mcc_process_default(SCDAT_MCC_ELASTIC, &p[0]);
    p[0].reference_energy_ev = 1.0;
    p[0].power_index = 0.5;
    scdat_mcc_process_default(SCDAT_MCC_EXCITATION, &p[1]);
    scdat_mcc_process_default(SCDAT_MCC_IONIZATION, &p[2]);
    for (auto& q : p) {
      q.species = se;
      q.target_density = c.neutral_density_m3;
    }
    p[2].product_electron_species = se;
    p[2].product_ion_species = sp;
    WCHK(scdat_mcc_configure(ctx, p, 3));
    WCHK(scdat_set_seed(ctx, c.seed ^ 0x9E3779B9ull));
  }
  WCHK(scdat_initialize(ctx));
  const int steps = std::max(1, c.window_steps);
  const int rc = scdat_step(ctx, steps);
  if (rc != SCDAT_OK && rc != SCDAT_ERR_SOLVER) WCHK(rc);
  scdat_ledger_entry led[6];
  WCHK(scdat_get_ledger(ctx, led, 6));
  scdat_event_counts ev;
  WCHK(scdat_get_event_counts(ctx, &ev));
  const double T = cfg.dt * steps;
  const double area = std::max(1e-12, Lx * Ly);
  scdat_current_sample& o = w.out;
  o.sampled_surface_potential_v = c.surface_potential_v;
  o.electron_collection_current_density_a_per_m2 = led[4].absorbed_electron_charge_c / (T * area);
  o.ion_collection_current_density_a_per_m2 = led[4].absorbed_ion_charge_c / (T * area);
  o.emitted_electron_current_density_a_per_m2 = -led[4].emitted_electron_charge_c / (T * area);
  o.net_collection_current_density_a_per_m2 = o.electron_collection_current_density_a_per_m2 +
                                              o.ion_collection_current_density_a_per_m2 +
                                              o.emitted_electron_current_density_a_per_m2;
  int64_t tot = 0;
  for (int k = 0; k < SCDAT_MAX_MCC_PROCESSES; ++k) tot += ev.collisions[k];
  o.total_collisions = tot;
  o.time_step = cfg.dt;
  o.macro_electrons = ne;
  o.macro_ions = ni;
  o.valid = std::isfinite(o.net_collection_current_density_a_per_m2) ? 1 : 0;
  scdat_destroy(ctx);
}

}  // namespace

extern "C" {

void scdat_sampler_config_default(scdat_sampler_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  // PicMccSurfaceSamplerConfig defaults (PicMccSurfaceCurrentSampler.h:20-49)
  c->surface_area_m2 = 1e-4;
  c->gap_distance_m = 1e-3;
  c->surface_aspect_ratio = 1.0;
  c->electron_density_m3 = 1e12;
  c->electron_temperature_ev = 1.0;
  c->ion_density_m3 = 1e12;
  c->ion_temperature_ev = 1.0;
  c->ion_mass_amu = 1.0;
  c->nx = 1;
  c->ny = 1;
  c->z_layers = 8;
  c->particles_per_element = 2;
  c->window_steps = 12;
  c->precision = 1;
  c->seed = 20260408u;
}

void scdat_sampler_plate(const scdat_sampler_config* c, scdat_box_spec* spec) {
  const double area = std::max(1e-12, c->surface_area_m2);
  const double ar = std::min(std::max(c->surface_aspect_ratio, 0.25), 4.0);
  std::memset(spec, 0, sizeof(*spec));
  spec->hi[0] = std::sqrt(area * ar);
  spec->hi[1] = std::sqrt(area / ar);
  spec->hi[2] = std::max(1e-6, c->gap_distance_m);
  spec->n[0] = std::max(1, c->nx);
  spec->n[1] = std::max(1, c->ny);
  spec->n[2] = std::max(4, c->z_layers);
}

void scdat_sampler_counts(const scdat_sampler_config* c, int64_t* n_electrons, int64_t* n_ions, double* weight) {
  scdat_box_spec spec;
  scdat_sampler_plate(c, &spec);
  const int64_t nt = 6LL * spec.n[0] * spec.n[1] * spec.n[2];
  const int64_t ne = nt * std::max(1, c->particles_per_element);
  const double vol = (spec.hi[0] - spec.lo[0]) * (spec.hi[1] - spec.lo[1]) * (spec.hi[2] - spec.lo[2]);
  const double w = std::max(0.0, c->electron_density_m3) * vol / (double)ne;
  *n_electrons = ne;
  *n_ions = w > 0.0 ? (int64_t)std::llround(std::max(0.0, c->ion_density_m3) * vol / w) : ne;
  *weight = w > 0.0 ? w : 1.0;
}

void scdat_sampler_direct(int64_t n, double* v) {
  for (int64_t p = 0; p < n; ++p) v[3 * p + 2] = -std::abs(v[3 * p + 2]);
}

int scdat_sample_surface_current(const scdat_sampler_config* cfg, double probe_delta_v,
                                 scdat_current_sample* out) {
  if (!cfg || !out) return SCDAT_ERR_ARG;
  if (!(cfg->surface_area_m2 > 0.0) || !(cfg->gap_distance_m > 0.0) || !(cfg->electron_density_m3 >= 0.0))
    return SCDAT_ERR_ARG;
  const bool deriv = probe_delta_v != 0.0;
  const double d = std::max(0.1, std::abs(probe_delta_v));
  std::vector<Window> w(deriv ? 3 : 1);
  for (auto& x : w) x.cfg = *cfg;
  if (deriv) {
    w[1].cfg.surface_potential_v += d;
    w[2].cfg.surface_potential_v -= d;
    std::vector<std::thread> th;
    for (auto& x : w) th.emplace_back(run_window, std::ref(x));
    for (auto& t : th) t.join();
  } else {
    run_window(w[0]);
  }
  for (auto& x : w)
    if (x.rc != SCDAT_OK) return x.rc;
  *out = w[0].out;
  if (deriv && w[1].out.valid && w[2].out.valid) {
    out->current_derivative_a_per_m2_per_v =
        (w[1].out.net_collection_current_density_a_per_m2 - w[2].out.net_collection_current_density_a_per_m2) /
        (2.0 * d);
    out->total_collisions = std::max({w[0].out.total_collisions, w[1].out.total_collisions,
                                      w[2].out.total_collisions});
  }
  return SCDAT_OK;
}

}  // extern "C"
