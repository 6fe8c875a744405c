// Harness exposing the reference's collision cross sections and surface
// emission yields through a C ABI.
//
// TEST INFRASTRUCTURE ONLY.  Compiled with the reference's own sources
// (oracle/ref/Makefile) into oracle/_ref/libscdat_ref_surface.so; used to
// produce golden vectors that pin the oracle.  Nothing in the product links
// or loads it.
//
// Reference routines used (paths relative to the SCDAT source tree):
//   Tools/Interactions/Collisions/src/CollisonTypes.cpp:60-198  cross
//       sections (elastic, ionisation, excitation, recombination, Coulomb)
//       and the Coulomb logarithm
//   Tools/Material/src/SurfaceMaterialModel.cpp:163-208,369-402
//       secondaryYield and photoCurrentDensity through
//       BasicSurfaceMaterialModel::evaluateCurrents
//   Tools/Coupling/src/SurfaceCircuitCoupling.cpp:801-1127  advanceImplicit
#include <cstdint>

#include "Tools/Interactions/Collisions/include/CollisonTypes.h"
#include "Tools/Coupling/include/SurfaceCircuitCoupling.h"
#include "Tools/Material/include/SurfaceMaterialModel.h"

namespace C = SCDAT::Collision;
namespace Mt = SCDAT::Material;

extern "C" {

// type: 0 elastic, 1 ionisation, 2 excitation, 3 recombination, 4 Coulomb.
// prm: threshold, sigma_max, reference_energy, power_index, alpha, beta,
//      electron_density, electron_temperature_K (unused entries ignored).
void ref_sigma(int type, const double* prm, int64_t n, const double* e_ev, double* out) {
  C::CollisionParameters p;
  p.energy_threshold = prm[0];
  p.cross_section_max = prm[1];
  p.additional_params["reference_energy"] = prm[2];
  p.additional_params["power_index"] = prm[3];
  p.additional_params["alpha"] = prm[4];
  p.additional_params["beta"] = prm[5];
  p.additional_params["electron_density"] = prm[6];
  p.additional_params["electron_temperature"] = prm[7];
  for (int64_t i = 0; i < n; ++i) {
    double s = 0.0;
    switch (type) {
      case 0: s = C::ElasticCollisionCrossSection(p).calculateCrossSection(e_ev[i]); break;
      case 1: s = C::IonizationCollisionCrossSection(p).calculateCrossSection(e_ev[i]); break;
      case 2: s = C::ExcitationCollisionCrossSection(p).calculateCrossSection(e_ev[i]); break;
      case 3: s = C::RecombinationCollisionCrossSection(p).calculateCrossSection(e_ev[i]); break;
      default: s = C::CoulombCollisionCrossSection(p).calculateCrossSection(e_ev[i]); break;
    }
    out[i] = s;
  }
}

double ref_coulomb_log(double temperature_k, double density_m3) {
  return C::CoulombCollisionCrossSection::calculateCoulombLogarithm(temperature_k, density_m3);
}

// Secondary-electron yield (emitted / collected) and photoelectron current
// density of BasicSurfaceMaterialModel for one material.
void ref_surface(double see_max, double see_peak_ev, double photo_yield, double photo_temp_ev,
                 int64_t n, const double* e_ev, const double* angle_rad, const double* surf_v,
                 double ref_v, double source_j, double* yield, double* photo_j) {
  Mt::MaterialProperty mat;
  mat.setSecondaryElectronYieldMax(see_max);
  mat.setSecondaryElectronPeakEnergyEv(see_peak_ev);
  mat.setPhotoelectronYield(photo_yield);
  mat.setPhotoelectronTemperatureEv(photo_temp_ev);
  Mt::BasicSurfaceMaterialModel model;
  for (int64_t i = 0; i < n; ++i) {
    Mt::SurfaceModelContext ctx;
    ctx.incident_particle = Mt::SurfaceIncidentParticle::Electron;
    ctx.incident_energy_ev = e_ev[i];
    ctx.incident_angle_rad = angle_rad[i];
    ctx.exposed_area_m2 = 1.0;
    ctx.surface_potential_v = surf_v[i];
    ctx.reference_potential_v = ref_v;
    ctx.source_current_density_a_per_m2 = source_j;
    const Mt::SurfaceComponentCurrents c = model.evaluateCurrents(mat, ctx);
    yield[i] = c.electron_collection_a_per_m2 > 0.0
                   ? c.secondary_electron_emission_a_per_m2 / c.electron_collection_a_per_m2
                   : 0.0;
    photo_j[i] = c.photoelectron_emission_a_per_m2;
  }
}

// One advanceImplicit call.  node: 5 per node (potential, capacitance,
// shunt conductance, fixed flag, fixed value); branch: 4 per branch (from,
// to, conductance, bias); entry: 3 per extra matrix entry (row, column,
// coefficient).  Returns 0, or 1 when the reference throws.
int ref_circuit_advance(int nn, const double* node, int nb, const double* branch, double dt,
                        const double* current, const double* didv, const double* rhs, const double* diag,
                        int ne, const double* entry, double max_delta, double* v_out, double* i_out,
                        double* b_out) {
  namespace Cp = SCDAT::Coupling;
  Cp::SurfaceCircuitCoupling circ;
  for (int i = 0; i < nn; ++i) {
    Cp::SurfaceCircuitNode n;
    n.potential_v = node[5 * i];
    n.capacitance_f = node[5 * i + 1];
    n.shunt_conductance_s = node[5 * i + 2];
    n.fixed_potential = node[5 * i + 3] != 0.0;
    n.fixed_value_v = node[5 * i + 4];
    circ.addNode(n);
  }
  for (int k = 0; k < nb; ++k) {
    Cp::SurfaceCircuitBranch b;
    b.from_node = (std::size_t)branch[4 * k];
    b.to_node = (std::size_t)branch[4 * k + 1];
    b.conductance_s = branch[4 * k + 2];
    b.bias_v = branch[4 * k + 3];
    circ.addBranch(b);
  }
  Cp::SurfaceCircuitLinearization lin;
  lin.plasma_current_a.assign(current, current + nn);
  lin.plasma_didv_a_per_v.assign(didv, didv + nn);
  lin.additional_rhs_a.assign(rhs, rhs + nn);
  lin.additional_diagonal_a_per_v.assign(diag, diag + nn);
  for (int e = 0; e < ne; ++e)
    lin.additional_off_diagonal_entries.push_back(
        {(std::size_t)entry[3 * e], (std::size_t)entry[3 * e + 1], entry[3 * e + 2]});
  try {
    const Cp::SurfaceCircuitAdvanceResult r = circ.advanceImplicit(dt, lin, max_delta);
    for (int i = 0; i < nn; ++i) {
      v_out[i] = r.node_potentials_v[i];
      i_out[i] = r.plasma_currents_a[i];
    }
    for (int k = 0; k < nb; ++k) b_out[k] = r.branch_currents_a[k];
  } catch (...) {
    return 1;
  }
  return 0;
}

}  // extern "C"
