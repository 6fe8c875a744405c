"""The reference-shaped collision interfaces (MonteCarloCollisionHandler,
executeBackgroundMccStep) on the host: the default process set, density
estimate and the per-step CollisionStepResult from cumulative counts, driven
by a stand-in context (no device)."""
import pytest

import scdat_amd as S
from scdat_amd import collisions as CL


class FakeCtx:
    """Records the configuration; event counts advance by hand."""

    def __init__(self, species):
        self.species = species  # (q, m, w)
        self.alive = [0] * len(species)
        self.procs = []
        self.seed = None
        self.collisions = [0] * S.MAX_MCC_PROCESSES
        self.created = 0
        self.time = 0.0

    def set_seed(self, s):
        self.seed = s

    def mcc_configure(self, procs):
        self.procs = list(procs)

    def mcc_set_density(self, k, d):
        self.procs[k].target_density = d

    def event_counts(self):
        return {"collisions": list(self.collisions), "created": self.created}

    def stats(self):
        return {"simulation_time": self.time}

    def particle_count(self, s):
        return self.alive[s], self.alive[s]


def test_parse_and_names():
    for s in CL.CollisionCrossSectionSetId:
        assert CL.parseCollisionCrossSectionSetId(CL.toString(s)) is s
    assert CL.parseCollisionCrossSectionSetId("BACKGROUND_MCC_V1") is CL.CollisionCrossSectionSetId.BackgroundMccV1
    assert CL.parseCollisionCrossSectionSetId("nope") is None


def test_default_configuration_is_the_reference_set():
    h = CL.MonteCarloCollisionHandler(electron_species=0, ion_species=1)
    h.initializeDefaultConfiguration(1e19, 2e16, 3e16)
    names = h.processNames()
    # CollisionAlgorithm.cpp:649-700, held in name order
    assert names == sorted(["elastic_e_n", "inelastic_e_n", "ionization_e_n", "excitation_e_n",
                            "charge_exchange_i_n", "recombination_e_i", "coulomb_e_i"])
    p = {n: h._procs[n].process for n in names}
    assert p["ionization_e_n"].type == S.MCC_IONIZATION and p["ionization_e_n"].threshold_ev == 15.76
    assert p["ionization_e_n"].product_ion_species == 1 and p["ionization_e_n"].product_electron_species == 0
    assert p["charge_exchange_i_n"].species == 1 and p["charge_exchange_i_n"].target_density == 1e19
    assert p["recombination_e_i"].target_density == 2e16
    assert p["coulomb_e_i"].target_density == 3e16
    assert p["elastic_e_n"].sigma_max_m2 == 2.0e-19 and p["elastic_e_n"].power_index == 0.3


def test_step_result_counts_and_reinitialisation():
    ctx = FakeCtx([(-S.ELEMENTARY_CHARGE, S.ELECTRON_MASS, 2.0), (S.ELEMENTARY_CHARGE, S.PROTON_MASS, 3.0)])
    h = CL.MonteCarloCollisionHandler(random_seed=7)
    h.initializeDefaultConfiguration(1e18, 1e15, 1e15)
    h.bind(ctx)
    assert ctx.seed == 7 and len(ctx.procs) == 7
    names = h.processNames()
    # one step: 5 ionisations, 2 charge exchanges, 4 elastic, 5 products
    ctx.collisions[names.index("ionization_e_n")] += 5
    ctx.collisions[names.index("charge_exchange_i_n")] += 2
    ctx.collisions[names.index("elastic_e_n")] += 4
    ctx.created += 10
    ctx.time = 1e-6
    ctx.alive = [1000, 500]
    floor = CL.CollisionBackgroundDensities(1e18, 1e15, 1e15)
    r = CL.executeBackgroundMccStep(ctx, h, 1e-9, 1e-12, floor, reinitialize_configuration=True)
    assert (r.collision_events, r.ionization_events, r.charge_exchange_events, r.excitation_events) == (11, 5, 2, 0)
    assert r.generated_particles == 10 and r.active_particles == 1500
    # weight x alive / volume where it exceeds the floor
    assert r.effective_electron_density_m3 == pytest.approx(2.0 * 1000 / 1e-12)
    assert r.effective_ion_density_m3 == pytest.approx(3.0 * 500 / 1e-12)
    assert r.effective_neutral_density_m3 == 1e18
    # re-initialised at the effective densities, same process set
    by = {n: q for n, q in zip(h.processNames(), ctx.procs)}
    assert by["coulomb_e_i"].target_density == r.effective_electron_density_m3
    # the next call sees only the next step's events; statistics accumulate
    ctx.collisions[names.index("excitation_e_n")] += 3
    r2 = CL.executeBackgroundMccStep(ctx, h, 1e-9, 1e-12, CL.BackgroundMccRuntimeConfig(floor))
    assert (r2.collision_events, r2.excitation_events, r2.generated_particles) == (3, 3, 0)
    st = h.getStatistics()
    assert st.total_collisions == 14 and st.particles_created == 10
    assert st.collision_counts["ionization_e_n"] == 5
    assert st.total_rates["ionization_e_n"] == pytest.approx(5 / 1e-6)
    h.resetStatistics()
    assert h.getStatistics().total_collisions == 0
    # disabled: no processes on the device
    h.setCollisionEnabled(False)
    h.getStatistics()
    assert ctx.procs == []
    # invalid step arguments give an empty result
    assert CL.executeBackgroundMccStep(ctx, h, 0.0, 1.0).collision_events == 0
    with pytest.raises(NotImplementedError):
        CL.initializeConfiguredCollisionHandler(h, CL.CollisionCrossSectionSetId.SurfacePicV1, 1, 1, 1)
