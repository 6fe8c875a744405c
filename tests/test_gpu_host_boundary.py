"""Caller-side boundary conditions (SCDAT_BC_HOST) through the PICCycle
mirror: a condition object with the reference's processParticle
(BoundaryConditions.h:116-121) runs on the host for every particle that
reaches its face, as PICCycle::applyRuntimeBoundaryCondition calls it
(PICCycle.cpp:1221-1293).  An absorbing condition written that way gives the
device absorbing kind's fields, particles and ledger; a reflecting,
secondary-emitting one returns its particles to the device and its ledger
counts the emitted electrons."""
import numpy as np
import pytest

import scdat_amd as S
from scdat_amd.piccycle import BoundaryCondition, BoundaryParticle, Parameters, PICCycle
from helpers import plasma_setup

pytestmark = pytest.mark.gpu


class AbsorbAll:
    def processParticle(self, particle, point, normal, dt):
        particle.markAsAbsorbed()
        return []


class SpecularWithSecondary:
    """Reflects the particle at the crossing point and, for electrons, emits
    one slow electron there along the inward normal."""

    def __init__(self):
        self.calls = 0
        self.emitted = 0

    def processParticle(self, particle, point, normal, dt):
        self.calls += 1
        v = particle.getVelocity()
        particle.setVelocity(v - 2.0 * float(v @ normal) * normal)
        particle.setPosition(point)
        if particle.getCharge() < 0:
            self.emitted += 1
            return [BoundaryParticle(point, 1e5 * normal, particle.getCharge(), particle.getMass(),
                                     particle.getWeight())]
        return []


def _cycle(st, wall):
    pc = PICCycle(Parameters(time_step=st["dt"], precision=0, pcg_tolerance=1e-8, sort_interval=0))
    m = st["mesh"]
    pc.setMesh(m["xyz"], m["tets"], m["kind"], face_group=m["group"])
    for f in range(6):
        pc.setBoundaryCondition(f, wall(f))
    for g in range(6, 12):
        pc.setBoundaryCondition(g, BoundaryCondition.Conductor)
    n = len(st["e"][0])
    xe, ve, ce = st["e"]
    xi, vi, ci = st["i"]
    pc.addElectrons(xe, ve, st["w"], capacity=4 * n, cell=ce)
    pc.addIons(xi, vi, st["w"], capacity=4 * n, cell=ci)
    pc.initialize()
    return pc


def test_absorbing_condition_on_the_host_equals_device_absorption():
    st = plasma_setup(n=8, npart=20000, seed=21, Te_eV=5.0)
    dev = _cycle(st, lambda f: BoundaryCondition.Absorbing)
    host = _cycle(st, lambda f: AbsorbAll())
    for _ in range(6):
        dev.executeTimeStep()
        host.executeTimeStep()
    np.testing.assert_array_equal(dev.getPotential(), host.getPotential())
    np.testing.assert_array_equal(dev.getElectricField(), host.getElectricField())
    sd, sh = dev.getStatistics(), host.getStatistics()
    assert sd["active_particles"] == sh["active_particles"]
    assert sd["floating_potential"] == sh["floating_potential"]
    ld, lh = dev.getSurfaceCurrentLedger(), host.getSurfaceCurrentLedger()
    hits = 0
    for f in range(6):
        for k in ("absorbed_electron_particles", "absorbed_ion_particles", "emitted_electron_particles"):
            assert ld[f][k] == lh[f][k], (f, k)
        hits += lh[f]["absorbed_electron_particles"] + lh[f]["absorbed_ion_particles"]
        for k in ("absorbed_electron_charge_c", "absorbed_ion_charge_c", "incident_energy_j"):
            assert lh[f][k] == pytest.approx(ld[f][k], rel=1e-12, abs=1e-300), (f, k)
    assert hits > 0
    assert host.getAndResetSurfaceCharge() == pytest.approx(dev.getAndResetSurfaceCharge(), rel=1e-12)


def test_reflecting_emitting_condition_returns_particles():
    st = plasma_setup(n=8, npart=20000, seed=22, Te_eV=5.0)
    cond = SpecularWithSecondary()
    pc = _cycle(st, lambda f: cond)
    n0 = 2 * len(st["e"][0])
    for _ in range(6):
        pc.executeTimeStep()
    led = pc.context.ledger(12)
    on_conductor = sum(led[g]["absorbed_electron_particles"] + led[g]["absorbed_ion_particles"] for g in range(6, 12))
    walls = pc.getSurfaceCurrentLedger()
    emitted = sum(walls[f]["emitted_electron_particles"] for f in range(6))
    assert cond.calls > 0 and emitted == cond.emitted > 0
    # nothing is lost at the walls: every particle that reached one came back
    assert sum(walls[f]["absorbed_electron_particles"] + walls[f]["absorbed_ion_particles"] for f in range(6)) == 0
    # particles on the device now (the statistics count the last push's,
    # before its impacts came back)
    alive = sum(pc.context.particle_count(s)[0] for s in (0, 1))
    assert alive == n0 - on_conductor + emitted


def test_host_kind_is_a_boundary_type():
    c = S.Context(precision=0)
    m = S.box_mesh(4)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    c.set_boundary(0, S.BC_HOST)
    with pytest.raises(S.ScdatError):
        c.set_boundary(0, 6)
    c.close()
