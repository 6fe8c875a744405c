"""PIC boundary window of the vacuum-arc toolkit on the device path.

The surface-discharge arc model couples its channel to a small PIC window at
the cathode (SurfaceDischargeArcAlgorithm, Toolkit/Vacuum Arc/src/
SurfaceDischargeArcAlgorithm.cpp):

* initializePiccoreBoundaryCoupling (:346-496): a plate of 2 r_channel x
  2 r_channel x gap, 1 x 1 x 6 hexahedra; nodes on the two z faces are
  Dirichlet, the rest free; x and y periodic; the cathode (ZMin) is a
  field-emission boundary (emission area = plate area, at most 8 field
  electrons per ion impact, the configured ion-induced secondaries); the
  anode (ZMax) absorbs.  Cycle time step 1e-10 s.
* applyPiccoreBoundaryPotentials (:498-526): cathode nodes at the surface
  potential, anode nodes at 0 V, every other node 0 V.
* seedBoundaryProbeParticles (:528-562): each advance seeds
  min(4, max(1, ceil(0.002 J A max(dt, 1e-12) / e))) protons of weight 1 at
  the top tetrahedron's centroid, 1e-9 m apart downwards, moving at
  (0, 0, -2e6) m/s.
* advancePiccoreBoundaryCoupling (:575-640): set the time step and the
  cathode's current density, apply the potentials, seed the probes, run one
  PIC step, and turn the cathode + anode ledger increments into absorbed
  electron / ion and emitted-electron currents and a surface-charge change.

Mesh: the reference splits each hexahedron into five tetrahedra
(Tools/Mesh/src/MeshParsing.cpp:1248-1254) with the same diagonal pattern in
every layer, so neighbouring layers meet on crossed face diagonals and the
mesh is not conforming.  The window here uses the conforming six-tetrahedron
split of the same hexahedra, which the tetrahedron walk needs.
"""
import math
from dataclasses import dataclass

import numpy as np

from . import (BC_ABSORB, BC_PERIODIC, ELECTRON_MASS, ELEMENTARY_CHARGE, NODE_DIRICHLET, NODE_FREE, PROTON_MASS,
               Context, box_mesh, field_emission)

ZMIN, ZMAX = 4, 5  # boundary groups of the plate's z faces


@dataclass
class WindowConfig:
    """The fields of SurfaceDischargeArcAlgorithm's configuration the window
    reads (include/SurfaceDischargeArcAlgorithm.h:60-122 defaults)."""
    channel_radius_m: float = 2.0e-5
    gap_distance_m: float = 5.0e-4
    surface_potential_v: float = 0.0
    surface_charge_density_c_per_m2: float = 0.0
    enable_secondary_electron_emission: bool = False
    secondary_electron_yield_per_ion: float = 0.0
    secondary_electron_speed_m_per_s: float = 8.0e5
    precision: int = 1            # 0 mixed, 1 FP64
    capacity: int = 1 << 16       # particle slots per species
    sort_interval: int = 8


@dataclass
class WindowStatus:
    """The window's share of SurfaceDischargeArcAlgorithm's status."""
    pic_absorbed_electron_current_a: float = 0.0
    pic_absorbed_ion_current_a: float = 0.0
    pic_emitted_electron_current_a: float = 0.0
    pic_boundary_net_current_a: float = 0.0
    pic_surface_charge_delta_c: float = 0.0


def plate_dimensions(cfg):
    """(length, width, thickness) of the window (:352-355)."""
    side = max(2.0 * cfg.channel_radius_m, 1.0e-4)
    return side, side, max(cfg.gap_distance_m, 1.0e-4)


def probe_count(j_emit, area_m2, dt):
    """Protons seeded per advance (:541-549)."""
    expected = max(0.0, j_emit) * area_m2 * max(dt, 1.0e-12) / ELEMENTARY_CHARGE
    scaled = max(1.0, math.ceil(expected * 0.002))
    return int(min(4.0, max(1.0, scaled)))


def plate_mesh(length, width, thickness, nz=6):
    """1 x 1 x nz plate from the origin; Dirichlet exactly on the two z faces
    (:384-400), boundary groups as the box mesher numbers them."""
    m = box_mesh((1, 1, nz), hi=(length, width, thickness))
    z = m["xyz"][:, 2]
    zmin, zmax = z.min(), z.max()
    on = (np.abs(z - zmin) < 1.0e-12) | (np.abs(z - zmax) < 1.0e-12)
    m["kind"] = np.where(on, NODE_DIRICHLET, NODE_FREE).astype(np.int32)
    return m


def top_tet(m):
    """findTopElementId (:70-91): the first tetrahedron of highest centroid z."""
    cz = m["xyz"][m["tets"], 2].mean(axis=1)
    return int(np.argmax(cz))


class BoundaryWindow:
    """The arc model's PIC boundary window; one step per `advance`."""

    def __init__(self, config=None):
        self.config = cfg = config or WindowConfig()
        L, W, T = plate_dimensions(cfg)
        self.area = max(L * W, 1.0e-12)
        self.mesh = m = plate_mesh(L, W, T)
        self.top = top_tet(m)
        self.top_centroid = m["xyz"][m["tets"][self.top]].mean(axis=0)
        z = m["xyz"][:, 2]
        self._cathode_nodes = np.abs(z - z.min()) < 1.0e-12
        self._anode_nodes = np.abs(z - z.max()) < 1.0e-12
        c = self.ctx = Context(precision=cfg.precision, dt=1.0e-10, sort_interval=cfg.sort_interval)
        c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
        c.set_boundary(0, BC_PERIODIC, 1, [L, 0, 0])
        c.set_boundary(1, BC_PERIODIC, 0, [-L, 0, 0])
        c.set_boundary(2, BC_PERIODIC, 3, [0, W, 0])
        c.set_boundary(3, BC_PERIODIC, 2, [0, -W, 0])
        c.set_boundary(ZMAX, BC_ABSORB)
        self.electrons = c.add_species(-ELEMENTARY_CHARGE, ELECTRON_MASS, 1.0, cfg.capacity)
        self.ions = c.add_species(ELEMENTARY_CHARGE, PROTON_MASS, 1.0, cfg.capacity)
        self.emission = field_emission(
            electron_species=self.electrons, emission_area_m2=self.area, max_emitted_particles_per_event=8,
            enable_secondary_electron_emission=int(bool(cfg.enable_secondary_electron_emission)),
            secondary_electron_yield_per_ion=max(0.0, cfg.secondary_electron_yield_per_ion),
            secondary_electron_speed_m_per_s=max(0.0, cfg.secondary_electron_speed_m_per_s))
        c.set_field_emission(ZMIN, self.emission)
        c.initialize()
        self.apply_potentials()
        self._last = (0.0, 0.0, 0.0)
        self._surface_mark = 0.0
        self.status = WindowStatus()

    def apply_potentials(self):
        phi = np.zeros(len(self.mesh["kind"]))
        phi[self._cathode_nodes] = self.config.surface_potential_v
        phi[self._anode_nodes] = 0.0
        self.ctx.set_potential(phi)

    def seed_probes(self, j_emit, dt):
        n = probe_count(j_emit, self.area, dt)
        x = np.repeat(self.top_centroid[None, :], n, axis=0)
        x[:, 2] -= 1.0e-9 * np.arange(1, n + 1)
        v = np.tile([0.0, 0.0, -2.0e6], (n, 1))
        self.ctx.append_particles(self.ions, x, v, np.full(n, self.top, np.int32))
        return n

    def advance(self, j_emit, dt):
        self.status = st = WindowStatus()
        if dt <= 0.0:
            return st
        c = self.ctx
        c.set_time_step(dt)
        c.set_emission_current_density(ZMIN, j_emit)
        self.apply_potentials()
        self.seed_probes(j_emit, dt)
        c.step(1)
        led = c.ledger(6)
        cat, ano = led[ZMIN], led[ZMAX]
        ae = cat["absorbed_electron_charge_c"] + ano["absorbed_electron_charge_c"]
        ai = cat["absorbed_ion_charge_c"] + ano["absorbed_ion_charge_c"]
        em = cat["emitted_electron_charge_c"] + ano["emitted_electron_charge_c"]
        d = (ae - self._last[0], ai - self._last[1], em - self._last[2])
        self._last = (ae, ai, em)
        st.pic_absorbed_electron_current_a = d[0] / dt
        st.pic_absorbed_ion_current_a = d[1] / dt
        st.pic_emitted_electron_current_a = -d[2] / dt
        st.pic_boundary_net_current_a = (st.pic_absorbed_electron_current_a + st.pic_absorbed_ion_current_a
                                         + st.pic_emitted_electron_current_a)
        q = c.stats()["surface_deposited_charge"]
        st.pic_surface_charge_delta_c = q - self._surface_mark
        self._surface_mark = q
        self.config.surface_charge_density_c_per_m2 += st.pic_surface_charge_delta_c / self.area
        return st

    def close(self):
        self.ctx.close()
