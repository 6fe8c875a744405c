"""Python mirror of the reference orchestrator SCDAT::PICcore::PICCycle.

Same public names and meaning as Tools/PICcore/include/PICCycle.h:33-185;
every call goes through the C ABI to the gfx950 path.  Errors surface as
``ScdatError`` (the reference returns ``false`` and prints to ``std::cerr``).
"""
from dataclasses import dataclass

import numpy as np

from . import (BC_ABSORB, BC_CONDUCTOR, BC_HOST, BC_PERIODIC, BC_REFLECT, Context, ELEMENTARY_CHARGE,
               ELECTRON_MASS, PROTON_MASS, SolverFailure, VACUUM_PERMITTIVITY, default_config)


class BoundaryFace:
    """PICCycle::BoundaryFace (PICCycle.h:36-44)."""
    XMin, XMax, YMin, YMax, ZMin, ZMax = range(6)


class BoundaryCondition:
    """Particle boundary condition kinds (Tools/Boundary/include/BoundaryConditions.h)."""
    Absorbing = BC_ABSORB
    Periodic = BC_PERIODIC
    Reflective = BC_REFLECT
    Conductor = BC_CONDUCTOR


class BoundaryParticle:
    """The particle a caller-side boundary condition receives: the fields of
    the reference's ParticleTypeDef that BoundaryConditionBase::
    processParticle reads and changes (Tools/Particle/include/
    ParticleDefinitions.h), with its status calls."""
    ACTIVE, ABSORBED, ESCAPED = "active", "absorbed", "escaped"

    def __init__(self, position, velocity, charge, mass, weight, species=-1, id=-1):
        self.position = np.array(position, dtype=np.float64)
        self.velocity = np.array(velocity, dtype=np.float64)
        self.charge, self.mass, self.weight = float(charge), float(mass), float(weight)
        self.species, self.id = species, id
        self.status = BoundaryParticle.ACTIVE

    def getPosition(self):
        return self.position

    def setPosition(self, x):
        self.position = np.array(x, dtype=np.float64)

    def getVelocity(self):
        return self.velocity

    def setVelocity(self, v):
        self.velocity = np.array(v, dtype=np.float64)

    def getCharge(self):
        return self.charge

    def getMass(self):
        return self.mass

    def getWeight(self):
        return self.weight

    def getKineticEnergy(self):
        return 0.5 * self.mass * float(self.velocity @ self.velocity)

    def isActive(self):
        return self.status == BoundaryParticle.ACTIVE

    def markAsAbsorbed(self):
        self.status = BoundaryParticle.ABSORBED

    def markAsEscaped(self):
        self.status = BoundaryParticle.ESCAPED


_LEDGER_KEYS = ("absorbed_electron_charge_c", "absorbed_ion_charge_c", "emitted_electron_charge_c",
                "incident_energy_j", "absorbed_electron_particles", "absorbed_ion_particles",
                "emitted_electron_particles")


@dataclass
class Parameters:
    """PICCycle::Parameters (PICCycle.h:46-58) plus device-path controls."""
    time_step: float = 1e-12
    max_iterations: int = 1000
    convergence_tolerance: float = 1e-6
    adaptive_time_step: bool = True
    charge_conservation: bool = True
    statistics_frequency: int = 10
    verbose: bool = False
    precision: int = 0            # 0 mixed, 1 FP64
    pcg_tolerance: float = 1e-8
    permittivity: float = VACUUM_PERMITTIVITY
    sort_interval: int = 20


class PICCycle:
    def __init__(self, params=None):
        self.params = params or Parameters()
        p = self.params
        self._cfg = default_config(precision=p.precision, dt=p.time_step,
                                   adaptive_substep=int(bool(p.adaptive_time_step)),
                                   pcg_tol=p.pcg_tolerance, permittivity=p.permittivity,
                                   sort_interval=p.sort_interval)
        self._ctx = Context(self._cfg)
        self._bc = {}
        self._host_bc = {}        # face -> caller-side condition (processParticle)
        self._host_ledger = {}    # face -> ledger fields recorded by those conditions
        self._host_surface_charge = 0.0
        self._species = []        # (charge, mass, weight) per device species
        self._ids = []
        self._surface_charge_mark = 0.0
        self._initialized = False

    # --- setup (PICCycle.h:63-69) -------------------------------------------
    def setMesh(self, xyz, tets, node_kind, dirichlet=None, face_group=None):
        self._ctx.set_mesh(xyz, tets, node_kind, dirichlet, face_group)
        self._n_nodes = len(node_kind)
        xyz = np.asarray(xyz, dtype=np.float64)
        self._lo, self._hi = xyz.min(0), xyz.max(0)

    def setBoundaryCondition(self, face, condition, partner=-1, shift=None):
        """A device kind (BoundaryCondition.*), or any object with the
        reference's processParticle(particle, intersection_point, normal, dt)
        (BoundaryConditions.h:116-121): that condition runs here, on the
        host, for every particle that reaches the face (the slow path)."""
        face = int(face)
        if hasattr(condition, "processParticle"):
            self._ctx.set_boundary(face, BC_HOST, partner, shift)
            self._host_bc[face] = condition
            self._bc[face] = BC_HOST
            return
        self._ctx.set_boundary(face, int(condition), partner, shift)
        self._host_bc.pop(face, None)
        self._bc[face] = int(condition)

    def clearBoundaryConditions(self):
        for g in list(self._bc):
            self._ctx.set_boundary(g, BC_ABSORB)
        self._bc.clear()
        self._host_bc.clear()

    def addSpecies(self, charge, mass, weight, capacity):
        s = self._ctx.add_species(charge, mass, weight, capacity)
        self._species.append((charge, mass, weight))
        return s

    def addElectrons(self, x, v, weight, capacity=None, cell=None):
        """ParticleManager::createElectron over arrays (ParticleManager.cpp:707-717)."""
        s = self.addSpecies(-ELEMENTARY_CHARGE, ELECTRON_MASS, weight, capacity if capacity else len(x))
        self._ctx.set_particles(s, x, v, cell)
        return s

    def addIons(self, x, v, weight, mass_number=1, charge_number=1, capacity=None, cell=None):
        """ParticleManager::createIon over arrays (ParticleManager.cpp:719-729).
        The reference uses m = A * m_p (ParticleDefinitions.cpp:386-388)."""
        s = self.addSpecies(charge_number * ELEMENTARY_CHARGE, mass_number * PROTON_MASS, weight,
                            capacity if capacity else len(x))
        self._ctx.set_particles(s, x, v, cell)
        return s

    def setConductorCharge(self, q):
        self._ctx.set_conductor_charge(q)

    # --- cycle (PICCycle.h:71-73) ---------------------------------------------
    def initialize(self):
        try:
            self._ctx.initialize()
        except SolverFailure:
            self._initialized = True
            return False
        self._initialized = True
        return True

    def executeTimeStep(self):
        """One step.  The reference returns !checkConvergence(), which is
        always true (PICCycle.cpp:410-416,968-972); here a step whose Poisson
        solve missed the tolerance returns False.  That step still ran and
        kept the previous potential and field (PICCycle.cpp:731-767);
        getStatistics()["pcg_failed_solves"] counts such steps."""
        return self.executeSteps(1)

    def executeSteps(self, n):
        """n steps; False when any of their solves missed the tolerance.
        With caller-side boundary conditions the steps run one at a time and
        each step's impacts go through processParticle after it."""
        if not self._host_bc:
            try:
                self._ctx.step(n)
            except SolverFailure:
                return False
            return True
        ok = True
        for _ in range(int(n)):
            try:
                self._ctx.step(1)
            except SolverFailure:
                ok = False
            self._process_host_impacts()
        return ok

    # --- caller-side boundary conditions (the slow path) ----------------------
    def _face_of(self, x):
        """PICCycle::detectBoundaryFace: the box face a position lies beyond
        (PICCycle.cpp:1074-1133), None inside."""
        span = np.maximum(self._hi - self._lo, 1e-12)
        tol = max(1e-10, 1e-8 * float(span.max()))
        best, face = 0.0, None
        for ax in range(3):
            for side, d in ((0, self._lo[ax] - x[ax]), (1, x[ax] - self._hi[ax])):
                if d > tol and d > best:
                    best, face = d, 2 * ax + side
        return face

    def _project(self, face, x):
        p = np.array(x, dtype=np.float64)
        ax = face // 2
        p[ax] = self._lo[ax] if face % 2 == 0 else self._hi[ax]
        return p

    @staticmethod
    def _inward_normal(face):
        n = np.zeros(3)
        n[face // 2] = 1.0 if face % 2 == 0 else -1.0
        return n

    def _species_of(self, charge, mass, weight):
        for k, (q, m, w) in enumerate(self._species):
            if q == charge and m == mass and w == weight:
                return k
        raise ValueError("an emitted particle's (charge, mass, weight) matches no species of the cycle")

    def _record(self, face, p, emitted):
        """PICCycle::recordSurfaceBoundaryInteraction (PICCycle.cpp:1295-1329)."""
        led = self._host_ledger.setdefault(face, dict.fromkeys(_LEDGER_KEYS, 0))
        qw = p.getCharge() * p.getWeight()
        if p.status == BoundaryParticle.ABSORBED:
            led["incident_energy_j"] += p.getKineticEnergy() * max(1.0, p.getWeight())
            if p.getCharge() < 0:
                led["absorbed_electron_charge_c"] += qw
                led["absorbed_electron_particles"] += 1
            else:
                led["absorbed_ion_charge_c"] += qw
                led["absorbed_ion_particles"] += 1
            self._host_surface_charge += qw
        for e in emitted:
            if e.getCharge() < 0:
                led["emitted_electron_charge_c"] += e.getCharge() * e.getWeight()
                led["emitted_electron_particles"] += 1

    def _process_host_impacts(self):
        """PICCycle::applyRuntimeBoundaryCondition (PICCycle.cpp:1221-1293)
        for the last step's impacts on caller-side faces, in (species, id)
        order: up to six passes of processParticle; particles that stay
        active and those the conditions emit return to the device (they
        enter the charge from the next step on)."""
        add = {}
        dt = self.params.time_step
        for imp in self._ctx.host_impacts():
            q, m, w = self._species[imp["species"]]
            p = BoundaryParticle(imp["x"], imp["v"], q, m, w, imp["species"], imp["id"])
            face = imp["group"]
            for _ in range(6):
                cond = self._host_bc.get(face)
                if cond is None:
                    p.markAsEscaped()
                    break
                emitted = list(cond.processParticle(p, self._project(face, p.getPosition()),
                                                    self._inward_normal(face), dt) or [])
                self._record(face, p, emitted)
                for e in emitted:
                    k = self._species_of(e.getCharge(), e.getMass(), e.getWeight())
                    add.setdefault(k, []).append((e.getPosition(), e.getVelocity()))
                if not p.isActive():
                    break
                nxt = self._face_of(p.getPosition())
                if nxt is None:
                    add.setdefault(p.species, []).append((p.getPosition(), p.getVelocity()))
                    break
                face = nxt
            else:
                if p.isActive() and self._face_of(p.getPosition()) is None:
                    add.setdefault(p.species, []).append((p.getPosition(), p.getVelocity()))
        # a position on a face moves a hair inside so that the host finds its tet
        span = float(np.max(self._hi - self._lo))
        for k, lst in add.items():
            x = np.array([a for a, _ in lst], dtype=np.float64)
            v = np.array([b for _, b in lst], dtype=np.float64)
            x = np.minimum(np.maximum(x, self._lo + 1e-12 * span), self._hi - 1e-12 * span)
            self._ctx.append_particles(k, x, v)

    def executeSimulation(self, total_time):
        n = int(round(total_time / self.params.time_step))
        return self.executeSteps(n) if n > 0 else True

    # --- field access (PICCycle.h:73-84,90-117) ------------------------------
    def getChargeDensity(self):
        return self._ctx.fields()[0]

    def getPotential(self):
        return self._ctx.fields()[1]

    def getElectricField(self):
        return self._ctx.fields()[2]

    def setPotential(self, phi):
        self._ctx.set_potential(phi)

    def setNodePotential(self, node_id, value):
        phi = self.getPotential()
        phi[node_id] = value
        self._ctx.set_potential(phi)

    def setTimeStep(self, dt):
        if dt > 0.0:
            self.params.time_step = dt
            self._ctx.set_time_step(dt)

    def getTimeStep(self):
        return self.params.time_step

    # --- statistics (PICCycle.h:135-185) -------------------------------------
    def getStatistics(self):
        return self._ctx.stats()

    def getSurfaceCurrentLedger(self, n_groups=6):
        """The device ledger plus what caller-side conditions recorded."""
        led = self._ctx.ledger(n_groups)
        for face, h in self._host_ledger.items():
            if face < n_groups:
                for k in _LEDGER_KEYS:
                    led[face][k] += h[k]
        return led

    def getAndResetSurfaceCharge(self):
        q = self._ctx.stats()["surface_deposited_charge"] + self._host_surface_charge
        d = q - self._surface_charge_mark
        self._surface_charge_mark = q
        return d

    @property
    def context(self):
        return self._ctx
