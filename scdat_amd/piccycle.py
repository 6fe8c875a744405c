"""Python mirror of the reference orchestrator SCDAT::PICcore::PICCycle.

Same public names and meaning as Tools/PICcore/include/PICCycle.h:33-185;
every call goes through the C ABI to the gfx950 path.  Errors surface as
``ScdatError`` (the reference returns ``false`` and prints to ``std::cerr``).
"""
from dataclasses import dataclass

import numpy as np

from . import (BC_ABSORB, BC_CONDUCTOR, BC_PERIODIC, BC_REFLECT, Context, ELEMENTARY_CHARGE,
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
        self._ids = []
        self._surface_charge_mark = 0.0
        self._initialized = False

    # --- setup (PICCycle.h:63-69) -------------------------------------------
    def setMesh(self, xyz, tets, node_kind, dirichlet=None, face_group=None):
        self._ctx.set_mesh(xyz, tets, node_kind, dirichlet, face_group)
        self._n_nodes = len(node_kind)

    def setBoundaryCondition(self, face, condition, partner=-1, shift=None):
        self._ctx.set_boundary(int(face), int(condition), partner, shift)
        self._bc[int(face)] = int(condition)

    def clearBoundaryConditions(self):
        for g in list(self._bc):
            self._ctx.set_boundary(g, BC_ABSORB)
        self._bc.clear()

    def addSpecies(self, charge, mass, weight, capacity):
        return self._ctx.add_species(charge, mass, weight, capacity)

    def addElectrons(self, x, v, weight, capacity=None, cell=None):
        """ParticleManager::createElectron over arrays (ParticleManager.cpp:707-717)."""
        s = self._ctx.add_species(-ELEMENTARY_CHARGE, ELECTRON_MASS, weight,
                                  capacity if capacity else len(x))
        self._ctx.set_particles(s, x, v, cell)
        return s

    def addIons(self, x, v, weight, mass_number=1, charge_number=1, capacity=None, cell=None):
        """ParticleManager::createIon over arrays (ParticleManager.cpp:719-729).
        The reference uses m = A * m_p (ParticleDefinitions.cpp:386-388)."""
        s = self._ctx.add_species(charge_number * ELEMENTARY_CHARGE, mass_number * PROTON_MASS,
                                  weight, capacity if capacity else len(x))
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
        """n steps; False when any of their solves missed the tolerance."""
        try:
            self._ctx.step(n)
        except SolverFailure:
            return False
        return True

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
        return self._ctx.ledger(n_groups)

    def getAndResetSurfaceCharge(self):
        q = self._ctx.stats()["surface_deposited_charge"]
        d = q - self._surface_charge_mark
        self._surface_charge_mark = q
        return d

    @property
    def context(self):
        return self._ctx
