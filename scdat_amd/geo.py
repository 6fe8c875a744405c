"""GEO-like charging case: hot plasma, neutral background, surface emission.

The spacecraft-charging geometry of the PIC benchmark (1 m box with the
0.2 m conductor cube [0.4, 0.6]^3 floating at its centre, phi = 0 on the
outer walls, Main/main_PICcore.cpp:151-229 builds the reference's case) in
a geosynchronous environment:

* Maxwellian electrons and protons, Te = Ti = 10 keV, n = 1e6 m^-3, loaded
  uniformly outside the conductor;
* an atomic-oxygen background at 1e12 m^-3: electron elastic scattering,
  electron-impact ionisation (O -> O+ + e-, threshold 13.62 eV) and proton
  elastic scattering, drawn by the null-collision method
  (MonteCarloCollisionHandler, CollisionAlgorithm.cpp:500-538, with the
  cross-section forms of CollisonTypes.cpp:56-132);
* secondary emission from electrons absorbed on the conductor, Sternglass
  yield delta_max = 2.0 at E_max = 400 eV, secondaries at 2 eV;
* photoemission of 20 uA/m^2 (2 eV photoelectrons) from the sunlit +x face
  of the conductor.

The time step resolves the electron transit of a cell: dt = 0.2 h / v_th,e
(h = 1 m / edge).  The process parameters are this workload's, chosen for
oxygen; the forms and their meaning are the reference's.
"""
import math

from . import (BC_CONDUCTOR, ELECTRON_MASS, ELEMENTARY_CHARGE, MCC_ELASTIC, MCC_IONIZATION, ORDER_TILED,
               PRECOND_TWO_LEVEL, PROTON_MASS, SEE_STERNGLASS, Context, box_mesh, box_spec, mcc_process)
from .loading import maxwellian_box

CONDUCTOR = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))
CONDUCTOR_GROUPS = range(6, 12)
SUNLIT_GROUP = 7  # the conductor's +x face
ELECTRON, PROTON, OXYGEN_ION = 0, 1, 2
O_MASS = 16.0 * PROTON_MASS

DEFAULTS = dict(edge=48, n_particles=8 * 1024 * 1024, te_ev=1.0e4, ti_ev=1.0e4, density=1.0e6,
                neutral_density=1.0e12, delta_max=2.0, energy_max_ev=400.0, see_temperature_ev=2.0,
                photo_current_density=20e-6, photo_temperature_ev=2.0, seed=20261205)


def processes(neutral_density):
    """The oxygen background's collision processes, as plain dicts whose keys
    name the fields of a process (type, incident species, product species,
    density, threshold, sigma_max, energy loss, reference energy, power
    index): e- + O elastic (sigma_max (E_ref/E)^p), e- + O ionisation
    (sigma_max ln(E/I)(1 - I/E)/E, I = 13.62 eV, peak about 1.4e-20 m^2),
    H+ + O elastic."""
    n = float(neutral_density)
    return [
        dict(type=MCC_ELASTIC, species=ELECTRON, density=n, sigma_max=1.0e-19, reference_energy=10.0,
             power_index=0.5),
        dict(type=MCC_IONIZATION, species=ELECTRON, e_species=ELECTRON, i_species=OXYGEN_ION, density=n,
             threshold=13.62, sigma_max=7.5e-19, energy_loss=13.62),
        dict(type=MCC_ELASTIC, species=PROTON, density=n, sigma_max=2.0e-19, reference_energy=10.0,
             power_index=0.3),
    ]


def device_process(p):
    """scdat_mcc_process of one process dict (unlisted fields: the reference defaults)."""
    kw = dict(target_density=p["density"], sigma_max_m2=p["sigma_max"])
    for src, dst in (("threshold", "threshold_ev"), ("energy_loss", "energy_loss_ev"),
                     ("reference_energy", "reference_energy_ev"), ("power_index", "power_index"),
                     ("e_species", "product_electron_species"), ("i_species", "product_ion_species")):
        if src in p:
            kw[dst] = p[src]
    return mcc_process(p["type"], p["species"], **kw)


def inputs(**kw):
    """Mesh, particles (n_particles / 2 electrons and protons), weight, time
    step and settings of the case; keyword arguments override DEFAULTS."""
    cfg = dict(DEFAULTS)
    cfg.update(kw)
    n = int(cfg["edge"])
    m = box_mesh(n, conductor=CONDUCTOR)
    spec = box_spec(n, conductor=CONDUCTOR)
    qe = ELEMENTARY_CHARGE
    vth_e = math.sqrt(cfg["te_ev"] * qe / ELECTRON_MASS)
    vth_i = math.sqrt(cfg["ti_ev"] * qe / PROTON_MASS)
    half = int(cfg["n_particles"]) // 2
    xe, ve, ce = maxwellian_box(spec, half, int(cfg["seed"]), 0, vth_e)
    xi, vi, ci = maxwellian_box(spec, half, int(cfg["seed"]), 1, vth_i)
    w = float(cfg["density"]) * (1.0 - 0.2 ** 3) / half
    dt = 0.2 * (1.0 / n) / vth_e
    return dict(cfg=cfg, mesh=m, spec=spec, e=(xe, ve, ce), i=(xi, vi, ci), w=w, dt=dt,
                procs=processes(cfg["neutral_density"]))


def species():
    """(charge, mass) of the case's species: e-, H+, O+."""
    qe = ELEMENTARY_CHARGE
    return [(-qe, ELECTRON_MASS), (qe, PROTON_MASS), (qe, O_MASS)]


def capacities(st, steps):
    """Species capacities for a run of `steps` steps: the loaded electrons
    plus the photoelectrons the sunlit face can emit (J A dt / (e w) a step)
    and as many secondaries, protons as loaded, and room for O+."""
    cfg = st["cfg"]
    half = len(st["e"][0])
    area = 0.2 * 0.2
    photo = cfg["photo_current_density"] * area * st["dt"] / (ELEMENTARY_CHARGE * st["w"])
    extra = int(2 * photo * steps) + 4096
    return [half + extra, half, max(4096, half // 8)]


def context(st, precision=0, tol=1e-8, sort_interval=20, steps=2000, seed=None, two_level=False):
    """The case's device context, initialized; two_level: the two-level
    Poisson preconditioner in the tiled node order."""
    cfg = st["cfg"]
    m = st["mesh"]
    c = Context(precision=precision, dt=st["dt"], pcg_tol=tol, sort_interval=sort_interval)
    if two_level:
        c.set_node_order(ORDER_TILED)
        c.set_preconditioner(PRECOND_TWO_LEVEL)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in CONDUCTOR_GROUPS:
        c.set_boundary(g, BC_CONDUCTOR)
    for (q, mass), cap in zip(species(), capacities(st, steps)):
        c.add_species(q, mass, st["w"], cap)
    c.set_particles(ELECTRON, *st["e"])
    c.set_particles(PROTON, *st["i"])
    c.mcc_configure([device_process(p) for p in st["procs"]])
    c.set_seed(int(cfg["seed"] if seed is None else seed))
    for g in CONDUCTOR_GROUPS:
        c.set_surface_emission(g, SEE_STERNGLASS, ELECTRON, cfg["delta_max"], cfg["energy_max_ev"],
                               cfg["see_temperature_ev"],
                               photo_j=cfg["photo_current_density"] if g == SUNLIT_GROUP else 0.0,
                               photo_temperature=cfg["photo_temperature_ev"])
    c.initialize()
    return c

