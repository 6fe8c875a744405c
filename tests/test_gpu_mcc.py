"""Collisions, secondary emission and photoemission of the gfx950 path through
the C ABI, against the oracle on identical seeded inputs.

Both sides draw every random number from Philox keyed by (seed, particle id,
step, species, stream), so in FP64 the device reproduces the oracle's events
one for one: the same collisions, the same generated particles (matched by
their generated ids) with the same positions and velocities.  The mixed
(FP32 storage) path is checked for its event statistics and bookkeeping.
"""
import numpy as np
import pytest

import oracle as O
import scdat_amd as S
from helpers import COND, EPS0, QE, rel_err

pytestmark = pytest.mark.gpu
ME, MP = S.ELECTRON_MASS, S.PROTON_MASS
SEED = 0x5EED2026


def _setup(n=6, npart=3000, te_ev=30.0, seed=31):
    m = S.box_mesh(n, conductor=COND)
    spec = S.box_spec(n, conductor=COND)
    xe, ve, ce = S.maxwellian_box(spec, npart, seed, 0, np.sqrt(te_ev * QE / ME))
    xi, vi, ci = S.maxwellian_box(spec, npart, seed, 1, np.sqrt(1.0 * QE / MP))
    w = 1e10 * (1 - 0.008) / npart
    wpe = np.sqrt(1e10 * QE ** 2 / (EPS0 * ME))
    return dict(mesh=m, spec=spec, e=(xe, ve, ce), i=(xi, vi, ci), w=w, dt=0.2 / wpe)


PROCS = [
    dict(type=O.MCC_ELASTIC, species=0, density=2e18, sigma_max=2e-19, reference_energy=10.0, power_index=0.3),
    dict(type=O.MCC_EXCITATION, species=0, density=2e18, threshold=11.5, sigma_max=2.5e-19, energy_loss=11.5),
    dict(type=O.MCC_IONIZATION, species=0, e_species=0, i_species=2, density=2e18, threshold=13.6,
         sigma_max=1e-17),
    dict(type=O.MCC_CHARGE_EXCHANGE, species=1, density=1e20, sigma_max=1.2e-19, reference_energy=1.0,
         power_index=0.5),
]


def _dev_proc(p):
    kw = dict(target_density=p["density"], sigma_max_m2=p["sigma_max"])
    for src, dst in (("threshold", "threshold_ev"), ("energy_loss", "energy_loss_ev"),
                     ("reference_energy", "reference_energy_ev"), ("power_index", "power_index"),
                     ("e_species", "product_electron_species"), ("i_species", "product_ion_species")):
        if src in p:
            kw[dst] = p[src]
    if p["type"] == O.MCC_EXCITATION:
        kw.update(alpha=1.2, beta=0.35)
    return S.mcc_process(p["type"], p["species"], **kw)


def _pair(st, precision, procs=PROCS, groups_bc=S.BC_CONDUCTOR):
    m = st["mesh"]
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        o.set_group(g, groups_bc)
    o.set_poisson(0, EPS0, 1, 1e-12, 50000)
    o.set_push(st["dt"], 1, 1e-3, 64, 0)
    for q, mass in ((-QE, ME), (QE, MP), (QE, 16 * MP)):
        o.add_species(q, mass, st["w"])
    o.set_particles(0, *st["e"])
    o.set_particles(1, *st["i"])
    o.set_particles(2, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    o.mcc_configure(procs)
    o.set_seed(SEED)
    c = S.Context(precision=precision, dt=st["dt"], pcg_tol=1e-12 if precision else 1e-8, sort_interval=0)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        c.set_boundary(g, groups_bc)
    n = len(st["e"][0])
    for q, mass in ((-QE, ME), (QE, MP), (QE, 16 * MP)):
        c.add_species(q, mass, st["w"], 4 * n)
    c.set_particles(0, *st["e"])
    c.set_particles(1, *st["i"])
    c.mcc_configure([_dev_proc(p) for p in procs])
    c.set_seed(SEED)
    return o, c


def _by_id(x, v, cell, ids):
    a = cell >= 0
    order = np.argsort(ids[a], kind="stable")
    return ids[a][order], x[a][order], v[a][order]


def _compare_species(o, c, s, vscale, xtol, vtol):
    xo, vo, co = o.get_particles(s)
    io_ = o.get_ids(s)
    xc, vc, cc, ic = c.get_particles(s)
    ia, xa, va = _by_id(xo, vo, co, io_)
    ib, xb, vb = _by_id(xc, vc, cc, ic)
    assert np.array_equal(ia, ib), (s, len(ia), len(ib), np.setdiff1d(ia, ib)[:5], np.setdiff1d(ib, ia)[:5])
    if len(ia):
        assert np.max(np.abs(xa - xb)) < xtol
        assert np.max(np.abs(va - vb)) < vtol * vscale
    return len(ia)


def test_collisions_fp64_match_oracle_event_for_event():
    st = _setup()
    o, c = _pair(st, precision=1)
    c.initialize()
    o.deposit()
    o.solve()
    o.efield()
    for _ in range(3):
        o.step(1)
        c.step(1)
    eo, ec = o.mcc_counts(), c.event_counts()
    print("oracle", eo, "device", ec)
    assert list(eo["collisions"][:4]) == ec["collisions"][:4]
    assert eo["collisions"][:3].sum() > 50 and eo["created"] > 10 and eo["removed"] > 10
    assert eo["created"] == ec["created"] and eo["removed"] == ec["removed"] and ec["overflow"] == 0
    # null-collision pre-selection: the same candidates on both sides (this
    # case is collisional, nu dt ~ 1, so most particles are candidates), and
    # the majorant held for every one
    assert ec["mcc_candidates"] == eo["candidates"] > eo["collisions"][:4].sum()
    assert ec["mcc_majorant_exceeded"] == 0 and eo["majorant_exceeded"] == 0
    vth = np.sqrt(30 * QE / ME)
    # velocities carry the 1e-12 solver tolerance of three field solves
    # through the pushes (relative 1e-9); collision outcomes themselves agree
    n_alive = [_compare_species(o, c, s, vth, 1e-9, 1e-7) for s in range(3)]
    assert n_alive[2] == eo["created"] // 2
    _, phi_o, _ = o.fields()
    _, phi_c, _ = c.fields()
    assert rel_err(phi_c, phi_o) < 1e-7


def test_collisions_mixed_precision_statistics_and_bookkeeping():
    st = _setup(npart=20000)
    o, c = _pair(st, precision=0)
    c.initialize()
    o.deposit()
    o.solve()
    o.efield()
    o.step(1)
    c.step(1)
    eo, ec = o.mcc_counts(), c.event_counts()
    print("oracle", eo, "device", ec)
    for k in range(4):
        a, b = eo["collisions"][k], ec["collisions"][k]
        # same random streams: FP32 velocities flip only decisions within
        # rounding of the probability, so counts agree to a few events
        assert abs(a - b) <= 3 + 0.01 * a, (k, a, b)
    assert ec["created"] % 2 == 0 and ec["overflow"] == 0
    _, _, cell2, ids2 = c.get_particles(2)
    assert (cell2 >= 0).sum() == ec["created"] // 2 and np.all(ids2 < 0)
    _, v2, _, _ = c.get_particles(2)
    assert np.all(v2[cell2 >= 0] == 0.0)  # ions born at rest (not yet pushed)


def _see_pair(precision, model=S.SEE_STERNGLASS, n=4000, e_ev=300.0):
    m = S.box_mesh(10, conductor=COND)
    spec = S.box_spec(10, conductor=COND)
    rng = np.random.default_rng(8)
    x = np.column_stack([np.full(n, 0.385), rng.uniform(0.41, 0.59, n), rng.uniform(0.41, 0.59, n)])
    sp = np.sqrt(2 * e_ev * QE / ME)
    v = np.column_stack([np.full(n, sp), rng.normal(0, 0.2 * sp, n), np.zeros(n)])
    cell = S.loading.box_locate(spec, x)
    dt = 0.03 / sp
    w = 10.0
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, EPS0, 1, 1e-12, 50000)
    o.set_push(dt, 0, 1e-3, 1, 0)
    o.add_species(-QE, ME, w)
    o.set_particles(0, x, v, cell)
    o.set_seed(SEED)
    c = S.Context(precision=precision, dt=dt, pcg_tol=1e-12 if precision else 1e-8, sort_interval=0,
                  adaptive_substep=0)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        c.set_boundary(g, S.BC_CONDUCTOR)
    c.add_species(-QE, ME, w, 8 * n)
    c.set_particles(0, x, v, cell)
    c.set_seed(SEED)
    for g in (6, 7):
        o.set_emission(g, see_model=model, delta_max=2.0, energy_max=400.0, see_temperature=2.0,
                       electron_species=0)
        c.set_surface_emission(g, see_model=model, electron_species=0, delta_max=2.0, energy_max=400.0,
                               see_temperature=2.0)
    return o, c, sp


@pytest.mark.parametrize("model", [S.SEE_STERNGLASS, S.SEE_REFERENCE])
def test_secondary_emission_fp64_matches_oracle(model):
    o, c, sp = _see_pair(1, model)
    c.initialize()
    o.deposit()
    o.solve()
    o.efield()
    o.step(1)
    c.step(1)
    eo, ec = o.mcc_counts(), c.event_counts()
    print("oracle", eo["see_emitted"], "device", ec["see_emitted"])
    assert eo["see_emitted"] == ec["see_emitted"] > 1000 and ec["impacts_dropped"] == 0
    _compare_species(o, c, 0, sp, 1e-9, 1e-9)
    led_o, led_c = o.ledger(12), c.ledger(12)
    assert led_c[6]["emitted_electron_particles"] == led_o[6, 6]
    assert abs(c.stats()["conductor_charge"] - o.conductor_charge()) <= 1e-9 * abs(o.conductor_charge())


def test_secondary_emission_mixed_precision_counts():
    o, c, sp = _see_pair(0, S.SEE_STERNGLASS)
    c.initialize()
    c.step(1)
    ec = c.event_counts()
    led = c.ledger(12)
    n_abs = led[6]["absorbed_electron_particles"]
    assert n_abs > 3000
    # Sternglass at 300 eV-ish with the spread of incidence energies: between
    # floor and ceil of the yield range per impact
    assert ec["see_emitted"] == led[6]["emitted_electron_particles"] + led[7]["emitted_electron_particles"]
    q_exp = -QE * 10.0 * (n_abs + led[7]["absorbed_electron_particles"]) + QE * 10.0 * ec["see_emitted"]
    assert abs(c.stats()["conductor_charge"] - q_exp) <= 1e-9 * abs(q_exp)


def test_photoemission_fp64_matches_oracle():
    st = _setup(n=6, npart=500)
    o, c = _pair(st, precision=1, procs=[])
    for g in (10, 11, 2):  # conductor y-min, z-min faces and the wall y-min (grounded)
        o.set_emission(g, photo_j=0.02, photo_temperature=2.0, photo_reference_v=0.0, electron_species=0)
        c.set_surface_emission(g, electron_species=0, photo_j=0.02, photo_temperature=2.0)
    c.initialize()
    o.deposit()
    o.solve()
    o.efield()
    for _ in range(3):
        o.step(1)
        c.step(1)
    eo, ec = o.mcc_counts(), c.event_counts()
    print("photo oracle", eo["photo_emitted"], "device", ec["photo_emitted"])
    assert eo["photo_emitted"] == ec["photo_emitted"] > 50
    # velocities carry the solver tolerance of three field solves (see above)
    _compare_species(o, c, 0, np.sqrt(2 * QE / ME), 1e-9, 1e-7)
    # same bound as the FP64 full-cycle parity test (1e-12 solves, three steps)
    assert abs(c.stats()["floating_potential"] - o.floating_potential()) <= 1e-6 * abs(o.floating_potential()) + 1e-12


def test_collision_handler_mirror_reports_the_device_steps():
    """MonteCarloCollisionHandler + executeBackgroundMccStep after each
    Context.step (the reference main loop's order): per-step counts are the
    device's event-count differences, and a re-initialisation sets the
    effective densities on the device's processes."""
    from scdat_amd import collisions as CL
    st = _setup(n=6, npart=3000, te_ev=30.0, seed=33)
    m = st["mesh"]
    c = S.Context(precision=0, dt=st["dt"], pcg_tol=1e-10, sort_interval=4)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        c.set_boundary(g, S.BC_CONDUCTOR)
    c.add_species(-QE, ME, st["w"], 20000)
    c.add_species(QE, MP, st["w"], 20000)
    c.set_particles(0, *st["e"])
    c.set_particles(1, *st["i"])
    h = CL.MonteCarloCollisionHandler(random_seed=SEED, electron_species=0, ion_species=1)
    h.initializeDefaultConfiguration(neutral_density=3e19, ion_density=1e16, electron_density=1e16)
    h.bind(c)
    c.initialize()
    cfg = CL.BackgroundMccRuntimeConfig(CL.CollisionBackgroundDensities(3e19, 1e16, 1e16))
    tot = 0
    for k in range(4):
        before = c.event_counts()
        c.step(1)
        after = c.event_counts()
        if k == 2:
            cfg.reinitialize_configuration = True
        r = CL.executeBackgroundMccStep(c, h, st["dt"], 1.0, cfg)
        names = h.processNames()
        d = [a - b for a, b in zip(after["collisions"], before["collisions"])]
        assert r.collision_events == sum(d[:len(names)])
        assert r.ionization_events == d[names.index("ionization_e_n")]
        assert r.excitation_events == d[names.index("excitation_e_n")]
        assert r.charge_exchange_events == d[names.index("charge_exchange_i_n")]
        assert r.generated_particles == after["created"] - before["created"]
        assert r.active_particles == sum(c.particle_count(s)[0] for s in range(2))
        tot += r.collision_events
    assert tot > 0
    stt = h.getStatistics()
    assert stt.total_collisions == tot
    # the re-initialisation moved the electron-ion processes to the effective densities
    ev = CL.estimateCollisionBackgroundDensities(c, 1.0, cfg.floor_densities)
    assert h._procs["recombination_e_i"].process.target_density == pytest.approx(ev.ion_density_m3, rel=0.2)
    c.close()
