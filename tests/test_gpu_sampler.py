"""Live-PIC surface current sampler (scdat_sample_surface_current) against
the oracle running the same window, the concurrency of its probe windows,
and a surface-charging loop closing it with the circuit integrator."""
import numpy as np
import pytest

import oracle as O
import scdat_amd as S
from helpers import EPS0, QE, periodic_partners

pytestmark = pytest.mark.gpu
AMU = 1.66053906660e-27


def _oracle_window(cfg):
    spec, ne, ni, w = S.sampler_inputs(cfg)
    n = tuple(spec.n)
    m = S.box_mesh(n, lo=tuple(spec.lo), hi=tuple(spec.hi))
    z = m["xyz"][:, 2]
    kind = np.zeros(len(z), np.int32)
    dval = np.zeros(len(z))
    bot, top = np.abs(z - spec.lo[2]) < 1e-15, np.abs(z - spec.hi[2]) < 1e-15
    kind[bot | top] = 1
    dval[bot] = cfg.surface_potential_v
    dval[top] = cfg.plasma_reference_potential_v
    o = O.Oracle(m["xyz"], m["tets"], kind, dirichlet=dval, group=m["group"])
    o.set_partner(periodic_partners(m))
    Lx, Ly = spec.hi[0] - spec.lo[0], spec.hi[1] - spec.lo[1]
    for g, sh in ((0, [Lx, 0, 0]), (1, [-Lx, 0, 0]), (2, [0, Ly, 0]), (3, [0, -Ly, 0])):
        o.set_group(g, S.BC_PERIODIC, sh)
    o.set_group(4, S.BC_ABSORB)
    o.set_group(5, S.BC_ABSORB)
    o.set_poisson(0, EPS0, 1, 1e-12, 50000)
    mi = max(1.0, cfg.ion_mass_amu) * AMU
    vth = (np.sqrt(max(1e-3, cfg.electron_temperature_ev) * QE / S.ELECTRON_MASS),
           np.sqrt(max(1e-3, cfg.ion_temperature_ev) * QE / mi))
    for s, (q, mass, cnt) in enumerate(((-QE, S.ELECTRON_MASS, ne), (QE, mi, ni))):
        o.add_species(q, mass, w)
        x, v, c = S.maxwellian_box(spec, cnt, cfg.seed, s, vth[s])
        o.set_particles(s, x, S.sampler_direct(v), c)
    return o, Lx * Ly


def test_sampler_window_matches_oracle():
    cfg = S.sampler_config(surface_potential_v=2.0, particles_per_element=40, window_steps=20)
    out = S.sample_surface_current(cfg)
    o, area = _oracle_window(cfg)
    o.set_push(out["time_step"], 1, 1e-3, 64, 0)
    o.deposit()
    o.solve()
    o.efield()
    o.step(cfg.window_steps)
    led = o.ledger(6)
    T = out["time_step"] * cfg.window_steps
    je, ji = led[4, 0] / (T * area), led[4, 1] / (T * area)
    unit = QE * S.sampler_inputs(cfg)[3] / (T * area)
    print(out, je, ji, led[4, 4], led[4, 5])
    assert led[4, 4] > 20
    # the device solves to 1e-8 (the oracle 1e-12): at most one grazing
    # particle may land on the other side of the surface
    assert abs(out["electron_collection_current_density_a_per_m2"] - je) <= unit * 1.01
    assert abs(out["ion_collection_current_density_a_per_m2"] - ji) <= unit * 1.01


def test_probe_windows_run_concurrently_and_match_single_windows():
    kw = dict(surface_potential_v=1.0, particles_per_element=20, window_steps=16)
    d = 0.5
    c3 = S.sample_surface_current(S.sampler_config(**kw), probe_delta_v=d)
    jp = S.sample_surface_current(S.sampler_config(**dict(kw, surface_potential_v=1.0 + d)))
    jm = S.sample_surface_current(S.sampler_config(**dict(kw, surface_potential_v=1.0 - d)))
    j0 = S.sample_surface_current(S.sampler_config(**kw))
    assert c3["net_collection_current_density_a_per_m2"] == j0["net_collection_current_density_a_per_m2"]
    want = (jp["net_collection_current_density_a_per_m2"] - jm["net_collection_current_density_a_per_m2"]) / (2 * d)
    assert c3["current_derivative_a_per_m2_per_v"] == want


def test_sampler_with_collisions_matches_oracle_counts():
    cfg = S.sampler_config(electron_temperature_ev=30.0, neutral_density_m3=5e21, enable_mcc=1,
                           particles_per_element=40, window_steps=20)
    out = S.sample_surface_current(cfg)
    o, area = _oracle_window(cfg)
    o.add_species(QE, 40 * S.PROTON_MASS, S.sampler_inputs(cfg)[3])
    o.set_particles(2, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    # the SurfacePicV1 set (CollisionPicAdapter.cpp:129-149)
    n = cfg.neutral_density_m3
    o.mcc_configure([
        dict(type=O.MCC_ELASTIC, species=0, density=n, sigma_max=2e-19, reference_energy=1.0, power_index=0.5),
        dict(type=O.MCC_EXCITATION, species=0, density=n, threshold=11.5, sigma_max=2.5e-20, energy_loss=11.5),
        dict(type=O.MCC_IONIZATION, species=0, e_species=0, i_species=2, density=n, threshold=15.76,
             sigma_max=3e-20, energy_loss=15.76)])
    o.set_seed(cfg.seed ^ 0x9E3779B9)
    o.set_push(out["time_step"], 1, 1e-3, 64, 0)
    o.deposit()
    o.solve()
    o.efield()
    o.step(cfg.window_steps)
    tot = int(o.mcc_counts()["collisions"].sum())
    led = o.ledger(6)
    print(out, tot, led[4])
    assert tot > 50
    # same Philox streams; the device's 1e-8 field solve moves velocities by
    # ~1e-8 relative, which flips at most a handful of decisions
    assert abs(out["total_collisions"] - tot) <= 3
    T = out["time_step"] * cfg.window_steps
    je = led[4, 0] / (T * area)
    unit = QE * S.sampler_inputs(cfg)[3] / (T * area)
    assert abs(out["electron_collection_current_density_a_per_m2"] - je) <= 3 * unit


def test_surface_charging_loop_reduces_net_current():
    """A floating patch (capacitance C to a grounded body) charged by the
    sampled plasma current: circuit.advanceImplicit with I = J A and
    dI/dV = dJ/dV A each step (PicMccSurfaceCurrentSampler + SurfaceCircuitCoupling,
    the toolkit's live-PIC charging loop)."""
    area = 1e-4
    circ = S.Circuit()
    body = circ.addNode(fixed=True)
    patch = circ.addNode(0.0, capacitance_f=1e-12)
    circ.addBranch(patch, body, 0.0)
    v = 0.0
    hist = []
    for _ in range(6):
        cfg = S.sampler_config(surface_potential_v=v, particles_per_element=30, window_steps=16)
        s = S.sample_surface_current(cfg, probe_delta_v=0.5)
        i_a = s["net_collection_current_density_a_per_m2"] * area
        didv = min(s["current_derivative_a_per_m2_per_v"] * area, 0.0)
        pv, _, _, _ = circ.advanceImplicit(1e-6, [0.0, i_a], [0.0, didv], max_delta_potential_v=2.0)
        hist.append((v, i_a))
        v = pv[patch]
    print(hist)
    # electrons (faster) charge the patch negative; the collected net current
    # shrinks in magnitude as the potential approaches floating
    assert hist[-1][0] < 0.0
    assert abs(hist[-1][1]) < abs(hist[0][1])
