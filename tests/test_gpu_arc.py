"""Field-emission cathode and the vacuum-arc PIC window on the device, through
the C ABI.  Ion impacts come from the device push; the emitted electrons are
checked against the oracle's restatement of
FieldEmissionBoundaryCondition::processParticle (count, id, position,
velocity), the ledger against the counts, and the window's currents against
the ledger identities."""
import numpy as np
import pytest

import oracle as O
import scdat_amd as S
from scdat_amd import arc

pytestmark = pytest.mark.gpu
QE = S.ELEMENTARY_CHARGE
L = W = 1.0e-4
T = 5.0e-4


def _cathode_ctx(precision, dt, cap=4096, **emission):
    m = arc.plate_mesh(L, W, T)
    c = S.Context(precision=precision, dt=dt, adaptive_substep=0, sort_interval=0)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    c.set_boundary(0, S.BC_PERIODIC, 1, [L, 0, 0])
    c.set_boundary(1, S.BC_PERIODIC, 0, [-L, 0, 0])
    c.set_boundary(2, S.BC_PERIODIC, 3, [0, W, 0])
    c.set_boundary(3, S.BC_PERIODIC, 2, [0, -W, 0])
    c.set_boundary(arc.ZMAX, S.BC_ABSORB)
    e = c.add_species(-QE, S.ELECTRON_MASS, 1.0, cap)
    i = c.add_species(QE, S.PROTON_MASS, 1.0, cap)
    c.set_field_emission(arc.ZMIN, S.field_emission(electron_species=e, **emission))
    c.initialize()
    return c, m, e, i


def _electrons(c, e):
    x, v, cell, ids = c.get_particles(e)
    o = np.argsort(ids)
    return x[o], v[o], cell[o], ids[o]


@pytest.mark.parametrize("precision", [1, 0])
def test_sub_particle_budget_on_device(precision):
    # VacuumArc_presets_smoke_test.cpp:197-244 driven by device impacts: 0.25
    # expected field electrons per ion, one ion per step; the fourth emits one
    dt = 1.0e-9
    j = 0.25 * QE / (1.0e-8 * dt)
    c, m, e, i = _cathode_ctx(precision, dt, emission_area_m2=1.0e-8, emitted_electron_speed_m_per_s=1.0e6,
                              max_emitted_particles_per_event=16, emission_current_density_a_per_m2=j)
    x0 = np.array([[3.1e-5, 6.7e-5, 2.0e-7]])
    for k in range(4):
        print("budget step", k, "precision", precision, flush=True)
        c.append_particles(i, x0, [[0.0, 0.0, -1.0e5]])
        c.step(1)
        n_e = c.particle_count(e)[0]
        assert n_e == (1 if k == 3 else 0)
    x, v, cell, ids = _electrons(c, e)
    tol = 1e-15 if precision == 1 else 2e-11
    np.testing.assert_allclose(x[0, :2], x0[0, :2], rtol=0, atol=tol)
    assert abs(x[0, 2] - 1.0e-10) < (1e-15 if precision == 1 else 1e-16)
    np.testing.assert_allclose(v[0], [0, 0, 1.0e6], rtol=0, atol=1e-6)
    assert ids[0] == 0 and cell[0] >= 0
    led = c.ledger(6)[arc.ZMIN]
    assert led["absorbed_ion_particles"] == 4
    assert led["emitted_electron_particles"] == 1
    assert led["emitted_electron_charge_c"] == -QE
    c.close()


@pytest.mark.parametrize("precision", [1, 0])
def test_many_impacts_match_oracle(precision):
    """40 ions hit the cathode in one step; field (0.37 per ion) and secondary
    (yield 0.75) electrons in ion-id order against the oracle."""
    dt = 1.0e-9
    kw = dict(emission_area_m2=1.0e-8, emission_current_density_a_per_m2=0.37 * QE / (1.0e-8 * dt),
              emitted_electron_speed_m_per_s=2.0e6, max_emitted_particles_per_event=64,
              enable_secondary_electron_emission=1, secondary_electron_yield_per_ion=0.75,
              secondary_electron_speed_m_per_s=5.0e5, max_secondary_particles_per_event=64)
    c, m, e, i = _cathode_ctx(precision, dt, **kw)
    rng = np.random.default_rng(11)
    n = 40
    x0 = np.column_stack([rng.uniform(0.05, 0.95, n) * L, rng.uniform(0.05, 0.95, n) * W,
                          rng.uniform(1e-7, 5e-7, n)])
    v0 = np.tile([0.0, 0.0, -2.0e6], (n, 1))
    print("many impacts precision", precision, flush=True)
    c.append_particles(i, x0, v0)
    c.step(1)
    if precision == 0:
        x0 = x0.astype(np.float32).astype(np.float64)
    hits = [((x0[k, 0], x0[k, 1], 0.0), (0.0, 0.0, 1.0), 1.0) for k in range(n)]
    p = {f: getattr(S.field_emission(**kw), f) for f, _ in S.FieldEmission._fields_}
    p["enable_secondary_electron_emission"] = True
    out, _ = O.field_emission(p, hits, dt)
    x, v, cell, ids = _electrons(c, e)
    assert len(ids) == len(out) and len(out) > 20
    np.testing.assert_array_equal(ids, [o[0] for o in out])
    xo = np.array([o[1] for o in out])
    vo = np.array([o[2] for o in out])
    tol = 1e-15 if precision == 1 else 1e-16
    np.testing.assert_allclose(x[:, :2], xo[:, :2], rtol=0, atol=tol)
    np.testing.assert_allclose(x[:, 2], xo[:, 2], rtol=0, atol=1e-16)
    np.testing.assert_allclose(v, vo, rtol=0, atol=1e-6)
    assert np.all(cell >= 0)
    led = c.ledger(6)[arc.ZMIN]
    assert led["absorbed_ion_particles"] == n
    assert led["emitted_electron_particles"] == len(out)
    assert abs(led["emitted_electron_charge_c"] + QE * len(out)) < 1e-30
    c.close()


def test_set_potential_moves_dirichlet_values():
    # PICCycle.cpp:674-688: a DIRICHLET node's value is the potential set
    # last; with no charge the solve is the linear profile between the plates
    m = arc.plate_mesh(L, W, T)
    c = S.Context(precision=1, dt=1e-10, pcg_tol=1e-12)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    c.add_species(-QE, S.ELECTRON_MASS, 1.0, 1)
    c.initialize()
    z = m["xyz"][:, 2]
    phi = np.zeros(len(z))
    phi[np.abs(z) < 1e-12] = -1000.0
    c.set_potential(phi)
    c.solve()
    c.compute_efield()
    _, phi_c, E_c = c.fields()
    np.testing.assert_allclose(phi_c, -1000.0 * (1.0 - z / T), rtol=0, atol=1e-8)
    np.testing.assert_allclose(E_c, np.tile([0.0, 0.0, -1000.0 / T], (len(z), 1)), rtol=0, atol=1e-3)
    c.close()


def test_window_advance_currents_follow_the_ledger():
    cfg = arc.WindowConfig(surface_potential_v=-200.0, enable_secondary_electron_emission=True,
                           secondary_electron_yield_per_ion=0.5)
    w = arc.BoundaryWindow(cfg)
    dt = 1.0e-10
    j = 0.8e0 * QE / (w.area * dt) * 5.0  # 4 field electrons expected per impact
    probes = 0
    sums = np.zeros(3)
    surf = 0.0
    for k in range(8):
        print("window advance", k, flush=True)
        probes += arc.probe_count(j, w.area, dt)
        st = w.advance(j, dt)
        assert np.isfinite(st.pic_boundary_net_current_a)
        assert st.pic_boundary_net_current_a == pytest.approx(
            st.pic_absorbed_electron_current_a + st.pic_absorbed_ion_current_a + st.pic_emitted_electron_current_a,
            rel=1e-12, abs=1e-30)
        sums += np.array([st.pic_absorbed_electron_current_a, st.pic_absorbed_ion_current_a,
                          st.pic_emitted_electron_current_a]) * dt
        surf += st.pic_surface_charge_delta_c
    led = w.ctx.ledger(6)
    cat, ano = led[arc.ZMIN], led[arc.ZMAX]
    hits = cat["absorbed_ion_particles"]
    assert 0 < hits <= probes
    # emitted count from the oracle's accumulator over the same impacts
    p = {f: getattr(w.emission, f) for f, _ in S.FieldEmission._fields_}
    p["emission_current_density_a_per_m2"] = j
    p["enable_secondary_electron_emission"] = True
    out, _ = O.field_emission(p, [((0, 0, 0), (0, 0, 1), 1.0)] * hits, dt)
    assert cat["emitted_electron_particles"] == len(out)
    assert sums[2] == pytest.approx(QE * len(out), rel=1e-12)
    assert sums[1] == pytest.approx(QE * hits, rel=1e-12)
    n_abs_e = cat["absorbed_electron_particles"] + ano["absorbed_electron_particles"]
    assert sums[0] == pytest.approx(-QE * n_abs_e, rel=1e-12, abs=1e-30)
    assert surf == pytest.approx(QE * (hits + ano["absorbed_ion_particles"]) - QE * n_abs_e, rel=1e-9, abs=1e-30)
    assert n_abs_e > 0  # emitted electrons reach the anode
    w.close()
