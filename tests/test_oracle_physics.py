"""The oracle against closed forms, invariants and the reference's own unit
tests (re-expressed: GoogleTest is not buildable offline).  Each check is
chosen so that a plausible mistake (sign, dropped term, transposed index)
fails it."""
import numpy as np
import pytest

import oracle as O
import scdat_amd as S

EPS0 = 8.8541878128e-12
QE, ME, MP = S.ELEMENTARY_CHARGE, S.ELECTRON_MASS, S.PROTON_MASS
COND = ((0.4, 0.4, 0.4), (0.6, 0.6, 0.6))


def test_boris_known_answer_reference_unit_test():
    # Tools/Particle/test/ParticleDefinitions_test.cpp:72-91: E = (2,0,0),
    # q = m = 1, dt = 0.5 from rest gives v = (1,0,0) (and x = 0.5).
    v = O.boris([0, 0, 0], [2, 0, 0], [0, 0, 0], 1.0, 0.5)
    assert np.allclose(v, [1, 0, 0], atol=1e-15)


def test_boris_pure_rotation_conserves_speed_and_angle():
    v0 = np.array([3.0, -1.0, 2.0])
    B = np.array([0.0, 0.0, 2.0])
    qm, dt = 1.5, 0.1
    v = O.boris(v0, [0, 0, 0], B, qm, dt)
    assert abs(np.linalg.norm(v) - np.linalg.norm(v0)) < 1e-14
    # rotation angle of the Boris scheme: 2 atan(qm |B| dt / 2)
    a, b = np.r_[v0[:2], 0.0], np.r_[v[:2], 0.0]
    th = np.arctan2(np.cross(a, b)[2], np.dot(a, b))
    assert abs(abs(th) - 2 * np.arctan(qm * 2.0 * dt / 2)) < 1e-13
    assert v[2] == v0[2]
    # positive charge gyrates clockwise about +z
    assert th < 0


def test_boris_exb_drift():
    # E x B drift: average velocity E x B / |B|^2 independent of charge sign
    E = np.array([0.0, 1e3, 0.0])
    B = np.array([0.0, 0.0, 0.1])
    drift = np.cross(E, B) / np.dot(B, B)
    for qm in (1e8, -1e8):
        v = np.zeros(3)
        acc = np.zeros(3)
        dt = 2 * np.pi / (abs(qm) * 0.1) / 200
        for _ in range(2000):
            v = O.boris(v, E, B, qm, dt)
            acc += v
        assert np.allclose(acc / 2000, drift, rtol=2e-3, atol=1e-4 * np.linalg.norm(drift))


def test_adaptive_dt_closed_forms():
    # PICParticleAlgorithms.cpp:750-801
    v0 = [1e5, 0, 0]
    qm = -1.7588e11
    # weak field: full step
    assert O.adaptive_dt(v0, [1e-3, 0, 0], [0, 0, 0], qm, 1e-9) == 1e-9
    # strong field: 0.1 |v| / |qE/m|
    E = [100.0, 0, 0]
    expect = 0.1 * 1e5 / (100.0 * abs(qm))
    assert abs(O.adaptive_dt(v0, E, [0, 0, 0], qm, 1e-8) - expect) < 1e-22
    # slow particle (|v| <= 1 m/s): no velocity limit
    assert O.adaptive_dt([0.5, 0, 0], E, [0, 0, 0], qm, 1e-6) == 1e-6
    # floor: 1e-3 of the remaining time
    assert O.adaptive_dt(v0, [1e12, 0, 0], [0, 0, 0], qm, 1e-6) == pytest.approx(1e-9)
    # cyclotron limit 0.2 pi / omega_c
    B = [0, 0, 1e-3]
    assert O.adaptive_dt(v0, [0, 0, 0], B, qm, 1e-8) == pytest.approx(0.2 * np.pi / (1e-3 * abs(qm)))


def _oracle(n=6, cond=None, dirichlet=None):
    m = S.box_mesh(n, conductor=cond)
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], dirichlet=dirichlet, group=m["group"])
    return m, o


def test_fem_patch_test_linear_potential_is_exact():
    m = S.box_mesh(5)
    phi = 3 * m["xyz"][:, 0] - 2 * m["xyz"][:, 1] + 0.5 * m["xyz"][:, 2] + 1
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], dirichlet=phi, group=m["group"])
    o.set_poisson(0, 1.0, 1, 1e-14, 10000)
    o.solve()
    o.efield()
    _, p, E = o.fields()
    assert np.max(np.abs(p - phi)) < 1e-12
    assert np.max(np.abs(E - [-3, 2, -0.5])) < 1e-11


def test_parallel_plate_reference_unit_test():
    # Tools/FieldSolver/test/PoissonSolver_test.cpp:48-117: 10 V / 0 V plates,
    # Dirichlet values exact, interior within [0, 10], E finite.  With P1 FEM
    # and no charge the solution is the exact linear profile.
    m = S.box_mesh((2, 2, 8))
    z = m["xyz"][:, 2]
    kind = np.where((z < 1e-12) | (z > 1 - 1e-12), 1, 0).astype(np.int32)
    dval = np.where(z < 1e-12, 10.0, 0.0)
    o = O.Oracle(m["xyz"], m["tets"], kind, dirichlet=dval, group=m["group"])
    o.set_poisson(0, EPS0, 1, 1e-14, 10000)
    o.solve()
    o.efield()
    _, p, E = o.fields()
    assert np.max(np.abs(p[kind == 1] - dval[kind == 1])) < 1e-8
    assert p.min() >= -1e-9 and p.max() <= 10 + 1e-9
    assert np.max(np.abs(p - 10 * (1 - z))) < 1e-10
    assert np.allclose(E[:, 2], 10.0, rtol=1e-9)


def test_floating_conductor_gauss_law_and_superposition():
    m, o = _oracle(8, COND)
    o.set_poisson(0, EPS0, 1, 1e-14, 100000)
    nn = len(m["kind"])
    # 1) bare conductor with charge Q: equipotential, V = Q / C, linear in Q
    o.set_rho(np.zeros(nn))
    o.set_conductor_charge(1e-12)
    o.solve()
    v1 = o.floating_potential()
    o.set_conductor_charge(2e-12)
    o.solve()
    v2 = o.floating_potential()
    assert v1 > 0 and abs(v2 - 2 * v1) < 1e-12 * v2
    # capacitance of a 0.2 m cube centred in a grounded 1 m box: between the
    # isolated cube (0.0734 eps... C = 0.6607 * 4 pi eps0 a, a = 0.2 -> 14.7 pF)
    # and twice that; the grounded box raises it moderately
    C = 1e-12 / v1
    assert 14e-12 < C < 30e-12
    # 2) Gauss: the charge induced on the conductor (K phi on conductor rows)
    # equals Q plus the charge deposited on conductor nodes
    K = o.kfull()
    rho = np.random.default_rng(1).normal(size=nn) * 1e-13
    rho[m["kind"] == 1] = 0
    o.set_rho(rho)
    o.set_conductor_charge(3e-13)
    o.solve()
    _, phi, _ = o.fields()
    cond = m["kind"] == 2
    assert np.ptp(phi[cond]) == 0.0
    induced = (K @ phi)[cond].sum()
    assert abs(induced - (3e-13 + rho[cond].sum())) < 1e-9 * abs(induced)
    # free rows satisfy K phi = rho
    free = m["kind"] == 0
    assert np.max(np.abs((K @ phi)[free] - rho[free])) < 1e-9 * np.max(np.abs(rho))


def test_lsq_gradient_exact_for_linear_fields_on_cut_mesh():
    m, o = _oracle(6, COND)
    o.set_poisson(0, EPS0, 1, 1e-12, 100)
    g = np.array([0.7, -1.3, 2.1])
    o.set_phi(m["xyz"] @ g + 0.25)
    o.efield()
    E = o.fields()[2]
    assert np.max(np.abs(E + g)) < 1e-11


def test_deposition_node_centroid_and_conservation():
    m, o = _oracle(4)
    t = 5
    vx = m["xyz"][m["tets"][t]]
    pts = np.vstack([vx[2], vx.mean(0)])
    s = o.add_species(-QE, ME, 3.0)
    o.set_particles(s, pts, np.zeros_like(pts), [t, t])
    o.deposit()
    rho = o.fields()[0]
    q = -QE * 3.0
    expect = np.zeros(len(m["xyz"]))
    expect[m["tets"][t]] += q / 4
    expect[m["tets"][t][2]] += q
    assert np.max(np.abs(rho - expect)) < 1e-12 * abs(q)


def test_fixed_point_deposit_conserves_charge_exactly_and_is_order_free():
    m, o = _oracle(6, COND)
    spec = S.box_spec(6, conductor=COND)
    x, v, c = S.maxwellian_box(spec, 20000, 9, 0, 1.0)
    A, X0, _ = o.geometry()
    maps = np.concatenate([A, X0], 1).astype(np.float32)
    x32 = x.astype(np.float32)
    cs = -256
    r1 = O.deposit_fixed_f32(maps, m["tets"], x32, c, cs, len(m["xyz"]))
    assert r1.sum() == cs * (1 << 24) * len(x)  # exact integer conservation
    perm = np.random.default_rng(2).permutation(len(x))
    r2 = O.deposit_fixed_f32(maps, m["tets"], x32[perm], c[perm], cs, len(m["xyz"]))
    assert np.array_equal(r1, r2)
    # and it equals the FP64 deposit to FP32 resolution (position rounding
    # 6e-8 m times |grad lambda| ~ 12/m, summed over ~230 particles per node)
    s = o.add_species(-QE, ME, 1.0)
    o.set_particles(s, x, v, c)
    o.deposit()
    unit = QE / 256 / (1 << 24)
    rho64 = o.fields()[0]
    assert np.max(np.abs(r1 * unit - rho64)) < 1e-5 * np.max(np.abs(rho64))


def test_walk_conserves_particles_and_periodic_wraps():
    # PICCycleBoundary_test.cpp:121-142 analogue: periodic x/y keep particles in
    m = S.box_mesh((2, 2, 6))
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    o.set_poisson(0, EPS0, 1, 1e-10, 100)
    partner = _periodic_partners(m)
    o.set_partner(partner)
    for g, sh in ((0, [1, 0, 0]), (1, [-1, 0, 0]), (2, [0, 1, 0]), (3, [0, -1, 0])):
        o.set_group(g, 1, sh)
    for g in (4, 5):
        o.set_group(g, 0)
    o.set_push(1e-6, 0, 1e-3, 1, 0)
    s = o.add_species(-QE, ME, 1.0)
    x = np.array([[0.95, 0.5, 0.5], [0.1, 0.05, 0.5]])
    v = np.array([[1e5, 0, 0], [-2e5, -3e5, 0]])
    cell = S.loading.box_locate(S.box_spec((2, 2, 6)), x)
    o.set_particles(s, x, v, cell)
    o.set_efield(np.zeros((len(m["xyz"]), 3)))
    o.push()
    xn, _, cn = o.get_particles(s)
    assert (cn >= 0).all()
    assert np.allclose(xn[0], [0.05, 0.5, 0.5], atol=1e-12)
    assert np.allclose(xn[1], [0.9, 0.75, 0.5], atol=1e-12)
    # absorbing z walls: one electron through ZMax records one electron,
    # negative charge (PICCycleBoundary_test.cpp:144-168)
    s2 = o.add_species(-QE, ME, 1.0)
    x = np.array([[0.5, 0.5, 0.99]])
    o.set_particles(s2, x, np.array([[0, 0, 1e5]]), S.loading.box_locate(S.box_spec((2, 2, 6)), x))
    o.push()
    assert o.get_particles(s2)[2][0] == -1
    led = o.ledger(6)
    assert led[5, 4] == 1 and led[5, 0] < 0


def _periodic_partners(m):
    """Partner tet across periodic x and y faces, by face-centroid matching."""
    xyz, tets, grp = m["xyz"], m["tets"], m["group"]
    partner = -np.ones_like(tets)
    cent = {}
    for t, k in zip(*np.nonzero(grp >= 0)):
        c = xyz[np.delete(tets[t], k)].mean(0)
        cent[(t, k)] = (grp[t, k], c)
    for ga, gb, ax in ((0, 1, 0), (2, 3, 1)):
        A = [(key, c) for key, (g, c) in cent.items() if g == ga]
        B = [(key, c) for key, (g, c) in cent.items() if g == gb]
        for (ka, ca) in A:
            for (kb, cb) in B:
                d = cb - ca
                d[ax] = 0
                if np.linalg.norm(d) < 1e-9:
                    partner[ka] = kb[0]
                    partner[kb] = ka[0]
    return partner


def _boundary_oracle(n, cond=None):
    m = S.box_mesh(n, conductor=cond)
    o = O.Oracle(m["xyz"], m["tets"], m["kind"], group=m["group"])
    o.set_poisson(0, EPS0, 1, 1e-10, 100)
    o.set_push(1e-6, 0, 1e-3, 1, 0)
    o.set_efield(np.zeros((len(m["xyz"]), 3)))
    return m, o, S.box_spec(n, conductor=cond)


def test_specular_reflection_closed_form():
    # Mirror image about the wall plane, normal velocity flipped; a corner hit
    # reflects about both planes (PICCycle.cpp:1221-1293 reflect branch).
    m, o, spec = _boundary_oracle((2, 2, 6))
    for g in range(6):
        o.set_group(g, S.BC_REFLECT)
    s = o.add_species(-QE, ME, 1.0)
    x = np.array([[0.5, 0.5, 0.95], [0.95, 0.5, 0.97], [0.3, 0.02, 0.5]])
    v = np.array([[1e5, 0, 1e5], [1e5, 2e4, 1e5], [0, -7e4, 3e4]])
    o.set_particles(s, x, v, S.loading.box_locate(spec, x))
    o.push()
    xn, vn, cn = o.get_particles(s)
    want_x = np.array([[0.6, 0.5, 0.95], [0.95, 0.52, 0.93], [0.3, 0.05, 0.53]])
    want_v = np.array([[1e5, 0, -1e5], [-1e5, 2e4, -1e5], [0, 7e4, 3e4]])
    assert (cn >= 0).all()
    np.testing.assert_allclose(xn, want_x, atol=1e-12)
    np.testing.assert_allclose(vn, want_v, rtol=1e-14)
    assert np.abs(o.ledger(6)).sum() == 0


def test_conductor_absorption_charges_conductor_and_ledger():
    # An electron and a proton reaching the conductor's x-min face (group 6) are
    # absorbed there; each adds q*w to the conductor charge and one count to
    # the face's ledger row (PICCycle.cpp:1295-1329).
    m, o, spec = _boundary_oracle(10, COND)
    for g in range(6, 12):
        o.set_group(g, S.BC_CONDUCTOR)
    w = 3.0
    se = o.add_species(-QE, ME, w)
    si = o.add_species(QE, MP, w)
    x = np.array([[0.35, 0.5, 0.5]])
    o.set_particles(se, x, np.array([[1e5, 0, 0]]), S.loading.box_locate(spec, x))
    xi = np.array([[0.65, 0.45, 0.55]])
    o.set_particles(si, xi, np.array([[-1e5, 0, 0]]), S.loading.box_locate(spec, xi))
    o.set_conductor_charge(1e-15)
    o.push()
    assert o.get_particles(se)[2][0] == -1 and o.get_particles(si)[2][0] == -1
    led = o.ledger(12)
    assert led[6, 4] == 1 and led[6, 5] == 0 and np.isclose(led[6, 0], -QE * w)
    assert led[7, 5] == 1 and led[7, 4] == 0 and np.isclose(led[7, 1], QE * w)
    ke_e, ke_i = 0.5 * ME * 1e10 * w, 0.5 * MP * 1e10 * w
    np.testing.assert_allclose([led[6, 3], led[7, 3]], [ke_e, ke_i], rtol=1e-12)
    assert np.abs(np.delete(led, [6, 7], axis=0)).sum() == 0
    np.testing.assert_allclose(o.conductor_charge(), 1e-15, rtol=1e-12)  # +e w - e w


def test_escape_default_removes_without_ledger():
    # A face with no boundary condition: the particle leaves the domain with no
    # ledger entry and no conductor charge (PICCycle.cpp:1106-1133 default).
    m, o, spec = _boundary_oracle(10, COND)
    s = o.add_species(-QE, ME, 1.0)
    x = np.array([[0.5, 0.5, 0.97], [0.35, 0.5, 0.5]])
    o.set_particles(s, x, np.array([[0, 0, 1e5], [1e5, 0, 0]]), S.loading.box_locate(spec, x))
    o.push()
    _, _, cn = o.get_particles(s)
    assert (cn == -1).all()
    assert np.abs(o.ledger(12)).sum() == 0
    assert o.conductor_charge() == 0.0 and o.lost() == 0


def test_kinetic_energy_statistic_closed_form():
    m, o, spec = _boundary_oracle(4)
    rng = np.random.default_rng(5)
    se = o.add_species(-QE, ME, 2.5)
    si = o.add_species(QE, MP, 0.5)
    x = rng.uniform(0.1, 0.9, (7, 3))
    ve, vi = rng.normal(0, 1e5, (7, 3)), rng.normal(0, 1e3, (7, 3))
    o.set_particles(se, x, ve, S.loading.box_locate(spec, x))
    o.set_particles(si, x, vi, S.loading.box_locate(spec, x))
    want = 0.5 * ME * 2.5 * (ve ** 2).sum() + 0.5 * MP * 0.5 * (vi ** 2).sum()
    np.testing.assert_allclose(o.kinetic_energy(), want, rtol=1e-13)


def _charging_params(**kw):
    p = S.charging_params(photo_current_density=0.0, surface_conductance_s=0.0)
    for k, v in kw.items():
        if k == "face_yield":
            for i in range(6):
                p.face_yield[i] = v[i]
        else:
            setattr(p, k, v)
    return p


def test_charging_forward_conductor_relaxation_closed_form():
    # No particles: Q_{n+1} = Q_n + dt (J0 A_sunlit - G V_n), V = Q / C.
    m, o, spec = _boundary_oracle(8, COND)
    for g in range(6, 12):
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, EPS0, 1, 1e-13, 5000)
    o.add_species(-QE, ME, 1.0)
    o.set_particles(0, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32))
    q0, dt = -2e-11, 1e-6
    o.set_push(dt, 0, 1e-3, 1, 0)
    o.set_conductor_charge(q0)
    o.solve()
    cap = q0 / o.floating_potential()
    assert 1e-11 < cap < 3e-11  # a 0.2 m cube in a 1 m box: ~ 4 pi eps0 0.13 m
    lit = 0b000100  # face of group 8
    area = 0.2 * 0.2
    p = _charging_params(photo_current_density=3e-6, photo_reference_v=10.0, surface_conductance_s=2e-6,
                         sunlit_faces=lit)
    V = q0 / cap
    for _ in range(12):
        V = V + dt * (3e-6 * area - 2e-6 * V) / cap
    assert o.charging_forward(p, 12) == pytest.approx(V, rel=1e-9)


def test_charging_forward_secondary_charge_of_one_impact():
    # One electron hits the conductor's x-min face (group 6) at 1e7 m/s: the
    # conductor gains -e w plus the Sternglass secondary charge e w delta.
    m, o, spec = _boundary_oracle(10, COND)
    for g in range(6, 12):
        o.set_group(g, S.BC_CONDUCTOR)
    o.set_poisson(0, EPS0, 1, 1e-13, 5000)
    w = 5.0
    se = o.add_species(-QE, ME, w)
    x = np.array([[0.35, 0.5, 0.5]])
    o.set_particles(se, x, np.array([[1e7, 0, 0]]), S.loading.box_locate(spec, x))
    o.set_push(1e-8, 0, 1e-3, 1, 0)
    o.solve()
    o.set_efield(np.zeros((len(m["xyz"]), 3)))
    yields = [0.7, 1.0, 1.0, 1.0, 1.0, 1.0]
    p = _charging_params(see_delta_max=2.5, see_energy_max_ev=400.0, face_yield=yields)
    o.charging_forward(p, 1)
    E = 0.5 * ME * 1e14 / QE
    r = E / 400.0
    delta = 2.5 * 0.7 * r * np.exp(2 - 2 * np.sqrt(r))
    assert o.get_particles(se)[2][0] == -1
    np.testing.assert_allclose(o.conductor_charge(), -QE * w + QE * w * delta, rtol=1e-12)
