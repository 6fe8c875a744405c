"""Conductor patches (six faces + body) held at the potentials of a surface
circuit, through the C ABI: the superposition field against the oracle's
direct solve with the patches as Dirichlet nodes, and the circuit's
backward-Euler update from the device's collected charge against the same
circuit advanced on the host."""
import numpy as np
import pytest

import oracle as O
import scdat_amd as S
from helpers import COND, EPS0, QE, make_ctx, plasma_setup, rel_err

pytestmark = pytest.mark.gpu


def _patch_map(m):
    """Default patch of every node: conductor faces (groups 6..11) -> 0..5,
    nodes on two or more faces -> 6 (body), -1 off the conductor."""
    bits = np.zeros(len(m["xyz"]), np.int64)
    for t, k in zip(*np.nonzero((m["group"] >= 6) & (m["group"] <= 11))):
        for q in range(4):
            if q != k:
                bits[m["tets"][t, q]] |= 1 << (m["group"][t, k] - 6)
    npatch = -np.ones(len(bits), np.int32)
    cond = m["kind"] == 2
    one = cond & (bits > 0) & ((bits & (bits - 1)) == 0)
    npatch[one] = np.log2(bits[one]).astype(np.int32)
    npatch[cond & ~one] = 6
    return npatch


def _circuit(v0, C=1e-12, G=1e-9, Cb=5e-12):
    """ground (0, fixed) -- body (7, Cb) -- six faces (1..6, C each, G to body)."""
    c = S.Circuit()
    c.addNode(0.0, fixed=True)
    faces = [c.addNode(v0[k], capacitance_f=C) for k in range(6)]
    body = c.addNode(v0[6], capacitance_f=Cb)
    for f in faces:
        c.addBranch(f, body, G)
    c.addBranch(body, 0, 1e-12)
    return c, faces + [body]


def test_patch_superposition_matches_direct_dirichlet_solve():
    st = plasma_setup(n=8, npart=3000, seed=41)
    m = st["mesh"]
    v0 = np.array([1.5, -0.5, 0.25, 2.0, -1.0, 0.75, 0.1])
    circ, nodes = _circuit(v0)
    for k, n in enumerate(nodes):  # hold every patch fixed for this check
        circ.setNode(n, v0[k], fixed=True, fixed_value_v=v0[k])
    c = make_ctx(st, precision=1, tol=1e-12, sort_interval=0)
    c.set_conductor_circuit(circ, nodes)
    c.initialize()
    rho_c, phi_c, E_c = c.fields()
    npatch = _patch_map(m)
    for k in range(7):
        sel = npatch == k
        assert sel.sum() > 0
        np.testing.assert_allclose(phi_c[sel], v0[k], rtol=1e-12, atol=1e-12)
        basis = c.patch_basis(k)
        np.testing.assert_allclose(basis[sel], 1.0, rtol=0, atol=1e-12)
    # the oracle solves with the patch nodes as Dirichlet nodes
    kind = m["kind"].copy()
    dval = np.zeros(len(kind))
    kind[npatch >= 0] = 1
    dval[npatch >= 0] = v0[npatch[npatch >= 0]]
    o = O.Oracle(m["xyz"], m["tets"], kind, dirichlet=dval, group=m["group"])
    o.set_poisson(0, EPS0, 1, 1e-13, 50000)
    o.add_species(-QE, S.ELECTRON_MASS, st["w"])
    o.add_species(QE, S.PROTON_MASS, st["w"])
    o.set_particles(0, *st["e"])
    o.set_particles(1, *st["i"])
    o.deposit()
    rho_o, _, _ = o.fields()
    # the device deposits in 2^-24 fixed point: the charge agrees to that
    # bound, and the solve is compared on the device's own charge
    assert rel_err(rho_c, rho_o) < 1e-6
    o.set_rho(rho_c)
    o.solve()
    o.efield()
    _, phi_o, E_o = o.fields()
    print("phi rel", rel_err(phi_c, phi_o), "E rel", rel_err(E_c, E_o))
    assert rel_err(phi_c, phi_o) < 1e-9 and rel_err(E_c, E_o) < 1e-8


def test_circuit_driven_patches_follow_collected_charge():
    st = plasma_setup(n=10, npart=20000, seed=43, Te_eV=5.0)
    m = st["mesh"]
    v0 = np.zeros(7)
    circ, nodes = _circuit(v0)
    host, hnodes = _circuit(v0)  # the same circuit advanced here
    c = make_ctx(st, precision=1, tol=1e-12, sort_interval=0)
    c.set_conductor_circuit(circ, nodes)
    c.initialize()
    dt = st["dt"]
    prev = np.zeros(7)
    for _ in range(4):
        c.step(1)
        led = c.ledger(12)
        q = np.zeros(7)
        for k in range(6):
            e = led[6 + k]
            q[k] = e["absorbed_electron_charge_c"] + e["absorbed_ion_charge_c"] - e["emitted_electron_charge_c"]
        cur = np.zeros(8)
        for k in range(7):
            cur[hnodes[k]] = (q[k] - prev[k]) / dt
        prev = q
        v, _, _, _ = host.advanceImplicit(dt, cur, np.zeros(8))
        vd = c.patch_potentials()
        print("device", vd, "host", v[hnodes])
        np.testing.assert_allclose(vd, v[hnodes], rtol=1e-12, atol=1e-15)
    assert np.abs(prev).max() > 0  # charge was collected
    _, phi, _ = c.fields()
    npatch = _patch_map(m)
    for k in range(7):
        np.testing.assert_allclose(phi[npatch == k], vd[k], rtol=1e-12, atol=1e-15)
