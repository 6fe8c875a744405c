"""Surface circuit (host FP64, scdat_circuit_*) against the compiled
reference's SurfaceCircuitCoupling::advanceImplicit golden vectors, the
reference's own Coupling tests (Tools/Coupling/test/Coupling_test.cpp:69-161)
and the closed form of backward-Euler RC relaxation."""
import os

import numpy as np
import pytest

import scdat_amd as S

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _build(nodes, branches):
    c = S.Circuit()
    for v, cap, g, fixed, fv in nodes:
        c.addNode(v, cap, g, bool(fixed), fv)
    for f, t, g, b in branches:
        c.addBranch(int(f), int(t), g, b)
    return c


def test_matches_reference_advance_implicit():
    g = np.load(os.path.join(GOLD, "circuit_cases.npz"))
    for k in range(int(g["n_cases"])):
        q = {key: g[f"c{k}_{key}"] for key in ("nodes", "branches", "entries", "current", "didv", "rhs", "diag",
                                                "dt", "max_delta", "v", "i", "b")}
        c = _build(q["nodes"], q["branches"])
        v, i, b, _ = c.advanceImplicit(float(q["dt"]), q["current"], q["didv"], q["rhs"], q["diag"],
                                       [tuple(e) for e in q["entries"]], float(q["max_delta"]))
        np.testing.assert_allclose(v, q["v"], rtol=1e-11, atol=1e-13)
        np.testing.assert_allclose(i, q["i"], rtol=1e-11, atol=1e-22)
        np.testing.assert_allclose(b, q["b"], rtol=1e-9, atol=1e-21)


def test_two_node_patch_moves_toward_injected_potential():
    # Coupling_test.cpp:69-101
    c = S.Circuit()
    body = c.addNode(0.0, fixed=True, fixed_value_v=0.0)
    patch = c.addNode(0.0, capacitance_f=1e-9)
    c.addBranch(patch, body, 1e-9)
    v, i, b, _ = c.advanceImplicit(1.0, [0.0, 1e-9], [0.0, -1e-11], max_delta_potential_v=50.0)
    assert v[patch] > 0.0 and b[0] < i[patch]


def test_potential_step_limit():
    # Coupling_test.cpp:103-128: a 1 uA current into 1 pF would move 1e6 V
    c = S.Circuit()
    body = c.addNode(fixed=True)
    patch = c.addNode(capacitance_f=1e-12)
    c.addBranch(patch, body, 0.0)
    v, _, _, md = c.advanceImplicit(1.0, [0.0, 1e-6], [0.0, 0.0], max_delta_potential_v=5.0)
    assert abs(v[patch] - 5.0) < 1e-9 and md == 5.0


def test_additional_matrix_entries():
    # Coupling_test.cpp:130-161: (1 + 1) V_l - V_r = 1, (1 + 1) V_r - V_l = 0
    c = S.Circuit()
    left = c.addNode(1.0, capacitance_f=1.0)
    right = c.addNode(0.0, capacitance_f=1.0)
    v, _, _, _ = c.advanceImplicit(1.0, [0.0, 0.0], [0.0, 0.0], [0.0, 0.0], [1.0, 1.0],
                                   entries=[(left, right, -1.0), (right, left, -1.0)])
    assert abs(v[left] - 2 / 3) < 1e-12 and abs(v[right] - 1 / 3) < 1e-12


def test_backward_euler_rc_relaxation_closed_form():
    # C dV/dt = I - G V by backward Euler: V_{n+1} = (C/dt V_n + I) / (C/dt + G)
    C_, G_, I_, dt = 2e-9, 5e-10, 3e-9, 0.7
    c = S.Circuit()
    gnd = c.addNode(fixed=True)
    node = c.addNode(0.0, capacitance_f=C_)
    c.addBranch(node, gnd, G_)
    v = 0.0
    for _ in range(40):
        out, _, _, _ = c.advanceImplicit(dt, [0.0, I_], [0.0, 0.0])
        v = (C_ / dt * v + I_) / (C_ / dt + G_)
        assert abs(out[node] - v) <= 1e-12 * abs(v)
    # towards the steady state I/G
    assert abs(v - I_ / G_) < 0.02 * I_ / G_


def test_singular_circuit_is_reported():
    c = S.Circuit()
    c.addNode(0.0)  # free node, no capacitance, no conductance
    with pytest.raises(S.ScdatError):
        c.advanceImplicit(1.0, [1e-9], [0.0])
