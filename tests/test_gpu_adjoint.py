"""Discrete adjoint of the floating potential through the C ABI.

* No particles: V_{n+1} = V_n + dt (J0 A_sunlit - G V_n) / C has a closed
  form; the adjoint's dV/dJ0 and dV/dG must match its derivative.
* With a plasma (FP64): every gradient component against central finite
  differences of the same forward (h = 1e-4 relative), to 1e-3, and the
  torch.autograd harness returns the same gradient.
* Mixed precision: the FP32 adjoint against the FP64 one (reported error).
"""
import numpy as np
import pytest
import torch

import scdat_amd as S
from scdat_amd.adjoint import FloatingPotential, central_difference
from helpers import COND, QE, plasma_setup

pytestmark = pytest.mark.gpu


def _ctx(st, precision, q_init=0.0, particles=True):
    m = st["mesh"]
    c = S.Context(precision=precision, dt=st["dt"], pcg_tol=1e-12, sort_interval=0, adaptive_substep=0)
    c.set_mesh(m["xyz"], m["tets"], m["kind"], group=m["group"])
    for g in range(6, 12):
        c.set_boundary(g, S.BC_CONDUCTOR)
    n = len(st["e"][0]) if particles else 0
    c.add_species(-QE, S.ELECTRON_MASS, st["w"], max(1, n))
    c.add_species(QE, S.PROTON_MASS, st["w"], max(1, n))
    if particles:
        c.set_particles(0, *st["e"])
        c.set_particles(1, *st["i"])
    c.set_conductor_charge(q_init)
    c.initialize()
    return c


def _sunlit_area(m, faces):
    """Area of the conductor faces of groups 6 + k, k in faces."""
    a = 0.0
    for t, k in zip(*np.nonzero(m["group"] >= 6)):
        if m["group"][t, k] - 6 in faces:
            p = m["xyz"][[m["tets"][t, q] for q in range(4) if q != k]]
            a += 0.5 * np.linalg.norm(np.cross(p[1] - p[0], p[2] - p[0]))
    return a


def test_adjoint_matches_closed_form_without_particles():
    st = plasma_setup(n=8, npart=10, seed=3)
    q0 = -2e-11
    c = _ctx(st, 1, q_init=q0, particles=False)
    v0 = c.stats()["floating_potential"]
    cap = q0 / v0
    steps, dt = 25, st["dt"]
    p = S.charging_params(photo_current_density=3e-4, photo_reference_v=50.0, surface_conductance_s=4e-5,
                          sunlit_faces=0b000010)
    area = _sunlit_area(st["mesh"], {1})
    # forward-mode derivative of the recurrence (V < V_ref throughout: J = J0)
    V, dJ, dG = v0, 0.0, 0.0
    for _ in range(steps):
        dJ, dG = dJ + dt * (area - p.surface_conductance_s * dJ) / cap, dG + dt * (-V - p.surface_conductance_s * dG) / cap
        V = V + dt * (p.photo_current_density * area - p.surface_conductance_s * V) / cap
    vf, g = c.charging_adjoint(p, steps, checkpoint_interval=6)
    print("V", vf, V, "grad", g[:3], dJ, dG)
    assert vf == pytest.approx(V, rel=1e-10)
    assert c.charging_forward(p, steps) == pytest.approx(V, rel=1e-10)
    assert g[S.GRAD_PHOTO_J] == pytest.approx(dJ, rel=1e-9)
    assert g[S.GRAD_CONDUCTANCE] == pytest.approx(dG, rel=1e-9)
    assert g[S.GRAD_DELTA_MAX] == 0.0 and np.all(g[S.GRAD_FACE_YIELD:] == 0.0)


CASE = dict(n=8, npart=4000, seed=29, Te_eV=150.0, density=2e9)
THETA_KW = dict(see_delta_max=2.0, see_energy_max_ev=400.0, face_yield=[1.0, 0.8, 1.2, 0.9, 1.1, 1.05],
                photo_current_density=2e-4, photo_temperature_ev=2.0, photo_reference_v=0.0,
                sunlit_faces=0b000010, surface_conductance_s=1e-5)


def _case():
    st = plasma_setup(n=CASE["n"], npart=CASE["npart"], seed=CASE["seed"], Te_eV=CASE["Te_eV"],
                      density=CASE["density"])
    vth = np.sqrt(CASE["Te_eV"] * QE / S.ELECTRON_MASS)
    st["dt"] = min(st["dt"], 0.2 * (1.0 / CASE["n"]) / vth)
    return st


def test_adjoint_gradient_matches_central_differences_fp64():
    st = _case()
    c = _ctx(st, 1)
    p = S.charging_params(**THETA_KW)
    steps = 30
    vf, g = c.charging_adjoint(p, steps, checkpoint_interval=7)
    v_fwd = c.charging_forward(p, steps)
    assert vf == pytest.approx(v_fwd, rel=1e-9, abs=1e-12)
    theta0 = p.vector()
    fd = central_difference(c, theta0, steps, rel=1e-4, base=p)
    # compare sensitivities theta_i dV/dtheta_i (volts): the components have
    # different units
    sa, sf = theta0 * g, theta0 * fd
    scale = np.abs(sf).max()
    print("V", vf, "\nadjoint", g, "\nfd     ", fd, "\nrel", np.abs(sa - sf) / np.maximum(np.abs(sf), 1e-30))
    for i in range(S.N_GRAD):
        assert abs(sa[i] - sf[i]) <= 1e-3 * max(abs(sf[i]), 1e-3 * scale), (i, g[i], fd[i])
    # the collected charge of this case does depend on the yields
    assert abs(sa[S.GRAD_DELTA_MAX]) > 0.05 * abs(vf)
    scale = np.abs(g).max()
    # torch.autograd harness
    theta = torch.tensor(p.vector(), dtype=torch.float64, requires_grad=True)
    V = FloatingPotential.apply(theta, c, steps, 7, p)
    V.backward()
    assert V.item() == pytest.approx(vf, rel=1e-9)
    np.testing.assert_allclose(theta.grad.numpy(), g, rtol=1e-9, atol=1e-12 * scale)
    # the context is left as it was: a second adjoint gives the same result
    vf2, g2 = c.charging_adjoint(p, steps, checkpoint_interval=10)
    assert vf2 == pytest.approx(vf, rel=1e-9)
    np.testing.assert_allclose(g2, g, rtol=1e-6, atol=1e-9 * scale)


def test_mixed_precision_adjoint_error_against_fp64():
    st = _case()
    p = S.charging_params(**THETA_KW)
    steps = 30
    v64, g64 = _ctx(st, 1).charging_adjoint(p, steps, checkpoint_interval=10)
    v32, g32 = _ctx(st, 0).charging_adjoint(p, steps, checkpoint_interval=10)
    theta = p.vector()
    err = np.abs(theta * (g32 - g64)).max() / np.abs(theta * g64).max()
    print("V fp64", v64, "mixed", v32, "sensitivity error (max abs / max |theta g|)", err)
    assert abs(v32 - v64) <= 1e-4 * abs(v64)
    assert err < 1e-2


def test_forward_matches_oracle_fp64():
    """The differentiable forward on the device against the oracle's FP64
    restatement (anchor-collapse floating conductor, CPU walk and deposit)."""
    from helpers import make_oracle
    st = _case()
    p = S.charging_params(**THETA_KW)
    steps = 30
    v_dev = _ctx(st, 1).charging_forward(p, steps)
    o = make_oracle(st, tol=1e-13, adaptive=0, max_sub=1)
    o.deposit()
    o.solve()
    o.efield()
    v_orc = o.charging_forward(p, steps)
    print("V device", v_dev, "oracle", v_orc)
    assert v_dev == pytest.approx(v_orc, rel=1e-9)
