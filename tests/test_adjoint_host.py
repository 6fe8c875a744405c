"""Host side of the charging adjoint (no GPU): the parameter vector layout,
the torch.autograd harness and the central-difference helper, driven by an
analytic stand-in for the device context."""
import numpy as np
import torch

import scdat_amd as S
from scdat_amd.adjoint import FloatingPotential, central_difference


class Analytic:
    """V(theta) = sum a_i theta_i + b theta_0 theta_1, with its gradient."""

    a = np.array([0.3, 2e3, -5e4, 0.01, 0.02, 0.03, 0.04, 0.05, 0.06])
    b = 7.0

    def _v(self, p):
        t = p.vector()
        return float(self.a @ t + self.b * t[0] * t[1])

    def __init__(self):
        self.modes = []

    def charging_forward(self, p, steps, events=S.EVENTS_FREE):
        self.modes.append(events)
        return self._v(p)

    def charging_adjoint(self, p, steps, interval):
        t = p.vector()
        g = self.a.copy()
        g[0] += self.b * t[1]
        g[1] += self.b * t[0]
        return self._v(p), g


def test_params_vector_layout_matches_gradient_indices():
    p = S.charging_params(see_delta_max=1.5, photo_current_density=3e-5, surface_conductance_s=2e-9,
                          face_yield=[1, 2, 3, 4, 5, 6])
    v = p.vector()
    assert v[S.GRAD_DELTA_MAX] == 1.5 and v[S.GRAD_PHOTO_J] == 3e-5 and v[S.GRAD_CONDUCTANCE] == 2e-9
    np.testing.assert_array_equal(v[S.GRAD_FACE_YIELD:], [1, 2, 3, 4, 5, 6])
    q = S.charging_params().set_vector(v)
    np.testing.assert_array_equal(q.vector(), v)
    d = S.charging_params()
    assert d.see_delta_max == 2.0 and d.see_energy_max_ev == 400.0 and list(d.face_yield) == [1.0] * 6


def test_autograd_harness_and_central_difference():
    ctx = Analytic()
    p = S.charging_params(see_delta_max=1.5, photo_current_density=3e-5, surface_conductance_s=2e-9)
    theta = torch.tensor(p.vector(), dtype=torch.float64, requires_grad=True)
    V = FloatingPotential.apply(theta, ctx, 10, 5, p)
    (2.0 * V).backward()
    _, g = ctx.charging_adjoint(p, 10, 5)
    np.testing.assert_allclose(theta.grad.numpy(), 2.0 * g, rtol=1e-15)
    ctx.modes.clear()
    fd = central_difference(ctx, p.vector(), 10, rel=1e-4, base=p)
    np.testing.assert_allclose(fd, g, rtol=1e-7, atol=1e-9)
    # the events of the run at theta are recorded once, then every perturbed
    # run replays them
    assert ctx.modes[0] == S.EVENTS_RECORD and set(ctx.modes[1:]) == {S.EVENTS_REPLAY}
    assert len(ctx.modes) == 1 + 2 * len(g)
    ctx.modes.clear()
    central_difference(ctx, p.vector(), 10, rel=1e-4, base=p, frozen=False)
    assert set(ctx.modes) == {S.EVENTS_FREE}
