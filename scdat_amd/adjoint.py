"""torch.autograd harness of the discrete charging adjoint.

`FloatingPotential.apply(theta, ctx, steps)` returns the floating conductor's
potential after `steps` steps as a differentiable function of
theta = (delta_max, J0, G, face_yield[0..5]); its backward is the device's
discrete adjoint (checkpointed), not autograd's tape.  `central_difference`
gives the finite-difference gradient the adjoint is checked against.
"""
import numpy as np
import torch

from . import EVENTS_FREE, EVENTS_RECORD, EVENTS_REPLAY, ChargingParams, charging_params


def _params(theta, base):
    p = ChargingParams.from_buffer_copy(base) if base is not None else charging_params()
    return p.set_vector(np.asarray(theta, dtype=np.float64))


class FloatingPotential(torch.autograd.Function):
    @staticmethod
    def forward(fctx, theta, ctx, steps, checkpoint_interval=20, base=None):
        p = _params(theta.detach().cpu().double().numpy(), base)
        v, g = ctx.charging_adjoint(p, steps, checkpoint_interval)
        fctx.save_for_backward(torch.as_tensor(g, dtype=theta.dtype))
        return theta.new_tensor(v)

    @staticmethod
    def backward(fctx, gout):
        (g,) = fctx.saved_tensors
        return gout * g, None, None, None, None


def central_difference(ctx, theta, steps, rel=1e-4, base=None, indices=None, frozen=True):
    """dV/dtheta_i by (V(theta + h e_i) - V(theta - h e_i)) / 2h, h = rel |theta_i|
    (rel when theta_i = 0), from the same initial state and streams.

    frozen: both perturbed runs replay the absorption events of the run at
    theta (EVENTS_RECORD, then EVENTS_REPLAY), so the quotient differences
    the map the adjoint differentiates; otherwise each run has its own
    events, and a macro-particle absorbed in one run and not in the other
    moves V by q w / C whatever h is."""
    theta = np.asarray(theta, dtype=np.float64)
    out = np.full(len(theta), np.nan)
    mode = EVENTS_FREE
    if frozen:
        ctx.charging_forward(_params(theta, base), steps, EVENTS_RECORD)
        mode = EVENTS_REPLAY
    for i in (range(len(theta)) if indices is None else indices):
        h = rel * abs(theta[i]) if theta[i] != 0.0 else rel
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        vp = ctx.charging_forward(_params(tp, base), steps, mode)
        vm = ctx.charging_forward(_params(tm, base), steps, mode)
        out[i] = (vp - vm) / (2.0 * h)
    return out
