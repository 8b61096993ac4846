"""Sound speed -> misfit as a differentiable torch function.

``misfit_from_velocity(c, plan, wavelet, dobs)`` returns the waveform misfit J
of the model ``c`` (m/s) as a scalar tensor whose backward is the exact
discrete-adjoint gradient computed by the HIP solver: the forward pass runs
misfit_and_gradient once (forward sweep, checkpoints, re-computation and
adjoint sweep all on the device) and keeps g = dJ/dm; backward applies the
chain rule through m = 1/c^2 (simulation/ringfwi/phantom.py:13-14):

    dJ/dc = g * dm/dc = -2 g / c^3.

The gradient is that of the solver's discrete scheme, so it matches a finite
difference of the same solver's misfit (not the continuous PDE).  Every
misfit type of fwi.py:134-243 is accepted: L2 and GCN run fused in the
library; the others form their adjoint source from the device traces
(openusct_amd.misfits) and run the library's adjoint sweep on it.

``traces(m, plan, wavelet)`` is the lower-level piece: the full-matrix traces
of the squared-slowness model m as a differentiable tensor whose backward
applies the exact discrete adjoint (Plan.gradient_from_source), so any misfit
written in torch on the traces gets its m-gradient from the device solver.
"""

from __future__ import annotations

import torch

from . import misfits as trace_misfits
from .plan import Plan


class Traces(torch.autograd.Function):
    """m -> traces (n_tx, nt, n_rec); backward = transpose of the discrete
    forward's Jacobian, applied by the device adjoint sweep."""

    @staticmethod
    def forward(ctx, m, plan: Plan, wavelet, sponge=None):
        m64 = m.detach().to(torch.float64)
        ctx.save_for_backward(m64)
        ctx.plan, ctx.wavelet, ctx.sponge, ctx.in_dtype = plan, wavelet, sponge, m.dtype
        d = plan.forward_fmc(m64, wavelet, sponge)
        return d.to(m.dtype) if m.dtype != torch.float64 else d

    @staticmethod
    def backward(ctx, grad_d):
        (m64,) = ctx.saved_tensors
        g = ctx.plan.gradient_from_source(m64, ctx.wavelet, grad_d.to(torch.float64).contiguous(),
                                          ctx.sponge)
        return g.to(ctx.in_dtype), None, None, None


def traces(m, plan, wavelet, sponge=None):
    """Differentiable full-matrix traces of the squared-slowness model ``m``."""
    return Traces.apply(m, plan, wavelet, sponge)


def _misfit_and_gradient(plan, m, wavelet, dobs, sponge, misfit_type):
    if misfit_type in trace_misfits.DEVICE_MISFITS:
        d = plan.forward_fmc(m, wavelet, sponge)
        J, adj = trace_misfits.misfit_and_source(d, dobs.to(d), misfit_type)
        return J.reshape(1), plan.gradient_from_source(m, wavelet, adj, sponge)
    return plan.misfit_and_gradient(m, wavelet, dobs, sponge, misfit_type)


def _misfit(plan, m, wavelet, dobs, misfit_type, sponge=None):
    if misfit_type in trace_misfits.DEVICE_MISFITS:
        d = plan.forward_fmc(m, wavelet, sponge)
        J, _ = trace_misfits.misfit_and_source(d, dobs.to(d), misfit_type)
        return J.reshape(1)
    return plan.misfit(m, wavelet, dobs, sponge, misfit_type=misfit_type)


class SoundSpeedMisfit(torch.autograd.Function):
    """J(c) with the adjoint-state gradient; non-tensor arguments get no gradient."""

    @staticmethod
    def forward(ctx, c, plan: Plan, wavelet, dobs, misfit_type="l2", sponge=None):
        c64 = c.detach().to(torch.float64)
        m = 1.0 / (c64 * c64)
        J, g = _misfit_and_gradient(plan, m, wavelet, dobs, sponge, misfit_type)
        ctx.save_for_backward(c64, g)
        ctx.in_dtype = c.dtype
        return J[0].to(c.dtype) if c.dtype != torch.float64 else J[0]

    @staticmethod
    def backward(ctx, grad_out):
        c64, g = ctx.saved_tensors
        dc = (-2.0 * g / (c64 * c64 * c64)) * grad_out.to(torch.float64)
        return dc.to(ctx.in_dtype), None, None, None, None, None


def misfit_from_velocity(c, plan, wavelet, dobs, misfit_type="l2", sponge=None):
    """Differentiable misfit of the velocity model ``c`` (a HIP tensor)."""
    return SoundSpeedMisfit.apply(c, plan, wavelet, dobs, misfit_type, sponge)


def gradient_descent(c0, plan, wavelet, dobs, steps=5, step_frac=0.02, c_bounds=None,
                     mask=None, misfit_type="l2", n_linesearch=6):
    """Steepest descent on c through :class:`SoundSpeedMisfit`, with the step rule
    and halving line search of simulation/ringfwi/fwi.py:637-660 applied to c.

    Each step: J = misfit(c); J.backward(); direction -grad (times ``mask``);
    step = step_frac * mean(c) / max|grad|; try c - step * grad (clipped to
    ``c_bounds``) and halve the step until the misfit decreases, at most
    ``n_linesearch`` tries (the trial misfits are forward-only evaluations); stop
    when no trial decreases it.  Returns (c, history of accepted misfits).
    """
    c = c0.detach().clone().to(torch.float64).requires_grad_(True)
    history = []
    J = misfit_from_velocity(c, plan, wavelet, dobs, misfit_type)
    J.backward()
    for _ in range(steps):
        history.append(float(J.item()))
        with torch.no_grad():
            d = c.grad if mask is None else c.grad * mask
            dmax = float(d.abs().max())
            if dmax == 0.0:
                break
            step = step_frac * float(c.mean()) / dmax
            accepted = False
            for _ in range(n_linesearch):
                c_try = c - step * d
                if c_bounds is not None:
                    c_try.clamp_(c_bounds[0], c_bounds[1])
                J_try = _misfit(plan, 1.0 / (c_try * c_try), wavelet, dobs, misfit_type)
                if float(J_try.item()) < float(J.item()):
                    accepted = True
                    break
                step *= 0.5
        if not accepted:
            break
        c = c_try.detach().requires_grad_(True)
        J = misfit_from_velocity(c, plan, wavelet, dobs, misfit_type)
        J.backward()
    history.append(float(J.item()))
    return c.detach(), history
