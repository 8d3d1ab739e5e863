"""Expected weighted lateness as a differentiable function of the DAG's times, for torch.

``expected_lateness(sim, minimal_durations, earliest, seeds, weights)`` returns the
sample mean over ``seeds`` of

    J = sum_e weights[e] * max(0, realized[e] - earliest[e])

as a 0-d float64 tensor, with ``minimal_durations`` (one value per link, indexed by
``Activity.idx``) and ``earliest`` (one per event) as differentiable inputs.  The forward
pass runs ``run_many`` on the GPU and the backward pass returns the pathwise gradient
that the reverse kernel computed alongside it (``mcdp_gradient_host``).  The seeds fix
the random draws: two calls with the same seeds use common random numbers.

``sim`` is a ``Simulator`` / ``MonteCarloPropagator`` of this package or a
``capi.Propagator``; the call replaces its times with the given ones.
"""

from __future__ import annotations

import numpy as np
import torch


class ExpectedLateness(torch.autograd.Function):
    """torch.autograd.Function over the forward walk and the reverse kernel."""

    @staticmethod
    def forward(ctx, minimal_durations, earliest, sim, seeds, weights, chunk):
        d = minimal_durations.detach().to("cpu", torch.float64).numpy()
        e = earliest.detach().to("cpu", torch.float64).numpy()
        w = np.ascontiguousarray(torch.as_tensor(weights, dtype=torch.float64).cpu().numpy())
        s = [int(x) for x in np.asarray(seeds).ravel()]
        sim.set_times(d, e)
        obj, gd, ge = sim.gradient(s, w, chunk)
        ctx.save_for_backward(torch.as_tensor(np.asarray(gd), dtype=minimal_durations.dtype,
                                              device=minimal_durations.device),
                              torch.as_tensor(np.asarray(ge), dtype=earliest.dtype, device=earliest.device))
        return torch.tensor(float(obj), dtype=torch.float64, device=minimal_durations.device)

    @staticmethod
    def backward(ctx, grad_out):
        gd, ge = ctx.saved_tensors
        g = grad_out.to(torch.float64)
        return (g * gd).to(gd.dtype), (g * ge).to(ge.dtype), None, None, None, None


def expected_lateness(sim, minimal_durations: torch.Tensor, earliest: torch.Tensor, seeds, weights,
                      chunk: int = 0) -> torch.Tensor:
    """Mean weighted lateness over ``seeds``, differentiable in ``minimal_durations`` and ``earliest``."""
    return ExpectedLateness.apply(minimal_durations, earliest, sim, seeds, weights, chunk)
