"""Monte Carlo engine (mirrors src/mc_dagprop/monte_carlo/__init__.py:3-17).

Every class comes from the compiled module ``_core``, which drives the gfx950
kernels of ``libmcdp.so``.  There is no CPU fallback: without a HIP device,
``MonteCarloPropagator.run`` / ``run_many`` raise ``RuntimeError``.
"""

from __future__ import annotations

from ._core import Activity, DagContext, Event, EventTimestamp, GenericDelayGenerator, MonteCarloPropagator, SimResult

Simulator = MonteCarloPropagator

__all__ = [
    "GenericDelayGenerator",
    "DagContext",
    "SimResult",
    "Event",
    "Activity",
    "MonteCarloPropagator",
    "Simulator",
    "EventTimestamp",
]
