"""Public interface of :mod:`mc_dagprop_amd`.

The Monte Carlo API of ``mc_dagprop`` (src/mc_dagprop/__init__.py:5-50, Monte
Carlo names only), computed on an MI355X by hand-written gfx950 HIP kernels.
``import mc_dagprop_amd as mc_dagprop`` is the switch for Monte Carlo users.
"""

from __future__ import annotations

try:
    from .monte_carlo import (
        Activity,
        DagContext,
        Event,
        EventTimestamp,
        GenericDelayGenerator,
        MonteCarloPropagator,
        SimResult,
        Simulator,
    )
except ImportError as exc:  # pragma: no cover - compiled module missing
    raise ImportError(
        "mc_dagprop_amd requires the compiled extension 'mc_dagprop_amd.monte_carlo._core' and libmcdp.so; "
        "build them with `python -c 'import __graft_entry__ as g; g.build()'` or `python mc_dagprop_amd/build.py`."
    ) from exc

__version__ = "0.1.0"

__all__ = [
    "GenericDelayGenerator",
    "DagContext",
    "SimResult",
    "Event",
    "Activity",
    "Simulator",
    "MonteCarloPropagator",
    "EventTimestamp",
    "__version__",
]
