"""Re-export of the core data classes (mirrors src/mc_dagprop/core.py:5-7)."""

from __future__ import annotations

from .monte_carlo import Activity, DagContext, Event, EventTimestamp

__all__ = ["EventTimestamp", "Event", "Activity", "DagContext"]
