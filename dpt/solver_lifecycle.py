"""Progress events and termination categories of the recovery drivers.

Same contract as python/dpt/solver_lifecycle.py of the reference: an
observer receives immutable scalar snapshots of committed solver states
(``initial`` or ``resumed``, then ``accepted`` / ``rejected`` per controller
step, one ``terminal``); its return value never alters a solve, and an
exception it raises propagates unchanged.  Every termination reason maps to
one common category.  When a solve ends by an exception, the exception
carries a terminal snapshot with reason ``"exception"`` in
``__dict__["solver_lifecycle"]``.
"""

from __future__ import annotations

from collections.abc import Callable, Mapping
from dataclasses import dataclass
from typing import Literal

from .contracts import ContractError

TerminationCategory = Literal["stationary", "stagnation", "budget", "cancelled", "failed", "uncertainty"]
LifecycleStatus = Literal["initial", "resumed", "accepted", "rejected", "terminal"]

# reason -> category (the reasons of the reference's drivers)
_CATEGORY: dict[str, TerminationCategory] = {}
for _category, _reasons in (
    ("stationary", ("gradient_tolerance", "projected_gradient_tolerance")),
    ("stagnation", ("step_tolerance", "loss_stagnation", "loss_resolution", "radius_limit")),
    ("budget", ("iteration_budget", "evaluation_budget", "history_budget")),
    ("cancelled", ("cancelled",)),
    ("failed", ("line_search_failed", "damping_limit", "evaluation_failed", "exception")),
    ("uncertainty", ("gradient_band", "sampling_unresolved")),
):
    for _reason in _reasons:
        _CATEGORY[_reason] = _category  # type: ignore[assignment]


@dataclass(frozen=True, slots=True)
class SolverTermination:
    """A solver's own termination reason and its common category."""

    reason: str
    category: TerminationCategory


def termination_reason(reason: str) -> SolverTermination:
    category = _CATEGORY.get(reason)
    if category is None:
        raise ContractError(f"unknown solver termination reason: {reason!r}")
    return SolverTermination(reason, category)


def solver_termination(result: object) -> SolverTermination:
    """The termination of a result object (``.reason``) or a report mapping
    (``["termination"]``)."""
    if isinstance(result, Mapping):
        reason = result.get("termination")
    else:
        reason = getattr(result, "reason", None)
    if not isinstance(reason, str):
        raise ContractError("solver result must provide a string termination reason")
    return termination_reason(reason)


@dataclass(frozen=True, slots=True)
class SolverEvent:
    """Scalar progress snapshot: ``objective`` is the last accepted loss (None
    for drivers that only estimate objective changes); ``detail`` names a
    rejected proposal or keeps a terminal diagnostic."""

    solver: str
    status: LifecycleStatus
    iteration: int
    accepted_steps: int
    objective: float | None
    reason: str | None = None
    detail: str | None = None
    termination: SolverTermination | None = None


LifecycleObserver = Callable[[SolverEvent], None]


class LifecycleRecorder:
    """Per-solve event source; a context manager that stamps a terminal
    snapshot on an exception leaving the solve."""

    def __init__(self, solver: str, observer: LifecycleObserver | None) -> None:
        self.solver = solver
        self.observer = observer
        self.iteration = 0
        self.accepted_steps = 0
        self.objective: float | None = None

    def _snapshot(self, status: LifecycleStatus, reason: str | None = None, detail: str | None = None) -> SolverEvent:
        return SolverEvent(self.solver, status, self.iteration, self.accepted_steps, self.objective, reason, detail,
                           None if reason is None else termination_reason(reason))

    def _send(self, event: SolverEvent) -> None:
        if self.observer is not None:
            self.observer(event)

    def start(self, *, objective: float | None, iteration: int = 0, accepted_steps: int = 0,
              resumed: bool = False) -> None:
        self.objective, self.iteration, self.accepted_steps = objective, iteration, accepted_steps
        self._send(self._snapshot("resumed" if resumed else "initial"))

    def accepted(self, iteration: int, objective: float | None) -> None:
        self.iteration = iteration
        self.accepted_steps += 1
        self.objective = objective
        self._send(self._snapshot("accepted"))

    def rejected(self, iteration: int, detail: str) -> None:
        self.iteration = iteration
        self._send(self._snapshot("rejected", detail=detail))

    def finish(self, result: object, *, detail: str | None = None) -> None:
        self._send(self._snapshot("terminal", solver_termination(result).reason, detail))

    def __enter__(self) -> "LifecycleRecorder":
        return self

    def __exit__(self, kind, error, traceback) -> bool:
        if error is not None and "solver_lifecycle" not in error.__dict__:
            error.__dict__["solver_lifecycle"] = self._snapshot("terminal", "exception")
        return False
