"""Safeguarded limited-memory quasi-Newton search in a fixed parameter chart.

Same contract as python/dpt/registration.py:318-369 of the reference
(``recover_parameters`` and its records): the evaluator maps a point of a
fixed, dimensionless chart to an :class:`Evaluation` (scalar loss and its
gradient in that chart; it runs the device work, e.g. a projector forward and
pose VJP, and returns host scalars only); the driver keeps the small host
state — L-BFGS curvature pairs, a line search seeking strong Wolfe steps with
a strict-decrease Armijo fallback, termination tests, optional checkpoints
and lifecycle events.  Rejected trials (``TrialDomainError`` from the
evaluator, non-finite points) never replace the accepted point.

The names, policy fields and their defaults, stop reasons and error types
follow the reference so that callers and their tests carry over; the search
itself is written here (bisection in the zoom phase keeps rejected-domain
endpoints unambiguous).
"""

from __future__ import annotations

import math
from collections.abc import Callable
from dataclasses import dataclass, replace
from typing import Literal, Protocol

from .contracts import ContractError, NumericalError, TrialDomainError, finite_scalar, integer
from .solver_lifecycle import LifecycleObserver, LifecycleRecorder

Vector = tuple[float, ...]


@dataclass(frozen=True, slots=True)
class Evaluation:
    """Objective and its gradient with respect to the chart coordinates."""

    loss: float
    gradient: Vector


class Evaluator(Protocol):
    def __call__(self, parameters: Vector, /) -> Evaluation:
        """Evaluate; raise TrialDomainError only for a rejectable trial point."""
        ...


@dataclass(frozen=True, slots=True)
class RecoveryPolicy:
    """Budgets, tolerances and line-search constants of a chart search.

    ``absolute_loss_resolution`` (opt-in) is the caller's bound on meaningful
    loss changes: a search that fails only among trials within it stops as
    ``loss_resolution``; it never accepts a trial by that allowance.
    """

    max_iterations: int = 200
    max_evaluations: int = 2000
    memory: int = 10
    line_search_evaluations: int = 32
    gradient_tolerance: float = 1e-6
    relative_loss_tolerance: float = 1e-12
    step_tolerance: float = 1e-10
    armijo: float = 1e-4
    curvature: float = 0.9
    maximum_step: float = 16.0
    minimum_step: float = 1e-14
    absolute_loss_resolution: float = 0.0

    def __post_init__(self) -> None:
        for name in ("max_iterations", "max_evaluations", "memory", "line_search_evaluations"):
            integer(getattr(self, name), name, minimum=1)
        for name in ("gradient_tolerance", "relative_loss_tolerance", "step_tolerance", "absolute_loss_resolution",
                     "armijo", "curvature", "minimum_step", "maximum_step"):
            finite_scalar(getattr(self, name), name, minimum=0.0)
        if not 0.0 < self.armijo < self.curvature < 1.0:
            raise ContractError("strong-Wolfe constants must satisfy 0 < c1 < c2 < 1")
        if not 0.0 < self.minimum_step < self.maximum_step:
            raise ContractError("line-search step interval must be positive and nonempty")


StopReason = Literal["gradient_tolerance", "step_tolerance", "loss_stagnation", "line_search_failed",
                     "evaluation_budget", "iteration_budget", "cancelled", "evaluation_failed", "loss_resolution"]
SearchKind = Literal["initial", "wolfe", "armijo_bound", "armijo_best"]


@dataclass(frozen=True, slots=True)
class IterationRecord:
    iteration: int
    loss: float
    gradient_infinity_norm: float
    step: float
    evaluations: int
    curvature_pairs: int
    line_search: SearchKind = "initial"


@dataclass(frozen=True, slots=True)
class RejectedTrial:
    parameters: Vector
    message: str


@dataclass(frozen=True, slots=True)
class RecoveryCheckpoint:
    """Host state of one solve for one immutable objective and chart
    (``objective_id`` names both).  Resuming re-evaluates the accepted point
    and compares it before reusing the curvature; only the budgets of the
    policy may change."""

    objective_id: str
    policy: RecoveryPolicy
    parameters: Vector
    evaluation: Evaluation
    curvature_pairs: tuple[tuple[Vector, Vector, float], ...]
    history: tuple[IterationRecord, ...]
    rejected_trials: tuple[RejectedTrial, ...]
    evaluations: int
    resolution_limited_trials: int


@dataclass(frozen=True, slots=True)
class RecoveryResult:
    parameters: Vector
    evaluation: Evaluation
    reason: StopReason
    evaluations: int
    history: tuple[IterationRecord, ...]
    rejected_trials: tuple[RejectedTrial, ...] = ()
    resolution_limited_trials: int = 0
    checkpoint: RecoveryCheckpoint | None = None

    @property
    def stationary(self) -> bool:
        """Only the gradient test establishes stationarity; small steps or loss
        changes are stagnation, not proof of a recovered minimum."""
        return self.reason == "gradient_tolerance"


# ------------------------------------------------------------ vector helpers

def _inner(a: Vector, b: Vector) -> float:
    """Correctly rounded sum of products; NaN when it cannot be represented."""
    try:
        return math.fsum(x * y for x, y in zip(a, b, strict=True))
    except (OverflowError, ValueError):
        return math.nan


def _linf(a: Vector) -> float:
    return max(abs(x) for x in a)


def _axpy(alpha: float, x: Vector, y: Vector) -> Vector:
    return tuple(b + alpha * a for a, b in zip(x, y, strict=True))


def _verified(evaluation: Evaluation, dimension: int) -> Evaluation:
    if len(evaluation.gradient) != dimension:
        raise ContractError("evaluator gradient dimension differs from its parameter chart")
    if not (math.isfinite(evaluation.loss) and all(math.isfinite(g) for g in evaluation.gradient)):
        raise NumericalError("non-finite objective or parameter gradient")
    return evaluation


def _descent(gradient: Vector) -> Vector:
    """Steepest descent, scaled so that no coordinate moves more than 1."""
    scale = max(1.0, _linf(gradient))
    return tuple(-g / scale for g in gradient)


class _Curvature:
    """L-BFGS memory: pairs (s, y, 1 / s.y), oldest first."""

    def __init__(self, pairs, memory: int) -> None:
        self.pairs = list(pairs)
        self.memory = memory

    def direction(self, gradient: Vector) -> Vector:
        if not self.pairs:
            return _descent(gradient)
        s_last, y_last, _ = self.pairs[-1]
        yy = _inner(y_last, y_last)
        gamma = _inner(s_last, y_last) / yy if math.isfinite(yy) and yy > 0.0 else math.nan
        if not math.isfinite(gamma) or gamma <= 0.0:
            return _descent(gradient)
        q = gradient
        alphas = []
        for s, y, rho in reversed(self.pairs):
            a = rho * _inner(s, q)
            alphas.append(a)
            q = _axpy(-a, y, q)
        r = tuple(gamma * v for v in q)
        for (s, y, rho), a in zip(self.pairs, reversed(alphas)):
            r = _axpy(a - rho * _inner(y, r), s, r)
        return tuple(-v for v in r)

    def update(self, step: Vector, change: Vector) -> None:
        sy = _inner(step, change)
        floor = 1e-12 * math.hypot(*step) * math.hypot(*change)
        if math.isfinite(sy) and sy > max(0.0, floor) and math.isfinite(1.0 / sy):
            self.pairs.append((step, change, 1.0 / sy))
            del self.pairs[:-self.memory]

    def clear(self) -> None:
        self.pairs.clear()


class _OutOfEvaluations(Exception):
    """The evaluation budget ran out inside a line search."""


class _LineSearch:
    """One line search from x along d (a descent direction): grow the step
    (doubling up to maximum_step) while the trials keep decreasing, then
    bisect a bracket; accept the first strong-Wolfe point, else the best
    strict-decrease Armijo trial seen ("armijo_best"), or a bound step that
    satisfies Armijo ("armijo_bound")."""

    def __init__(self, evaluate, x: Vector, base: Evaluation, d: Vector, policy: RecoveryPolicy, hooks) -> None:
        self.evaluate, self.x, self.base, self.d, self.policy = evaluate, x, base, d, policy
        self.on_reject, self.on_resolution = hooks
        self.slope = _inner(base.gradient, d)
        self.trials = 0
        self.best: tuple[float, Evaluation] | None = None

    def _sufficient(self, alpha: float, value: Evaluation) -> bool:
        return value.loss <= self.base.loss + self.policy.armijo * alpha * self.slope

    def _curved(self, value: Evaluation) -> bool:
        return abs(_inner(value.gradient, self.d)) <= -self.policy.curvature * self.slope

    def _try(self, alpha: float) -> Evaluation | None:
        self.trials += 1
        point = _axpy(alpha, self.d, self.x)
        if not all(math.isfinite(v) for v in point):
            return None
        try:
            value = _verified(self.evaluate(point), len(self.x))
        except TrialDomainError as error:
            self.on_reject(RejectedTrial(point, str(error)))
            return None
        except _OutOfEvaluations:
            if self.best is None:
                raise
            self.trials = self.policy.line_search_evaluations  # finish with the best so far
            return None
        resolution = self.policy.absolute_loss_resolution
        if resolution > 0 and abs(value.loss - self.base.loss) <= resolution and \
                -self.policy.armijo * alpha * self.slope <= resolution:
            self.on_resolution()
        if value.loss < self.base.loss and self._sufficient(alpha, value) and \
                (self.best is None or value.loss < self.best[1].loss):
            self.best = (alpha, value)
        return value

    def _left(self) -> bool:
        return self.trials < self.policy.line_search_evaluations

    def _fallback(self):
        return None if self.best is None else (*self.best, "armijo_best")

    def _bisect(self, lo: float, hi: float, lo_value: Evaluation):
        # invariant: lo satisfies Armijo with the lowest loss of the bracket
        while self._left():
            mid = 0.5 * (lo + hi)
            if abs(hi - lo) < self.policy.minimum_step or mid in (lo, hi):
                break
            value = self._try(mid)
            if value is None or not self._sufficient(mid, value) or value.loss >= lo_value.loss:
                hi = mid
                continue
            if self._curved(value):
                return mid, value, "wolfe"
            if _inner(value.gradient, self.d) * (hi - lo) >= 0.0:
                hi = lo
            lo, lo_value = mid, value
        return self._fallback()

    def run(self, first_step: float):
        if not math.isfinite(self.slope) or self.slope >= 0.0:
            return None
        prev_alpha, prev_value = 0.0, self.base
        alpha = min(first_step, self.policy.maximum_step)
        while self._left():
            value = self._try(alpha)
            if value is None or not self._sufficient(alpha, value) or \
                    (prev_alpha > 0.0 and value.loss >= prev_value.loss):
                return self._bisect(prev_alpha, alpha, prev_value)
            if self._curved(value):
                return alpha, value, "wolfe"
            if _inner(value.gradient, self.d) >= 0.0:
                return self._bisect(alpha, prev_alpha, value)
            if alpha == self.policy.maximum_step:
                return (alpha, value, "armijo_bound") if value.loss < self.base.loss else self._fallback()
            prev_alpha, prev_value = alpha, value
            alpha = min(2.0 * alpha, self.policy.maximum_step)
        return self._fallback()


# ------------------------------------------------------------------ driver

def recover_parameters(
    evaluator: Evaluator,
    initial: Vector | None = None,
    *,
    policy: RecoveryPolicy | None = None,
    observe: Callable[[IterationRecord], None] | None = None,
    cancelled: Callable[[], bool] | None = None,
    objective_id: str | None = None,
    checkpoint: RecoveryCheckpoint | None = None,
    save_checkpoint: Callable[[RecoveryCheckpoint], None] | None = None,
    checkpoint_relative_tolerance: float = 0.0,
    observe_lifecycle: LifecycleObserver | None = None,
) -> RecoveryResult:
    """Minimise the evaluator's loss from ``initial`` (or a checkpoint) in its
    chart.  Returns the accepted parameters, their evaluation, the stop
    reason, the evaluation count and per-iteration records.  An exception
    from the evaluator or an observer propagates with the last accepted state
    in ``__dict__["recovery_diagnostics"]``.

    >>> r = recover_parameters(lambda x: Evaluation(0.5 * (x[0] - 2.0) ** 2, (x[0] - 2.0,)), (0.0,))
    >>> abs(r.parameters[0] - 2.0) < 1e-8
    True
    """
    with LifecycleRecorder("registration", observe_lifecycle) as lifecycle:
        result = _Solve(evaluator, initial, policy or RecoveryPolicy(), objective_id, checkpoint,
                        checkpoint_relative_tolerance, lifecycle).run(observe, cancelled, save_checkpoint)
        lifecycle.finish(result)
        return result


class _Solve:
    def __init__(self, evaluator, initial, policy, objective_id, checkpoint, tolerance, lifecycle) -> None:
        self.policy, self.objective_id, self.lifecycle = policy, objective_id, lifecycle
        self.tolerance = finite_scalar(tolerance, "checkpoint_relative_tolerance", minimum=0)
        if objective_id is not None and (not isinstance(objective_id, str) or not objective_id.strip()):
            raise ContractError("objective_id must be a nonempty immutable model/chart identity")
        if checkpoint is not None:
            if not isinstance(checkpoint, RecoveryCheckpoint):
                raise ContractError("checkpoint must be a RecoveryCheckpoint")
            if objective_id is None or objective_id != checkpoint.objective_id:
                raise ContractError("checkpoint objective/chart identity differs")
            budgets = dict(max_iterations=policy.max_iterations, max_evaluations=policy.max_evaluations)
            if replace(checkpoint.policy, **budgets) != policy:
                raise ContractError("checkpoint numerical policy differs; only budgets may change")
            if initial is not None and tuple(initial) != checkpoint.parameters:
                raise ContractError("initial parameters differ from the checkpoint")
            initial = checkpoint.parameters
        if initial is None:
            raise ContractError("supply initial parameters or a checkpoint")
        self.x: Vector = tuple(finite_scalar(v, "initial parameter") for v in initial)
        if not self.x:
            raise ContractError("recovery requires at least one active parameter")
        self.evaluator = evaluator
        self.checkpoint = checkpoint
        self.calls = checkpoint.evaluations if checkpoint is not None else 0
        if checkpoint is not None and self.calls >= policy.max_evaluations:
            raise ContractError("increase max_evaluations to verify and resume this checkpoint")
        self.rejected = list(checkpoint.rejected_trials) if checkpoint is not None else []
        self.unresolved = checkpoint.resolution_limited_trials if checkpoint is not None else 0
        self.history = list(checkpoint.history) if checkpoint is not None else []
        pairs = checkpoint.curvature_pairs if checkpoint is not None else ()
        for s, y, rho in pairs:
            if len(s) != len(self.x) or len(y) != len(self.x) or \
                    not all(math.isfinite(v) for v in (*s, *y, rho)) or rho <= 0:
                raise ContractError("invalid checkpoint curvature pair")
        if len(pairs) > policy.memory:
            raise ContractError("checkpoint exceeds the curvature memory bound")
        self.memory = _Curvature(pairs, policy.memory)
        self.current: Evaluation | None = None
        self.iteration = 0

    def _evaluate(self, point: Vector) -> Evaluation:
        if self.calls >= self.policy.max_evaluations:
            raise _OutOfEvaluations
        self.calls += 1
        return _verified(self.evaluator(point), len(self.x))

    def _snapshot(self) -> RecoveryCheckpoint | None:
        if self.objective_id is None:
            return None
        return RecoveryCheckpoint(self.objective_id, self.policy, self.x, self.current, tuple(self.memory.pairs),
                                  tuple(self.history), tuple(self.rejected), self.calls, self.unresolved)

    def _result(self, reason: str) -> RecoveryResult:
        return RecoveryResult(self.x, self.current, reason, self.calls, tuple(self.history), tuple(self.rejected),
                              self.unresolved, self._snapshot())

    def _reject(self, trial: RejectedTrial) -> None:
        self.rejected.append(trial)
        self.lifecycle.rejected(self.iteration, "domain: " + trial.message)

    def _count_unresolved(self) -> None:
        self.unresolved += 1

    def _search(self, direction: Vector, first_step: float = 1.0):
        search = _LineSearch(self._evaluate, self.x, self.current, direction, self.policy,
                             (self._reject, self._count_unresolved))
        return search.run(first_step)

    def run(self, observe, cancelled, save_checkpoint) -> RecoveryResult:
        policy = self.policy
        try:
            self.current = self._evaluate(self.x)
            if self.checkpoint is not None:
                old = _verified(self.checkpoint.evaluation, len(self.x))
                if not all(math.isclose(a, b, rel_tol=self.tolerance, abs_tol=0)
                           for a, b in zip((self.current.loss, *self.current.gradient), (old.loss, *old.gradient))):
                    raise ContractError("checkpoint accepted value/gradient differs from the current oracle")
            if not self.history:
                self.history.append(IterationRecord(0, self.current.loss, _linf(self.current.gradient), 0.0,
                                                    self.calls, 0))
            if self.history[-1].iteration > policy.max_iterations:
                raise ContractError("max_iterations precedes the checkpoint iteration")
            if save_checkpoint is not None and self.objective_id is None:
                raise ContractError("save_checkpoint requires objective_id")
            self.lifecycle.start(objective=self.current.loss, iteration=self.history[-1].iteration,
                                 accepted_steps=sum(r.iteration > 0 for r in self.history),
                                 resumed=self.checkpoint is not None)
            reason = "iteration_budget"
            for self.iteration in range(self.history[-1].iteration + 1, policy.max_iterations + 1):
                if cancelled is not None and cancelled():
                    reason = "cancelled"
                    break
                if _linf(self.current.gradient) <= policy.gradient_tolerance:
                    reason = "gradient_tolerance"
                    break
                unresolved = self.unresolved
                d = self.memory.direction(self.current.gradient)
                slope = _inner(d, self.current.gradient)
                if not (all(math.isfinite(v) for v in d) and math.isfinite(slope) and slope < 0.0):
                    self.memory.clear()
                    d = _descent(self.current.gradient)
                try:
                    found = self._search(d)
                    if found is None:
                        # a poor curvature model need not end the solve: one
                        # steepest-descent search from a small step, empty memory
                        self.memory.clear()
                        first = max(policy.minimum_step, 2.0 ** -min(16, policy.line_search_evaluations))
                        d = _descent(self.current.gradient)
                        found = self._search(d, first)
                except _OutOfEvaluations:
                    reason = "evaluation_budget"
                    break
                if found is None:
                    reason = "loss_resolution" if self.unresolved > unresolved else "line_search_failed"
                    self.lifecycle.rejected(self.iteration, reason)
                    break
                alpha, value, kind = found
                step = tuple(alpha * v for v in d)
                self.memory.update(step, tuple(a - b for a, b in zip(value.gradient, self.current.gradient)))
                previous = self.current.loss
                self.x, self.current = _axpy(1.0, step, self.x), value
                record = IterationRecord(self.iteration, value.loss, _linf(value.gradient), alpha, self.calls,
                                         len(self.memory.pairs), kind)
                self.history.append(record)
                self.lifecycle.accepted(self.iteration, value.loss)
                if observe is not None:
                    observe(record)
                if save_checkpoint is not None:
                    save_checkpoint(self._snapshot())
                if _linf(value.gradient) <= policy.gradient_tolerance:
                    reason = "gradient_tolerance"
                    break
                if _linf(step) <= policy.step_tolerance * max(1.0, _linf(self.x)):
                    reason = "step_tolerance"
                    break
                if abs(previous - value.loss) <= policy.relative_loss_tolerance * max(1.0, abs(previous),
                                                                                      abs(value.loss)):
                    reason = "loss_stagnation"
                    break
            return self._result(reason)
        except Exception as error:
            if self.current is not None:
                error.__dict__["recovery_diagnostics"] = self._result("evaluation_failed")
                if hasattr(error, "add_note"):  # Python >= 3.11
                    error.add_note("Last accepted recovery state is in recovery_diagnostics.")
            raise
