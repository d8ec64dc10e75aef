"""The chart search (dpt.registration.recover_parameters) and its lifecycle
events, on host objectives (no GPU).

Cases follow the reference's own tests of the same interface
(tests/python/test_solver_lifecycle.py, test_recovery_trial_diagnostics.py):
a quadratic and a domain-restricted quadratic with rejected trials, the
last accepted state kept on a fatal evaluator error, a checkpointed solve
resumed to the same result as an uninterrupted one, lifecycle events that
change nothing and observer errors that propagate; plus the Rosenbrock
valley, budgets, cancellation and policy validation.
"""

from __future__ import annotations

import math
from dataclasses import FrozenInstanceError, asdict

import pytest

from dpt.contracts import ContractError, NumericalError, TrialDomainError
from dpt.registration import Evaluation, RecoveryPolicy, recover_parameters
from dpt.solver_lifecycle import solver_termination, termination_reason


def quadratic(x):
    return Evaluation(0.5 * ((x[0] - 1.0) ** 2 + 10.0 * (x[1] + 2.0) ** 2), (x[0] - 1.0, 10.0 * (x[1] + 2.0)))


def rosenbrock(x):
    a, b = x
    return Evaluation((1 - a) ** 2 + 100 * (b - a * a) ** 2, (-2 * (1 - a) - 400 * a * (b - a * a), 200 * (b - a * a)))


def test_quadratic_and_rosenbrock_reach_the_gradient_tolerance():
    r = recover_parameters(quadratic, (3.0, 2.0))
    assert r.reason == "gradient_tolerance" and r.stationary
    assert r.parameters == pytest.approx((1.0, -2.0), abs=1e-6)
    r = recover_parameters(rosenbrock, (-1.2, 1.0))
    assert r.reason == "gradient_tolerance"
    assert r.parameters == pytest.approx((1.0, 1.0), abs=1e-6)
    losses = [h.loss for h in r.history]
    assert all(b <= a for a, b in zip(losses, losses[1:]))  # accepted points never increase the loss
    assert r.history[0].iteration == 0 and r.history[0].line_search == "initial"
    assert r.evaluations == r.history[-1].evaluations


def test_rejected_domain_trials_are_recorded_and_never_accepted():
    def objective(x):
        if x[0] <= 0:
            raise TrialDomainError("positive chart")
        return Evaluation(20 * (x[0] - 0.2) ** 2, (40 * (x[0] - 0.2),))

    r = recover_parameters(objective, (1.0,))
    assert r.reason == "gradient_tolerance"
    assert r.parameters[0] == pytest.approx(0.2)
    assert r.rejected_trials and all(t.parameters[0] <= 0 for t in r.rejected_trials)
    assert all(h.loss >= 0 for h in r.history)


def test_fatal_numerical_failure_retains_last_accepted_result():
    def objective(x):
        if x[0] != 1.0:
            raise NumericalError("corrupt fixed data")
        return Evaluation(0.5, (1.0,))

    with pytest.raises(NumericalError, match="corrupt fixed data") as failure:
        recover_parameters(objective, (1.0,))
    snapshot = failure.value.__dict__["recovery_diagnostics"]
    assert snapshot.parameters == (1.0,)
    assert snapshot.evaluation == Evaluation(0.5, (1.0,))
    assert snapshot.evaluations == 2
    assert len(snapshot.history) == 1
    assert snapshot.reason == "evaluation_failed"


def test_checkpoint_resume_reproduces_the_uninterrupted_solve():
    rules = RecoveryPolicy(max_iterations=40)
    full = recover_parameters(rosenbrock, (-1.2, 1.0), policy=rules)
    saved = []
    first = recover_parameters(rosenbrock, (-1.2, 1.0), policy=RecoveryPolicy(max_iterations=5),
                               objective_id="rosenbrock", save_checkpoint=saved.append)
    assert first.reason == "iteration_budget" and len(saved) == 5
    resumed = recover_parameters(rosenbrock, policy=rules, objective_id="rosenbrock", checkpoint=saved[-1])
    assert resumed.parameters == full.parameters
    assert resumed.reason == full.reason
    assert [h.iteration for h in resumed.history] == [h.iteration for h in full.history]
    with pytest.raises(ContractError, match="identity"):
        recover_parameters(rosenbrock, policy=rules, objective_id="other", checkpoint=saved[-1])
    with pytest.raises(ContractError, match="only budgets"):
        recover_parameters(rosenbrock, policy=RecoveryPolicy(memory=3), objective_id="rosenbrock",
                           checkpoint=saved[-1])
    with pytest.raises(ContractError, match="differs from the current oracle"):
        recover_parameters(lambda x: quadratic(x), policy=rules, objective_id="rosenbrock", checkpoint=saved[-1])


def _solve(observer=None, reject=False):
    calls = []

    def oracle(x):
        calls.append(x)
        if reject and x[0] <= 0:
            raise TrialDomainError("positive chart")
        if reject:
            return Evaluation(20 * (x[0] - 0.2) ** 2, (40 * (x[0] - 0.2),))
        return quadratic(x)

    result = recover_parameters(oracle, (1.0,) if reject else (3.0, 2.0), observe_lifecycle=observer)
    return result, calls


@pytest.mark.parametrize("reject", [False, True])
def test_lifecycle_events_change_nothing(reject):
    expected, calls = _solve(reject=reject)
    events = []
    actual, observed = _solve(events.append, reject=reject)
    assert asdict(actual) == asdict(expected) and observed == calls
    assert events[0].status == "initial" and events[0].iteration == events[0].accepted_steps == 0
    assert events[-1].status == "terminal" and sum(e.status == "terminal" for e in events) == 1
    assert events[-1].termination == solver_termination(actual)
    assert events[-1].objective == actual.evaluation.loss
    assert events[-1].accepted_steps == sum(e.status == "accepted" for e in events)
    if reject:
        assert any(e.status == "rejected" for e in events)
    with pytest.raises(FrozenInstanceError):
        events[0].objective = 99


@pytest.mark.parametrize("error_type", [ValueError, NumericalError, KeyboardInterrupt])
@pytest.mark.parametrize("phase", ["accepted", "rejected", "terminal"])
def test_observer_errors_propagate_with_a_terminal_snapshot(phase, error_type):
    sentinel = error_type("recorder failed")
    events = []

    def observer(event):
        events.append(event)
        if event.status == phase:
            raise sentinel

    with pytest.raises(error_type) as raised:
        _solve(observer, reject=True)
    assert raised.value is sentinel
    snapshot = raised.value.__dict__["solver_lifecycle"]
    assert snapshot.status == "terminal" and snapshot.reason == "exception"
    assert snapshot.termination.category == "failed"
    assert snapshot.accepted_steps == events[-1].accepted_steps


def test_budgets_cancellation_and_policy_validation():
    r = recover_parameters(rosenbrock, (-1.2, 1.0), policy=RecoveryPolicy(max_evaluations=7))
    assert r.reason == "evaluation_budget" and r.evaluations <= 7
    r = recover_parameters(rosenbrock, (-1.2, 1.0), policy=RecoveryPolicy(max_iterations=3))
    assert r.reason == "iteration_budget" and r.history[-1].iteration == 3
    r = recover_parameters(rosenbrock, (-1.2, 1.0), cancelled=lambda: True)
    assert r.reason == "cancelled" and r.parameters == (-1.2, 1.0)
    assert termination_reason("loss_stagnation").category == "stagnation"
    for bad in (dict(armijo=0.95), dict(minimum_step=20.0), dict(memory=0), dict(gradient_tolerance=math.nan)):
        with pytest.raises(ContractError):
            RecoveryPolicy(**bad)
    with pytest.raises(ContractError):
        recover_parameters(quadratic)
    with pytest.raises(ContractError, match="dimension"):
        recover_parameters(lambda x: Evaluation(0.0, (0.0, 0.0)), (1.0,))
