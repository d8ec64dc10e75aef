"""Errors and scalar validation shared by the projector's host interfaces.

The three error types and the status-bit mapping mirror the reference's
contract (python/dpt/contracts.py:162-266): a violated value or layout rule is
a ``ContractError`` (a ``ValueError``), a result that cannot be represented is
a ``NumericalError`` (a ``FloatingPointError``), and a physical input outside
its domain is a ``TrialDomainError`` (a ``NumericalError``).
"""

from __future__ import annotations

import math
from numbers import Real
from typing import Any


class ContractError(ValueError):
    """A value, layout or ownership rule required by an operator was violated."""


class NumericalError(FloatingPointError):
    """An operation cannot represent its declared result; discard its outputs."""


class TrialDomainError(NumericalError):
    """A proposed parameter point lies outside its chart (e.g. negative attenuation)."""


# Device status bits, in delivery precedence (python/dpt/contracts.py:229-234).
STATUS_POSE = 1
STATUS_RANGE = 2
STATUS_FAILURE = 4
STATUS_DOMAIN = 8

_PRECEDENCE = (
    (STATUS_POSE, ContractError),
    (STATUS_DOMAIN, TrialDomainError),
    (STATUS_RANGE, NumericalError),
    (STATUS_FAILURE, NumericalError),
)

# projection.py:505-515 message table
PROJECTION_MESSAGES = {
    STATUS_POSE: "invalid rigid pose coordinate map",
    STATUS_RANGE: "non-finite projection or gradient; discard these outputs",
    STATUS_FAILURE: "numerical execution failure",
    STATUS_DOMAIN: "field must be finite and non-negative",
}


def status_error(code: int, operator: str, messages: dict[int, str] | None = None) -> Exception | None:
    """Map one status word to its exception (or None when the word is clear)."""
    code = int(code)
    if code == 0:
        return None
    table = PROJECTION_MESSAGES if messages is None else messages
    for bit, kind in _PRECEDENCE:
        if code & bit:
            return kind(f"{operator}: {table[bit]}")
    return NumericalError(f"{operator}: unknown device status {code}")


def raise_view_statuses(codes, operator: str = "projection") -> None:
    """Raise the first failing view's error, noting every other failing view.

    Mirrors StatusCheckpoint delivery (python/dpt/_runtime.py:391-445): words
    are visited in view order and the first typed error is raised.
    """
    first: Exception | None = None
    notes: list[str] = []
    for view, code in enumerate(codes):
        error = status_error(int(code), operator)
        if error is None:
            continue
        error.args = (f"{operator} view {view}: {error.args[0]}",)
        notes.append(f"view {view} status {int(code)}: {error.args[0]}")
        if first is None:
            first = error
    if first is not None:
        first.status_diagnostics = tuple(notes)  # type: ignore[attr-defined]
        if hasattr(first, "add_note"):  # Python >= 3.11
            for note in notes:
                first.add_note(note)
        raise first


def integer(value: Any, name: str, *, minimum: int = 0, maximum: int = 2**31 - 1) -> int:
    """Require a Python int (not bool) in [minimum, maximum]."""
    if type(value) is not int or not minimum <= value <= maximum:
        raise ContractError(f"{name} must be an integer in [{minimum}, {maximum}]")
    return value


def finite_scalar(value: Any, name: str, *, minimum: float | None = None) -> float:
    """Accept a finite real scalar, optionally with an inclusive lower bound."""
    if isinstance(value, bool) or not isinstance(value, Real):
        raise ContractError(f"{name} must be a real scalar")
    result = float(value)
    if not math.isfinite(result) or (minimum is not None and result < minimum):
        qualifier = "finite" if minimum is None else f"finite and at least {minimum}"
        raise ContractError(f"{name} must be {qualifier}")
    return result


def finite_tuple(
    values: Any,
    name: str,
    *,
    positive: bool = False,
    minimum: float | None = 0.0,
    length: int | None = None,
) -> tuple[float, ...]:
    """Validate a non-empty tuple of finite reals (non-negative by default)."""
    if not isinstance(values, tuple) or not values:
        raise ContractError(f"{name} must be a non-empty immutable tuple")
    if length is not None and len(values) != length:
        raise ContractError(f"{name} must contain {length} values")
    result = tuple(finite_scalar(value, name, minimum=minimum) for value in values)
    if positive and any(value <= 0.0 for value in result):
        raise ContractError(f"{name} must be positive")
    return result
