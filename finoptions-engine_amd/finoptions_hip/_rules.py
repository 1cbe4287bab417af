"""Declarative checks of the immutable request values.

Each value class states its checks as a table: one row per field, in the order in which problems
are reported, each row a label, the value and the checks it must pass. ``enforce`` walks the table
and raises ``InputError`` with the message of the first check that fails. A check is a predicate
paired with a message template; templates may use ``{field}``, ``{value!r}``, ``{value}`` and
``{kind}`` (the value's type name). The message texts are the public error contract of the
reference's facade, which callers match on; the checks that need the numerical core (finiteness
and ordering of grids and curves, amounts, volatility) run in libfohip through the
``fohip_validate_*`` entry points.
"""

from __future__ import annotations

from collections.abc import Callable, Iterable, Sequence
from dataclasses import dataclass
from math import copysign, isfinite
from typing import Any

from ._lib import InputError

U32_MAX = (2**32) - 1


@dataclass(frozen=True, slots=True)
class Check:
    """``holds(value)`` must be true, else ``message`` (a format template) is raised."""

    holds: Callable[[Any], bool]
    message: str

    def failure(self, field: str, value: Any) -> str | None:
        if self.holds(value):
            return None
        return self.message.format(field=field, value=value, kind=type(value).__name__)


@dataclass(frozen=True, slots=True)
class Each:
    """Every element of a sequence passes ``checks`` (reported under the sequence's label)."""

    checks: Sequence[Check]

    def failure(self, field: str, value: Any) -> str | None:
        for item in value:
            for check in self.checks:
                message = check.failure(field, item)
                if message is not None:
                    return message
        return None


@dataclass(frozen=True, slots=True)
class Lazy:
    """A check whose message is built only when it fails: ``holds(value)`` else
    ``describe(value)``."""

    holds: Callable[[Any], bool]
    describe: Callable[[Any], str]

    def failure(self, field: str, value: Any) -> str | None:
        return None if self.holds(value) else self.describe(value)


Row = tuple[str, Any, Sequence["Check | Each | Lazy"]]


def enforce(rows: Iterable[Row]) -> None:
    """Raise ``InputError`` for the first failing check of the first failing row."""
    for field, value, checks in rows:
        for check in checks:
            message = check.failure(field, value)
            if message is not None:
                raise InputError(message)


# ------------------------------------------------------------------------------ check builders

def exact_type(cls: type, message: str) -> Check:
    return Check(lambda v: type(v) is cls, message)


def canonical_nonnegative(v: Any) -> bool:
    """A finite float that is positive or +0.0 (never -0.0)."""
    return type(v) is float and isfinite(v) and v >= 0.0 and copysign(1.0, v) > 0.0


PLAIN_U32 = (
    exact_type(int, "{field} must be a plain int; observed {kind}"),
    Check(lambda v: 0 <= v <= U32_MAX,
          "{field} must be in the unsigned 32-bit node domain; observed {value}"),
)
FLOAT_TUPLE = (
    exact_type(tuple, "{field} must be an immutable tuple[float, ...]"),
    Check(lambda v: all(type(x) is float for x in v), "every {field} value must be a plain float"),
)
U32_TUPLE = (
    exact_type(tuple, "{field} must be an immutable tuple[int, ...]"),
    Each(PLAIN_U32),
)
PLAIN_FLOAT = (exact_type(float, "{field} must be a plain float; observed {kind}"),)


def instance_of(cls: type) -> tuple[Check]:
    """``{field} must be <cls>; observed <type>`` for an exact runtime type."""
    return (exact_type(cls, "{field} must be " + cls.__name__ + "; observed {kind}"),)


def int_in(low: int, high: int) -> tuple[Check]:
    return (Check(lambda v: type(v) is int and low <= v <= high,
                  "{field} must be a plain int in [" + f"{low}, {high}" + "]; observed {value!r}"),)


__all__ = ["Check", "Each", "Lazy", "FLOAT_TUPLE", "PLAIN_FLOAT", "PLAIN_U32", "U32_MAX", "U32_TUPLE",
           "canonical_nonnegative", "enforce", "exact_type", "instance_of", "int_in"]
