"""Point-cone source law and original-history batches.

PointConeSource mirrors python/dpt/transport/source.py:54-85: every history
aims at a uniform point of the detector rectangle,
``y = lower_corner + U*width*u + V*height*v``, direction ``(y - s)/|y - s|``,
score weight ``weight`` (no isotropic or inverse-square factor is implied).
HistoryBatch mirrors transport/rng.py:80-150: identities
first_history .. first_history+count-1 under one 64-bit seed.
"""

from __future__ import annotations

import math
import operator
from dataclasses import dataclass

from ..geometry import DetectorGeometry
from .model import TransportError, _physical


@dataclass(frozen=True)
class PointConeSource:
    geometry: DetectorGeometry
    weight: float = 1.0

    def __post_init__(self) -> None:
        _physical(self.weight, "source weight", nonnegative=True)
        if not isinstance(self.geometry, DetectorGeometry):
            raise TransportError("PointConeSource.geometry must be a DetectorGeometry")
        g = self.geometry
        corner = self.corner_mm
        for column in (0, g.shape[1]):
            for row in (0, g.shape[0]):
                point = tuple(
                    corner[i] + column * g.spacing_mm[0] * g.u[i] + row * g.spacing_mm[1] * g.v[i] - g.source_mm[i]
                    for i in range(3)
                )
                distance = math.hypot(*point)
                if not math.isfinite(distance) or distance == 0:
                    raise TransportError("source-detector segments must have finite positive lengths")

    @property
    def corner_mm(self) -> tuple[float, float, float]:
        """Lower detector corner, as transport/source.py:165-168 computes it."""
        g = self.geometry
        return tuple(g.origin_mm[i] - 0.5 * (g.spacing_mm[0] * g.u[i] + g.spacing_mm[1] * g.v[i]) for i in range(3))

    @property
    def extent_mm(self) -> tuple[float, float]:
        """(width along u, height along v) (transport/source.py:176-179)."""
        g = self.geometry
        return (g.spacing_mm[0] * g.shape[1], g.spacing_mm[1] * g.shape[0])


def _identity(value, name: str, maximum: int, positive: bool = False) -> int:
    if isinstance(value, bool):
        raise TransportError(f"{name} must be an integer history identity")
    try:
        result = operator.index(value)
    except TypeError as error:
        raise TransportError(f"{name} must be an integer history identity") from error
    if not int(positive) <= result < maximum:
        raise TransportError(f"{name} is outside its declared history identity range")
    return result


@dataclass(frozen=True)
class HistoryBatch:
    seed: int
    first_history: int
    count: int

    def __post_init__(self) -> None:
        object.__setattr__(self, "seed", _identity(self.seed, "seed", 2**64))
        object.__setattr__(self, "first_history", _identity(self.first_history, "first_history", 2**64))
        object.__setattr__(self, "count", _identity(self.count, "count", 2**31, positive=True))
        if self.first_history + self.count > 2**64:
            raise TransportError("the history range would wrap its 64-bit identity")

    def shard(self, parts: int, rank: int) -> HistoryBatch:
        """Balanced contiguous shard of this batch's identities (transport/rng.py:154-161)."""
        parts, rank = operator.index(parts), operator.index(rank)
        if not 1 <= parts <= self.count or not 0 <= rank < parts:
            raise TransportError("shard parts must be 1..count and rank in 0..parts-1")
        base, extra = divmod(self.count, parts)
        start = rank * base + min(rank, extra)
        return HistoryBatch(self.seed, self.first_history + start, base + (1 if rank < extra else 0))
