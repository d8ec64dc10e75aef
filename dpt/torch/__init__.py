"""Torch adapter: ``Projector``, ``compose_pose`` and the ``dpt::`` operators."""

from .geometry import compose_pose
from .modules import Projector

__all__ = ["Projector", "compose_pose"]
