"""Torch adapter: ``Projector``, ``compose_pose``, ``PoseObjective`` and the ``dpt::`` operators."""

from .fitting import PoseObjective
from .geometry import compose_pose
from .modules import Projector

__all__ = ["PoseObjective", "Projector", "compose_pose"]
