"""dpt: MI355X-native differentiable cone-beam X-ray projector.

Drop-in for the reference's ``dpt.project()`` / ``dpt.torch.Projector`` path:
``(views, rows, columns)`` optical depth with volume and 6-DoF pose VJPs, run
by hand-written gfx950 HIP kernels behind a C ABI (include/dpt_amd.h).
"""

from .api import ImagingVolume, project
from .contracts import ContractError, NumericalError, TrialDomainError
from .geometry import (
    BatchedGeometry,
    ConeBeamIntrinsics,
    CurvedDetectorGeometry,
    DetectorGeometry,
    ParallelGeometry,
    RayBundleGeometry,
    RigidTransform,
    compose_pose,
)
from .volumes import GridSpec

__all__ = [
    "BatchedGeometry",
    "ConeBeamIntrinsics",
    "ContractError",
    "CurvedDetectorGeometry",
    "DetectorGeometry",
    "GridSpec",
    "ImagingVolume",
    "NumericalError",
    "ParallelGeometry",
    "RayBundleGeometry",
    "RigidTransform",
    "TrialDomainError",
    "compose_pose",
    "project",
]
__version__ = "0.1.0"
