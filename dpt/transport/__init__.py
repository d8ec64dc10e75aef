"""Monte Carlo photon transport with analogue Woodcock tracking.

MI355X-native restatement of the reference's transport path
(python/dpt/transport): point-cone sources, Klein-Nishina and Rayleigh
scattering, primary and scatter detector tallies with variance of the mean.
The walk runs in the HIP kernel behind the dpt_transport_* C ABI; there is no
CPU execution path.
"""

from .materials import chest_transport_scene
from .model import (
    HistoryStatus,
    IncompleteHistoryError,
    MaterialGrid,
    OrientedDetector,
    PlanarDetector,
    TransportError,
    TransportSpec,
)
from .physical import AtomicScatteringTable, PhysicalTransport, SourceSpectrum
from .simulation import (
    HistoryStatusHistogram,
    SimulationResult,
    SimulationWorkspace,
    TransportWorkspace,
    prepare_simulation,
    prepare_transport,
    simulate,
    trace_histories,
)
from .source import HistoryBatch, PointConeSource

__all__ = [
    "AtomicScatteringTable",
    "HistoryBatch",
    "HistoryStatus",
    "HistoryStatusHistogram",
    "IncompleteHistoryError",
    "MaterialGrid",
    "OrientedDetector",
    "PhysicalTransport",
    "PlanarDetector",
    "PointConeSource",
    "SimulationResult",
    "SimulationWorkspace",
    "SourceSpectrum",
    "TransportError",
    "TransportSpec",
    "TransportWorkspace",
    "chest_transport_scene",
    "prepare_simulation",
    "prepare_transport",
    "simulate",
    "trace_histories",
]
