"""fpmhip — MI355X (gfx950) HIP implementation of fpm-rs's image-plane forward
model and per-LED reconstruction updates (AP, FPIE, mPIE, EPRY), behind the
C ABI of include/fpmhip.h. The device library is loaded lazily by
`fpmhip.runtime`; importing this package needs no GPU."""
from .model import ImagePlaneModel, MeasurementStack, Optics, PlanarLedArray, ReconstructionProblem  # noqa: F401
from .algorithms import AlternatingProjection, Epry, Fpie, Mpie  # noqa: F401
from .checkpoint import CheckpointEvery, ReconstructionCheckpoint  # noqa: F401

__all__ = [
    "ImagePlaneModel", "MeasurementStack", "Optics", "PlanarLedArray", "ReconstructionProblem",
    "AlternatingProjection", "Epry", "Fpie", "Mpie", "CheckpointEvery", "ReconstructionCheckpoint",
]
