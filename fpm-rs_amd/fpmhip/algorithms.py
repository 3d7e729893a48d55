"""Algorithm configurations with fpm-rs's defaults and the `.run(problem)`
surface of its Python bindings (python/src/reconstruction.rs:1107-1128,
2239 AlternatingProjection, 2501 Mpie, 2613 Epry)."""
from __future__ import annotations

from dataclasses import dataclass, fields

from . import abi


@dataclass
class _Base:
    loss_type: str = "amplitude_mse"

    kind = -1

    def to_c(self) -> abi.Algorithm:
        a = abi.Algorithm()
        a.kind = self.kind
        a.loss_type = abi.LOSS[self.loss_type]
        a.batch_size = 1
        a.constrain_pupil = 1
        for f in fields(self):
            if f.name == "loss_type":
                continue
            v = getattr(self, f.name)
            if f.name in ("recover_pupil", "constrain_pupil_support", "recover_frame_gains", "recover_background"):
                name = {
                    "recover_pupil": "recover_pupil",
                    "constrain_pupil_support": "constrain_pupil",
                    "recover_frame_gains": "recover_gain",
                    "recover_background": "recover_background",
                }[f.name]
                setattr(a, name, int(bool(v)))
            else:
                setattr(a, f.name, v)
        return a

    def run(self, problem, precision="c128", schedule="sequential", context=None, callbacks=None,
            resume_from=None, schedule_seed=0):
        """Runner::run on the GPU (one session); returns a ReconstructionResult.
        `callbacks` may hold CheckpointEvery; `resume_from` a
        ReconstructionCheckpoint (python/fpm_rs/__init__.pyi:3316-3336)."""
        from .runtime import run_algorithm

        return run_algorithm(self, problem, precision=precision, schedule=schedule, context=context,
                             callbacks=callbacks, resume_from=resume_from, schedule_seed=schedule_seed)


@dataclass
class AlternatingProjection(_Base):
    """alternating_projection.rs:51-60."""

    iterations: int = 50
    object_step: float = 1.0
    batch_size: int = 1
    epsilon: float = 1e-10
    kind = abi.AP


@dataclass
class Fpie(_Base):
    """fpie.rs:53-62 (stability clamped to [0, 1] by its setter, :80)."""

    iterations: int = 50
    object_step: float = 0.8
    stability: float = 0.1
    batch_size: int = 1
    epsilon: float = 1e-10
    kind = abi.FPIE


@dataclass
class Mpie(_Base):
    """mpie.rs:81-95."""

    iterations: int = 50
    object_step: float = 0.2
    stability: float = 0.05
    momentum_interval: int = 30
    momentum_friction: float = 0.9
    momentum_feedback: float = 0.9
    batch_size: int = 1
    epsilon: float = 1e-10
    kind = abi.MPIE


@dataclass
class Epry(_Base):
    """epry.rs:91-111."""

    iterations: int = 100
    object_step: float = 0.8
    pupil_step: float = 0.1
    batch_size: int = 1
    recover_pupil: bool = True
    constrain_pupil_support: bool = True
    recover_frame_gains: bool = False
    gain_step: float = 0.2
    minimum_gain: float = 1e-6
    maximum_gain: float = 1e6
    recover_background: bool = False
    background_step: float = 0.2
    minimum_background: float = 0.0
    maximum_background: float = 1e12
    epsilon: float = 1e-10
    kind = abi.EPRY
