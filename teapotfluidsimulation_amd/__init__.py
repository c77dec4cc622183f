"""MI355X (gfx950) implementation of PourLab's 3D PIC/FLIP free-surface pour.

The work runs in libpourlab3d.so (hand-written HIP kernels behind the C ABI of
include/pourlab3d.h); this package is the host-side mirror of the reference's solver API.
"""
from .sim3d import (  # noqa: F401
    AIR,
    DISCHARGED,
    Ensemble3D,
    FLUID,
    IN_POT,
    LOST,
    LearnedPreconditioner,
    SOLID,
    SPILLED,
    Pcg3D,
    PoissonSet,
    Recorder3D,
    Simulation3D,
    gentlePour3D,
    lipPour3D,
    lowSpoutPour3D,
    rampPour3D,
    variedPour3D,
)
