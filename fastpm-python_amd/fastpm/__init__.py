"""FastPM particle-mesh solver on AMD MI355X (gfx950).

Drop-in for the reference's ``fastpm.core`` path: ``Solver``
(whitenoise / linear / lpt / nbody), ``StateVector`` and the stepping helpers,
backed by hand-written HIP kernels in ``libfpm.so``."""
from .core import Solver, StateVector, FastPMStep, leapfrog, leapfrog_with_lpt, autostages  # noqa: F401
from .pm import ParticleMesh  # noqa: F401
from .background import MatterDominated  # noqa: F401

__version__ = "0.1.0"
