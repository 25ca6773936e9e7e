"""Glass particle distributions (reference: fastpm/glass.py).

A glass is made by evolving a randomly displaced lattice under the PM force
with its sign inverted (particles repel) and a damping term, so the
particles settle into a configuration with suppressed large-scale power and
no preferred lattice directions.  The stepping, damping, periodic drift and
force are those of fastpm/glass.py:8-118; the force runs on the library's
Force path (binned paint, fused transforms, stencil readout on float32
tile-divisible meshes) with factor -1 and no window compensation.
"""
import ctypes

import numpy as np
import torch

from . import core, _lib
from .core import leapfrog
from .cosmology import Planck15
from .force.kernels import laplace, gradient
from .pm import _stream

ctypes_f64x3 = ctypes.c_double * 3


def generate_glass_particle_grid(pm, seed, B=2, spread=3.0, N=3):
    """Positions [pm.Nmesh.prod(), 3] (device tensor, in [0, BoxSize)) of a
    glass made from the uniform grid of `pm` (fastpm/glass.py:8-47).

    spread : initial random displacement in units of the particle spacing.
    N : number of 2 pi periods the damped glass dynamics run for."""
    solver = Solver(pm, B)
    Q = pm.generate_uniform_particle_grid(shift=0)
    # the period is 2 pi; after N + 1/4 periods the power is near its minimum
    stages = np.linspace(0, np.pi * 2 * (N + 0.25), int(8 * (N + 0.25) + 1))
    glass = solver.run(seed, Q, stages, spread=spread, N=N)
    X = glass.X
    box = (ctypes_f64x3)(*[float(b) for b in pm.BoxSize])
    _lib.call("fpm_drift_wrap", pm.dtype, X.data_ptr(), X.data_ptr(), len(X), 0.0, box, _stream())
    return X


class Solver(core.Solver):
    """Solver of the glass dynamics (fastpm/glass.py:49-86); the cosmology
    only names the run (Planck15, as in the reference)."""

    def __init__(self, pm, B=2):
        super().__init__(pm, Planck15, B=B)

    @property
    def nbodystep(self):
        return GlassStep(self)

    def run(self, seed, Q, stages, spread=3.0, N=3):
        """Uniform random displacements of `spread` particle spacings drawn
        with numpy's RandomState(seed) (the reference's generator, so the
        start is the same), zero momentum and force, then the damped leapfrog
        over `stages`; the state is tagged a = 1 at the end."""
        rng = np.random.RandomState(seed + self.pm.comm.rank)
        n = len(Q)
        nbar = 1 / (float(np.prod(self.pm.BoxSize)) / self.pm.comm.allreduce(n))
        state = core.StateVector(self, Q)
        S = spread * (rng.uniform(size=(n, 3)) - 0.5) * (nbar ** -0.3333333)
        state.S.copy_(torch.from_numpy(S).to(dtype=state.S.dtype))
        state.P.zero_()
        state.F.zero_()
        state.a["S"] = 0
        state.a["P"] = 0
        self.nbody(state, leapfrog(stages))
        state.a["S"] = 1
        state.a["P"] = 1
        state.a["F"] = 1
        return state


class GlassStep(core.FastPMStep):
    """Time step with the PM force inverted and a damping term
    (fastpm/glass.py:88-118)."""

    def Kick(self, state, ai, ac, af):
        fac = af - ai
        # P = P + fac (F - P)
        with self._mark("kick"):
            _lib.call("fpm_axpby", state.dtype, state.P.data_ptr(), state.F.data_ptr(),
                      state.P.numel(), float(1.0 - fac), float(fac), _stream())
        state.a["P"] = af

    def Drift(self, state, ai, ac, af):
        fac = af - ai
        box = (ctypes_f64x3)(*[float(b) for b in self.solver.pm.BoxSize])
        # S = (S + fac P) mod BoxSize
        with self._mark("drift"):
            _lib.call("fpm_drift_wrap", state.dtype, state.S.data_ptr(), state.P.data_ptr(),
                      len(state.S), float(fac), box, _stream())
        state.a["S"] = af

    def force_common(self):
        return laplace * -1.0

    def force_ops(self):
        return [laplace * gradient(d, order=1) * -1.0 for d in range(3)]

    def Force(self, state, ai, ac, af):
        self.force_kernels(state, keep_delta_k=False)
        state.a["F"] = af

