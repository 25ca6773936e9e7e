"""Binned power and cross spectra of device fields (estimator semantics of
fastpm/utils.py:77-164; output columns of fastpm/main.py:81-92).

Accumulation is in double precision on the GPU with a fixed reduction order.
The zero mode is excluded; half-spectrum modes with 0 < k_z < N/2 count
twice; bins are linear in |k| from kmin to kmax (default: 0 to the Nyquist
wave number of the mesh)."""
import os

import numpy as np
import torch

from . import _lib
from .pm import ComplexField, RealField, _stream


def power_sums(f1, f2=None, nbins=32, kmin=0.0, kmax=None):
    """Per-bin sums [w, w|k|, w Re(f1 f2*), w|f1|^2, w|f2|^2] as a numpy array."""
    if isinstance(f1, RealField):
        f1 = f1.r2c()
    if isinstance(f2, RealField):
        f2 = f2.r2c()
    pm = f1.pm
    if kmax is None:
        kmax = np.pi * float(pm.Nmesh.min()) / float(pm.BoxSize.max())
    dk = (kmax - kmin) / nbins
    work = pm.power_work(nbins)
    out = torch.empty((nbins, 5), dtype=torch.float64, device=pm.device)
    _lib.call("fpm_power_bin", pm.dtype, f1._t.data_ptr(), None if f2 is None else f2._t.data_ptr(),
              pm._desc, pm.pitch2, int(nbins), float(kmin), float(dk), work.data_ptr(),
              work.numel(), out.data_ptr(), _stream())
    return out.cpu().numpy()


def power(f1, f2=None, nbins=32, kmin=0.0, kmax=None):
    """Returns (k, P, modes) for an auto spectrum or (k, P12, modes, r) for a
    cross spectrum, with P in (Mpc/h)^3 for fields normalised like r2c."""
    s = power_sums(f1, f2, nbins, kmin, kmax)
    V = float(np.prod(f1.pm.BoxSize))
    with np.errstate(invalid="ignore", divide="ignore"):
        k = s[:, 1] / s[:, 0]
        p = s[:, 2] / s[:, 0] * V
        if f2 is None:
            return k, p, s[:, 0]
        r = s[:, 2] / np.sqrt(s[:, 3] * s[:, 4])
    return k, p, s[:, 0], r


def write_power(field, path, a, cosmology, nbins=None):
    """The CLI's power file (fastpm/main.py:81-92): rows of k, P, modes and
    P / D1(a)^2, written with ``numpy.savetxt`` as the reference does.  Bins
    are linear in |k| from 0 to the Nyquist wave number; the default count,
    Nmesh / 2, makes the bin width the fundamental mode 2 pi / L (the bin
    edges of nbodykit's FFTPower cannot be checked here)."""
    pm = field.pm
    if nbins is None:
        nbins = max(1, int(pm.Nmesh.min()) // 2)
    k, p, n = power(field, nbins=nbins)
    d = float(cosmology.scale_independent_growth_factor(1.0 / a - 1))
    np.savetxt(path, np.array([k, p, n, p / d ** 2]).T, comments="# k p N p/D**2")


class PipelineMonitor:
    """The monitor of the reference CLI (fastpm/main.py:94-113) for
    ``Solver.nbody``: prints each action, writes ``power-<a>.txt`` from every
    Force's delta_k, and saves a snapshot ``fpm-<a>`` when the state is
    synchronized at an output epoch in ``aout``.  ``write_linear`` writes
    ``power-linear.txt`` (main.py:94)."""

    def __init__(self, solver, prefix, aout=(), attrs=None, nbins=None, verbose=True):
        self.solver = solver
        self.prefix = prefix
        self.aout = [float(a) for a in aout]
        self.attrs = {} if attrs is None else attrs
        self.nbins = nbins
        self.verbose = verbose
        os.makedirs(prefix, exist_ok=True)

    def path(self, name):
        return os.path.join(self.prefix, name)

    def write_linear(self, dlin):
        write_power(dlin, self.path("power-linear.txt"), 1.0, self.solver.cosmology, self.nbins)

    def __call__(self, action, ai, ac, af, state, event):
        if self.verbose:
            print("Step %s %06.4f - (%06.4f) -> %06.4f" % (action, ai, ac, af),
                  "S %(S)06.4f P %(P)06.4f F %(F)06.4f" % state.a)
        if action == "F":
            a = state.a["F"]
            write_power(event["delta_k"], self.path("power-%06.4f.txt" % a), a,
                        self.solver.cosmology, self.nbins)
        if state.synchronized:
            a = state.a["S"]
            if any(abs(a - b) <= 1e-12 * max(1.0, abs(b)) for b in self.aout):
                state.save(self.path("fpm-%06.4f" % a), attrs=self.attrs)
