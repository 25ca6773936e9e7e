"""Binned power and cross spectra of device fields (estimator semantics of
fastpm/utils.py:77-164; output columns of fastpm/main.py:81-92).

Accumulation is in double precision on the GPU with a fixed reduction order.
The zero mode is excluded; half-spectrum modes with 0 < k_z < N/2 count
twice; bins are linear in |k| from kmin to kmax (default: 0 to the Nyquist
wave number of the mesh)."""
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
