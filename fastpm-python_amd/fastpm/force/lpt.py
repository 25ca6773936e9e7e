"""Lagrangian perturbation theory on the device (reference:
fastpm/force/lpt.py:4-64)."""
import torch

from ..pm import c2r_many, c2r_grad3, RealField, _stream
from .. import _lib
from .kernels import laplace, gradient


def lpt1(dlin_k, q, resampler="cic"):
    """First-order displacement -grad laplace^-1 delta read out at q with CIC
    (lpt.py:4-23).  The three inverse transforms share one read of dlin_k and
    the three readouts are one fused gather.  Returns a [N, 3] device tensor in
    the mesh dtype."""
    pm = dlin_k.pm
    meshes = c2r_grad3(dlin_k, laplace, grad_order=1)
    parts = pm._particles(q)
    out = torch.empty((len(parts), 3), dtype=pm.tdtype, device=pm.device)
    pm.readout_many(meshes, parts, [out[:, 0], out[:, 1], out[:, 2]], resampler)
    return out


_PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]


def lpt2source(dlin_k):
    """Second-order source 3/7 (sum_i<j phi_ii phi_jj - phi_ij^2), returned in
    k-space (lpt.py:25-64).  The six second derivatives come from one
    multi-output inverse transform; the product is one elementwise kernel."""
    pm = dlin_k.pm
    ops = [laplace * gradient(i, 1) * gradient(j, 1) for i, j in _PAIRS]
    phi = c2r_many(dlin_k, ops)
    source = RealField(pm)
    ptrs = _lib.ptr_array([p.value.data_ptr() for p in phi])
    _lib.call("fpm_lpt2_source", pm.dtype, ptrs, source.value.data_ptr(),
              source.value.numel(), _stream())
    del phi
    return source.r2c()
