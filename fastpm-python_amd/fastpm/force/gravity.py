"""Particle-mesh long-range force (reference: fastpm/force/gravity.py:5-29)."""
import torch

from ..pm import c2r_many
from .kernels import laplace, gradient, longrange as k_longrange, compensation


def force_ops(split, factor, deconv=0):
    """The three fused force multipliers
    factor * i a_d(k) / k^2 * exp(-k^2 r^2) [/ W(k)^p]  (gravity.py:17-27)."""
    ops = []
    for d in range(3):
        op = laplace * k_longrange(split) * gradient(d, order=1) * factor
        if deconv:
            op = op * compensation(deconv)
        ops.append(op)
    return ops


def longrange(x, delta_k, split, factor, deconv=0):
    """Force at positions x from delta_k, factor = 3/2 Omega_m when delta_k is
    the transform of 1 + delta (gravity.py:5-9)."""
    return longrange_batch([x], delta_k, split, factor, deconv)[0]


def longrange_batch(x, delta_k, split, factor, deconv=0):
    """Like longrange, for a list of position sets (gravity.py:11-29)."""
    pm = delta_k.pm
    meshes = c2r_many(delta_k, force_ops(split, factor, deconv))
    out = []
    for xi in x:
        parts = pm._particles(xi)
        f = torch.empty((len(parts), 3), dtype=pm.tdtype, device=pm.device)
        pm.readout_many(meshes, parts, [f[:, 0], f[:, 1], f[:, 2]])
        out.append(f)
    return out
