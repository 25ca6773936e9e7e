"""Reverse-mode derivatives of the PM forward model (white noise -> linear ->
2LPT -> KDK n-body -> painted density), by hand-written adjoint kernels.

Reference semantics: fastpm/operators.py:68-98 (lpt1_gradient),
:140-193 (lpt2source_gradient), :218-253 (gravity_gradient), which chain the
pmesh gradient operators (readout_gradient, paint_gradient, c2r_gradient,
r2c_gradient) with conjugated kernels.  Here every step is a libfpm kernel
(`fpm_paint_weighted`, `fpm_window_grad`, `fpm_kspace_accumulate`,
`fpm_hermitian_scale`, `fpm_lpt2_source_vjp`, the transforms).  Unlike the
reference adjoints, which use the exact i k gradient, these differentiate
this build's forward operators (the 4-point gradient and the optional window
compensation), so the gradient is exact for the model that is run.

Conventions.  A complex half-spectrum field's gradient G holds
dL/dRe + i dL/dIm of each stored mode (stored modes are independent
variables).  With r2c normalised by 1/N and c2r unnormalised:
    r2c^T(G) = (1/N) c2r(w_half G),    c2r^T(g) = N w_c r2c(g),
with w_c = 2 (w_half = 1/2) for 0 < k_z < n2/2 and 1 on the k_z = 0, n2/2
planes.  A k-space multiplier's adjoint is its conjugate.

The time stepping is reversed action by action, recomputing each force from a
checkpoint of the displacement taken before it (per-step checkpointing):
    K: P += c F     ->  gF += c gP
    D: S += c P     ->  gP += c gS
    F: F = f(S)     ->  gS += f'(S)^T gF,  gF = 0
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .core import StateVector, leapfrog
from .pm import RealField, ComplexField, Particles, c2r_many, _stream
from .force.kernels import laplace, gradient, compensation, identity
from .background import MatterDominated


def _axpy(y, x, a):
    _lib.call("fpm_axpy", _np_dtype(y), y.data_ptr(), x.data_ptr(), y.numel(), float(a), _stream())


def _drift_kick_vjp(gS, gP, gF, drift, kicks, fresh):
    """gP += drift gS, then gF (zero if fresh) += sum_i kicks[i] gP in one pass."""
    kc = (ctypes.c_double * 2)(*(list(kicks) + [0.0] * (2 - len(kicks))))
    _lib.call("fpm_drift_kick_vjp", _np_dtype(gS), gS.data_ptr(), gP.data_ptr(), gF.data_ptr(),
              gS.numel(), kc, len(kicks), float(drift), int(fresh), _stream())


def _np_dtype(t):
    return np.float32 if t.dtype == torch.float32 else np.float64


def _kaccum(pm, acc, src, op, conj, accumulate):
    k = op.fpm_kop() if op is not None else identity.fpm_kop()
    _lib.call("fpm_kspace_accumulate", pm.dtype, acc._t.data_ptr(), src._t.data_ptr(), k,
              1 if conj else 0, 1 if accumulate else 0, pm._desc, pm.pitch2, _stream())


def _herm(pm, field, mid, edge):
    _lib.call("fpm_hermitian_scale", pm.dtype, field._t.data_ptr(), float(mid), float(edge),
              pm._desc, pm.pitch2, _stream())


def paint_weighted(pm, parts, weights, out=None, scale=1.0, resampler=None, accumulate=False):
    """Paint with per-particle mass scale * weights (a strided 1-D view)."""
    from .pm import resampler_order

    out = RealField(pm) if out is None else out
    _lib.call("fpm_paint_weighted", pm.dtype, parts.desc(pm.dtype), pm._desc,
              resampler_order(resampler or pm.resampler), weights.data_ptr(),
              int(weights.stride(0)), float(scale), out.value.data_ptr(),
              1 if accumulate else 0, _stream())
    return out


def window_grad(pm, parts, meshes, weights, gx, scale=1.0, accumulate=True, resampler=None):
    """gx (+)= scale * sum_c w_c * grad_x readout(mesh_c) (weights None -> 1)."""
    from .pm import resampler_order

    n = len(meshes)
    mp = _lib.ptr_array([m.value.data_ptr() for m in meshes])
    if weights is None:
        wp, ws = None, None
    else:
        wp = _lib.ptr_array([w.data_ptr() for w in weights])
        ws = (_lib.ctypes.c_int64 * n)(*[int(w.stride(0)) for w in weights])
    _lib.call("fpm_window_grad", pm.dtype, parts.desc(pm.dtype), pm._desc,
              resampler_order(resampler or pm.resampler), n, mp, wp, ws, float(scale),
              gx.data_ptr(), 1 if accumulate else 0, _stream())


def _table(pm, fn):
    """Device table of fn(k) at k = 2 pi sqrt(n^2) / L for integer n^2."""
    box = pm.BoxSize
    nmax = int(sum((int(n) // 2) ** 2 for n in pm.Nmesh))
    k = 2 * np.pi * np.sqrt(np.arange(nmax + 1, dtype="f8")) / box[0]
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.asarray(fn(k), dtype="f8")
    t = np.nan_to_num(t, nan=0.0, posinf=0.0, neginf=0.0)
    t[0] = 0.0
    return torch.from_numpy(t).to(pm.device)


def _scale_by_table(pm, field, tab):
    _lib.call("fpm_linear_scale", pm.dtype, field._t.data_ptr(), pm._desc, pm.pitch2,
              tab.data_ptr(), tab.numel(), _stream())


# --------------------------------------------------------------------------- force
def grad4_mesh(pm, phi, outs, scale=1.0):
    """outs[d] = scale * D_d phi, the four-point gradient as a stencil."""
    _lib.call("fpm_grad4_mesh", pm.dtype, phi.value.data_ptr(),
              _lib.ptr_array([o.value.data_ptr() for o in outs]), pm._desc, float(scale),
              _stream())


def grad4_div(pm, g, out, scale=1.0, accumulate=False):
    """out (+)= scale * sum_d D_d g[d]."""
    _lib.call("fpm_grad4_div", pm.dtype, _lib.ptr_array([m.value.data_ptr() for m in g]),
              out.value.data_ptr(), pm._desc, float(scale), 1 if accumulate else 0, _stream())


def force_vjp(step, S, Q, gF, gS):
    """gS += (d F / d S)^T gF for the PM force of FastPMStep.Force
    (gravity_gradient, fastpm/operators.py:218-253).

    F_d = readout(D_d phi) with phi = c2r(common r2c(rho)) and D_d the
    four-point gradient (a stencil, equal to the k-space kernel).  Its adjoint
    with respect to the mesh is g_phi = sum_d D_d^T paint(gF_d) =
    -sum_d D_d paint(gF_d), and since common is real and r2c^T c2r^T cancel
    their Hermitian weights, g_rho = c2r(common r2c(g_phi)): the forward's own
    fused transform.  The position gradient collects the window derivatives
    of the force readout and of the paint."""
    from .pm import r2c_c2r_many

    pm = step.pm
    parts = Particles(Q, S)
    nbar = 1.0 * len(Q) / float(np.prod(pm.Nmesh))
    common = step.force_common()
    tiled = pm.tiles_supported()
    if tiled:
        # float32 on tile-divisible meshes: the binned paints
        pm.bin_particles(parts)
        rho = pm.paint_tiled(parts, mass=1.0 / nbar)
    else:
        rho = pm.paint(parts, mass=1.0 / nbar)
    phi = RealField(pm)
    work = [pm._complex_tensor()]
    # the spectra themselves are not needed (kout=False)
    r2c_c2r_many(rho, [common], kout=False, outs=[phi], work=work)
    m = [RealField(pm) for _ in range(3)]
    # readout's position derivative
    if tiled and pm.grad4_supported():
        # stencil applied to the staged potential: the D_d phi meshes are
        # never formed
        pm.window_grad_tiled_grad4(phi, parts, gF, gS, 1.0, accumulate=True)
    else:
        grad4_mesh(pm, phi, m)
        window_grad(pm, parts, m, [gF[:, 0], gF[:, 1], gF[:, 2]], gS, 1.0, accumulate=True)
    # readout's mesh adjoint, the stencil's adjoint, then the multiplier back to rho
    if tiled:
        pm.paint_tiled_weighted(parts, [gF[:, 0], gF[:, 1], gF[:, 2]], m)
    else:
        for d in range(3):
            paint_weighted(pm, parts, gF[:, d], out=m[d])
    gphi = rho  # rho is no longer needed
    grad4_div(pm, m, gphi, scale=-1.0)
    grho = phi
    r2c_c2r_many(gphi, [common], kout=False, outs=[grho], work=work)
    # paint's position derivative (rho = paint(X) / nbar)
    window_grad(pm, parts, [grho], None, gS, 1.0 / nbar, accumulate=True)


# --------------------------------------------------------------------------- LPT
def lpt_vjp(solver, dlin, Q, a, gS, gP, gF, order=2):
    """Gradient with respect to the linear field dlin of the 2LPT state
    (core.py:120-145; lpt1_gradient and lpt2source_gradient of
    fastpm/operators.py).  Returns a ComplexField G (dL/dRe + i dL/dIm)."""
    pm = dlin.pm
    pt = MatterDominated(solver.cosmology.Om0, a=[a], a_normalize=solver.a_linear)
    D1, D2 = pt.D1(a), pt.D2(a)
    a2E = a ** 2 * pt.E(a)
    c = [D1, D2, a2E * pt.f1(a) * D1, a2E * pt.f2(a) * D2, a2E * pt.gf(a), a2E * pt.gf2(a)]
    g1 = torch.zeros_like(gS)
    _axpy(g1, gS, c[0])
    _axpy(g1, gP, c[2])
    _axpy(g1, gF, c[4])
    parts = Particles(Q, None)

    def lpt1_t(g):
        """sum_d conj(L_d) r2c(paint_Q(g_d)) with L_d = i a_d / k^2 (without
        the N w_c factor)."""
        acc = ComplexField(pm)
        for d in range(3):
            m = paint_weighted(pm, parts, g[:, d], resampler="cic")
            _kaccum(pm, acc, m.r2c(), laplace * gradient(d, 1), conj=True, accumulate=d > 0)
        return acc

    G = lpt1_t(g1)
    if order == 2:
        g2 = torch.zeros_like(gS)
        _axpy(g2, gS, c[1])
        _axpy(g2, gP, c[3])
        _axpy(g2, gF, c[5])
        H = lpt1_t(g2)
        gs = H.c2r()  # gradient with respect to the real-space source
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (1, 2), (2, 0)]
        ops = [laplace * gradient(i, 1) * gradient(j, 1) for i, j in pairs]
        phi = c2r_many(dlin, ops)
        gphi = [RealField(pm) for _ in range(6)]
        _lib.call("fpm_lpt2_source_vjp", pm.dtype, gs.value.data_ptr(),
                  _lib.ptr_array([f.value.data_ptr() for f in phi]),
                  _lib.ptr_array([f.value.data_ptr() for f in gphi]), gs.value.numel(), _stream())
        for op, gp in zip(ops, gphi):
            _kaccum(pm, G, gp.r2c(), op, conj=True, accumulate=True)
    # c2r^T = N w_c r2c
    N = float(np.prod(pm.Nmesh))
    _herm(pm, G, 2.0 * N, N)
    return G


# --------------------------------------------------------------------------- model
class ForwardModel:
    """w (complex white noise on the particle mesh) -> final StateVector, with
    the reverse pass.

    solver : fastpm.core.Solver
    Pk     : linear power callable
    stages : scale factors of the KDK steps (the first is the 2LPT epoch)
    """

    def __init__(self, solver, Pk, stages, order=2):
        self.solver = solver
        self.Pk = Pk
        self.stages = np.asarray(stages, dtype="f8")
        self.order = order
        self.Q = solver.pm.generate_uniform_particle_grid(shift=0.0)
        pm = solver.pm
        V = float(np.prod(pm.BoxSize))
        self._lin_tab = _table(pm, lambda k: np.sqrt(Pk(k) / V))

    def linear(self, w):
        out = w.copy()
        _scale_by_table(out.pm, out, self._lin_tab)
        return out

    def forward(self, w):
        """Runs the model; keeps the linear field and a displacement checkpoint
        before every force."""
        self.dlin = self.linear(w)
        state = self.solver.lpt(self.dlin, self.Q, a=self.stages[0], order=self.order)
        self.step = self.solver.nbodystep
        self.actions = list(leapfrog(self.stages))
        self.checkpoints = []
        pending = []
        for act in self.actions:
            if act[0] == "F":
                # the kicks and drift since the last Force in one pass (on the
                # tiled path the Force's binning pass); the Force leaves S as
                # the drift left it, so the checkpoint is taken after it
                self.step.run_actions(pending + [act], state)
                pending = []
                self.checkpoints.append(state.S.clone())
            else:
                pending.append(act)
        self.step.run_actions(pending, state)
        return state

    # the reverse pass applies a drift and the kicks before it (in the
    # forward order) as one fpm_drift_kick_vjp pass; False: one fpm_axpy per
    # action (the same result bit for bit)
    fuse_vjp = True

    def backward(self, gS, gP=None, gF=None):
        """Returns G_w, the gradient with respect to the white-noise modes,
        given the gradients with respect to the final S, P and F."""
        gS = gS.clone()
        gP = torch.zeros_like(gS) if gP is None else gP.clone()
        gF = torch.zeros_like(gS) if gF is None else gF.clone()
        pt = self.step._pt
        ck = list(self.checkpoints)
        acts = list(reversed(self.actions))
        fresh = False  # gF was consumed by a force VJP and not yet cleared
        i = 0
        while i < len(acts):
            action, ai, ac, af = acts[i]
            if action == "D" and self.fuse_vjp:
                kicks = []
                j = i + 1
                while j < len(acts) and acts[j][0] == "K" and len(kicks) < 2:
                    kicks.append(pt.kick_factor(*acts[j][1:]))
                    j += 1
                _drift_kick_vjp(gS, gP, gF, pt.drift_factor(ai, ac, af), kicks, fresh)
                fresh = fresh and not kicks
                i = j
                continue
            if fresh:
                gF.zero_()
                fresh = False
            if action == "K":
                _axpy(gF, gP, pt.kick_factor(ai, ac, af))
            elif action == "D":
                _axpy(gP, gS, pt.drift_factor(ai, ac, af))
            else:
                S = ck.pop()
                force_vjp(self.step, S, self.Q, gF, gS)
                fresh = True
            i += 1
        if fresh:
            gF.zero_()
        G = lpt_vjp(self.solver, self.dlin, self.Q, self.stages[0], gS, gP, gF, self.order)
        _scale_by_table(G.pm, G, self._lin_tab)
        return G


# --------------------------------------------------------------------------- loss
class DensityLoss:
    """L = sum over modes of |delta_k - t_k|^2 / P(k), with delta the painted
    overdensity of the final particles on the force mesh (CIC) and t the
    target's modes; the sum runs over the full spectrum (zero mode excluded)."""

    def __init__(self, solver, Pk, target_k):
        self.fpm = solver.fpm
        self.target = target_k
        self._inv_sqrt = _table(self.fpm, lambda k: 1.0 / np.sqrt(Pk(k)))
        N = float(np.prod(self.fpm.Nmesh))
        # dL/d rho = r2c^T(2 (delta - t) / P) = (1/N) c2r(2 (delta - t) / P)
        self._two_inv = _table(self.fpm, lambda k: 2.0 / (N * Pk(k)))

    def density_k(self, state):
        fpm = self.fpm
        nbar = 1.0 * state.csize / float(np.prod(fpm.Nmesh))
        parts = state.particles
        if fpm.tiles_supported():
            # binned fixed-point paint: deterministic, no float atomics
            fpm.bin_particles(parts, resampler="cic")
            rho = fpm.paint_tiled(parts, mass=1.0 / nbar, resampler="cic")
        else:
            rho = fpm.paint(parts, mass=1.0 / nbar, resampler="cic")
        return rho.r2c()

    def __call__(self, state, grad=True):
        fpm = self.fpm
        d = self.density_k(state)
        _kaccum(fpm, d, self.target, identity * -1.0, conj=False, accumulate=True)
        q = d.copy()
        _scale_by_table(fpm, q, self._inv_sqrt)
        from .utils import power_sums

        kmax = 10.0 * np.pi * float(fpm.Nmesh.max()) / float(fpm.BoxSize.min())
        L = float(power_sums(q, nbins=1, kmin=0.0, kmax=kmax)[0, 3])
        if not grad:
            return L, None
        _scale_by_table(fpm, d, self._two_inv)
        grho = d.c2r()
        nbar = 1.0 * state.csize / float(np.prod(fpm.Nmesh))
        gS = torch.zeros_like(state.S)
        window_grad(fpm, state.particles, [grho], None, gS, 1.0 / nbar, accumulate=False,
                    resampler="cic")
        return L, gS


class SimulateFunction(torch.autograd.Function):
    """torch.autograd entry point: white-noise modes (a complex tensor of the
    particle mesh's half spectrum, [n0, n1, n2/2+1]) -> final displacement S.
    backward runs the adjoint kernels."""

    @staticmethod
    def forward(ctx, w_modes, model):
        pm = model.solver.pm
        w = ComplexField(pm)
        w.value.copy_(w_modes)
        state = model.forward(w)
        ctx.model = model
        return state.S.clone()

    @staticmethod
    def backward(ctx, gS):
        G = ctx.model.backward(gS.contiguous())
        return G.value.clone(), None
