"""FastPM solver API on one MI355X (reference: fastpm/core.py).

``Solver``, ``StateVector``, ``FastPMStep``, ``leapfrog``,
``leapfrog_with_lpt`` and ``autostages`` keep the reference's names, argument
meanings, ownership rules and errors; the arrays of a state are device tensors
and every particle or mesh operation runs in libfpm's HIP kernels.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .background import MatterDominated
from .pm import (ParticleMesh, Particles, RealField, ComplexField, c2r_many, c2r_grad3,
                 r2c_c2r_many, _stream)
from .force.kernels import laplace, gradient, compensation


class StateVector:
    """Particle state: Q (Lagrangian lattice or explicit positions), displacement
    S, momentum P, force F ([N, 3] device tensors) and RHO ([N]); time tags in
    ``a`` (core.py:7-89)."""

    def __init__(self, solver, Q):
        self.solver = solver
        self.pm = solver.pm
        self.Q = Q
        self.csize = solver.pm.comm.allreduce(len(self.Q))
        self.dtype = solver.pm.dtype
        self.cosmology = solver.cosmology
        self.H0 = 100.0  # km/s / (Mpc/h)
        self.GM0 = (self.H0 ** 2 / (4 * np.pi) * 1.5 * self.cosmology.Om0
                    * float(np.prod(self.pm.BoxSize)) / self.csize)
        n = len(Q)
        dev, dt = self.pm.device, self.pm.tdtype
        self.S = torch.zeros((n, 3), dtype=dt, device=dev)
        self.P = torch.zeros((n, 3), dtype=dt, device=dev)
        self.F = torch.zeros((n, 3), dtype=dt, device=dev)
        self.RHO = torch.zeros((n,), dtype=dt, device=dev)
        self.a = dict(S=None, P=None, F=None)

    def copy(self):
        """Deep copy of S, P, F, RHO; Q is shared (core.py:26-34)."""
        obj = object.__new__(type(self))
        obj.__dict__.update(self.__dict__)
        obj.S = self.S.clone()
        obj.P = self.P.clone()
        obj.F = self.F.clone()
        obj.RHO = self.RHO.clone()
        obj.a = dict(self.a)
        return obj

    @property
    def synchronized(self):
        a = self.a["S"]
        return a == self.a["P"] and a == self.a["F"]

    @property
    def particles(self):
        return Particles(self.Q, self.S)

    @property
    def X(self):
        """Positions Q + S (a new tensor)."""
        return self.pm.positions(self.particles)

    @property
    def V(self):
        """Peculiar velocity P H0 / a_P (a new tensor)."""
        return self.P * (self.H0 / self.a["P"])

    def to_mesh(self):
        """Paint X on the particle mesh with unit mass (core.py:50-55)."""
        return self.pm.paint(self.particles, hold=False)

    def save(self, filename, attrs={}):
        """Snapshot in the bigfile directory layout (core.py:65-89): a
        ``Header`` block with the attributes Om0, Tcmb0, Neff, [m_nu,] Ob0,
        Ode0, Time, h, RSDFactor and `attrs` (those that encode), and the
        blocks ``1/Position``, ``1/Velocity`` (km/s) and ``1/Density``.  For
        restarts (``StateVector.load``) it also writes ``1/Displacement``,
        ``1/Momentum``, ``1/Force`` and the time tags ``a_S``, ``a_P``,
        ``a_F`` as Header attributes (blocks and attributes the reference
        does not write)."""
        from .snapshot import write_snapshot, _encode_attr

        a = self.a["S"]
        header = {}
        keylist = ["Om0", "Tcmb0", "Neff", "Ob0", "Ode0"]
        if getattr(self.cosmology, "m_nu", None) is not None:
            keylist.insert(3, "m_nu")
        for key in keylist:
            header[key] = getattr(self.cosmology, key)
        header["Time"] = a
        header["h"] = self.cosmology.H0 / self.H0  # relative h
        header["RSDFactor"] = 1.0 / (self.H0 * a * self.cosmology.efunc(1.0 / a - 1))
        for key in attrs:
            try:  # best effort, as in the reference
                _encode_attr(attrs[key])
                header[key] = attrs[key]
            except (TypeError, ValueError):
                pass
        for key in ("S", "P", "F"):
            if self.a[key] is not None:
                header["a_" + key] = self.a[key]
        blocks = {"1/Position": self.X, "1/Velocity": self.V, "1/Density": self.RHO,
                  "1/Displacement": self.S, "1/Momentum": self.P, "1/Force": self.F}
        write_snapshot(filename, header, {k: v.cpu().numpy() for k, v in blocks.items()})

    @classmethod
    def load(cls, solver, filename, Q):
        """Restores a state written by ``save`` (the reference has no loader):
        S, P, F, RHO and the time tags, on the Lagrangian positions Q.  Every
        block's file sizes and byte sums are checked against its header."""
        from .snapshot import read_snapshot

        names = {"S": "1/Displacement", "P": "1/Momentum", "F": "1/Force", "RHO": "1/Density"}
        header, arrs = read_snapshot(filename, list(names.values()))
        state = cls(solver, Q)
        for attr, name in names.items():
            arr = arrs[name]
            t = getattr(state, attr)
            if tuple(arr.shape) != tuple(t.shape):
                raise ValueError("%s has shape %s, expected %s" % (name, arr.shape, tuple(t.shape)))
            t.copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(dtype=t.dtype))
        state.a = {key: (float(header["a_" + key][0]) if "a_" + key in header else None)
                   for key in ("S", "P", "F")}
        return state


class Solver:
    """PM solver (core.py:91-152).

    pm : ParticleMesh of the particle lattice; the force mesh has Nmesh * B.
    cosmology : any object with ``Om0`` (``efunc``/``H0``/... for ``save``).
    a_linear : scale factor at which the linear field has D1 = 1.
    compensate : divide the force by the mass-assignment window (off in the
        reference; used by the window-compensated benchmark configurations).
    fused : True runs the force as one multi-output inverse transform and one
        four-component readout (on float32 tile-divisible meshes: binned
        paint, one transform of the potential and a stencil readout) and,
        in nbody without a monitor, each step boundary's kicks and drift as
        one pass; False runs the plain sequence of the reference (one inverse
        transform and one readout per component, separate updates).
    autotune : choose the force mesh's kernel variants by timing them
        (fastpm.tune): the transform variants here, the binning, paint and
        readout variants on the particles of the first force.
    """

    def __init__(self, pm, cosmology, B=1, a_linear=1.0, compensate=False, fused=True,
                 autotune=False):
        if not hasattr(cosmology, "Om0"):
            raise TypeError("cosmology must provide Om0")
        fpm = ParticleMesh(Nmesh=pm.Nmesh * B, BoxSize=pm.BoxSize, dtype=pm.dtype, comm=pm.comm,
                           resampler=pm.resampler, device=pm.device)
        self.pm = pm
        self.fpm = fpm
        self.B = B
        self.cosmology = cosmology
        self.a_linear = a_linear
        self.compensate = bool(compensate)
        self.fused = bool(fused)
        self.autotune = bool(autotune)
        self._particles_tuned = False
        if autotune:
            # transform variants of the force mesh (fastpm.tune)
            from .tune import autotune_fft

            autotune_fft(fpm)

    @property
    def nbodystep(self):
        return FastPMStep(self)

    def whitenoise(self, seed, unitary=False):
        return self.pm.generate_whitenoise(seed, type="complex", unitary=unitary)

    def linear(self, whitenoise, Pk):
        """delta_k = sqrt(P(|k|)/V) w_k (core.py:116-118).  P, a host callable,
        is evaluated once per distinct |k| of the mesh and applied by table
        lookup: on a cubic box the table is indexed by the integer |n|^2; when
        the sides differ the distinct k^2 are found on the device (sorted,
        with each mode's index into them)."""
        pm = whitenoise.pm
        box = pm.BoxSize
        out = whitenoise.copy()
        if np.all(box == box[0]):
            nmax = int(sum((int(n) // 2) ** 2 for n in pm.Nmesh))
            k = 2 * np.pi * np.sqrt(np.arange(nmax + 1, dtype="f8")) / box[0]
            tab_d = torch.from_numpy(_linear_table(Pk, k, box)).to(pm.device)
            _lib.call("fpm_linear_scale", pm.dtype, out._t.data_ptr(), pm._desc, pm.pitch2,
                      tab_d.data_ptr(), len(tab_d), _stream())
            out._table = tab_d  # the table lives as long as the field
            return out
        kk = None
        for d, k1 in enumerate(pm.k_axes()):
            shape = [1, 1, 1]
            shape[d] = len(k1)
            t = torch.from_numpy(k1 ** 2).to(pm.device).reshape(shape)
            kk = t if kk is None else kk + t
        uniq, index = torch.unique(kk.reshape(-1), sorted=True, return_inverse=True)
        del kk
        k = np.sqrt(uniq.cpu().numpy())
        tab_d = torch.from_numpy(_linear_table(Pk, k, box)).to(pm.device)
        index = index.to(torch.int32).contiguous()
        _lib.call("fpm_linear_scale_indexed", pm.dtype, out._t.data_ptr(), pm._desc, pm.pitch2,
                  tab_d.data_ptr(), len(tab_d), index.data_ptr(), _stream())
        out._table = (tab_d, index)
        return out

    def lpt(self, linear, Q, a, order=2):
        """2LPT initial state at a (core.py:120-145)."""
        assert order in (1, 2)
        from .force.lpt import lpt1, lpt2source

        state = StateVector(self, Q)
        pt = MatterDominated(self.cosmology.Om0, a=[a], a_normalize=self.a_linear)
        D1, D2 = pt.D1(a), pt.D2(a)
        a2E = a ** 2 * pt.E(a)
        dx1 = lpt1(linear, Q)
        if order == 2:
            dx2 = lpt1(lpt2source(linear), Q)
            coef = [D1, D2, a2E * pt.f1(a) * D1, a2E * pt.f2(a) * D2, a2E * pt.gf(a),
                    a2E * pt.gf2(a)]
        else:
            dx2 = None
            coef = [D1, 0.0, a2E * pt.f1(a) * D1, 0.0, a2E * pt.gf(a), 0.0]
        c = np.array(coef, dtype="f8")
        _lib.call("fpm_lpt_combine", self.pm.dtype, dx1.data_ptr(),
                  None if dx2 is None else dx2.data_ptr(), state.S.data_ptr(),
                  state.P.data_ptr(), state.F.data_ptr(), state.S.numel(), c.ctypes.data,
                  _stream())
        state.a["S"] = a
        state.a["P"] = a
        return state

    def nbody(self, state, stepping, monitor=None):
        """Apply an action list; mutates and returns ``state`` (core.py:147-152)."""
        self.nbodystep.run_actions(stepping, state, monitor)
        return state


def _linear_table(Pk, k, box):
    """sqrt(P(k) / V) at the wave numbers k (0 at k = 0 and where P is not
    finite)."""
    try:
        pk = np.asarray(Pk(k), dtype="f8")
    except Exception:
        pk = np.array([float(Pk(x)) for x in k])
    with np.errstate(invalid="ignore"):
        tab = np.sqrt(pk / float(np.prod(box)))
    tab[k == 0] = 0.0
    return np.ascontiguousarray(np.nan_to_num(tab, nan=0.0, posinf=0.0))


class FastPMStep:
    """Kick, drift and PM force (core.py:155-210).  Mesh buffers are allocated
    once per step object and reused across forces."""

    def __init__(self, solver):
        self.cosmology = solver.cosmology
        self.pm = solver.fpm
        self.solver = solver
        self._ws = None
        self._pt = MatterDominated(self.cosmology.Om0, a_normalize=solver.a_linear)
        self.timer = None  # optional KernelTimer recording each library call
        # fused float32 path on tile-divisible meshes: binned deterministic
        # paint and LDS-staged readout; the force's four-point gradient is
        # applied as a real-space stencil inside the readout (one inverse
        # transform of the potential instead of three)
        self.tiled = solver.fused and self.pm.tiles_supported()
        self.stencil = self.tiled and self.pm.grad4_supported()
        # whether Force stores delta_k; run() clears it for a Force whose
        # event no monitor receives
        self.keep_delta_k = True
        # a step boundary's kicks and drift (P, F, kicks, drift) that the
        # next Force's binning applies in its own pass (run_actions sets it)
        self._pending_kd = None

    def _mark(self, name):
        return self.timer.region(name) if self.timer is not None else _NullRegion()

    def run(self, action, ai, ac, af, state, monitor):
        actions = dict(K=self.Kick, D=self.Drift, F=self.Force)
        fn = actions[action]
        # without a monitor the Force's delta_k event is never seen, so the
        # fused transform need not store the spectrum
        self.keep_delta_k = monitor is not None
        try:
            event = fn(state, ai, ac, af)
        finally:
            self.keep_delta_k = True
        if monitor is not None:
            monitor(action, ai, ac, af, state, event)

    def run_actions(self, actions, state, monitor=None):
        """run() over an action list.  Without a monitor, on a fused solver and
        when Kick and Drift are this class's own, each run of Kick actions followed by a
        Drift (K, D at the first step; K, K, D at every later step boundary)
        is one fused pass (fpm_kick_drift) that gives the separate updates'
        result bit for bit; when a Force follows on the tiled path, that
        pass is the Force's binning pass (fpm_kick_drift_rebin, the same
        result again).  Every other action, and any sequence whose time
        tags would fail an assertion, goes through run()."""
        actions = list(actions)
        fuse = (monitor is None and self.solver.fused and type(self).Kick is FastPMStep.Kick
                and type(self).Drift is FastPMStep.Drift)
        fuse_force = fuse and self.tiled and type(self).Force is FastPMStep.Force
        i = 0
        while i < len(actions):
            j = i
            while fuse and j < len(actions) and j - i < 2 and actions[j][0] == "K":
                j += 1
            if j > i and j < len(actions) and actions[j][0] == "D":
                k = self._kick_drift_factors(state, actions[i:j], actions[j])
                if k is not None:
                    facs, fac, aP, aS = k
                    force = actions[j + 1] if j + 1 < len(actions) else None
                    if fuse_force and force is not None and force[0] == "F" and force[2] == aS:
                        state.a["P"] = aP
                        state.a["S"] = aS
                        self._pending_kd = (state.P, state.F, facs, fac)
                        try:
                            self.run(*force, state, monitor)
                        finally:
                            self._pending_kd = None
                        i = j + 2
                        continue
                    self._kick_drift(state, facs, fac)
                    state.a["P"] = aP
                    state.a["S"] = aS
                    i = j + 1
                    continue
            self.run(*actions[i], state, monitor)
            i += 1

    def _kick_drift_factors(self, state, kicks, drift):
        """(kick factors, drift factor, new P tag, new S tag) of Kick(s) +
        Drift, or None if a time tag would fail one of their assertions."""
        aF, aP = state.a["F"], state.a["P"]
        facs = []
        for _, ai, ac, af in kicks:
            if ac != aF:
                return None
            facs.append(self._pt.kick_factor(ai, ac, af))
            aP = af
        _, ai, ac, af = drift
        if ac != aP:
            return None
        return facs, self._pt.drift_factor(ai, ac, af), aP, af

    def _kick_drift(self, state, facs, fac):
        """Fused Kick(s) + Drift in one pass (fpm_kick_drift)."""
        kc = (ctypes.c_double * 2)(*(facs + [0.0] * (2 - len(facs))))
        with self._mark("kick_drift"):
            _lib.call("fpm_kick_drift", state.dtype, state.S.data_ptr(), state.P.data_ptr(),
                      state.F.data_ptr(), state.S.numel(), kc, len(facs), float(fac), _stream())

    def Kick(self, state, ai, ac, af):
        assert ac == state.a["F"]
        fac = self._pt.kick_factor(ai, ac, af)
        with self._mark("kick"):
            _lib.call("fpm_axpy", state.dtype, state.P.data_ptr(), state.F.data_ptr(),
                      state.P.numel(), float(fac), _stream())
        state.a["P"] = af

    def Drift(self, state, ai, ac, af):
        assert ac == state.a["P"]
        fac = self._pt.drift_factor(ai, ac, af)
        with self._mark("drift"):
            _lib.call("fpm_axpy", state.dtype, state.S.data_ptr(), state.P.data_ptr(),
                      state.S.numel(), float(fac), _stream())
        state.a["S"] = af

    def workspace(self):
        """Force-mesh buffers, allocated once: the stencil path needs the
        density, one real and one complex mesh; the k-space path three of
        each and the spectrum (reused by forces whose delta_k no monitor
        receives; a monitor gets a new field per force, see force_kernels)."""
        if self._ws is None:
            pm = self.pm
            n = 1 if self.stencil else 3
            self._ws = dict(
                rho=RealField(pm),
                delta_k=None if self.stencil else ComplexField(pm),
                force=[RealField(pm) for _ in range(n)],
                work=[pm._complex_tensor() for _ in range(n)],
            )
        return self._ws

    def delta_k_buffer(self):
        ws = self.workspace()
        if ws["delta_k"] is None:
            ws["delta_k"] = ComplexField(self.pm)
        return ws["delta_k"]

    def prepare_force(self, state, smoothing):
        """Paint 1 + delta of the particles on the force mesh (core.py:182-193)."""
        nbar = 1.0 * state.csize / float(np.prod(self.pm.Nmesh))
        X1 = state.particles
        layout = self.pm.decompose(X1, smoothing)
        rho = self.workspace()["rho"]
        if self.tiled and self.solver.autotune and not self.solver._particles_tuned:
            # binning, paint and readout variants on these particles (fastpm.tune)
            from .tune import autotune_particles

            autotune_particles(self.pm, X1)
            self.solver._particles_tuned = True
        if self.tiled:
            kd, self._pending_kd = self._pending_kd, None
            with self._mark("bin" if kd is None else "kick_drift_bin"):
                # one pass into the previous Force's capacity layout (the
                # particles moved little since), with the step boundary's
                # kicks and drift when run_actions handed them over
                self.pm.bin_particles(X1, rebin=True, kick_drift=kd)
            with self._mark("paint"):
                self.pm.paint_tiled(X1, mass=1.0 / nbar, out=rho)
        else:
            with self._mark("paint"):
                self.pm.paint(X1, mass=1.0 / nbar, out=rho)
        return layout, X1, rho

    def force_common(self):
        """The factor shared by the three force components:
        3/2 Om0 / k^2 [/ W(k)^p]."""
        op = laplace * (1.5 * self.cosmology.Om0)
        if self.solver.compensate:
            op = op * compensation(self.pm.order)
        return op

    def force_ops(self):
        factor = 1.5 * self.cosmology.Om0
        ops = []
        for d in range(3):
            op = laplace * gradient(d, order=1) * factor
            if self.solver.compensate:
                op = op * compensation(self.pm.order)
            ops.append(op)
        return ops

    def Force(self, state, ai, ac, af):
        """RHO and F at the particles (core.py:195-210): paint, forward
        transform, then either (float32 tiled) one inverse transform of the
        potential and a readout applying the four-point gradient stencil, or
        three fused multiplier + inverse transforms and one fused
        four-component readout."""
        assert ac == state.a["S"]
        delta_k = self.force_kernels(state, keep_delta_k=self.keep_delta_k)
        state.a["F"] = af
        return dict(delta_k=delta_k)

    def force_kernels(self, state, keep_delta_k=True):
        """Enqueues the force computation (no host-side bookkeeping, so it can
        be captured in a graph).  With keep_delta_k the spectrum goes to a new
        ComplexField, returned, which a monitor may keep (the reference makes
        a new field per Force, core.py:205); otherwise it is not stored on the
        stencil path (None is returned) and goes to the reused workspace
        buffer on the k-space path."""
        ws = self.workspace()
        layout, X1, rho = self.prepare_force(state, smoothing=None)
        F = state.F
        if self.stencil:
            # delta_k = r2c(rho) and phi = c2r(common * delta_k), sharing the x pass
            phi = ws["force"][0]
            with self._mark("fft"):
                kout = ComplexField(self.pm) if keep_delta_k else False
                delta_k, _ = r2c_c2r_many(rho, [self.force_common()], kout=kout,
                                          outs=[phi], work=ws["work"][:1])
            with self._mark("readout"):
                self.pm.readout_tiled_grad4(phi, X1, F, rho=rho, rho_out=state.RHO)
            return delta_k
        with self._mark("r2c"):
            delta_k = rho.r2c(out=ComplexField(self.pm) if keep_delta_k
                              else self.delta_k_buffer())
        ops = self.force_ops()
        if self.solver.fused:
            with self._mark("c2r"):
                c2r_grad3(delta_k, self.force_common(), outs=ws["force"], work=ws["work"])
            with self._mark("readout"):
                readout = self.pm.readout_tiled if self.tiled else self.pm.readout_many
                readout([rho] + ws["force"], X1, [state.RHO, F[:, 0], F[:, 1], F[:, 2]])
        else:
            with self._mark("readout"):
                self.pm.readout_many([rho], X1, [state.RHO])
            for d in range(3):
                with self._mark("c2r"):
                    c2r_many(delta_k, [ops[d]], outs=ws["force"][d:d + 1], work=ws["work"][d:d + 1])
                with self._mark("readout"):
                    self.pm.readout_many([ws["force"][d]], X1, [F[:, d]])
        return delta_k


class _NullRegion:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class KernelTimer:
    """Records HIP events around library calls on the current stream (the
    stream every libfpm call is enqueued on) and reports per-region averages."""

    def __init__(self):
        self.events = {}

    def region(self, name):
        timer = self

        class _R:
            def __enter__(self_):
                self_.a = torch.cuda.Event(enable_timing=True)
                self_.a.record()
                return self_

            def __exit__(self_, *exc):
                b = torch.cuda.Event(enable_timing=True)
                b.record()
                timer.events.setdefault(name, []).append((self_.a, b))
                return False

        return _R()

    def summary(self):
        """name -> (count, mean ms); synchronises the device."""
        torch.cuda.synchronize()
        return {k: (len(v), float(np.mean([a.elapsed_time(b) for a, b in v])))
                for k, v in self.events.items()}


def autostages(knots, N, astart=None, N0=None):
    """N stages that include every knot (core.py:212-259)."""
    knots = np.array(knots, dtype="f8")
    knots.sort()
    stages = np.array([], dtype="f8")
    if astart is not None and astart != knots.min():
        assert astart < knots.min()
        if N0 is None:
            N0 = 1
        knots = np.append([astart], knots)
    else:
        N0 = 1
    for i in range(0, len(knots) - 1):
        da = (knots[-1] - knots[i]) / (N - len(stages) - 1)
        n_span = int((knots[i + 1] - knots[i]) / da + 0.5)
        if i == 0 and n_span < N0:
            n_span = N0
        add = np.linspace(knots[i], knots[i + 1], n_span, endpoint=False)
        stages = np.append(stages, add)
    return np.append(stages, [knots[-1]])


def leapfrog(stages):
    """KDK action list: F at the first stage, then K, D, F, K per step
    (core.py:261-288)."""
    if len(stages) == 0:
        return
    ai = stages[0]
    yield "F", ai, ai, ai
    x, p, f = ai, ai, ai
    for i in range(len(stages) - 1):
        a0 = stages[i]
        a1 = stages[i + 1]
        ah = (a0 * a1) ** 0.5
        yield "K", p, f, ah
        p = ah
        yield "D", x, p, a1
        x = a1
        yield "F", f, x, a1
        f = a1
        yield "K", p, f, a1
        p = a1


def leapfrog_with_lpt(stages):
    """As leapfrog, without the first force (the LPT state carries it)
    (core.py:290-317)."""
    if len(stages) == 0:
        return
    ai = stages[0]
    x, p, f = ai, ai, ai
    for i in range(len(stages) - 1):
        a0 = stages[i]
        a1 = stages[i + 1]
        ah = (a0 * a1) ** 0.5
        yield "K", p, f, ah
        p = ah
        yield "D", x, p, a1
        x = a1
        yield "F", f, x, a1
        f = a1
        yield "K", p, f, a1
        p = a1
