"""Many independent boxes advanced together, one captured HIP graph per step.

Each box has its own particle mesh, force mesh (Nmesh * B), cosmology
(sigma8 may differ) and white-noise seed; all share the stage list.  The
kick/drift coefficients of every step are precomputed in FP64 into a device
table per box and read through a device step counter, so one step of all
boxes is captured once as a graph and replayed for every step.  The replayed
unit is a step boundary, as in Solver.nbody: the previous step's closing
kick, this step's opening kick and drift in one pass, then the force; the
last closing kick follows the final replay.

With `batched` (the default when every box takes the float32 stencil force)
the boxes are split into `nplans` groups of about equal mesh volume, and one
step of a group is a handful of launches (`fpm_batch_step`): one per
compiled kernel, each covering every box of the group that uses it, one
workgroup per tile, span or row group of each box, through a device table of
per-box arguments built once (`fpm_batch_create`).  The groups run on
parallel streams (branches of the captured graph), so one group's
latency-bound passes overlap another's bandwidth-bound ones.  Otherwise each
box enqueues its own kernels on one of `nstreams` streams.  Either way the per-box arithmetic
is the single-box Solver's (the tiled, deterministic paint, the fused
transform, the stencil readout), so a replayed ensemble reproduces the
per-box eager runs bit for bit.

The reference has no ensemble driver; this is the batched configuration of
the benchmark (64 boxes of {32, 48, 64, 96, 128}^3 particles, 20 steps).
"""
import numpy as np
import torch

from . import _lib
from .core import Solver, leapfrog
from .pm import ParticleMesh, _stream
from .cosmology import LinearPower


class Box:
    def __init__(self, nc, box, cosmo, seed, dtype="f4", B=2, resampler="cic", compensate=True):
        self.nc = int(nc)
        self.seed = int(seed)
        self.cosmo = cosmo
        self.pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=box, dtype=dtype, resampler=resampler)
        self.solver = Solver(self.pm, cosmo, B=B, compensate=compensate)
        self.step = self.solver.nbodystep
        self.state = None
        self.table = None
        self.rows = None


class Ensemble:
    def __init__(self, specs, stages, dtype="f4", B=2, resampler="cic", compensate=True,
                 nstreams=16, autotune=True, batched=True, nplans=2):
        """specs: iterable of dict(nc=, box=, cosmo=, seed=).  Boxes are spread
        over `nstreams` streams (largest first, each to the least loaded
        stream) so small boxes run concurrently; inside a captured graph the
        streams become parallel branches (64 boxes, 20 steps: 8.49 / 8.21 /
        8.12 / 8.02 ms per step with 2 / 4 / 8 / 16 streams).  autotune selects the kernel
        variants per force-mesh shape before the first step (fastpm.tune)."""
        self.stages = np.asarray(stages, dtype="f8")
        self.autotune = bool(autotune)
        self.batched = bool(batched)
        self.nplans = max(1, int(nplans))
        self.plans = []
        self.boxes = [Box(s["nc"], s["box"], s["cosmo"], s["seed"], dtype, B, resampler,
                          compensate) for s in specs]
        self.nsteps = len(self.stages) - 1
        dev = self.boxes[0].pm.device
        self.counter = torch.zeros(1, dtype=torch.int32, device=dev)
        self.graph = None
        self.nstreams = max(1, int(nstreams))
        self.streams = [torch.cuda.Stream(device=dev) for _ in range(self.nstreams)]
        load = [0] * self.nstreams
        self.lanes = [[] for _ in range(self.nstreams)]
        for b in sorted(self.boxes, key=lambda b: -b.nc ** 3):
            i = int(np.argmin(load))
            self.lanes[i].append(b)
            load[i] += b.solver.fpm.Nmesh.prod()

    @property
    def npart(self):
        return sum(b.nc ** 3 for b in self.boxes)

    def initialize(self, Pk_cache=None):
        """White noise, linear field, 2LPT at stages[0] and the first force of
        every box; builds the per-box coefficient tables."""
        Pk_cache = {} if Pk_cache is None else Pk_cache
        a0 = self.stages[0]
        if self.autotune:
            from .tune import autotune_fft, autotune_particles
        for b in self.boxes:
            key = (b.cosmo.sigma8, b.cosmo.Om0, b.cosmo.h, b.cosmo.n_s)
            if key not in Pk_cache:
                Pk_cache[key] = LinearPower(b.cosmo)
            dlin = b.solver.linear(b.solver.whitenoise(b.seed), Pk_cache[key])
            b.state = b.solver.lpt(dlin, b.pm.generate_uniform_particle_grid(), a=a0)
            if self.autotune:
                # kernel variants per force-mesh shape and lattice (fastpm.tune):
                # timed on the first box of each, cached for the others
                autotune_fft(b.solver.fpm)
                autotune_particles(b.solver.fpm, b.state.particles)
            b.step.run("F", a0, a0, a0, b.state, None)
            pt = b.step._pt
            rows = []
            for i in range(self.nsteps):
                s0, s1 = self.stages[i], self.stages[i + 1]
                ah = (s0 * s1) ** 0.5
                rows.append([pt.kick_factor(s0, s0, ah), pt.drift_factor(s0, ah, s1),
                             pt.kick_factor(ah, s1, s1)])
            b.rows = rows
            b.table = torch.tensor(np.array(rows, dtype="f8"), device=self.counter.device)
        self.counter.zero_()
        if self.batched and self.nsteps > 0 and all(b.step.stencil for b in self.boxes):
            # groups of about equal mesh volume, largest boxes first
            k = min(self.nplans, len(self.boxes), len(self.streams))
            groups, load = [[] for _ in range(k)], [0] * k
            for b in sorted(self.boxes, key=lambda b: -int(b.solver.fpm.Nmesh.prod())):
                i = int(np.argmin(load))
                groups[i].append(b)
                load[i] += int(b.solver.fpm.Nmesh.prod())
            self.plans = [self._build_plan(g) for g in groups if g]

    @property
    def plan(self):
        """The first batch plan (None when the boxes run one by one)."""
        return self.plans[0][0] if self.plans else None

    def _build_plan(self, group):
        """A batch plan of the boxes in `group`: every box's buffers, force
        multiplier and coefficient table in one fpm_batch_box array; returns
        (plan, its device table memory)."""
        import ctypes

        n = len(group)
        boxes = (_lib.FpmBatchBox * n)()
        for i, b in enumerate(group):
            s, step, fpm = b.state, b.step, b.solver.fpm
            ws = step.workspace()
            parts = s.particles
            bx = boxes[i]
            bx.part = parts.desc(fpm.dtype)
            bx.mesh = fpm._desc
            bx.order = fpm.order
            # the paint mass of FastPMStep.prepare_force, the same expression
            nbar = 1.0 * s.csize / float(np.prod(fpm.Nmesh))
            bx.mass = 1.0 / nbar
            bx.common = step.force_common().fpm_kop()
            bx.pitch2 = fpm.pitch2
            bx.rho = ws["rho"].value.data_ptr()
            bx.phi = ws["force"][0].value.data_ptr()
            bx.work = ws["work"][0].data_ptr()
            for d, t in enumerate(fpm.twiddles()):
                bx.tw[d] = t.data_ptr()
            tws = fpm.tile_workspace(len(parts), fpm.order)
            bx.ws, bx.ws_bytes = tws.data_ptr(), tws.numel()
            bx.rho_out, bx.rho_stride = s.RHO.data_ptr(), 1
            bx.force_out, bx.force_stride = s.F.data_ptr(), 3
            bx.rho_scale = bx.force_scale = 1.0
            bx.P = s.P.data_ptr()
            bx.table = b.table.data_ptr()
            bx.rows, bx.stride, bx.close, bx.open, bx.drift = self.nsteps, 3, 2, 0, 1
        L = _lib.lib()
        nbytes = int(L.fpm_batch_device_bytes(n, boxes))
        if nbytes < 0:
            raise RuntimeError("fpm_batch_device_bytes failed: %s" % L.fpm_strerror(nbytes).decode())
        mem = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.counter.device)
        st = ctypes.c_int(0)
        plan = L.fpm_batch_create(n, boxes, mem.data_ptr(), mem.numel(), ctypes.byref(st))
        if not plan:
            raise RuntimeError("fpm_batch_create failed: %s" % L.fpm_strerror(st.value).decode())
        return plan, mem

    def __del__(self):
        plans, self.plans = getattr(self, "plans", []), []
        for plan, _ in plans:
            try:
                _lib.lib().fpm_batch_destroy(plan)
            except Exception:  # interpreter shutdown: the library may be gone
                pass

    def _enqueue_box(self, b):
        st = _stream()
        cnt = self.counter.data_ptr()
        s = b.state
        n = s.P.numel()
        # previous closing kick (slot 2 of the previous row), opening kick
        # (slot 0) and drift (slot 1) of this row
        _lib.call("fpm_kick_drift_table", b.pm.dtype, s.S.data_ptr(), s.P.data_ptr(),
                  s.F.data_ptr(), n, b.table.data_ptr(), cnt, self.nsteps, 3, 2, 0, 1, st)
        b.step.force_kernels(s, keep_delta_k=False)

    def _enqueue_step(self):
        """[K,] K, D, F of every box for the step held in the device counter
        (one batched step, or fork to the box streams and join), then advance
        the counter."""
        main = torch.cuda.current_stream()
        if self.plans:
            for (plan, _), stream in zip(self.plans, self.streams):
                stream.wait_stream(main)
                with torch.cuda.stream(stream):
                    _lib.call_plain("fpm_batch_step", plan, self.counter.data_ptr(), _stream())
            for stream in self.streams[:len(self.plans)]:
                main.wait_stream(stream)
            _lib.call_plain("fpm_counter_add", self.counter.data_ptr(), 1, _stream())
            return
        for lane, stream in zip(self.lanes, self.streams):
            stream.wait_stream(main)
            with torch.cuda.stream(stream):
                for b in lane:
                    self._enqueue_box(b)
        for stream in self.streams:
            main.wait_stream(stream)
        _lib.call_plain("fpm_counter_add", self.counter.data_ptr(), 1, _stream())

    def step_eager(self):
        self._enqueue_step()

    def capture(self):
        """Captures one step of all boxes as a graph (run one eager step first
        so every buffer exists)."""
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self._enqueue_step()
        self.graph = g
        return g

    def finish(self):
        """The closing kick of the last step of every box (after the last
        step or replay), and the final time tags."""
        if self.nsteps > 0:
            for b in self.boxes:
                s = b.state
                _lib.call("fpm_axpy", b.pm.dtype, s.P.data_ptr(), s.F.data_ptr(), s.P.numel(),
                          float(b.rows[-1][2]), _stream())
        self._finish_tags()

    def run(self, use_graph=True):
        """Runs all steps: the first eagerly, the rest by graph replay."""
        done = 0
        if self.nsteps > 0:
            self.step_eager()
            done = 1
        if use_graph and self.nsteps > 1:
            torch.cuda.synchronize()
            if self.graph is None:
                self.capture()
            for _ in range(self.nsteps - done):
                self.graph.replay()
        else:
            for _ in range(self.nsteps - done):
                self.step_eager()
        self.finish()

    def _finish_tags(self):
        a = self.stages[-1]
        for b in self.boxes:
            b.state.a = dict(S=a, P=a, F=a)


def config_specs(n=64, seed=0, grids=(32, 48, 64, 96, 128), box_range=(100.0, 800.0),
                 sigma8_range=(0.7, 0.9), base=None):
    """Box specifications of the batched benchmark configuration: particle
    grids drawn from `grids`, sides and sigma8 uniform in their ranges,
    white-noise seeds 0..n-1."""
    from .cosmology import Cosmology

    base = Cosmology(h=0.7, Om0=0.3, n_s=0.96, sigma8=0.8) if base is None else base
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        specs.append(dict(nc=int(rng.choice(grids)), box=float(rng.uniform(*box_range)),
                          cosmo=base.clone(sigma8=float(rng.uniform(*sigma8_range))), seed=i))
    return specs
