"""Particle mesh on one MI355X: fields as device tensors, operations via libfpm.

This replaces the parts of ``pmesh.pm.ParticleMesh`` that the reference solver
calls (fastpm/core.py:51-54,102,114,187-205; fastpm/force/lpt.py:11-22;
fastpm/force/gravity.py:17-24): ``create``, ``paint``, ``readout``, ``r2c``,
``c2r``, ``apply`` with the reference's k-space kernels,
``generate_whitenoise``, ``generate_uniform_particle_grid`` and the (trivial,
single-device) ``decompose`` / ``exchange`` / ``gather``.

Storage: real fields are [n0, n1, n2] tensors; complex fields are
[n0, n1, pitch2] complex tensors with pitch2 = n2/2+1 rounded up to 128 bytes;
``ComplexField.value`` is the [n0, n1, n2/2+1] view.
"""
import numpy as np
import torch

from . import _lib


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("fastpm-mi355x runs on a ROCm GPU; none is visible (no CPU fallback)")


def _stream():
    return torch.cuda.current_stream().cuda_stream


_TORCH = {np.dtype("f4"): torch.float32, np.dtype("f8"): torch.float64}
_CTORCH = {np.dtype("f4"): torch.complex64, np.dtype("f8"): torch.complex128}


class SelfComm:
    """Stand-in communicator for one device (size 1, rank 0)."""

    size = 1
    rank = 0

    def allreduce(self, x, op=None):
        return x

    def allgather(self, x):
        return [x]

    def bcast(self, x, root=0):
        return x

    def barrier(self):
        pass


class LatticeGrid:
    """Uniform particle grid, one particle per cell of an n0 x n1 x n2 mesh, in
    C order: Q_p = (idx(p) + shift) * BoxSize / Nmesh.  Kept implicit (integer
    lattice index); ``value`` materialises positions on the device."""

    def __init__(self, pm, shift=0.0):
        self.pm = pm
        self.nlat = tuple(int(n) for n in pm.Nmesh)
        self.shift = float(shift)
        self.dtype = pm.dtype

    def __len__(self):
        return int(np.prod(self.nlat))

    @property
    def shape(self):
        return (len(self), 3)

    @property
    def size(self):
        return 3 * len(self)

    @property
    def value(self):
        return self.pm.positions(Particles(self, None))

    def __array__(self, dtype=None):
        a = self.value.cpu().numpy()
        return a if dtype is None else a.astype(dtype)


class Particles:
    """Positions X = Q + S with Q a LatticeGrid or an explicit [N, 3] device
    tensor, and S a [N, 3] device tensor (or None)."""

    def __init__(self, Q, S=None):
        self.Q = Q
        self.S = S

    def __len__(self):
        return len(self.Q)

    def desc(self, dtype):
        p = _lib.FpmParticles()
        p.n = len(self.Q)
        if isinstance(self.Q, LatticeGrid):
            p.q = None
            for d in range(3):
                p.lattice[d] = self.Q.nlat[d]
            p.lattice_shift = self.Q.shift
        else:
            _check_tensor(self.Q, dtype, (p.n, 3))
            p.q = self.Q.data_ptr()
        if self.S is not None:
            _check_tensor(self.S, dtype, (p.n, 3))
            p.s = self.S.data_ptr()
        return p


def _check_tensor(t, dtype, shape):
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a device tensor, got %s" % type(t))
    if not t.is_cuda:
        raise ValueError("tensor must live on the GPU")
    if t.dtype != _TORCH[np.dtype(dtype)]:
        raise TypeError("tensor dtype %s does not match the mesh dtype %s" % (t.dtype, dtype))
    if tuple(t.shape) != tuple(shape):
        raise ValueError("tensor shape %s, expected %s" % (tuple(t.shape), tuple(shape)))
    if not t.is_contiguous():
        raise ValueError("tensor must be contiguous")


class Layout:
    """Domain layout on one device: exchange and gather are identities
    (pmesh decompose/exchange/gather, fastpm/core.py:187-189,203,207)."""

    def exchange(self, x):
        return x

    def gather(self, x, mode="sum"):
        return x


class ParticleMesh:
    def __init__(self, Nmesh, BoxSize, dtype="f4", resampler="cic", comm=None, device=None):
        _require_gpu()
        nm = np.array(Nmesh, dtype="i8").reshape(-1)
        if nm.size == 1:
            nm = np.repeat(nm, 3)
        if nm.size != 3:
            raise ValueError("only 3-dimensional meshes are supported")
        box = np.array(BoxSize, dtype="f8").reshape(-1)
        if box.size == 1:
            box = np.repeat(box, 3)
        self.Nmesh = nm
        self.BoxSize = box
        self.ndim = 3
        self.dtype = np.dtype(dtype)
        if self.dtype not in _TORCH:
            raise TypeError("dtype must be f4 or f8")
        self.resampler = resampler
        self.comm = SelfComm() if comm is None else comm
        self.device = torch.device("cuda", torch.cuda.current_device()) if device is None \
            else torch.device(device)
        for d in range(3):
            if not _lib.fft_supported(int(nm[d])) and d < 2:
                raise ValueError("unsupported mesh size %d" % nm[d])
        if nm[2] % 2 or not _lib.fft_supported(int(nm[2] // 2)):
            raise ValueError("unsupported mesh size %d (last axis must be even, n/2 supported)"
                             % nm[2])
        self.nh = int(nm[2] // 2 + 1)
        align = 16 if self.dtype == np.float32 else 8
        self.pitch2 = (self.nh + align - 1) // align * align
        self._tw = None
        self._tile_binned = None  # (count, order, workspace) of the last bin call
        self._power_work = None
        self._desc = _lib.FpmMesh()
        for d in range(3):
            self._desc.nmesh[d] = int(nm[d])
            self._desc.box[d] = float(box[d])

    # ------------------------------------------------------------------ helpers
    @property
    def variant(self):
        """Kernel variants of this mesh ({field: choice}, fpm.h FPM_VAR_*;
        0 = default), passed with every library call on the mesh; set by
        fastpm.tune."""
        return _lib.unpack_variant(self._desc.variant)

    def set_variant(self, **choices):
        """Updates the named variant fields, e.g. set_variant(fft_cols=2)."""
        cur = self.variant
        cur.update(choices)
        self._desc.variant = _lib.pack_variant(cur)

    @property
    def order(self):
        return resampler_order(self.resampler)

    @property
    def tdtype(self):
        return _TORCH[self.dtype]

    def k_axes(self):
        """Wave numbers of the half-spectrum axes in h/Mpc as float64 numpy
        arrays: 2 pi fftfreq(n_d) n_d / L_d along x and y, 2 pi m / L_z with
        m = 0 .. n_z/2 along z."""
        out = []
        for d in range(3):
            n = int(self.Nmesh[d])
            m = np.arange(n // 2 + 1) if d == 2 else np.fft.fftfreq(n, 1.0 / n)
            out.append(2 * np.pi * m / float(self.BoxSize[d]))
        return out

    def twiddles(self):
        if self._tw is None:
            self._tw = [torch.from_numpy(_lib.twiddles(int(n), self.dtype)).to(self.device)
                        for n in self.Nmesh]
        return self._tw

    def _real_tensor(self):
        return torch.empty(tuple(int(n) for n in self.Nmesh), dtype=self.tdtype, device=self.device)

    def _complex_tensor(self):
        return torch.zeros((int(self.Nmesh[0]), int(self.Nmesh[1]), self.pitch2),
                           dtype=_CTORCH[self.dtype], device=self.device)

    def create(self, mode="real", base=None, value=None):
        if mode == "real":
            f = RealField(self)
        elif mode == "complex":
            f = ComplexField(self)
        else:
            raise ValueError("mode must be 'real' or 'complex'")
        if value is not None:
            f.value[...] = value
        return f

    def decompose(self, pos, smoothing=None):
        return Layout()

    def generate_uniform_particle_grid(self, shift=0.0, dtype=None):
        return LatticeGrid(self, shift)

    def _particles(self, pos):
        if isinstance(pos, Particles):
            return pos
        if isinstance(pos, LatticeGrid):
            return Particles(pos, None)
        if isinstance(pos, np.ndarray):
            pos = torch.from_numpy(np.ascontiguousarray(pos, dtype=self.dtype)).to(self.device)
        return Particles(pos.contiguous(), None)

    def positions(self, parts):
        """X = Q + S as a [N, 3] device tensor."""
        parts = self._particles(parts)
        out = torch.empty((len(parts), 3), dtype=self.tdtype, device=self.device)
        d = parts.desc(self.dtype)
        _lib.call("fpm_positions", self.dtype, d, self._desc, out.data_ptr(), _stream())
        return out

    # ------------------------------------------------------------------ operations
    def paint(self, pos, mass=1.0, resampler=None, hold=False, out=None):
        """pmesh ParticleMesh.paint (fastpm/core.py:191): mass assignment."""
        if out is None:
            out = RealField(self)
        parts = self._particles(pos)
        d = parts.desc(self.dtype)
        _lib.call("fpm_paint", self.dtype, d, self._desc, resampler_order(resampler or self.resampler),
                  float(mass), out.value.data_ptr(), 1 if hold else 0, _stream())
        return out

    def readout_many(self, fields, pos, outs, resampler=None, scales=None):
        """Fused readout of up to four real fields into strided outputs."""
        parts = self._particles(pos)
        d = parts.desc(self.dtype)
        n = len(fields)
        meshes = _lib.ptr_array([f.value.data_ptr() for f in fields])
        optrs = _lib.ptr_array([o.data_ptr() for o in outs])
        strides = (ctypes_i64 * n)(*[int(o.stride(0)) for o in outs])
        sc = (ctypes_f64 * n)(*([1.0] * n if scales is None else [float(s) for s in scales]))
        _lib.call("fpm_readout", self.dtype, d, self._desc,
                  resampler_order(resampler or self.resampler), n, meshes, optrs, strides, sc,
                  _stream())

    # ------------------------------------------------------------------ tiled path
    def tiles_supported(self):
        return self.dtype == np.float32 and bool(_lib.lib().fpm_tiles_supported(self._desc))

    def tile_workspace(self, n, order):
        """Device workspace for binning n particles (cached, grown on demand)."""
        need = int(_lib.lib().fpm_tile_workspace_size(self._desc, int(order), int(n)))
        if need < 0:
            raise ValueError("mesh does not support the tiled path")
        ws = getattr(self, "_tile_ws", None)
        if ws is None or ws.numel() < need:
            ws = torch.empty(need, dtype=torch.uint8, device=self.device)
            self._tile_ws = ws
        return ws

    def tile_misplaced(self, n, order):
        """Particles the last bin call's scatter could not store (a slot
        outside [0, n): 0 unless the count and scatter passes disagreed; then
        the tile lists are incomplete).  Synchronises."""
        off = int(_lib.lib().fpm_tile_misplaced_offset(self._desc, int(order), int(n)))
        ws = self._tile_ws
        return int(ws[off:off + 4].view(torch.int32).item())

    def tile_layout_exact(self, n, order):
        """Whether the last bin call placed its records exactly (count,
        scan and scatter passes) rather than in one pass into the capacity
        layout its previous call left.  Synchronises."""
        off = int(_lib.lib().fpm_tile_misplaced_offset(self._desc, int(order), int(n)))
        ws = self._tile_ws
        return bool(ws[off + 4:off + 8].view(torch.int32).item())

    def bin_particles(self, pos, resampler=None, rebin=False, kick_drift=None):
        """Bins `pos` into the mesh's tile workspace for the tiled paint and
        readouts.  rebin=True lets the call place the records in one pass
        into the capacity layout the previous bin call on this mesh left
        (it falls back to exact placement by itself when that layout does
        not fit); it is used only when that call binned as many particles
        with the same order into the same workspace.  kick_drift =
        (P, F, kicks, drift) first applies a step boundary to the momenta
        P and to `pos`'s displacement S (P += k F for k in kicks, then
        S += drift P, as fpm_kick_drift): inside the binning pass when it
        rebins (fpm_kick_drift_rebin), else as a pass of its own."""
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        key = (len(parts), order, ws.data_ptr())
        can_rebin = rebin and self._tile_binned == key
        if kick_drift is not None:
            P, F, kicks, drift = kick_drift
            kicks = [float(k) for k in kicks]
            if parts.S is None or len(kicks) > 2:
                raise ValueError("kick_drift needs a displacement and at most two kicks")
            kc = (ctypes.c_double * 2)(*(kicks + [0.0] * (2 - len(kicks))))
            if can_rebin:
                _check_tensor(P, self.dtype, tuple(parts.S.shape))
                _check_tensor(F, self.dtype, tuple(parts.S.shape))
                _lib.call_plain("fpm_kick_drift_rebin_f32", parts.desc(self.dtype), self._desc,
                                order, P.data_ptr(), F.data_ptr(), kc, len(kicks), float(drift),
                                ws.data_ptr(), ws.numel(), _stream())
                self._tile_binned = key
                return parts
            _lib.call("fpm_kick_drift", self.dtype, parts.S.data_ptr(), P.data_ptr(), F.data_ptr(),
                      parts.S.numel(), kc, len(kicks), float(drift), _stream())
        fn = "fpm_rebin_particles_f32" if can_rebin else "fpm_bin_particles_f32"
        _lib.call_plain(fn, parts.desc(self.dtype), self._desc, order, ws.data_ptr(), ws.numel(),
                        _stream())
        self._tile_binned = key
        return parts

    def paint_tiled(self, pos, mass=1.0, resampler=None, out=None):
        """Deterministic tiled paint; `pos` must have been binned with
        bin_particles (same particles and resampler)."""
        if out is None:
            out = RealField(self)
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        _lib.call_plain("fpm_paint_tiled_f32", parts.desc(self.dtype), self._desc, order,
                        float(mass), out.value.data_ptr(), ws.data_ptr(), ws.numel(), _stream())
        return out

    def readout_tiled(self, fields, pos, outs, resampler=None, scales=None):
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        n = len(fields)
        meshes = _lib.ptr_array([f.value.data_ptr() for f in fields])
        optrs = _lib.ptr_array([o.data_ptr() for o in outs])
        strides = (ctypes_i64 * n)(*[int(o.stride(0)) for o in outs])
        sc = (ctypes_f64 * n)(*([1.0] * n if scales is None else [float(s) for s in scales]))
        _lib.call_plain("fpm_readout_tiled_f32", parts.desc(self.dtype), self._desc, order, n,
                        meshes, optrs, strides, sc, ws.data_ptr(), ws.numel(), _stream())

    def paint_tiled_weighted(self, pos, weights, outs, scale=1.0, resampler=None):
        """outs[c] = scale * paint(pos, mass=weights[c]) for up to three 1-D
        weight views at once (the adjoint of a readout); `pos` must have been
        binned with bin_particles."""
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        n = len(weights)
        wp = _lib.ptr_array([w.data_ptr() for w in weights])
        st = (ctypes_i64 * n)(*[int(w.stride(0)) for w in weights])
        op = _lib.ptr_array([o.value.data_ptr() for o in outs])
        _lib.call_plain("fpm_paint_tiled_weighted_f32", parts.desc(self.dtype), self._desc, order,
                        n, wp, st, float(scale), op, ws.data_ptr(), ws.numel(), _stream())
        return outs

    def grad4_supported(self, resampler=None):
        """Whether readout_tiled_grad4 runs on this mesh (tiled, and at least
        16 + order + 3 cells per side for the stencil box)."""
        order = resampler_order(resampler or self.resampler)
        return self.tiles_supported() and int(self.Nmesh.min()) >= 16 + order + 3

    def readout_tiled_grad4(self, phi, pos, force, rho=None, rho_out=None, resampler=None,
                            force_scale=1.0, rho_scale=1.0):
        """force[:, d] = force_scale * readout(D_d phi) with D_d the four-point
        gradient stencil (equal to c2r of the order-1 gradient kernel times
        phi's spectrum), and rho_out = rho_scale * readout(rho) when given;
        `pos` must have been binned with bin_particles."""
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        if force.dim() != 2 or force.shape[1] != 3 or force.stride(1) != 1:
            raise ValueError("force must be an [N, 3] array with unit stride along the last axis")
        rp = rho.value.data_ptr() if rho is not None else None
        ro = rho_out.data_ptr() if rho_out is not None else None
        rs = int(rho_out.stride(0)) if rho_out is not None else 1
        _lib.call_plain("fpm_readout_tiled_grad4_f32", parts.desc(self.dtype), self._desc, order,
                        rp, phi.value.data_ptr(), ro, rs, force.data_ptr(), int(force.stride(0)),
                        float(rho_scale), float(force_scale), ws.data_ptr(), ws.numel(), _stream())

    def window_grad_tiled_grad4(self, phi, pos, gf, gx, scale=1.0, accumulate=True,
                                resampler=None):
        """gx (+)= scale * grad_x sum_d gf[:, d] readout(D_d phi) with D_d the
        four-point gradient stencil (the position part of the stencil force's
        adjoint); `pos` must have been binned with bin_particles."""
        parts = self._particles(pos)
        order = resampler_order(resampler or self.resampler)
        ws = self.tile_workspace(len(parts), order)
        if gf.dim() != 2 or gf.shape[1] != 3 or gf.stride(1) != 1:
            raise ValueError("gf must be an [N, 3] array with unit stride along the last axis")
        if not gx.is_contiguous() or gx.shape != (len(parts), 3):
            raise ValueError("gx must be a contiguous [N, 3] array")
        _lib.call_plain("fpm_window_grad_tiled_grad4_f32", parts.desc(self.dtype), self._desc,
                        order, phi.value.data_ptr(), gf.data_ptr(), int(gf.stride(0)),
                        gx.data_ptr(), float(scale), 1 if accumulate else 0, ws.data_ptr(),
                        ws.numel(), _stream())

    def generate_whitenoise(self, seed, type="complex", unitary=False, mean=0):
        """Counter-based complex white noise (replaces pmesh generate_whitenoise,
        fastpm/core.py:114)."""
        f = ComplexField(self)
        _lib.call("fpm_whitenoise", self.dtype, f._t.data_ptr(), self._desc, self.pitch2,
                  int(seed) & 0xFFFFFFFFFFFFFFFF, 1 if unitary else 0, _stream())
        if type == "complex":
            return f
        if type == "real":
            return f.c2r()
        raise ValueError("type must be 'complex' or 'real'")

    def power_work(self, nbins):
        n = int(_lib.lib().fpm_power_work_len(self._desc, int(nbins)))
        if self._power_work is None or self._power_work.numel() < n:
            self._power_work = torch.empty(n, dtype=torch.float64, device=self.device)
        return self._power_work


import ctypes  # noqa: E402

ctypes_i64 = ctypes.c_int64
ctypes_f64 = ctypes.c_double


def resampler_order(resampler):
    r = str(resampler).lower()
    if r == "cic":
        return 2
    if r == "tsc":
        return 3
    raise ValueError("resampler must be 'cic' or 'tsc', got %r" % (resampler,))


class RealField:
    def __init__(self, pm, tensor=None):
        self.pm = pm
        self.value = pm._real_tensor() if tensor is None else tensor
        self.Nmesh = pm.Nmesh
        self.BoxSize = pm.BoxSize
        self.ndim = 3

    @property
    def dtype(self):
        return self.pm.dtype

    def r2c(self, out=None):
        """Forward transform with 1/N (pmesh r2c, fastpm/core.py:205).  The real
        field is left intact; ``out=Ellipsis`` is accepted for API parity."""
        if out is None or out is Ellipsis:
            out = ComplexField(self.pm)
        tw = self.pm.twiddles()
        _lib.call("fpm_fft_r2c", self.pm.dtype, self.value.data_ptr(), out._t.data_ptr(),
                  self.pm._desc, self.pm.pitch2, tw[0].data_ptr(), tw[1].data_ptr(),
                  tw[2].data_ptr(), _stream())
        return out

    def readout(self, pos, out=None, resampler=None, layout=None):
        parts = self.pm._particles(pos)
        if out is None:
            out = torch.empty(len(parts), dtype=self.pm.tdtype, device=self.pm.device)
        self.pm.readout_many([self], parts, [out], resampler)
        return out

    def paint(self, pos, mass=1.0, resampler=None, hold=False, layout=None):
        self.pm.paint(pos, mass, resampler, hold, out=self)
        return self

    def copy(self):
        return RealField(self.pm, self.value.clone())


class ComplexField:
    def __init__(self, pm, tensor=None):
        self.pm = pm
        self._t = pm._complex_tensor() if tensor is None else tensor
        self.Nmesh = pm.Nmesh
        self.BoxSize = pm.BoxSize
        self.ndim = 3

    @property
    def value(self):
        return self._t[:, :, : self.pm.nh]

    @property
    def dtype(self):
        return np.dtype("c8") if self.pm.dtype == np.float32 else np.dtype("c16")

    def copy(self):
        return ComplexField(self.pm, self._t.clone())

    def apply(self, kernel, out=None, kind="wavenumber"):
        """pmesh ComplexField.apply (fastpm/core.py:116-118,
        fastpm/operators.py:5-22, fastpm/force/gravity.py:17-22):
        ``out=Ellipsis`` applies in place, ``out=None`` returns a new field.

        ``kernel`` is either an operator of ``fastpm.force.kernels`` (one
        fused multiplier kernel in libfpm) or a callable ``f(k, v)`` as in
        the reference.  The callable receives, on the device, ``k``: three
        broadcastable wave-number tensors of the field's real dtype (h/Mpc;
        fftfreq order along x and y, 0 .. n/2 along z; with kind='circular'
        k h in [-pi, pi), with kind='index' the integer wave numbers), and
        ``v``: the half-spectrum values, carrying ``v.i`` (broadcastable
        integer indices 0 .. n-1), ``v.Nmesh`` and ``v.BoxSize``; it returns
        the new values.  Arithmetic, comparisons, indexing and numpy ufuncs
        (numpy.sin, numpy.exp, ...) act on these tensors on the device, and
        ``nonzero()`` returns a tuple of index tensors as numpy's does."""
        from .force.kernels import KernelOp

        if out is Ellipsis:
            dst = self
        else:
            dst = self.copy() if out is None else out
            if out is not None:
                out._t.copy_(self._t)
        if isinstance(kernel, KernelOp):
            op = kernel.fpm_kop()
            _lib.call("fpm_kspace_apply", self.pm.dtype, dst._t.data_ptr(), op, self.pm._desc,
                      self.pm.pitch2, _stream())
            return dst
        if not callable(kernel):
            raise TypeError("apply() takes a fastpm.force.kernels operator or a callable "
                            "f(k, v), got %r" % (kernel,))
        k, v = _mode_arguments(self.pm, dst.value, kind)
        res = kernel(k, v)
        if not isinstance(res, torch.Tensor):
            raise TypeError("apply(): the callable must return a tensor, got %r" % type(res))
        dst.value.copy_(res.as_subclass(torch.Tensor).broadcast_to(dst.value.shape))
        return dst

    def c2r(self, out=None, ops=None):
        """Inverse transform (pmesh c2r).  ``ops`` optionally fuses a list of
        k-space operators into the first pass, giving one output per operator."""
        return c2r_many(self, [None] if ops is None else ops, None if out in (None, Ellipsis)
                        else [out])[0]


# numpy ufuncs on apply() arguments, by the torch function that computes them
_UFUNCS = {
    "add": torch.add, "subtract": torch.sub, "multiply": torch.mul, "divide": torch.div,
    "true_divide": torch.div, "power": torch.pow, "negative": torch.neg, "absolute": torch.abs,
    "sqrt": torch.sqrt, "square": torch.square, "exp": torch.exp, "expm1": torch.expm1,
    "log": torch.log, "log1p": torch.log1p, "log10": torch.log10, "sin": torch.sin,
    "cos": torch.cos, "tan": torch.tan, "arcsin": torch.asin, "arccos": torch.acos,
    "arctan": torch.atan, "arctan2": torch.atan2, "sinh": torch.sinh, "cosh": torch.cosh,
    "tanh": torch.tanh, "conjugate": torch.conj_physical, "sign": torch.sgn,
    "floor": torch.floor, "ceil": torch.ceil, "rint": torch.round, "maximum": torch.maximum,
    "minimum": torch.minimum, "equal": torch.eq, "not_equal": torch.ne, "less": torch.lt,
    "less_equal": torch.le, "greater": torch.gt, "greater_equal": torch.ge,
    "logical_and": torch.logical_and, "logical_or": torch.logical_or,
    "logical_not": torch.logical_not, "isfinite": torch.isfinite, "isnan": torch.isnan,
}


class ModeTensor(torch.Tensor):
    """A device tensor as handed to ``ComplexField.apply`` callables: a numpy
    ufunc applied to it runs as the torch function of the same name on the
    device (its result is again a ModeTensor), and ``nonzero()`` returns a
    tuple of index tensors, as numpy's does, so kernels written against numpy
    arrays (fastpm/operators.py:5-22, fastpm/force/kernels.py:3-35) run
    unchanged.  Functions that are not ufuncs (numpy.interp, numpy.sinc) are
    not translated."""

    def __array_ufunc__(self, ufunc, method, *inputs, **kwargs):
        fn = _UFUNCS.get(ufunc.__name__)
        if method != "__call__" or fn is None or kwargs:
            raise TypeError("numpy.%s on device mode arrays is not supported in apply()"
                            % ufunc.__name__)
        args = [x.as_subclass(torch.Tensor) if isinstance(x, torch.Tensor) else x
                for x in inputs]
        if not any(isinstance(x, torch.Tensor) for x in args[:1]):
            # a scalar first operand: torch functions want a tensor there
            t = next(x for x in args if isinstance(x, torch.Tensor))
            args[0] = torch.as_tensor(args[0], device=t.device)
        return fn(*args).as_subclass(ModeTensor)

    def nonzero(self, *args, **kwargs):
        return torch.Tensor.nonzero(self.as_subclass(torch.Tensor), as_tuple=True)


def _mode_arguments(pm, values, kind):
    """(k, v) of an apply() callable for the half-spectrum view `values`."""
    if kind not in ("wavenumber", "circular", "index"):
        raise ValueError("kind must be 'wavenumber', 'circular' or 'index', got %r" % (kind,))
    rdt = pm.tdtype
    k, idx = [], []
    for d, kd in enumerate(pm.k_axes()):
        n = int(pm.Nmesh[d])
        shape = [1, 1, 1]
        shape[d] = len(kd)
        if kind == "circular":
            kd = kd * float(pm.BoxSize[d]) / n
        elif kind == "index":
            kd = np.rint(kd * float(pm.BoxSize[d]) / (2 * np.pi)).astype("i8")
        t = torch.from_numpy(np.ascontiguousarray(kd)).to(pm.device)
        if kind != "index":
            t = t.to(rdt)
        k.append(t.reshape(shape).as_subclass(ModeTensor))
        idx.append(torch.arange(len(kd), device=pm.device).reshape(shape).as_subclass(ModeTensor))
    v = values.as_subclass(ModeTensor)
    v.i = idx
    v.Nmesh = np.array(pm.Nmesh)
    v.BoxSize = np.array(pm.BoxSize)
    return k, v


def c2r_grad3(cfield, common, outs=None, work=None, grad_order=1):
    """outs[d] = c2r(i a_d(k) * common(k) * cfield) for d = 0, 1, 2 in one
    call (the three force or displacement components); a_d is the exact
    (grad_order 0) or 4-point (1) gradient.  The input is left intact."""
    pm = cfield.pm
    if outs is None:
        outs = [RealField(pm) for _ in range(3)]
    if work is None:
        work = [pm._complex_tensor() for _ in range(3)]
    k = common.fpm_kop() if not isinstance(common, _lib.FpmKop) else common
    if k.ngrad != 0:
        raise ValueError("the common factor of c2r_grad3 carries no gradient")
    k.grad_order = int(grad_order)
    tw = pm.twiddles()
    _lib.call("fpm_fft_c2r_grad3", pm.dtype, cfield._t.data_ptr(), k,
              _lib.ptr_array([o.value.data_ptr() for o in outs]),
              _lib.ptr_array([w.data_ptr() for w in work]), pm._desc, pm.pitch2,
              tw[0].data_ptr(), tw[1].data_ptr(), tw[2].data_ptr(), _stream())
    return outs


def r2c_c2r_many(rfield, ops, kout=None, outs=None, work=None):
    """kout = rfield.r2c() and outs[c] = c2r(ops[c] * kout) in one call whose
    last forward and first inverse pass share a kernel; returns (kout, outs).
    kout=False skips storing the spectrum (returned as None)."""
    pm = rfield.pm
    n = len(ops)
    if kout is None:
        kout = ComplexField(pm)
    if outs is None:
        outs = [RealField(pm) for _ in range(n)]
    if work is None:
        work = [pm._complex_tensor() for _ in range(n)]
    kops = (_lib.FpmKop * n)()
    for c, o in enumerate(ops):
        if o is None:
            kops[c].factor = 1.0
        else:
            kops[c] = o.fpm_kop() if not isinstance(o, _lib.FpmKop) else o
    tw = pm.twiddles()
    if kout is False:
        kout = None
    _lib.call("fpm_fft_r2c_c2r", pm.dtype, rfield.value.data_ptr(),
              kout._t.data_ptr() if kout is not None else None, n, kops,
              _lib.ptr_array([o.value.data_ptr() for o in outs]),
              _lib.ptr_array([w.data_ptr() for w in work]), pm._desc, pm.pitch2,
              tw[0].data_ptr(), tw[1].data_ptr(), tw[2].data_ptr(), _stream())
    return kout, outs


def c2r_many(cfield, ops, outs=None, work=None):
    """outs[c] = c2r(ops[c] * cfield) for up to six operators; the input is read
    once and left intact."""
    pm = cfield.pm
    n = len(ops)
    if outs is None:
        outs = [RealField(pm) for _ in range(n)]
    if work is None:
        work = [pm._complex_tensor() for _ in range(n)]
    kops = (_lib.FpmKop * n)()
    for c, o in enumerate(ops):
        if o is None:
            kops[c].factor = 1.0
        else:
            kops[c] = o.fpm_kop() if not isinstance(o, _lib.FpmKop) else o
    tw = pm.twiddles()
    _lib.call("fpm_fft_c2r", pm.dtype, cfield._t.data_ptr(), n, kops,
              _lib.ptr_array([o.value.data_ptr() for o in outs]),
              _lib.ptr_array([w.data_ptr() for w in work]), pm._desc, pm.pitch2,
              tw[0].data_ptr(), tw[1].data_ptr(), tw[2].data_ptr(), _stream())
    return outs
