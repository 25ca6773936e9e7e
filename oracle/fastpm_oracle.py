"""FP64 CPU oracle for the FastPM particle-mesh path.

TEST INFRASTRUCTURE ONLY.  This module is the checker: it is imported by
``tests/``, by ``__graft_entry__.smoke()`` and by the ``cpu_baseline`` leg of
``bench.py``, and nowhere else.  The product (``fastpm-python_amd/``) never
imports, links or executes anything under ``oracle/``.

It restates, in double precision numpy/scipy, the operator sequence of the
reference (rainwoodman/fastpm-python) for the particle-mesh path.  The
reference's own arithmetic lives in pmesh/nbodykit, which are not available;
where the reference delegates to them the behaviour restated here is the one
its call sites rely on, and each function says how it is pinned:

  function           follows                                 pinned by (tests/test_oracle.py)
  ----------------   -------------------------------------   ------------------------------------
  philox4x32         Salmon et al. SC11 (Random123)           published known-answer vectors
  whitenoise         core.py:113-114 (pmesh generate_...)     Hermitian symmetry, moments, |w|=1,
                                                              real self-conjugate modes with <w^2>=1
  linear             core.py:116-118                          unitary noise -> P(k) exact per mode
  r2c / c2r          core.py:205 (pmesh r2c/c2r)              constant field, plane wave, round trip
  k_laplace/gradient force/kernels.py:3-35                    plane waves (closed forms), Hermiticity
                                                              (test_kernel.py:7-13)
  paint / readout    core.py:191,203 (pmesh window)           brute-force loops, mass, grid points,
                                                              adjoint dot-product identity
  lpt1               force/lpt.py:4-23                        ZA of a plane wave (closed form)
  lpt2source         force/lpt.py:25-64                       single wave -> 0; two waves closed form
  Growth             background.py:4-256                      Einstein-de Sitter closed forms of D1,
                                                              D2, gf, gf2 and the kick/drift factors;
                                                              independent DOP853 solve; the growth-ODE
                                                              identities gf = 3/2 Om(a) a E D1 and
                                                              gf2 = 3/2 Om(a) a E (D2 - c D1^2)
  lpt (Solver.lpt)   core.py:120-145                          kick identity of test_core.py:46-62
  leapfrog/autostages core.py:212-317                         counts of test_core.py:11-30
  kick/drift/force   core.py:168-210, force/gravity.py:10-29  JANUS reversibility (examples/janus.py),
                                                              kick identity, plane-wave force
  longrange          force/gravity.py:5-29                    plane-wave closed form with split and
                                                              compensation; equals Solver.force at
                                                              split 0
  power              utils.py:77-164 (estimator semantics)    brute-force binning loops

Deliberate, documented choices where the reference is silent or unreachable:
  * white noise uses this build's counter-based generator (pmesh's bit stream
    cannot be reproduced offline); oracle and GPU consume the same modes;
  * the 4-point gradient is set to exactly 0 on the Nyquist plane (the
    reference formula gives sin(pi) ~ 1e-16 there);
  * optional window compensation (divide by sinc^p of the paint window) is an
    extension used by the benchmark configurations; it is off by default,
    as in the reference.
"""

import numpy as np

# ----------------------------------------------------------------------------
# Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC11)
_M0 = np.uint64(0xD2511F53)
_M1 = np.uint64(0xCD9E8D57)
_W0 = np.uint64(0x9E3779B9)
_W1 = np.uint64(0xBB67AE85)
_MASK = np.uint64(0xFFFFFFFF)


def philox4x32(c0, c1, c2, c3, k0, k1, rounds=10):
    """Vectorised Philox4x32: arrays of 32-bit words (held as uint64)."""
    c0, c1, c2, c3 = (np.asarray(c, dtype=np.uint64) for c in (c0, c1, c2, c3))
    k0 = np.uint64(k0) & _MASK
    k1 = np.uint64(k1) & _MASK
    for _ in range(rounds):
        p0 = _M0 * c0
        p1 = _M1 * c2
        hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK
        hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK
        c0, c1, c2, c3 = hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0
        k0 = (k0 + _W0) & _MASK
        k1 = (k1 + _W1) & _MASK
    return c0, c1, c2, c3


def _u53(hi, lo):
    m = (hi << np.uint64(21)) | (lo >> np.uint64(11))
    return (m.astype(np.float64) + 0.5) * 2.0 ** -53


# ----------------------------------------------------------------------------
# mesh helpers
def fftfreq_int(n):
    return np.fft.fftfreq(n, 1.0 / n).round().astype(np.int64)


def kvectors(nmesh, box):
    """k components broadcastable over the half spectrum [n0, n1, n2//2+1]."""
    n0, n1, n2 = nmesh
    kx = 2 * np.pi * fftfreq_int(n0) / box[0]
    ky = 2 * np.pi * fftfreq_int(n1) / box[1]
    kz = 2 * np.pi * np.arange(n2 // 2 + 1) / box[2]
    return kx[:, None, None], ky[None, :, None], kz[None, None, :]


_FFT_WORKERS = 1


def set_fft_workers(n):
    """Threads used by the oracle's transforms (scipy.fft workers)."""
    global _FFT_WORKERS
    _FFT_WORKERS = int(n)


def r2c(real):
    """pmesh RealField.r2c: forward transform normalised by 1/N (core.py:205)."""
    import scipy.fft

    return scipy.fft.rfftn(real, workers=_FFT_WORKERS) / real.size


def c2r(cplx, nmesh):
    """pmesh ComplexField.c2r: unnormalised inverse of r2c."""
    import scipy.fft

    return scipy.fft.irfftn(cplx, s=tuple(nmesh), workers=_FFT_WORKERS) * float(np.prod(nmesh))


def whitenoise(nmesh, seed, unitary=False):
    """Complex white noise on the half spectrum (core.py:113-114 contract).

    <|w|^2> = 1, |w| = 1 if unitary, zero mode 0, Hermitian on the k_z = 0 and
    Nyquist planes.  Counter = canonical mode index, key = seed."""
    n0, n1, n2 = nmesh
    nh = n2 // 2 + 1
    i = np.arange(n0)[:, None, None] * np.ones((1, n1, nh), dtype=np.int64)
    j = np.arange(n1)[None, :, None] * np.ones((n0, 1, nh), dtype=np.int64)
    k = np.arange(nh)[None, None, :] * np.ones((n0, n1, 1), dtype=np.int64)
    plane = (k == 0) | ((n2 % 2 == 0) & (k == n2 // 2))
    ip = (n0 - i) % n0
    jp = (n1 - j) % n1
    lin = i * n1 + j
    linp = ip * n1 + jp
    conj = plane & (linp < lin)
    self_ = plane & (linp == lin)
    ic = np.where(conj, ip, i)
    jc = np.where(conj, jp, j)
    ctr = ((ic * n1 + jc) * nh + k).astype(np.uint64)
    seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    o0, o1, o2, o3 = philox4x32(ctr & _MASK, ctr >> np.uint64(32), 0 * ctr, 0 * ctr,
                                seed & 0xFFFFFFFF, seed >> 32)
    ua = _u53(o0, o1)
    ub = _u53(o2, o3)
    cs = np.cos(2 * np.pi * ub)
    sn = np.sin(2 * np.pi * ub)
    if unitary:
        re, im = cs, sn
        re = np.where(self_, np.where(cs >= 0, 1.0, -1.0), re)
    else:
        amp = np.sqrt(-np.log(ua))
        re, im = amp * cs, amp * sn
        re = np.where(self_, np.sqrt(2.0) * re, re)
    im = np.where(self_, 0.0, im)
    im = np.where(conj, -im, im)
    w = re + 1j * im
    w[0, 0, 0] = 0
    return w


def linear(wn, Pk, nmesh, box):
    """Solver.linear (core.py:116-118): sqrt(P(|k|)) w / sqrt(V)."""
    kx, ky, kz = kvectors(nmesh, box)
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    out = np.sqrt(Pk(kk.ravel()).reshape(kk.shape)) * wn / np.sqrt(np.prod(box))
    out[0, 0, 0] = 0
    return out


# ----------------------------------------------------------------------------
# k-space kernels (force/kernels.py:3-35)
def _axis_index(n, half):
    return np.arange(n // 2 + 1) if half else fftfreq_int(n)


def k_laplace(nmesh, box):
    """1/k^2 with the k=0 mode set to 0 (kernels.py:3-9)."""
    kx, ky, kz = kvectors(nmesh, box)
    kk = kx ** 2 + ky ** 2 + kz ** 2
    kk[0, 0, 0] = 1
    out = 1.0 / kk
    out[0, 0, 0] = 0
    return out


def k_gradient(d, nmesh, box, order=1):
    """i a_d(k) (kernels.py:11-25). order 0: i k_d, zero on the Nyquist plane;
    order 1: i (8 sin w - sin 2w) / (6h), w = k_d h (exactly 0 at Nyquist)."""
    n = nmesh[d]
    half = d == 2
    idx = _axis_index(n, half)
    k = 2 * np.pi * idx / box[d]
    nyq = (np.arange(len(idx)) == n // 2) & (n % 2 == 0)
    if order == 0:
        a = np.where(nyq, 0.0, k)
    else:
        h = box[d] / n
        w = k * h
        a = np.where(nyq, 0.0, (8 * np.sin(w) - np.sin(2 * w)) / (6 * h))
    shape = [1, 1, 1]
    shape[d] = len(idx)
    return 1j * a.reshape(shape)


def k_longrange(nmesh, box, r_split):
    """exp(-k^2 r^2), identity at r = 0 (kernels.py:27-35)."""
    if r_split == 0:
        return 1.0
    kx, ky, kz = kvectors(nmesh, box)
    return np.exp(-(kx ** 2 + ky ** 2 + kz ** 2) * r_split ** 2)


def k_deconv(nmesh, box, p):
    """1 / prod_d sinc(k_d h_d / 2)^p: compensation of an order-p window."""
    if not p:
        return 1.0
    out = 1.0
    for d in range(3):
        n = nmesh[d]
        idx = _axis_index(n, d == 2)
        x = idx / n  # (k h / 2) / pi
        s = np.sinc(x)  # numpy sinc is sin(pi x)/(pi x)
        shape = [1, 1, 1]
        shape[d] = len(idx)
        out = out * s.reshape(shape)
    return 1.0 / out ** p


# ----------------------------------------------------------------------------
# paint / readout (pmesh window semantics, core.py:191,203)
def _window(x, order):
    """Per-axis window support points and weights for positions in cell units."""
    if order == 2:
        i0 = np.floor(x)
        f = x - i0
        return i0.astype(np.int64), [1 - f, f]
    if order == 3:
        c = np.floor(x + 0.5)
        d = x - c
        return (c - 1).astype(np.int64), [0.5 * (0.5 - d) ** 2, 0.75 - d ** 2, 0.5 * (0.5 + d) ** 2]
    raise ValueError("order must be 2 (CIC) or 3 (TSC)")


def _cell_coords(pos, nmesh, box):
    return [np.asarray(pos[:, d], dtype=np.float64) * nmesh[d] / box[d] for d in range(3)]


def paint(pos, nmesh, box, order=2, mass=1.0):
    """Periodic mass assignment, no volume division (rho/nbar = 1 + delta)."""
    nmesh = tuple(int(n) for n in nmesh)
    x = _cell_coords(pos, nmesh, box)
    win = [_window(x[d], order) for d in range(3)]
    out = np.zeros(int(np.prod(nmesh)), dtype=np.float64)
    for a in range(order):
        ix = (win[0][0] + a) % nmesh[0]
        wa = win[0][1][a]
        for b in range(order):
            iy = (win[1][0] + b) % nmesh[1]
            wab = wa * win[1][1][b]
            for c in range(order):
                iz = (win[2][0] + c) % nmesh[2]
                flat = (ix * nmesh[1] + iy) * nmesh[2] + iz
                out += np.bincount(flat, weights=wab * win[2][1][c] * mass, minlength=out.size)
    return out.reshape(nmesh)


def readout(mesh, pos, box, order=2):
    """Window interpolation, the transpose of paint."""
    nmesh = mesh.shape
    x = _cell_coords(pos, nmesh, box)
    win = [_window(x[d], order) for d in range(3)]
    out = np.zeros(len(pos), dtype=np.float64)
    for a in range(order):
        ix = (win[0][0] + a) % nmesh[0]
        for b in range(order):
            iy = (win[1][0] + b) % nmesh[1]
            wab = win[0][1][a] * win[1][1][b]
            for c in range(order):
                iz = (win[2][0] + c) % nmesh[2]
                out += wab * win[2][1][c] * mesh[ix, iy, iz]
    return out


def lattice(nlat, box, shift=0.0):
    """pm.generate_uniform_particle_grid(shift): (i + shift) L / N in C order
    (restated in operators.py:24-42)."""
    nlat = tuple(int(n) for n in nlat)
    axes = [(np.arange(nlat[d]) + shift) * box[d] / nlat[d] for d in range(3)]
    g = np.meshgrid(*axes, indexing="ij")
    return np.stack([gi.ravel() for gi in g], axis=1)


# ----------------------------------------------------------------------------
# LPT (force/lpt.py)
def lpt1(dlin_k, q, nmesh, box, order=2):
    """First-order displacement, CIC readout at q (lpt.py:4-23)."""
    lap = k_laplace(nmesh, box)
    out = np.zeros((len(q), 3))
    for d in range(3):
        disp = c2r(dlin_k * lap * k_gradient(d, nmesh, box, 1), nmesh)
        out[:, d] = readout(disp, q, box, order)
    return out


def lpt2source(dlin_k, nmesh, box):
    """Second-order source (lpt.py:25-64)."""
    lap = k_laplace(nmesh, box)
    g = [k_gradient(d, nmesh, box, 1) for d in range(3)]
    phi_ii = [c2r(dlin_k * lap * g[d] * g[d], nmesh) for d in range(3)]
    D1 = [1, 2, 0]
    D2 = [2, 0, 1]
    source = np.zeros(tuple(nmesh))
    for d in range(3):
        source += phi_ii[D1[d]] * phi_ii[D2[d]]
    for d in range(3):
        phi_ij = c2r(dlin_k * lap * g[D1[d]] * g[D2[d]], nmesh)
        source -= phi_ij ** 2
    source *= 3.0 / 7
    return r2c(source)


# ----------------------------------------------------------------------------
# long-range force (force/gravity.py:5-29)
def longrange(x_list, delta_k, nmesh, box, split, factor, order=2, deconv=0):
    """longrange_batch (gravity.py:11-29): pot = delta_k / k^2 (0 at k = 0)
    * exp(-k^2 split^2) [/ W(k)^deconv, this build's compensation]; for each
    axis d, c2r(i a_d(k) pot) with the four-point a_d, read out at every
    position set; every force multiplied by factor.  One [N, 3] array per set
    (longrange(x, ...) of gravity.py:5-8 is the one-set case)."""
    pot = (delta_k * k_laplace(nmesh, box) * k_longrange(nmesh, box, split)
           * k_deconv(nmesh, box, deconv))
    out = [np.empty((len(x), 3)) for x in x_list]
    for d in range(3):
        fd = c2r(pot * k_gradient(d, nmesh, box, 1), nmesh)
        for x, o in zip(x_list, out):
            o[:, d] = readout(fd, x, box, order)
    for o in out:
        o *= factor
    return out


# ----------------------------------------------------------------------------
# growth (background.py:4-256), restated with odeint exactly as the reference
class Growth:
    """Matter-dominated growth: D1, D2 and derivatives in ln a at grid points
    {1e-7, a_normalize, a...}, linear interpolation in ln a elsewhere."""

    def __init__(self, Om0, a=None, a_normalize=1.0, Ok0=0.0, Ol0=None):
        from scipy.integrate import odeint

        self.Om0 = Om0
        self.Ok0 = Ok0
        self.Ol0 = 1 - Ok0 - Om0 if Ol0 is None else Ol0
        if a is None:
            lna = np.log(np.logspace(-7, 0, 1024 * 10, endpoint=True))
        else:
            a = np.array(a, dtype="f8").ravel()
            if a_normalize not in a:
                a = np.concatenate([[a_normalize], a])
            a.sort()
            if a[0] > 1e-7:
                a = np.concatenate([[1e-7], a])
            lna = np.log(a)
        self.lna = lna
        a0 = np.exp(lna[0])
        y0 = [a0, a0, -3.0 / 7 * a0 ** 2, -6.0 / 7 * a0 ** 2]
        y = odeint(self._ode, y0, lna, tcrit=[0.0], atol=0)
        v1, v2 = [], []
        for yi, li in zip(y, lna):
            d = self._ode(yi, li)
            v1.append((yi[0], yi[1], d[1]))
            v2.append((yi[2], yi[3], d[3]))
        v1 = np.array(v1)
        v2 = np.array(v2)
        ind = abs(lna - np.log(a_normalize)).argmin()
        self._D1 = v1 / v1[ind][0]
        self._D2 = v2 / v2[ind][0]

    def efunc(self, a):
        return (self.Om0 / a ** 3 + self.Ok0 / a ** 2 + self.Ol0) ** 0.5

    def efunc_prime(self, a):
        return 0.5 / self.efunc(a) * (-3 * self.Om0 / a ** 4 - 2 * self.Ok0 / a ** 3)

    def Om(self, a):
        return self.Om0 / a ** 3 / self.efunc(a) ** 2

    def E(self, a, order=0):
        return self.efunc(a) if order == 0 else self.efunc_prime(a) * a

    def _ode(self, y, lna):
        D1, F1, D2, F2 = y
        a = np.exp(lna)
        hfac = -2.0 - self.E(a, 1) / self.E(a)
        om = self.Om(a)
        return (F1, hfac * F1 + 1.5 * om * D1, F2, hfac * F2 + 1.5 * om * D2 - 1.5 * om * D1 ** 2)

    def D1(self, a, order=0):
        return np.interp(np.log(a), self.lna, self._D1[:, order])

    def D2(self, a, order=0):
        return np.interp(np.log(a), self.lna, self._D2[:, order])

    def f1(self, a):
        return self.D1(a, 1) / self.D1(a)

    def f2(self, a):
        return self.D2(a, 1) / self.D2(a)

    def Gp(self, a):
        return self.D1(a)

    def gp(self, a):
        return self.D1(a, 1) / a

    def Gf(self, a):
        return self.D1(a, 1) * a ** 2 * self.E(a)

    def gf(self, a):
        return 1 / a * (self.D1(a, 2) * a ** 2 * self.E(a)
                        + self.D1(a, 1) * (a ** 2 * self.E(a, 1) + 2 * a ** 2 * self.E(a)))

    def gf2(self, a):
        return 1 / a * (self.D2(a, 2) * a ** 2 * self.E(a)
                        + self.D2(a, 1) * (a ** 2 * self.E(a, 1) + 2 * a ** 2 * self.E(a)))


# ----------------------------------------------------------------------------
# stepping (core.py:212-317)
def leapfrog(stages):
    stages = list(stages)
    if len(stages) == 0:
        return
    ai = stages[0]
    yield "F", ai, ai, ai
    x = p = f = ai
    for i in range(len(stages) - 1):
        a0, a1 = stages[i], stages[i + 1]
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
    stages = list(stages)
    if len(stages) == 0:
        return
    x = p = f = stages[0]
    for i in range(len(stages) - 1):
        a0, a1 = stages[i], stages[i + 1]
        ah = (a0 * a1) ** 0.5
        yield "K", p, f, ah
        p = ah
        yield "D", x, p, a1
        x = a1
        yield "F", f, x, a1
        f = a1
        yield "K", p, f, a1
        p = a1


def autostages(knots, N, astart=None, N0=None):
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
    for i in range(len(knots) - 1):
        da = (knots[-1] - knots[i]) / (N - len(stages) - 1)
        n_span = int((knots[i + 1] - knots[i]) / da + 0.5)
        if i == 0 and n_span < N0:
            n_span = N0
        stages = np.append(stages, np.linspace(knots[i], knots[i + 1], n_span, endpoint=False))
    return np.append(stages, [knots[-1]])


# ----------------------------------------------------------------------------
# Solver (core.py:91-210)
class State:
    def __init__(self, Q, dtype="f8"):
        self.Q = np.asarray(Q, dtype=dtype)
        self.S = np.zeros_like(self.Q)
        self.P = np.zeros_like(self.Q)
        self.F = np.zeros_like(self.Q)
        self.RHO = np.zeros(len(self.Q), dtype=dtype)
        self.a = dict(S=None, P=None, F=None)

    @property
    def X(self):
        return self.S + self.Q

    def copy(self):
        o = object.__new__(State)
        o.Q = self.Q
        o.S, o.P, o.F, o.RHO = self.S.copy(), self.P.copy(), self.F.copy(), self.RHO.copy()
        o.a = dict(self.a)
        return o


class Solver:
    """FP64 restatement of fastpm.core.Solver for a particle mesh of ``nc``
    cells per side (an int, or three sides) and force mesh ``nc * B``, in a
    box of side ``box`` (a float, or three sides)."""

    def __init__(self, nc, box, Om0, B=1, a_linear=1.0, resampler="cic", deconv=0):
        nc = tuple(int(n) for n in np.broadcast_to(nc, (3,)))
        self.nmesh = nc
        self.fnmesh = tuple(n * B for n in nc)
        self.box = tuple(float(b) for b in np.broadcast_to(box, (3,)))
        self.Om0 = Om0
        self.a_linear = a_linear
        self.order = {"cic": 2, "tsc": 3}[resampler]
        self.deconv = deconv

    def whitenoise(self, seed, unitary=False):
        return whitenoise(self.nmesh, seed, unitary)

    def linear(self, wn, Pk):
        return linear(wn, Pk, self.nmesh, self.box)

    def lpt(self, dlin, Q, a, order=2):
        assert order in (1, 2)
        st = State(Q)
        pt = Growth(self.Om0, a=[a], a_normalize=self.a_linear)
        DX1 = pt.D1(a) * lpt1(dlin, st.Q, self.nmesh, self.box)
        V1 = a ** 2 * pt.f1(a) * pt.E(a) * DX1
        if order == 2:
            DX2 = pt.D2(a) * lpt1(lpt2source(dlin, self.nmesh, self.box), st.Q, self.nmesh, self.box)
            V2 = a ** 2 * pt.f2(a) * pt.E(a) * DX2
            st.S[...] = DX1 + DX2
            st.P[...] = V1 + V2
            st.F[...] = a ** 2 * pt.E(a) * (pt.gf(a) / pt.D1(a) * DX1 + pt.gf2(a) / pt.D2(a) * DX2)
        else:
            st.S[...] = DX1
            st.P[...] = V1
            st.F[...] = a ** 2 * pt.E(a) * (pt.gf(a) / pt.D1(a) * DX1)
        st.a["S"] = a
        st.a["P"] = a
        return st

    def kick_factor(self, ai, ac, af):
        pt = Growth(self.Om0, a=[ai, ac, af], a_normalize=self.a_linear)
        return 1 / (ac ** 2 * pt.E(ac)) * (pt.Gf(af) - pt.Gf(ai)) / pt.gf(ac)

    def drift_factor(self, ai, ac, af):
        pt = Growth(self.Om0, a=[ai, ac, af], a_normalize=self.a_linear)
        return 1 / (ac ** 3 * pt.E(ac)) * (pt.Gp(af) - pt.Gp(ai)) / pt.gp(ac)

    def kick(self, st, ai, ac, af):
        assert ac == st.a["F"]
        st.P[...] = st.P + self.kick_factor(ai, ac, af) * st.F
        st.a["P"] = af

    def drift(self, st, ai, ac, af):
        assert ac == st.a["P"]
        st.S[...] = st.S + self.drift_factor(ai, ac, af) * st.P
        st.a["S"] = af

    def force(self, st, ai, ac, af):
        assert ac == st.a["S"]
        X = st.X
        nbar = len(X) / np.prod(self.fnmesh)
        rho = paint(X, self.fnmesh, self.box, self.order) / nbar
        st.RHO[...] = readout(rho, X, self.box, self.order)
        delta_k = r2c(rho)
        pot = delta_k * k_laplace(self.fnmesh, self.box) * k_deconv(self.fnmesh, self.box, self.deconv)
        fac = 1.5 * self.Om0
        for d in range(3):
            fd = c2r(pot * k_gradient(d, self.fnmesh, self.box, 1), self.fnmesh)
            st.F[:, d] = readout(fd, X, self.box, self.order) * fac
        st.a["F"] = af
        return dict(delta_k=delta_k)

    def nbody(self, st, stepping, monitor=None):
        acts = dict(K=self.kick, D=self.drift, F=self.force)
        for action, ai, ac, af in stepping:
            ev = acts[action](st, ai, ac, af)
            if monitor is not None:
                monitor(action, ai, ac, af, st, ev)
        return st


# ----------------------------------------------------------------------------
# glass initial conditions (fastpm/glass.py:8-118)
def glass_stages(N=3):
    """fastpm/glass.py:33: N + 1/4 periods of 2 pi, 8 stages per period."""
    return np.linspace(0, np.pi * 2 * (N + 0.25), int(8 * (N + 0.25) + 1))


def glass(nc, box, seed, B=2, spread=3.0, N=3, order=2, stages=None):
    """Positions of a glass from the uniform grid of side nc (fastpm/glass.py:
    8-118): random start from numpy's RandomState(seed), then the leapfrog
    with the damped kick P += (af - ai)(F - P) (glass.py:89-93), the periodic
    drift S = (S + (af - ai) P) mod L (glass.py:95-98) and the inverted PM
    force F = -longrange(X, delta_k, 0, 1) (glass.py:100-118); returns X mod L."""
    nmesh = (nc,) * 3
    fnmesh = (nc * B,) * 3
    boxv = (float(box),) * 3
    Q = lattice(nmesh, boxv)
    n = len(Q)
    rng = np.random.RandomState(seed)
    nbar = 1 / (float(box) ** 3 / n)
    S = spread * (rng.uniform(size=(n, 3)) - 0.5) * (nbar ** -0.3333333)
    P = np.zeros_like(S)
    F = np.zeros_like(S)
    L = float(box)

    def force(S):
        X = Q + S
        nb = n / np.prod(fnmesh)
        rho = paint(X, fnmesh, boxv, order) / nb
        pot = r2c(rho) * k_laplace(fnmesh, boxv)
        out = np.empty_like(S)
        for d in range(3):
            fd = c2r(pot * k_gradient(d, fnmesh, boxv, 1), fnmesh)
            out[:, d] = -readout(fd, X, boxv, order)
        return out

    stages = glass_stages(N) if stages is None else stages
    for action, ai, ac, af in leapfrog(stages):
        fac = af - ai
        if action == "K":
            P = P + fac * (F - P)
        elif action == "D":
            S = (S + fac * P) % L
        else:
            F = force(S)
    return (Q + S) % L


# ----------------------------------------------------------------------------
# binned spectra (utils.py:77-164 semantics, linear bins)
def power_bins(f1, nmesh, box, nbins, kmin, dk, f2=None):
    """Sums per bin [w, w|k|, w Re(f1 f2*), w|f1|^2, w|f2|^2]; Hermitian weight 2
    for 0 < k_z index < n2/2; zero mode excluded."""
    if f2 is None:
        f2 = f1
    n2 = nmesh[2]
    kx, ky, kz = kvectors(nmesh, box)
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    fb = np.floor((kk - kmin) / dk)
    iz = np.arange(n2 // 2 + 1)[None, None, :]
    w = np.where((iz == 0) | ((n2 % 2 == 0) & (iz == n2 // 2)), 1.0, 2.0) * np.ones(kk.shape)
    valid = (fb >= 0) & (fb < nbins)
    valid[0, 0, 0] = False
    b = fb[valid].astype(np.int64)
    wv = w[valid]
    out = np.zeros((nbins, 5))
    out[:, 0] = np.bincount(b, weights=wv, minlength=nbins)
    out[:, 1] = np.bincount(b, weights=wv * kk[valid], minlength=nbins)
    out[:, 2] = np.bincount(b, weights=wv * (f1[valid] * np.conj(f2[valid])).real, minlength=nbins)
    out[:, 3] = np.bincount(b, weights=wv * abs(f1[valid]) ** 2, minlength=nbins)
    out[:, 4] = np.bincount(b, weights=wv * abs(f2[valid]) ** 2, minlength=nbins)
    return out


def power_spectrum(f1, nmesh, box, nbins=32, kmax=None, f2=None):
    """k, P(k), modes and r(k) from power_bins over [0, kmax) (kmax defaults to
    the Nyquist wave number)."""
    if kmax is None:
        kmax = np.pi * nmesh[0] / box[0]
    s = power_bins(f1, nmesh, box, nbins, 0.0, kmax / nbins, f2)
    V = float(np.prod(box))
    with np.errstate(invalid="ignore", divide="ignore"):
        k = s[:, 1] / s[:, 0]
        p = s[:, 2] / s[:, 0] * V
        r = s[:, 2] / np.sqrt(s[:, 3] * s[:, 4])
    return k, p, s[:, 0], r
