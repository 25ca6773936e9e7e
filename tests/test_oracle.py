"""Pins the FP64 oracle (CPU only): published known-answer vectors, closed
forms, brute force on tiny inputs, and the reference's own assertions
(test_core.py, test_kernel.py, examples/janus.py).  Each check is chosen so a
plausible mistake (a dropped term, a wrong sign or index) fails it."""
import itertools

import numpy as np
import pytest

import fastpm_oracle as O


# ---------------------------------------------------------------- Philox
@pytest.mark.parametrize("ctr,key,expect", [
    ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
    ((0xFFFFFFFF,) * 4, (0xFFFFFFFF, 0xFFFFFFFF), (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
    ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
     (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
])
def test_philox_known_answers(ctr, key, expect):
    """Random123 kat_vectors for philox4x32-10."""
    out = O.philox4x32(*[np.array([c], dtype=np.uint64) for c in ctr], *key)
    assert tuple(int(o[0]) for o in out) == expect


# ---------------------------------------------------------------- white noise
@pytest.mark.parametrize("nmesh", [(8, 8, 8), (16, 12, 6), (6, 10, 8)])
@pytest.mark.parametrize("unitary", [False, True])
def test_whitenoise_hermitian(nmesh, unitary):
    w = O.whitenoise(nmesh, 11, unitary)
    full = np.fft.irfftn(w, s=nmesh, axes=(0, 1, 2))
    # a Hermitian half spectrum survives a real round trip unchanged
    back = np.fft.rfftn(full)
    np.testing.assert_allclose(back, w, atol=1e-12)
    assert w[0, 0, 0] == 0
    if unitary:
        m = np.abs(w)
        m[0, 0, 0] = 1
        np.testing.assert_allclose(m, 1.0, rtol=1e-14)


def test_whitenoise_moments():
    w = O.whitenoise((64, 64, 64), 3)
    v = np.abs(w[1:, 1:, 1:32]) ** 2
    assert abs(v.mean() - 1) < 0.01
    # |w|^2 is exponential with mean 1: variance 1
    assert abs(v.var() - 1) < 0.03
    assert abs(w[1:, 1:, 1:32].real.mean()) < 0.01
    # seed changes the field; same seed reproduces it
    assert not np.allclose(O.whitenoise((8, 8, 8), 4), O.whitenoise((8, 8, 8), 5))
    np.testing.assert_array_equal(O.whitenoise((8, 8, 8), 4), O.whitenoise((8, 8, 8), 4))


@pytest.mark.parametrize("unitary", [False, True])
def test_whitenoise_self_conjugate_modes(unitary):
    """The seven nonzero self-conjugate modes of an even mesh (every index 0 or
    n/2) are their own conjugates, so they must be real; they must still carry
    <w^2> = 1 like every other mode (a real Gaussian with the complex modes'
    amplitude would give 1/2), be Gaussian (fourth moment 3) and symmetric in
    sign.  Unitary noise puts +1 or -1 there, each half the time."""
    vals = []
    for seed in range(1500):
        w = O.whitenoise((4, 4, 4), seed, unitary)
        for i, j, k in itertools.product((0, 2), repeat=3):
            if (i, j, k) != (0, 0, 0):
                vals.append(w[i, j, k])
    v = np.array(vals)
    assert np.all(v.imag == 0)
    x = v.real
    assert abs((x > 0).mean() - 0.5) < 0.03
    if unitary:
        np.testing.assert_array_equal(np.abs(x), 1.0)
    else:
        assert abs((x ** 2).mean() - 1) < 0.05
        assert abs((x ** 4).mean() - 3) < 0.4


def test_linear_unitary_power_exact():
    """With unitary noise, V |delta_k|^2 equals P(|k|) on every mode."""
    nmesh, box = (16, 16, 16), (100.0,) * 3
    Pk = lambda k: 1e3 * np.exp(-k) * k  # noqa: E731
    d = O.linear(O.whitenoise(nmesh, 5, unitary=True), Pk, nmesh, box)
    kx, ky, kz = O.kvectors(nmesh, box)
    kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
    m = kk > 0
    np.testing.assert_allclose((np.abs(d) ** 2 * 1e6)[m], Pk(kk[m]), rtol=1e-12)


# ---------------------------------------------------------------- transforms
def test_fft_normalisation():
    nmesh = (8, 8, 8)
    one = np.ones(nmesh)
    ck = O.r2c(one)
    assert ck[0, 0, 0] == pytest.approx(1.0)
    assert np.abs(ck).sum() == pytest.approx(1.0)
    # plane wave cos(2 pi x / L) -> 1/2 at n = 1
    x = np.arange(8)[:, None, None] * np.ones(nmesh)
    ck = O.r2c(np.cos(2 * np.pi * x / 8))
    assert ck[1, 0, 0] == pytest.approx(0.5) and ck[7, 0, 0] == pytest.approx(0.5)
    rng = np.random.default_rng(0)
    r = rng.standard_normal(nmesh)
    np.testing.assert_allclose(O.c2r(O.r2c(r), nmesh), r, atol=1e-13)


def test_gradient_hermitian():
    """test_kernel.py:7-13: c2r(g).r2c() == g for the gradient kernels."""
    nmesh, box = (8, 8, 8), (128.0,) * 3
    rng = np.random.default_rng(333)
    dk = O.r2c(rng.standard_normal(nmesh))
    for order in (0, 1):
        for d in range(3):
            g = dk * O.k_gradient(d, nmesh, box, order)
            np.testing.assert_allclose(O.r2c(O.c2r(g, nmesh)), g, atol=1e-9)


def test_kernels_plane_wave():
    """Closed forms: a plane wave cos(k0 x) through laplace and the gradients."""
    N, L = 16, 64.0
    nmesh, box = (N,) * 3, (L,) * 3
    h = L / N
    n = 3
    k0 = 2 * np.pi * n / L
    x = (np.arange(N) * h)[:, None, None] * np.ones(nmesh)
    f = np.cos(k0 * x)
    fk = O.r2c(f)
    # laplace: cos / k0^2
    np.testing.assert_allclose(O.c2r(fk * O.k_laplace(nmesh, box), nmesh), f / k0 ** 2, atol=1e-12)
    # exact gradient: -k0 sin
    np.testing.assert_allclose(O.c2r(fk * O.k_gradient(0, nmesh, box, 0), nmesh),
                               -k0 * np.sin(k0 * x), atol=1e-12)
    # 4-point gradient: -(8 sin(k0 h) - sin(2 k0 h))/(6h) sin
    a = (8 * np.sin(k0 * h) - np.sin(2 * k0 * h)) / (6 * h)
    np.testing.assert_allclose(O.c2r(fk * O.k_gradient(0, nmesh, box, 1), nmesh),
                               -a * np.sin(k0 * x), atol=1e-12)
    # gradient along another axis annihilates it
    assert np.abs(O.c2r(fk * O.k_gradient(1, nmesh, box, 1), nmesh)).max() < 1e-12
    # window compensation of order p divides by sinc(k0 h/2)^p
    w = np.sinc(n / N)
    np.testing.assert_allclose(O.c2r(fk * O.k_deconv(nmesh, box, 2), nmesh), f / w ** 2, atol=1e-12)
    # Gaussian split
    np.testing.assert_allclose(O.c2r(fk * O.k_longrange(nmesh, box, 2.0), nmesh),
                               f * np.exp(-k0 ** 2 * 4.0), atol=1e-12)


# ---------------------------------------------------------------- paint / readout
def _brute_paint(pos, nmesh, box, order):
    out = np.zeros(nmesh)
    for p in pos:
        x = [p[d] * nmesh[d] / box[d] for d in range(3)]
        pts = []
        for d in range(3):
            if order == 2:
                i0 = int(np.floor(x[d]))
                f = x[d] - i0
                pts.append([(i0, 1 - f), (i0 + 1, f)])
            else:
                c = int(np.floor(x[d] + 0.5))
                dd = x[d] - c
                pts.append([(c - 1, 0.5 * (0.5 - dd) ** 2), (c, 0.75 - dd ** 2),
                            (c + 1, 0.5 * (0.5 + dd) ** 2)])
        for (i, wi), (j, wj), (k, wk) in itertools.product(*pts):
            out[i % nmesh[0], j % nmesh[1], k % nmesh[2]] += wi * wj * wk
    return out


@pytest.mark.parametrize("order", [2, 3])
def test_paint_brute_force(order):
    rng = np.random.default_rng(7)
    nmesh, box = (6, 5, 4), (12.0, 10.0, 8.0)
    pos = rng.uniform(-15, 25, size=(40, 3))
    np.testing.assert_allclose(O.paint(pos, nmesh, box, order), _brute_paint(pos, nmesh, box, order),
                               atol=1e-13)


@pytest.mark.parametrize("order", [2, 3])
def test_paint_invariants(order):
    rng = np.random.default_rng(8)
    nmesh, box = (8, 8, 8), (16.0,) * 3
    pos = rng.uniform(0, 16, size=(100, 3))
    m = O.paint(pos, nmesh, box, order)
    assert m.sum() == pytest.approx(100.0)
    # a particle on a grid point: all mass there (CIC) or 1/8, 3/4, 1/8 (TSC)
    g = O.paint(np.array([[4.0, 6.0, 2.0]]), nmesh, box, order)
    if order == 2:
        assert g[2, 3, 1] == pytest.approx(1.0) and g.sum() == pytest.approx(1.0)
    else:
        assert g[2, 3, 1] == pytest.approx(0.75 ** 3)
        assert g[1, 3, 1] == pytest.approx(0.125 * 0.75 ** 2)
    # translation by whole cells rolls the mesh
    m2 = O.paint(pos + np.array([2.0, 0, 0]), nmesh, box, order)
    np.testing.assert_allclose(m2, np.roll(m, 1, axis=0), atol=1e-12)
    # adjoint: <paint(x), f> == sum readout(f)(x)
    f = rng.standard_normal(nmesh)
    assert (m * f).sum() == pytest.approx(O.readout(f, pos, box, order).sum(), rel=1e-12)


def test_lattice_matches_create_grid():
    q = O.lattice((2, 3, 4), (2.0, 3.0, 8.0), 0.5)
    assert q.shape == (24, 3)
    np.testing.assert_allclose(q[0], [0.5, 0.5, 1.0])
    np.testing.assert_allclose(q[1], [0.5, 0.5, 3.0])  # last axis fastest (C order)
    np.testing.assert_allclose(q[4], [0.5, 1.5, 1.0])


# ---------------------------------------------------------------- LPT
def test_lpt1_plane_wave():
    """ZA displacement of delta = A cos(k0 x): psi_x = -(a(k0)/k0^2) A sin(k0 x)
    with the 4-point gradient a; read out on the lattice."""
    N, L = 16, 100.0
    nmesh, box = (N,) * 3, (L,) * 3
    h = L / N
    k0 = 2 * np.pi * 2 / L
    q = O.lattice(nmesh, box)
    x = (np.arange(N) * h)[:, None, None] * np.ones(nmesh)
    dk = O.r2c(0.1 * np.cos(k0 * x))
    psi = O.lpt1(dk, q, nmesh, box)
    a = (8 * np.sin(k0 * h) - np.sin(2 * k0 * h)) / (6 * h)
    np.testing.assert_allclose(psi[:, 0], -0.1 * a / k0 ** 2 * np.sin(k0 * q[:, 0]), atol=1e-14)
    assert np.abs(psi[:, 1:]).max() < 1e-14


def test_lpt2source_single_wave_zero_and_two_waves():
    N, L = 16, 100.0
    nmesh, box = (N,) * 3, (L,) * 3
    h = L / N
    x = (np.arange(N) * h)[:, None, None] * np.ones(nmesh)
    y = (np.arange(N) * h)[None, :, None] * np.ones(nmesh)
    k1, k2 = 2 * np.pi * 1 / L, 2 * np.pi * 2 / L
    src = O.lpt2source(O.r2c(np.cos(k1 * x)), nmesh, box)
    assert np.abs(src).max() < 1e-15
    # delta = cos(k1 x) + cos(k2 y): phi_xx = -(a1^2/k1^2) cos(k1 x),
    # phi_yy = -(a2^2/k2^2) cos(k2 y); source = 3/7 phi_xx phi_yy
    a1 = (8 * np.sin(k1 * h) - np.sin(2 * k1 * h)) / (6 * h)
    a2 = (8 * np.sin(k2 * h) - np.sin(2 * k2 * h)) / (6 * h)
    src = O.lpt2source(O.r2c(np.cos(k1 * x) + np.cos(k2 * y)), nmesh, box)
    expect = 3.0 / 7 * (a1 ** 2 / k1 ** 2) * (a2 ** 2 / k2 ** 2) * np.cos(k1 * x) * np.cos(k2 * y)
    np.testing.assert_allclose(O.c2r(src, nmesh), expect, atol=1e-14)


# ---------------------------------------------------------------- growth
def test_growth_eds():
    """Omega_m = 1: D1 = a, D2 = a^2 (normalised), f1 = 1, f2 = 2."""
    a = np.array([0.1, 0.3, 0.7])
    g = O.Growth(1.0, a=a, a_normalize=1.0)
    np.testing.assert_allclose(g.D1(a), a, rtol=1e-7)
    np.testing.assert_allclose(g.D2(a), a ** 2, rtol=1e-6)
    np.testing.assert_allclose(g.f1(a), 1.0, rtol=1e-7)
    np.testing.assert_allclose(g.f2(a), 2.0, rtol=1e-6)
    np.testing.assert_allclose(g.gp(a), 1.0, rtol=1e-7)  # dD1/da


def test_growth_lcdm_independent_solver():
    from scipy.integrate import solve_ivp

    Om = 0.3
    a = np.array([0.05, 0.2, 0.6, 1.0])
    g = O.Growth(Om, a=a, a_normalize=1.0)

    def rhs(x, y):
        A = np.exp(x)
        E = np.sqrt(Om / A ** 3 + 1 - Om)
        Ep = A * 0.5 / E * (-3 * Om / A ** 4)
        h = -2 - Ep / E
        o = 1.5 * Om / A ** 3 / E ** 2
        return [y[1], h * y[1] + o * y[0], y[3], h * y[3] + o * y[2] - o * y[0] ** 2]

    a0 = 1e-7
    s = solve_ivp(rhs, [np.log(a0), 0.0], [a0, a0, -3 / 7 * a0 ** 2, -6 / 7 * a0 ** 2],
                  method="DOP853", rtol=1e-13, atol=1e-300, t_eval=np.log(a))
    np.testing.assert_allclose(g.D1(a), s.y[0] / s.y[0][-1], rtol=1e-7)
    np.testing.assert_allclose(g.D2(a), s.y[2] / s.y[2][-1], rtol=1e-6)
    np.testing.assert_allclose(g.D1(a, 1), s.y[1] / s.y[0][-1], rtol=1e-7)


def independent_growth(Om, a, a_normalize):
    """Raw (unnormalised) D1, D2 of the growth ODE at a and at a_normalize from
    a DOP853 solve that shares no code with the oracle (scipy.integrate.solve_ivp,
    rtol 1e-13), with E and Omega_m(a) written out for a flat matter +
    Lambda universe."""
    from scipy.integrate import solve_ivp

    def rhs(x, y):
        A = np.exp(x)
        E = np.sqrt(Om / A ** 3 + 1 - Om)
        Ep = A * 0.5 / E * (-3 * Om / A ** 4)
        h = -2 - Ep / E
        o = 1.5 * Om / A ** 3 / E ** 2
        return [y[1], h * y[1] + o * y[0], y[3], h * y[3] + o * y[2] - o * y[0] ** 2]

    a0 = 1e-7
    pts = np.unique(np.concatenate([np.atleast_1d(a), [a_normalize]]))
    s = solve_ivp(rhs, [np.log(a0), np.log(pts[-1])], [a0, a0, -3 / 7 * a0 ** 2, -6 / 7 * a0 ** 2],
                  method="DOP853", rtol=1e-13, atol=1e-300, t_eval=np.log(pts))
    at = {float(p): (s.y[0][i], s.y[2][i]) for i, p in enumerate(pts)}
    D1 = np.array([at[float(x)][0] for x in np.atleast_1d(a)])
    D2 = np.array([at[float(x)][1] for x in np.atleast_1d(a)])
    return D1, D2, at[float(a_normalize)][0], at[float(a_normalize)][1]


def test_growth_eds_force_and_step_factors():
    """Einstein-de Sitter closed forms of what the stepping uses.  With
    Omega_m = 1: D1 = a, D2 = a^2 (each normalised at a = 1), E = a^(-3/2), so
    Gf = D1' a^2 E = a^(3/2), gf = dGf/da = 3/2 a^(1/2), gf2 = 5 a^(3/2), and
    the FastPM factors of core.py:168-180 are
    kick = (af^(3/2) - ai^(3/2)) / (3/2 ac) and drift = (af - ai) / ac^(3/2).
    Dropping any of gf's three terms (D1'' a^2 E, D1' a^2 E', 2 D1' a^2 E)
    changes the 3/2 and fails here."""
    a = np.array([0.1, 0.3, 0.7, 1.0])
    g = O.Growth(1.0, a=a, a_normalize=1.0)
    np.testing.assert_allclose(g.Gf(a), a ** 1.5, rtol=1e-7)
    np.testing.assert_allclose(g.gf(a), 1.5 * a ** 0.5, rtol=1e-7)
    np.testing.assert_allclose(g.gf2(a), 5 * a ** 1.5, rtol=1e-7)
    s = O.Solver(8, 100.0, 1.0)
    for ai, ac, af in [(0.1, 0.1, (0.1 * 0.2) ** 0.5), (0.2, 0.25, 0.3), (0.5, 0.7, 1.0)]:
        assert s.kick_factor(ai, ac, af) == pytest.approx(
            (af ** 1.5 - ai ** 1.5) / (1.5 * ac), rel=1e-7)
        assert s.drift_factor(ai, ac, af) == pytest.approx((af - ai) / ac ** 1.5, rel=1e-7)


@pytest.mark.parametrize("Om,a_norm", [(0.3, 1.0), (0.3075, 0.5)])
def test_growth_lcdm_force_factor_identity(Om, a_norm):
    """gf = dGf/da, which the growth ODE (D1'' = -(2 + E'/E) D1' + 3/2 Om(a) D1)
    turns into gf(a) = 3/2 Om(a) a E(a) D1(a); with D1, D2 normalised
    separately by their raw values c1, c2 at a_normalize the same algebra gives
    gf2(a) = 3/2 Om(a) a E(a) (D2(a) - c1^2/c2 D1(a)^2).  The right-hand sides
    use neither second derivatives nor E', and c1, c2, D1, D2 come from the
    independent solve."""
    a = np.array([0.05, 0.1, 0.4, 0.8, 1.0])
    g = O.Growth(Om, a=a, a_normalize=a_norm)
    D1, D2, c1, c2 = independent_growth(Om, a, a_norm)
    D1, D2 = D1 / c1, D2 / c2
    E = np.sqrt(Om / a ** 3 + 1 - Om)
    Oma = Om / a ** 3 / E ** 2
    np.testing.assert_allclose(g.gf(a), 1.5 * Oma * a * E * D1, rtol=1e-7)
    np.testing.assert_allclose(g.gf2(a), 1.5 * Oma * a * E * (D2 - c1 ** 2 / c2 * D1 ** 2),
                               rtol=1e-7)


# ---------------------------------------------------------------- stepping
def test_autostages():
    """test_core.py:11-23."""
    for knots in [[0.1, 1.0], [1.0], [0.5], [0.1, 0.2, 1.0]]:
        stages = O.autostages(knots, astart=0.1, N=12)
        assert len(stages) == 12
        for k in knots:
            assert k in stages


def test_leapfrog_counts():
    """test_core.py:25-30."""
    assert len(list(O.leapfrog(np.linspace(0.1, 1.0, 2, endpoint=True)))) == 1 + 4
    assert len(list(O.leapfrog([1.0]))) == 1
    assert len(list(O.leapfrog_with_lpt([0.1, 0.5, 1.0]))) == 8


# ---------------------------------------------------------------- solver
def _pk(k):
    return 2e4 * k / (1 + (k / 0.02) ** 2) ** 1.3


def test_kick_identity():
    """test_core.py:46-62 on the oracle solver."""
    s = O.Solver(8, 128.0, 0.3075)
    dlin = s.linear(s.whitenoise(1234), _pk)
    q = O.lattice((8,) * 3, (128.0,) * 3)
    s1 = s.lpt(dlin, q, a=0.01, order=1)
    s2 = s.lpt(dlin, q, a=1.0, order=1)
    pt = O.Growth(0.3075, a=[0.01, 1.0], a_normalize=1.0)
    fac = 1 / (0.01 ** 2 * pt.E(0.01)) * (pt.Gf(1.0) - pt.Gf(0.01)) / pt.gf(0.01)
    np.testing.assert_allclose(s2.P, s1.P + fac * s1.F)


def test_janus_reversibility():
    """examples/janus.py:65-71: forward then leapfrog(stages[::-1]) recovers X."""
    s = O.Solver(8, 64.0, 0.3, B=2)
    dlin = s.linear(s.whitenoise(400), _pk)
    stages = np.linspace(0.1, 1.0, 6)
    lpt = s.lpt(dlin, O.lattice((8,) * 3, (64.0,) * 3), stages[0])
    st = s.nbody(lpt.copy(), O.leapfrog(stages))
    assert np.abs(st.X - lpt.X).max() > 0.1
    rev = s.nbody(st, O.leapfrog(stages[::-1]))
    np.testing.assert_allclose(rev.X, lpt.X, atol=1e-9)


def test_force_plane_wave():
    """Closed form of the PM force on a lattice shifted by h/4 along x and
    displaced by eps sin(k0 q): every particle stays in its cell, so the CIC
    density is delta_i = (s_{i-1} - s_i)/h exactly, a cosine about x_i - h/4 of
    amplitude A = -2 eps sin(k0 h/2)/h; the grid force is
    -1.5 Om (a(k0)/k0^2) A sin(k0 (x - h/4)) and the particle reads
    0.75 F(i) + 0.25 F(i+1) (to first order in eps)."""
    N, L, Om = 16, 64.0, 0.3
    h = L / N
    s = O.Solver(N, L, Om, B=1)
    q = O.lattice((N,) * 3, (L,) * 3)
    q[:, 0] += 0.25 * h
    st = O.State(q)
    k0 = 2 * np.pi * 2 / L
    eps = 1e-7
    st.S[:, 0] = eps * np.sin(k0 * q[:, 0])
    st.a = dict(S=1.0, P=1.0, F=None)
    s.force(st, 1.0, 1.0, 1.0)
    a = (8 * np.sin(k0 * h) - np.sin(2 * k0 * h)) / (6 * h)
    A = -2 * eps * np.sin(k0 * h / 2) / h
    i = np.round(q[:, 0] / h - 0.25).astype(int)
    Fg = lambda j: -1.5 * Om * (a / k0 ** 2) * A * np.sin(k0 * (j - 0.25) * h)  # noqa: E731
    expect = 0.75 * Fg(i) + 0.25 * Fg(i + 1)
    np.testing.assert_allclose(st.F[:, 0], expect, rtol=0, atol=1e-5 * np.abs(expect).max())
    assert np.abs(st.F[:, 1:]).max() < 1e-9 * np.abs(expect).max()
    # RHO reads 1 + delta at the particle
    d = lambda j: A * np.cos(k0 * (j - 0.25) * h)  # noqa: E731
    np.testing.assert_allclose(st.RHO - 1, 0.75 * d(i) + 0.25 * d(i + 1), rtol=0,
                               atol=1e-5 * abs(A))


def test_power_brute_force():
    rng = np.random.default_rng(5)
    nmesh, box = (8, 8, 6), (10.0, 10.0, 10.0)
    f1 = O.r2c(rng.standard_normal(nmesh))
    f2 = O.r2c(rng.standard_normal(nmesh))
    nb, kmin, dk = 5, 0.0, 0.4
    got = O.power_bins(f1, nmesh, box, nb, kmin, dk, f2)
    ref = np.zeros((nb, 5))
    for i, j, k in itertools.product(range(8), range(8), range(4)):
        if i == j == k == 0:
            continue
        ni = i if i < 4 else i - 8
        nj = j if j < 4 else j - 8
        kk = 2 * np.pi / 10.0 * np.sqrt(ni ** 2 + nj ** 2 + k ** 2)
        b = int(np.floor((kk - kmin) / dk))
        if not 0 <= b < nb:
            continue
        w = 1.0 if k in (0, 3) else 2.0
        a, c = f1[i, j, k], f2[i, j, k]
        ref[b] += [w, w * kk, w * (a * np.conj(c)).real, w * abs(a) ** 2, w * abs(c) ** 2]
    np.testing.assert_allclose(got, ref, rtol=1e-12)
    # the full-spectrum Parseval sum: sum over all modes of |f|^2 (weights
    # restore the missing half) equals the real-space variance / N
    allb = O.power_bins(f1, nmesh, box, 1, 0.0, 100.0)
    full = np.fft.fftn(np.fft.irfftn(f1, s=nmesh, axes=(0, 1, 2)))
    assert allb[0, 3] == pytest.approx((np.abs(full) ** 2).sum() - abs(full[0, 0, 0]) ** 2, rel=1e-12)


def _glass_power_ratio(X, nc, box):
    """P(k) of a point set over that of uniform random points, on a 2nc mesh."""
    nm = (2 * nc,) * 3
    def pk(Y):
        rho = O.paint(Y, nm, (box,) * 3, 2) / (len(Y) / np.prod(nm))
        return O.power_spectrum(O.r2c(rho), nm, (box,) * 3, nbins=8)[1]
    rng = np.random.default_rng(1)
    return pk(X) / pk(rng.uniform(0, box, size=X.shape))


def test_glass_oracle():
    """Glass oracle (fastpm/glass.py): as many points as grid cells
    (test_glass.py), all in [0, L), large-scale power far below a random
    set's (a repulsive, damped relaxation; an attractive force or a missing
    damping term fails this), the stage list of glass.py:33, and the
    unperturbed lattice as an exact fixed point of the dynamics."""
    X = O.glass(8, 128.0, 123)
    assert X.shape == (8 ** 3, 3)
    assert X.min() >= 0 and X.max() < 128.0
    ratio = _glass_power_ratio(X, 8, 128.0)
    assert np.all(ratio[1:4] < 0.02), ratio
    st = O.glass_stages(3)
    assert len(st) == 27 and st[0] == 0 and abs(st[-1] - 6.5 * np.pi) < 1e-12
    X0 = O.glass(8, 128.0, 123, spread=0.0)
    assert np.array_equal(X0, O.lattice((8,) * 3, (128.0,) * 3) % 128.0)


@pytest.mark.parametrize("split,deconv", [(0.0, 0), (3.0, 0), (2.5, 2)])
def test_longrange_plane_wave(split, deconv):
    """gravity.py:11-29 on a density cosine A cos(k0 x): pot = delta_k / k^2
    (no sign), so the x force on the mesh is
    -factor a(k0) A / k0^2 sin(k0 x) exp(-k0^2 r^2) / sinc(k0 h / 2)^deconv
    with the four-point a(k0) = (8 sin k0 h - sin 2 k0 h) / (6 h), and zero
    along y and z.  Read out at grid points (exact) and half a cell along x
    (CIC: the mean of the two neighbours), as two position sets."""
    N, L, factor = 16, 64.0, 0.45
    h = L / N
    k0 = 2 * np.pi * 3 / L
    x = np.arange(N) * h
    A = 0.3
    rho = A * np.cos(k0 * x)[:, None, None] * np.ones((N, N, N))
    dk = O.r2c(rho)
    a = (8 * np.sin(k0 * h) - np.sin(2 * k0 * h)) / (6 * h)
    amp = -factor * a * A / k0 ** 2 * np.exp(-k0 ** 2 * split ** 2)
    if deconv:
        amp /= (np.sin(k0 * h / 2) / (k0 * h / 2)) ** deconv
    q0 = O.lattice((N,) * 3, (L,) * 3)
    q1 = q0 + np.array([0.5 * h, 0.25 * h, 0.0])
    f0, f1 = O.longrange([q0, q1], dk, (N,) * 3, (L,) * 3, split, factor, 2, deconv)
    np.testing.assert_allclose(f0[:, 0], amp * np.sin(k0 * q0[:, 0]), atol=1e-13)
    mid = 0.5 * (np.sin(k0 * q0[:, 0]) + np.sin(k0 * (q0[:, 0] + h)))
    np.testing.assert_allclose(f1[:, 0], amp * mid, atol=1e-13)
    assert np.abs(f0[:, 1:]).max() < 1e-13 and np.abs(f1[:, 1:]).max() < 1e-13


def test_longrange_is_solver_force():
    """At split 0 with factor 3/2 Om0 and the painted 1 + delta, longrange is
    the Solver's Force (core.py:195-210)."""
    s = O.Solver(8, 64.0, 0.3, B=2)
    st = s.lpt(s.linear(s.whitenoise(5), _pk), O.lattice((8,) * 3, (64.0,) * 3), 0.5)
    X = st.X
    nm = (16,) * 3
    rho = O.paint(X, nm, (64.0,) * 3, 2) / (len(X) / 16 ** 3)
    f = O.longrange([X], O.r2c(rho), nm, (64.0,) * 3, 0.0, 1.5 * 0.3)[0]
    s.force(st, 0.5, 0.5, 0.5)
    np.testing.assert_allclose(f, st.F, rtol=1e-12, atol=1e-14 * np.abs(st.F).max())
