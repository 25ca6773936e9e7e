"""Host-side logic of the product that runs without a GPU: stepping
generators, kick/drift factors and k-space operator composition."""
import numpy as np
import pytest

import fastpm_oracle as O


def test_leapfrog_matches_oracle_and_reference_counts():
    from fastpm.core import leapfrog, leapfrog_with_lpt

    for stages in ([0.1, 1.0], [1.0], np.linspace(0.1, 1.0, 11), [1.0, 0.5, 0.1]):
        assert list(leapfrog(stages)) == list(O.leapfrog(stages))
        assert list(leapfrog_with_lpt(stages)) == list(O.leapfrog_with_lpt(stages))
    assert len(list(leapfrog(np.linspace(0.1, 1.0, 2)))) == 1 + 4 * (2 - 1)
    assert len(list(leapfrog([1.0]))) == 1
    assert list(leapfrog([])) == []


def test_autostages_matches():
    from fastpm.core import autostages

    for knots in [[0.1, 1.0], [1.0], [0.5], [0.1, 0.2, 1.0]]:
        a = autostages(knots, astart=0.1, N=12)
        assert len(a) == 12
        for k in knots:
            assert k in a
        np.testing.assert_array_equal(a, O.autostages(knots, astart=0.1, N=12))


def test_kick_drift_factors_match_oracle():
    from fastpm.background import MatterDominated

    pt = MatterDominated(0.3075, a_normalize=1.0)
    osol = O.Solver(8, 100.0, 0.3075)
    for ai, ac, af in [(0.1, 0.1, 0.1 ** 0.5 * 0.2 ** 0.5), (0.2, 0.25, 0.3), (0.5, 0.7, 1.0)]:
        assert pt.kick_factor(ai, ac, af) == pytest.approx(osol.kick_factor(ai, ac, af), rel=1e-7)
        assert pt.drift_factor(ai, ac, af) == pytest.approx(osol.drift_factor(ai, ac, af), rel=1e-7)


def test_growth_api_matches_oracle():
    from fastpm.background import MatterDominated

    a = np.array([0.05, 0.3, 1.0])
    p = MatterDominated(0.3, a=a, a_normalize=0.5)
    g = O.Growth(0.3, a=a, a_normalize=0.5)
    for name in ("D1", "D2", "f1", "f2", "Gp", "gp", "Gf", "gf", "gf2", "E"):
        np.testing.assert_allclose(getattr(p, name)(a), getattr(g, name)(a), rtol=2e-7,
                                   err_msg=name)


def test_growth_eds_closed_forms_product():
    """The product's growth and stepping factors in Einstein-de Sitter
    (Omega_m = 1): gf = 3/2 a^(1/2), gf2 = 5 a^(3/2), Gf = a^(3/2),
    kick = (af^(3/2) - ai^(3/2)) / (3/2 ac), drift = (af - ai) / ac^(3/2)."""
    from fastpm.background import MatterDominated

    a = np.array([0.1, 0.3, 0.7, 1.0])
    p = MatterDominated(1.0, a_normalize=1.0)
    np.testing.assert_allclose(p.D1(a), a, rtol=1e-9)
    np.testing.assert_allclose(p.D2(a), a ** 2, rtol=1e-9)
    np.testing.assert_allclose(p.Gf(a), a ** 1.5, rtol=1e-9)
    np.testing.assert_allclose(p.gf(a), 1.5 * a ** 0.5, rtol=1e-9)
    np.testing.assert_allclose(p.gf2(a), 5 * a ** 1.5, rtol=1e-9)
    for ai, ac, af in [(0.1, 0.1, (0.1 * 0.2) ** 0.5), (0.2, 0.25, 0.3), (0.5, 0.7, 1.0)]:
        assert p.kick_factor(ai, ac, af) == pytest.approx(
            (af ** 1.5 - ai ** 1.5) / (1.5 * ac), rel=1e-9)
        assert p.drift_factor(ai, ac, af) == pytest.approx((af - ai) / ac ** 1.5, rel=1e-9)


@pytest.mark.parametrize("Om,a_norm", [(0.3, 1.0), (0.3075, 0.5)])
def test_growth_lcdm_force_factor_identity_product(Om, a_norm):
    """gf = 3/2 Om(a) a E D1 and gf2 = 3/2 Om(a) a E (D2 - c1^2/c2 D1^2)
    (consequences of the growth ODE; c1, c2 the raw D1, D2 at a_normalize)
    against an independent DOP853 solve, for the product's MatterDominated."""
    from fastpm.background import MatterDominated
    from test_oracle import independent_growth

    a = np.array([0.05, 0.1, 0.4, 0.8, 1.0])
    p = MatterDominated(Om, a_normalize=a_norm)
    D1, D2, c1, c2 = independent_growth(Om, a, a_norm)
    D1, D2 = D1 / c1, D2 / c2
    E = np.sqrt(Om / a ** 3 + 1 - Om)
    Oma = Om / a ** 3 / E ** 2
    np.testing.assert_allclose(p.D1(a), D1, rtol=1e-9)
    np.testing.assert_allclose(p.D2(a), D2, rtol=1e-9)
    np.testing.assert_allclose(p.gf(a), 1.5 * Oma * a * E * D1, rtol=1e-9)
    np.testing.assert_allclose(p.gf2(a), 1.5 * Oma * a * E * (D2 - c1 ** 2 / c2 * D1 ** 2),
                               rtol=1e-9)


def test_kernel_composition():
    from fastpm.force.kernels import laplace, gradient, longrange, compensation, KernelOp

    op = laplace * gradient(0, 1) * 1.5 * gradient(2, 1) * compensation(2) * longrange(3.0)
    k = op.fpm_kop()
    assert k.laplace == 1 and k.ngrad == 2 and list(k.grad_dir) == [0, 2]
    assert k.grad_order == 1 and k.deconv == 2 and k.r_split == 3.0 and k.factor == 1.5
    with pytest.raises(ValueError):
        laplace * laplace
    with pytest.raises(ValueError):
        gradient(0, 0) * gradient(1, 1)
    with pytest.raises(ValueError):
        gradient(0) * gradient(1) * gradient(2)
    with pytest.raises(ValueError):
        gradient(0, 3)
    assert isinstance(2.0 * laplace, KernelOp)


def test_linear_power_normalisation():
    from fastpm.cosmology import Cosmology, LinearPower, eh_nowiggle_transfer
    from scipy.integrate import simpson

    c = Cosmology(sigma8=0.8)
    Pk = LinearPower(c)
    # independent sigma_8 by Simpson on a log grid
    lk = np.linspace(np.log(1e-5), np.log(1e2), 200001)
    k = np.exp(lk)
    x = 8.0 * k
    w = 3 * (np.sin(x) - x * np.cos(x)) / x ** 3
    s8 = np.sqrt(simpson(k ** 3 * Pk(k) * w * w / (2 * np.pi ** 2), x=lk))
    assert s8 == pytest.approx(0.8, rel=1e-6)
    # transfer -> 1 at k -> 0, decreasing, ~ ln(k)/k^2 at large k
    T = eh_nowiggle_transfer(np.array([1e-6, 1e-3, 1e-1, 1.0, 10.0]), c)
    assert T[0] == pytest.approx(1.0, abs=1e-4)
    assert np.all(np.diff(T) < 0)
    assert Pk(0.0) == 0.0


def test_gpu_required():
    import torch
    from fastpm.pm import ParticleMesh

    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    with pytest.raises(RuntimeError):
        ParticleMesh(Nmesh=8, BoxSize=1.0)


def test_run_actions_groups_kicks_and_drift():
    """FastPMStep.run_actions fuses each K[,K],D run when no monitor is given
    (K D at the first step, K K D at every later boundary, the last K alone);
    on the tiled path each such run goes into the following Force's binning
    (the Force runs with the run's factors pending).  It runs everything one
    by one with a monitor, on an unfused solver, or when a time tag would
    fail an assertion.  The step's methods are replaced by recorders, so no
    GPU is needed."""
    import types
    from fastpm.core import FastPMStep, leapfrog

    class Rec(FastPMStep):
        def __init__(self, fused=True, refuse=False, tiled=False):
            self.solver = type("S", (), {"fused": fused})()
            self.log = []
            self.refuse = refuse
            self.tiled = tiled
            self._pending_kd = None

        def run(self, action, ai, ac, af, state, monitor):
            kd = self._pending_kd
            self.log.append(action if kd is None else "".join(kd[2]) + kd[3] + "+" + action)

        def _kick_drift_factors(self, state, kicks, drift):
            if self.refuse:
                return None
            return [k[0] for k in kicks], drift[0], kicks[-1][3], drift[3]

        def _kick_drift(self, state, facs, fac):
            self.log.append("".join(facs) + fac)

    def state():
        return types.SimpleNamespace(a={}, P=None, F=None)

    acts = list(leapfrog([0.1, 0.4, 0.7, 1.0]))
    s = Rec()
    s.run_actions(acts, state())
    assert s.log == ["F", "KD", "F", "KKD", "F", "KKD", "F", "K"]
    s = Rec(tiled=True)
    st = state()
    s.run_actions(acts, st)
    assert s.log == ["F", "KD+F", "KKD+F", "KKD+F", "K"]
    d = max(i for i, a in enumerate(acts) if a[0] == "D")
    assert s._pending_kd is None and st.a == {"P": acts[d - 1][3], "S": acts[d][3]}
    for step in (Rec(fused=False), Rec(refuse=True), Rec(fused=False, tiled=True),
                 Rec(refuse=True, tiled=True)):
        step.run_actions(iter(acts), state())
        assert step.log == [a[0] for a in acts]
    for tiled in (False, True):
        s = Rec(tiled=tiled)
        s.run_actions(acts, state(), monitor=lambda *args: None)
        assert s.log == [a[0] for a in acts]


def test_reverse_pass_fuses_drift_and_kicks():
    """ForwardModel.backward applies each drift and the kicks before it as
    one fused VJP (gF counted as zero after a force VJP instead of cleared),
    with the same result as one axpy per action.  The device calls are
    replaced by float64 CPU stand-ins, so no GPU is needed."""
    import types
    import torch
    import fastpm.adjoint as A
    from fastpm.core import leapfrog

    log = []

    def axpy(y, x, a):
        log.append("axpy")
        y.add_(a * x)

    def dkv(gS, gP, gF, drift, kicks, fresh):
        log.append("D" + "K" * len(kicks) + ("*" if fresh else ""))
        gP.add_(drift * gS)
        if kicks:
            f = torch.zeros_like(gF) if fresh else gF.clone()
            for k in kicks:
                f = f + k * gP
            gF.copy_(f)

    def fvjp(step, S, Q, gF, gS):
        log.append("F")
        gS.add_(0.3 * gF + 0.01 * S)

    def lvjp(solver, dlin, Q, a, gS, gP, gF, order):
        return types.SimpleNamespace(pm=None, vals=(gS.clone(), gP.clone(), gF.clone()))

    pt = types.SimpleNamespace(kick_factor=lambda ai, ac, af: 1.0 + ai + 2 * ac + 3 * af,
                               drift_factor=lambda ai, ac, af: 0.5 - ai + ac * af)
    stages = [0.1, 0.4, 0.7, 1.0]
    saved = (A._axpy, A._drift_kick_vjp, A.force_vjp, A.lpt_vjp, A._scale_by_table)
    A._axpy, A._drift_kick_vjp, A.force_vjp, A.lpt_vjp = axpy, dkv, fvjp, lvjp
    A._scale_by_table = lambda pm, G, tab: None
    try:
        out, logs = {}, {}
        for fuse in (True, False):
            m = A.ForwardModel.__new__(A.ForwardModel)
            m.fuse_vjp = fuse
            m.actions = list(leapfrog(stages))
            m.step = types.SimpleNamespace(_pt=pt)
            g = torch.Generator().manual_seed(3)
            m.checkpoints = [torch.randn(12, generator=g, dtype=torch.float64) for _ in range(4)]
            m.solver = m.dlin = m.Q = m._lin_tab = None
            m.stages, m.order = stages, 2
            gS = torch.randn(12, generator=g, dtype=torch.float64)
            gP = torch.randn(12, generator=g, dtype=torch.float64)
            gF = torch.randn(12, generator=g, dtype=torch.float64)
            log.clear()
            out[fuse] = m.backward(gS, gP, gF).vals
            logs[fuse] = list(log)
    finally:
        A._axpy, A._drift_kick_vjp, A.force_vjp, A.lpt_vjp, A._scale_by_table = saved
    assert logs[True] == ["axpy", "F", "DKK*", "F", "DKK*", "F", "DK*", "F"]
    assert logs[False] == ["axpy", "F"] + ["axpy"] * 3 + ["F"] + ["axpy"] * 3 + ["F"] + \
        ["axpy"] * 2 + ["F"]
    for a, b in zip(out[True], out[False]):
        assert torch.equal(a, b)


def test_mode_tensor_runs_numpy_style_kernels():
    """apply() callables written against numpy arrays (the reference's
    laplace_kernel and diff_kernel, operators.py:5-22, and the order-1
    gradient of force/kernels.py:17-24) give numpy's results on ModeTensor
    arguments (CPU tensors here; the device path is the same code)."""
    import torch
    from fastpm.pm import ModeTensor

    def laplace_kernel(k, v):
        kk = sum(ki ** 2 for ki in k)
        mask = (kk == 0).nonzero()
        kk[mask] = 1
        b = v / kk
        b[mask] = 0
        return b

    def diff_kernel(dir, k, v):
        mask = (v.i[dir] != v.Nmesh[dir] // 2)
        return v * (1j * k[dir] * mask)

    def grad4(dir, k, v):
        cellsize = v.BoxSize[dir] / v.Nmesh[dir]
        w = k[dir] * cellsize
        a = 1 / (6.0 * cellsize) * (8 * np.sin(w) - np.sin(2 * w))
        return v * (1j * a)

    rng = np.random.default_rng(0)
    n, L = (8, 6, 4), np.array([10.0, 12.0, 8.0])
    kn = [2 * np.pi * np.fft.fftfreq(n[0], 1.0 / n[0]) / L[0],
          2 * np.pi * np.fft.fftfreq(n[1], 1.0 / n[1]) / L[1],
          2 * np.pi * np.arange(n[2] // 2 + 1) / L[2]]
    kb = [kn[0][:, None, None], kn[1][None, :, None], kn[2][None, None, :]]
    vals = rng.standard_normal((8, 6, 3)) + 1j * rng.standard_normal((8, 6, 3))

    class V(np.ndarray):
        pass

    def numpy_args():
        v = vals.copy().view(V)
        v.i = [np.arange(8)[:, None, None], np.arange(6)[None, :, None],
               np.arange(3)[None, None, :]]
        v.Nmesh, v.BoxSize = np.array(n), L
        return [x.copy() for x in kb], v

    def torch_args():
        v = torch.from_numpy(vals.copy()).as_subclass(ModeTensor)
        v.i = [torch.from_numpy(x).as_subclass(ModeTensor) for x in numpy_args()[1].i]
        v.Nmesh, v.BoxSize = np.array(n), L
        return [torch.from_numpy(x.copy()).as_subclass(ModeTensor) for x in kb], v

    cases = [laplace_kernel, lambda k, v: diff_kernel(0, k, v),
             lambda k, v: diff_kernel(1, k, v), lambda k, v: grad4(2, k, v),
             lambda k, v: v * np.exp(-sum(ki ** 2 for ki in k) * 0.3)]
    for f in cases:
        want = np.asarray(f(*numpy_args()))
        got = f(*torch_args())
        assert isinstance(got, ModeTensor)
        np.testing.assert_allclose(got.as_subclass(torch.Tensor).numpy(), want, rtol=1e-14,
                                   atol=1e-14)
    with pytest.raises(TypeError):
        np.add.reduce(torch_args()[0][0])
