"""End-to-end parity of the Solver path (whitenoise -> linear -> lpt -> nbody)
against the FP64 oracle Solver (GPU only).

Float64 runs agree with the oracle to the tolerance of the growth ODE: the
oracle restates the reference's odeint solve (relative tolerance ~1.5e-8)
while the product integrates with fixed-step Dormand-Prince 5, so end-to-end
agreement is checked at rtol 1e-6 of each array's scale.  Float32 runs are checked against the
north-star gates at small sizes: P(k) relative error <= 1e-3 for
k < 0.5 k_Nyq and r(k) >= 0.999."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _setup(nc, box, B, dtype, resampler="cic", compensate=False, seed=1234, sigma8=0.8):
    from fastpm.core import Solver
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=sigma8)
    Pk = LinearPower(cosmo)
    pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=box, dtype=dtype, resampler=resampler)
    solver = Solver(pm, cosmo, B=B, compensate=compensate)
    osolver = O.Solver(nc, box, cosmo.Om0, B=B, resampler=resampler,
                       deconv=(2 if resampler == "cic" else 3) if compensate else 0)
    return solver, osolver, Pk, pm


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("cfg", [
    dict(nc=8, box=128.0, B=1, resampler="cic", compensate=False, stages=[0.1, 0.5, 1.0]),
    dict(nc=16, box=100.0, B=2, resampler="tsc", compensate=True, stages=[0.1, 0.3, 0.6, 1.0]),
    dict(nc=12, box=90.0, B=2, resampler="cic", compensate=True, stages=[0.2, 0.5, 1.0]),
])
def test_solver_f64_matches_oracle(gpu, cfg):
    from fastpm.core import leapfrog

    print("solver f64", cfg, flush=True)
    solver, osolver, Pk, pm = _setup(cfg["nc"], cfg["box"], cfg["B"], "f8", cfg["resampler"],
                                     cfg["compensate"])
    wn = solver.whitenoise(1234)
    dlin = solver.linear(wn, Pk)
    Q = pm.generate_uniform_particle_grid(shift=0)
    state = solver.lpt(dlin, Q, a=cfg["stages"][0], order=2)

    owhite = osolver.whitenoise(1234)
    odlin = osolver.linear(owhite, Pk)
    q = O.lattice((cfg["nc"],) * 3, (cfg["box"],) * 3)
    ostate = osolver.lpt(odlin, q, a=cfg["stages"][0], order=2)
    for key in "SPF":
        assert _rel(getattr(state, key).cpu().numpy(), getattr(ostate, key)) < 1e-6, key

    events = []
    solver.nbody(state, leapfrog(cfg["stages"]), monitor=lambda *a: events.append(a[0]))
    osolver.nbody(ostate, leapfrog(cfg["stages"]))
    assert events == [a[0] for a in leapfrog(cfg["stages"])]
    X = state.X.cpu().numpy()
    dx = (X - ostate.X + cfg["box"] / 2) % cfg["box"] - cfg["box"] / 2
    assert np.abs(dx).max() < 1e-6 * cfg["box"] / cfg["nc"]
    assert _rel(state.P.cpu().numpy(), ostate.P) < 1e-6
    assert _rel(state.F.cpu().numpy(), ostate.F) < 1e-6
    assert _rel(state.RHO.cpu().numpy(), ostate.RHO) < 1e-6
    assert state.a == ostate.a


def test_kick_identity(gpu):
    """test_core.py:46-62: for 1LPT, P(a=1) == P(0.01) + fac F(0.01)."""
    from fastpm.background import MatterDominated

    solver, _, Pk, pm = _setup(8, 128.0, 1, "f8")
    dlin = solver.linear(solver.whitenoise(1234), Pk)
    Q = pm.generate_uniform_particle_grid(shift=0)
    s1 = solver.lpt(dlin, Q, a=0.01, order=1)
    s2 = solver.lpt(dlin, Q, a=1.0, order=1)
    pt = MatterDominated(solver.cosmology.Om0, a=[0.01, 1.0], a_normalize=1.0)
    fac = 1 / (0.01 ** 2 * pt.E(0.01)) * (pt.Gf(1.0) - pt.Gf(0.01)) / pt.gf(0.01)
    np.testing.assert_allclose(s2.P.cpu().numpy(), (s1.P + fac * s1.F).cpu().numpy(), rtol=1e-7,
                               atol=1e-12 * float(s2.P.abs().max()))


def test_janus_reversibility(gpu):
    """examples/janus.py:65-71: leapfrog forward then leapfrog(stages[::-1])
    recovers the initial positions (float64)."""
    from fastpm.core import leapfrog

    solver, _, Pk, pm = _setup(16, 64.0, 2, "f8")
    dlin = solver.linear(solver.whitenoise(400), Pk)
    Q = pm.generate_uniform_particle_grid()
    stages = np.linspace(0.1, 1.0, 10, endpoint=True)
    lpt = solver.lpt(dlin, Q, stages[0])
    state = solver.nbody(lpt.copy(), leapfrog(stages))
    moved = np.abs((state.X - lpt.X).cpu().numpy()).max()
    reverse = solver.nbody(state, leapfrog(stages[::-1]))
    err = np.abs((reverse.X - lpt.X).cpu().numpy()).max()
    assert moved > 0.1
    assert err < 1e-7 * 64.0


def test_fp32_gates(gpu):
    """Float32 vs float64 on the same inputs: P(k) rel err <= 1e-3 for
    k < 0.5 k_Nyq and r(k) >= 0.999 at the final step."""
    from fastpm.core import leapfrog
    from fastpm.utils import power

    res = {}
    for dt in ("f8", "f4"):
        solver, _, Pk, pm = _setup(32, 200.0, 2, dt, "cic", True, seed=7)
        dlin = solver.linear(solver.whitenoise(7), Pk)
        st = solver.lpt(dlin, pm.generate_uniform_particle_grid(), 0.1)
        solver.nbody(st, leapfrog(np.linspace(0.1, 1.0, 11)))
        res[dt] = (st.X.double(), st.to_mesh())
    Xd, md = res["f8"]
    Xs, ms = res["f4"]
    pmd = md.pm
    k, p64, _ = power(md.r2c())
    fs = ms.value.double()
    from fastpm.pm import RealField
    ms64 = RealField(pmd, fs.contiguous())
    k2, p32, _, r = power(ms64.r2c(), md.r2c())
    _, p32a, _ = power(ms64.r2c())
    knyq = np.pi * 32 / 200.0
    sel = k < 0.5 * knyq
    assert np.nanmax(np.abs(p32a[sel] / p64[sel] - 1)) < 1e-3
    assert np.nanmin(r[sel]) > 0.999
    dx = ((Xs - Xd + 100.0) % 200.0 - 100.0).abs().max().item()
    assert dx < 0.01 * 200.0 / 32


def test_state_api(gpu, tmp_path):
    from fastpm.core import leapfrog

    solver, _, Pk, pm = _setup(8, 64.0, 2, "f4")
    dlin = solver.linear(solver.whitenoise(5), Pk)
    st = solver.lpt(dlin, pm.generate_uniform_particle_grid(), 0.5)
    c = st.copy()
    assert c.Q is st.Q and c.S is not st.S
    P_alias = st.P
    solver.nbody(st, leapfrog([0.5, 1.0]))
    assert P_alias is st.P and not torch.equal(P_alias, c.P)
    assert st.synchronized
    with pytest.raises(AssertionError):
        solver.nbodystep.Kick(st, 0.5, 0.7, 1.0)
    with pytest.raises(KeyError):
        solver.nbodystep.run("X", 1, 1, 1, st, None)
    with pytest.raises(AssertionError):
        solver.lpt(dlin, pm.generate_uniform_particle_grid(), 0.5, order=3)
    m = st.to_mesh()
    assert abs(float(m.value.double().sum()) - 8 ** 3) < 1e-2
    st.save(str(tmp_path / "snap"))
    from fastpm.snapshot import Block

    X = Block(str(tmp_path / "snap"), "1/Position").read()
    assert X.shape == (512, 3) and X.dtype == np.float32
    assert np.array_equal(X, st.X.cpu().numpy())
    head = Block(str(tmp_path / "snap"), "Header").attrs()
    assert float(head["Time"][0]) == st.a["S"] and float(head["Om0"][0]) == 0.3


def test_pipeline_monitor_outputs(gpu, tmp_path):
    """The CLI outputs (main.py:81-113): power-linear.txt, one power-<a>.txt per
    Force with columns k, P, N, P/D^2 equal to the oracle estimator on the same
    delta_k, and a snapshot at the output epoch."""
    from fastpm.core import leapfrog
    from fastpm.utils import PipelineMonitor

    solver, _, Pk, pm = _setup(8, 64.0, 2, "f8")
    dlin = solver.linear(solver.whitenoise(11), Pk)
    st = solver.lpt(dlin, pm.generate_uniform_particle_grid(), 0.5)
    mon = PipelineMonitor(solver, str(tmp_path), aout=[1.0], verbose=False)
    mon.write_linear(dlin)
    seen = []

    def monitor(action, ai, ac, af, state, event):
        mon(action, ai, ac, af, state, event)
        if action == "F":
            seen.append((state.a["F"], event["delta_k"].value.cpu().numpy().copy()))

    solver.nbody(st, leapfrog([0.5, 0.75, 1.0]), monitor=monitor)
    assert len(seen) == 3
    nm = (16, 16, 16)
    growth = O.Growth(0.3, a=[0.5, 0.75, 1.0])
    for a, dk in seen:
        got = np.loadtxt(tmp_path / ("power-%06.4f.txt" % a))
        assert got.shape == (8, 4)
        k, p, n, _ = O.power_spectrum(dk, nm, (64.0,) * 3, nbins=8)
        ok = n > 0
        assert np.array_equal(got[:, 2], n)
        assert np.allclose(got[ok, 0], k[ok], rtol=1e-12)
        assert np.allclose(got[ok, 1], p[ok], rtol=1e-10)
        assert np.allclose(got[ok, 3], p[ok] / growth.D1(a) ** 2, rtol=1e-6)
    lin = np.loadtxt(tmp_path / "power-linear.txt")
    assert lin.shape == (4, 4)
    snap = tmp_path / "fpm-1.0000"
    assert (snap / "Header" / "attr-v2").exists() and (snap / "1" / "Position" / "000000").exists()


def test_state_restart_bitwise(gpu, tmp_path):
    """save -> load -> continue reproduces an uninterrupted float32 run bit for
    bit (tiled deterministic paint, same kernels)."""
    from fastpm.core import StateVector, leapfrog

    solver, _, Pk, pm = _setup(8, 64.0, 2, "f4")
    dlin = solver.linear(solver.whitenoise(3), Pk)
    Q = pm.generate_uniform_particle_grid()
    ref = solver.lpt(dlin, Q, 0.3)
    solver.nbody(ref, leapfrog([0.3, 0.5, 0.7, 1.0]))
    st = solver.lpt(dlin, Q, 0.3)
    solver.nbody(st, leapfrog([0.3, 0.5]))
    st.save(str(tmp_path / "ck"))
    back = StateVector.load(solver, str(tmp_path / "ck"), Q)
    assert back.a == st.a
    steps = list(leapfrog([0.5, 0.7, 1.0]))[1:]  # the first Force is in the checkpoint
    solver.nbody(back, steps)
    for key in "SPF":
        assert torch.equal(getattr(back, key), getattr(ref, key)), key


def test_monitor_sees_delta_k_and_unmonitored_run_matches(gpu):
    """Float32 stencil path: a run without a monitor (its Forces skip storing
    the spectrum) ends in the same state, bit for bit, as a run whose monitor
    receives each Force's delta_k; every delta_k is the forward transform of
    the painted 1 + delta, so its zero mode is 1 (core.py:161-166,195-210)."""
    from fastpm.core import leapfrog

    solver, _, Pk, pm = _setup(16, 100.0, 2, "f4")
    assert solver.nbodystep.stencil
    dlin = solver.linear(solver.whitenoise(5), Pk)
    Q = pm.generate_uniform_particle_grid()
    a = solver.lpt(dlin, Q, 0.1)
    b = solver.lpt(dlin, Q, 0.1)
    seen = []

    def monitor(action, ai, ac, af, state, event):
        if action == "F":
            seen.append(complex(event["delta_k"].value[0, 0, 0].item()))

    stages = [0.1, 0.4, 0.7, 1.0]
    solver.nbody(a, leapfrog(stages), monitor=monitor)
    solver.nbody(b, leapfrog(stages))
    for key in ("S", "P", "F", "RHO"):
        assert torch.equal(getattr(a, key), getattr(b, key)), key
    assert len(seen) == len(stages)
    assert all(abs(z - 1) < 1e-5 for z in seen), seen


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_monitor_may_keep_delta_k(gpu, dtype):
    """A monitor that keeps the event objects past the callback still holds
    each Force's own delta_k after the run: every Force makes a new field
    (core.py:205), none is overwritten by a later Force.  Float32 takes the
    stencil path, float64 the k-space path."""
    from fastpm.core import leapfrog

    solver, _, Pk, pm = _setup(16, 100.0, 2, dtype)
    assert solver.nbodystep.stencil == (dtype == "f4")
    st = solver.lpt(solver.linear(solver.whitenoise(8), Pk), pm.generate_uniform_particle_grid(),
                    0.1)
    kept, copies = [], []

    def monitor(action, ai, ac, af, state, event):
        if action == "F":
            kept.append(event)
            copies.append(event["delta_k"].value.clone())

    solver.nbody(st, leapfrog([0.1, 0.4, 0.7, 1.0]), monitor=monitor)
    assert len(kept) == 4
    fields = [e["delta_k"] for e in kept]
    assert len({id(f._t) for f in fields}) == 4
    for f, c in zip(fields, copies):
        assert torch.equal(f.value, c)
    for i in range(3):
        assert not torch.equal(copies[i], copies[i + 1])
