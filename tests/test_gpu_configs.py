"""The benchmark configurations as parity tests (GPU only).

* 256^3 particles on a 512^3 mesh (L = 1000 Mpc/h, CIC with window
  compensation, seed 7, 2LPT at a = 0.1, 40 steps) at full size: the
  autotuned float32 run against the float64 plain run on the GPU (the float64
  path is itself checked against the oracle below and in test_gpu_solver).
* The same configuration scaled to 32^3 particles on a 64^3 mesh (the box
  scaled with the grid, L = 125 Mpc/h, so the resolution is the same) and the
  large-box configuration (TSC with compensation, 2LPT at a = 0.05) scaled the
  same way: float64 against the FP64 oracle Solver (rtol 1e-6, the oracle's
  odeint tolerance) and float32 against the oracle through the gates
  P(k) relative error <= 1e-3 for k < k_Nyq / 2, r(k) >= 0.999.
* The batched ensemble (64 boxes of {32..128}^3 particles, sides U[100, 800],
  sigma8 U[0.7, 0.9], seeds 0-63, 20 steps, float32): sampled boxes of
  different size, side and sigma8 against the oracle through the same gates,
  and the full specification as a property test: every box advanced by graph
  replay equals its own single-box Solver run bit for bit (the reference's
  decomposition invariance, test_core.py:64-93, as batch invariance).
"""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _cosmo():
    from fastpm.cosmology import Cosmology

    return Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=0.8, Ob0=0.05, Tcmb0=2.7255)


_ORACLE = {}


def _oracle_run(nc, box, B, resampler, a0, nsteps, seed, cosmo, Pk):
    """Final oracle state of one configuration (cached per module)."""
    key = (nc, box, B, resampler, a0, nsteps, seed, cosmo.sigma8)
    if key not in _ORACLE:
        s = O.Solver(nc, box, cosmo.Om0, B=B, resampler=resampler,
                     deconv=2 if resampler == "cic" else 3)
        st = s.lpt(s.linear(s.whitenoise(seed), Pk), O.lattice((nc,) * 3, (box,) * 3), a0)
        s.nbody(st, O.leapfrog(np.linspace(a0, 1.0, nsteps + 1)))
        _ORACLE[key] = st
    return _ORACLE[key]


def _gates_vs_oracle(X, Xo, nmesh, box):
    """P(k) relative error (k < k_Nyq / 2), min r(k) and the largest periodic
    displacement difference in cells, both densities painted with CIC by the
    oracle and binned by the oracle estimator."""
    nm = (nmesh,) * 3
    bx = (box,) * 3
    ck = O.r2c(O.paint(X, nm, bx, 2))
    co = O.r2c(O.paint(Xo, nm, bx, 2))
    k, p, _, _ = O.power_spectrum(ck, nm, bx)
    _, po, n, _ = O.power_spectrum(co, nm, bx)
    _, _, _, r = O.power_spectrum(ck, nm, bx, f2=co)
    sel = (k < 0.5 * np.pi * nmesh / box) & (n > 0)
    dx = (X - Xo + box / 2) % box - box / 2
    return (float(np.abs(p[sel] / po[sel] - 1).max()), float(r[sel].min()),
            float(np.abs(dx).max() * nmesh / box))


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("cfg", [
    dict(name="256c_512m_cic_scaled", resampler="cic", a0=0.1),
    dict(name="512c_1024m_tsc_scaled", resampler="tsc", a0=0.05),
])
def test_config_scaled_vs_oracle(gpu, cfg, dtype):
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import LinearPower

    nc, box, nsteps, seed = 32, 125.0, 40, 7
    print("config", cfg["name"], dtype, flush=True)
    cosmo = _cosmo()
    Pk = LinearPower(cosmo)
    pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=box, dtype=dtype, resampler=cfg["resampler"])
    solver = Solver(pm, cosmo, B=2, compensate=True, autotune=True)
    if dtype == "f4":
        # the benchmark's float32 path: binned paint, fused transform, stencil readout
        assert solver.nbodystep.stencil
    st = solver.lpt(solver.linear(solver.whitenoise(seed), Pk),
                    pm.generate_uniform_particle_grid(), cfg["a0"])
    solver.nbody(st, leapfrog(np.linspace(cfg["a0"], 1.0, nsteps + 1)))
    ost = _oracle_run(nc, box, 2, cfg["resampler"], cfg["a0"], nsteps, seed, cosmo, Pk)
    X = st.X.cpu().numpy().astype("f8")
    if dtype == "f8":
        for key in "SP":
            got = getattr(st, key).cpu().numpy()
            ref = getattr(ost, key)
            assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max(), key
    perr, rmin, derr = _gates_vs_oracle(X, ost.X, 2 * nc, box)
    print("  P(k) err %.2e  r(k) min 1-%.1e  max |dx| %.2e cells" % (perr, 1 - rmin, derr),
          flush=True)
    assert perr <= 1e-3 and rmin >= 0.999
    assert derr < (1e-6 if dtype == "f8" else 1e-2)


def test_config_256c_512m_full_size_fp32_vs_fp64(gpu):
    """The headline configuration at full size: autotuned float32 stencil path
    against the float64 plain path, final density P(k) within 1e-3 for
    k < k_Nyq / 2 and r(k) >= 0.999 (product estimator, FP64 binning)."""
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh, RealField
    from fastpm.cosmology import LinearPower
    from fastpm.utils import power

    nc, box, a0, nsteps, seed = 256, 1000.0, 0.1, 40, 7
    cosmo = _cosmo()
    Pk = LinearPower(cosmo)
    states, solvers = {}, {}
    for dtype, fused in (("f4", True), ("f8", False)):
        print("full size", dtype, flush=True)
        pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=box, dtype=dtype)
        solver = Solver(pm, cosmo, B=2, compensate=True, fused=fused, autotune=fused)
        st = solver.lpt(solver.linear(solver.whitenoise(seed), Pk),
                        pm.generate_uniform_particle_grid(), a0)
        solver.nbody(st, leapfrog(np.linspace(a0, 1.0, nsteps + 1)))
        states[dtype], solvers[dtype] = st, solver
    fpm = solvers["f8"].fpm
    m64 = fpm.paint(states["f8"].particles, resampler="cic")
    m32 = solvers["f4"].fpm.paint(states["f4"].particles, resampler="cic")
    c64 = m64.r2c()
    c32 = RealField(fpm, m32.value.double().contiguous()).r2c()
    del m32, m64
    k, p64, _ = power(c64, nbins=32)
    _, p32, _ = power(c32, nbins=32)
    _, _, _, r = power(c32, c64, nbins=32)
    sel = (k < 0.5 * np.pi * 2 * nc / box) & np.isfinite(p64)
    perr = float(np.abs(p32[sel] / p64[sel] - 1).max())
    rmin = float(r[sel].min())
    dx = (states["f4"].X.double() - states["f8"].X + box / 2) % box - box / 2
    derr = float(dx.abs().max()) * 2 * nc / box
    print("  P(k) err %.2e  r(k) min 1-%.1e  max |dx| %.2e cells" % (perr, 1 - rmin, derr),
          flush=True)
    assert perr <= 1e-3 and rmin >= 0.999
    assert derr < 1e-2


def _ensemble_specs():
    from fastpm.ensemble import config_specs

    return config_specs(64)


def test_ensemble_sampled_boxes_vs_oracle(gpu):
    """Boxes drawn from the 64-box specification (two of 32^3 and one of 48^3
    particles, each with its own side, sigma8 and seed), advanced together by
    graph replay, each against the oracle Solver run with that box's
    parameters."""
    from fastpm.ensemble import Ensemble
    from fastpm.cosmology import LinearPower

    specs = _ensemble_specs()
    pick = [s for s in specs if s["nc"] == 32][:2] + [s for s in specs if s["nc"] == 48][:1]
    assert len(pick) == 3 and len({s["box"] for s in pick}) == 3
    assert len({s["cosmo"].sigma8 for s in pick}) == 3
    stages = np.linspace(0.1, 1.0, 21)
    ens = Ensemble(pick, stages)
    ens.initialize()
    ens.run(use_graph=True)
    for b, s in zip(ens.boxes, pick):
        print("ensemble box nc=%d L=%.1f sigma8=%.3f seed=%d" % (
            s["nc"], s["box"], s["cosmo"].sigma8, s["seed"]), flush=True)
        assert b.nc == s["nc"] and float(b.pm.BoxSize[0]) == s["box"]
        ost = _oracle_run(s["nc"], s["box"], 2, "cic", 0.1, 20, s["seed"], s["cosmo"],
                          LinearPower(s["cosmo"]))
        perr, rmin, derr = _gates_vs_oracle(b.state.X.cpu().numpy().astype("f8"), ost.X,
                                            2 * s["nc"], s["box"])
        print("  P(k) err %.2e  r(k) min 1-%.1e  max |dx| %.2e cells" % (perr, 1 - rmin, derr),
              flush=True)
        assert perr <= 1e-3 and rmin >= 0.999 and derr < 1e-2


def test_ensemble_full_spec_equals_single_box_runs(gpu):
    """All 64 boxes of the specification advanced by one captured graph per
    step: each box's final S, P, F equal its own single-box Solver run bit for
    bit."""
    from fastpm.ensemble import Ensemble
    from fastpm.core import leapfrog
    from fastpm.cosmology import LinearPower

    specs = _ensemble_specs()
    assert len(specs) == 64 and {s["seed"] for s in specs} == set(range(64))
    assert {s["nc"] for s in specs} == {32, 48, 64, 96, 128}
    assert all(100.0 <= s["box"] <= 800.0 and 0.7 <= s["cosmo"].sigma8 <= 0.9 for s in specs)
    stages = np.linspace(0.1, 1.0, 21)
    ens = Ensemble(specs, stages)
    ens.initialize()
    ens.run(use_graph=True)
    torch.cuda.synchronize()
    for i, b in enumerate(ens.boxes):
        ref = b.solver.lpt(b.solver.linear(b.solver.whitenoise(b.seed), LinearPower(b.cosmo)),
                           b.pm.generate_uniform_particle_grid(), stages[0])
        b.solver.nbody(ref, leapfrog(stages))
        for key in "SPF":
            assert torch.equal(getattr(b.state, key), getattr(ref, key)), (i, b.nc, key)
        del ref
