"""The reference's own CPU-runnable case at its full size (GPU): plain FastPM,
L = 256 Mpc/h, 64^3 particles on a 128^3 mesh, Eisenstein-Hu no-wiggle
P(k) (Om0 0.3, h 0.7, ns 0.96, sigma8 0.8), white-noise seed 42, 2LPT at
a = 0.1, 10 KDK steps linear in a to a = 1, CIC.  The outputs the run
produces -- positions, velocities, the density field and P(k) in 32 linear
k bins -- are compared with the FP64 oracle run on the same modes: float64
to the growth ODE's tolerance, float32 against the P(k) / r(k) gates."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu

NC, BOX, SEED = 64, 256.0, 42
STAGES = np.linspace(0.1, 1.0, 11)


def _cosmo():
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=0.8)
    return cosmo, LinearPower(cosmo)


@pytest.fixture(scope="module")
def oracle_run():
    """The oracle's final state, density field (unit-mass CIC paint of X on
    the particle mesh, as StateVector.to_mesh) and its spectrum sums."""
    from fastpm.core import leapfrog

    cosmo, Pk = _cosmo()
    osolver = O.Solver(NC, BOX, cosmo.Om0, B=2)
    st = osolver.lpt(osolver.linear(osolver.whitenoise(SEED), Pk),
                     O.lattice((NC,) * 3, (BOX,) * 3), a=STAGES[0])
    osolver.nbody(st, leapfrog(STAGES))
    dens = O.paint(st.X % BOX, (NC,) * 3, (BOX,) * 3, 2)
    return st, dens


def _run(dtype):
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh

    cosmo, Pk = _cosmo()
    pm = ParticleMesh(Nmesh=[NC] * 3, BoxSize=BOX, dtype=dtype)
    solver = Solver(pm, cosmo, B=2)
    state = solver.lpt(solver.linear(solver.whitenoise(SEED), Pk),
                       pm.generate_uniform_particle_grid(shift=0), a=STAGES[0])
    solver.nbody(state, leapfrog(STAGES))
    return state, state.to_mesh()


# 32 linear bins up to just below the Nyquist wave number: with the edges at
# exact multiples of the fundamental mode, modes on an edge would fall on
# either side by rounding (1 ulp) in any two implementations
KMAX = np.pi * NC / BOX * (1 - 1e-6)


def _spectrum_sums(dens, other=None):
    return O.power_bins(O.r2c(dens), (NC,) * 3, (BOX,) * 3, 32, 0.0, KMAX / 32,
                        None if other is None else O.r2c(other))


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


def test_config1_fp64_matches_oracle(gpu, oracle_run):
    from fastpm.utils import power_sums

    print("config1 f64", flush=True)
    ost, odens = oracle_run
    state, dens = _run("f8")
    X = state.X.cpu().numpy()
    dx = (X - ost.X + BOX / 2) % BOX - BOX / 2
    assert np.abs(dx).max() < 1e-6 * BOX / NC
    V = state.V.cpu().numpy()
    oV = ost.P * (100.0 / ost.a["P"])
    assert _rel(V, oV) < 1e-6
    assert _rel(state.RHO.cpu().numpy(), ost.RHO) < 1e-6
    assert _rel(dens.value.cpu().numpy(), odens) < 1e-6
    got = power_sums(dens.r2c(), nbins=32, kmin=0.0, kmax=KMAX)
    ref = _spectrum_sums(odens)
    assert np.array_equal(got[:, 0], ref[:, 0])  # mode counts per bin
    full = ref[:, 0] > 0  # the first bin holds no mode
    P, oP = got[full, 2] / got[full, 0], ref[full, 2] / ref[full, 0]
    assert np.abs(P / oP - 1).max() < 1e-6


def test_config1_fp32_gates(gpu, oracle_run):
    """Float32 (integer lattice + float32 displacement) against the FP64
    oracle: P(k) relative error <= 1e-3 for k < 0.5 k_Nyq and r(k) >= 0.999
    on the density field; the largest displacement error is printed."""
    print("config1 f32", flush=True)
    ost, odens = oracle_run
    state, dens = _run("f4")
    d32 = dens.value.double().cpu().numpy()
    s32 = _spectrum_sums(d32)
    cross = _spectrum_sums(d32, odens)
    ref = _spectrum_sums(odens)
    full = ref[:, 0] > 0
    k = np.where(full, ref[:, 1] / np.maximum(ref[:, 0], 1), np.inf)
    sel = k < 0.5 * np.pi * NC / BOX
    perr = np.abs((s32[sel, 2] / s32[sel, 0]) / (ref[sel, 2] / ref[sel, 0]) - 1)
    r = cross[sel, 2] / np.sqrt(cross[sel, 3] * cross[sel, 4])
    X = state.X.double().cpu().numpy()
    dx = np.abs((X - ost.X + BOX / 2) % BOX - BOX / 2).max()
    print("config1 f32: P(k) err %.3e, r(k) min %.12f, max |dx| %.3e cells"
          % (perr.max(), r.min(), dx * NC / BOX), flush=True)
    assert perr.max() < 1e-3
    assert r.min() > 0.999
    assert dx < 0.01 * BOX / NC
