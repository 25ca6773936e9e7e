"""ComplexField.apply with callables f(k, v) as the reference writes them
(fastpm/core.py:116-118, fastpm/operators.py:5-22, fastpm/force/kernels.py)
and Solver.linear / the whole Solver on boxes whose sides differ, against the
oracle (GPU only)."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def laplace_kernel(k, v):
    kk = sum(ki ** 2 for ki in k)
    mask = (kk == 0).nonzero()
    kk[mask] = 1
    b = v / kk
    b[mask] = 0
    return b


def diff_kernel(dir, conjugate=False):
    def kernel(k, v):
        factor = -1j if conjugate else 1j
        mask = (v.i[dir] != v.Nmesh[dir] // 2)
        return v * (factor * k[dir] * mask)
    return kernel


def gradient4(dir):
    def kernel(k, v):
        cellsize = v.BoxSize[dir] / v.Nmesh[dir]
        w = k[dir] * cellsize
        a = 1 / (6.0 * cellsize) * (8 * np.sin(w) - np.sin(2 * w))
        return v * (1j * a)
    return kernel


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_apply_callables_vs_oracle(gpu, dtype):
    from fastpm.pm import ParticleMesh
    from fastpm.force.kernels import laplace, gradient

    nmesh, box = (16, 12, 8), (64.0, 48.0, 40.0)
    rng = np.random.default_rng(2)
    x = rng.standard_normal(nmesh)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    ck = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu)).r2c()
    xk = O.r2c(x)
    tol = 1e-12 if dtype == "f8" else 2e-6
    lap = O.k_laplace(nmesh, box)
    cases = [
        (laplace_kernel, xk * lap),
        (diff_kernel(0), xk * O.k_gradient(0, nmesh, box, 0)),
        (diff_kernel(1, conjugate=True), xk * np.conj(O.k_gradient(1, nmesh, box, 0))),
        (gradient4(2), xk * O.k_gradient(2, nmesh, box, 1)),
        (lambda k, v: v * np.exp(-sum(ki ** 2 for ki in k) * 4.0),
         xk * O.k_longrange(nmesh, box, 2.0)),
    ]
    for f, ref in cases:
        got = ck.apply(f).value.cpu().numpy()
        assert np.abs(got - ref).max() < tol * np.abs(ref).max()
    # the reference's chained form (operators.py:61-62) equals the fused
    # device operator, and out=Ellipsis works in place
    chain = ck.apply(laplace_kernel).apply(gradient4(1), out=Ellipsis)
    fused = ck.apply(laplace * gradient(1, order=1))
    assert torch.allclose(chain.value, fused.value, rtol=0,
                          atol=tol * float(fused.value.abs().max()))
    # kind='index' gives the integer wave numbers, 'circular' k h
    seen = {}

    def probe(k, v):
        seen["k"] = [t.as_subclass(torch.Tensor).reshape(-1).cpu().numpy() for t in k]
        return v

    ck.apply(probe, kind="index")
    assert list(seen["k"][0]) == list(np.fft.fftfreq(16, 1 / 16))
    assert list(seen["k"][2]) == [0, 1, 2, 3, 4]
    ck.apply(probe, kind="circular")
    np.testing.assert_allclose(seen["k"][1], 2 * np.pi * np.fft.fftfreq(12), rtol=1e-6)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_linear_anisotropic_box(gpu, dtype):
    """Solver.linear on a box whose sides differ equals the oracle's
    sqrt(P(|k|)/V) w_k (core.py:116-118) on every mode, and a cubic box gives
    the same as the integer-table path."""
    from fastpm.core import Solver
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology()
    Pk = LinearPower(cosmo)
    nmesh, box = (16, 24, 12), (100.0, 150.0, 80.0)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    s = Solver(pm, cosmo)
    wn = s.whitenoise(9)
    got = s.linear(wn, Pk).value.cpu().numpy()
    ref = O.linear(wn.value.cpu().numpy().astype("c16"), Pk, nmesh, box)
    assert np.abs(got - ref).max() < (1e-12 if dtype == "f8" else 1e-6) * np.abs(ref).max()


def test_solver_anisotropic_box_f64_vs_oracle(gpu):
    """The whole Solver (2LPT + 3 KDK steps) on an 8 x 12 x 16 particle
    lattice in a 64 x 96 x 128 Mpc/h box against the oracle Solver."""
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology()
    Pk = LinearPower(cosmo)
    nc, box = (8, 12, 16), (64.0, 96.0, 128.0)
    pm = ParticleMesh(Nmesh=nc, BoxSize=box, dtype="f8")
    s = Solver(pm, cosmo, B=2)
    st = s.lpt(s.linear(s.whitenoise(4), Pk), pm.generate_uniform_particle_grid(), 0.1)
    stages = [0.1, 0.4, 0.7, 1.0]
    s.nbody(st, leapfrog(stages))
    os_ = O.Solver(nc, box, cosmo.Om0, B=2)
    ost = os_.lpt(os_.linear(os_.whitenoise(4), Pk), O.lattice(nc, box), 0.1)
    os_.nbody(ost, O.leapfrog(stages))
    for key in "SPF":
        got, ref = getattr(st, key).cpu().numpy(), getattr(ost, key)
        assert np.abs(got - ref).max() < 1e-6 * np.abs(ref).max(), key
