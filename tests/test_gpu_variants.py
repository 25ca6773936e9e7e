"""The compiled kernel variants a mesh can select (fpm.h FPM_VAR_*, chosen by
fastpm.tune) against the oracle and against each other (GPU only): the z
transform pass with 1024 / 2048 / 4096 elements per workgroup and both column
tilings to rounding against the oracle (on meshes whose Nyquist modes take
super-tiles in plane layout, in row layout with a partial last super-tile,
or whole-row tiles); the binning walks, paint and readout
thread counts and the two-tile readout bit for bit against the default
variant (which the other tests check against the oracle) and the paint
against the oracle; the tuner keeps a valid choice."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("nmesh", [(32, 32, 32), (64, 48, 96), (16, 128, 512), (96, 24, 8),
                                   (48, 24, 64), (96, 12, 32)])
def test_transform_variants_vs_oracle(gpu, nmesh, dtype):
    from fastpm.pm import ParticleMesh, r2c_c2r_many, c2r_many
    from fastpm.force.kernels import laplace, gradient, compensation

    print("transform variants", nmesh, dtype, flush=True)
    box = (100.0, 80.0, 60.0)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    rng = np.random.default_rng(4)
    x = rng.standard_normal(nmesh)
    rf = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu))
    xk = O.r2c(x)
    ops = [laplace * 0.45 * compensation(2), laplace * gradient(0, order=1)]
    refs = [O.c2r(xk * O.k_laplace(nmesh, box) * 0.45 * O.k_deconv(nmesh, box, 2), nmesh),
            O.c2r(xk * O.k_laplace(nmesh, box) * O.k_gradient(0, nmesh, box, 1), nmesh)]
    tol = 1e-11 if dtype == "f8" else 3e-6
    for cols in (1, 2):
        for zl in (0, 1, 2):
            pm.set_variant(fft_cols=cols, z_lines=zl)
            assert _rel(rf.r2c().value.cpu().numpy(), xk) < 4 * tol
            kout, outs = r2c_c2r_many(rf, ops)
            assert _rel(kout.value.cpu().numpy(), xk) < 8 * tol
            for o, r in zip(outs, refs):
                assert _rel(o.value.cpu().numpy(), r) < 8 * tol, (cols, zl)
            for o, r in zip(c2r_many(kout, ops), refs):
                assert _rel(o.value.cpu().numpy(), r) < 8 * tol, (cols, zl)


def _lattice_parts(pm_force, B, rng, spread, gpu):
    from fastpm.pm import ParticleMesh, Particles

    nlat = [int(n) // B for n in pm_force.Nmesh]
    pm_p = ParticleMesh(Nmesh=nlat, BoxSize=pm_force.BoxSize, dtype="f4")
    Q = pm_p.generate_uniform_particle_grid()
    S = torch.from_numpy(rng.normal(0, spread, size=(len(Q), 3)).astype("f4")).to(gpu)
    return Particles(Q, S)


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
@pytest.mark.parametrize("nmesh", [(64, 64, 64), (32, 48, 64)])
def test_particle_variants_bitwise(gpu, nmesh, resampler):
    """Every binning walk, paint and readout thread count gives the default
    variant's density, RHO and force bit for bit; the density also matches
    the oracle's paint."""
    from fastpm.pm import ParticleMesh

    print("particle variants", nmesh, resampler, flush=True)
    box = (96.0, 72.0, 128.0)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    order = 2 if resampler == "cic" else 3
    rng = np.random.default_rng(9)
    parts = _lattice_parts(pm, 2, rng, 3.0, gpu)
    phi = pm.create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    n = len(parts)

    def run():
        pm.bin_particles(parts)
        assert pm.tile_misplaced(n, order) == 0
        rho = pm.paint_tiled(parts, mass=0.7)
        F = torch.empty((n, 3), dtype=torch.float32, device=gpu)
        R = torch.empty((n,), dtype=torch.float32, device=gpu)
        pm.readout_tiled_grad4(phi, parts, F, rho=rho, rho_out=R)
        return rho.value.clone(), F, R

    ref = run()
    # the kernels place a lattice particle as an exact lattice index plus its
    # float32 displacement: the oracle gets the same position in float64
    X = O.lattice(parts.Q.nlat, box) + parts.S.cpu().numpy().astype("f8")
    dens = O.paint(X, nmesh, box, order, 0.7)
    assert np.abs(ref[0].cpu().numpy() - dens).max() < 2e-6 * dens.max()
    for field, opts in (("bin", (1, 2)), ("paint", (1, 2)), ("readout", (1, 2, 3)), ("rho", (1,)),
                        ("span", (1,))):
        for c in opts:
            pm.set_variant(**{field: c})
            got = run()
            for a, b in zip(got, ref):
                assert torch.equal(a, b), (field, c)
        pm.set_variant(**{field: 0})
    # the density read through the cache with every thread count
    for c in (1, 2, 3):
        pm.set_variant(rho=1, readout=c)
        got = run()
        for a, b in zip(got, ref):
            assert torch.equal(a, b), ("rho+readout", c)
    pm.set_variant(rho=0, readout=0)


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
@pytest.mark.parametrize("nmesh", [(32, 32, 32), (64, 48, 96), (128, 96, 64), (48, 64, 48)])
def test_two_tile_readout_bitwise(gpu, nmesh, resampler):
    """The stencil readout with two z-adjacent tiles per workgroup (FPM_VAR_SPAN
    1) gives the one-tile readout's force and RHO bit for bit, for every thread
    count and density source, with and without RHO, on a lattice set and on a
    clustered set whose densest tile pairs hold more than 2048 particles (slots
    past the first two per thread are loaded in the loop); an odd tile count
    along z (the last mesh) falls back to one tile per workgroup."""
    from fastpm.pm import ParticleMesh

    print("two-tile readout", nmesh, resampler, flush=True)
    box = (90.0, 75.0, 110.0)
    order = 2 if resampler == "cic" else 3
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    rng = np.random.default_rng(17)
    phi = pm.create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    rho = pm.create("real")
    rho.value.copy_(torch.from_numpy(rng.random(nmesh).astype("f4")))
    lat = _lattice_parts(pm, 2, rng, 2.0, gpu)
    # a clump of 6000 particles within a few cells of one point, plus a
    # uniform background
    clump = rng.normal(0.0, 1.5 * box[0] / nmesh[0], size=(6000, 3)) + np.array(box) * 0.37
    back = rng.random((20000, 3)) * np.array(box)
    clustered = (np.concatenate([clump, back]) % np.array(box)).astype("f4")
    for parts in (lat, clustered):
        n = len(parts)
        pm.bin_particles(parts)
        assert pm.tile_misplaced(n, order) == 0

        def run(with_rho):
            F = torch.full((n, 3), float("nan"), dtype=torch.float32, device=gpu)
            R = torch.full((n,), float("nan"), dtype=torch.float32, device=gpu)
            if with_rho:
                pm.readout_tiled_grad4(phi, parts, F, rho=rho, rho_out=R, force_scale=0.8,
                                       rho_scale=1.3)
            else:
                pm.readout_tiled_grad4(phi, parts, F, force_scale=0.8)
            return F, R

        for with_rho in (True, False):
            pm.set_variant(span=0, readout=0, rho=0)
            ref = run(with_rho)
            assert torch.isfinite(ref[0]).all()
            for nt in (0, 2, 3):
                for r in (0, 1):
                    pm.set_variant(span=1, readout=nt, rho=r)
                    got = run(with_rho)
                    assert torch.equal(got[0], ref[0]), (nmesh, with_rho, nt, r)
                    if with_rho:
                        assert torch.equal(got[1], ref[1]), (nmesh, nt, r)
    pm.set_variant(span=0, readout=0, rho=0)


def test_autotune_chooses_valid_variants(gpu):
    from fastpm.pm import ParticleMesh
    from fastpm import tune

    pm = ParticleMesh(Nmesh=(64, 64, 64), BoxSize=100.0, dtype="f4")
    parts = _lattice_parts(pm, 2, np.random.default_rng(1), 1.0, gpu)
    tune._CACHE.clear()
    v = tune.autotune(pm, parts)
    assert v["fft_cols"] in (1, 2) and v["z_lines"] in (0, 1, 2)
    assert v["bin"] in (0, 1, 2) and v["paint"] in (0, 1, 2) and v["readout"] in (1, 2, 3)
    assert v["rho"] in (0, 1) and v["span"] in (0, 1)
    # a second mesh of the same shape takes the cached choice without timing
    pm2 = ParticleMesh(Nmesh=(64, 64, 64), BoxSize=50.0, dtype="f4")
    assert tune.autotune(pm2, _lattice_parts(pm2, 2, np.random.default_rng(2), 1.0, gpu)) == v
