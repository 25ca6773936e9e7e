"""Per-kernel parity of the HIP path against the FP64 oracle (GPU only).

Tolerances: float64 storage must agree to rounding (rtol 1e-11 relative to the
field's scale); float32 storage to a few float32 ulps of the field's scale
(documented per test)."""
import ctypes

import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _pm(nmesh, box, dtype, resampler="cic"):
    from fastpm.pm import ParticleMesh

    return ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype, resampler=resampler)


def _tol(dtype):
    return 1e-11 if np.dtype(dtype) == np.float64 else 3e-6


def _rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


MESHES = [(8, 8, 8), (16, 16, 16), (32, 16, 8), (12, 24, 48), (64, 64, 64), (96, 48, 32),
          (128, 128, 128), (256, 8, 16), (512, 4, 8), (8, 512, 4), (4, 8, 1024)]


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nmesh", MESHES)
def test_r2c(gpu, nmesh, dtype):
    print("r2c", nmesh, dtype, flush=True)
    rng = np.random.default_rng(1)
    x = rng.standard_normal(nmesh)
    pm = _pm(nmesh, (100.0, 80.0, 60.0), dtype)
    f = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu))
    c = f.r2c()
    ref = O.r2c(x)
    got = c.value.cpu().numpy()
    assert _rel(got, ref) < _tol(dtype) * 4
    # the input is left intact
    assert np.array_equal(f.value.cpu().numpy(), x.astype(dtype))


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nmesh", MESHES)
def test_c2r_ops(gpu, nmesh, dtype):
    print("c2r", nmesh, dtype, flush=True)
    from fastpm.pm import c2r_many
    from fastpm.force.kernels import laplace, gradient, longrange, compensation, identity

    rng = np.random.default_rng(2)
    box = (100.0, 80.0, 60.0)
    x = rng.standard_normal(nmesh)
    xk = O.r2c(x)
    pm = _pm(nmesh, box, dtype)
    f = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu))
    ck = f.r2c()
    ops = [identity, laplace * gradient(0, 1) * 1.7, laplace * gradient(1, 1) * gradient(2, 1),
           gradient(2, 0) * compensation(2), laplace * longrange(3.0), compensation(3) * 0.5]
    refs = [
        O.c2r(xk, nmesh),
        O.c2r(xk * O.k_laplace(nmesh, box) * O.k_gradient(0, nmesh, box, 1) * 1.7, nmesh),
        O.c2r(xk * O.k_laplace(nmesh, box) * O.k_gradient(1, nmesh, box, 1)
              * O.k_gradient(2, nmesh, box, 1), nmesh),
        O.c2r(xk * O.k_gradient(2, nmesh, box, 0) * O.k_deconv(nmesh, box, 2), nmesh),
        O.c2r(xk * O.k_laplace(nmesh, box) * O.k_longrange(nmesh, box, 3.0), nmesh),
        O.c2r(xk * O.k_deconv(nmesh, box, 3) * 0.5, nmesh),
    ]
    outs = c2r_many(ck, ops)
    for o, r in zip(outs, refs):
        assert _rel(o.value.cpu().numpy(), r) < _tol(dtype) * 8
    # identity round trip
    assert _rel(outs[0].value.cpu().numpy(), x) < _tol(dtype) * 8


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_kspace_apply(gpu, dtype):
    from fastpm.force.kernels import laplace, gradient

    nmesh, box = (16, 16, 16), (50.0,) * 3
    rng = np.random.default_rng(3)
    x = rng.standard_normal(nmesh)
    pm = _pm(nmesh, box, dtype)
    ck = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu)).r2c()
    out = ck.apply(laplace).apply(gradient(1, order=1), out=Ellipsis)
    ref = O.r2c(x) * O.k_laplace(nmesh, box) * O.k_gradient(1, nmesh, box, 1)
    assert _rel(out.value.cpu().numpy(), ref) < _tol(dtype) * 4


def _particles(n, box, rng, spread):
    q = rng.uniform(0, box, size=(n, 3))
    s = rng.normal(0, spread, size=(n, 3))
    return q, s


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("order", [2, 3])
def test_paint_explicit(gpu, dtype, order):
    from fastpm.pm import Particles

    nmesh, box = (16, 16, 16), 64.0
    rng = np.random.default_rng(4)
    q, s = _particles(3000, box, rng, 20.0)  # displacements wrap across the box
    pm = _pm(nmesh, box, dtype, "cic" if order == 2 else "tsc")
    parts = Particles(torch.from_numpy(q.astype(dtype)).to(gpu), torch.from_numpy(s.astype(dtype)).to(gpu))
    rho = pm.paint(parts, mass=0.7)
    ref = O.paint(q + s, nmesh, (box,) * 3, order, 0.7)
    assert _rel(rho.value.cpu().numpy(), ref) < (1e-12 if dtype == "f8" else 2e-5)
    # hold=True accumulates
    pm.paint(parts, mass=0.3, hold=True, out=rho)
    ref2 = ref + O.paint(q + s, nmesh, (box,) * 3, order, 0.3)
    assert _rel(rho.value.cpu().numpy(), ref2) < (1e-12 if dtype == "f8" else 2e-5)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("B", [1, 2, 3])
def test_paint_readout_lattice(gpu, dtype, order, B):
    from fastpm.pm import Particles

    nc, box = 8, 32.0
    nmesh = (nc * B,) * 3
    rng = np.random.default_rng(5 + B)
    pm_p = _pm((nc,) * 3, box, dtype)
    pm = _pm(nmesh, box, dtype, "cic" if order == 2 else "tsc")
    Q = pm_p.generate_uniform_particle_grid(shift=0.5 if B == 3 else 0.0)
    s = rng.normal(0, 3.0, size=(nc ** 3, 3))
    parts = Particles(Q, torch.from_numpy(s.astype(dtype)).to(gpu))
    q = O.lattice((nc,) * 3, (box,) * 3, Q.shift)
    rho = pm.paint(parts)
    ref = O.paint(q + s, nmesh, (box,) * 3, order)
    assert _rel(rho.value.cpu().numpy(), ref) < (1e-12 if dtype == "f8" else 2e-5)
    # positions
    X = pm.positions(parts).cpu().numpy()
    assert _rel(X, q + s) < (1e-14 if dtype == "f8" else 1e-6)
    # fused 4-component readout with scales, into strided outputs
    meshes = [rng.standard_normal(nmesh) for _ in range(4)]
    fields = [pm.create("real", value=torch.from_numpy(m.astype(dtype)).to(gpu)) for m in meshes]
    F = torch.empty((nc ** 3, 3), dtype=pm.tdtype, device=gpu)
    R = torch.empty((nc ** 3,), dtype=pm.tdtype, device=gpu)
    pm.readout_many(fields, parts, [R, F[:, 0], F[:, 1], F[:, 2]], scales=[1.0, 2.0, -1.0, 0.5])
    sc = [1.0, 2.0, -1.0, 0.5]
    refs = [O.readout(m, q + s, (box,) * 3, order) * c for m, c in zip(meshes, sc)]
    got = [R.cpu().numpy()] + [F[:, d].cpu().numpy() for d in range(3)]
    for g, r in zip(got, refs):
        assert _rel(g, r) < (1e-12 if dtype == "f8" else 2e-5)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("unitary", [False, True])
@pytest.mark.parametrize("nmesh", [(8, 8, 8), (16, 12, 6), (32, 32, 32)])
def test_whitenoise(gpu, dtype, unitary, nmesh):
    pm = _pm(nmesh, 100.0, dtype)
    w = pm.generate_whitenoise(1234, unitary=unitary).value.cpu().numpy()
    ref = O.whitenoise(nmesh, 1234, unitary)
    assert _rel(w, ref) < (1e-13 if dtype == "f8" else 1e-6)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_linear(gpu, dtype):
    from fastpm.core import Solver
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology()
    Pk = LinearPower(cosmo)
    nmesh, box = (32,) * 3, 200.0
    pm = _pm(nmesh, box, dtype)
    solver = Solver(pm, cosmo, B=1)
    wn = solver.whitenoise(7)
    d = solver.linear(wn, Pk).value.cpu().numpy()
    ref = O.linear(O.whitenoise(nmesh, 7), Pk, nmesh, (box,) * 3)
    assert _rel(d, ref) < (1e-12 if dtype == "f8" else 1e-6)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_power_bins(gpu, dtype):
    from fastpm.utils import power_sums

    nmesh, box = (32, 32, 32), 100.0
    rng = np.random.default_rng(9)
    a = rng.standard_normal(nmesh)
    b = a + 0.3 * rng.standard_normal(nmesh)
    pm = _pm(nmesh, box, dtype)
    fa = pm.create("real", value=torch.from_numpy(a.astype(dtype)).to(gpu)).r2c()
    fb = pm.create("real", value=torch.from_numpy(b.astype(dtype)).to(gpu)).r2c()
    kmax = np.pi * 32 / box
    got = power_sums(fa, fb, nbins=20, kmin=0.0, kmax=kmax)
    ref = O.power_bins(fa.value.cpu().numpy().astype("c16"), nmesh, (box,) * 3, 20, 0.0, kmax / 20,
                       fb.value.cpu().numpy().astype("c16"))
    assert np.array_equal(got[:, 0], ref[:, 0])
    np.testing.assert_allclose(got, ref, rtol=1e-12, atol=0)
    # determinism: bitwise identical on a repeat
    got2 = power_sums(fa, fb, nbins=20, kmin=0.0, kmax=kmax)
    assert np.array_equal(got, got2)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_axpy_combine(gpu, dtype):
    from fastpm import _lib
    from fastpm.pm import _stream

    rng = np.random.default_rng(11)
    n = 1000
    x, y, z = (rng.standard_normal(n) for _ in range(3))
    tx, ty, tz = (torch.from_numpy(v.astype(dtype)).cuda() for v in (x, y, z))
    _lib.call("fpm_axpy", dtype, ty.data_ptr(), tx.data_ptr(), n, 0.37, _stream())
    assert _rel(ty.cpu().numpy(), y.astype(dtype) + 0.37 * x.astype(dtype)) < _tol(dtype)
    S, P, F = (torch.empty(n, dtype=tx.dtype, device="cuda") for _ in range(3))
    c = np.array([1.5, -2.0, 0.25, 3.0, 7.0, -1.0])
    _lib.call("fpm_lpt_combine", dtype, tx.data_ptr(), tz.data_ptr(), S.data_ptr(), P.data_ptr(),
              F.data_ptr(), n, c.ctypes.data, _stream())
    xs, zs = x.astype(dtype).astype("f8"), z.astype(dtype).astype("f8")
    for t, (c0, c1) in zip((S, P, F), [(c[0], c[1]), (c[2], c[3]), (c[4], c[5])]):
        assert _rel(t.cpu().numpy(), c0 * xs + c1 * zs) < _tol(dtype) * 4


@pytest.mark.parametrize("dtype", ["f8", "f4"])
def test_lpt_ops(gpu, dtype):
    from fastpm.force.lpt import lpt1, lpt2source

    nmesh, box = (16, 16, 16), 64.0
    rng = np.random.default_rng(12)
    x = rng.standard_normal(nmesh)
    pm = _pm(nmesh, box, dtype)
    dk = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu)).r2c()
    dref = O.r2c(x)
    src = lpt2source(dk).value.cpu().numpy()
    sref = O.lpt2source(dref, nmesh, (box,) * 3)
    assert _rel(src, sref) < (1e-11 if dtype == "f8" else 2e-5)
    Q = pm.generate_uniform_particle_grid(shift=0.25)
    q = O.lattice(nmesh, (box,) * 3, 0.25)
    disp = lpt1(dk, Q).cpu().numpy()
    ref = O.lpt1(dref, q, nmesh, (box,) * 3)
    assert _rel(disp, ref) < (1e-11 if dtype == "f8" else 1e-5)


def test_errors(gpu):
    from fastpm import _lib
    from fastpm.pm import ParticleMesh, _stream

    with pytest.raises(ValueError):
        ParticleMesh(Nmesh=[10, 10, 10], BoxSize=1.0)
    pm = ParticleMesh(Nmesh=[8, 8, 8], BoxSize=1.0, dtype="f4")
    with pytest.raises(ValueError):
        pm.paint(np.zeros((4, 3), "f4"), resampler="pcs")
    # bad order straight through the ABI
    f = pm.create("real")
    d = _lib.FpmParticles()
    d.n = 8
    d.lattice[0] = d.lattice[1] = d.lattice[2] = 2
    with pytest.raises(RuntimeError):
        _lib.call("fpm_paint", "f4", d, pm._desc, 5, 1.0, f.value.data_ptr(), 0, _stream())
    # lattice size inconsistent with n
    d.n = 9
    with pytest.raises(RuntimeError):
        _lib.call("fpm_paint", "f4", d, pm._desc, 2, 1.0, f.value.data_ptr(), 0, _stream())


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("nmesh,kind", [((16, 16, 16), "lattice"), ((32, 48, 64), "random"),
                                        ((64, 64, 64), "lattice"), ((96, 32, 16), "cluster"),
                                        ((32, 48, 48), "lattice4"), ((32, 48, 64), "lattice"),
                                        ((64, 64, 64), "lattice_spread")])
def test_tiled_paint_readout(gpu, order, nmesh, kind):
    """Binned paint/readout against the oracle; bitwise reproducible; no
    particle found outside its binned tile.  The lattices cover the three
    binning orders: 8^3 (8 x 8 patches), 32^3 and 16 x 24 x 32 (4 x 8 x 16
    blocks per workgroup), 8 x 12 x 12 (index order); lattice_spread moves
    every particle anywhere within one box either way, so a block's 512
    particles fall in up to 512 distinct sub-bins."""
    from fastpm.pm import Particles

    print("tiled", order, nmesh, kind, flush=True)
    box = (64.0, 96.0, 128.0)
    rng = np.random.default_rng(21)
    pm = _pm(nmesh, box, "f4", "cic" if order == 2 else "tsc")
    if kind.startswith("lattice"):
        # lattice4: sides 8 x 12 x 12, not multiples of 8 (index-order binning)
        nlat = tuple(n // (4 if kind == "lattice4" else 2) for n in nmesh)
        pm_p = _pm(nlat, box, "f4")
        Q = pm_p.generate_uniform_particle_grid(shift=0.0)
        if kind == "lattice_spread":
            s = rng.uniform(-1, 1, size=(len(Q), 3)) * np.array(box)
        else:
            s = rng.normal(0, 3.0, size=(len(Q), 3))
        q = O.lattice(nlat, box)
        parts = Particles(Q, torch.from_numpy(s.astype("f4")).to(gpu))
    else:
        n = 20000
        q = rng.uniform(0, 1, size=(n, 3)) * np.array(box)
        if kind == "cluster":  # most particles in one cell, some exactly on tile edges
            q[: n // 2] = np.array(box) * 0.5 + rng.normal(0, 0.01, size=(n // 2, 3))
            q[n // 2: n // 2 + 100] = np.array([16.0, 48.0, 0.0]) * np.array(box) / np.array(nmesh)
        s = rng.normal(0, 1.0, size=q.shape)
        parts = Particles(torch.from_numpy(q.astype("f4")).to(gpu),
                          torch.from_numpy(s.astype("f4")).to(gpu))
        q = q.astype("f4").astype("f8")
        s = s.astype("f4").astype("f8")
    X = pm.positions(parts).cpu().numpy().astype("f8")
    pm.bin_particles(parts)
    a = pm.paint_tiled(parts, mass=0.75).value.clone()
    b = pm.paint_tiled(parts, mass=0.75).value.clone()
    assert torch.equal(a, b)
    ref = O.paint(X, nmesh, box, order, 0.75)
    # displacements of up to a box side (lattice_spread) are rounded in
    # float32 to ~1.5e-5 cells, in the kernels and in X alike
    tol = 2e-5 if kind == "lattice_spread" else 5e-6
    assert _rel(a.cpu().numpy(), ref) < tol
    plain = pm.paint(parts, mass=0.75).value
    assert _rel(a.cpu().numpy(), plain.cpu().numpy().astype("f8")) < tol
    meshes = [rng.standard_normal(nmesh) for _ in range(4)]
    fields = [pm.create("real", value=torch.from_numpy(m.astype("f4")).to(gpu)) for m in meshes]
    F = torch.empty((len(parts), 3), dtype=torch.float32, device=gpu)
    R = torch.empty((len(parts),), dtype=torch.float32, device=gpu)
    pm.readout_tiled(fields, parts, [R, F[:, 0], F[:, 1], F[:, 2]], scales=[1.0, 2.0, -1.0, 0.5])
    sc = [1.0, 2.0, -1.0, 0.5]
    got = [R.cpu().numpy()] + [F[:, d].cpu().numpy() for d in range(3)]
    for g, m, c in zip(got, meshes, sc):
        r = O.readout(m.astype("f4").astype("f8"), X, box, order) * c
        assert np.abs(g - r).max() < 4 * tol * max(1.0, np.abs(r).max())
    assert pm.tile_misplaced(len(parts), order) == 0


def test_tiled_solver_matches_plain(gpu):
    """The fused float32 Force (tiled paint/readout) and the plain float32
    Force agree to float32 rounding on the same state."""
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology()
    Pk = LinearPower(cosmo)
    out = {}
    for fused in (True, False):
        pm = ParticleMesh(Nmesh=[32] * 3, BoxSize=200.0, dtype="f4")
        solver = Solver(pm, cosmo, B=2, fused=fused, compensate=True)
        st = solver.lpt(solver.linear(solver.whitenoise(3), Pk), pm.generate_uniform_particle_grid(), 0.1)
        step = solver.nbodystep
        assert step.tiled == fused
        for act in leapfrog([0.1, 0.4, 1.0]):
            step.run(*act, st, None)
        out[fused] = (st.X.cpu().numpy(), st.F.cpu().numpy(), st.RHO.cpu().numpy())
    for a, b in zip(out[True], out[False]):
        assert _rel(a, b) < 1e-4


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nmesh", [(8, 8, 8), (16, 32, 64), (96, 48, 32), (64, 64, 64), (8, 512, 4)])
@pytest.mark.parametrize("order", [0, 1])
def test_c2r_grad3(gpu, dtype, nmesh, order):
    from fastpm.pm import c2r_grad3
    from fastpm.force.kernels import laplace, compensation

    box = (100.0, 80.0, 60.0)
    rng = np.random.default_rng(31)
    x = rng.standard_normal(nmesh)
    xk = O.r2c(x)
    pm = _pm(nmesh, box, dtype)
    ck = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu)).r2c()
    outs = c2r_grad3(ck, laplace * 0.45 * compensation(2), grad_order=order)
    common = xk * O.k_laplace(nmesh, box) * 0.45 * O.k_deconv(nmesh, box, 2)
    for d in range(3):
        ref = O.c2r(common * O.k_gradient(d, nmesh, box, order), nmesh)
        assert _rel(outs[d].value.cpu().numpy(), ref) < _tol(dtype) * 8, d
    # input untouched
    assert _rel(ck.value.cpu().numpy(), xk) < _tol(dtype) * 4


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("nmesh", [(32, 32, 32), (32, 48, 64), (64, 32, 48)])
def test_readout_grad4_matches_kspace(gpu, order, nmesh):
    """readout(D_d phi) with the four-point stencil equals the readout of
    c2r(i a_d(k) phi_k) (order-1 gradient kernel): against the oracle in FP64
    and against the library's own k-space path; RHO rides along."""
    from fastpm.pm import Particles, c2r_grad3
    from fastpm.force.kernels import KernelOp

    print("grad4", order, nmesh, flush=True)
    box = (80.0, 96.0, 112.0)
    rng = np.random.default_rng(5)
    pm = _pm(nmesh, box, "f4", "cic" if order == 2 else "tsc")
    nlat = tuple(n // 2 for n in nmesh)
    pm_p = _pm(nlat, box, "f4")
    Q = pm_p.generate_uniform_particle_grid(shift=0.0)
    s = rng.normal(0, 2.0, size=(len(Q), 3)).astype("f4")
    parts = Particles(Q, torch.from_numpy(s).to(gpu))
    X = pm.positions(parts).cpu().numpy().astype("f8")
    phi_h = rng.standard_normal(nmesh).astype("f4")
    rho_h = rng.standard_normal(nmesh).astype("f4")
    phi = pm.create("real", value=torch.from_numpy(phi_h).to(gpu))
    rho = pm.create("real", value=torch.from_numpy(rho_h).to(gpu))
    pm.bin_particles(parts)
    F = torch.empty((len(Q), 3), dtype=torch.float32, device=gpu)
    R = torch.empty((len(Q),), dtype=torch.float32, device=gpu)
    pm.readout_tiled_grad4(phi, parts, F, rho=rho, rho_out=R, force_scale=1.5, rho_scale=0.5)
    got = F.cpu().numpy().astype("f8")
    xk = O.r2c(phi_h.astype("f8"))
    for d in range(3):
        ref = O.readout(O.c2r(xk * O.k_gradient(d, nmesh, box, 1), nmesh), X, box, order) * 1.5
        assert np.abs(got[:, d] - ref).max() < 2e-5 * np.abs(ref).max(), d
    rref = O.readout(rho_h.astype("f8"), X, box, order) * 0.5
    assert np.abs(R.cpu().numpy() - rref).max() < 2e-5 * np.abs(rref).max()
    # the library's k-space route: c2r_grad3 with an identity common factor
    outs = c2r_grad3(phi.r2c(), KernelOp(), grad_order=1)
    F2 = torch.empty_like(F)
    pm.readout_tiled(outs, parts, [F2[:, 0], F2[:, 1], F2[:, 2]], scales=[1.5] * 3)
    assert _rel(got, F2.cpu().numpy().astype("f8")) < 2e-5
    assert pm.tile_misplaced(len(Q), order) == 0


def test_readout_grad4_small_mesh_rejected(gpu):
    """Meshes below 16 + order + 3 cells per side are refused (the stencil
    box would wrap onto itself); the Force then keeps the k-space route."""
    from fastpm.pm import Particles
    from fastpm import _lib

    pm = _pm((16, 16, 16), (50.0,) * 3, "f4")
    assert pm.tiles_supported() and not pm.grad4_supported()
    Q = _pm((8, 8, 8), (50.0,) * 3, "f4").generate_uniform_particle_grid()
    parts = Particles(Q, None)
    pm.bin_particles(parts)
    phi = pm.create("real")
    F = torch.empty((len(Q), 3), dtype=torch.float32, device=gpu)
    with pytest.raises(RuntimeError, match="unsupported"):
        pm.readout_tiled_grad4(phi, parts, F)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nmesh", [(8, 8, 8), (16, 32, 64), (96, 48, 32), (64, 64, 64),
                                   (4, 16, 8), (24, 12, 16), (512, 8, 4)])
def test_r2c_c2r_fused(gpu, dtype, nmesh):
    """fpm_fft_r2c_c2r (forward x pass fused with the first inverse pass) gives
    the same spectrum as fpm_fft_r2c and the same outputs as fpm_fft_c2r, and
    matches the oracle, for one and for several outputs."""
    from fastpm.pm import r2c_c2r_many, c2r_many
    from fastpm.force.kernels import laplace, gradient, compensation

    print("r2c_c2r", dtype, nmesh, flush=True)
    box = (100.0, 80.0, 60.0)
    rng = np.random.default_rng(17)
    x = rng.standard_normal(nmesh)
    pm = _pm(nmesh, box, dtype)
    rf = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu))
    for ops in ([laplace * 0.45 * compensation(2)],
                [laplace * 0.45, laplace * gradient(1, order=1), gradient(2, order=0)]):
        kout, outs = r2c_c2r_many(rf, ops)
        ck = rf.r2c()
        ref_k = O.r2c(x)
        assert _rel(kout.value.cpu().numpy(), ck.value.cpu().numpy()) < _tol(dtype) * 4
        assert _rel(kout.value.cpu().numpy(), ref_k) < _tol(dtype) * 8
        sep = c2r_many(ck, ops)
        for o, s in zip(outs, sep):
            assert _rel(o.value.cpu().numpy(), s.value.cpu().numpy()) < _tol(dtype) * 8
        o0 = O.c2r(ref_k * O.k_laplace(nmesh, box) * 0.45 *
                   (O.k_deconv(nmesh, box, 2) if len(ops) == 1 else 1.0), nmesh)
        assert _rel(outs[0].value.cpu().numpy(), o0) < _tol(dtype) * 8
        # without storing the spectrum (kout NULL): the same arithmetic, with
        # the forward passes in work[0] and the x pass in place there
        none, outs2 = r2c_c2r_many(rf, ops, kout=False)
        assert none is None
        for o, o2 in zip(outs, outs2):
            assert torch.equal(o.value, o2.value)
        assert torch.equal(rf.value.cpu(), torch.from_numpy(x.astype(dtype)))


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("nmesh", [(64, 64, 64), (96, 48, 32), (32, 128, 16)])
def test_fft_tiling_bitwise(gpu, dtype, nmesh):
    """Both column tile widths give the same transforms to rounding (r2c, fused
    r2c/c2r with a multiplier, c2r_grad3: the per-column arithmetic is the
    same, up to the compiler's contraction of multiply-adds in each
    instantiation); the autotuner picks one of them."""
    from fastpm.pm import r2c_c2r_many, c2r_grad3
    from fastpm.force.kernels import laplace, compensation
    from fastpm import tune

    rng = np.random.default_rng(3)
    pm = _pm(nmesh, (90.0, 70.0, 50.0), dtype)
    x = pm.create("real", value=torch.from_numpy(rng.standard_normal(nmesh).astype(dtype)).to(gpu))
    res = []
    for wide in (0, 1):
        tune.set_tiling(pm, wide)
        assert set(tune.get_tiling(pm).values()) == {wide}
        k, outs = r2c_c2r_many(x, [laplace * 0.7 * compensation(2)])
        g = c2r_grad3(x.r2c(), laplace * 0.3, grad_order=1)
        res.append([k.value.clone(), outs[0].value.clone()] + [o.value.clone() for o in g])
    tol = 1e-6 if dtype == "f4" else 1e-14
    for a, b in zip(*res):
        assert (a - b).abs().max().item() <= tol * b.abs().max().item()
    tune._CACHE.clear()
    best = tune.autotune_fft(pm, reps=3)
    assert best["fft_cols"] in (1, 2) and best["z_lines"] in (0, 1, 2)
    assert set(tune.get_tiling(pm).values()) == {best["fft_cols"] - 1}
    assert pm.variant["fft_cols"] == best["fft_cols"]


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("nmesh,kind", [((32, 48, 64), "random"), ((16, 16, 16), "lattice"),
                                        ((96, 32, 16), "cluster")])
def test_paint_tiled_weighted(gpu, order, nmesh, kind):
    """Three weighted paints at once from the binned records equal the
    oracle's weighted paint of each weight array (the adjoint of readout)."""
    from fastpm.pm import Particles

    box = (64.0, 96.0, 128.0)
    rng = np.random.default_rng(4)
    pm = _pm(nmesh, box, "f4", "cic" if order == 2 else "tsc")
    if kind == "lattice":
        nlat = tuple(n // 2 for n in nmesh)
        Q = _pm(nlat, box, "f4").generate_uniform_particle_grid(shift=0.0)
        s = rng.normal(0, 2.0, size=(len(Q), 3)).astype("f4")
        parts = Particles(Q, torch.from_numpy(s).to(gpu))
    else:
        n = 20000
        q = rng.uniform(0, 1, size=(n, 3)) * np.array(box)
        if kind == "cluster":
            q[: n // 2] = np.array(box) * 0.5 + rng.normal(0, 0.01, size=(n // 2, 3))
        parts = Particles(torch.from_numpy(q.astype("f4")).to(gpu), None)
    X = pm.positions(parts).cpu().numpy().astype("f8")
    W = torch.from_numpy(rng.standard_normal((len(X), 3)).astype("f4")).to(gpu)
    pm.bin_particles(parts)
    outs = [pm.create("real") for _ in range(3)]
    pm.paint_tiled_weighted(parts, [W[:, 0], W[:, 1], W[:, 2]], outs, scale=0.5)
    Wh = W.cpu().numpy().astype("f8")
    for c in range(3):
        ref = O.paint(X, nmesh, box, order, mass=0.5 * Wh[:, c])
        # fixed-point sums, each deposit rounded once in float: bounded by the
        # paint of |w| (signed weights cancel in the clustered cell)
        scale = O.paint(X, nmesh, box, order, mass=0.5 * np.abs(Wh[:, c])).max()
        assert np.abs(outs[c].value.cpu().numpy() - ref).max() < 1e-6 * scale, c
    # exact integer sums: bitwise the same on a repeat, and component c of a
    # three-component call equals a one-component call of the same weights
    again = [pm.create("real") for _ in range(3)]
    pm.paint_tiled_weighted(parts, [W[:, 0], W[:, 1], W[:, 2]], again, scale=0.5)
    for c in range(3):
        assert torch.equal(outs[c].value, again[c].value), c
    one = [pm.create("real")]
    pm.paint_tiled_weighted(parts, [W[:, 1]], one, scale=0.5)
    assert torch.equal(one[0].value, outs[1].value)


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nkick", [0, 1, 2])
def test_kick_drift_fused_bitwise(gpu, dtype, nkick):
    """fpm_kick_drift equals nkick fpm_axpy kicks then an fpm_axpy drift bit
    for bit, and both match P += k F (per kick), S += d P in float64."""
    from fastpm import _lib
    from fastpm.pm import _stream

    rng = np.random.default_rng(21 + nkick)
    n = 3 * 12345
    S0, P0, F0 = (torch.from_numpy(rng.standard_normal(n).astype(dtype)).cuda() for _ in range(3))
    kicks, drift = [0.37, -1.25][:nkick], 0.81
    S1, P1 = S0.clone(), P0.clone()
    for k in kicks:
        _lib.call("fpm_axpy", dtype, P1.data_ptr(), F0.data_ptr(), n, k, _stream())
    _lib.call("fpm_axpy", dtype, S1.data_ptr(), P1.data_ptr(), n, drift, _stream())
    S2, P2 = S0.clone(), P0.clone()
    kc = (ctypes.c_double * 2)(*(kicks + [0.0] * (2 - nkick)))
    _lib.call("fpm_kick_drift", dtype, S2.data_ptr(), P2.data_ptr(), F0.data_ptr(), n, kc, nkick,
              drift, _stream())
    assert torch.equal(S1, S2) and torch.equal(P1, P2)
    p = P0.double().cpu().numpy()
    f = F0.double().cpu().numpy()
    for k in kicks:
        p = p + k * f
    s = S0.double().cpu().numpy() + drift * p
    assert _rel(P2.double().cpu().numpy(), p) < _tol(dtype) * 4
    assert _rel(S2.double().cpu().numpy(), s) < _tol(dtype) * 4
    # argument checks
    with pytest.raises(RuntimeError):
        _lib.call("fpm_kick_drift", dtype, S2.data_ptr(), P2.data_ptr(), F0.data_ptr(), n, kc, 3,
                  drift, _stream())


def test_nbody_fused_steps_bitwise(gpu):
    """Solver.nbody (kicks and drift of each step boundary fused, spectrum
    not stored) ends in the state of action-by-action run() calls: bit for
    bit on the float32 stencil path (deterministic binned paint), to rounding
    on the float64 path, whose atomic paint sums in arrival order."""
    from fastpm.core import Solver, leapfrog
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower

    cosmo = Cosmology(Om0=0.3, h=0.7, n_s=0.96, sigma8=0.8)
    Pk = LinearPower(cosmo)
    for dtype, nc in (("f4", 16), ("f8", 8)):
        pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=100.0, dtype=dtype)
        solver = Solver(pm, cosmo, B=2)
        dlin = solver.linear(solver.whitenoise(8), Pk)
        Q = pm.generate_uniform_particle_grid()
        a = solver.lpt(dlin, Q, 0.1)
        b = solver.lpt(dlin, Q, 0.1)
        stages = np.linspace(0.1, 1.0, 5)
        solver.nbody(a, leapfrog(stages))
        step = solver.nbodystep
        for act in leapfrog(stages):
            step.run(*act, b, None)
        assert a.a == b.a
        for key in ("S", "P", "F", "RHO"):
            x, y = getattr(a, key), getattr(b, key)
            if dtype == "f4":
                assert torch.equal(x, y), (dtype, key)
            else:
                assert ((x - y).abs().max() / y.abs().max()).item() < 1e-11, (dtype, key)


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_empty_particle_set(gpu, dtype):
    """No particles (a rank that owns none, in the reference's decomposed
    runs): paints give a zero mesh, readouts and gradients are no-ops, and
    no entry point rejects the empty (null) arrays."""
    from fastpm.pm import ParticleMesh, Particles
    from fastpm.adjoint import window_grad, paint_weighted

    pm = ParticleMesh(Nmesh=(32, 32, 32), BoxSize=50.0, dtype=dtype)
    td = torch.float32 if dtype == "f4" else torch.float64
    Q = torch.empty((0, 3), dtype=td, device=gpu)
    parts = Particles(Q, torch.empty((0, 3), dtype=td, device=gpu))
    field = pm.create("real", value=torch.ones(32, 32, 32, dtype=td, device=gpu))
    rho = pm.paint(parts, mass=1.0)
    assert float(rho.value.abs().max()) == 0.0
    out = torch.empty(0, dtype=td, device=gpu)
    pm.readout_many([field], parts, [out])
    g = torch.empty((0, 3), dtype=td, device=gpu)
    window_grad(pm, parts, [field], None, g, 1.0)
    assert float(paint_weighted(pm, parts, out).value.abs().max()) == 0.0
    assert pm.positions(parts).shape == (0, 3)
    if pm.tiles_supported():
        pm.bin_particles(parts)
        assert float(pm.paint_tiled(parts, mass=1.0).value.abs().max()) == 0.0
        pm.readout_tiled([field], parts, [out])
        F = torch.empty((0, 3), dtype=td, device=gpu)
        pm.readout_tiled_grad4(field, parts, F, rho=field, rho_out=out)
        outs = [pm.create("real") for _ in range(3)]
        pm.paint_tiled_weighted(parts, [F[:, 0], F[:, 1], F[:, 2]], outs)
        assert all(float(o.value.abs().max()) == 0.0 for o in outs)
        pm.window_grad_tiled_grad4(field, parts, F, g)
    torch.cuda.synchronize()
