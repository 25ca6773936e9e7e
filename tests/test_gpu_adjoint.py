"""Adjoint (reverse-mode) kernels and the differentiable forward model (GPU).

Pins: adjoint dot-product identities of each primitive, central finite
differences of the position derivative, finite differences of the whole
model's loss on random white-noise modes (float64, also against the FP64
oracle's forward model), and float32 vs float64 gradients."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _pm(nmesh, box, dtype, resampler="cic"):
    from fastpm.pm import ParticleMesh

    return ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype, resampler=resampler)


def _parts(pm, n, rng, gpu):
    from fastpm.pm import Particles

    q = rng.uniform(0, 1, size=(n, 3)) * pm.BoxSize
    s = rng.normal(0, 2.0, size=(n, 3))
    t = lambda a: torch.from_numpy(a.astype(pm.dtype)).to(gpu)  # noqa: E731
    return Particles(t(q), t(s)), q + s


@pytest.mark.parametrize("order", [2, 3])
def test_paint_weighted_is_readout_adjoint(gpu, order):
    from fastpm.adjoint import paint_weighted

    rng = np.random.default_rng(1)
    pm = _pm((16, 16, 16), 40.0, "f8", "cic" if order == 2 else "tsc")
    parts, X = _parts(pm, 500, rng, gpu)
    wts = torch.from_numpy(rng.standard_normal(500)).to(gpu)
    m = rng.standard_normal((16, 16, 16))
    mesh = paint_weighted(pm, parts, wts, scale=0.5)
    lhs = float((mesh.value.cpu().numpy() * m).sum())
    rhs = 0.5 * float((O.readout(m, X, (40.0,) * 3, order) * wts.cpu().numpy()).sum())
    assert lhs == pytest.approx(rhs, rel=1e-12)
    ref = O.paint(X, (16, 16, 16), (40.0,) * 3, order, 1.0)  # weights 1 sanity
    ones = torch.ones(500, dtype=torch.float64, device=gpu)
    np.testing.assert_allclose(paint_weighted(pm, parts, ones).value.cpu().numpy(), ref, atol=1e-12)


@pytest.mark.parametrize("order", [2, 3])
def test_window_grad_finite_difference(gpu, order):
    from fastpm.adjoint import window_grad
    from fastpm.pm import Particles

    rng = np.random.default_rng(2)
    pm = _pm((16, 16, 16), 40.0, "f8", "cic" if order == 2 else "tsc")
    n = 300
    # keep particles away from cell faces so the window is smooth around them
    cells = rng.integers(0, 16, size=(n, 3))
    X = (cells + rng.uniform(0.2, 0.8, size=(n, 3))) * 2.5
    if order == 3:
        X = (cells + 0.5 + rng.uniform(-0.3, 0.3, size=(n, 3))) * 2.5
    m = rng.standard_normal((16, 16, 16))
    field = pm.create("real", value=torch.from_numpy(m).to(gpu))
    w = torch.from_numpy(rng.standard_normal(n)).to(gpu)
    gx = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    parts = Particles(torch.from_numpy(X).to(gpu), None)
    window_grad(pm, parts, [field], [w], gx, scale=1.0, accumulate=False)
    h = 1e-5
    for d in range(3):
        e = np.zeros(3)
        e[d] = h
        fd = (O.readout(m, X + e, (40.0,) * 3, order) - O.readout(m, X - e, (40.0,) * 3, order)) / (2 * h)
        np.testing.assert_allclose(gx[:, d].cpu().numpy(), fd * w.cpu().numpy(), rtol=1e-6, atol=1e-9)


def test_fft_and_multiplier_adjoint_identities(gpu):
    """<c2r(M X), g> = <X, conj(M) c2r^T(g)> with c2r^T(g) = N w_c r2c(g)
    and the real inner product of stored modes."""
    from fastpm.adjoint import _kaccum, _herm
    from fastpm.pm import ComplexField
    from fastpm.force.kernels import laplace, gradient, compensation

    rng = np.random.default_rng(3)
    nm = (16, 12, 8)
    pm = _pm(nm, (50.0, 40.0, 30.0), "f8")
    X = pm.create("real", value=torch.from_numpy(rng.standard_normal(nm)).to(gpu)).r2c()
    g = rng.standard_normal(nm)
    gf = pm.create("real", value=torch.from_numpy(g).to(gpu))
    N = float(np.prod(nm))
    for op in [laplace * gradient(0, 1), gradient(2, 0) * compensation(2), laplace * 3.0]:
        y = X.apply(op).c2r().value.cpu().numpy()
        lhs = float((y * g).sum())
        G = gf.r2c()
        _herm(pm, G, 2 * N, N)
        acc = ComplexField(pm)
        _kaccum(pm, acc, G, op, conj=True, accumulate=False)
        a, b = X.value.cpu().numpy(), acc.value.cpu().numpy()
        rhs = float((a.real * b.real + a.imag * b.imag).sum())
        assert lhs == pytest.approx(rhs, rel=1e-10)


def _model(dtype, nc=16, B=2, box=100.0, nsteps=3, seed_t=99, deconv=True, resampler="cic"):
    from fastpm.core import Solver
    from fastpm.pm import ParticleMesh
    from fastpm.cosmology import Cosmology, LinearPower
    from fastpm.adjoint import ForwardModel, DensityLoss

    cosmo = Cosmology()
    Pk = LinearPower(cosmo)
    pm = ParticleMesh(Nmesh=[nc] * 3, BoxSize=box, dtype=dtype, resampler=resampler)
    solver = Solver(pm, cosmo, B=B, compensate=deconv)
    stages = np.linspace(0.1, 1.0, nsteps + 1)
    model = ForwardModel(solver, Pk, stages)
    tgt_state = model.forward(solver.whitenoise(seed_t))
    loss = DensityLoss(solver, Pk, None)
    loss.target = loss.density_k(tgt_state)
    return solver, model, loss, Pk


@pytest.mark.parametrize("resampler,eps,tol", [("tsc", 1e-3, 2e-2), ("cic", 1e-6, 2e-3)])
def test_model_gradient_finite_differences(gpu, resampler, eps, tol):
    """Central differences of the float64 loss on 16 random mode components
    with 0 < k_z < n/2 against the adjoint gradient.  TSC windows are C1, so
    eps = 1e-3 converges; CIC windows are only C0 (the loss has kinks where a
    particle crosses a cell face), so CIC uses a small eps.  Most components
    agree to ~1e-9; the few large-gradient modes carry the O(eps^2) truncation
    error of the difference quotient (~1e-3 relative at eps = 1e-3)."""
    solver, model, loss, _ = _model("f8", resampler=resampler)
    w = solver.whitenoise(7)
    st = model.forward(w)
    L0, gS = loss(st)
    G = model.backward(gS).value.cpu().numpy()
    rng = np.random.default_rng(5)
    nc = 16
    fds, gs = [], []
    for n in range(16):
        i, j, k = rng.integers(0, nc), rng.integers(0, nc), rng.integers(1, nc // 2)
        part = 1.0 if n % 2 == 0 else 1j
        wp = w.copy()
        wp.value[i, j, k] += eps * part
        wm = w.copy()
        wm.value[i, j, k] -= eps * part
        Lp, _ = loss(model.forward(wp), grad=False)
        Lm, _ = loss(model.forward(wm), grad=False)
        fds.append((Lp - Lm) / (2 * eps))
        gs.append(G[i, j, k].real if part == 1.0 else G[i, j, k].imag)
    fds, gs = np.array(fds), np.array(gs)
    print("fd", fds, "adjoint", gs)
    assert np.linalg.norm(fds - gs) / np.linalg.norm(gs) < tol
    assert np.median(np.abs(fds - gs) / np.abs(gs)) < 1e-6


def test_model_gradient_vs_oracle_forward(gpu):
    """The adjoint gradient against central differences of the FP64 oracle's
    forward model (an independent implementation of the same loss)."""
    solver, model, loss, Pk = _model("f8", nc=8, box=64.0, nsteps=2)
    w = solver.whitenoise(11)
    st = model.forward(w)
    _, gS = loss(st)
    G = model.backward(gS).value.cpu().numpy()
    osol = O.Solver(8, 64.0, solver.cosmology.Om0, B=2, deconv=2)
    q = O.lattice((8,) * 3, (64.0,) * 3)
    tgt = loss.target.value.cpu().numpy()
    fnm = (16, 16, 16)

    def oloss(wn):
        stt = osol.lpt(osol.linear(wn, Pk), q, 0.1)
        osol.nbody(stt, O.leapfrog(model.stages))
        d = O.r2c(O.paint(stt.X, fnm, (64.0,) * 3, 2) / (8 ** 3 / 16 ** 3))
        kx, ky, kz = O.kvectors(fnm, (64.0,) * 3)
        kk = np.sqrt(kx ** 2 + ky ** 2 + kz ** 2)
        with np.errstate(divide="ignore"):
            ip = np.where(kk > 0, 1.0 / Pk(np.where(kk > 0, kk, 1.0)), 0.0)
        wgt = np.where((np.arange(9) == 0) | (np.arange(9) == 8), 1.0, 2.0)[None, None, :]
        return float((wgt * np.abs(d - tgt) ** 2 * ip).sum())

    w0 = w.value.cpu().numpy()
    rng = np.random.default_rng(6)
    eps = 1e-4
    for _ in range(4):
        i, j, k = rng.integers(0, 8), rng.integers(0, 8), rng.integers(1, 4)
        wp, wm = w0.copy(), w0.copy()
        wp[i, j, k] += eps
        wm[i, j, k] -= eps
        fd = (oloss(wp) - oloss(wm)) / (2 * eps)
        assert G[i, j, k].real == pytest.approx(fd, rel=1e-4, abs=1e-8 * abs(fd) + 1e-12)


def test_fp32_gradient_matches_fp64(gpu):
    res = {}
    for dt in ("f8", "f4"):
        solver, model, loss, _ = _model(dt, nc=32, box=200.0, nsteps=5)
        st = model.forward(solver.whitenoise(7))
        _, gS = loss(st)
        res[dt] = model.backward(gS).value.cpu().numpy().astype("c16")
    a, b = res["f4"].ravel(), res["f8"].ravel()
    corr = np.vdot(a, b).real / np.sqrt(np.vdot(a, a).real * np.vdot(b, b).real)
    assert corr > 0.9999
    assert np.abs(a - b).max() < 1e-2 * np.abs(b).max()


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("nkick", [0, 1, 2])
@pytest.mark.parametrize("fresh", [0, 1])
def test_drift_kick_vjp_bitwise(gpu, dtype, nkick, fresh):
    """fpm_drift_kick_vjp equals the separate fpm_axpy calls bit for bit
    (gF cleared first when fresh, untouched without kicks)."""
    import ctypes
    from fastpm import _lib
    from fastpm.pm import _stream

    print("drift_kick_vjp", dtype, nkick, fresh, flush=True)
    tdt = torch.float32 if dtype == "f4" else torch.float64
    g = torch.Generator(device="cpu").manual_seed(5)
    n = 3 * 10007
    gS, gP, gF = (torch.randn(n, generator=g, dtype=tdt).to(gpu) for _ in range(3))
    kicks, drift = [0.37, -1.3e-2][:nkick], 0.61
    a = [gS.clone(), gP.clone(), gF.clone()]
    _lib.call("fpm_axpy", dtype, a[1].data_ptr(), a[0].data_ptr(), n, drift, _stream())
    if fresh and nkick:
        a[2].zero_()
    for k in kicks:
        _lib.call("fpm_axpy", dtype, a[2].data_ptr(), a[1].data_ptr(), n, k, _stream())
    b = [gS.clone(), gP.clone(), gF.clone()]
    kc = (ctypes.c_double * 2)(*(kicks + [0.0] * (2 - nkick)))
    _lib.call("fpm_drift_kick_vjp", dtype, b[0].data_ptr(), b[1].data_ptr(), b[2].data_ptr(), n,
              kc, nkick, drift, fresh, _stream())
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_fused_reverse_pass_bitwise(gpu, dtype):
    """The reverse pass with fused drift/kick VJPs equals the pass with one
    fpm_axpy per action: bit for bit in float32 (fixed-point tiled paints);
    in float64 the force VJP's atomic paints sum in arrival order, so two
    runs agree to rounding only."""
    print("fused reverse pass", dtype, flush=True)
    solver, model, loss, _ = _model(dtype, nc=16, box=100.0, nsteps=3)
    st = model.forward(solver.whitenoise(7))
    _, gS = loss(st)
    out = {}
    for fuse in (True, False):
        model.fuse_vjp = fuse
        out[fuse] = model.backward(gS).value.clone()
    model.fuse_vjp = True
    if dtype == "f4":
        assert torch.equal(out[True], out[False])
    else:
        d = (out[True] - out[False]).abs().max() / out[False].abs().max()
        assert float(d) < 1e-12


def test_autograd_function(gpu):
    from fastpm.adjoint import SimulateFunction

    solver, model, loss, _ = _model("f8", nc=8, box=64.0, nsteps=2)
    w = solver.whitenoise(3).value.clone().requires_grad_(True)
    S = SimulateFunction.apply(w, model)
    S.sum().backward()
    # d(sum S)/dw through the same adjoint path directly
    G = model.backward(torch.ones_like(S)).value
    assert torch.allclose(w.grad, G)


def test_config4_gradient(gpu):
    """The differentiable configuration: L = 500 Mpc/h, 128^3 particles on a
    256^3 mesh, 10 steps from a = 0.1, loss against the density of seed 99
    divided by P(k); float32 gradient with respect to the white-noise modes
    against the float64 gradient and against central differences (eps = 1e-3)
    of the float64 loss on 32 random modes."""
    print("config4", flush=True)
    res = {}
    for dt in ("f8", "f4"):
        solver, model, loss, _ = _model(dt, nc=128, box=500.0, nsteps=10, deconv=False)
        w = solver.whitenoise(7)
        st = model.forward(w)
        L0, gS = loss(st)
        res[dt] = (model.backward(gS).value.cpu().numpy().astype("c16"), solver, model, loss, w, L0)
    g32, g64 = res["f4"][0].ravel(), res["f8"][0].ravel()
    corr = np.vdot(g32, g64).real / np.sqrt(np.vdot(g32, g32).real * np.vdot(g64, g64).real)
    assert corr > 0.9999
    _, solver, model, loss, w, L0 = res["f8"]
    G = res["f8"][0]
    rng = np.random.default_rng(32)
    fds, gs = [], []
    for n in range(32):
        i, j, k = rng.integers(0, 128), rng.integers(0, 128), rng.integers(1, 64)
        part = 1.0 if n % 2 == 0 else 1j
        wp, wm = w.copy(), w.copy()
        wp.value[i, j, k] += 1e-3 * part
        wm.value[i, j, k] -= 1e-3 * part
        fd = (loss(model.forward(wp), grad=False)[0] - loss(model.forward(wm), grad=False)[0]) / 2e-3
        fds.append(fd)
        gs.append(G[i, j, k].real if part == 1.0 else G[i, j, k].imag)
    fds, gs = np.array(fds), np.array(gs)
    rel = np.abs(fds - gs) / np.abs(gs)
    print("config4 corr32/64 %.8f median rel %.3e max rel %.3e" % (corr, np.median(rel), rel.max()))
    assert np.median(rel) < 1e-3
    assert np.linalg.norm(fds - gs) / np.linalg.norm(gs) < 5e-2


@pytest.mark.parametrize("dtype", ["f8", "f4"])
@pytest.mark.parametrize("nmesh", [(16, 16, 16), (8, 12, 32), (32, 24, 6)])
def test_grad4_stencil(gpu, dtype, nmesh):
    """fpm_grad4_mesh equals the oracle's c2r(i a_d(k) r2c(f)) with the
    four-point kernel; fpm_grad4_div with scale -1 is its adjoint
    (<D f, g> = -<f, D g> summed over d)."""
    from fastpm.pm import ParticleMesh
    from fastpm.adjoint import grad4_mesh, grad4_div

    box = (50.0, 70.0, 90.0)
    rng = np.random.default_rng(12)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    f = rng.standard_normal(nmesh)
    fk = O.r2c(f)
    F = pm.create("real", value=torch.from_numpy(f.astype(dtype)).to(gpu))
    outs = [pm.create("real") for _ in range(3)]
    grad4_mesh(pm, F, outs, scale=2.0)
    tol = 1e-12 if dtype == "f8" else 2e-5
    for d in range(3):
        ref = 2.0 * O.c2r(fk * O.k_gradient(d, nmesh, box, 1), nmesh)
        got = outs[d].value.cpu().numpy()
        assert np.abs(got - ref).max() < tol * np.abs(ref).max(), d
    g = [rng.standard_normal(nmesh) for _ in range(3)]
    G = [pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu)) for x in g]
    div = pm.create("real")
    grad4_div(pm, G, div, scale=-1.0)
    lhs = sum(float((outs[d].value.double() * G[d].value.double()).sum()) for d in range(3)) / 2.0
    rhs = float((F.value.double() * div.value.double()).sum())
    assert abs(lhs - rhs) < (1e-10 if dtype == "f8" else 1e-3) * abs(lhs)


@pytest.mark.parametrize("order", [2, 3])
@pytest.mark.parametrize("nmesh,kind", [((32, 32, 32), "lattice"), ((32, 48, 64), "random"),
                                        ((48, 32, 32), "cluster")])
def test_window_grad_tiled_grad4(gpu, order, nmesh, kind):
    """fpm_window_grad_tiled_grad4_f32 (the stencil applied to the staged
    potential) equals fpm_window_grad of the three D_d phi meshes from
    fpm_grad4_mesh weighted by gf, accumulating into gx."""
    from fastpm.pm import ParticleMesh, Particles
    from fastpm.adjoint import grad4_mesh, window_grad

    print("pos_grad", order, nmesh, kind, flush=True)
    box = (60.0, 80.0, 100.0)
    rng = np.random.default_rng(40 + order)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler="cic" if order == 2 else "tsc")
    if kind == "lattice":
        Q = ParticleMesh(Nmesh=[n // 2 for n in nmesh], BoxSize=box,
                         dtype="f4").generate_uniform_particle_grid(shift=0.0)
        S = torch.from_numpy(rng.normal(0, 2.0, size=(len(Q), 3)).astype("f4")).to(gpu)
    else:
        n = 20000
        q = rng.uniform(0, 1, size=(n, 3)) * np.array(box)
        if kind == "cluster":
            q[: n // 2] = np.array(box) * 0.5 + rng.normal(0, 0.05, size=(n // 2, 3))
        Q = torch.from_numpy(q.astype("f4")).to(gpu)
        S = torch.from_numpy(rng.normal(0, 1.0, size=(n, 3)).astype("f4")).to(gpu)
    parts = Particles(Q, S)
    n = len(parts)
    phi = pm.create("real", value=torch.from_numpy(rng.standard_normal(nmesh).astype("f4")).to(gpu))
    gf = torch.from_numpy(rng.standard_normal((n, 3)).astype("f4")).to(gpu)
    g0 = torch.from_numpy(rng.standard_normal((n, 3)).astype("f4")).to(gpu)
    m = [pm.create("real") for _ in range(3)]
    grad4_mesh(pm, phi, m)
    ref = g0.clone()
    window_grad(pm, parts, m, [gf[:, 0], gf[:, 1], gf[:, 2]], ref, 0.7, accumulate=True)
    got = g0.clone()
    pm.bin_particles(parts)
    pm.window_grad_tiled_grad4(phi, parts, gf, got, 0.7, accumulate=True)
    d = (got - ref).abs().max().item()
    assert d < 2e-5 * (ref - g0).abs().max().item(), d
    # overwrite mode
    got2 = torch.full_like(g0, 123.0)
    pm.window_grad_tiled_grad4(phi, parts, gf, got2, 0.7, accumulate=False)
    assert (got2 - (ref - g0)).abs().max().item() < 2e-5 * (ref - g0).abs().max().item()
