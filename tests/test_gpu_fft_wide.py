"""The strided transform passes at the axis lengths of the 512^3 and 1024^3
force meshes, under both column tilings (64- and 128-byte row segments; the
autotuner picks the 128-byte tiles at 512^3), against the FP64 oracle (GPU
only).  The z extent of 32 (17 modes, a pitch of 32 complex values) gives
each tile row a full 16-column segment.  A fixture puts every tiling back
afterwards."""
import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu

_LENGTHS = (4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512, 768, 1024)


@pytest.fixture
def keep_tiling(gpu):
    from fastpm import _lib

    L = _lib.lib()
    saved = {(eb, n): L.fpm_fft_get_tiling(eb, n) for eb in (4, 8) for n in _LENGTHS
             if L.fpm_fft_get_tiling(eb, n) >= 0}
    yield
    for (eb, n), w in saved.items():
        L.fpm_fft_set_tiling(eb, n, w)


def _rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max()


@pytest.mark.parametrize("dtype", ["f4", "f8"])
@pytest.mark.parametrize("wide", [0, 1])
@pytest.mark.parametrize("nmesh", [(512, 16, 32), (16, 512, 32), (1024, 8, 32), (8, 1024, 32)])
def test_wide_axes_vs_oracle(gpu, keep_tiling, nmesh, wide, dtype):
    from fastpm.pm import ParticleMesh, r2c_c2r_many, c2r_grad3, c2r_many
    from fastpm.force.kernels import laplace, gradient, compensation
    from fastpm import tune

    print("wide axes", nmesh, "tiling", wide, dtype, flush=True)
    box = (100.0, 80.0, 60.0)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype=dtype)
    tune.set_tiling(pm, wide)
    assert set(tune.get_tiling(pm).values()) == {wide}
    rng = np.random.default_rng(11)
    x = rng.standard_normal(nmesh)
    rf = pm.create("real", value=torch.from_numpy(x.astype(dtype)).to(gpu))
    tol = 1e-11 if dtype == "f8" else 3e-6
    xk = O.r2c(x)
    assert _rel(rf.r2c().value.cpu().numpy(), xk) < 4 * tol
    lap, dec = O.k_laplace(nmesh, box), O.k_deconv(nmesh, box, 2)
    ops = [laplace * 0.45 * compensation(2), laplace * gradient(1, order=1)]
    refs = [O.c2r(xk * lap * 0.45 * dec, nmesh),
            O.c2r(xk * lap * O.k_gradient(1, nmesh, box, 1), nmesh)]
    kout, outs = r2c_c2r_many(rf, ops)
    assert _rel(kout.value.cpu().numpy(), xk) < 8 * tol
    for o, r in zip(outs, refs):
        assert _rel(o.value.cpu().numpy(), r) < 8 * tol
    none, outs2 = r2c_c2r_many(rf, ops, kout=False)
    assert none is None
    for o, r in zip(outs2, refs):
        assert _rel(o.value.cpu().numpy(), r) < 8 * tol
    sep = c2r_many(kout, ops)
    for o, r in zip(sep, refs):
        assert _rel(o.value.cpu().numpy(), r) < 8 * tol
    g = c2r_grad3(kout, laplace * 0.3, grad_order=1)
    for d in range(3):
        r = O.c2r(xk * lap * 0.3 * O.k_gradient(d, nmesh, box, 1), nmesh)
        assert _rel(g[d].value.cpu().numpy(), r) < 8 * tol, d


@pytest.mark.parametrize("wide", [0, 1])
def test_force_transform_512_cube_vs_oracle(gpu, keep_tiling, wide):
    """The headline force mesh itself (512^3, float32): the Force's fused
    transform (spectrum stored and not stored) with both tilings against the
    oracle's float64 transforms."""
    from fastpm.pm import ParticleMesh, r2c_c2r_many
    from fastpm.force.kernels import laplace, compensation
    from fastpm import tune

    print("512^3 force transform, tiling", wide, flush=True)
    n, box = 512, 1000.0
    pm = ParticleMesh(Nmesh=[n] * 3, BoxSize=box, dtype="f4")
    tune.set_tiling(pm, wide)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((n,) * 3, dtype=np.float32)
    rf = pm.create("real", value=torch.from_numpy(x).to(gpu))
    op = laplace * 0.45 * compensation(2)
    kout, outs = r2c_c2r_many(rf, [op])
    _, outs2 = r2c_c2r_many(rf, [op], kout=False)
    nm, bx = (n,) * 3, (box,) * 3
    xk = O.r2c(x.astype("f8"))
    assert _rel(kout.value.cpu().numpy(), xk) < 2e-5
    ref = O.c2r(xk * O.k_laplace(nm, bx) * 0.45 * O.k_deconv(nm, bx, 2), nm)
    del xk
    assert _rel(outs[0].value.cpu().numpy(), ref) < 2e-5
    assert _rel(outs2[0].value.cpu().numpy(), ref) < 2e-5
