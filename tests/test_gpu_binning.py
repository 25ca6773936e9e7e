"""One-pass binning into the capacity layout (GPU only): a bin call into a
workspace whose previous call had the same particle count and mesh places
every record straight at its sub-bin's capacity start plus its rank; a
sub-bin that outgrows its capacity, or a workspace without that layout,
falls back to exact placement.  Paint, weighted paint, the tiled readout and
the stencil force readout must give bit for bit what an exact binning gives
in every case, also when a step boundary's kicks and drift run inside the
binning pass."""
import ctypes

import numpy as np
import pytest
import torch

import fastpm_oracle as O

pytestmark = pytest.mark.gpu


def _outputs(pm, parts, phi, fields, w, gpu):
    n = len(parts)
    rho = pm.paint_tiled(parts, mass=0.7)
    F = torch.empty((n, 3), dtype=torch.float32, device=gpu)
    R = torch.empty((n,), dtype=torch.float32, device=gpu)
    if pm.grad4_supported(pm.resampler):
        pm.readout_tiled_grad4(phi, parts, F, rho=rho, rho_out=R)
    G = torch.empty((n, 2), dtype=torch.float32, device=gpu)
    pm.readout_tiled(fields, parts, [G[:, 0], G[:, 1]], scales=[1.0, -0.5])
    outs = [pm.create("real") for _ in range(2)]
    pm.paint_tiled_weighted(parts, [w[:, 0], w[:, 1]], outs, scale=0.3)
    return [rho.value.clone(), F, R, G] + [o.value.clone() for o in outs]


def _parts(pm, rng, spread, gpu, cluster=False):
    from fastpm.pm import ParticleMesh, Particles

    nlat = [int(n) // 2 for n in pm.Nmesh]
    pm_p = ParticleMesh(Nmesh=nlat, BoxSize=pm.BoxSize, dtype="f4")
    Q = pm_p.generate_uniform_particle_grid()
    S = rng.normal(0, spread, size=(len(Q), 3))
    if cluster:
        # most particles pulled into a few cells near the box centre: their
        # sub-bins overflow the capacity the spread-out set left
        X = O.lattice(nlat, pm.BoxSize)
        c = np.asarray(pm.BoxSize) * 0.5
        S[: len(Q) * 3 // 4] = (c - X[: len(Q) * 3 // 4]) + rng.normal(0, 0.3, size=(len(Q) * 3 // 4, 3))
    return Particles(Q, torch.from_numpy(S.astype("f4")).to(gpu))


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
@pytest.mark.parametrize("nmesh", [(64, 64, 64), (32, 48, 64)])
def test_capacity_layout_bitwise(gpu, nmesh, resampler):
    from fastpm.pm import ParticleMesh

    print("capacity layout", nmesh, resampler, flush=True)
    box = (96.0, 72.0, 128.0)
    order = 2 if resampler == "cic" else 3
    rng = np.random.default_rng(5)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    fresh = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    phi = pm.create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    fields = [pm.create("real", value=torch.from_numpy(rng.standard_normal(nmesh).astype("f4")).to(gpu))
              for _ in range(2)]
    A = _parts(pm, rng, 2.0, gpu)
    B = _parts(pm, rng, 2.5, gpu)  # the same count, moved a little
    n = len(A)
    w = torch.from_numpy(rng.standard_normal((n, 2)).astype("f4")).to(gpu)

    pm.bin_particles(A)
    assert pm.tile_layout_exact(n, order)  # bin: exact placement
    ref_a = _outputs(pm, A, phi, fields, w, gpu)
    # the same particles again: one pass into the capacity layout
    pm.bin_particles(A, rebin=True)
    assert not pm.tile_layout_exact(n, order)
    assert pm.tile_misplaced(n, order) == 0
    for x, y in zip(_outputs(pm, A, phi, fields, w, gpu), ref_a):
        assert torch.equal(x, y)
    # moved particles: capacity layout, the results of an exact binning
    pm.bin_particles(B, rebin=True)
    assert not pm.tile_layout_exact(n, order)
    got = _outputs(pm, B, phi, fields, w, gpu)
    fresh.bin_particles(B)
    assert fresh.tile_layout_exact(n, order)
    for x, y in zip(got, _outputs(fresh, B, phi, fields, w, gpu)):
        assert torch.equal(x, y)
    # the density against the oracle's paint
    X = O.lattice(B.Q.nlat, box) + B.S.cpu().numpy().astype("f8")
    dens = O.paint(X, nmesh, box, order, 0.7)
    assert np.abs(got[0].cpu().numpy() - dens).max() < 2e-6 * dens.max()


def _placement(pm, n, order):
    """The exact-placement word of the last bin call: 0 one pass into the
    capacity layout, 1 exact with the count pass's ranks, 2 exact with the
    particles ranked again by the scatter."""
    from fastpm import _lib

    off = int(_lib.lib().fpm_tile_misplaced_offset(pm._desc, int(order), int(n)))
    return int(pm._tile_ws[off + 4:off + 8].view(torch.int32).item())


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
@pytest.mark.parametrize("settle", [1, 5])
def test_capacity_overflow_falls_back(gpu, resampler, settle):
    """A set that crowds into a few sub-bins overflows the capacity the
    previous, spread-out set left: the call places its records exactly and
    gives an exact binning's results; binning it again uses its own
    capacity layout.  One rebin after an exact placement still has the wide
    margin, whose count pass keeps the ranks; after five the margin is
    narrow and the overflow is ranked again by the scatter."""
    from fastpm.pm import ParticleMesh

    print("capacity overflow", resampler, settle, flush=True)
    nmesh, box = (64, 64, 64), (100.0, 100.0, 100.0)
    order = 2 if resampler == "cic" else 3
    rng = np.random.default_rng(8)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    fresh = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
    phi = pm.create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    fields = [pm.create("real", value=torch.from_numpy(rng.standard_normal(nmesh).astype("f4")).to(gpu))
              for _ in range(2)]
    A = _parts(pm, rng, 1.0, gpu)
    C = _parts(pm, rng, 1.0, gpu, cluster=True)
    n = len(A)
    w = torch.from_numpy(rng.standard_normal((n, 2)).astype("f4")).to(gpu)
    pm.bin_particles(A)
    assert _placement(pm, n, order) == 1
    for _ in range(settle):
        pm.bin_particles(A, rebin=True)
        assert not pm.tile_layout_exact(n, order)
    pm.bin_particles(C, rebin=True)
    assert pm.tile_layout_exact(n, order)  # overflowed: exact placement
    assert _placement(pm, n, order) == (1 if settle == 1 else 2)
    assert pm.tile_misplaced(n, order) == 0
    got = _outputs(pm, C, phi, fields, w, gpu)
    fresh.bin_particles(C)
    ref = _outputs(fresh, C, phi, fields, w, gpu)
    for x, y in zip(got, ref):
        assert torch.equal(x, y)
    pm.bin_particles(C, rebin=True)
    assert not pm.tile_layout_exact(n, order)
    for x, y in zip(_outputs(pm, C, phi, fields, w, gpu), ref):
        assert torch.equal(x, y)


def test_capacity_other_count(gpu):
    """A workspace last binned for another particle count holds no layout
    for this one: the call places its records exactly."""
    from fastpm.pm import ParticleMesh, Particles

    print("capacity other count", flush=True)
    nmesh, box = (32, 32, 64), (50.0, 50.0, 100.0)
    rng = np.random.default_rng(2)
    pm = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4")
    big = _parts(pm, rng, 1.5, gpu)
    pm.bin_particles(big)
    pm.bin_particles(big, rebin=True)
    assert not pm.tile_layout_exact(len(big), 2)
    # fewer particles (explicit positions) reuse the larger workspace; the
    # C entry point is called directly, as a caller that ignores the
    # count would: the layout's header sends it to exact placement
    from fastpm import _lib
    from fastpm.pm import _stream
    m = len(big) // 3
    q = torch.from_numpy((rng.uniform(0, 1, size=(m, 3)) * np.array(box)).astype("f4")).to(gpu)
    s = torch.zeros_like(q)
    small = Particles(q, s)
    ws = pm.tile_workspace(m, 2)
    _lib.call_plain("fpm_rebin_particles_f32", small.desc(pm.dtype), pm._desc, 2, ws.data_ptr(),
                    ws.numel(), _stream())
    pm._tile_binned = (m, 2, ws.data_ptr())
    assert pm.tile_layout_exact(m, 2)
    rho = pm.paint_tiled(small, mass=1.0).value.cpu().numpy()
    ref = O.paint(q.cpu().numpy().astype("f8"), nmesh, box, 2, 1.0)
    # explicit float32 positions: each window fraction carries ~1e-6 of a
    # cell of rounding (the tolerance of the other explicit-position tests)
    assert np.abs(rho - ref).max() < 5e-6 * ref.max()
    # and bit for bit what a fresh exact binning gives
    fresh = ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4")
    fresh.bin_particles(small)
    assert np.array_equal(fresh.paint_tiled(small, mass=1.0).value.cpu().numpy(), rho)


@pytest.mark.parametrize("resampler", ["cic", "tsc"])
@pytest.mark.parametrize("explicit,walk", [(False, 0), (False, 1), (False, 2), (True, 0)])
def test_kick_drift_rebin_bitwise(gpu, resampler, explicit, walk):
    """fpm_kick_drift_rebin (a step boundary's kicks and drift inside the
    next binning's pass) leaves P, S and every tiled output equal bit for bit
    to fpm_kick_drift followed by a rebin: for 0, 1 and 2 kicks, on a layout
    that fits and on one that overflows (particles pulled into a cluster), on
    explicit Lagrangian positions and on the lattice in each binning walk
    (blocks, patches, index order)."""
    from fastpm import _lib
    from fastpm.pm import ParticleMesh, Particles

    print("kick-drift rebin", resampler, explicit, walk, flush=True)
    nmesh, box = (64, 64, 64), (100.0, 100.0, 100.0)
    order = 2 if resampler == "cic" else 3
    rng = np.random.default_rng(21)
    phi = None
    meshes = [ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4", resampler=resampler)
              for _ in range(2)]
    for m in meshes:
        m.set_variant(bin=walk)
    phi = meshes[0].create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    fields = [meshes[0].create("real", value=torch.from_numpy(
        rng.standard_normal(nmesh).astype("f4")).to(gpu)) for _ in range(2)]
    base = _parts(meshes[0], rng, 1.0, gpu)
    n = len(base)
    Q = base.Q
    if explicit:
        Q = meshes[0].positions(Particles(base.Q, None)).clone()
    S0 = base.S.clone()
    P0 = torch.from_numpy(rng.normal(0, 0.5, size=(n, 3)).astype("f4")).to(gpu)
    F0 = torch.from_numpy(rng.normal(0, 0.5, size=(n, 3)).astype("f4")).to(gpu)
    # a pull towards the box centre that crowds the particles on the last case
    X = O.lattice([int(m) // 2 for m in nmesh], box)
    pull = torch.from_numpy((np.asarray(box) * 0.5 - X).astype("f4")).to(gpu)
    w = torch.from_numpy(rng.standard_normal((n, 2)).astype("f4")).to(gpu)
    cases = [([], 0.7), ([0.3], 0.5), ([0.2, -0.4], 1.1), ([1.0], 1.0)]
    state = [[S0.clone(), P0.clone()] for _ in meshes]
    for m, (S, P) in zip(meshes, state):
        m.bin_particles(Particles(Q, S))  # the layout the rebins start from
    for c, (kicks, drift) in enumerate(cases):
        # last case: the kick sets P to the pull, the drift moves the
        # particles into a cluster about the box centre
        F = pull - state[0][1] if c == len(cases) - 1 else F0
        (S1, P1), (S2, P2) = state
        kc = (ctypes.c_double * 2)(*(kicks + [0.0] * (2 - len(kicks))))
        _lib.call("fpm_kick_drift", "f4", S1.data_ptr(), P1.data_ptr(), F.data_ptr(),
                  S1.numel(), kc, len(kicks), float(drift), None)
        meshes[0].bin_particles(Particles(Q, S1), rebin=True)
        meshes[1].bin_particles(Particles(Q, S2), rebin=True, kick_drift=(P2, F, kicks, drift))
        assert torch.equal(S1, S2) and torch.equal(P1, P2), c
        ex = [m.tile_layout_exact(n, order) for m in meshes]
        assert ex[0] == ex[1], c
        if c == len(cases) - 1:
            assert ex[1]  # the pull overflowed the layout: exact placement
        for m in meshes:
            assert m.tile_misplaced(n, order) == 0
        a = _outputs(meshes[0], Particles(Q, S1), phi, fields, w, gpu)
        b = _outputs(meshes[1], Particles(Q, S2), phi, fields, w, gpu)
        for x, y in zip(a, b):
            assert torch.equal(x, y), c


def test_kick_drift_bin_without_layout(gpu):
    """bin_particles(kick_drift=...) on a workspace that cannot rebin (no
    previous binning of these particles) applies the step boundary as a
    pass of its own before an exact binning: the same P, S and outputs as
    fpm_kick_drift followed by a bin."""
    from fastpm import _lib
    from fastpm.pm import ParticleMesh, Particles

    print("kick-drift without layout", flush=True)
    nmesh, box = (64, 64, 64), (100.0, 100.0, 100.0)
    rng = np.random.default_rng(4)
    meshes = [ParticleMesh(Nmesh=nmesh, BoxSize=box, dtype="f4") for _ in range(2)]
    phi = meshes[0].create("real")
    phi.value.copy_(torch.from_numpy(rng.standard_normal(nmesh).astype("f4")))
    fields = [meshes[0].create("real", value=torch.from_numpy(
        rng.standard_normal(nmesh).astype("f4")).to(gpu)) for _ in range(2)]
    base = _parts(meshes[0], rng, 1.0, gpu)
    n = len(base)
    P0 = torch.from_numpy(rng.normal(0, 0.5, size=(n, 3)).astype("f4")).to(gpu)
    F = torch.from_numpy(rng.normal(0, 0.5, size=(n, 3)).astype("f4")).to(gpu)
    w = torch.from_numpy(rng.standard_normal((n, 2)).astype("f4")).to(gpu)
    S1, S2, P1, P2 = base.S.clone(), base.S.clone(), P0.clone(), P0.clone()
    kc = (ctypes.c_double * 2)(0.25, -0.5)
    _lib.call("fpm_kick_drift", "f4", S1.data_ptr(), P1.data_ptr(), F.data_ptr(), S1.numel(),
              kc, 2, 0.8, None)
    meshes[0].bin_particles(Particles(base.Q, S1))
    meshes[1].bin_particles(Particles(base.Q, S2), rebin=True, kick_drift=(P2, F, [0.25, -0.5], 0.8))
    assert torch.equal(S1, S2) and torch.equal(P1, P2)
    assert meshes[1].tile_layout_exact(n, 2)
    a = _outputs(meshes[0], Particles(base.Q, S1), phi, fields, w, gpu)
    b = _outputs(meshes[1], Particles(base.Q, S2), phi, fields, w, gpu)
    for x, y in zip(a, b):
        assert torch.equal(x, y)
