"""Autotuner for the kernel variants of a mesh.

Every mesh descriptor carries a variant word (fpm.h FPM_VAR_*) that selects,
per mesh, among compiled variants of the hot kernels:

  fft_cols  the strided (y, x) transform passes: column tiles of 64- or
            128-byte row segments (measured on gfx950: 128-byte tiles win at
            512^3, 64-byte tiles at 192^3);
  z_lines   the z transform pass: 1024, 2048 or 4096 complex elements per
            workgroup (fewer lines per workgroup fill the chip on small
            meshes);
  bin       the binning walk over lattice particles: 4 x 8 x 16 blocks,
            8 x 8 patches or index order;
  paint     threads per tile of the tiled paint (256, 512, 1024);
  readout   threads per tile of the stencil force readout (256, 512, 1024);
  rho       the stencil readout's density: staged in LDS, or read through
            the cache (a third less LDS per tile);
  span      tiles per workgroup of the stencil readout: one, or two adjacent
            along z sharing their boxes (fewer cache-line requests per
            particle, twice the LDS).

Every choice gives the same results (bitwise for the particle kernels, to
rounding for the transform tilings).  `autotune` times them on the mesh by
coordinate descent (the transform fields together on the Force's transform,
then each particle-kernel field on the binning, paint and readout of a
particle set) and writes the choice into the mesh descriptor.  A particle
field takes the first choice, in an order of preference that starts with
its default, whose time is within MARGIN of the fastest: the particles are
timed as they are when tuning runs (usually the initial conditions), and the
preferred choices are those measured best on the clustered late-time
distribution (at a = 0.1 the binning walks differ by 1%, at a = 1 blocks
beat patches by 12%).
Choices are cached per (precision, mesh shape, window order, lattice) for the
process and, when FASTPM_AMD_TUNE_CACHE names a JSON file, across processes.
"""
import json
import os

import numpy as np
import torch

from . import _lib

_CACHE = {}

MARGIN = 0.05  # a particle field leaves a preferred choice only for a gain above this


def _key(pm, order=None, parts=None):
    k = "%s:%s" % (np.dtype(pm.dtype).str, "x".join(str(int(n)) for n in pm.Nmesh))
    if order is not None:
        k += ":o%d" % order
    if parts is not None:
        from .pm import LatticeGrid

        q = parts.Q
        k += ":lat%s" % ("x".join(str(n) for n in q.nlat) if isinstance(q, LatticeGrid)
                         else "explicit%d" % len(q))
    return k


def _load_file():
    path = os.environ.get("FASTPM_AMD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        data = json.load(f)
    return {k: v for k, v in data.items() if isinstance(v, dict)}


def _save_file(entries):
    path = os.environ.get("FASTPM_AMD_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data.update(entries)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def _lookup(key):
    if key not in _CACHE:
        _CACHE.update({k: v for k, v in _load_file().items() if k == key})
    return _CACHE.get(key)


def set_tiling(pm, wide):
    """Selects the column tile width of pm's strided transform passes
    (wide = 1: 128-byte, 0: 64-byte row segments) for this mesh only."""
    pm.set_variant(fft_cols=2 if wide else 1)


def get_tiling(pm):
    """{axis length: 1 if 128-byte tiles else 0} for pm's strided axes."""
    v = pm.variant["fft_cols"]
    eb = np.dtype(pm.dtype).itemsize
    out = {}
    for n in {int(pm.Nmesh[0]), int(pm.Nmesh[1])}:
        out[n] = (v - 1) if v else int(_lib.lib().fpm_fft_get_tiling(eb, int(n)))
    return out


def _time(fn, reps):
    for _ in range(2):
        fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    b.synchronize()
    return a.elapsed_time(b) / reps


def _time_opts(opts, select, fn, reps, rounds=3):
    """{option: seconds per call}: every option timed in each of `rounds`
    interleaved rounds (select(option) first), keeping its fastest round, so
    that a slow spell of the device does not land on one option only."""
    best = {}
    for _ in range(rounds):
        for c in opts:
            select(c)
            t = _time(fn, reps)
            best[c] = min(best.get(c, t), t)
    return best


def _reps(pm):
    return max(8, min(50, int(2 ** 28 // max(int(np.prod(pm.Nmesh)), 1))))


FFT_MARGIN = 0.02  # the transform choices often lie within a few percent


def autotune_fft(pm, reps=None, verbose=False):
    """Times the transform fields (fft_cols x z_lines) on the Force's fused
    transform of pm and keeps the first choice, in order of preference,
    within FFT_MARGIN of the fastest; returns the chosen fields."""
    from .pm import ComplexField, r2c_c2r_many
    from .force.kernels import laplace

    key = _key(pm) + ":fft"
    hit = _lookup(key)
    if hit is not None:
        pm.set_variant(**hit)
        return hit
    reps = _reps(pm) if reps is None else reps
    x = pm.create("real", value=torch.randn(tuple(int(n) for n in pm.Nmesh), device=pm.device,
                                            dtype=pm.tdtype))
    outs, work = [pm.create("real")], [pm._complex_tensor()]
    ops = [laplace * 1.0]
    # in order of preference (128-byte column tiles and 2048-element z groups
    # measured best pass by pass at 512^3), the first within FFT_MARGIN of the
    # fastest: the whole-transform timing cannot resolve smaller differences
    # from run to run
    opts = [(cols, zl) for cols in (2, 1) for zl in (0, 1, 2)]
    times = _time_opts(opts, lambda c: pm.set_variant(fft_cols=c[0], z_lines=c[1]),
                       lambda: r2c_c2r_many(x, ops, kout=False, outs=outs, work=work), reps)
    fastest = min(times.values())
    cols, zl = next(o for o in opts if times[o] <= (1 + FFT_MARGIN) * fastest)
    best = dict(fft_cols=cols, z_lines=zl)
    pm.set_variant(**best)
    _CACHE[key] = best
    _save_file({key: best})
    if verbose:
        print("autotune_fft %s: %s -> %s" % (key, times, best))
    return best


def autotune_particles(pm, parts, order=None, reps=None, verbose=False):
    """Times the particle-kernel fields (bin, paint, readout) of pm on the
    binning, tiled paint and stencil force readout of `parts` (a Particles
    or position set; nothing of it is modified) and keeps the fastest of
    each; returns the chosen fields.  Meshes without the tiled path are left
    as they are."""
    from .pm import resampler_order

    order = pm.order if order is None else order
    if not pm.tiles_supported():
        return {}
    parts = pm._particles(parts)
    key = _key(pm, order, parts)
    hit = _lookup(key)
    if hit is not None:
        pm.set_variant(**hit)
        return hit
    rs = {2: "cic", 3: "tsc"}[order]
    assert resampler_order(rs) == order
    reps = max(5, min(20, int(2 ** 25 // max(len(parts), 1)))) if reps is None else reps
    rho = pm.create("real")
    phi = pm.create("real")
    phi.value.normal_()
    F = torch.empty((len(parts), 3), dtype=pm.tdtype, device=pm.device)
    R = torch.empty((len(parts),), dtype=pm.tdtype, device=pm.device)
    stencil = pm.grad4_supported(rs)
    choices = {}
    times = {}

    def bin_():
        pm.bin_particles(parts, resampler=rs)

    def paint():
        pm.paint_tiled(parts, mass=1.0, resampler=rs, out=rho)

    def readout():
        pm.readout_tiled_grad4(phi, parts, F, rho=rho, rho_out=R, resampler=rs)

    # (fields, choices in order of preference with the default first, timed
    # call); the readout's thread count, density source and span are searched
    # together (two-tile spans with 512 or 1024 threads), the density box in
    # LDS preferred over reads through the cache (measured at a = 1 on 512^3:
    # 504 vs 596 us, while they tie on the initial conditions the tuner sees)
    fields = [(("bin",), [(0,), (1,), (2,)], bin_), (("paint",), [(0,), (1,), (2,)], paint)]
    if stencil:
        nt0 = 3 if order == 2 else 2
        opts = [(nt0, 0, 0)] + [(n, r, sp) for r in (0, 1) for sp in (1, 0) for n in (3, 2, 1)
                                if (n, r, sp) != (nt0, 0, 0) and not (sp and n == 1)]
        fields.append((("readout", "rho", "span"), opts, readout))
    for names, opts, fn in fields:
        def select(c):
            pm.set_variant(**dict(zip(names, c)))
            bin_()

        t = _time_opts(opts, select, fn, reps)
        # the first choice in order of preference within MARGIN of the fastest
        best = min(t.values())
        c = next(o for o in opts if t[o] <= (1 + MARGIN) * best)
        pm.set_variant(**dict(zip(names, c)))
        choices.update(dict(zip(names, c)))
        times["/".join(names)] = t
    bin_()
    _CACHE[key] = choices
    _save_file({key: choices})
    if verbose:
        print("autotune_particles %s: %s -> %s" % (key, times, choices))
    return choices


def autotune(pm, parts=None, order=None, verbose=False):
    """Transform fields, then (given a particle set) the particle fields of
    pm; returns pm.variant."""
    autotune_fft(pm, verbose=verbose)
    if parts is not None:
        autotune_particles(pm, parts, order=order, verbose=verbose)
    return pm.variant
