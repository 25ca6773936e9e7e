"""Autotuner for the transform tiling.

The strided (y, x) passes of the 3D transforms come in two compiled column
tile widths (64- or 128-byte row segments, `fpm_fft_set_tiling`).  Which is
faster depends on the axis length and on what else runs (measured on gfx950:
128-byte tiles win at 512^3, 64-byte tiles at 192^3).  `autotune_fft` times
the Force's transform (forward transform of the density and inverse
transform of the potential) for each choice on scratch fields of the mesh
and keeps the faster one for the mesh's axis lengths.  Choices are cached
per (precision, shape) for the process and, when FASTPM_AMD_TUNE_CACHE names
a JSON file, across processes.  The results agree to rounding either way.
"""
import json
import os

import numpy as np
import torch

from . import _lib

_CACHE = {}


def _key(pm):
    return "%s:%s" % (np.dtype(pm.dtype).str, "x".join(str(int(n)) for n in pm.Nmesh))


def _load_file():
    path = os.environ.get("FASTPM_AMD_TUNE_CACHE")
    if not path or not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _save_file(entries):
    path = os.environ.get("FASTPM_AMD_TUNE_CACHE")
    if not path:
        return
    data = _load_file()
    data.update(entries)
    with open(path, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def set_tiling(pm, wide):
    """Selects the column tile width for the strided axes of pm's mesh."""
    eb = np.dtype(pm.dtype).itemsize
    for n in {int(pm.Nmesh[0]), int(pm.Nmesh[1])}:
        _lib.call_plain("fpm_fft_set_tiling", eb, n, int(wide))


def get_tiling(pm):
    eb = np.dtype(pm.dtype).itemsize
    return {int(n): int(_lib.lib().fpm_fft_get_tiling(eb, int(n)))
            for n in {int(pm.Nmesh[0]), int(pm.Nmesh[1])}}


def autotune_fft(pm, reps=None, verbose=False):
    """Times both column tile widths on pm's mesh and selects the faster;
    returns the chosen width flag (0: 64-byte, 1: 128-byte segments)."""
    from .pm import ComplexField, r2c_c2r_many
    from .force.kernels import laplace

    key = _key(pm)
    if key not in _CACHE:
        _CACHE.update({k: v for k, v in _load_file().items() if k == key})
    if key in _CACHE:
        set_tiling(pm, _CACHE[key])
        return _CACHE[key]
    ncell = int(np.prod(pm.Nmesh))
    if reps is None:
        reps = max(3, min(50, int(2 ** 27 // max(ncell, 1))))
    x = pm.create("real", value=torch.randn(tuple(int(n) for n in pm.Nmesh), device=pm.device,
                                            dtype=pm.tdtype))
    kout, outs, work = ComplexField(pm), [pm.create("real")], [pm._complex_tensor()]
    ops = [laplace * 1.0]
    times = {}
    for wide in (0, 1):
        set_tiling(pm, wide)
        for _ in range(2):
            r2c_c2r_many(x, ops, kout=kout, outs=outs, work=work)
        torch.cuda.synchronize(pm.device)
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        for _ in range(reps):
            r2c_c2r_many(x, ops, kout=kout, outs=outs, work=work)
        b.record()
        b.synchronize()
        times[wide] = a.elapsed_time(b) / reps
    best = min(times, key=times.get)
    set_tiling(pm, best)
    _CACHE[key] = best
    _save_file({key: best})
    if verbose:
        print("autotune_fft %s: %s -> wide=%d" % (key, times, best))
    return best
