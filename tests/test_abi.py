"""C ABI checks that need no GPU: the library loads, exports every function
include/fpm.h declares, and its host-side functions (growth tables, twiddles,
error strings, FFT plan query) behave."""
import os
import re

import numpy as np
import pytest

import fastpm_oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "fpm.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fpm_[a-z0-9_]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    from fastpm import _lib

    L = _lib.lib()
    names = _declared()
    assert len(names) >= 30
    for n in names:
        assert hasattr(L, n), n
    # the ctypes table covers the header exactly
    assert sorted(_lib.exported_names()) == names


def test_strerror_and_version():
    from fastpm import _lib

    L = _lib.lib()
    assert L.fpm_version().startswith(b"fpm")
    assert L.fpm_strerror(0) == b"ok"
    assert b"unsupported" in L.fpm_strerror(-2)


def test_fft_plans():
    from fastpm import _lib

    for n in [2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 3, 6, 12, 24, 48, 96, 192, 384, 768]:
        assert _lib.fft_supported(n)
    for n in [5, 10, 7, 2048, 1000]:
        assert not _lib.fft_supported(n)


@pytest.mark.parametrize("dtype", ["f4", "f8"])
def test_twiddles(dtype):
    from fastpm import _lib

    for n in (8, 12, 512):
        t = _lib.twiddles(n, dtype).reshape(-1, 2)
        ref = np.exp(-2j * np.pi * np.arange(n) / n)
        tol = 1e-15 if dtype == "f8" else 6e-8
        assert np.abs(t[:, 0] + 1j * t[:, 1] - ref).max() < tol
    with pytest.raises(RuntimeError):
        _lib.call("fpm_twiddles", dtype, 0, np.zeros(2, dtype).ctypes.data)


def test_growth_matches_oracle_and_eds():
    from fastpm import _lib

    a = np.array([0.01, 0.1, 0.35, 1.0])
    tab = _lib.growth(0.3, 0.7, 0.0, 1.0, a)
    g = O.Growth(0.3, a=a, a_normalize=1.0)
    np.testing.assert_allclose(tab[:, 0], g.D1(a), rtol=1e-7)
    np.testing.assert_allclose(tab[:, 1], g.D1(a, 1), rtol=1e-7)
    np.testing.assert_allclose(tab[:, 2], g.D1(a, 2), rtol=1e-6)
    np.testing.assert_allclose(tab[:, 3], g.D2(a), rtol=1e-6)
    np.testing.assert_allclose(tab[:, 4], g.D2(a, 1), rtol=1e-6)
    np.testing.assert_allclose(tab[:, 6], g.E(a), rtol=1e-14)
    np.testing.assert_allclose(tab[:, 7], g.E(a, 1), rtol=1e-14)
    e = _lib.growth(1.0, 0.0, 0.0, 1.0, a)
    np.testing.assert_allclose(e[:, 0], a, rtol=1e-12)
    np.testing.assert_allclose(e[:, 3], a ** 2, rtol=1e-12)
    np.testing.assert_allclose(e[:, 4] / e[:, 3], 2.0, rtol=1e-12)
    with pytest.raises(RuntimeError):
        _lib.growth(0.3, 0.7, 0.0, 1.0, [-1.0])


def test_fft_tiling_table():
    """The run-time tiling table (host-side): valid lengths and element sizes
    are stored and read back, anything else is refused."""
    from fastpm import _lib

    L = _lib.lib()
    assert L.fpm_fft_get_tiling(4, 512) in (0, 1)
    assert L.fpm_fft_set_tiling(4, 96, 1) == 0 and L.fpm_fft_get_tiling(4, 96) == 1
    assert L.fpm_fft_set_tiling(4, 96, 0) == 0 and L.fpm_fft_get_tiling(4, 96) == 0
    assert L.fpm_fft_set_tiling(8, 256, 1) == 0 and L.fpm_fft_get_tiling(8, 256) == 1
    assert L.fpm_fft_set_tiling(8, 256, 0) == 0
    assert L.fpm_fft_set_tiling(4, 100, 1) < 0      # no plan for 100
    assert L.fpm_fft_set_tiling(2, 64, 1) < 0       # element size
    assert L.fpm_fft_set_tiling(4, 64, 2) < 0       # flag
    assert L.fpm_fft_get_tiling(4, 2048) < 0
