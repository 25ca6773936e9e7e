"""ctypes binding of libfpm.so, the C ABI over the gfx950 HIP kernels.

The library is built in-tree (``make -C fastpm-python_amd/csrc``).  There is no
fallback: if the library is missing, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfpm.so")

FPM_OK = 0


class FpmParticles(ctypes.Structure):
    _fields_ = [
        ("q", ctypes.c_void_p),
        ("s", ctypes.c_void_p),
        ("n", ctypes.c_int64),
        ("lattice", ctypes.c_int32 * 3),
        ("pad_", ctypes.c_int32),
        ("lattice_shift", ctypes.c_double),
    ]


class FpmMesh(ctypes.Structure):
    _fields_ = [
        ("nmesh", ctypes.c_int32 * 3),
        ("variant", ctypes.c_int32),
        ("box", ctypes.c_double * 3),
    ]


# fields of FpmMesh.variant (fpm.h FPM_VAR_*): name -> (bit shift, choices)
VARIANT_FIELDS = {
    "fft_cols": (0, (0, 1, 2)),
    "z_lines": (2, (0, 1, 2)),
    "bin": (4, (0, 1, 2)),
    "readout": (6, (0, 1, 2, 3)),
    "paint": (8, (0, 1, 2)),
    "rho": (10, (0, 1)),
    "span": (12, (0, 1)),
}


def pack_variant(choices):
    """FpmMesh.variant from {field: choice}; unnamed fields stay 0 (default)."""
    v = 0
    for name, c in choices.items():
        shift, allowed = VARIANT_FIELDS[name]
        if int(c) not in allowed:
            raise ValueError("variant %s=%r not in %s" % (name, c, allowed))
        v |= int(c) << shift
    return v


def unpack_variant(v):
    return {name: (int(v) >> shift) & 3 for name, (shift, _) in VARIANT_FIELDS.items()}


class FpmKop(ctypes.Structure):
    _fields_ = [
        ("factor", ctypes.c_double),
        ("r_split", ctypes.c_double),
        ("laplace", ctypes.c_int32),
        ("ngrad", ctypes.c_int32),
        ("grad_dir", ctypes.c_int32 * 2),
        ("grad_order", ctypes.c_int32),
        ("deconv", ctypes.c_int32),
    ]


class FpmBatchBox(ctypes.Structure):
    """fpm.h fpm_batch_box: one box of a batched step."""
    _fields_ = [
        ("part", FpmParticles),
        ("mesh", FpmMesh),
        ("order", ctypes.c_int32),
        ("pad_", ctypes.c_int32),
        ("mass", ctypes.c_double),
        ("common", FpmKop),
        ("pitch2", ctypes.c_int64),
        ("rho", ctypes.c_void_p),
        ("phi", ctypes.c_void_p),
        ("work", ctypes.c_void_p),
        ("tw", ctypes.c_void_p * 3),
        ("ws", ctypes.c_void_p),
        ("ws_bytes", ctypes.c_int64),
        ("rho_out", ctypes.c_void_p),
        ("rho_stride", ctypes.c_int64),
        ("force_out", ctypes.c_void_p),
        ("force_stride", ctypes.c_int64),
        ("rho_scale", ctypes.c_double),
        ("force_scale", ctypes.c_double),
        ("P", ctypes.c_void_p),
        ("table", ctypes.c_void_p),
        ("rows", ctypes.c_int32),
        ("stride", ctypes.c_int32),
        ("close", ctypes.c_int32),
        ("open", ctypes.c_int32),
        ("drift", ctypes.c_int32),
        ("pad2_", ctypes.c_int32),
    ]


_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_PP = ctypes.POINTER(FpmParticles)
_PM = ctypes.POINTER(FpmMesh)
_PK = ctypes.POINTER(FpmKop)

# name -> (restype, argtypes); every compute function exists as _f32 and _f64
_TYPED = {
    "fpm_paint": (_int, [_PP, _PM, _int, _dbl, _vp, _int, _vp]),
    "fpm_readout": (_int, [_PP, _PM, _int, _int, _vp, _vp, _vp, _vp, _vp]),
    "fpm_fft_r2c": (_int, [_vp, _vp, _PM, _i64, _vp, _vp, _vp, _vp]),
    "fpm_axpby": (_int, [_vp, _vp, _i64, _dbl, _dbl, _vp]),
    "fpm_grad4_mesh": (_int, [_vp, _vp, _PM, _dbl, _vp]),
    "fpm_grad4_div": (_int, [_vp, _vp, _PM, _dbl, _int, _vp]),
    "fpm_drift_wrap": (_int, [_vp, _vp, _i64, _dbl, _vp, _vp]),
    "fpm_kick_drift": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _dbl, _vp]),
    "fpm_drift_kick_vjp": (_int, [_vp, _vp, _vp, _i64, _vp, _int, _dbl, _int, _vp]),
    "fpm_fft_c2r": (_int, [_vp, _int, _PK, _vp, _vp, _PM, _i64, _vp, _vp, _vp, _vp]),
    "fpm_fft_r2c_c2r": (_int, [_vp, _vp, _int, _PK, _vp, _vp, _PM, _i64, _vp, _vp, _vp, _vp]),
    "fpm_fft_c2r_grad3": (_int, [_vp, _PK, _vp, _vp, _PM, _i64, _vp, _vp, _vp, _vp]),
    "fpm_kspace_apply": (_int, [_vp, _PK, _PM, _i64, _vp]),
    "fpm_whitenoise": (_int, [_vp, _PM, _i64, ctypes.c_uint64, _int, _vp]),
    "fpm_linear_scale": (_int, [_vp, _PM, _i64, _vp, _i64, _vp]),
    "fpm_linear_scale_indexed": (_int, [_vp, _PM, _i64, _vp, _i64, _vp, _vp]),
    "fpm_lpt2_source": (_int, [_vp, _vp, _i64, _vp]),
    "fpm_axpy": (_int, [_vp, _vp, _i64, _dbl, _vp]),
    "fpm_lpt_combine": (_int, [_vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "fpm_positions": (_int, [_PP, _PM, _vp, _vp]),
    "fpm_power_bin": (_int, [_vp, _vp, _PM, _i64, _int, _dbl, _dbl, _vp, _i64, _vp, _vp]),
    "fpm_power_bin_noalloc": (_int, [_vp, _vp, _PM, _i64, _int, _dbl, _dbl, _vp, _vp]),
    "fpm_twiddles": (_int, [_int, _vp]),
    "fpm_paint_weighted": (_int, [_PP, _PM, _int, _vp, _i64, _dbl, _vp, _int, _vp]),
    "fpm_window_grad": (_int, [_PP, _PM, _int, _int, _vp, _vp, _vp, _dbl, _vp, _int, _vp]),
    "fpm_kspace_accumulate": (_int, [_vp, _vp, _PK, _int, _int, _PM, _i64, _vp]),
    "fpm_hermitian_scale": (_int, [_vp, _dbl, _dbl, _PM, _i64, _vp]),
    "fpm_lpt2_source_vjp": (_int, [_vp, _vp, _vp, _i64, _vp]),
    "fpm_axpy_table": (_int, [_vp, _vp, _i64, _vp, _vp, _int, _int, _int, _vp]),
    "fpm_kick_drift_table": (_int, [_vp, _vp, _vp, _i64, _vp, _vp, _int, _int, _int, _int, _int,
                                    _vp]),
}
_PLAIN = {
    "fpm_version": (ctypes.c_char_p, []),
    "fpm_strerror": (ctypes.c_char_p, [_int]),
    "fpm_fft_supported": (_int, [_int]),
    "fpm_fft_set_tiling": (_int, [_int, _int, _int]),
    "fpm_fft_get_tiling": (_int, [_int, _int]),
    "fpm_power_work_len": (_i64, [_PM, _int]),
    "fpm_growth": (_int, [_dbl, _dbl, _dbl, _dbl, _vp, _int, _vp]),
    "fpm_tiles_supported": (_int, [_PM]),
    "fpm_tile_workspace_size": (_i64, [_PM, _int, _i64]),
    "fpm_tile_misplaced_offset": (_i64, [_PM, _int, _i64]),
    "fpm_bin_particles_f32": (_int, [_PP, _PM, _int, _vp, _i64, _vp]),
    "fpm_rebin_particles_f32": (_int, [_PP, _PM, _int, _vp, _i64, _vp]),
    "fpm_kick_drift_rebin_f32": (_int, [_PP, _PM, _int, _vp, _vp, _vp, _int, _dbl, _vp, _i64,
                                        _vp]),
    "fpm_paint_tiled_f32": (_int, [_PP, _PM, _int, _dbl, _vp, _vp, _i64, _vp]),
    "fpm_readout_tiled_f32": (_int, [_PP, _PM, _int, _int, _vp, _vp, _vp, _vp, _vp, _i64, _vp]),
    "fpm_paint_tiled_weighted_f32": (_int, [_PP, _PM, _int, _int, _vp, _vp, _dbl, _vp, _vp, _i64,
                                            _vp]),
    "fpm_readout_tiled_grad4_f32": (_int, [_PP, _PM, _int, _vp, _vp, _vp, _i64, _vp, _i64, _dbl,
                                           _dbl, _vp, _i64, _vp]),
    "fpm_window_grad_tiled_grad4_f32": (_int, [_PP, _PM, _int, _vp, _vp, _i64, _vp, _dbl, _int,
                                               _vp, _i64, _vp]),
    "fpm_counter_add": (_int, [_vp, ctypes.c_int32, _vp]),
    "fpm_batch_device_bytes": (_i64, [_int, _vp]),
    "fpm_batch_create": (_vp, [_int, _vp, _vp, _i64, ctypes.POINTER(ctypes.c_int)]),
    "fpm_batch_step": (_int, [_vp, _vp, _vp]),
    "fpm_batch_destroy": (None, [_vp]),
}


def exported_names():
    names = list(_PLAIN)
    for n in _TYPED:
        names += [n + "_f32", n + "_f64"]
    return names


_lib = None


def lib():
    """Load libfpm.so once; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                "libfpm.so not found at %s: build it with `make -C fastpm-python_amd/csrc` "
                "(there is no CPU fallback)" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        for n, (rt, at) in _PLAIN.items():
            f = getattr(L, n)
            f.restype, f.argtypes = rt, at
        for n, (rt, at) in _TYPED.items():
            for suf in ("_f32", "_f64"):
                f = getattr(L, n + suf)
                f.restype, f.argtypes = rt, at
        _lib = L
    return _lib


def suffix(dtype):
    dtype = np.dtype(dtype)
    if dtype == np.float32:
        return "_f32"
    if dtype == np.float64:
        return "_f64"
    raise TypeError("only float32 and float64 storage are supported, got %s" % dtype)


def call(name, dtype, *args):
    """Call a typed entry point and raise on a non-zero status."""
    f = getattr(lib(), name + suffix(dtype))
    rc = f(*args)
    if rc != FPM_OK:
        raise RuntimeError("%s%s failed: %s (%d)" % (name, suffix(dtype),
                                                      lib().fpm_strerror(rc).decode(), rc))


def call_plain(name, *args):
    f = getattr(lib(), name)
    rc = f(*args)
    if rc != FPM_OK:
        raise RuntimeError("%s failed: %s (%d)" % (name, lib().fpm_strerror(rc).decode(), rc))


def growth(omega_m, omega_lambda, omega_k, a_normalize, a):
    """Growth table (host C++): rows of D1, D1', D1'', D2, D2', D2'', E, E', Om."""
    a = np.ascontiguousarray(np.atleast_1d(np.asarray(a, dtype="f8")))
    out = np.zeros((len(a), 9), dtype="f8")
    call_plain("fpm_growth", float(omega_m), float(omega_lambda), float(omega_k),
               float(a_normalize), a.ctypes.data, len(a), out.ctypes.data)
    return out


def twiddles(n, dtype):
    out = np.zeros(2 * n, dtype=dtype)
    call("fpm_twiddles", dtype, int(n), out.ctypes.data)
    return out


def fft_supported(n):
    return bool(lib().fpm_fft_supported(int(n)))


def ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))(*[ctypes.c_void_p(int(p)) for p in ptrs])
    return arr
