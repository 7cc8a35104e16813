"""ctypes front end of the CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product path
(flagfft_amd) never imports this module.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_int = ctypes.c_int
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> str:
    """Compile oracle.c with gcc (make) into oracle/_build/liboracle.so."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "ora_rng_signed_unit": (None, [_u64, _i64, _dp]),
            "ora_rng_normal": (None, [_u64, _i64, _dp]),
            "ora_accuracy_seed": (_u64, [_int, _int, _int, _u64]),
            "ora_dft_ld": (None, [_dp, _dp, _i64, _i64, _int, _int]),
            "ora_fft_c2c_d": (_int, [_dp, _dp, _i64, _i64, _int, _int]),
            "ora_fft_c2c_f": (_int, [_fp, _fp, _i64, _i64, _int, _int]),
            "ora_fft_r2c_d": (_int, [_dp, _dp, _i64, _i64, _i64, _i64, _int]),
            "ora_fft_r2c_f": (_int, [_fp, _fp, _i64, _i64, _i64, _i64, _int]),
            "ora_fft_c2r_d": (_int, [_dp, _dp, _i64, _i64, _i64, _i64, _int]),
            "ora_fft_c2r_f": (_int, [_fp, _fp, _i64, _i64, _i64, _i64, _int]),
            "ora_fft2d_c2c_d": (_int, [_dp, _dp, _i64, _i64, _i64, _int, _int]),
            "ora_fft2d_r2c_d": (_int, [_dp, _dp, _i64, _i64, _i64, _int]),
            "ora_fft2d_c2r_d": (_int, [_dp, _dp, _i64, _i64, _i64, _int]),
            "ora_primitive_root": (_i64, [_i64]),
            "ora_rader_index_table": (None, [_i64, _i64, _i64p]),
            "ora_bluestein_chirp": (None, [_i64, _int, _dp]),
            "ora_rader_c2c_d": (_int, [_dp, _dp, _i64, _i64, _int]),
            "ora_bluestein_c2c_d": (_int, [_dp, _dp, _i64, _i64, _i64, _int]),
            "ora_rel_l2": (ctypes.c_double, [_dp, _dp, _i64, _i64, _int, _dp]),
            "ora_work_factor": (ctypes.c_double, [_i64]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _ptr(a: np.ndarray, kind=_dp):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(kind)


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


# ---- inputs ---------------------------------------------------------------
def signed_unit(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().ora_rng_signed_unit(seed & (2**64 - 1), count, _ptr(out))
    return out


def normal(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    lib().ora_rng_normal(seed & (2**64 - 1), count, _ptr(out))
    return out


def accuracy_seed(type_code: int, n: int, batch: int, variant: int = 0) -> int:
    return int(lib().ora_accuracy_seed(type_code, n, batch, variant))


def complex_normal(seed: int, shape) -> np.ndarray:
    count = int(np.prod(shape))
    v = normal(seed, 2 * count)
    return (v[0::2] + 1j * v[1::2]).reshape(shape)


def complex_signed_unit(seed: int, shape) -> np.ndarray:
    count = int(np.prod(shape))
    v = signed_unit(seed, 2 * count)
    return (v[0::2] + 1j * v[1::2]).reshape(shape)


# ---- transforms (complex128 / float64 numpy in, numpy out) -----------------
def dft(x, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """Naive long-double DFT along the last axis."""
    x = _c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    lib().ora_dft_ld(_ptr(x.view(np.float64)), _ptr(out.view(np.float64)), n, x.size // n, sign, nthreads)
    return out


def fft(x, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """Fast FP64 C2C along the last axis (contiguous rows)."""
    x = _c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    rc = lib().ora_fft_c2c_d(_ptr(x.view(np.float64)), _ptr(out.view(np.float64)), n, x.size // n, sign, nthreads)
    assert rc == 0
    return out


def fft_f32(x, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """Fast FP32 C2C along the last axis (the FP32 CPU baseline path)."""
    x = np.ascontiguousarray(x, dtype=np.complex64)
    n = x.shape[-1]
    out = np.empty_like(x)
    rc = lib().ora_fft_c2c_f(_ptr(x.view(np.float32), _fp), _ptr(out.view(np.float32), _fp), n, x.size // n,
                             sign, nthreads)
    assert rc == 0
    return out


def rfft_f32(x, nthreads: int = 0) -> np.ndarray:
    """Fast FP32 R2C (the FP32 CPU baseline path)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[-1]
    h = n // 2 + 1
    out = np.empty(x.shape[:-1] + (h,), dtype=np.complex64)
    rc = lib().ora_fft_r2c_f(_ptr(x, _fp), _ptr(out.view(np.float32), _fp), n, x.size // n, n, h, nthreads)
    assert rc == 0
    return out


def irfft_f32(X, n: int, nthreads: int = 0) -> np.ndarray:
    """Fast FP32 unnormalised C2R."""
    X = np.ascontiguousarray(X, dtype=np.complex64)
    h = n // 2 + 1
    out = np.empty(X.shape[:-1] + (n,), dtype=np.float32)
    rc = lib().ora_fft_c2r_f(_ptr(X.view(np.float32), _fp), _ptr(out, _fp), n, X.size // h, h, n, nthreads)
    assert rc == 0
    return out


def rfft(x, nthreads: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n = x.shape[-1]
    b = x.size // n
    h = n // 2 + 1
    out = np.empty(x.shape[:-1] + (h,), dtype=np.complex128)
    rc = lib().ora_fft_r2c_d(_ptr(x), _ptr(out.view(np.float64)), n, b, n, h, nthreads)
    assert rc == 0
    return out


def irfft(X, n: int, nthreads: int = 0) -> np.ndarray:
    """Unnormalised C2R ignoring Im of the DC / Nyquist bins."""
    X = _c128(X)
    h = n // 2 + 1
    assert X.shape[-1] == h
    b = X.size // h
    out = np.empty(X.shape[:-1] + (n,), dtype=np.float64)
    rc = lib().ora_fft_c2r_d(_ptr(X.view(np.float64)), _ptr(out), n, b, h, n, nthreads)
    assert rc == 0
    return out


def fft2(x, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    x = _c128(x)
    n0, n1 = x.shape[-2:]
    out = np.empty_like(x)
    rc = lib().ora_fft2d_c2c_d(_ptr(x.view(np.float64)), _ptr(out.view(np.float64)), n0, n1,
                               x.size // (n0 * n1), sign, nthreads)
    assert rc == 0
    return out


def rfft2(x, nthreads: int = 0) -> np.ndarray:
    x = np.ascontiguousarray(x, dtype=np.float64)
    n0, n1 = x.shape[-2:]
    out = np.empty(x.shape[:-1] + (n1 // 2 + 1,), dtype=np.complex128)
    rc = lib().ora_fft2d_r2c_d(_ptr(x), _ptr(out.view(np.float64)), n0, n1, x.size // (n0 * n1), nthreads)
    assert rc == 0
    return out


def irfft2(X, n1: int, nthreads: int = 0) -> np.ndarray:
    X = _c128(X)
    n0 = X.shape[-2]
    h = n1 // 2 + 1
    assert X.shape[-1] == h
    out = np.empty(X.shape[:-1] + (n1,), dtype=np.float64)
    rc = lib().ora_fft2d_c2r_d(_ptr(X.view(np.float64)), _ptr(out), n0, n1, X.size // (n0 * h), nthreads)
    assert rc == 0
    return out


def _along(x, axis: int, f):
    """Apply a last-axis transform f along `axis` (FP64, contiguous copies)."""
    y = f(np.ascontiguousarray(np.moveaxis(x, axis, -1)))
    return np.ascontiguousarray(np.moveaxis(y, -1, axis))


def fft3(x, sign: int = -1, nthreads: int = 0) -> np.ndarray:
    """C2C over the last three axes by separability: the 2D transform of the
    last two axes, then the 1D transform along axis -3."""
    return _along(fft2(x, sign, nthreads), -3, lambda v: fft(v, sign, nthreads))


def rfft3(x, nthreads: int = 0) -> np.ndarray:
    """R2C over the last three axes: rfft2 of the last two, then C2C along axis -3."""
    return _along(rfft2(x, nthreads), -3, lambda v: fft(v, -1, nthreads))


def irfft3(X, n2: int, nthreads: int = 0) -> np.ndarray:
    """Unnormalised C2R over the last three axes (Im of the DC / Nyquist bins
    of the last axis ignored): inverse C2C along axis -3, then irfft2."""
    return irfft2(_along(_c128(X), -3, lambda v: fft(v, 1, nthreads)), n2, nthreads)


def primitive_root(p: int) -> int:
    return int(lib().ora_primitive_root(p))


def rader_index_table(p: int, g: int) -> np.ndarray:
    out = np.empty(p - 1, dtype=np.int64)
    lib().ora_rader_index_table(p, g, _ptr(out, _i64p))
    return out


def bluestein_chirp(n: int, sign: int) -> np.ndarray:
    out = np.empty(n, dtype=np.complex128)
    lib().ora_bluestein_chirp(n, sign, _ptr(out.view(np.float64)))
    return out


def rader_fft(x, sign: int = -1) -> np.ndarray:
    x = _c128(x)
    p = x.shape[-1]
    out = np.empty_like(x)
    rc = lib().ora_rader_c2c_d(_ptr(x.view(np.float64)), _ptr(out.view(np.float64)), p, x.size // p, sign)
    assert rc == 0
    return out


def bluestein_fft(x, m: int, sign: int = -1) -> np.ndarray:
    x = _c128(x)
    n = x.shape[-1]
    out = np.empty_like(x)
    rc = lib().ora_bluestein_c2c_d(_ptr(x.view(np.float64)), _ptr(out.view(np.float64)), n, m, x.size // n, sign)
    assert rc == 0
    return out


# ---- metrics -------------------------------------------------------------
def rel_l2(value, ref, row_len: int | None = None):
    """Worst-row relative L2 and Linf (ctest/flagfft_test.h:222-262), long double."""
    ref = np.asarray(ref)
    value = np.asarray(value)
    cplx = np.iscomplexobj(ref) or np.iscomplexobj(value)
    if cplx:
        v = np.ascontiguousarray(value, dtype=np.complex128).view(np.float64)
        r = np.ascontiguousarray(ref, dtype=np.complex128).view(np.float64)
    else:
        v = np.ascontiguousarray(value, dtype=np.float64)
        r = np.ascontiguousarray(ref, dtype=np.float64)
    total = ref.size
    row_len = row_len or (ref.shape[-1] if ref.ndim else 1)
    linf = ctypes.c_double(0.0)
    l2 = lib().ora_rel_l2(_ptr(v), _ptr(r), row_len, total // row_len, 1 if cplx else 0, ctypes.byref(linf))
    return float(l2), float(linf.value)


def work_factor(n: int) -> float:
    return float(lib().ora_work_factor(n))
