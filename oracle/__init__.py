"""ctypes view of the FP64 checker library (oracle/tcfft_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the tcfft_amd product.  Every function
here forwards to the C restatement whose header cites the reference file:line
it follows.  Arrays are numpy complex128 (interleaved re/im, exactly the C
layout).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIBRARY selects another build of the checker (the sanitizer build of
# tests/run_sanitized.py); by default the gcc build of oracle/Makefile.
_LIB_PATH = os.environ.get("ORACLE_LIBRARY") or os.path.join(_HERE, "_build", "libtcfft_oracle.so")
_lib = None

_i32p = ctypes.POINTER(ctypes.c_int)
_f64p = ctypes.POINTER(ctypes.c_double)


def build() -> str:
    """Compile the checker with its Makefile (gcc only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.oracle_gen_rev.argtypes = [ctypes.c_int, _i32p, _i32p, ctypes.c_int]
        L.oracle_generate_data.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_get_error.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_get_error.restype = ctypes.c_double
        L.oracle_ref_plan_1d.argtypes = [ctypes.c_int, _i32p, _i32p, _i32p, _i32p]
        L.oracle_ref_plan_2d.argtypes = [ctypes.c_int, ctypes.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
        L.oracle_dit.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int]
        L.oracle_dit_2d.argtypes = [_f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    _i32p, ctypes.c_int, _i32p, ctypes.c_int, ctypes.c_int]
        L.oracle_fft.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_fft_2d.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.oracle_dft_ld.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_rfft.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_irfft.argtypes = [_f64p, _f64p, ctypes.c_int, ctypes.c_int]
        L.oracle_rel_l2.argtypes = [_f64p, _f64p, ctypes.c_size_t]
        L.oracle_rel_l2.restype = ctypes.c_double
        L.oracle_default_radices.argtypes = [ctypes.c_int, _i32p, ctypes.c_int]
        _lib = L
    return _lib


def _ip(a: np.ndarray):
    return a.ctypes.data_as(_i32p)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(_f64p)


def _c128(x) -> np.ndarray:
    return np.ascontiguousarray(x, dtype=np.complex128)


def gen_rev(n: int, radices) -> np.ndarray:
    """tcfft_doit_half.cpp:6-29."""
    r = np.ascontiguousarray(radices, dtype=np.int32)
    out = np.empty(n, dtype=np.int32)
    lib().oracle_gen_rev(n, _ip(out), _ip(r), len(r))
    return out


def generate_data(n: int, n_batch: int, seed: int = 42) -> np.ndarray:
    """doit.h:7-19 (glibc rand), as complex128 of shape (n_batch, n)."""
    out = np.empty((n_batch, n), dtype=np.complex128)
    lib().oracle_generate_data(_dp(out), n, n_batch, seed)
    return out


def get_error(tested, standard) -> float:
    """accuracy.cpp:25-41 over arrays of shape (n_batch, n)."""
    t = _c128(tested)
    s = _c128(standard)
    n = t.shape[-1]
    b = t.size // n
    return lib().oracle_get_error(_dp(t), _dp(s), n, b)


def ref_plan_1d(n: int):
    """tcfft_half.cu:1235-1330.  Returns (ok, radices[:n_radices], mergings[:n_mergings])."""
    rad = np.zeros(9, dtype=np.int32)
    mer = np.zeros(3, dtype=np.int32)
    nr = ctypes.c_int()
    nm = ctypes.c_int()
    rc = lib().oracle_ref_plan_1d(n, _ip(rad), ctypes.byref(nr), _ip(mer), ctypes.byref(nm))
    return rc == 0, list(rad[: nr.value]), list(mer[: nm.value])


def ref_plan_2d(nx: int, ny: int):
    """tcfft_half_2d.cu:1151-1206.  Returns (ok, radices_x, radices_y, mergings)."""
    rx = np.zeros(3, dtype=np.int32)
    ry = np.zeros(3, dtype=np.int32)
    mer = np.zeros(2, dtype=np.int32)
    nrx = ctypes.c_int()
    nry = ctypes.c_int()
    rc = lib().oracle_ref_plan_2d(nx, ny, _ip(rx), ctypes.byref(nrx), _ip(ry), ctypes.byref(nry), _ip(mer))
    return rc == 0, list(rx[: nrx.value]), list(ry[: nry.value]), list(mer)


def dit(x, radices, sign: int = -1) -> np.ndarray:
    """In-place DIT passes on digit-reversed input (tcfft_half.cu:1177-1233)."""
    a = _c128(x).copy()
    n = a.shape[-1]
    b = a.size // n
    r = np.ascontiguousarray(radices, dtype=np.int32)
    lib().oracle_dit(_dp(a), n, b, _ip(r), len(r), sign)
    return a


def dit_2d(x, radices_x, radices_y, sign: int = -1) -> np.ndarray:
    """tcfft_half_2d.cu:1128-1149 on input digit-reversed along both axes."""
    a = _c128(x).copy()
    nx, ny = a.shape[-2], a.shape[-1]
    b = a.size // (nx * ny)
    rx = np.ascontiguousarray(radices_x, dtype=np.int32)
    ry = np.ascontiguousarray(radices_y, dtype=np.int32)
    lib().oracle_dit_2d(_dp(a), nx, ny, b, _ip(rx), len(rx), _ip(ry), len(ry), sign)
    return a


def fft(x, sign: int = -1) -> np.ndarray:
    """Natural-order FP64 FFT along the last axis (FFTW's role, accuracy.cpp:11-23)."""
    a = _c128(x)
    n = a.shape[-1]
    b = a.size // n
    out = np.empty_like(a)
    lib().oracle_fft(_dp(a), _dp(out), n, b, sign)
    return out


def fft_2d(x, sign: int = -1) -> np.ndarray:
    """Natural-order FP64 2D FFT over the last two axes (accuracy_2d.cpp:11-23)."""
    a = _c128(x)
    nx, ny = a.shape[-2], a.shape[-1]
    b = a.size // (nx * ny)
    out = np.empty_like(a)
    lib().oracle_fft_2d(_dp(a), _dp(out), nx, ny, b, sign)
    return out


def rfft(x) -> np.ndarray:
    """Bins 0..n/2 of the DFT of real signals along the last axis (the
    real-as-complex transform the reference's callers run, truncated)."""
    a = np.ascontiguousarray(x, dtype=np.float64)
    n = a.shape[-1]
    b = a.size // n
    out = np.empty(a.shape[:-1] + (n // 2 + 1,), dtype=np.complex128)
    lib().oracle_rfft(_dp(a), _dp(out), n, b)
    return out


def irfft(X, n: int) -> np.ndarray:
    """Unnormalised inverse of rfft: n * numpy.fft.irfft(X, n)."""
    a = _c128(X)
    b = a.size // (n // 2 + 1)
    out = np.empty(a.shape[:-1] + (n,), dtype=np.float64)
    lib().oracle_irfft(_dp(a), _dp(out), n, b)
    return out


def dft_ld(x, sign: int = -1) -> np.ndarray:
    """Brute-force long-double DFT of one signal."""
    a = _c128(x).reshape(-1)
    out = np.empty_like(a)
    lib().oracle_dft_ld(_dp(a), _dp(out), a.size, sign)
    return out


def rel_l2(a, ref) -> float:
    x = _c128(a)
    y = _c128(ref)
    return lib().oracle_rel_l2(_dp(x), _dp(y), x.size * 2)


def default_radices(n: int):
    out = np.zeros(32, dtype=np.int32)
    k = lib().oracle_default_radices(n, _ip(out), 32)
    return list(out[:k])
