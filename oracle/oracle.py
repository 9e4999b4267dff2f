"""ctypes wrapper for the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` and by the
``cpu_baseline`` leg of ``bench.py`` -- never by the product package.

Arrays cross the boundary as numpy complex64/complex128 (interleaved re,im in memory, the
layout of GFFT's ``T* data``).  Sign convention: S=+1 forward (exp(-2 pi i nk/N)),
S=-1 inverse (exp(+), and the GFFT compositions divide by N).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build() -> str:
    """Compile liboracle.so with the Makefile next to this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp, lg, ci = ctypes.c_void_p, ctypes.c_long, ctypes.c_int
        for suf in ("_d", "_f"):
            getattr(L, "orc_swap" + suf).argtypes = [vp, lg, lg]
            getattr(L, "orc_gfft_inplace" + suf).argtypes = [vp, lg, ci]
            getattr(L, "orc_gfft_oop" + suf).argtypes = [vp, vp, lg, ci]
            getattr(L, "orc_gfft_real_inplace" + suf).argtypes = [vp, lg, ci]
            getattr(L, "orc_gfft_real_oop" + suf).argtypes = [vp, vp, lg, ci]
            getattr(L, "orc_gfft_inplace_omp" + suf).argtypes = [vp, lg, ci, ci]
            getattr(L, "orc_gfft_oop_omp" + suf).argtypes = [vp, vp, lg, ci, ci]
            getattr(L, "orc_separate" + suf).argtypes = [vp, lg, ci]
            getattr(L, "orc_backward" + suf).argtypes = [vp, lg]
            for name in ("orc_gfft_inplace", "orc_gfft_oop", "orc_gfft_real_inplace",
                         "orc_gfft_real_oop", "orc_gfft_inplace_omp", "orc_gfft_oop_omp"):
                getattr(L, name + suf).restype = ci
        L.orc_ref_fft.argtypes = [vp, vp, lg, ci]
        L.orc_ref_fft.restype = ci
        L.orc_dft_ld.argtypes = [vp, vp, lg, ci]
        L.orc_dft_ld.restype = ci
        L.orc_root_d.argtypes = [lg, lg, ci, ctypes.POINTER(ctypes.c_double),
                                 ctypes.POINTER(ctypes.c_double)]
        L.orc_digit_reverse_map.argtypes = [lg, lg, vp]
        L.orc_factorize.argtypes = [lg, vp]
        L.orc_factorize.restype = ci
        _lib = L
    return _lib


def _suffix(a: np.ndarray) -> str:
    if a.dtype == np.complex128:
        return "_d"
    if a.dtype == np.complex64:
        return "_f"
    raise TypeError(f"oracle takes complex64/complex128, got {a.dtype}")


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


def _check(rc: int, what: str):
    if rc != 0:
        raise ValueError(f"oracle {what} failed (rc={rc})")


# ---- GFFT restatement ---------------------------------------------------------------------

def gfft_inplace(x: np.ndarray, sign: int = 1, nthreads: int = 1) -> np.ndarray:
    """GFFT in-place DFT (sign=+1) / IDFT (sign=-1, scaled 1/N), N a prime power.

    nthreads > 1 selects the OpenMP split (power-of-two N and NT)."""
    y = np.ascontiguousarray(x).copy()
    s = _suffix(y)
    if nthreads > 1:
        rc = getattr(lib(), "orc_gfft_inplace_omp" + s)(_ptr(y), y.size, sign, nthreads)
    else:
        rc = getattr(lib(), "orc_gfft_inplace" + s)(_ptr(y), y.size, sign)
    _check(rc, "gfft_inplace")
    return y


def gfft_oop(x: np.ndarray, sign: int = 1, nthreads: int = 1) -> np.ndarray:
    """GFFT out-of-place DFT/IDFT for any N with prime factors < 64."""
    src = np.ascontiguousarray(x)
    dst = np.empty_like(src)
    s = _suffix(src)
    if nthreads > 1:
        rc = getattr(lib(), "orc_gfft_oop_omp" + s)(_ptr(src), _ptr(dst), src.size, sign, nthreads)
    else:
        rc = getattr(lib(), "orc_gfft_oop" + s)(_ptr(src), _ptr(dst), src.size, sign)
    _check(rc, "gfft_oop")
    return dst


def gfft_real_inplace(x: np.ndarray, sign: int = 1) -> np.ndarray:
    """RDFT (sign=+1) / IRDFT (sign=-1) on 2N reals viewed as N complex (NR packing)."""
    y = np.ascontiguousarray(x).copy()
    rc = getattr(lib(), "orc_gfft_real_inplace" + _suffix(y))(_ptr(y), y.size, sign)
    _check(rc, "gfft_real_inplace")
    return y


def gfft_real_oop(x: np.ndarray, sign: int = 1) -> np.ndarray:
    """RDFT/IRDFT through the out-of-place complex path (any N)."""
    src = np.ascontiguousarray(x)
    dst = np.empty_like(src)
    rc = getattr(lib(), "orc_gfft_real_oop" + _suffix(src))(_ptr(src), _ptr(dst), src.size, sign)
    _check(rc, "gfft_real_oop")
    return dst


def swap(x: np.ndarray, radix: int, digits: int) -> np.ndarray:
    """GFFTswap2<radix,digits>: in-place digit reversal."""
    y = np.ascontiguousarray(x).copy()
    getattr(lib(), "orc_swap" + _suffix(y))(_ptr(y), radix, digits)
    return y


def separate(x: np.ndarray, sign: int) -> np.ndarray:
    y = np.ascontiguousarray(x).copy()
    getattr(lib(), "orc_separate" + _suffix(y))(_ptr(y), y.size, sign)
    return y


# ---- accurate references ------------------------------------------------------------------

def ref_fft(x: np.ndarray, sign: int = 1) -> np.ndarray:
    """Long-double Stockham FFT (unnormalised), result rounded to complex128."""
    src = np.ascontiguousarray(x, dtype=np.complex128)
    dst = np.empty_like(src)
    _check(lib().orc_ref_fft(_ptr(src), _ptr(dst), src.size, sign), "ref_fft")
    return dst


def dft_ld(x: np.ndarray, sign: int = 1) -> np.ndarray:
    """Long-double O(N^2) DFT with exact index reduction (small N)."""
    src = np.ascontiguousarray(x, dtype=np.complex128)
    dst = np.empty_like(src)
    _check(lib().orc_dft_ld(_ptr(src), _ptr(dst), src.size, sign), "dft_ld")
    return dst


def root(m: int, n: int, sign: int = 1) -> complex:
    c, s = ctypes.c_double(), ctypes.c_double()
    lib().orc_root_d(m, n, sign, ctypes.byref(c), ctypes.byref(s))
    return complex(c.value, s.value)


def digit_reverse_map(radix: int, digits: int) -> np.ndarray:
    out = np.empty(radix ** digits, dtype=np.int64)
    lib().orc_digit_reverse_map(radix, digits, _ptr(out))
    return out


def factorize(n: int) -> list[int]:
    f = np.zeros(64, dtype=np.int64)
    k = lib().orc_factorize(n, _ptr(f))
    if k < 0:
        raise ValueError(f"{n} has a prime factor beyond the oracle's leaf limit")
    return [int(v) for v in f[:k]]


# ---- error norms (test/gfft_accuracy.h:30-89, :190-203) -----------------------------------

def errors(y: np.ndarray, ref: np.ndarray) -> dict:
    """relL2, max-abs and GFFT's max|err|_inf / |ref|_inf over all re/im components."""
    y = np.asarray(y, dtype=np.complex128).ravel()
    ref = np.asarray(ref, dtype=np.complex128).ravel()
    d = np.concatenate([(y - ref).real, (y - ref).imag])
    r = np.concatenate([ref.real, ref.imag])
    nref2 = float(np.sqrt(np.dot(r, r)))
    ninf = float(np.max(np.abs(r))) if r.size else 0.0
    maxabs = float(np.max(np.abs(d))) if d.size else 0.0
    return {
        "rel_l2": float(np.sqrt(np.dot(d, d))) / nref2 if nref2 > 0 else float(np.sqrt(np.dot(d, d))),
        "max_abs": maxabs,
        "rel_inf": maxabs / ninf if ninf > 0 else maxabs,
    }
