"""Seeded synthetic inputs for the GFFT transform workloads.

SplitMix64 (Steele, Lea & Flood, "Fast splittable pseudorandom number generators", 2014)
drives everything: draw i (1-based) is mix(seed + i * 0x9E3779B97F4A7C15).  Uniforms take the
top 53 bits; normals use the Box-Muller transform on consecutive pairs.  C ``rand()`` and
``std::normal_distribution`` are avoided because they are implementation-defined.

Interleaved layout: element i of a complex signal takes draw 2i for its real part and draw
2i+1 for its imaginary part (the (re,im) pairs of GFFT's ``T* data``).
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """Draws offset+1 .. offset+n of the SplitMix64 stream started at ``seed``."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + i * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """U[0,1) from the top 53 bits."""
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def uniform_pm1(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """U[-1,1)."""
    return 2.0 * uniform01(seed, n, offset) - 1.0


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """N(0,1) by Box-Muller over draw pairs (u1, u2): z0 = r cos(2 pi u2), z1 = r sin(2 pi u2)."""
    m = (n + 1) // 2
    u = uniform01(seed, 2 * m, offset)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    t = 2.0 * math.pi * u2
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(t)
    z[1::2] = r * np.sin(t)
    return z[:n]


def complex_uniform(n: int, seed: int) -> np.ndarray:
    """re, im i.i.d. U[-1,1) (complex128)."""
    v = uniform_pm1(seed, 2 * n)
    return v[0::2] + 1j * v[1::2]


def complex_normal(n: int, seed: int, offset: int = 0) -> np.ndarray:
    """re, im i.i.d. N(0,1) (complex128)."""
    v = normal(seed, 2 * n, offset)
    return v[0::2] + 1j * v[1::2]


def real_sinusoids(n_real: int, seed: int, count: int = 8, noise: float = 0.01):
    """Sum of ``count`` cosines with integer bins f in [1, n/2) and amplitudes U[0.1,1], plus
    N(0, noise^2).  Returns (signal float64, freqs, amplitudes)."""
    u = uniform01(seed, 2 * count)
    freqs = 1 + np.floor(u[0::2] * (n_real // 2 - 1)).astype(np.int64)
    amps = 0.1 + 0.9 * u[1::2]
    t = np.arange(n_real, dtype=np.float64)
    x = np.zeros(n_real)
    for f, a in zip(freqs, amps):
        # exact phase reduction: (f*t mod n) keeps the argument small
        x += a * np.cos(2.0 * math.pi * ((f * t) % n_real) / n_real)
    x += noise * normal(seed, n_real, offset=2 * count)
    return x, freqs, amps


def ragged_lengths(count: int, seed: int, lo: int = 256, hi: int = 65536) -> np.ndarray:
    """Lengths log-uniform over {2^k, 3*2^k, 5*2^k} within [lo, hi]: draw t ~ U[log2 lo,
    log2 hi] and take the member of the set nearest to 2^t in log space."""
    cand = sorted({m << k for m in (1, 3, 5) for k in range(0, 20) if lo <= (m << k) <= hi})
    cand = np.array(cand, dtype=np.int64)
    lc = np.log2(cand.astype(np.float64))
    t = math.log2(lo) + uniform01(seed, count) * (math.log2(hi) - math.log2(lo))
    idx = np.abs(t[:, None] - lc[None, :]).argmin(axis=1)
    return cand[idx]


def sinusoid_batch(batch: int, n: int, seed: int, amp: float = 1e-3) -> np.ndarray:
    """batch x n complex: N(0,1) plus amp * exp(2 pi i f m / n), f seeded per batch row."""
    x = complex_normal(batch * n, seed).reshape(batch, n)
    f = np.floor(uniform01(seed, batch, offset=4 * batch * n) * n).astype(np.int64)
    m = np.arange(n, dtype=np.int64)
    ph = ((f[:, None] * m[None, :]) % n).astype(np.float64) * (2.0 * math.pi / n)
    x += amp * np.exp(1j * ph)
    return x
