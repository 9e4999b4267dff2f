"""Writes tests/golden/dft_golden.npz: DFT input/output vectors computed with mpmath at
40 significant digits, independent of the oracle and of the engine.

The reference ships no golden vectors (its accuracy test computes a double-double DFT at run
time and prints norms, test/gfft_accuracy.cpp:66-95), so these fixtures are made from the
definition X_k = sum_n x_n exp(-2 pi i S n k / N) (S=+1 forward, unnormalised; S=-1 inverse,
scaled by 1/N as GFFT's Backward does, src/gfftswap.h:355-362), with (n*k mod N) reduced
exactly before the exponential.

Inputs are re,im ~ U[-1,1) from SplitMix64 (seed 17 + N), the C1 input distribution.
The real-transform vectors hold GFFT's RDFT output in NR packing for 2N reals:
[X_0, X_N, Re X_1, Im X_1, ..., Re X_{N-1}, Im X_{N-1}] (src/gfftswap.h:239-287).

Run:  python tests/golden/make_golden.py    (about a minute)
"""
import os
import sys

import mpmath
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.abspath(os.path.join(HERE, "..", "..")))
from gfft_amd.inputs import complex_uniform  # noqa: E402

mpmath.mp.dps = 40

C2C_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 25, 27, 30, 32, 49, 60, 64, 81, 125, 128,
             243, 256, 343, 768, 1024]
REAL_SIZES = [2, 4, 8, 16, 64, 256]  # complex half-lengths N (2N reals)


def mp_dft(x, sign):
    n = len(x)
    w = [mpmath.expjpi(-2 * sign * mpmath.mpf(m) / n) for m in range(n)]
    xs = [mpmath.mpc(float(v.real), float(v.imag)) for v in x]
    out = []
    for k in range(n):
        s = mpmath.mpc(0)
        m = 0
        for j in range(n):
            s += xs[j] * w[m]
            m += k
            if m >= n:
                m -= n
        out.append(s)
    return out


def to_c128(v):
    return np.array([complex(float(z.real), float(z.imag)) for z in v], dtype=np.complex128)


def main():
    arrays = {}
    for n in C2C_SIZES:
        x = complex_uniform(n, 17 + n)
        fwd = mp_dft(x, +1)
        inv = [z / n for z in mp_dft(x, -1)]
        arrays[f"c2c_in_{n}"] = x
        arrays[f"c2c_fwd_{n}"] = to_c128(fwd)
        arrays[f"c2c_inv_{n}"] = to_c128(inv)
        print("c2c", n, flush=True)
    for n in REAL_SIZES:
        v = complex_uniform(n, 1000 + n)  # 2N reals = the interleaved re/im draws
        reals = [mpmath.mpf(float(t)) for t in np.stack([v.real, v.imag], 1).ravel()]
        m = 2 * n
        X = []
        for k in range(n + 1):
            s = mpmath.mpc(0)
            for j in range(m):
                s += reals[j] * mpmath.expjpi(-2 * mpmath.mpf((j * k) % m) / m)
            X.append(s)
        packed = [mpmath.mpc(X[0].real, X[n].real)] + X[1:n]
        arrays[f"r2c_in_{n}"] = v
        arrays[f"r2c_fwd_{n}"] = to_c128(packed)
        print("real", n, flush=True)
    np.savez_compressed(os.path.join(HERE, "dft_golden.npz"), **arrays)


if __name__ == "__main__":
    main()
