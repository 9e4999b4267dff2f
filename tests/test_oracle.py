"""Pins the CPU oracle (a restatement of GFFT's algorithm plus long-double references)
against mpmath golden vectors, closed forms and invariants.  CPU only."""
import math

import numpy as np
import pytest

from conftest import rel_l2
from oracle import oracle as O
from gfft_amd import inputs

C2C_SIZES = [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15, 16, 25, 27, 30, 32, 49, 60, 64, 81, 125, 128,
             243, 256, 343, 768, 1024]


def is_prime_power(n):
    f = O.factorize(n) if n > 1 else []
    return n > 1 and len(set(f)) == 1


@pytest.mark.parametrize("n", C2C_SIZES)
def test_reference_fft_matches_golden(golden, n):
    x = golden[f"c2c_in_{n}"]
    assert rel_l2(O.ref_fft(x, 1), golden[f"c2c_fwd_{n}"]) < 2e-16 * max(1, math.log2(max(n, 2)))
    assert rel_l2(O.ref_fft(x, -1) / n, golden[f"c2c_inv_{n}"]) < 2e-16 * max(1, math.log2(max(n, 2)))
    if n <= 768:
        assert rel_l2(O.dft_ld(x, 1), golden[f"c2c_fwd_{n}"]) < 1e-16 * max(1, math.log2(max(n, 2)))


@pytest.mark.parametrize("n", C2C_SIZES)
def test_gfft_oop_restatement_matches_golden(golden, n):
    """InTimeOOP + Direction (src/gfftpolicy.h:193-201) in DOUBLE and FLOAT."""
    x = golden[f"c2c_in_{n}"]
    assert rel_l2(O.gfft_oop(x, 1), golden[f"c2c_fwd_{n}"]) < 1e-14
    assert rel_l2(O.gfft_oop(x, -1), golden[f"c2c_inv_{n}"]) < 1e-14
    # FLOAT: the radix-2 sweep's recurrence runs in float (src/gfftalg.h:249-261)
    assert rel_l2(O.gfft_oop(x.astype(np.complex64), 1), golden[f"c2c_fwd_{n}"]) < 5e-6


@pytest.mark.parametrize("n", [n for n in C2C_SIZES if is_prime_power(n)])
def test_gfft_inplace_restatement_matches_golden(golden, n):
    """Swap + InTime + Direction (src/gfftpolicy.h:151-161)."""
    x = golden[f"c2c_in_{n}"]
    assert rel_l2(O.gfft_inplace(x, 1), golden[f"c2c_fwd_{n}"]) < 1e-14
    assert rel_l2(O.gfft_inplace(x, -1), golden[f"c2c_inv_{n}"]) < 1e-14


def test_inplace_rejects_multi_factor_lengths():
    """GFFT's multi-factor in-place InTime is a private no-op (src/gfftalg.h:386-395)."""
    with pytest.raises(ValueError):
        O.gfft_inplace(np.zeros(12, np.complex128), 1)


@pytest.mark.parametrize("n", [2, 4, 8, 16, 64, 256])
def test_rdft_restatement_matches_golden(golden, n):
    """RDFT = [Swap, InTime, Forward, Separate]: NR packing (src/gfftswap.h:239-287)."""
    z = golden[f"r2c_in_{n}"]
    ref = golden[f"r2c_fwd_{n}"]
    assert rel_l2(O.gfft_real_inplace(z, 1), ref) < 1e-14
    assert rel_l2(O.gfft_real_oop(z, 1), ref) < 1e-14
    # IRDFT inverts RDFT exactly up to rounding (Backward divides by N)
    assert rel_l2(O.gfft_real_inplace(O.gfft_real_inplace(z, 1), -1), z) < 1e-14


def test_rdft_odd_length_is_a_reference_quirk():
    """Separate's loop runs i < N/2 and negates scalar N+1 (src/gfftswap.h:255, :279): for an
    odd complex length the result is not the DFT.  The engine rejects odd lengths."""
    z = inputs.complex_uniform(27, 3)
    r = np.stack([z.real, z.imag], 1).ravel()
    R = np.fft.rfft(r)
    packed = R[:27].copy()
    packed[0] = R[0].real + 1j * R[27].real
    assert rel_l2(O.gfft_real_inplace(z, 1), packed) > 1e-3


def test_seq_closed_form():
    """GenInput::seq x[n] = 2n + i(2n+1) (src/direct.h:52-56): X_0 = (2+2i)N(N-1)/2 + iN,
    X_k = (2+2i) N / (w^k - 1), w = exp(-2 pi i/N).  Evaluated without cancellation through
    w^k - 1 = -2i sin(pi k/N) exp(-i pi k/N)."""
    for n in [8, 27, 64, 125, 1024, 3 * 256]:
        m = np.arange(n)
        x = (2 * m) + 1j * (2 * m + 1)
        k = np.arange(1, n)
        half = np.pi * k / n
        ref = np.empty(n, np.complex128)
        ref[0] = (2 + 2j) * n * (n - 1) / 2 + 1j * n
        sin_half = np.sin(np.pi * np.minimum(k, n - k) / n)  # sin(pi k/N) = sin(pi (N-k)/N)
        ref[1:] = (2 + 2j) * n * 1j * np.exp(1j * half) / (2 * sin_half)
        assert rel_l2(O.ref_fft(x), ref) < 1e-14
        assert rel_l2(O.gfft_oop(x), ref) < 1e-13


@pytest.mark.parametrize("n", [16, 81, 625, 1024, 3 * 2 ** 7])
def test_closed_forms(n):
    # delta at 0 -> all ones; constant -> N delta; exp(2 pi i f m/N) -> N delta_{k-f}
    d = np.zeros(n, np.complex128)
    d[0] = 1
    assert np.allclose(O.gfft_oop(d), np.ones(n), atol=1e-15)
    c = np.ones(n, np.complex128)
    e = np.zeros(n, np.complex128)
    e[0] = n
    assert np.allclose(O.gfft_oop(c), e, atol=1e-12 * n)
    f = 5 % n
    s = np.exp(2j * np.pi * ((f * np.arange(n)) % n) / n)
    e = np.zeros(n, np.complex128)
    e[f] = n
    assert np.allclose(O.ref_fft(s), e, atol=1e-13 * n)
    # shift theorem: x[m-1] -> X_k w^k
    x = inputs.complex_normal(n, 7)
    X = O.ref_fft(x)
    Xs = O.ref_fft(np.roll(x, 1))
    assert rel_l2(Xs, X * np.exp(-2j * np.pi * np.arange(n) / n)) < 1e-14


def test_sign_convention_detects_flipped_exponent(golden):
    """A planted mistake (the conjugate kernel) must fail the golden comparison."""
    x = golden["c2c_in_64"]
    wrong = np.conj(O.ref_fft(np.conj(x)))  # exp(+) instead of exp(-)
    assert rel_l2(wrong, golden["c2c_fwd_64"]) > 0.5
    wrong_scale = O.gfft_oop(x, -1) * 64  # missing 1/N
    assert rel_l2(wrong_scale, golden["c2c_inv_64"]) > 10


@pytest.mark.parametrize("radix,digits", [(2, 1), (2, 2), (2, 10), (3, 5), (5, 3), (7, 3)])
def test_swap_is_digit_reversal(radix, digits):
    """GFFTswap2<M,P> (src/gfftswap.h:74-140) applies the base-M digit reversal."""
    n = radix ** digits
    x = np.arange(n) + 1j * np.arange(n, 2 * n)
    y = O.swap(x, radix, digits)
    rev = O.digit_reverse_map(radix, digits)
    expect = np.empty_like(x)
    expect[rev] = x
    assert np.array_equal(y, expect)
    # brute force on the digits themselves
    for k in range(min(n, 50)):
        ds = [(k // radix ** i) % radix for i in range(digits)]
        r = sum(d * radix ** (digits - 1 - i) for i, d in enumerate(ds))
        assert rev[k] == r


@pytest.mark.parametrize("n,nt", [(256, 2), (1024, 4), (4096, 8), (2 ** 16, 8)])
def test_openmp_split_matches_serial(n, nt):
    """InTime_omp / InTimeOOP_omp (src/gfftomp.h:105-144, :493-538) compute the same DFT."""
    x = inputs.complex_normal(n, 11)
    ref = O.ref_fft(x)
    tol = 1e-16 * n ** 0.5 * 40
    assert rel_l2(O.gfft_inplace(x, 1, nt), ref) < tol
    assert rel_l2(O.gfft_oop(x, 1, nt), ref) < tol
    assert rel_l2(O.gfft_oop(x, -1, nt), np.conj(O.ref_fft(np.conj(x))) / n) < tol


def test_mixed_radix_openmp():
    x = inputs.complex_normal(3 * 2 ** 10, 12)
    assert rel_l2(O.gfft_oop(x, 1, 4), O.ref_fft(x)) < 1e-13


def test_invariants_large():
    n = 2 ** 16
    x = inputs.complex_normal(n, 24)
    X = O.ref_fft(x)
    # Parseval
    assert abs(np.vdot(X, X).real / n - np.vdot(x, x).real) / np.vdot(x, x).real < 1e-14
    # inverse o forward
    assert rel_l2(O.ref_fft(X, -1) / n, x) < 1e-15
    # linearity
    y = inputs.complex_normal(n, 25)
    assert rel_l2(O.ref_fft(2 * x + 3j * y), 2 * X + 3j * O.ref_fft(y)) < 1e-15
    # GFFT's DOUBLE recurrence drifts at large N but stays far below FP32 error
    assert rel_l2(O.gfft_inplace(x), X) < 1e-11


def test_root_octant_reduction():
    import mpmath

    mpmath.mp.dps = 30
    for n in [8, 12, 1000, 2 ** 24, 3 * 2 ** 20]:
        for m in [0, 1, n // 8, n // 4, n // 2, 3 * n // 4, n - 1, 12345 % n, 7 * n // 8 + 3]:
            w = O.root(m, n)
            ref = complex(mpmath.expjpi(mpmath.mpf(-2 * m) / n))
            assert abs(w.real - ref.real) <= 1.2e-16 and abs(w.imag - ref.imag) <= 1.2e-16
    assert O.root(2, 8) == complex(0.0, -1.0)
    assert O.root(4, 8) == complex(-1.0, 0.0)


def test_input_generator_is_deterministic():
    a = inputs.complex_uniform(1024, 17)
    b = inputs.complex_uniform(1024, 17)
    assert np.array_equal(a, b)
    assert a.real.min() >= -1 and a.real.max() < 1
    z = inputs.normal(24, 200000)
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
    # SplitMix64 known value: first output for seed 0 is 0xE220A8397B1DCDAF
    assert int(inputs.splitmix64(0, 1)[0]) == 0xE220A8397B1DCDAF
