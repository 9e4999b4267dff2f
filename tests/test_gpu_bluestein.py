"""GPU parity of lengths with a prime factor above 5 (chirp-z path) through the C ABI.

GFFT's out-of-place path takes any N, with an O(p^2) leaf for each odd prime p
(src/gfftspec.h:42-103, src/gfftalg.h:499-507).  The engine runs lengths whose factors its
tiles carry (2, 3, 5, 7) natively and computes every other length as a circular
convolution of 2/3/5-smooth length m >= 2N-1.  The references below:
  - the 40-digit golden DFTs for 7, 49 and 343;
  - the oracle's restatement of GFFT's out-of-place path (prime factors below 64) and its
    long-double Stockham FFT;
  - the long-double O(N^2) DFT for primes above 63 up to a few thousand;
  - numpy's FFT (complex128) beyond that.

Bars: FP64 relL2 <= 1e-13; FP32 relL2 <= 5e-7 (the convolution adds two length-m
transforms and three chirp products to the rounding of a direct FFT).
"""
import json

import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-13, np.complex64: 5e-7}
VT = {np.complex128: api.COMPLEX_DOUBLE, np.complex64: api.COMPLEX_FLOAT}
RVT = {np.complex128: api.DOUBLE, np.complex64: api.FLOAT}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def largest_prime_factor(n):
    p, big = 2, 1
    while p * p <= n:
        while n % p == 0:
            n //= p
            big = p
        p += 1
    return max(big, n)


def reference(x, sign):
    """Unnormalised DFT (sign +1: exp(-), -1: exp(+)) of the rows of x in complex128."""
    x = np.atleast_2d(x.astype(np.complex128))
    n = x.shape[-1]
    if largest_prime_factor(n) < 64:
        return np.stack([O.ref_fft(r, sign) for r in x])
    if n <= 4100:
        return np.stack([O.dft_ld(r, sign) for r in x])
    return np.fft.fft(x, axis=-1) if sign == 1 else np.fft.ifft(x, axis=-1) * n


def test_golden_vectors_prime_powers_of_7(torch_cuda, golden):
    """7, 49 and 343 run natively (radix-7 butterflies); the chirp-z path is pinned against
    them through 7 * 11, 11 * 13 and the primes below."""
    for n in (7, 49, 343):
        x = golden[f"c2c_in_{n}"]
        with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE) as p:
            assert "bluestein" not in json.loads(p.describe())
            s = dev(torch_cuda, x)
            o = torch_cuda.empty_like(s)
            p.fft(s, o)
            assert rel_l2(host(o), golden[f"c2c_fwd_{n}"]) <= 1e-14
        with api.Plan(n, api.COMPLEX_DOUBLE, api.IDFT, api.IN_PLACE) as p:
            d = dev(torch_cuda, x)
            p.fft(d)
            assert rel_l2(host(d), golden[f"c2c_inv_{n}"]) <= 1e-14


SIZES = [11, 13, 19, 22, 17, 31, 77, 97, 11 * 64, 1001, 1009, 2310, 4099, 13 * 2 ** 12, 65537, 11 * 3 ** 7]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", SIZES)
def test_forward_inverse(torch_cuda, n, dtype):
    x = inputs.complex_normal(n, 7000 + n).astype(dtype)
    fwd = reference(x, 1)[0]
    inv = reference(x, -1)[0] / n
    print(f"case chirp-z n={n} dtype={np.dtype(dtype).name}", flush=True)
    with api.Plan(n, VT[dtype], api.DFT, api.OUT_OF_PLACE) as pf, \
            api.Plan(n, VT[dtype], api.IDFT, api.IN_PLACE) as pi:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        pf.fft(s, o)
        assert rel_l2(host(o), fwd) <= TOL[dtype]
        assert np.array_equal(host(s), x)
        pi.fft(s)
        assert rel_l2(host(s), inv) <= TOL[dtype]


def test_parity_with_gfft_out_of_place_restatement(torch_cuda):
    """Against the oracle's restatement of GFFT's own out-of-place path (O(p^2) leaves)."""
    for n in (7 * 11 * 13, 2 * 3 * 5 * 7 * 11, 61 * 4):
        x = inputs.complex_uniform(n, n)
        with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE) as p:
            s = dev(torch_cuda, x)
            o = torch_cuda.empty_like(s)
            p.fft(s, o)
            assert rel_l2(host(o), O.gfft_oop(x, 1)) <= 1e-13


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_batched(torch_cuda, dtype):
    n, batch = 1009, 6
    x = np.stack([inputs.complex_normal(n, 50 + b) for b in range(batch)]).astype(dtype)
    ref = reference(x, 1)
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE, batch=batch) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= TOL[dtype], f"item {b}"


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [22, 2 * 1009, 11 * 2 ** 10])
def test_real_pair(torch_cuda, n, dtype):
    """RDFT/IRDFT of 2n reals whose complex length n has a factor above 5: NR packing
    [X_0, X_n, X_1 .. X_{n-1}] of the length-2n real DFT."""
    z = inputs.complex_normal(n, 90 + n).astype(dtype)
    xr = z.view(np.float64 if dtype == np.complex128 else np.float32).astype(np.float64)
    X = reference(xr, 1)[0]
    ref = np.concatenate([[X[0].real + 1j * X[n].real], X[1:n]])
    with api.Plan(n, RVT[dtype], api.RDFT, api.OUT_OF_PLACE) as pf, \
            api.Plan(n, RVT[dtype], api.IRDFT, api.OUT_OF_PLACE) as pi:
        s = dev(torch_cuda, z)
        o = torch_cuda.empty_like(s)
        pf.fft(s, o)
        assert rel_l2(host(o), ref) <= TOL[dtype]
        back = torch_cuda.empty_like(s)
        pi.fft(o, back)
        assert rel_l2(host(back), z) <= 2 * TOL[dtype]


def test_host_pointers_and_graph_replays(torch_cuda):
    n = 1009
    x = inputs.complex_normal(n, 5).astype(np.complex64)
    ref = reference(x, 1)[0]
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as p:
        out = np.zeros_like(x)
        p.fft(x, out)  # host arrays: staged and synchronous
        assert rel_l2(out, ref) <= TOL[np.complex64]
        eager = out
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, flags=api.FLAG_GRAPH) as pg:
        s = dev(torch_cuda, x)
        og = torch_cuda.empty_like(s)
        for rep in range(4):
            og.zero_()
            pg.fft(s, og)
            torch_cuda.cuda.synchronize()
            _ = float(s.abs().sum())
            assert np.array_equal(host(og), eager), f"replay {rep}"


def test_profile_slots(torch_cuda):
    n = 4099
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, flags=api.FLAG_PROFILE) as p:
        d = dev(torch_cuda, inputs.complex_normal(n, 1).astype(np.complex64))
        for _ in range(3):
            p.fft(d)
        prof = p.profile()
    names = [name for name, _, _ in prof]
    assert names[0] == "bluestein_pre" and names[-1] == "bluestein_post" and "bluestein_mul" in names
    assert all(cnt == 3 for _, _, cnt in prof)


def test_large_prime_round_trip(torch_cuda):
    """A prime length near 2^22 in FP32: against numpy and through a round trip."""
    n = 4194301  # prime
    x = inputs.complex_normal(n, 77).astype(np.complex64)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as pf, \
            api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE) as pi:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        pf.fft(s, o)
        y = host(o)
        assert rel_l2(y, np.fft.fft(x.astype(np.complex128))) <= 1e-6
        pi.fft(o)
        assert rel_l2(host(o), x) <= 1e-6


def test_limits(torch_cuda):
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(33554467, api.COMPLEX_DOUBLE)  # prime above 2^25: convolution length above 2^26
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(1009, api.COMPLEX_FLOAT, flags=api.FLAG_SPLIT_MFMA)
