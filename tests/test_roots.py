"""The engine's roots of unity (gfft_amd/csrc/roots.h: every stage, pass, chirp and split table
is built from them) against 40-digit mpmath values.  CPU only.

W_N^m = exp(-2 pi i m / N) must be within 2^-62 of the exact value in long double, and
rounded to double it must be the correctly rounded double of the exact value (a value that
sits within 2^-62 of a rounding boundary is skipped: either neighbour is then correct)."""
import os
import random
import subprocess

import mpmath
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "roots_check.cpp")


@pytest.fixture(scope="module")
def roots_bin(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("roots") / "roots_check")
    subprocess.run(["g++", "-O2", "-std=c++17", SRC, "-o", exe], check=True)
    return exe


def engine_roots(exe, pairs):
    text = "".join(f"{m} {n}\n" for m, n in pairs)
    out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
    res = []
    with mpmath.workdps(40):
        for line in out.split("\n"):
            if line.strip():
                _, _, c, s = line.split()
                res.append((mpmath.mpf(c), mpmath.mpf(s)))
    return res


def sample_pairs():
    rnd = random.Random(7)
    ns = [2 ** k for k in range(1, 30)] + [3 * 2 ** k for k in range(1, 27)] + [5 * 2 ** k for k in range(1, 26)]
    ns += [7 * 2 ** 20, 2 * 1048573, 2 * 1009, 2 * 4194301, 45 * 2 ** 10, 2 * 3 * 2 ** 20]
    pairs = []
    for n in ns:
        for m in {0, 1, n // 8, n // 4, n // 2, n - 1, (3 * n) // 8, (5 * n) // 8, (7 * n) // 8, n // 3}:
            pairs.append((m, n))
        for _ in range(6):
            pairs.append((rnd.randrange(n), n))
    # arguments outside [0, N) are reduced first
    pairs += [(-1, 4096), (4096 + 5, 4096), (-(2 ** 24) - 3, 2 ** 24)]
    return pairs


def exact(m, n):
    with mpmath.workdps(40):
        a = -2 * mpmath.pi * mpmath.mpf(m) / n
        return mpmath.cos(a), mpmath.sin(a)


@mpmath.workdps(40)
def test_roots_match_40_digit_values(roots_bin):
    pairs = sample_pairs()
    got = engine_roots(roots_bin, pairs)
    assert len(got) == len(pairs)
    bound = mpmath.mpf(2) ** -62
    worst = mpmath.mpf(0)
    mismatches = []
    for (m, n), (c, s) in zip(pairs, got):
        ec, es = exact(m, n)
        worst = max(worst, abs(c - ec), abs(s - es))
        assert abs(c - ec) <= bound and abs(s - es) <= bound, (m, n, c, ec, s, es)
        for v, e in ((c, ec), (s, es)):
            dv, de = float(v), float(e)
            if dv != de:
                # allowed only next to a rounding boundary of doubles
                mid = (mpmath.mpf(dv) + mpmath.mpf(de)) / 2
                if abs(e - mid) > bound:
                    mismatches.append((m, n, v, e))
    print(f"{len(pairs)} roots, worst long-double error {mpmath.nstr(worst, 3)}")
    assert not mismatches, mismatches[:5]


def test_planted_sign_error_is_caught():
    """A root with the opposite sign convention (exp(+2 pi i m/N)) misses the exact value."""
    ec, es = exact(1, 16)
    assert abs(-es - es) > mpmath.mpf(2) ** -62
