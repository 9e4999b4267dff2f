"""GPU parity of real transforms (RDFT/IRDFT) through the C ABI.

GFFT's RDFT maps 2n reals to n complex values in NR packing [X_0, X_n, X_1 .. X_{n-1}]
(src/gfftswap.h:239-287); IRDFT inverts it including the 1/n of Backward.  The bins layout
(n+1 complex bins) is the engine's extra format for the 3*2^20 workload."""
import json

import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

VT = {np.complex128: api.DOUBLE, np.complex64: api.FLOAT}
TOL = {np.complex128: 1e-13, np.complex64: 2e-7}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def packed_ref(z):
    """Exact NR-packed RDFT of the 2n reals interleaved in z (long-double oracle)."""
    n = z.size
    r = np.stack([z.real, z.imag], 1).ravel().astype(np.float64)
    X = O.ref_fft(r.astype(np.complex128), 1)[: n + 1]
    out = X[:n].copy()
    out[0] = X[0].real + 1j * X[n].real
    return out, X


def test_golden_rdft(torch_cuda, golden):
    for n in [2, 4, 8, 16, 64, 256]:
        z = golden[f"r2c_in_{n}"]
        with api.Plan(n, api.DOUBLE, api.RDFT, api.IN_PLACE) as p:
            d = dev(torch_cuda, z)
            p.fft(d)
            assert rel_l2(host(d), golden[f"r2c_fwd_{n}"]) <= 1e-14


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [32, 512, 4096, 3 * 256, 5 * 128, 2 ** 16])
def test_rdft_irdft(torch_cuda, dtype, n):
    z = inputs.complex_normal(n, 40 + n).astype(dtype)
    ref, _ = packed_ref(z)
    print(f"case real n={n} dtype={np.dtype(dtype).name}", flush=True)
    try:
        pf = api.Plan(n, VT[dtype], api.RDFT, api.OUT_OF_PLACE)
    except api.UnsupportedTransform:
        pytest.skip(f"no kernel for n={n}")
    with pf, api.Plan(n, VT[dtype], api.IRDFT, api.OUT_OF_PLACE) as pi:
        s = dev(torch_cuda, z)
        o = torch_cuda.empty_like(s)
        pf.fft(s, o)
        y = host(o)
        assert rel_l2(y, ref) <= TOL[dtype]
        if n <= 4096 and dtype == np.complex128 and (n & (n - 1)) == 0:
            assert rel_l2(y, O.gfft_real_inplace(z, 1)) <= 1e-13
        back = torch_cuda.empty_like(s)
        pi.fft(o, back)
        assert rel_l2(host(back), z) <= 2 * TOL[dtype]
        assert np.array_equal(host(s), z)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,batch", [(64, 1), (4096, 3), (3 * 2 ** 10, 2), (2 ** 18, 1), (3 * 2 ** 19, 1)])
def test_rdft_irdft_in_place(torch_cuda, dtype, n, batch):
    """GFFT's real transforms are in place only (Separate, src/gfftswap.h:239-287; out of
    place exits): RDFT in place against the long-double packed reference for every batch
    item, then IRDFT in place back to the input."""
    z = inputs.complex_normal(n * batch, 60 + n).astype(dtype).reshape(batch, n)
    print(f"case real in place n={n} batch={batch} dtype={np.dtype(dtype).name}", flush=True)
    with api.Plan(n, VT[dtype], api.RDFT, api.IN_PLACE, batch=batch) as pf, \
            api.Plan(n, VT[dtype], api.IRDFT, api.IN_PLACE, batch=batch) as pi:
        d = dev(torch_cuda, z)
        pf.fft(d)
        y = host(d)
        for b in range(batch):
            ref, _ = packed_ref(z[b])
            assert rel_l2(y[b], ref) <= TOL[dtype], b
        pi.fft(d)
        assert rel_l2(host(d), z) <= 2 * TOL[dtype]


@pytest.mark.parametrize("dtype,n,passes,batch,layout", [
    (np.complex64, 2 ** 15, "128,256", 3, "nr"), (np.complex64, 49152, "192,256", 1, "bins"),
    (np.complex64, 46080, "45,1024", 2, "nr"), (np.complex64, 2 ** 18, "1024,256", 1, "bins"),
    (np.complex64, 3 * 2 ** 19, "1536,1024", 1, "bins"), (np.complex64, 20480, "5,4096", 1, "nr"),
    (np.complex128, 2 ** 14, "64,256", 2, "nr"), (np.complex128, 40960, "320,128", 1, "bins"),
    (np.complex128, 46080, "45,1024", 1, "nr"), (np.complex128, 12288, "3,4096", 1, "nr"),
    (np.complex128, 15360, "15,1024", 1, "bins")])
def test_fused_split_last_pass(torch_cuda, monkeypatch, dtype, n, passes, batch, layout):
    """Multi-pass real forward transforms fuse the R2C split into the last pass: its tiles
    pair sequence t with S - t, S = n / (last pass length), odd (3, 5, 15, 45) and even
    here.  Against the long-double oracle, against the separate split kernel, and in
    place."""
    monkeypatch.setenv("GFFTX_PASSES", passes)
    z = inputs.complex_normal(n * batch, 300 + n).astype(dtype).reshape(batch, n)
    flags = api.FLAG_R2C_BINS if layout == "bins" else 0
    w = n + 1 if layout == "bins" else n
    print(f"case fused split n={n} batch={batch} {layout} {np.dtype(dtype).name}", flush=True)
    outs = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", fused)
        with api.Plan(n, VT[dtype], api.RDFT, api.OUT_OF_PLACE, batch=batch, flags=flags) as p:
            assert ("_split" in p.describe()) == (fused == "1")
            s = dev(torch_cuda, z)
            o = torch_cuda.zeros((batch, w), dtype=s.dtype, device="cuda")
            p.fft(s, o)
            outs[fused] = host(o)
    for b in range(batch):
        ref, X = packed_ref(z[b])
        want = X if layout == "bins" else ref
        assert rel_l2(outs["1"][b], want) <= TOL[dtype], b
    assert rel_l2(outs["1"], outs["0"]) <= TOL[dtype]
    if layout == "nr":
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", "1")
        with api.Plan(n, VT[dtype], api.RDFT, api.IN_PLACE, batch=batch) as p:
            d = dev(torch_cuda, z)
            p.fft(d)
            assert rel_l2(host(d), outs["1"]) <= TOL[dtype]


def test_rdft_bins_layout_and_odd_rejected(torch_cuda):
    n = 1024
    z = inputs.complex_normal(n, 9)
    _, X = packed_ref(z)
    with api.Plan(n, api.DOUBLE, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as p:
        s = dev(torch_cuda, z)
        o = torch_cuda.zeros(n + 1, dtype=torch_cuda.complex128, device="cuda")
        p.fft(s, o)
        assert rel_l2(host(o), X) <= 1e-13
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(27, api.DOUBLE, api.RDFT, api.IN_PLACE)


def test_c3_config_r2c_3x2p20(torch_cuda):
    """N = 3*2^20 real FP32: 8 integer-bin cosines + N(0, 0.01^2) noise, seed 3; N/2+1 bins."""
    nr = 3 * 2 ** 20
    x, freqs, amps = inputs.real_sinusoids(nr, 3)
    n = nr // 2
    z = (x[0::2] + 1j * x[1::2]).astype(np.complex64)
    ref = O.ref_fft(np.stack([z.real, z.imag], 1).ravel().astype(np.complex128), 1)[: n + 1]
    with api.Plan(n, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as p:
        s = dev(torch_cuda, z)
        o = torch_cuda.zeros(n + 1, dtype=torch_cuda.complex64, device="cuda")
        p.fft(s, o)
        y = host(o)
    err = rel_l2(y, ref)
    print(f"C3 relL2 {err:.3e}")
    assert err <= 2e-7
    # peaks at the seeded bins with amplitude A*N/2
    for f, a in zip(freqs, amps):
        assert abs(abs(y[f]) - a * nr / 2) < 1e-3 * a * nr / 2 + 50


@pytest.mark.parametrize("tnb", [2, 4])
@pytest.mark.parametrize("passes", ["1536,1024", "1024,1536"])
def test_narrow_pass_tiles(torch_cuda, monkeypatch, tnb, passes):
    """Narrow pass tiles (2 or 4 sequences per workgroup, chosen with GFFTX_TNB) at the
    C3 shape n = 3*2^19: the real forward transform with the split fused into the last
    pass, against the packed reference; the complex forward and inverse against numpy."""
    monkeypatch.setenv("GFFTX_PASSES", passes)
    monkeypatch.setenv("GFFTX_TNB", str(tnb))
    n = 3 * 2 ** 19
    z = inputs.complex_normal(n, 70 + tnb).astype(np.complex64)
    print(f"case narrow tiles nb={tnb} passes={passes}", flush=True)
    with api.Plan(n, api.FLOAT, api.RDFT, api.OUT_OF_PLACE) as p:
        desc = json.loads(p.describe())
        assert all(f"_nb{tnb}_" in ps["kernel"] for ps in desc["passes"]), desc
        assert desc["passes"][-1]["kernel"].endswith("_split"), desc
        s = dev(torch_cuda, z)
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        ref, _ = packed_ref(z)
        assert rel_l2(host(o), ref) <= 2e-7
    x = z.astype(np.complex128)
    for tt, want in ((api.DFT, np.fft.fft(x)), (api.IDFT, np.fft.ifft(x))):
        with api.Plan(n, api.COMPLEX_FLOAT, tt, api.IN_PLACE) as p:
            d = dev(torch_cuda, z)
            p.fft(d)
            assert rel_l2(host(d), want) <= 2e-7


@pytest.mark.parametrize("n,passes,tnb,twg", [
    (1536 * 1536, "1536,1536", 4, 768), (1024 * 1024, "1024,1024", 8, 1024),
    (1024 * 1024, "1024,1024", 4, 512)])
def test_radix8_pass_tiles(torch_cuda, monkeypatch, n, passes, tnb, twg):
    """Pass tiles with eight points per thread (radix-8 stages), forced through GFFTX_TNB and
    GFFTX_TWG: the real forward transform with the fused split, and the complex forward and
    inverse, against numpy."""
    monkeypatch.setenv("GFFTX_PASSES", passes)
    monkeypatch.setenv("GFFTX_TNB", str(tnb))
    monkeypatch.setenv("GFFTX_TWG", str(twg))
    z = inputs.complex_normal(n, 80 + tnb).astype(np.complex64)
    print(f"case radix-8 tiles n={n} nb={tnb} wg={twg}", flush=True)
    with api.Plan(n, api.FLOAT, api.RDFT, api.OUT_OF_PLACE) as p:
        desc = json.loads(p.describe())
        assert all(f"_nb{tnb}_wg{twg}_" in ps["kernel"] for ps in desc["passes"]), desc
        assert desc["passes"][-1]["kernel"].endswith("_split"), desc
        s = dev(torch_cuda, z)
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        ref, _ = packed_ref(z)
        assert rel_l2(host(o), ref) <= 2e-7
    x = z.astype(np.complex128)
    for tt, want in ((api.DFT, np.fft.fft(x)), (api.IDFT, np.fft.ifft(x))):
        with api.Plan(n, api.COMPLEX_FLOAT, tt, api.IN_PLACE) as p:
            d = dev(torch_cuda, z)
            p.fft(d)
            assert rel_l2(host(d), want) <= 2e-7


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_real_every_even_length_to_256(torch_cuda, dtype):
    """RDFT and IRDFT in place for every even complex length n from 2 to 256 (2n reals; n
    with a prime factor above 7 through the chirp-z path), NR packing, against the
    long-double O(N^2) DFT of the 2n reals."""
    def smooth(m):
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        return m == 1

    worst = (0.0, 0)
    for n in range(2, 257, 2):
        tol = TOL[dtype] if dtype == np.complex128 or smooth(n) else 5e-7
        z = inputs.complex_normal(n, 950 + n).astype(dtype)
        r = np.stack([z.real, z.imag], 1).ravel().astype(np.complex128)
        X = O.dft_ld(r, 1)
        ref = X[:n].copy()
        ref[0] = X[0].real + 1j * X[n].real
        with api.Plan(n, VT[dtype], api.RDFT, api.IN_PLACE) as pf, \
                api.Plan(n, VT[dtype], api.IRDFT, api.IN_PLACE) as pi:
            d = dev(torch_cuda, z)
            pf.fft(d)
            e = rel_l2(host(d), ref)
            assert e <= tol, (n, e)
            worst = max(worst, (e, n))
            pi.fft(d)
            assert rel_l2(host(d), z) <= 2 * tol, n
    print(f"real every even length 2..256 {np.dtype(dtype).name}: worst relL2 {worst[0]:.2e} at n={worst[1]}")


@pytest.mark.parametrize("dtype,n,passes,batch,layout", [
    (np.complex64, 2 ** 15, "128,256", 3, "nr"), (np.complex64, 49152, "256,192", 1, "bins"),
    (np.complex64, 46080, "1024,45", 2, "nr"), (np.complex64, 2 ** 18, "256,1024", 1, "bins"),
    (np.complex64, 3 * 2 ** 19, "1536,1024", 1, "bins"), (np.complex64, 3 * 2 ** 19, "1536,1024", 2, "nr"),
    (np.complex64, 20480, "4096,5", 1, "nr"), (np.complex64, 2 ** 20, "1024,1024", 1, "nr"),
    (np.complex128, 2 ** 14, "64,256", 2, "nr"), (np.complex128, 40960, "320,128", 1, "bins"),
    (np.complex128, 46080, "1024,45", 1, "nr"), (np.complex128, 12288, "4096,3", 1, "nr"),
    (np.complex128, 2 ** 20, "256,16,256", 1, "bins")])
def test_fused_presplit_first_pass(torch_cuda, monkeypatch, dtype, n, passes, batch, layout):
    """Multi-pass real inverse transforms fuse the C2R pre-pass into the first pass: its tiles
    pair sequence t with S - t, S = n / (first pass length), even and odd (5, 45, 3).  The
    result must return the input of the forward transform (the spectrum comes from the
    long-double oracle) and agree with the separate split kernel; out of place the input is
    left unchanged, and in place gives the same result."""
    monkeypatch.setenv("GFFTX_PASSES", passes)
    z = inputs.complex_normal(n * batch, 500 + n).astype(dtype).reshape(batch, n)
    flags = api.FLAG_R2C_BINS if layout == "bins" else 0
    spec = []
    for b in range(batch):
        ref, X = packed_ref(z[b])
        spec.append(X if layout == "bins" else ref)
    spec = np.stack(spec).astype(dtype)
    print(f"case fused presplit n={n} batch={batch} {layout} {np.dtype(dtype).name}", flush=True)
    outs = {}
    for fused in ("1", "0"):
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", fused)
        with api.Plan(n, VT[dtype], api.IRDFT, api.OUT_OF_PLACE, batch=batch, flags=flags) as p:
            assert ("_presplit" in p.describe()) == (fused == "1")
            s = dev(torch_cuda, spec)
            o = torch_cuda.zeros((batch, n), dtype=s.dtype, device="cuda")
            p.fft(s, o)
            outs[fused] = host(o)
            assert np.array_equal(host(s), spec)
    for b in range(batch):
        assert rel_l2(outs["1"][b], z[b]) <= 2 * TOL[dtype], b
    assert rel_l2(outs["1"], outs["0"]) <= 2 * TOL[dtype]
    if layout == "nr":
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", "1")
        with api.Plan(n, VT[dtype], api.IRDFT, api.IN_PLACE, batch=batch) as p:
            d = dev(torch_cuda, spec)
            p.fft(d)
            assert rel_l2(host(d), outs["1"]) <= TOL[dtype]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,batch,layout", [(2, 5, "nr"), (16, 3, "nr"), (64, 1, "bins"), (256, 7, "nr"),
                                            (96, 4, "nr"), (1280, 2, "nr"), (2048, 1, "bins"), (4096, 3, "nr"),
                                            (6144, 1, "bins"), (8192, 2, "nr")])
def test_fused_split_single_pass(torch_cuda, monkeypatch, dtype, n, batch, layout):
    """Single-pass real transforms run the R2C post-pass and the C2R pre-pass inside the
    batched tile (one kernel per call): forward against the long-double packed reference,
    inverse back to the input, both against the separate split kernel, in and out of place."""
    z = inputs.complex_normal(n * batch, 700 + n).astype(dtype).reshape(batch, n)
    flags = api.FLAG_R2C_BINS if layout == "bins" else 0
    w = n + 1 if layout == "bins" else n
    refs = [packed_ref(z[b]) for b in range(batch)]
    want = np.stack([X if layout == "bins" else ref for ref, X in refs])
    print(f"case fused single-pass n={n} batch={batch} {layout} {np.dtype(dtype).name}", flush=True)
    fwd, inv = {}, {}
    for fused in ("1", "0"):
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", fused)
        with api.Plan(n, VT[dtype], api.RDFT, api.OUT_OF_PLACE, batch=batch, flags=flags) as pf, \
                api.Plan(n, VT[dtype], api.IRDFT, api.OUT_OF_PLACE, batch=batch, flags=flags) as pi:
            passes_f = json.loads(pf.describe())["passes"]
            passes_i = json.loads(pi.describe())["passes"]
            assert len(passes_f) == 1 and len(passes_i) == 1
            assert ("_e_split" in passes_f[0]["kernel"]) == (fused == "1")
            assert ("_e_presplit" in passes_i[0]["kernel"]) == (fused == "1")
            s = dev(torch_cuda, z)
            o = torch_cuda.zeros((batch, w), dtype=s.dtype, device="cuda")
            pf.fft(s, o)
            fwd[fused] = host(o)
            sp = dev(torch_cuda, want.astype(dtype))
            b = torch_cuda.zeros((batch, n), dtype=s.dtype, device="cuda")
            pi.fft(sp, b)
            inv[fused] = host(b)
            assert np.array_equal(host(sp), want.astype(dtype))
    for k in range(batch):
        assert rel_l2(fwd["1"][k], want[k]) <= TOL[dtype], k
        assert rel_l2(inv["1"][k], z[k]) <= 2 * TOL[dtype], k
    assert rel_l2(fwd["1"], fwd["0"]) <= TOL[dtype]
    assert rel_l2(inv["1"], inv["0"]) <= 2 * TOL[dtype]
    if layout == "nr":
        monkeypatch.setenv("GFFTX_FUSED_SPLIT", "1")
        with api.Plan(n, VT[dtype], api.RDFT, api.IN_PLACE, batch=batch) as pf, \
                api.Plan(n, VT[dtype], api.IRDFT, api.IN_PLACE, batch=batch) as pi:
            d = dev(torch_cuda, z)
            pf.fft(d)
            assert rel_l2(host(d), fwd["1"]) <= TOL[dtype]
            pi.fft(d)
            assert rel_l2(host(d), z) <= 2 * TOL[dtype]
