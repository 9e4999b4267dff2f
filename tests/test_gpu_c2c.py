"""GPU parity of complex transforms through the C ABI against the CPU oracle.

Bars: FP64 relL2 <= 1e-13, FP32 relL2 <= 2e-7 against the long-double oracle (FP64 result);
output order natural; inverse scaled by 1/N as GFFT's Backward (src/gfftswap.h:355-362).
"""
import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

TOL = {np.complex128: 1e-13, np.complex64: 2e-7}
VT = {np.complex128: api.COMPLEX_DOUBLE, np.complex64: api.COMPLEX_FLOAT}


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def oracle_fwd(x):
    return O.ref_fft(x.astype(np.complex128), 1)


def oracle_inv(x):
    n = x.shape[-1]
    return O.ref_fft(x.astype(np.complex128), -1) / n


POW2 = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", POW2)
def test_single_pass_forward_inverse(torch_cuda, dtype, n):
    x = inputs.complex_uniform(n, 17 + n).astype(dtype)
    ref_f, ref_i = oracle_fwd(x), oracle_inv(x)
    print(f"case n={n} dtype={np.dtype(dtype).name}", flush=True)
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        assert rel_l2(host(d), ref_f) <= TOL[dtype]
    with api.Plan(n, VT[dtype], api.IDFT, api.OUT_OF_PLACE) as p:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        assert rel_l2(host(o), ref_i) <= TOL[dtype]
        assert np.array_equal(host(s), x)  # src untouched


MIXED = [2, 4, 8, 48, 80, 96, 160, 192, 320, 384, 640, 768, 1280, 1536, 2560, 3072, 5120, 6144]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", MIXED)
def test_small_and_mixed_radix(torch_cuda, dtype, n):
    batch = 7
    x = inputs.complex_uniform(n * batch, 300 + n).astype(dtype).reshape(batch, n)
    ref = np.stack([oracle_fwd(r) for r in x])
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE, batch=batch) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= TOL[dtype]
    with api.Plan(n, VT[dtype], api.IDFT, api.OUT_OF_PLACE) as p:
        s = dev(torch_cuda, x[0])
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        assert rel_l2(host(o), oracle_inv(x[0])) <= TOL[dtype]


@pytest.mark.parametrize("n", [10240, 12288, 16384])
def test_large_single_pass_fp32(torch_cuda, n):
    x = inputs.complex_normal(n, n).astype(np.complex64)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        assert rel_l2(host(d), oracle_fwd(x)) <= 2e-7


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [3 * 2 ** 19, 1536 * 256, 5 * 2 ** 16, 15 * 2 ** 12, 9 * 2 ** 14, 25 * 2 ** 12,
                               45 * 2 ** 10, 3 * 5 * 2 ** 18])
def test_mixed_radix_multi_pass(torch_cuda, dtype, n):
    x = inputs.complex_normal(n, 3).astype(dtype)
    with api.Plan(n, VT[dtype], api.DFT, api.OUT_OF_PLACE) as p:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        assert rel_l2(host(o), oracle_fwd(x)) <= TOL[dtype]


SEVEN = [7, 14, 28, 49, 56, 112, 224, 448, 896, 1792, 3584, 343, 1680, 7 ** 3 * 2 ** 10, 7 * 2 ** 20]


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", SEVEN)
def test_radix7_native(torch_cuda, dtype, n):
    """Lengths with the factor 7 run on radix-7 butterflies (single-pass and multi-pass),
    not through the chirp-z path; against the long-double oracle, both directions."""
    import json

    x = inputs.complex_normal(n, 700 + n).astype(dtype)
    print(f"case radix7 n={n} dtype={np.dtype(dtype).name}", flush=True)
    with api.Plan(n, VT[dtype], api.DFT, api.OUT_OF_PLACE) as pf, api.Plan(n, VT[dtype], api.IDFT, api.IN_PLACE) as pi:
        assert "bluestein" not in json.loads(pf.describe())
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        pf.fft(s, o)
        assert rel_l2(host(o), oracle_fwd(x)) <= TOL[dtype]
        pi.fft(s)
        assert rel_l2(host(s), oracle_inv(x)) <= TOL[dtype]


def test_golden_vectors(torch_cuda, golden):
    for n in [2, 4, 8, 16, 32, 64, 128, 256, 768, 1024]:
        x = golden[f"c2c_in_{n}"]
        with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE) as p:
            s = dev(torch_cuda, x)
            o = torch_cuda.empty_like(s)
            p.fft(s, o)
            assert rel_l2(host(o), golden[f"c2c_fwd_{n}"]) <= 1e-15 * np.log2(n) * 4
        with api.Plan(n, api.COMPLEX_DOUBLE, api.IDFT, api.IN_PLACE) as p:
            d = dev(torch_cuda, x)
            p.fft(d)
            assert rel_l2(host(d), golden[f"c2c_inv_{n}"]) <= 1e-15 * np.log2(n) * 4


def test_c1_config_fp64_1024(torch_cuda):
    """N=2^10 FP64 in place forward, U[-1,1) seed 17; against the oracle and the GFFT restatement."""
    n = 1024
    x = inputs.complex_uniform(n, 17)
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    assert rel_l2(y, oracle_fwd(x)) <= 1e-13
    assert rel_l2(y, O.gfft_inplace(x, 1)) <= 1e-13
    # output order: the largest bin of a pure tone lands where GFFT puts it
    tone = np.exp(-2j * np.pi * ((37 * np.arange(n)) % n) / n)
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, tone)
        p.fft(d)
        assert int(np.argmax(np.abs(host(d)))) == n - 37
        assert int(np.argmax(np.abs(O.gfft_inplace(tone, 1)))) == n - 37


def test_seq_kat(torch_cuda):
    n = 4096
    m = np.arange(n)
    x = (2 * m) + 1j * (2 * m + 1)
    k = np.arange(1, n)
    ref = np.empty(n, np.complex128)
    ref[0] = (2 + 2j) * n * (n - 1) / 2 + 1j * n
    sin_half = np.sin(np.pi * np.minimum(k, n - k) / n)
    ref[1:] = (2 + 2j) * n * 1j * np.exp(1j * np.pi * k / n) / (2 * sin_half)
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        assert rel_l2(host(d), ref) <= 1e-14


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n,batch", [(16, 1000), (256, 333), (1024, 64), (4096, 17)])
def test_batched(torch_cuda, dtype, n, batch):
    x = inputs.complex_normal(n * batch, 5).astype(dtype).reshape(batch, n)
    ref = np.stack([oracle_fwd(r) for r in x])
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE, batch=batch) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    for b in range(batch):
        assert rel_l2(y[b], ref[b]) <= TOL[dtype]


def test_host_pointers_are_synchronous(torch_cuda):
    n = 2048
    x = inputs.complex_uniform(n, 3)
    ref = oracle_fwd(x)
    with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE) as p:
        h = x.copy()
        p.fft(h)  # numpy array: host pointer, staged and synchronous
        assert rel_l2(h, ref) <= 1e-13
    with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE) as p:
        out = np.zeros_like(x)
        p.fft(x, out)
        assert rel_l2(out, ref) <= 1e-13


def test_stream_synchronize_waits_for_device_calls(torch_cuda):
    """gfftx_stream_synchronize, which the C++ policy layer calls after every device-pointer
    fft() to keep GFFT's synchronous semantics, returns only when the stream is idle."""
    torch = torch_cuda
    n = 2 ** 24
    x = torch.randn(n, dtype=torch.complex64, device="cuda")
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        p.fft(x, stream=0)
        assert api.lib().gfftx_stream_synchronize(None) == api.SUCCESS
        for _ in range(4):
            p.fft(x, stream=0)  # about 0.5 ms of work on the null stream
        assert api.lib().gfftx_stream_synchronize(None) == api.SUCCESS
        assert torch.cuda.default_stream().query()


@pytest.mark.parametrize("n,vt,tt", [(2 ** 20, api.COMPLEX_FLOAT, api.DFT), (4096, api.COMPLEX_DOUBLE, api.IDFT),
                                     (3 * 2 ** 12, api.FLOAT, api.RDFT), (2 ** 16, api.DOUBLE, api.IRDFT)])
def test_graph_replays_match_eager(torch_cuda, n, vt, tt):
    """GFFTX_FLAG_GRAPH: several replays, with eager work and synchronisation between them,
    each equal to the eager result."""
    dt = np.complex64 if vt in (api.FLOAT, api.COMPLEX_FLOAT) else np.complex128
    x = inputs.complex_normal(n, 31).astype(dt)
    with api.Plan(n, vt, tt, api.OUT_OF_PLACE) as pe, api.Plan(n, vt, tt, api.OUT_OF_PLACE, flags=api.FLAG_GRAPH) as pg:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        pe.fft(s, o)
        ref = host(o)
        og = torch_cuda.empty_like(s)
        for rep in range(4):
            og.zero_()
            pg.fft(s, og)
            torch_cuda.cuda.synchronize()
            _ = float(s.abs().sum())
            assert np.array_equal(host(og), ref), f"replay {rep}"
            assert np.array_equal(host(s), x)


def test_graph_cache_rotating_buffers(torch_cuda):
    """GFFTX_FLAG_GRAPH with more (in, out) pairs than cached graphs: every call, cached,
    evicted or re-captured, equals the eager result for its own buffer."""
    n = 3 * 2 ** 14
    nbuf = 11
    xs = [inputs.complex_normal(n, 100 + i).astype(np.complex64) for i in range(nbuf)]
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as pe, \
            api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, flags=api.FLAG_GRAPH) as pg:
        srcs = [dev(torch_cuda, x) for x in xs]
        refs = []
        for s_ in srcs:
            o = torch_cuda.empty_like(s_)
            pe.fft(s_, o)
            refs.append(host(o))
        outs = [torch_cuda.empty_like(s_) for s_ in srcs]
        order = [0, 1, 2, 0, 1, 2] + list(range(nbuf)) + [10, 0, 5, 10]
        for k, i in enumerate(order):
            outs[i].zero_()
            pg.fft(srcs[i], outs[i])
            torch_cuda.cuda.synchronize()
            assert np.array_equal(host(outs[i]), refs[i]), f"call {k} buffer {i}"
        # the same rotation issued back to back, nothing waited for between the calls: an
        # evicted graph may still be running when its entry is reused
        for o in outs:
            o.zero_()
        for i in order:
            pg.fft(srcs[i], outs[i])
        torch_cuda.cuda.synchronize()
        for i in range(nbuf):
            assert np.array_equal(host(outs[i]), refs[i]), f"back to back, buffer {i}"


def test_autotuner_measures_then_caches(torch_cuda, tmp_path, monkeypatch):
    """A multi-pass plan times several pass splits on the device, records the winner in the
    JSON cache, and the next plan of the same shape and place reuses it (forward and inverse
    share an entry); an out-of-place plan of that length is tuned and cached on its own."""
    import json

    cache = tmp_path / "tune.json"
    monkeypatch.setenv("GFFTX_TUNE_CACHE", str(cache))
    n = 2 ** 20
    x = inputs.complex_normal(n, 8).astype(np.complex64)
    ref = oracle_fwd(x)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        d1 = json.loads(p.describe())
        d = dev(torch_cuda, x)
        p.fft(d)
        assert rel_l2(host(d), ref) <= 2e-7
    assert isinstance(d1["tuned"], list) and len(d1["tuned"]) >= 2
    text = cache.read_text()
    assert '"version": 2' in text and "n=1048576" in text
    with api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE) as p:
        d2 = json.loads(p.describe())
        assert d2["tuned"] == "cache"
        s = dev(torch_cuda, ref.astype(np.complex64))
        p.fft(s)
        assert rel_l2(host(s), x) <= 4e-7
    with api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.OUT_OF_PLACE) as p:
        d3 = json.loads(p.describe())
        assert isinstance(d3["tuned"], list)
        s = dev(torch_cuda, ref.astype(np.complex64))
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        assert rel_l2(host(o), x) <= 4e-7
    text = cache.read_text()
    assert "n=1048576|b=1|ip" in text and "n=1048576|b=1|oop" in text
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, flags=api.FLAG_NO_AUTOTUNE) as p:
        assert "tuned" not in json.loads(p.describe())


@pytest.mark.parametrize("dtype,n,batch,tiles", [
    (np.complex64, 4096, 64, ((1, 256), (1, 512))), (np.complex64, 2048, 40, ((2, 256), (2, 512))),
    (np.complex64, 5120, 24, ((1, 320), (1, 640))), (np.complex64, 6144, 24, ((1, 384), (1, 768))),
    (np.complex64, 8192, 16, ((1, 512), (1, 1024))),
    (np.complex128, 1024, 96, ((4, 256), (4, 512), (1, 64)))])
def test_single_pass_tiles_tuned_and_cached(torch_cuda, tmp_path, monkeypatch, dtype, n, batch, tiles):
    """A batched single-pass plan whose length has several compiled tiles (radix-16 stages,
    radix-8 stages with twice the threads, one sequence per wave) times them and caches the
    choice as [n, 1, 0, nb, wg].  Each tile, forced through its own cache file, matches the
    oracle, and every tile agrees with the first."""
    import json

    x = np.stack([inputs.complex_normal(n, 300 + b) for b in range(batch)]).astype(dtype)
    ref = np.stack([oracle_fwd(r) for r in x])
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("GFFTX_TUNE_CACHE", str(cache))
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE, batch=batch) as p:
        desc = json.loads(p.describe())
    assert isinstance(desc["tuned"], list) and len(desc["tuned"]) >= 2, desc
    assert all(t["split"].startswith("single/") for t in desc["tuned"])
    lines = cache.read_text().splitlines()
    key = [ln for ln in lines if f"n={n}|b={batch}" in ln]
    assert len(key) == 1, lines
    prefix = key[0].split("[")[0]
    outs = {}
    for nb, wg in tiles:
        forced = tmp_path / f"tune_{wg}.json"
        forced.write_text('{"version": 2, "entries": {\n' + prefix + f"[{n}, 1, 0, {nb}, {wg}]\n" + "}}\n")
        monkeypatch.setenv("GFFTX_TUNE_CACHE", str(forced))
        print(f"case single-pass tile n={n} nb={nb} wg={wg}", flush=True)
        with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE, batch=batch) as p:
            desc = json.loads(p.describe())
            assert desc["tuned"] == "cache" and f"_nb{nb}_wg{wg}_" in desc["passes"][0]["kernel"], desc
            d = dev(torch_cuda, x)
            p.fft(d)
            outs[wg] = host(d)
        for b in range(batch):
            assert rel_l2(outs[wg][b], ref[b]) <= TOL[dtype], (wg, b)
    for nb, wg in tiles[1:]:
        assert rel_l2(outs[tiles[0][1]], outs[wg]) <= TOL[dtype], wg


@pytest.mark.parametrize("dtype,n,passes,tnb,twg", [
    (np.complex128, 2 ** 16, "256,256", 16, 512), (np.complex128, 2 ** 24, "256,256,256", 16, 512),
    (np.complex64, 192 * 192, "192,192", 16, 192), (np.complex64, 192 * 192, "192,192", 32, 384),
    (np.complex64, 320 * 320, "320,320", 16, 320), (np.complex64, 80 * 80, "80,80", 32, 160),
    (np.complex64, 80 * 80, "80,80", 48, 240)])
def test_forced_pass_tiles(torch_cuda, monkeypatch, dtype, n, passes, tnb, twg):
    """Pass tiles forced through GFFTX_TNB / GFFTX_TWG (the FP64 256-point tile with radix
    8x8x4 stages and 512 threads; the FP32 192-, 320- and 80-point tiles whose rows are whole
    128-byte lines): forward against the oracle (numpy's FFT above 2^16), and the inverse back
    to the input."""
    monkeypatch.setenv("GFFTX_PASSES", passes)
    monkeypatch.setenv("GFFTX_TNB", str(tnb))
    monkeypatch.setenv("GFFTX_TWG", str(twg))
    x = inputs.complex_normal(n, 500 + n % 997).astype(dtype)
    ref = oracle_fwd(x) if n <= 2 ** 16 else np.fft.fft(x.astype(np.complex128))
    print(f"case forced pass tiles n={n} nb={tnb} wg={twg}", flush=True)
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE) as pf, api.Plan(n, VT[dtype], api.IDFT, api.IN_PLACE) as pi:
        import json

        assert all(f"_nb{tnb}_wg{twg}_" in ps["kernel"] for ps in json.loads(pf.describe())["passes"])
        d = dev(torch_cuda, x)
        pf.fft(d)
        assert rel_l2(host(d), ref) <= TOL[dtype]
        pi.fft(d)
        assert rel_l2(host(d), x) <= 2 * TOL[dtype]


def test_non_temporal_stores_change_no_result(torch_cuda, monkeypatch):
    """The last pass stores a large output non-temporally (GFFTX_NT_STORE_MIB, default 256):
    the store kind must not change a single bit.  Each plan runs with the hint forced on for
    every size (0) and off (-1): batched single-pass tiles, a multi-pass transform (transposed
    last-pass stores), the split-precision MFMA kernel, a real forward and inverse pair and a
    ragged batch."""
    def run(mib):
        monkeypatch.setenv("GFFTX_NT_STORE_MIB", str(mib))
        outs = []
        x = inputs.complex_normal(1024 * 48, 71).astype(np.complex64)
        with api.Plan(1024, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=48) as p:
            d, o = dev(torch_cuda, x), torch_cuda.empty(1024 * 48, dtype=torch_cuda.complex64, device="cuda")
            p.fft(d, o)
            outs.append(host(o))
        x = inputs.complex_normal(2 ** 20, 72)
        with api.Plan(2 ** 20, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE) as p:
            d = dev(torch_cuda, x)
            p.fft(d)
            outs.append(host(d))
        x = inputs.complex_normal(4096 * 32, 73).astype(np.complex64)
        with api.Plan(4096, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=32, flags=api.FLAG_SPLIT_MFMA) as p:
            d, o = dev(torch_cuda, x), torch_cuda.empty(4096 * 32, dtype=torch_cuda.complex64, device="cuda")
            p.fft(d, o)
            outs.append(host(o))
        xr = inputs.complex_normal(2 ** 19, 74).astype(np.complex64)  # 2^20 reals, interleaved
        with api.Plan(2 ** 19, api.FLOAT, api.RDFT, api.IN_PLACE) as pf, \
                api.Plan(2 ** 19, api.FLOAT, api.IRDFT, api.IN_PLACE) as pi:
            d = dev(torch_cuda, xr)
            pf.fft(d)
            outs.append(host(d))
            pi.fft(d)
            outs.append(host(d))
        lengths = [256, 5120, 65536, 384, 640, 12288, 4096]
        x = inputs.complex_normal(sum(lengths), 75).astype(np.complex64)
        rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
        d = dev(torch_cuda, x)
        rp.fft(d)
        outs.append(host(d))
        rp.close()
        return outs

    on, off = run(0), run(-1)
    for k, (a, b) in enumerate(zip(on, off)):
        assert np.array_equal(a, b), f"case {k}: non-temporal stores changed the result"


def test_out_of_place_alias_rejected(torch_cuda):
    with api.Plan(256, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as p:
        d = torch_cuda.zeros(256, dtype=torch_cuda.complex64, device="cuda")
        with pytest.raises(api.GfftxError) as e:
            p.fft(d, d)
        assert e.value.code == api.ERR_ALIAS


def test_unsupported_length_raises(torch_cuda):
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(1, api.COMPLEX_DOUBLE)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
@pytest.mark.parametrize("n", [2 ** 16, 2 ** 18, 2 ** 20])
def test_multi_pass(torch_cuda, dtype, n):
    x = inputs.complex_normal(n, 24).astype(dtype)
    ref = oracle_fwd(x)
    print(f"case multipass n={n} dtype={np.dtype(dtype).name}", flush=True)
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        assert rel_l2(host(d), ref) <= TOL[dtype]
    with api.Plan(n, VT[dtype], api.IDFT, api.OUT_OF_PLACE) as p:
        s = dev(torch_cuda, ref.astype(dtype))
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        assert rel_l2(host(o), x) <= 2 * TOL[dtype]


def test_very_long_fp32(torch_cuda):
    """2^26 FP32 forward against the oracle (four passes or fewer)."""
    n = 2 ** 26
    x = inputs.complex_normal(n, 26).astype(np.complex64)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    assert rel_l2(y, oracle_fwd(x)) <= 2e-7


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_c2_config_2p24_round_trip(torch_cuda, dtype):
    """N=2^24 forward then inverse in place, N(0,1): forward against the FP64 oracle, round
    trip against the input.  Seed 24 (the workload's), and for FP32, whose forward error sits
    near its 2e-7 bar, seeds 25 and 26 too; the worst seed is asserted."""
    n = 2 ** 24
    seeds = (24,) if dtype == np.complex128 else (24, 25, 26)
    worst_fwd = worst_rt = 0.0
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE) as pf, \
            api.Plan(n, VT[dtype], api.IDFT, api.IN_PLACE) as pi:
        for seed in seeds:
            x = inputs.complex_normal(n, seed).astype(dtype)
            ref = oracle_fwd(x)
            print(f"case 2^24 dtype={np.dtype(dtype).name} seed={seed}", flush=True)
            d = dev(torch_cuda, x)
            pf.fft(d)
            y = host(d)
            e_fwd = rel_l2(y, ref)
            pi.fft(d)
            e_rt = rel_l2(host(d), x)
            print(f"2^24 {np.dtype(dtype).name} seed {seed}: fwd relL2 {e_fwd:.3e} round-trip {e_rt:.3e}")
            worst_fwd, worst_rt = max(worst_fwd, e_fwd), max(worst_rt, e_rt)
    print(f"2^24 {np.dtype(dtype).name}: worst fwd relL2 {worst_fwd:.3e}, worst round trip {worst_rt:.3e}")
    assert worst_fwd <= TOL[dtype]
    assert worst_rt <= 2 * TOL[dtype]


@pytest.mark.parametrize("dtype,log2n", [(np.complex64, 28), (np.complex128, 26)])
def test_max_length_properties(torch_cuda, dtype, log2n):
    """The longest native lengths (FP32 2^28, FP64 2^26), where the oracle would take minutes,
    through size-independent properties:
      - a tone exp(2 pi i f j / N) transforms to N at bin f and zero elsewhere (sign, order
        and scale of the forward transform);
      - Parseval: sum |X|^2 = N sum |x|^2 on normal input;
      - the inverse transform returns the input.
    FP32 bar 5e-7: 2^28 adds four radix-2 levels of rounding to the 2^24 case (2e-7)."""
    import math

    torch = torch_cuda
    n = 1 << log2n
    tdt = torch.complex64 if dtype == np.complex64 else torch.complex128
    tol = 5e-7 if dtype == np.complex64 else 1e-13
    print(f"case max length 2^{log2n} {np.dtype(dtype).name}", flush=True)
    f = 123457
    j = torch.arange(n, dtype=torch.int64, device="cuda")
    ph = ((j * f) % n).to(torch.float64) * (2 * math.pi / n)  # f j mod N is exact
    del j
    x = torch.polar(torch.ones_like(ph), ph).to(tdt)
    del ph
    with api.Plan(n, VT[dtype], api.DFT, api.IN_PLACE) as pf, \
            api.Plan(n, VT[dtype], api.IDFT, api.IN_PLACE) as pi:
        pf.fft(x)
        torch.cuda.synchronize()
        peak = complex(x[f].item())
        x[f] -= n
        e_tone = torch.linalg.vector_norm(x.to(torch.complex128)).item() / n
        g = torch.Generator(device="cuda").manual_seed(log2n)
        x = torch.randn(n, dtype=tdt, device="cuda", generator=g)
        x0 = x.clone()
        pf.fft(x)
        e_pars = abs(torch.sum(torch.abs(x.to(torch.complex128)) ** 2).item()
                     / (n * torch.sum(torch.abs(x0.to(torch.complex128)) ** 2).item()) - 1.0)
        pi.fft(x)
        e_rt = (torch.linalg.vector_norm((x - x0).to(torch.complex128)) /
                torch.linalg.vector_norm(x0.to(torch.complex128))).item()
    print(f"2^{log2n}: tone peak {peak} relL2 {e_tone:.3e}; Parseval {e_pars:.3e}; round trip {e_rt:.3e}")
    assert abs(peak - n) <= tol * n
    assert e_tone <= tol
    assert e_pars <= 10 * tol
    assert e_rt <= 2 * tol


@pytest.mark.parametrize("radix,digits", [(2, 10), (2, 16), (3, 7), (5, 5), (7, 4)])
def test_digit_reverse_bit_exact(torch_cuda, radix, digits):
    """The permutation kernel equals GFFTswap2's in-place digit reversal bit for bit."""
    n = radix ** digits
    x = inputs.complex_normal(n, radix * 100 + digits)
    s = dev(torch_cuda, x)
    o = torch_cuda.empty_like(s)
    api.digit_reverse(s, o, radix, digits, api.COMPLEX_DOUBLE)
    assert np.array_equal(host(o), O.swap(x, radix, digits))


def test_cpp_policy_program(torch_cuda):
    """tests/cpp/gfft_check (built by __graft_entry__.build()) drives the C++ GFFT surface
    with the HIP<0> policy: Transform<...>::Instance objects, the GenerateTransform
    factory by run-time IDs, and the unknown-transform exception."""
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "gfft_check")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/gfft_check is not built (run __graft_entry__.build())")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout[-2000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_every_length_to_256(torch_cuda, dtype):
    """GFFT's out-of-place path takes any N (src/gfftalg.h:459-507): every length from 2 to 256
    (products of 2, 3, 5, 7 natively, the others through the chirp-z path), forward and
    inverse, three transforms per call, against the long-double O(N^2) DFT."""
    def smooth(m):
        for p in (2, 3, 5, 7):
            while m % p == 0:
                m //= p
        return m == 1

    worst = (0.0, 0)
    for n in range(2, 257):
        # chirp-z lengths: FP32 bar 5e-7 (two length-m transforms and three chirp products)
        tol = TOL[dtype] if dtype == np.complex128 or smooth(n) else 5e-7
        x = inputs.complex_normal(3 * n, 900 + n).astype(dtype).reshape(3, n)
        ref = np.stack([O.dft_ld(x[b].astype(np.complex128), 1) for b in range(3)])
        with api.Plan(n, VT[dtype], api.DFT, api.OUT_OF_PLACE, batch=3) as pf, \
                api.Plan(n, VT[dtype], api.IDFT, api.OUT_OF_PLACE, batch=3) as pi:
            s = dev(torch_cuda, x)
            o = torch_cuda.empty_like(s)
            pf.fft(s, o)
            y = host(o)
            e = rel_l2(y, ref)
            assert e <= tol, (n, e)
            worst = max(worst, (e, n))
            back = torch_cuda.empty_like(s)
            pi.fft(o, back)
            assert rel_l2(host(back), x) <= 2 * tol, n
    print(f"every length 2..256 {np.dtype(dtype).name}: worst relL2 {worst[0]:.2e} at n={worst[1]}")
