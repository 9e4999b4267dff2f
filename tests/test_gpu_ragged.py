"""GPU parity of the ragged batch (many lengths in one call) through the C ABI.

Every transform is compared with the long-double oracle at relL2 <= 2e-7 (FP32).  Lengths
above 4096 run as batched launches per length class (GFFTX_RAGGED_MODE=0, the default) or
inside the persistent kernel, fused (1) or split into two launches (2); all three modes
are checked."""
import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu

C4_LENGTHS = sorted({m << k for m in (1, 3, 5) for k in range(0, 17) if 256 <= (m << k) <= 65536})


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


def check_all(y, x, lengths, inverse=False, tol=2e-7):
    off = 0
    worst = 0.0
    for L in lengths:
        xs = x[off:off + L].astype(np.complex128)
        ref = O.ref_fft(xs, -1) / L if inverse else O.ref_fft(xs, 1)
        e = rel_l2(y[off:off + L], ref)
        worst = max(worst, e)
        assert e <= tol, (L, e)
        off += L
    return worst


MODES = ["0", "1", "2"]


@pytest.mark.parametrize("mode", MODES)
def test_all_lengths_once(torch_cuda, monkeypatch, mode):
    monkeypatch.setenv("GFFTX_RAGGED_MODE", mode)
    lengths = [16, 32, 48, 64, 80, 96, 128, 160, 192] + C4_LENGTHS + C4_LENGTHS[::-1]
    assert len(C4_LENGTHS) == 25
    x = inputs.complex_normal(sum(lengths), 77).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    d = dev(torch_cuda, x)
    p.fft(d)
    worst = check_all(host(d), x, lengths)
    print("ragged all lengths worst relL2", worst, p.describe()[:200])
    p.close()


def test_all_lengths_as_class_launches(torch_cuda, monkeypatch):
    """GFFTX_RAGGED_PERSISTENT_MAX=0: no persistent kernel, every length (16 to 65536) a
    batched class launch on the side streams; forward and inverse against the oracle."""
    monkeypatch.setenv("GFFTX_RAGGED_PERSISTENT_MAX", "0")
    lengths = [16, 48, 80, 192, 256] + C4_LENGTHS + [4096, 320, 1536]
    x = inputs.complex_normal(sum(lengths), 78).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    assert '"items":[0,0]' in p.describe().replace(" ", "")
    d = dev(torch_cuda, x)
    p.fft(d)
    y = host(d)
    check_all(y, x, lengths)
    p.close()
    q = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.IDFT)
    q.fft(d)
    check_all(host(d), y.astype(np.complex128), lengths, inverse=True)
    q.close()


def test_persistent_kernel_forced(torch_cuda, monkeypatch):
    """GFFTX_RAGGED_PERSISTENT_MAX=4096: the lengths up to 4096 run in the persistent kernel
    whatever the autotuner would choose; forward against the oracle."""
    import json

    monkeypatch.setenv("GFFTX_RAGGED_PERSISTENT_MAX", "4096")
    lengths = [16, 48, 80, 192, 256] + C4_LENGTHS + [4096, 320, 1536]
    x = inputs.complex_normal(sum(lengths), 79).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    desc = json.loads(p.describe())
    assert desc["items"][0] > 0 and desc["persistent_max"] == 4096 and desc["persistent_choice"] == "environment"
    d = dev(torch_cuda, x)
    p.fft(d)
    check_all(host(d), x, lengths)
    p.close()


def test_persistent_choice_measured_and_cached(torch_cuda, tmp_path, monkeypatch):
    """Without GFFTX_RAGGED_PERSISTENT_MAX the autotuner times the batch with the persistent
    kernel and with class launches for every length, keeps the faster and caches the limit
    under the batch's classes; the next plan of the same classes reads it.  Both plans match
    the oracle."""
    import json

    monkeypatch.delenv("GFFTX_RAGGED_PERSISTENT_MAX", raising=False)
    cache = tmp_path / "tune.json"
    monkeypatch.setenv("GFFTX_TUNE_CACHE", str(cache))
    lengths = [int(v) for v in inputs.ragged_lengths(600, 31)]
    x = inputs.complex_normal(sum(lengths), 80).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    desc = json.loads(p.describe())
    choice = desc["persistent_choice"]
    print("ragged persistent choice", desc["persistent_max"], choice)
    assert isinstance(choice, dict) and choice["persistent_ms"] > 0 and choice["class_launches_ms"] > 0
    faster_classes = choice["class_launches_ms"] < choice["persistent_ms"]
    assert desc["persistent_max"] == (0 if faster_classes else 4096)
    assert (desc["items"] == [0, 0]) == faster_classes
    entries = [ln for ln in cache.read_text().splitlines() if "|ragged|h=" in ln]
    assert len(entries) == 1 and entries[0].rstrip(",").endswith(f"[{desc['persistent_max']}]"), entries
    d = dev(torch_cuda, x)
    p.fft(d)
    check_all(host(d), x, lengths)
    p.close()
    q = api.RaggedPlan(lengths[::-1], api.COMPLEX_FLOAT, api.DFT)  # same classes, other order
    d2 = json.loads(q.describe())
    assert d2["persistent_choice"] == "cache" and d2["persistent_max"] == desc["persistent_max"]
    xr = np.concatenate([x[sum(lengths[:i]):sum(lengths[:i + 1])] for i in range(len(lengths) - 1, -1, -1)])
    e = dev(torch_cuda, xr)
    q.fft(e)
    check_all(host(e), xr, lengths[::-1])
    q.close()
    r = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT, flags=api.FLAG_NO_AUTOTUNE)
    d3 = json.loads(r.describe())
    r.close()
    assert d3["persistent_choice"] == "default" and d3["persistent_max"] == 4096


@pytest.mark.parametrize("mode", MODES)
def test_inverse_round_trip(torch_cuda, monkeypatch, mode):
    monkeypatch.setenv("GFFTX_RAGGED_MODE", mode)
    lengths = [256, 5120, 65536, 384, 640, 12288]
    x = inputs.complex_normal(sum(lengths), 78).astype(np.complex64)
    pf = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    pi = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.IDFT)
    d = dev(torch_cuda, x)
    pf.fft(d)
    y = host(d)
    pi.fft(d)
    back = host(d)
    check_all(y, x, lengths)
    assert rel_l2(back, x) <= 4e-7
    xi = y.copy()
    d2 = dev(torch_cuda, xi)
    pi.fft(d2)
    check_all(host(d2), xi, lengths, inverse=True)


def test_lengths_outside_the_persistent_tiles(torch_cuda):
    """Lengths the persistent kernel has no tile for (below 16, odd factors of 3 and 5 that
    are not 3*2^k or 5*2^k, above 65536) run as batched launches of their class."""
    lengths = [2, 4, 8, 12, 6, 10, 20, 100, 9, 45, 2 ** 17, 12, 3 * 2 ** 15 * 5, 100, 2]
    x = inputs.complex_normal(sum(lengths), 79).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    d = dev(torch_cuda, x)
    p.fft(d)
    check_all(host(d), x, lengths)
    pi = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.IDFT)
    pi.fft(d)
    assert rel_l2(host(d), x) <= 4e-7


def test_unsupported_lengths(torch_cuda):
    with pytest.raises(api.UnsupportedTransform):
        api.RaggedPlan([256, 1], api.COMPLEX_FLOAT)
    with pytest.raises(api.UnsupportedTransform):
        api.RaggedPlan([256, 1], api.COMPLEX_DOUBLE)


@pytest.mark.parametrize("dtype,tol", [(np.complex64, 5e-7), (np.complex128, 1e-13)])
def test_any_length_classes(torch_cuda, dtype, tol):
    """Lengths with a prime factor above 5 run as chirp-z classes whose element-wise kernels
    reach the transforms through the offset table; mixed with smooth lengths, both
    directions, and graph replays."""
    lengths = [7 * 64, 256, 1009, 7 * 64, 11 * 13, 4096, 1009, 97, 2 ** 17 + 1, 12]
    x = inputs.complex_normal(sum(lengths), 81).astype(dtype)
    vt = api.COMPLEX_FLOAT if dtype == np.complex64 else api.COMPLEX_DOUBLE
    pf = api.RaggedPlan(lengths, vt, api.DFT, flags=api.FLAG_GRAPH)
    pi = api.RaggedPlan(lengths, vt, api.IDFT)
    d = dev(torch_cuda, x)
    pf.fft(d)
    y = host(d)
    off = 0
    for L in lengths:
        xs = x[off:off + L].astype(np.complex128)
        ref = np.fft.fft(xs) if L > 4100 else O.dft_ld(xs, 1)
        assert rel_l2(y[off:off + L], ref) <= tol, L
        off += L
    g = dev(torch_cuda, x)
    pf.fft(g)
    assert np.array_equal(host(g), y)  # second replay of the captured graph
    pi.fft(d)
    assert rel_l2(host(d), x) <= 2 * tol


def test_fp64_batch(torch_cuda):
    """FP64 ragged batches run every length class as batched launches (relL2 <= 1e-13)."""
    lengths = [16, 256, 4096, 48, 5120, 65536, 12, 256, 2 ** 17, 3 * 2 ** 10, 16]
    x = inputs.complex_normal(sum(lengths), 80)
    pf = api.RaggedPlan(lengths, api.COMPLEX_DOUBLE, api.DFT)
    pg = api.RaggedPlan(lengths, api.COMPLEX_DOUBLE, api.IDFT, flags=api.FLAG_GRAPH)
    d = dev(torch_cuda, x)
    pf.fft(d)
    check_all(host(d), x, lengths, tol=1e-13)
    for _ in range(2):
        g = d.clone()
        pg.fft(g)
        assert rel_l2(host(g), x) <= 1e-13
    h = x.copy()
    pf.fft(h)  # host array: staged and synchronous
    assert np.array_equal(h, host(d))


@pytest.mark.parametrize("mode", ["0", "1"])
def test_graph_replays_match_eager(torch_cuda, monkeypatch, mode):
    monkeypatch.setenv("GFFTX_RAGGED_MODE", mode)
    lengths = list(inputs.ragged_lengths(300, 11))
    x = inputs.complex_normal(int(sum(lengths)), 12).astype(np.complex64)
    eager = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    graph = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT, flags=api.FLAG_GRAPH)
    d = dev(torch_cuda, x)
    eager.fft(d)
    ref = host(d)
    g = dev(torch_cuda, x)
    for rep in range(4):
        g.copy_(torch_cuda.from_numpy(x))
        graph.fft(g)
        torch_cuda.cuda.synchronize()
        _ = (g.abs().sum()).item()  # eager work between replays
        assert np.array_equal(host(g), ref), f"replay {rep} differs from the eager result"


def test_c4_config_4096_transforms(torch_cuda):
    """4096 FP32 transforms, lengths log-uniform over {2^k,3*2^k,5*2^k} in [2^8,2^16], seed 4096."""
    lengths = [int(v) for v in inputs.ragged_lengths(4096, 4096)]
    total = sum(lengths)
    x = inputs.complex_normal(total, 4096).astype(np.complex64)
    p = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    d = dev(torch_cuda, x)
    p.fft(d)
    worst = check_all(host(d), x, lengths)
    print(f"C4: {len(lengths)} transforms, {total} points, worst relL2 {worst:.3e}")
