"""Host-pointer calls: the pinned, chunked pipeline behind GFFT's synchronous fft(T*).

A host call runs the same kernels as a device call, item group by item group, so its result
is bit-identical to the device call's.  The staging chunk is set to 1 MiB here
(GFFTX_STAGE_CHUNK_MB) so that every case spans several groups or several chunks per item.
"""
import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def small_chunks(monkeypatch):
    monkeypatch.setenv("GFFTX_STAGE_CHUNK_MB", "1")


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


CASES = [
    # n, batch, value type, transform, place
    (2 ** 16, 7, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE),     # 2 items per 1 MiB group
    (2 ** 20, 3, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE),        # 8 chunks per item, multi-pass
    (3 * 2 ** 12, 40, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE),   # ragged last group
    (2 ** 15, 9, api.DOUBLE, api.RDFT, api.OUT_OF_PLACE),          # real forward
    (2 ** 15, 9, api.FLOAT, api.IRDFT, api.OUT_OF_PLACE),          # real inverse (copy of src)
    (1009, 300, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE),      # chirp-z, many groups
    (5 * 2 ** 17, 2, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE),
    (2 ** 14, 5, api.DOUBLE, api.RDFT, api.IN_PLACE),              # GFFT's real mode: in place
    (3 * 2 ** 18, 2, api.FLOAT, api.IRDFT, api.IN_PLACE),           # multi-pass real inverse
]


@pytest.mark.parametrize("n,batch,vt,tt,place", CASES)
def test_host_call_equals_device_call(torch_cuda, n, batch, vt, tt, place):
    dt = np.complex128 if vt in (api.DOUBLE, api.COMPLEX_DOUBLE) else np.complex64
    x = inputs.complex_normal(n * batch, 600 + n).astype(dt).reshape(batch, n)
    print(f"case host n={n} batch={batch} vt={vt} tt={tt} place={place}", flush=True)
    with api.Plan(n, vt, tt, place, batch=batch) as p:
        d = dev(torch_cuda, x)
        if place == api.IN_PLACE:
            p.fft(d)
            ref = host(d)
            h = x.copy()
            p.fft(h)
            assert np.array_equal(h, ref)
        else:
            o = torch_cuda.empty_like(d)
            p.fft(d, o)
            ref = host(o)
            ho = np.zeros_like(x)
            p.fft(x, ho)
            assert np.array_equal(ho, ref)
            # one side on the device, the other on the host
            ho2 = np.zeros_like(x)
            p.fft(d, ho2)
            assert np.array_equal(ho2, ref)
            o2 = torch_cuda.zeros_like(d)
            p.fft(x, o2)
            assert np.array_equal(host(o2), ref)
            assert np.array_equal(host(d), x)


def test_host_call_split_mfma(torch_cuda):
    """The split-precision MFMA plan called on host arrays returns what its device call
    returns, bit for bit (forward out of place, inverse in place)."""
    n, batch = 4096, 64
    x = inputs.complex_normal(n * batch, 650).astype(np.complex64).reshape(batch, n)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch, flags=api.FLAG_SPLIT_MFMA) as pf, \
            api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE, batch=batch, flags=api.FLAG_SPLIT_MFMA) as pi:
        d = dev(torch_cuda, x)
        o = torch_cuda.empty_like(d)
        pf.fft(d, o)
        ref = host(o)
        ho = np.zeros_like(x)
        pf.fft(x, ho)
        assert np.array_equal(ho, ref)
        pi.fft(o)
        back = host(o)
        hb = ho.copy()
        pi.fft(hb)
        assert np.array_equal(hb, back)
        assert rel_l2(hb, x) <= 4e-7


def test_host_call_matches_oracle_and_repeats(torch_cuda):
    """Repeated host calls on one plan (staging reused) stay equal to the oracle."""
    n, batch = 2 ** 14, 20
    x = inputs.complex_normal(n * batch, 3).reshape(batch, n)
    ref = np.stack([O.ref_fft(r, 1) for r in x])
    with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE, batch=batch) as p:
        for _ in range(3):
            out = np.zeros_like(x)
            p.fft(x, out)
            assert rel_l2(out, ref) <= 1e-13


def test_threads_share_one_plan(torch_cuda):
    """GFFT's fft() is reentrant: one transform object may be called from several threads.
    Host-pointer calls from four threads on one plan (and on one ragged plan) each return
    what a call alone returns, bit for bit."""
    import threading

    n, batch = 2 ** 16, 4
    xs = [inputs.complex_normal(n * batch, 700 + i).astype(np.complex64) for i in range(4)]
    lengths = [4096, 2 ** 15, 1536, 20480]
    rs = [inputs.complex_normal(sum(lengths), 710 + i).astype(np.complex64) for i in range(4)]
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch) as p:
        rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
        alone = []
        for x in xs:
            y = np.zeros_like(x)
            p.fft(x, y)
            alone.append(y)
        ragged_alone = []
        for r in rs:
            z = r.copy()
            rp.fft(z)
            ragged_alone.append(z)
        got = [[None] * 3 for _ in xs]
        got_r = [[None] * 3 for _ in rs]
        errors = []

        def work(i):
            try:
                for k in range(3):
                    y = np.zeros_like(xs[i])
                    p.fft(xs[i], y)
                    got[i][k] = y
                    z = rs[i].copy()
                    rp.fft(z)
                    got_r[i][k] = z
            except Exception as e:  # reported below; a thread's exception is not raised
                errors.append(repr(e))

        th = [threading.Thread(target=work, args=(i,)) for i in range(len(xs))]
        for t in th:
            t.start()
        for t in th:
            t.join()
        rp.close()
    assert not errors, errors
    for i in range(len(xs)):
        for k in range(3):
            assert np.array_equal(got[i][k], alone[i]), (i, k)
            assert np.array_equal(got_r[i][k], ragged_alone[i]), (i, k)


def test_exec_batch_equals_single_calls(torch_cuda):
    """gfftx_exec_batch runs plans[i] on data[i] in order: the same plan twice, a batched FP64
    plan and a multi-pass FP32 plan give bit for bit what separate calls give; an
    out-of-place plan in the list is refused before anything runs."""
    torch = torch_cuda
    xs = [inputs.complex_normal(4096, 60).astype(np.complex64), inputs.complex_normal(4096, 61).astype(np.complex64),
          inputs.complex_normal(1024 * 8, 62), inputs.complex_normal(2 ** 20, 63).astype(np.complex64)]
    with api.Plan(4096, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as a, \
            api.Plan(1024, api.COMPLEX_DOUBLE, api.IDFT, api.IN_PLACE, batch=8) as b, \
            api.Plan(2 ** 20, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as c, \
            api.Plan(4096, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as oop:
        plans = [a, a, b, c]
        refs = [torch.from_numpy(x).cuda() for x in xs]
        for p, r in zip(plans, refs):
            p.fft(r)
        bufs = [torch.from_numpy(x).cuda() for x in xs]
        api.exec_batch(plans, bufs)
        torch.cuda.synchronize()
        for r, o in zip(refs, bufs):
            assert torch.equal(r, o)
        keep = [x.clone() for x in bufs]
        with pytest.raises(api.GfftxError):
            api.exec_batch([a, oop], bufs[:2])
        torch.cuda.synchronize()
        assert torch.equal(keep[0], bufs[0])  # nothing ran
