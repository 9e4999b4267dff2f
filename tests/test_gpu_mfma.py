"""GPU parity of the split-precision MFMA path (radix-16 stages on the matrix cores).

Bar: relL2 <= 2e-7 against the long-double oracle, per transform."""
import numpy as np
import pytest

from conftest import rel_l2
from gfft_amd import api, inputs
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    import torch

    torch.cuda.synchronize()
    return t.cpu().numpy()


SCALES = ["stage", "column"]


@pytest.mark.parametrize("scale", SCALES)
@pytest.mark.parametrize("n,batch", [(256, 64), (4096, 16)])
def test_mfma_forward_inverse(torch_cuda, monkeypatch, scale, n, batch):
    """Both scalings of the FP16 split (per stage, the default; per column) against the
    long-double oracle, forward in place and inverse out of place."""
    monkeypatch.setenv("GFFTX_MFMA_SCALE", scale)
    x = inputs.complex_normal(n * batch, 500 + n).astype(np.complex64).reshape(batch, n)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, batch=batch, flags=api.FLAG_SPLIT_MFMA) as p:
        assert "mfma" in p.describe() and ("_col" in p.describe()) == (scale == "column")
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    with api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.OUT_OF_PLACE, batch=batch, flags=api.FLAG_SPLIT_MFMA) as p:
        s = dev(torch_cuda, x)
        o = torch_cuda.empty_like(s)
        p.fft(s, o)
        yi = host(o)
    for b in range(batch):
        xb = x[b].astype(np.complex128)
        assert rel_l2(y[b], O.ref_fft(xb, 1)) <= 2e-7
        assert rel_l2(yi[b], O.ref_fft(xb, -1) / n) <= 2e-7


@pytest.mark.parametrize("scale", SCALES)
@pytest.mark.parametrize("n", [4096, 256])
def test_mfma_scaling_edge_cases(torch_cuda, monkeypatch, scale, n):
    monkeypatch.setenv("GFFTX_MFMA_SCALE", scale)
    batch = 6
    x = inputs.complex_normal(n * batch, 9).astype(np.complex64).reshape(batch, n)
    x[0] = 0                        # all-zero transform (every column max is 0)
    x[1] *= 1e-20                   # tiny magnitudes
    x[2] *= 1e20                    # huge magnitudes (FP16 range needs the scaling)
    x[3, ::7] *= 1e6                # wide dynamic range inside columns
    x[4] *= 1e-36                   # scale exponent at its clamp (2^126)
    x[5] *= 1e33                    # near the top of FP32 (the transform still fits)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, batch=batch, flags=api.FLAG_SPLIT_MFMA) as p:
        d = dev(torch_cuda, x)
        p.fft(d)
        y = host(d)
    assert np.all(y[0] == 0)
    for b in (1, 2, 3, 4, 5):
        assert rel_l2(y[b], O.ref_fft(x[b].astype(np.complex128))) <= 2e-7


def test_mfma_unsupported_requests(torch_cuda):
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(4096, api.COMPLEX_DOUBLE, api.DFT, flags=api.FLAG_SPLIT_MFMA)
    with pytest.raises(api.UnsupportedTransform):
        api.Plan(1024, api.COMPLEX_FLOAT, api.DFT, flags=api.FLAG_SPLIT_MFMA)


def test_c5_config_16384x4096(torch_cuda):
    """16384 C2C transforms of N=4096: N(0,1) plus a 1e-3 seeded sinusoid, seed 5."""
    n, batch = 4096, 16384
    x = inputs.sinusoid_batch(batch, n, 5).astype(np.complex64)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch,
                  flags=api.FLAG_SPLIT_MFMA) as pm, \
            api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch) as pp:
        s = dev(torch_cuda, x)
        om = torch_cuda.empty_like(s)
        op = torch_cuda.empty_like(s)
        pm.fft(s, om)
        pp.fft(s, op)
        ym, yp = host(om), host(op)
    rng = np.random.default_rng(5)
    sample = rng.choice(batch, 256, replace=False)
    worst_m = worst_p = 0.0
    for b in sample:
        ref = O.ref_fft(x[b].astype(np.complex128))
        worst_m = max(worst_m, rel_l2(ym[b], ref))
        worst_p = max(worst_p, rel_l2(yp[b], ref))
    print(f"C5 sampled worst relL2: mfma {worst_m:.3e}, plain {worst_p:.3e}")
    assert worst_m <= 2e-7 and worst_p <= 2e-7
    assert rel_l2(ym, yp) <= 3e-7
    # every transform against numpy's complex128 FFT (pocketfft, an independent FP64 reference)
    ref = np.fft.fft(x.astype(np.complex128), axis=-1)
    rm = np.linalg.norm(ym - ref, axis=1) / np.linalg.norm(ref, axis=1)
    rp = np.linalg.norm(yp - ref, axis=1) / np.linalg.norm(ref, axis=1)
    print(f"C5 all 16384 transforms, worst relL2 vs numpy: mfma {rm.max():.3e}, plain {rp.max():.3e}")
    assert rm.max() <= 2e-7 and rp.max() <= 2e-7
