"""gfftx_exec_inplace_noalloc: the no-allocation variant of gfftx_exec_inplace, called through
the C ABI, gives bit for bit what gfftx_exec_inplace gives (same kernels; only the
intermediate arrays differ), on two calls in a row, the second with a smaller input; sizes
beyond its static scratch and host pointers are refused before anything runs."""
import ctypes

import numpy as np
import pytest

from gfft_amd import api, inputs

pytestmark = pytest.mark.gpu


def _noalloc(plan, t):
    L = api.lib()
    s = api._stream(t, None)
    return L.gfftx_exec_inplace_noalloc(plan._h, t.data_ptr(), s)


@pytest.mark.parametrize("case", [
    # (n, vtype, ttype, batch, numpy dtype): first call, then a smaller second call
    ((2 ** 22, api.COMPLEX_FLOAT, api.DFT, 1, np.complex64), (2 ** 20, api.COMPLEX_FLOAT, api.DFT, 1, np.complex64)),
    ((2 ** 21, api.COMPLEX_DOUBLE, api.IDFT, 1, np.complex128), (3 * 2 ** 12, api.COMPLEX_DOUBLE, api.IDFT, 4, np.complex128)),
    ((3 * 2 ** 19, api.FLOAT, api.RDFT, 1, np.complex64), (2 ** 12, api.FLOAT, api.IRDFT, 16, np.complex64)),
    ((4096, api.COMPLEX_FLOAT, api.DFT, 512, np.complex64), (1024, api.COMPLEX_DOUBLE, api.DFT, 64, np.complex128)),
])
def test_noalloc_equals_exec_inplace(torch_cuda, case):
    torch = torch_cuda
    for k, (n, vt, tt, batch, dt) in enumerate(case):
        print(f"case noalloc call {k}: n={n} vtype={vt} ttype={tt} batch={batch}", flush=True)
        x = inputs.complex_normal(n * batch, 1000 + k).astype(dt)
        with api.Plan(n, vt, tt, api.IN_PLACE, batch=batch) as p:
            ref = torch.from_numpy(x).cuda()
            p.fft(ref)
            got = torch.from_numpy(x).cuda()
            assert _noalloc(p, got) == api.SUCCESS, api.lib().gfftx_last_error()
            torch.cuda.synchronize()
            assert torch.equal(got, ref)


def test_noalloc_refusals(torch_cuda):
    torch = torch_cuda
    L = api.lib()
    # a multi-pass batch beyond 32 MiB: refused on the host, nothing launched
    with api.Plan(2 ** 23, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        t = torch.zeros(2 ** 23, dtype=torch.complex64, device="cuda")
        assert _noalloc(p, t) == api.ERR_INVALID_ARG
        assert b"static scratch" in L.gfftx_last_error()
    # host pointers and out-of-place plans are refused
    with api.Plan(1024, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE) as p:
        h = np.zeros(1024, np.complex128)
        assert L.gfftx_exec_inplace_noalloc(p._h, h.ctypes.data, None) == api.ERR_INVALID_ARG
    with api.Plan(1024, api.COMPLEX_DOUBLE, api.DFT, api.OUT_OF_PLACE) as p:
        t = torch.zeros(1024, dtype=torch.complex128, device="cuda")
        assert _noalloc(p, t) == api.ERR_INVALID_ARG
    # a chirp-z length is not supported without allocation
    with api.Plan(1009, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE) as p:
        t = torch.zeros(1009, dtype=torch.complex128, device="cuda")
        assert _noalloc(p, t) == api.ERR_UNSUPPORTED
