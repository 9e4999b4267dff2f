"""The exchange between two radix-16 stages through LDS, against cross-lane forms of it.

tests/cpp/exchange_probe (built by __graft_entry__.build()) runs a batched 256-point FP32 FFT
with the library's 256-point tile shape four ways, differing only in the stage exchange: LDS
(as the library does), DPP moves, ds_swizzle, and gfx950's permlane swaps.  The cross-lane
forms must match the LDS form bit for bit, and the LDS form a double-precision DFT.  The
timings it prints (HBM-bound and cache-resident) are the measurement behind the library's
choice of the LDS exchange.
"""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu


def test_cross_lane_exchanges_match_lds(torch_cuda):
    exe = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "exchange_probe")
    if not os.path.exists(exe):
        pytest.skip("tests/cpp/exchange_probe is not built (run __graft_entry__.build())")
    print("case exchange probe: 65536 x 256 FP32, four exchange forms", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout[-3000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout, r.stdout + r.stderr
    for form in ("DPP", "ds_swizzle", "permlane+DPP"):
        assert f"{form} exchange: 0 of 16777216 elements differ" in r.stdout
