"""Throughput of the GFFT transform path on one MI355X.

A step is one forward + inverse in-place C2C round trip of N = 2^24 FP32 points (the large
plain-method workload), inputs resident in HBM, rotating over buffers totalling more than
256 MB so that the Infinity Cache does not serve the input.  value = GFLOP/s with the
reference's flop count 5 N log2 N per direction (test/gfft_performance.h:77).

Also reported on the same line:
  roofline      dominant kernel's algorithmic bytes (one read + one write of the array,
                2 N b_c per pass) over its average duration from HIP events recorded on
                the launch stream (GFFTX_FLAG_PROFILE), against 8 TB/s;
  cpu_baseline  the GFFT restatement (OpenMP in-place path) on the host cores, bounded
                sample, same input distribution;
  fp64          the same round trip in FP64.

roofline.traffic is the HBM traffic of the dominant kernel per launch from two rocprofv3
--pmc passes (FETCH_SIZE, WRITE_SIZE) over `python bench.py --counters`, each scaled by what
the same counter reads for a device copy of 256 MiB in that run (FETCH_SIZE reads half of
the bytes read, WRITE_SIZE all of the bytes written).  roofline.copy_GBps is a 1 GiB device
copy timed in this run: the streaming rate HBM reaches in practice.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1]
       python bench.py --counters      (workloads only, for hardware-counter passes)
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s

# rocprofv3 --pmc over `python bench.py --counters` (N = 2^24 FP32 pass kernels): average
# FETCH_SIZE and WRITE_SIZE per dispatch in KiB, and the same counters for a 256 MiB device
# copy (__amd_rocclr_copyBuffer) in that run.
PMC_COPY_KIB = {"fetch": 131086.98, "write": 262144.0, "bytes_kib": 262144.0}
PMC_KIB = {
    "tile_float_L256_nb16_wg256_truefalse": {"fetch": 65616.38, "write": 131072.0},
    "tile_float_L256_nb16_wg256_truetrue": {"fetch": 65749.73, "write": 131072.0},
}


def hbm_traffic(kernel_name):
    """Calibrated HBM bytes per launch of a 2^24 FP32 pass kernel, or None if not measured."""
    c = PMC_KIB.get(kernel_name)
    if c is None:
        return None
    fetch = c["fetch"] * PMC_COPY_KIB["bytes_kib"] / PMC_COPY_KIB["fetch"]
    write = c["write"] * PMC_COPY_KIB["bytes_kib"] / PMC_COPY_KIB["write"]
    return int(round((fetch + write) * 1024))


def copy_rate(torch, nbytes=1 << 30, iters=20):
    """Read+write bytes per second of a device-to-device copy of `nbytes`."""
    a = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    sec = e0.elapsed_time(e1) * 1e-3 / iters
    del a, b
    torch.cuda.empty_cache()
    return 2 * nbytes / sec


def counters_run():
    """The workloads the counter passes read, a few calls each, after a 256 MiB copy."""
    import torch

    from gfft_amd import api, inputs

    api.lib()
    a = torch.empty(2 ** 26, dtype=torch.float32, device="cuda").fill_(1.0)
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    del a, b
    print("step: 256 MiB copy", flush=True)
    n = 2 ** 24
    x = torch.from_numpy(inputs.complex_normal(n, 24).astype(np.complex64)).cuda()
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as p:
        for _ in range(3):
            p.fft(x)
        torch.cuda.synchronize()
        print("step: 2^24 FP32 forward x3", p.describe(), flush=True)
    del x
    xb = torch.randn(4096 * 16384, dtype=torch.complex64, device="cuda")
    ob = torch.empty_like(xb)
    for fl in (0, api.FLAG_SPLIT_MFMA):
        with api.Plan(4096, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=16384, flags=fl) as p:
            for _ in range(3):
                p.fft(xb, ob)
            torch.cuda.synchronize()
    print("step: 16384 x 4096 FP32 plain and MFMA x3", flush=True)
    del xb, ob
    xd = torch.randn(1024 * 65536, dtype=torch.complex128, device="cuda")
    with api.Plan(1024, api.DOUBLE, api.DFT, api.IN_PLACE, batch=65536) as p:
        for _ in range(3):
            p.fft(xd)
        torch.cuda.synchronize()
    print("step: 65536 x 1024 FP64 x3", flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--log2n", type=int, default=24)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-others", action="store_true")
    ap.add_argument("--counters", action="store_true",
                    help="run the profiled workloads only (for rocprofv3 --pmc passes)")
    return ap.parse_args()


def _settle(torch, run, seconds=0.3):
    """Issues run() until `seconds` of wall time have passed, then waits: the device reaches
    its sustained clocks before a measurement (after a host-side pause, the first tens of
    milliseconds of work measured 5-10% slower)."""
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for _ in range(4):
            run()
        torch.cuda.synchronize()


def round_trip_run(torch, api, n, dtype, steps, warmup, nbuf, seed=24):
    """Times `steps` forward+inverse round trips (kernels issued directly on the stream, no
    per-kernel events inside the timed region), then the same loop with the plans replayed
    as hipGraphs, then per-kernel durations in a separate loop with GFFTX_FLAG_PROFILE plans
    of the same split."""
    from gfft_amd import inputs

    vt = api.COMPLEX_FLOAT if dtype == np.complex64 else api.COMPLEX_DOUBLE
    x = inputs.complex_normal(n, seed).astype(dtype)
    bufs = [torch.from_numpy(x).to("cuda") for _ in range(nbuf)]
    pf = api.Plan(n, vt, api.DFT, api.IN_PLACE)
    pi = api.Plan(n, vt, api.IDFT, api.IN_PLACE)
    stream = torch.cuda.current_stream()
    # one round trip: the identity up to rounding
    pf.fft(bufs[0])
    pi.fft(bufs[0])
    torch.cuda.synchronize()
    rt_err = float(np.linalg.norm(bufs[0].cpu().numpy() - x) / np.linalg.norm(x))
    bufs[0].copy_(torch.from_numpy(x))
    ctr = [0]

    def one(a, b):
        buf = bufs[ctr[0] % nbuf]
        ctr[0] += 1
        a.fft(buf)
        b.fft(buf)

    _settle(torch, lambda: one(pf, pi))
    for i in range(max(warmup, nbuf)):
        b = bufs[i % nbuf]
        pf.fft(b)
        pi.fft(b)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    for i in range(steps):
        b = bufs[i % nbuf]
        pf.fft(b)
        pi.fft(b)
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    desc = pf.describe()
    pf.close()
    pi.close()
    # the same loop with each direction replayed as one captured graph
    gf = api.Plan(n, vt, api.DFT, api.IN_PLACE, flags=api.FLAG_GRAPH)
    gi = api.Plan(n, vt, api.IDFT, api.IN_PLACE, flags=api.FLAG_GRAPH)
    graph_ms = _call_time_us(torch, lambda: one(gf, gi), iters=steps) * 1e-3
    gf.close()
    gi.close()
    # per-kernel durations (HIP events attached to every kernel dispatch), after settling
    qf = api.Plan(n, vt, api.DFT, api.IN_PLACE, flags=api.FLAG_PROFILE)
    qi = api.Plan(n, vt, api.IDFT, api.IN_PLACE, flags=api.FLAG_PROFILE)
    _settle(torch, lambda: one(qf, qi))
    qf.profile()
    qi.profile()
    for i in range(max(steps, 100)):
        one(qf, qi)
    torch.cuda.synchronize()
    prof = {"fwd": qf.profile(), "inv": qi.profile()}
    qf.close()
    qi.close()
    del bufs
    torch.cuda.empty_cache()
    return {"ms": ms, "wall_s": wall, "profile": prof, "rt_err": rt_err, "plan": desc,
            "graph_ms_per_step": graph_ms}


def _kernel_time_us(plan, run, iters=20, warm=3):
    """Average duration (us) of each launch slot of `plan` over `iters` calls of run()."""
    import torch

    for _ in range(warm):
        run()
    torch.cuda.synchronize()
    plan.profile()
    for _ in range(iters):
        run()
    torch.cuda.synchronize()
    return [(name, tot / cnt * 1e3) for name, tot, cnt in plan.profile()]


def _call_time_us(torch, run, iters=50):
    """Average time (us) of one call of run(), all of its launches, over `iters` calls issued
    back to back, from events around the whole loop on the calling stream.  Unlike the
    per-launch event pairs, this leaves out the launch overhead that small kernels would
    otherwise carry each."""
    _settle(torch, run)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(iters):
        run()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) * 1e3 / iters


def other_workloads(torch, api):
    """The remaining transform workloads, measured in the same run (one line each)."""
    from gfft_amd import inputs

    res = {}
    # C1: N=2^10 FP64 in place forward, 2^16 copies per call
    n, batch = 1024, 65536
    x = torch.from_numpy(inputs.complex_uniform(n * batch, 17)).cuda()
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE, batch=batch) as p:
        us = _call_time_us(torch, lambda: p.fft(x), iters=30)
        kernel = [k["kernel"] for k in json.loads(p.describe())["passes"]]
    res["c1_fp64_1024_x65536"] = {"us": round(us, 1), "GFLOPs": round(5 * n * 10 * batch / us / 1e3, 1),
                                  "GBps_alg": round(2 * n * batch * 16 / us / 1e3, 1), "kernels": kernel}
    del x
    # C1 single transform: time per call issued back to back, and the latency of one forward
    # call waited for (launch, kernel and synchronisation, host clock, median of 200); forward
    # and inverse calls alternate so that the values stay bounded
    x1 = torch.from_numpy(inputs.complex_uniform(n, 17)).cuda()
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE) as p, \
            api.Plan(n, api.DOUBLE, api.IDFT, api.IN_PLACE) as q:
        us1 = _call_time_us(torch, lambda: (p.fft(x1), q.fft(x1)), iters=200) / 2
        lat = []
        for _ in range(200):
            t0 = time.perf_counter()
            p.fft(x1)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
            q.fft(x1)
            torch.cuda.synchronize()
    res["c1_fp64_1024_single"] = {"us_per_call_back_to_back": round(us1, 2),
                                  "synchronous_call_us_median": round(float(np.median(lat)) * 1e6, 1)}
    del x1
    # C3: R2C N = 3*2^20 real FP32 -> N/2+1 bins
    nr = 3 * 2 ** 20
    xr, _, _ = inputs.real_sinusoids(nr, 3)
    z = torch.from_numpy(xr.astype(np.float32)).cuda()
    o = torch.zeros(nr // 2 + 1, dtype=torch.complex64, device="cuda")
    # whole-call time of calls issued back to back, then of graph replays; the per-kernel
    # event pairs (which each include a launch) give the breakdown
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as pd:
        us = _call_time_us(torch, lambda: pd.fft(z, o), iters=200)
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS | api.FLAG_GRAPH) as pg:
        us_graph = _call_time_us(torch, lambda: pg.fft(z, o), iters=200)
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS | api.FLAG_PROFILE) as p:
        ks = _kernel_time_us(p, lambda: p.fft(z, o))
    res["c3_r2c_3x2p20_f32"] = {"us": round(us, 1), "graph_us": round(us_graph, 1),
                                "GFLOPs_2.5NlogN": round(2.5 * nr * np.log2(nr) / us / 1e3, 1),
                                "GBps_alg": round((4 * nr + 8 * (nr // 2 + 1)) / us / 1e3, 1),
                                "kernels": [(k, round(t, 1)) for k, t in ks],
                                "kernel_events_sum_us": round(sum(t for _, t in ks), 1)}
    del z, o
    # C4: ragged batch of 4096 FP32 transforms, one call replayed as a captured hipGraph (the
    # workload's definition); the same calls issued directly are reported beside it
    lengths = [int(v) for v in inputs.ragged_lengths(4096, 4096)]
    tot = sum(lengths)
    xg = torch.from_numpy(inputs.complex_normal(tot, 4096).astype(np.complex64)).cuda()
    rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT, flags=api.FLAG_GRAPH)
    us = _call_time_us(torch, lambda: rp.fft(xg), iters=30)
    rp.close()
    rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    us_direct = _call_time_us(torch, lambda: rp.fft(xg), iters=30)
    rp.close()
    flops = sum(5 * L * np.log2(L) for L in lengths)
    res["c4_ragged_4096_f32"] = {"us": round(us, 1), "launch": "hipGraph replay", "direct_us": round(us_direct, 1),
                                 "GFLOPs": round(flops / us / 1e3, 1),
                                 "GBps_alg": round(2 * tot * 8 / us / 1e3, 1), "points": tot}
    del xg
    # C5: 16384 x 4096 FP32 plain vs split-precision MFMA; accuracy against the FP64 engine
    n, batch = 4096, 16384
    xs = inputs.sinusoid_batch(batch, n, 5)
    x32 = torch.from_numpy(xs.astype(np.complex64)).cuda()
    out = torch.empty_like(x32)
    c5 = {}
    for label, flags in (("plain", 0), ("mfma", api.FLAG_SPLIT_MFMA)):
        with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch,
                      flags=flags) as p:
            us = _call_time_us(torch, lambda: p.fft(x32, out), iters=30)
            kernel = json.loads(p.describe())["passes"][0]["kernel"]
            rows = out.cpu().numpy()[::64]
        c5[label] = {"kernel": kernel, "us": round(us, 1),
                     "GFLOPs": round(5 * n * 12 * batch / us / 1e3, 1),
                     "GBps_alg": round(2 * n * batch * 8 / us / 1e3, 1), "rows": rows}
    del x32, out
    x64 = torch.from_numpy(xs[::64].astype(np.complex64).astype(np.complex128)).cuda()
    with api.Plan(n, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE, batch=x64.shape[0]) as p:
        p.fft(x64)
    ref = x64.cpu().numpy()
    for label in ("plain", "mfma"):
        r = c5[label].pop("rows")
        errs = np.linalg.norm(r - ref, axis=1) / np.linalg.norm(ref, axis=1)
        c5[label]["relL2_max_vs_fp64"] = float(errs.max())
    c5["mfma_over_plain"] = round(c5["plain"]["us"] / c5["mfma"]["us"], 3)
    # matrix-core work per launch: 16384 transforms x 3 radix-16 stages x 16 blocks x 2 row
    # halves x 3 hi/lo products of 16x16x32 (the instruction count the MFMA counters read)
    mflops = 16384 * 3 * 16 * 2 * 3 * (16 * 16 * 32 * 2)
    c5["mfma"]["mfma_flops_per_launch"] = mflops
    c5["mfma"]["mfma_TFLOPs"] = round(mflops / c5["mfma"]["us"] / 1e6, 1)
    c5["mfma"]["mfma_frac_of_fp16_peak"] = round(mflops / c5["mfma"]["us"] / 1e6 / 2500.0, 4)
    res["c5_16384x4096_f32"] = c5
    res["host_pointers"] = host_pointer_workloads(torch, api)
    res["vs_torch_fft"] = vendor_fft_comparison(torch, api)
    # any-N path: a prime length (chirp-z convolution of smooth length m >= 2n-1), FP32
    n = 1048573  # prime
    xp = torch.from_numpy(inputs.complex_normal(n, 13).astype(np.complex64)).cuda()
    op = torch.empty_like(xp)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as pd:
        us = _call_time_us(torch, lambda: pd.fft(xp, op), iters=100)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, flags=api.FLAG_GRAPH) as pg:
        us_graph = _call_time_us(torch, lambda: pg.fft(xp, op), iters=100)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, flags=api.FLAG_PROFILE) as p:
        ks = _kernel_time_us(p, lambda: p.fft(xp, op))
        m = json.loads(p.describe())["bluestein"]["m"]
    res["chirp_z_prime_1048573_f32"] = {"us": round(us, 1), "graph_us": round(us_graph, 1), "m": m,
                                        "GFLOPs_5NlogN": round(5 * n * np.log2(n) / us / 1e3, 1),
                                        "GBps_caller": round(2 * n * 8 / us / 1e3, 1),
                                        "kernels": len(ks),
                                        "kernel_events_sum_us": round(sum(t for _, t in ks), 1)}
    del xp, op
    return res


def vendor_fft_comparison(torch, api):
    """The same transforms through torch.fft (hipFFT over the rocFFT that PyTorch ships), timed
    the same way (calls back to back between two events, after settling): the vendor library
    on the same device, as a reference point.  torch.fft writes out of place."""
    out = {}
    n = 1 << 24
    bufs = [torch.randn(n, dtype=torch.complex64, device="cuda") for _ in range(4)]
    tmp = torch.empty_like(bufs[0])
    ctr = [0]

    def torch_rt():
        b = bufs[ctr[0] % 4]
        ctr[0] += 1
        torch.fft.fft(b, out=tmp)
        torch.fft.ifft(tmp, out=b)

    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as pf, \
            api.Plan(n, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE) as pi:
        def ours_rt():
            b = bufs[ctr[0] % 4]
            ctr[0] += 1
            pf.fft(b)
            pi.fft(b)

        t_torch = _call_time_us(torch, torch_rt)
        t_ours = _call_time_us(torch, ours_rt)
    out["c2_f32_round_trip"] = {"torch_fft_us": round(t_torch, 1), "gfftx_us": round(t_ours, 1),
                                "speedup": round(t_torch / t_ours, 2)}
    del bufs, tmp
    torch.cuda.empty_cache()
    x5 = torch.randn(16384, 4096, dtype=torch.complex64, device="cuda")
    o5 = torch.empty_like(x5)
    t_torch = _call_time_us(torch, lambda: torch.fft.fft(x5, dim=-1, out=o5), iters=20)
    with api.Plan(4096, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=16384) as p:
        t_ours = _call_time_us(torch, lambda: p.fft(x5, o5), iters=20)
    out["c5_16384x4096_f32"] = {"torch_fft_us": round(t_torch, 1), "gfftx_us": round(t_ours, 1),
                                "speedup": round(t_torch / t_ours, 2)}
    del x5, o5
    torch.cuda.empty_cache()
    nr = 3 * 2 ** 20
    z = torch.randn(nr, dtype=torch.float32, device="cuda")
    oz = torch.empty(nr // 2 + 1, dtype=torch.complex64, device="cuda")
    t_torch = _call_time_us(torch, lambda: torch.fft.rfft(z, out=oz), iters=200)
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as p:
        t_ours = _call_time_us(torch, lambda: p.fft(z, oz), iters=200)
    out["c3_r2c_3x2p20_f32"] = {"torch_fft_us": round(t_torch, 1), "gfftx_us": round(t_ours, 1),
                                "speedup": round(t_torch / t_ours, 2)}
    return out


def host_pointer_workloads(torch, api, reps=5):
    """GFFT's fft(T*) on host arrays: the library's pinned, chunked pipeline against a plain
    sequence of pageable copies (torch .cuda(), transform, copy back), median of `reps`.
    Rate = host bytes in + out per second."""
    from gfft_amd import inputs

    out = {}
    for label, n, batch in (("single_2p24_f32", 2 ** 24, 1), ("batch32_2p20_f32", 2 ** 20, 32)):
        x = inputs.complex_normal(n * batch, 11).astype(np.complex64).reshape(batch, n)
        y = np.zeros_like(x)
        y2 = np.zeros_like(x)
        y2_t = torch.from_numpy(y2)
        with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch) as p:
            p.fft(x, y)
            d = torch.from_numpy(x).cuda()
            o = torch.empty_like(d)
            t_pipe, t_plain = [], []
            for _ in range(reps):
                t0 = time.perf_counter()
                p.fft(x, y)
                t_pipe.append(time.perf_counter() - t0)
                t0 = time.perf_counter()
                d.copy_(torch.from_numpy(x))
                p.fft(d, o)
                y2_t.copy_(o)
                t_plain.append(time.perf_counter() - t0)
        assert np.array_equal(y, y2)
        nbytes = 2 * x.nbytes
        tp, tq = float(np.median(t_pipe)), float(np.median(t_plain))
        out[label] = {"pipelined_ms": round(tp * 1e3, 2), "pipelined_GBps": round(nbytes / tp / 1e9, 1),
                      "pageable_copies_ms": round(tq * 1e3, 2), "pageable_GBps": round(nbytes / tq / 1e9, 1),
                      "speedup": round(tq / tp, 2)}
        del d, o
    torch.cuda.empty_cache()
    return out


def cpu_baseline(n_log2: int):
    """GFFT restatement, OpenMP in-place path (src/gfftomp.h:105-144), FP32 round trip of
    2^n_log2 points, best of 3 (test/gfft_performance.h:104-123), wall clock."""
    from gfft_amd import inputs
    from oracle import oracle as O

    ncores = os.cpu_count() or 1
    nt = 1
    while nt * 2 <= min(ncores, 16):
        nt *= 2
    n = 1 << n_log2
    x = inputs.complex_normal(n, 24).astype(np.complex64)
    best = 1e30
    for _ in range(3):
        t0 = time.perf_counter()
        y = O.gfft_inplace(x, 1, nt)
        O.gfft_inplace(y, -1, nt)
        best = min(best, time.perf_counter() - t0)
    flops = 2 * 5 * n * n_log2
    return {"value": round(flops / best / 1e9, 3), "unit": "GFLOP/s", "cores": nt, "kind": "port",
            "sample": f"3 round trips (forward+inverse, in place) of N=2^{n_log2} FP32 with the "
                      f"GFFT OpenMP restatement (NT={nt}, serial combine and swap), best of 3 "
                      f"= {best:.2f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported: every "
              "workload here runs on a single MI355X", file=sys.stderr)
        sys.exit(2)
    if args.counters:
        counters_run()
        return
    import torch

    from gfft_amd import api

    api.lib()
    n = 1 << args.log2n
    # > 256 MB of rotating input buffers
    nbuf32 = max(2, math.ceil(512 * 2 ** 20 / (n * 8)))
    r32 = round_trip_run(torch, api, n, np.complex64, args.steps, args.warmup, nbuf32)
    flops_step = 2 * 5 * n * args.log2n
    ms_step = r32["ms"] / args.steps
    value = flops_step / (ms_step * 1e-3) / 1e9

    # dominant kernel of the FP32 round trip
    kern = {}
    for direction in ("fwd", "inv"):
        for i, (name, tot, cnt) in enumerate(r32["profile"][direction]):
            key = f"{direction}:pass{i}:{name}"
            kern[key] = (tot, cnt)
    dom_key = max(kern, key=lambda k: kern[k][0])
    dom_tot, dom_cnt = kern[dom_key]
    dom_avg_s = dom_tot / dom_cnt * 1e-3
    bytes_per_launch = 2 * n * 8
    achieved = bytes_per_launch / dom_avg_s
    copy_bps = copy_rate(torch)
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": hbm_traffic(dom_key.split(":")[-1]) if args.log2n == 24 else None,
                "kernel": dom_key, "avg_us": round(dom_avg_s * 1e6, 2),
                "bytes_per_launch": bytes_per_launch,
                "copy_GBps": round(copy_bps / 1e9, 1), "frac_of_copy": round(achieved / copy_bps, 4)}
    # every pass of the timed round trips together: each pass reads and writes the array once
    npasses = 2 * len(json.loads(r32["plan"])["passes"])
    all_bps = npasses * bytes_per_launch / (ms_step * 1e-3)
    roofline["all_passes_GBps"] = round(all_bps / 1e9, 1)
    roofline["all_passes_frac"] = round(all_bps / HBM_PEAK, 4)
    per_kernel = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2),
                      "GBps_alg": round(bytes_per_launch / (v[0] / v[1] * 1e-3) / 1e9, 1)}
                  for k, v in kern.items()}

    out = {
        "metric": "C2C FFT GFLOP/s + %HBM peak: MFMA split-prec vs plain FP32 at relL2<=2e-7",
        "value": round(value, 2),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: re,im ~ N(0,1) from SplitMix64 seed 24",
        "config": {"workload": f"C2C forward+inverse round trip, N=2^{args.log2n}, FP32, in place, "
                               f"multi-pass Stockham, {nbuf32} rotating buffers",
                   "n": n, "passes": len(json.loads(r32["plan"])["passes"]),
                   "split": [p_["L"] for p_ in json.loads(r32["plan"])["passes"]],
                   "launch": "kernels issued directly on one stream",
                   "graph_replay_ms_per_step": round(r32["graph_ms_per_step"], 4),
                   "round_trip_relL2": r32["rt_err"]},
        "alg_GBps": round(2 * 2 * n * 8 / (ms_step * 1e-3) / 1e9, 1),
        "roofline": roofline,
        "kernels": per_kernel,
    }
    if not args.no_others:
        out["workloads"] = other_workloads(torch, api)
    if not args.no_fp64:
        nbuf64 = max(2, math.ceil(512 * 2 ** 20 / (n * 16)))
        r64 = round_trip_run(torch, api, n, np.complex128, max(10, args.steps // 2),
                             args.warmup, nbuf64)
        ms64 = r64["ms"] / max(10, args.steps // 2)
        out["fp64"] = {"GFLOPs": round(flops_step / (ms64 * 1e-3) / 1e9, 2),
                       "ms_per_step": round(ms64, 4), "round_trip_relL2": r64["rt_err"],
                       "alg_GBps": round(2 * 2 * n * 16 / (ms64 * 1e-3) / 1e9, 1)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.log2n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
