"""Throughput of the GFFT transform path on one MI355X.

A step is one forward + inverse in-place C2C round trip of N = 2^24 FP32 points (the large
plain-method workload), inputs resident in HBM, rotating over buffers totalling more than
256 MB so that the Infinity Cache does not serve the input.  value = GFLOP/s with the
reference's flop count 5 N log2 N per direction (test/gfft_performance.h:77).

Also reported on the same line:
  roofline      algorithmic bytes of the timed step (2 N b_c per transform: one read and one
                write of the array) over the step's time, against 8 TB/s; `dominant_kernel`
                gives the longest launch's share of those bytes (2 N b_c / passes) over its
                average duration from HIP events attached to its dispatches
                (GFFTX_FLAG_PROFILE), and `per_pass_*` the bytes one pass moves (2 N b_c)
                over the same duration; `dominant_kernel_all_launches` the same for the
                kernel with the most time per step over all its launches (forward and
                inverse passes that run the same kernel together), as a profiler's
                per-kernel statistics group them;
  errors        every workload's forward result against numpy's complex128 FFT of the same
                (rounded) input, computed outside the timed regions: relative L2, largest
                absolute component error and GFFT's max|err|/max|y| over re and im
                components (test/gfft_accuracy.h:190-203);
  cpu_baseline  the GFFT restatement (OpenMP in-place path) on the host cores, FP32 and FP64
                round trips, bounded sample, same input;
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
PMC_COPY_KIB = {"fetch": 131085.31, "write": 262144.0, "bytes_kib": 262144.0}
PMC_KIB = {
    "tile_float_L256_nb16_wg256_truefalse": {"fetch": 65616.73, "write": 131072.0},
    "tile_float_L256_nb16_wg256_truetrue": {"fetch": 65753.16, "write": 131072.0},
}


def hbm_traffic(kernel_name):
    """Calibrated HBM bytes per launch of a 2^24 FP32 pass kernel, or None if not measured."""
    c = PMC_KIB.get(kernel_name)
    if c is None:
        return None
    fetch = c["fetch"] * PMC_COPY_KIB["bytes_kib"] / PMC_COPY_KIB["fetch"]
    write = c["write"] * PMC_COPY_KIB["bytes_kib"] / PMC_COPY_KIB["write"]
    return int(round((fetch + write) * 1024))


def log(msg):
    """Progress line on stderr (the result line alone goes to stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def errors(y, ref):
    """Errors of a computed transform y against an FP64 reference: relL2 = |y-ref|_2/|ref|_2,
    max_abs = largest |re| or |im| error component, rel_inf = max_abs / (largest |re| or |im|
    component of y), GFFT's relative error (test/gfft_accuracy.h:190-203)."""
    y = np.asarray(y).astype(np.complex128).ravel()
    ref = np.asarray(ref).astype(np.complex128).ravel()
    d = y - ref
    dmax = float(max(np.abs(d.real).max(), np.abs(d.imag).max()))
    ymax = float(max(np.abs(y.real).max(), np.abs(y.imag).max()))
    return {"relL2": float(np.linalg.norm(d) / np.linalg.norm(ref)), "max_abs": dmax,
            "rel_inf": dmax / ymax if ymax > 0 else 0.0}


def np_fft_rows(x, inverse=False):
    """numpy (pocketfft) complex128 FFT of each row: the independent FP64 reference, with the
    engine's conventions (forward e^-, inverse e^+ scaled 1/N)."""
    x = np.asarray(x).astype(np.complex128)
    return np.fft.ifft(x, axis=-1) if inverse else np.fft.fft(x, axis=-1)


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
    # forward result against numpy's complex128 FFT of the same input, then the round trip
    # (the identity up to rounding)
    pf.fft(bufs[0])
    torch.cuda.synchronize()
    fwd_err = errors(bufs[0].cpu().numpy(), np_fft_rows(x))
    pi.fft(bufs[0])
    torch.cuda.synchronize()
    rt_err = errors(bufs[0].cpu().numpy(), x)
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
    return {"ms": ms, "wall_s": wall, "profile": prof, "rt_err": rt_err, "fwd_err": fwd_err, "plan": desc,
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
    x_h = inputs.complex_uniform(n * batch, 17)
    x = torch.from_numpy(x_h).cuda()
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE, batch=batch) as p:
        us = _call_time_us(torch, lambda: p.fft(x), iters=30)
        kernel = [k["kernel"] for k in json.loads(p.describe())["passes"]]
        x.copy_(torch.from_numpy(x_h))  # one call on the original input for the errors
        p.fft(x)
        rows = x.view(batch, n)[::16].cpu().numpy()
    res["c1_fp64_1024_x65536"] = {"us": round(us, 1), "GFLOPs": round(5 * n * 10 * batch / us / 1e3, 1),
                                  "GBps_alg": round(2 * n * batch * 16 / us / 1e3, 1), "kernels": kernel,
                                  "errors_vs_numpy_fp64": errors(rows, np_fft_rows(x_h.reshape(batch, n)[::16])),
                                  "errors_rows": "every 16th of the 65536 transforms"}
    del x, x_h, rows
    # C1 single transform: time per call issued back to back, and the latency of one forward
    # call waited for (launch, kernel and synchronisation, host clock, median of 200); forward
    # and inverse calls alternate so that the values stay bounded
    x1_h = inputs.complex_uniform(n, 17)
    x1 = torch.from_numpy(x1_h).cuda()
    with api.Plan(n, api.DOUBLE, api.DFT, api.IN_PLACE) as p, \
            api.Plan(n, api.DOUBLE, api.IDFT, api.IN_PLACE) as q:
        p.fft(x1)
        e1 = errors(x1.cpu().numpy(), np_fft_rows(x1_h))
        x1.copy_(torch.from_numpy(x1_h))
        us1 = _call_time_us(torch, lambda: (p.fft(x1), q.fft(x1)), iters=200) / 2
        lat = []
        for _ in range(200):
            t0 = time.perf_counter()
            p.fft(x1)
            torch.cuda.synchronize()
            lat.append(time.perf_counter() - t0)
            q.fft(x1)
            torch.cuda.synchronize()
    log("C1 done")
    res["c1_fp64_1024_single"] = {"us_per_call_back_to_back": round(us1, 2),
                                  "synchronous_call_us_median": round(float(np.median(lat)) * 1e6, 1),
                                  "errors_vs_numpy_fp64": e1}
    del x1
    # C3: R2C N = 3*2^20 real FP32 -> N/2+1 bins
    nr = 3 * 2 ** 20
    xr, _, _ = inputs.real_sinusoids(nr, 3)
    z = torch.from_numpy(xr.astype(np.float32)).cuda()
    o = torch.zeros(nr // 2 + 1, dtype=torch.complex64, device="cuda")
    # whole-call time of calls issued back to back, then of graph replays; the per-kernel
    # event pairs (which each include a launch) give the breakdown
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as pd:
        pd.fft(z, o)
        e3 = errors(o.cpu().numpy(), np.fft.rfft(xr.astype(np.float32).astype(np.float64)))
        us = _call_time_us(torch, lambda: pd.fft(z, o), iters=200)
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS | api.FLAG_GRAPH) as pg:
        us_graph = _call_time_us(torch, lambda: pg.fft(z, o), iters=200)
    with api.Plan(nr // 2, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS | api.FLAG_PROFILE) as p:
        ks = _kernel_time_us(p, lambda: p.fft(z, o))
    log("C3 done")
    res["c3_r2c_3x2p20_f32"] = {"us": round(us, 1), "graph_us": round(us_graph, 1),
                                "GFLOPs_2.5NlogN": round(2.5 * nr * np.log2(nr) / us / 1e3, 1),
                                "GBps_alg": round((4 * nr + 8 * (nr // 2 + 1)) / us / 1e3, 1),
                                "kernels": [(k, round(t, 1)) for k, t in ks],
                                "kernel_events_sum_us": round(sum(t for _, t in ks), 1),
                                "errors_vs_numpy_fp64": e3}
    # C2R, the inverse of C3: the n+1 bins back to 3*2^20 reals, the C2R pre-pass fused into
    # the first pass; errors against numpy's irfft of the same bins (IRDFT inverts RDFT, as
    # irfft inverts rfft) and against the C3 input (round trip)
    zb = torch.empty(nr, dtype=torch.float32, device="cuda")
    with api.Plan(nr // 2, api.FLOAT, api.IRDFT, api.OUT_OF_PLACE, flags=api.FLAG_R2C_BINS) as pd:
        pd.fft(o, zb)
        spec = o.cpu().numpy().astype(np.complex128)
        ec = errors(zb.cpu().numpy().astype(np.complex128), np.fft.irfft(spec, n=nr))
        ert = errors(zb.cpu().numpy().astype(np.complex128), xr.astype(np.float32).astype(np.complex128))
        us = _call_time_us(torch, lambda: pd.fft(o, zb), iters=200)
        kern = [k["kernel"] for k in json.loads(pd.describe())["passes"]]
    log("C2R done")
    res["c2r_3x2p20_f32"] = {"us": round(us, 1), "GBps_alg": round((8 * (nr // 2 + 1) + 4 * nr) / us / 1e3, 1),
                             "kernels": kern, "errors_vs_numpy_fp64": ec, "round_trip_errors": ert}
    del z, o, zb
    # C4: ragged batch of 4096 FP32 transforms, one call replayed as a captured hipGraph (the
    # workload's definition); the same calls issued directly are reported beside it
    lengths = [int(v) for v in inputs.ragged_lengths(4096, 4096)]
    tot = sum(lengths)
    xg_h = inputs.complex_normal(tot, 4096).astype(np.complex64)
    xg = torch.from_numpy(xg_h).cuda()
    rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT, flags=api.FLAG_GRAPH)
    d4 = json.loads(rp.describe())
    rp.fft(xg)
    e4 = ragged_errors(xg.cpu().numpy(), xg_h, lengths)
    xg.copy_(torch.from_numpy(xg_h))
    us = _call_time_us(torch, lambda: rp.fft(xg), iters=30)
    rp.close()
    rp = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    us_direct = _call_time_us(torch, lambda: rp.fft(xg), iters=30)
    rp.close()
    flops = sum(5 * L * np.log2(L) for L in lengths)
    log("C4 done")
    res["c4_ragged_4096_f32"] = {"us": round(us, 1), "launch": "hipGraph replay", "direct_us": round(us_direct, 1),
                                 "GFLOPs": round(flops / us / 1e3, 1),
                                 "GBps_alg": round(2 * tot * 8 / us / 1e3, 1), "points": tot,
                                 "persistent_max": d4.get("persistent_max"),
                                 "persistent_choice": d4.get("persistent_choice"),
                                 "errors_vs_numpy_fp64": e4}
    del xg, xg_h
    # C5: 16384 x 4096 FP32 plain vs split-precision MFMA; accuracy against the FP64 engine
    n, batch = 4096, 16384
    xs = inputs.sinusoid_batch(batch, n, 5)
    x32 = torch.from_numpy(xs.astype(np.complex64)).cuda()
    out = torch.empty_like(x32)
    c5 = {}
    # the two plans are timed in alternation (three rounds of 30 calls each), so that a drift
    # of the device's clocks or temperature during the run weighs on both alike; each
    # reports the median round
    plans = {label: api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, batch=batch, flags=flags)
             for label, flags in (("plain", 0), ("mfma", api.FLAG_SPLIT_MFMA))}
    samples = {label: [] for label in plans}
    for _ in range(3):
        for label, p in plans.items():
            samples[label].append(_call_time_us(torch, lambda: p.fft(x32, out), iters=30))
    for label, p in plans.items():
        p.fft(x32, out)
        rows = out.view(batch, n)[::16].cpu().numpy()
        kernel = json.loads(p.describe())["passes"][0]["kernel"]
        p.close()
        us = float(np.median(samples[label]))
        c5[label] = {"kernel": kernel, "us": round(us, 1), "us_rounds": [round(t, 1) for t in samples[label]],
                     "GFLOPs": round(5 * n * 12 * batch / us / 1e3, 1),
                     "GBps_alg": round(2 * n * batch * 8 / us / 1e3, 1), "rows": rows}
    del x32, out
    ref = np_fft_rows(xs[::16].astype(np.complex64))
    for label in ("plain", "mfma"):
        r = c5[label].pop("rows")
        errs = np.linalg.norm(r - ref, axis=1) / np.linalg.norm(ref, axis=1)
        c5[label]["errors_vs_numpy_fp64"] = errors(r, ref)
        c5[label]["relL2_max_per_transform"] = float(errs.max())
    c5["errors_rows"] = "every 16th of the 16384 transforms"
    c5["mfma_over_plain"] = round(c5["plain"]["us"] / c5["mfma"]["us"], 3)
    # matrix-core work per launch: 16384 transforms x 3 radix-16 stages x 16 blocks x 2 row
    # halves x 3 hi/lo products of 16x16x32 (the instruction count the MFMA counters read)
    mflops = 16384 * 3 * 16 * 2 * 3 * (16 * 16 * 32 * 2)
    c5["mfma"]["mfma_flops_per_launch"] = mflops
    c5["mfma"]["mfma_TFLOPs"] = round(mflops / c5["mfma"]["us"] / 1e6, 1)
    c5["mfma"]["mfma_frac_of_fp16_peak"] = round(mflops / c5["mfma"]["us"] / 1e6 / 2500.0, 4)
    log("C5 done")
    res["c5_16384x4096_f32"] = c5
    res["host_pointers"] = host_pointer_workloads(torch, api)
    log("host pointers done")
    res["vs_torch_fft"] = vendor_fft_comparison(torch, api)
    log("torch.fft comparison done")
    # any-N path: a prime length (chirp-z convolution of smooth length m >= 2n-1), FP32
    n = 1048573  # prime
    xp_h = inputs.complex_normal(n, 13).astype(np.complex64)
    xp = torch.from_numpy(xp_h).cuda()
    op = torch.empty_like(xp)
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE) as pd:
        pd.fft(xp, op)
        ez = errors(op.cpu().numpy(), np_fft_rows(xp_h))
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
                                        "kernel_events_sum_us": round(sum(t for _, t in ks), 1),
                                        "errors_vs_numpy_fp64": ez}
    del xp, op, xp_h
    return res


def ragged_errors(y, x, lengths):
    """Errors of a ragged batch against numpy's complex128 FFT of every transform: the worst
    per-transform relL2 and the errors over the whole batch."""
    offs = np.concatenate([[0], np.cumsum(lengths)])
    ref = np.empty(len(y), np.complex128)
    worst = 0.0
    by_len = {}
    for i, L in enumerate(lengths):
        by_len.setdefault(L, []).append(i)
    for L, idx in by_len.items():
        rows = np.stack([x[offs[i]:offs[i] + L] for i in idx])
        r = np_fft_rows(rows)
        got = np.stack([y[offs[i]:offs[i] + L] for i in idx]).astype(np.complex128)
        rel = np.linalg.norm(got - r, axis=1) / np.linalg.norm(r, axis=1)
        worst = max(worst, float(rel.max()))
        for j, i in enumerate(idx):
            ref[offs[i]:offs[i] + L] = r[j]
    out = errors(y, ref)
    out["relL2_max_per_transform"] = worst
    return out


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


def cpu_info():
    """(CPU model, logical CPUs this process may run on, CPUs the machine reports)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        allowed = len(os.sched_getaffinity(0))
    except AttributeError:
        allowed = os.cpu_count() or 1
    return model, allowed, os.cpu_count() or 1


def cpu_baseline(n_log2: int):
    """GFFT restatement, OpenMP in-place path (src/gfftomp.h:105-144): forward + inverse round
    trips of 2^n_log2 points in FP32 and in FP64, best of 3 each (test/gfft_performance.h:
    104-123), wall clock.  NT = the CPUs this process may run on, rounded down to a power of
    two (GFFT's split uses gcd(N, NT) threads), at most 16: one GPU's share of a GPU host."""
    from gfft_amd import inputs
    from oracle import oracle as O

    model, allowed, reported = cpu_info()
    nt = 1
    while nt * 2 <= min(allowed, 16):
        nt *= 2
    n = 1 << n_log2
    x = inputs.complex_normal(n, 24)
    flops = 2 * 5 * n * n_log2
    res = {}
    for label, dt in (("f32", np.complex64), ("f64", np.complex128)):
        xx = x.astype(dt)
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            y = O.gfft_inplace(xx, 1, nt)
            O.gfft_inplace(y, -1, nt)
            best = min(best, time.perf_counter() - t0)
        res[label] = (flops / best / 1e9, best)
    return {"value": round(res["f32"][0], 3), "unit": "GFLOP/s", "cores": nt, "kind": "port",
            "fp64": {"value": round(res["f64"][0], 3), "unit": "GFLOP/s", "round_trip_s": round(res["f64"][1], 3)},
            "cpu_model": model, "cpus_allowed": allowed, "cpus_reported": reported,
            "sample": f"3 round trips (forward+inverse, in place) of N=2^{n_log2}, FP32 (value) and FP64 "
                      f"(fp64), with the GFFT OpenMP restatement (NT={nt}, serial combine and swap), "
                      f"best of 3 = {res['f32'][1]:.2f} s / {res['f64'][1]:.2f} s per round trip"}


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
    log("FP32 round trip done")
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
    npasses = len(json.loads(r32["plan"])["passes"])
    bytes_transform = 2 * n * 8  # algorithmic bytes of one transform: one read + one write
    bytes_step = 2 * bytes_transform  # forward + inverse
    achieved = bytes_step / (ms_step * 1e-3)
    copy_bps = copy_rate(torch)
    # the longest launch: its share of the transform's algorithmic bytes (one of `npasses`
    # passes), and the bytes a pass itself reads and writes (the whole array once)
    dom_alg = bytes_transform / npasses / dom_avg_s
    dom_pass = bytes_transform / dom_avg_s
    roofline = {"bound": "hbm", "achieved": round(achieved / 1e9, 1), "peak": HBM_PEAK / 1e9,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK, 4),
                "traffic": hbm_traffic(dom_key.split(":")[-1]) if args.log2n == 24 else None,
                "definition": "2 N 8 B per transform (one read + one write of the array), forward + "
                              "inverse per step, over the timed step",
                "dominant_kernel": {"kernel": dom_key, "avg_us": round(dom_avg_s * 1e6, 2),
                                    "alg_bytes_per_launch": bytes_transform // npasses,
                                    "alg_GBps": round(dom_alg / 1e9, 1), "alg_frac": round(dom_alg / HBM_PEAK, 4),
                                    "pass_bytes_per_launch": bytes_transform,
                                    "per_pass_GBps": round(dom_pass / 1e9, 1),
                                    "per_pass_frac": round(dom_pass / HBM_PEAK, 4)},
                "copy_GBps": round(copy_bps / 1e9, 1), "frac_of_copy": round(achieved / copy_bps, 4)}
    # every pass of the timed round trips together: each pass reads and writes the array once
    all_bps = 2 * npasses * bytes_transform / (ms_step * 1e-3)
    roofline["all_passes_GBps"] = round(all_bps / 1e9, 1)
    roofline["all_passes_frac"] = round(all_bps / HBM_PEAK, 4)
    per_kernel = {k: {"avg_us": round(v[0] / v[1] * 1e3, 2),
                      "per_pass_GBps": round(bytes_transform / (v[0] / v[1] * 1e-3) / 1e9, 1)}
                  for k, v in kern.items()}
    # the same launches grouped by kernel (forward and inverse passes that run the same
    # kernel together), as a profiler's per-kernel statistics show them: the kernel with the
    # largest time per step
    by_name = {}
    for k, (tot, cnt) in kern.items():
        name = k.split(":")[-1]
        t, c, m = by_name.get(name, (0.0, 0, 0))
        by_name[name] = (t + tot, c + cnt, m + 1)
    top = max(by_name, key=lambda k: by_name[k][0])
    t_tot, t_cnt, t_per_step = by_name[top]
    t_avg = t_tot / t_cnt * 1e-3
    roofline["dominant_kernel_all_launches"] = {
        "kernel": top, "launches_per_step": t_per_step, "avg_us": round(t_avg * 1e6, 2),
        "alg_GBps": round(bytes_transform / npasses / t_avg / 1e9, 1),
        "per_pass_GBps": round(bytes_transform / t_avg / 1e9, 1),
        "per_pass_frac": round(bytes_transform / t_avg / HBM_PEAK, 4)}

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
                   "graph_replay_ms_per_step": round(r32["graph_ms_per_step"], 4)},
        "errors": {"forward_vs_numpy_fp64": r32["fwd_err"], "round_trip": r32["rt_err"],
                   "input": "seed 24, FP32-rounded, reference numpy complex128 FFT"},
        "alg_GBps": round(achieved / 1e9, 1),
        "roofline": roofline,
        "kernels": per_kernel,
    }
    if not args.no_others:
        out["workloads"] = other_workloads(torch, api)
        c5 = out["workloads"]["c5_16384x4096_f32"]
        # the metric's named comparison: split-precision MFMA against plain FP32 on the
        # 16384 x 4096 batch, both errors against numpy's complex128 FFT
        out["mfma_vs_plain_f32"] = {"speed_ratio": c5["mfma_over_plain"], "target_ratio": 1.5,
                                    "relL2_gate": 2e-7,
                                    "mfma_relL2": c5["mfma"]["errors_vs_numpy_fp64"]["relL2"],
                                    "plain_relL2": c5["plain"]["errors_vs_numpy_fp64"]["relL2"],
                                    "mfma_us": c5["mfma"]["us"], "plain_us": c5["plain"]["us"]}
    if not args.no_fp64:
        nbuf64 = max(2, math.ceil(512 * 2 ** 20 / (n * 16)))
        r64 = round_trip_run(torch, api, n, np.complex128, max(10, args.steps // 2),
                             args.warmup, nbuf64)
        ms64 = r64["ms"] / max(10, args.steps // 2)
        log("FP64 round trip done")
        out["fp64"] = {"GFLOPs": round(flops_step / (ms64 * 1e-3) / 1e9, 2),
                       "ms_per_step": round(ms64, 4),
                       "errors": {"forward_vs_numpy_fp64": r64["fwd_err"], "round_trip": r64["rt_err"]},
                       "alg_GBps": round(2 * 2 * n * 16 / (ms64 * 1e-3) / 1e9, 1),
                       "alg_frac": round(2 * 2 * n * 16 / (ms64 * 1e-3) / HBM_PEAK, 4)}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.log2n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
