"""Plans executed inside a capture the caller runs (torch.cuda.graph on the caller's stream),
and plans issued on several streams at once.

The library records nothing but kernel launches, event records and stream waits on the
stream it is given (no memsets, no allocations, no synchronisation for device pointers), so a
caller may capture a call into its own graph.  Each replay runs after eager work and a
synchronisation of its own and must equal the eager result bit for bit (same kernels)."""
import numpy as np
import pytest

from gfft_amd import api, inputs

pytestmark = pytest.mark.gpu


def _check_replays(torch, call, bufs, x_devs, refs, reps=3):
    """Captures call() once, then replays it `reps` times, each after refilling the inputs,
    some unrelated eager work and a synchronisation; every replay must give refs."""
    for b, x in zip(bufs, x_devs):
        b.copy_(x)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        call()
    junk = torch.ones(1 << 20, device="cuda")
    for r in range(reps):
        for b, x in zip(bufs, x_devs):
            b.copy_(x)
        junk.mul_(1.0001).add_(r)  # eager work between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for out, ref in zip(bufs[len(bufs) - len(refs):], refs):
            assert torch.equal(out, ref), f"replay {r}"
    del g


@pytest.mark.parametrize("case", ["multipass_f32", "batched_f64", "r2c_bins_f32", "chirp_z_f32", "ragged_f32"])
def test_caller_capture_replays_match_eager(torch_cuda, case):
    torch = torch_cuda
    print(f"case {case}", flush=True)
    if case == "ragged_f32":
        lengths = [256, 4096, 1009, 3 * 2 ** 12, 2 ** 15, 64, 5120]
        x = torch.from_numpy(inputs.complex_normal(sum(lengths), 90).astype(np.complex64)).cuda()
        plan = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
        ref = x.clone()
        plan.fft(ref)
        buf = torch.empty_like(x)
        _check_replays(torch, lambda: plan.fft(buf), [buf], [x], [ref])
        plan.close()
        return
    if case == "multipass_f32":
        n, vt, tt, place, batch, flags, dt = 2 ** 20, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, 1, 0, np.complex64
    elif case == "batched_f64":
        n, vt, tt, place, batch, flags, dt = 4096, api.COMPLEX_DOUBLE, api.IDFT, api.IN_PLACE, 64, 0, np.complex128
    elif case == "r2c_bins_f32":
        n, vt, tt, place, batch, flags, dt = 3 * 2 ** 12, api.FLOAT, api.RDFT, api.OUT_OF_PLACE, 1, \
            api.FLAG_R2C_BINS, np.float32
    else:
        n, vt, tt, place, batch, flags, dt = 1009, api.COMPLEX_FLOAT, api.DFT, api.OUT_OF_PLACE, 3, 0, np.complex64
    if dt == np.float32:
        xs = inputs.complex_normal(n * batch, 91).view(np.float64).astype(np.float32)
        out_shape, out_dt = (n + 1,), torch.complex64
    else:
        xs = inputs.complex_normal(n * batch, 91).astype(dt)
        out_shape, out_dt = None, None
    x = torch.from_numpy(xs).cuda()
    with api.Plan(n, vt, tt, place, batch=batch, flags=flags) as p:
        if place == api.IN_PLACE:
            ref = x.clone()
            p.fft(ref)
            buf = torch.empty_like(x)
            _check_replays(torch, lambda: p.fft(buf), [buf], [x], [ref])
        else:
            o_ref = torch.empty(out_shape, dtype=out_dt, device="cuda") if out_shape else torch.empty_like(x)
            p.fft(x, o_ref)
            src, dst = torch.empty_like(x), torch.empty_like(o_ref)
            _check_replays(torch, lambda: p.fft(src, dst), [src, dst], [x], [o_ref])


def test_caller_capture_of_profiled_plan(torch_cuda):
    """A GFFTX_FLAG_PROFILE plan records no timings while the caller captures its stream; its
    replays still equal the eager result, and eager calls are timed again afterwards."""
    torch = torch_cuda
    n = 2 ** 18
    x = torch.from_numpy(inputs.complex_normal(n, 92).astype(np.complex64)).cuda()
    with api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, flags=api.FLAG_PROFILE) as p:
        ref = x.clone()
        p.fft(ref)
        torch.cuda.synchronize()
        assert all(cnt == 1 for _, _, cnt in p.profile())
        buf = torch.empty_like(x)
        _check_replays(torch, lambda: p.fft(buf), [buf], [x], [ref])
        assert all(cnt == 0 for _, _, cnt in p.profile())
        p.fft(buf)
        torch.cuda.synchronize()
        assert all(cnt == 1 for _, _, cnt in p.profile())


def test_plans_on_two_streams_match_serial(torch_cuda):
    """Calls on different plans may run concurrently: two plans issued alternately on two
    streams (default priority) give, bit for bit, what the same calls give one after the
    other on one stream."""
    torch = torch_cuda
    a = torch.from_numpy(inputs.complex_normal(2 ** 20, 93).astype(np.complex64)).cuda()
    b = torch.from_numpy(inputs.complex_normal(4096 * 64, 94).astype(np.complex128)).cuda()
    with api.Plan(2 ** 20, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE) as fa, \
            api.Plan(2 ** 20, api.COMPLEX_FLOAT, api.IDFT, api.IN_PLACE) as ia, \
            api.Plan(4096, api.COMPLEX_DOUBLE, api.DFT, api.IN_PLACE, batch=64) as fb, \
            api.Plan(4096, api.COMPLEX_DOUBLE, api.IDFT, api.IN_PLACE, batch=64) as ib:
        ra, rb = a.clone(), b.clone()
        for _ in range(3):
            fa.fft(ra)
            ia.fft(ra)
            fb.fft(rb)
            ib.fft(rb)
        torch.cuda.synchronize()
        ca, cb = a.clone(), b.clone()
        torch.cuda.synchronize()
        s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
        for _ in range(3):
            with torch.cuda.stream(s1):
                fa.fft(ca)
            with torch.cuda.stream(s2):
                fb.fft(cb)
            with torch.cuda.stream(s1):
                ia.fft(ca)
            with torch.cuda.stream(s2):
                ib.fft(cb)
        torch.cuda.synchronize()
        assert torch.equal(ca, ra)
        assert torch.equal(cb, rb)


@pytest.mark.parametrize("case", ["multipass_f32", "graph_f32", "ragged_f32"])
def test_one_plan_on_two_streams_matches_serial(torch_cuda, case):
    """Calls on ONE plan issued alternately on two unordered streams, each on its own buffer,
    give bit for bit the serial results: the plan's scratch arrays serve every call, and a
    call on the other stream waits for the previous call's work on the device."""
    torch = torch_cuda
    print(f"case {case}", flush=True)
    if case == "ragged_f32":
        lengths = [4096, 2 ** 15, 3 * 2 ** 14, 1009, 256, 2 ** 16]
        tot = sum(lengths)
        xs = [torch.from_numpy(inputs.complex_normal(tot, 95 + i).astype(np.complex64)).cuda() for i in range(2)]
        plan = api.RaggedPlan(lengths, api.COMPLEX_FLOAT, api.DFT)
    else:
        n = 2 ** 22
        xs = [torch.from_numpy(inputs.complex_normal(n, 95 + i).astype(np.complex64)).cuda() for i in range(2)]
        flags = api.FLAG_GRAPH if case == "graph_f32" else 0
        plan = api.Plan(n, api.COMPLEX_FLOAT, api.DFT, api.IN_PLACE, flags=flags)
    refs = [x.clone() for x in xs]
    for _ in range(2):
        for r in refs:
            plan.fft(r)
    torch.cuda.synchronize()
    bufs = [x.clone() for x in xs]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    for _ in range(2):
        for b, s in zip(bufs, streams):
            with torch.cuda.stream(s):
                plan.fft(b)
    torch.cuda.synchronize()
    for b, r in zip(bufs, refs):
        assert torch.equal(b, r)
    plan.close()
