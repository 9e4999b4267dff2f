// host_pipeline.cpp -- host-pointer calls: pinned double-buffered staging with threaded host
// copies, both DMA directions at once, and the transform of batch-item groups in between.
#include "plan_internal.h"

namespace gfftx {
namespace detail {

// Copy between pageable and pinned host memory with up to GFFTX_HOST_THREADS threads
// (default 8; chunks of 1 MiB or more).
void host_copy(void* dst, const void* src, size_t len)
{
    unsigned t = std::thread::hardware_concurrency();
    const unsigned cap = (unsigned)std::max(1, std::min(64, env_int("GFFTX_HOST_THREADS", 8)));
    t = std::min<unsigned>(std::max(t, 1u), cap);
    t = (unsigned)std::min<size_t>(t, len >> 20);
    if (t <= 1) {
        memcpy(dst, src, len);
        return;
    }
    const size_t part = ((len + t - 1) / t + 63) & ~(size_t)63;
    std::vector<std::thread> th;
    for (unsigned i = 1; i < t && i * part < len; ++i) {
        const size_t o = i * part, l = std::min(part, len - o);
        th.emplace_back([=] { memcpy((char*)dst + o, (const char*)src + o, l); });
    }
    memcpy(dst, src, std::min(part, len));
    for (auto& x : th) x.join();
}

int ensure_host_pipeline(gfftx_plan_s* p, size_t want)
{
    if (p->pin_bytes >= want) return GFFTX_SUCCESS;
    for (int k = 0; k < 2; ++k) {
        if (p->pin_in[k]) (void)hipHostFree(p->pin_in[k]);
        if (p->pin_out[k]) (void)hipHostFree(p->pin_out[k]);
        p->pin_in[k] = p->pin_out[k] = nullptr;
    }
    p->pin_bytes = 0;
    for (int k = 0; k < 2; ++k) {
        if (hipHostMalloc(&p->pin_in[k], want, hipHostMallocDefault) != hipSuccess ||
            hipHostMalloc(&p->pin_out[k], want, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return fail(GFFTX_ERR_ALLOC, "pinned staging of %zu bytes failed", want);
        }
    }
    p->pin_bytes = want;
    if (!p->h2d_stream) HIP_TRY(hipStreamCreateWithFlags(&p->h2d_stream, hipStreamNonBlocking));
    if (!p->d2h_stream) HIP_TRY(hipStreamCreateWithFlags(&p->d2h_stream, hipStreamNonBlocking));
    for (int k = 0; k < 2; ++k) {
        if (!p->pin_in_ev[k]) HIP_TRY(hipEventCreateWithFlags(&p->pin_in_ev[k], hipEventDisableTiming));
        if (!p->pin_out_ev[k]) HIP_TRY(hipEventCreateWithFlags(&p->pin_out_ev[k], hipEventDisableTiming));
    }
    if (!p->h2d_ev) HIP_TRY(hipEventCreateWithFlags(&p->h2d_ev, hipEventDisableTiming));
    if (!p->done_ev) HIP_TRY(hipEventCreateWithFlags(&p->done_ev, hipEventDisableTiming));
    return GFFTX_SUCCESS;
}

// Host-pointer call.  Batch items go in groups of about one staging chunk; for each group:
// host -> pinned (threads) -> device on the upload stream, the transform of those items on
// the caller's stream, device -> pinned on the download stream -> host.  Two pinned buffers
// each way let the host copies overlap the DMA, and the two directions run at once.  A single
// transform is one group: its upload and download are still chunked and overlapped with the
// host copies.  Returns when `out` holds the result, as GFFT's fft() does.
int exec_host(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, bool din, bool dout, size_t in_bytes,
              size_t out_bytes)
{
    const size_t eb = elem_bytes(p->prec), full = (size_t)p->dist * p->batch * eb, item = (size_t)p->dist * eb;
    if (!din && !p->stage_in.p) {
        int rc = dev_alloc(p, &p->stage_in.p, full);
        if (rc) return rc;
        p->stage_in.bytes = full;
    }
    void* dev_in = din ? const_cast<void*>(in) : p->stage_in.p;
    void* dev_out = dev_in;
    if (in != out) {
        if (dout) {
            dev_out = out;
        } else {
            if (!p->stage_out.p) {
                int rc = dev_alloc(p, &p->stage_out.p, full);
                if (rc) return rc;
                p->stage_out.bytes = full;
            }
            dev_out = p->stage_out.p;
        }
    }
    const size_t chunk = (size_t)std::max(1, env_int("GFFTX_STAGE_CHUNK_MB", 16)) << 20;
    int rc = ensure_host_pipeline(p, std::min(chunk, std::max(in_bytes, out_bytes)));
    if (rc) return rc;
    const size_t pb = p->pin_bytes;
    const unsigned long long per = std::max<unsigned long long>(1, chunk / item);
    struct Pending {
        bool on = false;
        size_t off = 0, len = 0;
    } pend[2];
    unsigned ki = 0, ko = 0;
    auto retire = [&](int k) -> int {
        if (!pend[k].on) return GFFTX_SUCCESS;
        HIP_TRY(hipEventSynchronize(p->pin_out_ev[k]));
        host_copy((char*)out + pend[k].off, p->pin_out[k], pend[k].len);
        pend[k].on = false;
        return GFFTX_SUCCESS;
    };
    for (unsigned long long b0 = 0; b0 < p->batch; b0 += per) {
        const unsigned long long nb = std::min<unsigned long long>(per, p->batch - b0);
        if (!din) {
            const size_t lo = std::min<size_t>(b0 * item, in_bytes), hi = std::min<size_t>((b0 + nb) * item, in_bytes);
            for (size_t o = lo; o < hi; o += pb) {
                const size_t len = std::min(pb, hi - o);
                const int k = ki++ & 1;
                HIP_TRY(hipEventSynchronize(p->pin_in_ev[k]));  // its previous upload is done
                host_copy(p->pin_in[k], (const char*)in + o, len);
                HIP_TRY(hipMemcpyAsync((char*)dev_in + o, p->pin_in[k], len, hipMemcpyHostToDevice, p->h2d_stream));
                HIP_TRY(hipEventRecord(p->pin_in_ev[k], p->h2d_stream));
            }
            HIP_TRY(hipEventRecord(p->h2d_ev, p->h2d_stream));
            HIP_TRY(hipStreamWaitEvent(s, p->h2d_ev, 0));
        }
        rc = exec_device(p, dev_in, dev_out, s, Range{b0, nb});
        if (rc) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamSynchronize(p->h2d_stream);
            return rc;
        }
        if (!dout) {
            HIP_TRY(hipEventRecord(p->done_ev, s));
            HIP_TRY(hipStreamWaitEvent(p->d2h_stream, p->done_ev, 0));
            const size_t lo = std::min<size_t>(b0 * item, out_bytes), hi = std::min<size_t>((b0 + nb) * item, out_bytes);
            for (size_t o = lo; o < hi; o += pb) {
                const size_t len = std::min(pb, hi - o);
                const int k = ko++ & 1;
                rc = retire(k);
                if (rc) return rc;
                HIP_TRY(hipMemcpyAsync(p->pin_out[k], (const char*)dev_out + o, len, hipMemcpyDeviceToHost,
                                       p->d2h_stream));
                HIP_TRY(hipEventRecord(p->pin_out_ev[k], p->d2h_stream));
                pend[k].on = true;
                pend[k].off = o;
                pend[k].len = len;
            }
        }
    }
    rc = retire(ko & 1);  // the older of the two pending downloads first
    if (rc) return rc;
    rc = retire((ko + 1) & 1);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(s));
    HIP_TRY(hipStreamSynchronize(p->h2d_stream));
    return GFFTX_SUCCESS;
}

}  // namespace detail
}  // namespace gfftx
