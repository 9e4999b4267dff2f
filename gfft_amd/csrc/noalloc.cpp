// noalloc.cpp -- in-place execution that allocates nothing: the intermediate arrays of a
// multi-pass transform come from a static device array instead of the plan's scratch.
//
// For callers that cannot allocate during a call (code under graph capture, tight serving
// loops, harnesses that forbid it).  The array is a file-scope __device__ object of
// kNoAllocScratchBytes, which every GPU the library runs on holds until the process ends.  Its
// device address is looked up once per device with hipGetSymbolAddress (the name itself is
// only a host-side stand-in, never a device address) and that address is what the kernels
// receive.  Every call on one GPU shares the array: calls that overlap on one GPU (several
// streams, several threads) race on it, so the caller orders them.  Nothing needs to start
// at zero: each pass writes every element it hands to the next one.
#include "plan_internal.h"

namespace gfftx {
namespace detail {

// Two halves: the outputs of alternate passes (a transform of three or more passes needs
// both; see run_passes).
constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr size_t kNoAllocHalf = kNoAllocScratchBytes / 2;
constexpr int kMaxDevices = 64;

static __device__ __align__(256) double2 scratch[kNoAllocScratchBytes / sizeof(double2)];

namespace {
std::mutex g_noalloc_mu;
void* g_scratch_ptr[kMaxDevices] = {};  // device address of `scratch` on each device
}  // namespace

// Device address of the static scratch on the current device (set-up, first call per device).
int noalloc_scratch(int device, void** out)
{
    if (device < 0 || device >= kMaxDevices) return fail(GFFTX_ERR_DEVICE, "device %d beyond %d", device, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_scratch_ptr[device]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return fail(GFFTX_ERR_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return GFFTX_SUCCESS;
}

}  // namespace detail
}  // namespace gfftx

extern "C" {

int gfftx_exec_inplace_noalloc(gfftx_plan plan, void* data, void* stream)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan is NULL");
    if (!data) return fail(GFFTX_ERR_INVALID_ARG, "data is NULL");
    if (plan->place != GFFTX_IN_PLACE) return fail(GFFTX_ERR_INVALID_ARG, "plan is out-of-place");
    if (plan->blue_fwd)
        return fail(GFFTX_ERR_UNSUPPORTED, "no-allocation calls take lengths of the factors 2, 3, 5 and 7 only (n=%llu)",
                    plan->n);
    const size_t bytes = (size_t)plan->dist * plan->batch * elem_bytes(plan->prec);
    if (plan->passes.size() > 1 && bytes > kNoAllocHalf)
        return fail(GFFTX_ERR_INVALID_ARG,
                    "no-allocation call: the transform's %zu bytes exceed the static scratch (%zu bytes per "
                    "intermediate array)",
                    bytes, (size_t)kNoAllocHalf);
    std::lock_guard<std::mutex> lk(plan->mu);
    DeviceGuard g(plan->device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", plan->device);
    if (!is_device_ptr(data)) return fail(GFFTX_ERR_INVALID_ARG, "no-allocation calls take device pointers");
    void* base = nullptr;
    int rc = noalloc_scratch(plan->device, &base);
    if (rc) return rc;
    plan->call_scratch[0] = base;
    plan->call_scratch[1] = (char*)base + kNoAllocHalf;
    plan->no_events = true;
    rc = exec_device(plan, data, data, (hipStream_t)stream);
    plan->call_scratch[0] = plan->call_scratch[1] = nullptr;
    plan->no_events = false;
    return rc;
}

}  // extern "C"
