// plan_internal.h -- declarations shared by the host side of the engine: the planner
// (planner.cpp), plans and their execution (plan.cpp), host-pointer staging
// (host_pipeline.cpp), the autotuner (tuner.cpp) and ragged batches (ragged_plan.cpp).
// Included by those files only.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <utility>
#include <vector>
#include <sys/stat.h>
#include <unistd.h>

#include "../../include/gfftx.h"
#include "ragged.h"
#include "registry.h"

namespace gfftx {
namespace detail {

// Thread-local text of the last error (gfftx_last_error); fail() sets it and returns code.
extern thread_local std::string g_last_error;
int fail(int code, const char* fmt, ...);

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(GFFTX_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

// --- planner.cpp: roots, tile lookup, cost model, pass splits, tables
void root_ld(long long m, long long N, long double* c, long double* s);

struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Pass {
    const TileEntry* k = nullptr;
    const MfmaEntry* mfma = nullptr;  // split-precision matrix-core pass (then k is unused)
    void* amat = nullptr;
    long long ntrans = 0, its = 0, ies = 0, ns_pass = 1, bdist = 0;
    unsigned batch = 1;
    void* stage_tw = nullptr;
    double2* ptw_lo = nullptr;
    double2* ptw_hi = nullptr;
    int ptw_shift = 0;
    long long ptw_m = 0;
    int xcd = 0;
    bool items_in_ntrans = false;  // single-pass and MFMA: one sequence per batch item
    int real = REAL_NONE;          // RealCfg tile: the real-transform split fused in
    double2* sp_lo = nullptr;      // its W_{2N} table
    double2* sp_hi = nullptr;
    int sp_shift = 0;
};

// Batch items [b0, b0 + nb) of a plan; nb == 0 means every item.
struct Range {
    unsigned long long b0 = 0, nb = 0;
    bool nt = false;  // the last pass stores the call's output non-temporally
};

// Whether an output of `bytes` is stored non-temporally: larger than the 256 MB Infinity
// Cache (GFFTX_NT_STORE_MIB, default 256; negative: never).
bool nt_output(size_t bytes);

inline size_t elem_bytes(int prec) { return prec == PREC_F64 ? 16 : 8; }

const TileEntry* find_tile(int prec, int L, bool loadT, bool storeT, int want_nb = 0, int want_wg = 0);
const TileEntry* find_real_tile(const TileEntry* k, int kind);
int env_int(const char* name, int dflt);
int device_cus(int device);
double pass_bandwidth(int prec, int L);
double pass_occupancy(int prec, int L, bool first, unsigned long long n, unsigned long long batch, int cus);
std::vector<std::vector<int>> candidate_passes(int prec, unsigned long long n, size_t k, unsigned long long batch = 1,
                                               int cus = 256);
void split_choice(const std::vector<int>& c, std::vector<int>& lens, std::vector<std::pair<int, int>>& tiles);
int xcd_remap(const TileEntry* k, int prec);
bool native_length(int prec, unsigned long long n);
bool choose_passes(int prec, unsigned long long n, std::vector<int>& out, unsigned long long batch = 1, int cus = 256);
unsigned long long bluestein_max_m(int prec);
unsigned long long bluestein_length(int prec, unsigned long long n, unsigned long long batch);
int dev_alloc_list(std::vector<void*>& owned, void** out, size_t bytes);
int upload_stage_table_r(std::vector<void*>& owned, int prec, int nrad, const int* radices, void** out,
                         double scale = 1.0);
int upload_root_tables(std::vector<void*>& owned, long long M, double2** lo_out, double2** hi_out, int* shift);
int upload_mfma_matrix(std::vector<void*>& owned, void** out);
int upload_mfma_twiddles(std::vector<void*>& owned, double scale, void** out);

// Instantiated hipGraphs keyed by the (in, out) device pointers they were captured with.  A
// few entries let a caller rotate over several buffers without re-capturing; the least
// recently used entry is replaced when all are taken.
struct GraphCache {
    static constexpr int kSlots = 8;
    struct Entry {
        const void* in = nullptr;
        void* out = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        unsigned long long used = 0;
    };
    Entry e[kSlots];
    unsigned long long tick = 0;
    hipStream_t cap_stream = nullptr;   // plan-owned stream the captures run on

    Entry* find(const void* in, void* out)
    {
        for (auto& x : e)
            if (x.exec && x.in == in && x.out == out) {
                x.used = ++tick;
                return &x;
            }
        return nullptr;
    }
    // frees the least recently used entry (or picks an empty one) and returns it; an entry
    // that was launched is destroyed only after the device has finished its work (a rare
    // path: more than kSlots buffer pairs in rotation)
    Entry* take()
    {
        Entry* v = &e[0];
        for (auto& x : e) {
            if (!x.exec) {
                v = &x;
                break;
            }
            if (x.used < v->used) v = &x;
        }
        if (v->exec) (void)hipDeviceSynchronize();
        release(*v);
        return v;
    }
    static void release(Entry& x)
    {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        if (x.graph) (void)hipGraphDestroy(x.graph);
        (void)hipGetLastError();
        x = Entry{};
    }
    static hipError_t launch(Entry& x, hipStream_t s) { return hipGraphLaunch(x.exec, s); }
    // the plan is being destroyed: its launched graphs finish before they are released
    void clear()
    {
        bool launched = false;
        for (auto& x : e) launched = launched || x.exec;
        if (launched) (void)hipDeviceSynchronize();
        for (auto& x : e) release(x);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        cap_stream = nullptr;
    }
};

// Orders the calls of one plan that arrive on different streams.  A plan's scratch arrays,
// staging buffers and graphs serve every call, so a call issued on another stream than the
// previous call's first waits until the device has finished its earlier work, that call's
// included (hipDeviceSynchronize: the previous stream may no longer exist, so nothing is
// recorded on it).  Calls on one stream are ordered by the stream and cost nothing here: an
// event recorded after every call measured 3-6 us per call on the device timeline.  A call
// on a stream that the caller is capturing into a graph is not ordered this way (no host
// wait is allowed there); the caller orders the replays of its own graphs.
struct StreamOrder {
    hipStream_t last = nullptr;
    bool used = false;
    hipError_t before(hipStream_t s)
    {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        hipError_t e = hipStreamIsCapturing(s, &cs);
        if (e != hipSuccess) return e;
        if (cs != hipStreamCaptureStatusNone) return hipSuccess;
        if (used && s != last) {
            e = hipDeviceSynchronize();
            if (e != hipSuccess) return e;
        }
        last = s;
        used = true;
        return hipSuccess;
    }
};

struct DeviceGuard {
    int prev = -1;
    bool ok = false;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess) {
            (void)hipGetLastError();
            prev = -1;
        }
        ok = hipSetDevice(dev) == hipSuccess;
        if (!ok) (void)hipGetLastError();
    }
    ~DeviceGuard()
    {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
};

bool is_device_ptr(const void* ptr);

}  // namespace detail
}  // namespace gfftx

using namespace gfftx;
using namespace gfftx::detail;

struct gfftx_plan_s {
    int device = 0;
    unsigned long long n = 0;      // complex length of one transform
    unsigned long long batch = 1;
    int vtype = 0, ttype = 0, place = 0;
    uint32_t flags = 0;
    int prec = PREC_F64;
    bool real = false;
    bool inverse = false;
    unsigned long long dist = 0;   // complex elements between batch items
    std::vector<int> forced_lens;  // plan choice made by the autotuner (empty: cost model), see split_choice
    std::string tune_note;         // autotuner record for plan_describe
    std::vector<Pass> passes;
    std::vector<void*> allocs;     // device allocations owned by the plan
    void* scratch[2] = {nullptr, nullptr};
    void* split_tw = nullptr;
    bool split_fused = false;      // real transform: the split runs inside a pass (last
                                   // pass forward, first pass inverse)
    void* real_work = nullptr;     // IRDFT out-of-place: split runs on a copy of src
    DevBuf stage_in, stage_out;    // device arrays a host-pointer call runs on
    // host-pointer pipeline: two pinned buffers each way, a stream for each copy direction
    void* pin_in[2] = {nullptr, nullptr};
    void* pin_out[2] = {nullptr, nullptr};
    size_t pin_bytes = 0;
    hipStream_t h2d_stream = nullptr, d2h_stream = nullptr;
    hipEvent_t pin_in_ev[2] = {nullptr, nullptr}, pin_out_ev[2] = {nullptr, nullptr};
    hipEvent_t h2d_ev = nullptr, done_ev = nullptr;
    std::string desc;
    // GFFTX_FLAG_PROFILE: one (start, stop) event pair per launch slot and execution
    struct Slot {
        std::string name;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pending;
        std::vector<std::pair<hipEvent_t, hipEvent_t>> pool;
    };
    std::vector<Slot> slots;
    bool capturing = false;
    // a no-allocation call (gfftx_exec_inplace_noalloc) in progress: its passes use these
    // intermediate arrays instead of `scratch`, and no profiling events are created
    void* call_scratch[2] = {nullptr, nullptr};
    bool no_events = false;
    // GFFTX_FLAG_GRAPH: launches captured once per (in, out) pair and replayed
    GraphCache graphs;
    // chirp-z path (n has a prime factor above 5): convolution of length blue_m
    unsigned long long blue_m = 0;
    gfftx_plan_s* blue_fwd = nullptr;  // forward C2C of length blue_m, in place on blue_work
    gfftx_plan_s* blue_inv = nullptr;  // inverse C2C (1/blue_m), in place on blue_work
    void* blue_chirp = nullptr;        // n entries W_{2n}^{j^2}
    void* blue_filter = nullptr;       // blue_m entries: spectrum of the chirp filter
    void* blue_work = nullptr;         // batch x blue_m
    // calls from several threads run one at a time: the graph cache, the profiling slots and
    // the host staging are the plan's own
    std::mutex mu;
    StreamOrder order;          // calls from several streams wait for each other
    bool shared_state = false;  // calls write plan-owned device arrays (scratch, graphs)
};

// ---------------------------------------------------------------------------------------
// Ragged batches: one persistent launch over a list of work items (ragged.hip) for the
// lengths up to the plan's limit, batched class launches for the others.
// Side streams for the class launches: with the caller's stream they fill the four
// hardware queues a process has by default.
constexpr int kRaggedSides = 3;

struct gfftx_ragged_s {
    int device = 0;
    int prec = PREC_F32;   // FP64 batches run every length class as batched launches
    bool inverse = false;
    uint32_t flags = 0;
    unsigned long long total = 0;          // complex elements of the whole batch
    std::vector<void*> allocs;
    RaggedArgs args{};
    int grid = 0, grid1 = 0;  // workgroups of the phase-0 and phase-1 launches
    // Grouped mode (default): each length class above the persistent limit (4096 points, or
    // 0 where the autotuner found class launches faster) runs as its own batched launch
    // sequence (a single-pass tile through the sequence-offset table, or multi-pass tiles
    // through a batch-offset table) on a side stream, while the persistent kernel takes the
    // shorter lengths on the caller's stream.
    struct LongClass {
        long long L = 0;
        long long count = 0;
        const long long* offs = nullptr;  // element offsets of the class's transforms
        std::vector<Pass> passes;         // one pass: single-pass tile over all of them
        gfftx_plan_s* blue = nullptr;     // length with a prime factor above 5: chirp-z plan
    };
    std::vector<LongClass> longs;
    void* lscratch[2][kRaggedSides] = {};  // per side stream, the size of the whole batch: same offsets
    int nsides = 0;                        // side streams in use (classes spread by work)
    hipStream_t side[kRaggedSides] = {};
    hipEvent_t fork_ev = nullptr, join_ev[kRaggedSides] = {};
    std::vector<int> side_of;              // side stream of each long class
    std::string desc;
    // hipGraph replay (GFFTX_FLAG_GRAPH), keyed by the data pointer
    GraphCache graphs;
    // host staging
    void* stage = nullptr;
    // profiling (GFFTX_FLAG_PROFILE): events around the launch
    std::vector<std::pair<hipEvent_t, hipEvent_t>> pending, pool;
    std::mutex mu;  // calls from several threads run one at a time
    StreamOrder order;  // calls from several streams wait for each other
};

namespace gfftx {
namespace detail {

// --- plan.cpp: plan construction and execution
int dev_alloc(gfftx_plan_s* p, void** out, size_t bytes);
int create_plan_impl(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                     uint32_t flags, const std::vector<int>& forced, const std::string& note);
int exec_device(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r = Range());
int run_passes(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, gfftx_plan_s* prof = nullptr,
               size_t slot0 = 0);
int run_bluestein(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, const long long* boff = nullptr);

// --- host_pipeline.cpp: host-pointer calls
int exec_host(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, bool din, bool dout, size_t in_bytes,
              size_t out_bytes);

// --- tuner.cpp: plan creation with the measured autotuner
int create_plan(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                uint32_t flags);
bool tune_length_class(int device, int prec, long long L, long long count, bool inverse, bool* single,
                       std::vector<int>* lens, std::vector<std::pair<int, int>>* tiles);
// cache entries of other choices (the ragged batch's persistent-kernel limit)
std::string ragged_tune_key(int device, int prec, const uint64_t* lengths, uint64_t count);
bool tune_get(const std::string& key, std::vector<int>* v);
void tune_put(const std::string& key, const std::vector<int>& v);

}  // namespace detail
}  // namespace gfftx
