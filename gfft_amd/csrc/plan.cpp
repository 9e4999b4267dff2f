// plan.cpp -- planner and C ABI of the gfx950 FFT engine.
//
// A plan is a list of passes over device buffers.  A transform whose length fits one LDS
// tile is one batched pass; a longer one is a Stockham decomposition into "super radices"
// N = P_0 * P_1 * ... each executed as one pass of the tile kernel (see tile_kernel.h).
// Twiddles are generated here once per plan from long-double evaluations with octant
// argument reduction and stored in the working precision (stage tables) or FP64 (pass
// tables).  GFFT evaluates the same roots W_N = exp(-2 pi i/N) at compile time to 18
// digits (src/metaroot.h:104-120) and then runs a recurrence at run time
// (src/gfftalg.h:249-261); tables avoid the recurrence's error growth.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <sys/stat.h>
#include <unistd.h>
#include <cctype>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/gfftx.h"
#include "ragged.h"
#include "registry.h"

using namespace gfftx;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_TRY(expr)                                                                  \
    do {                                                                               \
        hipError_t e_ = (expr);                                                        \
        if (e_ != hipSuccess)                                                          \
            return fail(GFFTX_ERR_DEVICE, "%s: %s", #expr, hipGetErrorString(e_));      \
    } while (0)

// exp(-2 pi i m / N) in long double, m reduced mod N and folded into the first octant.
void root_ld(long long m, long long N, long double* c, long double* s)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    m %= N;
    if (m < 0) m += N;
    // angle = (pi/4) * (8m/N); octant o and remainder in units of N/8
    const long double x8 = 8.0L * (long double)m;
    long long o = (long long)(x8 / (long double)N);
    if (o > 7) o = 7;
    const long double r = x8 - (long double)o * (long double)N;
    long double cc, ss;
    const bool odd = (o & 1) != 0;
    const long double a = (PI / 4.0L) * (odd ? ((long double)N - r) : r) / (long double)N;
    const long double ca = cosl(a), sa = sinl(a);
    // cos/sin of the full angle from the folded one
    switch (o) {
        case 0: cc = ca; ss = sa; break;
        case 1: cc = sa; ss = ca; break;
        case 2: cc = -sa; ss = ca; break;
        case 3: cc = -ca; ss = sa; break;
        case 4: cc = -ca; ss = -sa; break;
        case 5: cc = -sa; ss = -ca; break;
        case 6: cc = sa; ss = -ca; break;
        default: cc = ca; ss = -sa; break;
    }
    *c = cc;
    *s = -ss;
}

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
    bool split = false;            // SplitCfg tile: the R2C split fused into this last pass
    double2* sp_lo = nullptr;      // its W_{2N} table
    double2* sp_hi = nullptr;
    int sp_shift = 0;
};

// Batch items [b0, b0 + nb) of a plan; nb == 0 means every item.
struct Range {
    unsigned long long b0 = 0, nb = 0;
};

size_t elem_bytes(int prec) { return prec == PREC_F64 ? 16 : 8; }

// First registered tile for (prec, L, modes), or the first one with nb == want_nb and
// wg == want_wg (each when nonzero) when there is one.
const TileEntry* find_tile(int prec, int L, bool loadT, bool storeT, int want_nb = 0, int want_wg = 0)
{
    const TileEntry* first = nullptr;
    for (int i = 0; i < kNumTiles; ++i) {
        const TileEntry& e = kTiles[i];
        if (e.prec == prec && e.L == L && e.loadT == loadT && e.storeT == storeT && !e.split) {
            if ((want_nb || want_wg) && (!want_nb || e.nb == want_nb) && (!want_wg || e.wg == want_wg)) return &e;
            if (!first) first = &e;
        }
    }
    return first;
}

// The fused-split variant of a T->T pass tile (same length and shape), or null.
const TileEntry* find_split_tile(const TileEntry* k)
{
    for (int i = 0; i < kNumTiles; ++i) {
        const TileEntry& e = kTiles[i];
        if (e.split && e.prec == k->prec && e.L == k->L && e.nb == k->nb && e.wg == k->wg) return &e;
    }
    return nullptr;
}

int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

// Compute units of a device (256 on MI355X when the query fails).
int device_cus(int device)
{
    static int cached[64] = {0};
    if (device >= 0 && device < 64 && cached[device]) return cached[device];
    hipDeviceProp_t prop;
    int cus = 256;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) cus = prop.multiProcessorCount;
    else (void)hipGetLastError();
    if (device >= 0 && device < 64) cached[device] = cus;
    return cus;
}

// Bandwidth (TB/s) a multi-pass tile of length L reaches on MI355X with XCD remapping,
// measured over N = 2^24 passes; lengths not listed are interpolated by size class.
double pass_bandwidth(int prec, int L)
{
    const bool f32 = prec == PREC_F32;
    if (L < 16) return 5.0;  // one extra pass over the data for a leftover factor
    if (L <= 16) return f32 ? 5.5 : 5.6;
    if (L <= 64) return f32 ? 4.6 : 4.7;
    if (L <= 256) return f32 ? 4.2 : 4.6;
    if (L <= 1536) return f32 ? 3.1 : 3.5;
    return f32 ? 2.6 : 2.5;
}

// Fraction of the device one pass of length L keeps busy: its tiles (n/L sequences per
// batch item, NB per tile) over the workgroups all CUs hold at once (LDS and threads
// limit a CU to 160 KB and 2048 threads).
double pass_occupancy(int prec, int L, bool first, unsigned long long n, unsigned long long batch, int cus)
{
    const TileEntry* k = find_tile(prec, L, true, !first);
    if (!k) return 1.0;
    const int by_lds = k->lds_bytes > 0 ? (160 * 1024) / k->lds_bytes : 8;
    const int per_cu = std::max(1, std::min(by_lds, 2048 / k->wg));
    const double tiles = std::ceil((double)(n / (unsigned long long)L) / k->nb) * (double)batch;
    return std::min(1.0, tiles / ((double)cus * per_cu));
}

// Pass splits of n into compiled T-mode tile lengths (2..4 passes), best `k` by the cost
// model: each pass moves 2 n b_c bytes at the bandwidth its tile reaches
// (pass_bandwidth()) times the fraction of the device its tiles occupy.  The first length
// of every split has a T->E tile.
std::vector<std::vector<int>> candidate_passes(int prec, unsigned long long n, size_t k, unsigned long long batch = 1,
                                               int cus = 256)
{
    std::vector<int> lens;
    for (int i = 0; i < kNumTiles; ++i)
        if (kTiles[i].prec == prec && kTiles[i].loadT && kTiles[i].storeT && !kTiles[i].split) lens.push_back(kTiles[i].L);
    std::sort(lens.begin(), lens.end());
    lens.erase(std::unique(lens.begin(), lens.end()), lens.end());
    std::vector<std::pair<double, std::vector<int>>> found;
    std::vector<int> cur;
    std::function<void(unsigned long long, int)> rec = [&](unsigned long long rem, int maxi) {
        if (rem == 1) {
            if (cur.size() < 2) return;
            std::vector<int> order = cur;
            for (size_t i = 0; i < order.size(); ++i)
                if (find_tile(prec, order[i], true, false)) {
                    std::swap(order[0], order[i]);
                    break;
                }
            if (!find_tile(prec, order[0], true, false)) return;
            double cost = 0;
            for (size_t i = 0; i < order.size(); ++i)
                cost += 1.0 / (pass_bandwidth(prec, order[i]) * pass_occupancy(prec, order[i], i == 0, n, batch, cus));
            found.push_back({cost, order});
            return;
        }
        if (cur.size() == 4) return;
        for (int i = maxi; i >= 0; --i) {
            if (rem % (unsigned long long)lens[i]) continue;
            cur.push_back(lens[i]);
            rec(rem / lens[i], i);
            cur.pop_back();
        }
    };
    rec(n, (int)lens.size() - 1);
    std::stable_sort(found.begin(), found.end(),
                     [](const std::pair<double, std::vector<int>>& x, const std::pair<double, std::vector<int>>& y) {
                         return x.first < y.first;
                     });
    std::vector<std::vector<int>> out;
    for (size_t i = 0; i < found.size() && out.size() < k; ++i) out.push_back(found[i].second);
    return out;
}

// A plan choice is the pass lengths, optionally followed by 0 and one (nb, wg) pair per
// pass naming the tile to use ((0, 0): the default tile of that length and mode).
void split_choice(const std::vector<int>& c, std::vector<int>& lens, std::vector<std::pair<int, int>>& tiles)
{
    lens.clear();
    tiles.clear();
    size_t i = 0;
    while (i < c.size() && c[i] != 0) lens.push_back(c[i++]);
    if (i < c.size()) ++i;
    while (i + 1 < c.size()) {
        tiles.push_back({c[i], c[i + 1]});
        i += 2;
    }
}

// Whether n runs natively: one compiled single-pass tile, or a split into compiled pass
// tiles (lengths whose factors the tiles carry: 2, 3, 5 and 7).  Any other n runs as a
// chirp-z convolution.
bool native_length(int prec, unsigned long long n)
{
    if (n <= (1ull << 30) && find_tile(prec, (int)n, false, false)) return true;
    return !candidate_passes(prec, n, 1).empty();
}

// Split n into pass lengths available as T-mode tiles (largest first).  An override list
// (GFFTX_PASSES="a,b,c") is honoured when it multiplies to n.
bool choose_passes(int prec, unsigned long long n, std::vector<int>& out, unsigned long long batch = 1, int cus = 256)
{
    out.clear();
    if (const char* env = getenv("GFFTX_PASSES")) {
        std::vector<int> v;
        unsigned long long prod = 1;
        const char* p = env;
        while (*p) {
            const long x = strtol(p, const_cast<char**>(&p), 10);
            if (x <= 1) break;
            v.push_back((int)x);
            prod *= (unsigned long long)x;
            if (*p == ',') ++p;
            else break;
        }
        bool ok = prod == n && v.size() >= 2;
        for (size_t i = 0; ok && i < v.size(); ++i)
            ok = find_tile(prec, v[i], true, i == 0 ? false : true) != nullptr;
        if (ok) {
            out = v;
            return true;
        }
    }
    std::vector<std::vector<int>> c = candidate_passes(prec, n, 1, batch, cus);
    if (c.empty()) return false;
    out = c[0];
    return true;
}

}  // namespace

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
    // frees the least recently used entry (or picks an empty one) and returns it
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
        release(*v);
        return v;
    }
    static void release(Entry& x)
    {
        if (x.exec) (void)hipGraphExecDestroy(x.exec);
        if (x.graph) (void)hipGraphDestroy(x.graph);
        x = Entry{};
    }
    void clear()
    {
        for (auto& x : e) release(x);
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
        cap_stream = nullptr;
    }
};

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
    bool split_fused = false;      // real forward: the split runs inside the last pass
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
    // GFFTX_FLAG_GRAPH: launches captured once per (in, out) pair and replayed
    GraphCache graphs;
    // chirp-z path (n has a prime factor above 5): convolution of length blue_m
    unsigned long long blue_m = 0;
    gfftx_plan_s* blue_fwd = nullptr;  // forward C2C of length blue_m, in place on blue_work
    gfftx_plan_s* blue_inv = nullptr;  // inverse C2C (1/blue_m), in place on blue_work
    void* blue_chirp = nullptr;        // n entries W_{2n}^{j^2}
    void* blue_filter = nullptr;       // blue_m entries: spectrum of the chirp filter
    void* blue_work = nullptr;         // batch x blue_m
};

namespace {

int dev_alloc(gfftx_plan_s* p, void** out, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GFFTX_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    p->allocs.push_back(*out);
    return GFFTX_SUCCESS;
}

template <typename T> void pack_stage_table(int nrad, const int* radices, std::vector<T>& out)
{
    // stage s >= 1: rows k in [0, Ns), entries r in [0, R): W_{Ns*R}^{k*r}
    out.clear();
    long long ns = radices[0];
    for (int s = 1; s < nrad; ++s) {
        const long long R = radices[s];
        for (long long kk = 0; kk < ns; ++kk)
            for (long long r = 0; r < R; ++r) {
                long double c, si;
                root_ld(kk * r, ns * R, &c, &si);
                out.push_back((T)c);
                out.push_back((T)si);
            }
        ns *= R;
    }
}

// Device allocation recorded in an owner's list (plans and ragged plans).
int dev_alloc_list(std::vector<void*>& owned, void** out, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    hipError_t e = hipMalloc(out, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(GFFTX_ERR_ALLOC, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
    }
    owned.push_back(*out);
    return GFFTX_SUCCESS;
}

int upload_stage_table_r(std::vector<void*>& owned, int prec, int nrad, const int* radices, void** out)
{
    *out = nullptr;
    if (nrad <= 1) return GFFTX_SUCCESS;
    if (prec == PREC_F64) {
        std::vector<double> t;
        pack_stage_table(nrad, radices, t);
        int rc = dev_alloc_list(owned, out, t.size() * sizeof(double));
        if (rc) return rc;
        HIP_TRY(hipMemcpy(*out, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> t;
        pack_stage_table(nrad, radices, t);
        int rc = dev_alloc_list(owned, out, t.size() * sizeof(float));
        if (rc) return rc;
        HIP_TRY(hipMemcpy(*out, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return GFFTX_SUCCESS;
}

int upload_stage_table(gfftx_plan_s* p, const TileEntry* k, void** out)
{
    return upload_stage_table_r(p->allocs, p->prec, k->nrad, k->radices, out);
}

// FP64 two-level table for W_M^m: lo[m mod 2^b], hi[m >> b], 2^b ~ sqrt(M).
int upload_root_tables(std::vector<void*>& owned, long long M, double2** lo_out, double2** hi_out, int* shift)
{
    int b = 0;
    while ((1LL << (2 * b)) < M) ++b;
    const long long nlo = 1LL << b, nhi = (M + nlo - 1) / nlo + 1;
    std::vector<double> lo(2 * nlo), hi(2 * nhi);
    for (long long m = 0; m < nlo; ++m) {
        long double c, s;
        root_ld(m, M, &c, &s);
        lo[2 * m] = (double)c;
        lo[2 * m + 1] = (double)s;
    }
    for (long long h = 0; h < nhi; ++h) {
        long double c, s;
        root_ld(h * nlo, M, &c, &s);
        hi[2 * h] = (double)c;
        hi[2 * h + 1] = (double)s;
    }
    void* dlo = nullptr;
    void* dhi = nullptr;
    int rc = dev_alloc_list(owned, &dlo, lo.size() * sizeof(double));
    if (rc) return rc;
    rc = dev_alloc_list(owned, &dhi, hi.size() * sizeof(double));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(dlo, lo.data(), lo.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(dhi, hi.data(), hi.size() * sizeof(double), hipMemcpyHostToDevice));
    *lo_out = (double2*)dlo;
    *hi_out = (double2*)dhi;
    *shift = b;
    return GFFTX_SUCCESS;
}

// A operands of the MFMA radix-16 stage (mfma_kernel.h): [mt][hi/lo][lane][8] halves.
int upload_mfma_matrix(std::vector<void*>& owned, void** out)
{
    std::vector<_Float16> h(MFMA_A_HALFS);
    for (int mt = 0; mt < 2; ++mt)
        for (int lane = 0; lane < 64; ++lane)
            for (int i = 0; i < 8; ++i) {
                const int q = lane % 16, k = 8 * (lane / 16) + i, r = k / 2, im = k % 2;
                long double c, sn;
                root_ld((long long)q * r, 16, &c, &sn);  // F[q][r] = W16^{qr} = c + i sn
                long double val;
                if (mt == 0) val = im ? -sn : c;   // Re Y row: Re F, -Im F
                else val = im ? c : sn;            // Im Y row: Im F,  Re F
                const _Float16 hi = (_Float16)(double)val;
                const _Float16 lo = (_Float16)(double)(val - (long double)(double)hi);
                h[((mt * 2 + 0) * 64 + lane) * 8 + i] = hi;
                h[((mt * 2 + 1) * 64 + lane) * 8 + i] = lo;
            }
    int rc = dev_alloc_list(owned, out, h.size() * sizeof(_Float16));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*out, h.data(), h.size() * sizeof(_Float16), hipMemcpyHostToDevice));
    return GFFTX_SUCCESS;
}

int upload_pass_table(gfftx_plan_s* p, Pass& ps, long long M)
{
    ps.ptw_m = M;
    return upload_root_tables(p->allocs, M, &ps.ptw_lo, &ps.ptw_hi, &ps.ptw_shift);
}

int build_c2c_passes(gfftx_plan_s* p)
{
    const unsigned long long n = p->n;
    char buf[256];
    p->desc = "{\"n\":" + std::to_string(n) + ",\"batch\":" + std::to_string(p->batch) +
              ",\"precision\":\"" + (p->prec == PREC_F64 ? "f64" : "f32") + "\",\"passes\":[";
    if (p->flags & GFFTX_FLAG_SPLIT_MFMA) {
        const MfmaEntry* me = nullptr;
        for (int i = 0; i < kNumMfma; ++i)
            if ((unsigned long long)kMfma[i].L == n) me = &kMfma[i];
        if (p->prec != PREC_F32 || p->real || !me)
            return fail(GFFTX_ERR_UNSUPPORTED, "split-precision MFMA path: FP32 C2C of length 256 or 4096 only (n=%llu)", n);
        Pass ps;
        ps.mfma = me;
        ps.items_in_ntrans = true;
        ps.ntrans = (long long)p->batch;
        const int rad[4] = {16, 16, 16, 16};
        int rc = upload_stage_table_r(p->allocs, PREC_F32, me->L == 256 ? 2 : 3, rad, &ps.stage_tw);
        if (rc) return rc;
        rc = upload_mfma_matrix(p->allocs, &ps.amat);
        if (rc) return rc;
        p->passes.push_back(ps);
        snprintf(buf, sizeof(buf), "{\"kernel\":\"%s\",\"L\":%d,\"ntrans\":%lld}", me->name, me->L, ps.ntrans);
        p->desc += buf;
        p->desc += "]}";
        return GFFTX_SUCCESS;
    }
    const TileEntry* single = n <= (1ull << 30) ? find_tile(p->prec, (int)n, false, false) : nullptr;
    if (single) {
        const TileEntry* k = single;
        Pass ps;
        ps.k = k;
        ps.items_in_ntrans = true;
        ps.ntrans = (long long)p->batch;
        ps.its = (long long)p->dist;
        ps.ies = 1;
        ps.ns_pass = 1;
        ps.bdist = 0;
        ps.batch = 1;
        int rc = upload_stage_table(p, k, &ps.stage_tw);
        if (rc) return rc;
        p->passes.push_back(ps);
        snprintf(buf, sizeof(buf), "{\"kernel\":\"%s\",\"L\":%d,\"ntrans\":%lld}", k->name, k->L, ps.ntrans);
        p->desc += buf;
    } else {
        std::vector<int> lens;
        std::vector<std::pair<int, int>> tiles;
        split_choice(p->forced_lens, lens, tiles);
        if (lens.empty() && !choose_passes(p->prec, n, lens, p->batch, device_cus(p->device)))
            return fail(GFFTX_ERR_UNSUPPORTED, "no multi-pass decomposition for n=%llu", n);
        long long ns = 1;
        for (size_t i = 0; i < lens.size(); ++i) {
            const int L = lens[i];
            const int want_nb = i < tiles.size() && tiles[i].first ? tiles[i].first : env_int("GFFTX_TNB", 0);
            const int want_wg = i < tiles.size() && tiles[i].first ? tiles[i].second : env_int("GFFTX_TWG", 0);
            const TileEntry* k = find_tile(p->prec, L, true, i > 0, want_nb, want_wg);
            if (!k) return fail(GFFTX_ERR_UNSUPPORTED, "no pass kernel L=%d", L);
            Pass ps;
            ps.k = k;
            ps.xcd = env_int("GFFTX_XCD", 1);
            ps.ntrans = (long long)(n / L);
            ps.its = 1;
            ps.ies = (long long)(n / L);
            ps.ns_pass = ns;
            ps.bdist = (long long)p->dist;
            ps.batch = (unsigned)p->batch;
            int rc = upload_stage_table(p, k, &ps.stage_tw);
            if (rc) return rc;
            if (ns > 1) {
                rc = upload_pass_table(p, ps, ns * L);
                if (rc) return rc;
            }
            p->passes.push_back(ps);
            // real forward transform: the last pass also computes the R2C split when its
            // tile has a fused variant
            if (i + 1 == lens.size() && p->real && !p->inverse && env_int("GFFTX_FUSED_SPLIT", 1)) {
                if (const TileEntry* ks = find_split_tile(k)) {
                    Pass& last = p->passes.back();
                    last.k = ks;
                    last.split = true;
                    rc = upload_root_tables(p->allocs, (long long)(2 * n), &last.sp_lo, &last.sp_hi, &last.sp_shift);
                    if (rc) return rc;
                    p->split_fused = true;
                }
            }
            snprintf(buf, sizeof(buf), "%s{\"kernel\":\"%s\",\"L\":%d,\"ns\":%lld,\"ntrans\":%lld,\"xcd\":%d}",
                     i ? "," : "", p->passes.back().k->name, L, ns, ps.ntrans, ps.xcd);
            p->desc += buf;
            ns *= L;
        }
    }
    p->desc += "]" + p->tune_note + "}";
    return GFFTX_SUCCESS;
}

int upload_split_table(gfftx_plan_s* p)
{
    const long long n = (long long)p->n;
    const long long cnt = n / 2 + 1;
    if (p->prec == PREC_F64) {
        std::vector<double> t(2 * cnt);
        for (long long k = 0; k < cnt; ++k) {
            long double c, s;
            root_ld(k, 2 * n, &c, &s);
            t[2 * k] = (double)c;
            t[2 * k + 1] = (double)s;
        }
        int rc = dev_alloc(p, &p->split_tw, t.size() * sizeof(double));
        if (rc) return rc;
        HIP_TRY(hipMemcpy(p->split_tw, t.data(), t.size() * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> t(2 * cnt);
        for (long long k = 0; k < cnt; ++k) {
            long double c, s;
            root_ld(k, 2 * n, &c, &s);
            t[2 * k] = (float)c;
            t[2 * k + 1] = (float)s;
        }
        int rc = dev_alloc(p, &p->split_tw, t.size() * sizeof(float));
        if (rc) return rc;
        HIP_TRY(hipMemcpy(p->split_tw, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
    }
    return GFFTX_SUCCESS;
}

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

bool is_device_ptr(const void* ptr)
{
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, ptr);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

// Event pair for launch slot `slot` (GFFTX_FLAG_PROFILE); returns false when not profiling.
bool prof_begin(gfftx_plan_s* p, size_t slot, const char* name, hipStream_t s, std::pair<hipEvent_t, hipEvent_t>* ev)
{
    if (!(p->flags & GFFTX_FLAG_PROFILE) || p->capturing) return false;
    if (p->slots.size() <= slot) p->slots.resize(slot + 1);
    gfftx_plan_s::Slot& sl = p->slots[slot];
    sl.name = name;
    if (sl.pool.empty()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        sl.pool.push_back({a, b});
    }
    *ev = sl.pool.back();
    sl.pool.pop_back();
    (void)hipEventRecord(ev->first, s);
    return true;
}

void prof_end(gfftx_plan_s* p, size_t slot, hipStream_t s, const std::pair<hipEvent_t, hipEvent_t>& ev)
{
    (void)hipEventRecord(ev.second, s);
    p->slots[slot].pending.push_back(ev);
}

// Runs the pass list of `p` on batch items `r` of the arrays `in`/`out` (base pointers of
// the whole batch).  Profiling events go to the slots of `prof` (the plan itself, or the
// plan that owns `p` as an inner plan) starting at `slot0`.
int run_passes(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, gfftx_plan_s* prof = nullptr,
               size_t slot0 = 0)
{
    if (!prof) prof = p;
    if (!r.nb) r.nb = p->batch;
    const size_t off = (size_t)r.b0 * p->dist * elem_bytes(p->prec);
    auto at = [off](const void* q) { return (void*)((char*)const_cast<void*>(q) + off); };
    const size_t m = p->passes.size();
    // buffer for the output of pass i (i < m-1): alternate so that no pass is in place and
    // the caller's input is never written unless it is also the output
    std::vector<void*> dst(m);
    dst[m - 1] = out;
    for (size_t i = m - 1; i-- > 0;) {
        void* next_src_dst = dst[i + 1];
        void* cand = (next_src_dst == p->scratch[0]) ? out : p->scratch[0];
        if (cand == in || cand == next_src_dst) cand = p->scratch[1];
        dst[i] = cand;
    }
    const void* src = in;
    for (size_t i = 0; i < m; ++i) {
        const Pass& ps = p->passes[i];
        if (ps.mfma) {
            MfmaArgs ma;
            ma.in = (const float2*)at(src);
            ma.out = (float2*)at(dst[i]);
            ma.ntrans = (long long)r.nb;
            ma.tw = (const float2*)ps.stage_tw;
            ma.amat = (const _Float16*)ps.amat;
            ma.conj_in = p->inverse ? 1 : 0;
            ma.conj_out = p->inverse ? 1 : 0;
            ma.scale = p->inverse ? (float)(1.0 / (double)p->n) : 1.0f;
            dim3 grid((unsigned)((ps.ntrans + ps.mfma->nb - 1) / ps.mfma->nb));
            std::pair<hipEvent_t, hipEvent_t> ev;
            const bool on = prof_begin(prof, slot0 + i, ps.mfma->name, s, &ev);
            hipError_t e = ps.mfma->launch(ma, grid, s);
            if (on) prof_end(prof, slot0 + i, s, ev);
            if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "%s: %s", ps.mfma->name, hipGetErrorString(e));
            src = dst[i];
            continue;
        }
        TileArgs a;
        memset(&a, 0, sizeof(a));
        a.in = at(src);
        a.out = at(dst[i]);
        a.ntrans = ps.items_in_ntrans ? (long long)r.nb : ps.ntrans;
        a.its = ps.its;
        a.ies = ps.ies;
        a.ns_pass = ps.ns_pass;
        a.bdist = ps.bdist;
        a.tw = ps.stage_tw;
        a.ptw_lo = ps.ptw_lo;
        a.ptw_hi = ps.ptw_hi;
        a.ptw_shift = ps.ptw_shift;
        a.conj_in = (p->inverse && i == 0) ? 1 : 0;
        a.conj_out = (p->inverse && i == m - 1) ? 1 : 0;
        a.scale = (p->inverse && i == m - 1) ? 1.0 / (double)p->n : 1.0;
        a.xcd = ps.xcd;
        const long long nb = ps.k->nb;
        unsigned ntiles = (unsigned)((a.ntrans + nb - 1) / nb);
        if (ps.split) {
            // tile 0: sequences 0 and S/2; then nb/2 pairs (p, S - p), 1 <= p <= (S-1)/2
            a.sp_lo = ps.sp_lo;
            a.sp_hi = ps.sp_hi;
            a.sp_shift = ps.sp_shift;
            a.split_bins = (p->flags & GFFTX_FLAG_R2C_BINS) ? 1 : 0;
            a.nfull = (long long)p->n;
            const long long pairs = (a.ntrans - 1) / 2, h = nb / 2;
            ntiles = (unsigned)(1 + (pairs + h - 1) / h);
        }
        dim3 grid(ntiles, ps.items_in_ntrans ? 1u : (unsigned)r.nb);
        std::pair<hipEvent_t, hipEvent_t> ev;
        const bool on = prof_begin(prof, slot0 + i, ps.k->name, s, &ev);
        hipError_t e = ps.k->launch(a, grid, s);
        if (on) prof_end(prof, slot0 + i, s, ev);
        if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "%s: %s", ps.k->name, hipGetErrorString(e));
        src = dst[i];
    }
    return GFFTX_SUCCESS;
}

// Launch slots the complex transform of `p` uses (profiling); the split pass comes next.
size_t c2c_slots(const gfftx_plan_s* p)
{
    if (!p->blue_fwd) return p->passes.size();
    return 3 + p->blue_fwd->passes.size() + p->blue_inv->passes.size();
}

int blue_launch(gfftx_plan_s* p, int which, const BluesteinArgs& a, size_t slot, const char* name, hipStream_t s)
{
    std::pair<hipEvent_t, hipEvent_t> ev;
    const bool on = prof_begin(p, slot, name, s, &ev);
    hipError_t e = launch_bluestein(p->prec, which, a, s);
    if (on) prof_end(p, slot, s, ev);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "%s: %s", name, hipGetErrorString(e));
    return GFFTX_SUCCESS;
}

// Chirp-z transform: pre-multiply into the zero-padded work buffer, forward FFT of length m,
// multiply by the filter spectrum, inverse FFT, post-multiply into `out`.
// `boff` (optional) gives the element offset of every batch item in `in` and `out`; the
// range must then cover the whole batch.
int run_bluestein(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, const long long* boff = nullptr)
{
    if (!r.nb) r.nb = p->batch;
    const size_t eb = elem_bytes(p->prec);
    BluesteinArgs a;
    memset(&a, 0, sizeof(a));
    a.boff = boff;
    a.in = (const char*)in + (boff ? 0 : (size_t)r.b0 * p->dist * eb);
    a.out = (char*)out + (boff ? 0 : (size_t)r.b0 * p->dist * eb);
    a.work = (char*)p->blue_work + (size_t)r.b0 * p->blue_m * eb;
    a.chirp = p->blue_chirp;
    a.filter = p->blue_filter;
    a.n = (long long)p->n;
    a.m = (long long)p->blue_m;
    a.dist = (long long)p->dist;
    a.batch = (long long)r.nb;
    a.conj = p->inverse ? 1 : 0;
    a.scale = p->inverse ? 1.0 / (double)p->n : 1.0;
    size_t slot = 0;
    int rc = blue_launch(p, BLUESTEIN_PRE, a, slot++, "bluestein_pre", s);
    if (rc) return rc;
    rc = run_passes(p->blue_fwd, p->blue_work, p->blue_work, s, r, p, slot);
    if (rc) return rc;
    slot += p->blue_fwd->passes.size();
    rc = blue_launch(p, BLUESTEIN_MUL, a, slot++, "bluestein_mul", s);
    if (rc) return rc;
    rc = run_passes(p->blue_inv, p->blue_work, p->blue_work, s, r, p, slot);
    if (rc) return rc;
    slot += p->blue_inv->passes.size();
    return blue_launch(p, BLUESTEIN_POST, a, slot, "bluestein_post", s);
}

int run_c2c(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r)
{
    return p->blue_fwd ? run_bluestein(p, in, out, s, r) : run_passes(p, in, out, s, r);
}

int run_split(gfftx_plan_s* p, void* data, hipStream_t s, Range r)
{
    if (!r.nb) r.nb = p->batch;
    SplitArgs a;
    a.data = (char*)data + (size_t)r.b0 * p->dist * elem_bytes(p->prec);
    a.n = (long long)p->n;
    a.bdist = (long long)p->dist;
    a.nbatch = (long long)r.nb;
    a.tw = p->split_tw;
    a.inverse = p->inverse ? 1 : 0;
    a.bins = (p->flags & GFFTX_FLAG_R2C_BINS) ? 1 : 0;
    std::pair<hipEvent_t, hipEvent_t> ev;
    const size_t slot = c2c_slots(p);
    const bool prof = prof_begin(p, slot, "split_kernel", s, &ev);
    hipError_t e = launch_split(p->prec, a, s);
    if (prof) prof_end(p, slot, s, ev);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "split: %s", hipGetErrorString(e));
    return GFFTX_SUCCESS;
}

// Full transform on device pointers.  Real forward: complex FFT then split on the output.
// Real inverse: split on a copy of the input (the input of an out-of-place call is left
// unchanged), then the inverse complex FFT.
// `r` selects batch items (all by default); in and out are the base pointers of the batch.
int exec_device(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r = Range())
{
    if (!r.nb) r.nb = p->batch;
    if (!p->real) return run_c2c(p, in, out, s, r);
    if (!p->inverse) {
        int rc = run_c2c(p, in, out, s, r);
        if (rc) return rc;
        return p->split_fused ? GFFTX_SUCCESS : run_split(p, out, s, r);
    }
    void* work = out;
    if (in != out) {
        work = p->real_work;
        const size_t eb = elem_bytes(p->prec), off = (size_t)r.b0 * p->dist * eb;
        HIP_TRY(hipMemcpyAsync((char*)work + off, (const char*)in + off, (size_t)p->dist * r.nb * eb,
                               hipMemcpyDeviceToDevice, s));
    }
    int rc = run_split(p, work, s, r);
    if (rc) return rc;
    return run_c2c(p, work, out, s, r);
}

// Replays the plan's launches as one captured hipGraph, one per (in, out) pair (the last
// eight kept).  Capture happens on a plan-owned stream; the graph is launched on the caller's.
int exec_graph(gfftx_plan_s* p, const void* in, void* out, hipStream_t s)
{
    GraphCache::Entry* g = p->graphs.find(in, out);
    if (!g) {
        if (!p->graphs.cap_stream) HIP_TRY(hipStreamCreateWithFlags(&p->graphs.cap_stream, hipStreamNonBlocking));
        HIP_TRY(hipStreamBeginCapture(p->graphs.cap_stream, hipStreamCaptureModeThreadLocal));
        p->capturing = true;
        const int rc = exec_device(p, in, out, p->graphs.cap_stream);
        p->capturing = false;
        hipGraph_t gr = nullptr;
        const hipError_t ce = hipStreamEndCapture(p->graphs.cap_stream, &gr);
        if (rc) {
            if (gr) (void)hipGraphDestroy(gr);
            return rc;
        }
        if (ce != hipSuccess) return fail(GFFTX_ERR_DEVICE, "end capture: %s", hipGetErrorString(ce));
        g = p->graphs.take();
        g->graph = gr;
        const hipError_t ie = hipGraphInstantiate(&g->exec, gr, nullptr, nullptr, 0);
        if (ie != hipSuccess) {
            GraphCache::release(*g);
            return fail(GFFTX_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
        }
        g->in = in;
        g->out = out;
        g->used = ++p->graphs.tick;
    }
    HIP_TRY(hipGraphLaunch(g->exec, s));
    return GFFTX_SUCCESS;
}

// Copy between pageable and pinned host memory with up to 8 threads (chunks of 1 MiB or more).
void host_copy(void* dst, const void* src, size_t len)
{
    unsigned t = std::thread::hardware_concurrency();
    t = std::min<unsigned>(std::max(t, 1u), 8u);
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

int exec_any(gfftx_plan_s* p, const void* in, void* out, hipStream_t s)
{
    DeviceGuard g(p->device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", p->device);
    const bool din = is_device_ptr(in), dout = is_device_ptr(out);
    if (din && dout && (p->flags & GFFTX_FLAG_GRAPH)) return exec_graph(p, in, out, s);
    // complex elements the caller's arrays hold: the real side of a real transform has n
    // per item, the complex side dist (n, or n+1 with the bins layout)
    const size_t full = (size_t)p->dist * p->batch, real_side = (size_t)p->dist * (p->batch - 1) + p->n;
    const size_t eb = elem_bytes(p->prec);
    size_t in_bytes = full * eb, out_bytes = full * eb;
    if (p->real && in != out) {
        if (p->inverse) out_bytes = real_side * eb;
        else in_bytes = real_side * eb;
    }
    if (din && dout) return exec_device(p, in, out, s);
    return exec_host(p, in, out, s, din, dout, in_bytes, out_bytes);
}

int create_plan(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                uint32_t flags);
int create_plan_impl(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                     uint32_t flags, const std::vector<int>& forced, const std::string& note);

// Largest complex length the chirp-z path takes: its convolution length m < 2n must stay
// within the multi-pass limit of the precision.
unsigned long long bluestein_max_m(int prec) { return prec == PREC_F64 ? (1ull << 26) : (1ull << 28); }

// Convolution length for the chirp-z path: the 2/3/5-smooth m >= 2n-1 (up to twice that)
// with the lowest modelled cost, m times the sum over its passes of 1/bandwidth (a
// single-pass tile counts as one pass at 4 TB/s).  0 when none can be planned.
unsigned long long bluestein_length(int prec, unsigned long long n, unsigned long long batch)
{
    const unsigned long long lo = 2 * n - 1, hi = 2 * lo, maxm = bluestein_max_m(prec);
    std::vector<unsigned long long> cand;
    for (unsigned long long a = 1; a <= hi; a *= 2)
        for (unsigned long long b = a; b <= hi; b *= 3)
            for (unsigned long long c = b; c <= hi; c *= 5)
                if (c >= lo && c <= maxm) cand.push_back(c);
    std::sort(cand.begin(), cand.end());
    unsigned long long best = 0;
    double best_cost = 0;
    for (unsigned long long m : cand) {
        double per = 0;
        if (m <= (1ull << 30) && find_tile(prec, (int)m, false, false)) {
            per = 1.0 / 4.0;
        } else {
            const std::vector<std::vector<int>> c = candidate_passes(prec, m, 1, batch);
            if (c.empty()) continue;
            for (int L : c[0]) per += 1.0 / pass_bandwidth(prec, L);
        }
        const double cost = per * (double)m;
        if (!best || cost < best_cost) {
            best = m;
            best_cost = cost;
        }
    }
    return best;
}

// Moves the scratch buffers of `src` into `dst` where both have one (same length, batch and
// precision): the two inner plans of a chirp-z plan run one after the other on one stream.
void share_scratch(gfftx_plan_s* dst, gfftx_plan_s* src)
{
    for (int k = 0; k < 2; ++k) {
        if (!dst->scratch[k] || !src->scratch[k]) continue;
        auto it = std::find(dst->allocs.begin(), dst->allocs.end(), dst->scratch[k]);
        if (it != dst->allocs.end()) dst->allocs.erase(it);
        (void)hipFree(dst->scratch[k]);
        dst->scratch[k] = src->scratch[k];
    }
}

int build_bluestein(gfftx_plan_s* p)
{
    const unsigned long long n = p->n;
    if (p->flags & GFFTX_FLAG_SPLIT_MFMA)
        return fail(GFFTX_ERR_UNSUPPORTED, "split-precision MFMA path: FP32 C2C of length 256 or 4096 only (n=%llu)", n);
    if (n < 2 || 2 * n - 1 > bluestein_max_m(p->prec))
        return fail(GFFTX_ERR_UNSUPPORTED, "length %llu with a prime factor above 5: the chirp-z path takes 2 <= n <= %llu",
                    n, (bluestein_max_m(p->prec) + 1) / 2);
    const unsigned long long m = bluestein_length(p->prec, n, p->batch);
    if (!m) return fail(GFFTX_ERR_UNSUPPORTED, "no convolution length for n=%llu", n);
    p->blue_m = m;
    const int cvt = p->prec == PREC_F64 ? GFFTX_COMPLEX_DOUBLE : GFFTX_COMPLEX_FLOAT;
    const uint32_t iflags = p->flags & GFFTX_FLAG_NO_AUTOTUNE;
    gfftx_plan f = nullptr, b = nullptr;
    int rc = create_plan(&f, m, p->batch, cvt, GFFTX_DFT, GFFTX_IN_PLACE, p->device, iflags);
    if (rc) return rc;
    p->blue_fwd = f;
    rc = create_plan(&b, m, p->batch, cvt, GFFTX_IDFT, GFFTX_IN_PLACE, p->device, iflags);
    if (rc) return rc;
    p->blue_inv = b;
    share_scratch(b, f);
    const size_t eb = elem_bytes(p->prec);
    rc = dev_alloc(p, &p->blue_work, (size_t)m * p->batch * eb);
    if (rc) return rc;
    rc = dev_alloc(p, &p->blue_chirp, (size_t)n * eb);
    if (rc) return rc;
    rc = dev_alloc(p, &p->blue_filter, (size_t)m * eb);
    if (rc) return rc;
    // chirp from the two-level root table of order 2n
    ChirpTableArgs ca;
    memset(&ca, 0, sizeof(ca));
    ca.chirp = p->blue_chirp;
    ca.n = (long long)n;
    rc = upload_root_tables(p->allocs, (long long)(2 * n), const_cast<double2**>(&ca.lo),
                            const_cast<double2**>(&ca.hi), &ca.shift);
    if (rc) return rc;
    hipError_t e = launch_chirp_table(p->prec, ca, nullptr);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "chirp table: %s", hipGetErrorString(e));
    // filter spectrum: b in the work buffer, one forward transform of length m into the filter
    BluesteinArgs ba;
    memset(&ba, 0, sizeof(ba));
    ba.work = p->blue_work;
    ba.chirp = p->blue_chirp;
    ba.n = (long long)n;
    ba.m = (long long)m;
    ba.batch = 1;
    e = launch_bluestein(p->prec, BLUESTEIN_FILTER, ba, nullptr);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "chirp filter: %s", hipGetErrorString(e));
    gfftx_plan fp = nullptr;
    rc = create_plan_impl(&fp, m, 1, cvt, GFFTX_DFT, GFFTX_OUT_OF_PLACE, p->device, 0, {}, "");
    if (rc) return rc;
    rc = exec_device(fp, p->blue_work, p->blue_filter, nullptr);
    const hipError_t se = hipStreamSynchronize(nullptr);
    gfftx_plan_destroy(fp);
    if (rc) return rc;
    if (se != hipSuccess) return fail(GFFTX_ERR_DEVICE, "filter transform: %s", hipGetErrorString(se));
    p->desc = "{\"n\":" + std::to_string(n) + ",\"batch\":" + std::to_string(p->batch) + ",\"precision\":\"" +
              (p->prec == PREC_F64 ? "f64" : "f32") + "\",\"passes\":[],\"bluestein\":{\"m\":" + std::to_string(m) +
              ",\"fwd\":" + f->desc + ",\"inv\":" + b->desc + "}}";
    return GFFTX_SUCCESS;
}

int create_plan_impl(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                     uint32_t flags, const std::vector<int>& forced, const std::string& note)
{
    if (!out) return fail(GFFTX_ERR_INVALID_ARG, "plan pointer is NULL");
    *out = nullptr;
    if (n == 0 || batch == 0) return fail(GFFTX_ERR_INVALID_ARG, "n and batch must be positive");
    if (vtype < 0 || vtype > 3) return fail(GFFTX_ERR_INVALID_ARG, "unknown value type %d", vtype);
    if (ttype < 0 || ttype > 3) return fail(GFFTX_ERR_INVALID_ARG, "unknown transform type %d", ttype);
    if (place != GFFTX_IN_PLACE && place != GFFTX_OUT_OF_PLACE)
        return fail(GFFTX_ERR_INVALID_ARG, "unknown place %d", place);
    if (flags & ~(GFFTX_FLAG_R2C_BINS | GFFTX_FLAG_SPLIT_MFMA | GFFTX_FLAG_NO_AUTOTUNE | GFFTX_FLAG_GRAPH |
                  GFFTX_FLAG_PROFILE))
        return fail(GFFTX_ERR_INVALID_ARG, "unknown flags 0x%x", flags);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(GFFTX_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(GFFTX_ERR_DEVICE, "device %d out of range (%d)", device, ndev);
    const bool real = ttype == GFFTX_RDFT || ttype == GFFTX_IRDFT;
    if (real && (n % 2))
        return fail(GFFTX_ERR_UNSUPPORTED, "real transforms need an even complex length (n=%llu)",
                    (unsigned long long)n);
    if ((flags & GFFTX_FLAG_R2C_BINS) && !real)
        return fail(GFFTX_ERR_INVALID_ARG, "GFFTX_FLAG_R2C_BINS applies to RDFT/IRDFT only");
    if ((flags & GFFTX_FLAG_R2C_BINS) && batch != 1)
        return fail(GFFTX_ERR_UNSUPPORTED, "the bins layout is available for batch 1 only");

    DeviceGuard g(device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", device);
    gfftx_plan_s* p = new (std::nothrow) gfftx_plan_s();
    if (!p) return fail(GFFTX_ERR_ALLOC, "host allocation failed");
    p->device = device;
    p->n = n;
    p->batch = batch;
    p->vtype = vtype;
    p->ttype = ttype;
    p->place = place;
    p->flags = flags;
    p->prec = (vtype == GFFTX_DOUBLE || vtype == GFFTX_COMPLEX_DOUBLE) ? PREC_F64 : PREC_F32;
    p->real = real;
    p->inverse = ttype == GFFTX_IDFT || ttype == GFFTX_IRDFT;
    p->dist = n + ((real && (flags & GFFTX_FLAG_R2C_BINS)) ? 1 : 0);
    p->forced_lens = forced;
    p->tune_note = note;
    int rc = native_length(p->prec, n) ? build_c2c_passes(p) : build_bluestein(p);
    if (rc == GFFTX_SUCCESS && real) rc = upload_split_table(p);
    if (rc == GFFTX_SUCCESS && real && p->inverse)  // split pass of an out-of-place IRDFT works on a copy
        rc = dev_alloc(p, &p->real_work, (size_t)p->dist * batch * elem_bytes(p->prec));
    if (rc == GFFTX_SUCCESS && p->passes.size() > 1) {
        const size_t bytes = (size_t)p->dist * batch * elem_bytes(p->prec);
        rc = dev_alloc(p, &p->scratch[0], bytes);
        if (rc == GFFTX_SUCCESS && p->passes.size() > 2) rc = dev_alloc(p, &p->scratch[1], bytes);
    }
    if (rc != GFFTX_SUCCESS) {
        gfftx_plan_destroy(p);  // frees inner plans too; leaves the error message alone
        return rc;
    }
    *out = p;
    return GFFTX_SUCCESS;
}

}  // namespace


// ---------------------------------------------------------------------------------------
// Ragged batches: one persistent launch over a list of work items (ragged.hip).
// Side streams for the long length classes: with the caller's stream they fill the four
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
    // Grouped mode (default): each length class above 4096 points runs as its own batched
    // launch sequence (a single-pass tile through the sequence-offset table, or multi-pass
    // tiles through a batch-offset table) on a side stream, while the persistent kernel
    // takes the shorter lengths on the caller's stream.
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
};

namespace {

int ragged_single_cfg(long long L)
{
    for (int id = 0; id < ragged_num_cfgs(); ++id) {
        const RaggedCfgInfo c = ragged_cfg_info(id);
        if (!c.loadT && !c.storeT && c.L == L) return id;
    }
    return -1;
}

int ragged_pass1_cfg(long long P)
{
    for (int id = 0; id < ragged_num_cfgs(); ++id) {
        const RaggedCfgInfo c = ragged_cfg_info(id);
        if (!c.loadT && c.storeT && c.L == P) return id;
    }
    return -1;
}

// A length class above the persistent kernel's tiles: the transforms' offsets on the
// device, and the pass list that computes all of them in one batched launch per pass.
// A single-pass tile (one workgroup per transform, 16 bytes of HBM traffic per point)
// is used when one exists and the class has at least `min_single` transforms; otherwise
// multi-pass tiles (many workgroups per transform, 32 bytes per point).  The classes run
// concurrently, so the lower traffic wins even for classes too small to fill the device
// (measured on the 4096-transform batch).  `nscratch` is raised to the scratch arrays the
// pass list needs.
int ragged_long_class(gfftx_ragged_s* r, long long L, const std::vector<long long>& offs, long long min_single,
                      int* nscratch)
{
    gfftx_ragged_s::LongClass lc;
    lc.L = L;
    lc.count = (long long)offs.size();
    void* d = nullptr;
    int e = dev_alloc_list(r->allocs, &d, offs.size() * sizeof(long long));
    if (e) return e;
    HIP_TRY(hipMemcpy(d, offs.data(), offs.size() * sizeof(long long), hipMemcpyHostToDevice));
    lc.offs = (const long long*)d;
    if (!native_length(r->prec, (unsigned long long)L)) {
        // chirp-z plan for the whole class; its element-wise kernels address the transforms
        // through the offset table
        gfftx_plan bp = nullptr;
        const int cvt = r->prec == PREC_F64 ? GFFTX_COMPLEX_DOUBLE : GFFTX_COMPLEX_FLOAT;
        e = create_plan(&bp, (uint64_t)L, (uint64_t)lc.count, cvt, r->inverse ? GFFTX_IDFT : GFFTX_DFT,
                        GFFTX_OUT_OF_PLACE, r->device, r->flags & GFFTX_FLAG_NO_AUTOTUNE);
        if (e) return e;
        lc.blue = bp;
        r->longs.push_back(std::move(lc));
        return GFFTX_SUCCESS;
    }
    std::vector<int> lens;
    const bool multi = choose_passes(r->prec, (unsigned long long)L, lens, (unsigned long long)lc.count);
    const TileEntry* k1 = (L <= (1 << 30)) ? find_tile(r->prec, (int)L, false, false) : nullptr;
    if (k1 && (!multi || lc.count >= min_single)) {
        const TileEntry* k = k1;
        Pass ps;
        ps.k = k;
        ps.items_in_ntrans = true;
        ps.ntrans = lc.count;
        ps.ies = 1;
        e = upload_stage_table_r(r->allocs, r->prec, k->nrad, k->radices, &ps.stage_tw);
        if (e) return e;
        lc.passes.push_back(ps);
    } else {
        if (!multi) return fail(GFFTX_ERR_UNSUPPORTED, "ragged batch: no kernel for length %lld", L);
        long long ns = 1;
        for (size_t i = 0; i < lens.size(); ++i) {
            const TileEntry* k = find_tile(r->prec, lens[i], true, i > 0);
            if (!k) return fail(GFFTX_ERR_UNSUPPORTED, "ragged batch: no pass kernel L=%d", lens[i]);
            Pass ps;
            ps.k = k;
            ps.xcd = 1;
            ps.ntrans = L / lens[i];
            ps.its = 1;
            ps.ies = L / lens[i];
            ps.ns_pass = ns;
            ps.batch = (unsigned)lc.count;
            e = upload_stage_table_r(r->allocs, r->prec, k->nrad, k->radices, &ps.stage_tw);
            if (e) return e;
            if (ns > 1) {
                ps.ptw_m = ns * lens[i];
                e = upload_root_tables(r->allocs, ps.ptw_m, &ps.ptw_lo, &ps.ptw_hi, &ps.ptw_shift);
                if (e) return e;
            }
            lc.passes.push_back(ps);
            ns *= lens[i];
        }
        *nscratch = std::max(*nscratch, lens.size() > 2 ? 2 : 1);  // S0, and S1 from three passes
    }
    r->longs.push_back(std::move(lc));
    return GFFTX_SUCCESS;
}

// Launches of one long class on stream s, in place on `data`.  Pass i reads the output of
// pass i-1 and writes scratch S[i % 2], the last pass writes `data` (data -> S0 -> data,
// data -> S0 -> S1 -> data, ...); the scratch arrays use the data offsets.
int ragged_run_long(gfftx_ragged_s* r, const gfftx_ragged_s::LongClass& lc, void* data, int side, hipStream_t s)
{
    if (lc.blue) return run_bluestein(lc.blue, data, data, s, Range(), lc.offs);
    const size_t m = lc.passes.size();
    const void* src = data;
    for (size_t i = 0; i < m; ++i) {
        const Pass& ps = lc.passes[i];
        void* dst = (i + 1 == m) ? data : r->lscratch[i % 2][side];
        TileArgs a;
        memset(&a, 0, sizeof(a));
        a.in = src;
        a.out = dst;
        src = dst;
        a.tw = ps.stage_tw;
        a.conj_in = (r->inverse && i == 0) ? 1 : 0;
        a.conj_out = (r->inverse && i == m - 1) ? 1 : 0;
        a.scale = (r->inverse && i == m - 1) ? 1.0 / (double)lc.L : 1.0;
        dim3 grid;
        if (ps.items_in_ntrans) {
            a.ntrans = lc.count;
            a.ies = 1;
            a.ns_pass = 1;
            a.seq_off = lc.offs;
            grid = dim3((unsigned)((lc.count + ps.k->nb - 1) / ps.k->nb));
        } else {
            a.ntrans = ps.ntrans;
            a.its = ps.its;
            a.ies = ps.ies;
            a.ns_pass = ps.ns_pass;
            a.boff = lc.offs;
            a.ptw_lo = ps.ptw_lo;
            a.ptw_hi = ps.ptw_hi;
            a.ptw_shift = ps.ptw_shift;
            a.xcd = ps.xcd;
            grid = dim3((unsigned)((ps.ntrans + ps.k->nb - 1) / ps.k->nb), (unsigned)lc.count);
        }
        hipError_t le = ps.k->launch(a, grid, s);
        if (le != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "%s: %s", ps.k->name, hipGetErrorString(le));
    }
    return GFFTX_SUCCESS;
}

// Every launch of a ragged batch on stream s: the long classes on the side streams (forked
// from and joined back into s; each side stream has its own scratch), the persistent
// kernel on s.
int ragged_launch_all(gfftx_ragged_s* r, const RaggedArgs& a, hipStream_t s)
{
    if (!r->longs.empty()) HIP_TRY(hipEventRecord(r->fork_ev, s));
    if (a.nitems0 > 0) {
        hipError_t le = launch_ragged(a, r->grid, r->grid1, s);
        if (le != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "ragged_kernel: %s", hipGetErrorString(le));
    }
    if (r->longs.empty()) return GFFTX_SUCCESS;
    for (int k = 0; k < r->nsides; ++k) HIP_TRY(hipStreamWaitEvent(r->side[k], r->fork_ev, 0));
    for (size_t i = 0; i < r->longs.size(); ++i) {
        const int k = r->side_of[i];
        int rc = ragged_run_long(r, r->longs[i], a.data, k, r->side[k]);
        if (rc) return rc;
    }
    for (int k = 0; k < r->nsides; ++k) {
        HIP_TRY(hipEventRecord(r->join_ev[k], r->side[k]));
        HIP_TRY(hipStreamWaitEvent(s, r->join_ev[k], 0));
    }
    return GFFTX_SUCCESS;
}

int ragged_build(gfftx_ragged_s* r, const uint64_t* lengths, uint64_t count)
{
    // classes by length, in order of first appearance
    std::vector<long long> lens;
    std::vector<std::vector<long long>> members;  // element offsets per class
    std::vector<long long> offsets(count);
    long long off = 0;
    for (uint64_t i = 0; i < count; ++i) {
        const long long L = (long long)lengths[i];
        offsets[i] = off;
        off += L;
        size_t c = 0;
        while (c < lens.size() && lens[c] != L) ++c;
        if (c == lens.size()) {
            lens.push_back(L);
            members.emplace_back();
        }
        members[c].push_back(offsets[i]);
    }
    r->total = (unsigned long long)off;
    std::vector<RaggedClass> classes(lens.size());
    struct HostItem {
        RaggedItem it;
        double cost;
    };
    std::vector<HostItem> items, items1;  // phase 0, phase 1
    // lengths above 4096: batched launches per class (default), or inside the persistent
    // kernel, both passes of a transform by one workgroup (GFFTX_RAGGED_MODE=1) or one
    // launch per pass (GFFTX_RAGGED_MODE=2)
    const int mode = env_int("GFFTX_RAGGED_MODE", 0);
    const bool split = mode == 2;
    int nscratch = 0;
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, r->device));
    long long scratch = 0;
    char buf[160];
    r->desc = "{\"transforms\":" + std::to_string(count) + ",\"points\":" + std::to_string(off) + ",\"classes\":[";
    for (size_t c = 0; c < lens.size(); ++c) {
        const long long L = lens[c];
        RaggedClass& rc = classes[c];
        memset(&rc, 0, sizeof(rc));
        rc.L = L;
        rc.count = (long long)members[c].size();
        const int single = r->prec == PREC_F32 ? ragged_single_cfg(L) : -1;
        if (single < 0 && (mode == 0 || r->prec != PREC_F32)) {
            int e = ragged_long_class(r, L, members[c], env_int("GFFTX_RAGGED_MIN_SINGLE", 0), &nscratch);
            if (e) return e;
            snprintf(buf, sizeof(buf), "%s{\"L\":%lld,\"count\":%lld,\"batched\":[", c ? "," : "", L, rc.count);
            r->desc += buf;
            for (size_t i = 0; i < r->longs.back().passes.size(); ++i)
                r->desc += std::string(i ? ",\"" : "\"") + r->longs.back().passes[i].k->name + "\"";
            r->desc += "]}";
            continue;
        }
        // two-pass split L = A * P: the one whose smaller tile row (NB of either pass)
        // is widest, so that both passes move long contiguous rows
        int p1 = -1, cfg0 = -1;
        long long A = 0;
        for (int a0 : {RAGGED_CFG_T256, RAGGED_CFG_T128, RAGGED_CFG_T64}) {
            const RaggedCfgInfo ca = ragged_cfg_info(a0);
            if (L % ca.L) continue;
            const int cp = ragged_pass1_cfg(L / ca.L);
            if (cp < 0) continue;
            const int w = std::min(ca.nb, ragged_cfg_info(cp).nb);
            if (p1 < 0 || w > std::min(ragged_cfg_info(cfg0).nb, ragged_cfg_info(p1).nb)) {
                p1 = cp;
                cfg0 = a0;
                A = ca.L;
            }
        }
        const double unit = (double)L * std::log2((double)L);
        if (single >= 0) {
            const RaggedCfgInfo ci = ragged_cfg_info(single);
            void* offs = nullptr;
            int e = dev_alloc_list(r->allocs, &offs, members[c].size() * sizeof(long long));
            if (e) return e;
            HIP_TRY(hipMemcpy(offs, members[c].data(), members[c].size() * sizeof(long long), hipMemcpyHostToDevice));
            rc.seq_off = (const long long*)offs;
            void* tw = nullptr;
            e = upload_stage_table_r(r->allocs, PREC_F32, ci.nrad, ci.radices, &tw);
            if (e) return e;
            rc.tw0 = tw;
            const long long ntiles = (rc.count + ci.nb - 1) / ci.nb;
            for (long long t = 0; t < ntiles; ++t) {
                HostItem h;
                memset(&h.it, 0, sizeof(h.it));
                h.it.cfg = single;
                h.it.kind = RAGGED_SINGLE;
                h.it.cls = (int)c;
                h.it.tile = (int)t;
                h.cost = unit * (double)std::min<long long>(ci.nb, rc.count - t * ci.nb);
                items.push_back(h);
            }
            snprintf(buf, sizeof(buf), "%s{\"L\":%lld,\"count\":%lld,\"cfg\":%d,\"nb\":%d}", c ? "," : "", L,
                     rc.count, single, ci.nb);
        } else if (p1 >= 0) {
            const RaggedCfgInfo c0 = ragged_cfg_info(cfg0), c1 = ragged_cfg_info(p1);
            rc.A = (int)A;
            void* tw = nullptr;
            int e = upload_stage_table_r(r->allocs, PREC_F32, c0.nrad, c0.radices, &tw);
            if (e) return e;
            rc.tw0 = tw;
            e = upload_stage_table_r(r->allocs, PREC_F32, c1.nrad, c1.radices, &tw);
            if (e) return e;
            rc.tw1 = tw;
            double2 *lo, *hi;
            e = upload_root_tables(r->allocs, L, &lo, &hi, &rc.ptw_shift);
            if (e) return e;
            rc.ptw_lo = lo;
            rc.ptw_hi = hi;
            const long long P = L / A;
            const int nt0 = (int)((P + c0.nb - 1) / c0.nb), nt1 = (int)((A + c1.nb - 1) / c1.nb);
            for (long long o : members[c]) {
                if (!split) {
                    HostItem h;
                    memset(&h.it, 0, sizeof(h.it));
                    h.it.cfg = cfg0;
                    h.it.cfg1 = p1;
                    h.it.kind = RAGGED_FUSED;
                    h.it.cls = (int)c;
                    h.it.nt0 = nt0;
                    h.it.nt1 = nt1;
                    h.it.off = o;
                    h.it.soff = scratch;
                    h.cost = unit * 1.3;  // two passes, the second through L2
                    items.push_back(h);
                } else {
                    for (int pass = 0; pass < 2; ++pass) {
                        const int nt = pass ? nt1 : nt0;
                        for (int t = 0; t < nt; ++t) {
                            HostItem h;
                            memset(&h.it, 0, sizeof(h.it));
                            h.it.cfg = pass ? p1 : cfg0;
                            h.it.kind = pass ? RAGGED_PASS1 : RAGGED_PASS0;
                            h.it.cls = (int)c;
                            h.it.tile = t;
                            h.it.off = o;
                            h.it.soff = scratch;
                            // points of the tile times log2 of the FFT length it runs
                            h.cost = pass ? (double)std::min<long long>(c1.nb, A - (long long)t * c1.nb) * P *
                                                std::log2((double)P)
                                          : (double)std::min<long long>(c0.nb, P - (long long)t * c0.nb) * A *
                                                std::log2((double)A);
                            (pass ? items1 : items).push_back(h);
                        }
                    }
                }
                scratch += L;
            }
            snprintf(buf, sizeof(buf), "%s{\"L\":%lld,\"count\":%lld,\"split\":[%lld,%lld],\"cfg\":[%d,%d]}", c ? "," : "",
                     L, rc.count, A, P, cfg0, p1);
        } else {
            return fail(GFFTX_ERR_UNSUPPORTED, "ragged batch: no kernel for length %lld", L);
        }
        r->desc += buf;
    }
    // longest processing time first, within each phase
    auto by_cost = [](const HostItem& x, const HostItem& y) { return x.cost > y.cost; };
    std::stable_sort(items.begin(), items.end(), by_cost);
    std::stable_sort(items1.begin(), items1.end(), by_cost);
    std::vector<RaggedItem> host_items;
    host_items.reserve(items.size() + items1.size());
    for (const HostItem& h : items) host_items.push_back(h.it);
    for (const HostItem& h : items1) host_items.push_back(h.it);
    void *d_items = nullptr, *d_classes = nullptr, *d_counter = nullptr, *d_scratch = nullptr;
    // (at least one element each: every class may run as batched launches, leaving the
    // persistent kernel without items)
    int e = dev_alloc_list(r->allocs, &d_items, std::max<size_t>(host_items.size(), 1) * sizeof(RaggedItem));
    if (!e) e = dev_alloc_list(r->allocs, &d_classes, std::max<size_t>(classes.size(), 1) * sizeof(RaggedClass));
    if (!e) e = dev_alloc_list(r->allocs, &d_counter, 2 * sizeof(unsigned long long));
    if (!e) e = dev_alloc_list(r->allocs, &d_scratch, (size_t)std::max<long long>(scratch, 1) * sizeof(float2));
    if (e) return e;
    HIP_TRY(hipMemcpy(d_items, host_items.data(), host_items.size() * sizeof(RaggedItem), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(d_classes, classes.data(), classes.size() * sizeof(RaggedClass), hipMemcpyHostToDevice));
    r->args.items = (const RaggedItem*)d_items;
    r->args.classes = (const RaggedClass*)d_classes;
    r->args.counter = (unsigned long long*)d_counter;
    r->args.scratch = (float2*)d_scratch;
    if (!r->longs.empty()) {
        // classes to side streams, largest work first onto the least loaded stream
        r->nsides = std::max(1, std::min<int>(env_int("GFFTX_RAGGED_STREAMS", kRaggedSides), kRaggedSides));
        r->nsides = std::min<int>(r->nsides, (int)r->longs.size());
        std::vector<size_t> order(r->longs.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        auto work = [&](size_t i) { return (double)r->longs[i].L * (double)r->longs[i].count * std::log2((double)r->longs[i].L); };
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return work(x) > work(y); });
        std::vector<double> load(r->nsides, 0.0);
        r->side_of.assign(r->longs.size(), 0);
        for (size_t i : order) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            r->side_of[i] = k;
            load[k] += work(i);
        }
        for (int k = 0; k < r->nsides; ++k) {
            HIP_TRY(hipStreamCreateWithFlags(&r->side[k], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&r->join_ev[k], hipEventDisableTiming));
            for (int b = 0; b < nscratch; ++b) {
                e = dev_alloc_list(r->allocs, &r->lscratch[b][k], (size_t)off * elem_bytes(r->prec));
                if (e) return e;
            }
        }
        HIP_TRY(hipEventCreateWithFlags(&r->fork_ev, hipEventDisableTiming));
    }
    r->args.nitems0 = (int)items.size();
    r->args.nitems1 = (int)items1.size();
    r->args.inverse = r->inverse ? 1 : 0;
    const int per_cu = env_int("GFFTX_RAGGED_WG_PER_CU", 3);
    const int cus = prop.multiProcessorCount;
    r->grid = std::max(1, std::min<int>((int)items.size(), cus * per_cu));
    r->grid1 = std::max(1, std::min<int>((int)items1.size(), cus * per_cu));
    r->desc += "],\"items\":[" + std::to_string(items.size()) + "," + std::to_string(items1.size()) +
               "],\"grid\":[" + std::to_string(r->grid) + "," + std::to_string(r->grid1) +
               "],\"scratch_points\":" + std::to_string(scratch) + "}";
    return GFFTX_SUCCESS;
}

int ragged_launch(gfftx_ragged_s* r, void* data, hipStream_t s)
{
    RaggedArgs a = r->args;
    a.data = (float2*)data;
    std::pair<hipEvent_t, hipEvent_t> ev{nullptr, nullptr};
    const bool prof = (r->flags & GFFTX_FLAG_PROFILE) != 0;
    if (prof) {
        if (r->pool.empty()) {
            hipEvent_t x, y;
            HIP_TRY(hipEventCreate(&x));
            HIP_TRY(hipEventCreate(&y));
            r->pool.push_back({x, y});
        }
        ev = r->pool.back();
        r->pool.pop_back();
        HIP_TRY(hipEventRecord(ev.first, s));
    }
    if (r->flags & GFFTX_FLAG_GRAPH) {
        GraphCache::Entry* ge = r->graphs.find(data, data);
        if (!ge) {
            if (!r->graphs.cap_stream) HIP_TRY(hipStreamCreateWithFlags(&r->graphs.cap_stream, hipStreamNonBlocking));
            HIP_TRY(hipStreamBeginCapture(r->graphs.cap_stream, hipStreamCaptureModeThreadLocal));
            const int lrc = ragged_launch_all(r, a, r->graphs.cap_stream);
            hipGraph_t g = nullptr;
            hipError_t ce = hipStreamEndCapture(r->graphs.cap_stream, &g);
            if (lrc) {
                if (g) (void)hipGraphDestroy(g);
                return lrc;
            }
            if (ce != hipSuccess) return fail(GFFTX_ERR_DEVICE, "end capture: %s", hipGetErrorString(ce));
            ge = r->graphs.take();
            ge->graph = g;
            const hipError_t ie = hipGraphInstantiate(&ge->exec, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                GraphCache::release(*ge);
                return fail(GFFTX_ERR_DEVICE, "graph instantiate: %s", hipGetErrorString(ie));
            }
            ge->in = data;
            ge->out = data;
            ge->used = ++r->graphs.tick;
        }
        HIP_TRY(hipGraphLaunch(ge->exec, s));
    } else {
        const int lrc = ragged_launch_all(r, a, s);
        if (lrc) return lrc;
    }
    if (prof) {
        HIP_TRY(hipEventRecord(ev.second, s));
        r->pending.push_back(ev);
    }
    return GFFTX_SUCCESS;
}


// ---------------------------------------------------------------------------------------
// Autotuner: for a multi-pass plan, time the best few pass splits of the cost model on this
// device and keep the fastest.  Results are cached per (device, precision, n, batch) in
// memory and in a versioned JSON file ($GFFTX_TUNE_CACHE, default ~/.cache/gfftx/tune.json).
std::mutex g_tune_mu;
std::map<std::string, std::vector<int>> g_tune;
std::string g_tune_loaded;  // cache file the in-memory entries belong to

std::string tune_path()
{
    if (const char* e = getenv("GFFTX_TUNE_CACHE")) return e;
    const char* h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.cache/gfftx/tune.json";
}

// One entry per line: "key": [pass lengths..., 0, nb, wg per pass] (version 1 files hold
// the pass lengths only, which still load)
void tune_load()
{
    const std::string path = tune_path();
    if (g_tune_loaded == path) return;
    g_tune_loaded = path;
    g_tune.clear();
    FILE* f = fopen(path.c_str(), "r");
    if (!f) return;
    char line[512];
    while (fgets(line, sizeof(line), f)) {
        char* q1 = strchr(line, '"');
        if (!q1) continue;
        char* q2 = strchr(q1 + 1, '"');
        char* lb = q2 ? strchr(q2, '[') : nullptr;
        if (!lb) continue;
        const std::string key(q1 + 1, q2);
        if (key == "version" || key == "entries") continue;
        std::vector<int> v;
        char* p = lb + 1;
        while (*p && *p != ']') {
            const long x = strtol(p, &p, 10);
            if (x < 0) break;
            v.push_back((int)x);
            while (*p == ',' || *p == ' ') ++p;
            if (!isdigit((unsigned char)*p) && *p != ']') break;
        }
        std::vector<int> lens;
        std::vector<std::pair<int, int>> tiles;
        split_choice(v, lens, tiles);
        if (lens.size() >= 2) g_tune[key] = v;
    }
    fclose(f);
}

void tune_store()
{
    const std::string path = tune_path();
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) {
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    }
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) return;
    fprintf(f, "{\"version\": 2, \"entries\": {\n");
    size_t i = 0;
    for (const auto& kv : g_tune) {
        fprintf(f, "\"%s\": [", kv.first.c_str());
        for (size_t j = 0; j < kv.second.size(); ++j) fprintf(f, "%s%d", j ? ", " : "", kv.second[j]);
        fprintf(f, "]%s\n", ++i < g_tune.size() ? "," : "");
    }
    fprintf(f, "}}\n");
    fclose(f);
    (void)rename(tmp.c_str(), path.c_str());
}

std::string tune_key(int device, int prec, uint64_t n, uint64_t batch)
{
    hipDeviceProp_t prop;
    std::string dev = "dev";
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) dev = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount) + "cu";
    else (void)hipGetLastError();
    return dev + "|" + (prec == PREC_F64 ? "f64" : "f32") + "|n=" + std::to_string(n) + "|b=" + std::to_string(batch);
}

// Median of `reps` timed executions (ms) of a plan on scratch buffers.
int time_plan(gfftx_plan_s* p, void* a, void* b, hipStream_t s, int reps, float* ms_out)
{
    for (int i = 0; i < 2; ++i) {
        int rc = exec_device(p, a, b, s);
        if (rc) return rc;
    }
    std::vector<float> t;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    for (int i = 0; i < reps; ++i) {
        (void)hipEventRecord(e0, s);
        int rc = exec_device(p, a, b, s);
        (void)hipEventRecord(e1, s);
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        (void)hipEventSynchronize(e1);
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.push_back(ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return GFFTX_SUCCESS;
}

int create_plan(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                uint32_t flags)
{
    int rc = create_plan_impl(out, n, batch, vtype, ttype, place, device, flags, {}, "");
    if (rc != GFFTX_SUCCESS) return rc;
    gfftx_plan_s* p = *out;
    if (p->passes.size() < 2 || (flags & GFFTX_FLAG_NO_AUTOTUNE) || env_int("GFFTX_AUTOTUNE", 1) == 0 ||
        getenv("GFFTX_PASSES"))
        return GFFTX_SUCCESS;
    const std::string key = tune_key(device, p->prec, n, batch);
    std::vector<int> chosen;
    std::string note;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) {
            chosen = it->second;
            note = ",\"tuned\":\"cache\"";
        }
    }
    if (chosen.empty()) {
        const std::vector<std::vector<int>> splits =
            candidate_passes(p->prec, n, (size_t)env_int("GFFTX_TUNE_CANDIDATES", 5), batch, device_cus(device));
        // each split with its default tiles, then with every other compiled tile shape of
        // its first pass (T->E) and of each length it uses after the first pass (T->T)
        std::vector<std::vector<int>> cands;
        for (const std::vector<int>& sp : splits) {
            cands.push_back(sp);
            std::vector<int> seen;
            for (size_t i = 0; i < sp.size(); ++i) {
                const int L = sp[i];
                const bool later = i > 0;
                if (later && std::find(seen.begin(), seen.end(), L) != seen.end()) continue;
                if (later) seen.push_back(L);
                const TileEntry* def = find_tile(p->prec, L, true, later);
                for (int t = 0; t < kNumTiles; ++t) {
                    const TileEntry& e = kTiles[t];
                    if (e.prec != p->prec || e.L != L || !e.loadT || e.storeT != later || e.split || &e == def) continue;
                    std::vector<int> c = sp;
                    c.push_back(0);
                    for (size_t j = 0; j < sp.size(); ++j) {
                        const bool alt = later ? (j > 0 && sp[j] == L) : j == 0;
                        c.push_back(alt ? e.nb : 0);
                        c.push_back(alt ? e.wg : 0);
                    }
                    cands.push_back(c);
                }
            }
        }
        if (cands.size() < 2) return GFFTX_SUCCESS;
        DeviceGuard g(device);
        const size_t bytes = (size_t)p->dist * batch * elem_bytes(p->prec);
        void *a = nullptr, *b = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
            (void)hipGetLastError();
            if (a) (void)hipFree(a);
            return GFFTX_SUCCESS;  // no room to tune: keep the cost-model plan
        }
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        (void)hipMemsetAsync(a, 0, bytes, s);
        float best = 1e30f;
        std::string timings = ",\"tuned\":[";
        for (size_t i = 0; i < cands.size(); ++i) {
            gfftx_plan cp = nullptr;
            // the candidates are the complex transforms of the plan (a real plan's bins
            // layout only widens the buffers, which `bytes` already covers)
            if (create_plan_impl(&cp, n, batch, vtype, GFFTX_DFT, GFFTX_OUT_OF_PLACE, device, 0, cands[i], "") !=
                GFFTX_SUCCESS)
                continue;
            float ms = 0;
            const int trc = time_plan(cp, a, b, s, 5, &ms);
            gfftx_plan_destroy(cp);
            if (trc) continue;
            std::vector<int> sl;
            std::vector<std::pair<int, int>> tl;
            split_choice(cands[i], sl, tl);
            std::string lens;
            for (size_t j = 0; j < sl.size(); ++j) {
                lens += (j ? "x" : "") + std::to_string(sl[j]);
                if (j < tl.size() && tl[j].first) lens += "/" + std::to_string(tl[j].first) + "-" + std::to_string(tl[j].second);
            }
            char buf[160];
            snprintf(buf, sizeof(buf), "%s{\"split\":\"%s\",\"ms\":%.4f}", i ? "," : "", lens.c_str(), ms);
            timings += buf;
            if (ms < best) {
                best = ms;
                chosen = cands[i];
            }
        }
        timings += "]";
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        (void)hipFree(a);
        (void)hipFree(b);
        (void)hipGetLastError();
        if (chosen.empty()) return GFFTX_SUCCESS;
        note = timings;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = chosen;
        tune_store();
    }
    gfftx_plan q = nullptr;
    rc = create_plan_impl(&q, n, batch, vtype, ttype, place, device, flags, chosen, note);
    if (rc != GFFTX_SUCCESS) return GFFTX_SUCCESS;  // keep the cost-model plan
    gfftx_plan_destroy(p);
    *out = q;
    return GFFTX_SUCCESS;
}
}  // namespace

extern "C" {

int gfftx_plan_create(gfftx_plan* plan, uint64_t n, int vtype, int ttype, int place, int device, uint32_t flags)
{
    return create_plan(plan, n, 1, vtype, ttype, place, device, flags);
}

int gfftx_plan_create_batched(gfftx_plan* plan, uint64_t n, uint64_t batch, int vtype, int ttype, int place,
                              int device, uint32_t flags)
{
    return create_plan(plan, n, batch, vtype, ttype, place, device, flags);
}

int gfftx_exec_inplace(gfftx_plan plan, void* data, void* stream)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan is NULL");
    if (!data) return fail(GFFTX_ERR_INVALID_ARG, "data is NULL");
    if (plan->place != GFFTX_IN_PLACE) return fail(GFFTX_ERR_INVALID_ARG, "plan is out-of-place");
    return exec_any(plan, data, data, (hipStream_t)stream);
}

int gfftx_exec_oop(gfftx_plan plan, const void* src, void* dst, void* stream)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan is NULL");
    if (!src || !dst) return fail(GFFTX_ERR_INVALID_ARG, "src/dst is NULL");
    if (plan->place != GFFTX_OUT_OF_PLACE) return fail(GFFTX_ERR_INVALID_ARG, "plan is in-place");
    if (src == dst) return fail(GFFTX_ERR_ALIAS, "out-of-place transform with src == dst");
    return exec_any(plan, src, dst, (hipStream_t)stream);
}

int gfftx_plan_profile(gfftx_plan plan, double* total_ms, uint64_t* count, const char** names, int max,
                       int* nslots)
{
    if (!plan || !total_ms || !count || max < 0) return fail(GFFTX_ERR_INVALID_ARG, "bad arguments");
    if (!(plan->flags & GFFTX_FLAG_PROFILE)) return fail(GFFTX_ERR_INVALID_ARG, "plan was not created with GFFTX_FLAG_PROFILE");
    DeviceGuard g(plan->device);
    const int n = (int)plan->slots.size();
    if (nslots) *nslots = n;
    for (int i = 0; i < n; ++i) {
        gfftx_plan_s::Slot& sl = plan->slots[i];
        double tot = 0;
        for (auto& ev : sl.pending) {
            HIP_TRY(hipEventSynchronize(ev.second));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
            tot += ms;
            sl.pool.push_back(ev);
        }
        if (i < max) {
            total_ms[i] = tot;
            count[i] = sl.pending.size();
            if (names) names[i] = sl.name.c_str();
        }
        sl.pending.clear();
    }
    return GFFTX_SUCCESS;
}

int gfftx_plan_destroy(gfftx_plan plan)
{
    if (!plan) return GFFTX_SUCCESS;
    DeviceGuard g(plan->device);
    plan->graphs.clear();
    for (int k = 0; k < 2; ++k) {
        if (plan->pin_in[k]) (void)hipHostFree(plan->pin_in[k]);
        if (plan->pin_out[k]) (void)hipHostFree(plan->pin_out[k]);
        if (plan->pin_in_ev[k]) (void)hipEventDestroy(plan->pin_in_ev[k]);
        if (plan->pin_out_ev[k]) (void)hipEventDestroy(plan->pin_out_ev[k]);
    }
    if (plan->h2d_ev) (void)hipEventDestroy(plan->h2d_ev);
    if (plan->done_ev) (void)hipEventDestroy(plan->done_ev);
    if (plan->h2d_stream) (void)hipStreamDestroy(plan->h2d_stream);
    if (plan->d2h_stream) (void)hipStreamDestroy(plan->d2h_stream);
    if (plan->blue_inv) gfftx_plan_destroy(plan->blue_inv);  // shares scratch owned by blue_fwd
    if (plan->blue_fwd) gfftx_plan_destroy(plan->blue_fwd);
    for (auto& sl : plan->slots) {
        for (auto& ev : sl.pending) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
        for (auto& ev : sl.pool) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    }
    for (void* a : plan->allocs) (void)hipFree(a);
    delete plan;
    return GFFTX_SUCCESS;
}

int gfftx_plan_describe(gfftx_plan plan, char* buf, size_t len)
{
    if (!plan || !buf || len == 0) return fail(GFFTX_ERR_INVALID_ARG, "bad arguments");
    snprintf(buf, len, "%s", plan->desc.c_str());
    return plan->desc.size() < len ? GFFTX_SUCCESS : fail(GFFTX_ERR_INVALID_ARG, "buffer too small");
}

int gfftx_digit_reverse(const void* src, void* dst, uint64_t radix, uint32_t digits, int vtype, void* stream)
{
    if (!src || !dst || src == dst) return fail(GFFTX_ERR_INVALID_ARG, "src/dst NULL or aliased");
    if (radix < 2 || digits == 0) return fail(GFFTX_ERR_INVALID_ARG, "radix >= 2 and digits >= 1 required");
    long double nn = 1;
    for (uint32_t i = 0; i < digits; ++i) nn *= (long double)radix;
    if (nn > 9.0e15L) return fail(GFFTX_ERR_INVALID_ARG, "radix^digits too large");
    if (vtype < 0 || vtype > 3) return fail(GFFTX_ERR_INVALID_ARG, "unknown value type %d", vtype);
    if (!is_device_ptr(src) || !is_device_ptr(dst))
        return fail(GFFTX_ERR_INVALID_ARG, "gfftx_digit_reverse takes device pointers");
    const int prec = (vtype == GFFTX_DOUBLE || vtype == GFFTX_COMPLEX_DOUBLE) ? PREC_F64 : PREC_F32;
    hipError_t e = launch_digit_reverse(prec, src, dst, (long long)radix, (int)digits, (long long)nn,
                                        (hipStream_t)stream);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "digit_reverse: %s", hipGetErrorString(e));
    return GFFTX_SUCCESS;
}


int gfftx_ragged_create(gfftx_ragged* plan, const uint64_t* lengths, uint64_t count, int vtype, int ttype,
                        int device, uint32_t flags)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan pointer is NULL");
    *plan = nullptr;
    if (!lengths || count == 0) return fail(GFFTX_ERR_INVALID_ARG, "empty ragged batch");
    if (count > (1ull << 24)) return fail(GFFTX_ERR_INVALID_ARG, "too many transforms");
    if (vtype < 0 || vtype > 3) return fail(GFFTX_ERR_INVALID_ARG, "unknown value type %d", vtype);
    if (ttype != GFFTX_DFT && ttype != GFFTX_IDFT)
        return fail(GFFTX_ERR_UNSUPPORTED, "ragged batches are complex transforms (DFT / IDFT)");
    if (flags & ~(GFFTX_FLAG_GRAPH | GFFTX_FLAG_PROFILE | GFFTX_FLAG_NO_AUTOTUNE))
        return fail(GFFTX_ERR_INVALID_ARG, "unsupported flags 0x%x for a ragged batch", flags);
    for (uint64_t i = 0; i < count; ++i)
        if (lengths[i] == 0) return fail(GFFTX_ERR_INVALID_ARG, "zero length at %llu", (unsigned long long)i);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        return fail(GFFTX_ERR_DEVICE, "no HIP device available");
    }
    if (device < 0 || device >= ndev) return fail(GFFTX_ERR_DEVICE, "device %d out of range (%d)", device, ndev);
    DeviceGuard g(device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", device);
    gfftx_ragged_s* r = new (std::nothrow) gfftx_ragged_s();
    if (!r) return fail(GFFTX_ERR_ALLOC, "host allocation failed");
    r->device = device;
    r->prec = (vtype == GFFTX_DOUBLE || vtype == GFFTX_COMPLEX_DOUBLE) ? PREC_F64 : PREC_F32;
    r->inverse = ttype == GFFTX_IDFT;
    r->flags = flags;
    int rc = ragged_build(r, lengths, count);
    if (rc != GFFTX_SUCCESS) {
        gfftx_ragged_destroy(r);  // leaves the error message alone
        return rc;
    }
    *plan = r;
    return GFFTX_SUCCESS;
}

int gfftx_ragged_exec(gfftx_ragged plan, void* data, void* stream)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan is NULL");
    if (!data) return fail(GFFTX_ERR_INVALID_ARG, "data is NULL");
    DeviceGuard g(plan->device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", plan->device);
    hipStream_t s = (hipStream_t)stream;
    if (is_device_ptr(data)) return ragged_launch(plan, data, s);
    const size_t bytes = (size_t)plan->total * elem_bytes(plan->prec);
    if (!plan->stage) {
        int rc = dev_alloc_list(plan->allocs, &plan->stage, bytes);
        if (rc) return rc;
    }
    HIP_TRY(hipMemcpyAsync(plan->stage, data, bytes, hipMemcpyHostToDevice, s));
    int rc = ragged_launch(plan, plan->stage, s);
    if (rc) return rc;
    HIP_TRY(hipMemcpyAsync(data, plan->stage, bytes, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return GFFTX_SUCCESS;
}

int gfftx_ragged_describe(gfftx_ragged plan, char* buf, size_t len)
{
    if (!plan || !buf || len == 0) return fail(GFFTX_ERR_INVALID_ARG, "bad arguments");
    snprintf(buf, len, "%s", plan->desc.c_str());
    return plan->desc.size() < len ? GFFTX_SUCCESS : fail(GFFTX_ERR_INVALID_ARG, "buffer too small");
}

int gfftx_ragged_profile(gfftx_ragged plan, double* total_ms, uint64_t* count)
{
    if (!plan || !total_ms || !count) return fail(GFFTX_ERR_INVALID_ARG, "bad arguments");
    if (!(plan->flags & GFFTX_FLAG_PROFILE)) return fail(GFFTX_ERR_INVALID_ARG, "plan was not created with GFFTX_FLAG_PROFILE");
    DeviceGuard g(plan->device);
    double tot = 0;
    for (auto& ev : plan->pending) {
        HIP_TRY(hipEventSynchronize(ev.second));
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, ev.first, ev.second));
        tot += ms;
        plan->pool.push_back(ev);
    }
    *total_ms = tot;
    *count = plan->pending.size();
    plan->pending.clear();
    return GFFTX_SUCCESS;
}

int gfftx_ragged_destroy(gfftx_ragged plan)
{
    if (!plan) return GFFTX_SUCCESS;
    DeviceGuard g(plan->device);
    plan->graphs.clear();
    for (auto& lc : plan->longs)
        if (lc.blue) gfftx_plan_destroy(lc.blue);
    for (int k = 0; k < kRaggedSides; ++k) {
        if (plan->side[k]) (void)hipStreamDestroy(plan->side[k]);
        if (plan->join_ev[k]) (void)hipEventDestroy(plan->join_ev[k]);
    }
    if (plan->fork_ev) (void)hipEventDestroy(plan->fork_ev);
    for (auto& ev : plan->pending) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    for (auto& ev : plan->pool) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
    for (void* a : plan->allocs) (void)hipFree(a);
    delete plan;
    return GFFTX_SUCCESS;
}

const char* gfftx_strerror(int status)
{
    switch (status) {
        case GFFTX_SUCCESS: return "success";
        case GFFTX_ERR_INVALID_ARG: return "invalid argument";
        case GFFTX_ERR_UNSUPPORTED: return "Requested transform is not compiled! (unsupported length/type/place)";
        case GFFTX_ERR_ALLOC: return "allocation failed";
        case GFFTX_ERR_DEVICE: return "HIP device error";
        case GFFTX_ERR_ALIAS: return "out-of-place transform with src == dst";
        case GFFTX_ERR_LAUNCH: return "kernel launch failed";
        default: return "unknown status";
    }
}

const char* gfftx_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
