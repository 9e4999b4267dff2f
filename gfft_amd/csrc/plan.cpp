// plan.cpp -- plans of the gfx950 FFT engine: construction, execution and C ABI.
//
// A plan is a list of passes over device buffers.  A transform whose length fits one LDS
// tile is one batched pass; a longer one is a Stockham decomposition into "super radices"
// N = P_0 * P_1 * ... each executed as one pass of the tile kernel (see tile_kernel.h).
// Twiddles are generated here once per plan from long-double evaluations with octant
// argument reduction and stored in the working precision (stage tables) or FP64 (pass
// tables).  GFFT evaluates the same roots W_N = exp(-2 pi i/N) at compile time to 18
// digits (src/metaroot.h:104-120) and then runs a recurrence at run time
// (src/gfftalg.h:249-261); tables avoid the recurrence's error growth.
#include "plan_internal.h"

namespace gfftx {
namespace detail {

thread_local std::string g_last_error;

}  // namespace detail
}  // namespace gfftx

namespace gfftx {
namespace detail {

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

bool nt_output(size_t bytes)
{
    const long long mib = env_int("GFFTX_NT_STORE_MIB", 256);
    return mib >= 0 && bytes > (size_t)mib << 20;
}

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

int upload_stage_table(gfftx_plan_s* p, const TileEntry* k, void** out)
{
    return upload_stage_table_r(p->allocs, p->prec, k->nrad, k->radices, out);
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
        const char* sm = getenv("GFFTX_MFMA_SCALE");
        const int want = (sm && !strcmp(sm, "column")) ? MFMA_SCALE_COLUMN : MFMA_SCALE_STAGE;
        for (int i = 0; i < kNumMfma && !me; ++i)
            if ((unsigned long long)kMfma[i].L == n && kMfma[i].scale == want) me = &kMfma[i];
        if (p->prec != PREC_F32 || p->real || !me)
            return fail(GFFTX_ERR_UNSUPPORTED, "split-precision MFMA path: FP32 C2C of length 256 or 4096 only (n=%llu)", n);
        Pass ps;
        ps.mfma = me;
        ps.items_in_ntrans = true;
        ps.ntrans = (long long)p->batch;
        // per-stage scaling: each stage after the first takes its input times 2^-4 through
        // the twiddles (exact: a power of two)
        int rc = upload_mfma_twiddles(p->allocs, me->scale == MFMA_SCALE_STAGE ? 0.0625 : 1.0, &ps.stage_tw);
        if (rc) return rc;
        rc = upload_mfma_matrix(p->allocs, &ps.amat);
        if (rc) return rc;
        p->passes.push_back(ps);
        snprintf(buf, sizeof(buf), "{\"kernel\":\"%s\",\"L\":%d,\"ntrans\":%lld}", me->name, me->L, ps.ntrans);
        p->desc += buf;
        p->desc += "]}";
        return GFFTX_SUCCESS;
    }
    // a forced choice (autotuner candidates, tuning cache): [n, 1, 0, nb, wg] names a
    // single-pass tile; any other list is a pass split, which overrides the single-pass tile
    const bool forced_single = p->forced_lens.size() >= 2 && p->forced_lens[0] == (int)n && p->forced_lens[1] == 1;
    const TileEntry* single = nullptr;
    if (n <= (1ull << 30) && (p->forced_lens.empty() || forced_single)) {
        std::vector<int> fl;
        std::vector<std::pair<int, int>> ft;
        split_choice(p->forced_lens, fl, ft);
        single = find_tile(p->prec, (int)n, false, false, ft.empty() ? 0 : ft[0].first, ft.empty() ? 0 : ft[0].second);
    }
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
        // real transform: the split runs inside the tile (after the last stage forward, before
        // the first stage inverse) when the tile has such a variant
        if (p->real && env_int("GFFTX_FUSED_SPLIT", 1)) {
            if (const TileEntry* kr = find_real_tile(k, p->inverse ? REAL_PRE : REAL_POST)) {
                ps.k = kr;
                ps.real = kr->real;
                rc = upload_root_tables(p->allocs, (long long)(2 * n), &ps.sp_lo, &ps.sp_hi, &ps.sp_shift);
                if (rc) return rc;
                p->split_fused = true;
            }
        }
        p->passes.push_back(ps);
        snprintf(buf, sizeof(buf), "{\"kernel\":\"%s\",\"L\":%d,\"ntrans\":%lld}", ps.k->name, k->L, ps.ntrans);
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
            ps.xcd = xcd_remap(k, p->prec);
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
            // real transform: the last pass of a forward transform computes the R2C split, the
            // first pass of an inverse one the C2R pre-pass, when the tile has a fused variant
            const bool fuse_here = p->inverse ? i == 0 : i + 1 == lens.size();
            if (fuse_here && p->real && env_int("GFFTX_FUSED_SPLIT", 1)) {
                if (const TileEntry* ks = find_real_tile(k, p->inverse ? REAL_PRE : REAL_POST)) {
                    Pass& pr = p->passes.back();
                    pr.k = ks;
                    pr.real = ks->real;
                    rc = upload_root_tables(p->allocs, (long long)(2 * n), &pr.sp_lo, &pr.sp_hi, &pr.sp_shift);
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
// With `attached` the pair goes to the launch itself (hipExtLaunchKernel stamps the kernel's
// start and end); otherwise the events are recorded on s around the launch.
bool prof_begin(gfftx_plan_s* p, size_t slot, const char* name, hipStream_t s, std::pair<hipEvent_t, hipEvent_t>* ev,
                bool attached)
{
    if (!(p->flags & GFFTX_FLAG_PROFILE) || p->capturing || p->no_events) return false;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;  // a capture the caller runs
    if (hipStreamIsCapturing(s, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    if (cs != hipStreamCaptureStatusNone) return false;
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
    if (!attached) (void)hipEventRecord(ev->first, s);
    return true;
}

void prof_end(gfftx_plan_s* p, size_t slot, hipStream_t s, const std::pair<hipEvent_t, hipEvent_t>& ev, bool attached)
{
    if (!attached) (void)hipEventRecord(ev.second, s);
    p->slots[slot].pending.push_back(ev);
}

// Runs the pass list of `p` on batch items `r` of the arrays `in`/`out` (base pointers of
// the whole batch).  Profiling events go to the slots of `prof` (the plan itself, or the
// plan that owns `p` as an inner plan) starting at `slot0`.
int run_passes(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, gfftx_plan_s* prof,
               size_t slot0)
{
    if (!prof) prof = p;
    if (!r.nb) r.nb = p->batch;
    const size_t off = (size_t)r.b0 * p->dist * elem_bytes(p->prec);
    auto at = [off](const void* q) { return (void*)((char*)const_cast<void*>(q) + off); };
    const size_t m = p->passes.size();
    // buffer for the output of pass i (i < m-1): alternate so that no pass is in place and
    // the caller's input is never written unless it is also the output
    void* const s0 = p->call_scratch[0] ? p->call_scratch[0] : p->scratch[0];
    void* const s1 = p->call_scratch[0] ? p->call_scratch[1] : p->scratch[1];
    void* dst[32];
    if (m == 0 || m > 32) return fail(GFFTX_ERR_INVALID_ARG, "plan with %zu passes", m);
    dst[m - 1] = out;
    for (size_t i = m - 1; i-- > 0;) {
        void* next_src_dst = dst[i + 1];
        void* cand = (next_src_dst == s0) ? out : s0;
        if (cand == in || cand == next_src_dst) cand = s1;
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
            ma.nt_out = (r.nt && i == m - 1) ? 1 : 0;
            dim3 grid((unsigned)((ma.ntrans + ps.mfma->nb - 1) / ps.mfma->nb));
            std::pair<hipEvent_t, hipEvent_t> ev;
            const bool on = prof_begin(prof, slot0 + i, ps.mfma->name, s, &ev, true);
            hipError_t e = ps.mfma->launch(ma, grid, s, on ? ev.first : nullptr, on ? ev.second : nullptr);
            if (on) prof_end(prof, slot0 + i, s, ev, true);
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
        a.nt_out = (r.nt && i == m - 1) ? 1 : 0;
        const long long nb = ps.k->nb;
        unsigned ntiles = (unsigned)((a.ntrans + nb - 1) / nb);
        if (ps.real) {
            a.sp_lo = ps.sp_lo;
            a.sp_hi = ps.sp_hi;
            a.sp_shift = ps.sp_shift;
            a.split_bins = (p->flags & GFFTX_FLAG_R2C_BINS) ? 1 : 0;
            a.nfull = (long long)p->n;
            if (ps.k->loadT) {
                // paired T-mode tiles: tile 0 takes sequences 0 and S/2, then nb/2 pairs
                // (p, S - p), 1 <= p <= (S-1)/2
                const long long pairs = (a.ntrans - 1) / 2, h = nb / 2;
                ntiles = (unsigned)(1 + (pairs + h - 1) / h);
            }
        }
        dim3 grid(ntiles, ps.items_in_ntrans ? 1u : (unsigned)r.nb);
        std::pair<hipEvent_t, hipEvent_t> ev;
        const bool on = prof_begin(prof, slot0 + i, ps.k->name, s, &ev, true);
        hipError_t e = ps.k->launch(a, grid, s, on ? ev.first : nullptr, on ? ev.second : nullptr);
        if (on) prof_end(prof, slot0 + i, s, ev, true);
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
    const bool on = prof_begin(p, slot, name, s, &ev, false);
    hipError_t e = launch_bluestein(p->prec, which, a, s);
    if (on) prof_end(p, slot, s, ev, false);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "%s: %s", name, hipGetErrorString(e));
    return GFFTX_SUCCESS;
}

// Chirp-z transform: pre-multiply into the zero-padded work buffer, forward FFT of length m,
// multiply by the filter spectrum, inverse FFT, post-multiply into `out`.
// `boff` (optional) gives the element offset of every batch item in `in` and `out`; the
// range must then cover the whole batch.
int run_bluestein(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r, const long long* boff)
{
    if (!r.nb) r.nb = p->batch;
    r.nt = false;  // the inner transforms write the work array, which the next kernel reads
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
    const bool prof = prof_begin(p, slot, "split_kernel", s, &ev, false);
    hipError_t e = launch_split(p->prec, a, s);
    if (prof) prof_end(p, slot, s, ev, false);
    if (e != hipSuccess) return fail(GFFTX_ERR_LAUNCH, "split: %s", hipGetErrorString(e));
    return GFFTX_SUCCESS;
}

// Full transform on device pointers.  Real forward: complex FFT then split on the output.
// Real inverse: split on a copy of the input (the input of an out-of-place call is left
// unchanged), then the inverse complex FFT.  A plan whose split is fused into one of its
// passes runs the complex passes only.
// `r` selects batch items (all by default); in and out are the base pointers of the batch.
int exec_device(gfftx_plan_s* p, const void* in, void* out, hipStream_t s, Range r)
{
    if (!r.nb) {
        // a call over the whole batch on device arrays: nothing in the call reads its output
        // back, so a large one is stored non-temporally (the host path's staged chunks are
        // copied out right after, and stay cached)
        r.nb = p->batch;
        r.nt = nt_output((size_t)p->dist * p->batch * elem_bytes(p->prec));
    }
    if (!p->real || p->split_fused) return run_c2c(p, in, out, s, r);
    if (!p->inverse) {
        Range rt = r;
        rt.nt = false;  // the split kernel reads the transform's output
        int rc = run_c2c(p, in, out, s, rt);
        if (rc) return rc;
        return run_split(p, out, s, r);
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
    HIP_TRY(GraphCache::launch(*g, s));
    return GFFTX_SUCCESS;
}

int exec_dispatch(gfftx_plan_s* p, const void* in, void* out, hipStream_t s)
{
    // one pointer query per distinct array (an in-place call asks once)
    const bool din = is_device_ptr(in), dout = out == in ? din : is_device_ptr(out);
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

// One call of the plan: calls from several threads run one at a time, and a call on another
// stream than the previous call's waits until that call's work is done when the calls share
// device arrays (scratch, chirp-z work, graphs).
int exec_any(gfftx_plan_s* p, const void* in, void* out, hipStream_t s)
{
    std::lock_guard<std::mutex> lk(p->mu);
    DeviceGuard g(p->device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", p->device);
    if (p->shared_state) HIP_TRY(p->order.before(s));
    return exec_dispatch(p, in, out, s);
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
    // an unfused split runs as a kernel of its own with a table; that of an out-of-place
    // IRDFT works on a copy of the input
    if (rc == GFFTX_SUCCESS && real && !p->split_fused) rc = upload_split_table(p);
    if (rc == GFFTX_SUCCESS && real && p->inverse && !p->split_fused)
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
    // device arrays that every call writes: calls from different streams must not overlap
    p->shared_state = p->scratch[0] || p->blue_fwd || p->real_work || (flags & GFFTX_FLAG_GRAPH);
    *out = p;
    return GFFTX_SUCCESS;
}

}  // namespace detail
}  // namespace gfftx

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

int gfftx_exec_batch(const gfftx_plan* plans, void* const* data, size_t count, void* stream)
{
    if (count == 0) return GFFTX_SUCCESS;
    if (!plans || !data) return fail(GFFTX_ERR_INVALID_ARG, "plans/data is NULL");
    for (size_t i = 0; i < count; ++i) {
        if (!plans[i]) return fail(GFFTX_ERR_INVALID_ARG, "plans[%zu] is NULL", i);
        if (!data[i]) return fail(GFFTX_ERR_INVALID_ARG, "data[%zu] is NULL", i);
        if (plans[i]->place != GFFTX_IN_PLACE) return fail(GFFTX_ERR_INVALID_ARG, "plans[%zu] is out-of-place", i);
    }
    for (size_t i = 0; i < count; ++i) {
        const int rc = exec_any(plans[i], data[i], data[i], (hipStream_t)stream);
        if (rc) {
            const std::string msg = g_last_error;
            return fail(rc, "exec_batch item %zu: %s", i, msg.c_str());
        }
    }
    return GFFTX_SUCCESS;
}

int gfftx_stream_synchronize(void* stream)
{
    const hipError_t e = hipStreamSynchronize((hipStream_t)stream);
    if (e != hipSuccess) return fail(GFFTX_ERR_DEVICE, "stream synchronize: %s", hipGetErrorString(e));
    return GFFTX_SUCCESS;
}

int gfftx_plan_profile(gfftx_plan plan, double* total_ms, uint64_t* count, const char** names, int max,
                       int* nslots)
{
    if (!plan || !total_ms || !count || max < 0) return fail(GFFTX_ERR_INVALID_ARG, "bad arguments");
    if (!(plan->flags & GFFTX_FLAG_PROFILE)) return fail(GFFTX_ERR_INVALID_ARG, "plan was not created with GFFTX_FLAG_PROFILE");
    std::lock_guard<std::mutex> lk(plan->mu);
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
