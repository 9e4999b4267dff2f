// ragged_plan.cpp -- ragged batches: length classes, the persistent kernel's work lists,
// batched launches of the long classes on side streams, and the ragged C ABI.
#include "plan_internal.h"

namespace gfftx {
namespace detail {

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
// The autotuner times the class's options (tune_length_class): a single-pass tile (one
// workgroup per transform, 16 bytes of HBM traffic per point) against the best pass splits
// (many workgroups per transform, 32 bytes per point).  Without a measurement the rule is:
// the single-pass tile when one exists and the class has at least `min_single` transforms,
// else the cost model's split.  `nscratch` is raised to the scratch arrays the pass list
// needs.
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
    bool multi = choose_passes(r->prec, (unsigned long long)L, lens, (unsigned long long)lc.count);
    const TileEntry* k1 = (L <= (1 << 30)) ? find_tile(r->prec, (int)L, false, false) : nullptr;
    bool use_single = k1 && (!multi || lc.count >= min_single);
    // measured choice between the single-pass tile and the pass splits with their tile shapes
    // (unless the caller turned autotuning off or fixed the rule with GFFTX_RAGGED_MIN_SINGLE)
    bool t_single = false;
    std::vector<int> t_lens;
    std::vector<std::pair<int, int>> t_tiles;  // (nb, wg) per pass, (0, 0) = the default tile
    if (!(r->flags & GFFTX_FLAG_NO_AUTOTUNE) && !getenv("GFFTX_RAGGED_MIN_SINGLE") && !getenv("GFFTX_PASSES") &&
        tune_length_class(r->device, r->prec, L, lc.count, r->inverse, &t_single, &t_lens, &t_tiles)) {
        if (t_single && k1) {
            use_single = true;
            k1 = find_tile(r->prec, (int)L, false, false, t_tiles.empty() ? 0 : t_tiles[0].first,
                           t_tiles.empty() ? 0 : t_tiles[0].second);
        } else if (!t_single && t_lens.size() >= 2) {
            use_single = false;
            multi = true;
            lens = t_lens;
        }
    }
    if (!multi || use_single || lens != t_lens) t_tiles.clear();
    if (use_single) {
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
            const int tnb = i < t_tiles.size() ? t_tiles[i].first : 0, twg = i < t_tiles.size() ? t_tiles[i].second : 0;
            const TileEntry* k = find_tile(r->prec, lens[i], true, i > 0, tnb, twg);
            if (!k) return fail(GFFTX_ERR_UNSUPPORTED, "ragged batch: no pass kernel L=%d", lens[i]);
            Pass ps;
            ps.k = k;
            ps.xcd = xcd_remap(k, r->prec);
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
        a.nt_out = (i + 1 == m) ? r->args.nt_out : 0;
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
        hipError_t le = ps.k->launch(a, grid, s, nullptr, nullptr);
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

// Time of every long class's launches alone (ms, two runs after one warm-up) on side stream
// 0, in place on a zeroed array of the batch's size (`points` elements): side stream 1's
// first scratch array when there is one, else a temporary allocation.  Leaves `cost`
// unchanged when the array cannot be allocated or a launch fails.
void measure_class_costs(gfftx_ragged_s* r, long long points, std::vector<double>& cost)
{
    const bool borrowed = r->nsides >= 2 && r->lscratch[0][1] != nullptr;
    void* tmp = borrowed ? r->lscratch[0][1] : nullptr;
    const size_t bytes = (size_t)points * elem_bytes(r->prec);
    if (!borrowed && hipMalloc(&tmp, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return;
    }
    hipStream_t s = r->side[0];
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<double> t(cost.size(), 0.0);
    bool ok = hipMemsetAsync(tmp, 0, bytes, s) == hipSuccess && hipEventCreate(&e0) == hipSuccess &&
              hipEventCreate(&e1) == hipSuccess;
    for (size_t i = 0; ok && i < r->longs.size(); ++i) {
        ok = ragged_run_long(r, r->longs[i], tmp, 0, s) == GFFTX_SUCCESS && hipEventRecord(e0, s) == hipSuccess;
        for (int rep = 0; ok && rep < 2; ++rep) ok = ragged_run_long(r, r->longs[i], tmp, 0, s) == GFFTX_SUCCESS;
        float ms = 0;
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        t[i] = ms / 2;
    }
    (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (!borrowed) (void)hipFree(tmp);
    (void)hipGetLastError();
    if (ok) cost = t;
}

// Builds the plan of a ragged batch.  FP32 lengths up to `pmax` (at most 4096, the largest
// persistent tile) run in the persistent kernel; longer ones, or all with pmax = 0, run as
// batched class launches on the side streams.
int ragged_build(gfftx_ragged_s* r, const uint64_t* lengths, uint64_t count, long long pmax)
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
    // the transforms' final outputs are stored non-temporally when the batch is larger than
    // the Infinity Cache (nothing in the call reads them back)
    r->args.nt_out = nt_output((size_t)off * elem_bytes(r->prec)) ? 1 : 0;
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
        const int single = (r->prec == PREC_F32 && L <= pmax) ? ragged_single_cfg(L) : -1;
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
    std::string sched;  // describe(): each long class's cost and side stream
    if (!r->longs.empty()) {
        // classes to side streams, largest cost first onto the least loaded stream.  The cost
        // of a class is its measured time on this device (its own launches, on a scratch array
        // of the batch's size); without a measurement, the points its passes move.
        r->nsides = std::max(1, std::min<int>(env_int("GFFTX_RAGGED_STREAMS", kRaggedSides), kRaggedSides));
        r->nsides = std::min<int>(r->nsides, (int)r->longs.size());
        for (int k = 0; k < r->nsides; ++k) {
            HIP_TRY(hipStreamCreateWithFlags(&r->side[k], hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&r->join_ev[k], hipEventDisableTiming));
            for (int b = 0; b < nscratch; ++b) {
                e = dev_alloc_list(r->allocs, &r->lscratch[b][k], (size_t)off * elem_bytes(r->prec));
                if (e) return e;
            }
        }
        std::vector<double> cost(r->longs.size());
        for (size_t i = 0; i < cost.size(); ++i) {
            const gfftx_ragged_s::LongClass& lc = r->longs[i];
            cost[i] = (double)lc.L * (double)lc.count * (lc.blue ? 6.0 : (double)std::max<size_t>(1, lc.passes.size()));
        }
        if (!(r->flags & GFFTX_FLAG_NO_AUTOTUNE) && env_int("GFFTX_AUTOTUNE", 1) != 0) measure_class_costs(r, off, cost);
        std::vector<size_t> order(r->longs.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = i;
        std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return cost[x] > cost[y]; });
        std::vector<double> load(r->nsides, 0.0);
        r->side_of.assign(r->longs.size(), 0);
        for (size_t i : order) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            r->side_of[i] = k;
            load[k] += cost[i];
        }
        for (size_t i = 0; i < r->longs.size(); ++i) {
            char cb[96];
            snprintf(cb, sizeof(cb), "%s{\"L\":%lld,\"cost\":%.6g,\"side\":%d}", i ? "," : "", r->longs[i].L, cost[i],
                     r->side_of[i]);
            sched += cb;
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
               "],\"scratch_points\":" + std::to_string(scratch) + ",\"long_classes\":[" + sched + "]}";
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
            hipEvent_t x = nullptr, y = nullptr;
            if (hipEventCreate(&x) != hipSuccess || hipEventCreate(&y) != hipSuccess) {
                if (x) (void)hipEventDestroy(x);
                (void)hipGetLastError();
                return fail(GFFTX_ERR_DEVICE, "cannot create profiling events");
            }
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
        HIP_TRY(GraphCache::launch(*ge, s));
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

// Time (ms) of one whole call of a ragged plan, direct launches issued back to back on `s`
// over `buf`: after two warm-up calls, three groups of three calls between events, the
// smallest group mean.  Negative when a launch fails.
double time_ragged(gfftx_ragged_s* r, void* buf, hipStream_t s)
{
    RaggedArgs a = r->args;
    a.data = (float2*)buf;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        (void)hipGetLastError();
        return -1.0;
    }
    double best = -1.0;
    bool ok = true;
    for (int w = 0; ok && w < 2; ++w) ok = ragged_launch_all(r, a, s) == GFFTX_SUCCESS;
    for (int g = 0; ok && g < 3; ++g) {
        ok = hipEventRecord(e0, s) == hipSuccess;
        for (int c = 0; ok && c < 3; ++c) ok = ragged_launch_all(r, a, s) == GFFTX_SUCCESS;
        float ms = 0;
        ok = ok && hipEventRecord(e1, s) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
             hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
        if (ok && (best < 0 || ms / 3 < best)) best = ms / 3;
    }
    (void)hipStreamSynchronize(s);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipGetLastError();
    return ok ? best : -1.0;
}

// Chooses between the persistent kernel for the lengths up to 4096 (plan `r`, built with
// that limit) and class launches for every length: builds the second form, times one call
// of each twice in turn on a zeroed array of the batch's size, keeps the faster in `*r` and
// records the limit under `key`.  Leaves `*r` as it is when the second form or the array
// cannot be had.  Returns the note for describe().
std::string choose_persistent(gfftx_ragged_s** r, const uint64_t* lengths, uint64_t count, const std::string& key)
{
    gfftx_ragged_s* p = *r;
    gfftx_ragged_s* q = new (std::nothrow) gfftx_ragged_s();
    if (!q) return "";
    q->device = p->device;
    q->prec = p->prec;
    q->inverse = p->inverse;
    q->flags = p->flags;
    if (ragged_build(q, lengths, count, 0) != GFFTX_SUCCESS) {
        gfftx_ragged_destroy(q);
        return "";
    }
    const size_t bytes = (size_t)p->total * elem_bytes(p->prec);
    void* buf = nullptr;
    hipStream_t s = nullptr;
    double tp = -1, tq = -1;
    if (hipMalloc(&buf, bytes) == hipSuccess && hipStreamCreateWithFlags(&s, hipStreamNonBlocking) == hipSuccess &&
        hipMemsetAsync(buf, 0, bytes, s) == hipSuccess) {
        for (int rep = 0; rep < 2; ++rep) {
            const double a = time_ragged(p, buf, s), b = time_ragged(q, buf, s);
            if (a < 0 || b < 0) {
                tp = tq = -1;
                break;
            }
            tp = (rep == 0 || a < tp) ? a : tp;
            tq = (rep == 0 || b < tq) ? b : tq;
        }
    }
    if (s) (void)hipStreamDestroy(s);
    if (buf) (void)hipFree(buf);
    (void)hipGetLastError();
    if (tp < 0 || tq < 0) {
        gfftx_ragged_destroy(q);
        return "";
    }
    const bool classes = tq < tp;
    tune_put(key, {classes ? 0 : 4096});
    if (classes) {
        gfftx_ragged_destroy(p);
        *r = q;
    } else {
        gfftx_ragged_destroy(q);
    }
    char note[128];
    snprintf(note, sizeof(note), "{\"persistent_ms\":%.4f,\"class_launches_ms\":%.4f}", tp, tq);
    return note;
}

int ragged_exec_any(gfftx_ragged_s* plan, void* data, hipStream_t s)
{
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

}  // namespace detail
}  // namespace gfftx

extern "C" {

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
    // FP32 lengths up to `pmax` run in the persistent kernel.  GFFTX_RAGGED_PERSISTENT_MAX
    // fixes the limit; otherwise the autotuner times the batch both ways (persistent kernel
    // for the lengths up to 4096, or class launches for every length) and caches the choice.
    // Without autotuning, and in the other grouping modes, the limit is 4096.
    long long pmax = 4096;
    std::string key, note = "\"default\"";
    bool measure = false;
    if (getenv("GFFTX_RAGGED_PERSISTENT_MAX")) {
        pmax = env_int("GFFTX_RAGGED_PERSISTENT_MAX", 4096);
        note = "\"environment\"";
    } else if (r->prec == PREC_F32 && !(flags & GFFTX_FLAG_NO_AUTOTUNE) && env_int("GFFTX_AUTOTUNE", 1) != 0 &&
               env_int("GFFTX_RAGGED_MODE", 0) == 0) {
        key = ragged_tune_key(device, r->prec, lengths, count);
        std::vector<int> v;
        if (tune_get(key, &v) && v.size() == 1) {
            pmax = v[0];
            note = "\"cache\"";
        } else {
            measure = true;
        }
    }
    int rc = ragged_build(r, lengths, count, pmax);
    if (rc != GFFTX_SUCCESS) {
        gfftx_ragged_destroy(r);  // leaves the error message alone
        return rc;
    }
    if (measure && r->args.nitems0 + r->args.nitems1 > 0) {
        const std::string m = choose_persistent(&r, lengths, count, key);
        if (!m.empty()) note = m;
    }
    // describe(): the limit in force and how it was set
    if (!r->desc.empty() && r->desc.back() == '}') {
        r->desc.pop_back();
        r->desc += ",\"persistent_max\":" + std::to_string(r->args.nitems0 + r->args.nitems1 > 0 ? pmax : 0) +
                   ",\"persistent_choice\":" + note + "}";
    }
    *plan = r;
    return GFFTX_SUCCESS;
}

// Calls from several threads run one at a time; a call on another stream than the previous
// call's waits until that call's work is done (the work-item counter and the side-stream
// scratch are shared).
int gfftx_ragged_exec(gfftx_ragged plan, void* data, void* stream)
{
    if (!plan) return fail(GFFTX_ERR_INVALID_ARG, "plan is NULL");
    if (!data) return fail(GFFTX_ERR_INVALID_ARG, "data is NULL");
    std::lock_guard<std::mutex> lk(plan->mu);
    DeviceGuard g(plan->device);
    if (!g.ok) return fail(GFFTX_ERR_DEVICE, "cannot select device %d", plan->device);
    hipStream_t s = (hipStream_t)stream;
    HIP_TRY(plan->order.before(s));
    return ragged_exec_any(plan, data, s);
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
    std::lock_guard<std::mutex> lk(plan->mu);
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

}  // extern "C"
