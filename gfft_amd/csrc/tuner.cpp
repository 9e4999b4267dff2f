// tuner.cpp -- plan creation with the measured autotuner and its versioned JSON cache.
#include "plan_internal.h"

#include <fcntl.h>
#include <sys/file.h>

namespace gfftx {
namespace detail {

// ---------------------------------------------------------------------------------------
// Autotuner: for a multi-pass plan, time the best few pass splits of the cost model on this
// device and keep the fastest.  Results are cached per (device, precision, n, batch, kind,
// place) in memory and in a versioned JSON file ($GFFTX_TUNE_CACHE, default
// ~/.cache/gfftx/tune.json).
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
void read_entries(const std::string& path, std::map<std::string, std::vector<int>>& out)
{
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
        // plan entries hold two pass lengths or more; a ragged batch's entry one limit
        const bool ragged = key.find("|ragged|") != std::string::npos;
        if (lens.size() >= 2 || (ragged && v.size() == 1)) out[key] = v;
    }
    fclose(f);
}

void tune_load()
{
    const std::string path = tune_path();
    if (g_tune_loaded == path) return;
    g_tune_loaded = path;
    g_tune.clear();
    read_entries(path, g_tune);
}

// Writes the in-memory entries merged with what the file holds now: another process may have
// stored entries since this one read it.  An exclusive lock on "<file>.lock" serialises the
// read-merge-rename of concurrent writers; entries of this process win on equal keys.
void tune_store()
{
    const std::string path = tune_path();
    const size_t slash = path.find_last_of('/');
    if (slash != std::string::npos) {
        std::string dir = path.substr(0, slash);
        for (size_t i = 1; i <= dir.size(); ++i)
            if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    }
    const int lk = open((path + ".lock").c_str(), O_CREAT | O_RDWR, 0644);
    if (lk >= 0) (void)flock(lk, LOCK_EX);
    std::map<std::string, std::vector<int>> disk;
    read_entries(path, disk);
    for (const auto& kv : disk) g_tune.insert(kv);  // keeps this process's entry on equal keys
    const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
    FILE* f = fopen(tmp.c_str(), "w");
    if (!f) {
        if (lk >= 0) close(lk);
        return;
    }
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
    if (lk >= 0) close(lk);  // releases the lock
}

std::string tune_key(int device, int prec, uint64_t n, uint64_t batch);

// Key of a ragged batch's choice: the device, the precision, the point and transform counts
// and an FNV-1a hash of the (length, count) classes in increasing length.
std::string ragged_tune_key(int device, int prec, const uint64_t* lengths, uint64_t count)
{
    std::map<uint64_t, uint64_t> classes;
    uint64_t points = 0;
    for (uint64_t i = 0; i < count; ++i) {
        ++classes[lengths[i]];
        points += lengths[i];
    }
    uint64_t h = 1469598103934665603ull;
    auto mix = [&h](uint64_t v) {
        for (int b = 0; b < 8; ++b) {
            h ^= (v >> (8 * b)) & 0xff;
            h *= 1099511628211ull;
        }
    };
    for (const auto& kv : classes) {
        mix(kv.first);
        mix(kv.second);
    }
    char hex[24];
    snprintf(hex, sizeof(hex), "%016llx", (unsigned long long)h);
    return tune_key(device, prec, points, count) + "|ragged|h=" + hex;
}

bool tune_get(const std::string& key, std::vector<int>* v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load();
    auto it = g_tune.find(key);
    if (it == g_tune.end()) return false;
    *v = it->second;
    return true;
}

void tune_put(const std::string& key, const std::vector<int>& v)
{
    std::lock_guard<std::mutex> lk(g_tune_mu);
    tune_load();
    g_tune[key] = v;
    tune_store();
}

std::string tune_key(int device, int prec, uint64_t n, uint64_t batch)
{
    hipDeviceProp_t prop;
    std::string dev = "dev";
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) dev = std::string(prop.gcnArchName) + ":" + std::to_string(prop.multiProcessorCount) + "cu";
    else (void)hipGetLastError();
    return dev + "|" + (prec == PREC_F64 ? "f64" : "f32") + "|n=" + std::to_string(n) + "|b=" + std::to_string(batch);
}

// Time (ms) of one execution of a plan on scratch buffers, as a throughput: three groups of
// `reps` executions issued back to back between two events, the median of the group means.
// Back-to-back calls are what a batch loop issues; a call timed alone after a synchronisation
// also carries the start of an idle device.
int time_plan(gfftx_plan_s* p, void* a, void* b, hipStream_t s, int reps, float* ms_out)
{
    for (int i = 0; i < 2; ++i) {
        int rc = exec_device(p, a, b, s);
        if (rc) return rc;
    }
    std::vector<float> t;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        (void)hipGetLastError();
        return fail(GFFTX_ERR_DEVICE, "cannot create timing events");
    }
    for (int g = 0; g < 3; ++g) {
        (void)hipEventRecord(e0, s);
        int rc = GFFTX_SUCCESS;
        for (int i = 0; i < reps && !rc; ++i) rc = exec_device(p, a, b, s);
        (void)hipEventRecord(e1, s);
        (void)hipEventSynchronize(e1);
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
        float ms = 0;
        (void)hipEventElapsedTime(&ms, e0, e1);
        t.push_back(ms / reps);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::sort(t.begin(), t.end());
    *ms_out = t[t.size() / 2];
    return GFFTX_SUCCESS;
}

// Tuning candidates from pass splits: each split with its default tiles, then with every
// other compiled tile shape of its first pass (T->E) and of each length it uses after the
// first pass (T->T), one position at a time ([lens..., 0, nb, wg per pass], 0 = default).
static std::vector<std::vector<int>> split_variants(int prec, const std::vector<std::vector<int>>& splits)
{
    std::vector<std::vector<int>> cands;
    for (const std::vector<int>& sp : splits) {
        cands.push_back(sp);
        std::vector<int> seen;
        for (size_t i = 0; i < sp.size(); ++i) {
            const int L = sp[i];
            const bool later = i > 0;
            if (later && std::find(seen.begin(), seen.end(), L) != seen.end()) continue;
            if (later) seen.push_back(L);
            const TileEntry* def = find_tile(prec, L, true, later);
            for (int t = 0; t < kNumTiles; ++t) {
                const TileEntry& e = kTiles[t];
                if (e.prec != prec || e.L != L || !e.loadT || e.storeT != later || &e == def) continue;
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
    return cands;
}

// The candidates vary one pass position at a time around the default tiles.  For the winning
// split, the fastest variant timed at each position combined (ms[i] < 0: candidate i failed);
// empty when that is no new candidate.
static std::vector<int> best_combination(const std::vector<int>& chosen, const std::vector<std::vector<int>>& cands,
                                         const std::vector<float>& ms)
{
    std::vector<int> wl;
    std::vector<std::pair<int, int>> wt;
    split_choice(chosen, wl, wt);
    std::vector<std::pair<int, int>> bt(wl.size(), {0, 0});
    std::vector<float> bms(wl.size(), 1e30f);
    for (size_t i = 0; i < cands.size(); ++i) {
        if (ms[i] < 0) continue;
        std::vector<int> cl;
        std::vector<std::pair<int, int>> ct;
        split_choice(cands[i], cl, ct);
        if (cl != wl) continue;
        for (size_t j = 0; j < ct.size() && j < wl.size(); ++j)
            if (ct[j].first && ms[i] < bms[j]) {
                bms[j] = ms[i];
                bt[j] = ct[j];
            }
    }
    std::vector<int> combo = wl;
    combo.push_back(0);
    for (const auto& t : bt) {
        combo.push_back(t.first);
        combo.push_back(t.second);
    }
    const bool varied = std::any_of(bt.begin(), bt.end(), [](const std::pair<int, int>& t) { return t.first != 0; });
    if (!varied || combo == chosen || std::find(cands.begin(), cands.end(), combo) != cands.end()) return {};
    return combo;
}

int create_plan(gfftx_plan* out, uint64_t n, uint64_t batch, int vtype, int ttype, int place, int device,
                uint32_t flags)
{
    int rc = create_plan_impl(out, n, batch, vtype, ttype, place, device, flags, {}, "");
    if (rc != GFFTX_SUCCESS) return rc;
    gfftx_plan_s* p = *out;
    // a batched single-pass plan is tuned over the compiled tiles of its length (shape and
    // radices), a multi-pass plan over pass splits and tile shapes
    const bool single_pass = p->passes.size() == 1 && p->passes[0].k && p->passes[0].items_in_ntrans;
    if ((p->passes.size() < 2 && !single_pass) || (flags & GFFTX_FLAG_NO_AUTOTUNE) ||
        env_int("GFFTX_AUTOTUNE", 1) == 0 || getenv("GFFTX_PASSES"))
        return GFFTX_SUCCESS;
    // a real plan is timed with the split fused into its last (forward) or first (inverse)
    // pass, and every candidate runs with the plan's own placement (one buffer or two), so
    // the choice is kept per transform kind and place; forward and inverse complex transforms
    // share it
    const std::string key = tune_key(device, p->prec, n, batch) +
                            (ttype == GFFTX_RDFT ? "|rdft" : ttype == GFFTX_IRDFT ? "|irdft" : "") +
                            (place == GFFTX_IN_PLACE ? "|ip" : "|oop");
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
        std::vector<std::vector<int>> splits;
        if (!single_pass)
            splits = candidate_passes(p->prec, n, (size_t)env_int("GFFTX_TUNE_CANDIDATES", 5), batch, device_cus(device));
        std::vector<std::vector<int>> cands = split_variants(p->prec, splits);
        if (single_pass)
            for (int t = 0; t < kNumTiles; ++t) {
                const TileEntry& e = kTiles[t];
                if (e.prec == p->prec && (uint64_t)e.L == n && !e.loadT && !e.storeT)
                    cands.push_back({(int)n, 1, 0, e.nb, e.wg});
            }
        if (cands.size() < 2) return GFFTX_SUCCESS;
        DeviceGuard g(device);
        // the candidates run on at most 1 GiB of data (a larger batch is timed on a prefix of
        // its items), so tuning never doubles a large plan's memory; the prefix still exceeds
        // the 256 MB Infinity Cache, so the timings see HBM as the plan will
        const size_t item_bytes = (size_t)p->dist * elem_bytes(p->prec);
        const uint64_t tbatch = std::min<uint64_t>(batch, std::max<uint64_t>(1, (1ull << 30) / item_bytes));
        const size_t bytes = item_bytes * tbatch;
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
        size_t ntimed = 0;
        std::vector<float> cms(cands.size(), -1.f);  // time of each candidate (-1: failed)
        // time one candidate, record it in `timings`, keep it if it is the fastest so far
        auto try_cand = [&](const std::vector<int>& c) -> float {
            gfftx_plan cp = nullptr;
            // the candidates are the plan's own transform (a real forward plan times its last
            // pass with the R2C split fused in) with its own placement: an in-place plan runs
            // on one buffer, which the caches treat differently from two; `bytes` covers the
            // bins layout
            if (create_plan_impl(&cp, n, tbatch, vtype, ttype, place, device, flags & GFFTX_FLAG_R2C_BINS, c, "") !=
                GFFTX_SUCCESS)
                return -1.f;
            float ms = 0;
            const int trc = time_plan(cp, a, place == GFFTX_IN_PLACE ? a : b, s, 5, &ms);
            gfftx_plan_destroy(cp);
            if (trc) return -1.f;
            std::vector<int> sl;
            std::vector<std::pair<int, int>> tl;
            split_choice(c, sl, tl);
            std::string lens;
            if (single_pass) {
                lens = "single/" + std::to_string(tl[0].first) + "-" + std::to_string(tl[0].second);
            } else {
                for (size_t j = 0; j < sl.size(); ++j) {
                    lens += (j ? "x" : "") + std::to_string(sl[j]);
                    if (j < tl.size() && tl[j].first) lens += "/" + std::to_string(tl[j].first) + "-" + std::to_string(tl[j].second);
                }
            }
            char buf[160];
            snprintf(buf, sizeof(buf), "%s{\"split\":\"%s\",\"ms\":%.4f}", ntimed++ ? "," : "", lens.c_str(), ms);
            timings += buf;
            if (ms < best) {
                best = ms;
                chosen = c;
            }
            return ms;
        };
        for (size_t i = 0; i < cands.size(); ++i) cms[i] = try_cand(cands[i]);
        if (!single_pass && !chosen.empty()) {
            const std::vector<int> combo = best_combination(chosen, cands, cms);
            if (!combo.empty()) try_cand(combo);
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

// Plan choice for one length class of a ragged batch (count transforms of length L): every
// compiled single-pass tile of the length against the cost model's best three pass splits,
// each timed on a contiguous batch of the class's size.  Returns false when nothing was
// measured (autotuning off, or no buffers), and the caller keeps its built-in rule.  Cached
// like plans, under the class's own key; a single-pass choice is stored as [L, 1, 0, nb, wg]
// and returned with *single set and the tile's shape in *nb, *wg.
bool tune_length_class(int device, int prec, long long L, long long count, bool inverse, bool* single,
                       std::vector<int>* lens, std::vector<std::pair<int, int>>* tiles)
{
    if (env_int("GFFTX_AUTOTUNE", 1) == 0) return false;
    const bool has_single = L <= (1 << 30) && find_tile(prec, (int)L, false, false) != nullptr;
    const std::vector<std::vector<int>> splits = candidate_passes(prec, (unsigned long long)L, 3,
                                                                  (unsigned long long)count, device_cus(device));
    if ((has_single ? 1 : 0) + splits.size() < 2) return false;
    const std::string key = tune_key(device, prec, (uint64_t)L, (uint64_t)count) + "|class";
    std::vector<int> chosen;
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        tune_load();
        auto it = g_tune.find(key);
        if (it != g_tune.end()) chosen = it->second;
    }
    if (chosen.empty()) {
        std::vector<std::vector<int>> cands;
        for (int t = 0; has_single && t < kNumTiles; ++t) {
            const TileEntry& e = kTiles[t];
            if (e.prec == prec && e.L == L && !e.loadT && !e.storeT) cands.push_back({(int)L, 1, 0, e.nb, e.wg});
        }
        for (const std::vector<int>& sp : split_variants(prec, splits)) cands.push_back(sp);
        DeviceGuard g(device);
        // at most 1 GiB of data, like plan tuning
        const uint64_t tcount = std::min<uint64_t>((uint64_t)count,
                                                   std::max<uint64_t>(1, (1ull << 30) / ((size_t)L * elem_bytes(prec))));
        const size_t bytes = (size_t)L * tcount * elem_bytes(prec);
        void* a = nullptr;
        if (hipMalloc(&a, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        (void)hipMemsetAsync(a, 0, bytes, s);
        const int cvt = prec == PREC_F64 ? GFFTX_COMPLEX_DOUBLE : GFFTX_COMPLEX_FLOAT;
        float best = 1e30f;
        bool found = false;
        auto try_cand = [&](const std::vector<int>& c) -> float {
            gfftx_plan cp = nullptr;
            // in place, as the classes of a ragged batch run
            if (create_plan_impl(&cp, (uint64_t)L, tcount, cvt, inverse ? GFFTX_IDFT : GFFTX_DFT, GFFTX_IN_PLACE,
                                 device, GFFTX_FLAG_NO_AUTOTUNE, c, "") != GFFTX_SUCCESS)
                return -1.f;
            float ms = 0;
            const int trc = time_plan(cp, a, a, s, 5, &ms);
            gfftx_plan_destroy(cp);
            if (trc) return -1.f;
            if (ms < best) {
                best = ms;
                chosen = c;
                found = true;
            }
            return ms;
        };
        std::vector<float> cms(cands.size(), -1.f);
        for (size_t i = 0; i < cands.size(); ++i) cms[i] = try_cand(cands[i]);
        if (found && !(chosen.size() >= 2 && chosen[1] == 1 && chosen[0] == (int)L)) {
            const std::vector<int> combo = best_combination(chosen, cands, cms);
            if (!combo.empty()) try_cand(combo);
        }
        (void)hipStreamSynchronize(s);
        (void)hipStreamDestroy(s);
        (void)hipFree(a);
        (void)hipGetLastError();
        if (!found) return false;
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[key] = chosen;
        tune_store();
    }
    split_choice(chosen, *lens, *tiles);
    *single = lens->size() == 2 && (*lens)[1] == 1;
    if (*single) lens->clear();
    return true;
}

}  // namespace detail
}  // namespace gfftx
