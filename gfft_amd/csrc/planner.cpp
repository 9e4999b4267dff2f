// planner.cpp -- twiddle roots, compiled-tile lookup, the pass cost model and pass splits,
// the chirp-z convolution length, and the device tables plans upload.
#include "plan_internal.h"

#include <cstring>
#include "roots.h"

namespace gfftx {
namespace detail {

void root_ld(long long m, long long N, long double* c, long double* s) { root_ld_eval(m, N, c, s); }

// First registered tile for (prec, L, modes), or the first one with nb == want_nb and
// wg == want_wg (each when nonzero) when there is one.
const TileEntry* find_tile(int prec, int L, bool loadT, bool storeT, int want_nb, int want_wg)
{
    const TileEntry* first = nullptr;
    for (int i = 0; i < kNumTiles; ++i) {
        const TileEntry& e = kTiles[i];
        if (e.prec == prec && e.L == L && e.loadT == loadT && e.storeT == storeT) {
            if ((want_nb || want_wg) && (!want_nb || e.nb == want_nb) && (!want_wg || e.wg == want_wg)) return &e;
            if (!first) first = &e;
        }
    }
    return first;
}

// The variant of tile k (same length, shape, radices and access modes) with the real-transform
// split `kind` (REAL_POST, REAL_PRE) fused in, or null.
const TileEntry* find_real_tile(const TileEntry* k, int kind)
{
    for (int i = 0; i < kNumRealTiles; ++i) {
        const TileEntry& e = kRealTiles[i];
        if (e.real != kind || e.prec != k->prec || e.L != k->L || e.nb != k->nb || e.wg != k->wg ||
            e.loadT != k->loadT || e.storeT != k->storeT || e.nrad != k->nrad)
            continue;
        bool same = true;
        for (int r = 0; r < e.nrad; ++r) same = same && e.radices[r] == k->radices[r];
        if (same) return &e;
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
std::vector<std::vector<int>> candidate_passes(int prec, unsigned long long n, size_t k, unsigned long long batch,
                                               int cus)
{
    std::vector<int> lens;
    for (int i = 0; i < kNumTiles; ++i)
        if (kTiles[i].prec == prec && kTiles[i].loadT && kTiles[i].storeT) lens.push_back(kTiles[i].L);
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

// XCD-aware tile order for a pass tile (GFFTX_XCD overrides): on when the tile's rows are not
// whole 128-byte lines, so that the neighbouring tiles that share the lines at their ends run
// on one XCD and meet in its L2.  Tiles with whole-line rows share no lines; they load their
// inputs non-temporally and run fastest in dispatch order.
int xcd_remap(const TileEntry* k, int prec)
{
    const int forced = env_int("GFFTX_XCD", -1);
    if (forced >= 0) return forced ? 1 : 0;
    return (size_t)k->nb * elem_bytes(prec) % 128 != 0 ? 1 : 0;
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
bool choose_passes(int prec, unsigned long long n, std::vector<int>& out, unsigned long long batch, int cus)
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

int upload_stage_table_r(std::vector<void*>& owned, int prec, int nrad, const int* radices, void** out,
                         double scale)
{
    *out = nullptr;
    if (nrad <= 1) return GFFTX_SUCCESS;
    // the rows in the working precision, then, 16-byte aligned, the FP64 roots W_{Ns R}^k
    // (k < Ns) of every stage s >= 1, from which the tiles form their twiddles
    // (TileCfg::TW_POW)
    std::vector<double> roots;
    long long ns = radices[0];
    for (int st = 1; st < nrad; ++st) {
        for (long long kk = 0; kk < ns; ++kk) {
            long double c, si;
            root_ld(kk, ns * radices[st], &c, &si);
            roots.push_back((double)c);
            roots.push_back((double)si);
        }
        ns *= radices[st];
    }
    std::vector<unsigned char> bytes;
    if (prec == PREC_F64) {
        std::vector<double> t;
        pack_stage_table(nrad, radices, t);
        for (double& v : t) v *= scale;
        bytes.resize(t.size() * sizeof(double));
        memcpy(bytes.data(), t.data(), bytes.size());
    } else {
        std::vector<float> t;
        pack_stage_table(nrad, radices, t);
        for (float& v : t) v = (float)(v * scale);
        while (t.size() % 4) t.push_back(0.f);
        bytes.resize(t.size() * sizeof(float));
        memcpy(bytes.data(), t.data(), bytes.size());
    }
    const size_t rows = bytes.size();
    bytes.resize(rows + roots.size() * sizeof(double));
    memcpy(bytes.data() + rows, roots.data(), roots.size() * sizeof(double));
    int rc = dev_alloc_list(owned, out, bytes.size());
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*out, bytes.data(), bytes.size(), hipMemcpyHostToDevice));
    return GFFTX_SUCCESS;
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

// Twiddles of the MFMA kernel (MfmaArgs::tw): W_256^m times `scale` (the per-stage 2^-4 or 1),
// then W_4096^m, m < 256.
int upload_mfma_twiddles(std::vector<void*>& owned, double scale, void** out)
{
    std::vector<float> t;
    for (int m = 0; m < 256; ++m) {
        long double c, si;
        root_ld(m, 256, &c, &si);
        t.push_back((float)((double)c * scale));
        t.push_back((float)((double)si * scale));
    }
    for (int m = 0; m < 256; ++m) {
        long double c, si;
        root_ld(m, 4096, &c, &si);
        t.push_back((float)c);
        t.push_back((float)si);
    }
    int rc = dev_alloc_list(owned, out, t.size() * sizeof(float));
    if (rc) return rc;
    HIP_TRY(hipMemcpy(*out, t.data(), t.size() * sizeof(float), hipMemcpyHostToDevice));
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

}  // namespace detail
}  // namespace gfftx
