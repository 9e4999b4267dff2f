// tile_kernel.h -- batched Stockham FFT of one length L held in LDS (gfx950).
//
// One workgroup transforms NB independent sequences of length L.  Each sequence is worked
// on by TPT = WG/NB threads.  Stage s (radix R_s, Ns = R_0*..*R_{s-1}) is the self-sorting
// (Stockham, decimation in time) step
//     for j in [0, L/R):  k = j mod Ns
//         v[r] = x[j + r*L/R] * W_{Ns*R}^{k*r}          r = 0..R-1
//         y[(j/Ns)*Ns*R + k + r*Ns] = DFT_R(v)[r]
// so the output is in natural order and no digit-reversal pass exists.  The first stage
// reads global memory straight into registers and the last stage writes global memory
// straight from registers when the load and store access modes agree; otherwise the last
// stage goes through LDS once more so that the global store is coalesced.
//
// The same kernel runs every pass of a multi-pass transform of length N = P_0*P_1*..: pass i
// is the Stockham step above with radix P_i applied to the whole array, i.e. a batch of
// N/P_i "sequences" j whose element r sits at j + r*N/P_i, followed by the store rule with
// Ns = P_0*..*P_{i-1}.  The pass twiddle W_{Ns*P_i}^{k*r} is produced in FP64 from two
// small FP64 tables (W^m = lo[m mod 2^b] * hi[m >> b]).
//
// Access modes: "T" (transform-contiguous) means consecutive sequences are adjacent in
// memory (its == 1 on load, Ns > 1 on store); "E" (element-contiguous) means each sequence
// is contiguous.  Lanes are mapped so that consecutive lanes touch consecutive addresses.
#pragma once
#include <type_traits>

#include "fft_common.h"

namespace gfftx {

struct TileArgs {
    const void* in;
    void* out;
    long long ntrans;   // sequences per batch item
    long long its;      // input: distance between sequences (elements)
    long long ies;      // input: distance between elements of a sequence
    long long ns_pass;  // store rule Ns (1 = each sequence written contiguously)
    long long bdist;    // distance between batch items (grid.y), elements
    const void* tw;     // stage twiddles, working precision, rows of R entries
    const double2* ptw_lo;  // pass twiddles (null: none)
    const double2* ptw_hi;
    int ptw_shift;
    int conj_in;
    int conj_out;
    int xcd;            // 1: remap workgroups so neighbouring tiles share one XCD's L2
    double scale;
    const long long* seq_off;  // optional: element offset of every sequence (E-mode in/out)
    const long long* boff;     // optional: element offset of every batch item (else by*bdist)
    // fused real-transform split (RealCfg tiles): the R2C post-pass in the last pass of a real
    // forward transform, or the C2R pre-pass in the first pass of a real inverse transform
    const double2* sp_lo;      // two-level FP64 table of W_{2N}^k
    const double2* sp_hi;
    int sp_shift;
    int split_bins;            // 1: N+1 bins, 0: NR packing
    long long nfull;           // N, the complex length of the transform
    int nt_out;                // 1: the output is stored non-temporally (see put())
};

__device__ __forceinline__ long long batch_offset(const TileArgs& a, unsigned by)
{
    return a.boff ? a.boff[by] : (long long)by * a.bdist;
}

// Workgroups are dispatched round-robin over the 8 XCDs, so consecutive blockIdx values land
// on different L2 caches.  With xcd remapping, tile (b%8)*(full/8) + b/8 is processed by
// block b: the tiles one XCD receives are consecutive, and T-mode tiles that share cache
// lines are fetched into the same L2.  The tail (nbx % 8 blocks) keeps the identity map.
__device__ __forceinline__ unsigned tile_index(unsigned b, unsigned nbx, int remap)
{
    if (!remap) return b;
    const unsigned full = nbx - (nbx & 7u);
    return b < full ? (b & 7u) * (full >> 3) + (b >> 3) : b;
}

template <typename T_, int L_, int NB_, int WG_, bool LOADT_, bool STORET_, int... Rs>
struct TileCfg {
    using T = T_;
    using V = V2<T_>;
    static constexpr int L = L_, NB = NB_, WG = WG_;
    static constexpr bool LOADT = LOADT_, STORET = STORET_;
    static constexpr int NS = sizeof...(Rs);
    static constexpr int R[NS] = {Rs...};
    static constexpr int TPT = WG / NB;
    static_assert(WG % NB == 0, "WG must be a multiple of NB");
    static constexpr int ns(int s)
    {
        int n = 1;
        for (int i = 0; i < s; ++i) n *= R[i];
        return n;
    }
    static_assert(ns(NS) == L, "radices must multiply to L");
    static constexpr int nbf(int s) { return (L / R[s] + TPT - 1) / TPT; }
    static constexpr int twoff(int s)
    {
        int o = 0;
        for (int i = 1; i < s; ++i) o += ns(i) * R[i];
        return o;
    }
    static constexpr int twsize() { return twoff(NS); }
    // FP32 tiles form a stage's twiddles W^{k r}, r = 1..R-1, as FP64 powers of one FP64 root
    // W_{Ns R}^k, rounded to FP32 once: one 16-byte load per butterfly instead of R-1 table
    // entries.  The roots follow the rows, 16-byte aligned, stage s at twbase(s) = the sum of
    // Ns over the earlier stages.  FP64 tiles read the rows (the powers measured no faster).
    static constexpr bool TW_POW = sizeof(T_) == 4;
    static constexpr int twbase(int s)
    {
        int o = 0;
        for (int i = 1; i < s; ++i) o += ns(i);
        return o;
    }
    static constexpr int TW_ROOT_OFF = (twoff(NS) * (int)sizeof(V2<T_>) + 15) / 16 * 16;  // bytes
    // LDS: one padded element every 16, odd distance between sequences (bank spread).
    static constexpr int LP = L + L / 16;
    static constexpr int TS = (LP % 2 == 0) ? LP + 1 : LP;
    static constexpr bool DIRECT = (LOADT == STORET);
    static constexpr bool NEED_LDS = (NS > 1) || !DIRECT;
    static constexpr int LDS_ELEMS = NEED_LDS ? NB * TS : 1;
    static constexpr int REAL = 0;          // REAL_NONE: no real-transform split in this tile
    static constexpr bool PAIRED = false;   // lanes pair sequences p and S - p
    static constexpr int RUN = LOADT ? NB : NB * L;  // contiguous elements per loaded row
    // E->E tiles whose TPT divides the wave size: every sequence is worked on by the lanes of
    // one wave, so the exchanges between stages need no workgroup barrier
    static constexpr bool WAVE_LOCAL = !LOADT && !STORET && (64 % TPT == 0);
    __device__ static constexpr int lidx(int t, int e) { return t * TS + e + (e >> 4); }
};

// Orders the LDS accesses of the stages before it against those after it, for the threads
// that share a sequence.  A wave's LDS instructions complete in program order, so when a
// sequence lies in one wave a compiler-level fence suffices and the waves of a workgroup run
// their stages without waiting for each other; otherwise the workgroup barrier.
template <class C> __device__ __forceinline__ void stage_sync()
{
    if constexpr (C::WAVE_LOCAL) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

enum { REAL_NONE = 0, REAL_POST = 1, REAL_PRE = 2 };

// A tile with the real-transform split fused in.  REAL_POST: the R2C post-pass after the
// last stage of a real forward transform (T->T last pass, or E->E single pass); REAL_PRE: the
// C2R pre-pass before the first stage of a real inverse transform (T->E first pass, or E->E
// single pass).  The split needs Z[k] and Z[N-k].  In an E->E tile a sequence is the whole
// transform, so both are in the tile.  In a pass over S = ntrans sequences, element q of
// sequence t is t + q S, and N - (t + q S) = (S - t) + (L - 1 - q) S for t > 0 (sequence 0
// pairs with itself, element q with L - q): a T-mode tile then takes H = NB/2 sequences p
// and their partners S - p (lanes i and i + H), and tile 0 the self-paired sequences 0 and,
// for even S, S/2.
template <class Base, int KIND> struct RealCfg : Base {
    static_assert(KIND == REAL_POST || KIND == REAL_PRE, "split kind");
    static_assert(KIND != REAL_POST || Base::LOADT == Base::STORET, "post-pass: T->T or E->E tiles");
    static_assert(KIND != REAL_PRE || !Base::STORET, "pre-pass: T->E or E->E tiles");
    static constexpr int REAL = KIND;
    static constexpr bool PAIRED = Base::LOADT;
    static_assert(!PAIRED || (Base::NB % 2 == 0 && Base::NB >= 2), "sequences pair up inside a tile");
    static constexpr int RUN = PAIRED ? Base::NB / 2 : Base::RUN;
    static constexpr int LDS_ELEMS = Base::NB * Base::TS;  // the two halves meet in LDS
};
template <class Base> using SplitCfg = RealCfg<Base, REAL_POST>;

// Sequence of lane t in tile `tile` under the split pairing (PAIRED tiles), or -1 for none.
template <class C> __device__ __forceinline__ long long split_seq(long long S, unsigned tile, int t)
{
    constexpr int H = C::NB / 2;
    if (tile == 0) return t == 0 ? 0 : (t == 1 && S % 2 == 0 && S > 1) ? S / 2 : -1;
    const long long p = 1 + (long long)(tile - 1) * H + (t % H);
    if (p > (S - 1) / 2) return -1;
    return t < H ? p : S - p;
}

template <class C> struct Ctx {
    using V = typename C::V;
    V* lds;
    const V* in;         // input of this thread's sequence: element e at in[e * ies]
    V* out;              // output of this thread's sequence: element q at out[q * ns]
    const V* tw;
    int t, jt;
    long long tg;
    unsigned ies;        // input element stride
    unsigned ons;        // output element stride (the store rule's Ns)
    bool valid;          // sequence exists (loads are clamped to a valid one regardless)
    bool active;         // thread takes part (threads beyond C::WG idle through the barriers)
    long long bo;        // element offset of the batch item
    TileArgs a;
};

template <typename V> __device__ __forceinline__ V to_v(double2 w);
template <> __device__ __forceinline__ float2 to_v<float2>(double2 w) { return float2{(float)w.x, (float)w.y}; }
template <> __device__ __forceinline__ double2 to_v<double2>(double2 w) { return w; }

__device__ __forceinline__ double2 dmul(double2 a, double2 b)
{
    return double2{__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}

// Non-temporal load of one complex element (the cache lines are marked for early eviction).
template <class V> __device__ __forceinline__ V nt_load(const V* p)
{
    using E = decltype(V{}.x);
    typedef E vec __attribute__((ext_vector_type(2)));
    const vec r = __builtin_nontemporal_load(reinterpret_cast<const vec*>(p));
    return V{r.x, r.y};
}

// Store of one complex element of a launch's output: non-temporal when `nt` (the host sets it
// for the last pass of a call whose output is larger than the 256 MB Infinity Cache and is
// not read back by the call: caching those lines would only evict what the next loads need).
template <bool NT, class V> __device__ __forceinline__ void put(V* p, V v)
{
    if constexpr (NT) {
        using E = decltype(V{}.x);
        typedef E vec __attribute__((ext_vector_type(2)));
        __builtin_nontemporal_store(vec{v.x, v.y}, reinterpret_cast<vec*>(p));
    } else {
        *p = v;
    }
}

// Runs f with the store kind as a compile-time constant, so that each store loop is emitted
// once per kind and issues its stores back to back (a run-time test per store would put a
// branch between them).
template <class F> __device__ __forceinline__ void with_nt(bool nt, F&& f)
{
    if (nt) f(std::true_type{});
    else f(std::false_type{});
}

// Conjugation as a sign-bit flip of the imaginary part: mask is 0 or the sign bit.
__device__ __forceinline__ float2 flip_im(float2 u, unsigned mask)
{
    u.y = __uint_as_float(__float_as_uint(u.y) ^ mask);
    return u;
}
__device__ __forceinline__ double2 flip_im(double2 u, unsigned mask)
{
    u.y = __longlong_as_double(__double_as_longlong(u.y) ^ ((long long)mask << 32));
    return u;
}

// First-stage inputs of one thread, as loaded (the first stage conjugates them for inverse
// transforms, so that nothing waits on a load before the stage needs its value).
template <class C> struct FirstIn {
    typename C::V v[C::nbf(0)][C::R[0]];
    // pass twiddle of a later pass: the two table halves of W^{kp*j} per group and of the
    // step W^{kp*L/R}, fetched with the inputs (rot: the sequence has kp != 0)
    double2 blo[C::nbf(0)], bhi[C::nbf(0)];
    double2 slo, shi;
    bool rot;
};

// The input of a pass is dead once read (a scratch array, or the caller's input that the
// transform consumes), while its output is what the next pass reads.  Loading the input
// non-temporally lets the caches (L2 and the 256 MB Infinity Cache) keep the output instead:
// a 2^24 FP32 round trip runs about 30% faster.  Used where the tile's rows are whole
// 128-byte lines (T mode: NB adjacent sequences per row; E mode: NB whole sequences), so
// that no other tile needs a line this one evicts early; a row of, say, 21 sequences
// (168 bytes) shares its end lines with the neighbouring tiles.
template <class C, bool IDLE> __device__ __forceinline__ void load_first(const Ctx<C>& x, FirstIn<C>& f)
{
    using V = typename C::V;
    constexpr int R = C::R[0];
    constexpr int B = C::nbf(0);
    constexpr int LR = C::L / R;
    constexpr bool GUARD = (B * C::TPT > LR);
    constexpr bool NT = C::RUN * (int)sizeof(V) % 128 == 0;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = x.jt + b * C::TPT;
        if ((IDLE && !x.active) || (GUARD && j >= LR)) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const V* p = &x.in[(unsigned)(j + r * LR) * x.ies];
            f.v[b][r] = NT ? nt_load(p) : *p;
        }
    }
    const TileArgs& a = x.a;
    f.rot = false;
    if (a.ptw_lo != nullptr) {
        const long long kp = x.tg % a.ns_pass;
        f.rot = kp != 0;
        if (f.rot) {
            const long long mask = (1LL << a.ptw_shift) - 1;
            f.slo = a.ptw_lo[(kp * LR) & mask];
            f.shi = a.ptw_hi[(kp * LR) >> a.ptw_shift];
#pragma unroll
            for (int b = 0; b < B; ++b) {
                const long long m = kp * (x.jt + b * C::TPT);
                f.blo[b] = a.ptw_lo[m & mask];
                f.bhi[b] = a.ptw_hi[m >> a.ptw_shift];
            }
        }
    }
}

template <class C, bool IDLE, int S> __device__ __forceinline__ void tile_stage(Ctx<C>& x, FirstIn<C>& f)
{
    using V = typename C::V;
    constexpr int R = C::R[S];
    constexpr int NSS = C::ns(S);
    constexpr int B = C::nbf(S);
    constexpr int LR = C::L / R;
    constexpr bool FIRST = (S == 0);
    constexpr bool LAST = (S == C::NS - 1);
    constexpr bool GUARD = (B * C::TPT > LR);
    V v[B][R];

#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = x.jt + b * C::TPT;
        if ((IDLE && !x.active) || (GUARD && j >= LR)) continue;
        if constexpr (FIRST) {
            const TileArgs& a = x.a;
            const unsigned mask = a.conj_in ? 0x80000000u : 0u;
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = flip_im(f.v[b][r], mask);
            if (f.rot) {
                // W_M^{kp*e}, e = j + r*LR: base W^{kp*j} times step^r, step = W^{kp*LR}
                const double2 step = dmul(f.slo, f.shi);
                double2 w = dmul(f.blo[b], f.bhi[b]);
                v[b][0] = cmul(v[b][0], to_v<V>(w));
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    w = dmul(w, step);
                    v[b][r] = cmul(v[b][r], to_v<V>(w));
                }
            }
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[b][r] = x.lds[C::lidx(x.t, j + r * LR)];
            const int k = j % NSS;
            if constexpr (C::TW_POW) {
                const double2 w1 = reinterpret_cast<const double2*>(reinterpret_cast<const char*>(x.tw) +
                                                                    C::TW_ROOT_OFF)[C::twbase(S) + k];
                if constexpr (R >= 8 && R % 4 == 0) {
                    // W^{k (4a + c)} = (W^{4k})^a W^{k c}: short dependency chains (the
                    // stage's latency); each twiddle is at most five FP64 products from the
                    // root, a relative error near 1e-15 before the rounding to FP32
                    const double2 w2 = dmul(w1, w1), w3 = dmul(w2, w1), w4 = dmul(w2, w2);
                    v[b][1] = cmul(v[b][1], to_v<V>(w1));
                    v[b][2] = cmul(v[b][2], to_v<V>(w2));
                    v[b][3] = cmul(v[b][3], to_v<V>(w3));
                    double2 wa = w4;
#pragma unroll
                    for (int a = 1; a < R / 4; ++a) {
                        if (a > 1) wa = dmul(wa, w4);
                        v[b][4 * a] = cmul(v[b][4 * a], to_v<V>(wa));
                        v[b][4 * a + 1] = cmul(v[b][4 * a + 1], to_v<V>(dmul(wa, w1)));
                        v[b][4 * a + 2] = cmul(v[b][4 * a + 2], to_v<V>(dmul(wa, w2)));
                        v[b][4 * a + 3] = cmul(v[b][4 * a + 3], to_v<V>(dmul(wa, w3)));
                    }
                } else {
                    double2 w = w1;
                    v[b][1] = cmul(v[b][1], to_v<V>(w));
#pragma unroll
                    for (int r = 2; r < R; ++r) {
                        w = dmul(w, w1);
                        v[b][r] = cmul(v[b][r], to_v<V>(w));
                    }
                }
            } else {
                const V* row = x.tw + C::twoff(S) + k * R;
#pragma unroll
                for (int r = 1; r < R; ++r) v[b][r] = cmul(v[b][r], row[r]);
            }
        }
        Dft<R>::run(v[b]);
    }

    if constexpr (!FIRST) stage_sync<C>();

    if constexpr (LAST && C::DIRECT && C::REAL != REAL_POST) {
        if (x.valid && (!IDLE || x.active)) {
            const TileArgs& a = x.a;
            const unsigned mask = a.conj_out ? 0x80000000u : 0u;
            const bool nt = a.nt_out != 0;
            with_nt(nt, [&](auto ntc) {
                constexpr bool NT = decltype(ntc)::value;
                if (a.scale == 1.0) {
#pragma unroll
                    for (int b = 0; b < B; ++b) {
                        const int j = x.jt + b * C::TPT;
                        if (GUARD && j >= LR) continue;
#pragma unroll
                        for (int r = 0; r < R; ++r) put<NT>(&x.out[(unsigned)(j + r * LR) * x.ons], flip_im(v[b][r], mask));
                    }
                } else {
                    const auto sc = (decltype(v[0][0].x))a.scale;
#pragma unroll
                    for (int b = 0; b < B; ++b) {
                        const int j = x.jt + b * C::TPT;
                        if (GUARD && j >= LR) continue;
#pragma unroll
                        for (int r = 0; r < R; ++r)
                            put<NT>(&x.out[(unsigned)(j + r * LR) * x.ons], flip_im(cscale(v[b][r], sc), mask));
                    }
                }
            });
        }
    } else {
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = x.jt + b * C::TPT;
            if ((IDLE && !x.active) || (GUARD && j >= LR)) continue;
            const int k = j % NSS;
            const int base = (j / NSS) * NSS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) x.lds[C::lidx(x.t, base + r * NSS)] = v[b][r];
        }
        // after the last stage an E->E tile reads only its own sequence back (split_finish_e)
        stage_sync<C>();
    }
}

template <class C, bool IDLE, int S> __device__ __forceinline__ void stage_loop(Ctx<C>& x, FirstIn<C>& f)
{
    if constexpr (S < C::NS) {
        tile_stage<C, IDLE, S>(x, f);
        stage_loop<C, IDLE, S + 1>(x, f);
    }
}

// Output element offset of sequence tg under the store rule with Ns = ns (relative to the
// batch item): (tg/ns)*ns*L + tg%ns, or the sequence's own offset in the ragged layout.
template <class C> __device__ __forceinline__ long long out_base(const TileArgs& a, long long tg)
{
    if (a.seq_off) return a.seq_off[tg];
    const unsigned ns = (unsigned)a.ns_pass;
    if (ns == 1) return tg * C::L;
    const unsigned t32 = (unsigned)tg;  // sequences per batch item stay below 2^32
    return (long long)(t32 / ns) * ns * C::L + (t32 % ns);
}

// Thread context for one tile (NB sequences) of the pass described by `a`; `tile` selects
// the sequences, `by` the batch item.  Loads of sequences past the end are clamped to the
// last one (their results are never stored).
template <class C, bool IDLE>
__device__ __forceinline__ Ctx<C> make_ctx(const TileArgs& a, unsigned tile, unsigned by, typename C::V* lds)
{
    using V = typename C::V;
    const int tid = threadIdx.x;
    Ctx<C> x;
    x.lds = lds;
    x.a = a;
    x.active = !IDLE || tid < C::WG;
    if (C::LOADT) {
        x.t = tid % C::NB;
        x.jt = tid / C::NB;
    } else {
        x.t = tid / C::TPT;
        x.jt = tid % C::TPT;
    }
    if (IDLE && !x.active) {
        x.t = 0;
        x.jt = 0;
    }
    long long tl;
    if constexpr (C::PAIRED) {
        x.tg = split_seq<C>(a.ntrans, tile, x.t);
        x.valid = x.tg >= 0;
        tl = x.valid ? x.tg : 0;
    } else {
        x.tg = (long long)tile * C::NB + x.t;
        x.valid = x.tg < a.ntrans;
        tl = x.valid ? x.tg : a.ntrans - 1;
    }
    const long long bo = batch_offset(a, by);
    const V* in = reinterpret_cast<const V*>(a.in) + bo;
    V* out = reinterpret_cast<V*>(a.out) + bo;
    x.bo = bo;
    x.in = in + (a.seq_off ? a.seq_off[tl] : tl * a.its);
    x.out = out + out_base<C>(a, tl);
    x.ies = (unsigned)a.ies;
    x.ons = (unsigned)a.ns_pass;
    x.tw = reinterpret_cast<const V*>(a.tw);
    return x;
}

// Transposed store of a tile whose load and store modes differ: the LDS holds the
// natural-order outputs of the NB sequences.
template <class C, bool IDLE> __device__ __forceinline__ void tile_finish(const Ctx<C>& x, unsigned tile)
{
    using V = typename C::V;
    if constexpr (!C::DIRECT) {
        constexpr int PER = C::L / C::TPT;
        const TileArgs& a = x.a;
        const unsigned ns = (unsigned)a.ns_pass;
        const unsigned mask = a.conj_out ? 0x80000000u : 0u;
        const auto sc = (decltype(V{}.x))a.scale;
        const bool scale = a.scale != 1.0;
        const bool nt = a.nt_out != 0;
        const int tid = threadIdx.x;
        V* out = reinterpret_cast<V*>(a.out) + x.bo;
        if constexpr (C::PAIRED) {
            // paired sequences p and S - p (the C2R pre-pass of a first pass): element lin of
            // the tile is element q = lin % L of lane lin / L's sequence, stored by the rule
            if (!IDLE || x.active) {
                with_nt(nt, [&](auto ntc) {
#pragma unroll
                    for (int i = 0; i < PER; ++i) {
                        const int lin = tid + i * C::WG;
                        const int ts = lin / C::L, q = lin % C::L;
                        const long long tgs = split_seq<C>(a.ntrans, tile, ts);
                        if (tgs >= 0) {
                            V u = x.lds[C::lidx(ts, q)];
                            if (scale) u = cscale(u, sc);
                            put<decltype(ntc)::value>(&out[out_base<C>(a, tgs) + (unsigned)q * ns], flip_im(u, mask));
                        }
                    }
                });
            }
            return;
        }
        if (!C::STORET && ns == 1 && a.seq_off == nullptr) {
            // Contiguous sequences (the first pass of a multi-pass transform, batched E-mode
            // output): the tile's NB*L outputs form one contiguous run, and element lin of the
            // run is element q = lin % L of sequence lin / L.  The address needs no per-element
            // store rule, so the stores issue back to back.
            V* o = out + (long long)tile * C::NB * C::L;
            const long long left = a.ntrans - (long long)tile * C::NB;
            const int nvalid = left < C::NB ? (int)left : C::NB;
            if ((!IDLE || x.active) && nvalid == C::NB) {
                // full tile: no guard, so the LDS reads issue together
                V u[PER];
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int lin = tid + i * C::WG;
                    u[i] = x.lds[C::lidx(lin / C::L, lin % C::L)];
                }
                with_nt(nt, [&](auto ntc) {
#pragma unroll
                    for (int i = 0; i < PER; ++i)
                        put<decltype(ntc)::value>(&o[tid + i * C::WG], flip_im(scale ? cscale(u[i], sc) : u[i], mask));
                });
            } else if (!IDLE || x.active) {
                with_nt(nt, [&](auto ntc) {
#pragma unroll
                    for (int i = 0; i < PER; ++i) {
                        const int lin = tid + i * C::WG;
                        const int ts = lin / C::L, q = lin % C::L;
                        if (ts < nvalid) {
                            V u = x.lds[C::lidx(ts, q)];
                            if (scale) u = cscale(u, sc);
                            put<decltype(ntc)::value>(&o[lin], flip_im(u, mask));
                        }
                    }
                });
            }
            return;
        }
        if constexpr (C::STORET) {
            // WG is a multiple of NB, so element lin = tid + i*WG of the tile belongs to
            // sequence tid % NB for every i: the sequence's output base is computed once.
            const int ts = tid % C::NB;
            const long long tgs = (long long)tile * C::NB + ts;
            if ((!IDLE || x.active) && tgs < a.ntrans) {
                V* o = out + out_base<C>(a, tgs);
                with_nt(nt, [&](auto ntc) {
#pragma unroll
                    for (int i = 0; i < PER; ++i) {
                        const int q = (tid + i * C::WG) / C::NB;
                        V u = x.lds[C::lidx(ts, q)];
                        if (scale) u = cscale(u, sc);
                        put<decltype(ntc)::value>(&o[(unsigned)q * ns], flip_im(u, mask));
                    }
                });
            }
            return;
        }
        if (!IDLE || x.active) {
            with_nt(nt, [&](auto ntc) {
#pragma unroll
                for (int i = 0; i < PER; ++i) {
                    const int lin = tid + i * C::WG;
                    const int ts = lin / C::L, q = lin % C::L;
                    const long long tgs = (long long)tile * C::NB + ts;
                    if (tgs < a.ntrans) {
                        V u = x.lds[C::lidx(ts, q)];
                        if (scale) u = cscale(u, sc);
                        put<decltype(ntc)::value>(&out[out_base<C>(a, tgs) + (unsigned)q * ns], flip_im(u, mask));
                    }
                }
            });
        }
    }
}

// R2C split of a SplitCfg tile from the natural-order outputs in LDS (forward only):
//   X[k] = (Z[k] + conj Z[N-k]) / 2 - i W_{2N}^k (Z[k] - conj Z[N-k]) / 2,
// each element written by the thread that stores it, at its own position k = t + q S.
template <class C, bool IDLE> __device__ __forceinline__ void split_finish(const Ctx<C>& x, unsigned tile)
{
    using V = typename C::V;
    using S_ = decltype(V{}.x);
    constexpr int H = C::NB / 2;
    constexpr int PER = C::L * C::NB / C::WG;
    constexpr int QS = C::WG / C::NB;  // q step between a thread's elements
    static_assert(C::L * C::NB % C::WG == 0, "whole elements per thread");
    const TileArgs& a = x.a;
    const long long S = a.ntrans, N = a.nfull;
    const S_ half = (S_)0.5;
    V* out = reinterpret_cast<V*>(a.out) + x.bo;
    if (IDLE && !x.active) return;
    // WG is a multiple of NB: element i of this thread is element q = q0 + i*QS of sequence ts
    // (T-mode store order), so its k = tgs + q*S steps by QS*S and the twiddles W_{2N}^k form
    // a geometric sequence: an FP64 recurrence from two table roots, and W_{2N}^{N-k} =
    // -conj(W_{2N}^k) for the mirrored half.
    const int ts = (int)threadIdx.x % C::NB, q0 = (int)threadIdx.x / C::NB;
    const long long tgs = split_seq<C>(S, tile, ts);
    if (tgs < 0) return;
    const int pl = tile == 0 ? ts : (ts < H ? ts + H : ts - H);
    const long long mask = (1LL << a.sp_shift) - 1;
    const long long k0 = tgs + (long long)q0 * S, kst = (long long)QS * S;
    double2 w = dmul(a.sp_lo[k0 & mask], a.sp_hi[k0 >> a.sp_shift]);
    const double2 step = dmul(a.sp_lo[kst & mask], a.sp_hi[kst >> a.sp_shift]);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int q = q0 + i * QS;
        const int qp = tgs == 0 ? (q == 0 ? 0 : C::L - q) : C::L - 1 - q;
        const long long k = tgs + (long long)q * S;
        const V zk = x.lds[C::lidx(ts, q)], zp = x.lds[C::lidx(pl, qp)];
        if (k == 0) {
            const S_ x0 = zk.x + zk.y, xn = zk.x - zk.y;
            if (a.split_bins) {
                out[0] = V{x0, 0};
                out[N] = V{xn, 0};
            } else {
                out[0] = V{x0, xn};
            }
        } else {
            const bool lo = k <= N - k;
            const V A = lo ? zk : zp, B = lo ? zp : zk;
            const V wv = to_v<V>(lo ? w : double2{-w.x, w.y});  // W^k, or W^{N-k} = -conj(W^k)
            const V h1 = V{half * (A.x + B.x), half * (A.y - B.y)};
            const V dd = V{half * (A.x - B.x), half * (A.y + B.y)};
            const V p = cmul(wv, V{dd.y, -dd.x});
            out[k] = lo ? V{h1.x + p.x, h1.y + p.y} : V{h1.x - p.x, p.y - h1.y};
        }
        w = dmul(w, step);
    }
}

// R2C split of an E->E tile (REAL_POST, single pass): the LDS holds each sequence's Z in
// natural order and the sequence is the whole transform, so X[k] takes Z[k] and Z[N-k] from
// the same sequence.  Thread (t, i0) writes the elements k = i0 + i*TPT of sequence t; their
// twiddles W_{2N}^k step by W_{2N}^TPT (FP64 recurrence from two table roots).
template <class C, bool IDLE> __device__ __forceinline__ void split_finish_e(const Ctx<C>& x)
{
    using V = typename C::V;
    using S_ = decltype(V{}.x);
    constexpr int L = C::L, PER = L / C::TPT;
    static_assert(L % C::TPT == 0, "whole elements per thread");
    const TileArgs& a = x.a;
    if ((IDLE && !x.active) || !x.valid) return;
    const S_ half = (S_)0.5;
    V* out = x.out;
    const int i0 = x.jt;
    const long long mask = (1LL << a.sp_shift) - 1;
    double2 w = dmul(a.sp_lo[i0 & mask], a.sp_hi[i0 >> a.sp_shift]);
    const double2 step = dmul(a.sp_lo[C::TPT & mask], a.sp_hi[C::TPT >> a.sp_shift]);
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int k = i0 + i * C::TPT;
        const V zk = x.lds[C::lidx(x.t, k)], zp = x.lds[C::lidx(x.t, k == 0 ? 0 : L - k)];
        if (k == 0) {
            const S_ x0 = zk.x + zk.y, xn = zk.x - zk.y;
            if (a.split_bins) {
                out[0] = V{x0, 0};
                out[L] = V{xn, 0};
            } else {
                out[0] = V{x0, xn};
            }
        } else {
            const bool lo = k <= L - k;
            const V A = lo ? zk : zp, B = lo ? zp : zk;
            const V wv = to_v<V>(lo ? w : double2{-w.x, w.y});  // W^k, or W^{N-k} = -conj(W^k)
            const V h1 = V{half * (A.x + B.x), half * (A.y - B.y)};
            const V dd = V{half * (A.x - B.x), half * (A.y + B.y)};
            const V p = cmul(wv, V{dd.y, -dd.x});
            out[k] = lo ? V{h1.x + p.x, h1.y + p.y} : V{h1.x - p.x, p.y - h1.y};
        }
        w = dmul(w, step);
    }
}

// C2R pre-pass of a REAL_PRE tile, between the loads and the first stage: every loaded X[e]
// goes to LDS, each thread reads the partner X[N-e] of each of its elements and replaces
// X[e] by Z[e], the input of the inverse complex transform:
//   h1 = (A + conj B)/2,  d = (A - conj B)/2,  Z[k] = h1 + i conj(W_{2N}^k) d,
//   Z[N-k] = conj(h1 - i conj(W_{2N}^k) d)      (A = X[k], B = X[N-k], k <= N-k)
//   Z[0] = ((X_0 + X_N)/2, (X_0 - X_N)/2)
// with X_N in the imaginary part of element 0 (NR packing) or at element N (bins).  These are
// the relations of the split kernel's inverse branch.  E->E tiles: the partner is element
// N-e of the same sequence.  Paired T->E tiles (first pass, element q of sequence t is
// e = t + q S): the partner of (t, q) is (S - t, L - 1 - q), or (0, L - q) for t = 0.
template <class C, bool IDLE> __device__ __forceinline__ void pre_split(const Ctx<C>& x, FirstIn<C>& f)
{
    using V = typename C::V;
    using S_ = decltype(V{}.x);
    constexpr int R = C::R[0];
    constexpr int B = C::nbf(0);
    constexpr int LR = C::L / R;
    constexpr int H = C::NB / 2;
    constexpr bool GUARD = (B * C::TPT > LR);
    const TileArgs& a = x.a;
    const bool act = !IDLE || x.active;
#pragma unroll
    for (int b = 0; b < B; ++b) {
        const int j = x.jt + b * C::TPT;
        if (!act || (GUARD && j >= LR)) continue;
#pragma unroll
        for (int r = 0; r < R; ++r) x.lds[C::lidx(x.t, j + r * LR)] = f.v[b][r];
    }
    __syncthreads();
    if (act && x.valid) {
        const S_ half = (S_)0.5;
        const long long N = a.nfull;
        // lane of the partner sequence, the sequence's first element e0 and the element step
        int pl = x.t;
        long long e0 = 0, es = 1;
        if constexpr (C::PAIRED) {
            const unsigned tile_lane = (unsigned)x.t;
            pl = (x.tg == 0 || 2 * x.tg == a.ntrans) ? x.t : (tile_lane < (unsigned)H ? x.t + H : x.t - H);
            e0 = x.tg;
            es = a.ntrans;
        }
        const long long mask = (1LL << a.sp_shift) - 1;
        const long long rst = (long long)LR * es;  // e steps by LR S between a thread's r
        const double2 step = dmul(a.sp_lo[rst & mask], a.sp_hi[rst >> a.sp_shift]);
#pragma unroll
        for (int b = 0; b < B; ++b) {
            const int j = x.jt + b * C::TPT;
            if (GUARD && j >= LR) continue;
            const long long eb = e0 + (long long)j * es;
            double2 w = dmul(a.sp_lo[eb & mask], a.sp_hi[eb >> a.sp_shift]);
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int q = j + r * LR;
                const long long e = eb + (long long)r * rst;
                int qp;
                if constexpr (C::PAIRED) qp = x.tg == 0 ? (q == 0 ? 0 : C::L - q) : C::L - 1 - q;
                else qp = q == 0 ? 0 : C::L - q;
                const V xe = f.v[b][r], xp = x.lds[C::lidx(pl, qp)];
                V z;
                if (e == 0) {
                    const S_ x0 = xe.x;
                    // e == 0 is element 0 of sequence 0, so x.in is the item's start
                    const S_ xn = a.split_bins ? x.in[N].x : xe.y;
                    z = V{half * (x0 + xn), half * (x0 - xn)};
                } else {
                    const bool lo = e <= N - e;
                    const V A = lo ? xe : xp, Bv = lo ? xp : xe;
                    // conj(W^k) for k = min(e, N - e): conj(W^{N-e}) = -W^e
                    const V cw = to_v<V>(lo ? double2{w.x, -w.y} : double2{-w.x, -w.y});
                    const V h1 = V{half * (A.x + Bv.x), half * (A.y - Bv.y)};
                    const V dd = V{half * (A.x - Bv.x), half * (A.y + Bv.y)};
                    const V t = cmul(cw, dd);
                    const V p = V{-t.y, t.x};  // i conj(W^k) d
                    z = lo ? V{h1.x + p.x, h1.y + p.y} : V{h1.x - p.x, p.y - h1.y};
                }
                f.v[b][r] = z;
                w = dmul(w, step);
            }
        }
    }
    __syncthreads();  // the first stage writes the LDS next
}

// One tile, loaded and transformed.  Callable from any kernel whose workgroup has at least
// C::WG threads (IDLE: more than C::WG, the extra ones only join the barriers) and which
// provides C::LDS_ELEMS elements of LDS.
template <class C, bool IDLE = false>
__device__ __forceinline__ void tile_body(const TileArgs& a, unsigned tile, unsigned by, typename C::V* lds)
{
    Ctx<C> x = make_ctx<C, IDLE>(a, tile, by, lds);
    FirstIn<C> f;
    load_first<C, IDLE>(x, f);
    if constexpr (C::REAL == REAL_PRE) pre_split<C, IDLE>(x, f);
    stage_loop<C, IDLE, 0>(x, f);
    if constexpr (C::REAL == REAL_POST && C::PAIRED) split_finish<C, IDLE>(x, tile);
    else if constexpr (C::REAL == REAL_POST) split_finish_e<C, IDLE>(x);
    else tile_finish<C, IDLE>(x, tile);
}

template <class C> __global__ __launch_bounds__(C::WG) void tile_kernel(TileArgs a)
{
    using V = typename C::V;
    __shared__ V lds[C::LDS_ELEMS];
    tile_body<C, false>(a, tile_index(blockIdx.x, gridDim.x, a.xcd), blockIdx.y, lds);
}

}  // namespace gfftx
