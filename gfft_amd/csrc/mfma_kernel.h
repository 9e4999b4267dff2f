// mfma_kernel.h -- FP32 FFT of length 16^S whose radix-16 stages run on the matrix cores.
//
// Stage s of the Stockham sweep computes, for each of the L/16 columns j (k = j mod Ns),
//     Y[q][j] = sum_r F16[q][r] * (x[j + r L/16] * W_{16 Ns}^{k r}),
// i.e. a 16x16 complex DFT matrix times a 16 x (L/16) complex matrix.  In real form the DFT
// matrix is 32x32; with the K index interleaved as (Re v_r, Im v_r) = k = 2r, 2r+1:
//     rows of Re Y:  A0[q][2r] =  Re F[q][r],  A0[q][2r+1] = -Im F[q][r]
//     rows of Im Y:  A1[q][2r] =  Im F[q][r],  A1[q][2r+1] =  Re F[q][r].
// Each 16-column block is two v_mfma_f32_16x16x32_f16 products (A0 and A1 against one
// B operand).  FP32 accuracy is recovered with a hi/lo split of both operands,
//     A B ~= Ah Bh + Ah Bl + Al Bh        (FP32 accumulation, smallest terms first),
// with the data scaled by powers of two (exact) so that the FP16 parts stay in range:
//   - per stage (MFMA_SCALE_STAGE, the default): the first stage scales the transform by
//     s0 = 2^(14-e), e the exponent of its largest component, so every input component is
//     below 2^14; a radix-16 DFT grows a component at most 16 sqrt(2) times, so each later
//     stage takes its input times 2^-4 (below 2^14.5 and 2^15 for the second and third
//     stages), folded into its twiddle table; the last store multiplies by
//     2^(4 (stages-1)) / s0.  Per point and stage this costs the split alone (the scale
//     rides on the twiddle product);
//   - per column (MFMA_SCALE_COLUMN): every column of every stage is scaled so that its
//     largest entry sits just below 2^14 and unscaled after the product.
//
// Lane layout of the 16x16x32 f16 product on gfx950 (lane l, i = 0..7 / 0..3):
//     A: A[l%16][8(l/16)+i]   B: B[8(l/16)+i][l%16]   C: C[4(l/16)+i][l%16]
// so lane (c = l%16, g = l/16) of a block owns column j = 16*block + c with inputs
// r = 4g..4g+3 and, after the product, outputs q = 4g..4g+3 of the same column.
#pragma once
#include "tile_kernel.h"

namespace gfftx {

typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef _Float16 half2_t __attribute__((ext_vector_type(2)));
typedef float float2_t __attribute__((ext_vector_type(2)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// A operands in device memory: [mt][part][lane][8] with mt = 0 (Re rows) / 1 (Im rows) and
// part = 0 (hi) / 1 (lo).
constexpr int MFMA_A_HALFS = 2 * 2 * 64 * 8;

enum { MFMA_SCALE_STAGE = 0, MFMA_SCALE_COLUMN = 1 };

template <int L_, int NB_, int SCALE_ = MFMA_SCALE_STAGE, int WG_ = 256> struct MfmaCfg {
    static constexpr int L = L_, NB = NB_, WG = WG_;
    static constexpr int SCALE = SCALE_;
    static constexpr int WAVES = WG / 64;
    static constexpr int COLS = L / 16;                 // columns per sequence and stage
    static constexpr int BLOCKS = NB * COLS / 16;       // 16-column blocks per stage (all sequences)
    static constexpr int BPW = BLOCKS / WAVES;          // blocks per wave
    static constexpr int NS = (L == 256) ? 2 : (L == 4096) ? 3 : (L == 65536) ? 4 : 0;
    static_assert(NS > 0, "length must be 16^S");
    static_assert(BLOCKS % WAVES == 0 && BPW >= 1, "the waves share the blocks evenly");
    // the wave's four blocks are consecutive columns of one sequence (the last stage's Ns is
    // L/16 = COLS, so it writes element (q, j) to q COLS + j): the first loads and the last
    // stores then move whole rows of the wave's 64 columns
    static constexpr bool ROWS = BPW == 4 && COLS / 16 >= BPW;
    static constexpr int LP = L + L / 16;
    static constexpr int TS = (LP % 2 == 0) ? LP + 1 : LP;
    static constexpr int LDS_ELEMS = NB * TS;
    __device__ static constexpr int lidx(int t, int e) { return t * TS + e + (e >> 4); }
    static constexpr int ns(int s)
    {
        int n = 1;
        for (int i = 0; i < s; ++i) n *= 16;
        return n;
    }
    static constexpr int twoff(int s)
    {
        int o = 0;
        for (int i = 1; i < s; ++i) o += ns(i) * 16;
        return o;
    }
};

// Twiddle entries the kernel keeps in LDS.  Stage 1 (Ns = 16) multiplies row m of column j
// by W_256^{(j mod 16) m}; stage 2 (Ns = 256) by W_4096^{j m} = W_256^{jb m} W_4096^{c m} for
// j = 16 jb + c (c, m < 16), the product of two table entries in FP32.
constexpr int MFMA_TW = 512;

struct MfmaArgs {
    const float2* in;
    float2* out;
    long long ntrans;
    const float2* tw;       // MFMA_TW entries: W_256^m (m < 256), times 2^-4 for MFMA_SCALE_STAGE
                            // kernels, then W_4096^m (m < 256); the kernel keeps them in LDS
    const _Float16* amat;   // MFMA_A_HALFS halves
    int conj_in;
    int conj_out;
    float scale;
    int nt_out;             // 1: the output is stored non-temporally (put())
};

__device__ __forceinline__ float col_max4(float m)
{
    // Lanes c, c+16, c+32, c+48 hold the same column.  m >= 0, so the order of its bit
    // patterns as unsigned integers is the order of the values; the two lane swaps
    // (gfx950 v_permlane32_swap / v_permlane16_swap) exchange halves of the wave on the
    // VALU, without a round trip through LDS.
    unsigned u = __float_as_uint(m);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = max(a[0], a[1]);  // max(m[l], m[l ^ 32])
    const auto b = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = max(b[0], b[1]);  // and of m[l ^ 16], m[l ^ 48]
    return __uint_as_float(u);
}

// Largest value over the wave (m >= 0: compared as unsigned bit patterns).
__device__ __forceinline__ float wave_max(float m)
{
    unsigned u = __float_as_uint(m);
    const auto a = __builtin_amdgcn_permlane32_swap(u, u, false, false);
    u = max(a[0], a[1]);
    const auto b = __builtin_amdgcn_permlane16_swap(u, u, false, false);
    u = max(b[0], b[1]);
#pragma unroll
    for (int d = 8; d >= 1; d >>= 1) u = max(u, (unsigned)__shfl_xor((int)u, d, 64));
    return __uint_as_float(u);
}

// 4 x 4 transpose between the lane bits 4-5 (g = lane >> 4) and the index of v: lane (c, g)
// ends with v[b] = the old v[g] of lane (c, b).  gfx950's v_permlane16_swap / v_permlane32_swap
// exchange two registers between the lanes whose bit 4 / bit 5 differ, which is one step of
// the transpose by itself (no selects).
template <int D> __device__ __forceinline__ void swap_lanes(float2& u, float2& w)
{
    int ux = __float_as_int(u.x), wx = __float_as_int(w.x), uy = __float_as_int(u.y), wy = __float_as_int(w.y);
    if constexpr (D == 16) {
        const auto rx = __builtin_amdgcn_permlane16_swap(ux, wx, false, false);
        const auto ry = __builtin_amdgcn_permlane16_swap(uy, wy, false, false);
        ux = rx[0];
        wx = rx[1];
        uy = ry[0];
        wy = ry[1];
    } else {
        const auto rx = __builtin_amdgcn_permlane32_swap(ux, wx, false, false);
        const auto ry = __builtin_amdgcn_permlane32_swap(uy, wy, false, false);
        ux = rx[0];
        wx = rx[1];
        uy = ry[0];
        wy = ry[1];
    }
    u = float2{__int_as_float(ux), __int_as_float(uy)};
    w = float2{__int_as_float(wx), __int_as_float(wy)};
}

__device__ __forceinline__ void transpose_lane_block(float2& v0, float2& v1, float2& v2, float2& v3)
{
    swap_lanes<16>(v0, v1);  // lane bit 4 with index bit 0
    swap_lanes<16>(v2, v3);
    swap_lanes<32>(v0, v2);  // lane bit 5 with index bit 1
    swap_lanes<32>(v1, v3);
}

// s0 = 2^(14 - e) for the largest component m = f 2^e (f in [0.5, 1)) of a transform, the
// exponent clamped so that s0 and its inverse are normal FP32 numbers; returns log2 s0.
__device__ __forceinline__ int stage_shift(float m)
{
    int ex = 0;
    (void)frexpf(m, &ex);
    return min(max(14 - ex, -126), 126);
}

template <class C, int S>
__device__ __forceinline__ void mfma_stage(const MfmaArgs& a, float2* lds, float2* twl, const half8_t (&A)[2][2],
                                           long long t0, int (&sh)[C::BPW])
{
    constexpr int NSS = C::ns(S);
    constexpr bool FIRST = (S == 0), LAST = (S == C::NS - 1);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int c = lane & 15, g = lane >> 4;
    float2 y[C::BPW][4];
    // first stage of per-stage scaling: all inputs are loaded first, then the largest
    // component of each transform sets its scale
    float2 x0[FIRST ? C::BPW : 1][4];
    if constexpr (FIRST) {
        if constexpr (C::ROWS) {
            // lane (c, g) loads column 16 g + c of the wave's 64 for rows 4 b + i (every load
            // a contiguous 512-byte row piece); the transpose between g and b then gives it
            // column 16 b + c of block b for rows 4 g + i, the layout of the products
            const int blk0 = wave * C::BPW;
            const long long tg = t0 + blk0 / (C::COLS / 16);
            const int j0 = (blk0 % (C::COLS / 16)) * 16;
            const float2* src = a.in + (tg < a.ntrans ? tg : a.ntrans - 1) * C::L;
#pragma unroll
            for (int b = 0; b < C::BPW; ++b)
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    x0[b][i] = nt_load(&src[(unsigned)(j0 + 16 * g + c + (4 * b + i) * C::COLS)]);
#pragma unroll
            for (int i = 0; i < 4; ++i) transpose_lane_block(x0[0][i], x0[1][i], x0[2][i], x0[3][i]);
        } else {
#pragma unroll
            for (int b = 0; b < C::BPW; ++b) {
                const int blk = wave * C::BPW + b;
                const int t = blk / (C::COLS / 16);
                const int j = (blk % (C::COLS / 16)) * 16 + c;
                const long long tg = t0 + t;
                // loads past the last sequence are clamped to it; their results are not
                // stored.  The input is read once, so it is loaded non-temporally.
                const float2* src = a.in + (tg < a.ntrans ? tg : a.ntrans - 1) * C::L;
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    x0[b][i] = nt_load(&src[(unsigned)(j + (4 * g + i) * C::COLS)]);
            }
        }
        // the twiddle table is loaded behind the sequences, so that a workgroup issues its
        // data loads first, and stored to the LDS at the end of this block; the later stages
        // read it behind the barrier that ends this stage
        static_assert(MFMA_TW % C::WG == 0, "whole rounds of the table copy");
        float2 tv[MFMA_TW / C::WG];
#pragma unroll
        for (int k = 0; k < MFMA_TW / C::WG; ++k) tv[k] = a.tw[threadIdx.x + k * C::WG];
        if constexpr (C::SCALE == MFMA_SCALE_STAGE) {
            // blocks per sequence: COLS/16; a wave holds BPW blocks, so a sequence spans
            // (COLS/16)/BPW waves (one: the wave's blocks are whole sequences)
            constexpr int BPS = C::COLS / 16;
            if constexpr (BPS >= C::BPW) {
                static_assert(BPS % C::BPW == 0, "whole waves per sequence");
                float m = 0.f;
#pragma unroll
                for (int b = 0; b < C::BPW; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i) m = fmaxf(fmaxf(m, fabsf(x0[b][i].x)), fabsf(x0[b][i].y));
                m = wave_max(m);
                constexpr int WPS = BPS / C::BPW;  // waves per sequence
                if constexpr (WPS > 1) {
                    // the waves of a sequence meet in the LDS, before the stage writes it
                    float* red = reinterpret_cast<float*>(lds);
                    if (lane == 0) red[wave] = m;
                    __syncthreads();
                    const int w0 = (wave / WPS) * WPS;
#pragma unroll
                    for (int w = 0; w < WPS; ++w) m = fmaxf(m, red[w0 + w]);
                    __syncthreads();
                }
                const int e = stage_shift(m);
#pragma unroll
                for (int b = 0; b < C::BPW; ++b) sh[b] = e;
            } else {
                // several sequences per wave, a block each: the largest component of each block
                static_assert(BPS == 1, "a block is a whole sequence");
#pragma unroll
                for (int b = 0; b < C::BPW; ++b) {
                    float m = 0.f;
#pragma unroll
                    for (int i = 0; i < 4; ++i) m = fmaxf(fmaxf(m, fabsf(x0[b][i].x)), fabsf(x0[b][i].y));
                    sh[b] = stage_shift(wave_max(m));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < MFMA_TW / C::WG; ++k) twl[threadIdx.x + k * C::WG] = tv[k];
    }

    half8_t bh[C::BPW], bl[C::BPW];  // B operands of every block (hi and lo parts)
    float invs[C::BPW];              // per-column scaling: each block's unscale
    // the twiddle factors that do not depend on the block, read once: W_256^{c m} (stage 1),
    // W_4096^{c m} (stage 2)
    float2 wc[4];
    if constexpr (!FIRST) {
#pragma unroll
        for (int i = 0; i < 4; ++i) wc[i] = twl[(S == 1 ? 0 : 256) + c * (4 * g + i)];
    }
#pragma unroll
    for (int b = 0; b < C::BPW; ++b) {
        const int blk = wave * C::BPW + b;             // block index over all NB sequences
        const int t = blk / (C::COLS / 16);            // sequence within the workgroup
        const int j = (blk % (C::COLS / 16)) * 16 + c; // column
        const long long tg = t0 + t;
        float2 v[4];
        if constexpr (FIRST) {
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = x0[b][i];
        } else {
            // lidx(t, j + m COLS) with m = 4g + i: COLS is a multiple of 16 and j = 16 jb + c,
            // so the padding term (e >> 4) is jb + m COLS/16 and the rows are RS apart
            constexpr int RS = C::COLS + C::COLS / 16;
            const float2* rd = lds + t * C::TS + j + (j >> 4) + 4 * g * RS;
#pragma unroll
            for (int i = 0; i < 4; ++i) v[i] = rd[i * RS];
        }
        if constexpr (!FIRST) {
            static_assert(S <= 2, "lengths up to 4096");
            float2 w[4];
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int m = 4 * g + i;
                if constexpr (S == 1) {
                    w[i] = wc[i];
                } else {
                    w[i] = cmul(twl[((j >> 4) * m) & 255], wc[i]);
                }
            }
            if constexpr (C::SCALE == MFMA_SCALE_STAGE) {
                // the entries carry the stage's 2^-4, so W^0 is not 1 here
#pragma unroll
                for (int i = 0; i < 4; ++i) v[i] = cmul(v[i], w[i]);
            } else {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    if (g != 0 || i != 0) v[i] = cmul(v[i], w[i]);
            }
        }
        float sc = 1.0f, inv = 1.0f;
        if constexpr (C::SCALE == MFMA_SCALE_COLUMN) {
            // per-column power-of-two scale: largest entry just below 2^14 (the exponent is
            // clamped so that the scale and its inverse are normal FP32 numbers)
            float m = 0.f;
#pragma unroll
            for (int i = 0; i < 4; ++i) m = fmaxf(fmaxf(m, fabsf(v[i].x)), fabsf(v[i].y));  // one v_max3 per point
            m = col_max4(m);
            int ex = 0;
            (void)frexpf(m, &ex);
            const int csh = min(max(14 - ex, -126), 126);
            sc = ldexpf(1.0f, csh);
            inv = ldexpf(1.0f, -csh);
        } else if constexpr (FIRST) {
            sc = ldexpf(1.0f, sh[b]);
        } else {
            // s = 1, opaque to the compiler: x - hi then stays one mixed-precision FMA per
            // component instead of a conversion of hi and a subtraction
            asm volatile("" : "+v"(sc));
        }
        invs[b] = inv;
        // the inverse's conjugation of the input rides on the first stage's scale of Im
        const float scy = (FIRST && a.conj_in) ? -sc : sc;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            // hi = FP16(x s); lo = FP16(x s - hi): x s - hi is formed by one mixed-precision
            // FMA from x, s and the FP16 hi (the product by a power of two is exact; s = 1
            // after the first stage of per-stage scaling)
            const float2_t xs = (C::SCALE == MFMA_SCALE_STAGE && !FIRST) ? float2_t{v[i].x, v[i].y}
                                                                         : float2_t{v[i].x * sc, v[i].y * scy};
            const half2_t h = __builtin_convertvector(xs, half2_t);
            bh[b][2 * i] = h.x;
            bh[b][2 * i + 1] = h.y;
            bl[b][2 * i] = (_Float16)__builtin_fmaf(v[i].x, sc, -(float)h.x);
            bl[b][2 * i + 1] = (_Float16)__builtin_fmaf(v[i].y, scy, -(float)h.y);
        }
    }
    // the products of all blocks, each round over every block before the next, so that no
    // product waits for the one before it: A B ~= Al Bh + Ah Bl + Ah Bh (smallest terms
    // first) for the Re (mt = 0) and Im (mt = 1) rows
    float4_t acc[C::BPW][2];
#pragma unroll
    for (int b = 0; b < C::BPW; ++b)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt)
            acc[b][mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[mt][1], bh[b], float4_t{0.f, 0.f, 0.f, 0.f}, 0, 0, 0);
#pragma unroll
    for (int b = 0; b < C::BPW; ++b)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) acc[b][mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[mt][0], bl[b], acc[b][mt], 0, 0, 0);
#pragma unroll
    for (int b = 0; b < C::BPW; ++b)
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) acc[b][mt] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A[mt][0], bh[b], acc[b][mt], 0, 0, 0);
#pragma unroll
    for (int b = 0; b < C::BPW; ++b) {
        if constexpr (C::SCALE == MFMA_SCALE_COLUMN) {
#pragma unroll
            for (int i = 0; i < 4; ++i) y[b][i] = float2{acc[b][0][i] * invs[b], acc[b][1][i] * invs[b]};
        } else {
#pragma unroll
            for (int i = 0; i < 4; ++i) y[b][i] = float2{acc[b][0][i], acc[b][1][i]};
        }
    }

    if constexpr (!FIRST) __syncthreads();

    if constexpr (LAST && C::ROWS) {
        // the transpose between g and b gives lane (c, g) column 16 g + c of the wave's 64
        // for rows 4 b + i: every store writes a contiguous 512-byte row piece
        const int blk0 = wave * C::BPW;
        const long long tg = t0 + blk0 / (C::COLS / 16);
        const int j0 = (blk0 % (C::COLS / 16)) * 16;
#pragma unroll
        for (int i = 0; i < 4; ++i) transpose_lane_block(y[0][i], y[1][i], y[2][i], y[3][i]);
        if (tg < a.ntrans) {
            float2* dst = a.out + tg * C::L;
            // per-stage scaling: undo s0 and the 2^-4 of each later stage (one s0 per
            // sequence); the inverse's conjugation of the output rides on the scale of Im
            const float sc = C::SCALE == MFMA_SCALE_STAGE ? ldexpf(a.scale, 4 * (C::NS - 1) - sh[0]) : a.scale;
            const float scy = a.conj_out ? -sc : sc;
            with_nt(a.nt_out != 0, [&](auto ntc) {
#pragma unroll
                for (int b = 0; b < C::BPW; ++b)
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        put<decltype(ntc)::value>(&dst[(unsigned)(j0 + 16 * g + c + (4 * b + i) * C::COLS)],
                                                  float2{y[b][i].x * sc, y[b][i].y * scy});
            });
        }
        return;
    }

#pragma unroll
    for (int b = 0; b < C::BPW; ++b) {
        const int blk = wave * C::BPW + b;
        const int t = blk / (C::COLS / 16);
        const int j = (blk % (C::COLS / 16)) * 16 + c;
        const int k = j % NSS;
        const int base = (j / NSS) * NSS * 16 + k;
        if constexpr (LAST) {
            const long long tg = t0 + t;
            if (tg < a.ntrans) {
                float2* dst = a.out + tg * C::L;
                // per-stage scaling: undo s0 and the 2^-4 of each later stage
                // (the inverse's conjugation of the output rides on the scale of Im)
                const float sc = C::SCALE == MFMA_SCALE_STAGE ? ldexpf(a.scale, 4 * (C::NS - 1) - sh[b]) : a.scale;
                const float scy = a.conj_out ? -sc : sc;
                with_nt(a.nt_out != 0, [&](auto ntc) {
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        put<decltype(ntc)::value>(&dst[(unsigned)(base + (4 * g + i) * NSS)],
                                                  float2{y[b][i].x * sc, y[b][i].y * scy});
                });
            }
        } else {
            // lidx(t, base + m NSS), m = 4g + i: for NSS = 1, base = 16 j and m < 16; for NSS
            // a multiple of 16 the padding term grows by NSS/16 per m.  Either way the
            // elements are WS apart from t TS + base + (base >> 4)
            constexpr int WS = NSS == 1 ? 1 : NSS + NSS / 16;
            static_assert(NSS == 1 || NSS % 16 == 0, "radix-16 stages");
            float2* wr = lds + t * C::TS + base + (base >> 4) + 4 * g * WS;
#pragma unroll
            for (int i = 0; i < 4; ++i) wr[i * WS] = y[b][i];
        }
    }
    if constexpr (!LAST) __syncthreads();
}

template <class C, int S>
__device__ __forceinline__ void mfma_loop(const MfmaArgs& a, float2* lds, float2* twl, const half8_t (&A)[2][2],
                                          long long t0, int (&sh)[C::BPW])
{
    if constexpr (S < C::NS) {
        mfma_stage<C, S>(a, lds, twl, A, t0, sh);
        mfma_loop<C, S + 1>(a, lds, twl, A, t0, sh);
    }
}

template <class C> __global__ __launch_bounds__(C::WG) void mfma_fft_kernel(MfmaArgs a)
{
    __shared__ float2 lds[C::LDS_ELEMS];
    __shared__ float2 twl[MFMA_TW];  // filled by the first stage, read by the later ones
    const int lane = threadIdx.x & 63;
    half8_t A[2][2];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int i = 0; i < 8; ++i) A[mt][p][i] = a.amat[((mt * 2 + p) * 64 + lane) * 8 + i];
    int sh[C::BPW] = {};  // log2 of each block's first-stage scale (per-stage scaling)
    mfma_loop<C, 0>(a, lds, twl, A, (long long)blockIdx.x * C::NB, sh);
}

}  // namespace gfftx
