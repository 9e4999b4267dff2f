// kernels.hip -- instantiated FFT kernels for gfx950 and their host-visible registry.
#include "tile_table.h"

namespace gfftx {

// Batched single-pass (E->E) tiles.  TPT = L/16 threads per sequence (one radix-16
// butterfly each), NB sequences per workgroup with NB*TPT close to 256.  Radix-2/4/8/3/5
// stages spread their butterflies over the same TPT threads.
#define GFFTX_BATCH_POW2(PREC, TYPE)                                         \
    GFFTX_TILE(PREC, TYPE, 2, 64, 64, false, false, 2),                     \
    GFFTX_TILE(PREC, TYPE, 4, 64, 64, false, false, 4),                     \
    GFFTX_TILE(PREC, TYPE, 8, 64, 64, false, false, 8),                     \
    GFFTX_TILE(PREC, TYPE, 16, 64, 64, false, false, 16),                   \
    GFFTX_TILE(PREC, TYPE, 32, 64, 128, false, false, 16, 2),               \
    GFFTX_TILE(PREC, TYPE, 64, 64, 256, false, false, 16, 4),               \
    GFFTX_TILE(PREC, TYPE, 128, 32, 256, false, false, 16, 8),              \
    GFFTX_TILE(PREC, TYPE, 256, 16, 256, false, false, 16, 16),             \
    GFFTX_TILE(PREC, TYPE, 512, 8, 256, false, false, 16, 16, 2),           \
    GFFTX_TILE(PREC, TYPE, 1024, 4, 256, false, false, 16, 16, 4),          \
    GFFTX_TILE(PREC, TYPE, 2048, 2, 256, false, false, 16, 16, 8),          \
    GFFTX_TILE(PREC, TYPE, 4096, 1, 256, false, false, 16, 16, 16),         \
    GFFTX_TILE(PREC, TYPE, 8192, 1, 512, false, false, 16, 16, 16, 2)

#define GFFTX_BATCH_MIXED(PREC, TYPE)                                        \
    GFFTX_TILE(PREC, TYPE, 48, 85, 255, false, false, 16, 3),               \
    GFFTX_TILE(PREC, TYPE, 80, 51, 255, false, false, 16, 5),               \
    GFFTX_TILE(PREC, TYPE, 96, 42, 252, false, false, 16, 2, 3),            \
    GFFTX_TILE(PREC, TYPE, 160, 25, 250, false, false, 16, 2, 5),           \
    GFFTX_TILE(PREC, TYPE, 192, 21, 252, false, false, 16, 4, 3),           \
    GFFTX_TILE(PREC, TYPE, 320, 12, 240, false, false, 16, 4, 5),           \
    GFFTX_TILE(PREC, TYPE, 384, 10, 240, false, false, 16, 8, 3),           \
    GFFTX_TILE(PREC, TYPE, 640, 6, 240, false, false, 16, 8, 5),            \
    GFFTX_TILE(PREC, TYPE, 768, 5, 240, false, false, 16, 16, 3),           \
    GFFTX_TILE(PREC, TYPE, 1280, 3, 240, false, false, 16, 16, 5),          \
    GFFTX_TILE(PREC, TYPE, 1536, 2, 192, false, false, 16, 16, 2, 3),       \
    GFFTX_TILE(PREC, TYPE, 2560, 2, 320, false, false, 16, 16, 2, 5),       \
    GFFTX_TILE(PREC, TYPE, 3072, 2, 384, false, false, 16, 16, 4, 3),       \
    GFFTX_TILE(PREC, TYPE, 5120, 1, 320, false, false, 16, 16, 4, 5),       \
    GFFTX_TILE(PREC, TYPE, 6144, 1, 384, false, false, 16, 16, 8, 3)

// Pass tiles carrying the factors 3 and 5 (T->E and T->T), so that every 2^a 3^b 5^c
// length splits into compiled passes.
#define GFFTX_PASS_MIXED(PREC, TYPE)                                          \
    GFFTX_TILE(PREC, TYPE, 48, 85, 255, true, false, 16, 3),                 \
    GFFTX_TILE(PREC, TYPE, 48, 85, 255, true, true, 16, 3),                  \
    GFFTX_TILE(PREC, TYPE, 80, 51, 255, true, false, 16, 5),                 \
    GFFTX_TILE(PREC, TYPE, 80, 51, 255, true, true, 16, 5),                  \
    GFFTX_TILE(PREC, TYPE, 192, 21, 252, true, false, 16, 4, 3),             \
    GFFTX_TILE(PREC, TYPE, 192, 21, 252, true, true, 16, 4, 3),              \
    GFFTX_TILE(PREC, TYPE, 320, 12, 240, true, false, 16, 4, 5),             \
    GFFTX_TILE(PREC, TYPE, 320, 12, 240, true, true, 16, 4, 5),              \
    GFFTX_TILE(PREC, TYPE, 768, 5, 240, true, false, 16, 16, 3),             \
    GFFTX_TILE(PREC, TYPE, 768, 5, 240, true, true, 16, 16, 3),              \
    GFFTX_TILE(PREC, TYPE, 1280, 3, 240, true, false, 16, 16, 5),            \
    GFFTX_TILE(PREC, TYPE, 1280, 3, 240, true, true, 16, 16, 5)

// Small pass tiles (one or two stages) that absorb leftover factors in one extra pass.
#define GFFTX_PASS_SMALL_ONE(PREC, TYPE, ST)                                  \
    GFFTX_TILE(PREC, TYPE, 2, 256, 256, true, ST, 2),                        \
    GFFTX_TILE(PREC, TYPE, 3, 256, 256, true, ST, 3),                        \
    GFFTX_TILE(PREC, TYPE, 4, 256, 256, true, ST, 4),                        \
    GFFTX_TILE(PREC, TYPE, 5, 256, 256, true, ST, 5),                        \
    GFFTX_TILE(PREC, TYPE, 8, 256, 256, true, ST, 8),                        \
    GFFTX_TILE(PREC, TYPE, 9, 85, 255, true, ST, 3, 3),                      \
    GFFTX_TILE(PREC, TYPE, 15, 85, 255, true, ST, 5, 3),                     \
    GFFTX_TILE(PREC, TYPE, 25, 51, 255, true, ST, 5, 5),                     \
    GFFTX_TILE(PREC, TYPE, 45, 28, 252, true, ST, 5, 3, 3)
#define GFFTX_PASS_SMALL(PREC, TYPE) GFFTX_PASS_SMALL_ONE(PREC, TYPE, false), GFFTX_PASS_SMALL_ONE(PREC, TYPE, true)

// Pass tiles for multi-pass transforms: first pass (T->E) and later passes (T->T).
#define GFFTX_PASS_SET(PREC, TYPE, NB4096, WG4096, NB1024, WG1024, NB1536, WG1536)         \
    GFFTX_TILE(PREC, TYPE, 256, 16, 256, true, false, 16, 16),                             \
    GFFTX_TILE(PREC, TYPE, 256, 16, 256, true, true, 16, 16),                              \
    GFFTX_TILE(PREC, TYPE, 128, 32, 256, true, false, 16, 8),                              \
    GFFTX_TILE(PREC, TYPE, 128, 32, 256, true, true, 16, 8),                               \
    GFFTX_TILE(PREC, TYPE, 1024, NB1024, WG1024, true, false, 16, 16, 4),                  \
    GFFTX_TILE(PREC, TYPE, 1024, NB1024, WG1024, true, true, 16, 16, 4),                   \
    GFFTX_TILE(PREC, TYPE, 1536, NB1536, WG1536, true, false, 16, 16, 2, 3),               \
    GFFTX_TILE(PREC, TYPE, 1536, NB1536, WG1536, true, true, 16, 16, 2, 3),                \
    GFFTX_TILE(PREC, TYPE, 4096, NB4096, WG4096, true, false, 16, 16, 16),                 \
    GFFTX_TILE(PREC, TYPE, 4096, NB4096, WG4096, true, true, 16, 16, 16)

// Tiles carrying a factor 7 (batched single-pass, and T->E / T->T passes), so that lengths
// 2^a 3^b 5^c 7^d split into compiled passes.
#define GFFTX_BATCH_7(PREC, TYPE)                                            \
    GFFTX_TILE(PREC, TYPE, 7, 64, 64, false, false, 7),                     \
    GFFTX_TILE(PREC, TYPE, 14, 64, 128, false, false, 7, 2),                \
    GFFTX_TILE(PREC, TYPE, 28, 64, 256, false, false, 7, 4),                \
    GFFTX_TILE(PREC, TYPE, 49, 36, 252, false, false, 7, 7),                \
    GFFTX_TILE(PREC, TYPE, 56, 32, 256, false, false, 7, 8),                \
    GFFTX_TILE(PREC, TYPE, 112, 36, 252, false, false, 16, 7),              \
    GFFTX_TILE(PREC, TYPE, 224, 18, 252, false, false, 16, 2, 7),           \
    GFFTX_TILE(PREC, TYPE, 448, 9, 252, false, false, 16, 4, 7),            \
    GFFTX_TILE(PREC, TYPE, 896, 4, 224, false, false, 16, 8, 7),            \
    GFFTX_TILE(PREC, TYPE, 1792, 2, 224, false, false, 16, 16, 7),          \
    GFFTX_TILE(PREC, TYPE, 3584, 1, 224, false, false, 16, 16, 2, 7)
#define GFFTX_PASS_7_ONE(PREC, TYPE, ST)                                     \
    GFFTX_TILE(PREC, TYPE, 7, 256, 256, true, ST, 7),                        \
    GFFTX_TILE(PREC, TYPE, 49, 36, 252, true, ST, 7, 7),                     \
    GFFTX_TILE(PREC, TYPE, 112, 36, 252, true, ST, 16, 7),                   \
    GFFTX_TILE(PREC, TYPE, 448, 9, 252, true, ST, 16, 4, 7)
#define GFFTX_PASS_7(PREC, TYPE) GFFTX_PASS_7_ONE(PREC, TYPE, false), GFFTX_PASS_7_ONE(PREC, TYPE, true)

// Host table.  Not const: a const namespace-scope table would be emitted into the device
// image as well; both compilation passes still instantiate every kernel listed here.
TileEntry kTiles[] = {
    GFFTX_BATCH_POW2(PREC_F32, float),
    GFFTX_BATCH_MIXED(PREC_F32, float),
    GFFTX_TILE(PREC_F32, float, 10240, 1, 640, false, false, 16, 16, 8, 5),
    // 4096 points with eight per thread (radix-8 stages, 512 threads): the autotuner weighs
    // it against the radix-16 tile for batched transforms
    GFFTX_TILE(PREC_F32, float, 4096, 1, 512, false, false, 8, 8, 8, 8),
    GFFTX_TILE(PREC_F32, float, 2048, 2, 512, false, false, 8, 8, 8, 4),
    GFFTX_TILE(PREC_F32, float, 5120, 1, 640, false, false, 8, 8, 8, 2, 5),
    GFFTX_TILE(PREC_F32, float, 6144, 1, 768, false, false, 8, 8, 8, 4, 3),
    GFFTX_TILE(PREC_F32, float, 8192, 1, 1024, false, false, 8, 8, 8, 16),
    GFFTX_TILE(PREC_F32, float, 12288, 1, 768, false, false, 16, 16, 16, 3),
    GFFTX_TILE(PREC_F32, float, 16384, 1, 1024, false, false, 16, 16, 16, 4),
    GFFTX_PASS_SET(PREC_F32, float, 4, 1024, 8, 512, 8, 768),
    GFFTX_TILE(PREC_F32, float, 4096, 2, 512, true, false, 16, 16, 16),
    GFFTX_TILE(PREC_F32, float, 4096, 2, 512, true, true, 16, 16, 16),
    GFFTX_TILE(PREC_F32, float, 4096, 1, 256, true, false, 16, 16, 16),
    GFFTX_TILE(PREC_F32, float, 4096, 1, 256, true, true, 16, 16, 16),
    GFFTX_TILE(PREC_F32, float, 256, 32, 512, true, false, 16, 16),
    GFFTX_TILE(PREC_F32, float, 256, 32, 512, true, true, 16, 16),
    GFFTX_TILE(PREC_F32, float, 256, 32, 256, true, false, 16, 16),
    GFFTX_TILE(PREC_F32, float, 256, 32, 256, true, true, 16, 16),
    GFFTX_TILE(PREC_F32, float, 1024, 4, 256, true, false, 16, 16, 4),
    GFFTX_TILE(PREC_F32, float, 1024, 4, 256, true, true, 16, 16, 4),
    GFFTX_TILE(PREC_F32, float, 1536, 4, 384, true, false, 16, 16, 2, 3),
    GFFTX_TILE(PREC_F32, float, 1536, 4, 384, true, true, 16, 16, 2, 3),
    // narrow tiles: more workgroups for transforms of a few million points, whose passes
    // are latency-bound (the autotuner weighs them against the wider rows above)
    GFFTX_TILE(PREC_F32, float, 1536, 2, 192, true, false, 16, 16, 2, 3),
    GFFTX_TILE(PREC_F32, float, 1536, 2, 192, true, true, 16, 16, 2, 3),
    GFFTX_TILE(PREC_F32, float, 1024, 2, 128, true, false, 16, 16, 4),
    GFFTX_TILE(PREC_F32, float, 1024, 2, 128, true, true, 16, 16, 4),
    // eight points per thread (radix-8 stages): twice the waves of the radix-16 tiles
    GFFTX_TILE(PREC_F32, float, 1536, 4, 768, true, false, 8, 8, 8, 3),
    GFFTX_TILE(PREC_F32, float, 1536, 4, 768, true, true, 8, 8, 8, 3),
    GFFTX_TILE(PREC_F32, float, 1024, 8, 1024, true, false, 8, 8, 8, 2),
    GFFTX_TILE(PREC_F32, float, 1024, 8, 1024, true, true, 8, 8, 8, 2),
    GFFTX_TILE(PREC_F32, float, 1024, 4, 512, true, false, 8, 8, 8, 2),
    GFFTX_TILE(PREC_F32, float, 1024, 4, 512, true, true, 8, 8, 8, 2),
    GFFTX_TILE(PREC_F32, float, 16, 64, 64, true, true, 16),
    GFFTX_TILE(PREC_F32, float, 64, 64, 256, true, true, 16, 4),
    GFFTX_TILE(PREC_F32, float, 16, 64, 64, true, false, 16),
    GFFTX_TILE(PREC_F32, float, 64, 64, 256, true, false, 16, 4),
    GFFTX_PASS_MIXED(PREC_F32, float),
    GFFTX_PASS_SMALL(PREC_F32, float),
    GFFTX_BATCH_7(PREC_F32, float),
    GFFTX_PASS_7(PREC_F32, float),
    // pass tiles of 192, 320 and 80 points whose rows are whole 128-byte lines (16, 32 or 48
    // sequences), so that their inputs load non-temporally; the autotuner weighs them against
    // the default shapes above
    GFFTX_TILE(PREC_F32, float, 192, 16, 192, true, false, 16, 4, 3),
    GFFTX_TILE(PREC_F32, float, 192, 16, 192, true, true, 16, 4, 3),
    GFFTX_TILE(PREC_F32, float, 192, 32, 384, true, false, 16, 4, 3),
    GFFTX_TILE(PREC_F32, float, 192, 32, 384, true, true, 16, 4, 3),
    GFFTX_TILE(PREC_F32, float, 320, 16, 320, true, false, 16, 4, 5),
    GFFTX_TILE(PREC_F32, float, 320, 16, 320, true, true, 16, 4, 5),
    GFFTX_TILE(PREC_F32, float, 80, 32, 160, true, false, 16, 5),
    GFFTX_TILE(PREC_F32, float, 80, 32, 160, true, true, 16, 5),
    GFFTX_TILE(PREC_F32, float, 80, 48, 240, true, false, 16, 5),
    GFFTX_TILE(PREC_F32, float, 80, 48, 240, true, true, 16, 5),
    GFFTX_BATCH_POW2(PREC_F64, double),
    GFFTX_BATCH_MIXED(PREC_F64, double),
    GFFTX_PASS_SET(PREC_F64, double, 2, 512, 4, 256, 4, 384),
    // 1024 points with eight per thread in the first two stages (512 threads): the
    // autotuner weighs it against the radix-16 tile for batched FP64 transforms
    GFFTX_TILE(PREC_F64, double, 1024, 4, 512, false, false, 8, 8, 16),
    // one sequence per 64-thread workgroup: 17 KB of LDS each, nine sequences per CU where
    // the four-sequence tile fits eight
    GFFTX_TILE(PREC_F64, double, 1024, 1, 64, false, false, 16, 16, 4),
    GFFTX_TILE(PREC_F64, double, 256, 16, 512, true, false, 8, 8, 4),
    GFFTX_TILE(PREC_F64, double, 256, 16, 512, true, true, 8, 8, 4),
    GFFTX_TILE(PREC_F64, double, 256, 8, 128, true, false, 16, 16),
    GFFTX_TILE(PREC_F64, double, 256, 8, 128, true, true, 16, 16),
    GFFTX_TILE(PREC_F64, double, 4096, 1, 256, true, false, 16, 16, 16),
    GFFTX_TILE(PREC_F64, double, 4096, 1, 256, true, true, 16, 16, 16),
    GFFTX_TILE(PREC_F64, double, 16, 64, 64, true, true, 16),
    GFFTX_TILE(PREC_F64, double, 64, 64, 256, true, true, 16, 4),
    GFFTX_TILE(PREC_F64, double, 16, 64, 64, true, false, 16),
    GFFTX_TILE(PREC_F64, double, 64, 64, 256, true, false, 16, 4),
    GFFTX_PASS_MIXED(PREC_F64, double),
    GFFTX_PASS_SMALL(PREC_F64, double),
    GFFTX_BATCH_7(PREC_F64, double),
    GFFTX_PASS_7(PREC_F64, double),
};
const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

template <class C>
static hipError_t launch_mfma(const MfmaArgs& a, dim3 grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (t0) hipExtLaunchKernelGGL(mfma_fft_kernel<C>, grid, dim3(C::WG), 0, s, t0, t1, 0, a);
    else hipLaunchKernelGGL(mfma_fft_kernel<C>, grid, dim3(C::WG), 0, s, a);
    return hipGetLastError();
}

// Per-stage scaling first (the default); the per-column variants remain for comparison
// (GFFTX_MFMA_SCALE=column).
MfmaEntry kMfma[] = {
    {256, 16, MFMA_SCALE_STAGE, 256, &launch_mfma<MfmaCfg<256, 16>>, "mfma_split_f32_L256_nb16"},
    {4096, 1, MFMA_SCALE_STAGE, 256, &launch_mfma<MfmaCfg<4096, 1>>, "mfma_split_f32_L4096_nb1"},
    {256, 16, MFMA_SCALE_COLUMN, 256, &launch_mfma<MfmaCfg<256, 16, MFMA_SCALE_COLUMN>>, "mfma_split_f32_L256_nb16_col"},
    {4096, 1, MFMA_SCALE_COLUMN, 256, &launch_mfma<MfmaCfg<4096, 1, MFMA_SCALE_COLUMN>>, "mfma_split_f32_L4096_nb1_col"},
};
const int kNumMfma = sizeof(kMfma) / sizeof(kMfma[0]);

// ---------------------------------------------------------------------------------------
// Real-transform split pass.  Forward (post-pass of RDFT) for k in [1, n/2]:
//   h1 = (Z_k + conj Z_{n-k})/2,  h2 = -i (Z_k - conj Z_{n-k})/2,  w = W_{2n}^k
//   X_k = h1 + w h2,  X_{n-k} = conj(h1 - w h2)
// and X_0 = Re Z_0 + Im Z_0, X_n = Re Z_0 - Im Z_0.  Inverse (pre-pass of IRDFT) rebuilds
// Z_k = h1 + i conj(w) (X_k - conj X_{n-k})/2, Z_{n-k} = conj(h1 - i conj(w) (..)/2) with
// h1 = (X_k + conj X_{n-k})/2, and Z_0 = ((X_0 + X_n)/2, (X_0 - X_n)/2).  These are the
// relations GFFT's Separate evaluates with a twiddle recurrence (src/gfftswap.h:246-280).
template <typename T> __global__ __launch_bounds__(256) void split_kernel(SplitArgs a)
{
    using V = V2<T>;
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n = a.n;
    if (k > n / 2) return;
    V* d = reinterpret_cast<V*>(a.data) + (long long)blockIdx.y * a.bdist;
    const V* tw = reinterpret_cast<const V*>(a.tw);
    const T half = (T)0.5;
    if (k == 0) {
        const V z0 = d[0];
        if (!a.inverse) {
            const T x0 = z0.x + z0.y, xn = z0.x - z0.y;
            if (a.bins) {
                d[0] = V{x0, 0};
                d[n] = V{xn, 0};
            } else {
                d[0] = V{x0, xn};
            }
        } else {
            const T x0 = z0.x;
            const T xn = a.bins ? d[n].x : z0.y;
            d[0] = V{half * (x0 + xn), half * (x0 - xn)};
        }
        return;
    }
    const V A = d[k];
    const V B = d[n - k];
    const V w = tw[k];
    const V h1 = V{half * (A.x + B.x), half * (A.y - B.y)};
    const V dd = V{half * (A.x - B.x), half * (A.y + B.y)};  // (A - conj B)/2
    V p;
    if (!a.inverse) {
        const V h2 = V{dd.y, -dd.x};  // -i * dd
        p = cmul(w, h2);
    } else {
        const V cw = V{w.x, -w.y};
        const V t = cmul(cw, dd);
        p = V{-t.y, t.x};  // i * conj(w) * dd
    }
    d[k] = V{h1.x + p.x, h1.y + p.y};
    d[n - k] = V{h1.x - p.x, -(h1.y - p.y)};
}

hipError_t launch_split(int prec, const SplitArgs& a, hipStream_t s)
{
    const long long work = a.n / 2 + 1;
    dim3 grid((unsigned)((work + 255) / 256), (unsigned)a.nbatch);
    if (prec == PREC_F64) hipLaunchKernelGGL(split_kernel<double>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(split_kernel<float>, grid, dim3(256), 0, s, a);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------------
// Digit reversal dst[rev(k)] = src[k] (the permutation GFFTswap2 applies in place).
template <typename T>
__global__ __launch_bounds__(256) void digit_reverse_kernel(const V2<T>* src, V2<T>* dst, long long radix,
                                                            int digits, long long n)
{
    const long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    long long x = k, r = 0;
    for (int i = 0; i < digits; ++i) {
        r = r * radix + x % radix;
        x /= radix;
    }
    dst[r] = src[k];
}

hipError_t launch_digit_reverse(int prec, const void* src, void* dst, long long radix, int digits, long long n,
                                hipStream_t s)
{
    dim3 grid((unsigned)((n + 255) / 256));
    if (prec == PREC_F64)
        hipLaunchKernelGGL(digit_reverse_kernel<double>, grid, dim3(256), 0, s, (const double2*)src,
                           (double2*)dst, radix, digits, n);
    else
        hipLaunchKernelGGL(digit_reverse_kernel<float>, grid, dim3(256), 0, s, (const float2*)src,
                           (float2*)dst, radix, digits, n);
    return hipGetLastError();
}

}  // namespace gfftx
