// registry.h -- host-visible table of compiled FFT kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "mfma_kernel.h"
#include "tile_kernel.h"

namespace gfftx {

enum Prec { PREC_F64 = 0, PREC_F32 = 1 };

// Launchers take an optional event pair (both null or both created): when given, the
// launch goes through hipExtLaunchKernel, which stamps the events with the kernel's own
// start and end instead of recording them as separate stream commands.
typedef hipError_t (*TileLaunch)(const TileArgs&, dim3, hipStream_t, hipEvent_t, hipEvent_t);

struct TileEntry {
    int prec;
    int L;
    int nb;
    int wg;
    bool loadT;
    bool storeT;
    int nrad;
    const int* radices;
    int lds_bytes;
    TileLaunch launch;
    const char* name;
    int real;  // REAL_NONE, or the real-transform split fused in (REAL_POST, REAL_PRE: RealCfg)
};

// Complex tiles (kernels.hip) and tiles with a real-transform split fused in (real_tiles.hip).
extern TileEntry kTiles[];
extern const int kNumTiles;
extern TileEntry kRealTiles[];
extern const int kNumRealTiles;

// Split-precision MFMA kernels (FP32, lengths 16^S).
typedef hipError_t (*MfmaLaunch)(const MfmaArgs&, dim3, hipStream_t, hipEvent_t, hipEvent_t);
struct MfmaEntry {
    int L;
    int nb;
    int scale;  // MFMA_SCALE_STAGE (stage twiddles carry 2^-4) or MFMA_SCALE_COLUMN
    int wg;     // threads per workgroup
    MfmaLaunch launch;
    const char* name;
};
extern MfmaEntry kMfma[];
extern const int kNumMfma;

// Real-transform split pass (Separate, src/gfftswap.h:239-336) as one kernel.
struct SplitArgs {
    void* data;            // n complex (+1 when bins)
    long long n;           // complex half length
    long long bdist;       // distance between batch items, complex elements
    long long nbatch;
    const void* tw;        // n/2+1 entries W_{2n}^k in working precision (forward sign)
    int inverse;           // 0: post-pass of RDFT, 1: pre-pass of IRDFT
    int bins;              // 1: n+1 bins layout, 0: NR packing
};
hipError_t launch_split(int prec, const SplitArgs& a, hipStream_t s);

// Chirp-z (Bluestein) path for lengths with a prime factor above 5 (bluestein.hip).
struct ChirpTableArgs {
    void* chirp;            // n entries c_j = W_{2n}^{j^2}, working precision
    long long n;
    const double2* lo;      // two-level FP64 root table of order 2n
    const double2* hi;
    int shift;
};
struct BluesteinArgs {
    const void* in;         // caller's input (pre), batch items `dist` apart
    void* out;              // caller's output (post)
    void* work;             // batch x m convolution buffer
    const void* chirp;      // n entries
    const void* filter;     // m entries: FFT_m of the chirp filter
    long long n, m, dist, batch;
    int conj;               // inverse transform: conjugate on input and on output
    double scale;           // post: 1 or 1/n
    const long long* boff;  // optional: element offset of every batch item (else b * dist)
};
enum { BLUESTEIN_PRE = 0, BLUESTEIN_MUL = 1, BLUESTEIN_POST = 2, BLUESTEIN_FILTER = 3 };
hipError_t launch_bluestein(int prec, int which, const BluesteinArgs& a, hipStream_t s);
hipError_t launch_chirp_table(int prec, const ChirpTableArgs& a, hipStream_t s);

// Digit reversal dst[rev(k)] = src[k] over radix^digits elements.
hipError_t launch_digit_reverse(int prec, const void* src, void* dst, long long radix, int digits,
                                long long n, hipStream_t s);

}  // namespace gfftx
