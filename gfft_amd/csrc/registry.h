// registry.h -- host-visible table of compiled FFT kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "mfma_kernel.h"
#include "tile_kernel.h"

namespace gfftx {

enum Prec { PREC_F64 = 0, PREC_F32 = 1 };

typedef hipError_t (*TileLaunch)(const TileArgs&, dim3, hipStream_t);

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
};

extern TileEntry kTiles[];
extern const int kNumTiles;

// Split-precision MFMA kernels (FP32, lengths 16^S).
typedef hipError_t (*MfmaLaunch)(const MfmaArgs&, dim3, hipStream_t);
struct MfmaEntry {
    int L;
    int nb;
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

// Digit reversal dst[rev(k)] = src[k] over radix^digits elements.
hipError_t launch_digit_reverse(int prec, const void* src, void* dst, long long radix, int digits,
                                long long n, hipStream_t s);

}  // namespace gfftx
