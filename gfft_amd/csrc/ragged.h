// ragged.h -- work-list structures of the persistent ragged-batch kernel (FP32).
#pragma once
#include <hip/hip_runtime.h>

#include "tile_kernel.h"

namespace gfftx {

constexpr int RAGGED_WG = 256;
constexpr int RAGGED_NUM_CFGS = 37;
// configuration ids (ragged.hip, RAGGED_CONFIGS): single-pass E-mode tiles for lengths up
// to 4096, E->T second-pass tiles for the two-pass lengths, T->T first-pass tiles of 64,
// 128 and 256 points
constexpr int RAGGED_CFG_T64 = 34;
constexpr int RAGGED_CFG_T128 = 35;
constexpr int RAGGED_CFG_T256 = 36;

enum RaggedKind { RAGGED_SINGLE = 0, RAGGED_PASS0 = 1, RAGGED_PASS1 = 2, RAGGED_FUSED = 3 };

// One work item.  A two-pass transform of length L = A*P is computed by the four-step
// method: A-point FFTs of the P sequences x[j + P r], stored transposed to scratch, then
// P-point FFTs of the scratch rows with the twiddle W_L^{j r}, stored to X[j + A q].
//   RAGGED_SINGLE: one tile of NB transforms of one length class, single pass, in place;
//   RAGGED_PASS0:  one tile of the first pass of a two-pass transform (phase 0);
//   RAGGED_PASS1:  one tile of the second pass (phase 1);
//   RAGGED_FUSED:  a whole two-pass transform by one workgroup (both passes in turn, the
//                  scratch region stays in L2).
struct RaggedItem {
    int cfg;         // tile configuration (RAGGED_CONFIGS id); first-pass config when fused
    int kind;        // RaggedKind
    int cls;         // length class
    int tile;        // tile index within the class (single) or within the pass (two-pass)
    int cfg1;        // fused: second-pass config
    int nt0, nt1;    // fused: tiles of the first and second pass
    int pad_;
    long long off;   // two-pass: element offset of the transform in the data buffer
    long long soff;  // two-pass: element offset of its scratch region
};

struct RaggedClass {
    long long L;
    long long count;            // transforms of this length
    const long long* seq_off;   // their element offsets (single-pass classes)
    const void* tw0;            // stage twiddles of cfg / cfg1
    const void* tw1;
    const double2* ptw_lo;      // pass twiddles W_L (two-pass classes)
    const double2* ptw_hi;
    int ptw_shift;
    int A;                      // two-pass classes: first-pass FFT length (L = A * P)
};

// Phase 0 runs the single-pass tiles and the pass-0 tiles of every two-pass transform;
// phase 1, a separate launch, the pass-1 tiles.  The launch boundary orders the two
// passes of a transform, so no workgroup waits for another.
struct RaggedArgs {
    float2* data;
    float2* scratch;
    const RaggedItem* items;      // phase-0 items, then phase-1 items
    const RaggedClass* classes;
    unsigned long long* counter;  // two queue counters, one per phase
    int nitems0, nitems1;
    int inverse;
    int phase;                    // set per launch
    int nt_out;                   // 1: final outputs stored non-temporally (a batch larger
                                  // than the Infinity Cache)
};

struct RaggedCfgInfo {
    int L, nb, wg;
    bool loadT, storeT;
    int nrad;
    const int* radices;
};

hipError_t launch_ragged(const RaggedArgs& ra, int grid0, int grid1, hipStream_t s);
RaggedCfgInfo ragged_cfg_info(int id);
int ragged_num_cfgs();

}  // namespace gfftx
