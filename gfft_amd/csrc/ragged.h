// ragged.h -- work-list structures of the persistent ragged-batch kernel (FP32).
#pragma once
#include <hip/hip_runtime.h>

#include "tile_kernel.h"

namespace gfftx {

constexpr int RAGGED_WG = 256;
constexpr int RAGGED_NUM_CFGS = 35;
// configuration ids (ragged.hip, RAGGED_CONFIGS): single-pass E-mode tiles for lengths up
// to 4096, E->T pass-1 tiles for the two-pass lengths, the 64-point T->T pass-0 tile
constexpr int RAGGED_CFG_T64 = 34;

struct RaggedItem {
    int cfg;     // single-pass tile config, or pass-0 config of a two-pass transform
    int cfg1;    // pass-1 config of a two-pass transform, -1 for a single-pass tile
    int cls;     // length class
    int tile;    // single-pass: tile index within the class
    int ntiles0; // two-pass: tiles of pass 0 / pass 1
    int ntiles1;
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
    int pad_;
};

struct RaggedArgs {
    float2* data;
    float2* scratch;
    const RaggedItem* items;
    const RaggedClass* classes;
    unsigned long long* counter;
    int nitems;
    int inverse;
};

struct RaggedCfgInfo {
    int L, nb, wg;
    bool loadT, storeT;
    int nrad;
    const int* radices;
};

hipError_t launch_ragged(const RaggedArgs& ra, int grid, hipStream_t s);
RaggedCfgInfo ragged_cfg_info(int id);
int ragged_num_cfgs();

}  // namespace gfftx
