// real_tiles.hip -- tile kernels with the real-transform split fused in (RealCfg) and their
// table: the R2C post-pass in the last pass of a real forward transform (T->T) or after a
// single-pass one (E->E), and the C2R pre-pass in the first pass of a real inverse transform
// (T->E) or before a single-pass one (E->E).  A plan without a matching entry runs the split
// kernel as a pass of its own.
#include "tile_table.h"

namespace gfftx {

#define GFFTX_REAL_TILE(PREC, TYPE, KIND, SUFFIX, LT, ST, L, NB, WG, ...)                                 \
    {PREC,                                                                                                 \
     L,                                                                                                    \
     NB,                                                                                                   \
     WG,                                                                                                   \
     LT,                                                                                                   \
     ST,                                                                                                   \
     GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)::NS,                                                \
     GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)::R,                                                 \
     (int)(RealCfg<GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__), KIND>::LDS_ELEMS * sizeof(V2<TYPE>)), \
     &launch_tile<RealCfg<GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__), KIND>>,                       \
     "tile_" #TYPE "_L" #L "_nb" #NB "_wg" #WG "_" SUFFIX,                                               \
     KIND}

// last pass of a multi-pass forward transform (T->T, NB even)
#define GFFTX_POST_T(PREC, TYPE, L, NB, WG, ...) GFFTX_REAL_TILE(PREC, TYPE, REAL_POST, "split", true, true, L, NB, WG, __VA_ARGS__)
// first pass of a multi-pass inverse transform (T->E, NB even)
#define GFFTX_PRE_T(PREC, TYPE, L, NB, WG, ...) GFFTX_REAL_TILE(PREC, TYPE, REAL_PRE, "presplit", true, false, L, NB, WG, __VA_ARGS__)
// single-pass transforms (E->E), forward and inverse
#define GFFTX_REAL_E(PREC, TYPE, L, NB, WG, ...)                                                            \
    GFFTX_REAL_TILE(PREC, TYPE, REAL_POST, "e_split", false, false, L, NB, WG, __VA_ARGS__),               \
    GFFTX_REAL_TILE(PREC, TYPE, REAL_PRE, "e_presplit", false, false, L, NB, WG, __VA_ARGS__)

// The pass-tile shapes of kernels.hip (GFFTX_PASS_SET and the FP32 alternatives) that can
// end a forward or start an inverse real transform.
#define GFFTX_REAL_PASS_SET(PREC, TYPE, KIND_MACRO, NB4096, WG4096, NB1024, WG1024, NB1536, WG1536) \
    KIND_MACRO(PREC, TYPE, 16, 64, 64, 16),                                                            \
    KIND_MACRO(PREC, TYPE, 64, 64, 256, 16, 4),                                                        \
    KIND_MACRO(PREC, TYPE, 128, 32, 256, 16, 8),                                                       \
    KIND_MACRO(PREC, TYPE, 256, 16, 256, 16, 16),                                                      \
    KIND_MACRO(PREC, TYPE, 320, 12, 240, 16, 4, 5),                                                    \
    KIND_MACRO(PREC, TYPE, 1024, NB1024, WG1024, 16, 16, 4),                                           \
    KIND_MACRO(PREC, TYPE, 1536, NB1536, WG1536, 16, 16, 2, 3),                                        \
    KIND_MACRO(PREC, TYPE, 4096, NB4096, WG4096, 16, 16, 16)
#define GFFTX_REAL_PASS_F32(KIND_MACRO)                                                                 \
    KIND_MACRO(PREC_F32, float, 1536, 2, 192, 16, 16, 2, 3),                                           \
    KIND_MACRO(PREC_F32, float, 1536, 4, 384, 16, 16, 2, 3),                                           \
    KIND_MACRO(PREC_F32, float, 1024, 4, 256, 16, 16, 4),                                              \
    KIND_MACRO(PREC_F32, float, 1024, 2, 128, 16, 16, 4),                                              \
    KIND_MACRO(PREC_F32, float, 1536, 4, 768, 8, 8, 8, 3),                                             \
    KIND_MACRO(PREC_F32, float, 1024, 8, 1024, 8, 8, 8, 2),                                            \
    KIND_MACRO(PREC_F32, float, 1024, 4, 512, 8, 8, 8, 2)

// The batched single-pass shapes of kernels.hip with an even length.
#define GFFTX_REAL_E_SET(PREC, TYPE)                                        \
    GFFTX_REAL_E(PREC, TYPE, 2, 64, 64, 2),                                 \
    GFFTX_REAL_E(PREC, TYPE, 4, 64, 64, 4),                                 \
    GFFTX_REAL_E(PREC, TYPE, 8, 64, 64, 8),                                 \
    GFFTX_REAL_E(PREC, TYPE, 16, 64, 64, 16),                               \
    GFFTX_REAL_E(PREC, TYPE, 32, 64, 128, 16, 2),                           \
    GFFTX_REAL_E(PREC, TYPE, 64, 64, 256, 16, 4),                           \
    GFFTX_REAL_E(PREC, TYPE, 128, 32, 256, 16, 8),                          \
    GFFTX_REAL_E(PREC, TYPE, 256, 16, 256, 16, 16),                         \
    GFFTX_REAL_E(PREC, TYPE, 512, 8, 256, 16, 16, 2),                       \
    GFFTX_REAL_E(PREC, TYPE, 1024, 4, 256, 16, 16, 4),                      \
    GFFTX_REAL_E(PREC, TYPE, 2048, 2, 256, 16, 16, 8),                      \
    GFFTX_REAL_E(PREC, TYPE, 4096, 1, 256, 16, 16, 16),                     \
    GFFTX_REAL_E(PREC, TYPE, 8192, 1, 512, 16, 16, 16, 2),                  \
    GFFTX_REAL_E(PREC, TYPE, 48, 85, 255, 16, 3),                           \
    GFFTX_REAL_E(PREC, TYPE, 80, 51, 255, 16, 5),                           \
    GFFTX_REAL_E(PREC, TYPE, 96, 42, 252, 16, 2, 3),                        \
    GFFTX_REAL_E(PREC, TYPE, 160, 25, 250, 16, 2, 5),                       \
    GFFTX_REAL_E(PREC, TYPE, 192, 21, 252, 16, 4, 3),                       \
    GFFTX_REAL_E(PREC, TYPE, 320, 12, 240, 16, 4, 5),                       \
    GFFTX_REAL_E(PREC, TYPE, 384, 10, 240, 16, 8, 3),                       \
    GFFTX_REAL_E(PREC, TYPE, 640, 6, 240, 16, 8, 5),                        \
    GFFTX_REAL_E(PREC, TYPE, 768, 5, 240, 16, 16, 3),                       \
    GFFTX_REAL_E(PREC, TYPE, 1280, 3, 240, 16, 16, 5),                      \
    GFFTX_REAL_E(PREC, TYPE, 1536, 2, 192, 16, 16, 2, 3),                   \
    GFFTX_REAL_E(PREC, TYPE, 2560, 2, 320, 16, 16, 2, 5),                   \
    GFFTX_REAL_E(PREC, TYPE, 3072, 2, 384, 16, 16, 4, 3),                   \
    GFFTX_REAL_E(PREC, TYPE, 5120, 1, 320, 16, 16, 4, 5),                   \
    GFFTX_REAL_E(PREC, TYPE, 6144, 1, 384, 16, 16, 8, 3)

// Host table (not const, as the complex table in kernels.hip).
TileEntry kRealTiles[] = {
    GFFTX_REAL_PASS_SET(PREC_F32, float, GFFTX_POST_T, 4, 1024, 8, 512, 8, 768),
    GFFTX_REAL_PASS_F32(GFFTX_POST_T),
    GFFTX_REAL_PASS_SET(PREC_F32, float, GFFTX_PRE_T, 4, 1024, 8, 512, 8, 768),
    GFFTX_REAL_PASS_F32(GFFTX_PRE_T),
    GFFTX_REAL_E_SET(PREC_F32, float),
    GFFTX_REAL_E(PREC_F32, float, 10240, 1, 640, 16, 16, 8, 5),
    GFFTX_REAL_E(PREC_F32, float, 4096, 1, 512, 8, 8, 8, 8),
    GFFTX_REAL_E(PREC_F32, float, 2048, 2, 512, 8, 8, 8, 4),
    GFFTX_REAL_E(PREC_F32, float, 8192, 1, 1024, 8, 8, 8, 16),
    GFFTX_REAL_E(PREC_F32, float, 5120, 1, 640, 8, 8, 8, 2, 5),
    GFFTX_REAL_E(PREC_F32, float, 6144, 1, 768, 8, 8, 8, 4, 3),
    GFFTX_REAL_E(PREC_F32, float, 12288, 1, 768, 16, 16, 16, 3),
    GFFTX_REAL_E(PREC_F32, float, 16384, 1, 1024, 16, 16, 16, 4),
    GFFTX_REAL_PASS_SET(PREC_F64, double, GFFTX_POST_T, 2, 512, 4, 256, 4, 384),
    GFFTX_REAL_PASS_SET(PREC_F64, double, GFFTX_PRE_T, 2, 512, 4, 256, 4, 384),
    GFFTX_REAL_E_SET(PREC_F64, double),
    GFFTX_REAL_E(PREC_F64, double, 1024, 4, 512, 8, 8, 16),
};
const int kNumRealTiles = sizeof(kRealTiles) / sizeof(kRealTiles[0]);

}  // namespace gfftx
