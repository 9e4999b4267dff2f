// tile_table.h -- launcher and table-entry macros for the instantiated tile kernels
// (kernels.hip, real_tiles.hip).
#pragma once
#include <hip/hip_ext.h>

#include "registry.h"

namespace gfftx {

template <class C>
inline hipError_t launch_tile(const TileArgs& a, dim3 grid, hipStream_t s, hipEvent_t t0, hipEvent_t t1)
{
    if (t0) hipExtLaunchKernelGGL(tile_kernel<C>, grid, dim3(C::WG), 0, s, t0, t1, 0, a);
    else hipLaunchKernelGGL(tile_kernel<C>, grid, dim3(C::WG), 0, s, a);
    return hipGetLastError();
}

#define GFFTX_CFG(TYPE, L, NB, WG, LT, ST, ...) TileCfg<TYPE, L, NB, WG, LT, ST, __VA_ARGS__>
#define GFFTX_TILE(PREC, TYPE, L, NB, WG, LT, ST, ...)                                         \
    {PREC,                                                                                     \
     L,                                                                                        \
     NB,                                                                                       \
     WG,                                                                                       \
     LT,                                                                                       \
     ST,                                                                                       \
     GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)::NS,                                     \
     GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)::R,                                      \
     (int)(GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)::LDS_ELEMS * sizeof(V2<TYPE>)),    \
     &launch_tile<GFFTX_CFG(TYPE, L, NB, WG, LT, ST, __VA_ARGS__)>,                           \
     "tile_" #TYPE "_L" #L "_nb" #NB "_wg" #WG "_" #LT #ST,                                    \
     REAL_NONE}

}  // namespace gfftx
