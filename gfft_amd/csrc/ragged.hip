// ragged.hip -- many transforms of different lengths in one or two persistent launches.
//
// The host turns a ragged batch into work lists sorted longest first:
//   * lengths up to 4096: tiles of NB transforms of one length (the transforms stay where
//     the caller put them and are reached through an offset table);
//   * longer lengths L = A*P (A in {64, 128, 256}, chosen so that both passes move rows of
//     128 bytes or more): the four-step method, A-point FFTs of the P sequences
//     x[j + P r] stored transposed into the transform's scratch region, then P-point FFTs
//     of the contiguous scratch rows with the twiddle W_L^{j r}, stored to X[j + A q].
//     Either one workgroup runs both passes of a transform in turn (fused: the scratch
//     region stays in L2), or the first-pass tiles of every transform run in the first
//     launch and the second-pass tiles in a second one (split).
// Workgroups take items from their launch's counter until the list is exhausted, so every
// workgroup ends after at most (items + 1) fetches; the launch boundary orders the two
// passes of a split transform, and no workgroup ever waits for another.
#include "ragged.h"

namespace gfftx {

#define RCFG(L, NB, WG, LT, ST, ...) TileCfg<float, L, NB, WG, LT, ST, __VA_ARGS__>

// Configurations compiled into the ragged kernel, in RAGGED_CFG_* order (ragged.h).
#define RAGGED_CONFIGS(X)                              \
    X(0, RCFG(16, 256, 256, false, false, 16))         \
    X(1, RCFG(32, 128, 256, false, false, 16, 2))      \
    X(2, RCFG(48, 85, 255, false, false, 16, 3))       \
    X(3, RCFG(64, 64, 256, false, false, 16, 4))       \
    X(4, RCFG(80, 51, 255, false, false, 16, 5))       \
    X(5, RCFG(96, 42, 252, false, false, 16, 2, 3))    \
    X(6, RCFG(128, 32, 256, false, false, 16, 8))      \
    X(7, RCFG(160, 25, 250, false, false, 16, 2, 5))   \
    X(8, RCFG(192, 21, 252, false, false, 16, 4, 3))   \
    X(9, RCFG(256, 16, 256, false, false, 16, 16))     \
    X(10, RCFG(320, 12, 240, false, false, 16, 4, 5))  \
    X(11, RCFG(384, 10, 240, false, false, 16, 8, 3))  \
    X(12, RCFG(512, 8, 256, false, false, 16, 16, 2))  \
    X(13, RCFG(640, 6, 240, false, false, 16, 8, 5))   \
    X(14, RCFG(768, 5, 240, false, false, 16, 16, 3))  \
    X(15, RCFG(1024, 4, 256, false, false, 16, 16, 4)) \
    X(16, RCFG(1280, 3, 240, false, false, 16, 16, 5)) \
    X(17, RCFG(1536, 2, 192, false, false, 16, 16, 2, 3)) \
    X(18, RCFG(2048, 2, 256, false, false, 16, 16, 8)) \
    X(19, RCFG(2560, 1, 160, false, false, 16, 16, 2, 5)) \
    X(20, RCFG(3072, 1, 192, false, false, 16, 16, 4, 3)) \
    X(21, RCFG(4096, 1, 256, false, false, 16, 16, 16)) \
    X(22, RCFG(80, 51, 255, false, true, 16, 5))       \
    X(23, RCFG(96, 42, 252, false, true, 16, 2, 3))    \
    X(24, RCFG(128, 32, 256, false, true, 16, 8))      \
    X(25, RCFG(160, 25, 250, false, true, 16, 2, 5))   \
    X(26, RCFG(192, 21, 252, false, true, 16, 4, 3))   \
    X(27, RCFG(256, 16, 256, false, true, 16, 16))     \
    X(28, RCFG(320, 12, 240, false, true, 16, 4, 5))   \
    X(29, RCFG(384, 10, 240, false, true, 16, 8, 3))   \
    X(30, RCFG(512, 8, 256, false, true, 16, 16, 2))   \
    X(31, RCFG(640, 6, 240, false, true, 16, 8, 5))    \
    X(32, RCFG(768, 5, 240, false, true, 16, 16, 3))   \
    X(33, RCFG(1024, 4, 256, false, true, 16, 16, 4))  \
    X(34, RCFG(64, 64, 256, true, true, 16, 4))        \
    X(35, RCFG(128, 32, 256, true, true, 16, 8))       \
    X(36, RCFG(256, 16, 256, true, true, 16, 16))

constexpr int ragged_lds_elems()
{
    int m = 1;
#define X(ID, CFG) m = (CFG::LDS_ELEMS > m ? CFG::LDS_ELEMS : m);
    RAGGED_CONFIGS(X)
#undef X
    return m;
}
constexpr int kRaggedLds = ragged_lds_elems();

// One out-of-line function per configuration: each is unrolled and register-allocated on
// its own instead of as part of one very large switch.
template <class CFG> __device__ __noinline__ void ragged_tile(const TileArgs& a, unsigned tile, float2* lds)
{
    tile_body<CFG, CFG::WG < RAGGED_WG>(a, tile, 0, lds);
}

__device__ __forceinline__ void run_cfg(int cfg, const TileArgs& a, unsigned tile, float2* lds)
{
    switch (cfg) {
#define X(ID, CFG) \
    case ID: ragged_tile<CFG>(a, tile, lds); break;
        RAGGED_CONFIGS(X)
#undef X
        default: break;
    }
}

__global__ __launch_bounds__(RAGGED_WG) void ragged_reset_kernel(unsigned long long* counter)
{
    counter[0] = 0;
    counter[1] = 0;
}

__global__ __launch_bounds__(RAGGED_WG) void ragged_kernel(RaggedArgs ra)
{
    __shared__ float2 lds[kRaggedLds];
    __shared__ int cur;
    const RaggedItem* list = ra.phase ? ra.items + ra.nitems0 : ra.items;
    const int nitems = ra.phase ? ra.nitems1 : ra.nitems0;
    unsigned long long* counter = ra.counter + ra.phase;
    for (;;) {
        if (threadIdx.x == 0) cur = (int)atomicAdd(counter, 1ull);
        __syncthreads();
        const int it = cur;
        __syncthreads();
        if (it >= nitems) break;
        const RaggedItem item = list[it];
        const RaggedClass cls = ra.classes[item.cls];
        TileArgs a;
        a.its = 0;
        a.bdist = 0;
        a.xcd = 0;
        a.conj_in = 0;
        a.conj_out = 0;
        a.ptw_lo = nullptr;
        a.ptw_hi = nullptr;
        a.ptw_shift = 0;
        a.seq_off = nullptr;
        a.boff = nullptr;
        a.scale = 1.0;
        a.nt_out = 0;
        if (item.kind == RAGGED_SINGLE) {
            // tile of same-length transforms, in place through the offset table
            a.in = ra.data;
            a.out = ra.data;
            a.ntrans = cls.count;
            a.ies = 1;
            a.ns_pass = 1;
            a.tw = cls.tw0;
            a.seq_off = cls.seq_off;
            a.conj_in = ra.inverse;
            a.conj_out = ra.inverse;
            a.scale = ra.inverse ? 1.0 / (double)cls.L : 1.0;
            a.nt_out = ra.nt_out;
        } else {
            // first pass of a two-pass transform: A-point FFTs of x[j + P r] -> z[q P + j]
            const long long A = cls.A, P = cls.L / A;
            a.in = ra.data + item.off;
            a.out = ra.scratch + item.soff;
            a.ntrans = P;
            a.its = 1;
            a.ies = P;
            a.ns_pass = P;
            a.tw = cls.tw0;
            a.conj_in = ra.inverse;
            if (item.kind == RAGGED_FUSED) {
                for (int t = 0; t < item.nt0; ++t) {
                    run_cfg(item.cfg, a, (unsigned)t, lds);
                    __syncthreads();
                }
            }
            if (item.kind != RAGGED_PASS0) {
                // second pass: P-point FFTs of the rows z[j P + r] times W_L^{j r} -> X[j + A q]
                a.in = ra.scratch + item.soff;
                a.out = ra.data + item.off;
                a.ntrans = A;
                a.its = P;
                a.ies = 1;
                a.ns_pass = A;
                a.tw = cls.tw1;
                a.ptw_lo = cls.ptw_lo;
                a.ptw_hi = cls.ptw_hi;
                a.ptw_shift = cls.ptw_shift;
                a.conj_in = 0;
                a.conj_out = ra.inverse;
                a.scale = ra.inverse ? 1.0 / (double)cls.L : 1.0;
                a.nt_out = ra.nt_out;
            }
            if (item.kind == RAGGED_FUSED) {
                for (int t = 0; t < item.nt1; ++t) {
                    run_cfg(item.cfg1, a, (unsigned)t, lds);
                    __syncthreads();
                }
                continue;
            }
        }
        run_cfg(item.cfg, a, (unsigned)item.tile, lds);
        __syncthreads();
    }
}

hipError_t launch_ragged(const RaggedArgs& ra, int grid0, int grid1, hipStream_t s)
{
    hipLaunchKernelGGL(ragged_reset_kernel, dim3(1), dim3(1), 0, s, ra.counter);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    RaggedArgs a = ra;
    a.phase = 0;
    hipLaunchKernelGGL(ragged_kernel, dim3(grid0), dim3(RAGGED_WG), 0, s, a);
    e = hipGetLastError();
    if (e != hipSuccess || ra.nitems1 == 0) return e;
    a.phase = 1;
    hipLaunchKernelGGL(ragged_kernel, dim3(grid1), dim3(RAGGED_WG), 0, s, a);
    return hipGetLastError();
}

// Host view of the compiled configurations (for twiddle tables and planning).
RaggedCfgInfo ragged_cfg_info(int id)
{
    RaggedCfgInfo r{0, 0, 0, false, false, 0, nullptr};
    switch (id) {
#define X(ID, CFG) \
    case ID: r = RaggedCfgInfo{CFG::L, CFG::NB, CFG::WG, CFG::LOADT, CFG::STORET, CFG::NS, CFG::R}; break;
        RAGGED_CONFIGS(X)
#undef X
        default: break;
    }
    return r;
}

int ragged_num_cfgs() { return RAGGED_NUM_CFGS; }

}  // namespace gfftx
