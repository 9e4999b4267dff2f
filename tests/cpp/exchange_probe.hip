// exchange_probe.hip -- the exchange between two radix-16 stages, through LDS or across lanes.
//
// A batched 256-point FP32 FFT with the shape of the library's 256-point tile: 16 threads per
// sequence, 16 points per thread, a radix-16 stage, the W_256 twiddle, a second radix-16
// stage, non-temporal loads.  Between the stages every thread needs one element from each
// of the other 15 threads of its sequence: a 16 x 16 transpose between lanes and registers.
// Four kernels differ only in that exchange:
//   0  LDS: padded write, wave barrier, read (the library's wave-local tiles);
//   1  DPP: a butterfly transpose in four steps, lane distances 8, 4, 2, 1 (row_ror:8,
//      row_half_mirror then quad_perm, quad_perm), each a DPP move and three selects per
//      32-bit word;
//   2  ds_swizzle: the same steps with xor swizzles;
//   3  permlane: the sequence's 16 threads sit at lane bits 0, 1, 4 and 5, so that gfx950's
//      v_permlane32_swap / v_permlane16_swap do two of the steps with no selects (the swap of
//      two registers between lane halves is the butterfly step itself); DPP does the others.
// The arithmetic is the same in all four, so their outputs must be bit-identical.  The
// program checks that and the LDS kernel against a double-precision DFT, then times the four
// on 65536 sequences from HBM (four rotating 128 MiB buffers) and on 4096 sequences that stay
// in cache.  It exits non-zero on a mismatch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#define CK(x)                                                                         \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            printf("HIP error %s at line %d\n", hipGetErrorString(e_), __LINE__);     \
            exit(2);                                                                  \
        }                                                                             \
    } while (0)

__device__ __forceinline__ float2 add(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }  // -i a

__device__ __forceinline__ void dft4(float2& a, float2& b, float2& c, float2& d)
{
    const float2 t0 = add(a, c), t1 = sub(a, c), t2 = add(b, d), t3 = mul_mi(sub(b, d));
    a = add(t0, t2);
    c = sub(t0, t2);
    b = add(t1, t3);
    d = sub(t1, t3);
}

// forward DFT of 16 points in registers (4 x 4), natural order in and out
__device__ __forceinline__ void dft16(float2 (&x)[16])
{
    constexpr float C1 = 0.92387953251128674f, S1 = 0.38268343236508978f, R2 = 0.70710678118654752f;
#pragma unroll
    for (int n1 = 0; n1 < 4; ++n1) dft4(x[n1], x[n1 + 4], x[n1 + 8], x[n1 + 12]);
    // Y[n1][k1] sits at x[n1 + 4 k1]; times W_16^{n1 k1}
    x[5] = cmul(x[5], make_float2(C1, -S1));    // W^1
    x[9] = cmul(x[9], make_float2(R2, -R2));    // W^2
    x[13] = cmul(x[13], make_float2(S1, -C1));  // W^3
    x[6] = cmul(x[6], make_float2(R2, -R2));    // W^2
    x[10] = cmul(x[10], make_float2(0.f, -1.f)); // W^4
    x[14] = cmul(x[14], make_float2(-R2, -R2)); // W^6
    x[7] = cmul(x[7], make_float2(S1, -C1));    // W^3
    x[11] = cmul(x[11], make_float2(-R2, -R2)); // W^6
    x[15] = cmul(x[15], make_float2(-C1, S1));  // W^9
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) dft4(x[4 * k1], x[4 * k1 + 1], x[4 * k1 + 2], x[4 * k1 + 3]);
    // X[k1 + 4 k2] sits at x[4 k1 + k2]
    float2 y[16];
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1)
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) y[k1 + 4 * k2] = x[4 * k1 + k2];
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = y[i];
}

// value of lane (lane ^ D) within a row of 16 lanes
template <int D> __device__ __forceinline__ float xor_dpp(float v)
{
    int u = __float_as_int(v);
    if constexpr (D == 1) u = __builtin_amdgcn_mov_dpp(u, 0xB1, 0xF, 0xF, false);  // quad_perm 1,0,3,2
    else if constexpr (D == 2) u = __builtin_amdgcn_mov_dpp(u, 0x4E, 0xF, 0xF, false);  // quad_perm 2,3,0,1
    else if constexpr (D == 4)  // row_half_mirror (lane ^ 7), then quad_perm 3,2,1,0 (lane ^ 3)
        u = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(u, 0x141, 0xF, 0xF, false), 0x1B, 0xF, 0xF, false);
    else u = __builtin_amdgcn_mov_dpp(u, 0x128, 0xF, 0xF, false);  // row_ror:8
    return __int_as_float(u);
}

// value of lane (lane ^ D) within 32 lanes (ds_swizzle bit mode: and 0x1F, or 0, xor D)
template <int D> __device__ __forceinline__ float xor_swz(float v)
{
    return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F | (D << 10)));
}

// One butterfly-transpose step: lane bit D is exchanged with register bit RB.  Lane l takes,
// into each register i whose bit RB differs from l's bit D, register i ^ RB of lane l ^ D.
// MODE 1 moves words by DPP, MODE 2 by ds_swizzle.  The selects are written on the float
// components: a select of float2 lvalues would select addresses and put the array in memory.
template <int D, int RB, int MODE> __device__ __forceinline__ void transpose_step(float2 (&b)[16], int lane)
{
    const bool hi = (lane & D) != 0;
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int i0 = ((p & ~(RB - 1)) << 1) | (p & (RB - 1));  // the indices with bit RB clear
        const int i1 = i0 | RB;
        const float a0x = b[i0].x, a0y = b[i0].y, a1x = b[i1].x, a1y = b[i1].y;
        const float sx = hi ? a0x : a1x, sy = hi ? a0y : a1y;
        float rx, ry;
        if constexpr (MODE == 2) {
            rx = xor_swz<D>(sx);
            ry = xor_swz<D>(sy);
        } else {
            rx = xor_dpp<D>(sx);
            ry = xor_dpp<D>(sy);
        }
        b[i0] = make_float2(hi ? rx : a0x, hi ? ry : a0y);
        b[i1] = make_float2(hi ? a1x : rx, hi ? a1y : ry);
    }
}

// Steps for lane bits 5 (D = 32) and 4 (D = 16).  v_permlane32_swap(u, v) exchanges the upper
// 32 lanes of u with the lower 32 lanes of v, v_permlane16_swap the odd 16-lane rows of u with
// the even rows of v: the lower lanes keep u and take the upper lanes' u into v, the upper
// lanes keep v and take the lower lanes' v into u -- the butterfly step, with no selects.
template <int D, int RB> __device__ __forceinline__ void transpose_swap(float2 (&b)[16])
{
#pragma unroll
    for (int p = 0; p < 8; ++p) {
        const int i0 = ((p & ~(RB - 1)) << 1) | (p & (RB - 1));
        const int i1 = i0 | RB;
        int ux = __float_as_int(b[i0].x), vx = __float_as_int(b[i1].x);
        int uy = __float_as_int(b[i0].y), vy = __float_as_int(b[i1].y);
        if constexpr (D == 16) {
            const auto rx = __builtin_amdgcn_permlane16_swap(ux, vx, false, false);
            const auto ry = __builtin_amdgcn_permlane16_swap(uy, vy, false, false);
            ux = rx[0];
            vx = rx[1];
            uy = ry[0];
            vy = ry[1];
        } else {
            const auto rx = __builtin_amdgcn_permlane32_swap(ux, vx, false, false);
            const auto ry = __builtin_amdgcn_permlane32_swap(uy, vy, false, false);
            ux = rx[0];
            vx = rx[1];
            uy = ry[0];
            vy = ry[1];
        }
        b[i0] = make_float2(__int_as_float(ux), __int_as_float(uy));
        b[i1] = make_float2(__int_as_float(vx), __int_as_float(vy));
    }
}

// 16 sequences per 256-thread workgroup; sequence g = 16 blockIdx.x + s.  Thread t of a
// sequence loads x[t + 16 r] (r < 16), and after the exchange holds outputs t + 16 k2.
template <int MODE>
__global__ __launch_bounds__(256) void fft256(const float2* __restrict__ in, float2* __restrict__ out,
                                              const float2* __restrict__ tw)
{
    __shared__ float2 lds[16 * 16 * 17];
    const int lane = threadIdx.x & 63;
    // MODE 3: sequence sq of a wave on lanes 16 r + 4 sq + q, thread t = 4 r + q; otherwise
    // 16 consecutive lanes per sequence
    const int t = MODE == 3 ? 4 * (lane >> 4) + (lane & 3) : (threadIdx.x & 15);
    const int s = MODE == 3 ? 4 * (threadIdx.x >> 6) + ((lane >> 2) & 3) : (threadIdx.x >> 4);
    const long long g = (long long)blockIdx.x * 16 + s;
    const float2* src = in + g * 256;
    typedef float f2v __attribute__((ext_vector_type(2)));
    float2 x[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const f2v v = __builtin_nontemporal_load(reinterpret_cast<const f2v*>(&src[t + 16 * r]));
        x[r] = make_float2(v.x, v.y);
    }
    dft16(x);
#pragma unroll
    for (int k1 = 1; k1 < 16; ++k1) x[k1] = cmul(x[k1], tw[(t * k1) & 255]);
    if constexpr (MODE == 0) {
        float2* row = lds + s * 272;
#pragma unroll
        for (int k1 = 0; k1 < 16; ++k1) row[k1 * 17 + t] = x[k1];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int i = 0; i < 16; ++i) x[i] = row[t * 17 + i];
    } else if constexpr (MODE == 3) {
        transpose_swap<32, 8>(x);
        transpose_swap<16, 4>(x);
        transpose_step<2, 2, 1>(x, lane);
        transpose_step<1, 1, 1>(x, lane);
    } else {
        transpose_step<8, 8, MODE>(x, lane);
        transpose_step<4, 4, MODE>(x, lane);
        transpose_step<2, 2, MODE>(x, lane);
        transpose_step<1, 1, MODE>(x, lane);
    }
    dft16(x);
    float2* dst = out + g * 256;
#pragma unroll
    for (int k2 = 0; k2 < 16; ++k2) dst[t + 16 * k2] = x[k2];
}

static void fill(std::vector<float2>& h, unsigned long long st)
{
    for (size_t i = 0; i < h.size(); ++i) {
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        const float a = (float)((st >> 11) & 0xFFFFF) / 1048576.0f - 0.5f;
        st = st * 6364136223846793005ULL + 1442695040888963407ULL;
        const float b = (float)((st >> 11) & 0xFFFFF) / 1048576.0f - 0.5f;
        h[i] = make_float2(a, b);
    }
}

int main()
{
    const long long BMAX = 65536, N = 256, total = BMAX * N;
    const int NBUF = 4;
    const char* name[4] = {"LDS", "DPP", "ds_swizzle", "permlane+DPP"};
    std::vector<float2> h(total), tw(256);
    for (int m = 0; m < 256; ++m) tw[m] = make_float2((float)cos(-2 * M_PI * m / 256.0), (float)sin(-2 * M_PI * m / 256.0));
    float2 *din[NBUF], *dout[NBUF], *dtw;
    for (int k = 0; k < NBUF; ++k) {
        CK(hipMalloc(&din[k], total * sizeof(float2)));
        CK(hipMalloc(&dout[k], total * sizeof(float2)));
        fill(h, 12345 + k);
        CK(hipMemcpy(din[k], h.data(), total * sizeof(float2), hipMemcpyHostToDevice));
    }
    CK(hipMalloc(&dtw, 256 * sizeof(float2)));
    CK(hipMemcpy(dtw, tw.data(), 256 * sizeof(float2), hipMemcpyHostToDevice));
    // grid: B / 16 workgroups; B is 65536 or 4096, a multiple of 16, and every workgroup's
    // 16 sequences lie inside the buffers of BMAX sequences
    auto launch = [&](int mode, long long B, int k) {
        const dim3 grid((unsigned)(B / 16)), blk(256);
        if (mode == 0) hipLaunchKernelGGL(fft256<0>, grid, blk, 0, 0, din[k], dout[k], dtw);
        else if (mode == 1) hipLaunchKernelGGL(fft256<1>, grid, blk, 0, 0, din[k], dout[k], dtw);
        else if (mode == 2) hipLaunchKernelGGL(fft256<2>, grid, blk, 0, 0, din[k], dout[k], dtw);
        else hipLaunchKernelGGL(fft256<3>, grid, blk, 0, 0, din[k], dout[k], dtw);
    };
    fill(h, 12345);  // buffer 0's input
    std::vector<float2> o[4];
    for (int m = 0; m < 4; ++m) {
        launch(m, BMAX, 0);
        CK(hipGetLastError());
        CK(hipDeviceSynchronize());
        o[m].resize(total);
        CK(hipMemcpy(o[m].data(), dout[0], total * sizeof(float2), hipMemcpyDeviceToHost));
    }
    double num = 0, den = 0;
    for (int q = 0; q < 64; ++q)
        for (int k = 0; k < 256; ++k) {
            std::complex<double> acc = 0;
            for (int n = 0; n < 256; ++n) {
                const double ph = -2 * M_PI * (double)((n * k) % 256) / 256.0;
                acc += std::complex<double>(h[q * 256 + n].x, h[q * 256 + n].y) * std::complex<double>(cos(ph), sin(ph));
            }
            const std::complex<double> y(o[0][q * 256 + k].x, o[0][q * 256 + k].y);
            num += std::norm(y - acc);
            den += std::norm(acc);
        }
    const double rel = sqrt(num / den);
    printf("LDS exchange against a double DFT (64 sequences): relL2 %.3e\n", rel);
    bool ok = rel < 1e-6;
    for (int m = 1; m < 4; ++m) {
        long long diff = 0;
        for (long long i = 0; i < total; ++i)
            if (o[m][i].x != o[0][i].x || o[m][i].y != o[0][i].y) ++diff;
        printf("%s exchange: %lld of %lld elements differ from the LDS exchange\n", name[m], diff, total);
        ok = ok && diff == 0;
    }
    fflush(stdout);
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    struct Case {
        long long B;
        int nbuf;
        const char* what;
    };
    const Case cases[2] = {{BMAX, NBUF, "65536 x 256, four rotating 128 MiB buffers (HBM)"},
                           {4096, 1, "4096 x 256, one 8 MiB buffer (cache-resident)"}};
    for (const Case& c : cases) {
        printf("%s\n", c.what);
        double best[4] = {1e30, 1e30, 1e30, 1e30};
        for (int rep = 0; rep < 3; ++rep)
            for (int m = 0; m < 4; ++m) {
                for (int i = 0; i < 8; ++i) launch(m, c.B, i % c.nbuf);
                CK(hipEventRecord(e0, 0));
                const int iters = 100;
                for (int i = 0; i < iters; ++i) launch(m, c.B, i % c.nbuf);
                CK(hipEventRecord(e1, 0));
                CK(hipEventSynchronize(e1));
                float ms = 0;
                CK(hipEventElapsedTime(&ms, e0, e1));
                best[m] = std::min(best[m], ms * 1e3 / iters);
            }
        for (int m = 0; m < 4; ++m)
            printf("  %-13s %8.2f us  %6.0f GB/s  %+6.1f%% against LDS\n", name[m], best[m], 2.0 * c.B * N * 8 / best[m] / 1e3,
                   100.0 * (best[m] / best[0] - 1.0));
        fflush(stdout);
    }
    for (int k = 0; k < NBUF; ++k) {
        CK(hipFree(din[k]));
        CK(hipFree(dout[k]));
    }
    CK(hipFree(dtw));
    printf(ok ? "ALL OK\n" : "MISMATCH\n");
    return ok ? 0 : 1;
}
