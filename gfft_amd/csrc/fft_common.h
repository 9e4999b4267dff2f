// fft_common.h -- complex arithmetic and register butterflies for the gfx950 FFT kernels.
//
// Every butterfly computes a forward DFT (exp(-2 pi i nk/R)) on values held in registers.
// Inverse transforms are produced by conjugating on load and on store, so one set of
// butterflies and one set of twiddle tables serve both directions.
#pragma once
#include <hip/hip_runtime.h>

namespace gfftx {

template <typename T> struct vec2;
template <> struct vec2<float> { using type = float2; };
template <> struct vec2<double> { using type = double2; };
template <typename T> using V2 = typename vec2<T>::type;

template <typename V> __device__ __forceinline__ V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __device__ __forceinline__ V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V> __device__ __forceinline__ V cmul(V a, V b)
{
    return V{__builtin_fma(a.x, b.x, -(a.y * b.y)), __builtin_fma(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return float2{__builtin_fmaf(a.x, b.x, -(a.y * b.y)), __builtin_fmaf(a.x, b.y, a.y * b.x)};
}
// multiply by -i and +i
template <typename V> __device__ __forceinline__ V mul_mi(V a) { return V{a.y, -a.x}; }
template <typename V> __device__ __forceinline__ V mul_pi(V a) { return V{-a.y, a.x}; }
template <typename V> __device__ __forceinline__ V cconj(V a) { return V{a.x, -a.y}; }
template <typename V, typename S> __device__ __forceinline__ V cscale(V a, S s) { return V{a.x * s, a.y * s}; }

// Mathematical constants (exact decimal expansions, rounded by the compiler).
#define GFFTX_C8 0.70710678118654752440084436210484903928   // cos(pi/4)
#define GFFTX_C16 0.92387953251128675612818318939678828682  // cos(pi/8)
#define GFFTX_S16 0.38268343236508977172845998403039886676  // sin(pi/8)
#define GFFTX_S3 0.86602540378443864676372317075293618347   // sin(pi/3)
#define GFFTX_C5A 0.30901699437494742410229341718281905886  // cos(2pi/5)
#define GFFTX_C5B 0.80901699437494742410229341718281905886  // -cos(4pi/5)
#define GFFTX_S5A 0.95105651629515357211643933337938214340  // sin(2pi/5)
#define GFFTX_S5B 0.58778525229247312916870595463907276860  // sin(4pi/5)
#define GFFTX_C7A 0.62348980185873353052500488400423981063   // cos(2pi/7)
#define GFFTX_C7B -0.22252093395631440428890256449679475947  // cos(4pi/7)
#define GFFTX_C7C -0.90096886790241912623610231950744505117  // cos(6pi/7)
#define GFFTX_S7A 0.78183148246802980870844452667405775023   // sin(2pi/7)
#define GFFTX_S7B 0.97492791218182360701813168299393121723   // sin(4pi/7)
#define GFFTX_S7C 0.43388373911755812047576833284835875461   // sin(6pi/7)

template <int R> struct Dft;

template <> struct Dft<2> {
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        const V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <> struct Dft<3> {
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        using S = decltype(v[0].x);
        const V t1 = cadd(v[1], v[2]);
        const V t2 = V{v[0].x - (S)0.5 * t1.x, v[0].y - (S)0.5 * t1.y};
        const V t3 = cscale(csub(v[1], v[2]), (S)GFFTX_S3);
        v[0] = cadd(v[0], t1);
        // X1 = t2 - i t3, X2 = t2 + i t3
        v[1] = cadd(t2, mul_mi(t3));
        v[2] = csub(t2, mul_mi(t3));
    }
};

template <> struct Dft<4> {
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        const V s0 = cadd(v[0], v[2]), d0 = csub(v[0], v[2]);
        const V s1 = cadd(v[1], v[3]), d1 = mul_mi(csub(v[1], v[3]));
        v[0] = cadd(s0, s1);
        v[2] = csub(s0, s1);
        v[1] = cadd(d0, d1);
        v[3] = csub(d0, d1);
    }
};

template <> struct Dft<5> {
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        using S = decltype(v[0].x);
        const S ca = (S)GFFTX_C5A, cb = -(S)GFFTX_C5B, sa = (S)GFFTX_S5A, sb = (S)GFFTX_S5B;
        const V s1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        const V s2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        const V a1 = V{v[0].x + ca * s1.x + cb * s2.x, v[0].y + ca * s1.y + cb * s2.y};
        const V a2 = V{v[0].x + cb * s1.x + ca * s2.x, v[0].y + cb * s1.y + ca * s2.y};
        // b1 = sa*d1 + sb*d2, b2 = sb*d1 - sa*d2 ; X1 = a1 - i b1, X4 = a1 + i b1, X2 = a2 - i b2, X3 = a2 + i b2
        const V b1 = V{sa * d1.x + sb * d2.x, sa * d1.y + sb * d2.y};
        const V b2 = V{sb * d1.x - sa * d2.x, sb * d1.y - sa * d2.y};
        v[0] = cadd(v[0], cadd(s1, s2));
        v[1] = cadd(a1, mul_mi(b1));
        v[4] = csub(a1, mul_mi(b1));
        v[2] = cadd(a2, mul_mi(b2));
        v[3] = csub(a2, mul_mi(b2));
    }
};

template <> struct Dft<7> {
    // s_k = x_k + x_{7-k}, d_k = x_k - x_{7-k} (k = 1..3);
    // X_j = x_0 + sum_k cos(2 pi jk/7) s_k - i sum_k sin(2 pi jk/7) d_k, X_{7-j} its partner
    // with +i; jk mod 7 picks the constant (cos is even, sin odd in jk mod 7).
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        using S = decltype(v[0].x);
        const S c1 = (S)GFFTX_C7A, c2 = (S)GFFTX_C7B, c3 = (S)GFFTX_C7C;
        const S s1 = (S)GFFTX_S7A, s2 = (S)GFFTX_S7B, s3 = (S)GFFTX_S7C;
        const V p1 = cadd(v[1], v[6]), m1 = csub(v[1], v[6]);
        const V p2 = cadd(v[2], v[5]), m2 = csub(v[2], v[5]);
        const V p3 = cadd(v[3], v[4]), m3 = csub(v[3], v[4]);
        const V x0 = v[0];
        const V a1 = V{x0.x + c1 * p1.x + c2 * p2.x + c3 * p3.x, x0.y + c1 * p1.y + c2 * p2.y + c3 * p3.y};
        const V a2 = V{x0.x + c2 * p1.x + c3 * p2.x + c1 * p3.x, x0.y + c2 * p1.y + c3 * p2.y + c1 * p3.y};
        const V a3 = V{x0.x + c3 * p1.x + c1 * p2.x + c2 * p3.x, x0.y + c3 * p1.y + c1 * p2.y + c2 * p3.y};
        // jk mod 7 for j = 2: 2, 4, 6 -> s2, -s3, -s1; for j = 3: 3, 6, 2 -> s3, -s1, s2
        const V b1 = V{s1 * m1.x + s2 * m2.x + s3 * m3.x, s1 * m1.y + s2 * m2.y + s3 * m3.y};
        const V b2 = V{s2 * m1.x - s3 * m2.x - s1 * m3.x, s2 * m1.y - s3 * m2.y - s1 * m3.y};
        const V b3 = V{s3 * m1.x - s1 * m2.x + s2 * m3.x, s3 * m1.y - s1 * m2.y + s2 * m3.y};
        v[0] = cadd(x0, cadd(p1, cadd(p2, p3)));
        v[1] = cadd(a1, mul_mi(b1));
        v[6] = csub(a1, mul_mi(b1));
        v[2] = cadd(a2, mul_mi(b2));
        v[5] = csub(a2, mul_mi(b2));
        v[3] = cadd(a3, mul_mi(b3));
        v[4] = csub(a3, mul_mi(b3));
    }
};

template <> struct Dft<8> {
    // n = i + 2m, k = q + 4p:  X[q+4p] = sum_i W2^{ip} W8^{iq} DFT4_m(v[i+2m])[q]
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        using S = decltype(v[0].x);
        const S c = (S)GFFTX_C8;
        V e[4] = {v[0], v[2], v[4], v[6]};
        V o[4] = {v[1], v[3], v[5], v[7]};
        Dft<4>::run(e);
        Dft<4>::run(o);
        // twiddles W8^q on odd part: q=1: (c,-c), q=2: -i, q=3: (-c,-c)
        o[1] = V{c * (o[1].x + o[1].y), c * (o[1].y - o[1].x)};
        o[2] = mul_mi(o[2]);
        o[3] = V{c * (o[3].y - o[3].x), -c * (o[3].x + o[3].y)};
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            v[q] = cadd(e[q], o[q]);
            v[q + 4] = csub(e[q], o[q]);
        }
    }
};

template <> struct Dft<16> {
    // n = i + 4m, k = q + 4p:  X[q+4p] = DFT4_i( W16^{iq} DFT4_m(v[i+4m])[q] )[p]
    template <typename V> __device__ __forceinline__ static void run(V* v)
    {
        using S = decltype(v[0].x);
        const S c8 = (S)GFFTX_C8, c = (S)GFFTX_C16, s = (S)GFFTX_S16;
        V a[4][4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            a[i][0] = v[i];
            a[i][1] = v[i + 4];
            a[i][2] = v[i + 8];
            a[i][3] = v[i + 12];
            Dft<4>::run(a[i]);
        }
        // W16^{iq}: (1,1)=W^1=(c,-s) (1,2)=W^2=(c8,-c8) (1,3)=W^3=(s,-c)
        //           (2,1)=W^2 (2,2)=W^4=-i (2,3)=W^6=(-c8,-c8)
        //           (3,1)=W^3 (3,2)=W^6 (3,3)=W^9=(-c,s)
        const V w1 = V{c, -s}, w2 = V{c8, -c8}, w3 = V{s, -c}, w6 = V{-c8, -c8}, w9 = V{-c, s};
        a[1][1] = cmul(a[1][1], w1);
        a[1][2] = cmul(a[1][2], w2);
        a[1][3] = cmul(a[1][3], w3);
        a[2][1] = cmul(a[2][1], w2);
        a[2][2] = mul_mi(a[2][2]);
        a[2][3] = cmul(a[2][3], w6);
        a[3][1] = cmul(a[3][1], w3);
        a[3][2] = cmul(a[3][2], w6);
        a[3][3] = cmul(a[3][3], w9);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            V b[4] = {a[0][q], a[1][q], a[2][q], a[3][q]};
            Dft<4>::run(b);
            v[q] = b[0];
            v[q + 4] = b[1];
            v[q + 8] = b[2];
            v[q + 12] = b[3];
        }
    }
};

}  // namespace gfftx
