// bluestein.hip -- element-wise kernels of the chirp-z (Bluestein) path for lengths with a
// prime factor above 5.
//
// With c_j = W_{2n}^{j^2} = exp(-i pi j^2 / n) and jk = (j^2 + k^2 - (k-j)^2) / 2,
//     X_k = sum_j x_j W_n^{jk} = c_k * sum_j (x_j c_j) conj(c_{k-j}),
// a linear convolution of length 2n-1, computed as a circular one of a 2/3/5-smooth length
// m >= 2n-1 with two FFTs of length m and the precomputed spectrum of the chirp filter
//     b_j = conj(c_j) (0 <= j < n),  b_{m-j} = conj(c_j) (0 < j < n),  0 elsewhere.
// The inverse transform is conj(DFT(conj x)) / n, so one filter serves both directions.
//
// All kernels are streaming: one complex element per thread and iteration, batch items on
// the y dimension of the grid (grid-stride in both dimensions).
#include "registry.h"

namespace gfftx {

namespace {

// c_j = W_{2n}^{j^2 mod 2n} from the two-level FP64 root table of order 2n.
template <class T> __global__ __launch_bounds__(256) void chirp_table_kernel(ChirpTableArgs a)
{
    V2<T>* c = (V2<T>*)a.chirp;
    const unsigned long long two_n = 2ull * (unsigned long long)a.n;
    const unsigned long long mask = (1ull << a.shift) - 1;
    for (unsigned long long j = blockIdx.x * 256ull + threadIdx.x; j < (unsigned long long)a.n;
         j += (unsigned long long)gridDim.x * 256ull) {
        const unsigned long long m = (j * j) % two_n;  // j < 2^32: j^2 is exact in 64 bits
        const double2 lo = a.lo[m & mask], hi = a.hi[m >> a.shift];
        const double2 w = cmul(lo, hi);
        c[j] = V2<T>{(T)w.x, (T)w.y};
    }
}

// b over one length-m item: conj(c_j) at j < n and at m - j for 0 < j < n, zero elsewhere.
template <class T> __global__ __launch_bounds__(256) void chirp_filter_kernel(BluesteinArgs a)
{
    const V2<T>* c = (const V2<T>*)a.chirp;
    V2<T>* b = (V2<T>*)a.work;
    for (unsigned long long j = blockIdx.x * 256ull + threadIdx.x; j < (unsigned long long)a.m;
         j += (unsigned long long)gridDim.x * 256ull) {
        V2<T> v{0, 0};
        if (j < (unsigned long long)a.n) v = cconj(c[j]);
        else if (j > (unsigned long long)(a.m - a.n)) v = cconj(c[a.m - j]);
        b[j] = v;
    }
}

// work[b][j] = (conj? conj(x[b][j]) : x[b][j]) * c_j for j < n, zero up to m.
template <class T> __global__ __launch_bounds__(256) void bluestein_pre_kernel(BluesteinArgs a)
{
    const V2<T>* c = (const V2<T>*)a.chirp;
    const unsigned flip = a.conj ? 0x80000000u : 0u;
    for (unsigned long long b = blockIdx.y; b < (unsigned long long)a.batch; b += gridDim.y) {
        const V2<T>* x = (const V2<T>*)a.in + (a.boff ? (unsigned long long)a.boff[b] : b * (unsigned long long)a.dist);
        V2<T>* w = (V2<T>*)a.work + b * (unsigned long long)a.m;
        for (unsigned long long j = blockIdx.x * 256ull + threadIdx.x; j < (unsigned long long)a.m;
             j += (unsigned long long)gridDim.x * 256ull) {
            V2<T> v{0, 0};
            if (j < (unsigned long long)a.n) {
                V2<T> u = x[j];
                if (flip) u.y = -u.y;
                v = cmul(u, c[j]);
            }
            w[j] = v;
        }
    }
}

// work[b][k] *= filter[k] over the whole length m.
template <class T> __global__ __launch_bounds__(256) void bluestein_mul_kernel(BluesteinArgs a)
{
    const V2<T>* f = (const V2<T>*)a.filter;
    for (unsigned long long b = blockIdx.y; b < (unsigned long long)a.batch; b += gridDim.y) {
        V2<T>* w = (V2<T>*)a.work + b * (unsigned long long)a.m;
        for (unsigned long long j = blockIdx.x * 256ull + threadIdx.x; j < (unsigned long long)a.m;
             j += (unsigned long long)gridDim.x * 256ull)
            w[j] = cmul(w[j], f[j]);
    }
}

// out[b][k] = scale * (conj? conj(c_k y_k) : c_k y_k) for k < n, y = work[b].
template <class T> __global__ __launch_bounds__(256) void bluestein_post_kernel(BluesteinArgs a)
{
    const V2<T>* c = (const V2<T>*)a.chirp;
    const T sc = (T)a.scale;
    for (unsigned long long b = blockIdx.y; b < (unsigned long long)a.batch; b += gridDim.y) {
        const V2<T>* y = (const V2<T>*)a.work + b * (unsigned long long)a.m;
        V2<T>* o = (V2<T>*)a.out + (a.boff ? (unsigned long long)a.boff[b] : b * (unsigned long long)a.dist);
        for (unsigned long long k = blockIdx.x * 256ull + threadIdx.x; k < (unsigned long long)a.n;
             k += (unsigned long long)gridDim.x * 256ull) {
            V2<T> v = cmul(c[k], y[k]);
            if (a.conj) v.y = -v.y;
            o[k] = V2<T>{v.x * sc, v.y * sc};
        }
    }
}

dim3 stream_grid(long long len, long long batch)
{
    // about 16K workgroups in all (every CU many times over); the kernels stride over the rest
    const long long gy = batch < 65535 ? batch : 65535;
    long long gx = (len + 255) / 256;
    const long long cap = 16384 / gy > 1 ? 16384 / gy : 1;
    if (gx > cap) gx = cap;
    return dim3((unsigned)gx, (unsigned)gy);
}

template <class T> hipError_t launch_bluestein_t(int which, const BluesteinArgs& a, hipStream_t s)
{
    switch (which) {
        case BLUESTEIN_PRE:
            bluestein_pre_kernel<T><<<stream_grid(a.m, a.batch), 256, 0, s>>>(a);
            break;
        case BLUESTEIN_MUL:
            bluestein_mul_kernel<T><<<stream_grid(a.m, a.batch), 256, 0, s>>>(a);
            break;
        case BLUESTEIN_POST:
            bluestein_post_kernel<T><<<stream_grid(a.n, a.batch), 256, 0, s>>>(a);
            break;
        case BLUESTEIN_FILTER:
            chirp_filter_kernel<T><<<stream_grid(a.m, 1), 256, 0, s>>>(a);
            break;
        default:
            return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_bluestein(int prec, int which, const BluesteinArgs& a, hipStream_t s)
{
    return prec == PREC_F64 ? launch_bluestein_t<double>(which, a, s) : launch_bluestein_t<float>(which, a, s);
}

hipError_t launch_chirp_table(int prec, const ChirpTableArgs& a, hipStream_t s)
{
    const dim3 g = stream_grid(a.n, 1);
    if (prec == PREC_F64) chirp_table_kernel<double><<<g, 256, 0, s>>>(a);
    else chirp_table_kernel<float><<<g, 256, 0, s>>>(a);
    return hipGetLastError();
}

}  // namespace gfftx
