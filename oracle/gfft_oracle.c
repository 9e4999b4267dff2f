/*
 * gfft_oracle.c -- CPU oracle for the GFFT transform path.
 *
 * TEST INFRASTRUCTURE ONLY: this library is the checker (and the timed CPU baseline) for the
 * GPU engine.  Nothing in the product links or calls it.
 *
 * Contents
 *   1. A restatement of GFFT 0.32's algorithm (reference: scpp/gfft, src/), precision-generic
 *      in gfft_oracle_impl.h and instantiated for DOUBLE (T=double, TT=long double) and FLOAT
 *      (T=float, TT=double): GFFTswap2 digit reversal, in-place InTime, out-of-place InTimeOOP,
 *      DFTk_x_Im_T sweeps with GFFT's twiddle recurrences, Separate, Backward and the OpenMP
 *      split.  Functions cite the reference lines they follow.
 *   2. Accurate references that are *not* GFFT: a long-double mixed-radix Stockham FFT with
 *      octant-reduced long-double twiddles (orc_ref_fft) and a long-double O(N^2) DFT with
 *      exact index reduction (orc_dft_ld).  These are the FP64 oracle for large N.
 *
 * The reference cannot be compiled here (its metafactor submodule is absent), so the
 * restatement is pinned by mpmath golden vectors, closed forms and invariants in tests/.
 * The factor order for mixed radices (metafactor's Factorization<>) is taken as ascending
 * primes; this affects rounding only and is parity-unpinned.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>

#define ORC_PI 3.141592653589793238462643383279502884L

/* Prime factorisation, ascending, each prime repeated by multiplicity (the order GFFT's
 * in-file trial division used, src/gfftfactor.h:55-136).  Returns count, or -1 when a prime
 * factor exceeds the O(p^2) leaf limit. */
int orc_factorize(long n, long* f)
{
    int k = 0;
    if (n < 1) return -1;
    for (long p = 2; p * p <= n; ++p)
        while (n % p == 0) { f[k++] = p; n /= p; }
    if (n > 1) f[k++] = n;
    for (int i = 0; i < k; ++i)
        if (f[i] > 63) return -1;
    return k;
}

/* Permutation<K,[(p0,..)]>::value(ii) = (ii % p0)*(K/p0) + Next::value(ii/p0)
 * (src/gfftomp.h:192-221) */
static long orc_perm(long ii, const long* kf, int nk, long K)
{
    if (nk == 0) return ii;
    const long p = kf[0];
    return (ii % p) * (K / p) + orc_perm(ii / p, kf + 1, nk - 1, K / p);
}

#define T double
#define TT long double
#define SUF(x) x##_d
#include "gfft_oracle_impl.h"
#undef T
#undef TT
#undef SUF
#undef BF2W

#define T float
#define TT double
#define SUF(x) x##_f
#include "gfft_oracle_impl.h"
#undef T
#undef TT
#undef SUF
#undef BF2W

/* ======================================================================================
 * Accurate long-double references
 * ====================================================================================== */

/* exp(-2*pi*i*S*m/N) in long double with octant reduction: m is reduced mod N and mapped to
 * [0, N/8] by the symmetries of sin/cos so that sinl/cosl only see |angle| <= pi/4. */
static void orc_root_ld(long m, long N, int S, long double* c, long double* s)
{
    m %= N;
    if (m < 0) m += N;
    /* work with 8m against N: octant o = floor(8m/N) */
    const long double x8 = 8.0L * (long double)m; /* exact for m < 2^60 */
    long o = (long)(x8 / (long double)N);
    if (o > 7) o = 7;
    /* angle = 2pi m/N = (pi/4)*(8m/N); r = 8m - o*N in [0, N) */
    long double r = x8 - (long double)o * (long double)N;
    long double cc, ss;
    /* for odd octants reflect: use (N - r) measured from the next octant boundary */
    if (o & 1) {
        const long double a = (ORC_PI / 4.0L) * ((long double)N - r) / (long double)N;
        /* angle = (o+1)*pi/4 - a */
        const long double ca = cosl(a), sa = sinl(a);
        switch (o) {
            case 1: cc = sa; ss = ca; break;   /* pi/2 - a */
            case 3: cc = -ca; ss = sa; break;  /* pi - a   */
            case 5: cc = -sa; ss = -ca; break; /* 3pi/2 - a */
            default: cc = ca; ss = -sa; break; /* 2pi - a  */
        }
    } else {
        const long double a = (ORC_PI / 4.0L) * r / (long double)N;
        const long double ca = cosl(a), sa = sinl(a);
        switch (o) {
            case 0: cc = ca; ss = sa; break;
            case 2: cc = -sa; ss = ca; break;  /* pi/2 + a */
            case 4: cc = -ca; ss = -sa; break; /* pi + a   */
            default: cc = sa; ss = -ca; break; /* 3pi/2 + a */
        }
    }
    *c = cc;
    *s = -(long double)S * ss;
}

/* Twiddle exp(-2 pi i S m / N) rounded to double: the per-plan table generator is checked
 * against this. */
void orc_root_d(long m, long N, int S, double* c, double* s)
{
    long double cc, ss;
    orc_root_ld(m, N, S, &cc, &ss);
    *c = (double)cc;
    *s = (double)ss;
}

typedef struct { long double re, im; } cld;

/* Long-double Stockham FFT, radices 4/2/3/5 and generic primes; twiddles from a two-level
 * long-double table. src/dst are interleaved doubles; sign S=+1 forward (exp(-)), -1 inverse
 * (exp(+), unnormalised). Returns 0 on success. */
int orc_ref_fft(const double* src, double* dst, long N, int S)
{
    if (N < 1) return -1;
    long f[64];
    int nf = orc_factorize(N, f);
    if (nf < 0) return -1;
    /* radix schedule: pair 2s into 4s */
    long rad[64];
    int nr = 0, n2 = 0;
    for (int i = 0; i < nf; ++i) n2 += (f[i] == 2);
    for (int i = 0; i + 1 < n2; i += 2) rad[nr++] = 4;
    if (n2 & 1) rad[nr++] = 2;
    for (int i = 0; i < nf; ++i)
        if (f[i] != 2) rad[nr++] = f[i];

    cld* a = (cld*)malloc(sizeof(cld) * (size_t)N);
    cld* b = (cld*)malloc(sizeof(cld) * (size_t)N);
    long B = 1;
    while (B * B < N) B <<= 1;
    const long nhi = (N + B - 1) / B + 1;
    cld* tlo = (cld*)malloc(sizeof(cld) * (size_t)B);
    cld* thi = (cld*)malloc(sizeof(cld) * (size_t)nhi);
    if (!a || !b || !tlo || !thi) { free(a); free(b); free(tlo); free(thi); return -2; }
    for (long m = 0; m < B; ++m) orc_root_ld(m, N, S, &tlo[m].re, &tlo[m].im);
    for (long m = 0; m < nhi; ++m) orc_root_ld(m * B, N, S, &thi[m].re, &thi[m].im);
    for (long i = 0; i < N; ++i) { a[i].re = src[2 * i]; a[i].im = src[2 * i + 1]; }

    long Ns = 1;
    for (int s = 0; s < nr; ++s) {
        const long R = rad[s], LR = N / R, step = N / (Ns * R);
        cld wp[64];
        for (long q = 0; q < R; ++q) orc_root_ld(q, R, S, &wp[q].re, &wp[q].im);
        for (long j = 0; j < LR; ++j) {
            const long k = j % Ns;
            cld v[64], o[64];
            for (long r = 0; r < R; ++r) {
                cld x = a[j + r * LR];
                if (r && k) {
                    const long m = (k * r * step) % N;
                    const cld wl = tlo[m % B], wh = thi[m / B];
                    const long double wr = wl.re * wh.re - wl.im * wh.im;
                    const long double wi = wl.re * wh.im + wl.im * wh.re;
                    const long double xr = x.re * wr - x.im * wi;
                    x.im = x.re * wi + x.im * wr;
                    x.re = xr;
                }
                v[r] = x;
            }
            if (R == 2) {
                o[0].re = v[0].re + v[1].re; o[0].im = v[0].im + v[1].im;
                o[1].re = v[0].re - v[1].re; o[1].im = v[0].im - v[1].im;
            } else if (R == 4) {
                const long double s0r = v[0].re + v[2].re, s0i = v[0].im + v[2].im;
                const long double d0r = v[0].re - v[2].re, d0i = v[0].im - v[2].im;
                const long double s1r = v[1].re + v[3].re, s1i = v[1].im + v[3].im;
                const long double d1r = v[1].re - v[3].re, d1i = v[1].im - v[3].im;
                /* (-iS) * d1 */
                const long double rr = S * d1i, ri = -S * d1r;
                o[0].re = s0r + s1r; o[0].im = s0i + s1i;
                o[2].re = s0r - s1r; o[2].im = s0i - s1i;
                o[1].re = d0r + rr; o[1].im = d0i + ri;
                o[3].re = d0r - rr; o[3].im = d0i - ri;
            } else {
                for (long q = 0; q < R; ++q) {
                    long double sr = 0, si = 0;
                    for (long r = 0; r < R; ++r) {
                        const cld w = wp[(q * r) % R];
                        sr += v[r].re * w.re - v[r].im * w.im;
                        si += v[r].re * w.im + v[r].im * w.re;
                    }
                    o[q].re = sr; o[q].im = si;
                }
            }
            const long base = (j / Ns) * Ns * R + k;
            for (long r = 0; r < R; ++r) b[base + r * Ns] = o[r];
        }
        cld* t = a; a = b; b = t;
        Ns *= R;
    }
    for (long i = 0; i < N; ++i) { dst[2 * i] = (double)a[i].re; dst[2 * i + 1] = (double)a[i].im; }
    free(a); free(b); free(tlo); free(thi);
    return 0;
}

/* Long-double O(N^2) DFT with exact (n*k mod N) index reduction into a long-double root
 * table.  Output rounded to double.  For N up to a few thousand. */
int orc_dft_ld(const double* src, double* dst, long N, int S)
{
    cld* w = (cld*)malloc(sizeof(cld) * (size_t)N);
    if (!w) return -2;
    for (long m = 0; m < N; ++m) orc_root_ld(m, N, S, &w[m].re, &w[m].im);
    for (long k = 0; k < N; ++k) {
        long double sr = 0, si = 0;
        long m = 0;
        for (long n = 0; n < N; ++n) {
            const long double xr = src[2 * n], xi = src[2 * n + 1];
            sr += xr * w[m].re - xi * w[m].im;
            si += xr * w[m].im + xi * w[m].re;
            m += k;
            if (m >= N) m -= N;
        }
        dst[2 * k] = (double)sr;
        dst[2 * k + 1] = (double)si;
    }
    free(w);
    return 0;
}

/* Base-M digit reversal of P digits as an index map: out[n] = rev(n).  The closed form that
 * GFFTswap2 implements (src/gfftswap.h:74-140); used to check the swap and the GPU
 * permutation kernel. */
void orc_digit_reverse_map(long M, long P, long* out)
{
    long N = 1;
    for (long i = 0; i < P; ++i) N *= M;
    for (long n = 0; n < N; ++n) {
        long x = n, r = 0;
        for (long i = 0; i < P; ++i) { r = r * M + x % M; x /= M; }
        out[n] = r;
    }
}
