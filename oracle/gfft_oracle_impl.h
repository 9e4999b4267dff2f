/*
 * gfft_oracle_impl.h -- precision-generic body of the GFFT restatement.
 *
 * TEST INFRASTRUCTURE ONLY.  Included twice by gfft_oracle.c:
 *   T = double, TT = long double   (GFFT DOUBLE: ValueType/TempType,  src/gfftpolicy.h:78-89)
 *   T = float,  TT = double        (GFFT FLOAT:  ValueType/TempType,  src/gfftpolicy.h:94-105)
 * SUF(x) appends the precision suffix (_d / _f) to every exported name.
 *
 * Data are interleaved (re,im) scalars, exactly as GFFT's `T* data`.
 * Index arithmetic below is written in *complex* units unless a name ends in 2.
 */

/* ---- digit reversal: GFFTswap2<M,P,T> (src/gfftswap.h:74-140) ------------------------------
 * Recursive enumeration of (n, rev(n)) pairs: at level I digit I of n and digit P-1-I of r
 * take the same value; the leaf swaps complex n and r when n > r (src/gfftswap.h:94-101). */
static void SUF(swap_rec)(T* data, long M, long P, long I, long n, long r, long BN, long BR)
{
    if (I == P) {
        if (n > r) {
            T t0 = data[2 * n], t1 = data[2 * n + 1];
            data[2 * n] = data[2 * r];
            data[2 * n + 1] = data[2 * r + 1];
            data[2 * r] = t0;
            data[2 * r + 1] = t1;
        }
        return;
    }
    /* BN = M^I, BR = M^(P-I-1) (src/gfftswap.h:76-77) */
    const long qn = n / BN, rn = n % BN, qr = r / BR, rr = r % BR;
    for (long i = 0; i < M; ++i)
        SUF(swap_rec)(data, M, P, I + 1, (qn + i) * BN + rn, (qr + i) * BR + rr, BN * M,
                      (BR > 1) ? BR / M : 1);
}

void SUF(orc_swap)(T* data, long M, long P)
{
    if (P <= 1) return; /* GFFTswap2<2,1> is a no-op (src/gfftswap.h:136-140) */
    long BR = 1;
    for (long i = 0; i < P - 1; ++i) BR *= M;
    SUF(swap_rec)(data, M, P, 0, 0, 0, 1, BR);
}

/* ---- butterflies: DFTk_inp<K,MS,VType,S> (src/gfftspec_inp.h) -----------------------------
 * MS is the scalar stride between the K inputs (the template parameter M there). */

/* radix 2, no twiddle (src/gfftspec_inp.h:272-287) */
static inline void SUF(bf2)(T* d, long MS)
{
    const T tr = d[MS], ti = d[MS + 1];
    d[MS] = d[0] - tr;
    d[MS + 1] = d[1] - ti;
    d[0] += tr;
    d[1] += ti;
}
/* radix 2 with twiddle (src/gfftspec_inp.h:289-298); LT is the twiddle type */
#define BF2W(d, MS, wr, wi)                                   \
    do {                                                      \
        const T tr_ = (T)((d)[MS] * (wr) - (d)[(MS) + 1] * (wi)); \
        const T ti_ = (T)((d)[MS] * (wi) + (d)[(MS) + 1] * (wr)); \
        (d)[MS] = (d)[0] - tr_;                               \
        (d)[(MS) + 1] = (d)[1] - ti_;                         \
        (d)[0] += tr_;                                        \
        (d)[1] += ti_;                                        \
    } while (0)
/* radix 2 times (-iS) (src/gfftspec_inp.h:300-315) */
static inline void SUF(bf2_1)(T* d, long MS, int S)
{
    const T tr = (S < 0) ? -d[MS + 1] : d[MS + 1];
    const T ti = (S < 0) ? d[MS] : -d[MS];
    d[MS] = d[0] - tr;
    d[MS + 1] = d[1] - ti;
    d[0] += tr;
    d[1] += ti;
}

/* radix 3 (src/gfftspec_inp.h:179-263); coef = S*sqrt(3)/2 in T (:195) */
static inline void SUF(bf3_core)(T* d, long MS, T coef, T r1, T i1, T r2, T i2)
{
    const T sum_r = r1 + r2;
    const T dif_r = coef * (r1 - r2);
    const T sum_i = i1 + i2;
    const T dif_i = coef * (i1 - i2);
    const T tr = d[0] - (T)0.5 * sum_r;
    const T ti = d[1] - (T)0.5 * sum_i;
    d[0] += sum_r;
    d[1] += sum_i;
    d[MS] = tr + dif_i;
    d[MS + 1] = ti - dif_r;
    d[2 * MS] = tr - dif_i;
    d[2 * MS + 1] = ti + dif_r;
}

/* odd prime K > 3: O(K^2) in-place DFT (src/gfftspec_inp.h:41-122).
 * c[k] = cos(2pi(k+1)/K), s[k] = S*sin(2pi(k+1)/K) in TT (ComputeTwiddles, src/twiddles.h:177-198) */
static void SUF(bfp_transform)(T* d, long K, long MS, const TT* c, const TT* s,
                               const T* sr, const T* si, const T* dr, const T* di)
{
    const long H = (K - 1) / 2, NM = K * MS;
    for (long i = 1; i < H + 1; ++i) {
        T re1 = 0, re2 = 0, im1 = 0, im2 = 0;
        for (long j = 0; j < H; ++j) {
            const int sign_change = (i * (j + 1) % K) > H;
            const long kk = (i + j * i) % K;
            const long k = (kk > H) ? K - kk - 1 : kk - 1;
            const T s1 = (T)(s[k] * di[j]);
            const T s2 = (T)(s[k] * dr[j]);
            /* the product stays in TempType and the sum is rounded to T, as GFFT's
               `re1 += m_c[k]*sr[j]` with re1 of type T (src/gfftspec_inp.h:63-64) */
            re1 += c[k] * sr[j];
            im1 += c[k] * si[j];
            re2 += sign_change ? -s1 : s1;
            im2 -= sign_change ? -s2 : s2;
        }
        const long k = i * MS;
        d[k] = d[0] + re1 + re2;
        d[k + 1] = d[1] + im1 + im2;
        d[NM - k] = d[0] + re1 - re2;
        d[NM - k + 1] = d[1] + im1 - im2;
    }
    for (long i = 0; i < H; ++i) {
        d[0] += sr[i];
        d[1] += si[i];
    }
}

#define ORC_MAXP 64 /* largest odd prime factor handled by the O(p^2) leaves */

typedef struct {
    TT c[ORC_MAXP], s[ORC_MAXP];
} SUF(ptw_t);

static void SUF(ptw_init)(SUF(ptw_t) * t, long K, int S)
{
    for (long k = 1; k <= (K - 1) / 2; ++k) {
        t->c[k - 1] = (TT)cosl(2.0L * ORC_PI * (long double)k / (long double)K);
        t->s[k - 1] = (TT)(S * sinl(2.0L * ORC_PI * (long double)k / (long double)K));
    }
}

/* DFTk_inp<K>::apply(data) / apply(data, wr, wi) for any prime K (twiddles optional).
 * wr/wi (TT) hold the K-1 twiddles for inputs 1..K-1 (src/gfftspec_inp.h:106-122, :212-232). */
static void SUF(bfk)(T* d, long K, long MS, int S, const TT* wr, const TT* wi)
{
    if (K == 2) {
        if (wr) BF2W(d, MS, wr[0], wi[0]);
        else SUF(bf2)(d, MS);
        return;
    }
    if (K == 3) {
        /* m_coef = S*sqrt(3)*0.5 evaluated in T (src/gfftspec_inp.h:195) */
        const T coef = (T)(S * (T)sqrtl(3.0L) * 0.5);
        T r1 = d[MS], i1 = d[MS + 1], r2 = d[2 * MS], i2 = d[2 * MS + 1];
        if (wr) {
            const T tr1 = (T)(d[MS] * wr[0] - d[MS + 1] * wi[0]);
            const T ti1 = (T)(d[MS] * wi[0] + d[MS + 1] * wr[0]);
            const T tr2 = (T)(d[2 * MS] * wr[1] - d[2 * MS + 1] * wi[1]);
            const T ti2 = (T)(d[2 * MS] * wi[1] + d[2 * MS + 1] * wr[1]);
            r1 = tr1; i1 = ti1; r2 = tr2; i2 = ti2;
        }
        SUF(bf3_core)(d, MS, coef, r1, i1, r2, i2);
        return;
    }
    SUF(ptw_t) tw;
    SUF(ptw_init)(&tw, K, S);
    const long H = (K - 1) / 2, NM = K * MS;
    T sr[ORC_MAXP], si[ORC_MAXP], dr[ORC_MAXP], di[ORC_MAXP];
    for (long i = 0; i < H; ++i) {
        const long k = (i + 1) * MS;
        if (wr) {
            const T tr1 = (T)(d[k] * wr[i] - d[k + 1] * wi[i]);
            const T ti1 = (T)(d[k] * wi[i] + d[k + 1] * wr[i]);
            const T tr2 = (T)(d[NM - k] * wr[K - i - 2] - d[NM - k + 1] * wi[K - i - 2]);
            const T ti2 = (T)(d[NM - k] * wi[K - i - 2] + d[NM - k + 1] * wr[K - i - 2]);
            sr[i] = tr1 + tr2; si[i] = ti1 + ti2; dr[i] = tr1 - tr2; di[i] = ti1 - ti2;
        } else {
            sr[i] = d[k] + d[NM - k];
            si[i] = d[k + 1] + d[NM - k + 1];
            dr[i] = d[k] - d[NM - k];
            di[i] = d[k + 1] - d[NM - k + 1];
        }
    }
    SUF(bfp_transform)(d, K, MS, tw.c, tw.s, sr, si, dr, di);
}

/* apply_m: twiddle on input 0 too, then apply(data, wr+1, wi+1) (src/gfftspec_inp.h:124-132,
 * :233-241, :316-327). */
static void SUF(bfk_m)(T* d, long K, long MS, int S, const TT* wr, const TT* wi)
{
    if (K == 2) {
        /* radix-2 apply_m computes both products before the butterfly (:316-327) */
        const T tr0 = (T)(d[0] * wr[0] - d[1] * wi[0]);
        const T ti0 = (T)(d[0] * wi[0] + d[1] * wr[0]);
        const T tr1 = (T)(d[MS] * wr[1] - d[MS + 1] * wi[1]);
        const T ti1 = (T)(d[MS] * wi[1] + d[MS + 1] * wr[1]);
        d[MS] = tr0 - tr1;
        d[MS + 1] = ti0 - ti1;
        d[0] = tr0 + tr1;
        d[1] = ti0 + ti1;
        return;
    }
    const T t = d[0];
    d[0] = (T)(t * wr[0] - d[1] * wi[0]);
    d[1] = (T)(t * wi[0] + d[1] * wr[0]);
    SUF(bfk)(d, K, MS, S, wr + 1, wi + 1);
}

/* ---- roots of unity ---------------------------------------------------------------------
 * GetFirstRoot<N,S> = cos(2pi/N) - i*S*sin(2pi/N) (src/metaroot.h:104-120), evaluated to
 * 18 digits and rounded to double by Compute<> (src/metafunc.h:239-280).  Restated as the
 * long-double value rounded to double. */
static void SUF(first_root)(long N, int S, double* wr, double* wi)
{
    const long double a = 2.0L * ORC_PI / (long double)N;
    *wr = (double)cosl(a);
    *wi = (double)(-S * sinl(a));
}

/* ---- DFTk_x_Im_T: twiddled radix-K sweep ------------------------------------------------
 * General K (src/gfftalg.h:169-192) with ComputeRoots<K> recurrence in TT
 * (src/twiddles.h:212-231, :300-339).  MC = M (complex) = span of the sweep; Step in complex. */
static void SUF(xim_general)(T* d, long K, long MC, long Step, int S, double w1r, double w1i)
{
    TT wr[ORC_MAXP], wi[ORC_MAXP], wpr[ORC_MAXP], wpi[ORC_MAXP];
    const long MS = 2 * MC;
    SUF(bfk)(d, K, MS, S, NULL, NULL);
    wpr[0] = (TT)w1r;
    wpi[0] = (TT)w1i;
    for (long i = 0; i < K - 2; ++i) {
        wpr[i + 1] = wpr[i] * wpr[0] - wpi[i] * wpi[0];
        wpi[i + 1] = wpr[i] * wpi[0] + wpr[0] * wpi[i];
    }
    for (long i = 0; i < K - 1; ++i) { wr[i] = wpr[i]; wi[i] = wpi[i]; }
    SUF(bfk)(d + 2 * Step, K, MS, S, wr, wi);
    for (long j = 2 * Step; j < MC; j += Step) {
        for (long i = 0; i < K - 1; ++i) { /* RootsContainer::step (src/twiddles.h:222-230) */
            const TT t = wr[i];
            wr[i] = t * wpr[i] - wi[i] * wpi[i];
            wi[i] = wi[i] * wpr[i] + t * wpi[i];
        }
        SUF(bfk)(d + 2 * j, K, MS, S, wr, wi);
    }
}

/* Radix 2 (src/gfftalg.h:228-267): mirrored recurrence in T; M = MC complex, S2 = 2*Step.
 * M/Step == 2 takes the short specialisation (src/gfftalg.h:355-367). */
static void SUF(xim_radix2)(T* d, long MC, long Step, int S, double w1r, double w1i)
{
    const long MS = 2 * MC; /* scalar stride of the butterfly == "N" of the class in scalars */
    const long S2 = 2 * Step;
    SUF(bf2)(d, MS);
    if (MC / Step == 2) {
        SUF(bf2_1)(d + S2, MS, S);
        return;
    }
    if (MC % 2 == 0) SUF(bf2_1)(d + MC, MS, S);
    T wr, wi, t;
    const T wpr = (T)w1r, wpi = (T)w1i;
    wr = wpr;
    wi = wpi;
    BF2W(d + S2, MS, wr, wi);
    t = -wr;
    BF2W(d + MS - S2, MS, t, wi);
    for (long i = S2 + S2; i < MC; i += S2) {
        t = wr;
        wr = wr * wpr - wi * wpi;
        wi = wi * wpr + t * wpi;
        BF2W(d + i, MS, wr, wi);
        t = -wr;
        BF2W(d + MS - i, MS, t, wi);
    }
}

static void SUF(xim)(T* d, long K, long MC, long Step, int S, double w1r, double w1i)
{
    if (K == 2) SUF(xim_radix2)(d, MC, Step, S, w1r, w1i);
    else SUF(xim_general)(d, K, MC, Step, S, w1r, w1i);
}

/* ---- InTime: in-place recursive DIT for N = p^P (src/gfftalg.h:383-442) ----------------- */
static void SUF(intime)(T* d, long N, long p, int S)
{
    if (N == p) { /* prime leaf DFTk_inp<N, C> (src/gfftalg.h:431-442) */
        SUF(bfk)(d, p, 2, S, NULL, NULL);
        return;
    }
    const long M = N / p;
    for (long m = 0; m < p; ++m) SUF(intime)(d + 2 * m * M, M, p, S);
    double wr, wi;
    SUF(first_root)(N, S, &wr, &wi);
    SUF(xim)(d, p, M, 1, S, wr, wi);
}

/* ---- InTimeOOP: out-of-place recursive DIT, any N (src/gfftalg.h:459-507) -------------
 * fac[] lists the prime factors in the order they are consumed (ascending primes, each
 * repeated by its multiplicity); nf is how many remain. */
static void SUF(intime_oop)(const T* src, T* dst, long N, const long* fac, int nf, long LastK, int S)
{
    if (nf == 1) { /* leaf DFTk<N, LastK*2, 2> (src/gfftspec.h:42-188) */
        const long K = N, SI = 2 * LastK;
        if (K == 2) {
            const T tr = src[0] - src[SI], ti = src[1] - src[SI + 1];
            dst[0] = src[0] + src[SI];
            dst[1] = src[1] + src[SI + 1];
            dst[2] = tr;
            dst[3] = ti;
            return;
        }
        if (K == 3) {
            const T coef = (T)(S * (T)sqrtl(3.0L) * 0.5);
            const T sr = src[SI] + src[2 * SI];
            const T dr = coef * (src[SI] - src[2 * SI]);
            const T si = src[SI + 1] + src[2 * SI + 1];
            const T di = coef * (src[SI + 1] - src[2 * SI + 1]);
            const T tr = src[0] - (T)0.5 * sr;
            const T ti = src[1] - (T)0.5 * si;
            dst[0] = src[0] + sr;
            dst[1] = src[1] + si;
            dst[2] = tr + di;
            dst[3] = ti - dr;
            dst[4] = tr - di;
            dst[5] = ti + dr;
            return;
        }
        SUF(ptw_t) tw;
        SUF(ptw_init)(&tw, K, S);
        const long H = (K - 1) / 2, NSI = K * SI, NDI = 2 * K;
        T sr[ORC_MAXP], si[ORC_MAXP], dr[ORC_MAXP], di[ORC_MAXP];
        for (long i = 0; i < H; ++i) {
            const long k = (i + 1) * SI;
            sr[i] = src[k] + src[NSI - k];
            si[i] = src[k + 1] + src[NSI - k + 1];
            dr[i] = src[k] - src[NSI - k];
            di[i] = src[k + 1] - src[NSI - k + 1];
        }
        for (long i = 1; i < H + 1; ++i) {
            T re1 = 0, re2 = 0, im1 = 0, im2 = 0;
            for (long j = 0; j < H; ++j) {
                const int sign_change = (i * (j + 1) % K) > H;
                const long kk = (i + j * i) % K;
                const long kx = (kk > H) ? K - kk - 1 : kk - 1;
                const T s1 = (T)(tw.s[kx] * di[j]);
                const T s2 = (T)(tw.s[kx] * dr[j]);
                re1 += tw.c[kx] * sr[j];  /* TempType product and sum, rounded to T */
                im1 += tw.c[kx] * si[j];
                re2 += sign_change ? -s1 : s1;
                im2 -= sign_change ? -s2 : s2;
            }
            const long k = 2 * i;
            dst[k] = src[0] + re1 + re2;
            dst[k + 1] = src[1] + im1 + im2;
            dst[NDI - k] = src[0] + re1 - re2;
            dst[NDI - k + 1] = src[1] + im1 - im2;
        }
        dst[0] = src[0];
        dst[1] = src[1];
        for (long i = 0; i < H; ++i) {
            dst[0] += sr[i];
            dst[1] += si[i];
        }
        return;
    }
    const long K = fac[0], M = N / K;
    for (long m = 0; m < K; ++m)
        SUF(intime_oop)(src + 2 * m * LastK, dst + 2 * m * M, M, fac + 1, nf - 1, K * LastK, S);
    double wr, wi;
    SUF(first_root)(N, S, &wr, &wi);
    SUF(xim)(dst, K, M, 1, S, wr, wi);
}

/* ---- Separate: real-transform split pass (src/gfftswap.h:239-287), N complex ---------- */
void SUF(orc_separate)(T* data, long N, int S)
{
    const int MF = (S == 1) ? 2 : 1;
    TT wtemp, wr, wi, wpr, wpi, h1r, h1i, h2r, h2i, h3r, h3i;
    /* Sin<2N,1> = sin(pi/(2N)), Sin<N,1> = sin(pi/N) (src/pseudometafunc.h:88-110) */
    wtemp = (TT)sinl(ORC_PI / (2.0L * (long double)N));
    wpr = (TT)(-2.) * wtemp * wtemp;
    wpi = (TT)(-S) * (TT)sinl(ORC_PI / (long double)N);
    wr = (TT)1. + wpr;
    wi = wpi;
    for (long i = 1; i < N / 2; ++i) {
        const long i1 = i + i, i2 = i1 + 1, i3 = 2 * N - i1, i4 = i3 + 1;
        h1r = (TT)0.5 * (data[i1] + data[i3]);
        h1i = (TT)0.5 * (data[i2] - data[i4]);
        h2r = (TT)S * (TT)0.5 * (data[i2] + data[i4]);
        h2i = -(TT)S * (TT)0.5 * (data[i1] - data[i3]);
        h3r = wr * h2r - wi * h2i;
        h3i = wr * h2i + wi * h2r;
        data[i1] = (T)(h1r + h3r);
        data[i2] = (T)(h1i + h3i);
        data[i3] = (T)(h1r - h3r);
        data[i4] = (T)(-h1i + h3i);
        wtemp = wr;
        wr += wr * wpr - wi * wpi;
        wi += wi * wpr + wtemp * wpi;
    }
    h1r = data[0];
    data[0] = (T)(MF * 0.5 * (h1r + data[1]));
    data[1] = (T)(MF * 0.5 * (h1r - data[1]));
    if (N > 1) data[N + 1] = -data[N + 1];
}

/* Backward: every scalar divided by N (src/gfftswap.h:355-362) */
void SUF(orc_backward)(T* data, long N)
{
    for (long i = 0; i < 2 * N; ++i) data[i] /= (T)N;
}

/* ---- public compositions (src/gfftpolicy.h:151-161, :193-201, :262-297) ---------------- */

/* In-place DFT/IDFT for N = p^P: [Swap, InTime, Direction]. Returns -1 if N is not a prime power. */
int SUF(orc_gfft_inplace)(T* data, long N, int S)
{
    long f[64];
    int nf = orc_factorize(N, f);
    if (nf < 0) return -1;
    if (N == 1) return 0;
    for (int i = 1; i < nf; ++i)
        if (f[i] != f[0]) return -1; /* multi-factor in-place is not allowed (src/gfftalg.h:386-395) */
    SUF(orc_swap)(data, f[0], nf);
    SUF(intime)(data, N, f[0], S);
    if (S < 0) SUF(orc_backward)(data, N);
    return 0;
}

/* Out-of-place DFT/IDFT, any N with prime factors <= ORC_MAXP: [InTimeOOP, Direction]. */
int SUF(orc_gfft_oop)(const T* src, T* dst, long N, int S)
{
    long f[64];
    int nf = orc_factorize(N, f);
    if (nf < 0) return -1;
    if (N == 1) { dst[0] = src[0]; dst[1] = src[1]; return 0; }
    SUF(intime_oop)(src, dst, N, f, nf, 1, S);
    if (S < 0) SUF(orc_backward)(dst, N);
    return 0;
}

/* RDFT (S=+1) / IRDFT (S=-1) in place on 2N reals, N = p^P (src/gfftpolicy.h:262-297):
 * RDFT = [Swap, InTime, Forward, Separate], IRDFT = [Separate, Swap, InTime, Backward]. */
int SUF(orc_gfft_real_inplace)(T* data, long N, int S)
{
    if (S < 0) SUF(orc_separate)(data, N, S);
    int rc = SUF(orc_gfft_inplace)(data, N, S);
    if (rc) return rc;
    if (S > 0) SUF(orc_separate)(data, N, S);
    return 0;
}

/* Real transform for any N through the out-of-place complex path; GFFT itself stops here
 * ("Not implemented!", src/gfftswap.h:282-286).  RDFT: InTimeOOP then Separate on dst;
 * IRDFT: Separate on a copy of src, then InTimeOOP and Backward. */
int SUF(orc_gfft_real_oop)(const T* src, T* dst, long N, int S)
{
    if (S > 0) {
        int rc = SUF(orc_gfft_oop)(src, dst, N, S);
        if (rc) return rc;
        SUF(orc_separate)(dst, N, S);
        return 0;
    }
    T* tmp = (T*)malloc(sizeof(T) * 2 * (size_t)N);
    if (!tmp) return -2;
    memcpy(tmp, src, sizeof(T) * 2 * (size_t)N);
    SUF(orc_separate)(tmp, N, S);
    int rc = SUF(orc_gfft_oop)(tmp, dst, N, S);
    free(tmp);
    return rc;
}

/* ---- OpenMP split (src/gfftomp.h) --------------------------------------------------------
 * Active when N > NT^2 and N >= 256 (src/gfftpolicy.h:408-412).  The factor list becomes
 * [(NT,1), Factorize(N/NT)...] (src/gfftpolicy.h:424-435, NT | N); NT sub-transforms run in
 * parallel (ParallLoop, src/gfftomp.h:41-80) and one thread runs the radix-NT combine
 * DFTk_x_Im_T_omp (src/gfftomp.h:364-422) whose inner K-point DFT is DFTk_inp_adapter
 * (src/gfftomp.h:224-293) over the digit-reversed sub-blocks, with roots permuted by
 * Permutation<K,KFact> (src/gfftomp.h:192-221). */


/* K-point DFT over the strided elements d[0], d[MC], .. d[(K-1)MC] stored in digit-reversed
 * order, with input twiddles wr/wi (NULL for j == 0): DFTk_inp_adapter::apply. */
static void SUF(adapter)(T* d, long K, const long* kf, int nk, long MC, int S, double w1r,
                         double w1i, const TT* wr, const TT* wi, int with_m)
{
    if (nk == 1) { /* leaf: DFTk_inp<K, MC*2> (src/gfftomp.h:291-293) */
        if (with_m) SUF(bfk_m)(d, K, 2 * MC, S, wr, wi);
        else SUF(bfk)(d, K, 2 * MC, S, wr, wi);
        return;
    }
    const long KF = kf[0], KN = K / KF;
    /* root for the next level is W1^KF, i.e. the first root of K/KF */
    double nr, ni;
    SUF(first_root)(KN, S, &nr, &ni);
    if (!wr) {
        for (long i = 0; i < KF; ++i)
            SUF(adapter)(d + 2 * i * MC * KN, KN, kf + 1, nk - 1, MC, S, nr, ni, NULL, NULL, 0);
    } else if (!with_m) {
        SUF(adapter)(d, KN, kf + 1, nk - 1, MC, S, nr, ni, wr, wi, 0);
        for (long i = 1; i < KF; ++i)
            SUF(adapter)(d + 2 * i * MC * KN, KN, kf + 1, nk - 1, MC, S, nr, ni, wr + i * KN - 1,
                         wi + i * KN - 1, 1);
    } else {
        for (long i = 0; i < KF; ++i)
            SUF(adapter)(d + 2 * i * MC * KN, KN, kf + 1, nk - 1, MC, S, nr, ni, wr + i * KN,
                         wi + i * KN, 1);
    }
    /* dft_scaled = DFTk_x_Im_T<KF, KF*KN, KN*MC, Step=MC> with root W1 of this level */
    SUF(xim)(d, KF, KN * MC, MC, S, w1r, w1i);
}

/* DFTk_x_Im_T_omp::apply (src/gfftomp.h:409-420): sequential radix-K combine of M columns */
static void SUF(xim_omp)(T* d, long K, const long* kf, int nk, long M, int S, long N)
{
    double w1r, w1i, wkr, wki;
    SUF(first_root)(N, S, &w1r, &w1i); /* W1: root of this level's length N = K*M */
    SUF(first_root)(K, S, &wkr, &wki); /* W = GetFirstRoot<K,S> for the adapter */
    TT wr[ORC_MAXP * 4], wi[ORC_MAXP * 4], wpr[ORC_MAXP * 4], wpi[ORC_MAXP * 4];
    /* ComputeRoots<K,VType,W1,Perm>::init (src/twiddles.h:274-296) */
    wpr[0] = (TT)w1r;
    wpi[0] = (TT)w1i;
    for (long i = 0; i < K - 2; ++i) {
        wpr[i + 1] = wpr[i] * wpr[0] - wpi[i] * wpi[0];
        wpi[i + 1] = wpr[i] * wpi[0] + wpr[0] * wpi[i];
    }
    for (long i = 0; i < K - 1; ++i) {
        const long ii = orc_perm(i + 1, kf, nk, K) - 1;
        wr[ii] = wpr[i];
        wi[ii] = wpi[i];
    }
    for (long i = 0; i < K - 1; ++i) { wpr[i] = wr[i]; wpi[i] = wi[i]; }
    SUF(adapter)(d, K, kf, nk, M, S, wkr, wki, NULL, NULL, 0);
    SUF(adapter)(d + 2, K, kf, nk, M, S, wkr, wki, wr, wi, 0);
    for (long j = 2; j < M; ++j) {
        for (long i = 0; i < K - 1; ++i) {
            const TT t = wr[i];
            wr[i] = t * wpr[i] - wi[i] * wpi[i];
            wi[i] = wi[i] * wpr[i] + t * wpi[i];
        }
        SUF(adapter)(d + 2 * j, K, kf, nk, M, S, wkr, wki, wr, wi, 0);
    }
}

/* In-place OpenMP path: [Swap(2, log2 N), InTime_omp, Direction] for N = 2^n, NT = 2^t
 * (src/gfftpolicy.h:414-422 gives the full-length bit reversal). */
int SUF(orc_gfft_inplace_omp)(T* data, long N, int S, int nt)
{
    long f[64];
    int nf = orc_factorize(N, f);
    if (nf < 0) return -1;
    if (nt <= 1 || !(N > (long)nt * nt && N >= 256) || (N % nt)) return SUF(orc_gfft_inplace)(data, N, S);
    for (int i = 0; i < nf; ++i)
        if (f[i] != 2) return -1;
    long kf[64];
    int nk = orc_factorize(nt, kf);
    if (nk < 0) return -1;
    for (int i = 0; i < nk; ++i)
        if (kf[i] != 2) return -1;
    const long K = nt, M = N / K;
    SUF(orc_swap)(data, 2, nf);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (long tid = 0; tid < K; ++tid) SUF(intime)(data + 2 * tid * M, M, 2, S);
    SUF(xim_omp)(data, K, kf, nk, M, S, N);
    if (S < 0) SUF(orc_backward)(data, N);
    return 0;
}

/* Out-of-place OpenMP path: InTimeOOP_omp (src/gfftomp.h:493-538).  Sub-transform tid reads
 * the decimated sequence src[Perm(tid) + K*n] where Perm uses the reversed factor list of K. */
int SUF(orc_gfft_oop_omp)(const T* src, T* dst, long N, int S, int nt)
{
    long f[64];
    int nf = orc_factorize(N, f);
    if (nf < 0) return -1;
    if (nt <= 1 || !(N > (long)nt * nt && N >= 256) || (N % nt)) return SUF(orc_gfft_oop)(src, dst, N, S);
    long kf[64], rkf[64];
    int nk = orc_factorize(nt, kf);
    if (nk < 0) return -1;
    for (int i = 0; i < nk; ++i) rkf[i] = kf[nk - 1 - i];
    const long K = nt, M = N / K;
    /* remaining factors of M = N/K (ascending) */
    long mf[64];
    int nm = orc_factorize(M, mf);
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (long tid = 0; tid < K; ++tid) {
        const long p = orc_perm(tid, rkf, nk, K);
        if (M == 1) {
            dst[2 * tid] = src[2 * p];
            dst[2 * tid + 1] = src[2 * p + 1];
        } else {
            SUF(intime_oop)(src + 2 * p, dst + 2 * tid * M, M, mf, nm, K, S);
        }
    }
    SUF(xim_omp)(dst, K, kf, nk, M, S, N);
    if (S < 0) SUF(orc_backward)(dst, N);
    return 0;
}
