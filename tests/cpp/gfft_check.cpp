// gfft_check.cpp -- accuracy check through GFFT's C++ surface with the HIP<0> policy.
//
// Counterpart of the reference's GFFTcheck (test/gfft_accuracy.h:148-277): for each
// transform of a generated set, draw input, call fft(), compare with a direct DFT and print
// ||err||_2, ||err||_inf and ||err||_inf/||y||_inf.  Unlike the reference it also fails
// (exit 1) when an error exceeds the stated bound: relL2 1e-13 (double), 2e-7 (float).
// Input: re,im ~ U[-1,1) from SplitMix64.  Reference: long-double DFT with exact index
// reduction.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "gfftx/gfft.hpp"

using namespace gfftx;

static uint64_t splitmix(uint64_t seed, uint64_t i)
{
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static double upm1(uint64_t seed, uint64_t i) { return 2.0 * (double)(splitmix(seed, i) >> 11) * 0x1p-53 - 1.0; }

static void dft_ref(const std::vector<double>& x, std::vector<double>& y, long n, int sign)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    std::vector<long double> c(n), s(n);
    for (long m = 0; m < n; ++m) {
        c[m] = cosl(2 * PI * m / n);
        s[m] = -sign * sinl(2 * PI * m / n);
    }
    y.assign(2 * n, 0.0);
    for (long k = 0; k < n; ++k) {
        long double re = 0, im = 0;
        long m = 0;
        for (long j = 0; j < n; ++j) {
            re += x[2 * j] * c[m] - x[2 * j + 1] * s[m];
            im += x[2 * j] * s[m] + x[2 * j + 1] * c[m];
            m += k;
            if (m >= n) m -= n;
        }
        const long double sc = sign < 0 ? 1.0L / n : 1.0L;
        y[2 * k] = (double)(re * sc);
        y[2 * k + 1] = (double)(im * sc);
    }
}

template <typename T> static bool report(const char* what, long n, const T* got, const std::vector<double>& ref,
                                         double bound)
{
    double e2 = 0, r2 = 0, einf = 0, rinf = 0;
    for (long i = 0; i < 2 * n; ++i) {
        const double d = (double)got[i] - ref[i];
        e2 += d * d;
        r2 += ref[i] * ref[i];
        einf = std::fmax(einf, std::fabs(d));
        rinf = std::fmax(rinf, std::fabs(ref[i]));
    }
    const double rel = std::sqrt(e2 / r2);
    std::printf("%-28s %7ld  relL2 %.3e  |err|2 %.3e  |err|inf %.3e  rel_inf %.3e  %s\n", what, n, rel,
                std::sqrt(e2), einf, einf / rinf, rel <= bound ? "ok" : "FAIL");
    return rel <= bound;
}

template <class VType, class Type, class Place, unsigned long N> static bool check_one(const char* what, double bound)
{
    typedef typename VType::base_type B;
    const int sign = Type::Sign;
    std::vector<double> x(2 * N), ref;
    for (unsigned long i = 0; i < 2 * N; ++i) x[i] = upm1(17 + N, i);
    std::vector<B> in(x.begin(), x.end()), out(2 * N);
    std::vector<double> xr(in.begin(), in.end());  // reference sees the rounded input
    dft_ref(xr, ref, (long)N, sign);
    typename Transform<ulong_<N>, VType, Type, ulong_<1>, HIP<0>, Place>::Instance t;
    if constexpr (std::is_same<Place, IN_PLACE>::value) {
        t.fft(reinterpret_cast<typename VType::ValueType*>(in.data()));
        return report(what, (long)N, in.data(), ref, bound);
    } else {
        t.fft(reinterpret_cast<const typename VType::ValueType*>(in.data()),
              reinterpret_cast<typename VType::ValueType*>(out.data()));
        return report(what, (long)N, out.data(), ref, bound);
    }
}

int main()
{
    bool ok = true;
    ok &= check_one<DOUBLE, DFT, IN_PLACE, 1024>("DOUBLE DFT in-place", 1e-13);
    ok &= check_one<DOUBLE, IDFT, IN_PLACE, 1024>("DOUBLE IDFT in-place", 1e-13);
    ok &= check_one<DOUBLE, DFT, OUT_OF_PLACE, 4096>("DOUBLE DFT out-of-place", 1e-13);
    ok &= check_one<COMPLEX_DOUBLE, DFT, OUT_OF_PLACE, 256>("COMPLEX_DOUBLE DFT oop", 1e-13);
    ok &= check_one<FLOAT, DFT, IN_PLACE, 2048>("FLOAT DFT in-place", 2e-7);
    ok &= check_one<COMPLEX_FLOAT, IDFT, OUT_OF_PLACE, 512>("COMPLEX_FLOAT IDFT oop", 2e-7);
    ok &= check_one<DOUBLE, DFT, OUT_OF_PLACE, 768>("DOUBLE DFT 3*2^8 oop", 1e-13);
    ok &= check_one<FLOAT, DFT, IN_PLACE, 640>("FLOAT DFT 5*2^7 in-place", 2e-7);
    // a prime length: GFFT's out-of-place path takes any N (chirp-z here)
    ok &= check_one<DOUBLE, DFT, OUT_OF_PLACE, 1009>("DOUBLE DFT prime oop", 1e-13);
    ok &= check_one<COMPLEX_FLOAT, IDFT, IN_PLACE, 1001>("COMPLEX_FLOAT IDFT 7*11*13", 5e-7);

    // Factory path: objects created by run-time IDs, as src/gfft.cpp:65-67 does.
    typedef GenerateTransform<GenPowerList<4, 10>, DOUBLE, List<DFT, IDFT>, ulong_<1>, HIP<0>, OUT_OF_PLACE> Set;
    Set gfft;
    for (unsigned long n : {16ul, 256ul, 1024ul}) {
        AbstractFFT_oop<double>* obj = gfft.CreateTransformObject(n, DOUBLE::ID, DFT::ID, 1, HIP<0>::ID,
                                                                  OUT_OF_PLACE::ID);
        std::vector<double> x(2 * n), y(2 * n), ref;
        for (unsigned long i = 0; i < 2 * n; ++i) x[i] = upm1(99, i);
        obj->fft(x.data(), y.data());
        dft_ref(x, ref, (long)n, 1);
        ok &= report("factory DOUBLE DFT oop", (long)n, y.data(), ref, 1e-13);
        delete obj;
    }
    bool threw = false;
    try {
        gfft.CreateTransformObject(100, DOUBLE::ID, DFT::ID, 1, HIP<0>::ID, OUT_OF_PLACE::ID);
    } catch (const TransformFactoryError::Exception& e) {
        threw = true;
        std::printf("unknown transform -> %s\n", e.what());
    }
    ok &= threw;
    std::printf("%s\n", ok ? "ALL OK" : "FAILURES");
    return ok ? 0 : 1;
}
