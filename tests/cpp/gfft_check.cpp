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
#include <type_traits>
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

// The sample program's pattern with GFFT's seq input x_m = (2m, 2m+1) (src/direct.h:52-56),
// whose transform has the closed form X_0 = (2+2i) n(n-1)/2 + i n and
// X_k = (2+2i) n i e^{i pi k/n} / (2 sin(pi k/n)) for k > 0.
template <class ValueType, class Place, class Parall> static bool sample_program(const char* what, double bound)
{
    static const long_t N = 1024;
    typedef TYPELIST_1(ulong_<N>) NList;
    typedef GenerateTransform<NList, ValueType, TransformTypeGroup::FullList, ulong_<1>, Parall, Place> TransformSet;
    typedef DFT TransformType;
    typedef typename ValueType::ValueType T;
    typedef typename ValueType::base_type BT;
    static const int C = Loki::TypeTraits<T>::isStdFundamental ? 2 : 1;

    TransformSet gfft;
    typename TransformSet::ObjectType* fftobj =
        gfft.CreateTransformObject(N, ValueType::ID, TransformType::ID, 1, Parall::ID, Place::ID);
    std::vector<T> data(C * N), dataout(C * N);
    BT* d = reinterpret_cast<BT*>(data.data());
    for (long_t m = 0; m < N; ++m) {
        d[2 * m] = (BT)(2 * m);
        d[2 * m + 1] = (BT)(2 * m + 1);
    }
    BT* out = d;
    if constexpr (std::is_same<Place, OUT_OF_PLACE>::value) {
        fftobj->fft(data.data(), dataout.data());
        out = reinterpret_cast<BT*>(dataout.data());
    } else {
        fftobj->fft(data.data());
    }
    delete fftobj;
    const long double PI = 3.141592653589793238462643383279502884L;
    std::vector<double> ref(2 * N);
    ref[0] = 2.0 * N * (N - 1) / 2;
    ref[1] = 2.0 * N * (N - 1) / 2 + N;
    for (long_t k = 1; k < N; ++k) {
        const long double a = PI * k / N, den = 2 * sinl(a), r = (long double)N / den;
        // (2+2i) i e^{ia} n / (2 sin a) = r (2i - 2)(cos a + i sin a), r = n / (2 sin a)
        ref[2 * k] = (double)(r * (-2 * cosl(a) - 2 * sinl(a)));
        ref[2 * k + 1] = (double)(r * (2 * cosl(a) - 2 * sinl(a)));
    }
    return report(what, (long)N, out, ref, bound);
}

// The groups' lists, defaults and lengths as the reference states them
// (src/gfftparamgroups.h:35-78), and the IDs of the parallelisation policies
// (src/gfftpolicy.h:371-453, src/gfftgen.h:78).
static bool groups_match_reference()
{
    bool ok = true;
    ok &= std::is_same<ValueTypeGroup::Default, DOUBLE>::value && ValueTypeGroup::Length == 4;
    ok &= std::is_same<ValueTypeGroup::FullList::Tail::Tail::Tail::Head, COMPLEX_FLOAT>::value;
    ok &= std::is_same<TransformTypeGroup::Default, DFT>::value && TransformTypeGroup::Length == 4;
    ok &= std::is_same<TransformTypeGroup::FullList::Tail::Tail::Head, RDFT>::value;
    ok &= std::is_same<ParallelizationGroup::Default, Serial>::value && ParallelizationGroup::Length == 3;
    ok &= std::is_same<ParallelizationGroup::FullList::Tail::Tail::Tail::Head, OpenMP<8>>::value;
    ok &= std::is_same<PlaceGroup::Default, OUT_OF_PLACE>::value && PlaceGroup::Length == 2;
    ok &= Serial::ID == 0 && OpenMP<1>::ID == 0 && OpenMP<4>::ID == 3 && OpenMP<8>::NParProc == 8;
    typedef Transform<ulong_<256>, FLOAT, DFT, ulong_<1>, Serial, IN_PLACE> T256;
    typedef Transform<ulong_<256>, FLOAT, DFT, ulong_<1>, Serial, IN_PLACE, 77> T256id;
    ok &= (int)T256::ID == 255 && (int)T256id::ID == 77 && T256::Len == 256;
    typedef GenPowerList<3, 5, 2>::Result P;
    ok &= P::Head::value == 8 && P::Tail::Tail::Head::value == 32 && P::Head::ID == 7;
    typedef GenNumList<20, 22>::Result Q;
    ok &= Q::Tail::Head::value == 21;
    std::printf("parameter groups and IDs %s\n", ok ? "ok" : "FAIL");
    return ok;
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
    typedef GenerateTransform<GenPowerList<4, 10, 2>::Result, DOUBLE, TYPELIST_2(DFT, IDFT), ulong_<1>, HIP<0>,
                              OUT_OF_PLACE>
        Set;
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

    // GFFT's sample program (src/gfft.cpp:41-67) in the reference's own forms: a one-length
    // TYPELIST, the full transform-type group, the value type's ValueType/base_type and
    // Loki::TypeTraits, and CreateTransformObject with the policy's ID, for HIP<0> and for
    // an unchanged OpenMP<4> (which runs on GPU 0 here), in place and out of place.
    ok &= sample_program<DOUBLE, OUT_OF_PLACE, HIP<0>>("sample DOUBLE oop HIP<0>", 1e-13);
    ok &= sample_program<DOUBLE, IN_PLACE, OpenMP<4>>("sample DOUBLE in-place OpenMP<4>", 1e-13);
    ok &= sample_program<COMPLEX_FLOAT, OUT_OF_PLACE, Serial>("sample COMPLEX_FLOAT oop Serial", 2e-7);
    ok &= groups_match_reference();
    std::printf("%s\n", ok ? "ALL OK" : "FAILURES");
    return ok ? 0 : 1;
}
