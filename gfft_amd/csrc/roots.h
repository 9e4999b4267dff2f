// roots.h -- the roots of unity every twiddle table of the engine is built from.
//
// Host-only C++ (no HIP): planner.cpp evaluates the stage, pass, chirp and split tables with
// it, and a CPU test checks it against 40-digit values.
#pragma once
#include <cmath>

namespace gfftx {
namespace detail {

// exp(-2 pi i m / N) in long double, m reduced mod N and folded into the first octant.
inline void root_ld_eval(long long m, long long N, long double* c, long double* s)
{
    const long double PI = 3.141592653589793238462643383279502884L;
    m %= N;
    if (m < 0) m += N;
    // angle = (pi/4) * (8m/N); octant o and remainder in units of N/8
    const long double x8 = 8.0L * (long double)m;
    long long o = (long long)(x8 / (long double)N);
    if (o > 7) o = 7;
    const long double r = x8 - (long double)o * (long double)N;
    long double cc, ss;
    const bool odd = (o & 1) != 0;
    const long double a = (PI / 4.0L) * (odd ? ((long double)N - r) : r) / (long double)N;
    const long double ca = cosl(a), sa = sinl(a);
    // cos/sin of the full angle from the folded one
    switch (o) {
        case 0: cc = ca; ss = sa; break;
        case 1: cc = sa; ss = ca; break;
        case 2: cc = -sa; ss = ca; break;
        case 3: cc = -ca; ss = sa; break;
        case 4: cc = -ca; ss = -sa; break;
        case 5: cc = -sa; ss = -ca; break;
        case 6: cc = sa; ss = -ca; break;
        default: cc = ca; ss = -sa; break;
    }
    *c = cc;
    *s = -ss;
}

}  // namespace detail
}  // namespace gfftx
