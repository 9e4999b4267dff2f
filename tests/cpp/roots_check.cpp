// roots_check.cpp -- prints the engine's roots of unity for (m, N) pairs read from stdin
// ("m N" per line) as "m N cos sin" with 21 significant digits, for the CPU test that
// compares them with 40-digit values.
#include <cstdio>

#include "../../gfft_amd/csrc/roots.h"

int main()
{
    long long m, n;
    while (std::scanf("%lld %lld", &m, &n) == 2) {
        long double c, s;
        gfftx::detail::root_ld_eval(m, n, &c, &s);
        std::printf("%lld %lld %.21Le %.21Le\n", m, n, c, s);
    }
    return 0;
}
