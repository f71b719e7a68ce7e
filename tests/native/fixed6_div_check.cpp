// fixed6_div_check.cpp — div_1e6 (falx_amd/csrc/num_swar.h), the divide-free
// m / 10^6 of the fixed-6 cells, against the IEEE quotient for every
// m < 10^9: the whole domain of `[-]d{1,3}.d{6}` mantissas.
#include <cstdio>

#include "../../falx_amd/csrc/num_swar.h"

int main() {
    unsigned long long bad = 0;
    for (uint32_t m = 0; m < 1000000000u; ++m) {
        const double want = (double)m / 1e6;
        if (fx::f64_to_bits(fx::div_1e6(m)) != fx::f64_to_bits(want)) {
            if (bad < 5) std::printf("m=%u: %.17g, IEEE %.17g\n", m, fx::div_1e6(m), want);
            ++bad;
        }
    }
    if (bad) {
        std::printf("fixed6_div_check: %llu mismatches\n", bad);
        return 1;
    }
    std::printf("fixed6_div_check: all 1000000000 quotients agree\n");
    return 0;
}
