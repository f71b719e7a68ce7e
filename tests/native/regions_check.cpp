// regions_check.cpp — host build of falx_amd/csrc/regions.h (the quoted and
// comment regions of a 64-byte block, as the gfx950 scan kernels compute
// them) swept against a byte-at-a-time automaton over the same block:
// outside, a quote opens a quoted region and a comment byte a comment;
// inside quotes a quote closes it; inside a comment a newline ends it. A
// region covers its opening byte and not its closing one. Blocks are random
// with varying density, for every entry state.
//
// usage: regions_check <seed> <count>
#include <cstdint>
#include <cstdio>
#include <cstdlib>

#include "../../falx_amd/csrc/regions.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
};

uint64_t automaton(uint64_t q, uint64_t s, uint64_t n, uint32_t& state) {
    uint64_t inert = 0;
    for (int i = 0; i < 64; ++i) {
        const uint64_t bit = 1ULL << i;
        bool in = false;
        if (state == 0) {
            if (q & bit) { state = 1; in = true; }
            else if (s & bit) { state = 2; in = true; }
        } else if (state == 1) {
            if (q & bit) state = 0;
            else in = true;
        } else {
            if (n & bit) state = 0;
            else in = true;
        }
        if (in) inert |= bit;
    }
    return inert;
}

}  // namespace

int main(int argc, char** argv) {
    Rng g{argc > 1 ? strtoull(argv[1], nullptr, 0) | 1 : 1};
    const long count = argc > 2 ? atol(argv[2]) : 1000000;
    long bad = 0;
    for (long it = 0; it < count; ++it) {
        const uint32_t span = 4 + (uint32_t)(g.next() % 40);   // lower = denser
        uint64_t q = 0, s = 0, n = 0;
        for (int b = 0; b < 64; ++b) {
            const uint32_t r = (uint32_t)(g.next() % span);
            if (r == 0) q |= 1ULL << b;
            else if (r == 1) s |= 1ULL << b;
            else if (r == 2) n |= 1ULL << b;
        }
        if (it & 1) s &= (n << 1) | 1ULL;   // comment bytes at line starts only
        for (uint32_t st = 0; st < 3; ++st) {
            uint32_t a = st, b = st;
            const uint64_t want = automaton(q, s, n, a), got = fx::resolve_regions(q, s, n, b);
            if (want != got || a != b) {
                if (bad < 5)
                    printf("q=%016llx s=%016llx n=%016llx entry %u: want %016llx/%u got %016llx/%u\n",
                           (unsigned long long)q, (unsigned long long)s, (unsigned long long)n, st,
                           (unsigned long long)want, a, (unsigned long long)got, b);
                ++bad;
            }
        }
    }
    if (bad) {
        printf("%ld mismatches\n", bad);
        return 1;
    }
    printf("%ld blocks x 3 entry states agree\n", count);
    return 0;
}
