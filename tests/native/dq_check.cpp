// dq_check.cpp — host build of dq_dropped (falx_amd/csrc/cells.h), the
// doubled-quote drops the gfx950 string passes read from a tile's quote
// bitmap, swept against DoubledStream (the byte-at-a-time cleaning stream
// the same passes use otherwise) on random insides of quoted cells: quote
// runs of every length, at every alignment to the 32-byte bitmap words and
// across them. The kept bytes, in order, must be the stream's output.
//
// usage: dq_check <seed> <count>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../falx_amd/csrc/cells.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
    uint64_t below(uint64_t n) { return next() % n; }
};

struct BufSrc {
    const uint8_t* p;
    uint8_t at(uint32_t i) const { return p[i]; }
};

}  // namespace

int main(int argc, char** argv) {
    Rng r{argc > 1 ? strtoull(argv[1], nullptr, 0) | 1 : 1};
    const long count = argc > 2 ? atol(argv[2]) : 200000;
    static uint8_t buf[512];
    uint32_t qb[16];
    long bad = 0;
    for (long it = 0; it < count; ++it) {
        // A quoted cell at a random offset: quote, inside, quote.
        const uint32_t off = (uint32_t)r.below(64);
        const uint32_t n = (uint32_t)r.below(160);
        const uint32_t dens = 1 + (uint32_t)r.below(6);   // quote density
        std::memset(buf, 'x', sizeof buf);
        buf[off] = '"';
        for (uint32_t i = 0; i < n; ++i) buf[off + 1 + i] = r.below(8) < dens ? '"' : (uint8_t)('a' + r.below(26));
        buf[off + 1 + n] = '"';
        // The bitmap of quote bytes, as the tile's classification writes it.
        std::memset(qb, 0, sizeof qb);
        for (uint32_t p = 0; p < 512; ++p)
            if (buf[p] == '"') qb[p >> 5] |= 1u << (p & 31);
        // The stream's output over the inside.
        std::string want;
        {
            fx::DoubledStream<BufSrc> st{nullptr, off + 1, off + 1 + n, (uint8_t)'"'};
            BufSrc src{buf};
            st.s = &src;
            for (int c = st.next(); c >= 0; c = st.next()) want.push_back((char)c);
        }
        // The kept bytes by the bitmap, word by word, and their count.
        std::string got;
        uint32_t drops = 0;
        if (n > 0) {
            const uint32_t a = off + 1, z = off + n;   // inside, inclusive
            uint32_t in_run = 0, odd_next = 0;
            for (uint32_t k = a >> 5; k <= (z >> 5); ++k) {
                uint32_t wv = qb[k];
                if (k == (a >> 5)) wv &= ~0u << (a & 31u);
                if (k == (z >> 5)) wv &= (z & 31u) == 31u ? ~0u : (2u << (z & 31u)) - 1u;
                const uint32_t d = fx::dq_dropped(wv, in_run, odd_next);
                drops += (uint32_t)__builtin_popcount(d);
                const uint32_t p0 = a > (k << 5) ? a : (k << 5), p1 = z < (k << 5) + 31u ? z : (k << 5) + 31u;
                for (uint32_t p = p0; p <= p1; ++p)
                    if (!((d >> (p & 31u)) & 1u)) got.push_back((char)buf[p]);
            }
        }
        if (got != want || n - drops != want.size()) {
            if (bad < 5)
                std::printf("off %u inside [%.*s]: stream [%s] bitmap [%s] drops %u\n", off, (int)n,
                            (const char*)buf + off + 1, want.c_str(), got.c_str(), drops);
            ++bad;
        }
    }
    if (bad) {
        std::printf("%ld mismatches\n", bad);
        return 1;
    }
    std::printf("%ld quoted cells agree with the doubled-quote stream\n", count);
    return 0;
}
