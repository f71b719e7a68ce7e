// swar_check.cpp — host build of falx_amd/csrc/num_swar.h (the register
// parser the gfx950 column kernels try first) swept against the CPU oracle's
// cell parsers. Every cell it decides must agree bit-for-bit; the cells it
// leaves undecided go to the stream parser, which cells_check covers. The
// bytes below a cell in its window are random digits and grammar bytes, so a
// masking mistake shows up as a wrong value.
//
// usage: swar_check <seed> <count>
#include <cinttypes>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <string>

#include "../../falx_amd/csrc/num_swar.h"
#include "../../oracle/oracle.h"

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

void digits(Rng& r, std::string& s, int n) {
    for (int i = 0; i < n; ++i) s.push_back((char)('0' + r.below(10)));
}

// Numeric-looking cells of 0..28 bytes, weighted to the shapes the register
// parser decides and to the edges of its limits.
std::string cell(Rng& r) {
    std::string s;
    const uint64_t k = r.below(100);
    if (k < 90 && r.below(3) == 0) s.push_back(r.below(4) == 0 ? '+' : '-');
    if (k < 30) {  // d.ddd shapes, fixed-6 and near misses
        digits(r, s, (int)r.below(6));
        s.push_back('.');
        digits(r, s, r.below(2) ? 6 : (int)r.below(12));
    } else if (k < 45) {  // integers up to and past 19 digits, leading zeros
        if (r.below(4) == 0) s += std::string(1 + r.below(4), '0');
        digits(r, s, 1 + (int)r.below(21));
    } else if (k < 50) {  // the i64 limits
        s += r.below(2) ? "9223372036854775807" : "9223372036854775808";
    } else if (k < 75) {  // exponents of 0..6 digits
        digits(r, s, (int)r.below(10));
        if (r.below(2)) { s.push_back('.'); digits(r, s, (int)r.below(10)); }
        s.push_back(r.below(2) ? 'e' : 'E');
        const uint64_t es = r.below(3);
        if (es == 1) s.push_back('-'); else if (es == 2) s.push_back('+');
        digits(r, s, (int)r.below(7));
    } else if (k < 80) {  // 19-digit significands around the double range
        digits(r, s, 1 + (int)r.below(19));
        s.push_back('e');
        if (r.below(2)) s.push_back('-');
        s += std::to_string(r.below(340));
    } else if (k < 85) {  // specials
        const char* sp[] = {"inf", "-inf", "NaN", "", ".", "e5", "1e", "1e+", "-", "+", "0", "-0",
                            "0.0", "-0.000000", "1.", ".5", "+.5e-3", "\"12\"", "\"1.5\"", "1\"2"};
        s += sp[r.below(sizeof sp / sizeof sp[0])];
    } else {  // garbage over the grammar's bytes
        const char alpha[] = "0123456789.eE+-_ x\"";
        for (int i = 0, n = (int)r.below(26); i < n; ++i) s.push_back(alpha[r.below(sizeof alpha - 1)]);
    }
    return s;
}

// The window of a cell: 4 NW bytes ending at its last byte, random bytes below.
template <int NW>
fx::Win<NW> window(Rng& r, const std::string& c, uint8_t* b) {
    constexpr size_t B = 4 * NW;
    const char noise[] = "0123456789.eE+-,\"";
    for (size_t i = 0; i < B; ++i) b[i] = (uint8_t)noise[r.below(sizeof noise - 1)];
    const size_t n = c.size() < B ? c.size() : B;
    for (size_t i = 0; i < n; ++i) b[B - n + i] = (uint8_t)c[c.size() - n + i];
    fx::Win<NW> W;
    for (int k = 0; k < NW; ++k)
        W.w[k] = b[4 * k] | (uint32_t)b[4 * k + 1] << 8 | (uint32_t)b[4 * k + 2] << 16 | (uint32_t)b[4 * k + 3] << 24;
    return W;
}

int fail(const char* what, const std::string& c) {
    std::fprintf(stderr, "MISMATCH (%s) on cell [%s]\n", what, c.c_str());
    return 1;
}

}  // namespace

template <int NW>
int sweep(uint64_t seed, uint64_t count, uint64_t& f_dec, uint64_t& f_ok, uint64_t& i_dec, uint64_t& i_ok,
          uint64_t& f6_dec) {
    constexpr uint32_t B = 4 * NW;
    Rng r{seed | 1};
    static fx::BigDec big;
    for (uint64_t i = 0; i < count; ++i) {
        const std::string c = cell(r);
        const uint8_t* p = (const uint8_t*)c.data();
        const uint32_t n = (uint32_t)c.size();
        const int32_t quote = r.below(2) ? '"' : -1;
        uint8_t wb[B];
        const fx::Win<NW> W = window<NW>(r, c, wb);
        auto at = [&](uint32_t p) { return (uint32_t)wb[p]; };
        fx::F64Result fr;
        bool iok;
        int64_t iv;
        // f64
        if (fx::num_swar(W, n, true, quote, at, fr, iok, iv)) {
            ++f_dec;
            double ov;
            int ofx;
            int64_t om;
            const int ook = ora_parse_f64(p, n, &ov, &ofx, &om);
            if ((int)fr.ok != ook) return fail("f64 ok", c);
            if (ook) {
                ++f_ok;
                if (fr.deferred) {
                    BufSrc src{p};
                    fr.value = fx::bits_to_f64(fx::f64_field_exact_bits(src, 0, n, fx::CellDialect{-1, -1}, big));
                }
                if (fx::f64_to_bits(fr.value) != fx::f64_to_bits(ov)) return fail("f64 bits", c);
                if ((int)fr.fixed6 != ofx) return fail("f64 fixed6", c);
                if (ofx && fr.mantissa != om) return fail("f64 mantissa", c);
            }
        } else if (n <= B && n > 0 && c[0] != '"') {
            // Undecided cells must be ones the grammar check cannot take.
            const std::string allowed = "0123456789+-.eE";
            bool plain = c.find_first_not_of(allowed) == std::string::npos;
            size_t nd = 0;
            for (char ch : c) nd += ch >= '0' && ch <= '9';
            if (plain && nd <= 19 && std::count_if(c.begin(), c.end(), [](char ch) { return ch < 48 || ch > 57; }) <= 4 && c.find_first_of("eE") == std::string::npos &&
                std::count(c.begin(), c.end(), '.') <= 1 && c.find_first_of("+-", 1) == std::string::npos)
                return fail("f64 undecided", c);
        }
        // The fixed-6 shape from the last 12 bytes (fixed6_win): what it
        // decides agrees with the oracle, and it decides every fixed-6 cell.
        if (NW == 3) {
            fx::F64Result f6{};
            double ov;
            int ofx;
            int64_t om;
            const int ook = ora_parse_f64(p, n, &ov, &ofx, &om);
            if (fx::fixed6_win(W.w[0], W.w[1], W.w[2], n, f6)) {
                ++f6_dec;
                if (!ook || !ofx) return fail("fixed6_win decided a cell of another shape", c);
                if (fx::f64_to_bits(f6.value) != fx::f64_to_bits(ov) || !f6.fixed6 || f6.mantissa != om)
                    return fail("fixed6_win value", c);
            } else if (ook && ofx && n <= 11) {
                return fail("fixed6_win missed a fixed-6 cell", c);
            }
        }
        // i64
        if (fx::num_swar(W, n, false, quote, at, fr, iok, iv)) {
            ++i_dec;
            int64_t ov = 0;
            const int ook = ora_parse_i64(p, n, &ov);
            if ((int)iok != ook) return fail("i64 ok", c);
            if (ook) {
                ++i_ok;
                if (iv != ov) return fail("i64 value", c);
            }
        }
    }
    return 0;
}

int main(int argc, char** argv) {
    const uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 0) : 1;
    const uint64_t count = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 200000;
    uint64_t f_dec = 0, i_dec = 0, f_ok = 0, i_ok = 0, f6_dec = 0;
    if (sweep<6>(seed, count, f_dec, f_ok, i_dec, i_ok, f6_dec)) return 1;
    if (sweep<3>(seed + 1, count, f_dec, f_ok, i_dec, i_ok, f6_dec)) return 1;
    std::printf("swar_check: %" PRIu64 " cells agree (f64 decided %" PRIu64 ", %" PRIu64
                " valid; i64 decided %" PRIu64 ", %" PRIu64 " valid; fixed-6 windows decided %" PRIu64 ")\n",
                2 * count, f_dec, f_ok, i_dec, i_ok, f6_dec);
    return 0;
}
