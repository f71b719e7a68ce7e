// cells_check.cpp — host build of falx_amd/csrc/cells.h (the parsers the
// gfx950 kernels run) swept against the CPU oracle on randomized cells.
// Exit status 0 = every cell agreed bit-for-bit.
//
// usage: cells_check <seed> <count>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../falx_amd/csrc/cells.h"
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

std::string random_number(Rng& r) {
    std::string s;
    uint64_t k = r.below(100);
    if (r.below(3) == 0) s.push_back(r.below(4) == 0 ? '+' : '-');
    if (k < 25) {  // fixed-6 shapes and near misses
        digits(r, s, 1 + (int)r.below(4));
        s.push_back('.');
        digits(r, s, 5 + (int)r.below(3));
    } else if (k < 40) {  // Clinger range
        digits(r, s, 1 + (int)r.below(10));
        if (r.below(2)) { s.push_back('.'); digits(r, s, (int)r.below(10)); }
        if (r.below(3) == 0) { s.push_back('e'); if (r.below(2)) s.push_back('-'); s += std::to_string(r.below(30)); }
    } else if (k < 60) {  // long mantissas, big exponents
        digits(r, s, 1 + (int)r.below(40));
        if (r.below(2)) { s.push_back('.'); digits(r, s, (int)r.below(40)); }
        if (r.below(2)) {
            s.push_back(r.below(2) ? 'e' : 'E');
            uint64_t es = r.below(3);
            if (es == 1) s.push_back('-'); else if (es == 2) s.push_back('+');
            s += std::to_string(r.below(400));
        }
    } else if (k < 70) {  // near subnormal / overflow boundaries
        digits(r, s, 1 + (int)r.below(20));
        s.push_back('e');
        if (r.below(2)) { s.push_back('-'); s += std::to_string(300 + r.below(40)); }
        else s += std::to_string(290 + r.below(30));
    } else if (k < 75) {  // specials and spellings
        const char* sp[] = {"inf", "Inf", "INFINITY", "infinity", "nan", "NaN", "infinit", "in",
                            "nana", "infinityy", "e5", ".", "..", "1e", "1e+", "-", "+"};
        s += sp[r.below(sizeof sp / sizeof sp[0])];
    } else if (k < 85) {  // leading-dot / trailing-dot / zeros
        if (r.below(2)) digits(r, s, (int)r.below(5));
        s.push_back('.');
        if (r.below(2)) digits(r, s, (int)r.below(25));
        if (r.below(3) == 0) { s.push_back('e'); s += std::to_string(r.below(30)); }
    } else if (k < 92) {  // halfway-ish: 2^-k style long expansions
        s += "0.";
        for (int i = 0; i < 5 + (int)r.below(30); ++i) s.push_back('0');
        digits(r, s, 15 + (int)r.below(30));
        if (r.below(2)) s.push_back('5');
    } else {  // garbage
        int n = (int)r.below(12);
        const char alpha[] = "0123456789.eE+-_ xa,";
        for (int i = 0; i < n; ++i) s.push_back(alpha[r.below(sizeof alpha - 1)]);
    }
    return s;
}

std::string random_int(Rng& r) {
    std::string s;
    if (r.below(3) == 0) s.push_back(r.below(3) == 0 ? '+' : '-');
    uint64_t k = r.below(10);
    if (k < 6) digits(r, s, 1 + (int)r.below(22));
    else if (k < 8) s += r.below(2) ? "9223372036854775807" : "9223372036854775808";
    else if (k < 9) s += "0000000000000000000000012";
    else { digits(r, s, (int)r.below(4)); s.push_back("x. _"[r.below(4)]); }
    return s;
}

int fail(const char* what, const std::string& cell) {
    std::fprintf(stderr, "MISMATCH %s on cell [%s]\n", what, cell.c_str());
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 0) : 1;
    uint64_t count = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 200000;
    Rng r{seed | 1};
    // Fixed known-answer cells first (columns.rs:408-451 hand cases and IEEE edges).
    const char* fixed[] = {"1.5", "-0.0", "0.1", "0.0", "1_000", "1.5e308", "5e-324", "NaN", "inf",
                           "-inf", "", "2.2250738585072011e-308", "2.2250738585072014e-308",
                           "4.9406564584124654e-324", "2.4703282292062327e-324",
                           "2.4703282292062328e-324", "1.7976931348623157e308",
                           "1.7976931348623158e308", "1.7976931348623159e308",
                           "9007199254740993", "9007199254740992.5", "0.30000000000000004",
                           "-12.345678", "+12.345678", "999.999999", "-0.000000",
                           "123456789012345678901234567890", "1e23", "8.98846567431158e307"};
    uint64_t checked = 0;
    falx_spec dq{};
    dq.quote = '"';
    dq.escape = -1;
    dq.comment = -1;
    falx_spec bs = dq;
    bs.escape = '\\';
    for (uint64_t i = 0; i < count + sizeof fixed / sizeof fixed[0]; ++i) {
        std::string cell = i < sizeof fixed / sizeof fixed[0] ? std::string(fixed[i]) : random_number(r);
        const uint8_t* p = (const uint8_t*)cell.data();
        BufSrc src{p};
        double ov;
        int ofx;
        int64_t om;
        int ook = ora_parse_f64(p, cell.size(), &ov, &ofx, &om);
        fx::F64Result g = fx::parse_f64_stream(fx::RawStream<BufSrc>{&src, 0, (uint32_t)cell.size()});
        if ((int)g.ok != ook) return fail("f64 ok", cell);
        if (ook) {
            if (fx::f64_to_bits(g.value) != fx::f64_to_bits(ov)) return fail("f64 bits", cell);
            if ((int)g.fixed6 != ofx) return fail("f64 fixed6", cell);
            if (ofx && g.mantissa != om) return fail("f64 mantissa", cell);
        }
        // Quoted forms through both cleaning dialects.
        std::string quoted = "\"" + cell + "\"";
        const uint8_t* qp = (const uint8_t*)quoted.data();
        BufSrc qs{qp};
        for (int dial = 0; dial < 2; ++dial) {
            fx::CellDialect cd{'"', dial ? '\\' : -1};
            fx::F64Result gq = fx::parse_f64_field(qs, 0, (uint32_t)quoted.size(), cd);
            if ((int)gq.ok != ook) return fail("quoted f64 ok", quoted);
            if (ook && fx::f64_to_bits(gq.value) != fx::f64_to_bits(ov)) return fail("quoted f64", quoted);
        }
        ++checked;
    }
    for (uint64_t i = 0; i < count; ++i) {
        std::string cell = random_int(r);
        const uint8_t* p = (const uint8_t*)cell.data();
        BufSrc src{p};
        int64_t ov = 0, gv = 0;
        int ook = ora_parse_i64(p, cell.size(), &ov);
        bool gok = fx::parse_i64_stream(fx::RawStream<BufSrc>{&src, 0, (uint32_t)cell.size()}, gv);
        if ((int)gok != ook) return fail("i64 ok", cell);
        if (ook && gv != ov) return fail("i64 value", cell);
        ++checked;
    }
    // Cleaning: random quote/escape soup.
    for (uint64_t i = 0; i < count; ++i) {
        std::string cell;
        int n = (int)r.below(12);
        const char alpha[] = "\"\"\\ab,";
        for (int k = 0; k < n; ++k) cell.push_back(alpha[r.below(sizeof alpha - 1)]);
        const uint8_t* p = (const uint8_t*)cell.data();
        BufSrc src{p};
        for (int dial = 0; dial < 2; ++dial) {
            uint8_t out[64];
            uint64_t on = ora_clean(dial ? &bs : &dq, p, cell.size(), out);
            std::string got;
            fx::CellDialect cd{'"', dial ? '\\' : -1};
            uint32_t gn = fx::clean_cell(src, 0, (uint32_t)cell.size(), cd,
                                         [&](uint8_t b) { got.push_back((char)b); });
            if (gn != on || got.size() != on || std::memcmp(got.data(), out, on) != 0)
                return fail(dial ? "clean backslash" : "clean doubled", cell);
        }
        ++checked;
    }
    std::printf("cells_check: %" PRIu64 " cells agree\n", checked);
    return 0;
}
