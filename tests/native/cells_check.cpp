// cells_check.cpp — host build of falx_amd/csrc/cells.h (the parsers the
// gfx950 kernels run) swept against the CPU oracle on randomized cells.
// Exit status 0 = every cell agreed bit-for-bit.
//
// usage: cells_check <seed> <count> [near-halfway count]
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

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

// Exact decimal digits of m * 2^k as (digit string D, exponent x) with
// value = D * 10^x.
std::pair<std::string, int> exact_decimal(uint64_t m, int k) {
    std::vector<int> d;  // little-endian base-10
    for (uint64_t v = m; v; v /= 10) d.push_back((int)(v % 10));
    int mul = k >= 0 ? 2 : 5, reps = k >= 0 ? k : -k;
    for (int i = 0; i < reps; ++i) {
        int carry = 0;
        for (int& x : d) {
            int y = x * mul + carry;
            x = y % 10;
            carry = y / 10;
        }
        while (carry) { d.push_back(carry % 10); carry /= 10; }
    }
    std::string D;
    for (size_t i = d.size(); i-- > 0;) D.push_back((char)('0' + d[i]));
    return {D, k >= 0 ? 0 : k};
}

std::string with_exp(const std::string& D, int x) { return D + "e" + std::to_string(x); }

// Cells at, just above, just below and truncated near the midpoint between
// a random double and its successor (normal and subnormal).
std::string halfway_cell(Rng& r) {
    uint64_t M, E;
    int kind = (int)r.below(4);
    if (kind == 0) {  // subnormal
        M = r.next() & ((1ULL << 52) - 1);
        E = 0;
    } else {
        M = (1ULL << 52) | (r.next() & ((1ULL << 52) - 1));
        E = kind == 1 ? r.below(2046) + 1 : 1023 + r.below(140) - 70;
    }
    int e2 = (E == 0 ? -1074 : (int)E - 1075);  // value = M * 2^e2
    auto [D, x] = exact_decimal(2 * M + 1, e2 - 1);
    std::string s;
    switch (r.below(5)) {
        case 0: s = with_exp(D, x); break;
        case 1: s = with_exp(D + "0001", x - 4); break;
        case 2: {  // just below: D*10 - 1
            std::string B = D + "0";
            int i = (int)B.size() - 1;
            while (B[i] == '0') { B[i] = '9'; --i; }
            B[i] -= 1;
            s = with_exp(B, x - 1);
            break;
        }
        case 3: {  // truncated to 17..40 significant digits
            size_t n = 17 + r.below(24);
            if (n < D.size()) s = with_exp(D.substr(0, n), x + (int)(D.size() - n));
            else s = with_exp(D, x);
            break;
        }
        default: {  // same value in positional form when short enough
            if (x == 0) s = D;
            else if ((int)D.size() + x > 0) s = D.substr(0, D.size() + x) + "." + D.substr(D.size() + x);
            else s = "0." + std::string((size_t)(-x - (int)D.size()), '0') + D;
        }
    }
    if (r.below(2)) s = "-" + s;
    return s;
}

// Fixed-6 shapes with one random mutation (sign, digit count, separator,
// stray byte) to probe the edges of the fast path.
std::string near_fixed6(Rng& r) {
    std::string s;
    if (r.below(3) == 0) s.push_back("-+"[r.below(2)]);
    digits(r, s, 1 + (int)r.below(4));
    s.push_back('.');
    digits(r, s, 5 + (int)r.below(3));
    if (r.below(2)) {
        const char alpha[] = "0123456789.-+eE x/:";
        s[r.below(s.size())] = alpha[r.below(sizeof alpha - 1)];
    }
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
    uint64_t n_halfway = argc > 3 ? std::strtoull(argv[3], nullptr, 0) : count / 10;
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
    uint64_t checked = 0, deferred_plain = 0, deferred_halfway = 0, plain = 0, halfway = 0;
    static fx::BigDec big;
    falx_spec dq{};
    dq.quote = '"';
    dq.escape = -1;
    dq.comment = -1;
    falx_spec bs = dq;
    bs.escape = '\\';
    const uint64_t n_fixed = sizeof fixed / sizeof fixed[0];
    for (uint64_t i = 0; i < count + n_halfway + n_fixed; ++i) {
        const bool hw = i >= count + n_fixed;
        std::string cell = i < n_fixed ? std::string(fixed[i]) : hw ? halfway_cell(r) : random_number(r);
        const uint8_t* p = (const uint8_t*)cell.data();
        BufSrc src{p};
        double ov;
        int ofx;
        int64_t om;
        int ook = ora_parse_f64(p, cell.size(), &ov, &ofx, &om);
        fx::F64Result fast = fx::parse_f64_stream(fx::RawStream<BufSrc>{&src, 0, (uint32_t)cell.size()});
        if (fast.ok && fast.deferred) ++(hw ? deferred_halfway : deferred_plain);
        ++(hw ? halfway : plain);
        fx::F64Result g = fx::parse_f64_field_exact(src, 0, (uint32_t)cell.size(), fx::CellDialect{-1, -1}, big);
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
            fx::F64Result gq = fx::parse_f64_field_exact(qs, 0, (uint32_t)quoted.size(), cd, big);
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
    // The fixed-6 tail fast path accepts exactly the oracle's fixed-6 cells.
    uint64_t fx_checked = 0;
    for (uint64_t i = 0; i < count; ++i) {
        std::string cell = (i & 1) ? near_fixed6(r) : random_number(r);
        if (cell.size() < 8 || cell.size() > 11) continue;
        const uint8_t* p = (const uint8_t*)cell.data();
        uint64_t x = 0;
        std::memcpy(&x, p + cell.size() - 8, 8);
        fx::F64Result g{};
        bool got = fx::fixed6_tail(x, p, (uint32_t)cell.size() - 8, g);
        double ov;
        int ofx;
        int64_t om;
        int ook = ora_parse_f64(p, cell.size(), &ov, &ofx, &om);
        if (got != (ook && ofx)) return fail("fixed6 tail shape", cell);
        if (got && (fx::f64_to_bits(g.value) != fx::f64_to_bits(ov) || g.mantissa != om))
            return fail("fixed6 tail value", cell);
        ++fx_checked;
    }
    checked += fx_checked;
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
    std::printf("cells_check: %" PRIu64 " cells agree; deferred to the exact path: %" PRIu64 " of %" PRIu64
                " ordinary f64 cells, %" PRIu64 " of %" PRIu64 " near-halfway cells\n",
                checked, deferred_plain, plain, deferred_halfway, halfway);
    return 0;
}
