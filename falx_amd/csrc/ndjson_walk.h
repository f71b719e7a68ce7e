// ndjson_walk.h — per-record walker of the NDJSON `sum(id + nested.score)`
// sink, shared by the gfx950 kernel and its host build (for CPU sweeps).
//
// The reference's sink (ndjson.rs:679-723) splits records at raw '\n',
// reads a leading `{"id":` integer and a `],"nested":{"score":` integer, and
// panics on any other shape. This walker keeps the record split and the
// wrapping i64 id sum, and reads any JSON object layout: top-level key `id`
// (integer `-?[0-9]+`) and the `score` key of the top-level `nested` object
// (a JSON number, correctly rounded to f64), with whitespace anywhere JSON
// allows it and backslash escapes inside strings. The first occurrence of a
// key wins; keys are compared as raw bytes.
#pragma once

#include "cells.h"

namespace fx {

FX_HD bool jws(int b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }

template <class Src>
struct JCur {
    const Src* s;
    uint32_t i, n;
    FX_HD int peek() const { return i < n ? (int)s->at(i) : -1; }
    FX_HD void ws() { while (i < n && jws(s->at(i))) ++i; }
    FX_HD bool delim() const {
        int c = peek();
        return c < 0 || c == ',' || c == '}' || c == ']' || jws(c);
    }
};

// Bytes of x equal to zero, as the top bit of each byte (exact: no borrow
// runs from one byte into the next).
FX_HD uint64_t zero_bytes64(uint64_t x) {
    return ~(((x & 0x7F7F7F7F7F7F7F7FULL) + 0x7F7F7F7F7F7F7F7FULL) | x | 0x7F7F7F7F7F7F7F7FULL);
}

// A source whose bytes can also be read as aligned 8-byte words (word8(p)
// for p a multiple of 8, bytes past the record readable) lets strings be
// skipped a word at a time.
template <class Src, class = void>
struct has_word8 { static constexpr bool value = false; };
template <class Src>
struct has_word8<Src, decltype((void)Src::kWord8)> { static constexpr bool value = Src::kWord8; };

template <class Src>
FX_HD bool j_skip_string(JCur<Src>& c) {  // at the opening quote
    ++c.i;
    if constexpr (has_word8<Src>::value) {
        // The byte loop below, eight bytes per step: plain bytes are passed
        // over, a backslash skips the byte after it, a quote ends the string.
        uint32_t i = c.i;
        while (i < c.n) {
            const uint32_t a = i & ~7u;
            const uint64_t x = c.s->word8(a);
            const uint64_t zq = zero_bytes64(x ^ 0x2222222222222222ULL);
            const uint64_t hit = (zq | zero_bytes64(x ^ 0x5C5C5C5C5C5C5C5CULL)) & (~0ULL << (8u * (i - a)));
            if (!hit) {
                i = a + 8u;
                continue;
            }
            const uint32_t sb = (uint32_t)__builtin_ctzll(hit);
            const uint32_t p = a + (sb >> 3);
            if (p >= c.n) break;
            if ((zq >> sb) & 1ULL) {
                c.i = p + 1u;
                return true;
            }
            i = p + 2u;
        }
        c.i = i;
        return false;
    } else {
        while (c.i < c.n) {
            uint8_t b = c.s->at(c.i);
            if (b == '\\') { c.i += 2; continue; }
            ++c.i;
            if (b == '"') return true;
        }
        return false;
    }
}

template <class Src>
FX_HD bool j_skip_value(JCur<Src>& c) {
    c.ws();
    int b = c.peek();
    if (b < 0) return false;
    if (b == '"') return j_skip_string(c);
    if (b == '{' || b == '[') {
        int depth = 0;
        while (c.i < c.n) {
            uint8_t x = c.s->at(c.i);
            if (x == '"') {
                if (!j_skip_string(c)) return false;
                continue;
            }
            if (x == '{' || x == '[') ++depth;
            else if (x == '}' || x == ']') {
                --depth;
                if (depth == 0) { ++c.i; return true; }
            }
            ++c.i;
        }
        return false;
    }
    uint32_t s0 = c.i;
    while (!c.delim()) ++c.i;
    return c.i > s0;
}

// Key at '"': compares the raw key bytes with `lit` (len n).
template <class Src>
FX_HD bool j_key(JCur<Src>& c, uint32_t& ks, uint32_t& ke) {
    if (c.peek() != '"') return false;
    ks = c.i + 1;
    if (!j_skip_string(c)) return false;
    ke = c.i - 1;
    return true;
}

template <class Src>
FX_HD bool key_eq(const JCur<Src>& c, uint32_t ks, uint32_t ke, const char* lit, uint32_t n) {
    if (ke - ks != n) return false;
    for (uint32_t k = 0; k < n; ++k)
        if (c.s->at(ks + k) != (uint8_t)lit[k]) return false;
    return true;
}

// Bytes of x that are not ASCII digits, as the top bit of each byte (exact).
FX_HD uint64_t nondigit_bytes64(uint64_t x) {
    const uint64_t t = x ^ 0x3030303030303030ULL;
    return (((t & 0x7F7F7F7F7F7F7F7FULL) + 0x7676767676767676ULL) | t) & 0x8080808080808080ULL;
}

// The value of the first k (1..8) bytes of x, all digits, the first byte the
// most significant: Lemire's 8-digit fold (pairs, then quads, then the
// halves) after moving the k digits to the top (the bytes below read as 0).
FX_HD uint64_t fold_digits8(uint64_t x, uint32_t k) {
    uint64_t v = (x & 0x0F0F0F0F0F0F0F0FULL) << (8u * (8u - k));
    v = (v * 2561ULL) >> 8;
    v = ((v & 0x00FF00FF00FF00FFULL) * 6553601ULL) >> 16;
    return ((v & 0x0000FFFF0000FFFFULL) * 42949672960001ULL) >> 32;
}

FX_HD uint64_t pow10_u64(uint32_t k) {
    uint64_t p = 1;
    for (uint32_t j = 0; j < k; ++j) p *= 10;
    return p;
}

// The run of decimal digits at c.i (before c.n): its length and, when it is
// at most 16 digits long, its value. A word source tests 16 bytes at once;
// len 17 stands for "more than 16" (no value).
struct DigitRun {
    uint32_t len;
    uint64_t val;
};

template <class Src>
FX_HD DigitRun j_digits(const JCur<Src>& c) {
    DigitRun d{0, 0};
    const uint32_t i = c.i;
    if (i >= c.n) return d;
    const uint32_t a = i & ~7u, sh = 8u * (i - a);
    const uint64_t w0 = c.s->word8(a), w1 = c.s->word8(a + 8), w2 = c.s->word8(a + 16);
    const uint64_t x = sh ? (w0 >> sh) | (w1 << (64u - sh)) : w0;
    const uint64_t y = sh ? (w1 >> sh) | (w2 << (64u - sh)) : w1;
    const uint64_t ndx = nondigit_bytes64(x), ndy = nondigit_bytes64(y);
    uint32_t len = ndx ? ((uint32_t)__builtin_ctzll(ndx) >> 3) : (ndy ? 8u + ((uint32_t)__builtin_ctzll(ndy) >> 3) : 17u);
    if (len > c.n - i) len = c.n - i;
    d.len = len;
    if (len == 0 || len > 16) return d;
    d.val = len <= 8 ? fold_digits8(x, len) : fold_digits8(x, 8) * pow10_u64(len - 8) + fold_digits8(y, len - 8);
    return d;
}

template <class Src>
FX_HD bool j_int(JCur<Src>& c, int64_t& v) {
    bool neg = false;
    if (c.peek() == '-') { neg = true; ++c.i; }
    if (!is_digit(c.peek())) return false;
    uint64_t acc = 0;
    bool done = false;
    if constexpr (has_word8<Src>::value) {
        const DigitRun d = j_digits(c);
        if (d.len <= 16) {
            acc = d.val;
            c.i += d.len;
            done = true;
        }
    }
    if (!done)
        while (is_digit(c.peek())) { acc = acc * 10 + (uint64_t)(c.s->at(c.i) - '0'); ++c.i; }
    if (!c.delim()) return false;
    v = neg ? (int64_t)(0 - acc) : (int64_t)acc;
    return true;
}

// JSON number: -?(0|[1-9]d*)(.d+)?([eE][+-]?d+)? -> correctly rounded f64,
// plus the exact 1e-6 mantissa when it has no exponent, <= 6 decimals and
// <= 12 integer digits.
template <class Src>
FX_HD bool j_number(JCur<Src>& c, double& v, bool& fx6, int64_t& mant, bool& deferred, uint32_t& from,
                    uint32_t& to) {
    uint32_t s0 = c.i;
    bool neg = false;
    if (c.peek() == '-') { neg = true; ++c.i; }
    if (!is_digit(c.peek())) return false;
    uint32_t ip = c.i;
    // Digit runs by word where the source allows it; the values of runs of
    // at most 16 digits come with them (ival, fval; 0 otherwise).
    uint64_t ival = 0, fval = 0;
    bool vals = true;
    if (c.peek() == '0') {
        ++c.i;
    } else if constexpr (has_word8<Src>::value) {
        const DigitRun d = j_digits(c);
        if (d.len <= 16) { ival = d.val; c.i += d.len; }
        else { vals = false; while (is_digit(c.peek())) ++c.i; }
    } else {
        vals = false;
        while (is_digit(c.peek())) ++c.i;
    }
    uint32_t int_digits = c.i - ip;
    uint32_t frac_digits = 0, fs = 0;
    if (c.peek() == '.') {
        ++c.i;
        fs = c.i;
        if constexpr (has_word8<Src>::value) {
            const DigitRun d = j_digits(c);
            if (d.len <= 16) { fval = d.val; c.i += d.len; }
            else { vals = false; while (is_digit(c.peek())) ++c.i; }
        } else {
            vals = false;
            while (is_digit(c.peek())) ++c.i;
        }
        frac_digits = c.i - fs;
        if (frac_digits == 0) return false;
    }
    bool has_exp = false;
    int e = c.peek();
    if (e == 'e' || e == 'E') {
        has_exp = true;
        ++c.i;
        if (c.peek() == '+' || c.peek() == '-') ++c.i;
        uint32_t es = c.i;
        while (is_digit(c.peek())) ++c.i;
        if (c.i == es) return false;
    }
    if (!c.delim()) return false;
    from = s0;
    to = c.i;
    fx6 = false;
    if (!has_exp && int_digits + frac_digits <= 15) {
        // At most 15 digits and no exponent: the digits as an integer m
        // (< 2^53) over 10^frac_digits (<= 10^15, exact) is one correctly
        // rounded IEEE division (Clinger 1990), the value the general
        // parser gives.
        int64_t m = 0;
        if (vals) {
            m = (int64_t)(ival * pow10_u64(frac_digits) + fval);
        } else {
            for (uint32_t k = ip; k < ip + int_digits; ++k) m = m * 10 + (c.s->at(k) - '0');
            for (uint32_t k = 0; k < frac_digits; ++k) m = m * 10 + (c.s->at(fs + k) - '0');
        }
        const double mag = (double)m / pow10_exact((int)frac_digits);
        v = neg ? -mag : mag;
        deferred = false;
        if (frac_digits <= 6 && int_digits <= 12) {
            for (uint32_t k = frac_digits; k < 6; ++k) m *= 10;
            mant = neg ? -m : m;
            fx6 = true;
        }
        return true;
    }
    // Any other JSON number (it is a valid number of the general grammar):
    // its value is computed after the walk, from [from, to).
    v = 0.0;
    deferred = true;
    if (!has_exp && frac_digits <= 6 && int_digits <= 12) {
        int64_t m = 0;
        for (uint32_t k = ip; k < ip + int_digits; ++k) m = m * 10 + (c.s->at(k) - '0');
        for (uint32_t k = 0; k < 6; ++k)
            m = m * 10 + (k < frac_digits ? (c.s->at(fs + k) - '0') : 0);
        mant = neg ? -m : m;
        fx6 = true;
    }
    return true;
}

struct NdRecord {
    int64_t id;
    double score;
    bool fx6;
    bool score_deferred;          // score is computed after the walk from [sfrom, sto)
    uint32_t sfrom, sto;
    int64_t mant;
};

// One record data[from, to) (its '\n' excluded). 1 = valid, 0 = invalid.
template <class Src>
FX_HD int ndjson_record(const Src& s, uint32_t from, uint32_t to, NdRecord& out) {
    JCur<Src> c{&s, from, to};
    out.id = 0;
    out.score = 0;
    out.fx6 = false;
    out.score_deferred = false;
    out.sfrom = out.sto = 0;
    out.mant = 0;
    c.ws();
    if (c.peek() != '{') return 0;
    ++c.i;
    bool have_id = false, have_score = false;
    c.ws();
    if (c.peek() == '}') return 0;
    while (true) {
        c.ws();
        uint32_t ks, ke;
        if (!j_key(c, ks, ke)) return 0;
        c.ws();
        if (c.peek() != ':') return 0;
        ++c.i;
        c.ws();
        if (!have_id && key_eq(c, ks, ke, "id", 2)) {
            if (!j_int(c, out.id)) return 0;
            have_id = true;
        } else if (!have_score && key_eq(c, ks, ke, "nested", 6) && c.peek() == '{') {
            ++c.i;
            c.ws();
            if (c.peek() == '}') {
                ++c.i;
            } else {
                while (true) {
                    c.ws();
                    uint32_t ks2, ke2;
                    if (!j_key(c, ks2, ke2)) return 0;
                    c.ws();
                    if (c.peek() != ':') return 0;
                    ++c.i;
                    c.ws();
                    if (!have_score && key_eq(c, ks2, ke2, "score", 5)) {
                        if (!j_number(c, out.score, out.fx6, out.mant, out.score_deferred, out.sfrom, out.sto)) return 0;
                        have_score = true;
                    } else if (!j_skip_value(c)) {
                        return 0;
                    }
                    c.ws();
                    if (c.peek() == ',') { ++c.i; continue; }
                    if (c.peek() == '}') { ++c.i; break; }
                    return 0;
                }
            }
        } else if (!j_skip_value(c)) {
            return 0;
        }
        c.ws();
        if (c.peek() == ',') { ++c.i; continue; }
        if (c.peek() == '}') { ++c.i; break; }
        return 0;
    }
    return (have_id && have_score) ? 1 : 0;
}

// A deferred score by the general parser; true when that leaves it to the
// exact path (ndjson_resolve_exact).
template <class Src>
FX_HD bool ndjson_score_general(const Src& s, NdRecord& r) {
    const F64Result f = parse_f64_stream(RawStream<Src>{&s, r.sfrom, r.sto});
    r.score = f.value;
    return f.deferred;
}

// The exact path for a score (storage in `b`).
template <class Src>
FX_HD void ndjson_resolve_exact(const Src& s, NdRecord& r, BigDec& b) {
    r.score = bits_to_f64(exact_f64_bits(RawStream<Src>{&s, r.sfrom, r.sto}, b));
    r.score_deferred = false;
}

// Resolves a deferred score: the general parser, then the exact path if it
// needs it.
template <class Src>
FX_HD void ndjson_resolve(const Src& s, NdRecord& r, BigDec& b) {
    if (!r.score_deferred) return;
    if (ndjson_score_general(s, r)) ndjson_resolve_exact(s, r, b);
    r.score_deferred = false;
}

template <class Src>
FX_HD bool blank_span(const Src& s, uint32_t from, uint32_t to) {
    for (uint32_t p = from; p < to; ++p)
        if (!jws(s.at(p))) return false;
    return true;
}

}  // namespace fx
