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

// ---------------------------------------------------------------------------
// The common record layout, read in straight-line steps.
//
// The documents of this sink repeat one key order (the layout ndjson.rs:696-711
// reads by position):
//   {"id":INT,"name":STRING,"nested":{"score":NUMBER,"tags":VALUE}}
// with spaces around the punctuation. ndjson_common reads a record of that
// layout in a fixed sequence of steps, the same for every record, so the lanes
// of a wave take them together instead of each running the key loop of
// ndjson_record on its own. Each step is one of ndjson_record's steps on that
// layout; the first byte a step does not expect clears `ok` and freezes the
// cursor, and the record then goes to ndjson_record. Whitespace other than
// ' ', exponents, numbers of more than 15 digits and any other key order are
// left to ndjson_record. The source reads aligned 8-byte words (word8) up to
// 24 bytes past `to`, and the byte at `to` is '\n' or 0, which no step takes,
// so the cursor never passes `to`.
// ---------------------------------------------------------------------------

FX_HD uint32_t ctz64(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }

// Bytes sh/8 .. sh/8 + 7 of the 16 bytes hi:lo (sh a multiple of 8, 0..64).
FX_HD uint64_t nd_funnel(uint64_t lo, uint64_t hi, uint32_t sh) {
    return sh == 0 ? lo : (sh >= 64 ? hi : (lo >> sh) | (hi << (64u - sh)));
}

template <class Src>
FX_HD uint64_t nd_bytes8(const Src& s, uint32_t i) {
    const uint32_t a = i & ~7u;
    return nd_funnel(s.word8(a), s.word8(a + 8), 8u * (i - a));
}

// A little-endian word of up to 8 literal bytes.
constexpr uint64_t nd_lit8(const char* t, uint32_t n) {
    return n == 0 ? 0ULL : (uint64_t)(uint8_t)t[0] | (nd_lit8(t + 1, n - 1) << 8);
}

constexpr uint64_t ND_HI = 0x8080808080808080ULL;
// The keys of the common layout with their quotes.
constexpr uint64_t ND_K_ID = nd_lit8("\"id\"", 4), ND_K_NAME = nd_lit8("\"name\"", 6),
                   ND_K_NESTED = nd_lit8("\"nested\"", 8), ND_K_SCORE = nd_lit8("\"score\"", 7),
                   ND_K_TAGS = nd_lit8("\"tags\"", 6);

// Bit 7 of every byte of x that is not ' '.
FX_HD uint64_t nd_nonspace(uint64_t x) { return ~zero_bytes64(x ^ 0x2020202020202020ULL) & ND_HI; }

// ' '* P ' '* at i: the position of the byte after the spaces that follow P.
// Both runs of spaces and P lie in 8 bytes.
template <class Src>
FX_HD uint32_t nd_punct(const Src& s, uint32_t i, uint32_t P, bool& ok) {
    const uint64_t x = nd_bytes8(s, i);
    const uint64_t ns = nd_nonspace(x);
    const uint32_t k = ns ? ctz64(ns) >> 3 : 0u;
    const uint64_t after = (ns >> (8u * k)) & ~0xFFULL;
    const bool go = ok && ns != 0 && ((uint32_t)(x >> (8u * k)) & 0xFFu) == P && after != 0;
    ok = go;
    return go ? i + k + (ctz64(after | (1ULL << 63)) >> 3) : i;
}

// ' '* '}' at i: the position of the '}'.
template <class Src>
FX_HD uint32_t nd_close(const Src& s, uint32_t i, bool& ok) {
    const uint64_t x = nd_bytes8(s, i);
    const uint64_t ns = nd_nonspace(x);
    const uint32_t k = ns ? ctz64(ns) >> 3 : 0u;
    const bool go = ok && ns != 0 && ((uint32_t)(x >> (8u * k)) & 0xFFu) == '}';
    ok = go;
    return go ? i + k : i;
}

// The n (1..8) bytes of `lit` at i (a key with its quotes).
template <class Src>
FX_HD uint32_t nd_key(const Src& s, uint32_t i, uint64_t lit, uint32_t n, bool& ok) {
    const uint64_t mask = n >= 8 ? ~0ULL : (1ULL << (8u * n)) - 1u;
    const bool go = ok && ((nd_bytes8(s, i) ^ lit) & mask) == 0;
    ok = go;
    return go ? i + n : i;
}

// The first byte at or after byte `from` of a 16-byte window whose flags
// (bit 7 of each byte) are fx (bytes 0-7) and fy (bytes 8-15); 16 when none.
FX_HD uint32_t nd_first(uint64_t fx, uint64_t fy, uint32_t from) {
    if (from >= 16) return 16;
    const uint64_t mx = from >= 8 ? 0ULL : fx & (~0ULL << (8u * from));
    const uint64_t my = from <= 8 ? fy : fy & (~0ULL << (8u * (from - 8u)));
    return mx ? ctz64(mx) >> 3 : (my ? 8u + (ctz64(my) >> 3) : 16u);
}

// 10^k for k < 16, exact.
FX_HD uint64_t nd_pow10(uint32_t k) {
    uint64_t p = (k & 1u) ? 10ULL : 1ULL;
    if (k & 2u) p *= 100ULL;
    if (k & 4u) p *= 10000ULL;
    if (k & 8u) p *= 100000000ULL;
    return p;
}

// The value of the first n (1..16) bytes of the window x:y, all digits, the
// first the most significant: the digits are moved to the top of the 16
// bytes (the zero bytes below them read as leading zeros) and both halves
// folded (fold_digits8).
FX_HD uint64_t nd_value16(uint64_t x, uint64_t y, uint32_t n) {
    const uint32_t s = 8u * (16u - n);
    uint64_t lo, hi;
    if (s >= 64) {
        lo = 0;
        hi = x << (s - 64u);
    } else if (s == 0) {
        lo = x;
        hi = y;
    } else {
        lo = x << s;
        hi = (y << s) | (x >> (64u - s));
    }
    return fold_digits8(lo, 8) * 100000000ULL + fold_digits8(hi, 8);
}

// A number at i: its sign and the 16 bytes after the sign.
template <class Src>
FX_HD bool nd_signed16(const Src& s, uint32_t i, uint64_t& x, uint64_t& y) {
    const uint32_t a = i & ~7u, sh = 8u * (i - a);
    const uint64_t w0 = s.word8(a), w1 = s.word8(a + 8), w2 = s.word8(a + 16);
    const bool neg = (nd_funnel(w0, w1, sh) & 0xFFu) == '-';
    const uint32_t sh2 = sh + (neg ? 8u : 0u);
    x = nd_funnel(w0, w1, sh2);
    y = nd_funnel(w1, w2, sh2);
    return neg;
}

// j_int on -?[0-9]{1,15}: the position after the digits (the byte there is
// checked by the next step, which takes only ' ' or ',').
template <class Src>
FX_HD uint32_t nd_int(const Src& s, uint32_t i, int64_t& v, bool& ok) {
    uint64_t x, y;
    const bool neg = nd_signed16(s, i, x, y);
    const uint32_t n = nd_first(nondigit_bytes64(x), nondigit_bytes64(y), 0);
    const bool go = ok && n >= 1 && n <= 15;
    const uint64_t val = nd_value16(x, y, go ? n : 1u);
    v = neg ? (int64_t)(0 - val) : (int64_t)val;
    ok = go;
    return go ? i + (neg ? 1u : 0u) + n : i;
}

// j_number on -?(0|[1-9][0-9]*)(.[0-9]+)? with at most 15 digits in all
// (its fast branch: m / 10^frac_digits, one correctly rounded division, and
// the 1e-6 mantissa for <= 6 decimals and <= 12 integer digits).
template <class Src>
FX_HD uint32_t nd_number(const Src& s, uint32_t i, NdRecord& r, bool& ok) {
    uint64_t x, y;
    const bool neg = nd_signed16(s, i, x, y);
    const uint64_t fx = nondigit_bytes64(x), fy = nondigit_bytes64(y);
    const uint32_t p1 = nd_first(fx, fy, 0);   // integer digits
    const uint32_t b1 = p1 < 8 ? (uint32_t)(x >> (8u * p1)) & 0xFFu : (p1 < 16 ? (uint32_t)(y >> (8u * (p1 - 8u))) & 0xFFu : 0u);
    const bool dot = b1 == '.';
    const uint32_t p2 = dot ? nd_first(fx, fy, p1 + 1u) : p1;   // end of the number
    const uint32_t fd = dot ? p2 - p1 - 1u : 0u;                 // fraction digits
    const uint32_t nd = p1 + fd;
    const bool go = ok && p1 >= 1 && !((x & 0xFFu) == '0' && p1 > 1) && (!dot || fd >= 1) && p2 < 16 && nd <= 15;
    // The digits without the '.': bytes [0, p1), then those from p1 + 1 on.
    const uint64_t xs = (x >> 8) | (y << 56), ys = y >> 8;
    const uint64_t kx = dot ? (p1 >= 8 ? ~0ULL : (1ULL << (8u * p1)) - 1u) : ~0ULL;
    const uint64_t ky = dot ? (p1 <= 8 ? 0ULL : (1ULL << (8u * (p1 - 8u))) - 1u) : ~0ULL;
    const uint64_t dx = (x & kx) | (xs & ~kx), dy = (y & ky) | (ys & ~ky);
    const uint64_t m = nd_value16(dx, dy, go ? nd : 1u);
    const double mag = (double)m / (double)nd_pow10(go ? fd : 0u);
    r.score = neg ? -mag : mag;
    r.score_deferred = false;
    r.sfrom = r.sto = 0;
    r.fx6 = fd <= 6 && p1 <= 12;
    const int64_t mant = (int64_t)(m * nd_pow10(r.fx6 ? 6u - fd : 0u));
    r.mant = r.fx6 ? (neg ? -mant : mant) : 0;
    ok = go;
    return go ? i + (neg ? 1u : 0u) + p2 : i;
}

// j_skip_value on a string, array or object at i: strings by their quotes
// and backslashes, brackets by depth, 8 bytes per step between the bytes
// that matter. The position after the value.
template <class Src>
FX_HD uint32_t nd_skip(const Src& s, uint32_t i, uint32_t to, bool& ok) {
    const uint32_t b0 = (uint32_t)(nd_bytes8(s, i) & 0xFFu);
    bool bad = !ok || !(b0 == '"' || b0 == '[' || b0 == '{');
    bool instr = false, done = false;
    int depth = 0;
    uint32_t p = i;
    while (!done && !bad) {
        const uint32_t a = p & ~7u;
        const uint64_t x = s.word8(a);
        const uint64_t zq = zero_bytes64(x ^ 0x2222222222222222ULL);
        // In a string: '\\'. Outside: '[', ']', '{', '}' (and 'Y', '_',
        // 'y', 0x7F, which share their bits under 0xD9 and pass as plain).
        const uint64_t zo = instr ? zero_bytes64(x ^ 0x5C5C5C5C5C5C5C5CULL)
                                  : zero_bytes64((x & 0xD9D9D9D9D9D9D9D9ULL) ^ 0x5959595959595959ULL);
        const uint64_t hit = (zq | zo) & (~0ULL << (8u * (p - a)));
        const uint32_t sb = hit ? ctz64(hit) : 71u;
        const uint32_t pos = a + (sb >> 3);   // a + 8 when the word has none
        if (pos >= to) {
            bad = true;
        } else if (hit) {
            const uint32_t ch = (uint32_t)(x >> (sb & ~7u)) & 0xFFu;
            if (instr) {
                const bool q = ch == '"';
                instr = !q;
                done = q && depth == 0;
                p = pos + (q ? 1u : 2u);
            } else {
                const bool open = ch == '[' || ch == '{', close = ch == ']' || ch == '}';
                instr = ch == '"';
                depth += open ? 1 : (close ? -1 : 0);
                done = close && depth == 0;
                p = pos + 1u;
            }
        } else {
            p = pos;
        }
    }
    ok = !bad;
    return bad ? i : p;
}

// nd_skip on an array or object whose strings hold no backslash and which
// holds no array or object: 8 bytes per step whatever the number of strings
// in them (a string's quotes pair up by a prefix XOR over the word's quote
// bytes), ending at the first ']' or '}' outside a string. On any other
// value it returns nd_skip's result: the steps above take a byte the same
// way nd_skip does whenever they do not hand over to it.
template <class Src>
FX_HD uint32_t nd_skip_flat(const Src& s, uint32_t i, uint32_t to, bool& ok) {
    const uint32_t b0 = (uint32_t)(nd_bytes8(s, i) & 0xFFu);
    if (!ok || !(b0 == '[' || b0 == '{')) return nd_skip(s, i, to, ok);
    uint32_t p = i + 1u;
    bool instr = false, flat = true, done = false;
    uint32_t end = i;
    while (flat && !done) {
        const uint32_t a = p & ~7u;
        const uint64_t x = s.word8(a);
        const uint64_t vm = ND_HI & (~0ULL << (8u * (p - a)));
        const uint64_t zq = zero_bytes64(x ^ 0x2222222222222222ULL) & vm;
        const uint64_t zb = zero_bytes64(x ^ 0x5C5C5C5C5C5C5C5CULL) & vm;
        const uint64_t xm = x & 0xDFDFDFDFDFDFDFDFULL;   // '{' -> '[', '}' -> ']'
        const uint64_t zo = zero_bytes64(xm ^ 0x5B5B5B5B5B5B5B5BULL) & vm;
        const uint64_t zc = zero_bytes64(xm ^ 0x5D5D5D5D5D5D5D5DULL) & vm;
        // Inside a string before each byte: the quotes before it in the
        // word, and `instr` on entry.
        uint64_t v = zq;
        v ^= v << 8;
        v ^= v << 16;
        v ^= v << 32;
        const uint64_t in = (v ^ zq) ^ (instr ? vm : 0ULL);
        const uint64_t close = zc & ~in;
        const uint32_t cpos = close ? a + (ctz64(close) >> 3) : 0xFFFFFFFFu;
        // Bytes up to the close (or the whole word) must hold no backslash in
        // a string and no open outside one.
        const uint64_t upto = close ? (close | (close - 1u)) : ~0ULL;
        flat = ((zb & in) | (zo & ~in)) & upto ? false : true;
        if (flat && close) {
            done = true;
            flat = cpos < to;
            end = cpos + 1u;
        } else if (flat) {
            instr = instr != ((__builtin_popcountll(zq) & 1) != 0);
            p = a + 8u;
            flat = p < to;
        }
    }
    if (!flat) return nd_skip(s, i, to, ok);
    return end;
}

// The record data[from, to) in the common layout: true with `out` filled as
// ndjson_record fills it (which then returns 1), false when the record is of
// another form.
template <class Src>
FX_HD bool ndjson_common(const Src& s, uint32_t from, uint32_t to, NdRecord& out) {
    bool ok = true;
    uint32_t i = from;
    out.id = 0;
    i = nd_punct(s, i, '{', ok);
    i = nd_key(s, i, ND_K_ID, 4, ok);
    i = nd_punct(s, i, ':', ok);
    i = nd_int(s, i, out.id, ok);
    i = nd_punct(s, i, ',', ok);
    i = nd_key(s, i, ND_K_NAME, 6, ok);
    i = nd_punct(s, i, ':', ok);
    i = nd_skip(s, i, to, ok);
    i = nd_punct(s, i, ',', ok);
    i = nd_key(s, i, ND_K_NESTED, 8, ok);
    i = nd_punct(s, i, ':', ok);
    i = nd_punct(s, i, '{', ok);
    i = nd_key(s, i, ND_K_SCORE, 7, ok);
    i = nd_punct(s, i, ':', ok);
    i = nd_number(s, i, out, ok);
    i = nd_punct(s, i, ',', ok);
    i = nd_key(s, i, ND_K_TAGS, 6, ok);
    i = nd_punct(s, i, ':', ok);
    i = nd_skip_flat(s, i, to, ok);
    i = nd_punct(s, i, '}', ok);
    i = nd_close(s, i, ok);
    return ok && i < to;
}

template <class Src>
FX_HD bool blank_span(const Src& s, uint32_t from, uint32_t to) {
    for (uint32_t p = from; p < to; ++p)
        if (!jws(s.at(p))) return false;
    return true;
}

}  // namespace fx
