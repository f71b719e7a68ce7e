// cells.h — cell-level parsing shared by the gfx950 kernels and their host
// build (the host build exists so the CPU test-suite can sweep these exact
// functions against the oracle).
//
// Semantics follow the reference's generated cell parsers:
//  - i64: exactly `str::parse::<i64>` (codegen.rs INT_CELL_TPL 7500-7565):
//    `[+-]?[0-9]+`, overflow invalid, i64::MIN accepted.
//  - f64: exactly `str::parse::<f64>` (codegen.rs FLOAT_CELL_TPL 7566-7744,
//    csv_geo.rs:1313-1486): grammar `[+-]?(inf|infinity|nan|(d+.?d*|.d+)
//    ([eE][+-]?d+)?)` (inf/nan case-insensitive), correctly rounded, ties to
//    even. Paths, all exact: the fixed `[-]d{1,3}.d{6}` shape (m/1e6);
//    Clinger's fast path (exact mantissa <= 2^53, |e10| <= 22); Eisel-Lemire
//    (Lemire 2021) with a 128-bit 5^q table, exact integer rounding where a
//    tie is possible and a margin test elsewhere; and, for the cells that
//    leaves undecided (subnormal results, >19-digit mantissas whose
//    truncation bounds disagree, products within 2 units of a rounding
//    boundary), an exact big-decimal conversion (shift the decimal by powers
//    of two into [1/2, 1), then round the 53-bit significand half-to-even).
//    The kernels hand those rare cells to a separate fixup pass.
//  - cleaning: doubled-quote collapse or backslash unescape (codegen.rs
//    append_clean_fn / clean_fn 8008-8185).
//
// Cells are read through small stream objects so quoted cells are cleaned
// on the fly, without a buffer.
#pragma once

#include <stdint.h>

#include "pow5_table.h"

#if defined(__HIPCC__)
#define FX_HD __host__ __device__ __forceinline__
// Out of line on the device: for paths taken rarely, whose inlined code
// would only bloat the kernels that reach them.
#define FX_HD_COLD __host__ __device__ __attribute__((noinline))
#else
#define FX_HD inline
#define FX_HD_COLD inline
#endif

namespace fx {

FX_HD uint64_t f64_to_bits(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(v);
#else
    uint64_t b;
    __builtin_memcpy(&b, &v, 8);
    return b;
#endif
}

FX_HD double bits_to_f64(uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __longlong_as_double((long long)b);
#else
    double v;
    __builtin_memcpy(&v, &b, 8);
    return v;
#endif
}

FX_HD bool is_digit(int c) { return (unsigned)(c - '0') < 10u; }

// ---------------------------------------------------------------------------
// Streams. `Src` provides `uint8_t at(uint32_t pos) const`.
// next() returns the next byte of the (cleaned) cell, or -1 at its end.
// ---------------------------------------------------------------------------
template <class Src>
struct RawStream {
    const Src* s;
    uint32_t pos, end;
    FX_HD int next() { return pos < end ? (int)s->at(pos++) : -1; }
};

// Doubled-quote dialect: a cell that starts and ends with the quote byte
// (length >= 2) loses its outer quotes and each `QQ` collapses to `Q`.
template <class Src>
struct DoubledStream {
    const Src* s;
    uint32_t pos, end;
    uint8_t q;
    FX_HD int next() {
        if (pos >= end) return -1;
        uint8_t b = s->at(pos);
        if (b == q && pos + 1 < end && s->at(pos + 1) == q) pos += 2; else pos += 1;
        return b;
    }
};

// The second quotes of the doubled pairs among the quote bytes Q of a
// 32-byte word (bit i: byte i): the odd offsets within each run of quotes,
// counted from the run's first byte, which is what a left-to-right
// doubled-quote stream drops (DoubledStream in cells.h pairs a quote with the
// next one, then goes on after the pair). A run's start parity is found by
// adding its first bit to Q (the carry clears the run), so runs starting at
// even bytes drop their odd bytes and the others their even bytes.
// `in_run` and `odd_next` carry a run from the previous word: whether that
// word ended in a quote, and whether the offset of this word's first byte in
// that run is odd.
FX_HD uint32_t dq_dropped(uint32_t Q, uint32_t& in_run, uint32_t& odd_next) {
    constexpr uint32_t EVEN32 = 0x55555555u, ODD32 = 0xAAAAAAAAu;
    const uint32_t cont = in_run & Q & 1u;                        // a run goes on into byte 0
    const uint32_t starts = Q & ~((Q << 1) | cont);
    const uint32_t s_even = (starts & EVEN32) | (cont & (odd_next ^ 1u));
    const uint32_t r_even = Q & ~(Q + s_even);                    // runs of even start parity
    const uint32_t r_odd = Q & ~r_even;
    in_run = Q >> 31;
    odd_next = (r_odd >> 31) & 1u;
    return (r_even & ODD32) | (r_odd & EVEN32);
}

// Backslash dialect: outer quotes stripped when present at both ends, then
// `E x` -> `x` (a trailing lone escape byte is kept).
template <class Src>
struct BackslashStream {
    const Src* s;
    uint32_t pos, end;
    uint8_t e;
    FX_HD int next() {
        if (pos >= end) return -1;
        uint8_t b = s->at(pos);
        if (b == e && pos + 1 < end) { b = s->at(pos + 1); pos += 2; } else pos += 1;
        return b;
    }
};

// ---------------------------------------------------------------------------
// i64
// ---------------------------------------------------------------------------
template <class St>
FX_HD bool parse_i64_stream(St st, int64_t& out) {
    int c = st.next();
    bool neg = false;
    if (c == '-' || c == '+') { neg = c == '-'; c = st.next(); }
    if (c < 0) return false;
    uint64_t v = 0;
    // v * 10 + d must stay within 2^63 - 1 (2^63 when negative); both
    // limits are 922337203685477580 * 10 + 7 (+ 8), so no division is needed.
    const uint64_t lim10 = 922337203685477580ULL, lim_d = neg ? 8u : 7u;
    while (c >= 0) {
        if (!is_digit(c)) return false;
        uint64_t d = (uint64_t)(c - '0');
        if (v > lim10 || (v == lim10 && d > lim_d)) return false;
        v = v * 10 + d;
        c = st.next();
    }
    out = neg ? (int64_t)(0 - v) : (int64_t)v;
    return true;
}

// ---------------------------------------------------------------------------
// f64: exact big-decimal conversion (the slow path)
// ---------------------------------------------------------------------------
constexpr int DEC_MAX = 800;

struct BigDec {
    uint32_t nd;
    int32_t dp;      // value = 0.d0 d1 d2 ... * 10^dp
    bool trunc;      // nonzero digits were dropped beyond DEC_MAX
    uint8_t d[DEC_MAX + 24];

    FX_HD void trim() { while (nd > 0 && d[nd - 1] == 0) --nd; }

    // Multiply by 2^k, 1 <= k <= 60 (at most 19 new leading digits).
    FX_HD void left_shift(int k) {
        if (nd == 0) return;
        int32_t r = (int32_t)nd - 1;
        int32_t w = (int32_t)nd - 1 + 19;
        uint64_t n = 0;
        while (r >= 0) {
            n += (uint64_t)d[r] << k;
            uint64_t q = n / 10;
            d[w] = (uint8_t)(n - 10 * q);
            n = q;
            --w;
            --r;
        }
        while (n > 0) {
            uint64_t q = n / 10;
            d[w] = (uint8_t)(n - 10 * q);
            n = q;
            --w;
        }
        int32_t shift = w + 1;
        int32_t new_nd = (int32_t)nd + 19 - shift;
        dp += 19 - shift;
        for (int32_t i = 0; i < new_nd; ++i) d[i] = d[i + shift];
        if (new_nd > DEC_MAX) {
            for (int32_t i = DEC_MAX; i < new_nd; ++i)
                if (d[i]) trunc = true;
            new_nd = DEC_MAX;
        }
        nd = (uint32_t)new_nd;
        trim();
    }

    // Divide by 2^k, 1 <= k <= 60 (truncating; dropped digits are sticky).
    FX_HD void right_shift(int k) {
        uint32_t r = 0, w = 0;
        uint64_t n = 0;
        while ((n >> k) == 0) {
            if (r < nd) {
                n = 10 * n + d[r++];
            } else if (n == 0) {
                nd = 0;
                return;
            } else {
                while ((n >> k) == 0) { n *= 10; ++r; }
                break;
            }
        }
        dp -= (int32_t)r - 1;
        const uint64_t mask = (1ULL << k) - 1;
        while (r < nd) {
            uint8_t nd_digit = (uint8_t)(n >> k);
            n = 10 * (n & mask) + d[r++];
            d[w++] = nd_digit;
        }
        while (n > 0) {
            uint8_t nd_digit = (uint8_t)(n >> k);
            n = 10 * (n & mask);
            if (w < DEC_MAX) d[w++] = nd_digit;
            else if (nd_digit > 0) trunc = true;
        }
        nd = w;
        trim();
    }

    // Integer part, rounded half-to-even (sticky `trunc` counts as above half).
    FX_HD uint64_t round_int() const {
        if (dp < 0 || nd == 0) return 0;
        uint64_t n = 0;
        for (int32_t i = 0; i < dp; ++i) n = n * 10 + (i < (int32_t)nd ? d[i] : 0);
        bool up = false;
        if (dp < (int32_t)nd) {
            up = d[dp] >= 5;
            if (d[dp] == 5 && dp + 1 == (int32_t)nd) up = trunc || (n & 1);
        }
        return n + (up ? 1 : 0);
    }
};

// floor(k * log2(10)) for k < 19: shifts that move the decimal point by ~k.
FX_HD int pow2_shift(int k) {
    const uint8_t t[19] = {0, 3, 6, 9, 13, 16, 19, 23, 26, 29, 33, 36, 39, 43, 46, 49, 53, 56, 59};
    return k < 19 ? t[k] : 60;
}

// Convert a BigDec (already loaded, sign aside) to IEEE double bits.
FX_HD uint64_t bigdec_to_bits(BigDec& b) {
    b.trim();
    if (b.nd == 0) return 0;
    if (b.dp > 310) return 0x7FF0000000000000ULL;
    if (b.dp < -330) return 0;
    int32_t exp2 = 0;
    while (b.dp > 0) {
        int n = pow2_shift(b.dp);
        b.right_shift(n);
        exp2 += n;
    }
    while (b.dp < 0 || (b.dp == 0 && b.d[0] < 5)) {
        int n = b.dp == 0 ? 1 : pow2_shift(-b.dp);
        b.left_shift(n);
        exp2 -= n;
    }
    // value = dec * 2^exp2 with dec in [1/2, 1): unbiased exponent E.
    int32_t E = exp2 - 1;
    if (E > 1023) return 0x7FF0000000000000ULL;
    if (E < -1022) {
        int32_t n = -1022 - E;
        if (n > 60 * 20) return 0;
        while (n > 0) {
            int k = n > 60 ? 60 : n;
            b.right_shift(k);
            n -= k;
        }
        E = -1022;
    }
    // Significand: round(dec * 2^53).
    b.left_shift(53);
    uint64_t m = b.round_int();
    if (m >= (1ULL << 53)) { m >>= 1; E += 1; }
    if (E > 1023) return 0x7FF0000000000000ULL;
    if (m < (1ULL << 52)) return m;  // subnormal (E == -1022)
    return ((uint64_t)(E + 1023) << 52) | (m & ((1ULL << 52) - 1));
}

// ---------------------------------------------------------------------------
// Eisel-Lemire with exact tie handling
// ---------------------------------------------------------------------------
#if defined(__HIPCC__)
__device__ static const uint64_t kPow5Dev[FX_POW5_QMAX - FX_POW5_QMIN + 1][2] = FX_POW5_TABLE_INIT;
#endif
static const uint64_t kPow5Host[FX_POW5_QMAX - FX_POW5_QMIN + 1][2] = FX_POW5_TABLE_INIT;

FX_HD void pow5_entry(int32_t q, uint64_t& hi, uint64_t& lo) {
#if defined(__HIP_DEVICE_COMPILE__)
    hi = kPow5Dev[q - FX_POW5_QMIN][0];
    lo = kPow5Dev[q - FX_POW5_QMIN][1];
#else
    hi = kPow5Host[q - FX_POW5_QMIN][0];
    lo = kPow5Host[q - FX_POW5_QMIN][1];
#endif
}

FX_HD int clz64(uint64_t v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __clzll((long long)v);
#else
    return __builtin_clzll(v);
#endif
}

constexpr uint64_t F64_INF_BITS = 0x7FF0000000000000ULL;

struct ElOut {
    uint64_t bits;  // positive IEEE bits
    bool ok;        // false: undecided, use the exact big-decimal path
};

// value = v * 2^e2 for an integer v >= 1, rounded half-to-even to 53 bits.
// The caller keeps the result in the normal range.
FX_HD uint64_t round_dyadic(uint64_t v, int32_t e2) {
    int nbits = 64 - clz64(v);
    uint64_t M;
    int32_t e = e2;
    if (nbits <= 53) {
        M = v << (53 - nbits);
        e -= 53 - nbits;
    } else {
        int sh = nbits - 53;
        uint64_t rem = v & ((1ULL << sh) - 1), half = 1ULL << (sh - 1);
        M = v >> sh;
        e += sh;
        if (rem > half || (rem == half && (M & 1))) ++M;
        if (M == (1ULL << 53)) { M >>= 1; ++e; }
    }
    return ((uint64_t)(e + 52 + 1023) << 52) | (M & ((1ULL << 52) - 1));
}

// w * 10^q (w exact, w != 0) -> double bits, or undecided.
FX_HD ElOut eisel_lemire(uint64_t w, int32_t q) {
    if (w == 0 || q < FX_POW5_QMIN) return ElOut{0, true};
    if (q > FX_POW5_QMAX) return ElOut{F64_INF_BITS, true};
    if (q < 0 && q >= -27) {
        // w * 10^q is a dyadic rational iff 5^-q divides w: round exactly.
        uint64_t p = 1;
        for (int32_t i = 0; i < -q; ++i) p *= 5;
        if (w % p == 0) return ElOut{round_dyadic(w / p, q), true};
    }
    const int lz = clz64(w);
    const uint64_t wn = w << lz;
    uint64_t th, tl;
    pow5_entry(q, th, tl);
    typedef unsigned __int128 u128;
    const u128 a = (u128)wn * th, b = (u128)wn * tl;
    const uint64_t p0 = (uint64_t)b;
    const u128 mid = (u128)(uint64_t)a + (uint64_t)(b >> 64);
    const uint64_t p1 = (uint64_t)mid;
    const uint64_t p2 = (uint64_t)(a >> 64) + (uint64_t)(mid >> 64);
    const int upper = (int)(p2 >> 63);
    const int sh = upper + 9;
    const uint64_t m54 = p2 >> sh;                    // 53 bits + rounding bit
    const uint64_t rtop = p2 & ((1ULL << sh) - 1);    // discarded bits above p1
    int32_t biased = ((q * 217706) >> 16) + 63 + upper - lz + 1023;
    if (biased <= 0) return ElOut{0, false};          // subnormal: exact path
    uint64_t M;
    if (q >= 0 && q <= 55) {
        // 5^q is exact in the table: the 192-bit product is exact.
        if ((m54 & 1) && rtop == 0 && p1 == 0 && p0 == 0) {
            M = m54 >> 1;
            M += M & 1;                               // tie: to even
        } else {
            M = (m54 + (m54 & 1)) >> 1;
        }
    } else {
        // No tie is possible here; the product (p0 dropped) is within 2
        // units of p1's last place of the true value.
        const bool near_low = rtop == 0 && p1 <= 2;
        const bool near_high = rtop == ((1ULL << sh) - 1) && p1 >= ~0ULL - 2;
        if (((m54 & 1) && near_low) || (!(m54 & 1) && near_high)) return ElOut{0, false};
        M = (m54 + (m54 & 1)) >> 1;
    }
    if (M == (1ULL << 53)) { M >>= 1; ++biased; }
    if (biased >= 2047) return ElOut{F64_INF_BITS, true};
    return ElOut{((uint64_t)biased << 52) | (M & ((1ULL << 52) - 1)), true};
}

// ---------------------------------------------------------------------------
// f64 scan + fast paths
// ---------------------------------------------------------------------------
struct F64Result {
    bool ok;
    bool fixed6;
    bool deferred;      // valid cell whose value needs the exact slow path
    int64_t mantissa;   // signed 1e-6 units when fixed6
    double value;
};

FX_HD double pow10_exact(int e) {
    const double t[23] = {1e0, 1e1, 1e2, 1e3, 1e4, 1e5, 1e6, 1e7, 1e8, 1e9, 1e10, 1e11,
                          1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};
    return t[e];
}

// Slow path: reload the digits of an already-validated numeric cell into
// `b` and convert exactly. Returns positive bits (sign applied by caller).
template <class St>
FX_HD uint64_t slow_f64_bits_into(BigDec& b, St st) {
    b.nd = 0;
    b.dp = 0;
    b.trunc = false;
    int c = st.next();
    if (c == '+' || c == '-') c = st.next();
    bool seen_nonzero = false;
    while (is_digit(c)) {
        int dg = c - '0';
        if (dg != 0 || seen_nonzero) {
            seen_nonzero = true;
            if (b.nd < DEC_MAX) b.d[b.nd++] = (uint8_t)dg; else if (dg) b.trunc = true;
            ++b.dp;
        }
        c = st.next();
    }
    if (c == '.') {
        c = st.next();
        while (is_digit(c)) {
            int dg = c - '0';
            if (dg != 0 || seen_nonzero) {
                seen_nonzero = true;
                if (b.nd < DEC_MAX) b.d[b.nd++] = (uint8_t)dg; else if (dg) b.trunc = true;
            } else {
                --b.dp;
            }
            c = st.next();
        }
    }
    if (c == 'e' || c == 'E') {
        c = st.next();
        bool eneg = false;
        if (c == '+' || c == '-') { eneg = c == '-'; c = st.next(); }
        int64_t e = 0;
        while (is_digit(c)) {
            if (e < 100000000) e = e * 10 + (c - '0');
            c = st.next();
        }
        int64_t dp = (int64_t)b.dp + (eneg ? -e : e);
        if (dp > 100000) dp = 100000;
        if (dp < -100000) dp = -100000;
        b.dp = (int32_t)dp;
    }
    return bigdec_to_bits(b);
}

FX_HD F64Result f64_decide(bool neg, bool plus, bool dot, bool has_exp, int int_digits, int frac_digits,
                           uint64_t mant, int32_t exp10, bool dropped);

// Scan + fast paths. A cell the fast paths cannot decide comes back with
// ok && deferred (its value unset); see parse_f64_exact.
template <class St>
FX_HD F64Result parse_f64_stream(St st) {
    F64Result r{false, false, false, 0, 0.0};
    int c = st.next();
    bool neg = false, plus = false;
    if (c == '-' || c == '+') { neg = c == '-'; plus = c == '+'; c = st.next(); }
    if (c >= 0 && ((c | 0x20) >= 'a' && (c | 0x20) <= 'z')) {
        // inf / infinity / nan, case-insensitive, nothing else.
        uint64_t w = 0;
        int n = 0;
        while (c >= 0 && n < 9) {
            w |= (uint64_t)(c | 0x20) << (8 * n);
            ++n;
            c = st.next();
        }
        if (c >= 0) return r;
        const uint64_t INF = 0x666E69ULL, NAN_ = 0x6E616EULL, INFINITY_ = 0x7974696E69666E69ULL;
        uint64_t bits;
        if (n == 3 && w == INF) bits = F64_INF_BITS;
        else if (n == 8 && w == INFINITY_) bits = F64_INF_BITS;
        else if (n == 3 && w == NAN_) bits = 0x7FF8000000000000ULL;
        else return r;
        r.ok = true;
        r.value = bits_to_f64(bits | ((uint64_t)neg << 63));
        return r;
    }
    uint64_t mant = 0;
    int sig = 0;            // significant digits folded into mant
    int int_digits = 0, frac_digits = 0;
    int32_t exp10 = 0;
    bool dropped = false;   // nonzero significant digits beyond 19 were dropped
    bool dot = false;
    while (is_digit(c)) {
        ++int_digits;
        int dg = c - '0';
        if (sig < 19) {
            if (mant != 0 || dg != 0) { mant = mant * 10 + (uint64_t)dg; ++sig; }
        } else {
            ++exp10;
            if (dg) dropped = true;
        }
        c = st.next();
    }
    if (c == '.') {
        dot = true;
        c = st.next();
        while (is_digit(c)) {
            ++frac_digits;
            int dg = c - '0';
            if (sig < 19) {
                if (mant != 0 || dg != 0) { mant = mant * 10 + (uint64_t)dg; ++sig; }
                --exp10;
            } else if (dg) {
                dropped = true;
            }
            c = st.next();
        }
    }
    if (int_digits + frac_digits == 0) return r;
    bool has_exp = false;
    if (c == 'e' || c == 'E') {
        has_exp = true;
        c = st.next();
        bool eneg = false;
        if (c == '+' || c == '-') { eneg = c == '-'; c = st.next(); }
        if (!is_digit(c)) return r;
        int32_t e = 0;
        while (is_digit(c)) {
            if (e < 100000) e = e * 10 + (c - '0');
            c = st.next();
        }
        exp10 += eneg ? -e : e;
    }
    if (c >= 0) return r;
    return f64_decide(neg, plus, dot, has_exp, int_digits, frac_digits, mant, exp10, dropped);
}

// The value of a cell that matched the decimal grammar: `mant` holds its
// first 19 significant digits (`dropped` when nonzero digits followed),
// scaled by 10^exp10. Shared by the byte stream above and the register
// parser in num_swar.h, so both decide every cell the same way.
FX_HD F64Result f64_decide(bool neg, bool plus, bool dot, bool has_exp, int int_digits, int frac_digits,
                           uint64_t mant, int32_t exp10, bool dropped) {
    F64Result r{true, false, false, 0, 0.0};
    // Fixed `[-]d{1,3}.d{6}` shape (csv_geo.rs:1414-1453): m / 1e6, sign ORed.
    if (!plus && !has_exp && dot && frac_digits == 6 && int_digits >= 1 && int_digits <= 3) {
        double v = (double)mant / 1000000.0;
        r.value = bits_to_f64(f64_to_bits(v) | ((uint64_t)neg << 63));
        r.fixed6 = true;
        r.mantissa = neg ? -(int64_t)mant : (int64_t)mant;
        return r;
    }
    const uint64_t sign = (uint64_t)neg << 63;
    if (mant == 0) {
        r.value = bits_to_f64(sign);
        return r;
    }
    // Clinger (1990): exact mantissa and exact power of ten, one rounding.
    if (!dropped && mant <= (1ULL << 53) && exp10 >= -22 && exp10 <= 22) {
        double v = exp10 < 0 ? (double)mant / pow10_exact(-exp10) : (double)mant * pow10_exact(exp10);
        r.value = bits_to_f64(f64_to_bits(v) | sign);
        return r;
    }
    ElOut e = eisel_lemire(mant, exp10);
    if (e.ok && dropped) {
        // The true value lies strictly between mant and mant + 1 (times 10^q).
        ElOut e2 = eisel_lemire(mant + 1, exp10);
        if (!e2.ok || e2.bits != e.bits) e.ok = false;
    }
    if (!e.ok) {
        r.deferred = true;
        return r;
    }
    r.value = bits_to_f64(e.bits | sign);
    return r;
}

// The fixed `[-]d{1,3}.d{6}` shape (csv_geo.rs:1414-1453) from its last 8
// bytes `x` (low byte first) and the 0..3 bytes before them: the '.' and the
// digits are checked eight bytes at a time and the value is m / 1e6 with the
// sign bit ORed in, exactly as parse_f64_stream's fixed-6 path. False for
// any other shape (the general parser then decides).
FX_HD bool fixed6_tail(uint64_t x, const uint8_t* pre, uint32_t npre, F64Result& r) {
    if (npre > 3 || ((x >> 8) & 0xFFu) != '.') return false;
    const uint64_t y = (x & ~0xFF00ULL) | 0x3000ULL;    // '.' replaced by '0'
    if ((y & 0xF0F0F0F0F0F0F0F0ULL) != 0x3030303030303030ULL) return false;
    if (((y & 0x0F0F0F0F0F0F0F0FULL) + 0x0606060606060606ULL) & 0xF0F0F0F0F0F0F0F0ULL) return false;
    uint32_t frac = 0;
    for (int k = 2; k < 8; ++k) frac = frac * 10u + (uint32_t)((x >> (8 * k)) & 15u);
    uint32_t p = 0;
    bool neg = false;
    if (p < npre && pre[p] == '-') { neg = true; ++p; }
    if (npre - p > 2) return false;
    uint32_t iv = 0;
    for (; p < npre; ++p) {
        const uint32_t c = (uint32_t)pre[p] - '0';
        if (c > 9u) return false;
        iv = iv * 10u + c;
    }
    iv = iv * 10u + (uint32_t)(x & 15u);
    const uint64_t m = (uint64_t)iv * 1000000u + frac;
    const double v = (double)m / 1000000.0;
    r.ok = true;
    r.fixed6 = true;
    r.deferred = false;
    r.mantissa = neg ? -(int64_t)m : (int64_t)m;
    r.value = bits_to_f64(f64_to_bits(v) | ((uint64_t)neg << 63));
    return true;
}

// ---------------------------------------------------------------------------
// Field helpers with the dialect's cleaning (codegen.rs:6060-6085: numeric
// cells go through `clean` only when they start with the quote byte).
// ---------------------------------------------------------------------------
struct CellDialect {
    int32_t quote;    // -1 none
    int32_t escape;   // -1 none (doubled quotes)
};

template <class Src>
FX_HD bool quoted_both(const Src& s, uint32_t from, uint32_t to, uint8_t q) {
    return to - from >= 2 && s.at(from) == q && s.at(to - 1) == q;
}

// Calls fn(stream) with the stream that cleans the cell as the dialect
// says; numeric cells are cleaned only when they start with the quote byte.
template <class Src, class Fn>
FX_HD auto with_cell_stream(const Src& s, uint32_t from, uint32_t to, CellDialect d, Fn&& fn) {
    if (d.quote >= 0 && to > from && s.at(from) == (uint8_t)d.quote) {
        uint8_t q = (uint8_t)d.quote;
        bool both = quoted_both(s, from, to, q);
        if (d.escape < 0) {
            if (both) return fn(DoubledStream<Src>{&s, from + 1, to - 1, q});
            return fn(RawStream<Src>{&s, from, to});
        }
        if (both) return fn(BackslashStream<Src>{&s, from + 1, to - 1, (uint8_t)d.escape});
        return fn(BackslashStream<Src>{&s, from, to, (uint8_t)d.escape});
    }
    return fn(RawStream<Src>{&s, from, to});
}

// Fast paths only: a result with `deferred` set is a valid cell whose value
// comes from f64_field_exact_bits.
template <class Src>
FX_HD F64Result parse_f64_field(const Src& s, uint32_t from, uint32_t to, CellDialect d) {
    return with_cell_stream(s, from, to, d, [](auto st) { return parse_f64_stream(st); });
}

// Signed IEEE bits of a valid numeric stream by the exact big-decimal path.
template <class St>
FX_HD uint64_t exact_f64_bits(St st, BigDec& b) {
    St first = st;
    const uint64_t sign = (uint64_t)(first.next() == '-') << 63;
    return slow_f64_bits_into(b, st) | sign;
}

template <class Src>
FX_HD uint64_t f64_field_exact_bits(const Src& s, uint32_t from, uint32_t to, CellDialect d, BigDec& b) {
    return with_cell_stream(s, from, to, d, [&](auto st) { return exact_f64_bits(st, b); });
}

template <class Src>
FX_HD bool parse_i64_field(const Src& s, uint32_t from, uint32_t to, CellDialect d, int64_t& out) {
    if (d.quote >= 0 && to > from && s.at(from) == (uint8_t)d.quote) {
        uint8_t q = (uint8_t)d.quote;
        bool both = quoted_both(s, from, to, q);
        if (d.escape < 0) {
            if (both) return parse_i64_stream(DoubledStream<Src>{&s, from + 1, to - 1, q}, out);
            return parse_i64_stream(RawStream<Src>{&s, from, to}, out);
        }
        if (both) return parse_i64_stream(BackslashStream<Src>{&s, from + 1, to - 1, (uint8_t)d.escape}, out);
        return parse_i64_stream(BackslashStream<Src>{&s, from, to, (uint8_t)d.escape}, out);
    }
    return parse_i64_stream(RawStream<Src>{&s, from, to}, out);
}

// Cleaned cell for string columns: visit(byte) for each output byte; returns
// the cleaned length. No quote: verbatim. Doubled: strip + collapse when
// quoted at both ends, else verbatim. Backslash: strip when quoted at both
// ends, then unescape.
template <class Src, class Visit>
FX_HD uint32_t clean_cell(const Src& s, uint32_t from, uint32_t to, CellDialect d, Visit&& visit) {
    uint32_t n = 0;
    if (d.quote < 0) {
        for (uint32_t p = from; p < to; ++p) { visit(s.at(p)); ++n; }
        return n;
    }
    uint8_t q = (uint8_t)d.quote;
    bool both = quoted_both(s, from, to, q);
    if (d.escape < 0) {
        if (!both) {
            for (uint32_t p = from; p < to; ++p) { visit(s.at(p)); ++n; }
            return n;
        }
        DoubledStream<Src> st{&s, from + 1, to - 1, q};
        for (int c = st.next(); c >= 0; c = st.next()) { visit((uint8_t)c); ++n; }
        return n;
    }
    BackslashStream<Src> st{&s, both ? from + 1 : from, both ? to - 1 : to, (uint8_t)d.escape};
    for (int c = st.next(); c >= 0; c = st.next()) { visit((uint8_t)c); ++n; }
    return n;
}

// field_byte_len (codegen.rs:6391-6425): doubled-quote collapse whenever the
// dialect has a quote byte, raw length otherwise.
template <class Src>
FX_HD uint32_t field_byte_len(const Src& s, uint32_t from, uint32_t to, int32_t quote) {
    if (quote < 0) return to - from;
    uint8_t q = (uint8_t)quote;
    if (!quoted_both(s, from, to, q)) return to - from;
    DoubledStream<Src> st{&s, from + 1, to - 1, q};
    uint32_t n = 0;
    while (st.next() >= 0) ++n;
    return n;
}

// Complete parse with deferred cells resolved in `b` (host sweeps, fixups).
template <class Src>
FX_HD F64Result parse_f64_field_exact(const Src& s, uint32_t from, uint32_t to, CellDialect d, BigDec& b) {
    F64Result r = parse_f64_field(s, from, to, d);
    if (r.ok && r.deferred) {
        r.value = bits_to_f64(f64_field_exact_bits(s, from, to, d, b));
        r.deferred = false;
    }
    return r;
}

}  // namespace fx
