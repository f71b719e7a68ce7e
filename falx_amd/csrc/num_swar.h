// num_swar.h — typed numeric cells decided from registers.
//
// A cell is held as the 4*NW bytes that end at its last byte, in NW
// little-endian words (NW = 6 for cells of up to 24 bytes, 3 for up to 12):
// window byte i is byte (i & 3) of w[i >> 2], the cell is window bytes
// [4 NW - n, 4 NW) and the bytes below it are whatever preceded it. Then:
//  - its non-digit bytes come out of one SWAR test per word as a bit mask,
//    and the first four of them are fetched together;
//  - the decimal grammar is checked on those (sign, '.', 'e'/'E', exponent
//    sign);
//  - an exponent is read from the last word, then cut off by a funnel
//    shift that moves the significand's last digit to the window's end;
//  - the '.' is squeezed out by a second funnel shift of the bytes below it;
//  - the digits, now right-aligned, are folded four at a time: per 32-bit
//    word two multiply-adds (the byte-pair then pair-pair combination of
//    Lemire's "Quickly parsing eight digits", 2018), then 10^4 steps.
// Straight-line code: no loop over bytes and no branch per byte, so a
// wavefront's lanes stay together whatever their cells' lengths and
// columns (an i64 and an f64 cell take the same instructions).
//
// A cell is decided exactly as the byte-stream parsers in cells.h decide it
// (the value comes from the same f64_decide), or it is reported undecided:
// longer than the window, starting with the quote byte, more than 19
// digits, more than four non-digits, an exponent of more than four digits,
// or any byte outside `[+-]?d*.?d*([eE][+-]?d+)?` (inf, nan, stray
// characters). The stream parser then decides those.
#pragma once

#include "cells.h"

namespace fx {

// The NW words of bytes that end at a cell's last byte (the cell is window
// bytes [4 NW - n, 4 NW)). Six words take any cell the parser decides; three
// are enough for cells of up to 12 bytes, at half the work.
template <int NW>
struct Win {
    static constexpr uint32_t BYTES = 4 * NW;
    uint32_t w[NW];
};
using Win24 = Win<6>;
using Win12 = Win<3>;
constexpr uint32_t NUM_WIN = Win24::BYTES;

// Bit j set when byte j of w is not an ASCII digit.
FX_HD uint32_t nondigit4(uint32_t w) {
    const uint32_t x = w ^ 0x30303030u;   // digits become 0..9
    // (x & 0x7F) + 0x76 sets bit 7 exactly when the low seven bits are >= 10;
    // no byte carries into the next one.
    const uint32_t f = ((((x & 0x7F7F7F7Fu) + 0x76767676u) | x) & 0x80808080u) >> 7;
    // Gather bits 0, 8, 16, 24 into bits 24..27: the four partial products
    // land on distinct bits, so nothing carries.
    return (f * 0x01020408u) >> 24;
}

// Four digit values (0..9 per byte, the most significant in the low byte) as
// one number 0..9999.
FX_HD uint32_t digits4(uint32_t d) {
    const uint32_t t = (d * 10u + (d >> 8)) & 0x00FF00FFu;   // two 2-digit numbers
    return (t & 0xFFu) * 100u + (t >> 16);
}

// m / 10^6 correctly rounded, for m < 10^9, without a divide: the product
// q = m * RN(10^-6), then one Markstein correction q + RN(m - 10^6 q) * RN(10^-6)
// with the remainder exact by FMA. Equal to the IEEE quotient for every
// m < 10^9 (tests/native/fixed6_div_check.cpp runs them all).
FX_HD double div_1e6(uint32_t m) {
    const double a = (double)m;
    const double q = a * 1e-6;
    const double r = __builtin_fma(-q, 1e6, a);
    return __builtin_fma(r, 1e-6, q);
}

// Bit 7 of a byte of w set when the byte is not an ASCII digit.
FX_HD uint32_t nondigit_bits(uint32_t w) {
    const uint32_t x = w ^ 0x30303030u;
    return (((x & 0x7F7F7F7Fu) + 0x76767676u) | x) & 0x80808080u;
}

// The fixed `[-]d{1,3}.d{6}` shape (csv_geo.rs:1414-1453) of an n-byte cell
// from the 12 bytes that end at its last byte (w0 w1 w2, little-endian):
// w1 = [d '.' f1 f2], w2 = [f3 f4 f5 f6], and the n - 8 bytes before them
// are the top of w0. Decides exactly the cells fixed6_tail decides, with
// the same value, flag and mantissa; false for any other shape. Straight-
// line code: every condition is evaluated and the value computed whatever
// the shape, so a wavefront's lanes never branch apart here.
FX_HD bool fixed6_win(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t n, F64Result& r) {
    const uint32_t npre = n - 8u;   // 0..3 for a fixed-6 cell
    // the cell's first npre bytes, first one lowest (npre <= 3 here, so the shift is 8..32)
    const uint32_t pre = npre - 1u < 3u ? w0 >> (32u - 8u * npre) : 0u;
    const bool neg = npre - 1u < 3u && (pre & 0xFFu) == '-';
    const uint32_t dg = neg ? pre >> 8 : pre;
    const uint32_t nd = npre - (neg ? 1u : 0u);
    const uint32_t p0 = (dg & 0xFFu) - '0', p1 = ((dg >> 8) & 0xFFu) - '0';
    const bool ok = npre <= 3u && (w1 & 0xFF00u) == 0x2E00u &&
                    (nondigit_bits((w1 & 0xFFFF00FFu) | 0x3000u) | nondigit_bits(w2)) == 0u && nd <= 2u &&
                    (nd < 1u || p0 <= 9u) && (nd < 2u || p1 <= 9u);
    const uint32_t d = (w1 & 0xFFu) - '0';
    const uint32_t iv = (nd == 2u ? p0 * 100u + p1 * 10u : (nd == 1u ? p0 * 10u : 0u)) + d;
    const uint32_t f12 = (((w1 >> 16) & 0xFFu) - '0') * 10u + ((w1 >> 24) - '0');
    const uint32_t m = (iv & 0x3FFu) * 1000000u + f12 * 10000u + digits4(w2 ^ 0x30303030u);
    r.ok = ok;
    r.fixed6 = ok;
    r.deferred = false;
    r.mantissa = neg ? -(int64_t)m : (int64_t)m;
    r.value = bits_to_f64(f64_to_bits(div_1e6(m)) | ((uint64_t)neg << 63));
    return ok;
}

// 0x0F in each byte of a word that is at or above `below` bytes into it.
FX_HD uint32_t digit_mask(int32_t below) {
    return below <= 0 ? 0x0F0F0F0Fu : below >= 4 ? 0u : 0x0F0F0F0Fu << (8 * below);
}

// The number written by window bytes [4 NW - nd, 4 NW), all digits, nd <= 19
// (so at most the last five words hold digits).
template <int NW>
FX_HD uint64_t fold_digits(const Win<NW>& W, uint32_t nd) {
    constexpr int K0 = NW > 5 ? NW - 5 : 0;
    const int32_t rs = (int32_t)Win<NW>::BYTES - (int32_t)nd;
    uint32_t g[NW];
#pragma unroll
    for (int k = K0; k < NW; ++k) g[k] = digits4(W.w[k] & digit_mask(rs - 4 * k));
    uint64_t v = g[K0] * 10000u + g[K0 + 1];   // < 10^8
#pragma unroll
    for (int k = K0 + 2; k < NW; ++k) v = v * 10000u + g[k];
    return v;
}

// Window bytes moved up by s in [0, 8) positions: byte i takes byte i - s.
template <int NW>
FX_HD Win<NW> shift_up(const Win<NW>& W, uint32_t s) {
    Win<NW> R;
    const bool ws = s >= 4;
    const uint32_t rs = 32u - 8u * (s & 3u);
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t a0 = W.w[k];
        const uint32_t a1 = k >= 1 ? W.w[k - 1] : 0u;
        const uint32_t a2 = k >= 2 ? W.w[k - 2] : 0u;
        const uint32_t hi = ws ? a1 : a0, lo = ws ? a2 : a1;
        R.w[k] = (uint32_t)((((uint64_t)hi << 32) | lo) >> rs);
    }
    return R;
}

// Removes window byte dp (the '.'): the bytes below it move up by one.
template <int NW>
FX_HD Win<NW> squeeze_out(const Win<NW>& W, int32_t dp) {
    Win<NW> R;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const uint32_t up = (W.w[k] << 8) | (k ? W.w[k - 1] >> 24 : 0u);
        const int32_t r = dp + 1 - 4 * k;   // bytes of word k at or below the '.'
        const uint32_t keep = r <= 0 ? ~0u : r >= 4 ? 0u : ~0u << (8 * r);
        R.w[k] = (W.w[k] & keep) | (up & ~keep);
    }
    return R;
}

FX_HD bool is_sign(uint32_t b) { return b == '-' || b == '+'; }

// Decides a cell of n bytes held in W as f64 (as_f64) or as i64. Returns
// false when undecided (see above); otherwise sets fr or (iok, iv) exactly
// as parse_f64_field / parse_i64_field would. quote < 0: no quote byte;
// the quote byte must not be an ASCII digit (the caller checks).
// at(p) returns window byte p, p in [4 NW - n, 4 NW): the up to four
// non-digit bytes are read through it (from LDS in the kernels, all four
// reads issued together) rather than picked out of W by a varying index,
// which would put W in scratch memory.
template <int NW, class ByteAt>
FX_HD bool num_swar(const Win<NW>& W, uint32_t n, bool as_f64, int32_t quote, ByteAt&& at, F64Result& fr,
                    bool& iok, int64_t& iv) {
    constexpr uint32_t B = Win<NW>::BYTES;
    fr = F64Result{false, false, false, 0, 0.0};
    iok = false;
    iv = 0;
    if (n == 0) return true;
    if (n > B) return false;
    const uint32_t start = B - n;
    uint32_t nd = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) nd |= nondigit4(W.w[k]) << (4 * k);
    nd &= ((1u << B) - 1u) & ~((1u << start) - 1u);
    // The first four non-digits, in order (positions past the last repeat
    // `start`, and are ignored).
    uint32_t pos[4], byte[4];
    bool has[4];
    uint32_t m = nd;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        has[i] = m != 0;
        pos[i] = has[i] ? (uint32_t)__builtin_ctz(m) : start;
        m &= m - 1u;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) byte[i] = at(pos[i]);
    const bool lead = has[0] && pos[0] == start;   // the cell starts with a non-digit
    if (lead && quote >= 0 && byte[0] == (uint32_t)quote) return false;
    const bool sgn = lead && is_sign(byte[0]);
    const bool neg = sgn && byte[0] == '-', plus = sgn && byte[0] == '+';

    if (!as_f64) {
        const uint32_t digits = n - (uint32_t)sgn;
        if ((sgn ? has[1] : has[0]) || digits == 0) return true;
        if (digits > 19) return false;
        const uint64_t v = fold_digits(W, digits);
        if (v > 0x7FFFFFFFFFFFFFFFULL + (uint64_t)neg) return true;
        iok = true;
        iv = neg ? (int64_t)(0 - v) : (int64_t)v;
        return true;
    }

    // After the sign: at most '.', 'e'/'E' and the exponent's sign, in order.
    if (m != 0) return false;
    bool dot = false, has_exp = false, esign = false, eneg = false, bad = false;
    int32_t dotpos = -1, epos = 0;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        if (!has[i] || (i == 0 && sgn)) continue;
        const uint32_t b = byte[i];
        const int32_t p = (int32_t)pos[i];
        if (b == '.' && !dot && !has_exp) {
            dot = true;
            dotpos = p;
        } else if ((b | 0x20u) == 'e' && !has_exp) {
            has_exp = true;
            epos = p;
        } else if (is_sign(b) && has_exp && !esign && p == epos + 1) {
            esign = true;
            eneg = b == '-';
        } else {
            bad = true;
        }
    }
    if (bad) return false;
    const int32_t mstart = (int32_t)start + (int32_t)sgn;
    const int32_t mend = has_exp ? epos : (int32_t)B;
    const int32_t int_digits = (dot ? dotpos : mend) - mstart;
    const int32_t frac_digits = dot ? mend - dotpos - 1 : 0;
    if (int_digits + frac_digits == 0) return true;
    if (int_digits + frac_digits > 19) return false;
    Win<NW> M = W;
    int32_t exp10 = -frac_digits;
    if (has_exp) {
        const int32_t es = epos + 1 + (int32_t)esign;   // first exponent digit
        const int32_t ed = (int32_t)B - es;
        if (ed == 0) return true;
        if (ed > 4) return false;
        const int32_t e = (int32_t)digits4(W.w[NW - 1] & digit_mask(es - (int32_t)B + 4));
        exp10 += eneg ? -e : e;
        M = shift_up(W, (uint32_t)((int32_t)B - epos));   // 2..6: the significand now ends the window
        dotpos += (int32_t)B - epos;
    }
    if (dot) M = squeeze_out(M, dotpos);
    const uint64_t mant = fold_digits(M, (uint32_t)(int_digits + frac_digits));
    fr = f64_decide(neg, plus, dot, has_exp, int_digits, frac_digits, mant, exp10, false);
    return true;
}

}  // namespace fx
