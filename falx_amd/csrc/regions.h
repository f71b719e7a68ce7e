// regions.h — quoted and comment regions of one 64-byte block as bit masks,
// written once for the gfx950 kernels and for host-side checks.
//
// A block is given as three disjoint masks over its bytes: q (quote bytes),
// s (comment bytes at a line start) and n (newlines), plus the region state
// it is entered in: 0 outside, 1 inside quotes, 2 inside a comment. A quoted
// region runs from its opening quote up to, not including, the closing
// quote; a comment from its comment byte up to, not including, the newline
// that ends it. Outside quotes and comments, a quote opens a quoted region
// and a comment byte a comment; inside quotes only a quote counts, inside a
// comment only a newline. This is the three-state resolver of the
// reference's csv_hash dialect (csv_hash.rs:2149-2220; interp.rs:163-203).
#pragma once

#include <stdint.h>

#ifdef __HIPCC__
#define FX_RG __host__ __device__ __forceinline__
#else
#define FX_RG inline
#endif

namespace fx {

// Bit i = parity of the set bits at positions <= i.
FX_RG uint64_t prefix_xor(uint64_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    x ^= x << 16;
    x ^= x << 32;
    return x;
}

// Bits [from, to), to <= 64.
FX_RG uint64_t range_mask(uint32_t from, uint32_t to) {
    uint64_t hi = to >= 64 ? ~0ULL : ((1ULL << to) - 1);
    return hi & ~((1ULL << from) - 1);
}

FX_RG uint32_t low_bit(uint64_t m) { return (uint32_t)__builtin_ctzll(m); }

// Bits strictly above position p.
FX_RG uint64_t above(uint32_t p) { return p >= 63 ? 0ULL : (~0ULL << (p + 1)); }

// The block's inert bytes (inside a quoted region or a comment); `state`
// goes from the entry state to the exit state. Work is per comment, not per
// byte: the quoted regions between comments come from one prefix parity,
// the first comment byte outside them ends that stretch, and the comment
// ends at the next newline.
FX_RG uint64_t resolve_regions(uint64_t q, uint64_t s, uint64_t n, uint32_t& state) {
    if (state == 0 && (q | s) == 0) return 0;
    uint64_t inert = 0, from = ~0ULL;   // bytes not yet resolved
    uint32_t st = state;
    if (st == 2) {
        if (!n) return ~0ULL;   // the comment runs through the block
        const uint32_t e = low_bit(n);
        inert = range_mask(0, e);
        from = above(e);
        st = 0;
    }
    while (from) {
        const uint64_t inside = (prefix_xor(q & from) ^ (st ? ~0ULL : 0ULL)) & from;
        const uint64_t cs = s & from & ~inside;
        if (!cs) {
            inert |= inside;
            state = (uint32_t)(inside >> 63);
            return inert;
        }
        const uint32_t p = low_bit(cs);
        inert |= inside & range_mask(0, p);
        const uint64_t ends = n & above(p);
        if (!ends) {
            state = 2;
            return inert | range_mask(p, 64);
        }
        const uint32_t e = low_bit(ends);
        inert |= range_mask(p, e);
        from = above(e);
        st = 0;
    }
    state = 0;   // the block's last byte ends a comment
    return inert;
}

}  // namespace fx
