// inflate.h — BGZF blocks inflated on gfx950 (bgzf.rs:194-290; DEFLATE,
// RFC 1951), one wave per block, so compressed input is uploaded and the
// parsers read the inflated bytes in HBM.
//
// A BGZF block is an independent raw-DEFLATE stream of at most 64 KiB of
// output whose size (ISIZE) is known before inflation, so every block owns
// a disjoint slot of the output and blocks inflate in parallel. Within a
// block the Huffman decoding is serial: every lane of the wave runs the same
// decode on the same bits (wave-uniform values, scalar where the compiler
// can), lane 0 stores literal bytes, and the whole wave copies the bytes of
// each back-reference (the source of byte i of a copy of distance d is
// byte i mod d of the window before it, so overlapping copies need no
// sequencing). Decoding tables live in LDS: a 10-bit primary table for
// literal/length codes and a 9-bit one for distance codes; the rare longer
// codes are decoded canonically from the per-length counts.
//
// Symbols are decoded a window at a time: with 96 or more bits buffered,
// lane i looks up both tables at bit offset i (one LDS read each for the
// whole wave), and the chain of symbols is then followed from offset 0 by
// register reads (v_readlane at the running offset): a literal advances by
// its code length and is staged in lane k of a register, a length code
// reads its extra bits and the distance entry at the offset after them. So
// one window costs two LDS round trips however many symbols it holds. A
// code longer than the primary tables (or a length whose distance would
// start past the window) ends the window early; the longer codes are
// decoded canonically from the per-length counts.
//
// Neither the input nor recent output is read from HBM: the payload streams
// through a 512-byte LDS ring (256 bytes loaded by the wave at a time), and
// the output through a 2 KiB LDS ring flushed to the block's slot in 1 KiB
// steps. A back-reference reads the ring for bytes still in it and the slot
// (already flushed, behind a fence) for older ones. About 9 KiB of LDS and
// 70 VGPRs per block let 20 or more blocks share a CU: the decode of one
// block is a serial chain of dependent instructions, so the throughput
// comes from the number of blocks in flight.
//
// Every read stays inside the block's payload (missing bits read as zero and
// are an error if used), every write inside its output slot, every
// back-reference inside the bytes already written. A malformed block reports
// its index (the smallest one) and its slot is left incomplete.
#pragma once

namespace fx {

// falx_bgzf_block (include/falx_gpu.h).
struct InfBlock {
    unsigned long long payload;   // offset of the raw-DEFLATE payload in the compressed bytes
    uint32_t payload_len;
    uint32_t isize;               // inflated size
    unsigned long long out;       // offset of the block's output slot
};

constexpr int INF_WG = 64;
constexpr int INF_LB = 10;   // literal/length primary table bits
constexpr int INF_DB = 9;    // distance primary table bits
constexpr uint32_t INF_NONE = 0xFFFFFFFFu;
constexpr uint32_t INF_IN = 512;       // input ring
constexpr uint32_t INF_CHUNK = 256;    // input loaded per step: 4 bytes per lane
constexpr uint32_t INF_OUT = 2048;     // output ring
constexpr uint32_t INF_FLUSH = 1024;   // output flushed once this much is pending

// RFC 1951 3.2.5: base values and extra bits of the length codes 257..285
// and distance codes 0..29.
__constant__ uint16_t kLenBase[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                      31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
__constant__ uint16_t kDistBase[30] = {1,   2,   3,   4,   5,   7,    9,    13,   17,   25,   33,   49,    65,    97,    129,
                                       193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
// Order of the code-length code lengths (RFC 1951 3.2.7).
__constant__ uint8_t kClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// Primary table entries (32 bits): the symbol's kind (bits 28-31: 0 = a
// code longer than the table, 1 = literal, 2 = end of block, 3 = length or
// distance, 4 = a symbol the format forbids), its code length (bits 24-27),
// its extra bits (bits 16-23) and its value (bits 0-15: the literal byte,
// or the base length or distance).
enum : uint32_t { INF_K_LONG = 0, INF_K_LIT = 1, INF_K_EOB = 2, INF_K_BASE = 3, INF_K_BAD = 4 };
DEV uint32_t inf_entry(uint32_t kind, uint32_t len, uint32_t extra, uint32_t val) {
    return (kind << 28) | (len << 24) | (extra << 16) | val;
}

struct InfLds {
    uint32_t lt[1 << INF_LB];
    uint32_t dt[1 << INF_DB];
    uint16_t lcount[16], dcount[16], ccount[16], offs[16];
    uint16_t lsym[288], dsym[32], csym[19];
    uint8_t lens[288 + 32];
    uint8_t cl[19];             // code-length code lengths
    uint32_t ok;
    uint16_t lbase[29], dbase[30];      // kLenBase, kDistBase (LDS: indexed per symbol)
    uint8_t lextra[29], dextra[30];
    __align__(4) uint8_t in[INF_IN];   // payload ring
    uint8_t outb[INF_OUT];      // output ring
};

// Canonical code of `n` symbols with lengths `len`: per-length counts and
// the symbols in code order (count[0] counts unused symbols). Returns false
// for an over-subscribed set, and for an incomplete one unless `single` and
// it is one code of length 1 (literal/length and distance codes; an empty
// distance set is allowed too: decoding from it fails). Lane 0 only;
// `offs` is 16 words of LDS.
DEV bool inf_canon(const uint8_t* len, uint32_t n, uint16_t* count, uint16_t* sym, uint16_t* offs, bool single) {
    for (int l = 0; l < 16; ++l) count[l] = 0;
    for (uint32_t s = 0; s < n; ++s) ++count[len[s]];
    if (count[0] == n) return single;
    int left = 1;
    for (int l = 1; l < 16; ++l) {
        left <<= 1;
        left -= count[l];
        if (left < 0) return false;
    }
    if (left > 0 && !(single && n - count[0] == 1 && count[1] == 1)) return false;
    offs[1] = 0;
    for (int l = 1; l < 15; ++l) offs[l + 1] = offs[l] + count[l];
    for (uint32_t s = 0; s < n; ++s)
        if (len[s]) sym[offs[len[s]]++] = (uint16_t)s;
    return true;
}

// The primary table of a canonical code: every lane fills the entries of
// the symbols j = lane, lane + 64, ... in code order (`dist` selects the
// distance code's entries).
DEV void inf_fill(uint32_t* table, int bits, const uint16_t* count, const uint16_t* sym, bool dist) {
    const uint32_t lane = __lane_id();
    for (uint32_t i = lane; i < (1u << bits); i += INF_WG) table[i] = 0;
    __syncthreads();
    uint32_t nsym = 0;
    for (int l = 1; l < 16; ++l) nsym += count[l];
    for (uint32_t j = lane; j < nsym; j += INF_WG) {
        // code of the j-th symbol in code order
        uint32_t code = 0, first = 0, idx = 0;
        int l = 1;
        for (; l < 16; ++l) {
            if (j < idx + count[l]) break;
            idx += count[l];
            first = (first + count[l]) << 1;
        }
        if (l > bits) continue;
        code = first + (j - idx);
        uint32_t rev = 0;
        for (int b = 0; b < l; ++b) rev |= ((code >> b) & 1u) << (l - 1 - b);
        const uint32_t s = sym[j];
        uint32_t e;
        if (dist) e = s < 30 ? inf_entry(INF_K_BASE, l, kDistExtra[s], kDistBase[s]) : inf_entry(INF_K_BAD, l, 0, 0);
        else if (s < 256) e = inf_entry(INF_K_LIT, l, 0, s);
        else if (s == 256) e = inf_entry(INF_K_EOB, l, 0, 0);
        else if (s < 286) e = inf_entry(INF_K_BASE, l, kLenExtra[s - 257], kLenBase[s - 257]);
        else e = inf_entry(INF_K_BAD, l, 0, 0);
        for (uint32_t k = 0; k < (1u << (bits - l)); ++k) table[rev | (k << l)] = e;
    }
    __syncthreads();
}

// The bit stream: up to 128 buffered bits (b0 the first 64), refilled from
// the payload ring 32 bits at a time; every lane holds the same state.
struct InfBits {
    const uint8_t* src;
    uint8_t* ring;         // LDS, INF_IN bytes: payload byte p at ring[p % INF_IN] for p in [hi - INF_IN, hi)
    uint32_t n, pos, hi;   // payload bytes, bytes taken into the buffer, bytes loaded into the ring
    uint64_t b0, b1;
    uint32_t bc;           // valid bits in (b0, b1)
    uint32_t over;         // zero bytes taken past the payload

    // The next INF_CHUNK payload bytes into the ring (every lane, 4 each).
    DEV void load() {
        const uint32_t base = hi + 4u * __lane_id();
#pragma unroll
        for (uint32_t k = 0; k < 4; ++k) {
            const uint32_t p = base + k;
            ring[p % INF_IN] = p < n ? src[p] : (uint8_t)0;
        }
        hi += INF_CHUNK;
        __syncthreads();
    }
    // 32 more bits; missing bytes read as zero. The four bytes come from one
    // aligned ring word when pos is a multiple of 4 (always, except after a
    // stored block).
    DEV void add32() {
        if (pos + 4u > hi && hi < n) load();
        uint32_t w;
        if ((pos & 3u) == 0 && pos + 4u <= n) {
            w = reinterpret_cast<const uint32_t*>(ring)[(pos % INF_IN) >> 2];
        } else {
            w = 0;
#pragma unroll
            for (uint32_t k = 0; k < 4; ++k) {
                const uint32_t p = pos + k;
                w |= (p < n ? (uint32_t)ring[p % INF_IN] : 0u) << (8u * k);
            }
        }
        if (bc < 64) {
            b0 |= (uint64_t)w << bc;
            if (bc > 32) b1 |= (uint64_t)w >> (64u - bc);
        } else {
            b1 |= (uint64_t)w << (bc - 64u);
        }
        const uint32_t real = pos < n ? min(4u, n - pos) : 0u;
        over += 4u - real;
        pos += 4;
        bc += 32;
    }
    DEV void fill() {
        while (bc <= 96) add32();
    }
    // k <= 32 bits at offset off (off + k <= 128).
    DEV uint32_t at(uint32_t off, uint32_t k) const {
        uint64_t v;
        if (off == 0) v = b0;
        else if (off < 64) v = (b0 >> off) | (b1 << (64u - off));
        else v = b1 >> (off - 64u);
        return (uint32_t)(v & ((1ull << k) - 1ull));
    }
    DEV void drop(uint32_t k) {   // k <= 128
        if (k == 0) return;
        if (k >= 64) {
            b0 = k == 64 ? b1 : b1 >> (k - 64u);
            b1 = 0;
        } else {
            b0 = (b0 >> k) | (b1 << (64u - k));
            b1 >>= k;
        }
        bc -= k;
    }
    DEV uint32_t take(uint32_t k) {
        if (bc < 32) fill();
        const uint32_t v = at(0, k);
        drop(k);
        return v;
    }
    // Bits consumed beyond the payload.
    DEV bool overrun() const { return 8ull * (uint64_t)over > (uint64_t)bc; }
};

// One symbol of a canonical code, serially and bit by bit from the
// per-length counts (codes are sent most significant bit first): the codes
// longer than the primary tables. INF_NONE for a bit pattern that is not a
// code.
DEV uint32_t inf_decode(InfBits& r, const uint16_t* count, const uint16_t* sym) {
    if (r.bc < 32) r.fill();
    uint32_t code = 0, first = 0, idx = 0;
    for (int l = 1; l < 16; ++l) {
        code |= (uint32_t)((r.b0 >> (l - 1)) & 1ull);
        const uint32_t c = count[l];
        if (code - first < c) {
            r.drop((uint32_t)l);
            return sym[idx + (code - first)];
        }
        idx += c;
        first = (first + c) << 1;
        code <<= 1;
    }
    return INF_NONE;
}

__global__ __launch_bounds__(INF_WG) __attribute__((amdgpu_waves_per_eu(8))) void k_inflate(const uint8_t* comp, const InfBlock* blocks, uint8_t* out,
                                                    unsigned long long* bad) {
    __shared__ InfLds T;
    const uint32_t b = blockIdx.x, lane = __lane_id();
    const InfBlock B = blocks[b];
    uint8_t* dst = out + B.out;
    const uint32_t isize = B.isize;
    InfBits r{comp + B.payload, T.in, B.payload_len, 0, 0, 0ull, 0ull, 0, 0};
    if (lane < 29) {
        T.lbase[lane] = kLenBase[lane];
        T.lextra[lane] = kLenExtra[lane];
    }
    if (lane < 30) {
        T.dbase[lane] = kDistBase[lane];
        T.dextra[lane] = kDistExtra[lane];
    }
    __syncthreads();
    uint32_t o = 0;    // bytes produced
    uint32_t fl = 0;   // bytes flushed to the slot; [fl, o) is only in the ring
    // Ring bytes [fl, upto) to the slot; visible to later reads of the slot.
    auto flush = [&](uint32_t upto) {
#pragma unroll 1
        for (uint32_t x = fl + lane; x < upto; x += INF_WG) dst[x] = T.outb[x % INF_OUT];
        fl = upto;
        __threadfence_block();
    };
    // A back-reference (wave-wide): sources are bytes before o, in the ring
    // unless this copy's writes reach their slots, else in the slot (flushed:
    // pending output o - fl stays below INF_FLUSH + 64 + 258, so every
    // source older than the ring is below fl).
    auto copy = [&](uint32_t len, uint32_t dist) {
        const uint32_t from = o - dist;
        if (dist >= len && from + INF_OUT >= o + len) {
            // no overlap, every source in the ring
#pragma unroll 1
            for (uint32_t i = lane; i < len; i += INF_WG) T.outb[(o + i) % INF_OUT] = T.outb[(from + i) % INF_OUT];
        } else {
#pragma unroll 1
            for (uint32_t i = lane; i < len; i += INF_WG) {
                const uint32_t s = from + (dist < len ? i % dist : i);
                const uint8_t v = s + INF_OUT >= o + len ? T.outb[s % INF_OUT] : dst[s];
                T.outb[(o + i) % INF_OUT] = v;
            }
        }
        o += len;
    };
    bool ok = true, last = false;
    while (ok && !last) {
        last = r.take(1) != 0;
        const uint32_t type = r.take(2);
        if (type == 0) {
            // Stored: to the byte boundary, LEN and NLEN, LEN raw bytes
            // (into the slot and the ring, which keeps the last INF_OUT).
            const uint32_t p = r.pos - r.bc / 8u;
            r.b0 = r.b1 = 0;
            r.bc = 0;
            r.over = 0;
            if (p + 4u > r.n) { ok = false; break; }
            const uint32_t ln = (uint32_t)r.src[p] | ((uint32_t)r.src[p + 1] << 8);
            const uint32_t nl = (uint32_t)r.src[p + 2] | ((uint32_t)r.src[p + 3] << 8);
            if ((ln ^ 0xFFFFu) != nl || p + 4u + ln > r.n || o + ln > isize) { ok = false; break; }
            flush(o);
#pragma unroll 1
            for (uint32_t i = lane; i < ln; i += INF_WG) {
                const uint8_t v = r.src[p + 4u + i];
                dst[o + i] = v;
                T.outb[(o + i) % INF_OUT] = v;
            }
            o += ln;
            fl = o;
            __threadfence_block();
            r.pos = p + 4u + ln;
            r.hi = r.pos;   // the ring is loaded again from here
            continue;
        }
        if (type == 3) { ok = false; break; }
        // Code lengths: fixed (RFC 1951 3.2.6) or sent (3.2.7).
        uint32_t nlen = 288, ndist = 32;
        if (type == 1) {
            for (uint32_t s = lane; s < 320; s += INF_WG)
                T.lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : s < 288 ? 8 : 5;
        } else {
            nlen = r.take(5) + 257u;
            ndist = r.take(5) + 1u;
            const uint32_t ncl = r.take(4) + 4u;
            if (nlen > 286 || ndist > 30) { ok = false; break; }
            if (lane < 19) T.cl[lane] = 0;
            __syncthreads();
            for (uint32_t k = 0; k < ncl; ++k) {
                const uint32_t v = r.take(3);
                if (lane == 0) T.cl[kClOrder[k]] = (uint8_t)v;
            }
            __syncthreads();
            // The code-length code (7-bit codes at most), decoded bit by bit.
            if (lane == 0) T.ok = inf_canon(T.cl, 19, T.ccount, T.csym, T.offs, false) ? 1u : 0u;
            __syncthreads();
            if (!T.ok) { ok = false; break; }
            uint32_t i = 0;
            uint32_t prev = 0;
            while (i < nlen + ndist) {
                if (r.bc < 32) r.fill();
                uint32_t code = 0, first = 0, idx = 0, sym = INF_NONE;
                for (int l = 1; l < 8; ++l) {
                    code |= (uint32_t)((r.b0 >> (l - 1)) & 1ull);
                    if (code - first < T.ccount[l]) {
                        r.drop((uint32_t)l);
                        sym = T.csym[idx + (code - first)];
                        break;
                    }
                    idx += T.ccount[l];
                    first = (first + T.ccount[l]) << 1;
                    code <<= 1;
                }
                if (sym == INF_NONE) { ok = false; break; }
                uint32_t rep = 1, val = sym;
                if (sym == 16) {
                    if (i == 0) { ok = false; break; }
                    val = prev;
                    rep = 3 + r.take(2);
                } else if (sym == 17) {
                    val = 0;
                    rep = 3 + r.take(3);
                } else if (sym == 18) {
                    val = 0;
                    rep = 11 + r.take(7);
                }
                if (i + rep > nlen + ndist) { ok = false; break; }
                // lengths are stored at [0, nlen) and [288, 288 + ndist)
                for (uint32_t k = 0; k < rep; ++k) {
                    const uint32_t at = i + k < nlen ? i + k : 288u + (i + k - nlen);
                    if (lane == 0) T.lens[at] = (uint8_t)val;
                }
                prev = val;
                i += rep;
            }
            if (!ok) break;
            if (lane == 0) {
                for (uint32_t s = nlen; s < 288; ++s) T.lens[s] = 0;
                for (uint32_t s = 288 + ndist; s < 320; ++s) T.lens[s] = 0;
            }
            __syncthreads();
            if (T.lens[256] == 0) { ok = false; break; }   // no end-of-block code
        }
        __syncthreads();
        if (lane == 0) {
            T.ok = (inf_canon(T.lens, 288, T.lcount, T.lsym, T.offs, true) &&
                    inf_canon(T.lens + 288, 32, T.dcount, T.dsym, T.offs, true)) ? 1u : 0u;
        }
        __syncthreads();
        if (!T.ok) { ok = false; break; }
        inf_fill(T.lt, INF_LB, T.lcount, T.lsym, false);
        inf_fill(T.dt, INF_DB, T.dcount, T.dsym, true);
        // Symbols until the end of the block, a window at a time.
        bool eob = false;
        while (ok && !eob) {
            r.fill();   // 96+ bits (zeros past the payload are counted in `over`)
            if (o - fl >= INF_FLUSH) flush(o);
            const uint64_t v = lane ? (r.b0 >> lane) | (r.b1 << (64u - lane)) : r.b0;
            const uint32_t E = T.lt[v & ((1u << INF_LB) - 1u)];
            const uint32_t D = T.dt[v & ((1u << INF_DB) - 1u)];
            uint32_t p = 0, nlit = 0, litv = 0;
            bool slow = false;
            while (p < 64) {
                const uint32_t e = lane_value(E, p);
                const uint32_t kind = e >> 28, l = (e >> 24) & 15u;
                if (kind == INF_K_LIT) {
                    litv = (uint32_t)fx_writelane((int)(e & 0xFFu), (int)nlit, (int)litv);
                    ++nlit;
                    p += l;
                    continue;
                }
                if (kind == INF_K_EOB) {
                    p += l;
                    eob = true;
                    break;
                }
                if (kind == INF_K_LONG) { slow = true; break; }
                if (kind != INF_K_BASE) { ok = false; break; }
                const uint32_t le = (e >> 16) & 0xFFu, q = p + l + le;
                if (q >= 64) break;   // the next window starts at this length code
                const uint32_t d = lane_value(D, q);
                const uint32_t dkind = d >> 28;
                if (dkind == INF_K_LONG) { slow = true; break; }
                if (dkind != INF_K_BASE) { ok = false; break; }
                const uint32_t dl = (d >> 24) & 15u, de = (d >> 16) & 0xFFu;
                const uint32_t len = (e & 0xFFFFu) + r.at(p + l, le);
                const uint32_t dist = (d & 0xFFFFu) + r.at(q + dl, de);
                // the staged literals, then the copy
                if (o + nlit > isize) { ok = false; break; }
                if (lane < nlit) T.outb[(o + lane) % INF_OUT] = (uint8_t)litv;
                o += nlit;
                nlit = 0;
                if (dist > o || o + len > isize) { ok = false; break; }
                copy(len, dist);
                p = q + dl + de;
                if (o - fl >= INF_FLUSH) flush(o);
            }
            if (!ok) break;
            if (o + nlit > isize) { ok = false; break; }
            if (lane < nlit) T.outb[(o + lane) % INF_OUT] = (uint8_t)litv;
            o += nlit;
            r.drop(p);
            if (slow && !eob) {
                // One symbol with a code longer than the tables, serially.
                const uint32_t sym = inf_decode(r, T.lcount, T.lsym);
                if (sym < 256) {
                    if (o >= isize) { ok = false; break; }
                    if (lane == 0) T.outb[o % INF_OUT] = (uint8_t)sym;
                    ++o;
                    continue;
                }
                if (sym == 256) { eob = true; break; }
                if (sym == INF_NONE || sym > 285) { ok = false; break; }
                const uint32_t li = sym - 257;
                const uint32_t len = T.lbase[li] + r.take(T.lextra[li]);
                const uint32_t ds = inf_decode(r, T.dcount, T.dsym);
                if (ds == INF_NONE || ds > 29) { ok = false; break; }
                const uint32_t dist = T.dbase[ds] + r.take(T.dextra[ds]);
                if (dist > o || o + len > isize) { ok = false; break; }
                copy(len, dist);
            }
        }
        if (ok && r.overrun()) ok = false;
    }
    if (ok && o != isize) ok = false;
    if (ok) flush(o);
    if (!ok && lane == 0) atomicMin(bad, (unsigned long long)b);
}

}  // namespace fx
