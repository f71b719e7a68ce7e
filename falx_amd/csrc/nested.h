// nested.h — the JSON nested tape (json.rs:599-760, parse_nested and
// parse_nested_par) from the structural index, on gfx950.
//
// The index (scan.hip M_INDEX in its nested form) leaves one 64-bit entry
// per structural byte outside strings: the position in the low word and,
// for a bracket, the bracket byte in the high word. Brackets are matched by
// depth instead of a stack: with depth(i) the number of opens minus closes
// up to and including entry i, an open at depth L is closed by the first
// later entry whose depth is below L. That is the stack's answer for every
// close before the first entry of negative depth, where the stack runs
// empty (an unmatched close); past it the reference stops, and so does the
// tape. So the depths (one scan), a search for the next smaller depth, and
// the first negative depth and first mismatched pair give the reference's
// tape and error exactly.
//
// Three kernels over tiles of NT index entries:
//   k_nest_local — depths within the tile; brackets whose partner lies in
//     the same 64-entry block are paired by level (an open at level L pairs
//     with the next bracket of level L in the block, its close); the rest
//     search the tile's later blocks through the blocks' minimum depths;
//     opens still unpaired are listed for k_nest_far.
//   k_nest_prefix (one workgroup) — each tile's depth before it, the tile
//     minima as the leaves of a 64-ary min tree (k_nest_tree builds the
//     levels above), the first mismatch found so far.
//   k_nest_far — one wave per listed open: up the min tree to the first
//     later subtree whose minimum is below the open's depth, down to its
//     tile, block and entry; and the tile's first entry of negative depth.
// A document with an error then runs k_nest_cut over the tape before the
// error: an open whose close lies at or past it is unclosed there.
#pragma once

namespace fx {

constexpr uint32_t NT = 4096;                 // index entries per tile
constexpr uint32_t NT_BLOCKS = NT / 64;       // 64-entry blocks per tile
constexpr uint32_t NT_NONE = 0xFFFFFFFFu;     // "no partner" in the tape's high word
constexpr int NT_LEVELS = 4;                  // min-tree levels above the blocks: tiles, 64, 64^2, 64^3 tiles
constexpr int NT_FAR_WG = 256;

// 0 separator, 1 '{', 2 '[', 3 '}', 4 ']'.
DEV uint32_t nt_kind(uint32_t b) { return b == '{' ? 1u : b == '[' ? 2u : b == '}' ? 3u : b == ']' ? 4u : 0u; }
DEV int32_t nt_delta(uint32_t k) { return k == 0 ? 0 : (k <= 2 ? 1 : -1); }
DEV bool nt_open(uint32_t k) { return k == 1 || k == 2; }
DEV bool nt_pair_ok(uint32_t open_kind, uint32_t close_kind) { return close_kind == open_kind + 2; }

// Block statistics, two int16 in a word: the depth before the block's first
// entry (low) and the minimum depth over its entries (high), both relative
// to the tile's start. Tile-relative depths lie in [-NT, NT].
DEV uint32_t nt_bstat(int32_t start, int32_t mn) { return (uint32_t)(uint16_t)(int16_t)start | ((uint32_t)(uint16_t)(int16_t)mn << 16); }
DEV int32_t nt_bstart(uint32_t w) { return (int32_t)(int16_t)(uint16_t)(w & 0xFFFFu); }
DEV int32_t nt_bmin(uint32_t w) { return (int32_t)(int16_t)(uint16_t)(w >> 16); }

// A listed open: tile-local index (12 bits), '[' flag, depth + 2 NT (14 bits).
DEV uint32_t nt_res(uint32_t k, uint32_t kind, int32_t depth) {
    return k | (kind == 2 ? 1u << 12 : 0u) | ((uint32_t)(depth + 2 * (int32_t)NT) << 13);
}

struct NestScratch {
    int32_t* tsum;             // per tile: depth change over the tile
    int32_t* tmin;             // per tile: minimum tile-relative depth
    uint32_t* tmis;            // per tile: first mismatched close paired in the tile (global index) or NT_NONE
    uint32_t* rcount;          // per tile: opens listed for k_nest_far
    uint32_t* res;             // NT per tile: the listed opens (nt_res)
    uint32_t* bstat;           // NT_BLOCKS per tile (nt_bstat)
    long long* base;           // per tile: depth before the tile
    long long* lv[NT_LEVELS];  // min tree: lv[0] per tile (base + tmin), lv[k] per 64^k tiles
    uint32_t lv_n[NT_LEVELS];
    unsigned long long* result;  // R_NT_* slots
};

DEV uint32_t nt_blocks_of(uint32_t tile, uint32_t n) { return (min(NT, n - tile * NT) + 63u) / 64u; }

__global__ __launch_bounds__(WG) void k_nest_local(unsigned long long* tape, uint32_t n, NestScratch S) {
    __shared__ __align__(16) uint32_t shi[NT];     // high words: the bracket byte as indexed, then the partner
    __shared__ __align__(16) uint8_t skind[NT];
    __shared__ __align__(16) int16_t sdep[NT];     // tile-relative depth after each entry
    __shared__ int32_t sbmin[NT_BLOCKS];
    __shared__ uint32_t scan_sh[NWAVE];
    __shared__ uint32_t nres, smis;
    const uint32_t tile = blockIdx.x, tid = threadIdx.x, lane = __lane_id(), wave = uni(tid >> 6);
    const uint32_t e0 = tile * NT;
    const uint32_t cnt = min(NT, n - e0);
    const uint32_t nb = (cnt + 63u) / 64u;
    if (tid == 0) {
        nres = 0;
        smis = NT_NONE;
    }
    // The thread's entries k = tid + WG j stay in registers for the final
    // write; LDS holds their high words and kinds.
    constexpr uint32_t PER = NT / WG;
    unsigned long long word[PER];
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t k = tid + WG * j;
        word[j] = k < cnt ? tape[e0 + k] : 0ull;
    }
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t k = tid + WG * j;
        const uint32_t hi = (uint32_t)(word[j] >> 32);
        shi[k] = hi;
        skind[k] = (uint8_t)nt_kind(hi);
    }
    __syncthreads();

    // Depths: thread t owns entries [16t, 16t + 16), one block scan.
    const uint4 kw = *reinterpret_cast<const uint4*>(&skind[tid * 16]);
    const uint32_t kwords[4] = {kw.x, kw.y, kw.z, kw.w};
    int32_t sum = 0;
#pragma unroll
    for (uint32_t q = 0; q < 16; ++q) sum += nt_delta((kwords[q >> 2] >> (8 * (q & 3))) & 0xFFu);
    uint32_t tot;
    const int32_t before = (int32_t)block_excl_sum((uint32_t)sum, scan_sh, tot);
    int32_t run = before, mn = 0x7FFF;
    uint32_t dw[8];
#pragma unroll
    for (uint32_t q = 0; q < 16; ++q) {
        run += nt_delta((kwords[q >> 2] >> (8 * (q & 3))) & 0xFFu);
        if (tid * 16 + q < cnt) mn = min(mn, run);
        const uint32_t h = (uint32_t)(uint16_t)(int16_t)run;
        if (q & 1) dw[q >> 1] |= h << 16;
        else dw[q >> 1] = h;
    }
    *reinterpret_cast<uint4*>(&sdep[tid * 16]) = make_uint4(dw[0], dw[1], dw[2], dw[3]);
    *reinterpret_cast<uint4*>(&sdep[tid * 16 + 8]) = make_uint4(dw[4], dw[5], dw[6], dw[7]);
    // A 64-entry block is four consecutive threads.
    mn = min(mn, __shfl_xor(mn, 1));
    mn = min(mn, __shfl_xor(mn, 2));
    if ((tid & 3u) == 0) {
        const uint32_t b = tid >> 2;
        sbmin[b] = mn;
        S.bstat[(size_t)tile * NT_BLOCKS + b] = nt_bstat(before, mn);
    }
    __syncthreads();

    uint32_t mis = NT_NONE;
    auto pair = [&](uint32_t ki, uint32_t kj, uint32_t open_kind) {
        shi[ki] = e0 + kj;
        shi[kj] = e0 + ki;
        if (!nt_pair_ok(open_kind, skind[kj])) mis = min(mis, e0 + kj);
    };
    for (uint32_t b = wave; b < nb; b += NWAVE) {
        const uint32_t k = b * 64u + lane;
        const uint32_t kind = k < cnt ? (uint32_t)skind[k] : 0u;
        const bool open = nt_open(kind);
        const int32_t dep = sdep[k];
        const int32_t lvl = open ? dep : dep + 1;   // a close's level is the depth before it
        // Pair by level: all brackets of the first unpaired lane's level at once.
        bool done = kind == 0, far = false;
        uint64_t m;
        while ((m = ballot(!done)) != 0) {
            const uint32_t f = (uint32_t)__ffsll((long long)m) - 1u;
            const int32_t lf = (int32_t)lane_value((uint32_t)lvl, f);
            const bool mine = !done && lvl == lf;
            const uint64_t B = ballot(mine);
            if (mine) {
                done = true;
                if (open) {
                    const uint64_t above = B & ~((2ull << lane) - 1ull);
                    if (above) pair(k, b * 64u + (uint32_t)__ffsll((long long)above) - 1u, kind);
                    else far = true;
                }
            }
        }
        // Opens without a close in the block: the first later entry of the
        // tile whose depth is below theirs, found block then entry.
        uint64_t R = ballot(far);
        while (R) {
            const uint32_t r = (uint32_t)__ffsll((long long)R) - 1u;
            R &= R - 1ull;
            const uint32_t ki = b * 64u + r;
            const int32_t L = (int32_t)lane_value((uint32_t)lvl, r);
            const uint32_t okind = lane_value(kind, r);
            const uint64_t cb = ballot(lane > b && lane < nb && sbmin[lane] < L);
            if (cb) {
                const uint32_t b2 = (uint32_t)__ffsll((long long)cb) - 1u;
                const uint32_t k2 = b2 * 64u + lane;
                const uint64_t ce = ballot(k2 < cnt && (int32_t)sdep[k2] < L);
                if (ce && lane == 0) pair(ki, b2 * 64u + (uint32_t)__ffsll((long long)ce) - 1u, okind);
            } else if (lane == 0) {
                shi[ki] = NT_NONE;
                const uint32_t slot = atomicAdd(&nres, 1u);
                S.res[(size_t)tile * NT + slot] = nt_res(ki, okind, L);
            }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mis = min(mis, (uint32_t)__shfl_xor((int)mis, d));
    if (lane == 0 && mis != NT_NONE) atomicMin(&smis, mis);
    __syncthreads();
    // The words that changed (paired or listed brackets); a close paired
    // outside the tile keeps its bracket byte for k_nest_far's kind check.
#pragma unroll
    for (uint32_t j = 0; j < PER; ++j) {
        const uint32_t k = tid + WG * j;
        const uint32_t hi = shi[k];
        if (k < cnt && hi != (uint32_t)(word[j] >> 32)) tape[e0 + k] = ((unsigned long long)hi << 32) | (uint32_t)word[j];
    }
    if (tid == 0) {
        int32_t tm = 0x7FFF;
        for (uint32_t b = 0; b < nb; ++b) tm = min(tm, sbmin[b]);
        S.tsum[tile] = (int32_t)tot;
        S.tmin[tile] = tm;
        S.tmis[tile] = smis;
        S.rcount[tile] = nres;
    }
}

// One workgroup: depth before each tile (an exclusive scan in rounds of
// 4096 tiles, four per thread), the tile minima as min-tree leaves, and the
// result slots.
constexpr int NT_PREFIX_WG = 1024;

__global__ __launch_bounds__(NT_PREFIX_WG) void k_nest_prefix(uint32_t n_tiles, NestScratch S) {
    __shared__ long long wsum[NT_PREFIX_WG / 64];
    __shared__ uint32_t wmis[NT_PREFIX_WG / 64];
    const uint32_t tid = threadIdx.x, lane = __lane_id(), w = tid >> 6;
    constexpr uint32_t ROUND = 4 * NT_PREFIX_WG;   // four consecutive tiles per thread per round
    long long carry = 0;
    uint32_t mis = NT_NONE;
    for (uint32_t r0 = 0; r0 < n_tiles; r0 += ROUND) {
        const uint32_t t = r0 + 4 * tid;
        int32_t v[4] = {0, 0, 0, 0}, mn[4] = {0, 0, 0, 0};
        if (t + 4 <= n_tiles) {
            const int4 a = *reinterpret_cast<const int4*>(S.tsum + t);
            const int4 b = *reinterpret_cast<const int4*>(S.tmin + t);
            const uint4 c = *reinterpret_cast<const uint4*>(S.tmis + t);
            v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w;
            mn[0] = b.x; mn[1] = b.y; mn[2] = b.z; mn[3] = b.w;
            mis = min(mis, min(min(c.x, c.y), min(c.z, c.w)));
        } else {
            for (uint32_t q = 0; q < 4; ++q)
                if (t + q < n_tiles) {
                    v[q] = S.tsum[t + q];
                    mn[q] = S.tmin[t + q];
                    mis = min(mis, S.tmis[t + q]);
                }
        }
        const long long sum = (long long)v[0] + v[1] + v[2] + v[3];
        long long inc = sum;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const long long u = __shfl_up(inc, d);
            if ((int)lane >= d) inc += u;
        }
        if (lane == 63) wsum[w] = inc;
        __syncthreads();
        long long run = carry + inc - sum, all = 0;
        for (uint32_t k = 0; k < NT_PREFIX_WG / 64; ++k) {
            if (k < w) run += wsum[k];
            all += wsum[k];
        }
        __syncthreads();
        for (uint32_t q = 0; q < 4; ++q) {
            if (t + q < n_tiles) {
                S.base[t + q] = run;
                S.lv[0][t + q] = run + mn[q];
            }
            run += v[q];
        }
        carry += all;
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) mis = min(mis, (uint32_t)__shfl_xor((int)mis, d));
    if (lane == 0) wmis[w] = mis;
    __syncthreads();
    if (tid == 0) {
        uint32_t m = NT_NONE;
        for (uint32_t k = 0; k < NT_PREFIX_WG / 64; ++k) m = min(m, wmis[k]);
        S.result[R_NT_NEG] = NT_NONE;
        S.result[R_NT_MIS] = m;
        S.result[R_NT_OPEN] = 0;
        S.result[R_NT_DEPTH] = (unsigned long long)carry;
    }
}

// One min-tree level: out[i] = min of in[64 i, 64 i + 64).
__global__ __launch_bounds__(WG) void k_nest_tree(const long long* in, uint32_t n_in, long long* out, uint32_t n_out) {
    const uint32_t i = blockIdx.x * WG + threadIdx.x;
    if (i >= n_out) return;
    long long m = LLONG_MAX;
    const uint32_t hi = min(n_in, (i + 1u) * 64u);
    for (uint32_t c = i * 64u; c < hi; ++c) m = min(m, in[c]);
    out[i] = m;
}

// First entry of block (t, b) whose depth is below L, or NT_NONE; the
// depths come from the block's start depth and its bracket bytes. The
// whole wave calls it with uniform arguments.
DEV uint32_t nt_first_below(const unsigned long long* tape, const uint8_t* data, uint32_t n, const NestScratch& S,
                            uint32_t t, uint32_t b, long long L) {
    const uint32_t lane = __lane_id();
    const uint32_t e = t * NT + b * 64u + lane;
    uint32_t kind = 0;
    if (e < n) kind = nt_kind(data[(uint32_t)tape[e]]);
    const int32_t dl = (int32_t)wave_incl_sum((uint32_t)nt_delta(kind));
    const long long D = S.base[t] + nt_bstart(S.bstat[(size_t)t * NT_BLOCKS + b]) + dl;
    const uint64_t c = ballot(e < n && D < L);
    return c ? t * NT + b * 64u + (uint32_t)__ffsll((long long)c) - 1u : NT_NONE;
}

// First block of tile t after block `after` (or from block 0 when after is
// NT_NONE... passed as -1) whose minimum depth is below L, or NT_NONE.
DEV uint32_t nt_block_below(uint32_t n, const NestScratch& S, uint32_t t, int32_t after, long long L) {
    const uint32_t lane = __lane_id();
    const uint32_t nb = nt_blocks_of(t, n);
    const bool ok = (int32_t)lane > after && lane < nb &&
                    S.base[t] + nt_bmin(S.bstat[(size_t)t * NT_BLOCKS + lane]) < L;
    const uint64_t c = ballot(ok);
    return c ? (uint32_t)__ffsll((long long)c) - 1u : NT_NONE;
}

// The close of the open at tile t, block b whose depth is L: the first
// later entry with depth below L, or NT_NONE.
DEV uint32_t nt_search(const unsigned long long* tape, const uint8_t* data, uint32_t n, const NestScratch& S,
                       uint32_t t, uint32_t b, long long L) {
    const uint32_t lane = __lane_id();
    uint32_t b2 = nt_block_below(n, S, t, (int32_t)b, L);
    if (b2 != NT_NONE) return nt_first_below(tape, data, n, S, t, b2, L);
    // Up the tree: the first later sibling subtree with a minimum below L.
    uint32_t node = NT_NONE;
    int at = -1;
    for (int k = 0; k < NT_LEVELS && node == NT_NONE; ++k) {
        const uint32_t idx = t >> (6 * k);
        const uint32_t c_i = (idx & ~63u) + lane;
        const bool ok = lane > (idx & 63u) && c_i < S.lv_n[k] && S.lv[k][c_i] < L;
        const uint64_t c = ballot(ok);
        if (c) {
            node = (idx & ~63u) + (uint32_t)__ffsll((long long)c) - 1u;
            at = k;
        }
    }
    if (node == NT_NONE) return NT_NONE;
    // Down to the tile.
    for (int k = at; k > 0; --k) {
        const uint32_t c_i = node * 64u + lane;
        const uint64_t c = ballot(c_i < S.lv_n[k - 1] && S.lv[k - 1][c_i] < L);
        if (!c) return NT_NONE;
        node = node * 64u + (uint32_t)__ffsll((long long)c) - 1u;
    }
    b2 = nt_block_below(n, S, node, -1, L);
    if (b2 == NT_NONE) return NT_NONE;
    return nt_first_below(tape, data, n, S, node, b2, L);
}

__global__ __launch_bounds__(NT_FAR_WG) void k_nest_far(const uint8_t* data, unsigned long long* tape, uint32_t n,
                                                        NestScratch S) {
    const uint32_t t = blockIdx.x, lane = __lane_id(), wave = uni(threadIdx.x >> 6);
    // The tile's first entry of negative depth (the stack runs empty there).
    if (wave == 0 && S.lv[0][t] < 0) {
        const uint32_t b = nt_block_below(n, S, t, -1, 0);
        const uint32_t e = b == NT_NONE ? NT_NONE : nt_first_below(tape, data, n, S, t, b, 0);
        if (lane == 0 && e != NT_NONE) atomicMin(&S.result[R_NT_NEG], (unsigned long long)e);
    }
    const uint32_t nr = S.rcount[t];
    for (uint32_t r = wave; r < nr; r += NT_FAR_WG / 64) {
        const uint32_t w = S.res[(size_t)t * NT + r];
        const uint32_t ki = w & (NT - 1u);
        const uint32_t okind = (w >> 12) & 1u ? 2u : 1u;
        const long long L = S.base[t] + (long long)((int32_t)(w >> 13) - 2 * (int32_t)NT);
        const uint32_t gi = t * NT + ki;
        const uint32_t j = nt_search(tape, data, n, S, t, ki / 64u, L);
        if (lane == 0) {
            if (j != NT_NONE) {
                // The close was left as indexed: its bracket byte is in the high word.
                const unsigned long long vj = tape[j];
                const uint32_t pos_i = (uint32_t)tape[gi];
                tape[gi] = ((unsigned long long)j << 32) | pos_i;
                tape[j] = ((unsigned long long)gi << 32) | (uint32_t)vj;
                if (!nt_pair_ok(okind, nt_kind((uint32_t)(vj >> 32)))) atomicMin(&S.result[R_NT_MIS], (unsigned long long)j);
            } else {
                atomicMax(&S.result[R_NT_OPEN], (unsigned long long)gi + 1ull);
            }
        }
    }
}

// The tape before the first error at entry `cut`: an open whose close is at
// or past the cut was still open when the reference stopped.
__global__ __launch_bounds__(WG) void k_nest_cut(unsigned long long* tape, uint32_t cut) {
    for (uint32_t i = blockIdx.x * WG + threadIdx.x; i < cut; i += gridDim.x * WG) {
        const unsigned long long v = tape[i];
        const uint32_t p = (uint32_t)(v >> 32);
        if (p != NT_NONE && p > i && p >= cut) tape[i] = ((unsigned long long)NT_NONE << 32) | (uint32_t)v;
    }
}

}  // namespace fx
