// scan.hip — gfx950 kernels of the structural-indexing / column-parsing path.
//
// The input is cut into 8 KiB tiles; one workgroup (4 wave64s) takes a tile,
// reads it from HBM with coalesced 16-byte loads into LDS, and classifies
// it: each wave compares one byte per lane and __ballot gives the
// reference's per-block class masks (quote, separator, newline, escape,
// comment) as 64-bit values; block b's masks then go to lane b, so the
// dialect's mask algebra (quote parity / escape add-carry / NORMAL-QUOTE-
// COMMENT regions, csv_geo.rs:2116-2127, csv_hash.rs:2124-2138,
// ndjson.rs:1016-1032) runs for 32 blocks at once on the vector unit.
//
// A call is a short pipeline of passes over all tiles, with no waiting
// between tiles:
//   summary  per tile: the transfer function of its dialect state and, for
//            every possible entry state, its record/field counters;
//   scan     three small kernels compose the summaries (the algebra of the
//            reference's chunk transfer functions, csv_hash.rs:184-191) into
//            each tile's entry state and exclusive counters;
//   emit     per tile, with its entry state known: structural/terminator
//            masks, popcount/mbcnt compaction into an LDS list, then a walk
//            of the list with record context: index/tape offsets, typed
//            cells parsed in place, fused statistics (columns and stats run
//            first in a fast form with only the register number parsers;
//            tiles holding a cell it leaves undecided are redone in full);
//   strings  (string columns only) after a scan of the tiles' cleaned byte
//            counts, the cells are copied into the Arrow arena.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <type_traits>

#include "../../include/falx_gpu.h"
#include "cells.h"
#include "ndjson_walk.h"
#include "num_swar.h"
#include "regions.h"

#define DEV __device__ __forceinline__

namespace fx {

constexpr int WG = 256;
constexpr int NWAVE = WG / 64;
constexpr int BPW = 32;                   // 64-byte blocks per wave
constexpr int TILE = NWAVE * BPW * 64;    // 8192 bytes per workgroup
// The LDS copy of a tile also holds PRE bytes before it and HALO bytes
// after it, so that cells crossing the tile's edges (a tile's first entry
// usually ends a cell that began in the previous tile) are read from LDS;
// it has a few more bytes of room after that for word reads.
constexpr int PRE = 64;
constexpr int HALO = 64;
constexpr int TILE_LDS = PRE + TILE + HALO + 16;
// Terminators the numeric fast forms list per tile (more: the full pass).
constexpr int TL_CAP = 1024;
constexpr uint64_t EVEN_BITS = 0x5555555555555555ULL;
constexpr int STAT_GROUP = 2;             // stat columns accumulated per STATS pass
constexpr int COL_STAT_GROUP = 2;         // stat columns fused into a COLUMNS pass

enum Mode : int {
    M_INDEX = 0, M_TAPE, M_COLUMNS, M_STATS, M_FIELDBYTES, M_LOGFMT, M_SUMMARY, M_STRINGS,
    // COLUMNS and STATS with the register number parsers only: a tile holding
    // a cell they leave undecided is flagged in Params::fixup and redone by
    // the full pass, whose launch skips the unflagged tiles.
    M_COLUMNS_FAST, M_STATS_FAST,
    M_COLUMNS_FAST_STATS,  // M_COLUMNS_FAST with the fused statistics (the plain fast form has none)
    // The fast forms for dialects without escape or comment when every
    // column of the pass is numeric (i64/f64): the terminators are listed
    // while the tile is compacted, and records go straight to threads.
    M_NUM_FAST, M_NUM_FAST_STATS, M_NUM_STATS_FAST
};

// Record/field counters of a byte range; `compose(a, b)` = a then b.
struct Counts {
    uint32_t n_struct, n_seps, n_terms, n_rows;
    uint32_t has_term, ord, last_struct, prev_struct;  // positions stored +1, 0 = none
    uint32_t last_term, pad0, pad1, pad2;
};

struct ColDesc {
    void* values;
    unsigned long long* valid;
    uint8_t* data;
    uint64_t data_cap;
    uint32_t index, type, str_slot, pad;
};

struct StatPart {
    unsigned long long values, checksum, bytes;
    long long fixed6_sum;
    unsigned long long fixed6_cells;
    double sum;
};

// result[] slots written by the last tile / reduce kernels
enum : int {
    R_N_STRUCT = 0, R_N_SEPS, R_N_TERMS, R_N_ROWS,
    R_STR_LEN = 16,   // + str_slot (32)
    // nested tape: first entry below depth 0, first mismatched close, innermost
    // open never closed + 1, final depth
    R_NT_NEG = 56, R_NT_MIS, R_NT_OPEN, R_NT_DEPTH,
    R_INF_BAD = 60,   // first malformed BGZF block
    R_SLOTS = 64
};

struct Params {
    const uint8_t* data;
    uint32_t len;
    uint32_t n_tiles;
    uint32_t cand[8];
    uint32_t sepb[4];         // the candidate bytes other than '\n' when at most 4 (0x100 = unused)
    uint32_t n_sepb;          // number of candidate bytes other than '\n'; more than 4 uses the cand[] table
    uint32_t nl_cand;         // '\n' is a candidate
    int32_t quote, escape, comment;
    uint32_t n_cols, n_str;
    uint32_t col_group;       // STATS: first column of this launch's group
    uint64_t rows_cap;
    ColDesc cols[FALX_MAX_COLUMNS];
    uint32_t* tsum;           // M_SUMMARY out / emit in: per tile ts_stride words (tile summary)
    const uint32_t* tpre;     // emit modes: per tile 8 words, entry state + exclusive counters
    uint32_t ts_stride;
    uint32_t* str_len;        // COLUMNS out: per tile n_str cleaned byte counts
    const uint32_t* str_base; // STRINGS in: per tile n_str arena offsets
    uint32_t* out;
    uint64_t out_cap;
    unsigned long long* ends;
    uint64_t ends_cap;
    void* partials;
    unsigned long long* result;
    uint32_t* fixup;          // per tile: written by the *_FAST passes; a full pass given it skips zeros
};

// --------------------------------------------------------------------------
// wave / bit helpers
// --------------------------------------------------------------------------
DEV uint64_t ballot(bool p) { return (uint64_t)__ballot(p); }

DEV uint32_t mbcnt(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

DEV uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// v_writelane_b32 (the LLVM intrinsic; clang has no builtin for it here):
// `old` with lane `lane` replaced by the wave-uniform `val`.
extern "C" __device__ int fx_writelane(int val, int lane, int old) __asm("llvm.amdgcn.writelane.i32");

// v with lane `lane` (wave-uniform) replaced by the wave-uniform value s.
DEV uint64_t put_lane(uint64_t v, uint64_t s, uint32_t lane) {
    const uint32_t lo = (uint32_t)fx_writelane((int)(uint32_t)s, (int)lane, (int)(uint32_t)v);
    const uint32_t hi = (uint32_t)fx_writelane((int)(uint32_t)(s >> 32), (int)lane, (int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

DEV uint32_t popc(uint64_t m) { return (uint32_t)__popcll(m); }

DEV uint32_t hibit(uint64_t m) { return 63u - (uint32_t)__clzll(m); }

// Transfer functions on <= 8 states, 4 bits per entry.
DEV uint32_t fn_apply(uint32_t f, uint32_t s) { return (f >> (4 * s)) & 15u; }

template <int NS>
DEV uint32_t fn_then(uint32_t f, uint32_t g) {  // g after f
    uint32_t h = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) h |= fn_apply(g, fn_apply(f, (uint32_t)s)) << (4 * s);
    return h;
}

template <int NS>
DEV uint32_t fn_ident() {
    uint32_t h = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) h |= (uint32_t)s << (4 * s);
    return h;
}

DEV Counts counts_zero() {
    Counts c;
    c.n_struct = c.n_seps = c.n_terms = c.n_rows = 0;
    c.has_term = c.ord = c.last_struct = c.prev_struct = 0;
    c.last_term = c.pad0 = c.pad1 = c.pad2 = 0;
    return c;
}

// Operands by value: conditional expressions over references into
// different address spaces (LDS and registers) would force a private copy.
DEV Counts compose(Counts a, Counts b) {
    Counts r;
    r.n_struct = a.n_struct + b.n_struct;
    r.n_seps = a.n_seps + b.n_seps;
    r.n_terms = a.n_terms + b.n_terms;
    r.n_rows = a.n_rows + b.n_rows;
    r.has_term = a.has_term | b.has_term;
    const uint32_t a_ls = a.last_struct, a_ps = a.prev_struct, a_lt = a.last_term, a_ord = a.ord;
    const uint32_t b_ls = b.last_struct, b_ps = b.prev_struct, b_lt = b.last_term, b_ord = b.ord;
    r.ord = b.has_term ? b_ord : a_ord + b_ord;
    r.last_struct = b_ls ? b_ls : a_ls;
    r.prev_struct = b.n_struct >= 2 ? b_ps : (b.n_struct == 1 ? a_ls : a_ps);
    r.last_term = b_lt ? b_lt : a_lt;
    r.pad0 = r.pad1 = r.pad2 = 0;
    return r;
}

// Byte source for cells.h: LDS inside the tile, HBM outside it.
struct TileSrc {
    const uint8_t* g;
    const uint8_t* lds;
    uint32_t lo, hi;
    DEV uint8_t at(uint32_t p) const { return (p - lo) < (hi - lo) ? lds[p - lo] : g[p]; }
};

// --------------------------------------------------------------------------
// counters and tile summaries
// --------------------------------------------------------------------------
// Counts travel between kernels as 7 words (n_seps and has_term follow from
// them): n_struct n_terms n_rows ord last_struct prev_struct last_term.
constexpr int C7 = 7;

DEV Counts c7_load(const uint32_t* p) {
    Counts r;
    r.n_struct = p[0];
    r.n_terms = p[1];
    r.n_rows = p[2];
    r.ord = p[3];
    r.last_struct = p[4];
    r.prev_struct = p[5];
    r.last_term = p[6];
    r.n_seps = r.n_struct - r.n_terms;
    r.has_term = r.last_term != 0;
    r.pad0 = r.pad1 = r.pad2 = 0;
    return r;
}

DEV void c7_store(uint32_t* p, const Counts& c) {
    p[0] = c.n_struct;
    p[1] = c.n_terms;
    p[2] = c.n_rows;
    p[3] = c.ord;
    p[4] = c.last_struct;
    p[5] = c.prev_struct;
    p[6] = c.last_term;
}

// Tile summary layout (u32 words, stride 8 + 7*NS): [0] tile transfer
// function, [1..4] the four waves' functions, [8 + 7s ..] counters of the
// tile entered in state s. Functions pack 4 bits per state.
constexpr int TS_COUNTS = 8;
constexpr int NS_MAX = 6;

DEV uint32_t fn_ident_rt(uint32_t ns) {
    uint32_t h = 0;
    for (uint32_t s = 0; s < ns; ++s) h |= s << (4 * s);
    return h;
}


// f then g, for `ns` states.
DEV uint32_t fn_then_rt(uint32_t f, uint32_t g, uint32_t ns) {
    uint32_t h = 0;
    for (uint32_t s = 0; s < ns; ++s) h |= fn_apply(g, fn_apply(f, s)) << (4 * s);
    return h;
}

// --------------------------------------------------------------------------
// dialect step (one 64-byte block, wave-uniform)
// --------------------------------------------------------------------------
struct BM {
    uint64_t cand, nl, q, esc, cm;
};


// v + the value `CTRL` selects by DPP (0 where no lane is selected or the
// row is outside ROWM). All lanes of the wave must be active.
template <int CTRL, int ROWM>
DEV uint32_t dpp_plus(uint32_t v) {
    return v + (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWM, 0xF, false);
}

// Inclusive sum over lanes 0..ln of the wave: shifts within rows of 16
// lanes (row_shr 1, 2, 4, 8), then row 0's total into row 1 and row 2's into
// row 3 (row_bcast 15), then rows 0-1's total into rows 2-3 (row_bcast 31).
DEV uint32_t wave_incl_sum(uint32_t v) {
    v = dpp_plus<0x111, 0xF>(v);
    v = dpp_plus<0x112, 0xF>(v);
    v = dpp_plus<0x114, 0xF>(v);
    v = dpp_plus<0x118, 0xF>(v);
    v = dpp_plus<0x142, 0xA>(v);
    v = dpp_plus<0x143, 0xC>(v);
    return v;
}

// The value of the lane DPP control CTRL selects, for controls that select
// a lane of the row for every lane (quad permutes, row rotations).
template <int CTRL>
DEV uint32_t dpp_get(uint32_t v) { return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, 0xF, 0xF, false); }
template <int CTRL>
DEV unsigned long long dpp_get(unsigned long long v) {
    return ((unsigned long long)dpp_get<CTRL>((uint32_t)(v >> 32)) << 32) | dpp_get<CTRL>((uint32_t)v);
}
template <int CTRL>
DEV double dpp_get(double v) {
    return __longlong_as_double((long long)dpp_get<CTRL>((unsigned long long)__double_as_longlong(v)));
}

// Sum over the lane's row of 16 (quad swaps 1-0-3-2 and 2-3-0-1, then
// rotations by 4 and 8); lane 0 of the row has the row's sum in the order
// ((l0 + l1) + (l2 + l3)) + ... All lanes of the wave must be active.
template <class T>
DEV T row_sum16(T v) {
    v += dpp_get<0xB1>(v);
    v += dpp_get<0x4E>(v);
    v += dpp_get<0x124>(v);
    v += dpp_get<0x128>(v);
    return v;
}

template <int CTRL, int ROWM>
DEV uint32_t dpp_max(uint32_t v) {
    return max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, CTRL, ROWM, 0xF, false));
}

// Inclusive max over lanes 0..ln (values >= 0), the same DPP steps as
// wave_incl_sum.
DEV uint32_t wave_incl_max(uint32_t v) {
    v = dpp_max<0x111, 0xF>(v);
    v = dpp_max<0x112, 0xF>(v);
    v = dpp_max<0x114, 0xF>(v);
    v = dpp_max<0x118, 0xF>(v);
    v = dpp_max<0x142, 0xA>(v);
    v = dpp_max<0x143, 0xC>(v);
    return v;
}

// Lane ln - 1's value (0 in lane 0): DPP wave_shr 1.
DEV uint32_t dpp_prev_lane(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xF, 0xF, false);
}

DEV uint32_t lane_value(uint32_t v, uint32_t lane) { return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)lane); }

// The counters of a wave's blocks composed in lane order: lane b holds block
// b's structurals S, terminators T and comment-record terminators F at byte
// `base` (lanes without a block hold zero masks); the result is in every
// lane. Sums go through one scan as two 16-bit halves per word (a wave
// holds at most 4096 bytes); the positions come from the highest lanes with
// a structural or a terminator.
DEV Counts wave_counts(uint64_t S, uint64_t T, uint64_t F, uint32_t base) {
    Counts c = counts_zero();
    const uint64_t has_s = ballot(S != 0ULL);
    if (!has_s) return c;
    const uint64_t has_t = ballot(T != 0ULL);
    const uint32_t ln = __lane_id();
    const uint64_t seps = S & ~T;
    // Separators after the wave's last terminator: all of a lane's above
    // the terminator lane, those past its last terminator in that lane.
    uint32_t ord = popc(seps);
    if (has_t) {
        const uint32_t L = hibit(has_t);
        const uint32_t hb = T ? hibit(T) : 0u;
        if (ln == L) ord = hb == 63u ? 0u : popc(seps & ~((2ULL << hb) - 1ULL));
        else if (ln < L) ord = 0u;
        c.has_term = 1;
        c.last_term = lane_value(base + hb + 1u, L);
    }
    const uint32_t a = lane_value(wave_incl_sum(popc(S) | (popc(T) << 16)), 63);
    const uint32_t b = lane_value(wave_incl_sum(popc(T & ~F) | (ord << 16)), 63);
    c.n_struct = a & 0xFFFFu;
    c.n_terms = a >> 16;
    c.n_seps = c.n_struct - c.n_terms;
    c.n_rows = b & 0xFFFFu;
    c.ord = b >> 16;
    // Last and second-to-last structural of the wave.
    const uint32_t H = hibit(has_s);
    const uint32_t h = S ? hibit(S) : 0u;
    const uint64_t S2 = S & ~(1ULL << h);
    c.last_struct = lane_value(base + h + 1u, H);
    const uint32_t p2 = lane_value(S2 ? base + hibit(S2) + 1u : 0u, H);
    const uint64_t below = has_s & ~(1ULL << H);
    c.prev_struct = p2 ? p2 : (below ? lane_value(base + h + 1u, hibit(below)) : 0u);
    return c;
}

// The counters of a wave entered in state 0 and in state 1 of a quote-only
// dialect (the two entry states are complementary: what is inside quotes
// from one is outside from the other). Lane l holds a run of up to 32 bytes
// at `base`: its candidates C, newline candidates N (within C) and `inside`,
// the bytes inside quotes when the wave is entered in state 0. The same
// results as wave_counts on each state's masks, with the wave sums of both
// states shared (three scans instead of four) and 32-bit masks.
DEV void wave_counts_pair(uint32_t C, uint32_t N, uint32_t inside, uint32_t base, Counts& c0, Counts& c1) {
    const uint32_t ln = __lane_id();
    const uint32_t S0 = C & ~inside, S1 = C & inside, T0 = S0 & N, T1 = S1 & N;
    // Separators after the wave's last terminator, per state.
    auto ord_of = [&](uint32_t S, uint32_t T, uint64_t has_t) -> uint32_t {
        const uint32_t seps = S & ~T;
        if (!has_t) return (uint32_t)__popc(seps);
        const uint32_t L = hibit(has_t);
        if (ln < L) return 0u;
        if (ln > L) return (uint32_t)__popc(seps);
        const uint32_t hb = 31u - (uint32_t)__clz(T);
        return hb == 31u ? 0u : (uint32_t)__popc(seps >> (hb + 1u));
    };
    const uint64_t ht0 = ballot(T0 != 0u), ht1 = ballot(T1 != 0u);
    const uint32_t a0 = lane_value(wave_incl_sum((uint32_t)__popc(S0) | ((uint32_t)__popc(T0) << 16)), 63);
    const uint32_t a1 = lane_value(wave_incl_sum((uint32_t)__popc(S1) | ((uint32_t)__popc(T1) << 16)), 63);
    const uint32_t od = lane_value(wave_incl_sum(ord_of(S0, T0, ht0) | (ord_of(S1, T1, ht1) << 16)), 63);
    auto finish = [&](uint32_t S, uint32_t T, uint64_t ht, uint32_t a, uint32_t ord, Counts& c) {
        c = counts_zero();
        const uint64_t hs = ballot(S != 0u);
        if (!hs) return;
        c.n_struct = a & 0xFFFFu;
        c.n_terms = a >> 16;
        c.n_seps = c.n_struct - c.n_terms;
        c.n_rows = c.n_terms;
        c.ord = ord;
        if (ht) {
            const uint32_t L = hibit(ht);
            c.has_term = 1;
            c.last_term = lane_value(base + (T ? 31u - (uint32_t)__clz(T) : 0u) + 1u, L);
        }
        // Last and second-to-last structural of the wave.
        const uint32_t H = hibit(hs);
        const uint32_t h = S ? 31u - (uint32_t)__clz(S) : 0u;
        const uint32_t S2 = S & ~(1u << h);
        c.last_struct = lane_value(base + h + 1u, H);
        const uint32_t p2 = lane_value(S2 ? base + (31u - (uint32_t)__clz(S2)) + 1u : 0u, H);
        const uint64_t below = hs & ~(1ULL << H);
        c.prev_struct = p2 ? p2 : (below ? lane_value(base + h + 1u, hibit(below)) : 0u);
    };
    finish(S0, T0, ht0, a0, od & 0xFFFFu, c0);
    finish(S1, T1, ht1, a1, od >> 16, c1);
}

// --------------------------------------------------------------------------
// block-wide scans over 256 threads
// --------------------------------------------------------------------------
// A walk round's scan: terminators (a) and rows (b), summed, and the last
// terminator's list index (c, max; -1 = none).
struct Scan3 {
    uint32_t a, b;
    int32_t c;
};

// Exclusive scan across the workgroup; `tot` gets the inclusive total. a
// and b travel as the 16-bit halves of one word (a round holds at most
// WG * ITEMS_MAX = 2048 entries), c + 1 >= 0 as a max scan; DPP within each
// wave, then the wave totals through `sh` (NWAVE words of each).
DEV Scan3 block_excl(Scan3 v, uint32_t* sh, Scan3& tot) {
    const uint32_t ln = __lane_id(), w = uni(threadIdx.x >> 6);
    const uint32_t p = v.a | (v.b << 16), m = (uint32_t)(v.c + 1);
    const uint32_t pi = wave_incl_sum(p), mi = wave_incl_max(m);
    const uint32_t px = pi - p, mx = dpp_prev_lane(mi);   // exclusive
    if (ln == 63) { sh[w] = pi; sh[NWAVE + w] = mi; }
    __syncthreads();
    uint32_t pb = 0, mb = 0, pt = 0, mt = 0;
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k) {
        const uint32_t pk = sh[k], mk = sh[NWAVE + k];
        if (k < w) { pb += pk; mb = max(mb, mk); }
        pt += pk;
        mt = max(mt, mk);
    }
    __syncthreads();
    tot = Scan3{pt & 0xFFFFu, pt >> 16, (int32_t)mt - 1};
    const uint32_t pe = pb + px, me = max(mb, mx);
    return Scan3{pe & 0xFFFFu, pe >> 16, (int32_t)me - 1};
}

// Exclusive sum across the workgroup; `tot` gets the total. DPP within
// each wave, the wave totals through `sh` (NWAVE words).
DEV uint32_t block_excl_sum(uint32_t v, uint32_t* sh, uint32_t& tot) {
    const uint32_t ln = __lane_id(), w = uni(threadIdx.x >> 6);
    const uint32_t inc = wave_incl_sum(v);
    if (ln == 63) sh[w] = inc;
    __syncthreads();
    uint32_t base = 0;
    tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k) {
        const uint32_t t = sh[k];
        if (k < w) base += t;
        tot += t;
    }
    __syncthreads();
    return base + inc - v;
}

template <class T, int NW = NWAVE>
DEV T block_reduce_sum(T v, T* sh) {
    // fixed-order tree: wave butterfly, then waves in order
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d);
    const uint32_t ln = __lane_id(), w = uni(threadIdx.x >> 6);
    if (ln == 0) sh[w] = v;
    __syncthreads();
    T t = sh[0];
    for (int k = 1; k < NW; ++k) t += sh[k];
    __syncthreads();
    return t;
}

template <bool HQ, bool ESC, bool CMT>
struct Dial {
    static constexpr int NC = ESC ? 2 : 1;
    static constexpr int NR = CMT ? 3 : (HQ ? 2 : 1);
    static constexpr int NS = NC * NR;
};

struct StepOut {
    uint64_t S, T, F;
    uint32_t s_out;
};

// Apply one block with known entry state s_in (state = c + NC * r).
template <bool HQ, bool ESC, bool CMT>
DEV StepOut block_step(const BM& m, uint32_t s_in, uint64_t ls) {
    constexpr int NC = Dial<HQ, ESC, CMT>::NC;
    uint32_t c = ESC ? (s_in % NC) : 0u;
    uint32_t r = s_in / NC;
    uint64_t rq = m.q;
    uint32_t c_out = 0;
    if (ESC) {
        // escaped_positions (formats.rs:334-343) with the Add carry.
        uint64_t phase = EVEN_BITS ^ m.esc;
        uint64_t partial = EVEN_BITS + phase;
        uint32_t c1 = partial < EVEN_BITS;
        uint64_t sum = partial + c;
        uint32_t c2 = sum < partial;
        c_out = c1 | c2;
        rq &= EVEN_BITS ^ sum;
    }
    StepOut o;
    o.F = 0;
    if (CMT) {
        uint64_t starts = m.cm & ((m.nl << 1) | ls);
        uint32_t rr = r;
        uint64_t inert = resolve_regions(HQ ? rq : 0ULL, starts, m.nl, rr);
        o.S = HQ ? (m.cand & ~inert) : m.cand;
        o.T = o.S & m.nl;
        uint64_t inert_prev = (inert << 1) | (uint64_t)(r != 0);
        o.F = (o.T & inert_prev) | (o.S & ~m.nl & inert);
        r = rr;
    } else if (HQ) {
        uint64_t inside = prefix_xor(rq) ^ (r ? ~0ULL : 0ULL);
        o.S = m.cand & ~inside;
        o.T = o.S & m.nl;
        r = (uint32_t)(inside >> 63);
    } else {
        o.S = m.cand;
        o.T = o.S & m.nl;
    }
    o.s_out = c_out + NC * r;
    return o;
}

// a or b, field by field: a conditional over two objects would select
// between their addresses and keep them in memory.
DEV StepOut pick(bool first, const StepOut& a, const StepOut& b) {
    StepOut o;
    o.S = first ? a.S : b.S;
    o.T = first ? a.T : b.T;
    o.F = first ? a.F : b.F;
    o.s_out = first ? a.s_out : b.s_out;
    return o;
}

// One block's step from each of N entry states, held in registers (the
// members are named, so nothing is indexed at run time).
template <int N>
struct BlockSteps;
template <>
struct BlockSteps<1> {
    StepOut s0;
    template <bool HQ, bool ESC, bool CMT>
    DEV void fill(const BM& m, uint64_t ls) { s0 = block_step<HQ, ESC, CMT>(m, 0u, ls); }
    DEV uint32_t fn() const { return s0.s_out; }
    DEV StepOut at(uint32_t) const { return s0; }
};
template <>
struct BlockSteps<3> {
    StepOut s0, s1, s2;
    template <bool HQ, bool ESC, bool CMT>
    DEV void fill(const BM& m, uint64_t ls) {
        s0 = block_step<HQ, ESC, CMT>(m, 0u, ls);
        s1 = block_step<HQ, ESC, CMT>(m, 1u, ls);
        s2 = block_step<HQ, ESC, CMT>(m, 2u, ls);
    }
    DEV uint32_t fn() const { return s0.s_out | (s1.s_out << 4) | (s2.s_out << 8); }
    DEV StepOut at(uint32_t e) const { return pick(e == 0u, s0, pick(e == 1u, s1, s2)); }
};
template <>
struct BlockSteps<6> {
    StepOut s0, s1, s2, s3, s4, s5;
    template <bool HQ, bool ESC, bool CMT>
    DEV void fill(const BM& m, uint64_t ls) {
        s0 = block_step<HQ, ESC, CMT>(m, 0u, ls);
        s1 = block_step<HQ, ESC, CMT>(m, 1u, ls);
        s2 = block_step<HQ, ESC, CMT>(m, 2u, ls);
        s3 = block_step<HQ, ESC, CMT>(m, 3u, ls);
        s4 = block_step<HQ, ESC, CMT>(m, 4u, ls);
        s5 = block_step<HQ, ESC, CMT>(m, 5u, ls);
    }
    DEV uint32_t fn() const {
        return s0.s_out | (s1.s_out << 4) | (s2.s_out << 8) | (s3.s_out << 12) | (s4.s_out << 16) | (s5.s_out << 20);
    }
    DEV StepOut at(uint32_t e) const {
        return pick(e < 3u, pick(e == 0u, s0, pick(e == 1u, s1, s2)), pick(e == 3u, s3, pick(e == 4u, s4, s5)));
    }
};

// Transfer function of one block over all entry states.
template <bool HQ, bool ESC, bool CMT>
DEV uint32_t block_fn(const BM& m, uint64_t ls) {
    constexpr int NS = Dial<HQ, ESC, CMT>::NS;
    if (!HQ && !CMT) return 0u;
    if (HQ && !ESC && !CMT) {
        uint32_t p = popc(m.q) & 1u;
        return p | ((1u ^ p) << 4);
    }
    if (HQ && ESC && !CMT) {
        uint32_t f = 0;
#pragma unroll
        for (uint32_t c = 0; c < 2; ++c) {
            uint64_t phase = EVEN_BITS ^ m.esc;
            uint64_t partial = EVEN_BITS + phase;
            uint32_t c1 = partial < EVEN_BITS;
            uint64_t sum = partial + c;
            uint32_t c2 = sum < partial;
            uint32_t p = popc(m.q & (EVEN_BITS ^ sum)) & 1u;
            uint32_t co = c1 | c2;
            f |= (co + 2 * (0u ^ p)) << (4 * (c + 0));
            f |= (co + 2 * (1u ^ p)) << (4 * (c + 2));
        }
        return f;
    }
    uint32_t f = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) f |= block_step<HQ, ESC, CMT>(m, (uint32_t)s, ls).s_out << (4 * s);
    return f;
}

// --------------------------------------------------------------------------
// SWAR classification of 32 bytes (a thread's run in the SWAR front)
// --------------------------------------------------------------------------
// Quote, candidate, newline, comment and escape bytes of a run, bit i for
// byte i (comment and escape bytes only for the dialects that have them).
struct Masks32 {
    uint32_t q, c, n, h, e;
};

// Byte transpose of four words: byte k of o_j is byte j of a_k (v_perm_b32:
// selector bytes 0-3 pick bytes of the second operand, 4-7 of the first).
DEV void transpose4(uint32_t a0, uint32_t a1, uint32_t a2, uint32_t a3, uint32_t& o0, uint32_t& o1, uint32_t& o2,
                    uint32_t& o3) {
    const uint32_t u0 = __builtin_amdgcn_perm(a1, a0, 0x05010400u);   // a0.0 a1.0 a0.1 a1.1
    const uint32_t u1 = __builtin_amdgcn_perm(a1, a0, 0x07030602u);   // a0.2 a1.2 a0.3 a1.3
    const uint32_t u2 = __builtin_amdgcn_perm(a3, a2, 0x05010400u);   // a2.0 a3.0 a2.1 a3.1
    const uint32_t u3 = __builtin_amdgcn_perm(a3, a2, 0x07030602u);   // a2.2 a3.2 a2.3 a3.3
    o0 = __builtin_amdgcn_perm(u2, u0, 0x05040100u);
    o1 = __builtin_amdgcn_perm(u2, u0, 0x07060302u);
    o2 = __builtin_amdgcn_perm(u3, u1, 0x05040100u);
    o3 = __builtin_amdgcn_perm(u3, u1, 0x07060302u);
}

// Inclusive prefix XOR over the bits of a word (bit i = parity of bits 0..i).
DEV uint32_t prefix_xor32(uint32_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    x ^= x << 16;
    return x;
}

// Bit 7 of each byte of y set when the byte equals c (< 0x80), from the
// word's low seven bits of each byte (y7) and its bytes below 0x80 (lo):
// (b ^ c) + 0x7F carries into bit 7 exactly when the low bits differ, and
// no byte carries into the next.
DEV uint32_t eq_ascii(uint32_t y7, uint32_t lo, uint32_t c4) { return lo & ~((y7 ^ c4) + 0x7F7F7F7Fu); }

// Words x[0..7] hold the run's bytes in order (byte i = byte i & 3 of
// x[i >> 2]); NSEP separator bytes besides '\n' (all below 0x80, as is the
// quote, comment and escape bytes).
template <bool HQ, int NSEP, bool CMT = false, bool ESC = false>
DEV Masks32 classify32(const uint32_t (&x)[8], const Params& P) {
    // y[j] byte k = run byte 8k + j: words x0 x2 x4 x6 hold bytes 8k + 0..3,
    // x1 x3 x5 x7 bytes 8k + 4..7.
    uint32_t y[8];
    transpose4(x[0], x[2], x[4], x[6], y[0], y[1], y[2], y[3]);
    transpose4(x[1], x[3], x[5], x[7], y[4], y[5], y[6], y[7]);
    const uint32_t q4 = HQ ? (uint32_t)P.quote * 0x01010101u : 0u;
    const uint32_t s0 = (NSEP > 0 ? P.sepb[0] : 0u) * 0x01010101u, s1 = (NSEP > 1 ? P.sepb[1] : 0u) * 0x01010101u;
    const uint32_t s2 = (NSEP > 2 ? P.sepb[2] : 0u) * 0x01010101u, s3 = (NSEP > 3 ? P.sepb[3] : 0u) * 0x01010101u;
    const uint32_t nl_keep = P.nl_cand != 0u ? ~0u : 0u;   // '\n' flags kept among the candidates
    const uint32_t h4 = CMT ? (uint32_t)P.comment * 0x01010101u : 0u;
    const uint32_t e4 = ESC ? (uint32_t)P.escape * 0x01010101u : 0u;
    Masks32 m{0u, 0u, 0u, 0u, 0u};
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const uint32_t y7 = y[j] & 0x7F7F7F7Fu, lo = ~y[j] & 0x80808080u;
        const uint32_t en = eq_ascii(y7, lo, 0x0A0A0A0Au);
        uint32_t ec = en & nl_keep;
        if (NSEP > 0) ec |= eq_ascii(y7, lo, s0);
        if (NSEP > 1) ec |= eq_ascii(y7, lo, s1);
        if (NSEP > 2) ec |= eq_ascii(y7, lo, s2);
        if (NSEP > 3) ec |= eq_ascii(y7, lo, s3);
        m.n |= (en >> 7) << j;
        m.c |= (ec >> 7) << j;
        if (HQ) m.q |= (eq_ascii(y7, lo, q4) >> 7) << j;
        if (CMT) m.h |= (eq_ascii(y7, lo, h4) >> 7) << j;
        if (ESC) m.e |= (eq_ascii(y7, lo, e4) >> 7) << j;
    }
    return m;
}

// The same masks from the 256-entry candidate table, a byte at a time (more
// than four separator bytes, or a byte of 0x80 or above).
template <bool HQ, bool CMT = false, bool ESC = false>
DEV Masks32 classify32_table(const uint32_t (&x)[8], const Params& P, const uint8_t* cls) {
    Masks32 m{0u, 0u, 0u, 0u, 0u};
    // A word at a time (not unrolled: four table reads in flight, not 32).
#pragma unroll 1
    for (int k = 0; k < 8; ++k) {
        uint32_t v = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j)
            if (j == k) v = x[j];
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const uint32_t c = (v >> (8 * b)) & 0xFFu, i = 4u * (uint32_t)k + (uint32_t)b;
            m.c |= (uint32_t)cls[c] << i;
            m.n |= (uint32_t)(c == '\n') << i;
            if (HQ) m.q |= (uint32_t)(c == (uint32_t)P.quote) << i;
            if (CMT) m.h |= (uint32_t)(c == (uint32_t)P.comment) << i;
            if (ESC) m.e |= (uint32_t)(c == (uint32_t)P.escape) << i;
        }
    }
    return m;
}

// The escape-dialect step (quote parity with escapes, no comment) of a
// thread's 32-byte run from each of its four entry states s = c + 2r (c:
// the previous byte escapes this run's first, r: inside quotes). The escape
// add-carry of escaped_positions (formats.rs:334-343) and the quote parity
// prefix are block_step's on 64-byte blocks cut at bit 32: both are chains
// of carries and parities that run upward only, so two runs stepped in turn
// give the block's masks, carry and state.
struct EscRun {
    uint32_t px0, px1, co0, co1;   // quote parity prefix and carry out, for c = 0 and c = 1
    DEV void fill(const Masks32& m) {
        constexpr uint32_t EVEN32 = 0x55555555u;
        const uint32_t partial = EVEN32 + (EVEN32 ^ m.e);
        const uint32_t c1 = partial < EVEN32 ? 1u : 0u;
        const uint32_t sum1 = partial + 1u;
        px0 = prefix_xor32(m.q & (EVEN32 ^ partial));
        px1 = prefix_xor32(m.q & (EVEN32 ^ sum1));
        co0 = c1;
        co1 = c1 | (sum1 < partial ? 1u : 0u);
    }
    DEV uint32_t out(uint32_t s) const {
        const uint32_t c = s & 1u, r = s >> 1;
        return (c ? co1 : co0) + 2u * (((c ? px1 : px0) >> 31) ^ r);
    }
    DEV uint32_t fn() const { return out(0u) | (out(1u) << 4) | (out(2u) << 8) | (out(3u) << 12); }
    // Structurals and terminators of the run entered in state s.
    DEV void masks(const Masks32& m, uint32_t s, uint32_t& S, uint32_t& T) const {
        const uint32_t inside = ((s & 1u) ? px1 : px0) ^ ((s >> 1) ? ~0u : 0u);
        S = m.c & ~inside;
        T = S & m.n;
    }
};

// --------------------------------------------------------------------------
// the tile kernel
// --------------------------------------------------------------------------
// Exact conversion of the f64 cells the fast paths leave undecided (see
// cells.h): one big decimal per workgroup in LDS, used by one lane at a
// time. Lanes of a wave take turns by ballot; waves take an LDS lock.
struct SlowF64 {
    BigDec big;
    uint32_t lock;
};

DEV void slow_lock(SlowF64& s) {
    while (atomicCAS(&s.lock, 0u, 1u) != 0u) __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

DEV void slow_unlock(SlowF64& s) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    atomicExch(&s.lock, 0u);
}

// Called by exactly the lanes whose cell is deferred.
template <class Src>
DEV double resolve_f64(SlowF64& s, const Src& src, uint32_t from, uint32_t to, CellDialect cd) {
    uint64_t m = ballot(true);
    double v = 0.0;
    while (m) {
        const uint32_t l = (uint32_t)__ffsll((long long)m) - 1u;
        if (__lane_id() == l) {
            slow_lock(s);
            v = bits_to_f64(f64_field_exact_bits(src, from, to, cd, s.big));
            slow_unlock(s);
        }
        m &= m - 1;
    }
    return v;
}

template <class Src>
DEV void resolve_score(SlowF64& s, const Src& src, NdRecord& r) {
    uint64_t m = ballot(true);
    while (m) {
        const uint32_t l = (uint32_t)__ffsll((long long)m) - 1u;
        if (__lane_id() == l) {
            slow_lock(s);
            ndjson_resolve_exact(src, r, s.big);
            slow_unlock(s);
        }
        m &= m - 1;
    }
}

struct Shared {
    uint8_t lds_bytes[TILE_LDS];   // from tile start - PRE
    uint16_t list[TILE];
    uint8_t cls[256];
    uint32_t wfn[NWAVE];
    uint32_t wentry[NWAVE];
    uint32_t wcount[NWAVE];
    uint32_t wbase[NWAVE + 1];
    Counts pre;       // tile exclusive prefix
    Counts inc;       // tile inclusive
    uint32_t tile;
    uint32_t scan_sh[2 * NWAVE];
    uint32_t sum_sh[NWAVE];
    double dsum_sh[NWAVE];
    unsigned long long usum_sh[NWAVE];
    long long lsum_sh[NWAVE];
    uint32_t str_len[FALX_MAX_COLUMNS];   // tile cleaned bytes per str slot
    SlowF64 slow;                         // exact path for undecided f64 cells
    uint32_t fix_bad;                     // *_FAST passes: a cell here needs the full parsers
};

DEV uint32_t list_at(const Shared& sh, uint32_t k) { return sh.list[k]; }

// Per-cell outcome folded into the column statistics.
struct CellStat {
    bool ok, fx;
    unsigned long long csum, bytes;
    long long fxm;
    double d;
};

// One thread's share of one tile: counts and byte totals fit 32 bits.
struct Acc {
    uint32_t values, bytes, fx_cells;
    unsigned long long csum;
    long long fx_sum;
    double dsum;
    DEV void add(const CellStat& o) {
        if (!o.ok) return;
        values += 1; csum += o.csum; bytes += (uint32_t)o.bytes; dsum += o.d;
        if (o.fx) { fx_sum += o.fxm; fx_cells += 1; }
    }
};

// Lane value v summed over the wave's four rows of 16 lanes (DPP within a
// row, then the rows in order); the result is wave-uniform. All lanes must
// be active.
DEV uint32_t rows_sum_u32(uint32_t v) {
    v = row_sum16(v);
    return ((lane_value(v, 0) + lane_value(v, 16)) + lane_value(v, 32)) + lane_value(v, 48);
}
DEV unsigned long long lane_u64(unsigned long long v, uint32_t l) {
    return ((unsigned long long)lane_value((uint32_t)(v >> 32), l) << 32) | lane_value((uint32_t)v, l);
}
DEV unsigned long long rows_sum_u64(unsigned long long v) {
    v = row_sum16(v);
    return ((lane_u64(v, 0) + lane_u64(v, 16)) + lane_u64(v, 32)) + lane_u64(v, 48);
}
DEV double rows_sum_f64(double v) {
    v = row_sum16(v);
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const auto at = [&](uint32_t l) { return __longlong_as_double((long long)lane_u64(b, l)); };
    return ((at(0) + at(16)) + at(32)) + at(48);
}

// Per-tile statistics of up to N columns (accumulator j holds column
// cols[j], j < nc) plus the record count (with_records), reduced in a fixed order (the same in every pass, so a tile's
// partials are bit-identical whichever pass wrote them): waves 1-3 leave
// their lanes' accumulators in the LDS exchange area xs, wave 0 adds them to
// its own lane by lane (waves in order), sums each row of 16 lanes by DPP and
// then the four row sums in order, and writes the StatParts. Only the fields
// a column's type can make non-zero are reduced; counts of values and of
// fixed-6 cells travel as the 16-bit halves of one word (a tile holds at most
// 8193 records). Called by every thread; xs must not be in use.
template <uint32_t N>
DEV void publish_stats(const Params& P, uint32_t tile, const Acc (&acc)[N], const uint32_t (&cols)[N], uint32_t nc,
                       unsigned long long records, bool with_records, unsigned long long* xs) {
    constexpr uint32_t NSLOT = 4 * N + 1;   // per column: counts | bytes, checksum, fixed-6 sum, FP64 sum; records
    const uint32_t ln = __lane_id(), w = uni(threadIdx.x >> 6);
    __syncthreads();   // every thread is done with what xs held
    if (w > 0) {
        unsigned long long* o = xs + (w - 1) * NSLOT * 64 + ln;
#pragma unroll
        for (uint32_t j = 0; j < N; ++j) {
            if (j >= nc) break;
            o[(4 * j + 0) * 64] =
                (unsigned long long)(acc[j].values | (acc[j].fx_cells << 16)) | ((unsigned long long)acc[j].bytes << 32);
            o[(4 * j + 1) * 64] = acc[j].csum;
            o[(4 * j + 2) * 64] = (unsigned long long)acc[j].fx_sum;
            o[(4 * j + 3) * 64] = (unsigned long long)__double_as_longlong(acc[j].dsum);
        }
        o[4 * N * 64] = records;
    }
    __syncthreads();
    if (w != 0) return;
    const unsigned long long* in = xs + ln;
    StatPart* parts = reinterpret_cast<StatPart*>(P.partials);
#pragma unroll
    for (uint32_t j = 0; j < N; ++j) {
        if (j >= nc) break;
        const uint32_t type = P.cols[cols[j]].type;
        const bool is_f64 = type == FALX_COL_F64, is_num = is_f64 || type == FALX_COL_I64;
        uint32_t vc = acc[j].values | (acc[j].fx_cells << 16), by = acc[j].bytes;
        unsigned long long cs = acc[j].csum, fx = (unsigned long long)acc[j].fx_sum;
        double ds = acc[j].dsum;
#pragma unroll 1
        for (uint32_t k = 0; k < (uint32_t)NWAVE - 1; ++k) {
            const unsigned long long* q = in + k * NSLOT * 64;
            const unsigned long long c0 = q[(4 * j + 0) * 64];
            vc += (uint32_t)c0;
            by += (uint32_t)(c0 >> 32);
            cs += q[(4 * j + 1) * 64];
            if (is_f64) fx += q[(4 * j + 2) * 64];
            if (is_num) ds += __longlong_as_double((long long)q[(4 * j + 3) * 64]);
        }
        const uint32_t tvc = rows_sum_u32(vc);
        const unsigned long long tcs = rows_sum_u64(cs);
        const uint32_t tby = is_num ? 0u : rows_sum_u32(by);
        const unsigned long long tfx = is_f64 ? rows_sum_u64(fx) : 0ULL;
        const double tds = is_num ? rows_sum_f64(ds) : 0.0;
        if (ln == 0) {
            StatPart& o = parts[(size_t)tile * P.n_cols + cols[j]];
            o.values = tvc & 0xFFFFu;
            o.checksum = tcs;
            o.bytes = tby;
            o.fixed6_sum = (long long)tfx;
            o.fixed6_cells = tvc >> 16;
            o.sum = tds;
        }
    }
    if (with_records) {
        unsigned long long rec = records;
#pragma unroll
        for (uint32_t k = 0; k < (uint32_t)NWAVE - 1; ++k) rec += in[k * NSLOT * 64 + 4 * N * 64];
        const unsigned long long t = rows_sum_u64(rec);
        if (ln == 0) reinterpret_cast<unsigned long long*>(parts + (size_t)P.n_tiles * P.n_cols)[tile] = t;
    }
}

// Fixed-6 cells read straight from the tile's LDS copy: the last 8 bytes
// come from three aligned words (fixed6_tail in cells.h decides). [lf, lt)
// are offsets into the copy.
DEV bool fixed6_lds(const uint8_t* lds, uint32_t lf, uint32_t lt, F64Result& r) {
    const uint32_t len = lt - lf;
    if (len < 8 || len > 11) return false;
    const uint32_t s8 = lt - 8, a = s8 & ~3u, sh8 = (s8 & 3u) * 8u;
    // Up to 3 bytes past byte lt are read: inside the copy's spare room,
    // and never used.
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(lds + a);
    const uint32_t w0 = wp[0], w1 = wp[1], w2 = wp[2];
    const uint32_t lo = sh8 ? (w0 >> sh8) | (w1 << (32 - sh8)) : w0;
    const uint32_t hi = sh8 ? (w1 >> sh8) | (w2 << (32 - sh8)) : w1;
    return fixed6_tail(((uint64_t)hi << 32) | lo, lds + lf, len - 8, r);
}

// The 4 NW bytes of the LDS copy that end at offset lt, as NW words (the
// window num_swar reads): NW + 1 aligned LDS words funnel-shifted by lt mod
// 4. Words below the copy's first byte read as word 0 and only feed window
// bytes below the cell, which num_swar masks off. The highest word read
// ends at most 3 bytes past lt, inside the copy's spare room.
template <int NW>
DEV Win<NW> win_lds(const uint8_t* lds, uint32_t lt) {
    const uint32_t* wp = reinterpret_cast<const uint32_t*>(lds);
    const int32_t q = (int32_t)((lt + 32u) >> 2) - 8 - NW;   // word of local byte lt - 4 NW (may be < 0)
    const uint32_t sh8 = (lt & 3u) * 8u;
    uint32_t r[NW + 1];
#pragma unroll
    for (int k = 0; k <= NW; ++k) r[k] = wp[max(q + k, 0)];
    Win<NW> W;
#pragma unroll
    for (int k = 0; k < NW; ++k) W.w[k] = (uint32_t)((((uint64_t)r[k + 1] << 32) | r[k]) >> sh8);
    return W;
}

// A numeric cell read as f64 (f) or as i64 (iok, iv).
struct NumVal {
    F64Result f;
    bool iok;
    int64_t iv;
};

// One list entry with its record context.
struct Entry {
    bool valid, is_term, flag;
    uint32_t pos;   // byte offset of the structural (len for the virtual terminator)
    uint32_t fs;    // first byte of the field it ends
    uint32_t rs;    // first byte of its record
    uint32_t ord;   // field ordinal within the record
    uint32_t row;   // row of its record
    uint32_t k;     // index in the tile's list
};

// Consecutive list entries per thread in a walk: the tile's mean number of
// fields per record, 1..ITEMS_MAX, so that with regular records every lane
// is at the same field of its record at each step.
constexpr uint32_t ITEMS_MAX = 8;

// Walk kinds: plain, or with a per-entry weight scanned before the visits.
struct NoWeight { static constexpr bool value = false; };
struct Weighted { static constexpr bool value = true; };

// One tile of one pass (MODE) over the document P describes. Called by every
// thread of the workgroup; returns with the workgroup converged (every
// early return is uniform).
template <bool HQ, bool ESC, bool CMT, int KMODE>
__device__ __forceinline__ void scan_tile(const Params& P, const uint32_t tile, Shared& sh) {
    constexpr int NS = Dial<HQ, ESC, CMT>::NS;
    // A *_FAST pass is its pass without the stream parsers (and their
    // register and code footprint): cells they would take flag the tile.
    constexpr bool NUMFAST = KMODE == M_NUM_FAST || KMODE == M_NUM_FAST_STATS || KMODE == M_NUM_STATS_FAST;
    static_assert(!NUMFAST || (!ESC && !CMT), "numeric fast forms: dialects without escape or comment");
    constexpr bool FAST =
        NUMFAST || KMODE == M_COLUMNS_FAST || KMODE == M_COLUMNS_FAST_STATS || KMODE == M_STATS_FAST;
    constexpr int MODE = (KMODE == M_COLUMNS_FAST || KMODE == M_COLUMNS_FAST_STATS || KMODE == M_NUM_FAST ||
                          KMODE == M_NUM_FAST_STATS)
                             ? M_COLUMNS
                             : ((KMODE == M_STATS_FAST || KMODE == M_NUM_STATS_FAST) ? M_STATS : KMODE);
    const uint32_t tid = threadIdx.x, ln = tid & 63u, w = uni(tid >> 6);  // w is wave-uniform

    if (tid == 0) {
        sh.slow.lock = 0;
        sh.fix_bad = 0;
    }
    {
        uint32_t b = tid;
        sh.cls[b] = (uint8_t)((P.cand[b >> 5] >> (b & 31)) & 1u);
    }
    __syncthreads();
    const uint32_t tile_start = tile * (uint32_t)TILE;
    const uint32_t tile_len = min((uint32_t)TILE, P.len - tile_start);
    const bool last_tile = tile + 1 == P.n_tiles;
    // Bytes [lds_lo, lds_hi) of the document are in LDS; tile-local byte i
    // is tb[i], document byte x is sh.lds_bytes[x - tile_start + PRE]. The
    // passes that read cells load the bytes around the tile too.
    constexpr bool HALOED = MODE == M_COLUMNS || MODE == M_STATS || MODE == M_STRINGS;
    const uint32_t lds_lo = HALOED ? tile_start - min((uint32_t)PRE, tile_start) : tile_start;
    const uint32_t lds_hi = tile_start + min((uint32_t)(HALOED ? TILE + HALO : TILE), P.len - tile_start);
    uint8_t* const tb = sh.lds_bytes + PRE;
    // The tile's quote bytes as a bitmap (bit i of word k: tile byte 32k + i)
    // for the cell passes of doubled-quote dialects: a cell quoted at both
    // ends with no other quote byte cleans to its inner bytes, which are then
    // copied like an unquoted cell's instead of streamed byte by byte.
    constexpr bool QBITS = HQ && !ESC && HALOED;
    uint32_t* qbits = nullptr;
    if constexpr (QBITS) {
        __shared__ uint32_t qbits_arr[TILE / 32];
        qbits = qbits_arr;
    }

    // Summary mode's end: the tile summary from the waves' counters
    // (wsum[w * NS + s]: wave w entered in state s) and functions (sh.wfn).
    Counts* const wsum = reinterpret_cast<Counts*>(sh.list);   // the list is unused in summary mode
    auto publish_summary = [&]() __attribute__((always_inline)) {
        __syncthreads();
        uint32_t* out = P.tsum + (size_t)tile * P.ts_stride;
        if (tid < (uint32_t)NS) {
            uint32_t s = tid;
            Counts c = counts_zero();
            for (int k = 0; k < NWAVE; ++k) {
                c = compose(c, wsum[k * NS + s]);
                s = fn_apply(sh.wfn[k], s);
            }
            c7_store(out + TS_COUNTS + C7 * tid, c);
        }
        if (tid == 0) {
            uint32_t f = fn_ident<NS>();
            for (int k = 0; k < NWAVE; ++k) {
                f = fn_then<NS>(f, sh.wfn[k]);
                out[1 + k] = sh.wfn[k];
            }
            out[0] = f;
        }
    };
    // Emit modes: each wave's entry state from the tile prefix and the wave
    // functions of the summary, and the tile's exclusive and inclusive
    // counters (the prefix, then the tile's own counters for its entry state).
    auto load_entry_states = [&]() __attribute__((always_inline)) {
        if (tid == 0) {
            const uint32_t* tp = P.tpre + (size_t)tile * 8;
            const uint32_t* ts = P.tsum + (size_t)tile * P.ts_stride;
            const uint32_t s0 = tp[0];
            uint32_t s = s0;
            for (int k = 0; k < NWAVE; ++k) {
                sh.wentry[k] = s;
                s = fn_apply(ts[1 + k], s);
            }
            sh.pre = c7_load(tp + 1);
            sh.inc = compose(sh.pre, c7_load(ts + TS_COUNTS + C7 * s0));
        }
        __syncthreads();
    };

    // Dialects with neither escape nor comment take the SWAR front below; the
    // others classify by ballots with one lane per 64-byte block.
    constexpr bool SWAR = !(ESC && CMT);
    if constexpr (SWAR) {
        // ---- SWAR front: thread t holds tile bytes [32t, 32t + 32) ---------
        // The bytes come from HBM into registers (and into the LDS copy the
        // cell readers use) and are classified four to a word. A byte
        // transpose puts bytes 8k + j of the thread's run into byte k of
        // word j, so a candidate's per-byte equality flags (bit 7 of each
        // byte) fold into one 32-bit mask in byte order with a shift-or per
        // word. Each wave covers the same 2 KiB as in the ballot form, so
        // the two forms' tile summaries are interchangeable.
        const uint32_t t0 = tid * 32u;
        uint32_t x[8];
        if (t0 + 32u <= tile_len) {
            const uint4 a = *reinterpret_cast<const uint4*>(P.data + tile_start + t0);
            const uint4 b = *reinterpret_cast<const uint4*>(P.data + tile_start + t0 + 16u);
            x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
            x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                uint32_t v = 0;
                for (uint32_t i = 0; i < 4; ++i) {
                    const uint32_t at = t0 + 4u * (uint32_t)k + i;
                    if (at < tile_len) v |= (uint32_t)P.data[tile_start + at] << (8 * i);
                }
                x[k] = v;
            }
        }
        reinterpret_cast<uint4*>(tb + t0)[0] = make_uint4(x[0], x[1], x[2], x[3]);
        reinterpret_cast<uint4*>(tb + t0)[1] = make_uint4(x[4], x[5], x[6], x[7]);
        if (HALOED && tid < (uint32_t)((PRE + HALO) / 16)) {
            // PRE bytes before the tile and HALO bytes after it, zero outside
            // the document. LDS index i holds document byte tile_start - PRE + i.
            const uint32_t lo_i = lds_lo + PRE - tile_start, hi_i = lds_hi + PRE - tile_start;
            const uint32_t i = tid < (uint32_t)(PRE / 16) ? tid * 16u : (uint32_t)(PRE + TILE) + (tid - PRE / 16) * 16u;
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i >= lo_i && i + 16 <= hi_i) {
                v = *reinterpret_cast<const uint4*>(P.data + (tile_start - PRE + i));
            } else if (i >= lo_i && i < hi_i) {
                uint8_t tmp[16];
                for (int k = 0; k < 16; ++k) tmp[k] = (i + k < hi_i) ? P.data[tile_start - PRE + i + k] : 0;
                memcpy(&v, tmp, 16);
            }
            *reinterpret_cast<uint4*>(&sh.lds_bytes[i]) = v;
        }
        const uint32_t live32 = t0 + 32u <= tile_len ? ~0u : (t0 < tile_len ? (1u << (tile_len - t0)) - 1u : 0u);
        Masks32 m{0u, 0u, 0u};
        // Candidate and quote bytes below 0x80 with at most four separator
        // bytes take the word compares; other dialects the byte table.
        const uint32_t qa = HQ ? (uint32_t)P.quote : 0u, ha = CMT ? (uint32_t)P.comment : 0u;
        const uint32_t ea = ESC ? (uint32_t)P.escape : 0u;
        const bool ascii = qa < 0x80u && ha < 0x80u && ea < 0x80u && (P.n_sepb < 1 || P.sepb[0] < 0x80u) &&
                           (P.n_sepb < 2 || P.sepb[1] < 0x80u) && (P.n_sepb < 3 || P.sepb[2] < 0x80u) &&
                           (P.n_sepb < 4 || P.sepb[3] < 0x80u);
        switch (ascii ? P.n_sepb : 5u) {
        case 0: m = classify32<HQ, 0, CMT, ESC>(x, P); break;
        case 1: m = classify32<HQ, 1, CMT, ESC>(x, P); break;
        case 2: m = classify32<HQ, 2, CMT, ESC>(x, P); break;
        case 3: m = classify32<HQ, 3, CMT, ESC>(x, P); break;
        case 4: m = classify32<HQ, 4, CMT, ESC>(x, P); break;
        default: m = classify32_table<HQ, CMT, ESC>(x, P, sh.cls); break;
        }
        m.q &= live32;
        m.c &= live32;
        m.n &= live32;
        m.h &= live32;
        m.e &= live32;
        if constexpr (QBITS) qbits[tid] = m.q;
        if constexpr (CMT) {
            // Comment dialects: the NORMAL/QUOTE/COMMENT machine runs on the
            // thread's run as the ballot form runs it on a 64-byte block (the
            // run's masks are the low halves of the block masks; the high half
            // holds no byte and leaves the state as it is), then the runs'
            // transfer functions are scanned over the wave's lanes.
            const BM mine{(uint64_t)m.c, (uint64_t)m.n, (uint64_t)m.q, 0ULL, (uint64_t)m.h};
            // Line start at the run's first byte: the previous run ends in
            // '\n' (the wave's first run reads the byte before it).
            const uint32_t prev_n = __shfl_up(m.n, 1);
            const uint32_t at0 = tile_start + t0;
            // (A run past the tile's end holds no byte: its line start is moot.)
            const bool peek = ln == 0 && at0 > 0u && t0 <= tile_len;
            const uint32_t ls_first = at0 == 0u ? 1u : (peek ? (uint32_t)(P.data[at0 - 1u] == '\n') : 0u);
            const uint64_t ls = ln == 0 ? (uint64_t)ls_first : (uint64_t)(prev_n >> 31);
            BlockSteps<NS> steps;
            steps.template fill<HQ, false, true>(mine, ls);
            uint32_t f_incl = steps.fn();
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t g = __shfl_up(f_incl, d);
                if (ln >= (uint32_t)d) f_incl = fn_then<NS>(g, f_incl);
            }
            uint32_t f_excl = __shfl_up(f_incl, 1);
            if (ln == 0) f_excl = fn_ident<NS>();
            __syncthreads();   // the LDS copy is complete
            if (MODE == M_SUMMARY) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    const StepOut o = steps.at(fn_apply(f_excl, (uint32_t)s));
                    const Counts c = wave_counts(o.S, o.T, o.F, tile_start + t0);
                    if (ln == 0) wsum[w * NS + s] = c;
                }
                if (ln == 63) sh.wfn[w] = f_incl;
                publish_summary();
                return;
            }
            load_entry_states();
            {
                const StepOut o = steps.at(fn_apply(f_excl, uni(sh.wentry[w])));
                const uint32_t S = (uint32_t)o.S, T = (uint32_t)o.T, F = (uint32_t)o.F;
                const uint32_t n = (uint32_t)__popc(S);
                const uint32_t incl = wave_incl_sum(n);
                if (ln == 63) sh.wcount[w] = incl;
                __syncthreads();
                uint32_t wb = 0;
#pragma unroll
                for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k)
                    if (k < w) wb += sh.wcount[k];
                uint16_t* lst = sh.list + uni(wb) + (incl - n);
                for (uint32_t rem = S; rem; rem &= rem - 1u) {
                    const uint32_t bit = (uint32_t)__ffs((int)rem) - 1u;
                    *lst++ = (uint16_t)((t0 + bit) | (((T >> bit) & 1u) << 14) | (((F >> bit) & 1u) << 15));
                }
            }
        } else if constexpr (ESC) {
            // Escape dialects (no comment): each run's four entry states
            // (EscRun), the runs' functions scanned over the wave's lanes.
            static_assert(NS == 4, "escape dialects without comment: four entry states");
            EscRun er;
            er.fill(m);
            uint32_t f_incl = er.fn();
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t g = __shfl_up(f_incl, d);
                if (ln >= (uint32_t)d) f_incl = fn_then<NS>(g, f_incl);
            }
            uint32_t f_excl = __shfl_up(f_incl, 1);
            if (ln == 0) f_excl = fn_ident<NS>();
            __syncthreads();   // the LDS copy is complete
            if (MODE == M_SUMMARY) {
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    uint32_t S, T;
                    er.masks(m, fn_apply(f_excl, (uint32_t)s), S, T);
                    const Counts c = wave_counts((uint64_t)S, (uint64_t)T, 0ULL, tile_start + t0);
                    if (ln == 0) wsum[w * NS + s] = c;
                }
                if (ln == 63) sh.wfn[w] = f_incl;
                publish_summary();
                return;
            }
            load_entry_states();
            {
                uint32_t S, T;
                er.masks(m, fn_apply(f_excl, uni(sh.wentry[w])), S, T);
                const uint32_t n = (uint32_t)__popc(S);
                const uint32_t incl = wave_incl_sum(n);
                if (ln == 63) sh.wcount[w] = incl;
                __syncthreads();
                uint32_t wb = 0;
#pragma unroll
                for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k)
                    if (k < w) wb += sh.wcount[k];
                uint16_t* lst = sh.list + uni(wb) + (incl - n);
                for (uint32_t rem = S; rem; rem &= rem - 1u) {
                    const uint32_t bit = (uint32_t)__ffs((int)rem) - 1u;
                    *lst++ = (uint16_t)((t0 + bit) | (((T >> bit) & 1u) << 14));
                }
            }
        } else {
        // Quote parity: inclusive prefix from the thread's first byte, the
        // quotes of the wave's earlier threads (mbcnt over the odd ones) and
        // the wave's function (swap when its quote count is odd).
        const uint32_t pref = HQ ? prefix_xor32(m.q) : 0u;
        const uint64_t odd = ballot(HQ && (__popc(m.q) & 1));
        const uint32_t ex = mbcnt(odd) & 1u;
        const uint32_t wave_fn = NS == 2 ? ((popc(odd) & 1u) ? 0x01u : 0x10u) : 0u;
        __syncthreads();   // the LDS copy is complete
        if (MODE == M_SUMMARY) {
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const uint32_t inside = pref ^ (((uint32_t)s ^ ex) ? ~0u : 0u);
                const uint32_t S = HQ ? (m.c & ~inside) : m.c, T = S & m.n;
                const Counts c = wave_counts((uint64_t)S, (uint64_t)T, 0ULL, tile_start + t0);
                if (ln == 0) wsum[w * NS + s] = c;
            }
            if (ln == 0) sh.wfn[w] = wave_fn;
            publish_summary();
            return;
        }
        load_entry_states();
        {
            const uint32_t entry = HQ ? (uni(sh.wentry[w]) ^ ex) : 0u;
            const uint32_t inside = pref ^ (entry ? ~0u : 0u);
            const uint32_t S = HQ ? (m.c & ~inside) : m.c, T = S & m.n;
            // Structurals and terminators counted in the two halves of one
            // word (a tile holds at most 8192 of each).
            const uint32_t n = (uint32_t)__popc(S) | ((uint32_t)__popc(T) << 16);
            const uint32_t incl = wave_incl_sum(n);
            if (ln == 63) sh.wcount[w] = incl;
            __syncthreads();
            uint32_t wb = 0, tot = 0;
#pragma unroll
            for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k) {
                const uint32_t c = sh.wcount[k];
                if (k < w) wb += c;
                tot += c;
            }
            const uint32_t at = uni(wb) + (incl - n);
            uint16_t* lst = sh.list + (at & 0xFFFFu);
            if constexpr (NUMFAST) {
                // The numeric fast forms also list the terminators (list
                // index by rank) at the list's far end; a tile whose lists
                // would meet is left to the full pass.
                const bool fits = (tot & 0xFFFFu) <= (uint32_t)(TILE - TL_CAP) && (tot >> 16) <= (uint32_t)TL_CAP;
                uint16_t* tl = sh.list + (TILE - TL_CAP);
                uint32_t rank = at >> 16, idx = at & 0xFFFFu;
                for (uint32_t rem = S; rem; rem &= rem - 1u) {
                    const uint32_t bit = (uint32_t)__ffs((int)rem) - 1u;
                    const uint32_t t = (T >> bit) & 1u;
                    *lst++ = (uint16_t)((t0 + bit) | (t << 14));
                    if (t && fits) tl[rank] = (uint16_t)idx;
                    rank += t;
                    idx += 1u;
                }
                if (tid == 0 && !fits) sh.fix_bad = 1u;
            } else {
                for (uint32_t rem = S; rem; rem &= rem - 1u) {
                    const uint32_t bit = (uint32_t)__ffs((int)rem) - 1u;
                    *lst++ = (uint16_t)((t0 + bit) | (((T >> bit) & 1u) << 14));
                }
            }
            __syncthreads();   // every thread has read sh.wcount
            if (ln == 63) sh.wcount[w] = incl & 0xFFFFu;
        }
        }
    } else {
        // ---- load: 16-byte coalesced reads into LDS, zero padded --------------
        // LDS index i holds document byte tile_start - PRE + i for i in [lo_i, hi_i).
        const uint32_t lo_i = lds_lo + PRE - tile_start, hi_i = lds_hi + PRE - tile_start;
        for (uint32_t i = (HALOED ? 0u : (uint32_t)PRE) + tid * 16; i < (uint32_t)(HALOED ? PRE + TILE + HALO : PRE + TILE);
             i += WG * 16) {
            uint4 v = make_uint4(0, 0, 0, 0);
            if (i >= lo_i && i + 16 <= hi_i) {
                v = *reinterpret_cast<const uint4*>(P.data + (tile_start - PRE + i));
            } else if (i >= lo_i && i < hi_i) {
                uint8_t tmp[16];
                for (int k = 0; k < 16; ++k) tmp[k] = (i + k < hi_i) ? P.data[tile_start - PRE + i + k] : 0;
                memcpy(&v, tmp, 16);
            }
            *reinterpret_cast<uint4*>(&sh.lds_bytes[i]) = v;
        }
        __syncthreads();

        // line-start carry entering this wave's first block
        uint64_t ls0;
        {
            uint32_t wb = w * BPW * 64;
            uint32_t v;
            if (wb > 0) v = tb[wb - 1] == '\n';
            else v = tile_start == 0 ? 1u : (uint32_t)(P.data[tile_start - 1] == '\n');
            ls0 = uni(v);
        }
        const uint8_t qb = (uint8_t)P.quote, eb = (uint8_t)P.escape, cb = (uint8_t)P.comment;

        const auto live_mask = [&](uint32_t off) -> uint64_t {
            return off + 64 <= tile_len ? ~0ULL : (off < tile_len ? ((1ULL << (tile_len - off)) - 1) : 0ULL);
        };

        // ---- classification: ballots per block, then lane b owns block b ------
        // Each wave classifies its BPW blocks with one byte per lane and
        // __ballot, and writes block b's masks into lane b: the mask algebra,
        // the counters and the compaction then run for all blocks at once, one
        // per lane, instead of block by block on the scalar unit.
        // The candidate bytes other than '\n' are tested by up to four compares,
        // each its own ballot ORed as scalar masks (NSEP = their number; '\n'
        // reuses the newline ballot), or by the 256-entry table (NSEP < 0);
        // eight blocks' bytes are read before any is used.
        BM mine{0ULL, 0ULL, 0ULL, 0ULL, 0ULL};
        const bool nl_cand = P.nl_cand != 0u;
        auto classify_all = [&](auto nsep) __attribute__((always_inline)) {
            constexpr int NSEP = decltype(nsep)::value;
            const uint32_t s0 = P.sepb[0], s1 = P.sepb[1], s2 = P.sepb[2], s3 = P.sepb[3];
    #pragma unroll
            for (int b0 = 0; b0 < BPW; b0 += 8) {
                uint32_t cc[8];
    #pragma unroll
                for (int i = 0; i < 8; ++i) cc[i] = tb[(w * BPW + b0 + i) * 64 + ln];
    #pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const uint32_t b = (uint32_t)(b0 + i);
                    const uint32_t c = cc[i];
                    const uint64_t nl = ballot(c == '\n');
                    uint64_t cand;
                    if (NSEP < 0) {
                        cand = ballot(sh.cls[c] != 0);
                    } else {
                        cand = nl_cand ? nl : 0ULL;
                        if (NSEP > 0) cand |= ballot(c == s0);
                        if (NSEP > 1) cand |= ballot(c == s1);
                        if (NSEP > 2) cand |= ballot(c == s2);
                        if (NSEP > 3) cand |= ballot(c == s3);
                    }
                    mine.cand = put_lane(mine.cand, cand, b);
                    mine.nl = put_lane(mine.nl, nl, b);
                    if (HQ) mine.q = put_lane(mine.q, ballot(c == qb), b);
                    if (ESC) mine.esc = put_lane(mine.esc, ballot(c == eb), b);
                    if (CMT) mine.cm = put_lane(mine.cm, ballot(c == cb), b);
                }
            }
        };
        switch (P.n_sepb) {
        case 0: classify_all(std::integral_constant<int, 0>{}); break;
        case 1: classify_all(std::integral_constant<int, 1>{}); break;
        case 2: classify_all(std::integral_constant<int, 2>{}); break;
        case 3: classify_all(std::integral_constant<int, 3>{}); break;
        case 4: classify_all(std::integral_constant<int, 4>{}); break;
        default: classify_all(std::integral_constant<int, -1>{}); break;
        }
        const bool own = ln < (uint32_t)BPW;
        const uint32_t off = (w * BPW + (own ? ln : 0u)) * 64;      // this lane's block in the tile
        const uint64_t live = own ? live_mask(off) : 0ULL;
        if constexpr (QBITS) {
            if (own) {
                const uint64_t qm = mine.q & live;
                qbits[off >> 5] = (uint32_t)qm;
                qbits[(off >> 5) + 1] = (uint32_t)(qm >> 32);
            }
        }
        // Line-start carry: the previous block ends in '\n' (ls0 for the first).
        const uint32_t prev_hi = __shfl_up((uint32_t)(mine.nl >> 32), 1);
        const uint64_t ls = ln == 0 ? ls0 : (uint64_t)(prev_hi >> 31);
        // Exclusive scan of the blocks' transfer functions over the lanes.
        // Comment dialects resolve regions per entry state: each state's step of
        // this lane's block is computed once, for its transfer function and for
        // the passes below, which pick the step of the block's entry state.
        constexpr bool STEPS = CMT;
        BlockSteps<STEPS ? NS : 1> steps;
        const auto step_of = [&](uint32_t entry) __attribute__((always_inline)) -> StepOut {
            if constexpr (STEPS) return steps.at(entry);
            else return block_step<HQ, ESC, CMT>(mine, entry, ls);
        };
        // Exclusive scan of the blocks' functions over the lanes (f_excl), and
        // the wave's whole function (f_incl in lane 63).
        uint32_t f_incl, f_excl;
        if constexpr (NS == 1) {
            f_incl = f_excl = 0u;
        } else if constexpr (HQ && !ESC && !CMT) {
            // Outside and inside quotes: a block flips the state when it holds
            // an odd number of quotes, so a lane's entry function is the parity
            // of the odd blocks below it.
            const uint64_t odd = ballot(own && (popc(mine.q) & 1u));
            f_excl = (mbcnt(odd) & 1u) ? 0x01u : 0x10u;
            f_incl = (popc(odd) & 1u) ? 0x01u : 0x10u;
        } else {
            if constexpr (STEPS) {
                steps.template fill<HQ, ESC, CMT>(mine, ls);
                f_incl = own ? steps.fn() : fn_ident<NS>();
            } else {
                f_incl = own ? block_fn<HQ, ESC, CMT>(mine, ls) : fn_ident<NS>();
            }
    #pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const uint32_t g = __shfl_up(f_incl, d);
                if (ln >= (uint32_t)d) f_incl = fn_then<NS>(g, f_incl);
            }
            f_excl = __shfl_up(f_incl, 1);
            if (ln == 0) f_excl = fn_ident<NS>();
        }

        // ---- summary mode: every entry state's trajectory through the tile ----
        if (MODE == M_SUMMARY) {
    #pragma unroll
            for (int s = 0; s < NS; ++s) {
                uint64_t S = 0, T = 0, F = 0;
                if (own) {
                    const StepOut o = step_of(fn_apply(f_excl, (uint32_t)s));
                    S = o.S & live;
                    T = o.T & live;
                    F = o.F & live;
                }
                const Counts c = wave_counts(S, T, F, tile_start + off);
                if (ln == 0) wsum[w * NS + s] = c;
            }
            if (ln == 63) sh.wfn[w] = f_incl;
            publish_summary();
            return;
        }

        // ---- entry states: tile prefix from the scan, then through the waves --
        load_entry_states();

        // ---- pass 2: masks, compaction, counters (one block per lane) ---------
        {
            const uint32_t entry = fn_apply(f_excl, uni(sh.wentry[w]));
            uint64_t S = 0, T = 0, F = 0;
            if (own) {
                const StepOut o = step_of(entry);
                S = o.S & live;
                T = o.T & live;
                F = o.F & live;
            }
            const uint32_t n = popc(S);
            const uint32_t incl = wave_incl_sum(n);
            if (ln == 63) sh.wcount[w] = incl;
            __syncthreads();
            // One list for the tile, in byte order: this lane's structurals go
            // after the earlier waves' and the earlier lanes' entries.
            uint32_t wb = 0;
            for (uint32_t k = 0; k < w; ++k) wb += sh.wcount[k];
            uint16_t* lst = sh.list + uni(wb) + (incl - n);
            for (uint64_t rem = S; rem; rem &= rem - 1) {
                const uint32_t bit = (uint32_t)__ffsll((long long)rem) - 1u;
                *lst++ = (uint16_t)((off + bit) | (uint32_t)(((T >> bit) & 1ULL) << 14) |
                                    (uint32_t)(((F >> bit) & 1ULL) << 15));
            }
        }
    }
    __syncthreads();

    // ---- tile counters: list offsets of the waves; totals of the last tile -
    if (tid == 0) {
        const Counts inc = sh.inc;
        sh.wbase[0] = 0;
        for (int k = 0; k < NWAVE; ++k) sh.wbase[k + 1] = sh.wbase[k] + sh.wcount[k];
        if (last_tile) {
            P.result[R_N_STRUCT] = inc.n_struct;
            P.result[R_N_SEPS] = inc.n_seps;
            P.result[R_N_TERMS] = inc.n_terms;
        }
    }
    __syncthreads();

    const uint32_t n_list = uni(sh.wbase[NWAVE]);
    const Counts pre = sh.pre;

    if (MODE == M_INDEX) {
        if (P.ends) {
            // Nested-tape form: 64-bit entries, the position in the low
            // word and, for a bracket, its byte in the high word (the
            // bracket matcher replaces it by the partner's tape index).
            for (uint32_t k = tid; k < n_list; k += WG) {
                const uint64_t idx = (uint64_t)pre.n_struct + k;
                const uint32_t at = list_at(sh, k) & 0x3FFFu;
                const uint32_t b = tb[at];
                const uint32_t hi = (b == '{' || b == '[' || b == '}' || b == ']') ? b : 0u;
                if (idx < P.ends_cap) P.ends[idx] = ((unsigned long long)hi << 32) | (tile_start + at);
            }
            return;
        }
        for (uint32_t k = tid; k < n_list; k += WG) {
            uint64_t idx = (uint64_t)pre.n_struct + k;
            if (idx < P.out_cap) P.out[idx] = tile_start + (list_at(sh, k) & 0x3FFFu);
        }
        return;
    }

    TileSrc src{P.data, sh.lds_bytes + (lds_lo + PRE - tile_start), lds_lo, lds_hi};
    const CellDialect cd{P.quote, HQ ? P.escape : -1};

    // Walk the list with record context: each thread takes `items`
    // consecutive entries (WG*items per round); one block scan per round
    // gives every thread the terminators, rows and last terminator before
    // its first entry. Then all threads visit their entries in step:
    // visit(e, terms_before, weight_before) with counts over all earlier
    // entries of the tile. A weighted walk (WEIGHTED) first evaluates
    // weight(e) with context and scans those sums too, so each visit knows
    // the weight of everything before it (string arena offsets).
    // `vword` != 0 appends one virtual entry after the list: the terminator
    // at `len` that closes an unterminated last record (last tile only).
    // Entries per thread in a walk round: the tile's mean fields per record.
    auto walk_items = [&]() -> uint32_t {
        const uint32_t t_terms = sh.inc.n_terms - pre.n_terms;
        return uni(t_terms ? min(ITEMS_MAX, max(1u, (n_list + t_terms / 2) / t_terms)) : ITEMS_MAX);
    };
    auto walk_w = [&](auto weighted, auto&& weight, auto&& visit, uint32_t vword, auto&& after)
        __attribute__((always_inline)) {
        constexpr bool WEIGHTED = decltype(weighted)::value;
        uint32_t terms_before = 0, rows_before = 0, w_before = 0;
        int32_t lt_before = -1;
        const uint32_t n_walk = n_list + (vword ? 1u : 0u);
        const uint32_t items = walk_items();
        const auto word_at = [&](uint32_t k) -> uint32_t { return k < n_list ? (uint32_t)sh.list[k] : vword; };
        // Entry k without record context; `fs` is the end of entry k-1 plus 1.
        const auto entry_at = [&](uint32_t k, uint32_t fs) -> Entry {
            Entry e;
            e.k = k;
            e.valid = k < n_walk;
            const uint32_t v = e.valid ? word_at(k) : 0u;
            e.is_term = e.valid && (v & 0x4000u);
            e.flag = e.valid && (v & 0x8000u);
            e.pos = tile_start + (v & 0x3FFFu);
            e.fs = fs;
            e.rs = 0;
            e.ord = 0;
            e.row = 0;
            return e;
        };
        for (uint32_t base = 0; base < n_walk; base += WG * items) {
            const uint32_t k0 = base + tid * items;
            const uint32_t fs0 = k0 == 0 ? pre.last_struct
                                         : (k0 - 1 < n_walk ? tile_start + (word_at(k0 - 1) & 0x3FFFu) + 1 : 0u);
            // This thread's terminators, rows and last terminator.
            Scan3 agg{0, 0, -1};
            uint32_t fs = fs0;
            for (uint32_t i = 0; i < items; ++i) {
                const Entry e = entry_at(k0 + i, fs);
                fs = e.pos + 1;
                if (e.is_term) {
                    agg.a += 1;
                    agg.b += e.flag ? 0u : 1u;
                    agg.c = (int32_t)e.k;
                }
            }
            Scan3 tot;
            const Scan3 ex = block_excl(agg, sh.scan_sh, tot);
            const int32_t lt0 = ex.c > lt_before ? ex.c : lt_before;
            const uint32_t rs0 = lt0 >= 0 ? tile_start + (word_at((uint32_t)lt0) & 0x3FFFu) + 1 : pre.last_term;
            const uint32_t row0_t = pre.n_rows + rows_before + ex.b;
            // The thread's entries with record context, in order.
            const auto each = [&](auto&& f) {
                int32_t lt = lt0;
                uint32_t rs = rs0, row = row0_t, terms = terms_before + ex.a, fs1 = fs0;
                for (uint32_t i = 0; i < items; ++i) {
                    Entry e = entry_at(k0 + i, fs1);
                    fs1 = e.pos + 1;
                    e.rs = rs;
                    e.ord = lt >= 0 ? (e.k - (uint32_t)lt - 1) : (pre.ord + e.k);
                    e.row = row;
                    f(e, terms);
                    if (e.is_term) {
                        lt = (int32_t)e.k;
                        rs = e.pos + 1;
                        terms += 1;
                        if (!e.flag) row += 1;
                    }
                }
            };
            uint32_t wsum = 0;
            if (WEIGHTED) {
                uint32_t mine = 0;
                each([&](const Entry& e, uint32_t) { if (e.valid) mine += weight(e); });
                uint32_t wt;
                wsum = w_before + block_excl_sum(mine, sh.sum_sh, wt);
                w_before += wt;
            }
            // A weighted walk's visit returns the entry's weight (it has just
            // computed it), so the weight runs once more only for the scan.
            each([&](const Entry& e, uint32_t terms) {
                if constexpr (WEIGHTED) wsum += visit(e, terms, wsum);
                else visit(e, terms, wsum);
            });
            after(base, min(WG * items, n_walk - base));
            terms_before += tot.a;
            rows_before += tot.b;
            lt_before = tot.c > lt_before ? tot.c : lt_before;
        }
    };
    auto walk = [&](auto&& visit, uint32_t vword = 0u) __attribute__((always_inline)) {
        walk_w(NoWeight{}, [](const Entry&) { return 0u; }, visit, vword, [](uint32_t, uint32_t) {});
    };
    // Field end without the record's trailing '\r' (csv_geo.rs:1056-1060).
    // A string cell that cleans to itself: no quote convention, or doubled
    // quotes and not quoted at both ends (backslash dialects unescape
    // anywhere, so they always run the cleaning stream).
    auto raw_cell = [&](uint32_t from, uint32_t to) -> bool {
        if (cd.quote < 0) return true;
        if (cd.escape >= 0) return false;
        return !quoted_both(src, from, to, (uint8_t)cd.quote);
    };
    // A cell of the tile quoted at both ends with no other quote byte (read
    // from the tile's quote bitmap, a word per 32 bytes): it cleans to
    // [from + 1, to - 1).
    // The inner quote bytes of the tile bytes [a, z] (tile-relative, a <= z)
    // in word k, from the bitmap.
    auto qword = [&](uint32_t k, uint32_t a, uint32_t z) -> uint32_t {
        uint32_t wv = qbits[k];
        if (k == (a >> 5)) wv &= ~0u << (a & 31u);
        if (k == (z >> 5)) wv &= (z & 31u) == 31u ? ~0u : (2u << (z & 31u)) - 1u;
        return wv;
    };
    // A cell of the tile quoted at both ends: `drops` is the number of quote
    // bytes its doubled pairs lose (dq_dropped over its inside), so it cleans
    // to its inside less those; with none it cleans to [from + 1, to - 1).
    auto quoted_in_tile = [&](uint32_t from, uint32_t to, uint32_t& drops) -> bool {
        if constexpr (!QBITS) {
            return false;
        } else {
            if (from < tile_start || to > tile_start + tile_len || to - from < 2u) return false;
            const uint32_t a = from - tile_start, z = to - 1u - tile_start;   // the outer quotes
            if (((qbits[a >> 5] >> (a & 31u)) & (qbits[z >> 5] >> (z & 31u)) & 1u) == 0u) return false;
            drops = 0;
            if (z > a + 1u) {
                uint32_t in_run = 0, odd_next = 0;
                for (uint32_t k = (a + 1u) >> 5; k <= ((z - 1u) >> 5); ++k)
                    drops += (uint32_t)__popc(dq_dropped(qword(k, a + 1u, z - 1u), in_run, odd_next));
            }
            return true;
        }
    };
    // The bytes a cell cleans to when they are a run of the document: the
    // cell itself (raw), or its inside (quoted with no doubled quote). False:
    // `drops` tells whether the doubled-quote copy (quoted in the tile) or
    // the cleaning stream (drops = ~0) makes it.
    auto run_cell = [&](uint32_t from, uint32_t to, uint32_t& rf, uint32_t& rt, uint32_t& drops) -> bool {
        drops = ~0u;
        if (raw_cell(from, to)) {
            rf = from;
            rt = to;
            return true;
        }
        if (quoted_in_tile(from, to, drops) && drops == 0u) {
            rf = from + 1u;
            rt = to - 1u;
            return true;
        }
        return false;
    };
    auto clean_len = [&](uint32_t from, uint32_t to) -> uint32_t {
        uint32_t rf, rt, drops;
        if (run_cell(from, to, rf, rt, drops)) return rt - rf;
        if (drops != ~0u) return to - from - 2u - drops;
        return clean_cell(src, from, to, cd, [](uint8_t) {});
    };
    auto field_end = [&](const Entry& e) -> uint32_t {
        return (e.is_term && e.pos > e.rs && src.at(e.pos - 1) == '\r') ? e.pos - 1 : e.pos;
    };

    const Counts inc = sh.inc;
    const bool trailer = last_tile && inc.last_term < P.len;
    const bool trailer_comment = trailer && CMT && P.data[inc.last_term] == (uint8_t)P.comment;
    // Virtual terminator word (local position, term bit, comment flag).
    const uint32_t trailer_word =
        trailer ? ((P.len - tile_start) | 0x4000u | (trailer_comment ? 0x8000u : 0u)) : 0u;

    if (MODE == M_TAPE) {
        walk([&](const Entry& e, uint32_t terms_excl, uint32_t) {
            if (!e.valid) return;
            uint32_t seps_before = e.k - terms_excl;
            if (e.is_term) {
                uint64_t idx = (uint64_t)pre.n_terms + terms_excl;
                if (idx < P.ends_cap)
                    P.ends[idx] = ((unsigned long long)(pre.n_seps + seps_before) << 32) | e.pos;
            } else {
                uint64_t idx = (uint64_t)pre.n_seps + seps_before;
                if (idx < P.out_cap) P.out[idx] = e.pos;
            }
        });
        return;
    }

    if (MODE == M_FIELDBYTES) {
        unsigned long long fields = 0, bytes = 0;
        walk([&](const Entry& e, uint32_t, uint32_t) {
            if (!e.valid || e.flag) return;
            fields += 1;
            bytes += field_byte_len(src, e.fs, field_end(e), P.quote);
        });
        if (tid == 0 && trailer && !trailer_comment) {
            uint32_t to = (P.len > inc.last_term && P.data[P.len - 1] == '\r') ? P.len - 1 : P.len;
            fields += 1;
            bytes += field_byte_len(src, inc.last_struct, to, P.quote);
        }
        fields = block_reduce_sum(fields, sh.usum_sh);
        bytes = block_reduce_sum(bytes, sh.usum_sh);
        if (tid == 0) {
            unsigned long long* pp = reinterpret_cast<unsigned long long*>(P.partials) + 2 * (size_t)tile;
            pp[0] = fields;
            pp[1] = bytes;
        }
        return;
    }

    if (MODE == M_LOGFMT) {
        unsigned long long pairs = 0, kbytes = 0, vbytes = 0, csum = 0;
        auto pair = [&](uint32_t ks, uint32_t ke, uint32_t vs, uint32_t ve) {
            pairs += 1;
            kbytes += ke - ks;
            for (uint32_t p = ks; p < ke; ++p) csum += src.at(p);
            // add_logfmt_value (logfmt.rs:826-858)
            if (ve == vs || src.at(vs) != '"') {
                vbytes += ve - vs;
                for (uint32_t p = vs; p < ve; ++p) csum += src.at(p);
                return;
            }
            uint32_t bs = vs, be = ve;
            if (ve - vs >= 2 && src.at(ve - 1) == '"') { bs = vs + 1; be = ve - 1; }
            uint32_t i = bs;
            while (i < be) {
                uint8_t b = src.at(i);
                if (b == '\\' && i + 1 < be) { csum += src.at(i + 1); i += 2; }
                else { csum += b; i += 1; }
                vbytes += 1;
            }
        };
        walk([&](const Entry& e, uint32_t, uint32_t) {
            if (!e.valid || (e.ord & 1u) == 0) return;
            const uint32_t to = field_end(e);
            uint32_t prev1 = e.fs - 1;  // structural that ended the key field
            uint32_t ks;
            if (e.ord == 1) ks = e.rs;
            else if (e.k >= 2) ks = tile_start + (list_at(sh, e.k - 2) & 0x3FFFu) + 1;
            else if (e.k == 1) ks = pre.last_struct;
            else ks = pre.prev_struct;
            pair(ks, prev1, e.fs, to);
        });
        if (tid == 0 && last_tile) {
            uint32_t fs = inc.last_struct;
            if (fs < P.len || (P.len > 0 && P.data[P.len - 1] != '\n')) {
                if (inc.ord & 1u) {
                    uint32_t e_ = (P.len > inc.last_term && P.data[P.len - 1] == '\r') ? P.len - 1 : P.len;
                    uint32_t ks = inc.ord == 1 ? inc.last_term : inc.prev_struct;
                    pair(ks, fs - 1, fs, e_);
                }
            }
        }
        pairs = block_reduce_sum(pairs, sh.usum_sh);
        kbytes = block_reduce_sum(kbytes, sh.usum_sh);
        vbytes = block_reduce_sum(vbytes, sh.usum_sh);
        csum = block_reduce_sum(csum, sh.usum_sh);
        if (tid == 0) {
            unsigned long long* pp = reinterpret_cast<unsigned long long*>(P.partials) + 4 * (size_t)tile;
            pp[0] = pairs; pp[1] = kbytes; pp[2] = vbytes; pp[3] = csum;
        }
        return;
    }

    // acc[j] += o for a uniform j < N, without indexing registers.
    auto acc_add = [](auto& acc, uint32_t j, const CellStat& o) __attribute__((always_inline)) {
        constexpr uint32_t N = sizeof(acc) / sizeof(acc[0]);
#pragma unroll
        for (uint32_t jj = 0; jj < N; ++jj)
            if (jj == j) acc[jj].add(o);
    };
    auto acc_publish = [&](auto& acc, const auto& cols, uint32_t nc, unsigned long long records, bool with_records) {
        static_assert((NWAVE - 1) * (4 * STAT_GROUP + 1) * 64 * 8 <= sizeof(sh.list), "statistics exchange fits the list");
        publish_stats(P, tile, acc, cols, nc, records, with_records, reinterpret_cast<unsigned long long*>(sh.list));
    };
    // Typed value of a present cell as f64 (as_f64) or i64: for cells inside
    // the tile the fixed-6 test, then the register parser; the stream parser
    // for the cells those leave undecided; undecided f64 values are then
    // resolved exactly.
    const bool swar_ok = cd.quote < '0' || cd.quote > '9';   // num_swar's quote condition
    auto num_value = [&](uint32_t from, uint32_t to, bool as_f64) __attribute__((always_inline)) {
        NumVal v;
        bool done = false;
        // The fixed `[-]d{1,3}.d{6}` shape first: three word reads decide it.
        // Offsets into sh.lds_bytes (the cell is inside [lds_lo, lds_hi)).
        const bool in_lds = from >= lds_lo && to <= lds_hi;
        uint32_t lf = from - tile_start + PRE, lt = to - tile_start + PRE;
        // A cell quoted at both ends cleans to its inside with doubled
        // quotes collapsed; the register parsers take the inside, and an
        // inside with a quote byte left in it is not a number either way.
        if (HQ && !ESC && in_lds && lt - lf >= 2 && sh.lds_bytes[lf] == (uint8_t)cd.quote &&
            sh.lds_bytes[lt - 1] == (uint8_t)cd.quote) {
            lf += 1;
            lt -= 1;
        }
        const uint32_t n = lt - lf;
        if (as_f64 && in_lds && fixed6_lds(sh.lds_bytes, lf, lt, v.f)) {
            done = true;
        } else if (swar_ok && in_lds && n <= NUM_WIN) {
            // The narrow window when every lane here can use it.
            if (__all(n <= Win12::BYTES)) {
                auto at = [&](uint32_t p) { return (uint32_t)sh.lds_bytes[lt - Win12::BYTES + p]; };
                done = num_swar(win_lds<3>(sh.lds_bytes, lt), n, as_f64, cd.quote, at, v.f, v.iok, v.iv);
            } else {
                auto at = [&](uint32_t p) { return (uint32_t)sh.lds_bytes[lt - Win24::BYTES + p]; };
                done = num_swar(win_lds<6>(sh.lds_bytes, lt), n, as_f64, cd.quote, at, v.f, v.iok, v.iv);
            }
        }
        if constexpr (FAST) {
            if (!done || (as_f64 && v.f.ok && v.f.deferred)) {
                sh.fix_bad = 1u;
                v.f = F64Result{};
                v.iok = false;
                v.iv = 0;
            }
        } else {
            if (!done) {
                if (as_f64) {
                    v.f = parse_f64_field(src, from, to, cd);
                } else {
                    v.iok = parse_i64_field(src, from, to, cd, v.iv);
                    if (!v.iok) v.iv = 0;
                }
            }
            if (as_f64 && v.f.ok && v.f.deferred) v.f.value = resolve_f64(sh.slow, src, from, to, cd);
        }
        return v;
    };
    // ---- numeric fast forms: records straight to threads -------------------
    // Record r of the tile ends at its r-th terminator, list entry tl[r] (the
    // SWAR front listed them); the record after the last terminator ends at
    // the virtual terminator (last tile) or is still open, and then only its
    // fields that end here are written (the next tile writes the rest).
    // Thread r takes record r: its fields that end in this tile are entries
    // e0 .. e0 + n - 1, so column c's cell is entry e0 + (c.index - ord0).
    // Rows are consecutive across a wave's lanes, so a column's validity bits
    // are one ballot, ORed into at most two words.
    if constexpr (NUMFAST) {
        constexpr bool IS_COLS = MODE == M_COLUMNS;
        constexpr uint32_t NACC = 2;
        static_assert(NACC == STAT_GROUP && NACC == COL_STAT_GROUP, "accumulators per pass");
        constexpr bool with_stats = KMODE != M_NUM_FAST;
        const uint16_t* const tl = sh.list + (TILE - TL_CAP);
        const uint32_t c0 = IS_COLS ? 0u : P.col_group;
        const uint32_t nc = IS_COLS ? P.n_cols : min((uint32_t)STAT_GROUP, P.n_cols - c0);
        const uint64_t rows_cap = P.rows_cap;
        const auto word_at = [&](uint32_t k) -> uint32_t { return k < n_list ? (uint32_t)sh.list[k] : trailer_word; };
        const auto pos_of = [&](uint32_t wd) -> uint32_t { return tile_start + (wd & 0x3FFFu); };
        Acc acc[NACC];
#pragma unroll
        for (uint32_t j = 0; j < NACC; ++j) acc[j] = Acc{0, 0, 0, 0, 0, 0.0};
        unsigned long long records = 0;
        if (!uni(sh.fix_bad)) {
            const uint32_t nt = inc.n_terms - pre.n_terms;   // the tile's real terminators
            const uint32_t e_after = nt ? (uint32_t)tl[nt - 1] + 1u : 0u;
            const uint32_t n_rec = nt + ((trailer || e_after < n_list) ? 1u : 0u);
            for (uint32_t r0 = 0; r0 < n_rec; r0 += WG) {
                const uint32_t r = r0 + tid;
                const bool has = r < n_rec;
                uint32_t e0 = 0, n = 0, fs0 = 0, rs = 0, ord0 = 0;
                bool term = false;
                if (has) {
                    e0 = r == 0 ? 0u : (uint32_t)tl[r - 1] + 1u;
                    if (r < nt) {
                        term = true;
                        n = (uint32_t)tl[r] - e0 + 1u;
                    } else {
                        term = trailer;
                        n = n_list - e0 + (trailer ? 1u : 0u);
                    }
                    if (e0 == 0) {
                        fs0 = pre.last_struct;
                        rs = pre.last_term;
                        ord0 = pre.ord;
                    } else {
                        fs0 = pos_of(sh.list[e0 - 1]) + 1u;
                        rs = fs0;
                    }
                }
                if (has && term) records += 1;
                const uint32_t row = pre.n_rows + r;
                const bool live = row < rows_cap;
                // The wave's first row (lanes with a record are the lowest).
                const uint32_t rfirst = pre.n_rows + r0 + 64u * w;
                for (uint32_t jj = 0; jj < nc; ++jj) {
                    const ColDesc& c = P.cols[c0 + jj];
                    const int32_t i = (int32_t)c.index - (int32_t)ord0;
                    const bool here = has && i >= 0 && (uint32_t)i < n;
                    const bool missing = has && term && i >= (int32_t)n;
                    uint32_t fs = 0, fe = 0;
                    bool odd = false;   // a '\r' check outside the LDS copy, or a trim past the field start
                    if (here) {
                        const uint32_t k = e0 + (uint32_t)i;
                        fe = pos_of(word_at(k));
                        fs = i == 0 ? fs0 : pos_of(word_at(k - 1)) + 1u;
                        if (term && (uint32_t)i + 1u == n && fe > rs) {
                            if (fe - 1u < lds_lo) odd = true;
                            else if (sh.lds_bytes[fe - 1u - tile_start + PRE] == '\r') fe -= 1;
                        }
                        odd = odd || fe < fs;
                    }
                    if (odd) sh.fix_bad = 1u;
                    const bool as_f64 = c.type == FALX_COL_F64;
                    NumVal nv{};
                    // The fixed-6 shape from the 12 bytes ending at the cell,
                    // then the general register parsers.
                    bool done = false;
                    if (here && !odd && as_f64 && fs >= lds_lo && fe <= lds_hi) {
                        const Win<3> W = win_lds<3>(sh.lds_bytes, fe - tile_start + PRE);
                        done = fixed6_win(W.w[0], W.w[1], W.w[2], fe - fs, nv.f);
                    }
                    if (here && !odd && !done) nv = num_value(fs, fe, as_f64);
                    const bool vb = here && !odd && (as_f64 ? nv.f.ok : nv.iok);
                    if (IS_COLS && (here || missing) && live) {
                        if (as_f64) reinterpret_cast<double*>(c.values)[row] = vb ? nv.f.value : 0.0;
                        else reinterpret_cast<long long*>(c.values)[row] = vb ? nv.iv : 0;
                    }
                    if (with_stats && jj < NACC && (IS_COLS ? live : true)) {
                        CellStat o{false, false, 0, 0, 0, 0.0};
                        if (vb) {
                            o = as_f64 ? CellStat{true, nv.f.fixed6, f64_to_bits(nv.f.value), 0, nv.f.mantissa, nv.f.value}
                                       : CellStat{true, false, (unsigned long long)nv.iv, 0, 0, (double)nv.iv};
                        }
                        acc_add(acc, jj, o);
                    }
                    if (IS_COLS) {
                        const uint64_t W = ballot(vb && live);
                        if (W && ln == 0) {
                            const uint32_t wi = rfirst >> 6, sb = rfirst & 63u;
                            atomicOr(&c.valid[wi], (unsigned long long)(W << sb));
                            if (sb && (W >> (64u - sb))) atomicOr(&c.valid[wi + 1], (unsigned long long)(W >> (64u - sb)));
                        }
                    }
                }
            }
        }
        if (IS_COLS && tid == 0 && last_tile) P.result[R_N_ROWS] = inc.n_rows + (trailer ? 1 : 0);
        __syncthreads();
        if (tid == 0) P.fixup[tile] = sh.fix_bad;
        if (with_stats) {
            const uint32_t cols[NACC] = {c0, c0 + 1u};
            acc_publish(acc, cols, min(NACC, nc), records, c0 == 0u);
        }
        return;
    }

    // ---- COLUMNS and STATS: one record per thread --------------------------
    // The tile's records end at its terminator entries (and at the virtual
    // terminator of an unterminated last record); a record still open after
    // the tile's last terminator is finished by the next tile. A thread takes
    // one record: its fields that end in this tile are the list entries
    // e0 .. e0 + n - 1, so column c's cell is entry e0 + (c.index - ord0),
    // found by index, and the lanes of a wave run the same steps for each
    // column. Terminators are listed per segment of up to TL_MAX records (one
    // block scan over the list entries per segment). Validity words are built
    // per wave (rows of consecutive records are consecutive) and ORed into
    // the column with one atomic per word.
    if constexpr (MODE == M_COLUMNS || MODE == M_STATS || MODE == M_STRINGS) {
        constexpr uint32_t TL_MAX = 512;
        __shared__ uint32_t tl[TL_MAX];   // list index | (row rank in the segment << 16)
        const uint32_t n_ent = n_list + (trailer ? 1u : 0u);   // the virtual terminator is entry n_list
        const auto word_at = [&](uint32_t k) -> uint32_t { return k < n_list ? (uint32_t)sh.list[k] : trailer_word; };
        const auto pos_of = [&](uint32_t wd) -> uint32_t { return tile_start + (wd & 0x3FFFu); };
        const uint64_t rows_cap = P.rows_cap;

        // A record of the tile: its fields that end here are list entries
        // e0 .. e0 + n - 1; `term` when its terminator is here (then it has a
        // row unless it is a comment record); ord0 is the ordinal of field
        // e0 (non-zero only for the record open at the tile's start).
        struct Rec {
            bool has, term, comment;
            uint32_t e0, n, fs0, rs, ord0, row;
        };
        // Calls round(rec) once per round of WG records, with every thread
        // (a thread without a record gets has = false). Terminators are
        // listed per segment of up to TL_MAX records (one block scan over
        // the list entries per segment).
        auto for_records = [&](auto&& round) __attribute__((always_inline)) {
            uint32_t s0 = 0, row_base = pre.n_rows;
            while (true) {
                // The segment's terminators: entries [s0, n_ent) in runs of
                // `per` per thread, counted, scanned, then written by rank.
                const uint32_t rem = n_ent - s0;
                const uint32_t per = (rem + WG - 1) / WG;
                const uint32_t k0 = min(n_ent, s0 + tid * per), k1 = min(n_ent, k0 + per);
                Scan3 cnt{0, 0, -1};
                for (uint32_t k = k0; k < k1; ++k) {
                    const uint32_t wd = word_at(k);
                    if (wd & 0x4000u) {
                        cnt.a += 1;
                        cnt.b += (wd & 0x8000u) ? 0u : 1u;
                    }
                }
                Scan3 tot;
                const Scan3 ex = block_excl(cnt, sh.scan_sh, tot);
                {
                    uint32_t ra = ex.a, rb = ex.b;
                    for (uint32_t k = k0; k < k1; ++k) {
                        const uint32_t wd = word_at(k);
                        if (wd & 0x4000u) {
                            if (ra < TL_MAX) tl[ra] = k | (rb << 16);
                            ra += 1;
                            rb += (wd & 0x8000u) ? 0u : 1u;
                        }
                    }
                }
                __syncthreads();
                const bool last_seg = tot.a <= TL_MAX;
                const uint32_t n_seg = uni(min(tot.a, TL_MAX));
                const uint32_t t_last = n_seg ? uni(tl[n_seg - 1]) : 0u;
                const uint32_t last_k = t_last & 0xFFFFu;
                // An unfinished record after the tile's last terminator (never
                // with a trailer: the virtual terminator is the last entry).
                const uint32_t after = n_seg ? last_k + 1u : s0;
                const bool partial = last_seg && after < n_ent;
                const uint32_t n_rec = n_seg + (partial ? 1u : 0u);
                const uint32_t seg_rows = n_seg ? (t_last >> 16) + ((word_at(last_k) & 0x8000u) ? 0u : 1u) : 0u;
                for (uint32_t r0 = 0; r0 < n_rec; r0 += WG) {
                    const uint32_t r = r0 + tid;
                    Rec R{r < n_rec, false, false, 0, 0, 0, 0, 0, 0};
                    if (R.has) {
                        R.e0 = r == 0 ? s0 : (tl[r - 1] & 0xFFFFu) + 1u;
                        if (r < n_seg) {
                            const uint32_t t = tl[r];
                            const uint32_t e1 = t & 0xFFFFu;
                            R.term = true;
                            R.comment = (word_at(e1) & 0x8000u) != 0u;
                            R.n = e1 - R.e0 + 1u;
                            R.row = row_base + (t >> 16);
                        } else {
                            R.n = n_ent - R.e0;
                            R.comment = (word_at(R.e0) & 0x8000u) != 0u;
                            R.row = row_base + seg_rows;
                        }
                        if (R.e0 == 0) {
                            R.fs0 = pre.last_struct;
                            R.rs = pre.last_term;
                            R.ord0 = pre.ord;
                        } else {
                            R.fs0 = pos_of(word_at(R.e0 - 1)) + 1u;
                            R.rs = R.fs0;
                        }
                    }
                    round(R);
                }
                row_base += seg_rows;
                if (last_seg) break;
                s0 = last_k + 1u;
                __syncthreads();   // every thread has read tl before the next segment rewrites it
            }
        };
        // The cell of field ordinal `index` of record R: `here` when the field
        // ends in this tile (its bytes [fs, fe), the record's last field
        // without a trailing '\r', csv_geo.rs:1056-1060), `missing` when the
        // record ends before it.
        auto cell_of = [&](const Rec& R, uint32_t index, bool& here, bool& missing, uint32_t& fs, uint32_t& fe)
            __attribute__((always_inline)) {
            const bool live_rec = R.has && !R.comment;
            const int32_t i = (int32_t)index - (int32_t)R.ord0;
            here = live_rec && i >= 0 && (uint32_t)i < R.n;
            missing = live_rec && R.term && i >= (int32_t)R.n;
            fs = fe = 0;
            if (here) {
                const uint32_t k = R.e0 + (uint32_t)i;
                fe = pos_of(word_at(k));
                fs = i == 0 ? R.fs0 : pos_of(word_at(k - 1)) + 1u;
                if (R.term && (uint32_t)i + 1u == R.n && fe > R.rs && src.at(fe - 1) == '\r') fe -= 1;
            }
        };
        // Validity of one column for the wave's rows: lanes with a row (RL)
        // hold consecutive rows in lane order, so bit (row - first row) of
        // the wave's word is the lane's cell; words are ORed in by atomics
        // (the entry point cleared them).
        auto wave_valid = [&](const ColDesc& col, bool vb, uint32_t row, uint64_t RL) __attribute__((always_inline)) {
            if (!RL) return;
            const uint32_t f = (uint32_t)__ffsll((long long)RL) - 1u;
            const uint32_t rfirst = lane_value(row, f);
            const uint64_t run = RL >> f;
            uint64_t W;
            if ((run & (run + 1ULL)) == 0ULL) {
                W = ballot(vb) >> f;
            } else {
                unsigned long long v = vb ? (1ULL << ((row - rfirst) & 63u)) : 0ULL;
                v |= dpp_get<0xB1>(v);
                v |= dpp_get<0x4E>(v);
                v |= dpp_get<0x124>(v);
                v |= dpp_get<0x128>(v);
                const uint64_t r0 = ((uint64_t)lane_value((uint32_t)(v >> 32), 0) << 32) | lane_value((uint32_t)v, 0);
                const uint64_t r1 = ((uint64_t)lane_value((uint32_t)(v >> 32), 16) << 32) | lane_value((uint32_t)v, 16);
                const uint64_t r2 = ((uint64_t)lane_value((uint32_t)(v >> 32), 32) << 32) | lane_value((uint32_t)v, 32);
                const uint64_t r3 = ((uint64_t)lane_value((uint32_t)(v >> 32), 48) << 32) | lane_value((uint32_t)v, 48);
                W = r0 | r1 | r2 | r3;
            }
            if (W && ln == 0) {
                const uint32_t wi = rfirst >> 6, sb = rfirst & 63u;
                atomicOr(&col.valid[wi], (unsigned long long)(W << sb));
                if (sb && (W >> (64u - sb))) atomicOr(&col.valid[wi + 1], (unsigned long long)(W >> (64u - sb)));
            }
        };

        // ---- STRINGS: arena placement of the string columns ---------------
        // Per string column, a block scan over each round's cleaned lengths
        // gives every cell its arena offset (the tile's cells of a column fill
        // one contiguous range, whose base the slot scan computed). Cells are
        // written into an LDS stage when the tile's range fits, then stored
        // out contiguously; offsets and validity go straight to the column.
        if constexpr (MODE == M_STRINGS) {
            constexpr uint32_t SSTAGE = 8192;
            __shared__ uint8_t sstage[SSTAGE];
            // Per record of a round: end of its cell in the tile's range, and
            // the offset from output position to LDS source byte of a raw cell
            // copied by all threads together (RAW_SKIP: its owner wrote it).
            constexpr int32_t RAW_SKIP = -32768;
            __shared__ uint16_t rt_end[WG];
            __shared__ int16_t rt_delta[WG];
            const uint32_t n_cols = P.n_cols;
            for (uint32_t j = 0; j < n_cols; ++j) {
                const ColDesc& c = P.cols[j];
                if (c.type != FALX_COL_STR) continue;
                int32_t* offs = reinterpret_cast<int32_t*>(c.values);
                const uint32_t arena = P.str_base[(size_t)tile * P.n_str + c.str_slot];
                const uint32_t total = P.str_len[(size_t)tile * P.n_str + c.str_slot];
                const bool gather = total <= SSTAGE;
                // The shared copy pays for its two barriers per round when the
                // tile's cells are long: its mean cleaned length decides.
                const bool coop_tile = gather && total > 13u * (inc.n_rows - pre.n_rows + 1u);
                if (tile == 0 && tid == 0) offs[0] = 0;
                // Fused statistics of every string column (present cells,
                // cleaned bytes, their byte sum): no statistics pass runs for
                // a group of string columns.
                const bool stat_col = P.partials != nullptr;
                unsigned long long s_cells = 0, s_bytes = 0, s_sum = 0;
                uint32_t run = 0;   // bytes of this tile's range placed by earlier rounds
                for_records([&](const Rec& R) {
                    bool here, missing;
                    uint32_t fs, fe;
                    cell_of(R, c.index, here, missing, fs, fe);
                    // A cell that cleans to a run of the document [rf, rt)
                    // is copied as it stands.
                    uint32_t rf = 0, rt = 0, drops = ~0u;
                    const bool raw = here && run_cell(fs, fe, rf, rt, drops);
                    const uint32_t L = here ? (raw ? rt - rf : (drops != ~0u ? fe - fs - 2u - drops : clean_len(fs, fe))) : 0u;
                    uint32_t rtot;
                    const uint32_t at = run + block_excl_sum(L, sh.sum_sh, rtot);
                    const bool live = R.row < rows_cap;
                    // A live raw cell inside the tile's LDS copy joins the
                    // shared copy below: work per byte, not per longest cell.
                    const bool coop = coop_tile && raw && live && rf >= lds_lo && rt <= lds_hi;
                    if (coop_tile) {
                        rt_end[tid] = (uint16_t)(at + L);
                        rt_delta[tid] = coop ? (int16_t)((int32_t)(rf - tile_start + PRE) - (int32_t)at) : (int16_t)RAW_SKIP;
                    }
                    if (here && !coop) {
                        // Bytes of rows past rows_cap are not placed: the call
                        // reports CAPACITY then.
                        auto put = [&](uint32_t q, uint8_t b) {
                            if (!live) return;
                            if (gather) sstage[q] = b;
                            else if ((uint64_t)arena + q < c.data_cap) c.data[arena + q] = b;
                        };
                        if (raw) {
                            for (uint32_t k = 0; k < L; ++k) {
                                const uint8_t b = src.at(rf + k);
                                put(at + k, b);
                                s_sum += b;
                            }
                        } else if (drops != ~0u) {
                            // Doubled quotes, the cell in the tile: its inside
                            // less the dropped quotes, a bitmap word at a time.
                            const uint32_t a = fs + 1u - tile_start, z = fe - 2u - tile_start;
                            uint32_t q = at, in_run = 0, odd_next = 0;
                            for (uint32_t k = a >> 5; k <= (z >> 5); ++k) {
                                const uint32_t d = dq_dropped(qword(k, a, z), in_run, odd_next);
                                const uint32_t p0 = max(a, k << 5), p1 = min(z, (k << 5) + 31u);
                                for (uint32_t p = p0; p <= p1; ++p) {
                                    if ((d >> (p & 31u)) & 1u) continue;
                                    const uint8_t b = tb[p];
                                    put(q++, b);
                                    s_sum += b;
                                }
                            }
                        } else {
                            uint32_t q = at;
                            clean_cell(src, fs, fe, cd, [&](uint8_t b) {
                                put(q++, b);
                                s_sum += b;
                            });
                        }
                    }
                    if (here) {
                        s_cells += 1;
                        s_bytes += L;
                    }
                    if (R.term && !R.comment && R.has && live) offs[R.row + 1] = (int32_t)(arena + at + L);
                    wave_valid(c, here && live, R.row, ballot(R.has && !R.comment));
                    if (coop_tile) {
                        // The round's raw cells: each thread copies a run of
                        // the round's output range, finding its first cell by
                        // binary search over the cells' ends.
                        __syncthreads();
                        const uint32_t per = (rtot + WG - 1) / WG;
                        const uint32_t q0 = run + tid * per, q1 = min(q0 + per, run + rtot);
                        if (q0 < q1) {
                            uint32_t lo = 0, up = WG - 1;   // first record whose cell ends past q0
                            while (lo < up) {
                                const uint32_t mid = (lo + up) >> 1;
                                if (rt_end[mid] > q0) up = mid; else lo = mid + 1;
                            }
                            uint32_t i = lo;
                            for (uint32_t q = q0; q < q1; ++q) {
                                while (rt_end[i] <= q) ++i;
                                const int32_t d = rt_delta[i];
                                if (d != RAW_SKIP) {
                                    const uint8_t b = sh.lds_bytes[(int32_t)q + d];
                                    sstage[q] = b;
                                    s_sum += b;
                                }
                            }
                        }
                        __syncthreads();
                    }
                    run += rtot;
                });
                if (gather) {
                    __syncthreads();
                    const uint64_t end = min((uint64_t)arena + total, c.data_cap);
                    for (uint64_t q = (uint64_t)arena + tid; q < end; q += WG) c.data[q] = sstage[q - arena];
                    __syncthreads();
                }
                if (stat_col) {
                    // The column's StatPart of this tile (the COLUMNS pass
                    // wrote zeros there), in the fixed reduction order.
                    const unsigned long long v0 = block_reduce_sum(s_cells, sh.usum_sh);
                    const unsigned long long v1 = block_reduce_sum(s_sum, sh.usum_sh);
                    const unsigned long long v2 = block_reduce_sum(s_bytes, sh.usum_sh);
                    if (tid == 0) {
                        StatPart& o = reinterpret_cast<StatPart*>(P.partials)[(size_t)tile * P.n_cols + j];
                        o.values = v0;
                        o.checksum = v1;
                        o.bytes = v2;
                        o.fixed6_sum = 0;
                        o.fixed6_cells = 0;
                        o.sum = 0.0;
                    }
                }
            }
            return;
        }

        // ---- COLUMNS and STATS: one record per thread ----------------------
        // A thread takes one record; column c's cell is its entry
        // e0 + (c.index - ord0), found by index, so the lanes of a wave run
        // the same steps for each column.
        constexpr uint32_t NACC = 2;
        static_assert(NACC == STAT_GROUP && NACC == COL_STAT_GROUP, "accumulators per pass");
        constexpr bool IS_COLS = MODE == M_COLUMNS;
        const uint32_t c0 = IS_COLS ? 0u : P.col_group;
        const uint32_t nc = IS_COLS ? P.n_cols : min((uint32_t)STAT_GROUP, P.n_cols - c0);
        // Known at compile time in the fast forms, which then carry no
        // statistics registers when there are none.
        const bool with_stats = !IS_COLS                         ? true
                                : KMODE == M_COLUMNS_FAST       ? false
                                : KMODE == M_COLUMNS_FAST_STATS ? true
                                                                : P.partials != nullptr;
        const bool trailer_row_live = trailer && !trailer_comment;
        Acc acc[NACC];
#pragma unroll
        for (uint32_t j = 0; j < NACC; ++j) acc[j] = Acc{0, 0, 0, 0, 0, 0.0};
        // Accumulator j of a COLUMNS pass holds the j-th column that is not a
        // string (string columns' statistics come from the STRINGS pass); a
        // STATS pass holds its group's columns in order.
        uint32_t acc_col[NACC];
        uint32_t n_acc = 0;
        if (IS_COLS) {
#pragma unroll
            for (uint32_t j = 0; j < NACC; ++j) acc_col[j] = 0u;
            for (uint32_t k = 0; k < P.n_cols && n_acc < NACC; ++k)
                if (P.cols[k].type != FALX_COL_STR) {
#pragma unroll
                    for (uint32_t j = 0; j < NACC; ++j)
                        if (j == n_acc) acc_col[j] = k;
                    ++n_acc;
                }
        } else {
#pragma unroll
            for (uint32_t j = 0; j < NACC; ++j) acc_col[j] = c0 + j;
            n_acc = min(NACC, nc);
        }
        unsigned long long records = 0;
        if (IS_COLS && tid < FALX_MAX_COLUMNS) sh.str_len[tid] = 0;   // read after the segment scans' barriers
        for_records([&](const Rec& R) {
            const bool live_rec = R.has && !R.comment;
            if (live_rec && R.term) records += 1;
            const uint32_t row = R.row;
            const bool live = row < rows_cap;
            const uint64_t RL = IS_COLS ? ballot(live_rec) : 0ULL;
            uint32_t slot = 0;   // accumulator of the next column that is not a string
            for (uint32_t jj = 0; jj < nc; ++jj) {
                const uint32_t j = c0 + jj;
                const ColDesc& c = P.cols[j];
                bool here, missing;
                uint32_t fs, fe;
                cell_of(R, c.index, here, missing, fs, fe);
                const uint32_t acc_j = IS_COLS ? (c.type == FALX_COL_STR ? NACC : slot) : jj;
                if (IS_COLS && c.type != FALX_COL_STR) ++slot;
                // Statistics of COLUMNS cover the rows that fit rows_cap.
                const bool stat_ok = with_stats && acc_j < NACC && (IS_COLS ? live : true);
                bool vb = false;
                if (c.type == FALX_COL_F64 || c.type == FALX_COL_I64) {
                    const bool as_f64 = c.type == FALX_COL_F64;
                    NumVal nv{};
                    if (here) nv = num_value(fs, fe, as_f64);
                    CellStat o{false, false, 0, 0, 0, 0.0};
                    if (as_f64) {
                        vb = here && nv.f.ok;
                        if (IS_COLS && (here || missing) && live)
                            reinterpret_cast<double*>(c.values)[row] = vb ? nv.f.value : 0.0;
                        if (vb) o = CellStat{true, nv.f.fixed6, f64_to_bits(nv.f.value), 0, nv.f.mantissa, nv.f.value};
                    } else {
                        vb = here && nv.iok;
                        if (IS_COLS && (here || missing) && live)
                            reinterpret_cast<long long*>(c.values)[row] = vb ? nv.iv : 0;
                        if (vb) o = CellStat{true, false, (unsigned long long)nv.iv, 0, 0, (double)nv.iv};
                    }
                    if (stat_ok) acc_add(acc, acc_j, o);
                } else if (c.type == FALX_COL_STR) {
                    if (IS_COLS) {
                        // Cleaned byte counts only; the STRINGS pass places
                        // the bytes and writes the validity.
                        uint32_t len = here ? clean_len(fs, fe) : 0u;
                        len = row_sum16(len);
                        const uint32_t wsum = lane_value(len, 0) + lane_value(len, 16) + lane_value(len, 32) +
                                              lane_value(len, 48);
                        if (ln == 0 && wsum) atomicAdd(&sh.str_len[c.str_slot], wsum);
                    } else if (here) {
                        unsigned long long cs = 0;
                        const uint32_t nb = clean_cell(src, fs, fe, cd, [&](uint8_t b) { cs += b; });
                        if (stat_ok) acc_add(acc, acc_j, CellStat{true, false, cs, nb, 0, 0.0});
                    }
                } else {
                    vb = here && fe != fs;
                    if (IS_COLS && (here || missing) && live)
                        reinterpret_cast<uint2*>(c.values)[row] = vb ? make_uint2(fs, fe) : make_uint2(0, 0);
                    if (vb && stat_ok) {
                        unsigned long long cs = 0;
                        for (uint32_t q = fs; q < fe; ++q) cs += src.at(q);
                        acc_add(acc, acc_j, CellStat{true, false, cs, fe - fs, 0, 0.0});
                    }
                }
                if (IS_COLS && c.type != FALX_COL_STR) wave_valid(c, vb && live, row, RL);
            }
        });
        if (IS_COLS && tid == 0 && last_tile) P.result[R_N_ROWS] = inc.n_rows + (trailer_row_live ? 1 : 0);
        __syncthreads();
        if (FAST && tid == 0) P.fixup[tile] = sh.fix_bad;
        if (with_stats) acc_publish(acc, acc_col, n_acc, records, IS_COLS || c0 == 0);
        if (IS_COLS && P.n_str > 0 && tid < P.n_str) P.str_len[(size_t)tile * P.n_str + tid] = sh.str_len[tid];
    }
}

// Occupancy the STATS pass is held to (its statistics accumulators sit just
// over the register budget of 4 waves/SIMD otherwise); 1 = no constraint.
template <int MODE>
constexpr int plain_waves() {
    return MODE == M_STATS ? 4
                           : ((MODE == M_STATS_FAST || MODE == M_COLUMNS_FAST_STATS || MODE == M_NUM_FAST_STATS ||
                               MODE == M_NUM_STATS_FAST)
                                  ? 5
                                  : 1);
}

// A document's Params read through the constant address space: nothing
// writes them during a batch, and the compiler then loads their fields with
// scalar loads, as it does for kernel arguments.
typedef __attribute__((address_space(4))) const Params ConstParams;
DEV const Params& const_params(const Params* base, uint32_t i) {
    ConstParams* cp = reinterpret_cast<ConstParams*>(reinterpret_cast<uintptr_t>(base + i));
    return *(const Params*)cp;
}

// One workgroup per tile. Tiles depend on nothing but the tile prefix
// (tpre) and their own bytes, so the block index is the tile and its
// loads start at once.
template <bool HQ, bool ESC, bool CMT, int MODE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(plain_waves<MODE>())))
void k_scan(Params P) {
    __shared__ Shared sh;
    scan_tile<HQ, ESC, CMT, MODE>(P, blockIdx.x, sh);
}

// The full pass after a fast one (P.fixup holds the fast pass's flags): a
// workgroup per 64 tiles runs the flagged ones in order. Usually there are
// none, and the workgroup ends after one read of the flags.
template <bool HQ, bool ESC, bool CMT, int MODE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(4)))
void k_redo(Params P) {
    __shared__ Shared sh;
    __shared__ uint32_t s_mask[2];
    const uint32_t t0 = blockIdx.x * 64u;
    if (threadIdx.x < 64) {
        const uint32_t t = t0 + threadIdx.x;
        const uint64_t m = ballot(t < P.n_tiles && P.fixup[t] != 0u);
        if (threadIdx.x == 0) {
            s_mask[0] = (uint32_t)m;
            s_mask[1] = (uint32_t)(m >> 32);
        }
    }
    __syncthreads();
    uint64_t m = ((uint64_t)uni(s_mask[1]) << 32) | uni(s_mask[0]);
    while (m) {
        // Params are read from the kernel arguments through an address that
        // passes an empty asm each tile, so their loads stay in the loop
        // (hoisted, they would stay live across the tile body and spill).
        uintptr_t pa = reinterpret_cast<uintptr_t>(__builtin_amdgcn_kernarg_segment_ptr());
        asm volatile("" : "+s"(pa));
        scan_tile<HQ, ESC, CMT, MODE>(const_params(reinterpret_cast<const Params*>(pa), 0),
                                      t0 + (uint32_t)__ffsll((long long)m) - 1u, sh);
        __syncthreads();
        m &= m - 1;
    }
}

// --------------------------------------------------------------------------
// reductions, fills, NDJSON id+score
// --------------------------------------------------------------------------
#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_fill_u64(unsigned long long* p, uint64_t n, unsigned long long v) {
    for (uint64_t i = (uint64_t)blockIdx.x * WG + threadIdx.x; i < n; i += (uint64_t)gridDim.x * WG) p[i] = v;
}
#endif

// --------------------------------------------------------------------------
// tile-prefix scan: tile summaries -> each tile's entry state and exclusive
// counters. Summaries compose like the reference's chunk transfer functions
// (csv_hash.rs:184-191): (A then B).fn = B.fn o A.fn and
// (A then B).counts[s] = A.counts[s] then B.counts[A.fn(s)].
// --------------------------------------------------------------------------
constexpr uint32_t SG = 128;   // tiles per scan group

// Loads summary `i` (fn + ns counters) into LDS slots, or the identity.
DEV void sum_to_lds(const uint32_t* src, bool have, uint32_t ns, uint32_t* fn, uint32_t* cs) {
    *fn = have ? src[0] : fn_ident_rt(ns);
    for (uint32_t k = 0; k < ns * C7; ++k) cs[k] = have ? src[TS_COUNTS + k] : 0u;
}

// Level 1: one summary per group of SG tiles (tree reduction in LDS).
DEV void scan_reduce_body(const uint32_t* ts, uint32_t stride, uint32_t n_tiles, uint32_t ns, uint32_t* gs,
                          uint32_t blk) {
    __shared__ uint32_t fn[SG];
    __shared__ uint32_t cs[SG][NS_MAX * C7];
    const uint32_t t = threadIdx.x, tile = blk * SG + t;
    sum_to_lds(ts + (size_t)tile * stride, tile < n_tiles, ns, &fn[t], cs[t]);
    __syncthreads();
    for (uint32_t d = 1; d < SG; d <<= 1) {
        if ((t & (2 * d - 1)) == 0) {
            const uint32_t fa = fn[t], fb = fn[t + d];
            Counts out[NS_MAX];
#pragma unroll
            for (uint32_t s = 0; s < (uint32_t)NS_MAX; ++s)
                if (s < ns) out[s] = compose(c7_load(&cs[t][C7 * s]), c7_load(&cs[t + d][C7 * fn_apply(fa, s)]));
#pragma unroll
            for (uint32_t s = 0; s < (uint32_t)NS_MAX; ++s)
                if (s < ns) c7_store(&cs[t][C7 * s], out[s]);
            fn[t] = fn_then_rt(fa, fb, ns);
        }
        __syncthreads();
    }
    if (t < ns * C7 + 1) {
        uint32_t* dst = gs + (size_t)blk * stride;
        if (t == 0) dst[0] = fn[0];
        else dst[TS_COUNTS + t - 1] = cs[0][t - 1];
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SG) void k_scan_reduce(const uint32_t* ts, uint32_t stride, uint32_t n_tiles,
                                                     uint32_t ns, uint32_t* gs) {
    scan_reduce_body(ts, stride, n_tiles, ns, gs, blockIdx.x);
}
#endif

// Level 2: one wave over the group summaries. Each lane folds a run of
// groups, lane 0 chains the lanes from state 0, and each lane then walks its
// groups with the concrete state, writing every group's entry (state + 7
// counter words). The grand totals go to the result slots.
DEV void scan_groups_body(const uint32_t* gs, uint32_t stride, uint32_t n_groups, uint32_t ns, uint32_t* gpre,
                          unsigned long long* result) {
    __shared__ uint32_t lfn[64];
    __shared__ uint32_t lcs[64][NS_MAX * C7];
    __shared__ uint32_t lent[65][8];
    const uint32_t l = threadIdx.x;
    const uint32_t per = (n_groups + 63) / 64;
    const uint32_t g0 = min(n_groups, l * per), g1 = min(n_groups, g0 + per);
    uint32_t f = fn_ident_rt(ns);
    Counts c[NS_MAX];
#pragma unroll
    for (int s = 0; s < NS_MAX; ++s) c[s] = counts_zero();
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t* p = gs + (size_t)g * stride;
#pragma unroll
        for (uint32_t s = 0; s < (uint32_t)NS_MAX; ++s)
            if (s < ns) c[s] = compose(c[s], c7_load(p + TS_COUNTS + C7 * fn_apply(f, s)));
        f = fn_then_rt(f, p[0], ns);
    }
    lfn[l] = f;
#pragma unroll
    for (uint32_t s = 0; s < (uint32_t)NS_MAX; ++s)
        if (s < ns) c7_store(&lcs[l][C7 * s], c[s]);
    __syncthreads();
    if (l == 0) {
        uint32_t s = 0;
        Counts C = counts_zero();
        for (uint32_t k = 0; k < 64; ++k) {
            lent[k][0] = s;
            c7_store(&lent[k][1], C);
            C = compose(C, c7_load(&lcs[k][C7 * s]));
            s = fn_apply(lfn[k], s);
        }
        lent[64][0] = s;
        c7_store(&lent[64][1], C);
        result[R_N_STRUCT] = C.n_struct;
        result[R_N_SEPS] = C.n_seps;
        result[R_N_TERMS] = C.n_terms;
        result[R_N_ROWS] = C.n_rows;
    }
    __syncthreads();
    uint32_t s = lent[l][0];
    Counts C = c7_load(&lent[l][1]);
    for (uint32_t g = g0; g < g1; ++g) {
        const uint32_t* p = gs + (size_t)g * stride;
        uint32_t* o = gpre + (size_t)g * 8;
        o[0] = s;
        c7_store(o + 1, C);
        C = compose(C, c7_load(p + TS_COUNTS + C7 * s));
        s = fn_apply(p[0], s);
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(64) void k_scan_groups(const uint32_t* gs, uint32_t stride, uint32_t n_groups,
                                                     uint32_t ns, uint32_t* gpre, unsigned long long* result) {
    scan_groups_body(gs, stride, n_groups, ns, gpre, result);
}
#endif

// Levels 2 and 3 as parallel scans, for a known number of entry states NS:
// a summary (transfer function + NS counter sets) is held in registers and
// the workgroup scans them by Kogge-Stone steps through LDS. Level 2 takes
// up to GB_PER groups per thread in one workgroup; level 3 one workgroup per
// group of SG tiles. The results equal those of the serial walks above.
template <int NS>
struct Summ {
    uint32_t fn;
    uint32_t c[NS * C7];
};

template <int NS>
DEV Summ<NS> summ_ident() {
    Summ<NS> r;
    r.fn = fn_ident<NS>();
#pragma unroll
    for (int k = 0; k < NS * C7; ++k) r.c[k] = 0u;
    return r;
}

template <int NS>
DEV Summ<NS> summ_load(const uint32_t* p) {
    Summ<NS> r;
    r.fn = p[0];
#pragma unroll
    for (int k = 0; k < NS * C7; ++k) r.c[k] = p[TS_COUNTS + k];
    return r;
}

// The counters of entry state s (run time) by selects, not register indexing.
template <int NS>
DEV Counts summ_counts(const Summ<NS>& a, uint32_t s) {
    uint32_t w[C7];
#pragma unroll
    for (int k = 0; k < C7; ++k) w[k] = a.c[k];
#pragma unroll
    for (int t = 1; t < NS; ++t)
#pragma unroll
        for (int k = 0; k < C7; ++k) w[k] = s == (uint32_t)t ? a.c[C7 * t + k] : w[k];
    return c7_load(w);
}

// a then b.
template <int NS>
DEV Summ<NS> summ_then(const Summ<NS>& a, const Summ<NS>& b) {
    Summ<NS> r;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        uint32_t w[C7];
#pragma unroll
        for (int k = 0; k < C7; ++k) w[k] = a.c[C7 * s + k];
        const Counts x = compose(c7_load(w), summ_counts(b, fn_apply(a.fn, (uint32_t)s)));
        c7_store(w, x);
#pragma unroll
        for (int k = 0; k < C7; ++k) r.c[C7 * s + k] = w[k];
    }
    r.fn = fn_then<NS>(a.fn, b.fn);
    return r;
}

// Inclusive scan of the workgroup's summaries (thread t holds element t);
// `buf` holds blockDim.x summaries.
template <int NS>
DEV Summ<NS> summ_scan(Summ<NS> e, uint32_t (*buf)[1 + NS * C7]) {
    const uint32_t t = threadIdx.x, n = blockDim.x;
    for (uint32_t d = 1; d < n; d <<= 1) {
        buf[t][0] = e.fn;
#pragma unroll
        for (int k = 0; k < NS * C7; ++k) buf[t][1 + k] = e.c[k];
        __syncthreads();
        if (t >= d) {
            Summ<NS> p;
            p.fn = buf[t - d][0];
#pragma unroll
            for (int k = 0; k < NS * C7; ++k) p.c[k] = buf[t - d][1 + k];
            e = summ_then<NS>(p, e);
        }
        __syncthreads();
    }
    return e;
}

// The summary before thread t's (the identity for thread 0), given the
// inclusive scan.
template <int NS>
DEV Summ<NS> summ_exclusive(const Summ<NS>& incl, uint32_t (*buf)[1 + NS * C7]) {
    const uint32_t t = threadIdx.x;
    buf[t][0] = incl.fn;
#pragma unroll
    for (int k = 0; k < NS * C7; ++k) buf[t][1 + k] = incl.c[k];
    __syncthreads();
    Summ<NS> x = summ_ident<NS>();
    if (t > 0) {
        x.fn = buf[t - 1][0];
#pragma unroll
        for (int k = 0; k < NS * C7; ++k) x.c[k] = buf[t - 1][1 + k];
    }
    return x;
}

constexpr uint32_t GB_THREADS = 256;   // level 2 workgroup

#ifndef FX_KIND   // defined once, in the host translation unit
template <int NS>
__global__ __launch_bounds__(GB_THREADS) void k_scan_groups_par(const uint32_t* gs, uint32_t stride, uint32_t n_groups,
                                                                uint32_t* gpre, unsigned long long* result) {
    __shared__ uint32_t buf[GB_THREADS][1 + NS * C7];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (n_groups + GB_THREADS - 1) / GB_THREADS;
    const uint32_t g0 = min(n_groups, t * per), g1 = min(n_groups, g0 + per);
    Summ<NS> e = summ_ident<NS>();
    for (uint32_t g = g0; g < g1; ++g) e = summ_then<NS>(e, summ_load<NS>(gs + (size_t)g * stride));
    const Summ<NS> incl = summ_scan<NS>(e, buf);
    const Summ<NS> ex = summ_exclusive<NS>(incl, buf);
    // Walk this thread's groups from the state entering them (state 0 at
    // the document's start).
    uint32_t s = fn_apply(ex.fn, 0u);
    Counts C = summ_counts<NS>(ex, 0u);
    for (uint32_t g = g0; g < g1; ++g) {
        const Summ<NS> q = summ_load<NS>(gs + (size_t)g * stride);
        uint32_t* o = gpre + (size_t)g * 8;
        o[0] = s;
        c7_store(o + 1, C);
        C = compose(C, summ_counts<NS>(q, s));
        s = fn_apply(q.fn, s);
    }
    if (t == GB_THREADS - 1) {
        const Counts T = summ_counts<NS>(incl, 0u);
        result[R_N_STRUCT] = T.n_struct;
        result[R_N_SEPS] = T.n_seps;
        result[R_N_TERMS] = T.n_terms;
        result[R_N_ROWS] = T.n_rows;
    }
}

template <int NS>
__global__ __launch_bounds__(SG) void k_scan_tiles_par(const uint32_t* ts, uint32_t stride, uint32_t n_tiles,
                                                        const uint32_t* gpre, uint32_t* tpre) {
    __shared__ uint32_t buf[SG][1 + NS * C7];
    const uint32_t t = threadIdx.x, tile = blockIdx.x * SG + t;
    const bool have = tile < n_tiles;
    const Summ<NS> e = have ? summ_load<NS>(ts + (size_t)tile * stride) : summ_ident<NS>();
    const Summ<NS> incl = summ_scan<NS>(e, buf);
    const Summ<NS> ex = summ_exclusive<NS>(incl, buf);
    const uint32_t* g = gpre + (size_t)blockIdx.x * 8;
    const uint32_t s0 = g[0];
    const Counts C = compose(c7_load(g + 1), summ_counts<NS>(ex, s0));
    if (have) {
        uint32_t* o = tpre + (size_t)tile * 8;
        o[0] = fn_apply(ex.fn, s0);
        c7_store(o + 1, C);
    }
}
#endif

// String arena bases: exclusive prefix over tiles of one string slot's
// cleaned byte counts (blockIdx.x = slot). Each thread sums a run of tiles,
// the block scans the runs, and each thread writes its tiles' offsets; the
// total goes to result slot R_STR_LEN + slot.
constexpr uint32_t SLOT_WG = 1024;

DEV void slot_scan_body(const uint32_t* len, uint32_t n_tiles, uint32_t ns, uint32_t* base,
                        unsigned long long* result, uint32_t slot) {
    __shared__ unsigned long long wtot[SLOT_WG / 64];
    const uint32_t t = threadIdx.x, ln = __lane_id(), w = t >> 6;
    const uint32_t per = (n_tiles + SLOT_WG - 1) / SLOT_WG;
    const uint32_t t0 = min(n_tiles, t * per), t1 = min(n_tiles, t0 + per);
    unsigned long long sum = 0;
    for (uint32_t i = t0; i < t1; ++i) sum += len[(size_t)i * ns + slot];
    unsigned long long incl = sum;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const unsigned long long u = __shfl_up(incl, d);
        if (ln >= (uint32_t)d) incl += u;
    }
    if (ln == 63) wtot[w] = incl;
    __syncthreads();
    unsigned long long run = incl - sum;
    for (uint32_t k = 0; k < w; ++k) run += wtot[k];
    for (uint32_t i = t0; i < t1; ++i) {
        base[(size_t)i * ns + slot] = (uint32_t)run;
        run += len[(size_t)i * ns + slot];
    }
    if (t == SLOT_WG - 1) result[R_STR_LEN + slot] = run;
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SLOT_WG) void k_slot_scan(const uint32_t* len, uint32_t n_tiles, uint32_t ns,
                                                       uint32_t* base, unsigned long long* result) {
    slot_scan_body(len, n_tiles, ns, base, result, blockIdx.x);
}

// The same scan for a whole document in many workgroups: workgroup
// (chunk, slot) takes SLOT_CHUNK tiles, SLOT_PER per thread. Its base is the
// sum of the lengths of every tile before the chunk, which it reads itself
// (coalesced, at most n_tiles words), so no workgroup waits for another.
constexpr uint32_t SLOT_PER = 4;
constexpr uint32_t SLOT_CHUNK = SLOT_WG * SLOT_PER;

__global__ __launch_bounds__(SLOT_WG) void k_slot_scan_wide(const uint32_t* len, uint32_t n_tiles, uint32_t ns,
                                                            uint32_t* base, unsigned long long* result) {
    __shared__ unsigned long long wsh[SLOT_WG / 64];
    __shared__ uint32_t wsh32[SLOT_WG / 64];
    const uint32_t t = threadIdx.x, ln = __lane_id(), w = t >> 6, slot = blockIdx.y;
    const uint32_t c0 = blockIdx.x * SLOT_CHUNK;
    // Lengths of the tiles before the chunk.
    unsigned long long pre = 0;
    for (uint32_t i = t; i < c0; i += SLOT_WG) pre += len[(size_t)i * ns + slot];
    for (int d = 32; d > 0; d >>= 1) pre += __shfl_xor(pre, d);
    if (ln == 0) wsh[w] = pre;
    // The chunk: SLOT_PER consecutive tiles per thread, scanned in the block.
    const uint32_t i0 = c0 + t * SLOT_PER;
    uint32_t v[SLOT_PER], sum = 0;
#pragma unroll
    for (uint32_t k = 0; k < SLOT_PER; ++k) {
        v[k] = i0 + k < n_tiles ? len[(size_t)(i0 + k) * ns + slot] : 0u;
        sum += v[k];
    }
    const uint32_t incl = wave_incl_sum(sum);
    if (ln == 63) wsh32[w] = incl;
    __syncthreads();
    unsigned long long run = 0;
    uint32_t wb = 0;
#pragma unroll
    for (uint32_t k = 0; k < SLOT_WG / 64; ++k) {
        run += wsh[k];
        if (k < w) wb += wsh32[k];
    }
    run += wb + incl - sum;
#pragma unroll
    for (uint32_t k = 0; k < SLOT_PER; ++k) {
        const uint32_t i = i0 + k;
        if (i < n_tiles) {
            base[(size_t)i * ns + slot] = (uint32_t)run;
            run += v[k];
            if (i + 1 == n_tiles) result[R_STR_LEN + slot] = run;
        }
    }
}
#endif

// Level 3: per group, thread 0 walks its SG tiles from the group's entry.
DEV void scan_tiles_body(const uint32_t* ts, uint32_t stride, uint32_t n_tiles, uint32_t ns, const uint32_t* gpre,
                         uint32_t* tpre, uint32_t blk) {
    __shared__ uint32_t fn[SG];
    __shared__ uint32_t cs[SG][NS_MAX * C7];
    __shared__ uint32_t out[SG][8];
    const uint32_t t = threadIdx.x, tile = blk * SG + t;
    const bool have = tile < n_tiles;
    sum_to_lds(ts + (size_t)tile * stride, have, ns, &fn[t], cs[t]);
    __syncthreads();
    if (t == 0) {
        const uint32_t* g = gpre + (size_t)blk * 8;
        uint32_t s = g[0];
        Counts C = c7_load(g + 1);
        for (uint32_t k = 0; k < SG; ++k) {
            out[k][0] = s;
            c7_store(&out[k][1], C);
            C = compose(C, c7_load(&cs[k][C7 * s]));
            s = fn_apply(fn[k], s);
        }
    }
    __syncthreads();
    if (have)
        for (int i = 0; i < 8; ++i) tpre[(size_t)tile * 8 + i] = out[t][i];
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SG) void k_scan_tiles(const uint32_t* ts, uint32_t stride, uint32_t n_tiles,
                                                    uint32_t ns, const uint32_t* gpre, uint32_t* tpre) {
    scan_tiles_body(ts, stride, n_tiles, ns, gpre, tpre, blockIdx.x);
}
#endif

// Deterministic two-level reductions over per-tile partials: level 1 folds
// RED_BLOCKS contiguous tile ranges in a fixed order, level 2 folds those.
constexpr uint32_t RED_BLOCKS = 256;

// Sums `width` u64 lanes over `n` rows (row-major) into out[0..width).
// grid.x = number of row chunks (1 = single pass).
DEV void reduce_u64_body(const unsigned long long* parts, uint32_t n, uint32_t width, unsigned long long* out,
                         uint32_t blk, uint32_t nblk) {
    __shared__ unsigned long long sh[NWAVE];
    const uint32_t chunk = (n + nblk - 1) / nblk;
    const uint32_t lo = blk * chunk, hi = min(n, lo + chunk);
    for (uint32_t c = 0; c < width; ++c) {
        unsigned long long v = 0;
        for (uint32_t i = lo + threadIdx.x; i < hi; i += WG) v += parts[(size_t)i * width + c];
        v = block_reduce_sum(v, sh);
        if (threadIdx.x == 0) out[(size_t)blk * width + c] = v;
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_reduce_u64(const unsigned long long* parts, uint32_t n, uint32_t width,
                                                   unsigned long long* out) {
    reduce_u64_body(parts, n, width, out, blockIdx.x, gridDim.x);
}
#endif

// Per-column stat reduction. grid = (n_cols, chunks); fixed-order FP64 folds.
DEV void reduce_stats_body(const StatPart* parts, uint32_t n_tiles, uint32_t n_cols, StatPart* out, uint32_t c,
                           uint32_t blk, uint32_t nblk) {
    __shared__ unsigned long long ush[NWAVE];
    __shared__ long long lsh[NWAVE];
    __shared__ double dsh[NWAVE];
    const uint32_t chunk = (n_tiles + nblk - 1) / nblk;
    const uint32_t lo = blk * chunk, hi = min(n_tiles, lo + chunk);
    unsigned long long v = 0, cs = 0, by = 0, fc = 0;
    long long fs = 0;
    double ds = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += WG) {
        const StatPart& p = parts[(size_t)i * n_cols + c];
        v += p.values; cs += p.checksum; by += p.bytes; fs += p.fixed6_sum; fc += p.fixed6_cells; ds += p.sum;
    }
    v = block_reduce_sum(v, ush);
    cs = block_reduce_sum(cs, ush);
    by = block_reduce_sum(by, ush);
    fc = block_reduce_sum(fc, ush);
    fs = block_reduce_sum(fs, lsh);
    ds = block_reduce_sum(ds, dsh);
    if (threadIdx.x == 0) {
        StatPart& o = out[(size_t)blk * n_cols + c];
        o.values = v; o.checksum = cs; o.bytes = by;
        o.fixed6_sum = fs; o.fixed6_cells = fc; o.sum = ds;
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_reduce_stats(const StatPart* parts, uint32_t n_tiles, uint32_t n_cols,
                                                     StatPart* out) {
    reduce_stats_body(parts, n_tiles, n_cols, out, blockIdx.x, blockIdx.y, gridDim.y);
}
#endif

struct NdPart {
    unsigned long long records, invalid;
    long long id_sum, fx_sum;
    unsigned long long fx_cells;
    double score_sum;
    unsigned long long pad[2];
};

// NDJSON id+score (ndjson.rs:679-723): records are the bytes between raw
// '\n's. A block of ND_WAVES waves owns ND_WAVES 8 KiB tiles and holds in LDS
// their bytes plus ND_HALO bytes after them, so a record is read from LDS
// alone unless it runs past the halo (then one lane walks it with HBM reads).
// Each wave lists the newlines that end the records starting in its tile, in
// order, 1 KiB per step (16 bytes per lane, SWAR byte tests, ranks by a wave
// scan); then the block's 256 lanes read the records of all its tiles, lane t
// records t, t + 256, ... of the block's list: first in the straight-line
// steps of the common layout (ndjson_common), which every lane takes together,
// and with the record walker (ndjson_record) when a record has another form.
// Pooling the tiles keeps lanes busy: a tile holds about 1.2 records per lane
// of one wave, so a wave per tile would run a second round mostly idle. The
// fixed assignment also fixes the order of the f64 score sum.
constexpr uint32_t ND_WAVES = 4;
constexpr int ND_WG = 64 * ND_WAVES;
constexpr uint32_t ND_SPAN = ND_WAVES * TILE;     // bytes of the tiles of a block
constexpr uint32_t ND_HALO = 512;
constexpr uint32_t ND_WIN = ND_SPAN + ND_HALO;
constexpr uint32_t ND_MAXNL = 512;      // newlines listed per wave and round
// Zero bytes after the window: the walker's word reads (three aligned words
// from a byte of the record) end at most 24 bytes past its last byte.
constexpr uint32_t ND_PAD = 32;
constexpr uint32_t ND_NONE = 0xFFFFFFFFu;

// The block's LDS copy as a walker source (ndjson_walk.h): bytes, and
// aligned 8-byte words for skipping strings (block starts are multiples of 8).
struct NdLds {
    static constexpr bool kWord8 = true;
    const uint8_t* b;   // document bytes [lo, lo + ND_WIN), then ND_PAD zero bytes
    uint32_t lo;
    DEV uint8_t at(uint32_t p) const { return b[p - lo]; }
    DEV uint64_t word8(uint32_t p) const { return *reinterpret_cast<const uint64_t*>(b + (p - lo)); }
};

struct NdAcc {
    unsigned long long recs, inval, fxc, ids;
    long long fxs;
    double ss;
};

// One record's outcome: 0 blank, 1 valid, 2 invalid (returned by value and
// folded with selects, so the accumulators stay in registers).
struct NdOne {
    uint32_t kind;
    bool fx6;
    int64_t id, mant;
    double score;
};

template <class Src>
DEV NdOne nd_record(const Src& src, uint32_t from, uint32_t to, SlowF64& slow) {
    NdOne o{0, false, 0, 0, 0.0};
    if (blank_span(src, from, to)) return o;
    NdRecord r;
    if (ndjson_record(src, from, to, r)) {
        if (r.score_deferred) {
            if (ndjson_score_general(src, r)) resolve_score(slow, src, r);
            r.score_deferred = false;
        }
        o.kind = 1;
        o.fx6 = r.fx6;
        o.id = r.id;
        o.mant = r.mant;
        o.score = r.score;
    } else {
        o.kind = 2;
    }
    return o;
}

DEV void nd_add(NdAcc& a, const NdOne& o) {
    const bool v = o.kind == 1;
    a.recs += v ? 1u : 0u;
    a.inval += o.kind == 2 ? 1u : 0u;
    a.ids += v ? (unsigned long long)o.id : 0ULL;
    a.ss += v ? o.score : 0.0;
    a.fxs += (v && o.fx6) ? o.mant : 0;
    a.fxc += (v && o.fx6) ? 1u : 0u;
}

// Document bytes read from HBM (records running past a tile's LDS window).
struct GlobalSrc {
    const uint8_t* g;
    DEV uint8_t at(uint32_t p) const { return g[p]; }
};

// Newline bytes of 32 bytes (x[0..7], little-endian words) as a 32-bit mask,
// bit i for byte i (byte transpose, then one byte test per word).
DEV uint32_t nl_mask32(const uint32_t (&x)[8]) {
    uint32_t y[8];   // y[j] byte k = byte 8k + j
    transpose4(x[0], x[2], x[4], x[6], y[0], y[1], y[2], y[3]);
    transpose4(x[1], x[3], x[5], x[7], y[4], y[5], y[6], y[7]);
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) m |= (eq_ascii(y[j] & 0x7F7F7F7Fu, ~y[j] & 0x80808080u, 0x0A0A0A0Au) >> 7) << j;
    return m;
}

// far[tile] receives the one record of the tile that runs past the block's
// LDS window, (start, end), or (ND_NONE, 0); k_ndjson_far walks those
// records. The block's sums go to parts[first tile of the block], zeros to
// the parts of its other tiles.
DEV void ndjson_block(const uint8_t* data, uint32_t len, uint32_t n_tiles, NdPart* parts, uint2* far, uint32_t blk) {
    __shared__ __align__(16) uint8_t bytes[ND_WIN + ND_PAD];
    __shared__ uint16_t nls[ND_WAVES][ND_MAXNL];   // per wave: newline positions - block start, in order
    __shared__ uint32_t w_n[ND_WAVES], w_items[ND_WAVES], w_cur[ND_WAVES];
    __shared__ SlowF64 slow;
    __shared__ unsigned long long ush[ND_WAVES];
    __shared__ long long lsh[ND_WAVES];
    __shared__ double dsh[ND_WAVES];
    const uint32_t tid = threadIdx.x, ln = tid & 63u, w = uni(tid >> 6);
    const uint32_t blk_start = blk * ND_SPAN;
    const uint32_t win_hi = (uint32_t)min((uint64_t)blk_start + ND_WIN, (uint64_t)len);
    const uint32_t win_len = win_hi - blk_start;
    for (uint32_t i = tid * 16; i < ND_WIN + ND_PAD; i += ND_WG * 16) {
        uint4 v = make_uint4(0, 0, 0, 0);
        if (i + 16 <= win_len) {
            v = *reinterpret_cast<const uint4*>(data + blk_start + i);
        } else if (i < win_len) {
            uint8_t tmp[16];
            for (int k = 0; k < 16; ++k) tmp[k] = (i + k < win_len) ? data[blk_start + i + k] : 0;
            memcpy(&v, tmp, 16);
        }
        *reinterpret_cast<uint4*>(&bytes[i]) = v;
    }
    if (tid == 0) slow.lock = 0;
    __syncthreads();
    const NdLds src{bytes, blk_start};
    NdAcc acc{0, 0, 0, 0, 0, 0.0};
    uint2 far_rec = make_uint2(ND_NONE, 0u);
    // This wave's tile.
    const uint32_t tile = blk * ND_WAVES + w;
    const bool live = tile < n_tiles;
    const uint32_t tile_start = blk_start + w * (uint32_t)TILE;
    const uint32_t tile_end = live ? (uint32_t)min((uint64_t)tile_start + TILE, (uint64_t)len) : tile_start;
    // Start of the record being listed: the tile's first byte when a record
    // starts there, else it is the byte after the tile's first newline.
    uint32_t cur = (live && (tile_start == 0 || data[tile_start - 1] == '\n')) ? tile_start : ND_NONE;
    uint32_t scan = tile_start;     // first byte not yet searched for newlines
    bool done = !live;
    while (true) {
        // List newlines from `scan` on, stopping after the first one at or
        // past tile_end - 1 (the record after it is the next tile's), or
        // when the list is full. 32 bytes per lane and step.
        uint32_t n = 0;
        bool stop = false, full = false;
        while (!done && !stop && !full && scan < win_hi) {
            const uint32_t base = scan & ~31u;
            const uint32_t off = base - blk_start + 32 * ln;   // LDS offset of this lane's 32 bytes
            const uint32_t p0 = base + 32 * ln;
            uint32_t m = 0;
            if (off + 32 <= ND_WIN + ND_PAD) {
                const uint4 v0 = *reinterpret_cast<const uint4*>(&bytes[off]);
                const uint4 v1 = *reinterpret_cast<const uint4*>(&bytes[off + 16]);
                const uint32_t x[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
                m = nl_mask32(x);
            }
            // Positions >= scan and < win_hi only.
            if (p0 < scan) m &= scan - p0 >= 32 ? 0u : ~0u << (scan - p0);
            if (p0 + 32 > win_hi) m &= win_hi <= p0 ? 0u : (1u << (win_hi - p0)) - 1u;
            // The stopping newline: the first one at or past tile_end - 1.
            const uint32_t lim = tile_end - 1;
            uint32_t cand = m;
            if (p0 + 32 <= lim) cand = 0;
            else if (p0 < lim) cand &= ~0u << (lim - p0);
            const uint64_t cl = ballot(cand != 0);
            uint32_t stop_pos = ND_NONE;
            if (cl) {
                const uint32_t L = (uint32_t)__ffsll((long long)cl) - 1u;
                stop_pos = lane_value(p0 + (uint32_t)__ffs((int)cand) - 1u, L);
                if (p0 > stop_pos) m = 0;
                else if (p0 + 32 > stop_pos) m &= (2u << (stop_pos - p0)) - 1u;   // 2u << 31 wraps to 0: all bits
                stop = true;
            }
            const uint32_t cnt = popc((uint64_t)m);
            const uint32_t incl = wave_incl_sum(cnt);
            const uint32_t total = lane_value(incl, 63);
            uint32_t rank = n + incl - cnt;
            for (uint32_t mm = m; mm; mm &= mm - 1) {
                if (rank < ND_MAXNL) nls[w][rank] = (uint16_t)(p0 + (uint32_t)__ffs((int)mm) - 1u - blk_start);
                ++rank;
            }
            if (n + total >= ND_MAXNL && (n + total > ND_MAXNL || !stop)) {
                // Full: the next round resumes after the last listed newline.
                n = ND_MAXNL;
                full = true;
                stop = false;
            } else {
                n += total;
                scan = base + 32 * 64;
            }
        }
        __syncthreads();
        // Item k of a wave ends at its newline k (at len for the data's last
        // record, which has none); its start is the wave's `cur` for k = 0,
        // else the byte after newline k - 1.
        uint32_t items = 0, next_cur = cur;
        bool open_end = false;
        if (!done) {
            if (full) scan = blk_start + nls[w][ND_MAXNL - 1] + 1u;
            next_cur = n ? blk_start + nls[w][n - 1] + 1u : cur;
            open_end = !stop && !full && scan >= win_hi;   // no newline ends the record at next_cur
            const bool last_rec = open_end && win_hi == len && next_cur != ND_NONE && next_cur < tile_end &&
                                  next_cur < len;
            items = n + (last_rec ? 1u : 0u);
        }
        if (ln == 0) {
            w_n[w] = n;
            w_items[w] = items;
            w_cur[w] = cur;
        }
        __syncthreads();
        // The items of all waves, in wave order, ND_WG lanes at a time.
        uint32_t pre[ND_WAVES + 1];
        pre[0] = 0;
#pragma unroll
        for (uint32_t v = 0; v < ND_WAVES; ++v) pre[v + 1] = pre[v] + w_items[v];
        for (uint32_t g = tid; g < pre[ND_WAVES]; g += ND_WG) {
            uint32_t v = 0;
#pragma unroll
            for (uint32_t q = 1; q < ND_WAVES; ++q) v += g >= pre[q] ? 1u : 0u;
            const uint32_t k = g - pre[v];
            const uint32_t from = k == 0 ? w_cur[v] : blk_start + nls[v][k - 1] + 1u;
            const uint32_t to = k < w_n[v] ? blk_start + nls[v][k] : len;
            const bool have = from != ND_NONE;
            NdRecord r;
            const bool common = ndjson_common(src, have ? from : to, to, r);
            if (common) nd_add(acc, NdOne{1u, r.fx6, r.id, r.mant, r.score});
            else if (have) nd_add(acc, nd_record(src, from, to, slow));
        }
        __syncthreads();   // the lists are read before the next round writes them
        if (!done) {
            cur = next_cur;
            if (stop) {
                done = true;
            } else if (open_end) {
                if (cur != ND_NONE && cur < tile_end && win_hi < len) {
                    // A record running past the halo: its end is searched in
                    // HBM 512 bytes per step, and one lane walks it there.
                    uint32_t end = len;
                    for (uint32_t lo = win_hi & ~7u; lo < len; lo += 512) {
                        const uint32_t p = lo + 8 * ln;
                        uint64_t z = 0;
                        if (p + 8 <= len) {
                            z = zero_bytes64(*reinterpret_cast<const uint64_t*>(data + p) ^ 0x0A0A0A0A0A0A0A0AULL);
                        } else if (p < len) {
                            for (uint32_t q = p; q < len; ++q)
                                if (data[q] == '\n') z |= 0x80ULL << (8 * (q - p));
                        }
                        // Bytes below win_hi were searched in the window.
                        if (p < win_hi) z = win_hi - p >= 8 ? 0ULL : z & (~0ULL << (8 * (win_hi - p)));
                        const uint64_t hit = ballot(z != 0ULL);
                        if (hit) {
                            const uint32_t L = (uint32_t)__ffsll((long long)hit) - 1u;
                            end = lane_value(p + ((uint32_t)__builtin_ctzll(z) >> 3), L);
                            break;
                        }
                    }
                    far_rec = make_uint2(cur, end);
                }
                done = true;
            }
        }
        if (!__syncthreads_or(done ? 0 : 1)) break;
    }
    acc.recs = block_reduce_sum<unsigned long long, ND_WAVES>(acc.recs, ush);
    acc.inval = block_reduce_sum<unsigned long long, ND_WAVES>(acc.inval, ush);
    acc.fxc = block_reduce_sum<unsigned long long, ND_WAVES>(acc.fxc, ush);
    acc.ids = block_reduce_sum<unsigned long long, ND_WAVES>(acc.ids, ush);
    acc.fxs = block_reduce_sum<long long, ND_WAVES>(acc.fxs, lsh);
    acc.ss = block_reduce_sum<double, ND_WAVES>(acc.ss, dsh);
    if (ln == 0 && live) {
        NdPart& p = parts[tile];
        const bool first = w == 0;
        p.records = first ? acc.recs : 0; p.invalid = first ? acc.inval : 0;
        p.id_sum = first ? (long long)acc.ids : 0;
        p.fx_sum = first ? acc.fxs : 0; p.fx_cells = first ? acc.fxc : 0; p.score_sum = first ? acc.ss : 0.0;
        far[tile] = far_rec;
    }
}

// The records that ran past their tile's LDS window, one lane per tile,
// read from HBM; results go to fparts[tile] (zeros for tiles without one).
DEV void ndjson_far_tile(const uint8_t* data, const uint2* far, NdPart* fparts, uint32_t tile, bool have,
                         SlowF64& slow) {
    NdAcc acc{0, 0, 0, 0, 0, 0.0};
    const uint2 f = have ? far[tile] : make_uint2(ND_NONE, 0u);
    if (f.x != ND_NONE) nd_add(acc, nd_record(GlobalSrc{data}, f.x, f.y, slow));
    if (have) {
        NdPart& p = fparts[tile];
        p.records = acc.recs; p.invalid = acc.inval; p.id_sum = (long long)acc.ids;
        p.fx_sum = acc.fxs; p.fx_cells = acc.fxc; p.score_sum = acc.ss;
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(ND_WG) void k_ndjson_id_score(const uint8_t* data, uint32_t len, uint32_t n_tiles,
                                                        NdPart* parts, uint2* far) {
    ndjson_block(data, len, n_tiles, parts, far, blockIdx.x);
}

__global__ __launch_bounds__(64) void k_ndjson_far(const uint8_t* data, uint32_t n_tiles, const uint2* far,
                                                   NdPart* fparts) {
    __shared__ SlowF64 slow;
    if (threadIdx.x == 0) slow.lock = 0;
    __syncthreads();
    const uint32_t t = blockIdx.x * 64u + threadIdx.x;
    ndjson_far_tile(data, far, fparts, t, t < n_tiles, slow);
}
#endif

DEV void reduce_nd_body(const NdPart* parts, uint32_t n, NdPart* out, uint32_t blk, uint32_t nblk) {
    __shared__ unsigned long long ush[NWAVE];
    __shared__ long long lsh[NWAVE];
    __shared__ double dsh[NWAVE];
    const uint32_t chunk = (n + nblk - 1) / nblk;
    const uint32_t lo = blk * chunk, hi = min(n, lo + chunk);
    unsigned long long r = 0, iv = 0, fc = 0, id = 0;
    long long fs = 0;
    double ss = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += WG) {
        r += parts[i].records; iv += parts[i].invalid; fc += parts[i].fx_cells;
        id += (unsigned long long)parts[i].id_sum;
        fs += parts[i].fx_sum; ss += parts[i].score_sum;
    }
    r = block_reduce_sum(r, ush);
    iv = block_reduce_sum(iv, ush);
    fc = block_reduce_sum(fc, ush);
    id = block_reduce_sum(id, ush);
    fs = block_reduce_sum(fs, lsh);
    ss = block_reduce_sum(ss, dsh);
    if (threadIdx.x == 0) {
        NdPart& o = out[blk];
        o.records = r; o.invalid = iv; o.fx_cells = fc; o.id_sum = (long long)id; o.fx_sum = fs; o.score_sum = ss;
    }
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_reduce_nd(const NdPart* parts, uint32_t n, NdPart* out) {
    reduce_nd_body(parts, n, out, blockIdx.x, gridDim.x);
}
#endif

// --------------------------------------------------------------------------
// batch stages over the other kernels' bodies
// --------------------------------------------------------------------------
// Per-document arguments of the tile-prefix scan.
struct BPrefix {
    uint32_t* tsum;
    uint32_t* gsum;
    uint32_t* gpre;
    uint32_t* tpre;
    unsigned long long* result;
    uint32_t stride, n_tiles, ns, groups;
};

// One fixed-order reduction job (the plain entry points' two-level folds):
// level 1 folds `blocks` contiguous runs of n rows, level 2 folds those.
enum : uint32_t { RJ_STATS = 0, RJ_U64 = 1, RJ_ND = 2 };
struct BRed {
    const void* parts;
    void* tmp;     // level-1 outputs when blocks > 1
    void* out;
    uint32_t kind, n, width, blocks;   // width: columns (stats) or u64 lanes
};

struct BNd {
    const uint8_t* data;
    uint32_t len, n_tiles;
    NdPart* parts;    // 2 n_tiles: the tiles', then the far records'
    uint2* far;
};

struct BFill {
    unsigned long long* p;
    uint64_t n;
    unsigned long long v;
};

// A segment table by pointers into the batch's uploaded descriptor blob.
struct SegT {
    uint32_t n, total;
    const uint32_t* req;
    const uint32_t* chunk;
    const uint32_t* start;
};

// Segment of unit u < T.total: the last i with start[i] <= u, by a 64-way
// search (each lane tests one start per round: two rounds up to 4096
// segments). Called by all lanes of a wave with the same u.
DEV uint32_t seg_find(const SegT& T, uint32_t u) {
    const uint32_t ln = __lane_id();
    uint32_t lo = 0, hi = T.n;   // start[lo] <= u < start[hi] (start[n] = total)
    while (hi - lo > 1) {
        const uint32_t step = (hi - lo + 63u) >> 6;
        const uint32_t i = lo + ln * step;
        const uint64_t le = ballot(i < hi && T.start[i] <= u);   // lane 0 (i = lo) always
        lo += hibit(le) * step;
        hi = min(hi, lo + step);
    }
    return lo;
}

// Stage tables of a batch. Tile stages: summary per dialect kind, emit per
// (kind, emit mode), NDJSON tiles. Unit stages: prefix groups, documents,
// string slots, reduction levels, fills (4096 words a unit).
enum : uint32_t { E_INDEX = 0, E_COLUMNS, E_STATS, E_STRINGS, E_LOGFMT, E_FIELDBYTES, N_EMIT };
enum : uint32_t {
    T_SUM = 0, T_GROUPS = 6, T_PREQ = 7, T_SLOT = 8, T_NDJ = 9, T_RED1 = 10, T_RED2 = 11, T_FILL = 12,
    T_EMIT = 13, N_TAB = T_EMIT + 6 * N_EMIT
};
constexpr uint32_t FILL_UNIT = 4096;

struct BatchHdr {
    const Params* params;
    const BPrefix* pre;
    const BRed* red;
    const BNd* nd;
    const BFill* fill;
    SegT tab[N_TAB];
};


// Persistent tile pass of table `tab` (a summary or emit table).
template <int MODE>
constexpr int batch_waves() { return 4; }

// Units go to workgroups by stride (unit u to workgroup u mod grid): no
// shared counter, whose atomics serialise at one address.
template <bool HQ, bool ESC, bool CMT, int MODE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(batch_waves<MODE>())))
void k_tiles_batch(const BatchHdr* H, uint32_t tab) {
    __shared__ Shared sh;
    __shared__ uint32_t s_p, s_t0, s_t1;
    const SegT& T = H->tab[tab];
    const uint32_t total = T.total;
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        if (threadIdx.x < 64) {
            const uint32_t i = seg_find(T, u);
            if (threadIdx.x == 0) {
                const uint32_t pi = T.req[i], c = T.chunk[i];
                const uint32_t t0 = (u - T.start[i]) * c;
                s_p = pi;
                s_t0 = t0;
                s_t1 = min(t0 + c, H->params[pi].n_tiles);
            }
        }
        __syncthreads();
        const uint32_t pi = uni(s_p), t0 = uni(s_t0), t1 = uni(s_t1);
        __syncthreads();
        for (uint32_t t = t0; t < t1; ++t) {
            // The Params address passes an empty asm each tile, so their
            // loads stay inside the loop: hoisted, they would stay live
            // across the whole tile body and spill.
            const uintptr_t pv = reinterpret_cast<uintptr_t>(H->params);
            uintptr_t pa = (uintptr_t)uni((uint32_t)pv) | ((uintptr_t)uni((uint32_t)(pv >> 32)) << 32);
            asm volatile("" : "+s"(pa));
            scan_tile<HQ, ESC, CMT, MODE>(const_params(reinterpret_cast<const Params*>(pa), pi), t, sh);
            __syncthreads();
        }
    }
}

// Unit u of table T as (document, local unit).
struct Unit {
    uint32_t req, k;
};
DEV Unit unit_at(const SegT& T, uint32_t u) {
    const uint32_t i = seg_find(T, u);
    return Unit{T.req[i], u - T.start[i]};
}

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SG) void k_prefix_reduce_batch(const BatchHdr* H) {
    const SegT& T = H->tab[T_GROUPS];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const Unit x = unit_at(T, u);
        const BPrefix& A = H->pre[x.req];
        scan_reduce_body(A.tsum, A.stride, A.n_tiles, A.ns, A.gsum, x.k);
        __syncthreads();
    }
}
#endif

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(64) void k_prefix_groups_batch(const BatchHdr* H) {
    const SegT& T = H->tab[T_PREQ];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const BPrefix& A = H->pre[unit_at(T, u).req];
        scan_groups_body(A.gsum, A.stride, A.groups, A.ns, A.gpre, A.result);
        __syncthreads();
    }
}
#endif

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SG) void k_prefix_tiles_batch(const BatchHdr* H) {
    const SegT& T = H->tab[T_GROUPS];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const Unit x = unit_at(T, u);
        const BPrefix& A = H->pre[x.req];
        scan_tiles_body(A.tsum, A.stride, A.n_tiles, A.ns, A.gpre, A.tpre, x.k);
        __syncthreads();
    }
}
#endif

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(SLOT_WG) void k_slot_scan_batch(const BatchHdr* H) {
    const SegT& T = H->tab[T_SLOT];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const Unit x = unit_at(T, u);
        const Params& P = H->params[x.req];
        slot_scan_body(P.str_len, P.n_tiles, P.n_str, const_cast<uint32_t*>(P.str_base), P.result, x.k);
        __syncthreads();
    }
}
#endif

// Reduction level 1 (units: stats jobs width x blocks, others blocks) or
// level 2 (units: stats jobs width, others 1; jobs with blocks > 1 only).
#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_reduce_batch(const BatchHdr* H, uint32_t level) {
    const SegT& T = H->tab[level == 1 ? T_RED1 : T_RED2];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const Unit x = unit_at(T, u);
        const BRed& J = H->red[x.req];
        const bool l1 = level == 1;
        const void* src = l1 ? J.parts : J.tmp;
        void* dst = (l1 && J.blocks > 1) ? J.tmp : J.out;
        const uint32_t n = l1 ? J.n : J.blocks, nblk = l1 ? J.blocks : 1u;
        if (J.kind == RJ_STATS) {
            const uint32_t c = l1 ? x.k / J.blocks : x.k, b = l1 ? x.k % J.blocks : 0u;
            reduce_stats_body(static_cast<const StatPart*>(src), n, J.width, static_cast<StatPart*>(dst), c, b, nblk);
        } else if (J.kind == RJ_U64) {
            reduce_u64_body(static_cast<const unsigned long long*>(src), n, J.width,
                            static_cast<unsigned long long*>(dst), l1 ? x.k : 0u, nblk);
        } else {
            reduce_nd_body(static_cast<const NdPart*>(src), n, static_cast<NdPart*>(dst), l1 ? x.k : 0u, nblk);
        }
        __syncthreads();
    }
}
#endif

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(ND_WG) void k_ndjson_batch(const BatchHdr* H) {
    __shared__ uint32_t s_p, s_t0, s_t1;
    const SegT& T = H->tab[T_NDJ];
    const uint32_t total = T.total;
    for (uint32_t u = blockIdx.x; u < total; u += gridDim.x) {
        if (threadIdx.x < 64) {
            const uint32_t i = seg_find(T, u);
            if (threadIdx.x == 0) {
                const uint32_t pi = T.req[i], c = T.chunk[i];
                const uint32_t t0 = (u - T.start[i]) * c;
                s_p = pi;
                s_t0 = t0;
                s_t1 = min(t0 + c, H->nd[pi].n_tiles);
            }
        }
        __syncthreads();
        const uint32_t pi = uni(s_p), t0 = uni(s_t0), t1 = uni(s_t1);
        __syncthreads();
        const BNd& A = H->nd[pi];
        for (uint32_t t = t0; t < t1; t += ND_WAVES) {   // t0 and the chunk are multiples of ND_WAVES
            ndjson_block(A.data, A.len, A.n_tiles, A.parts, A.far, t / ND_WAVES);
            __syncthreads();
        }
    }
}

// The far records of the NDJSON documents, over the same units (a lane per
// tile of the unit).
__global__ __launch_bounds__(64) void k_ndjson_far_batch(const BatchHdr* H) {
    __shared__ SlowF64 slow;
    const SegT& T = H->tab[T_NDJ];
    if (threadIdx.x == 0) slow.lock = 0;
    __syncthreads();
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const uint32_t i = seg_find(T, u);
        const uint32_t pi = T.req[i], c = T.chunk[i];
        const BNd& A = H->nd[pi];
        const uint32_t t0 = (u - T.start[i]) * c, t1 = min(t0 + c, A.n_tiles);
        for (uint32_t t = t0; t < t1; t += 64) {
            const uint32_t tt = t + threadIdx.x;
            ndjson_far_tile(A.data, A.far, A.parts + A.n_tiles, tt, tt < t1, slow);
        }
    }
}
#endif

#ifndef FX_KIND   // defined once, in the host translation unit
__global__ __launch_bounds__(WG) void k_fill_batch(const BatchHdr* H) {
    const SegT& T = H->tab[T_FILL];
    for (uint32_t u = blockIdx.x; u < T.total; u += gridDim.x) {
        const Unit x = unit_at(T, u);
        const BFill& F = H->fill[x.req];
        const uint64_t lo = (uint64_t)x.k * FILL_UNIT, hi = min(F.n, lo + FILL_UNIT);
        for (uint64_t i = lo + threadIdx.x; i < hi; i += WG) F.p[i] = F.v;
    }
}
#endif

// --------------------------------------------------------------------------
// tile-kernel launches, one translation unit per dialect kind (kinds.hip)
// --------------------------------------------------------------------------
// Dialect kinds: 0 no quote, 1 comment only, 2 quote, 3 quote + backslash
// escape, 4 quote + comment, 5 quote + backslash escape + comment.
enum Form : int {
    F_PLAIN = 0,   // k_scan, a workgroup per tile
    F_REDO = 1,    // k_redo over the tiles a fast pass flagged
    F_BATCH = 2    // k_tiles_batch over table `tab` of the batch H
};

// Launches the tile kernel of dialect kind K for (form, mode); defined in
// kinds.hip, which is compiled once per kind. hipErrorInvalidValue for a
// (form, mode) that is never launched for this kind.
template <int K>
hipError_t launch_kind(int form, int mode, dim3 g, hipStream_t s, const Params* P, const BatchHdr* H, uint32_t tab);
template <>
hipError_t launch_kind<0>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);
template <>
hipError_t launch_kind<1>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);
template <>
hipError_t launch_kind<2>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);
template <>
hipError_t launch_kind<3>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);
template <>
hipError_t launch_kind<4>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);
template <>
hipError_t launch_kind<5>(int, int, dim3, hipStream_t, const Params*, const BatchHdr*, uint32_t);

inline hipError_t launch_tile_kernel(int kind, int form, int mode, dim3 g, hipStream_t s, const Params* P,
                                     const BatchHdr* H, uint32_t tab) {
    switch (kind) {
    case 0: return launch_kind<0>(form, mode, g, s, P, H, tab);
    case 1: return launch_kind<1>(form, mode, g, s, P, H, tab);
    case 2: return launch_kind<2>(form, mode, g, s, P, H, tab);
    case 3: return launch_kind<3>(form, mode, g, s, P, H, tab);
    case 4: return launch_kind<4>(form, mode, g, s, P, H, tab);
    default: return launch_kind<5>(form, mode, g, s, P, H, tab);
    }
}

}  // namespace fx

#include "swar_tile.h"
