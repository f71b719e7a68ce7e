// swar_tile.h — the dedicated tile kernels of the dialects with neither an
// escape nor a comment byte (quote-only dialects such as CSV, and dialects
// with no quote such as TSV): the tile summary pass and the numeric fast
// column passes. Included by scan.hip after the generic tile pass.
//
// Both classify a thread's 32 bytes with the SWAR front (classify32) and
// keep only the state their pass needs, so they hold less LDS and fewer
// registers than the generic pass and run more workgroups per CU:
//  - k_swar_summary writes the same tile summary as the generic M_SUMMARY
//    pass (the transfer function, the four waves' functions and the counters
//    for every entry state) without copying the tile into LDS;
//  - k_numfast is the fast form of the columns and statistics passes when
//    every column of the pass is i64 or f64. The tile's entry state and
//    counters come from the tile prefix by scalar loads (no exchange between
//    waves), the structurals and terminators are listed in one compaction,
//    and record r of the tile goes to thread r, which reads its cells by
//    index and decides them in registers (fixed6_win, then num_swar on the
//    12 bytes ending at the cell). A tile holding a cell those leave
//    undecided, a cell outside its LDS copy, or more structurals or records
//    than the lists hold is flagged in Params::fixup and redone by the full
//    pass; its results are then the full pass's.
#pragma once

namespace fx {

typedef __attribute__((address_space(4))) const uint32_t ConstU32;

// A buffer no thread of the launch writes, read through the constant address
// space: the compiler then loads wave-uniform words with scalar loads.
DEV ConstU32* as_const(const uint32_t* p) { return reinterpret_cast<ConstU32*>(reinterpret_cast<uintptr_t>(p)); }

DEV Counts c7_load(ConstU32* p) {
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

// Word compares apply when the quote byte and the separator bytes are below
// 0x80 and there are at most four separators; otherwise the byte table.
template <bool HQ>
DEV bool dialect_ascii(const Params& P) {
    const uint32_t qa = HQ ? (uint32_t)P.quote : 0u;
    return qa < 0x80u && P.n_sepb <= 4u && (P.n_sepb < 1 || P.sepb[0] < 0x80u) && (P.n_sepb < 2 || P.sepb[1] < 0x80u) &&
           (P.n_sepb < 3 || P.sepb[2] < 0x80u) && (P.n_sepb < 4 || P.sepb[3] < 0x80u);
}

// The thread's 32 tile bytes [t0, t0 + 32) in words (zero past the tile).
DEV void load_run(const Params& P, uint32_t tile_start, uint32_t tile_len, uint32_t t0, uint32_t (&x)[8]) {
    if (t0 + 32u <= tile_len) {
        const uint4 a = *reinterpret_cast<const uint4*>(P.data + tile_start + t0);
        const uint4 b = *reinterpret_cast<const uint4*>(P.data + tile_start + t0 + 16u);
        x[0] = a.x; x[1] = a.y; x[2] = a.z; x[3] = a.w;
        x[4] = b.x; x[5] = b.y; x[6] = b.z; x[7] = b.w;
        return;
    }
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

// Quote, candidate and newline masks of the run, zero past the tile.
template <bool HQ>
DEV Masks32 classify_run(const uint32_t (&x)[8], const Params& P, const uint8_t* cls, bool ascii, uint32_t live32) {
    Masks32 m{0u, 0u, 0u};
    switch (ascii ? P.n_sepb : 5u) {
    case 0: m = classify32<HQ, 0>(x, P); break;
    case 1: m = classify32<HQ, 1>(x, P); break;
    case 2: m = classify32<HQ, 2>(x, P); break;
    case 3: m = classify32<HQ, 3>(x, P); break;
    case 4: m = classify32<HQ, 4>(x, P); break;
    default: m = classify32_table<HQ>(x, P, cls); break;
    }
    m.q &= live32;
    m.c &= live32;
    m.n &= live32;
    return m;
}

DEV uint32_t live_run(uint32_t t0, uint32_t tile_len) {
    return t0 + 32u <= tile_len ? ~0u : (t0 < tile_len ? (1u << (tile_len - t0)) - 1u : 0u);
}

// ---------------------------------------------------------------------------
// Tile summaries
// ---------------------------------------------------------------------------
template <bool HQ>
__global__ __launch_bounds__(WG) void k_swar_summary(Params P) {
    constexpr int NS = HQ ? 2 : 1;
    __shared__ Counts wsum[NWAVE * NS];
    __shared__ uint32_t wfn[NWAVE];
    __shared__ uint8_t cls[256];
    const uint32_t tile = blockIdx.x, tid = threadIdx.x, ln = tid & 63u, w = uni(tid >> 6);
    const uint32_t tile_start = tile * (uint32_t)TILE, tile_len = min((uint32_t)TILE, P.len - tile_start);
    const bool ascii = dialect_ascii<HQ>(P);
    if (!ascii) {
        cls[tid] = (uint8_t)((P.cand[tid >> 5] >> (tid & 31)) & 1u);
        __syncthreads();
    }
    const uint32_t t0 = tid * 32u;
    uint32_t x[8];
    load_run(P, tile_start, tile_len, t0, x);
    const Masks32 m = classify_run<HQ>(x, P, cls, ascii, live_run(t0, tile_len));
    // Quote parity from the thread's first byte; the quotes of the wave's
    // earlier threads by mbcnt over the odd ones.
    const uint32_t pref = HQ ? prefix_xor32(m.q) : 0u;
    const uint64_t odd = ballot(HQ && (__popc(m.q) & 1));
    const uint32_t ex = mbcnt(odd) & 1u;
    if constexpr (HQ) {
        Counts c0, c1;
        wave_counts_pair(m.c, m.c & m.n, pref ^ (ex ? ~0u : 0u), tile_start + t0, c0, c1);
        if (ln == 0) {
            wsum[w * NS + 0] = c0;
            wsum[w * NS + 1] = c1;
        }
    } else {
        const Counts c = wave_counts((uint64_t)m.c, (uint64_t)(m.c & m.n), 0ULL, tile_start + t0);
        if (ln == 0) wsum[w * NS] = c;
    }
    if (ln == 0) wfn[w] = NS == 2 ? ((popc(odd) & 1u) ? 0x01u : 0x10u) : 0u;
    __syncthreads();
    uint32_t* out = P.tsum + (size_t)tile * P.ts_stride;
    if (tid < (uint32_t)NS) {
        uint32_t s = tid;
        Counts c = counts_zero();
        for (int k = 0; k < NWAVE; ++k) {
            c = compose(c, wsum[k * NS + s]);
            s = fn_apply(wfn[k], s);
        }
        c7_store(out + TS_COUNTS + C7 * tid, c);
    }
    if (tid == 0) {
        uint32_t f = fn_ident<NS>();
        for (int k = 0; k < NWAVE; ++k) {
            f = fn_then<NS>(f, wfn[k]);
            out[1 + k] = wfn[k];
        }
        out[0] = f;
    }
}

// ---------------------------------------------------------------------------
// Numeric fast column passes
// ---------------------------------------------------------------------------
constexpr uint32_t NF_LIST = 4096;   // structurals a numeric fast tile lists (more: the full pass)

struct NumTile {
    uint8_t bytes[TILE_LDS];   // document bytes from tile start - PRE
    uint16_t list[NF_LIST];    // tile-local position | terminator bit << 14, in byte order
    uint16_t tl[TL_CAP];       // list index of the tile's r-th terminator
};

// After the record walk the tile's bytes and lists are no longer read and
// the same LDS holds publish_stats's exchange.
struct NumShared {
    union {
        NumTile t;
        unsigned long long xs[(NWAVE - 1) * (4 * STAT_GROUP + 1) * 64];
    };
    uint32_t wcount[NWAVE];
    uint32_t fix_bad;
    uint8_t cls[256];
};

// Held to 8 waves/SIMD: the loop-invariant state of the record walk then
// spills a few words to scratch, which costs less than the latency the
// two extra waves hide (about 7% on the 1 GiB geo columns).
template <bool HQ, int KMODE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_numfast(Params P) {
    static_assert(KMODE == M_NUM_FAST || KMODE == M_NUM_FAST_STATS || KMODE == M_NUM_STATS_FAST, "numeric fast forms");
    constexpr bool IS_COLS = KMODE != M_NUM_STATS_FAST;
    constexpr bool STATS = KMODE != M_NUM_FAST;
    constexpr uint32_t NACC = STAT_GROUP;
    __shared__ NumShared sh;
    const uint32_t tile = blockIdx.x, tid = threadIdx.x, ln = tid & 63u, w = uni(tid >> 6);
    const uint32_t tile_start = tile * (uint32_t)TILE, tile_len = min((uint32_t)TILE, P.len - tile_start);
    const bool last_tile = tile + 1 == P.n_tiles;

    // The tile's entry state and exclusive counters (tile prefix), its own
    // counters and the waves' functions (tile summary): scalar loads, every
    // wave its own copy.
    ConstU32* const tp = as_const(P.tpre) + (size_t)tile * 8;
    ConstU32* const ts = as_const(P.tsum) + (size_t)tile * P.ts_stride;
    const uint32_t s0 = tp[0];
    const Counts pre = c7_load(tp + 1);
    const Counts inc = compose(pre, c7_load(ts + TS_COUNTS + C7 * s0));
    uint32_t went = s0;
    for (uint32_t k = 0; k < w; ++k) went = fn_apply(ts[1 + k], went);

    if (tid == 0) sh.fix_bad = 0u;
    const bool ascii = dialect_ascii<HQ>(P);
    if (!ascii) {
        sh.cls[tid] = (uint8_t)((P.cand[tid >> 5] >> (tid & 31)) & 1u);
        __syncthreads();
    }
    // The tile and PRE / HALO bytes around it into LDS (zero outside the
    // document); LDS byte i is document byte tile_start - PRE + i.
    const uint32_t t0 = tid * 32u;
    uint32_t x[8];
    load_run(P, tile_start, tile_len, t0, x);
    uint8_t* const lds = sh.t.bytes;
    reinterpret_cast<uint4*>(lds + PRE + t0)[0] = make_uint4(x[0], x[1], x[2], x[3]);
    reinterpret_cast<uint4*>(lds + PRE + t0)[1] = make_uint4(x[4], x[5], x[6], x[7]);
    const uint32_t lds_lo = tile_start - min((uint32_t)PRE, tile_start);
    const uint32_t lds_hi = tile_start + min((uint32_t)(TILE + HALO), P.len - tile_start);
    if (tid < (uint32_t)((PRE + HALO) / 16)) {
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
        *reinterpret_cast<uint4*>(lds + i) = v;
    }

    // Structurals and terminators of the run from its entry state.
    const Masks32 m = classify_run<HQ>(x, P, sh.cls, ascii, live_run(t0, tile_len));
    const uint32_t pref = HQ ? prefix_xor32(m.q) : 0u;
    const uint64_t odd = ballot(HQ && (__popc(m.q) & 1));
    const uint32_t entry = HQ ? (went ^ (mbcnt(odd) & 1u)) : 0u;
    const uint32_t inside = pref ^ (entry ? ~0u : 0u);
    const uint32_t S = HQ ? (m.c & ~inside) : m.c, T = S & m.n;
    // Counted in the two halves of one word (a wave holds at most 2048 of each).
    const uint32_t cnt = (uint32_t)__popc(S) | ((uint32_t)__popc(T) << 16);
    const uint32_t incl = wave_incl_sum(cnt);
    if (ln == 63) sh.wcount[w] = incl;
    __syncthreads();   // the LDS copy and the waves' counts
    uint32_t wb = 0, tot = 0;
#pragma unroll
    for (uint32_t k = 0; k < (uint32_t)NWAVE; ++k) {
        const uint32_t c = sh.wcount[k];
        if (k < w) wb += c;
        tot += c;
    }
    tot = uni(tot);
    const uint32_t n_list = tot & 0xFFFFu, nt = tot >> 16;
    const bool fits = n_list <= NF_LIST && nt <= (uint32_t)TL_CAP;
    if (fits) {
        const uint32_t at = uni(wb) + incl - cnt;
        uint32_t idx = at & 0xFFFFu, rank = at >> 16;
        for (uint32_t rem = S; rem; rem &= rem - 1u) {
            const uint32_t bit = (uint32_t)__ffs((int)rem) - 1u;
            const uint32_t t = (T >> bit) & 1u;
            sh.t.list[idx] = (uint16_t)((t0 + bit) | (t << 14));
            if (t) sh.t.tl[rank] = (uint16_t)idx;
            rank += t;
            idx += 1u;
        }
    } else if (tid == 0) {
        sh.fix_bad = 1u;
    }
    __syncthreads();   // the lists

    // ---- records: thread r takes record r ---------------------------------
    const bool trailer = last_tile && inc.last_term < P.len;
    const uint32_t trailer_word = trailer ? ((P.len - tile_start) | 0x4000u) : 0u;
    const auto word_at = [&](uint32_t k) -> uint32_t { return k < n_list ? (uint32_t)sh.t.list[k] : trailer_word; };
    const auto pos_of = [&](uint32_t wd) -> uint32_t { return tile_start + (wd & 0x3FFFu); };
    const uint32_t c0 = IS_COLS ? 0u : P.col_group;
    const uint32_t nc = IS_COLS ? P.n_cols : min((uint32_t)STAT_GROUP, P.n_cols - c0);
    const int32_t quote = HQ ? P.quote : -1;
    const bool swar_ok = quote < '0' || quote > '9';
    const uint64_t rows_cap = P.rows_cap;
    Acc acc[NACC];
#pragma unroll
    for (uint32_t j = 0; j < NACC; ++j) acc[j] = Acc{0, 0, 0, 0, 0, 0.0};
    unsigned long long records = 0;
    if (fits) {
        const uint32_t e_after = nt ? uni((uint32_t)sh.t.tl[nt - 1]) + 1u : 0u;
        const uint32_t n_rec = nt + ((trailer || e_after < n_list) ? 1u : 0u);
        for (uint32_t r0 = 0; r0 < n_rec; r0 += WG) {
            // Record r ends at terminator r, or is the trailer's or the open
            // one after the tile's last terminator; its fields that end here
            // are entries e0 .. e0 + n - 1, the first of ordinal ord0. The
            // steps below are straight-line (selects, no branches) except
            // for rare cells, so the lanes of a wave stay together.
            // A wave whose lanes all lie past the tile's records skips the
            // round (the steps below are straight-line and would otherwise
            // be issued for no lane).
            if (r0 + 64u * w >= n_rec) continue;
            const uint32_t r = r0 + tid;
            const bool has = r < n_rec;
            const uint32_t rr = has ? r : 0u;
            const uint32_t e0 = rr == 0 ? 0u : (uint32_t)sh.t.tl[rr - 1] + 1u;
            const bool real = rr < nt;
            const uint32_t e1 = real ? (uint32_t)sh.t.tl[rr] : n_list;   // the terminator entry (n_list: the trailer's)
            const bool term = has && (real || trailer);
            const uint32_t n = has ? e1 - e0 + ((real || trailer) ? 1u : 0u) : 0u;
            const uint32_t fs0 = e0 == 0 ? pre.last_struct : pos_of(sh.t.list[e0 - 1]) + 1u;
            const uint32_t rs = e0 == 0 ? pre.last_term : fs0;
            const uint32_t ord0 = e0 == 0 ? pre.ord : 0u;
            records += (has && term) ? 1u : 0u;
            const uint32_t row = pre.n_rows + r;
            const bool live = row < rows_cap;
            const uint32_t rfirst = pre.n_rows + r0 + 64u * w;   // the wave's first row
            for (uint32_t jj = 0; jj < nc; ++jj) {
                const ColDesc& c = P.cols[c0 + jj];
                const int32_t i = (int32_t)c.index - (int32_t)ord0;
                const bool here = has && i >= 0 && (uint32_t)i < n;
                const bool missing = has && term && i >= (int32_t)n;
                const bool as_f64 = c.type == FALX_COL_F64;
                // The cell [fs, fe): the record's last field without its
                // trailing '\r' (csv_geo.rs:1056-1060).
                const uint32_t k = here ? e0 + (uint32_t)i : 0u;
                uint32_t fe = pos_of(word_at(k));
                const uint32_t fs = i == 0 ? fs0 : pos_of(word_at(k - (k ? 1u : 0u))) + 1u;
                const bool trim = here && term && (uint32_t)i + 1u == n && fe > rs;
                const bool cr_out = trim && fe - 1u < lds_lo;   // the byte before lies outside the LDS copy
                const uint32_t crb = lds[(trim && !cr_out) ? fe - 1u - tile_start + PRE : 0u];
                fe -= (trim && !cr_out && crb == '\r') ? 1u : 0u;
                bool flag = here && (cr_out || fe < fs || fs < lds_lo || fe > lds_hi);
                const bool go = here && !flag;
                uint32_t lf = go ? fs - tile_start + PRE : PRE, lt = go ? fe - tile_start + PRE : PRE;
                Win<3> W = win_lds<3>(lds, lt);
                // A cell quoted at both ends is read inside the quotes.
                if (HQ && go && lt - lf >= 2u && (W.w[2] >> 24) == (uint32_t)quote && lds[lf] == (uint8_t)quote) {
                    lf += 1;
                    lt -= 1;
                    W = win_lds<3>(lds, lt);
                }
                const uint32_t cn = lt - lf;
                NumVal nv{};
                bool done = as_f64 && fixed6_win(W.w[0], W.w[1], W.w[2], cn, nv.f);
                if (go && !done) {
                    if (swar_ok && cn <= Win12::BYTES) {
                        auto at = [&](uint32_t p) { return (uint32_t)lds[lt - Win12::BYTES + p]; };
                        done = num_swar(W, cn, as_f64, quote, at, nv.f, nv.iok, nv.iv) &&
                               !(as_f64 && nv.f.ok && nv.f.deferred);
                    }
                    flag = !done;
                }
                if (flag) sh.fix_bad = 1u;
                const bool vb = go && done && (as_f64 ? nv.f.ok : nv.iok);
                const double fval = vb ? nv.f.value : 0.0;
                const long long ival = vb ? nv.iv : 0;
                if (IS_COLS && (here || missing) && live) {
                    if (as_f64) reinterpret_cast<double*>(c.values)[row] = fval;
                    else reinterpret_cast<long long*>(c.values)[row] = ival;
                }
                if (STATS && jj < NACC && (IS_COLS ? live : true)) {
                    // The cell's share of the column statistics (adding +0.0
                    // for an invalid cell leaves the FP64 sum's bits as they
                    // are: it starts at +0.0 and cannot become -0.0).
                    const bool fx = vb && as_f64 && nv.f.fixed6;
                    const unsigned long long bits = as_f64 ? f64_to_bits(fval) : (unsigned long long)ival;
                    const double dv = as_f64 ? fval : (double)ival;
#pragma unroll
                    for (uint32_t j = 0; j < NACC; ++j) {
                        if (j != jj) continue;
                        acc[j].values += vb ? 1u : 0u;
                        acc[j].csum += vb ? bits : 0ULL;
                        acc[j].dsum += vb ? dv : 0.0;
                        acc[j].fx_sum += fx ? nv.f.mantissa : 0LL;
                        acc[j].fx_cells += fx ? 1u : 0u;
                    }
                }
                if (IS_COLS) {
                    // Rows are consecutive across the lanes holding a record
                    // (the lowest ones): one ballot, at most two words.
                    const uint64_t Wv = ballot(vb && live);
                    if (Wv && ln == 0) {
                        const uint32_t wi = rfirst >> 6, sb = rfirst & 63u;
                        atomicOr(&c.valid[wi], (unsigned long long)(Wv << sb));
                        if (sb && (Wv >> (64u - sb))) atomicOr(&c.valid[wi + 1], (unsigned long long)(Wv >> (64u - sb)));
                    }
                }
            }
        }
    }
    if (IS_COLS && tid == 0 && last_tile) P.result[R_N_ROWS] = inc.n_rows + (trailer ? 1 : 0);
    if (STATS) {
        const uint32_t cols[NACC] = {c0, c0 + 1u};
        publish_stats(P, tile, acc, cols, min(NACC, nc), records, c0 == 0u, sh.xs);
    }
    else __syncthreads();   // every thread's flag is in
    if (tid == 0) P.fixup[tile] = sh.fix_bad;
}

}  // namespace fx
