// oracle.cpp — CPU restatement of the falx reference path (structural index,
// tape, typed columns, fused sinks), written from the reference's Rust source
// text. TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke()
// and bench.py's cpu_baseline leg; never by the product library.
//
// Each function cites the reference code it restates. The serial paths are
// the semantic definition; the `threads > 1` paths restate the reference's
// parallel drivers (they are what the CPU baseline times).
#include "oracle.h"

#include <algorithm>
#include <array>
#include <clocale>
#include <cstdlib>
#include <cstring>
#include <locale.h>
#include <memory>
#include <thread>
#include <vector>

#if defined(__x86_64__)
#include <immintrin.h>
#endif

namespace {

constexpr uint64_t EVEN = 0x5555555555555555ULL;

// ---------------------------------------------------------------------------
// Dialect (formats.rs:30-53) decoded from the runtime spec.
// ---------------------------------------------------------------------------
struct Dialect {
    bool cand[256];      // structural candidate bytes (formats.rs:208-224)
    uint8_t cand_list[16];
    int n_cand = 0;
    bool nl_structural = false;
    int quote = -1, escape = -1, comment = -1;
};

Dialect make_dialect(const falx_spec* s) {
    Dialect d;
    std::memset(d.cand, 0, sizeof d.cand);
    bool has_sep = false;
    for (uint32_t i = 0; i < s->n_structural && i < 16; ++i) {
        uint8_t b = s->structural[i];
        if (b == '\n') d.nl_structural = true; else has_sep = true;
    }
    // candidates = seps | terminators (if '\n' structural); no seps -> '\n'.
    for (uint32_t i = 0; i < s->n_structural && i < 16; ++i) {
        uint8_t b = s->structural[i];
        if (b != '\n') d.cand[b] = true;
    }
    if (d.nl_structural || !has_sep) d.cand['\n'] = true;
    for (int b = 0; b < 256; ++b)
        if (d.cand[b] && d.n_cand < 16) d.cand_list[d.n_cand++] = (uint8_t)b;
    d.quote = s->quote;
    d.escape = s->quote >= 0 ? s->escape : -1;
    d.comment = s->comment;
    return d;
}

// ---------------------------------------------------------------------------
// 64-byte block classification (the eq_mask/class nodes of every `step`).
// ---------------------------------------------------------------------------
struct Masks {
    uint64_t cand, nl, q, esc, cm;
};

void classify_scalar(const uint8_t* p, const Dialect& d, Masks& m) {
    m = Masks{0, 0, 0, 0, 0};
    for (int i = 0; i < 64; ++i) {
        uint8_t b = p[i];
        uint64_t bit = 1ULL << i;
        if (d.cand[b]) m.cand |= bit;
        if (b == '\n') m.nl |= bit;
        if ((int)b == d.quote) m.q |= bit;
        if ((int)b == d.escape) m.esc |= bit;
        if ((int)b == d.comment) m.cm |= bit;
    }
}

#if defined(__x86_64__)
__attribute__((target("avx512f,avx512bw")))
void classify_avx512(const uint8_t* p, const Dialect& d, Masks& m) {
    __m512i v = _mm512_loadu_si512((const void*)p);
    uint64_t cand = 0;
    for (int i = 0; i < d.n_cand; ++i)
        cand |= _mm512_cmpeq_epi8_mask(v, _mm512_set1_epi8((char)d.cand_list[i]));
    m.cand = cand;
    m.nl = _mm512_cmpeq_epi8_mask(v, _mm512_set1_epi8('\n'));
    m.q = d.quote >= 0 ? _mm512_cmpeq_epi8_mask(v, _mm512_set1_epi8((char)d.quote)) : 0;
    m.esc = d.escape >= 0 ? _mm512_cmpeq_epi8_mask(v, _mm512_set1_epi8((char)d.escape)) : 0;
    m.cm = d.comment >= 0 ? _mm512_cmpeq_epi8_mask(v, _mm512_set1_epi8((char)d.comment)) : 0;
}

__attribute__((target("pclmul,sse2")))
uint64_t prefix_xor_clmul(uint64_t x) {
    __m128i a = _mm_set_epi64x(0, (long long)x);
    __m128i ones = _mm_set1_epi8((char)0xFF);
    return (uint64_t)_mm_cvtsi128_si64(_mm_clmulepi64_si128(a, ones, 0));
}
#endif

// interp.rs:143-151.
inline uint64_t prefix_xor_shift(uint64_t x) {
    x ^= x << 1;
    x ^= x << 2;
    x ^= x << 4;
    x ^= x << 8;
    x ^= x << 16;
    x ^= x << 32;
    return x;
}

bool g_avx512 = false;
bool g_pclmul = false;
struct IsaInit {
    IsaInit() {
#if defined(__x86_64__)
        __builtin_cpu_init();
        g_avx512 = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw");
        g_pclmul = __builtin_cpu_supports("pclmul");
#endif
    }
} g_isa_init;

inline void classify(const uint8_t* p, const Dialect& d, Masks& m) {
#if defined(__x86_64__)
    if (g_avx512) { classify_avx512(p, d, m); return; }
#endif
    classify_scalar(p, d, m);
}

inline uint64_t prefix_xor(uint64_t x) {
#if defined(__x86_64__)
    if (g_pclmul) return prefix_xor_clmul(x);
#endif
    return prefix_xor_shift(x);
}

// Bits [from, to) set (interp.rs:205-209).
inline uint64_t range_mask(uint32_t from, uint32_t to) {
    uint64_t hi = to >= 64 ? ~0ULL : ((1ULL << to) - 1);
    return hi & ~((1ULL << from) - 1);
}

// Three-state region resolver with the generated kernels' two fast paths
// (csv_hash.rs:2149-2220; slow walk = interp.rs:163-203).
uint64_t resolve_regions(uint64_t q, uint64_t s, uint64_t n, uint64_t& state) {
    const uint64_t NORMAL = 0, QUOTE = 1, COMMENT = 2;
    if (state == NORMAL && (q | s) == 0) return 0;
    if (state != COMMENT && s == 0) {
        uint64_t inert = prefix_xor(q);
        if (state == QUOTE) inert = ~inert;
        state = inert >> 63;
        return inert;
    }
    uint64_t inert = 0;
    uint32_t run_start = 0;
    uint64_t events = q | s | n;
    while (events) {
        uint32_t p = (uint32_t)__builtin_ctzll(events);
        uint64_t bit = 1ULL << p;
        if (state == QUOTE) {
            if (q & bit) { inert |= range_mask(run_start, p); state = NORMAL; }
        } else if (state == COMMENT) {
            if (n & bit) { inert |= range_mask(run_start, p); state = NORMAL; }
        } else {
            if (q & bit) { state = QUOTE; run_start = p; }
            else if (s & bit) { state = COMMENT; run_start = p; }
        }
        events &= events - 1;
    }
    if (state != NORMAL) inert |= range_mask(run_start, 64);
    return inert;
}

// Carried state across blocks. CARRY_INIT per dialect: parity [0]
// (csv_geo.rs:21), regions [line_start=1, region=0] (csv_hash.rs:21),
// escapes [add=0, parity=0] (ndjson.rs:21).
struct Carry {
    uint64_t parity = 0;   // 0 or all-ones
    uint64_t add = 0;      // escape Add carry (0/1)
    uint64_t ls = 1;       // line-start carry (ShiftLeft1Seeded)
    uint64_t region = 0;   // 0 normal, 1 quote, 2 comment
};

// One block `step`: formats.rs:206-297 graph semantics.
inline void step(const Dialect& d, const Masks& m, Carry& c, uint64_t& S, uint64_t& T) {
    uint64_t rq = 0;
    if (d.quote >= 0) {
        rq = m.q;
        if (d.escape >= 0) {
            // escaped_positions (formats.rs:334-343): quote is real where
            // EVEN ^ (EVEN + (EVEN ^ esc) + carry) is set (ndjson.rs:1020-1026).
            uint64_t phase = EVEN ^ m.esc;
            uint64_t partial = EVEN + phase;
            uint64_t c1 = partial < EVEN;
            uint64_t sum = partial + c.add;
            uint64_t c2 = sum < partial;
            c.add = c1 | c2;
            rq &= EVEN ^ sum;
        }
    }
    uint64_t out;
    if (d.quote < 0) {
        out = m.cand;  // comment-without-quote: no kernel masking (formats.rs:256-265)
    } else if (d.comment < 0) {
        uint64_t inside = prefix_xor(rq) ^ c.parity;
        c.parity = (uint64_t)((int64_t)inside >> 63);
        out = m.cand & ~inside;
    } else {
        uint64_t ls = (m.nl << 1) | c.ls;
        c.ls = m.nl >> 63;
        uint64_t starts = m.cm & ls;
        uint64_t inert = resolve_regions(rq, starts, m.nl, c.region);
        out = m.cand & ~inert;
    }
    S = out;
    T = out & m.nl;
}

// Walk blocks of data[start..len) (start 64-aligned), tail zero-padded and
// masked (interp.rs:225-253). `f(S, T, base)` returns false to stop.
template <class F>
void run_blocks(const Dialect& d, const uint8_t* data, uint64_t len, uint64_t start,
                uint64_t stop, Carry& c, F&& f) {
    uint64_t off = start;
    Masks m;
    uint64_t S, T;
    while (off + 64 <= len && off < stop) {
        classify(data + off, d, m);
        step(d, m, c, S, T);
        if (!f(S, T, (uint32_t)off)) return;
        off += 64;
    }
    if (off < len && off < stop) {
        uint8_t blk[64] = {0};
        uint64_t rem = len - off;
        std::memcpy(blk, data + off, rem);
        classify(blk, d, m);
        step(d, m, c, S, T);
        uint64_t live = (1ULL << rem) - 1;
        f(S & live, T & live, (uint32_t)off);
    }
}

int clamp_threads(int threads, uint64_t len) {
    int64_t t = threads < 1 ? 1 : threads;
    int64_t cap = (int64_t)(len / 64 + 1);
    return (int)std::min<int64_t>(t, cap);
}

struct Chunks {
    int threads;
    std::vector<uint64_t> bounds;
};

// Chunking shared by every parallel driver: chunk = (len/threads + 63) & !63
// (csv_geo.rs:96-101).
bool make_chunks(uint64_t len, int threads, Chunks& ch) {
    threads = clamp_threads(threads, len);
    uint64_t chunk = (len / threads + 63) & ~63ULL;
    if (threads == 1 || chunk == 0) return false;
    ch.threads = threads;
    ch.bounds.resize(threads + 1);
    for (int t = 0; t <= threads; ++t)
        ch.bounds[t] = t == threads ? len : std::min<uint64_t>((uint64_t)t * chunk, len);
    return true;
}

enum class ParKind { Parity, Regions, Serial };

ParKind par_kind(const Dialect& d) {
    if (d.escape >= 0) return ParKind::Serial;
    if (d.quote >= 0 && d.comment >= 0) return ParKind::Regions;
    return ParKind::Parity;
}

// region_scan3 (csv_hash.rs:1920-1951): chunk transfer function of the
// region state, all three entry states advanced in one pass.
void region_scan3(const Dialect& d, const uint8_t* data, uint64_t from, uint64_t to,
                  uint64_t ls, uint64_t states[3]) {
    states[0] = 0; states[1] = 1; states[2] = 2;
    uint64_t lsc = ls;
    uint64_t off = from;
    Masks m;
    auto one = [&](const Masks& mm) {
        uint64_t l = (mm.nl << 1) | lsc;
        lsc = mm.nl >> 63;
        uint64_t s = mm.cm & l;
        for (int k = 0; k < 3; ++k) resolve_regions(mm.q, s, mm.nl, states[k]);
    };
    while (off + 64 <= to) { classify(data + off, d, m); one(m); off += 64; }
    if (off < to) {
        uint8_t blk[64] = {0};
        std::memcpy(blk, data + off, to - off);
        classify(blk, d, m);
        one(m);
    }
}

// Entry carries for every chunk of a parallel driver. Parity dialects use
// the quote-parity prepass (csv_geo.rs:1176-1205); region dialects the
// 3-phase transfer-function composition (csv_hash.rs:152-191).
std::vector<Carry> chunk_entry_carries(const Dialect& d, const uint8_t* data,
                                       const Chunks& ch, ParKind kind) {
    int T = ch.threads;
    std::vector<Carry> entry(T);
    if (kind == ParKind::Parity) {
        std::vector<uint64_t> parity(T, 0);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                Carry c;
                run_blocks(d, data + ch.bounds[t], ch.bounds[t + 1] - ch.bounds[t], 0,
                           ~0ULL, c, [](uint64_t, uint64_t, uint32_t) { return true; });
                parity[t] = c.parity & 1;
            });
        for (auto& x : th) x.join();
        uint64_t acc = 0;
        for (int t = 0; t < T; ++t) {
            entry[t].parity = (uint64_t)(-(int64_t)acc);
            acc ^= parity[t];
        }
    } else if (kind == ParKind::Regions) {
        std::vector<std::array<uint64_t, 3>> tr(T);
        auto line_start = [&](uint64_t b) -> uint64_t {
            return b == 0 ? 1 : (uint64_t)(data[b - 1] == '\n');
        };
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                uint64_t st[3];
                region_scan3(d, data, ch.bounds[t], ch.bounds[t + 1], line_start(ch.bounds[t]), st);
                tr[t] = {st[0], st[1], st[2]};
            });
        for (auto& x : th) x.join();
        uint64_t region = 0;
        for (int t = 0; t < T; ++t) {
            entry[t].region = region;
            entry[t].ls = line_start(ch.bounds[t]);
            region = tr[t][region];
        }
    }
    return entry;
}

// ---------------------------------------------------------------------------
// Cell parsers: codegen.rs INT_CELL_TPL (7500-7565), FLOAT_CELL_TPL
// (7566-7744), csv_geo.rs:1313-1486.
// ---------------------------------------------------------------------------
inline bool is_digit(uint8_t b) { return b >= '0' && b <= '9'; }

bool parse_i64_cell(const uint8_t* s, uint64_t n, int64_t& out) {
    bool neg = false;
    const uint8_t* dg = s;
    uint64_t nd = n;
    if (n > 0 && (s[0] == '-' || s[0] == '+')) { neg = s[0] == '-'; dg = s + 1; nd = n - 1; }
    if (nd == 0) return false;
    if (nd <= 16) {
        uint64_t v = 0;
        for (uint64_t i = 0; i < nd; ++i) {
            if (!is_digit(dg[i])) return false;
            v = v * 10 + (dg[i] - '0');
        }
        out = neg ? -(int64_t)v : (int64_t)v;
        return true;
    }
    // 17+ digits: accumulate negative with overflow checks so i64::MIN parses.
    int64_t v = 0;
    for (uint64_t i = 0; i < nd; ++i) {
        if (!is_digit(dg[i])) return false;
        int64_t t;
        if (__builtin_mul_overflow(v, (int64_t)10, &t)) return false;
        if (__builtin_sub_overflow(t, (int64_t)(dg[i] - '0'), &v)) return false;
    }
    if (neg) { out = v; return true; }
    if (v == INT64_MIN) return false;
    out = -v;
    return true;
}

// `[-]d{1,3}.d{6}` (csv_geo.rs:1414-1453): returns mantissa in 1e-6 units.
bool fixed6(const uint8_t* s, uint64_t n, bool& neg, uint64_t& m) {
    if (n < 8 || n > 11) return false;
    if (s[n - 7] != '.') return false;
    for (uint64_t i = n - 6; i < n; ++i) if (!is_digit(s[i])) return false;
    if (!is_digit(s[n - 8])) return false;
    uint64_t pre = n - 8;  // bytes before the last whole digit
    uint64_t i = 0;
    neg = false;
    if (pre > 0 && s[0] == '-') { neg = true; i = 1; }
    if (pre - i > 2) return false;
    uint64_t v = 0;
    for (uint64_t k = i; k < n - 7; ++k) {
        if (!is_digit(s[k])) return false;
        v = v * 10 + (s[k] - '0');
    }
    for (uint64_t k = n - 6; k < n; ++k) v = v * 10 + (s[k] - '0');
    m = v;
    return true;
}

locale_t c_locale() {
    static locale_t loc = newlocale(LC_NUMERIC_MASK, "C", (locale_t)0);
    return loc;
}

inline uint64_t f64_bits(double v) { uint64_t b; std::memcpy(&b, &v, 8); return b; }
inline double bits_f64(uint64_t b) { double v; std::memcpy(&v, &b, 8); return v; }

bool ieq(const uint8_t* s, uint64_t n, const char* lit) {
    uint64_t l = std::strlen(lit);
    if (n != l) return false;
    for (uint64_t i = 0; i < n; ++i)
        if ((s[i] | 0x20) != (uint8_t)lit[i]) return false;
    return true;
}

// Exactly `str::parse::<f64>`: the Rust grammar
// `[+-]?(inf|infinity|nan|(d+\.?d*|\.d+)([eE][+-]?d+)?)`, correctly rounded
// (ties to even). Fast paths as in the reference; the general conversion is
// glibc strtod on the grammar-validated text (correctly rounded).
bool parse_f64_cell(const uint8_t* s, uint64_t n, double& out, bool* is_fixed6 = nullptr,
                    int64_t* mant = nullptr) {
    bool neg;
    uint64_t m;
    if (fixed6(s, n, neg, m)) {
        double v = (double)m / 1000000.0;
        out = bits_f64(f64_bits(v) | ((uint64_t)neg << 63));
        if (is_fixed6) *is_fixed6 = true;
        if (mant) *mant = neg ? -(int64_t)m : (int64_t)m;
        return true;
    }
    if (is_fixed6) *is_fixed6 = false;
    uint64_t i = 0;
    neg = false;
    if (n > 0 && (s[0] == '-' || s[0] == '+')) { neg = s[0] == '-'; i = 1; }
    const uint8_t* r = s + i;
    uint64_t rn = n - i;
    if (ieq(r, rn, "inf") || ieq(r, rn, "infinity")) {
        out = bits_f64(0x7FF0000000000000ULL | ((uint64_t)neg << 63));
        return true;
    }
    if (ieq(r, rn, "nan")) {
        out = bits_f64(0x7FF8000000000000ULL | ((uint64_t)neg << 63));
        return true;
    }
    uint64_t k = 0, int_digits = 0, frac_digits = 0;
    while (k < rn && is_digit(r[k])) { ++k; ++int_digits; }
    if (k < rn && r[k] == '.') {
        ++k;
        while (k < rn && is_digit(r[k])) { ++k; ++frac_digits; }
    }
    if (int_digits + frac_digits == 0) return false;
    if (k < rn && (r[k] | 0x20) == 'e') {
        ++k;
        if (k < rn && (r[k] == '+' || r[k] == '-')) ++k;
        uint64_t ed = 0;
        while (k < rn && is_digit(r[k])) { ++k; ++ed; }
        if (ed == 0) return false;
    }
    if (k != rn) return false;
    std::vector<char> buf(n + 1);
    std::memcpy(buf.data(), s, n);
    buf[n] = 0;
    out = strtod_l(buf.data(), nullptr, c_locale());
    return true;
}

// ---------------------------------------------------------------------------
// Cleaning (codegen.rs:8008-8185; csv_geo_text.rs:1563-1593).
// ---------------------------------------------------------------------------
template <class Emit>
void clean_cell(const Dialect& d, const uint8_t* raw, uint64_t n, Emit&& emit) {
    if (d.quote < 0) {
        for (uint64_t i = 0; i < n; ++i) emit(raw[i]);
        return;
    }
    const uint8_t Q = (uint8_t)d.quote;
    if (d.escape < 0) {
        if (n >= 2 && raw[0] == Q && raw[n - 1] == Q) {
            const uint8_t* in = raw + 1;
            uint64_t m = n - 2, i = 0;
            while (i < m) {
                emit(in[i]);
                if (in[i] == Q && i + 1 < m && in[i + 1] == Q) i += 2; else i += 1;
            }
        } else {
            for (uint64_t i = 0; i < n; ++i) emit(raw[i]);
        }
        return;
    }
    const uint8_t E = (uint8_t)d.escape;
    const uint8_t* body = raw;
    uint64_t m = n;
    if (n >= 2 && raw[0] == Q && raw[n - 1] == Q) { body = raw + 1; m = n - 2; }
    uint64_t i = 0;
    while (i < m) {
        if (body[i] == E && i + 1 < m) { emit(body[i + 1]); i += 2; }
        else { emit(body[i]); i += 1; }
    }
}

// Numeric field helpers (codegen.rs:6060-6085): only a leading quote byte
// sends the cell through `clean`.
bool parse_f64_field(const Dialect& d, const uint8_t* raw, uint64_t n, double& v,
                     bool* fx = nullptr, int64_t* mant = nullptr) {
    if (d.quote >= 0 && n > 0 && raw[0] == (uint8_t)d.quote) {
        std::vector<uint8_t> c;
        c.reserve(n);
        clean_cell(d, raw, n, [&](uint8_t b) { c.push_back(b); });
        return parse_f64_cell(c.data(), c.size(), v, fx, mant);
    }
    return parse_f64_cell(raw, n, v, fx, mant);
}

bool parse_i64_field(const Dialect& d, const uint8_t* raw, uint64_t n, int64_t& v) {
    if (d.quote >= 0 && n > 0 && raw[0] == (uint8_t)d.quote) {
        std::vector<uint8_t> c;
        c.reserve(n);
        clean_cell(d, raw, n, [&](uint8_t b) { c.push_back(b); });
        return parse_i64_cell(c.data(), c.size(), v);
    }
    return parse_i64_cell(raw, n, v);
}

// field_byte_len (codegen.rs:6391-6425): doubled-quote collapse whenever the
// dialect has a quote byte.
uint64_t field_byte_len(const Dialect& d, const uint8_t* raw, uint64_t n) {
    if (d.quote < 0) return n;
    const uint8_t Q = (uint8_t)d.quote;
    if (n >= 2 && raw[0] == Q && raw[n - 1] == Q) {
        const uint8_t* in = raw + 1;
        uint64_t m = n - 2, i = 0, len = 0;
        while (i < m) {
            ++len;
            if (in[i] == Q && i + 1 < m && in[i + 1] == Q) i += 2; else i += 1;
        }
        return len;
    }
    return n;
}

// ---------------------------------------------------------------------------
// Record sinks. All share the reference's record-ownership contract
// (csv_geo.rs:952-1101): a sink over [start, end) starts emitting after the
// first terminator >= start (unless it is the first), and finishes the
// record it is in at the first terminator >= end.
// ---------------------------------------------------------------------------
struct RecordWalker {
    const uint8_t* data;
    uint32_t len;
    uint32_t field_start, record_start, ordinal = 0, end;
    bool emitting, done = false;
    RecordWalker(const uint8_t* d, uint32_t l, uint32_t start, uint32_t e, bool em)
        : data(d), len(l), field_start(start), record_start(start), end(e), emitting(em) {}
};

struct ColumnsHost {
    uint64_t rows = 0;
    struct Col {
        std::vector<int64_t> i64;
        std::vector<double> f64;
        std::vector<uint32_t> spans;
        std::vector<int32_t> offsets{0};
        std::vector<uint8_t> data;
        std::vector<uint64_t> valid;
    } cols[FALX_MAX_COLUMNS];
};

struct ColumnSink : RecordWalker {
    const Dialect& d;
    const falx_spec& spec;
    uint32_t pfrom[FALX_MAX_COLUMNS], pto[FALX_MAX_COLUMNS];
    uint64_t found = 0;
    ColumnsHost out;
    ColumnSink(const Dialect& dd, const falx_spec& sp, const uint8_t* data, uint32_t len,
               uint32_t start, uint32_t end, bool em)
        : RecordWalker(data, len, start, end, em), d(dd), spec(sp) {}

    void mark(uint32_t from, uint32_t to) {
        for (uint32_t k = 0; k < spec.n_columns; ++k)
            if (spec.columns[k].index == ordinal) { pfrom[k] = from; pto[k] = to; found |= 1ULL << k; }
    }
    bool on(uint32_t p, bool term) {
        if (term) {
            if (emitting) flush(p); else emitting = true;
            record_start = p + 1; field_start = p + 1; ordinal = 0; found = 0;
            if (p >= end) { done = true; return false; }
        } else if (emitting) {
            mark(field_start, p);
            ordinal += 1;
            field_start = p + 1;
        }
        return true;
    }
    void flush(uint32_t e) {
        if (d.comment >= 0 && e > record_start && data[record_start] == (uint8_t)d.comment) return;
        uint32_t to = (e > record_start && data[e - 1] == '\r') ? e - 1 : e;
        mark(field_start, to);
        uint64_t row = out.rows;
        for (uint32_t k = 0; k < spec.n_columns; ++k) {
            auto& c = out.cols[k];
            if ((row & 63) == 0) c.valid.push_back(0);
            bool f = (found >> k) & 1;
            const uint8_t* raw = data + pfrom[k];
            uint64_t n = f ? pto[k] - pfrom[k] : 0;
            bool ok = false;
            switch (spec.columns[k].type) {
            case FALX_COL_I64: {
                int64_t v = 0;
                ok = f && parse_i64_field(d, raw, n, v);
                c.i64.push_back(ok ? v : 0);
                break;
            }
            case FALX_COL_F64: {
                double v = 0;
                ok = f && parse_f64_field(d, raw, n, v);
                c.f64.push_back(ok ? v : 0.0);
                break;
            }
            case FALX_COL_STR: {
                ok = f;
                if (ok) clean_cell(d, raw, n, [&](uint8_t b) { c.data.push_back(b); });
                c.offsets.push_back((int32_t)c.data.size());
                break;
            }
            default: {  // BYTES: raw span, valid only when non-empty
                ok = f && pfrom[k] != pto[k];
                c.spans.push_back(ok ? pfrom[k] : 0);
                c.spans.push_back(ok ? pto[k] : 0);
                break;
            }
            }
            if (ok) c.valid[row >> 6] |= 1ULL << (row & 63);
        }
        out.rows = row + 1;
    }
    void finish() {
        if (emitting && !done && record_start < len) flush(len);
    }
};

// Fused column statistics (CsvGeoStatsSink csv_geo.rs:1602-1770 and
// CsvGeoTextStatsSink csv_geo_text.rs:1609-1913, generalised by spec).
struct StatsSink : RecordWalker {
    const Dialect& d;
    const falx_spec& spec;
    uint32_t pfrom[FALX_MAX_COLUMNS], pto[FALX_MAX_COLUMNS];
    uint64_t found = 0;
    falx_csv_stats st;
    StatsSink(const Dialect& dd, const falx_spec& sp, const uint8_t* data, uint32_t len,
              uint32_t start, uint32_t end, bool em)
        : RecordWalker(data, len, start, end, em), d(dd), spec(sp) {
        std::memset(&st, 0, sizeof st);
    }
    void mark(uint32_t from, uint32_t to) {
        for (uint32_t k = 0; k < spec.n_columns; ++k)
            if (spec.columns[k].index == ordinal) { pfrom[k] = from; pto[k] = to; found |= 1ULL << k; }
    }
    bool on(uint32_t p, bool term) {
        if (term) {
            if (emitting) flush(p); else emitting = true;
            record_start = p + 1; field_start = p + 1; ordinal = 0; found = 0;
            if (p >= end) { done = true; return false; }
        } else if (emitting) {
            mark(field_start, p);
            ordinal += 1;
            field_start = p + 1;
        }
        return true;
    }
    void flush(uint32_t e) {
        if (d.comment >= 0 && e > record_start && data[record_start] == (uint8_t)d.comment) return;
        uint32_t to = (e > record_start && data[e - 1] == '\r') ? e - 1 : e;
        mark(field_start, to);
        st.records += 1;
        for (uint32_t k = 0; k < spec.n_columns; ++k) {
            if (!((found >> k) & 1)) continue;
            auto& c = st.cols[k];
            const uint8_t* raw = data + pfrom[k];
            uint64_t n = pto[k] - pfrom[k];
            switch (spec.columns[k].type) {
            case FALX_COL_F64: {
                double v;
                bool fx = false;
                int64_t mant = 0;
                if (parse_f64_field(d, raw, n, v, &fx, &mant)) {
                    c.values += 1;
                    c.checksum += f64_bits(v);
                    c.sum += v;
                    if (fx) { c.fixed6_sum += mant; c.fixed6_cells += 1; }
                }
                break;
            }
            case FALX_COL_I64: {
                int64_t v;
                if (parse_i64_field(d, raw, n, v)) {
                    c.values += 1;
                    c.checksum += (uint64_t)v;
                    c.sum += (double)v;
                }
                break;
            }
            case FALX_COL_STR: {
                c.values += 1;
                clean_cell(d, raw, n, [&](uint8_t b) { c.bytes += 1; c.checksum += b; });
                break;
            }
            default: {
                if (n > 0) {
                    c.values += 1;
                    c.bytes += n;
                    for (uint64_t i = 0; i < n; ++i) c.checksum += raw[i];
                }
                break;
            }
            }
        }
    }
    void finish() {
        if (emitting && !done && record_start < len) flush(len);
    }
};

// FieldByteSink (csv_hash.rs:1123-1208).
struct FieldByteSink : RecordWalker {
    const Dialect& d;
    falx_field_byte_stats st{0, 0};
    FieldByteSink(const Dialect& dd, const uint8_t* data, uint32_t len, uint32_t start,
                  uint32_t end, bool em)
        : RecordWalker(data, len, start, end, em), d(dd) {}
    void flush_field(uint32_t e, bool record_end) {
        if (d.comment >= 0 && e > record_start && data[record_start] == (uint8_t)d.comment) return;
        uint32_t to = (record_end && e > record_start && data[e - 1] == '\r') ? e - 1 : e;
        st.fields += 1;
        st.bytes += field_byte_len(d, data + field_start, to - field_start);
    }
    bool on(uint32_t p, bool term) {
        if (term) {
            if (emitting) flush_field(p, true); else emitting = true;
            record_start = p + 1; field_start = p + 1;
            if (p >= end) { done = true; return false; }
        } else if (emitting) {
            flush_field(p, false);
            field_start = p + 1;
        }
        return true;
    }
    void finish() {
        if (emitting && !done && record_start < len) flush_field(len, true);
    }
};

template <class Sink>
void drive_sink(const Dialect& d, const uint8_t* data, uint64_t len, uint64_t start,
                Carry c, Sink& sink) {
    run_blocks(d, data, len, start, ~0ULL, c, [&](uint64_t S, uint64_t T, uint32_t base) {
        while (S) {
            uint32_t i = (uint32_t)__builtin_ctzll(S);
            if (!sink.on(base + i, (T >> i) & 1)) return false;
            S &= S - 1;
        }
        return true;
    });
}

// Serial or chunked (`*_par` / `*_chunks_par`) run of a record sink.
template <class Sink, class Make>
std::vector<Sink> run_sinks(const Dialect& d, const uint8_t* data, uint64_t len, int threads,
                            Make&& make) {
    std::vector<Sink> parts;
    Chunks ch;
    ParKind kind = par_kind(d);
    if (kind == ParKind::Serial || !make_chunks(len, threads, ch)) {
        parts.push_back(make(0u, (uint32_t)len, true));
        drive_sink(d, data, len, 0, Carry{}, parts.back());
        parts.back().finish();
        return parts;
    }
    std::vector<Carry> entry = chunk_entry_carries(d, data, ch, kind);
    for (int t = 0; t < ch.threads; ++t)
        parts.push_back(make((uint32_t)ch.bounds[t], (uint32_t)ch.bounds[t + 1], t == 0));
    std::vector<std::thread> th;
    for (int t = 0; t < ch.threads; ++t)
        th.emplace_back([&, t] {
            drive_sink(d, data, len, ch.bounds[t], entry[t], parts[t]);
            parts[t].finish();
        });
    for (auto& x : th) x.join();
    return parts;
}

// append_bitmap (csv_geo.rs:1289-1301).
void append_bitmap(std::vector<uint64_t>& dst, uint64_t dst_rows, const std::vector<uint64_t>& src,
                   uint64_t src_rows) {
    uint64_t words = (src_rows + 63) / 64;
    uint64_t shift = dst_rows & 63;
    if (shift == 0) {
        dst.insert(dst.end(), src.begin(), src.begin() + words);
    } else {
        for (uint64_t w = 0; w < words; ++w) {
            dst.back() |= src[w] << shift;
            dst.push_back(src[w] >> (64 - shift));
        }
        dst.resize((dst_rows + src_rows + 63) / 64);
    }
}

// ---------------------------------------------------------------------------
// NDJSON schema walker for `sum(id + nested.score)`: top-level integer `id`
// and `nested.score` number (JSON grammar), escapes respected, depth tracked.
// The reference's own walker (ndjson.rs:679-723) requires a literal
// `{"id":` prefix, an integer score and a fixed key order; this generalises
// it to any key order and JSON whitespace.
// ---------------------------------------------------------------------------
inline bool jws(uint8_t b) { return b == ' ' || b == '\t' || b == '\r' || b == '\n'; }

struct JCur {
    const uint8_t* p;
    uint64_t i, n;
    void ws() { while (i < n && jws(p[i])) ++i; }
};

bool j_skip_string(JCur& c) {  // at opening quote
    ++c.i;
    while (c.i < c.n) {
        uint8_t b = c.p[c.i];
        if (b == '\\') { c.i += 2; continue; }
        ++c.i;
        if (b == '"') return true;
    }
    return false;
}

bool j_skip_value(JCur& c) {
    c.ws();
    if (c.i >= c.n) return false;
    uint8_t b = c.p[c.i];
    if (b == '"') return j_skip_string(c);
    if (b == '{' || b == '[') {
        int depth = 0;
        while (c.i < c.n) {
            uint8_t x = c.p[c.i];
            if (x == '"') { if (!j_skip_string(c)) return false; continue; }
            if (x == '{' || x == '[') ++depth;
            else if (x == '}' || x == ']') { --depth; if (depth == 0) { ++c.i; return true; } }
            ++c.i;
        }
        return false;
    }
    uint64_t s = c.i;
    while (c.i < c.n) {
        uint8_t x = c.p[c.i];
        if (x == ',' || x == '}' || x == ']' || jws(x)) break;
        ++c.i;
    }
    return c.i > s;
}

// Key at a '"' position; returns raw key bytes span.
bool j_key(JCur& c, uint64_t& ks, uint64_t& ke) {
    if (c.i >= c.n || c.p[c.i] != '"') return false;
    ks = c.i + 1;
    if (!j_skip_string(c)) return false;
    ke = c.i - 1;
    return true;
}

bool key_is(const JCur& c, uint64_t ks, uint64_t ke, const char* lit) {
    uint64_t l = std::strlen(lit);
    return ke - ks == l && std::memcmp(c.p + ks, lit, l) == 0;
}

inline bool j_delim(const JCur& c) {
    return c.i >= c.n || c.p[c.i] == ',' || c.p[c.i] == '}' || c.p[c.i] == ']' || jws(c.p[c.i]);
}

bool j_int(JCur& c, int64_t& v) {
    bool neg = false;
    if (c.i < c.n && c.p[c.i] == '-') { neg = true; ++c.i; }
    if (c.i >= c.n || !is_digit(c.p[c.i])) return false;
    uint64_t acc = 0;
    while (c.i < c.n && is_digit(c.p[c.i])) { acc = acc * 10 + (c.p[c.i] - '0'); ++c.i; }
    if (!j_delim(c)) return false;
    v = neg ? (int64_t)(0 - acc) : (int64_t)acc;
    return true;
}

// JSON number -> f64 (correctly rounded) plus exact 1e-6 mantissa when the
// number has no exponent, at most six decimals and fits.
bool j_number(JCur& c, double& v, bool& fx, int64_t& mant) {
    uint64_t s = c.i;
    bool neg = false;
    if (c.i < c.n && c.p[c.i] == '-') { neg = true; ++c.i; }
    if (c.i >= c.n || !is_digit(c.p[c.i])) return false;
    uint64_t ip = c.i;
    if (c.p[c.i] == '0') ++c.i;
    else while (c.i < c.n && is_digit(c.p[c.i])) ++c.i;
    uint64_t int_digits = c.i - ip;
    uint64_t frac_digits = 0;
    uint64_t fs = 0;
    if (c.i < c.n && c.p[c.i] == '.') {
        ++c.i;
        fs = c.i;
        while (c.i < c.n && is_digit(c.p[c.i])) ++c.i;
        frac_digits = c.i - fs;
        if (frac_digits == 0) return false;
    }
    bool has_exp = false;
    if (c.i < c.n && (c.p[c.i] | 0x20) == 'e') {
        has_exp = true;
        ++c.i;
        if (c.i < c.n && (c.p[c.i] == '+' || c.p[c.i] == '-')) ++c.i;
        uint64_t es = c.i;
        while (c.i < c.n && is_digit(c.p[c.i])) ++c.i;
        if (c.i == es) return false;
    }
    if (!j_delim(c)) return false;
    if (!parse_f64_cell(c.p + s, c.i - s, v)) return false;
    fx = false;
    if (!has_exp && frac_digits <= 6 && int_digits <= 12) {
        int64_t m = 0;
        for (uint64_t k = ip; k < ip + int_digits; ++k) m = m * 10 + (c.p[k] - '0');
        for (uint64_t k = 0; k < 6; ++k)
            m = m * 10 + (k < frac_digits ? (c.p[fs + k] - '0') : 0);
        mant = neg ? -m : m;
        fx = true;
    }
    return true;
}

// One record (a line without its '\n'): returns 1 valid, 0 invalid.
int ndjson_record(const uint8_t* p, uint64_t n, int64_t& id, double& score, bool& fx,
                  int64_t& mant) {
    JCur c{p, 0, n};
    c.ws();
    if (c.i >= c.n || c.p[c.i] != '{') return 0;
    ++c.i;
    bool have_id = false, have_score = false;
    c.ws();
    if (c.i < c.n && c.p[c.i] == '}') return 0;
    while (true) {
        c.ws();
        uint64_t ks, ke;
        if (!j_key(c, ks, ke)) return 0;
        c.ws();
        if (c.i >= c.n || c.p[c.i] != ':') return 0;
        ++c.i;
        c.ws();
        if (!have_id && key_is(c, ks, ke, "id")) {
            if (!j_int(c, id)) return 0;
            have_id = true;
        } else if (!have_score && key_is(c, ks, ke, "nested") && c.i < c.n && c.p[c.i] == '{') {
            ++c.i;
            c.ws();
            if (c.i < c.n && c.p[c.i] == '}') { ++c.i; }
            else {
                while (true) {
                    c.ws();
                    uint64_t ks2, ke2;
                    if (!j_key(c, ks2, ke2)) return 0;
                    c.ws();
                    if (c.i >= c.n || c.p[c.i] != ':') return 0;
                    ++c.i;
                    c.ws();
                    if (!have_score && key_is(c, ks2, ke2, "score")) {
                        if (!j_number(c, score, fx, mant)) return 0;
                        have_score = true;
                    } else if (!j_skip_value(c)) {
                        return 0;
                    }
                    c.ws();
                    if (c.i < c.n && c.p[c.i] == ',') { ++c.i; continue; }
                    if (c.i < c.n && c.p[c.i] == '}') { ++c.i; break; }
                    return 0;
                }
            }
        } else if (!j_skip_value(c)) {
            return 0;
        }
        c.ws();
        if (c.i < c.n && c.p[c.i] == ',') { ++c.i; continue; }
        if (c.i < c.n && c.p[c.i] == '}') { ++c.i; break; }
        return 0;
    }
    return (have_id && have_score) ? 1 : 0;
}

bool blank_line(const uint8_t* p, uint64_t n) {
    for (uint64_t i = 0; i < n; ++i)
        if (!jws(p[i])) return false;
    return true;
}

void ndjson_id_score_range(const uint8_t* data, uint64_t from, uint64_t to,
                           falx_ndjson_id_score_stats& st) {
    uint64_t i = from;
    while (i < to) {
        const uint8_t* nl = (const uint8_t*)std::memchr(data + i, '\n', to - i);
        uint64_t e = nl ? (uint64_t)(nl - data) : to;
        if (!blank_line(data + i, e - i)) {
            int64_t id = 0, mant = 0;
            double score = 0;
            bool fx = false;
            if (ndjson_record(data + i, e - i, id, score, fx, mant)) {
                st.records += 1;
                st.id_sum = (int64_t)((uint64_t)st.id_sum + (uint64_t)id);
                st.score_sum += score;
                if (fx) { st.score_fixed6_sum += mant; st.score_fixed6_cells += 1; }
            } else {
                st.invalid += 1;
            }
        }
        i = e + 1;
    }
}

// ndjson_line_bounds (ndjson.rs:857-879): snap to the byte after a raw '\n'.
std::vector<uint64_t> ndjson_line_bounds(const uint8_t* data, uint64_t len, int threads) {
    if (len == 0) return {0};
    threads = clamp_threads(threads, len);
    uint64_t chunk = (len / threads + 63) & ~63ULL;
    if (threads == 1 || chunk == 0) return {0, len};
    std::vector<uint64_t> b{0};
    for (int t = 1; t < threads; ++t) {
        uint64_t bound = std::min<uint64_t>((uint64_t)t * chunk, len);
        while (bound < len && data[bound - 1] != '\n') ++bound;
        if (bound > b.back() && bound < len) b.push_back(bound);
    }
    b.push_back(len);
    return b;
}

// LogfmtSink (logfmt.rs:717-812).
struct LogfmtSink {
    uint64_t data_len, record_start = 0, field_start = 0, key_start = 0, key_end = 0;
    bool has_key = false;
    falx_logfmt_stats st{0, 0, 0, 0};
    explicit LogfmtSink(uint64_t n) : data_len(n) {}
    static void add_value(falx_logfmt_stats& st, const uint8_t* v, uint64_t n) {
        const uint8_t Q = '"', E = '\\';
        if (n == 0 || v[0] != Q) {
            st.value_bytes += n;
            for (uint64_t i = 0; i < n; ++i) st.checksum += v[i];
            return;
        }
        const uint8_t* body = v;
        uint64_t m = n;
        if (n >= 2 && v[n - 1] == Q) { body = v + 1; m = n - 2; }
        uint64_t i = 0;
        while (i < m) {
            if (body[i] == E && i + 1 < m) { st.value_bytes += 1; st.checksum += body[i + 1]; i += 2; }
            else { st.value_bytes += 1; st.checksum += body[i]; i += 1; }
        }
    }
    void field(const uint8_t* data, uint64_t e) {
        if (has_key) {
            st.pairs += 1;
            st.key_bytes += key_end - key_start;
            for (uint64_t i = key_start; i < key_end; ++i) st.checksum += data[i];
            add_value(st, data + field_start, e - field_start);
            has_key = false;
        } else {
            key_start = field_start;
            key_end = e;
            has_key = true;
        }
    }
    void on(const uint8_t* data, uint64_t p, bool term) {
        uint64_t fe = (term && p > record_start && data[p - 1] == '\r') ? p - 1 : p;
        field(data, fe);
        field_start = p + 1;
        if (term) { has_key = false; record_start = p + 1; }
    }
    void finish(const uint8_t* data) {
        if (field_start < data_len || (data_len > 0 && data[data_len - 1] != '\n')) {
            uint64_t e = (data_len > record_start && data[data_len - 1] == '\r') ? data_len - 1 : data_len;
            field(data, e);
        }
    }
};

falx_spec logfmt_spec() {
    falx_spec s;
    std::memset(&s, 0, sizeof s);
    s.structural[0] = ' '; s.structural[1] = '='; s.structural[2] = '\n';
    s.n_structural = 3;
    s.quote = '"'; s.escape = '\\'; s.comment = -1;
    return s;
}

falx_spec ndjson_spec() {
    falx_spec s;
    std::memset(&s, 0, sizeof s);
    s.structural[0] = '\n';
    s.n_structural = 1;
    s.quote = '"'; s.escape = '\\'; s.comment = -1;
    return s;
}

falx_logfmt_stats logfmt_serial(const uint8_t* data, uint64_t len) {
    falx_spec sp = logfmt_spec();
    Dialect d = make_dialect(&sp);
    LogfmtSink sink(len);
    Carry c;
    run_blocks(d, data, len, 0, ~0ULL, c, [&](uint64_t S, uint64_t T, uint32_t base) {
        while (S) {
            uint32_t i = (uint32_t)__builtin_ctzll(S);
            sink.on(data, base + i, (T >> i) & 1);
            S &= S - 1;
        }
        return true;
    });
    sink.finish(data);
    return sink.st;
}

uint64_t ndjson_terms(const uint8_t* data, uint64_t len) {
    falx_spec sp = ndjson_spec();
    Dialect d = make_dialect(&sp);
    Carry c;
    uint64_t count = 0;
    run_blocks(d, data, len, 0, ~0ULL, c, [&](uint64_t, uint64_t T, uint32_t) {
        count += (uint64_t)__builtin_popcountll(T);
        return true;
    });
    return count;
}

// ---- JSON nested tape (kernels/json.rs:636-760, 996-1108) -----------------
// json_dialect (formats.rs:164-173).
falx_spec json_spec() {
    falx_spec s;
    std::memset(&s, 0, sizeof s);
    const char* b = "{}[],:";
    for (int i = 0; i < 6; ++i) s.structural[i] = (uint8_t)b[i];
    s.n_structural = 6;
    s.quote = '"'; s.escape = '\\'; s.comment = -1;
    return s;
}

constexpr uint64_t kNoPartner = 0xFFFFFFFFull;

uint8_t close_of(uint8_t b) { return b == '{' ? '}' : b == '[' ? ']' : 0; }

// The tape is written into caller memory of at least idx.size() entries.
struct Nest {
    uint64_t* tape = nullptr;
    uint64_t n = 0;          // tape entries
    int32_t error = 0;       // falx_nest_error
    uint64_t error_pos = 0;
};

// push_nested (json.rs:692-760) over the structural positions, serially:
// separators append their position; an open pushes (tape index << 8) | the
// close byte it expects and appends (no partner << 32) | pos; a close pops
// and patches both entries, or stops the tape at an empty stack or a wrong
// kind. Then the innermost open left on the stack (json.rs:651-653).
void nest_serial(const uint8_t* data, const uint32_t* idx, uint64_t n_idx, Nest& out) {
    uint64_t* tape = out.tape;
    uint64_t t = 0;
    out.error = 0;
    out.error_pos = 0;
    std::vector<uint64_t> stack;
    for (uint64_t k = 0; k < n_idx; ++k) {
        const uint32_t pos = idx[k];
        const uint8_t b = data[pos];
        if (b == '{' || b == '[') {
            stack.push_back((t << 8) | close_of(b));
            tape[t++] = (kNoPartner << 32) | pos;
        } else if (b == '}' || b == ']') {
            if (stack.empty() || (uint8_t)stack.back() != b) {
                out.error = FALX_NEST_UNMATCHED_CLOSE;
                out.error_pos = pos;
                out.n = t;
                return;
            }
            const uint64_t open = stack.back() >> 8;
            stack.pop_back();
            tape[open] = (t << 32) | (uint32_t)tape[open];
            tape[t] = (open << 32) | pos;
            ++t;
        } else {
            tape[t++] = pos;
        }
    }
    out.n = t;
    if (!stack.empty()) {
        out.error = FALX_NEST_UNCLOSED_OPEN;
        out.error_pos = (uint32_t)tape[stack.back() >> 8];
    }
}

// parse_nested_par (json.rs:1010-1108): chunks of the tape matched
// concurrently into their own slot ranges (here chunks of the structural
// index, whose entry counts are the slot counts); a close with no local open
// is left pending, opens left on a chunk's stack are carried; a pending
// close is matched in order against the opens of earlier chunks (the
// parenthesis reduction). Any definite mismatch falls back to the serial
// pass, as the reference does.
void nest_par(const uint8_t* data, const uint32_t* idx, uint64_t n, int threads, Nest& out) {
    const int T = std::max(1, std::min<int>(threads, (int)(n / 64 + 1)));
    if (T == 1) return nest_serial(data, idx, n, out);
    uint64_t* tape = out.tape;
    struct Part {
        bool bad = false;
        std::vector<uint64_t> opens;     // (tape index << 8) | close byte, bottom to top
        std::vector<uint64_t> pending;   // (tape index << 32) | pos, in order
    };
    std::vector<Part> parts(T);
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            const uint64_t lo = n * t / T, hi = n * (t + 1) / T;
            Part& P = parts[t];
            for (uint64_t i = lo; i < hi; ++i) {
                const uint32_t pos = idx[i];
                const uint8_t b = data[pos];
                if (b == '{' || b == '[') {
                    P.opens.push_back((i << 8) | close_of(b));
                    tape[i] = (kNoPartner << 32) | pos;
                } else if (b == '}' || b == ']') {
                    if (P.opens.empty()) {
                        P.pending.push_back((i << 32) | pos);
                        tape[i] = pos;
                        continue;
                    }
                    if ((uint8_t)P.opens.back() != b) {
                        P.bad = true;
                        return;
                    }
                    const uint64_t open = P.opens.back() >> 8;
                    P.opens.pop_back();
                    tape[open] = (i << 32) | (uint32_t)tape[open];
                    tape[i] = (open << 32) | pos;
                } else {
                    tape[i] = pos;
                }
            }
        });
    for (auto& x : th) x.join();
    for (const Part& P : parts)
        if (P.bad) return nest_serial(data, idx, n, out);
    std::vector<uint64_t> g;
    for (const Part& P : parts) {
        for (uint64_t pend : P.pending) {
            const uint64_t ci = pend >> 32;
            const uint32_t cpos = (uint32_t)pend;
            if (g.empty() || (uint8_t)g.back() != data[cpos]) return nest_serial(data, idx, n, out);
            const uint64_t oi = g.back() >> 8;
            g.pop_back();
            tape[oi] = (ci << 32) | (uint32_t)tape[oi];
            tape[ci] = (oi << 32) | cpos;
        }
        g.insert(g.end(), P.opens.begin(), P.opens.end());
    }
    out.n = n;
    out.error = 0;
    out.error_pos = 0;
    if (!g.empty()) {
        out.error = FALX_NEST_UNCLOSED_OPEN;
        out.error_pos = (uint32_t)tape[g.back() >> 8];
    }
}

// ---- DEFLATE (RFC 1951) and BGZF framing (bgzf.rs:61-290) ---------------
// The reference inflates with flate2 1.x (miniz_oxide) or libdeflater
// 1.25.2, neither in this tree; this is the published algorithm restated:
// block headers, stored/fixed/dynamic blocks, canonical Huffman codes read
// most significant bit first, length/distance pairs with the RFC's tables.
struct Inflater {
    const uint8_t* src;
    uint64_t n, pos = 0;
    uint64_t bb = 0;
    int bc = 0;
    bool err = false;

    uint32_t bits(int k) {
        while (bc < k) {
            if (pos >= n) { err = true; return 0; }
            bb |= (uint64_t)src[pos++] << bc;
            bc += 8;
        }
        const uint32_t v = (uint32_t)(bb & ((1ull << k) - 1));
        bb >>= k;
        bc -= k;
        return v;
    }
};

struct Canon {
    int count[16];
    int sym[288];
    // Per length counts and symbols in code order; false when over-subscribed
    // or incomplete (unless `single` and it is one code of length 1, or empty).
    bool build(const uint8_t* len, int n, bool single) {
        for (int l = 0; l < 16; ++l) count[l] = 0;
        for (int s = 0; s < n; ++s) ++count[len[s]];
        if (count[0] == n) return single;
        int left = 1;
        for (int l = 1; l < 16; ++l) {
            left = 2 * left - count[l];
            if (left < 0) return false;
        }
        if (left > 0 && !(single && n - count[0] == 1 && count[1] == 1)) return false;
        int offs[16];
        offs[1] = 0;
        for (int l = 1; l < 15; ++l) offs[l + 1] = offs[l] + count[l];
        for (int s = 0; s < n; ++s)
            if (len[s]) sym[offs[len[s]]++] = s;
        return true;
    }
    // Next symbol: codes are canonical, first bit sent is the top bit.
    int decode(Inflater& in) const {
        int code = 0, first = 0, idx = 0;
        for (int l = 1; l < 16; ++l) {
            code |= (int)in.bits(1);
            if (in.err) return -1;
            if (code - first < count[l]) return sym[idx + code - first];
            idx += count[l];
            first = (first + count[l]) << 1;
            code <<= 1;
        }
        return -1;
    }
};

const uint16_t kOraLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
const uint8_t kOraLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
const uint16_t kOraDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129, 193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
const uint8_t kOraDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
const uint8_t kOraClOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// Raw DEFLATE into out[0, cap); the output length, or -1 when malformed or
// longer than cap.
int64_t inflate_raw(const uint8_t* src, uint64_t n, uint8_t* out, uint64_t cap) {
    Inflater in{src, n};
    uint64_t o = 0;
    bool last = false;
    Canon lit, dist;
    while (!last) {
        last = in.bits(1) != 0;
        const uint32_t type = in.bits(2);
        if (in.err) return -1;
        if (type == 0) {
            in.bb = 0;   // to the byte boundary (whole bytes are never buffered ahead)
            in.bc = 0;
            if (in.pos + 4 > n) return -1;
            const uint32_t ln = src[in.pos] | (src[in.pos + 1] << 8), nl = src[in.pos + 2] | (src[in.pos + 3] << 8);
            in.pos += 4;
            if ((ln ^ 0xFFFFu) != nl || in.pos + ln > n || o + ln > cap) return -1;
            std::memcpy(out + o, src + in.pos, ln);
            in.pos += ln;
            o += ln;
            continue;
        }
        if (type == 3) return -1;
        uint8_t lens[320];
        if (type == 1) {
            for (int s = 0; s < 320; ++s) lens[s] = s < 144 ? 8 : s < 256 ? 9 : s < 280 ? 7 : s < 288 ? 8 : 5;
        } else {
            const int nlen = (int)in.bits(5) + 257, ndist = (int)in.bits(5) + 1, ncl = (int)in.bits(4) + 4;
            if (in.err || nlen > 286 || ndist > 30) return -1;
            uint8_t cl[19] = {0};
            for (int k = 0; k < ncl; ++k) cl[kOraClOrder[k]] = (uint8_t)in.bits(3);
            Canon cc;
            if (in.err || !cc.build(cl, 19, false)) return -1;
            std::memset(lens, 0, sizeof lens);
            int i = 0, prev = 0;
            while (i < nlen + ndist) {
                const int sym = cc.decode(in);
                if (sym < 0) return -1;
                int rep = 1, val = sym;
                if (sym == 16) {
                    if (i == 0) return -1;
                    val = prev;
                    rep = 3 + (int)in.bits(2);
                } else if (sym == 17) {
                    val = 0;
                    rep = 3 + (int)in.bits(3);
                } else if (sym == 18) {
                    val = 0;
                    rep = 11 + (int)in.bits(7);
                }
                if (in.err || i + rep > nlen + ndist) return -1;
                for (int k = 0; k < rep; ++k, ++i) lens[i < nlen ? i : 288 + (i - nlen)] = (uint8_t)val;
                prev = val;
            }
            if (lens[256] == 0) return -1;
        }
        if (!lit.build(lens, 288, true) || !dist.build(lens + 288, 32, true)) return -1;
        while (true) {
            const int sym = lit.decode(in);
            if (sym < 0) return -1;
            if (sym < 256) {
                if (o >= cap) return -1;
                out[o++] = (uint8_t)sym;
                continue;
            }
            if (sym == 256) break;
            if (sym > 285) return -1;
            const uint32_t len = kOraLenBase[sym - 257] + in.bits(kOraLenExtra[sym - 257]);
            const int ds = dist.decode(in);
            if (ds < 0 || ds > 29) return -1;
            const uint32_t d = kOraDistBase[ds] + in.bits(kOraDistExtra[ds]);
            if (in.err || d > o || o + len > cap) return -1;
            for (uint32_t k = 0; k < len; ++k, ++o) out[o] = out[o - d];
        }
    }
    return (int64_t)o;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* ora_isa(void) { return g_avx512 ? "avx512bw" : "scalar"; }

uint64_t ora_index_scalar(const falx_spec* spec, const uint8_t* data, uint64_t len,
                          uint32_t* out) {
    bool structural[256] = {false};
    for (uint32_t i = 0; i < spec->n_structural && i < 16; ++i) structural[spec->structural[i]] = true;
    int quote = spec->quote, escape = spec->quote >= 0 ? spec->escape : -1, comment = spec->comment;
    bool in_quotes = false, in_comment = false, line_start = true;
    uint64_t escape_run = 0, n = 0;
    for (uint64_t i = 0; i < len; ++i) {
        uint8_t b = data[i];
        bool escaped = (escape_run & 1) == 1;
        if ((int)b == escape) ++escape_run; else escape_run = 0;
        if (in_comment) {
            if (b == '\n') {
                in_comment = false;
                if (structural['\n']) out[n++] = (uint32_t)i;
            }
        } else if ((int)b == quote) {
            if (!escaped) in_quotes = !in_quotes;
        } else if (!in_quotes && (int)b == comment && line_start) {
            in_comment = true;
        } else if (!in_quotes && structural[b]) {
            out[n++] = (uint32_t)i;
        }
        line_start = b == '\n';
    }
    return n;
}

uint64_t ora_index_structurals(const falx_spec* spec, const uint8_t* data, uint64_t len,
                               int threads, uint32_t* out) {
    Dialect d = make_dialect(spec);
    Chunks ch;
    ParKind kind = par_kind(d);
    auto index_range = [&](uint64_t from, uint64_t to, Carry c, std::vector<uint32_t>& part) {
        // Index data[from..to) as its own buffer (chunk-local zero padding).
        run_blocks(d, data + from, to - from, 0, ~0ULL, c, [&](uint64_t S, uint64_t, uint32_t base) {
            while (S) { part.push_back((uint32_t)from + base + (uint32_t)__builtin_ctzll(S)); S &= S - 1; }
            return true;
        });
        return c;
    };
    if (kind == ParKind::Serial || !make_chunks(len, threads, ch)) {
        uint64_t n = 0;
        Carry c;
        run_blocks(d, data, len, 0, ~0ULL, c, [&](uint64_t S, uint64_t, uint32_t base) {
            while (S) { out[n++] = base + (uint32_t)__builtin_ctzll(S); S &= S - 1; }
            return true;
        });
        return n;
    }
    int T = ch.threads;
    std::vector<std::vector<uint32_t>> parts(T);
    if (kind == ParKind::Parity) {
        // Speculative seed 0, redo chunks entered mid-quote (csv_geo.rs:95-149).
        std::vector<uint64_t> carry(T);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] {
                Carry c = index_range(ch.bounds[t], ch.bounds[t + 1], Carry{}, parts[t]);
                carry[t] = c.parity;
            });
        for (auto& x : th) x.join();
        std::vector<int> redo;
        uint64_t entry = 0;
        for (int t = 0; t < T; ++t) {
            if (entry != 0) redo.push_back(t);
            entry ^= carry[t];
        }
        th.clear();
        for (int t : redo)
            th.emplace_back([&, t] {
                parts[t].clear();
                Carry c;
                c.parity = ~0ULL;
                index_range(ch.bounds[t], ch.bounds[t + 1], c, parts[t]);
            });
        for (auto& x : th) x.join();
    } else {
        std::vector<Carry> entry = chunk_entry_carries(d, data, ch, kind);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] { index_range(ch.bounds[t], ch.bounds[t + 1], entry[t], parts[t]); });
        for (auto& x : th) x.join();
    }
    // scatter_u32 (csv_geo.rs:154-184): disjoint slots, copied in parallel.
    std::vector<uint64_t> at(T + 1, 0);
    for (int t = 0; t < T; ++t) at[t + 1] = at[t] + parts[t].size();
    std::vector<std::thread> th;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] { std::memcpy(out + at[t], parts[t].data(), parts[t].size() * 4); });
    for (auto& x : th) x.join();
    return at[T];
}

void ora_parse(const falx_spec* spec, const uint8_t* data, uint64_t len, int threads,
               uint32_t* seps, uint64_t* n_seps, uint64_t* ends, uint64_t* n_ends) {
    Dialect d = make_dialect(spec);
    Chunks ch;
    ParKind kind = par_kind(d);
    struct Part { std::vector<uint32_t> seps; std::vector<uint64_t> ends; };
    auto tape_range = [&](uint64_t from, uint64_t to, Carry c, Part& part) {
        run_blocks(d, data + from, to - from, 0, ~0ULL, c, [&](uint64_t S, uint64_t T, uint32_t base) {
            uint64_t sep = S & ~T;
            uint64_t base_count = part.seps.size();
            uint32_t b = (uint32_t)from + base;
            for (uint64_t m = sep; m; m &= m - 1) part.seps.push_back(b + (uint32_t)__builtin_ctzll(m));
            for (uint64_t t = T; t; t &= t - 1) {
                uint32_t idx = (uint32_t)__builtin_ctzll(t);
                uint64_t below = (uint64_t)__builtin_popcountll(sep & ((1ULL << idx) - 1));
                part.ends.push_back(((base_count + below) << 32) | (uint64_t)(b + idx));
            }
            return true;
        });
        return c;
    };
    std::vector<Part> parts;
    if (kind == ParKind::Serial || !make_chunks(len, threads, ch)) {
        parts.resize(1);
        tape_range(0, len, Carry{}, parts[0]);
    } else {
        int T = ch.threads;
        parts.resize(T);
        std::vector<Carry> entry = chunk_entry_carries(d, data, ch, kind);
        std::vector<std::thread> th;
        for (int t = 0; t < T; ++t)
            th.emplace_back([&, t] { tape_range(ch.bounds[t], ch.bounds[t + 1], entry[t], parts[t]); });
        for (auto& x : th) x.join();
    }
    // scatter_tape (csv_geo.rs:277-334): rebase the ends' high word.
    uint64_t sp = 0, ep = 0;
    for (auto& p : parts) {
        std::memcpy(seps + sp, p.seps.data(), p.seps.size() * 4);
        for (size_t i = 0; i < p.ends.size(); ++i) ends[ep + i] = p.ends[i] + (sp << 32);
        sp += p.seps.size();
        ep += p.ends.size();
    }
    *n_seps = sp;
    *n_ends = ep;
}

int ora_parse_columns(const falx_spec* spec, const uint8_t* data, uint64_t len, int threads,
                      falx_columns* out) {
    Dialect d = make_dialect(spec);
    const falx_spec& sp = *spec;
    auto parts = run_sinks<ColumnSink>(d, data, len, threads, [&](uint32_t s, uint32_t e, bool em) {
        return ColumnSink(d, sp, data, (uint32_t)len, s, e, em);
    });
    // parse_columns_par concatenation (csv_geo.rs:1216-1284): values copied,
    // string offsets rebased by the previous parts' data length, bitmaps joined.
    uint64_t rows = 0;
    for (auto& p : parts) rows += p.out.rows;
    out->rows = rows;
    if (rows > out->rows_cap) return FALX_E_CAPACITY;
    for (uint32_t k = 0; k < sp.n_columns; ++k) {
        falx_column_buf& b = out->cols[k];
        std::vector<uint64_t> valid;
        uint64_t r = 0, dlen = 0;
        if (sp.columns[k].type == FALX_COL_STR) ((int32_t*)b.values)[0] = 0;
        for (auto& p : parts) {
            auto& c = p.out.cols[k];
            uint64_t pr = p.out.rows;
            switch (sp.columns[k].type) {
            case FALX_COL_I64: std::memcpy((int64_t*)b.values + r, c.i64.data(), pr * 8); break;
            case FALX_COL_F64: std::memcpy((double*)b.values + r, c.f64.data(), pr * 8); break;
            case FALX_COL_BYTES: std::memcpy((uint32_t*)b.values + 2 * r, c.spans.data(), pr * 8); break;
            default: {
                if (dlen + c.data.size() > b.data_cap) return FALX_E_CAPACITY;
                for (uint64_t i = 0; i < pr; ++i)
                    ((int32_t*)b.values)[r + i + 1] = (int32_t)(c.offsets[i + 1] + dlen);
                std::memcpy(b.data + dlen, c.data.data(), c.data.size());
                dlen += c.data.size();
            }
            }
            append_bitmap(valid, r, c.valid, pr);
            r += pr;
        }
        if (sp.columns[k].type == FALX_COL_STR) b.data_len = dlen;
        std::memcpy(b.valid, valid.data(), valid.size() * 8);
    }
    return FALX_OK;
}

void ora_csv_stats(const falx_spec* spec, const uint8_t* data, uint64_t len, int threads,
                   falx_csv_stats* out) {
    Dialect d = make_dialect(spec);
    const falx_spec& sp = *spec;
    auto parts = run_sinks<StatsSink>(d, data, len, threads, [&](uint32_t s, uint32_t e, bool em) {
        return StatsSink(d, sp, data, (uint32_t)len, s, e, em);
    });
    std::memset(out, 0, sizeof *out);
    for (auto& p : parts) {  // CsvGeoStats::merge (csv_geo.rs:1762-1769), wrapping
        out->records += p.st.records;
        for (uint32_t k = 0; k < sp.n_columns; ++k) {
            auto& a = out->cols[k];
            auto& b = p.st.cols[k];
            a.values += b.values;
            a.checksum += b.checksum;
            a.bytes += b.bytes;
            a.fixed6_sum += b.fixed6_sum;
            a.fixed6_cells += b.fixed6_cells;
            a.sum += b.sum;
        }
    }
}

void ora_field_bytes(const falx_spec* spec, const uint8_t* data, uint64_t len, int threads,
                     falx_field_byte_stats* out) {
    Dialect d = make_dialect(spec);
    auto parts = run_sinks<FieldByteSink>(d, data, len, threads, [&](uint32_t s, uint32_t e, bool em) {
        return FieldByteSink(d, data, (uint32_t)len, s, e, em);
    });
    out->fields = 0;
    out->bytes = 0;
    for (auto& p : parts) { out->fields += p.st.fields; out->bytes += p.st.bytes; }
}

uint64_t ora_ndjson_lines(const uint8_t* data, uint64_t len, int threads) {
    // ndjson.rs:618-657: terminators per line-aligned chunk, each chunk from
    // fresh carries, plus the trailing unterminated record.
    uint64_t trailing = (len > 0 && data[len - 1] != '\n') ? 1 : 0;
    std::vector<uint64_t> b = threads > 1 ? ndjson_line_bounds(data, len, threads)
                                          : std::vector<uint64_t>{0, len};
    if (b.size() <= 2) return ndjson_terms(data, len) + trailing;
    std::vector<uint64_t> counts(b.size() - 1);
    std::vector<std::thread> th;
    for (size_t t = 0; t + 1 < b.size(); ++t)
        th.emplace_back([&, t] { counts[t] = ndjson_terms(data + b[t], b[t + 1] - b[t]); });
    for (auto& x : th) x.join();
    uint64_t total = trailing;
    for (auto c : counts) total += c;
    return total;
}

void ora_ndjson_id_score(const uint8_t* data, uint64_t len, int threads,
                         falx_ndjson_id_score_stats* out) {
    std::memset(out, 0, sizeof *out);
    std::vector<uint64_t> b = threads > 1 ? ndjson_line_bounds(data, len, threads)
                                          : std::vector<uint64_t>{0, len};
    if (b.size() <= 2) { ndjson_id_score_range(data, 0, len, *out); return; }
    std::vector<falx_ndjson_id_score_stats> parts(b.size() - 1);
    std::vector<std::thread> th;
    for (size_t t = 0; t + 1 < b.size(); ++t)
        th.emplace_back([&, t] {
            std::memset(&parts[t], 0, sizeof parts[t]);
            ndjson_id_score_range(data, b[t], b[t + 1], parts[t]);
        });
    for (auto& x : th) x.join();
    for (auto& p : parts) {
        out->records += p.records;
        out->invalid += p.invalid;
        out->id_sum = (int64_t)((uint64_t)out->id_sum + (uint64_t)p.id_sum);
        out->score_fixed6_sum += p.score_fixed6_sum;
        out->score_fixed6_cells += p.score_fixed6_cells;
        out->score_sum += p.score_sum;
    }
}

void ora_logfmt_pairs(const uint8_t* data, uint64_t len, int threads, falx_logfmt_stats* out) {
    // logfmt.rs:632-687: serial below 1 MiB; else newline-snapped chunks.
    int t = clamp_threads(threads, len);
    if (t == 1 || len < (1u << 20)) { *out = logfmt_serial(data, len); return; }
    uint64_t approx = len / t;
    std::vector<uint64_t> b(t + 1, 0);
    b[t] = len;
    for (int i = 1; i < t; ++i) {
        uint64_t from = std::min<uint64_t>((uint64_t)i * approx, len);
        const uint8_t* nl = (const uint8_t*)std::memchr(data + from, '\n', len - from);
        b[i] = nl ? (uint64_t)(nl - data) + 1 : len;
    }
    std::vector<falx_logfmt_stats> parts(t, falx_logfmt_stats{0, 0, 0, 0});
    std::vector<std::thread> th;
    for (int i = 0; i < t; ++i)
        th.emplace_back([&, i] {
            if (b[i] < b[i + 1]) parts[i] = logfmt_serial(data + b[i], b[i + 1] - b[i]);
        });
    for (auto& x : th) x.join();
    *out = falx_logfmt_stats{0, 0, 0, 0};
    for (auto& p : parts) {
        out->pairs += p.pairs;
        out->key_bytes += p.key_bytes;
        out->value_bytes += p.value_bytes;
        out->checksum += p.checksum;
    }
}

int ora_parse_f64(const uint8_t* s, uint64_t n, double* out, int* fx, int64_t* mant) {
    bool f = false;
    int64_t m = 0;
    double v = 0;
    bool ok = parse_f64_cell(s, n, v, &f, &m);
    *out = ok ? v : 0.0;
    if (fx) *fx = f;
    if (mant) *mant = m;
    return ok;
}

int ora_parse_i64(const uint8_t* s, uint64_t n, int64_t* out) {
    int64_t v = 0;
    bool ok = parse_i64_cell(s, n, v);
    *out = ok ? v : 0;
    return ok;
}

uint64_t ora_clean(const falx_spec* spec, const uint8_t* s, uint64_t n, uint8_t* out) {
    Dialect d = make_dialect(spec);
    uint64_t k = 0;
    clean_cell(d, s, n, [&](uint8_t b) { out[k++] = b; });
    return k;
}

}  // extern "C"

extern "C" {

uint64_t ora_parse_nested(const uint8_t* data, uint64_t len, int threads, uint64_t* tape, uint64_t cap,
                          int32_t* error, uint64_t* error_pos) {
    falx_spec sp = json_spec();
    // Uninitialised: pages are touched only as the index fills them.
    std::unique_ptr<uint32_t[]> idx(new uint32_t[len + 1]);
    const uint64_t n_idx = ora_index_structurals(&sp, data, len, threads, idx.get());
    *error = 0;
    *error_pos = 0;
    if (n_idx > cap) return n_idx;   // the tape needs at most n_idx entries
    Nest out;
    out.tape = tape;
    if (threads <= 1) nest_serial(data, idx.get(), n_idx, out);
    else nest_par(data, idx.get(), n_idx, threads, out);
    *error = out.error;
    *error_pos = out.error_pos;
    return out.n;
}

}  // extern "C"

extern "C" {

int64_t ora_inflate(const uint8_t* src, uint64_t n, uint8_t* out, uint64_t cap) {
    return inflate_raw(src, n, out, cap);
}

// bgzf::decompress_par (bgzf.rs:209-260) over a block table from
// falx_bgzf_scan's rules: the blocks in contiguous runs per thread, each
// block into its slot; -1 when any block is malformed or its inflated size
// differs from ISIZE.
int64_t ora_bgzf_inflate(const uint8_t* comp, const falx_bgzf_block* blocks, uint64_t n_blocks, int threads,
                         uint8_t* out) {
    const int T = std::max(1, std::min<int>(threads, (int)std::max<uint64_t>(1, n_blocks)));
    std::vector<int> bad(T, 0);
    std::vector<std::thread> th;
    const uint64_t per = (n_blocks + T - 1) / T;
    for (int t = 0; t < T; ++t)
        th.emplace_back([&, t] {
            for (uint64_t k = t * per; k < std::min(n_blocks, (t + 1) * per); ++k) {
                const falx_bgzf_block& b = blocks[k];
                if (inflate_raw(comp + b.payload, b.payload_len, out + b.out, b.isize) != (int64_t)b.isize) {
                    bad[t] = 1;
                    return;
                }
            }
        });
    for (auto& x : th) x.join();
    for (int t = 0; t < T; ++t)
        if (bad[t]) return -1;
    uint64_t total = 0;
    for (uint64_t k = 0; k < n_blocks; ++k) total += blocks[k].isize;
    return (int64_t)total;
}

}  // extern "C"
