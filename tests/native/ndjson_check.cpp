// ndjson_check.cpp — host build of falx_amd/csrc/ndjson_walk.h (the record
// walker the gfx950 NDJSON kernel runs) swept against the CPU oracle on
// randomized records: valid shapes, decoy keys inside strings, escapes,
// whitespace, missing keys and malformed JSON.
// usage: ndjson_check <seed> <count>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../falx_amd/csrc/ndjson_walk.h"
#include "../../oracle/oracle.h"

namespace {

struct Rng {
    uint64_t s;
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 0x2545F4914F6CDD1DULL;
    }
    uint64_t below(uint64_t n) { return next() % n; }
};

struct BufSrc {
    const uint8_t* p;
    uint8_t at(uint32_t i) const { return p[i]; }
};

// The kernel's LDS source: bytes also read as aligned 8-byte words (p is
// 8-aligned, 32 bytes past the record are readable and hold decoys, digits
// among them).
struct WordSrc {
    static constexpr bool kWord8 = true;
    const uint8_t* p;
    uint8_t at(uint32_t i) const { return p[i]; }
    uint64_t word8(uint32_t a) const { uint64_t w; std::memcpy(&w, p + a, 8); return w; }
};

// One record through the walker: 1 valid, 0 invalid, -1 blank.
template <class Src>
int walk(const Src& src, uint32_t from, uint32_t to, fx::NdRecord& rec) {
    if (fx::blank_span(src, from, to)) return -1;
    int ok = fx::ndjson_record(src, from, to, rec);
    static fx::BigDec big;
    if (ok == 1) fx::ndjson_resolve(src, rec, big);
    return ok;
}

std::string ws(Rng& r) {
    const char* w[] = {"", "", "", " ", "  ", "\t", " \r "};
    return w[r.below(7)];
}

std::string jstring(Rng& r) {
    std::string s = "\"";
    int n = (int)r.below(12);
    for (int i = 0; i < n; ++i) {
        uint64_t k = r.below(14);
        if (k == 0) s += "\\\"";
        else if (k == 1) s += "\\\\";
        else if (k == 2) s += "id\\\":9";
        else if (k == 3) s += "{[";
        else if (k == 4) s += "}]";
        else s.push_back((char)('a' + r.below(26)));
    }
    return s + "\"";
}

std::string jnumber(Rng& r) {
    std::string s;
    if (r.below(3) == 0) s += "-";
    if (r.below(25) == 0) {  // long mantissas near the subnormal range: exact path
        s.push_back((char)('1' + r.below(9)));
        for (int i = 0, n = 15 + (int)r.below(30); i < n; ++i) s.push_back((char)('0' + r.below(10)));
        s += "e-" + std::to_string(280 + r.below(60));
        return s;
    }
    uint64_t k = r.below(8);
    if (k == 0) s += "0";
    else s += std::to_string(r.below(1000000));
    if (r.below(2)) { s += "."; int d = 1 + (int)r.below(8); for (int i = 0; i < d; ++i) s.push_back((char)('0' + r.below(10))); }
    if (r.below(6) == 0) { s += r.below(2) ? "e" : "E"; if (r.below(2)) s += "-"; s += std::to_string(r.below(30)); }
    if (r.below(40) == 0) s += "x";  // malformed
    return s;
}

std::string jvalue(Rng& r, int depth);

std::string jobject(Rng& r, int depth, bool with_keys) {
    std::string s = "{" + ws(r);
    int n = (int)r.below(4);
    bool id_first = with_keys && r.below(2);
    std::string id = "\"id\"" + ws(r) + ":" + ws(r) + (r.below(20) ? std::to_string((int64_t)r.below(1ULL << 40) - (r.below(4) == 0 ? 5000 : 0)) : jnumber(r));
    std::string nested = "\"nested\"" + ws(r) + ":" + ws(r) + "{" + ws(r) +
                         (r.below(3) ? "\"tags\":[" + jstring(r) + "]," + ws(r) : "") +
                         (r.below(15) ? "\"score\"" + ws(r) + ":" + ws(r) + jnumber(r) : "\"scor\":1") + ws(r) + "}";
    bool first = true;
    auto add = [&](const std::string& kv) {
        if (!first) s += "," + ws(r);
        s += kv;
        first = false;
    };
    if (id_first && r.below(10)) add(id);
    for (int i = 0; i < n; ++i) add(jstring(r) + ws(r) + ":" + ws(r) + jvalue(r, depth + 1));
    if (with_keys && r.below(10)) add(nested);
    if (!id_first && with_keys && r.below(10)) add(id);
    s += ws(r) + "}";
    return s;
}

std::string jvalue(Rng& r, int depth) {
    uint64_t k = r.below(depth > 2 ? 3 : 5);
    if (k == 0) return jstring(r);
    if (k == 1) return jnumber(r);
    if (k == 2) return r.below(2) ? "true" : "null";
    if (k == 3) return "[" + jvalue(r, depth + 1) + "," + ws(r) + jvalue(r, depth + 1) + "]";
    return jobject(r, depth + 1, false);
}

// Records in the common layout (as the data generator writes them). With
// probability 1/2 one part is changed to a form the straight-line reader
// must refuse or still read right; `plain` says no part was changed.
std::string canon(Rng& r, bool& plain) {
    auto sp = [&]() {
        const uint64_t k = r.below(10);
        return k == 0 ? std::string(" ") : (k == 1 ? std::string("  ") : std::string());
    };
    const int mut = r.below(2) ? -1 : (int)r.below(32);
    plain = mut < 0;
    auto ws_at = [&](int slot) {
        if (mut == 0 && slot == 3) return std::string("\t");
        if (mut == 1 && slot == 7) return std::string(" \r");
        if (mut == 2 && slot == 5) return std::string(8, ' ');
        if (mut == 3 && slot == 9) return std::string(5, ' ');
        return sp();
    };
    std::string id = std::to_string(r.below(1ULL << 40));
    if (mut == 4) id = "-" + id;
    if (mut == 5) id = "000" + id;
    if (mut == 6) id = std::string(16, '7');
    if (mut == 7) id = "123456789012345";
    if (mut == 8) id = "-";
    if (mut == 9) id = "12x";
    std::string name = "\"";
    for (int i = 0, n = 4 + (int)r.below(29); i < n; ++i) {
        const uint64_t k = r.below(40);
        if (k == 0) name += "\\\"";
        else if (k == 1) name += "\\\\";
        else if (k == 2) name += "[{";
        else name.push_back((char)('a' + r.below(26)));
    }
    name += "\"";
    if (mut == 10) name = "12";
    if (mut == 11) name = "{\"a\":[1,{\"b\":\"}\"}],\"c\":\"]\"}";
    if (mut == 12) name = "\"abc\\\"";
    const int dec = 1 + (int)r.below(6);
    std::string score = std::to_string(r.below(1000)) + ".";
    for (int i = 0; i < dec; ++i) score.push_back((char)('0' + r.below(10)));
    if (mut == 13) score += "e" + std::to_string(r.below(20));
    if (mut == 14) score = "012.5";
    if (mut == 15) score = "0.25";
    if (mut == 16) score = ".5";
    if (mut == 17) score = "5.";
    if (mut == 18) score = "123456789.1234567";
    if (mut == 19) score = "-" + score;
    if (mut == 20) score = std::to_string(r.below(100000));
    if (mut == 21) score = "1234567890123.5";
    std::string tags = "[";
    for (int k = 0, n = (int)r.below(6); k < n; ++k) {
        if (k) tags += "," + sp();
        tags += "\"";
        for (int i = 0, tl = 1 + (int)r.below(8); i < tl; ++i) {
            // Now and then a byte that matters to a skip: brackets and braces
            // inside the string, an escaped quote or backslash.
            const uint64_t t = r.below(60);
            if (t == 0) tags += "]";
            else if (t == 1) tags += "}";
            else if (t == 2) tags += "[";
            else if (t == 3) tags += "\\\"";
            else if (t == 4) tags += "\\\\";
            else tags.push_back((char)('a' + r.below(26)));
        }
        tags += "\"";
    }
    tags += "]";
    if (mut == 22) tags = "[[1, 2], {\"x\": [\"]\"]}, \"y\\\\\"]";
    if (mut == 23) tags = "\"t\"";
    if (mut == 30) tags = "[1, \"a]\", null,true ]";
    if (mut == 31) tags = "[\"" + std::string(r.below(40), 'x') + "\\\"\"]";
    std::string k_id = "\"id\"", k_name = "\"name\"", k_nested = "\"nested\"", k_score = "\"score\"";
    if (mut == 24) k_id = "\"Id\"";
    if (mut == 25) k_score = "\"scores\"";
    if (mut == 26) k_name = "\"id\"";
    std::string s = ws_at(0) + "{" + ws_at(1) + k_id + ws_at(2) + ":" + ws_at(3) + id + ws_at(4) + "," + ws_at(5) +
                    k_name + ws_at(6) + ":" + ws_at(7) + name + ws_at(8) + "," + ws_at(9) + k_nested + sp() + ":" +
                    sp() + "{" + sp() + k_score + sp() + ":" + sp() + score + sp() + "," + sp() + "\"tags\"" +
                    sp() + ":" + sp() + tags + sp() + "}" + sp() + "}";
    if (mut == 27) s += " trailing";
    if (mut == 28) s = s.substr(0, r.below(s.size() + 1));
    if (mut == 29) s.erase(s.size() - 1);
    return s;
}

// The straight-line reader on `line` at a random alignment of an 8-aligned
// buffer, with '\n' after the record (as the kernel sees it) and quotes,
// backslashes, digits and spaces after that. When it accepts, the walker must
// return 1 with the same fields. Returns -1 on a mismatch, else whether it
// accepted.
int common_agrees(Rng& r, const std::string& line, int ok, const fx::NdRecord& rec) {
    alignas(8) static uint8_t buf[1 << 16];
    const uint32_t off = (uint32_t)r.below(8);
    if (off + line.size() + 40 > sizeof buf) return 0;
    std::memcpy(buf + off, line.data(), line.size());
    buf[off + line.size()] = '\n';
    for (int k = 1; k < 40; ++k) buf[off + line.size() + k] = " \"\\7,}:"[r.below(7)];
    fx::NdRecord c{};
    if (!fx::ndjson_common(WordSrc{buf}, off, off + (uint32_t)line.size(), c)) return 0;
    if (ok != 1 || c.id != rec.id || c.fx6 != rec.fx6 || c.mant != rec.mant || c.score_deferred ||
        fx::f64_to_bits(c.score) != fx::f64_to_bits(rec.score)) {
        std::fprintf(stderr, "COMMON LAYOUT MISMATCH on record [%s]: walker %d id %" PRId64 "/%" PRId64
                             " mant %" PRId64 "/%" PRId64 "\n",
                     line.c_str(), ok, c.id, rec.id, c.mant, rec.mant);
        return -1;
    }
    return 1;
}

}  // namespace

int main(int argc, char** argv) {
    uint64_t seed = argc > 1 ? std::strtoull(argv[1], nullptr, 0) : 1;
    uint64_t count = argc > 2 ? std::strtoull(argv[2], nullptr, 0) : 100000;
    Rng r{seed | 1};
    uint64_t valid = 0;
    for (uint64_t i = 0; i < count; ++i) {
        std::string line = jobject(r, 0, true);
        if (r.below(50) == 0) line = line.substr(0, r.below(line.size() + 1));  // truncated record
        std::string withnl = line + "\n";
        falx_ndjson_id_score_stats want;
        ora_ndjson_id_score((const uint8_t*)withnl.data(), withnl.size(), 1, &want);
        BufSrc src{(const uint8_t*)line.data()};
        fx::NdRecord rec;
        int ok = walk(src, 0, (uint32_t)line.size(), rec);
        // The same record at every alignment in an 8-aligned buffer, with
        // quotes and backslashes after it: the word walker must agree.
        {
            alignas(8) static uint8_t wbuf[1 << 16];
            const uint32_t off = (uint32_t)r.below(8);
            if (off + line.size() + 32 <= sizeof wbuf) {
                std::memcpy(wbuf + off, line.data(), line.size());
                for (int k = 0; k < 32; ++k) wbuf[off + line.size() + k] = "\"\\x7"[r.below(4)];
                fx::NdRecord wrec;
                int wok = walk(WordSrc{wbuf}, off, off + (uint32_t)line.size(), wrec);
                if (wok != ok || (ok == 1 && (wrec.id != rec.id || wrec.fx6 != rec.fx6 || wrec.mant != rec.mant ||
                                             fx::f64_to_bits(wrec.score) != fx::f64_to_bits(rec.score)))) {
                    std::fprintf(stderr, "WORD WALKER MISMATCH on record [%s]\n", line.c_str());
                    return 1;
                }
            }
        }
        uint64_t g_rec = ok == 1, g_inv = ok == 0;
        int64_t g_id = ok == 1 ? rec.id : 0;
        int64_t g_fx = ok == 1 && rec.fx6 ? rec.mant : 0;
        uint64_t g_fc = ok == 1 && rec.fx6;
        double g_s = ok == 1 ? rec.score : 0.0;
        if (g_rec != want.records || g_inv != want.invalid || g_id != want.id_sum ||
            g_fx != want.score_fixed6_sum || g_fc != want.score_fixed6_cells ||
            fx::f64_to_bits(0.0 + g_s) != fx::f64_to_bits(want.score_sum)) {
            std::fprintf(stderr, "MISMATCH on record [%s]: got rec=%" PRIu64 " inv=%" PRIu64 " id=%" PRId64
                                 " want rec=%" PRIu64 " inv=%" PRIu64 " id=%" PRId64 "\n",
                         line.c_str(), g_rec, g_inv, g_id, want.records, want.invalid, want.id_sum);
            return 1;
        }
        valid += g_rec;
        if (common_agrees(r, line, ok, rec) < 0) return 1;
    }
    // The common layout: every plain record is taken by the straight-line
    // reader, and whatever it takes agrees with the walker.
    uint64_t plain_n = 0, taken = 0;
    for (uint64_t i = 0; i < count; ++i) {
        bool plain = false;
        const std::string line = canon(r, plain);
        fx::NdRecord rec;
        const int ok = walk(BufSrc{(const uint8_t*)line.data()}, 0, (uint32_t)line.size(), rec);
        const int c = common_agrees(r, line, ok, rec);
        if (c < 0) return 1;
        if (plain && c != 1) {
            std::fprintf(stderr, "COMMON LAYOUT REFUSED a plain record [%s]\n", line.c_str());
            return 1;
        }
        plain_n += plain;
        taken += (uint64_t)c;
    }
    std::printf("ndjson_check: %" PRIu64 " records agree (%" PRIu64 " valid); common layout: %" PRIu64
                " records, %" PRIu64 " plain, %" PRIu64 " taken by the straight-line reader\n",
                count, valid, count, plain_n, taken);
    return 0;
}
