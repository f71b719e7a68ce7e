// datagen.cpp — seeded synthetic datasets for the benchmark lanes, the
// counterpart of the reference's examples/make_datasets.rs (xorshift64*,
// make_datasets.rs:218-272). Every generator writes whole records into a
// caller buffer until at least `target` bytes and reports generator-side
// truths (exact mantissa sums, bit checksums, counts) that the tests use as
// closed-form expectations.
#include <cmath>
#include <cstdint>
#include <cstring>

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
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Out {
    uint8_t* p;
    uint64_t n, cap;
    bool ok = true;
    void put(uint8_t b) { if (n < cap) p[n++] = b; else ok = false; }
    void put(const char* s) { while (*s) put((uint8_t)*s++); }
    void put_u64(uint64_t v) {
        char buf[24];
        int k = 0;
        do { buf[k++] = (char)('0' + v % 10); v /= 10; } while (v);
        while (k) put((uint8_t)buf[--k]);
    }
    void put_i64(int64_t v) {
        if (v < 0) { put('-'); put_u64((uint64_t)0 - (uint64_t)v); } else put_u64((uint64_t)v);
    }
    // Signed fixed-point with `dec` decimals: m / 10^dec.
    void put_fixed(int64_t m, int dec) {
        uint64_t a = m < 0 ? (uint64_t)0 - (uint64_t)m : (uint64_t)m;
        if (m < 0) put('-');
        uint64_t p10 = 1;
        for (int i = 0; i < dec; ++i) p10 *= 10;
        put_u64(a / p10);
        if (dec > 0) {
            put('.');
            uint64_t f = a % p10;
            char buf[20];
            for (int i = dec - 1; i >= 0; --i) { buf[i] = (char)('0' + f % 10); f /= 10; }
            for (int i = 0; i < dec; ++i) put((uint8_t)buf[i]);
        }
    }
};

inline uint64_t f64_bits(double v) { uint64_t b; std::memcpy(&b, &v, 8); return b; }

}  // namespace

extern "C" {

struct geo_truth {
    uint64_t rows;           // records incl. header
    uint64_t data_rows;
    int64_t lat_fixed6_sum;  // exact sum of signed 1e-6 mantissas
    int64_t lon_fixed6_sum;
    uint64_t lat_checksum;   // wrapping sum of f64 bit patterns
    uint64_t lon_checksum;
    uint64_t city_bytes;     // cleaned bytes over all rows (header "City" included)
    uint64_t city_checksum;  // wrapping byte sum of cleaned City bytes
    uint64_t quoted_cities;
    int64_t id_sum;
};

// csv-geo-style CSV: header `id,City,Latitude,Longitude`, sequential u64 id,
// City of ASCII letters/spaces (length 4-20, 5% double-quoted holding a comma
// or a doubled quote), Latitude U[-90,90] and Longitude U[-180,180] printed
// with six decimals.
uint64_t datagen_geo_csv(uint8_t* buf, uint64_t target, uint64_t cap, uint64_t seed,
                         geo_truth* t) {
    Rng r{seed};
    Out o{buf, 0, cap};
    std::memset(t, 0, sizeof *t);
    o.put("id,City,Latitude,Longitude\n");
    t->rows = 1;
    t->city_bytes = 4;
    t->city_checksum = 'C' + 'i' + 't' + 'y';
    static const char alpha[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz";
    uint64_t id = 1;
    uint8_t city[32];
    while (o.n < target && o.ok) {
        o.put_u64(id);
        t->id_sum += (int64_t)id;
        o.put(',');
        int len = 4 + (int)r.below(17);
        for (int i = 0; i < len; ++i) {
            uint64_t c = r.below(60);
            city[i] = c < 52 ? (uint8_t)alpha[c] : (uint8_t)' ';
        }
        city[0] = (uint8_t)alpha[r.below(52)];
        if (r.below(20) == 0) {
            int at = 1 + (int)r.below((uint64_t)len - 1);
            bool comma = r.below(2) == 0;
            o.put('"');
            for (int i = 0; i < len; ++i) {
                if (i == at) { if (comma) o.put(','); else { o.put('"'); o.put('"'); } }
                o.put(city[i]);
                t->city_checksum += city[i];
            }
            o.put('"');
            t->city_bytes += (uint64_t)len + 1;
            t->city_checksum += comma ? ',' : '"';
            t->quoted_cities += 1;
        } else {
            for (int i = 0; i < len; ++i) { o.put(city[i]); t->city_checksum += city[i]; }
            t->city_bytes += (uint64_t)len;
        }
        o.put(',');
        int64_t lat = (int64_t)r.below(180000001ULL) - 90000000LL;
        int64_t lon = (int64_t)r.below(360000001ULL) - 180000000LL;
        o.put_fixed(lat, 6);
        o.put(',');
        o.put_fixed(lon, 6);
        o.put('\n');
        t->lat_fixed6_sum += lat;
        t->lon_fixed6_sum += lon;
        double lv = (double)(lat < 0 ? -lat : lat) / 1000000.0;
        double ov = (double)(lon < 0 ? -lon : lon) / 1000000.0;
        t->lat_checksum += f64_bits(lv) | ((uint64_t)(lat < 0) << 63);
        t->lon_checksum += f64_bits(ov) | ((uint64_t)(lon < 0) << 63);
        t->rows += 1;
        t->data_rows += 1;
        ++id;
    }
    return o.ok ? o.n : 0;
}

struct hash_truth {
    uint64_t lines;          // all lines, comments included
    uint64_t comment_lines;
    uint64_t records;        // non-comment records incl. header
    uint64_t hazard_fields;
    int64_t i64_sum;         // wrapping sum of the i64 column
};

// csv_hash dialect: header `amount,value,label`; 3% of lines are `#`
// comments holding stray quotes and commas; i64 U(+-2^40); f64 with 1-9
// decimals and an occasional exponent; string length log-uniform 1-64; 14%
// of string fields are quoted, and 4% of the quoted ones hold `#`, `,`, `""`
// and an embedded LF.
uint64_t datagen_hash_csv(uint8_t* buf, uint64_t target, uint64_t cap, uint64_t seed,
                          hash_truth* t) {
    Rng r{seed};
    Out o{buf, 0, cap};
    std::memset(t, 0, sizeof *t);
    o.put("amount,value,label\n");
    t->lines = 1;
    t->records = 1;
    static const char alnum[] = "abcdefghijklmnopqrstuvwxyz0123456789";
    while (o.n < target && o.ok) {
        t->lines += 1;
        if (r.below(100) < 3) {
            o.put("# note ");
            int n = 3 + (int)r.below(20);
            for (int i = 0; i < n; ++i) {
                uint64_t c = r.below(10);
                o.put(c == 0 ? '"' : c == 1 ? ',' : (uint8_t)alnum[r.below(36)]);
            }
            o.put('\n');
            t->comment_lines += 1;
            continue;
        }
        t->records += 1;
        int64_t a = (int64_t)r.below((1ULL << 41) + 1) - (int64_t)(1ULL << 40);
        o.put_i64(a);
        t->i64_sum = (int64_t)((uint64_t)t->i64_sum + (uint64_t)a);
        o.put(',');
        int dec = 1 + (int)r.below(9);
        uint64_t ip = r.below(100000);
        uint64_t p10 = 1;
        for (int i = 0; i < dec; ++i) p10 *= 10;
        int64_t m = (int64_t)(ip * p10 + r.below(p10));
        if (r.below(2)) m = -m;
        o.put_fixed(m, dec);
        if (r.below(20) == 0) {
            o.put(r.below(2) ? 'e' : 'E');
            uint64_t s = r.below(3);
            if (s == 1) o.put('-'); else if (s == 2) o.put('+');
            o.put_u64(r.below(30));
        }
        o.put(',');
        int len = (int)std::floor(std::exp(r.uniform() * std::log(64.0)));
        if (len < 1) len = 1;
        if (len > 64) len = 64;
        const bool quoted = r.below(100) < 14;
        if (quoted && r.below(100) < 4) {
            t->hazard_fields += 1;
            o.put('"');
            for (int i = 0; i < len; ++i) {
                uint64_t c = r.below(12);
                if (c == 0) o.put('#');
                else if (c == 1) o.put(',');
                else if (c == 2) { o.put('"'); o.put('"'); }
                else if (c == 3) o.put('\n');
                else o.put((uint8_t)alnum[r.below(36)]);
            }
            if (len < 4) { o.put('#'); o.put(','); o.put('\n'); o.put('"'); o.put('"'); }
            o.put('"');
        } else if (quoted) {
            o.put('"');
            for (int i = 0; i < len; ++i) o.put((uint8_t)alnum[r.below(36)]);
            o.put('"');
        } else {
            for (int i = 0; i < len; ++i) o.put((uint8_t)alnum[r.below(36)]);
        }
        o.put('\n');
    }
    return o.ok ? o.n : 0;
}

struct ndjson_truth {
    uint64_t records;
    int64_t id_sum;
    int64_t score_fixed6_sum;
    uint64_t escaped_names;
};

static void ws(Rng& r, Out& o) {
    uint64_t k = r.below(10);
    if (k == 0) o.put(' ');
    else if (k == 1) { o.put(' '); o.put(' '); }
}

// NDJSON: {"id":u64<2^40,"name":str(4-32, 2% with \" and \\),
// "nested":{"score":f64 in [0,1000) with 1-6 decimals,"tags":[0-5 strings]}}
// with random whitespace around keys and separators.
uint64_t datagen_ndjson(uint8_t* buf, uint64_t target, uint64_t cap, uint64_t seed,
                        ndjson_truth* t) {
    Rng r{seed};
    Out o{buf, 0, cap};
    std::memset(t, 0, sizeof *t);
    static const char alnum[] = "abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ0123456789 ";
    while (o.n < target && o.ok) {
        uint64_t id = r.below(1ULL << 40);
        o.put('{'); ws(r, o);
        o.put("\"id\""); ws(r, o); o.put(':'); ws(r, o);
        o.put_u64(id);
        ws(r, o); o.put(','); ws(r, o);
        o.put("\"name\""); ws(r, o); o.put(':'); ws(r, o);
        o.put('"');
        int len = 4 + (int)r.below(29);
        bool esc = r.below(50) == 0;
        int at = esc ? (int)r.below((uint64_t)len) : -1;
        for (int i = 0; i < len; ++i) {
            if (i == at) {
                if (r.below(2)) { o.put('\\'); o.put('"'); } else { o.put('\\'); o.put('\\'); }
            }
            o.put((uint8_t)alnum[r.below(63)]);
        }
        o.put('"');
        if (esc) t->escaped_names += 1;
        ws(r, o); o.put(','); ws(r, o);
        o.put("\"nested\""); ws(r, o); o.put(':'); ws(r, o);
        o.put('{'); ws(r, o);
        o.put("\"score\""); ws(r, o); o.put(':'); ws(r, o);
        int dec = 1 + (int)r.below(6);
        uint64_t p10 = 1;
        for (int i = 0; i < dec; ++i) p10 *= 10;
        uint64_t m = r.below(1000) * p10 + r.below(p10);
        o.put_fixed((int64_t)m, dec);
        uint64_t scale = 1;
        for (int i = dec; i < 6; ++i) scale *= 10;
        ws(r, o); o.put(','); ws(r, o);
        o.put("\"tags\""); ws(r, o); o.put(':'); ws(r, o);
        o.put('[');
        int tags = (int)r.below(6);
        for (int k = 0; k < tags; ++k) {
            if (k) { o.put(','); ws(r, o); }
            o.put('"');
            int tl = 1 + (int)r.below(8);
            for (int i = 0; i < tl; ++i) o.put((uint8_t)alnum[r.below(26)]);
            o.put('"');
        }
        o.put(']');
        ws(r, o); o.put('}'); ws(r, o);
        o.put('}');
        o.put('\n');
        t->records += 1;
        t->id_sum += (int64_t)id;
        t->score_fixed6_sum += (int64_t)(m * scale);
    }
    return o.ok ? o.n : 0;
}

struct tsv_truth {
    uint64_t rows;
    int64_t id_sum;
};

// TSV variant of the served mix: `id\tname\tvalue` (unquoted).
uint64_t datagen_tsv(uint8_t* buf, uint64_t target, uint64_t cap, uint64_t seed, tsv_truth* t) {
    Rng r{seed};
    Out o{buf, 0, cap};
    std::memset(t, 0, sizeof *t);
    static const char alnum[] = "abcdefghijklmnopqrstuvwxyz0123456789";
    uint64_t id = 1;
    while (o.n < target && o.ok) {
        o.put_u64(id);
        o.put('\t');
        int len = 3 + (int)r.below(20);
        for (int i = 0; i < len; ++i) o.put((uint8_t)alnum[r.below(36)]);
        o.put('\t');
        o.put_fixed((int64_t)r.below(100000000), 3);
        o.put('\n');
        t->rows += 1;
        t->id_sum += (int64_t)id;
        ++id;
    }
    return o.ok ? o.n : 0;
}

struct logfmt_truth {
    uint64_t lines;
    uint64_t pairs;        // key=value pairs written
    uint64_t key_bytes;    // key lengths
    uint64_t value_bytes;  // value lengths after quote stripping and unescaping
    uint64_t checksum;     // wrapping byte sum of the keys and the unescaped values
};

// logfmt `k=v` lines with quoted values holding escaped quotes/backslashes.
// The truths count what each pair decodes to: a quoted value without its
// quotes, and each `\x` of it as the one byte x.
uint64_t datagen_logfmt(uint8_t* buf, uint64_t target, uint64_t cap, uint64_t seed,
                        logfmt_truth* t) {
    Rng r{seed};
    Out o{buf, 0, cap};
    std::memset(t, 0, sizeof *t);
    static const char* levels[] = {"debug", "info", "warn", "error"};
    static const char alnum[] = "abcdefghijklmnopqrstuvwxyz0123456789";
    // Byte sum and length of what was written since `from`.
    auto since = [&](uint64_t from, uint64_t& bytes, uint64_t& sum) {
        for (uint64_t i = from; i < o.n; ++i) sum += o.p[i];
        bytes += o.n - from;
    };
    // A pair whose key is `k` and whose (unquoted) value is written by `value`.
    auto pair = [&](const char* k, auto&& value) {
        const uint64_t k0 = o.n;
        o.put(k);
        since(k0, t->key_bytes, t->checksum);
        o.put('=');
        const uint64_t v0 = o.n;
        value();
        since(v0, t->value_bytes, t->checksum);
        t->pairs += 1;
    };
    while (o.n < target && o.ok) {
        pair("ts", [&] { o.put_u64(1700000000 + r.below(10000000)); });
        o.put(' ');
        pair("level", [&] { o.put(levels[r.below(4)]); });
        o.put(" msg=\"");
        t->key_bytes += 3;
        t->checksum += 'm' + 's' + 'g';
        int len = 4 + (int)r.below(40);
        for (int i = 0; i < len; ++i) {
            uint64_t c = r.below(40);
            uint8_t b;
            if (c == 0) { o.put('\\'); b = '"'; }
            else if (c == 1) { o.put('\\'); b = '\\'; }
            else if (c < 6) b = ' ';
            else if (c == 6) b = '=';
            else b = (uint8_t)alnum[r.below(36)];
            o.put(b);
            t->value_bytes += 1;
            t->checksum += b;
        }
        o.put('"');
        t->pairs += 1;
        o.put(' ');
        pair("dur", [&] {
            o.put_fixed((int64_t)r.below(100000), 2);
            o.put("ms");
        });
        if (r.below(4) == 0) {
            o.put(' ');
            pair("user", [&] { for (int i = 0; i < 6; ++i) o.put((uint8_t)alnum[r.below(36)]); });
        }
        o.put('\n');
        t->lines += 1;
    }
    return o.ok ? o.n : 0;
}

}  // extern "C"
