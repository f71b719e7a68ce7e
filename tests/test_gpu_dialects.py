"""GPU parity beyond the doubled-quote CSV dialects: every dialect kind's
column, statistics and field-byte passes against the oracle (no quote,
comment only, backslash escape, backslash escape with comments), the field
byte and logfmt sinks against their independent pins (records() walk, hand
answers, generator truths), byte-by-byte seam sweeps around the 8 KiB tile
edge, NDJSON records longer than a tile, and the C ABI's error statuses."""
import ctypes as C

import numpy as np
import pytest

from falx_amd import datagen, gpu
from falx_amd import spec as S
from helpers import Rng, assert_columns_equal, assert_stats_equal

pytestmark = pytest.mark.gpu

TILE = 8192

# One typed projection per dialect kind (kinds 0-5 of the kernels).
TSV_TYPED = S.TSV_TYPED                                                          # kind 0
CMT_TYPED = S.Spec("comment_tsv_typed", b"\t\n", None, None, ord("#"),
                   (S.Column(0, S.I64), S.Column(1, S.F64), S.Column(2, S.STR), S.Column(3, S.BYTES)))  # kind 1
CSV_TYPED4 = S.CSV.with_columns([(0, S.I64), (1, S.F64), (2, S.STR), (3, S.BYTES)], "csv_typed4")     # kind 2
BS_TYPED = S.Spec("backslash_csv_typed", b",\n", ord('"'), ord("\\"), None,
                  (S.Column(0, S.I64), S.Column(1, S.F64), S.Column(2, S.STR), S.Column(3, S.STR)))   # kind 3
HASH4 = S.CSV_HASH_DIALECT.with_columns([(0, S.I64), (1, S.F64), (2, S.STR), (3, S.BYTES)], "hash4")  # kind 4
BS_CMT_TYPED = S.Spec("backslash_comment_typed", b",\n", ord('"'), ord("\\"), ord("#"),
                      (S.Column(0, S.I64), S.Column(1, S.F64), S.Column(2, S.STR), S.Column(3, S.STR)))  # kind 5
KIND_SPECS = [TSV_TYPED, CMT_TYPED, CSV_TYPED4, BS_TYPED, HASH4, BS_CMT_TYPED]
# Numeric-only projections of the kinds with neither escape nor comment (the
# numeric fast forms), with a third column so statistics take a second pass.
TSV_NUM = S.TSV.with_columns([(0, S.I64), (1, S.F64), (4, S.F64)], "tsv_num")                          # kind 0
CSV_NUM = S.CSV.with_columns([(4, S.I64), (1, S.F64), (0, S.I64)], "csv_num")                          # kind 2
NUM_SPECS = [TSV_NUM, CSV_NUM]
LOGFMT_ALPHA = b'ab= =  ""\\\\\r\n\n'


def _digits(rng, n):
    return bytes(48 + rng.below(10) for _ in range(n))


def _number(rng) -> bytes:
    k = rng.below(8)
    s = b"-" if rng.below(4) == 0 else b""
    if k < 3:
        return s + _digits(rng, 1 + rng.below(4)) + b"." + _digits(rng, rng.below(10))
    if k < 5:
        return s + _digits(rng, 1 + rng.below(19))
    if k == 5:
        return s + _digits(rng, 1 + rng.below(5)) + b"e" + (b"-" if rng.below(2) else b"") + _digits(rng, 1 + rng.below(3))
    if k == 6:
        return [b"inf", b"nan", b"", b"12x", b"+7", b"0007", b"9223372036854775808"][rng.below(7)]
    return s + b"." + _digits(rng, 1 + rng.below(8))


def _text(rng, spec, n) -> bytes:
    """A string cell holding the dialect's hazards: separators, newlines and
    comment bytes inside quotes, escaped or doubled quotes, escape runs."""
    sep = bytes([spec.structural[0]])
    q = spec.quote
    plain = bytes(97 + rng.below(26) for _ in range(n))
    if q is None:
        return plain
    if rng.below(3):
        return plain
    body = bytearray()
    for _ in range(n):
        k = rng.below(9)
        if k == 0:
            body += sep
        elif k == 1:
            body += b"\n"
        elif k == 2 and spec.comment is not None:
            body.append(spec.comment)
        elif k == 3:
            body += bytes([spec.escape, q]) if spec.escape is not None else bytes([q, q])
        elif k == 4 and spec.escape is not None:
            body += bytes([spec.escape]) * (1 + rng.below(4)) + (b"x" if rng.below(2) else b"")
        else:
            body.append(97 + rng.below(26))
    if spec.escape is not None and len(body) and body[-1] == spec.escape:
        # An odd escape run before the closing quote would escape it.
        run = len(body) - len(bytes(body).rstrip(bytes([spec.escape])))
        if run % 2:
            body.append(spec.escape)
    return bytes([q]) + bytes(body) + bytes([q])


def dialect_doc(rng: Rng, spec: S.Spec, records: int) -> bytes:
    """Typed records for `spec` (i64, f64, string, string/bytes fields), with
    comment lines carrying stray quotes and separators, empty lines, CRLF,
    short and long records, and sometimes no final newline."""
    sep = bytes([spec.structural[0]])
    out = bytearray()
    for _ in range(records):
        if spec.comment is not None and rng.below(20) == 0:
            junk = bytes(b'x ,\t"\\'[rng.below(6)] for _ in range(rng.below(30)))
            out += bytes([spec.comment]) + junk + b"\n"
            continue
        if rng.below(40) == 0:
            out += b"\n"
            continue
        nf = 1 + rng.below(6)
        cells = [_number(rng), _number(rng), _text(rng, spec, rng.below(20)), _text(rng, spec, rng.below(12)),
                 _number(rng), _text(rng, spec, 3)][:nf]
        out += sep.join(cells) + (b"\r\n" if rng.below(15) == 0 else b"\n")
    if rng.below(3) == 0 and out.endswith(b"\n"):
        out = out[:-1]
    return bytes(out)


def _check_all(ctx, ora, spec, doc, label):
    buf = ctx.upload(doc)
    n = len(doc)
    want = ora.parse_columns(spec, doc)
    assert_columns_equal(ctx.parse_columns(spec, buf, n), want, spec, label)
    want_st = ora.csv_stats(spec, doc)
    assert_stats_equal(ctx.csv_stats(spec, buf, n), want_st, spec, label + " stats")
    cols, _keep = ctx.alloc_columns(spec, want["rows"], max(16, n))
    assert_stats_equal(ctx.parse_columns_with_stats_into(spec, buf, n, cols), want_st, spec, label + " fused")
    assert ctx.field_bytes(spec, buf, n) == ora.field_bytes(spec, doc), label + " field bytes"
    assert np.array_equal(ctx.index_structurals(spec, buf, n), ora.index_structurals(spec, doc)), label + " index"


@pytest.mark.parametrize("spec", KIND_SPECS + NUM_SPECS, ids=lambda s: s.name)
def test_dialect_kinds_columns_stats_field_bytes(ctx, ora, spec):
    """Columns, statistics (separate and fused), field bytes and the index of
    every dialect kind against the oracle: small documents and ones of ~40
    tiles."""
    rng = Rng(0xD1A1_EC70 + len(spec.name))
    for i in range(40):
        doc = dialect_doc(rng, spec, 1 + rng.below(30))
        if doc:
            _check_all(ctx, ora, spec, doc, f"{spec.name} small {i}")
    for i in range(3):
        doc = dialect_doc(rng, spec, 6000 + rng.below(3000))
        _check_all(ctx, ora, spec, doc, f"{spec.name} long {i}")


def test_tsv_generator_columns_and_truths(ctx, ora):
    """The served mix's TSV (`id\\tname\\tvalue`) against the oracle and the
    generator's row count and id sum."""
    data, truth = datagen.tsv(3 * 1024 * 1024 + 17)
    _check_all(ctx, ora, S.TSV_TYPED, data, "tsv generated")
    st = ctx.csv_stats(S.TSV_TYPED, ctx.upload(data), data.size)
    assert st["records"] == truth["rows"] and st["cols"][0]["checksum"] == truth["id_sum"] % 2**64


def test_backslash_string_columns_on_logfmt(ctx, ora):
    """Backslash-escaped string cells through the GPU string arena: logfmt
    keys and (quoted, escaped) values as string columns."""
    spec = S.LOGFMT.with_columns([(0, S.STR), (1, S.STR), (3, S.STR), (5, S.STR), (7, S.STR), (1, S.I64)],
                                 "logfmt_strings")
    data, _ = datagen.logfmt(2 * 1024 * 1024 + 9)
    _check_all(ctx, ora, spec, data, "logfmt strings")
    rng = Rng(0x10F3_57A1)
    for i in range(60):
        doc = bytes(LOGFMT_ALPHA[rng.below(len(LOGFMT_ALPHA))] for _ in range(rng.below(200)))
        if doc:
            _check_all(ctx, ora, spec, doc, f"logfmt hazards {i}")


# ---- field bytes and logfmt against their independent pins ---------------
def test_field_bytes_records_reference(ctx, ora):
    """csv_hash_fused_par.rs:186-233 on the GPU: the 12 hazard cases and the
    120 random documents, against the records() walk and the oracle."""
    from test_oracle import FIELD_BYTE_HAND_CASES, field_byte_random_docs, records_field_bytes
    spec = S.CSV_HASH_DIALECT
    for i, doc in enumerate(FIELD_BYTE_HAND_CASES + field_byte_random_docs()):
        got = ctx.field_bytes(spec, ctx.upload(doc), len(doc))
        assert (got["fields"], got["bytes"]) == records_field_bytes(ora, spec, doc), f"doc {i} {doc[:40]!r}"


def test_logfmt_pins(ctx, ora):
    """Hand answers, random hazard documents (records() walk) and the
    generator's truths for falx_parse_logfmt_pairs."""
    from test_oracle import records_logfmt
    hand = [b'a="x \\" y" b=plain\n', b"k=v\n\nk2=\n", b"lonely\n", b"a=1 b\r\n", b'x="q\\\\z" y="" z=w\n',
            b'k="abc', b"a=1 b=2", b"a=1 ", b"a=1 b=2 c\n", b"p=a\\b\n", b'q="\\\\"\n', b'e="a=b c" f=1\n']
    rng = Rng(0x10F3_A7C0_FFEE_0002)
    docs = hand + [bytes(LOGFMT_ALPHA[rng.below(len(LOGFMT_ALPHA))] for _ in range(rng.below(300)))
                   for _ in range(200)]
    for i, doc in enumerate(d for d in docs if d):
        got = ctx.logfmt_pairs(ctx.upload(doc), len(doc))
        assert got == records_logfmt(ora, doc) == ora.logfmt_pairs(doc), f"doc {i} {doc!r}"
    data, t = datagen.logfmt(5 * 1024 * 1024 + 1)
    got = ctx.logfmt_pairs(ctx.upload(data), data.size)
    assert got == {"pairs": t["pairs"], "key_bytes": t["key_bytes"], "value_bytes": t["value_bytes"],
                   "checksum": t["checksum"] % 2**64}


# ---- tile seams, byte by byte ------------------------------------------------
SEAM_HAZARDS = [b'"x\ny",7\n', b'# c " c\n', b'"#",7\n', b'"a""b",7\n', b'"' + b"q,\n" * 40 + b'",1\n',
                b"\\\\\\\"" + b'"a\\"b",2\n']


@pytest.mark.parametrize("spec", [S.CSV, S.CSV_HASH_DIALECT, S.LOGFMT, S.NDJSON], ids=lambda s: s.name)
def test_index_hazards_every_seam_offset(ctx, ora, spec):
    """Each hazard starts at every byte offset from 70 before to 6 after the
    8 KiB tile seam (csv_hash_fused_par.rs:88-118 at tile scale)."""
    for hazard in SEAM_HAZARDS:
        for pad in range(TILE - 70, TILE + 6):
            doc = b"k" * pad + b"\n" + hazard + b"tail,99\n" + hazard
            got = ctx.index_structurals(spec, ctx.upload(doc), len(doc))
            assert np.array_equal(got, ora.index_structurals(spec, doc)), f"{spec.name} pad {pad} {hazard!r}"


def _geo_prefix(n: int) -> bytes:
    """Geo rows of exactly n bytes (the last row's City padded to fit)."""
    out = bytearray(b"id,City,Latitude,Longitude\n")
    i = 1
    while len(out) < n:
        row = b"%d,Cityname,%d.123456,-%d.654321\n" % (i, i % 90, i % 180)
        if len(out) + len(row) > n:
            room = n - len(out)
            row = b"%d,C,1.000000,2.000000\n" % i
            if room < len(row):
                return bytes(out) + b"x" * (room - 1) + b"\n" if room > 0 else bytes(out)
            row = row[:row.index(b",C,") + 2] + b"C" * (room - len(row)) + row[row.index(b",C,") + 2:]
        out += row
        i += 1
    return bytes(out)


SEAM_RECORDS = [b'777,"Q,x",-12.5,"33.25"\n', b'778,"a""b",1e3,-0.000001\n', b"779,,,\n",
                b'780,"multi\nline",12.345678,"-1.5"\r\n', b"781,Z,123.456789,-179.999999\n"]


@pytest.mark.parametrize("spec", [S.GEO_ALL, S.GEO_LATLON], ids=lambda s: s.name)
def test_columns_every_seam_offset(ctx, ora, spec):
    """Numeric, quoted numeric, empty and quoted string cells crossing the
    tile seam at every byte offset (the cells read through the 64-byte LDS
    halo on each side of a tile); the rest of the document is fixed-6 geo
    rows, so the fast pass's own output is what is compared."""
    suffix = b"".join(b"%d,Tail,%d.500000,%d.250000\n" % (900 + k, k, k) for k in range(300))
    for rec in SEAM_RECORDS:
        for pad in range(TILE - 80, TILE + 16):
            doc = _geo_prefix(pad) + rec + suffix
            buf = ctx.upload(doc)
            want = ora.parse_columns(spec, doc)
            cols, keep = ctx.alloc_columns(spec, want["rows"], len(doc))
            st = ctx.parse_columns_with_stats_into(spec, buf, len(doc), cols)
            assert_stats_equal(st, ora.csv_stats(spec, doc), spec, f"pad {pad} {rec!r}")
            assert_columns_equal(ctx.parse_columns(spec, buf, len(doc)), want, spec, f"pad {pad} {rec!r}")


TSV_OPEN = S.TSV.with_columns([(0, S.I64), (2, S.F64), (3, S.F64)], "tsv_open")                      # kind 0


def _open_record_doc(rng: Rng, lo: int, hi: int, sep: bytes = b",") -> bytes:
    """Geo-schema rows whose second field is lo..hi bytes: with sep ',' half
    of them quoted and holding commas, newlines and doubled quotes; with a
    tab, plain text. Tiles open inside such records."""
    quoted = sep == b","
    out = bytearray(sep.join([b"id", b"City", b"Latitude", b"Longitude"]) + b"\n")
    k = 0
    while len(out) < 3 * TILE + 777:
        n = lo + rng.below(hi - lo + 1)
        if quoted and rng.below(2):
            body = bytearray()
            while len(body) < n:
                t = rng.below(12)
                body += b"," if t == 0 else (b"\n" if t == 1 else (b'""' if t == 2 else bytes([97 + rng.below(26)])))
            city = b'"' + bytes(body) + b'"'
        else:
            city = bytes(97 + rng.below(26) if rng.below(8) else 32 for _ in range(n))
        lat = b"%d.%06d" % (rng.below(90), rng.below(10 ** 6))
        lon = b"-%d.%06d" % (rng.below(180), rng.below(10 ** 6))
        out += sep.join([b"%d" % k, city, lat, lon]) + b"\n"
        k += 1
    return bytes(out)


@pytest.mark.parametrize("lo,hi", [(20, 200), (300, 2500)], ids=["within_reach", "beyond_reach"])
def test_numeric_columns_open_records(ctx, ora, lo, hi):
    """Numeric columns of records open across tile starts: quoted fields with
    commas, newlines and doubled quotes just before the seam (CSV, kind 2),
    plain long fields (TSV, kind 0), records of up to 256 bytes and records
    reaching several KiB back. Columns and fused statistics against the
    oracle."""
    rng = Rng(0x0BE2 + lo)
    for spec, sep in ((S.GEO_LATLON, b","), (TSV_OPEN, b"\t")):
        for i in range(6):
            doc = _open_record_doc(rng, lo, hi, sep)
            buf = ctx.upload(doc)
            want = ora.parse_columns(spec, doc)
            assert_columns_equal(ctx.parse_columns(spec, buf, len(doc)), want, spec, f"doc {i}")
            cols, keep = ctx.alloc_columns(spec, want["rows"], len(doc))
            st = ctx.parse_columns_with_stats_into(spec, buf, len(doc), cols)
            assert_stats_equal(st, ora.csv_stats(spec, doc), spec, f"doc {i}")


# ---- NDJSON records longer than a tile ---------------------------------------
def _nd_record(rng: Rng, name_len: int) -> bytes:
    name = bytearray()
    while len(name) < name_len:
        k = rng.below(50)
        name += b'\\"' if k == 0 else (b"\\\\" if k == 1 else bytes([97 + rng.below(26)]))
    return (b'{"id":%d,"name":"%s","nested":{"score":%d.%02d,"tags":["a","b"]}}' %
            (rng.below(1 << 40), bytes(name), rng.below(1000), rng.below(100)))


def test_ndjson_records_longer_than_a_tile(ctx, ora):
    """Records of 20-100 KiB spanning several tiles, mixed with short ones;
    a document that is one newline-free record of 3 MiB; and one whose last
    long record has no newline."""
    rng = Rng(0x0B16_2EC0)
    recs = []
    for i in range(80):
        recs.append(_nd_record(rng, 20_000 + rng.below(80_000) if i % 3 == 0 else rng.below(60)))
    docs = [b"\n".join(recs) + b"\n", b"\n".join(recs), _nd_record(rng, 3 << 20),
            b"\n" * 5 + _nd_record(rng, 50_000) + b"\n\n" + _nd_record(rng, 9000)]
    for i, doc in enumerate(docs):
        got = ctx.ndjson_id_score(ctx.upload(doc), len(doc))
        want = ora.ndjson_id_score(doc)
        for k in ("records", "invalid", "id_sum", "score_fixed6_sum", "score_fixed6_cells"):
            assert got[k] == want[k], f"doc {i}: {k} {got[k]} != {want[k]}"
        assert got["score_sum"] == pytest.approx(want["score_sum"], rel=1e-12, abs=1e-9)
        assert ctx.ndjson_lines(ctx.upload(doc), len(doc)) == ora.ndjson_lines(doc)


def _nd_common(rng, mut):
    """A record in the common layout {"id","name","nested":{"score","tags"}}
    with spaces, or (mut >= 0) one part changed to a form the kernel's
    straight-line reader must hand to the walker or still read right."""
    def sp():
        k = rng.below(10)
        return b" " if k == 0 else (b"  " if k == 1 else b"")
    ws = [sp() for _ in range(10)]
    if mut == 0:
        ws[3] = b"\t"
    elif mut == 1:
        ws[7] = b" \r"
    elif mut == 2:
        ws[5] = b" " * 8
    ident = b"%d" % rng.below(1 << 40)
    ident = {3: b"-" + ident, 4: b"000" + ident, 5: b"7" * 16, 6: b"123456789012345", 7: b"-", 8: b"12x"}.get(mut, ident)
    name = bytearray(b'"')
    for _ in range(4 + rng.below(29)):
        k = rng.below(40)
        name += b'\\"' if k == 0 else (b"\\\\" if k == 1 else (b"[{" if k == 2 else bytes([97 + rng.below(26)])))
    name = {9: b"12", 10: b'{"a":[1,{"b":"}"}],"c":"]"}', 11: b'"abc\\"'}.get(mut, bytes(name + b'"'))
    score = b"%d." % rng.below(1000) + _digits(rng, 1 + rng.below(6))
    score = {12: score + b"e%d" % rng.below(20), 13: b"012.5", 14: b"0.25", 15: b".5", 16: b"5.",
             17: b"123456789.1234567", 18: b"-" + score, 19: b"%d" % rng.below(100000),
             20: b"1234567890123.5"}.get(mut, score)
    tags = b"[" + b",".join(b'"' + bytes(97 + rng.below(26) for _ in range(1 + rng.below(8))) + b'"'
                            for _ in range(rng.below(6))) + b"]"
    tags = {21: b'[[1, 2], {"x": ["]"]}, "y\\\\"]', 22: b'"t"'}.get(mut, tags)
    k_id = b'"Id"' if mut == 23 else b'"id"'
    k_score = b'"scores"' if mut == 24 else b'"score"'
    k_name = b'"id"' if mut == 25 else b'"name"'
    s = (ws[0] + b"{" + ws[1] + k_id + ws[2] + b":" + ws[3] + ident + ws[4] + b"," + ws[5] + k_name + ws[6] + b":" +
         ws[7] + name + ws[8] + b"," + ws[9] + b'"nested"' + sp() + b":" + sp() + b"{" + sp() + k_score + sp() +
         b":" + sp() + score + sp() + b"," + sp() + b'"tags"' + sp() + b":" + sp() + tags + sp() + b"}" + sp() + b"}")
    if mut == 26:
        s += b" trailing"
    elif mut == 27:
        s = s[:rng.below(len(s) + 1)]
    elif mut == 28:
        s = s[:-1]
    return s


def test_ndjson_common_layout_and_near_misses(ctx, ora):
    """Documents of records in the common layout, a quarter of them changed
    in one part (whitespace other than spaces, signs, leading zeros, 15 and 16
    digit numbers, exponents, nested tags, other keys, truncation): the
    kernel reads the plain ones in straight-line steps and hands the others
    to its record walker, and every field agrees with the oracle. A mutation
    class per document, so a wrong reading of one class shows in its sums."""
    rng = Rng(0xC0_4417_0A7E)
    for mut in range(-1, 29):
        recs = [_nd_common(rng, mut if rng.below(4) == 0 else -1) for _ in range(700)]
        doc = b"\n".join(recs) + (b"\n" if mut % 2 else b"")
        got = ctx.ndjson_id_score(ctx.upload(doc), len(doc))
        want = ora.ndjson_id_score(doc)
        for k in ("records", "invalid", "id_sum", "score_fixed6_sum", "score_fixed6_cells"):
            assert got[k] == want[k], f"mutation {mut}: {k} {got[k]} != {want[k]}"
        assert got["score_sum"] == pytest.approx(want["score_sum"], rel=1e-12, abs=1e-9), f"mutation {mut}"


# ---- error statuses through the C ABI ----------------------------------------
def test_error_statuses(ctx, ora):
    """FALX_E_TOO_LARGE for inputs of 2^32 bytes or more (checked before any
    device access), FALX_E_INVALID for null or misaligned inputs and bad
    specs, FALX_E_SCHEMA for NDJSON records without the schema keys (with
    the invalid count equal to the oracle's)."""
    L = gpu.load()
    buf = ctx.upload(b"a,b\n" * 16)
    cs = S.CSV.to_c()
    cnt = C.c_uint64()
    big = 1 << 32
    assert L.falx_index_structurals(ctx.h, C.byref(cs), buf.ptr, big, buf.ptr, 16, C.byref(cnt)) == S.E_TOO_LARGE
    st = S.CCsvStats()
    assert L.falx_parse_csv_stats(ctx.h, C.byref(S.GEO_LATLON.to_c()), buf.ptr, big, C.byref(st)) == S.E_TOO_LARGE
    nd = S.CNdjsonIdScore()
    assert L.falx_parse_ndjson_id_score(ctx.h, buf.ptr, big, C.byref(nd)) == S.E_TOO_LARGE
    assert L.falx_index_structurals(ctx.h, C.byref(cs), None, 64, buf.ptr, 16, C.byref(cnt)) == S.E_INVALID
    assert L.falx_index_structurals(ctx.h, C.byref(cs), buf.ptr + 1, 32, buf.ptr, 16, C.byref(cnt)) == S.E_INVALID
    bad = S.Spec("bad", b",\n", ord('"'), ord('"'), None).to_c()     # escape == quote
    assert L.falx_index_structurals(ctx.h, C.byref(bad), buf.ptr, 64, buf.ptr, 16, C.byref(cnt)) == S.E_INVALID
    malformed = [b'{"id":1,"nested":{"score":2}}\n{"id":2}\n{"nested":{"score":1}}\n{"id":3,"nested":{"score":x}}\n'
                 b'not json\n{"id":4,"nested":{"score":1.5}}\n   \n{"id":5 "nested":{"score":1}}\n',
                 b'{"id":1,"nested":{"score":2}}', b'{"id":"7","nested":{"score":2}}\n{}\n[1,2]\n']
    for doc in malformed:
        b = ctx.upload(doc)
        r = S.CNdjsonIdScore()
        status = L.falx_parse_ndjson_id_score(ctx.h, b.ptr, len(doc), C.byref(r))
        want = ora.ndjson_id_score(doc)
        assert status == (S.E_SCHEMA if want["invalid"] else S.OK), doc
        assert (r.records, r.invalid, r.id_sum) == (want["records"], want["invalid"], want["id_sum"]), doc


@pytest.mark.parametrize("spec", NUM_SPECS, ids=lambda s: s.name)
def test_numeric_fast_forms_crowded_tiles(ctx, ora, spec):
    """Tiles with more records (over 1024) or structurals (over 7168) than
    the numeric fast forms list go to the full pass; runs of them between
    ordinary tiles, records crossing the seams, and no final newline."""
    sep = bytes([spec.structural[0]])
    rng = Rng(0xC20D_7ED + len(spec.name))
    for i in range(6):
        parts = []
        for _ in range(3 + rng.below(4)):
            k = rng.below(3)
            if k == 0:
                parts.append(b"".join(b"%d\n" % rng.below(10) for _ in range(2000 + rng.below(6000))))
            elif k == 1:
                parts.append((sep * 4 + b"\n") * (1500 + rng.below(3000)))
            else:
                parts.append(dialect_doc(rng, spec, 200 + rng.below(400)))
        doc = b"".join(parts)
        if i % 2:
            doc = doc.rstrip(b"\n")
        _check_all(ctx, ora, spec, doc, f"{spec.name} crowded {i}")
