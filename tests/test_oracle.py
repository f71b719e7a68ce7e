"""CPU tests of the oracle (the checker the GPU parity tests trust).

Pins: the reference's known-answer tests and fixtures, brute force over tiny
inputs, scalar (scalar.rs) vs block (generated-kernel) semantics, parallel
drivers == serial, f64 goldens from an independent correctly rounded
conversion, and generator closed forms."""
import itertools
import json
import os

import numpy as np
import pytest

from falx_amd import datagen
from falx_amd import spec as S
from falx_amd.records import Parsed
from falx_amd.records import clean as clean_field
from helpers import Rng, assert_columns_equal, assert_stats_equal, dialect_alphabet, random_doc

HERE = os.path.dirname(os.path.abspath(__file__))
SIX = [S.CSV, S.TSV, S.LOGFMT, S.NDJSON, S.MULTI, S.CSV_HASH_DIALECT]


def test_ir_interpreter_matches_scalar(ora):
    """lib.rs:51-67: the block (graph) semantics equal the scalar reference,
    2000 random docs over `",\\n\\rxy`."""
    rng = Rng(0xD1B54A32D192ED03)
    for _ in range(2000):
        doc = random_doc(rng, b'",\n\rxy', rng.below(300))
        assert np.array_equal(ora.index_structurals(S.CSV, doc), ora.index_scalar(S.CSV, doc)), doc


@pytest.mark.parametrize("spec", SIX, ids=lambda s: s.name)
def test_generated_semantics_differential(ora, spec):
    """codegen.rs:391-468: 800 random docs per dialect, blocks == scalar."""
    rng = Rng(0xDEADBEEFCAFEBABE)
    alpha = dialect_alphabet(spec)
    for _ in range(800):
        doc = random_doc(rng, alpha, rng.below(300))
        assert np.array_equal(ora.index_structurals(spec, doc), ora.index_scalar(spec, doc)), doc


@pytest.mark.parametrize("spec,alpha,n", [(S.CSV_HASH_DIALECT, b'",\n#x', 7), (S.NDJSON, b'\\"\nx', 8),
                                          (S.LOGFMT, b'\\" =\n', 7), (S.CSV, b'",\nx', 8)],
                         ids=["csv_hash", "ndjson", "logfmt", "csv"])
def test_brute_force_tiny_inputs(ora, spec, alpha, n):
    """Every string up to length n over the dialect's hazard alphabet."""
    for length in range(n + 1):
        for tup in itertools.product(alpha, repeat=length):
            doc = bytes(tup)
            assert np.array_equal(ora.index_structurals(spec, doc), ora.index_scalar(spec, doc)), doc


def test_block_seams_and_brute_force_long_runs(ora):
    """Escape runs and quoted regions across 64-byte block seams
    (tests/ir.rs:24-120, 169-212)."""
    for pad in range(50, 80):
        for run in range(0, 6):
            doc = b"x" * pad + b"\\" * run + b'"a,b\n"' + b"\\" * run + b'"\n'
            assert np.array_equal(ora.index_structurals(S.NDJSON, doc), ora.index_scalar(S.NDJSON, doc))
            doc2 = b"x" * pad + b'"' + b"#," * run + b'\n"#c\n'
            assert np.array_equal(ora.index_structurals(S.CSV_HASH_DIALECT, doc2),
                                  ora.index_scalar(S.CSV_HASH_DIALECT, doc2))


@pytest.mark.parametrize("spec", [S.CSV, S.TSV, S.CSV_HASH_DIALECT, S.MULTI], ids=lambda s: s.name)
def test_parallel_drivers_match_serial(ora, spec):
    """codegen.rs:581-609, spans.rs:367-394, comments.rs:59-126: par == serial
    at thread counts 1,2,3,7,16,48 (more threads than chunks included)."""
    rng = Rng(0x0DDBA1150DDBA115)
    alpha = dialect_alphabet(spec) + b"\rxy"
    for _ in range(6):
        doc = random_doc(rng, alpha, 4096 + rng.below(60_000))
        s_idx = ora.index_structurals(spec, doc)
        s_tape = ora.parse(spec, doc)
        for t in (2, 3, 7, 16, 48):
            assert np.array_equal(ora.index_structurals(spec, doc, t), s_idx), t
            p = ora.parse(spec, doc, t)
            assert np.array_equal(p[0], s_tape[0]) and np.array_equal(p[1], s_tape[1]), t


def test_columns_and_sinks_par_match_serial(ora):
    """columns.rs:694-827 and csv_hash_fused_par.rs: columns, stats and field
    bytes are identical serial and chunked."""
    data, _ = datagen.hash_csv(400_000, seed=99)
    geo, _ = datagen.geo_csv(400_000, seed=7)
    for spec, d in [(S.HASH_TYPED, data), (S.GEO_ALL, geo), (S.CSV_HASH, data)]:
        s_cols = ora.parse_columns(spec, d)
        s_stats = ora.csv_stats(spec, d)
        s_fb = ora.field_bytes(spec, d)
        for t in (2, 3, 7, 16):
            assert_columns_equal(ora.parse_columns(spec, d, t), s_cols, spec, f"{spec.name} t{t}")
            p = ora.csv_stats(spec, d, t)
            for k in range(len(spec.columns)):
                p["cols"][k]["sum"] = s_stats["cols"][k]["sum"]  # fold order differs
            assert p == s_stats, t
            assert ora.field_bytes(spec, d, t) == s_fb


def test_comments_golden(ora):
    """tests/comments.rs:13-53."""
    data = (b"# header comment, with \"quotes\" and, commas\nfoo,42\n#mid, comment\nbar,7\n"
            b"\"#not,a comment\",9\n# trailing comment")
    seps, ends = ora.parse(S.CSV_HASH_DIALECT, data)
    recs = [r.as_bytes() for r in Parsed(data, seps, ends, S.CSV_HASH_DIALECT).records()]
    assert recs == [b"foo,42", b"bar,7", b'"#not,a comment",9']
    cols = ora.parse_columns(S.CSV_HASH, data)
    assert cols["rows"] == 3
    o, d = cols["cols"][0]["offsets"], cols["cols"][0]["data"].tobytes()
    assert [d[o[r]:o[r + 1]] for r in range(3)] == [b"foo", b"bar", b"#not,a comment"]
    assert cols["cols"][1]["values"].tolist() == [42, 7, 9]
    tricky = b"#comment with one \" quote\na,\"multi\nline\"\n#another \" stray\nb,2\n"
    seps, ends = ora.parse(S.CSV_HASH_DIALECT, tricky)
    assert [r.as_bytes() for r in Parsed(tricky, seps, ends, S.CSV_HASH_DIALECT).records()] == \
        [b'a,"multi\nline"', b"b,2"]


def test_spans_hand_cases(ora):
    """tests/spans.rs:123-177 and 289-365."""
    data = b'a,"b,c","d""e"\r\nlast,row'
    p = Parsed(data, *ora.parse(S.CSV, data), S.CSV)
    r = list(p.records())
    assert len(r) == 2 and r[0].field_count() == 3
    assert r[0].fields() == [b"a", b"b,c", b'd"e'] and r[0].field(3) is None
    assert r[1].fields() == [b"last", b"row"]
    for doc, want in [(b"", []), (b"\n", [[b""]]), (b"a,b", [[b"a", b"b"]]), (b"a\r\nb\r\n", [[b"a"], [b"b"]])]:
        assert [x.fields() for x in Parsed(doc, *ora.parse(S.CSV, doc), S.CSV).records()] == want
    lf = b'a="x \\" y" b=plain\n'
    rec = list(Parsed(lf, *ora.parse(S.LOGFMT, lf), S.LOGFMT).records())
    assert rec[0].fields() == [b"a", b'x " y', b"b", b"plain"]
    # records_range split invariance (spans.rs:345-365)
    data = b'a,b\nc,"d\ne"\nf\r\ng,h,i\nlast,tail'
    p = Parsed(data, *ora.parse(S.CSV, data), S.CSV)
    full = [x.fields() for x in p.records()]
    n = p.terminated_record_count()
    for split in range(n + 1):
        got = [x.fields() for x in p.records_range(0, split)] + [x.fields() for x in p.records_range(split, n)]
        assert got == full, split


def test_ndjson_golden(ora):
    """tests/ndjson_id_score.rs: records 3, sum 62, serial == par."""
    data = (b'{"id":10,"name":"a","tags":["x","y"],"nested":{"score":7,"ok":true},"note":"plain"}\n'
            b'{"id":-3,"name":"quoted \\" id\\":999","tags":["n","m"],"nested":{"score":11,"ok":true},'
            b'"note":"brace } inside"}\n'
            b'{"id":42,"name":"slash \\\\ nested score\\":123","tags":["p","q"],"nested":{"score":-5,'
            b'"ok":true},"note":"end"}\n')
    for t in (1, 2, 8):
        r = ora.ndjson_id_score(data, t)
        assert r["records"] == 3 and r["invalid"] == 0
        assert r["id_sum"] + r["score_fixed6_sum"] // 1_000_000 == 62


def test_ndjson_lines_random_escapes(ora):
    """spans.rs:179-258: escape-aware record count equals a line split for
    NDJSON whose strings escape quotes and backslashes."""
    rng = Rng(0xFEEDFACECAFEBEEF)
    out = bytearray()
    for _ in range(200):
        out += b'{"k":"'
        for _ in range(rng.below(30)):
            c = rng.below(5)
            b = b"\\" if c == 0 else b'"' if c == 1 else bytes([97 + rng.below(26)])
            if b in (b"\\", b'"'):
                out += b"\\"
            out += b
        out += b'"}\n'
    assert ora.ndjson_lines(bytes(out)) == 200 == ora.ndjson_lines(bytes(out), 8)


def test_f64_goldens(ora):
    """Independent pin of `str::parse::<f64>`: tests/golden/f64_cells.json."""
    with open(os.path.join(HERE, "golden", "f64_cells.json")) as f:
        cells = json.load(f)["cells"]
    assert len(cells) > 500
    for cell, ok, bits, fixed6 in cells:
        got = ora.parse_f64(cell.encode())
        assert got[0] == ok, cell
        if ok:
            assert "%016x" % np.float64(got[1]).view(np.uint64) == bits, cell
            assert got[2] == fixed6, cell


def test_i64_cells(ora):
    """tests/columns.rs:414-444 integer cases; `str::parse::<i64>` closed form."""
    cases = {b"9223372036854775807": 9223372036854775807, b"-9223372036854775808": -9223372036854775808,
             b"9223372036854775808": None, b"+5": 5, b"-0": 0, b"0007": 7, b"": None, b"-": None,
             b"1_000": None, b"00000000000000000000000042": 42, b"-00000000000000009223372036854775808":
             -9223372036854775808, b" 1": None, b"1 ": None}
    for cell, want in cases.items():
        ok, v = ora.parse_i64(cell)
        assert ok == (want is not None) and (want is None or v == want), cell


def test_generator_closed_forms(ora):
    data, t = datagen.geo_csv(2_000_000)
    st = ora.csv_stats(S.GEO_CITY_LATLON, data)
    assert st["records"] == t["rows"]
    assert st["cols"][0]["bytes"] == t["city_bytes"] and st["cols"][0]["checksum"] == t["city_checksum"]
    for k, pre in ((1, "lat"), (2, "lon")):
        assert st["cols"][k]["values"] == t["data_rows"] == st["cols"][k]["fixed6_cells"]
        assert st["cols"][k]["fixed6_sum"] == t[pre + "_fixed6_sum"]
        assert st["cols"][k]["checksum"] == t[pre + "_checksum"]
    nd, nt = datagen.ndjson(1_000_000)
    r = ora.ndjson_id_score(nd)
    assert (r["records"], r["id_sum"], r["score_fixed6_sum"]) == (nt["records"], nt["id_sum"], nt["score_fixed6_sum"])
    h, ht = datagen.hash_csv(1_000_000)
    hs = ora.csv_stats(S.HASH_TYPED, h)
    assert hs["records"] == ht["records"] and hs["cols"][0]["checksum"] == ht["i64_sum"] % 2**64


# ---- field bytes: pinned by the independent records() walk ----------------
FIELD_BYTE_HAND_CASES = [
    b"a,1\nb,2\nc,3\n", b"# header\na,1\nb,2\n", b"a,1\n# tail\n", b"# just a comment\n",
    b"\"a\nb\",1\nc,2\n", b"\"a#b\",1\n\"#\",2\n", b"# a \" b\na,1\n", b"# \"\na,1\nb,2\n",
    b"\"a\"\"b\",1\n# c\n", b",\n,\n", b"a,1\nb,2", b"a,1\r\n# c\r\nb,2\r\n",
]


def records_field_bytes(ora, spec, data):
    """(fields, bytes) from the tape's record iterator: every field of every
    non-comment record, cleaned (csv_hash_fused_par.rs:153-167)."""
    p = Parsed(data, *ora.parse(spec, data), spec)
    fields = nbytes = 0
    for rec in p.records():
        for f in rec.fields():
            fields += 1
            nbytes += len(f)
    return fields, nbytes


def field_byte_random_docs(seed=0x0FB72C419E55D003, n=120):
    """The random documents of csv_hash_fused_par.rs:203-229 (same seed and
    generator, so the same documents)."""
    rng = Rng(seed)
    docs = []
    for _ in range(n):
        rows = 1 + rng.next() % 50
        d = bytearray()
        for _ in range(rows):
            c = rng.next() % 7
            if c == 0:
                d += b"# comment with \" quote\n"
            elif c == 1:
                d += b"\"quoted\nnewline\",42\n"
            elif c == 2:
                d += b"\"has#hash\",7\n"
            elif c == 3:
                d += b"\"a\"\"b\",13\n"
            elif c == 4:
                d += b",\n"
            elif c == 5:
                d += b"\"a,b\",5\n"
            else:
                k = rng.next() % 1000
                d += b"key%d,%d\n" % (k, k)
        docs.append(bytes(d))
    return docs


def test_field_bytes_match_records_reference(ora):
    """csv_hash_fused_par.rs:186-233: the fused field-byte sink equals the
    tape's records() walk on the 12 hazard cases and the 120 random
    documents, serial and at 1, 2, 3, 5, 8 and 17 threads."""
    spec = S.CSV_HASH_DIALECT
    for doc in FIELD_BYTE_HAND_CASES + field_byte_random_docs():
        want = records_field_bytes(ora, spec, doc)
        for t in (1, 2, 3, 5, 8, 17):
            got = ora.field_bytes(spec, doc, t)
            assert (got["fields"], got["bytes"]) == want, (doc, t)


def test_field_bytes_known_answers(ora):
    """Hand-counted answers: cleaned lengths with doubled quotes collapsed,
    comment records skipped, the record's trailing CR trimmed."""
    spec = S.CSV_HASH_DIALECT
    cases = {
        b"a,1\nb,2\nc,3\n": (6, 6),
        b"\"a\"\"b\",1\n# c\n": (2, 4),        # a"b + 1
        b",\n,\n": (4, 0),
        b"a,1\r\n# c\r\nb,2\r\n": (4, 4),       # CR trimmed from the last field
        b"# just a comment\n": (0, 0),
        b"\"a\nb\",1\nc,2\n": (4, 6),           # a\nb is 3 bytes
        b"a,1\nb,2": (4, 4),                    # trailing unterminated record
        b"\n": (1, 0),                          # an empty line is one empty field
    }
    for doc, want in cases.items():
        got = ora.field_bytes(spec, doc)
        assert (got["fields"], got["bytes"]) == want, doc


@pytest.mark.parametrize("spec", [S.CSV, S.TSV, S.CSV_HASH_DIALECT], ids=lambda s: s.name)
def test_field_bytes_random_hazard_docs(ora, spec):
    """Random documents over each dialect's hazard alphabet (quotes, comment
    bytes, CR, blank lines): the sink equals the records() walk."""
    rng = Rng(0x5EED_F1E1_DB17_E500)
    alpha = dialect_alphabet(spec) + b"\r\n"
    for _ in range(400):
        doc = random_doc(rng, alpha, rng.below(200))
        want = records_field_bytes(ora, spec, doc)
        for t in (1, 3):
            got = ora.field_bytes(spec, doc, t)
            assert (got["fields"], got["bytes"]) == want, (doc, t)


# ---- logfmt pairs: records() walk, hand answers, generator truths ---------
def records_logfmt(ora, data):
    """logfmt.rs:686-858 restated over the tape: consecutive fields of a
    record pair up (key, value); a value that starts with a quote loses its
    quotes and escapes as the dialect's cleaning does, any other value is
    counted raw."""
    spec = S.LOGFMT
    p = Parsed(data, *ora.parse(spec, data), spec)
    pairs = kb = vb = cs = 0
    for rec in p.records():
        raw = rec.fields_raw()
        for i in range(0, len(raw) - 1, 2):
            k, v = raw[i], raw[i + 1]
            if v[:1] == b'"':
                v = clean_field(spec, v)
            pairs += 1
            kb += len(k)
            vb += len(v)
            cs += sum(k) + sum(v)
    return {"pairs": pairs, "key_bytes": kb, "value_bytes": vb, "checksum": cs % 2**64}


def test_logfmt_known_answers(ora):
    """Hand-counted logfmt pairs (keys and values counted after unquoting and
    unescaping quoted values, logfmt.rs:826-858)."""
    s = lambda b: sum(b)  # noqa: E731
    cases = {
        b'a="x \\" y" b=plain\n': (2, 2, 10, s(b"a") + s(b'x " y') + s(b"b") + s(b"plain")),
        b"k=v\n\nk2=\n": (2, 3, 1, s(b"kvk2")),
        b"lonely\n": (0, 0, 0, 0),
        b"a=1 b\r\n": (1, 1, 1, s(b"a1")),
        b'x="q\\\\z" y="" z=w\n': (3, 3, 4, s(b"xq\\zyzw")),
        b'k="abc': (1, 1, 4, s(b'k"abc')),         # unterminated quote: counted raw
        b"a=1 b=2": (2, 2, 2, s(b"a1b2")),
        b"a=1 ": (1, 1, 1, s(b"a1")),
        b"a=1 b=2 c\n": (2, 2, 2, s(b"a1b2")),
        b"p=a\\b\n": (1, 1, 3, s(b"pa\\b")),       # unquoted value: escapes kept
        b'q="\\\\"\n': (1, 1, 1, s(b"q\\")),
        b'e="a=b c" f=1\n': (2, 2, 6, s(b"ea=b cf1")),
    }
    for doc, (pairs, kb, vb, cs) in cases.items():
        got = ora.logfmt_pairs(doc)
        assert (got["pairs"], got["key_bytes"], got["value_bytes"], got["checksum"]) == (pairs, kb, vb, cs), doc
        assert records_logfmt(ora, doc) == got, doc


def test_logfmt_random_hazard_docs(ora):
    """Random logfmt documents (quotes, escapes, '=', spaces, CR, blank lines,
    odd field counts, no final newline) against the records() walk."""
    rng = Rng(0x10F3_A7C0_FFEE_0001)
    alpha = b'ab= =  ""\\\\\r\n\n'
    for _ in range(600):
        doc = random_doc(rng, alpha, rng.below(160))
        assert ora.logfmt_pairs(doc) == records_logfmt(ora, doc), doc


def test_logfmt_generator_closed_forms(ora):
    """The generator's own count of what it wrote: pairs, key and unescaped
    value bytes, byte checksum; serial and newline-chunked alike."""
    data, t = datagen.logfmt(3_000_000)
    want = {"pairs": t["pairs"], "key_bytes": t["key_bytes"], "value_bytes": t["value_bytes"],
            "checksum": t["checksum"] % 2**64}
    for threads in (1, 4, 16):
        assert ora.logfmt_pairs(data, threads) == want, threads
    assert records_logfmt(ora, bytes(data[:200_000]).rsplit(b"\n", 1)[0] + b"\n")["pairs"] > 0


# ---- string cleaning, backslash dialects ------------------------------------
def test_backslash_clean_known_answers(ora):
    """codegen.rs:8078-8110 (append_clean, backslash escape): outer quotes
    stripped, each escape byte followed by a byte becomes that byte, a final
    lone escape byte is kept; the same rule unquoted."""
    cases = {
        b'"a\\"b"': b'a"b', b"plain": b"plain", b'"x"': b"x", b"a\\\\b": b"a\\b",
        b'"ab\\"': b"ab\\", b"\\": b"\\", b'"': b'"', b'""': b"", b'"\\\\"': b"\\",
        b"x\\y\\": b"xy\\", b'"unterminated': b'"unterminated', b"\\n\\t": b"nt",
        b'"a,b"': b"a,b", b"": b"",
    }
    for spec in (S.LOGFMT, S.NDJSON):
        for cell, want in cases.items():
            assert ora.clean(spec, cell) == want, (spec.name, cell)
            assert clean_field(spec, cell) == want, (spec.name, cell)


def test_doubled_quote_clean_known_answers(ora):
    """csv_hash.rs:1749-1779: quoted at both ends -> inside with each doubled
    quote collapsed (a lone quote copied); otherwise raw."""
    cases = {b'"a""b"': b'a"b', b'"a"b"': b'a"b', b'""': b"", b'"': b'"', b'x"y': b'x"y',
             b'""""': b'"', b'"a,b"': b"a,b", b"plain": b"plain", b'"""': b'"'}
    for cell, want in cases.items():
        assert ora.clean(S.CSV, cell) == want, cell
        assert clean_field(S.CSV, cell) == want, cell
