"""GPU parity: every entry point of libfalx_gpu.so (called through the C ABI)
against the CPU oracle on the same inputs, bit-exact. Inputs follow the
reference's own test strategy (tests/codegen.rs differential alphabets,
tests/columns.rs cells, tests/csv_hash_fused_par.rs hazards, block/chunk
boundary sweeps) plus generator data with closed-form truths."""
import numpy as np
import pytest

from falx_amd import datagen
from falx_amd import spec as S
from helpers import Rng, assert_columns_equal, assert_stats_equal, dialect_alphabet, random_doc

pytestmark = pytest.mark.gpu

TILE = 8192
COMMENT_PLAIN = S.Spec("vcf_like", b"\t\n", None, None, ord("#"))
INDEX_DIALECTS = [S.CSV, S.TSV, S.LOGFMT, S.NDJSON, S.MULTI, S.CSV_HASH_DIALECT, COMMENT_PLAIN,
                  S.Spec("esc_regions", b",\n", ord('"'), ord("\\"), ord("#"))]


def _docs(spec, seed, count, max_len):
    rng = Rng(seed)
    alpha = dialect_alphabet(spec)
    return [random_doc(rng, alpha, rng.below(max_len)) for _ in range(count)]


@pytest.mark.parametrize("spec", INDEX_DIALECTS, ids=lambda s: s.name)
def test_index_structurals_random(ctx, ora, spec):
    """codegen.rs:391-468: randomized docs over the dialect alphabet."""
    for i, doc in enumerate(_docs(spec, 0xDEADBEEFCAFEBABE, 150, 300) + [b"", b"\n", b'"', b"#"]):
        want = ora.index_structurals(spec, doc)
        if doc:
            got = ctx.index_structurals(spec, ctx.upload(doc), len(doc))
        else:
            got = np.empty(0, np.uint32)
        assert np.array_equal(got, want), f"{spec.name} doc {i}: {doc!r}"
        if spec.name != "vcf_like":  # comment-without-quote: kernel semantics differ from scalar.rs
            assert np.array_equal(want, ora.index_scalar(spec, doc))


@pytest.mark.parametrize("spec", INDEX_DIALECTS, ids=lambda s: s.name)
def test_index_structurals_long(ctx, ora, spec):
    """codegen.rs:471 long inputs, spanning many 8 KiB tiles and look-back."""
    rng = Rng(0x1234_5678_9ABC_DEF1)
    alpha = dialect_alphabet(spec)
    for length in (8191, 8192, 8193, 100_000, 300_001):
        doc = random_doc(rng, alpha, length)
        want = ora.index_structurals(spec, doc)
        got = ctx.index_structurals(spec, ctx.upload(doc), len(doc))
        assert np.array_equal(got, want), f"{spec.name} len {length}"


HAZARDS = [b'"x\ny",7\n', b'# c " c\n', b'"#",7\n', b'"a""b",7\n', b'"' + b"q,\n" * 40 + b'",1\n',
           b"\\\\\\\"" + b'"a\\"b",2\n']


@pytest.mark.parametrize("spec", [S.CSV, S.CSV_HASH_DIALECT, S.LOGFMT, S.NDJSON], ids=lambda s: s.name)
def test_hazards_across_tile_boundaries(ctx, ora, spec):
    """csv_hash_fused_par.rs:88-118, at tile scale: each hazard lands at every
    offset around the 8 KiB tile seam (and thus every 64-byte block seam)."""
    for hazard in HAZARDS:
        for pad in range(TILE - 70, TILE + 6, 7):
            doc = b"k" * pad + b"\n" + hazard + b"tail,99\n" + hazard
            want = ora.index_structurals(spec, doc)
            got = ctx.index_structurals(spec, ctx.upload(doc), len(doc))
            assert np.array_equal(got, want), f"{spec.name} pad {pad} hazard {hazard!r}"


@pytest.mark.parametrize("spec", [S.CSV, S.CSV_HASH_DIALECT, S.NDJSON, S.TSV], ids=lambda s: s.name)
def test_parse_tape(ctx, ora, spec):
    rng = Rng(0x5EED_5EED_5EED_5EED)
    alpha = dialect_alphabet(spec) + b"\rxy"
    for length in (0, 1, 63, 64, 65, 5000, 40_000):
        doc = random_doc(rng, alpha, length)
        ws, we = ora.parse(spec, doc)
        if length == 0:
            continue
        gs, ge = ctx.parse(spec, ctx.upload(doc), len(doc))
        assert np.array_equal(gs, ws) and np.array_equal(ge, we), f"{spec.name} len {length}"


def ref_parse_records(data: bytes):
    """tests/columns.rs:14-74: toggle on every quote, split at unquoted ',' and '\\n'."""
    records, cur, fs, inq = [], [], 0, False
    for pos, b in enumerate(data):
        if b == 34:
            inq = not inq
        elif b == 44 and not inq:
            cur.append(data[fs:pos]); fs = pos + 1
        elif b == 10 and not inq:
            end = pos - 1 if pos > fs and data[pos - 1] == 13 else pos
            cur.append(data[fs:end]); records.append(cur); cur = []; fs = pos + 1
    if fs < len(data):
        end = len(data) - 1 if len(data) > fs and data[-1] == 13 else len(data)
        if end > fs or cur:
            cur.append(data[fs:end]); records.append(cur)
    elif cur:
        cur.append(b""); records.append(cur)
    return records


def columns_hand_doc() -> bytes:
    # tests/columns.rs:408-451 hand-picked cells
    rows = [
        "9223372036854775807,hello,1.5,x,label0\n",
        "-9223372036854775808,,-0.0,x,\"lab,el1\"\n",
        "9223372036854775808,\"a,\"\"b\"\",0.1,x,\"\"\n",
        "+5,\"quo\nted\",0.0,x,lab3\n",
        "-0,\"\",1_000,x,lab4\n",
        "0007,world,1.5e308,x,\"123\"\n",
        ",,,5e-324,x,\"abc\"\n",
        "17\n",
        "18,test,NaN,x,label8\n",
        "19,\"a\"\"b\",inf,x,label9\n",
        "20,simple,-inf,x,label10\n",
    ]
    return "".join(rows).encode()


def typed_random_doc(rng: Rng) -> bytes:
    """tests/columns.rs:453-620 randomized record generator."""
    out = bytearray()
    for _ in range(rng.below(40) + 1):
        nf = rng.below(8)
        crlf = rng.below(10) < 1
        for f in range(nf):
            if f:
                out += b","
            t = rng.below(7)
            if t == 0:
                if rng.below(3) == 0:
                    out += b"-" if rng.below(2) == 0 else b"+"
                out += bytes(48 + rng.below(10) for _ in range(rng.below(22) + 1))
            elif t == 1:
                if rng.below(2) == 0:
                    out += b"-"
                style = rng.below(5)
                dg = lambda n: bytes(48 + rng.below(10) for _ in range(n))
                if style == 0:
                    out += dg(rng.below(18) + 1) + b"." + dg(rng.below(18))
                elif style == 1:
                    out += dg(rng.below(15) + 1)
                elif style == 2:
                    out += dg(rng.below(8) + 1) + b"e" + (b"-" if rng.below(2) == 0 else b"") + \
                        str(abs(rng.below(700) - 350)).encode()
                elif style == 3:
                    out += b"." + dg(rng.below(15) + 1)
                else:
                    out += dg(rng.below(15) + 1) + b"."
            elif t == 2:
                out += bytes(97 + rng.below(26) for _ in range(rng.below(10)))
            elif t == 3:
                pass
            elif t == 4:
                out += [b"inf", b"nan", b"NaN"][rng.below(3)]
            else:
                out += b'"'
                for _ in range(rng.below(50)):
                    it = rng.below(5)
                    out += [b",", b'""', b"\n", b" ", bytes([97 + rng.below(26)])][it]
                out += b'"'
        out += b"\r\n" if crlf else b"\n"
    if rng.below(3) == 0:
        for f in range(rng.below(8)):
            if f:
                out += b","
            out += bytes(48 + rng.below(10) for _ in range(rng.below(20)))
    return bytes(out)


def test_columns_hand_cells(ctx, ora):
    doc = columns_hand_doc()
    want = ora.parse_columns(S.CSV_TYPED, doc)
    got = ctx.parse_columns(S.CSV_TYPED, ctx.upload(doc), len(doc))
    assert_columns_equal(got, want, S.CSV_TYPED, "hand cells")
    # Row 2's title `"a,""b""` holds an odd number of quotes, so the quoted
    # region runs on; the reference's dumb model agrees on 4 records.
    assert want["rows"] == len(ref_parse_records(doc)) == 4


def test_columns_randomized(ctx, ora):
    rng = Rng(0xDEADBEEFCAFEBABE)
    for i in range(300):
        doc = typed_random_doc(rng)
        want = ora.parse_columns(S.CSV_TYPED, doc)
        got = ctx.parse_columns(S.CSV_TYPED, ctx.upload(doc), len(doc))
        assert_columns_equal(got, want, S.CSV_TYPED, f"doc {i}")


def test_columns_comments_golden(ctx, ora):
    """tests/comments.rs:13-53 golden: keys foo/bar/#not,a comment, amounts [42,7,9]."""
    doc = (b"# header comment, with \"quotes\" and, commas\nfoo,42\n#mid, comment\nbar,7\n"
           b"\"#not,a comment\",9\n# trailing comment")
    got = ctx.parse_columns(S.CSV_HASH, ctx.upload(doc), len(doc))
    assert got["rows"] == 3
    offs, data = got["cols"][0]["offsets"], got["cols"][0]["data"].tobytes()
    assert [data[offs[r]:offs[r + 1]] for r in range(3)] == [b"foo", b"bar", b"#not,a comment"]
    assert got["cols"][1]["values"].tolist() == [42, 7, 9]
    assert_columns_equal(got, ora.parse_columns(S.CSV_HASH, doc), S.CSV_HASH, "comments")


def test_columns_region_hazard_docs(ctx, ora):
    """tests/csv_hash_fused_par.rs:61-80 cases and random documents (120-151)."""
    cases = [b"a,1\nb,2\nc,3\n", b"# header\na,1\nb,2\n", b"a,1\n# tail\n", b"# just a comment\n",
             b"\"a\nb\",1\nc,2\n", b"\"a#b\",1\n\"#\",2\n", b"# a \" b\na,1\n", b"# \"\na,1\nb,2\n",
             b"\"a,b\",1\n", b"\"a\"\"b\",1\n", b",\n,\n", b"a,1\nb,2", b"a,1\r\nb,2\r\n", b"a,1\n# mid\nb,2\n"]
    rng = Rng(0x5EED_C5A5_1234_ABCD)
    for _ in range(60):
        d = bytearray()
        for _ in range(1 + rng.below(60)):
            k = rng.below(8)
            d += [b"# comment with \" quote\n", b"# plain comment\n", b"\"quoted\nnewline\",42\n",
                  b"\"has#hash\",7\n", b"\"a\"\"b\",13\n", b",\n", b"\"a,b\",5\n",
                  f"key{rng.below(1000)},{rng.below(100000)}\n".encode()][k]
        if rng.below(4) == 0:
            d = d[:-1]
        cases.append(bytes(d))
    for i, doc in enumerate(cases):
        want = ora.parse_columns(S.CSV_HASH, doc)
        got = ctx.parse_columns(S.CSV_HASH, ctx.upload(doc), len(doc))
        assert_columns_equal(got, want, S.CSV_HASH, f"case {i} {doc[:40]!r}")
        fb = ctx.field_bytes(S.CSV_HASH, ctx.upload(doc), len(doc))
        assert fb == ora.field_bytes(S.CSV_HASH, doc), f"field bytes case {i}"


@pytest.mark.parametrize("spec", [S.CSV_HASH, S.GEO_CITY_LATLON], ids=lambda s: s.name)
def test_quoted_string_cells_across_tile_seam(ctx, ora, spec):
    """String cells quoted at both ends, with and without doubled quotes, at
    every offset from 80 bytes before the 8 KiB tile seam to 8 after it. A
    cell that lies in one tile and holds no quote but its outer two is
    cleaned as a run of the document; one crossing the seam, starting
    before its record's tile or holding a doubled quote takes the cleaning
    stream. Both kinds of tile front run: csv_hash (comment dialect, one
    byte per lane and ballots) and the geo CSV (32 bytes per thread)."""
    cells = [b'"ab,cd"', b'"a""b"', b'""', b'"x"', b'"' + b"q" * 70 + b'"', b'"' + b"r" * 40 + b'""s"']
    for cell in cells:
        for off in range(TILE - 80, TILE + 8):
            if spec is S.CSV_HASH:
                doc = b"k" * (off - 1) + b"\n" + cell + b",7\n" + cell + b",8\n"
            else:
                doc = b"9" * (off - 3) + b"\n1," + cell + b",1.500000,-2.250000\n2," + cell + b",3.000000,4.000000\n"
            want = ora.parse_columns(spec, doc)
            got = ctx.parse_columns(spec, ctx.upload(doc), len(doc))
            assert_columns_equal(got, want, spec, f"cell {cell[:12]!r} at {off}")


@pytest.mark.parametrize("spec", [S.GEO_LATLON, S.GEO_CITY_LATLON, S.GEO_ALL], ids=lambda s: s.name)
def test_geo_columns_and_stats(ctx, ora, spec):
    data, truth = datagen.geo_csv(3 * 1024 * 1024 + 777)
    buf = ctx.upload(data)
    want = ora.parse_columns(spec, data)
    got = ctx.parse_columns(spec, buf, data.size)
    assert_columns_equal(got, want, spec, spec.name)
    st = ctx.csv_stats(spec, buf, data.size)
    assert_stats_equal(st, ora.csv_stats(spec, data), spec, spec.name)
    assert st["records"] == truth["rows"]
    lat = [k for k, c in enumerate(spec.columns) if c.index == 2][0]
    assert st["cols"][lat]["fixed6_sum"] == truth["lat_fixed6_sum"]
    assert st["cols"][lat]["checksum"] == truth["lat_checksum"]


@pytest.mark.parametrize("spec", [S.GEO_LATLON, S.GEO_ALL], ids=lambda s: s.name)
def test_columns_fast_pass_redoes_flagged_tiles(ctx, ora, spec):
    """A multi-tile geo document whose Latitude/Longitude cells are fixed-6
    except in a few scattered rows (quoted, exponent, 25-digit, inf, empty,
    and cells across tile edges): the fast columns and stats passes leave
    those tiles to the full pass, and every cell matches the oracle."""
    data, _ = datagen.geo_csv(700_000, seed=77)
    lines = bytes(data).split(b"\n")
    odd = [b'"12.5"', b"1e5", b"-0000000000000000000000012.75", b"inf", b"", b"  7.25", b"1.0000000000000000000001"]
    rng = Rng(0xF1A6)
    for t in range(40):
        i = 1 + rng.below(len(lines) - 2)
        f = lines[i].split(b",")
        if len(f) == 4:
            f[2 + rng.below(2)] = odd[t % len(odd)]
            lines[i] = b",".join(f)
    doc = np.frombuffer(b"\n".join(lines), dtype=np.uint8)
    buf = ctx.upload(doc)
    assert_columns_equal(ctx.parse_columns(spec, buf, doc.size), ora.parse_columns(spec, doc), spec, spec.name)
    assert_stats_equal(ctx.csv_stats(spec, buf, doc.size), ora.csv_stats(spec, doc), spec, spec.name)
    cols, keep = ctx.alloc_columns(spec, ora.parse_columns(spec, doc)["rows"], doc.size)
    st = ctx.parse_columns_with_stats_into(spec, buf, doc.size, cols)
    assert_stats_equal(st, ora.csv_stats(spec, doc), spec, spec.name + " fused")


def test_columns_every_tile_flagged(ctx, ora):
    """Every Latitude quoted and every 7th in exponent form: the register
    parsers take the quoted cells, the exponent cells flag every tile, and
    the full pass redoes all of them (many tiles per redo workgroup)."""
    data, _ = datagen.geo_csv(900_000, seed=5)
    lines = bytes(data).split(b"\n")
    for i in range(1, len(lines)):
        f = lines[i].split(b",")
        if len(f) == 4:
            f[2] = b'"' + (f[2].replace(b".", b"e") if i % 7 == 0 else f[2]) + b'"'
            lines[i] = b",".join(f)
    doc = np.frombuffer(b"\n".join(lines), dtype=np.uint8)
    buf = ctx.upload(doc)
    spec = S.GEO_LATLON
    assert_columns_equal(ctx.parse_columns(spec, buf, doc.size), ora.parse_columns(spec, doc), spec, "flagged")
    assert_stats_equal(ctx.csv_stats(spec, buf, doc.size), ora.csv_stats(spec, doc), spec, "flagged stats")


def test_hash_typed_columns_stats(ctx, ora):
    data, truth = datagen.hash_csv(2 * 1024 * 1024 + 333)
    buf = ctx.upload(data)
    spec = S.HASH_TYPED
    assert_columns_equal(ctx.parse_columns(spec, buf, data.size), ora.parse_columns(spec, data), spec, "hash")
    assert_stats_equal(ctx.csv_stats(spec, buf, data.size), ora.csv_stats(spec, data), spec, "hash stats")
    assert ctx.field_bytes(spec, buf, data.size) == ora.field_bytes(spec, data)
    assert np.array_equal(ctx.index_structurals(spec, buf, data.size), ora.index_structurals(spec, data))


def test_ndjson_golden_and_generated(ctx, ora):
    golden = (b'{"id":10,"name":"a","tags":["x","y"],"nested":{"score":7,"ok":true},"note":"plain"}\n'
              b'{"id":-3,"name":"quoted \\" id\\":999","tags":["n","m"],"nested":{"score":11,"ok":true},'
              b'"note":"brace } inside"}\n'
              b'{"id":42,"name":"slash \\\\ nested score\\":123","tags":["p","q"],"nested":{"score":-5,'
              b'"ok":true},"note":"end"}\n')
    r = ctx.ndjson_id_score(ctx.upload(golden), len(golden))
    assert r["records"] == 3 and r["id_sum"] + r["score_fixed6_sum"] // 1_000_000 == 62  # ndjson_id_score.rs
    data, truth = datagen.ndjson(2 * 1024 * 1024 + 5)
    buf = ctx.upload(data)
    got = ctx.ndjson_id_score(buf, data.size)
    want = ora.ndjson_id_score(data)
    for k in ("records", "invalid", "id_sum", "score_fixed6_sum", "score_fixed6_cells"):
        assert got[k] == want[k], k
    assert got["records"] == truth["records"] and got["id_sum"] == truth["id_sum"]
    assert got["score_fixed6_sum"] == truth["score_fixed6_sum"]
    assert ctx.ndjson_lines(buf, data.size) == ora.ndjson_lines(data)


def test_logfmt_pairs(ctx, ora):
    data, _ = datagen.logfmt(1024 * 1024 + 3)
    buf = ctx.upload(data)
    assert ctx.logfmt_pairs(buf, data.size) == ora.logfmt_pairs(data)
    rng = Rng(0xE7037ED1A0B428DB)
    for i in range(80):
        doc = random_doc(rng, b' ="\\\nab\r', rng.below(400))
        if not doc:
            continue
        assert ctx.logfmt_pairs(ctx.upload(doc), len(doc)) == ora.logfmt_pairs(doc), f"doc {i} {doc!r}"


WIDE = S.GEO_ALL.with_columns([(0, S.I64), (1, S.STR), (2, S.F64), (3, S.F64), (2, S.F64), (0, S.BYTES)],
                              name="geo_wide")


@pytest.mark.parametrize("spec", [S.GEO_LATLON, S.GEO_CITY_LATLON, WIDE], ids=lambda s: s.name)
def test_columns_with_stats_one_pass(ctx, ora, spec):
    """falx_parse_columns_with_stats == parse_columns + csv_stats (oracle)."""
    data, truth = datagen.geo_csv(1024 * 1024 + 99, seed=42)
    buf = ctx.upload(data)
    rows = truth["rows"]
    cols, keep = ctx.alloc_columns(spec, rows, data.size)
    st = ctx.parse_columns_with_stats_into(spec, buf, data.size, cols)
    assert cols.rows == rows
    assert_stats_equal(st, ora.csv_stats(spec, data), spec, spec.name)
    want = ora.parse_columns(spec, data)
    got = {"rows": rows, "cols": []}
    for k, c in enumerate(spec.columns):
        vals, valid, dbuf = keep[k]
        e = {"valid": valid.download(np.uint64, (rows + 63) // 64)}
        if c.type == S.STR:
            e["offsets"] = vals.download(np.int32, rows + 1)
            e["data"] = dbuf.download(np.uint8, cols.cols[k].data_len)
        elif c.type == S.BYTES:
            e["values"] = vals.download(np.uint32, 2 * rows).reshape(-1, 2)
        else:
            e["values"] = vals.download(np.int64 if c.type == S.I64 else np.float64, rows)
        got["cols"].append(e)
    assert_columns_equal(got, want, spec, spec.name)


def test_capacity_errors(ctx, ora):
    """Too-small outputs report CAPACITY with the required sizes."""
    from falx_amd.gpu import FalxError
    data, truth = datagen.geo_csv(200_000, seed=3)
    buf = ctx.upload(data)
    cols, _ = ctx.alloc_columns(S.GEO_CITY_LATLON, truth["rows"] - 5, 100)
    assert ctx.parse_columns_into(S.GEO_CITY_LATLON, buf, data.size, cols) == S.E_CAPACITY
    assert cols.rows == truth["rows"] and cols.cols[0].data_len == truth["city_bytes"]
    small = ctx.alloc(4 * 10)
    with pytest.raises(FalxError):
        ctx.index_structurals(S.CSV, buf, data.size, out=small)


def test_columns_dense_rows_and_fixed6_edges(ctx, ora):
    """Tiles holding thousands of short records with six columns, so a tile's
    row span does not fit the LDS validity stage (global atomics per cell),
    mixed with fixed-6 edge cells at tile seams (fast path vs general path)."""
    spec = S.CSV.with_columns([(0, S.F64), (1, S.I64), (2, S.STR), (0, S.F64), (3, S.BYTES), (5, S.F64)],
                              "dense6")
    rng = Rng(0xF1E6_0D5E)
    edge = [b"1.234567", b"-1.234567", b"+1.234567", b"123.456789", b"-123.456789", b"1234.567890",
            b"0.000000", b"-0.000000", b"1.23456", b"1.2345678", b"12.345e67", b"1.234567\r", b'"1.234567"',
            b".1234567", b"-.123456", b"1..23456", b"999.999999", b"-999.999999"]
    parts = []
    for i in range(30000):
        k = rng.below(10)
        if k < 6:
            parts.append(b"\n")                                  # empty record
        elif k < 8:
            parts.append(edge[rng.below(len(edge))] + b"\n")
        else:
            parts.append(edge[rng.below(len(edge))] + b"," + str(rng.below(1000)).encode() + b",x,y,z,"
                         + edge[rng.below(len(edge))] + b"\n")
    doc = b"".join(parts)
    want = ora.parse_columns(spec, doc)
    got = ctx.parse_columns(spec, ctx.upload(doc), len(doc))
    assert_columns_equal(got, want, spec, "dense rows")
    assert_stats_equal(ctx.csv_stats(spec, ctx.upload(doc), len(doc)), ora.csv_stats(spec, doc), spec, "dense")


def _numeric_cell(rng: Rng) -> bytes:
    """Numeric-looking cells around the register parser's limits: 12- and
    24-byte windows, 19 digits, four non-digits, exponents of 0-6 digits,
    signs, quotes and stray bytes (the shapes it must hand to the stream
    parser)."""
    dg = lambda n: bytes(48 + rng.below(10) for _ in range(n))
    s = b"-" if rng.below(4) == 0 else (b"+" if rng.below(12) == 0 else b"")
    k = rng.below(10)
    if k < 3:
        s += dg(rng.below(6)) + b"." + dg(rng.below(14))
    elif k < 5:
        s += dg(1 + rng.below(21))
    elif k < 7:
        s += dg(rng.below(9)) + (b"." + dg(rng.below(9)) if rng.below(2) else b"") + (b"e" if rng.below(2) else b"E")
        s += [b"", b"-", b"+"][rng.below(3)] + dg(rng.below(7))
    elif k < 8:
        s += [b"9223372036854775807", b"9223372036854775808", b"00000000000000000000012", b"inf", b"nan"][rng.below(5)]
    elif k < 9:
        s = b'"' + s + dg(1 + rng.below(8)) + (b"." + dg(3) if rng.below(2) else b"") + b'"'
    else:
        s += bytes(b"0123456789.eE+-x "[rng.below(17)] for _ in range(rng.below(26)))
    return s


def test_columns_register_parser_edges(ctx, ora):
    """i64 and f64 columns over the same cells, records of varying length so
    the cells fall at every offset of a tile and across tile seams."""
    spec = S.CSV.with_columns([(0, S.I64), (0, S.F64), (1, S.F64), (2, S.I64), (3, S.STR)], "numedges")
    rng = Rng(0x5A4_C311)
    lines = []
    for _ in range(40000):
        cells = [_numeric_cell(rng) for _ in range(1 + rng.below(3))]
        if rng.below(4) == 0:
            cells.append(b"pad" * rng.below(9))
        lines.append(b",".join(cells) + (b"\r\n" if rng.below(20) == 0 else b"\n"))
    doc = b"".join(lines)
    want = ora.parse_columns(spec, doc)
    got = ctx.parse_columns(spec, ctx.upload(doc), len(doc))
    assert_columns_equal(got, want, spec, "numeric edges")
    assert_stats_equal(ctx.csv_stats(spec, ctx.upload(doc), len(doc)), ora.csv_stats(spec, doc), spec, "numeric edges")
