"""The batched mixed stream against the oracle: one batch holding every spec
of the served mix (geo CSV, csv_hash, TSV, logfmt, NDJSON) and every op, with
documents from 4 KiB up to the largest size classes (class 6: 32-128 MiB,
class 7: 128 MiB and more, where the tuned chunks are largest). Each
document's outputs and result are compared with the CPU oracle, not only
with the plain entry points."""
import numpy as np
import pytest

from falx_amd import datagen, gpu
from falx_amd import spec as S
from helpers import Rng, assert_columns_equal, assert_stats_equal
from test_gpu_batch import Doc

pytestmark = pytest.mark.gpu

MIB = 1 << 20


def _host_columns(d: Doc):
    rows = d.cols.rows
    res = {"rows": rows, "cols": []}
    for k, c in enumerate(d.spec.columns):
        vals, valid, dbuf = d.keep[k]
        e = {"valid": valid.download(np.uint64, (rows + 63) // 64)}
        if c.type == S.STR:
            e["offsets"] = vals.download(np.int32, rows + 1)
            e["data"] = dbuf.download(np.uint8, d.cols.cols[k].data_len)
        elif c.type == S.BYTES:
            e["values"] = vals.download(np.uint32, 2 * rows).reshape(-1, 2)
        else:
            e["values"] = vals.download(np.int64 if c.type == S.I64 else np.float64, rows)
        res["cols"].append(e)
    return res


def _check_doc(ora, d: Doc, r, label):
    data = d.data
    if d.op in (S.OP_COLUMNS, S.OP_COLUMNS_STATS):
        assert r["status"] == S.OK, label
        assert_columns_equal(_host_columns(d), ora.parse_columns(d.spec, data, 16), d.spec, label)
        if d.op == S.OP_COLUMNS_STATS:
            assert_stats_equal(r["stats"], ora.csv_stats(d.spec, data, 16), d.spec, label)
    elif d.op == S.OP_CSV_STATS:
        assert_stats_equal(r["stats"], ora.csv_stats(d.spec, data, 16), d.spec, label)
    elif d.op == S.OP_INDEX:
        got = d.index.download(np.uint32, r["n_index"])
        assert np.array_equal(got, ora.index_structurals(d.spec, data, 16)), label
    elif d.op == S.OP_FIELD_BYTES:
        assert r["field_bytes"] == ora.field_bytes(d.spec, data, 16), label
    elif d.op == S.OP_LOGFMT_PAIRS:
        assert r["logfmt"] == ora.logfmt_pairs(data), label
    else:
        want = ora.ndjson_id_score(data, 16)
        got = r["ndjson"]
        for k in ("records", "invalid", "id_sum", "score_fixed6_sum", "score_fixed6_cells"):
            assert got[k] == want[k], f"{label}: {k}"
        assert got["score_sum"] == pytest.approx(want["score_sum"], rel=1e-12)


def test_mixed_batch_all_size_classes_vs_oracle(ctx, ora):
    rng = Rng(0xC0F1_64A7)
    plan = [
        (S.OP_COLUMNS_STATS, S.GEO_CITY_LATLON, datagen.geo_csv, 140 * MIB + 333),   # class 7
        (S.OP_COLUMNS_STATS, S.HASH_TYPED, datagen.hash_csv, 40 * MIB + 5),          # class 6
        (S.OP_COLUMNS_STATS, S.TSV_TYPED, datagen.tsv, 36 * MIB + 77),               # class 6
        (S.OP_LOGFMT_PAIRS, None, datagen.logfmt, 34 * MIB + 1),                     # class 6
        (S.OP_NDJSON_ID_SCORE, None, datagen.ndjson, 40 * MIB + 9),                  # class 6
        (S.OP_INDEX, S.HASH_TYPED, datagen.hash_csv, 9 * MIB),                       # class 5
        (S.OP_FIELD_BYTES, S.HASH_TYPED, datagen.hash_csv, 3 * MIB),
        (S.OP_CSV_STATS, S.GEO_CITY_LATLON, datagen.geo_csv, 2 * MIB + 1),
        (S.OP_COLUMNS, S.GEO_LATLON, datagen.geo_csv, 700 * 1024),
        (S.OP_COLUMNS_STATS, S.TSV_TYPED, datagen.tsv, 300 * 1024),
        (S.OP_CSV_STATS, S.TSV_TYPED, datagen.tsv, 64 * 1024),
        (S.OP_LOGFMT_PAIRS, None, datagen.logfmt, 20 * 1024),
        (S.OP_NDJSON_ID_SCORE, None, datagen.ndjson, 4096),
        (S.OP_COLUMNS_STATS, S.HASH_TYPED, datagen.hash_csv, 4096),
        (S.OP_INDEX, S.GEO_CITY_LATLON, datagen.geo_csv, 5000),
    ]
    docs = []
    for op, spec, gen, n in plan:
        data = gen(n, seed=rng.next())[0]
        docs.append(Doc(ctx, op, spec, np.ascontiguousarray(data[:n])))
    reqs = [d.request() for d in docs]
    scratch = sum(r.scratch_bytes() for r in reqs) + (64 << 20)
    b = gpu.Batch(ctx, scratch)
    b.submit(reqs)
    results = b.wait()
    for i, (d, r) in enumerate(zip(docs, results)):
        print(f"doc {i}: op {d.op} {d.spec.name if d.spec else '-'} {d.n} B", flush=True)
        _check_doc(ora, d, r, f"doc {i} op {d.op} ({d.n} B)")
