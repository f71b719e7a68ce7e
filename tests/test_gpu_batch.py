"""GPU parity of batched execution (include/falx_gpu.h falx_batch_*): every
document of a mixed batch (five specs, every op, sizes from empty to ~1000
tiles) gets exactly the plain entry point's outputs and result, bit for
bit (the fixed-order FP64 sums included), across graph replays with fresh
data, and a document alone in a batch matches the same document in the mix."""
import numpy as np
import pytest

from falx_amd import datagen
from falx_amd import gpu
from falx_amd import spec as S
from helpers import Rng

pytestmark = pytest.mark.gpu

SCRATCH = 256 << 20


def _docs(rng: Rng, scale: int):
    """(op, spec, bytes) triples of the served mix at sizes 0 .. ~scale."""
    sizes = [0, 100, 4096, 8191, 8192, 8193, 70000, scale // 3, scale]
    out = []
    gens = [
        (S.OP_COLUMNS_STATS, S.GEO_CITY_LATLON, datagen.geo_csv),
        (S.OP_COLUMNS, S.GEO_LATLON, datagen.geo_csv),
        (S.OP_COLUMNS_STATS, S.HASH_TYPED, datagen.hash_csv),
        (S.OP_INDEX, S.HASH_TYPED, datagen.hash_csv),
        (S.OP_FIELD_BYTES, S.HASH_TYPED, datagen.hash_csv),
        (S.OP_COLUMNS_STATS, S.TSV_TYPED, datagen.tsv),
        (S.OP_CSV_STATS, S.TSV_TYPED, datagen.tsv),
        (S.OP_CSV_STATS, S.GEO_CITY_LATLON, datagen.geo_csv),
        (S.OP_LOGFMT_PAIRS, None, datagen.logfmt),
        (S.OP_NDJSON_ID_SCORE, None, datagen.ndjson),
    ]
    for op, spec, gen in gens:
        for _ in range(2):
            n = sizes[rng.below(len(sizes))]
            data = gen(n, seed=rng.next())[0] if n else np.zeros(0, np.uint8)
            out.append((op, spec, np.ascontiguousarray(data[:n]) if n else data))
    return out


class Doc:
    def __init__(self, ctx, op, spec, data):
        self.op, self.spec, self.data = op, spec, data
        n = int(data.size)
        self.buf = ctx.upload(data if n else np.zeros(16, np.uint8))
        self.n = n
        self.cols = self.keep = self.index = None
        if op in (S.OP_COLUMNS, S.OP_COLUMNS_STATS):
            self.cols, self.keep = ctx.alloc_columns(spec, n + 1, max(16, n))
        if op == S.OP_INDEX:
            self.index = ctx.alloc(4 * (n + 16))

    def request(self):
        return gpu.Request(self.op, self.buf, self.n, spec=self.spec, columns=self.cols, index=self.index,
                           index_cap=self.n + 16 if self.index is not None else 0)

    def outputs(self):
        """Device outputs as host arrays (columns, index)."""
        if self.cols is not None:
            rows = self.cols.rows
            out = [rows]
            for k, c in enumerate(self.spec.columns):
                vals, valid, dbuf = self.keep[k]
                out.append(valid.download(np.uint64, (rows + 63) // 64).tobytes())
                if c.type == S.STR:
                    out.append(vals.download(np.int32, rows + 1).tobytes())
                    out.append(dbuf.download(np.uint8, self.cols.cols[k].data_len).tobytes())
                else:
                    w = 2 if c.type == S.BYTES else 1
                    out.append(vals.download(np.uint64, w * rows).tobytes())
            return out
        return None


def plain(ctx, d: Doc):
    """The plain entry point's result dict and outputs for one document."""
    n = d.n
    if d.op == S.OP_INDEX:
        cnt = ctx.index_structurals(d.spec, d.buf, n, out=d.index, download=False)[1] if n else 0
        return {"status": S.OK, "n_index": cnt}, d.index.download(np.uint32, cnt).tobytes()
    if d.op == S.OP_COLUMNS:
        ctx.parse_columns_into(d.spec, d.buf, n, d.cols)
        return {"status": S.OK}, d.outputs()
    if d.op == S.OP_COLUMNS_STATS:
        st = ctx.parse_columns_with_stats_into(d.spec, d.buf, n, d.cols)
        return {"status": S.OK, "stats": st}, d.outputs()
    if d.op == S.OP_CSV_STATS:
        return {"status": S.OK, "stats": ctx.csv_stats(d.spec, d.buf, n)}, None
    if d.op == S.OP_FIELD_BYTES:
        return {"status": S.OK, "field_bytes": ctx.field_bytes(d.spec, d.buf, n)}, None
    if d.op == S.OP_LOGFMT_PAIRS:
        return {"status": S.OK, "logfmt": ctx.logfmt_pairs(d.buf, n)}, None
    nd = ctx.ndjson_id_score(d.buf, n)   # FALX_E_SCHEMA is returned as invalid > 0
    return {"status": S.E_SCHEMA if nd["invalid"] else S.OK, "ndjson": nd}, None


def _same(a, b):
    """Exact equality, FP64 compared by bits (NaN included)."""
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if isinstance(a, float):
        return np.float64(a).tobytes() == np.float64(b).tobytes()
    return a == b


def _index_out(d: Doc, r):
    return d.index.download(np.uint32, r["n_index"]).tobytes()


def _check(ctx, docs, results, label):
    for i, (d, r) in enumerate(zip(docs, results)):
        got_out = d.outputs() if d.cols is not None else (_index_out(d, r) if d.op == S.OP_INDEX else None)
        want, want_out = plain(ctx, d)
        assert _same(r, want), f"{label} doc {i} (op {d.op}, {d.n} B): {r} != {want}"
        assert _same(got_out, want_out), f"{label} doc {i} (op {d.op}, {d.n} B): outputs differ"


def test_mixed_batch_matches_plain_entry_points(ctx):
    rng = Rng(0xBA7C4)
    b = gpu.Batch(ctx, SCRATCH)
    # Replays of cached launch sequences with fresh documents each time
    # (use_graph = 1: every batch runs as a graph, captured the first time).
    for rep in range(4):
        docs = [Doc(ctx, op, spec, data) for op, spec, data in _docs(rng, 6 << 20)]
        b.submit([d.request() for d in docs])
        results = b.wait()
        _check(ctx, docs, results, f"replay {rep}")


def test_batch_without_graph_and_with_chunks(ctx):
    rng = Rng(0xC4A2)
    b = gpu.Batch(ctx, SCRATCH)
    t = b.tuning()
    t.use_graph = 0
    t.tile_workgroups = 37
    t.unit_workgroups = 5
    for k in range(S.N_SIZE_CLASS):
        t.chunk[k] = 1 + (k * 3) % 7
    b.set_tuning(t)
    docs = [Doc(ctx, op, spec, data) for op, spec, data in _docs(rng, 3 << 20)]
    b.submit([d.request() for d in docs])
    _check(ctx, docs, b.wait(), "no graph")


def test_document_alone_equals_document_in_mix(ctx):
    rng = Rng(0xA10E)
    b = gpu.Batch(ctx, SCRATCH)
    mix = [Doc(ctx, op, spec, data) for op, spec, data in _docs(rng, 2 << 20)]
    b.submit([d.request() for d in mix])
    in_mix = b.wait()
    for i, d in enumerate(mix):
        out_mix = d.outputs() if d.cols is not None else None
        b.submit([d.request()])
        alone = b.wait()[0]
        assert _same(alone, in_mix[i]), f"doc {i}: alone {alone} != in mix {in_mix[i]}"
        if d.cols is not None:
            assert _same(d.outputs(), out_mix), f"doc {i}: outputs differ alone vs in mix"


def test_scratch_budget_refuses_what_does_not_fit(ctx):
    data = datagen.geo_csv(1 << 20)[0]
    buf = ctx.upload(data)
    req = gpu.Request(S.OP_CSV_STATS, buf, int(data.size), spec=S.GEO_LATLON)
    need = req.scratch_bytes()
    b = gpu.Batch(ctx, need + (need // 2))
    b.submit([req, req, req])
    res = b.wait()
    assert res[0]["status"] == S.OK
    assert res[1]["status"] == S.E_BUSY and res[2]["status"] == S.E_BUSY
    assert _same(res[0]["stats"], ctx.csv_stats(S.GEO_LATLON, buf, int(data.size)))
