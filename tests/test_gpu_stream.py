"""Streaming ingestion (falx_stream_*, the reference's StreamParser,
csv_geo.rs:688-807): feeding a document in arbitrary chunks gives exactly
the records of parsing it whole (tests/spans.rs:393-490,
tests/comments.rs:169-183), checked against the oracle's whole-document
tape; and the stream's tapes, rebased and joined, are the whole document's
tape word for word at 64 MiB."""
import numpy as np
import pytest

from falx_amd import datagen, gpu
from falx_amd import spec as S
from falx_amd.records import Parsed
from helpers import Rng, random_doc

pytestmark = pytest.mark.gpu


def batch_records(ora, spec, data: bytes):
    seps, ends = ora.parse(spec, data)
    return [r.fields() for r in Parsed(data, seps, ends, spec).records()]


def stream_records(ctx, spec, data: bytes, cuts, cap=1 << 20):
    st = gpu.Stream(ctx, spec, cap)
    out, prev = [], 0
    try:
        for c in list(cuts) + [len(data)]:
            chunk = data[prev:c]
            buf = ctx.upload(np.frombuffer(chunk, dtype=np.uint8) if chunk else np.zeros(1, np.uint8))
            out += [r.fields() for r in st.feed(buf, len(chunk)).records()]
            buf.free()
            prev = c
        out += [r.fields() for r in st.finish().records()]
    finally:
        st.close()
    return out


@pytest.mark.parametrize("spec,alpha,seed", [
    (S.CSV, b'",\n\rxy', 0x57AE57AE57AE57A1),
    (S.NDJSON, b'\\\\"\n{}x', 0x57AE57AE57AE57A2),
    (S.LOGFMT, b'\\" =\nxy', 0x57AE57AE57AE57A3),
    (S.CSV_HASH_DIALECT, b'"",,\n\n##xy\r', 0x57AE57AE57AE57A4),
], ids=["csv", "ndjson", "logfmt", "csv_hash"])
def test_streaming_matches_batch(ctx, ora, spec, alpha, seed):
    """tests/spans.rs:393-461: random documents, random cut points (chunks
    down to 0 and 1 bytes)."""
    rng = Rng(seed)
    for i in range(150):
        data = random_doc(rng, alpha, rng.below(600))
        cuts = sorted(rng.below(max(1, len(data))) for _ in range(rng.below(12)))
        assert stream_records(ctx, spec, data, cuts) == batch_records(ora, spec, data), (i, data, cuts)


def test_streaming_compaction_large_input(ctx, ora):
    """tests/spans.rs:463-490: 300 000 bytes in feeds of 1024, 4097 and
    65536 bytes."""
    rng = Rng(0xC0DEC0DEC0DEC0DE)
    data = random_doc(rng, b'\\",\n{}xyz0', 300_000)
    for spec in (S.CSV, S.NDJSON):
        want = batch_records(ora, spec, data)
        for feed in (1024, 4097, 65536):
            cuts = list(range(feed, len(data), feed))
            assert stream_records(ctx, spec, data, cuts) == want, (spec.name, feed)


def test_streaming_skips_comments_like_batch(ctx, ora):
    """tests/comments.rs:169-183."""
    data = b"# c1\nfoo,1\n#c2,x\nbar,2\n# trailing"
    spec = S.CSV_HASH_DIALECT
    seps, ends = ora.parse(spec, data)
    want = [r.as_bytes() for r in Parsed(data, seps, ends, spec).records()]
    for feed in (1, 3, 7, 64, 1024):
        st = gpu.Stream(ctx, spec, 4096)
        got = []
        for k in range(0, len(data), feed):
            chunk = data[k:k + feed]
            got += [r.as_bytes() for r in st.feed(ctx.upload(np.frombuffer(chunk, np.uint8)), len(chunk)).records()]
        got += [r.as_bytes() for r in st.finish().records()]
        st.close()
        assert got == want, feed


def test_stream_buffer_capacity(ctx, ora):
    """A chunk that does not fit beside the kept record is refused and the
    stream goes on unchanged."""
    spec = S.CSV
    st = gpu.Stream(ctx, spec, 64)
    try:
        a = b"a,b\nc,"
        assert [r.fields() for r in st.feed(ctx.upload(np.frombuffer(a, np.uint8))).records()] == [[b"a", b"b"]]
        big = np.frombuffer(b"x" * 100, np.uint8)
        with pytest.raises(gpu.FalxError):
            st.feed(ctx.upload(big))
        rest = b"d\ne,f"
        got = [r.fields() for r in st.feed(ctx.upload(np.frombuffer(rest, np.uint8))).records()]
        got += [r.fields() for r in st.finish().records()]
        assert got == [[b"c", b"d"], [b"e", b"f"]]
    finally:
        st.close()


def test_stream_tape_equals_whole_document_tape(ctx):
    """64 MiB of geo CSV in 4 MiB feeds: the feeds' tapes rebased to stream
    offsets and joined equal falx_parse of the whole document."""
    data, truth = datagen.geo_csv(64 * 1024 * 1024 + 7)
    whole = ctx.upload(data)
    seps_w, ends_w = ctx.parse(S.GEO_LATLON, whole, data.size)
    lib = gpu.load()
    cs = S.GEO_LATLON.to_c()
    import ctypes as C
    h = C.c_void_p()
    assert lib.falx_stream_create(ctx.h, C.byref(cs), 16 << 20, C.byref(h)) == S.OK
    d_seps, d_ends = ctx.alloc(4 << 20), ctx.alloc(8 << 20)
    seps, ends = [], []
    base, sep_base = 0, 0
    feed = 4 << 20
    try:
        for k in range(0, data.size + 1, feed):
            out = S.CStreamOut()
            if k < data.size:
                n = min(feed, data.size - k)
                st = lib.falx_stream_feed(h, whole.ptr + k, n, d_seps.ptr, d_seps.nbytes // 4, d_ends.ptr,
                                          d_ends.nbytes // 8, C.byref(out))
            else:
                st = lib.falx_stream_finish(h, d_seps.ptr, d_seps.nbytes // 4, d_ends.ptr, d_ends.nbytes // 8,
                                            C.byref(out))
            assert st == S.OK
            s_ = d_seps.download(np.uint32, out.n_seps).astype(np.uint64) + np.uint64(base)
            e_ = d_ends.download(np.uint64, out.n_ends)
            e_ = ((e_ >> np.uint64(32)) + np.uint64(sep_base)) << np.uint64(32) | ((e_ & np.uint64(0xFFFFFFFF))
                                                                                  + np.uint64(base))
            seps.append(s_)
            ends.append(e_)
            base += out.data_end
            sep_base += out.n_seps
    finally:
        lib.falx_stream_destroy(h)
    assert np.array_equal(np.concatenate(seps), seps_w.astype(np.uint64))
    assert np.array_equal(np.concatenate(ends), ends_w)
    assert base == data.size
