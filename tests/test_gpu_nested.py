"""falx_parse_nested on the GPU against the oracle's nested tape
(ora_parse_nested, kernels/json.rs:636-760): the same tape words, tape
length and error on the reference's cases, random documents, documents
whose brackets span many tiles and every level of the matcher's min tree,
malformed documents (errors far from their opens), and generated NDJSON."""
import numpy as np
import pytest

from falx_amd import datagen
from falx_amd import spec as S
from falx_amd.nested import Nested
from helpers import Rng, json_text, json_value

pytestmark = pytest.mark.gpu


def check(ctx, ora, data: bytes, label=""):
    buf = ctx.upload(np.frombuffer(data, dtype=np.uint8) if data else np.zeros(1, np.uint8))
    got, n, err, pos = ctx.parse_nested(buf, len(data))
    want, werr, wpos = ora.parse_nested(data)
    assert (n, err, pos) == (want.size, werr, wpos), label
    assert np.array_equal(got, want), label
    buf.free()
    return got, err, pos


def test_reference_cases(ctx, ora):
    """tests/nested.rs:152-233 inputs."""
    cases = [b'{"a": [1, 2}', b"[1, 2", b"1]", b"}", b"[[]", b"[ {", b' {"a": 1} [2, 3]\n{"b": [4]} 5 ', b"",
             b"   ", b" 42 ", b"[[], {}, [{}, [[]]]]", rb'["a}b]{", "\"["]', b"[" * 100_000 + b"]" * 100_000]
    for c in cases:
        check(ctx, ora, c, c[:40])


def test_random_documents_and_seams(ctx, ora):
    """Random compact and pretty documents one by one, every 64-byte seam
    offset of one document, and all of them concatenated (many tiles)."""
    rng = Rng(0xBEEFCAFE12345678)
    docs = [json_text(json_value(rng, 6), pretty=(i % 4 == 3)) for i in range(200)]
    for i, d in enumerate(docs):
        check(ctx, ora, d, f"doc {i}")
    seam = json_text(json_value(Rng(7), 5))
    for pad in range(0, 130, 3):
        check(ctx, ora, b" " * pad + seam, f"pad {pad}")
    big, _, _ = check(ctx, ora, b" \n".join(docs * 200), "concatenated")
    assert big.size > 16 * 4096


def test_spanning_and_malformed(ctx, ora):
    """tests/nested.rs:283-398 documents: containers across every tile, pure
    depth, escapes across tiles, errors in mid-stream, a wrong kind."""
    unit = rb'{"a": "text [with] {brackets} and \\\\ runs \" inside", "b": [1]}'
    filler = b'{"k": [1, 2, 3]} ' * 5000
    docs = {
        "cross_tile": b"[" + b",".join([b'[{"k": [1, 2]}]'] * 50_000) + b"]",
        "deep": b"[" * 200_000 + b"]" * 200_000,
        "boundary": b"\n".join([unit] * 4000),
        "mid_close": filler + b"]" + filler,
        "mid_open": filler + b"[ " + filler,
        "mismatch": filler + b"[1, 2}} " + filler,
        "far_mismatch": b"[" + filler + b"}" + filler,
        "far_mismatch_deep": b"{" * 70_000 + b"[" + b"]" * 1 + b"}" * 69_999 + b"]",
        "tiny": b"]",
    }
    for name, d in docs.items():
        _, err, _ = check(ctx, ora, d, name)
        if "mid" in name or "mismatch" in name or name == "tiny":
            assert err != 0, name


@pytest.mark.parametrize("depth", [3_000_000, 10_000_000])
def test_deep_tree_levels(ctx, ora, depth):
    """Opens whose closes lie thousands of tiles away (the min tree's upper
    levels): balanced, one short (unclosed), and closed by the wrong kind
    at the far end."""
    check(ctx, ora, b"[" * depth + b"]" * depth, f"balanced {depth}")
    check(ctx, ora, b"[" * depth + b"]" * (depth - 1), f"unclosed {depth}")
    check(ctx, ora, b"[" * depth + b"]" * (depth - 1) + b"}", f"wrong kind {depth}")


def test_random_bracket_streams(ctx, ora):
    """Random streams of brackets, separators and string runs, 1 MiB each:
    random depths across tiles, errors at random places."""
    for seed in range(12):
        r = np.random.default_rng(seed)
        n = 1 << 20
        if seed % 3 == 0:
            # balanced random walk: errors only if a close goes below zero
            steps = r.integers(0, 2, n // 2)
            data = np.where(steps == 1, ord("["), ord("]")).astype(np.uint8)
            data[r.integers(0, data.size, data.size // 8)] = ord(",")
        else:
            alpha = np.frombuffer(b'[[[{{,,: x"]]]}}', dtype=np.uint8)
            data = alpha[r.integers(0, alpha.size, n)]
        d = data.tobytes()
        check(ctx, ora, d, f"stream {seed}")


def test_generated_ndjson(ctx, ora):
    """The NDJSON generator's documents: one top-level object per line."""
    data, truth = datagen.ndjson(8 * 1024 * 1024 + 11)
    got, err, _ = check(ctx, ora, data.tobytes(), "ndjson")
    assert err == 0
    small, _ = datagen.ndjson(64 * 1024)
    t, _, _, _ = ctx.parse_nested(ctx.upload(small), small.size)
    items = list(Nested(small.tobytes(), t).items())
    assert len(items) == ora.ndjson_lines(small) and all(it.open() == ord("{") for it in items)


def test_capacity(ctx, ora):
    data = b'[{"a": 1}, {"b": [2, 3]}]' * 100
    want, _, _ = ora.parse_nested(data)
    buf = ctx.upload(np.frombuffer(data, dtype=np.uint8))
    small = ctx.alloc(8 * 10)
    st, n, _, _ = ctx.parse_nested_into(buf, len(data), small)
    assert st == S.E_CAPACITY and n == want.size
