"""CPU tests of the nested-tape oracle (ora_parse_nested, the stack builder of
kernels/json.rs:636-760 and the chunked driver of json.rs:996-1108) and of
the host navigation over a tape (falx_amd/nested.py, json.rs:762-993).

Pins: the reference's own cases (tests/nested.rs: error positions, multiple
top-level documents, empty inputs, brackets in strings, deep nesting), a
structural differential against Python's json module on random documents
(compact and pretty), the block-seam sweep, partner invariants, brute force
over every short bracket string against an independent algorithm (depth
arithmetic instead of a stack), and parallel == serial."""
import itertools
import json

import numpy as np
import pytest

from falx_amd import spec as S
from falx_amd.nested import NO_PARTNER, UNCLOSED_OPEN, UNMATCHED_CLOSE, Nested
from helpers import Rng, json_text, json_value, nest_by_depth

JSON = S.Spec("json", b"{}[],:", ord('"'), ord("\\"), None, ())


def doc_of(ora, data: bytes, threads: int = 1) -> Nested:
    tape, err, pos = ora.parse_nested(data, threads)
    return Nested(data, tape, err, pos)


def assert_matches(node, value, ctx):
    """tests/nested.rs:70-111: containers match kind and arity, keys and
    values alternate in objects, scalar spans re-parse to the value."""
    if isinstance(value, dict):
        assert node.open() == ord("{"), ctx
        items = list(node.items())
        assert len(items) == 2 * len(value), ctx
        for i, (k, v) in enumerate(value.items()):
            assert items[2 * i].open() is None, ctx
            assert json.loads(items[2 * i].bytes()) == k, ctx
            assert_matches(items[2 * i + 1], v, ctx)
    elif isinstance(value, list):
        assert node.open() == ord("["), ctx
        items = list(node.items())
        assert len(items) == len(value), ctx
        for it, v in zip(items, value):
            assert_matches(it, v, ctx)
    else:
        assert node.open() is None, ctx
        assert json.loads(node.bytes()) == value, ctx


def assert_doc_matches(ora, text: bytes, value, ctx):
    doc = doc_of(ora, text)
    assert doc.error is None, ctx
    items = list(doc.items())
    assert len(items) == 1, ctx
    assert_matches(items[0], value, ctx)


def test_error_positions(ora):
    """tests/nested.rs:204-224."""
    cases = [(b'{"a": [1, 2}', UNMATCHED_CLOSE, 11), (b"[1, 2", UNCLOSED_OPEN, 0), (b"1]", UNMATCHED_CLOSE, 1),
             (b"}", UNMATCHED_CLOSE, 0), (b"[[]", UNCLOSED_OPEN, 0), (b"[ {", UNCLOSED_OPEN, 2)]
    for data, kind, pos in cases:
        assert doc_of(ora, data).error == (kind, pos), data


def test_multiple_top_level_documents(ora):
    """tests/nested.rs:152-164."""
    doc = doc_of(ora, b' {"a": 1} [2, 3]\n{"b": [4]} 5 ')
    assert doc.error is None
    items = list(doc.items())
    assert [it.open() for it in items] == [ord("{"), ord("["), ord("{"), None]
    assert items[3].bytes() == b"5"


def test_empty_and_scalar_inputs(ora):
    """tests/nested.rs:166-180."""
    for empty in [b"", b"   ", b" \t\r\n "]:
        doc = doc_of(ora, empty)
        assert doc.error is None and len(doc.tape) == 0 and list(doc.items()) == []
    doc = doc_of(ora, b" 42 ")
    items = list(doc.items())
    assert doc.error is None and len(items) == 1 and items[0].bytes() == b"42"


def test_empty_containers_and_inert_brackets(ora):
    """tests/nested.rs:182-186, 226-233."""
    text = b"[[], {}, [{}, [[]]]]"
    assert_doc_matches(ora, text, json.loads(text), "empty containers")
    doc = doc_of(ora, rb'["a}b]{", "\"["]')
    assert doc.error is None
    arr = next(doc.items())
    assert [n.bytes() for n in arr.items()] == [rb'"a}b]{"', rb'"\"["']


def test_deep_nesting(ora):
    """tests/nested.rs:188-210: the outermost pair is the first and last tape
    entry; an iterative walk reaches the full depth."""
    depth = 100_000
    doc = doc_of(ora, b"[" * depth + b"]" * depth)
    assert doc.error is None
    t = doc.tape
    assert len(t) == 2 * depth and t[0] >> 32 == len(t) - 1 and t[-1] >> 32 == 0
    node, d = next(doc.items()), 1
    while True:
        child = next(node.items(), None)
        if child is None:
            break
        assert child.open() == ord("[")
        node, d = child, d + 1
    assert d == depth


def test_differential_vs_json_module(ora):
    """tests/nested.rs:113-131 with Python's json in place of serde_json:
    400 compact and 200 pretty random documents."""
    rng = Rng(0xBEEFCAFE12345678)
    for r in range(400):
        v = json_value(rng, 5)
        assert_doc_matches(ora, json_text(v), v, f"compact {r}")
    rng = Rng(0x00DDBA11F00D5EED)
    for r in range(200):
        v = json_value(rng, 5)
        assert_doc_matches(ora, json_text(v, pretty=True), v, f"pretty {r}")


def test_block_seam_sweep(ora):
    """tests/nested.rs:133-150: quotes, escapes and brackets on every 64-byte
    seam position."""
    value = json.loads(r'{"k\\\\": ["va\"l", -3.5, {"x": [], "y\\": [null, true]}], "": {}}')
    text = json_text(value)
    for pad in range(130):
        assert_doc_matches(ora, b" " * pad + text, value, f"pad {pad}")


def test_tape_partner_invariants(ora):
    """tests/nested.rs:235-265: partners are mutual, opens precede closes,
    and every pair encloses a range that nests with the others."""
    rng = Rng(0x7777AAAA3333DDDD)
    for _ in range(100):
        text = json_text(json_value(rng, 6))
        doc = doc_of(ora, text)
        t = doc.tape
        for i, e in enumerate(t):
            pos, p = e & 0xFFFFFFFF, e >> 32
            if text[pos] in b"{[":
                assert p != NO_PARTNER and p > i and t[p] >> 32 == i
                assert text[t[p] & 0xFFFFFFFF] == {ord("{"): ord("}"), ord("["): ord("]")}[text[pos]]
                # nothing between them pairs outside them
                for k in range(i + 1, p):
                    if text[t[k] & 0xFFFFFFFF] in b"{[]}":
                        assert i < (t[k] >> 32) < p
            elif text[pos] in b"}]":
                assert t[p] >> 32 == i and p < i


@pytest.mark.parametrize("alpha,n", [(b"[]{},", 7), (b'[]{"\\x', 6)], ids=["brackets", "strings"])
def test_brute_force_against_depth_arithmetic(ora, alpha, n):
    """Every string over the alphabet up to length n: the oracle's stack
    tape and error equal the depth-arithmetic construction (helpers
    nest_by_depth) over the scalar index (scalar.rs semantics)."""
    for length in range(n + 1):
        for tup in itertools.product(alpha, repeat=length):
            data = bytes(tup)
            tape, err, pos = ora.parse_nested(data)
            want = nest_by_depth(data, ora.index_scalar(JSON, data))
            assert ([int(x) for x in tape], err, pos) == want, data


def test_random_malformed_against_depth_arithmetic(ora):
    """Longer random bracket documents (most malformed somewhere) against the
    depth-arithmetic construction."""
    rng = Rng(0x5EED0FBADC0FFEE5)
    for _ in range(300):
        data = bytes(b"[]{},x "[rng.below(7)] for _ in range(rng.below(400)))
        tape, err, pos = ora.parse_nested(data)
        assert ([int(x) for x in tape], err, pos) == nest_by_depth(data, ora.index_scalar(JSON, data)), data


def par_docs():
    rng = Rng(0x13579BDF2468ACE0)
    yield "repeated", b" \n".join([json_text(json_value(rng, 6))] * 300)
    yield "cross_chunk", b"[" + b",".join([b'[{"k": [1, 2]}]'] * 50_000) + b"]"
    yield "deep", b"[" * 200_000 + b"]" * 200_000
    unit = rb'{"a": "text [with] {brackets} and \\\\ runs \" inside", "b": [1]}'
    yield "boundary", b"\n".join([unit] * 4000)
    filler = b'{"k": [1, 2, 3]} ' * 5000
    yield "mid_close", filler + b"]" + filler
    yield "mid_open", filler + b"[ " + filler
    yield "mismatch", filler + b"[1, 2}} " + filler
    yield "tiny", b"]"


@pytest.mark.parametrize("name,data", list(par_docs()), ids=lambda x: x if isinstance(x, str) else "")
def test_parallel_matches_serial(ora, name, data):
    """tests/nested.rs:283-398: the chunked driver gives the serial tape and
    error at every thread count, malformed input included."""
    t1, e1, p1 = ora.parse_nested(data, 1)
    for threads in (2, 3, 5, 8, 16, 17):
        t, e, p = ora.parse_nested(data, threads)
        assert (e, p) == (e1, p1), (name, threads)
        assert np.array_equal(t, t1), (name, threads)
    if name.startswith("mid") or name in ("mismatch", "tiny"):
        assert e1 != 0


def collect_scalars(node, out):
    if node.open() is not None:
        for c in node.items():
            collect_scalars(c, out)
    else:
        out.append(node.bytes())


def test_scalars_flat_matches_recursive(ora):
    """tests/nested.rs:413-431."""
    rng = Rng(0x5CA1AB1E5EED0001)
    cases = [b"", b" ", b"42", b"  7  ", b"{}", b"[]", b"[1,[2,[3,[4]]],5]", rb'{"a":1,"b":[2,3],"c":{"d":4}}',
             rb'{"x":{},"y":[],"z":"s"}']
    cases += [json_text(json_value(rng, 4)) for _ in range(600)]
    for text in cases:
        doc = doc_of(ora, text)
        if doc.error is not None:
            continue
        rec = []
        for it in doc.items():
            collect_scalars(it, rec)
        assert list(doc.scalars()) == rec, text
