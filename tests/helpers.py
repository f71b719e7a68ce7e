"""Test helpers: seeded generators (xorshift64*, as tests/common/mod.rs:1-16
of the reference) and exact comparisons of column results."""
from __future__ import annotations

import json

import numpy as np


class Rng:
    def __init__(self, seed: int):
        self.s = seed & (2**64 - 1)

    def next(self) -> int:
        s = self.s
        s ^= s >> 12
        s ^= (s << 25) & (2**64 - 1)
        s ^= s >> 27
        self.s = s
        return (s * 0x2545F4914F6CDD1D) & (2**64 - 1)

    def below(self, n: int) -> int:
        return self.next() % n


def random_doc(rng: Rng, alphabet: bytes, length: int) -> bytes:
    return bytes(alphabet[rng.below(len(alphabet))] for _ in range(length))


def dialect_alphabet(spec) -> bytes:
    """codegen.rs:402-423: structural bytes, quote x2, escape x3, comment x2, filler."""
    a = bytearray(spec.structural)
    if spec.quote is not None:
        a += bytes([spec.quote]) * 2
    if spec.escape is not None:
        a += bytes([spec.escape]) * 3
    if spec.comment is not None:
        a += bytes([spec.comment]) * 2
    a += b"xy"
    return bytes(a)


def assert_columns_equal(got, want, spec, label=""):
    assert got["rows"] == want["rows"], f"{label}: rows {got['rows']} != {want['rows']}"
    for k, c in enumerate(spec.columns):
        g, w = got["cols"][k], want["cols"][k]
        assert np.array_equal(g["valid"], w["valid"]), f"{label}: col {k} validity differs"
        if "values" in w:
            gv, wv = g["values"], w["values"]
            if gv.dtype == np.float64:
                gv, wv = gv.view(np.uint64), wv.view(np.uint64)
            if not np.array_equal(gv, wv):
                bad = np.nonzero(gv != wv)[0][:5] if gv.shape == wv.shape else "shape"
                raise AssertionError(f"{label}: col {k} values differ at {bad}")
        else:
            assert np.array_equal(g["offsets"], w["offsets"]), f"{label}: col {k} offsets differ"
            assert np.array_equal(g["data"], w["data"]), f"{label}: col {k} data differs"


def assert_stats_equal(got, want, spec, label=""):
    assert got["records"] == want["records"], f"{label}: records {got['records']} != {want['records']}"
    for k in range(len(spec.columns)):
        g, w = got["cols"][k], want["cols"][k]
        for key in ("values", "checksum", "bytes", "fixed6_sum", "fixed6_cells"):
            assert g[key] == w[key], f"{label}: col {k} {key} {g[key]} != {w[key]}"
        # FP64 sums: blocked tree (GPU) vs sequential fold (oracle) differ
        # only by rounding; the exact integer sums above pin the terms.
        if not np.isfinite(w["sum"]) or not np.isfinite(g["sum"]):
            # A NaN or infinite cell makes both sums NaN or the same infinity.
            assert (np.isnan(g["sum"]) and np.isnan(w["sum"])) or g["sum"] == w["sum"], \
                f"{label}: col {k} sum {g['sum']} vs {w['sum']}"
            continue
        n = max(1, w["values"])
        tol = 1e-7 * max(1.0, abs(w["sum"])) + 1e-6 * n ** 0.5
        assert abs(g["sum"] - w["sum"]) <= tol, f"{label}: col {k} sum {g['sum']} vs {w['sum']}"


# ---- JSON documents for the nested tape (tests/nested.rs:13-67) -------------
JSON_STRING_POOL = [
    "plain", "", "with space", "br{ack}ets", "[inside]", "a,b:c,d", 'quote"inside',
    "back\\slash", "run\\\\\\of\\\\backslashes", "ends with \\", "\\", "\\\\",
    "tab\tand newline\n", "unicode é文\U0001F980", '{"looks":["like","json"]}',
]


def json_value(rng: Rng, depth: int):
    """A random JSON value: scalars, strings from the pool, arrays and objects
    of up to 4 members with unique keys."""
    choice = rng.below(5) if depth == 0 else rng.below(7)
    if choice == 0:
        return None
    if choice == 1:
        return rng.below(2) == 0
    if choice == 2:
        return (rng.next() % (2 ** 63)) % 100000 - (50000 if rng.below(2) else 0)
    if choice == 3:
        return ((rng.next() % (2 ** 63)) % 100000) / 128.0
    if choice == 4:
        return JSON_STRING_POOL[rng.below(len(JSON_STRING_POOL))]
    if choice == 5:
        return [json_value(rng, depth - 1) for _ in range(rng.below(5))]
    return {f"k{i} {JSON_STRING_POOL[rng.below(len(JSON_STRING_POOL))]}": json_value(rng, depth - 1)
            for i in range(rng.below(5))}


def json_text(value, pretty: bool = False) -> bytes:
    if pretty:
        return json.dumps(value, indent=2, ensure_ascii=False).encode()
    return json.dumps(value, separators=(",", ":"), ensure_ascii=False).encode()


def nest_by_depth(data: bytes, positions) -> tuple:
    """The nested tape by depth arithmetic, an algorithm independent of the
    reference's stack: with depth(i) = opens - closes through entry i, an open
    of depth L is closed by the first later entry of depth below L; the tape
    stops at the first entry of negative depth or the first close of the
    wrong kind; an open with no such entry is unclosed. Returns (tape list,
    error, error_pos) in falx_parse_nested's terms."""
    pos = [int(p) for p in positions]
    n = len(pos)
    kind = [data[p] for p in pos]
    depth, d = [], 0
    for b in kind:
        d += 1 if b in b"{[" else -1 if b in b"}]" else 0
        depth.append(d)
    cut = next((i for i in range(n) if depth[i] < 0), n)
    partner = {}
    for i in range(n):
        if kind[i] in b"{[":
            j = next((k for k in range(i + 1, n) if depth[k] < depth[i]), None)
            if j is not None:
                partner[i] = j
                if (kind[i], kind[j]) not in ((ord("{"), ord("}")), (ord("["), ord("]"))):
                    cut = min(cut, j)
    tape = []
    for i in range(min(cut, n)):
        if kind[i] in b"{[":
            j = partner.get(i)
            tape.append(((j if j is not None and j < cut else 0xFFFFFFFF) << 32) | pos[i])
        elif kind[i] in b"}]":
            o = max(k for k, v in partner.items() if v == i)
            tape.append((o << 32) | pos[i])
        else:
            tape.append(pos[i])
    if cut < n:
        return tape, 1, pos[cut]
    unclosed = [i for i in range(n) if kind[i] in b"{[" and i not in partner]
    if unclosed:
        return tape, 2, pos[max(unclosed)]
    return tape, 0, 0
