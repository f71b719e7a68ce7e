"""Test helpers: seeded generators (xorshift64*, as tests/common/mod.rs:1-16
of the reference) and exact comparisons of column results."""
from __future__ import annotations

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
