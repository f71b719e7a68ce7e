"""BGZF inflated on the GPU (falx_bgzf_inflate) against Python's zlib and the
oracle's DEFLATE restatement: every block type zlib writes, edge lengths,
long overlapping copies, corrupted blocks (reported by index, as the oracle
finds them), and the parse of the inflated bytes in HBM against the oracle's
parse of the original document."""
import os
import zlib

import numpy as np
import pytest

from falx_amd import bgzf, datagen, gpu
from falx_amd import spec as S
from helpers import Rng, assert_stats_equal, random_doc

pytestmark = pytest.mark.gpu

STRATEGIES = [(0, zlib.Z_DEFAULT_STRATEGY), (1, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_DEFAULT_STRATEGY),
              (9, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_FIXED), (6, zlib.Z_HUFFMAN_ONLY), (6, zlib.Z_RLE)]


def gpu_inflate(ctx, comp: bytes):
    blocks, total = bgzf.scan(comp)
    d_comp = ctx.upload(np.frombuffer(comp, dtype=np.uint8))
    d_out = ctx.alloc(max(1, total))
    st, bad = bgzf.inflate(ctx, d_comp, len(comp), blocks, d_out)
    out = d_out.download(np.uint8, total).tobytes() if st == S.OK else None
    d_comp.free()
    d_out.free()
    return st, bad, out


def docs():
    rng = Rng(0xB62F0000D0D0CAFE)
    yield b"a"
    yield b"ab" * 60000
    yield bytes(range(256)) * 600
    yield os.urandom(150000)
    yield random_doc(rng, b'",\n\rxyz0123456789.', 300_000)
    d, _ = datagen.geo_csv(1 << 20)
    yield d.tobytes()


@pytest.mark.parametrize("level,strategy", STRATEGIES)
def test_inflate_matches_zlib(ctx, level, strategy):
    for i, data in enumerate(docs()):
        st, bad, out = gpu_inflate(ctx, bgzf.bgzip(data, level, strategy))
        assert st == S.OK and out == data, (level, strategy, i, bad)


def test_block_sizes(ctx):
    """Blocks of 1..64 KiB of output, and a stream of one-byte blocks."""
    rng = Rng(0x5EED5EED5EED0003)
    data = random_doc(rng, b"abcabcabd,\n", 400_000)
    for block in (1, 7, 1000, 0xFF00, 65536):
        st, bad, out = gpu_inflate(ctx, bgzf.bgzip(data[: 40_000 if block < 100 else None], 6, block=block))
        assert st == S.OK and out == data[: 40_000 if block < 100 else None], block


def test_corrupted_blocks_reported(ctx, ora):
    """Bit flips in some blocks: the GPU fails exactly when the oracle finds a
    bad block (a malformed stream or an inflated size other than ISIZE), and
    reports the first one."""
    rng = Rng(0x0BADF11E5EED0002)
    data = b"".join(b"%d,%d;" % (i, i * i % 1000) for i in range(100_000))
    comp = bgzf.bgzip(data, 6)
    blocks, total = bgzf.scan(comp)
    for trial in range(40):
        q = bytearray(comp)
        for _ in range(1 + rng.below(3)):
            k = rng.below(len(blocks))
            b = blocks[k]
            bit = rng.below(8 * int(b["payload_len"]))
            q[int(b["payload"]) + bit // 8] ^= 1 << (bit % 8)
        first = next((k for k, b in enumerate(blocks) if not block_ok(ora, q, b)), None)
        st, bad, out = gpu_inflate(ctx, bytes(q))
        if first is None:
            want = ora.bgzf_inflate(bytes(q), blocks, total, 1)
            assert st == S.OK and out == want.tobytes(), trial
        else:
            assert st == S.E_INVALID and bad == first, (trial, bad, first)


def block_ok(ora, comp, b) -> bool:
    p0 = int(b["payload"])
    out = ora.inflate(bytes(comp[p0:p0 + int(b["payload_len"])]), int(b["isize"]))
    return out is not None and len(out) == int(b["isize"])


def test_parse_of_inflated_bytes(ctx, ora):
    """bgzip'd geo CSV: inflate in HBM, then the headline columns + sums and
    the structural index of the inflated bytes equal the oracle's on the
    original document."""
    data, truth = datagen.geo_csv(16 * 1024 * 1024 + 5)
    comp = bgzf.bgzip(data.tobytes(), 6, threads=8)
    d = bgzf.decompress(ctx, comp)
    spec = S.GEO_LATLON
    cols, keep = ctx.alloc_columns(spec, truth["rows"], 16)
    got = ctx.parse_columns_with_stats_into(spec, d, data.size, cols)
    want = ora.csv_stats(spec, data)
    assert_stats_equal(got, want, spec)
    assert np.array_equal(ctx.index_structurals(spec, d, data.size), ora.index_structurals(spec, data, 4))
