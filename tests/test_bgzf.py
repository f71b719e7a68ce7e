"""CPU tests of BGZF: the header scan (falx_bgzf_scan, bgzf.rs:61-106, a host
function of the library) and the oracle's DEFLATE restatement, pinned
against Python's zlib (the reference's flate2 backend family) on every
block type zlib writes, hand-made stored blocks, edge lengths and
corrupted streams."""
import os
import struct
import zlib

import numpy as np
import pytest

from falx_amd import bgzf, gpu
from falx_amd import spec as S
from helpers import Rng, random_doc

STRATEGIES = [(0, zlib.Z_DEFAULT_STRATEGY), (1, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_DEFAULT_STRATEGY),
              (9, zlib.Z_DEFAULT_STRATEGY), (6, zlib.Z_FIXED), (6, zlib.Z_HUFFMAN_ONLY), (6, zlib.Z_RLE),
              (9, zlib.Z_FILTERED)]


def corpus():
    rng = Rng(0xB62F0000D0D0CAFE)
    yield b""
    yield b"a"
    yield b"ab" * 40000                          # long overlapping copies (distance 2)
    yield bytes(range(256)) * 300
    yield os.urandom(70000)                      # incompressible: stored blocks
    yield random_doc(rng, b'",\n\rxyz0123456789.', 200_000)
    yield b"".join(b"%d,city %d,%.6f,%.6f\n" % (i, i % 97, (i * 7919 % 180000) / 1000 - 90, i / 3.0)
                   for i in range(20000))


def raw(data: bytes, level: int, strategy: int) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    return c.compress(data) + c.flush()


@pytest.mark.parametrize("level,strategy", STRATEGIES)
def test_inflate_matches_zlib(ora, level, strategy):
    for data in corpus():
        p = raw(data, level, strategy)
        assert zlib.decompress(p, -15) == data
        assert ora.inflate(p, len(data)) == data, (level, strategy, len(data))


def test_inflate_hand_made_stored_blocks(ora):
    """Non-final then final stored blocks (RFC 1951 3.2.4), an empty one
    between them, LEN/NLEN checked."""
    a, b = b"hello ", b"world"
    s = bytes([0]) + struct.pack("<HH", len(a), len(a) ^ 0xFFFF) + a
    s += bytes([0]) + struct.pack("<HH", 0, 0xFFFF)
    s += bytes([1]) + struct.pack("<HH", len(b), len(b) ^ 0xFFFF) + b
    assert zlib.decompress(s, -15) == a + b
    assert ora.inflate(s, 64) == a + b
    bad = bytes([1]) + struct.pack("<HH", 3, 3) + b"abc"
    assert ora.inflate(bad, 64) is None


def test_inflate_corrupted_streams_agree_with_zlib(ora):
    """Single bit flips and truncations: the oracle fails exactly when zlib
    does, and otherwise gives zlib's bytes."""
    rng = Rng(0x0BADF11E5EED0001)
    data = b"".join(b"%d,%d;" % (i, i * i % 1000) for i in range(3000))
    for level, strategy in STRATEGIES:
        p = bytearray(raw(data, level, strategy))
        for _ in range(60):
            q = bytearray(p)
            bit = rng.below(8 * len(q))
            q[bit // 8] ^= 1 << (bit % 8)
            if rng.below(4) == 0:
                q = q[: rng.below(len(q))]
            try:
                want = zlib.decompress(bytes(q), -15)
            except zlib.error:
                want = None
            got = ora.inflate(bytes(q), 1 << 20)
            assert got == want, (level, strategy, bit)


def test_scan_blocks_and_eof():
    data = b"x" * 200_000
    comp = bgzf.bgzip(data)
    blocks, total = bgzf.scan(comp)
    assert total == len(data) and len(blocks) == (len(data) + 0xFEFF) // 0xFF00
    assert list(blocks["isize"][:-1]) == [0xFF00] * (len(blocks) - 1)
    assert np.array_equal(blocks["out"], np.concatenate([[0], np.cumsum(blocks["isize"])[:-1]]))
    for b in blocks:
        assert zlib.decompress(comp[b["payload"]:b["payload"] + b["payload_len"]], -15) == b"x" * int(b["isize"])
    assert bgzf.scan(bgzf.EOF_MEMBER)[1] == 0 and len(bgzf.scan(bgzf.EOF_MEMBER)[0]) == 0


@pytest.mark.parametrize("mutate", ["magic", "no_extra", "no_bc", "truncated", "short_bsize"])
def test_scan_rejects_malformed_headers(mutate):
    comp = bytearray(bgzf.bgzip(b"abc,def\n" * 1000))
    if mutate == "magic":
        comp[1] = 0x8C
    elif mutate == "no_extra":
        comp[3] = 0
    elif mutate == "no_bc":
        comp[12] = ord("X")
    elif mutate == "truncated":
        comp = comp[:-40]
    else:
        comp[16:18] = struct.pack("<H", 10)
    with pytest.raises(gpu.FalxError) as e:
        bgzf.scan(bytes(comp))
    assert e.value.status == S.E_INVALID


def test_oracle_bgzf_inflate_blocks(ora):
    data = b"".join(corpus())
    comp = bgzf.bgzip(data, threads=4)
    blocks, total = bgzf.scan(comp)
    for threads in (1, 3, 8):
        assert ora.bgzf_inflate(comp, blocks, total, threads).tobytes() == data
