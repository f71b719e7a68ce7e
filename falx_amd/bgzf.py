"""BGZF on the device: block boundaries (falx_bgzf_scan, bgzf.rs:61-106),
inflation into HBM (falx_bgzf_inflate, bgzf.rs:194-290) and the parse of
the inflated bytes; plus `bgzip`, which writes BGZF the way bgzip does
(gzip members of at most 0xFF00 input bytes with the `BC` extra subfield,
then the empty EOF member) using Python's zlib for the DEFLATE payloads.
"""
from __future__ import annotations

import ctypes as C
import struct
import zlib
from typing import Optional, Tuple

import numpy as np

from . import gpu
from . import spec as S

BLOCK_DTYPE = np.dtype([("payload", "<u8"), ("payload_len", "<u4"), ("isize", "<u4"), ("out", "<u8")])
EOF_MEMBER = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")


def _member(chunk: bytes, level: int, strategy: int) -> bytes:
    c = zlib.compressobj(level, zlib.DEFLATED, -15, 9, strategy)
    payload = c.compress(chunk) + c.flush()
    bsize = 12 + 6 + len(payload) + 8 - 1
    return (struct.pack("<BBBBIBBH", 0x1F, 0x8B, 8, 4, 0, 0, 0xFF, 6) + struct.pack("<BBHH", ord("B"), ord("C"), 2, bsize)
            + payload + struct.pack("<II", zlib.crc32(chunk) & 0xFFFFFFFF, len(chunk)))


def bgzip(data, level: int = 6, strategy: int = zlib.Z_DEFAULT_STRATEGY, block: int = 0xFF00,
          eof: bool = True, threads: int = 1) -> bytes:
    """BGZF members of `block` input bytes each (zlib releases the GIL, so
    `threads` compress members concurrently), then the EOF member."""
    mv = memoryview(data)
    chunks = [bytes(mv[k:k + block]) for k in range(0, len(mv), block)]
    if threads > 1 and len(chunks) > 1:
        from concurrent.futures import ThreadPoolExecutor
        with ThreadPoolExecutor(threads) as ex:
            members = list(ex.map(lambda ch: _member(ch, level, strategy), chunks))
    else:
        members = [_member(ch, level, strategy) for ch in chunks]
    return b"".join(members) + (EOF_MEMBER if eof else b"")


def scan(comp: bytes) -> Tuple[np.ndarray, int]:
    """The block table and the inflated size (FalxError on a bad header)."""
    arr = np.frombuffer(comp, dtype=np.uint8)
    cap = max(16, arr.size // 18 + 1)
    blocks = np.zeros(cap, dtype=BLOCK_DTYPE)
    n, total = C.c_uint64(), C.c_uint64()
    st = gpu.load().falx_bgzf_scan(arr.ctypes.data if arr.size else None, arr.size, blocks.ctypes.data, cap,
                                   C.byref(n), C.byref(total))
    gpu._check("falx_bgzf_scan", st)
    return blocks[: n.value].copy(), total.value


def inflate(ctx: gpu.Context, d_comp: gpu.DeviceBuffer, comp_len: int, blocks: np.ndarray,
            d_out: gpu.DeviceBuffer) -> Tuple[int, Optional[int]]:
    """(status, first bad block or None): the blocks inflated into d_out."""
    b = np.ascontiguousarray(blocks, dtype=BLOCK_DTYPE)
    bad = C.c_uint64()
    st = gpu.load().falx_bgzf_inflate(ctx.h, d_comp.ptr, comp_len, b.ctypes.data if b.size else None, b.size,
                                      d_out.ptr, d_out.nbytes, C.byref(bad))
    ctx._ck("falx_bgzf_inflate", st, allow=(S.OK, S.E_INVALID))
    return st, (None if st == S.OK else bad.value)


def decompress(ctx: gpu.Context, comp: bytes) -> gpu.DeviceBuffer:
    """bgzf::decompress on the device: the inflated bytes in HBM (16 bytes
    of slack so the parsers' 16-byte alignment and reads are satisfied)."""
    blocks, total = scan(comp)
    d_comp = ctx.upload(np.frombuffer(comp, dtype=np.uint8))
    d_out = ctx.alloc(total + 16)
    st, bad = inflate(ctx, d_comp, len(comp), blocks, d_out)
    d_comp.free()
    if st != S.OK:
        d_out.free()
        raise gpu.FalxError("falx_bgzf_inflate", st, f"block {bad}")
    return d_out
