"""Record view over a structural tape: the host mirror of the reference's
`Parsed` / `Records` / `Record` API (src/kernels/csv_geo.rs:414-690,
csv_hash.rs:431-557).

A tape is what ``falx_parse`` returns: separator offsets (u32) and one end
entry per terminated record, ``(separators before it << 32) | terminator
offset``. Iteration yields the trailing unterminated record too, trims a
record's trailing ``\\r`` lazily, and skips comment records when the dialect
has a comment byte.
"""
from __future__ import annotations

from typing import Iterator, List, Optional, Sequence, Tuple

import numpy as np

from .spec import Spec


def clean(spec: Spec, raw: bytes) -> bytes:
    """Dialect field cleaning (codegen.rs clean_fn 8112-8185)."""
    if spec.quote is None:
        return raw
    q = spec.quote
    if spec.escape is None:
        if len(raw) >= 2 and raw[0] == q and raw[-1] == q:
            inner, out, i = raw[1:-1], bytearray(), 0
            while i < len(inner):
                out.append(inner[i])
                i += 2 if inner[i] == q and i + 1 < len(inner) and inner[i + 1] == q else 1
            return bytes(out)
        return raw
    e = spec.escape
    body = raw[1:-1] if len(raw) >= 2 and raw[0] == q and raw[-1] == q else raw
    out, i = bytearray(), 0
    while i < len(body):
        if body[i] == e and i + 1 < len(body):
            out.append(body[i + 1])
            i += 2
        else:
            out.append(body[i])
            i += 1
    return bytes(out)


class Record:
    __slots__ = ("data", "start", "end", "seps", "spec")

    def __init__(self, data: bytes, start: int, end: int, seps: Sequence[int], spec: Spec):
        self.data, self.start, self.end, self.seps, self.spec = data, start, end, seps, spec

    def trimmed_end(self) -> int:
        if self.end > self.start and self.data[self.end - 1] == 0x0D:
            return self.end - 1
        return self.end

    def as_bytes(self) -> bytes:
        return self.data[self.start:self.trimmed_end()]

    def field_count(self) -> int:
        return len(self.seps) + 1

    def field_span(self, i: int) -> Optional[Tuple[int, int]]:
        if i > len(self.seps):
            return None
        frm = self.start if i == 0 else int(self.seps[i - 1]) + 1
        to = self.trimmed_end() if i == len(self.seps) else int(self.seps[i])
        return frm, to

    def field_raw(self, i: int) -> Optional[bytes]:
        sp = self.field_span(i)
        return None if sp is None else self.data[sp[0]:sp[1]]

    def field(self, i: int) -> Optional[bytes]:
        raw = self.field_raw(i)
        return None if raw is None else clean(self.spec, raw)

    def fields(self) -> List[bytes]:
        return [self.field(i) for i in range(self.field_count())]

    def fields_raw(self) -> List[bytes]:
        return [self.field_raw(i) for i in range(self.field_count())]


class Parsed:
    """`Parsed<'a>` (csv_geo.rs:414-466) over host copies of the tape."""

    def __init__(self, data: bytes, seps: np.ndarray, ends: np.ndarray, spec: Spec):
        self.data = bytes(data)
        self.seps = np.asarray(seps, dtype=np.uint32)
        self.ends = np.asarray(ends, dtype=np.uint64)
        self.spec = spec

    def terminated_record_count(self) -> int:
        return int(self.ends.size)

    def records(self) -> Iterator[Record]:
        return self.records_range(0, self.terminated_record_count())

    def records_range(self, start: int, end: int) -> Iterator[Record]:
        n = self.terminated_record_count()
        if start == 0:
            byte_pos, sep_pos = 0, 0
        else:
            prev = int(self.ends[start - 1])
            byte_pos, sep_pos = (prev & 0xFFFFFFFF) + 1, prev >> 32
        if start >= end and n > 0:
            data_end = byte_pos
        elif end == n:
            data_end = len(self.data)
        else:
            data_end = (int(self.ends[end - 1]) & 0xFFFFFFFF) + 1
        comment = self.spec.comment
        k = start
        while True:
            rec_start = byte_pos
            if k < end:
                entry = int(self.ends[k])
                k += 1
                rec_end, cum = entry & 0xFFFFFFFF, entry >> 32
                seps = self.seps[sep_pos:cum]
                sep_pos, byte_pos = cum, rec_end + 1
            else:
                if rec_start >= data_end:
                    return
                seps = self.seps[sep_pos:]
                sep_pos, byte_pos = self.seps.size, data_end
                rec_end = data_end
            if comment is not None and rec_end > rec_start and self.data[rec_start] == comment:
                continue
            yield Record(self.data, rec_start, rec_end, seps, self.spec)
