"""Navigation over a JSON nested tape, the reference's `Nested` API
(kernels/json.rs:762-993): `items()`, `scalars()` and `Node.{open, bytes,
items}` over the tape that falx_parse_nested builds in HBM.

The tape is one uint64 per structural byte outside strings: the byte
position in the low word and, for a bracket, the partner's tape index in the
high word (0xFFFFFFFF for an open never closed). Containers are skipped in
O(1) through the partner; scalars are the whitespace-trimmed gaps between
consecutive structural positions, zero-copy views of the input.
"""
from __future__ import annotations

from typing import Iterator, Optional

import numpy as np

NO_PARTNER = 0xFFFFFFFF
NEST_OK, UNMATCHED_CLOSE, UNCLOSED_OPEN = 0, 1, 2
_WS = b" \t\r\n"


def _trim(data: bytes, s: int, e: int):
    """trim_ws (json.rs:763-771)."""
    while s < e and data[s] in _WS:
        s += 1
    while e > s and data[e - 1] in _WS:
        e -= 1
    return s, e


class Nested:
    """A document and its tape. `error` is None or (kind, position) with kind
    UNMATCHED_CLOSE or UNCLOSED_OPEN (NestError, json.rs:608-615)."""

    def __init__(self, data: bytes, tape: np.ndarray, error: int = NEST_OK, error_pos: int = 0):
        self.data = bytes(data)
        self.tape = [int(x) for x in np.asarray(tape, dtype=np.uint64)]
        self.error = None if error == NEST_OK else (int(error), int(error_pos))

    def _pos(self, i: int) -> int:
        return self.tape[i] & 0xFFFFFFFF

    def _partner(self, i: int) -> int:
        return self.tape[i] >> 32

    def items(self) -> "Items":
        """The top-level items (json.rs:778-788)."""
        return Items(self, 0, len(self.tape), 0, len(self.data))

    def scalars(self) -> Iterator[bytes]:
        """Every scalar span in input order, flat (json.rs:790-806, 956-993)."""
        data, cursor = self.data, 0
        for entry in self.tape:
            pos = entry & 0xFFFFFFFF
            s, e = _trim(data, cursor, pos)
            cursor = pos + 1
            if s < e:
                yield data[s:e]
        if cursor <= len(data):
            s, e = _trim(data, cursor, len(data))
            if s < e:
                yield data[s:e]


class Node:
    """A bracketed container (tape index of its open) or a trimmed scalar span."""

    def __init__(self, nested: Nested, container: Optional[int] = None, span=None):
        self.nested = nested
        self.container = container
        self.span = span

    def open(self) -> Optional[int]:
        """The opening bracket byte; None for scalars (json.rs:832-841)."""
        if self.container is None:
            return None
        return self.nested.data[self.nested._pos(self.container)]

    def bytes(self) -> bytes:
        """The full span, brackets included (json.rs:843-861)."""
        nd = self.nested
        if self.container is None:
            return nd.data[self.span[0]:self.span[1]]
        start = nd._pos(self.container)
        close = nd._partner(self.container)
        end = len(nd.data) if close == NO_PARTNER else nd._pos(close) + 1
        return nd.data[start:end]

    def items(self) -> "Items":
        """The container's items in input order; keys and values of an object
        alternate (json.rs:863-896)."""
        nd = self.nested
        if self.container is None:
            return Items(nd, 0, 0, 0, 0)
        i = self.container
        close = nd._partner(i)
        if close == NO_PARTNER:
            end, limit = len(nd.tape), len(nd.data)
        else:
            end, limit = close, nd._pos(close)
        return Items(nd, i + 1, end, nd._pos(i) + 1, limit)


class Items:
    """Iterator over one nesting level (json.rs:899-954)."""

    def __init__(self, nested: Nested, nxt: int, end: int, cursor: int, limit: int):
        self.nested, self.next, self.end, self.cursor, self.limit = nested, nxt, end, cursor, limit

    def __iter__(self):
        return self

    def __next__(self) -> Node:
        nd = self.nested
        data = nd.data
        while True:
            if self.next >= self.end:
                s, e = _trim(data, self.cursor, self.limit)
                self.cursor = self.limit
                if s < e:
                    return Node(nd, span=(s, e))
                raise StopIteration
            entry = nd.tape[self.next]
            pos = entry & 0xFFFFFFFF
            if data[pos] in b"{[":
                s, e = _trim(data, self.cursor, pos)
                if s < e:
                    self.cursor = pos
                    return Node(nd, span=(s, e))
                node = Node(nd, container=self.next)
                close = entry >> 32
                if close == NO_PARTNER or close >= self.end:
                    self.next = self.end
                    self.cursor = self.limit
                else:
                    self.next = close + 1
                    self.cursor = nd._pos(close) + 1
                return node
            self.next += 1
            s, e = _trim(data, self.cursor, pos)
            self.cursor = pos + 1
            if s < e:
                return Node(nd, span=(s, e))
