"""Runtime format specs: the ctypes mirror of ``falx_spec`` and friends in
include/falx_gpu.h, plus the dialects the reference ships as spec files.

A reference spec file (specs/*.toml, parsed by src/spec.rs:44-276) becomes a
:class:`Spec`; ``Spec.to_c()`` produces the C struct the library takes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple

MAX_STRUCTURAL = 16
MAX_COLUMNS = 32

I64, F64, STR, BYTES = 0, 1, 2, 3
_TYPE_NAMES = {"i64": I64, "f64": F64, "string": STR, "str": STR, "bytes": BYTES}

# status codes (falx_status)
OK, E_INVALID, E_TOO_LARGE, E_CAPACITY, E_DEVICE, E_NO_DEVICE, E_SCHEMA, E_BUSY = range(8)
STATUS_NAMES = {
    OK: "OK", E_INVALID: "INVALID", E_TOO_LARGE: "TOO_LARGE", E_CAPACITY: "CAPACITY",
    E_DEVICE: "DEVICE", E_NO_DEVICE: "NO_DEVICE", E_SCHEMA: "SCHEMA", E_BUSY: "BUSY",
}


class CColumn(C.Structure):
    _fields_ = [("index", C.c_uint32), ("type", C.c_uint32)]


class CSpec(C.Structure):
    _fields_ = [
        ("structural", C.c_uint8 * MAX_STRUCTURAL),
        ("n_structural", C.c_uint32),
        ("quote", C.c_int32),
        ("escape", C.c_int32),
        ("comment", C.c_int32),
        ("n_columns", C.c_uint32),
        ("columns", CColumn * MAX_COLUMNS),
    ]


class CColumnBuf(C.Structure):
    _fields_ = [
        ("values", C.c_void_p),
        ("valid", C.c_void_p),
        ("data", C.c_void_p),
        ("data_cap", C.c_uint64),
        ("data_len", C.c_uint64),
    ]


class CColumns(C.Structure):
    _fields_ = [
        ("rows_cap", C.c_uint64),
        ("rows", C.c_uint64),
        ("cols", CColumnBuf * MAX_COLUMNS),
    ]


class CColumnStats(C.Structure):
    _fields_ = [
        ("values", C.c_uint64),
        ("checksum", C.c_uint64),
        ("bytes", C.c_uint64),
        ("fixed6_sum", C.c_int64),
        ("fixed6_cells", C.c_uint64),
        ("sum", C.c_double),
    ]


class CCsvStats(C.Structure):
    _fields_ = [("records", C.c_uint64), ("cols", CColumnStats * MAX_COLUMNS)]


class CFieldByteStats(C.Structure):
    _fields_ = [("fields", C.c_uint64), ("bytes", C.c_uint64)]


class CLogfmtStats(C.Structure):
    _fields_ = [
        ("pairs", C.c_uint64),
        ("key_bytes", C.c_uint64),
        ("value_bytes", C.c_uint64),
        ("checksum", C.c_uint64),
    ]


class CNdjsonIdScore(C.Structure):
    _fields_ = [
        ("records", C.c_uint64),
        ("invalid", C.c_uint64),
        ("id_sum", C.c_int64),
        ("score_fixed6_sum", C.c_int64),
        ("score_fixed6_cells", C.c_uint64),
        ("score_sum", C.c_double),
    ]


def _byte(x) -> int:
    if isinstance(x, int):
        return x
    b = x.encode() if isinstance(x, str) else bytes(x)
    if len(b) != 1:
        raise ValueError(f"expected a single byte, got {x!r}")
    return b[0]


# Batched execution (include/falx_gpu.h, falx_batch_*)
OP_INDEX, OP_COLUMNS, OP_COLUMNS_STATS, OP_CSV_STATS, OP_FIELD_BYTES, OP_LOGFMT_PAIRS, OP_NDJSON_ID_SCORE = range(1, 8)
N_SIZE_CLASS = 8


class CRequest(C.Structure):
    _fields_ = [
        ("op", C.c_uint32), ("chunk", C.c_uint32), ("spec", C.POINTER(CSpec)),
        ("d_data", C.c_void_p), ("len", C.c_uint64), ("columns", C.POINTER(CColumns)),
        ("d_index", C.c_void_p), ("index_cap", C.c_uint64),
    ]


class CResult(C.Structure):
    _fields_ = [
        ("status", C.c_int32), ("reserved", C.c_uint32), ("n_index", C.c_uint64),
        ("stats", CCsvStats), ("field_bytes", CFieldByteStats), ("logfmt", CLogfmtStats),
        ("ndjson", CNdjsonIdScore),
    ]


class CStreamOut(C.Structure):
    _fields_ = [("d_buf", C.c_void_p), ("n_seps", C.c_uint64), ("n_ends", C.c_uint64), ("data_end", C.c_uint64),
                ("need_bytes", C.c_uint64)]


class CBatchTuning(C.Structure):
    _fields_ = [
        ("tile_workgroups", C.c_uint32), ("unit_workgroups", C.c_uint32),
        ("chunk", C.c_uint32 * N_SIZE_CLASS), ("use_graph", C.c_uint32),
    ]


@dataclass(frozen=True)
class Column:
    index: int
    type: int
    name: str = ""


@dataclass(frozen=True)
class Spec:
    """A dialect (formats.rs:30-53) plus its typed projection."""

    name: str
    structural: bytes
    quote: Optional[int] = None
    escape: Optional[int] = None     # backslash-style escape byte, None = none/doubled
    comment: Optional[int] = None
    columns: Tuple[Column, ...] = field(default_factory=tuple)

    def with_columns(self, cols: Sequence[Tuple[int, int]], name: Optional[str] = None) -> "Spec":
        return Spec(name or self.name, self.structural, self.quote, self.escape, self.comment,
                    tuple(Column(i, t) for i, t in cols))

    def to_c(self) -> CSpec:
        s = CSpec()
        if len(self.structural) > MAX_STRUCTURAL:
            raise ValueError("too many structural bytes")
        for i, b in enumerate(self.structural):
            s.structural[i] = b
        s.n_structural = len(self.structural)
        s.quote = -1 if self.quote is None else self.quote
        s.escape = -1 if self.escape is None else self.escape
        s.comment = -1 if self.comment is None else self.comment
        if len(self.columns) > MAX_COLUMNS:
            raise ValueError("at most 32 columns")
        s.n_columns = len(self.columns)
        for i, c in enumerate(self.columns):
            s.columns[i].index = c.index
            s.columns[i].type = c.type
        return s

    @property
    def kind(self) -> str:
        if self.quote is None:
            return "plain"
        if self.escape is not None:
            return "escape"
        if self.comment is not None:
            return "regions"
        return "parity"


def _mk(name, structural, quote=None, escape=None, comment=None, columns=()):
    return Spec(name, bytes(structural), quote, escape, comment,
                tuple(Column(i, t) for i, t in columns))


# The reference's dialects (formats.rs:56-173) and spec files (specs/*.toml).
CSV = _mk("csv", b",\n", quote=ord('"'))
TSV = _mk("tsv", b"\t\n")
LOGFMT = _mk("logfmt", b" =\n", quote=ord('"'), escape=ord("\\"))
MULTI = _mk("multi", b",;|\t: /=&\n", quote=ord('"'))
CSV_HASH_DIALECT = _mk("csv_hash", b",\n", quote=ord('"'), comment=ord("#"))
NDJSON = _mk("ndjson", b"\n", quote=ord('"'), escape=ord("\\"))
LINES = _mk("lines", b"\n")
CSV_SEMICOLON = _mk("csv_semicolon", b";\n", quote=ord('"'))
PSV = _mk("psv", b"|\n")

# specs/csv-hash.toml: key:string@0, amount:i64@1
CSV_HASH = _mk("csv_hash", b",\n", quote=ord('"'), comment=ord("#"),
               columns=[(0, STR), (1, I64)])
# specs/csv-typed.toml: id:i64@0, title:string@1, value:f64@2, label:bytes@4
CSV_TYPED = _mk("csv_typed", b",\n", quote=ord('"'),
                columns=[(0, I64), (1, STR), (2, F64), (4, BYTES)])
# specs/csv-geo.toml and csv-geo-text.toml (worldcitiespop ordinals 1/5/6)
CSV_GEO_WCP = _mk("csv_geo", b",\n", quote=ord('"'), columns=[(5, F64), (6, F64)])
CSV_GEO_TEXT_WCP = _mk("csv_geo_text", b",\n", quote=ord('"'),
                       columns=[(1, STR), (5, F64), (6, F64)])

# Benchmark schema `id,City,Latitude,Longitude`: lat/lon lane, City+lat/lon
# lane, and the full projection (u64 id read as i64, as str::parse would).
GEO_LATLON = _mk("geo_latlon", b",\n", quote=ord('"'), columns=[(2, F64), (3, F64)])
GEO_CITY_LATLON = _mk("geo_city_latlon", b",\n", quote=ord('"'),
                      columns=[(1, STR), (2, F64), (3, F64)])
GEO_ALL = _mk("geo_all", b",\n", quote=ord('"'),
              columns=[(0, I64), (1, STR), (2, F64), (3, F64)])
# csv_hash benchmark schema: i64@0, f64@1, string@2
HASH_TYPED = _mk("hash_typed", b",\n", quote=ord('"'), comment=ord("#"),
                 columns=[(0, I64), (1, F64), (2, STR)])
# TSV variant of the served mix: id, name, value
TSV_TYPED = _mk("tsv_typed", b"\t\n", columns=[(0, I64), (1, STR), (2, F64)])

DIALECTS = {s.name: s for s in [CSV, TSV, LOGFMT, MULTI, CSV_HASH_DIALECT, NDJSON, LINES,
                                CSV_SEMICOLON, PSV]}


def parse_spec_toml(text: str) -> Spec:
    """Parse a reference spec file (docs/spec.md) into a Spec.

    Applies the validation rules of src/spec.rs:44-276 that concern this
    path: single-byte strings, escape mode, comment constraints, <= 32
    columns, column types.
    """
    try:
        import tomllib  # type: ignore
    except ModuleNotFoundError:  # python 3.10
        import tomli as tomllib  # type: ignore
    t = tomllib.loads(text)
    if "name" not in t or not isinstance(t["name"], str):
        raise ValueError("Missing required field: 'name' (must be a string)")
    if "structural" not in t or not isinstance(t["structural"], list):
        raise ValueError("Missing required field: 'structural' (must be an array of strings)")
    structural = bytes(_byte(s) for s in t["structural"])
    quote = _byte(t["quote"]) if "quote" in t else None
    esc_mode = t.get("escape", "none")
    esc_char = _byte(t.get("escape_char", "\\"))
    if esc_mode in ("none", "doubled"):
        escape = None
    elif esc_mode == "backslash":
        escape = esc_char
    else:
        raise ValueError(f"Invalid 'escape' value: '{esc_mode}'")
    comment = None
    if "comment" in t:
        comment = _byte(t["comment"])
        if comment in (10, 13):
            raise ValueError("'comment' must not be a line terminator")
        if comment == quote:
            raise ValueError("'comment' must differ from 'quote'")
    for pair in t.get("nesting", []):
        for b in pair.encode():
            if b not in structural:
                structural += bytes([b])
    cols: List[Column] = []
    for c in t.get("columns", []):
        ty = _TYPE_NAMES.get(c.get("type"))
        if ty is None:
            raise ValueError(f"invalid column type {c.get('type')!r}")
        cols.append(Column(int(c["index"]), ty, c.get("name", f"c{c['index']}")))
    if len(cols) > MAX_COLUMNS:
        raise ValueError("at most 32 columns")
    return Spec(t["name"], structural, quote, escape, comment, tuple(cols))
