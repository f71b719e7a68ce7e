"""ctypes wrapper of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py. The product package never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Dict

import numpy as np

from falx_amd import spec as S

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER
        u8p, u32p, u64p = C.c_void_p, C.c_void_p, C.c_void_p
        specp = P(S.CSpec)
        L.ora_index_scalar.restype = C.c_uint64
        L.ora_index_scalar.argtypes = [specp, u8p, C.c_uint64, u32p]
        L.ora_index_structurals.restype = C.c_uint64
        L.ora_index_structurals.argtypes = [specp, u8p, C.c_uint64, C.c_int, u32p]
        L.ora_parse.restype = None
        L.ora_parse.argtypes = [specp, u8p, C.c_uint64, C.c_int, u32p, P(C.c_uint64), u64p,
                                P(C.c_uint64)]
        L.ora_parse_columns.restype = C.c_int
        L.ora_parse_columns.argtypes = [specp, u8p, C.c_uint64, C.c_int, P(S.CColumns)]
        L.ora_csv_stats.restype = None
        L.ora_csv_stats.argtypes = [specp, u8p, C.c_uint64, C.c_int, P(S.CCsvStats)]
        L.ora_field_bytes.restype = None
        L.ora_field_bytes.argtypes = [specp, u8p, C.c_uint64, C.c_int, P(S.CFieldByteStats)]
        L.ora_ndjson_lines.restype = C.c_uint64
        L.ora_ndjson_lines.argtypes = [u8p, C.c_uint64, C.c_int]
        L.ora_ndjson_id_score.restype = None
        L.ora_ndjson_id_score.argtypes = [u8p, C.c_uint64, C.c_int, P(S.CNdjsonIdScore)]
        L.ora_logfmt_pairs.restype = None
        L.ora_logfmt_pairs.argtypes = [u8p, C.c_uint64, C.c_int, P(S.CLogfmtStats)]
        L.ora_parse_f64.restype = C.c_int
        L.ora_parse_f64.argtypes = [u8p, C.c_uint64, P(C.c_double), P(C.c_int), P(C.c_int64)]
        L.ora_parse_i64.restype = C.c_int
        L.ora_parse_i64.argtypes = [u8p, C.c_uint64, P(C.c_int64)]
        L.ora_clean.restype = C.c_uint64
        L.ora_clean.argtypes = [specp, u8p, C.c_uint64, u8p]
        L.ora_parse_nested.restype = C.c_uint64
        L.ora_parse_nested.argtypes = [u8p, C.c_uint64, C.c_int, u64p, C.c_uint64, P(C.c_int32),
                                       P(C.c_uint64)]
        L.ora_inflate.restype = C.c_int64
        L.ora_inflate.argtypes = [u8p, C.c_uint64, u8p, C.c_uint64]
        L.ora_bgzf_inflate.restype = C.c_int64
        L.ora_bgzf_inflate.argtypes = [u8p, C.c_void_p, C.c_uint64, C.c_int, u8p]
        L.ora_isa.restype = C.c_char_p
        _lib = L
    return _lib


def _buf(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def _ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data) if a.size else C.c_void_p(0)


def isa() -> str:
    return lib().ora_isa().decode()


def index_scalar(spec: S.Spec, data) -> np.ndarray:
    d = _buf(data)
    out = np.empty(max(1, d.size), dtype=np.uint32)
    cs = spec.to_c()
    n = lib().ora_index_scalar(C.byref(cs), _ptr(d), d.size, _ptr(out))
    return out[:n].copy()


def index_structurals(spec: S.Spec, data, threads: int = 1) -> np.ndarray:
    d = _buf(data)
    out = np.empty(max(1, d.size), dtype=np.uint32)
    cs = spec.to_c()
    n = lib().ora_index_structurals(C.byref(cs), _ptr(d), d.size, threads, _ptr(out))
    return out[:n].copy()


def parse(spec: S.Spec, data, threads: int = 1):
    d = _buf(data)
    seps = np.empty(max(1, d.size), dtype=np.uint32)
    ends = np.empty(max(1, d.size), dtype=np.uint64)
    ns, ne = C.c_uint64(), C.c_uint64()
    cs = spec.to_c()
    lib().ora_parse(C.byref(cs), _ptr(d), d.size, threads, _ptr(seps), C.byref(ns), _ptr(ends),
                    C.byref(ne))
    return seps[: ns.value].copy(), ends[: ne.value].copy()


def alloc_columns(spec: S.Spec, rows_cap: int, data_cap: int):
    """Host numpy buffers + the CColumns struct pointing at them."""
    cols = S.CColumns()
    cols.rows_cap = rows_cap
    keep = []
    words = (rows_cap + 63) // 64
    for k, c in enumerate(spec.columns):
        valid = np.zeros(max(1, words), dtype=np.uint64)
        if c.type == S.I64:
            vals = np.zeros(max(1, rows_cap), dtype=np.int64)
            data = None
        elif c.type == S.F64:
            vals = np.zeros(max(1, rows_cap), dtype=np.float64)
            data = None
        elif c.type == S.BYTES:
            vals = np.zeros(max(1, 2 * rows_cap), dtype=np.uint32)
            data = None
        else:
            vals = np.zeros(rows_cap + 1, dtype=np.int32)
            data = np.zeros(max(1, data_cap), dtype=np.uint8)
        cols.cols[k].values = vals.ctypes.data
        cols.cols[k].valid = valid.ctypes.data
        if data is not None:
            cols.cols[k].data = data.ctypes.data
            cols.cols[k].data_cap = data_cap
        keep.append((vals, valid, data))
    return cols, keep


def columns_result(spec: S.Spec, cols: S.CColumns, keep) -> Dict:
    rows = cols.rows
    words = (rows + 63) // 64
    res = {"rows": rows, "cols": []}
    for k, c in enumerate(spec.columns):
        vals, valid, data = keep[k]
        entry = {"valid": valid[:words].copy()}
        if c.type in (S.I64, S.F64):
            entry["values"] = vals[:rows].copy()
        elif c.type == S.BYTES:
            entry["values"] = vals[: 2 * rows].reshape(-1, 2).copy() if rows else vals[:0].reshape(0, 2)
        else:
            entry["offsets"] = vals[: rows + 1].copy()
            entry["data"] = data[: cols.cols[k].data_len].copy()
        res["cols"].append(entry)
    return res


def parse_columns(spec: S.Spec, data, threads: int = 1) -> Dict:
    d = _buf(data)
    cols, keep = alloc_columns(spec, d.size + 1, d.size + 16)
    cs = spec.to_c()
    st = lib().ora_parse_columns(C.byref(cs), _ptr(d), d.size, threads, C.byref(cols))
    if st != 0:
        raise RuntimeError(f"oracle parse_columns status {st}")
    return columns_result(spec, cols, keep)


def stats_dict(spec: S.Spec, st: S.CCsvStats) -> Dict:
    out = {"records": st.records, "cols": []}
    for k in range(len(spec.columns)):
        c = st.cols[k]
        out["cols"].append({
            "values": c.values, "checksum": c.checksum, "bytes": c.bytes,
            "fixed6_sum": c.fixed6_sum, "fixed6_cells": c.fixed6_cells, "sum": c.sum,
        })
    return out


def csv_stats(spec: S.Spec, data, threads: int = 1) -> Dict:
    d = _buf(data)
    st = S.CCsvStats()
    cs = spec.to_c()
    lib().ora_csv_stats(C.byref(cs), _ptr(d), d.size, threads, C.byref(st))
    return stats_dict(spec, st)


def field_bytes(spec: S.Spec, data, threads: int = 1) -> Dict:
    d = _buf(data)
    st = S.CFieldByteStats()
    cs = spec.to_c()
    lib().ora_field_bytes(C.byref(cs), _ptr(d), d.size, threads, C.byref(st))
    return {"fields": st.fields, "bytes": st.bytes}


def ndjson_lines(data, threads: int = 1) -> int:
    d = _buf(data)
    return lib().ora_ndjson_lines(_ptr(d), d.size, threads)


def ndjson_id_score(data, threads: int = 1) -> Dict:
    d = _buf(data)
    st = S.CNdjsonIdScore()
    lib().ora_ndjson_id_score(_ptr(d), d.size, threads, C.byref(st))
    return {"records": st.records, "invalid": st.invalid, "id_sum": st.id_sum,
            "score_fixed6_sum": st.score_fixed6_sum,
            "score_fixed6_cells": st.score_fixed6_cells, "score_sum": st.score_sum}


def logfmt_pairs(data, threads: int = 1) -> Dict:
    d = _buf(data)
    st = S.CLogfmtStats()
    lib().ora_logfmt_pairs(_ptr(d), d.size, threads, C.byref(st))
    return {"pairs": st.pairs, "key_bytes": st.key_bytes, "value_bytes": st.value_bytes,
            "checksum": st.checksum}


def parse_f64(cell: bytes):
    """(ok, value, fixed6, mantissa)"""
    d = _buf(cell)
    v, fx, m = C.c_double(), C.c_int(), C.c_int64()
    ok = lib().ora_parse_f64(_ptr(d), d.size, C.byref(v), C.byref(fx), C.byref(m))
    return bool(ok), v.value, bool(fx.value), m.value


def parse_i64(cell: bytes):
    d = _buf(cell)
    v = C.c_int64()
    ok = lib().ora_parse_i64(_ptr(d), d.size, C.byref(v))
    return bool(ok), v.value


def clean(spec: S.Spec, cell: bytes) -> bytes:
    d = _buf(cell)
    out = np.zeros(max(1, d.size), dtype=np.uint8)
    cs = spec.to_c()
    n = lib().ora_clean(C.byref(cs), _ptr(d), d.size, _ptr(out))
    return out[:n].tobytes()


def parse_nested(data, threads: int = 1):
    """(tape uint64 array, error, error_pos): json::parse_nested[_par]."""
    d = _buf(data)
    # At most one entry per input byte; np.empty leaves untouched pages
    # uncommitted.
    tape = np.empty(max(1, d.size), dtype=np.uint64)
    err, pos = C.c_int32(), C.c_uint64()
    n = lib().ora_parse_nested(_ptr(d), d.size, threads, _ptr(tape), tape.size, C.byref(err), C.byref(pos))
    return tape[:n].copy(), err.value, pos.value


def inflate(payload: bytes, cap: int):
    """Raw DEFLATE -> bytes, or None when malformed."""
    d = _buf(payload)
    out = np.zeros(max(1, cap), dtype=np.uint8)
    n = lib().ora_inflate(_ptr(d), d.size, _ptr(out), cap)
    return None if n < 0 else out[:n].tobytes()


def bgzf_inflate(comp, blocks: np.ndarray, total: int, threads: int = 1):
    """Inflated bytes of a block table (falx_bgzf_scan's), or None."""
    d = _buf(comp)
    out = np.zeros(max(1, total), dtype=np.uint8)
    b = np.ascontiguousarray(blocks)
    n = lib().ora_bgzf_inflate(_ptr(d), C.c_void_p(b.ctypes.data) if b.size else None, b.size, threads, _ptr(out))
    return None if n < 0 else out[:total]
