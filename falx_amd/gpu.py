"""Host-side mirror of the reference's generated-parser API over the C ABI of
libfalx_gpu.so (include/falx_gpu.h).

The reference's functions take a `&[u8]` and return owned vectors
(src/kernels/<fmt>.rs); here a :class:`Context` owns a device and a stream,
inputs live in HBM as :class:`DeviceBuffer` objects, and each method returns
numpy arrays copied back from the device (or leaves results on the device
when ``download=False``). The native library is required: importing this
module without a built library raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Dict, Optional

import numpy as np

from . import build as _build
from . import spec as S

_lib = None

# (name, restype, argtypes) for every entry point of include/falx_gpu.h
_VP, _U64, _U64P = C.c_void_p, C.c_uint64, C.POINTER(C.c_uint64)
_SIGS = [
    ("falx_ctx_create", C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    ("falx_ctx_destroy", None, [_VP]),
    ("falx_ctx_stream", _VP, [_VP]),
    ("falx_ctx_last_kernel_ms", C.c_double, [_VP]),
    ("falx_ctx_last_error", C.c_int, [_VP, C.POINTER(C.c_int32)]),
    ("falx_device_alloc", C.c_int, [_VP, _U64, C.POINTER(C.c_void_p)]),
    ("falx_device_free", C.c_int, [_VP, _VP]),
    ("falx_host_alloc_pinned", C.c_int, [_VP, _U64, C.POINTER(C.c_void_p)]),
    ("falx_host_free_pinned", C.c_int, [_VP, _VP]),
    ("falx_upload", C.c_int, [_VP, _VP, _VP, _U64]),
    ("falx_download", C.c_int, [_VP, _VP, _VP, _U64]),
    ("falx_spec_validate", C.c_int, [C.POINTER(S.CSpec)]),
    ("falx_index_structurals", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, _VP, _U64, _U64P]),
    ("falx_index_structurals_static", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, _VP, _U64, _U64P]),
    ("falx_parse", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, _VP, _U64, _VP, _U64, _U64P, _U64P]),
    ("falx_parse_columns", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, C.POINTER(S.CColumns)]),
    ("falx_parse_columns_with_stats", C.c_int,
     [_VP, C.POINTER(S.CSpec), _VP, _U64, C.POINTER(S.CColumns), C.POINTER(S.CCsvStats)]),
    ("falx_parse_csv_stats", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, C.POINTER(S.CCsvStats)]),
    ("falx_parse_field_bytes", C.c_int, [_VP, C.POINTER(S.CSpec), _VP, _U64, C.POINTER(S.CFieldByteStats)]),
    ("falx_parse_ndjson_lines", C.c_int, [_VP, _VP, _U64, _U64P]),
    ("falx_parse_ndjson_id_score", C.c_int, [_VP, _VP, _U64, C.POINTER(S.CNdjsonIdScore)]),
    ("falx_parse_logfmt_pairs", C.c_int, [_VP, _VP, _U64, C.POINTER(S.CLogfmtStats)]),
    ("falx_parse_nested", C.c_int, [_VP, _VP, _U64, _VP, _U64, _U64P, C.POINTER(C.c_int32), _U64P]),
    ("falx_stream_create", C.c_int, [_VP, C.POINTER(S.CSpec), _U64, C.POINTER(C.c_void_p)]),
    ("falx_stream_destroy", None, [_VP]),
    ("falx_stream_reset", C.c_int, [_VP]),
    ("falx_stream_feed", C.c_int, [_VP, _VP, _U64, _VP, _U64, _VP, _U64, C.POINTER(S.CStreamOut)]),
    ("falx_stream_finish", C.c_int, [_VP, _VP, _U64, _VP, _U64, C.POINTER(S.CStreamOut)]),
    ("falx_bgzf_scan", C.c_int, [_VP, _U64, _VP, _U64, _U64P, _U64P]),
    ("falx_bgzf_inflate", C.c_int, [_VP, _VP, _U64, _VP, _U64, _VP, _U64, _U64P]),
    ("falx_export_arrow", C.c_int, [C.POINTER(S.CSpec), C.POINTER(S.CColumns), C.c_uint32, C.c_int32, C.c_int64,
                                    _VP, _VP]),
    ("falx_batch_create", C.c_int, [_VP, _U64, C.POINTER(C.c_void_p)]),
    ("falx_batch_destroy", None, [_VP]),
    ("falx_request_scratch_bytes", _U64, [C.POINTER(S.CRequest)]),
    ("falx_batch_get_tuning", C.c_int, [_VP, C.POINTER(S.CBatchTuning)]),
    ("falx_batch_set_tuning", C.c_int, [_VP, C.POINTER(S.CBatchTuning)]),
    ("falx_batch_submit", C.c_int, [_VP, C.POINTER(S.CRequest), C.c_uint32]),
    ("falx_batch_query", C.c_int, [_VP]),
    ("falx_batch_wait", C.c_int, [_VP, C.POINTER(S.CResult)]),
    ("falx_batch_last_ms", C.c_double, [_VP]),
]
EXPORTED = [name for name, _, _ in _SIGS]


class FalxError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str = ""):
        super().__init__(f"{fn}: {S.STATUS_NAMES.get(status, status)}{detail}")
        self.status = status


def lib_path() -> str:
    return _build.gpu_lib_path()


def load():
    """Load libfalx_gpu.so (built in-tree by __graft_entry__.build())."""
    global _lib
    if _lib is None:
        path = lib_path()
        if not os.path.exists(path):
            raise ImportError(f"libfalx_gpu.so not built ({path}); run __graft_entry__.build()")
        L = C.CDLL(path)
        for name, res, args in _SIGS:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(fn: str, st: int, allow=(S.OK,), ctx=None):
    if st not in allow:
        detail = ""
        if st == S.E_DEVICE and ctx is not None:
            hip = C.c_int32()
            if load().falx_ctx_last_error(ctx, C.byref(hip)) == S.OK:
                detail = f" (hip error {hip.value})"
        raise FalxError(fn, st, detail)
    return st


def _stats_dict(spec: S.Spec, st: S.CCsvStats) -> Dict:
    out = {"records": st.records, "cols": []}
    for k in range(len(spec.columns)):
        c = st.cols[k]
        out["cols"].append({"values": c.values, "checksum": c.checksum, "bytes": c.bytes,
                            "fixed6_sum": c.fixed6_sum, "fixed6_cells": c.fixed6_cells, "sum": c.sum})
    return out


class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check("falx_device_alloc", load().falx_device_alloc(ctx.h, max(1, self.nbytes), C.byref(p)))
        self.ptr = p.value

    def free(self):
        if self.ptr:
            load().falx_device_free(self.ctx.h, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass

    def download(self, dtype=np.uint8, count: Optional[int] = None, offset: int = 0) -> np.ndarray:
        dt = np.dtype(dtype)
        n = (self.nbytes - offset) // dt.itemsize if count is None else int(count)
        out = np.empty(n, dtype=dt)
        if n:
            _check("falx_download", load().falx_download(self.ctx.h, out.ctypes.data,
                                                         self.ptr + offset, n * dt.itemsize))
        return out


class Context:
    def __init__(self, device: int = 0):
        L = load()
        h = C.c_void_p()
        _check("falx_ctx_create", L.falx_ctx_create(device, C.byref(h)))
        self.h = h.value

    def close(self):
        if self.h:
            load().falx_ctx_destroy(self.h)
            self.h = None

    def _ck(self, fn: str, st: int, allow=(S.OK,)):
        return _check(fn, st, allow, self.h)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def last_kernel_ms(self) -> float:
        return load().falx_ctx_last_kernel_ms(self.h)

    def alloc(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def upload(self, data) -> DeviceBuffer:
        arr = np.ascontiguousarray(data if isinstance(data, np.ndarray) else
                                   np.frombuffer(bytes(data), dtype=np.uint8))
        buf = DeviceBuffer(self, arr.nbytes)
        if arr.nbytes:
            self._ck("falx_upload", load().falx_upload(self.h, buf.ptr, arr.ctypes.data, arr.nbytes))
        return buf

    # -- structural index -------------------------------------------------
    def index_structurals(self, spec: S.Spec, data: DeviceBuffer, length: Optional[int] = None,
                          out: Optional[DeviceBuffer] = None, download: bool = True):
        n = data.nbytes if length is None else length
        cs = spec.to_c()
        if out is None:
            out = self.alloc(4 * max(1, n))
        cnt = C.c_uint64()
        self._ck("falx_index_structurals",
               load().falx_index_structurals(self.h, C.byref(cs), data.ptr, n, out.ptr, out.nbytes // 4,
                                             C.byref(cnt)))
        return out.download(np.uint32, cnt.value) if download else (out, cnt.value)

    def parse(self, spec: S.Spec, data: DeviceBuffer, length: Optional[int] = None):
        n = data.nbytes if length is None else length
        cs = spec.to_c()
        seps = self.alloc(4 * max(1, n))
        ends = self.alloc(8 * max(1, n))
        ns, ne = C.c_uint64(), C.c_uint64()
        self._ck("falx_parse", load().falx_parse(self.h, C.byref(cs), data.ptr, n, seps.ptr, max(1, n), ends.ptr,
                                               max(1, n), C.byref(ns), C.byref(ne)))
        return seps.download(np.uint32, ns.value), ends.download(np.uint64, ne.value)

    # -- columns ----------------------------------------------------------
    def alloc_columns(self, spec: S.Spec, rows_cap: int, data_cap: int):
        cols = S.CColumns()
        cols.rows_cap = rows_cap
        keep = []
        words = (rows_cap + 63) // 64
        for k, c in enumerate(spec.columns):
            if c.type in (S.I64, S.F64, S.BYTES):
                vals = self.alloc(8 * max(1, rows_cap))
                data = None
            else:
                vals = self.alloc(4 * (rows_cap + 1))
                data = self.alloc(max(1, data_cap))
            valid = self.alloc(8 * max(1, words))
            cols.cols[k].values = vals.ptr
            cols.cols[k].valid = valid.ptr
            if data is not None:
                cols.cols[k].data = data.ptr
                cols.cols[k].data_cap = data_cap
            keep.append((vals, valid, data))
        return cols, keep

    def parse_columns_into(self, spec: S.Spec, data: DeviceBuffer, length: int, cols: S.CColumns) -> int:
        cs = spec.to_c()
        return self._ck("falx_parse_columns",
                      load().falx_parse_columns(self.h, C.byref(cs), data.ptr, length, C.byref(cols)),
                      allow=(S.OK, S.E_CAPACITY))

    def parse_columns_with_stats_into(self, spec: S.Spec, data: DeviceBuffer, length: int,
                                      cols: S.CColumns) -> Dict:
        """One pass: columns into `cols` and the fused statistics returned."""
        cs = spec.to_c()
        st = S.CCsvStats()
        self._ck("falx_parse_columns_with_stats",
               load().falx_parse_columns_with_stats(self.h, C.byref(cs), data.ptr, length, C.byref(cols),
                                                    C.byref(st)), allow=(S.OK, S.E_CAPACITY))
        return _stats_dict(spec, st)

    def parse_columns(self, spec: S.Spec, data: DeviceBuffer, length: Optional[int] = None,
                      rows_cap: Optional[int] = None, data_cap: Optional[int] = None) -> Dict:
        n = data.nbytes if length is None else length
        rows_cap = n + 1 if rows_cap is None else rows_cap
        data_cap = max(16, n) if data_cap is None else data_cap
        cols, keep = self.alloc_columns(spec, rows_cap, data_cap)
        st = self.parse_columns_into(spec, data, n, cols)
        if st == S.E_CAPACITY:
            raise FalxError("falx_parse_columns", st)
        rows = cols.rows
        words = (rows + 63) // 64
        res = {"rows": rows, "cols": []}
        for k, c in enumerate(spec.columns):
            vals, valid, dbuf = keep[k]
            entry = {"valid": valid.download(np.uint64, words)}
            if c.type == S.I64:
                entry["values"] = vals.download(np.int64, rows)
            elif c.type == S.F64:
                entry["values"] = vals.download(np.float64, rows)
            elif c.type == S.BYTES:
                entry["values"] = vals.download(np.uint32, 2 * rows).reshape(-1, 2)
            else:
                entry["offsets"] = vals.download(np.int32, rows + 1)
                entry["data"] = dbuf.download(np.uint8, cols.cols[k].data_len)
            res["cols"].append(entry)
        return res

    # -- fused sinks ------------------------------------------------------
    def csv_stats(self, spec: S.Spec, data: DeviceBuffer, length: Optional[int] = None) -> Dict:
        n = data.nbytes if length is None else length
        cs = spec.to_c()
        st = S.CCsvStats()
        self._ck("falx_parse_csv_stats", load().falx_parse_csv_stats(self.h, C.byref(cs), data.ptr, n, C.byref(st)))
        return _stats_dict(spec, st)

    def field_bytes(self, spec: S.Spec, data: DeviceBuffer, length: Optional[int] = None) -> Dict:
        n = data.nbytes if length is None else length
        cs = spec.to_c()
        st = S.CFieldByteStats()
        self._ck("falx_parse_field_bytes",
               load().falx_parse_field_bytes(self.h, C.byref(cs), data.ptr, n, C.byref(st)))
        return {"fields": st.fields, "bytes": st.bytes}

    def ndjson_lines(self, data: DeviceBuffer, length: Optional[int] = None) -> int:
        n = data.nbytes if length is None else length
        r = C.c_uint64()
        self._ck("falx_parse_ndjson_lines", load().falx_parse_ndjson_lines(self.h, data.ptr, n, C.byref(r)))
        return r.value

    def ndjson_id_score(self, data: DeviceBuffer, length: Optional[int] = None) -> Dict:
        n = data.nbytes if length is None else length
        st = S.CNdjsonIdScore()
        self._ck("falx_parse_ndjson_id_score",
               load().falx_parse_ndjson_id_score(self.h, data.ptr, n, C.byref(st)), allow=(S.OK, S.E_SCHEMA))
        return {"records": st.records, "invalid": st.invalid, "id_sum": st.id_sum,
                "score_fixed6_sum": st.score_fixed6_sum, "score_fixed6_cells": st.score_fixed6_cells,
                "score_sum": st.score_sum}

    def logfmt_pairs(self, data: DeviceBuffer, length: Optional[int] = None) -> Dict:
        n = data.nbytes if length is None else length
        st = S.CLogfmtStats()
        self._ck("falx_parse_logfmt_pairs", load().falx_parse_logfmt_pairs(self.h, data.ptr, n, C.byref(st)))
        return {"pairs": st.pairs, "key_bytes": st.key_bytes, "value_bytes": st.value_bytes,
                "checksum": st.checksum}

    # -- JSON nested tape -----------------------------------------------------
    def parse_nested_into(self, data: DeviceBuffer, length: int, tape: DeviceBuffer):
        """falx_parse_nested into `tape` (8-byte entries): (status, n_tape,
        error, error_pos); status is OK or CAPACITY (n_tape = entries needed)."""
        n, err, pos = C.c_uint64(), C.c_int32(), C.c_uint64()
        st = load().falx_parse_nested(self.h, data.ptr, length, tape.ptr, tape.nbytes // 8, C.byref(n),
                                      C.byref(err), C.byref(pos))
        self._ck("falx_parse_nested", st, allow=(S.OK, S.E_CAPACITY))
        return st, n.value, err.value, pos.value

    def parse_nested(self, data: DeviceBuffer, length: Optional[int] = None, download: bool = True):
        """json::parse_nested on the device: (tape, n_tape, error, error_pos),
        the tape downloaded as uint64 (or left as a DeviceBuffer). The tape
        is sized for one entry per 4 input bytes and grown once if needed."""
        n = data.nbytes if length is None else length
        tape = self.alloc(8 * max(1, n // 4 + 64))
        st, cnt, err, pos = self.parse_nested_into(data, n, tape)
        if st == S.E_CAPACITY:
            tape.free()
            tape = self.alloc(8 * max(1, cnt))
            st, cnt, err, pos = self.parse_nested_into(data, n, tape)
            self._ck("falx_parse_nested", st)
        if download:
            out = tape.download(np.uint64, cnt)
            tape.free()
            return out, cnt, err, pos
        return tape, cnt, err, pos


class Stream:
    """A device StreamParser (falx_stream_*): feed() and finish() return the
    records each call completes as a records.Parsed over a host copy of the
    stream buffer (its `records()` yields them, comment records skipped)."""

    def __init__(self, ctx: Context, spec: S.Spec, buf_cap: int):
        self.ctx, self.spec = ctx, spec
        self._cs = spec.to_c()
        h = C.c_void_p()
        ctx._ck("falx_stream_create", load().falx_stream_create(ctx.h, C.byref(self._cs), buf_cap, C.byref(h)))
        self.h = h.value
        self._seps = ctx.alloc(4 * 1024)
        self._ends = ctx.alloc(8 * 1024)

    def close(self):
        if self.h:
            load().falx_stream_destroy(self.h)
            self.h = None

    def _call(self, fn: str, *args):
        from .records import Parsed
        while True:
            out = S.CStreamOut()
            st = getattr(load(), fn)(self.h, *args, self._seps.ptr, self._seps.nbytes // 4, self._ends.ptr,
                                     self._ends.nbytes // 8, C.byref(out))
            if st == S.E_CAPACITY and not out.need_bytes:
                if out.n_seps > self._seps.nbytes // 4:
                    self._seps.free()
                    self._seps = self.ctx.alloc(4 * (2 * out.n_seps + 1024))
                if out.n_ends > self._ends.nbytes // 8:
                    self._ends.free()
                    self._ends = self.ctx.alloc(8 * (2 * out.n_ends + 1024))
                continue
            self.ctx._ck(fn, st)
            n = out.data_end
            data = np.empty(n, dtype=np.uint8)
            if n:
                _check("falx_download", load().falx_download(self.ctx.h, data.ctypes.data, out.d_buf, n))
            seps = self._seps.download(np.uint32, out.n_seps)
            ends = self._ends.download(np.uint64, out.n_ends)
            return Parsed(data.tobytes(), seps, ends, self.spec)

    def feed(self, chunk: "DeviceBuffer", length: Optional[int] = None):
        n = chunk.nbytes if length is None else length
        return self._call("falx_stream_feed", chunk.ptr, n)

    def finish(self):
        return self._call("falx_stream_finish")


def _result_dict(op: int, spec: Optional[S.Spec], r: S.CResult) -> Dict:
    out = {"status": r.status}
    if op == S.OP_INDEX:
        out["n_index"] = r.n_index
    elif op in (S.OP_COLUMNS_STATS, S.OP_CSV_STATS):
        out["stats"] = _stats_dict(spec, r.stats)
    elif op == S.OP_FIELD_BYTES:
        out["field_bytes"] = {"fields": r.field_bytes.fields, "bytes": r.field_bytes.bytes}
    elif op == S.OP_LOGFMT_PAIRS:
        L = r.logfmt
        out["logfmt"] = {"pairs": L.pairs, "key_bytes": L.key_bytes, "value_bytes": L.value_bytes,
                         "checksum": L.checksum}
    elif op == S.OP_NDJSON_ID_SCORE:
        N = r.ndjson
        out["ndjson"] = {"records": N.records, "invalid": N.invalid, "id_sum": N.id_sum,
                         "score_fixed6_sum": N.score_fixed6_sum, "score_fixed6_cells": N.score_fixed6_cells,
                         "score_sum": N.score_sum}
    return out


class Request:
    """One document of a batch: an op of falx_op, its spec (CSV-family ops),
    its device input and, for COLUMNS ops, device column buffers
    (:meth:`Context.alloc_columns`) or, for INDEX, an offsets buffer."""

    def __init__(self, op: int, data: DeviceBuffer, length: int, spec: Optional[S.Spec] = None,
                 columns: Optional[S.CColumns] = None, index: Optional[DeviceBuffer] = None,
                 index_cap: int = 0, chunk: int = 0):
        self.op, self.data, self.length, self.spec = op, data, length, spec
        self.columns, self.index, self.index_cap, self.chunk = columns, index, index_cap, chunk
        self._cspec = spec.to_c() if spec is not None else None

    def to_c(self) -> S.CRequest:
        r = S.CRequest()
        r.op = self.op
        r.chunk = self.chunk
        r.spec = C.pointer(self._cspec) if self._cspec is not None else None
        r.d_data = self.data.ptr if self.data is not None else None
        r.len = self.length
        r.columns = C.pointer(self.columns) if self.columns is not None else None
        r.d_index = self.index.ptr if self.index is not None else None
        r.index_cap = self.index_cap
        return r

    def scratch_bytes(self) -> int:
        r = self.to_c()
        return load().falx_request_scratch_bytes(C.byref(r))


class Batch:
    """Batched execution over one context (include/falx_gpu.h falx_batch_*):
    submit() enqueues many documents at once and returns; wait() returns one
    result dict per document, each equal to the plain entry point's."""

    def __init__(self, ctx: Context, scratch_bytes: int):
        self.ctx = ctx
        h = C.c_void_p()
        ctx._ck("falx_batch_create", load().falx_batch_create(ctx.h, scratch_bytes, C.byref(h)))
        self.h = h
        self._pending = None

    def close(self):
        if self.h:
            load().falx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def tuning(self) -> S.CBatchTuning:
        t = S.CBatchTuning()
        self.ctx._ck("falx_batch_get_tuning", load().falx_batch_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, t: S.CBatchTuning):
        self.ctx._ck("falx_batch_set_tuning", load().falx_batch_set_tuning(self.h, C.byref(t)))

    def submit(self, reqs):
        arr = (S.CRequest * max(1, len(reqs)))(*[r.to_c() for r in reqs])
        self.ctx._ck("falx_batch_submit", load().falx_batch_submit(self.h, arr, len(reqs)))
        self._pending = (list(reqs), arr)

    def done(self) -> bool:
        st = load().falx_batch_query(self.h)
        if st == S.E_BUSY:
            return False
        self.ctx._ck("falx_batch_query", st)
        return True

    def wait(self):
        reqs, arr = self._pending
        res = (S.CResult * max(1, len(reqs)))()
        self.ctx._ck("falx_batch_wait", load().falx_batch_wait(self.h, res))
        self._pending = None
        return [_result_dict(r.op, r.spec, res[i]) for i, r in enumerate(reqs)]

    @property
    def last_ms(self) -> float:
        return load().falx_batch_last_ms(self.h)
