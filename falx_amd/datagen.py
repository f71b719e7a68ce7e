"""Seeded synthetic datasets for the benchmark lanes (the counterpart of the
reference's examples/make_datasets.rs). Generation runs in native code
(falx_amd/datagen/datagen.cpp); each generator also returns closed-form
truths (exact mantissa sums, bit checksums, counts) about what it wrote.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import build as _build

# Seeds of the reference's dataset pack (make_datasets.rs:221-230).
SEED_CSV_GEO = 0xDEADBEEFCAFEBABE
SEED_CSV_HASH = 0xC0FFEE15600DF00D
SEED_TSV = 0xA0761D6478BD642F
SEED_LOGFMT = 0xE7037ED1A0B428DB
SEED_NDJSON = 0x853C49E6748FEA9B
GENERATOR_VERSION = 2


class GeoTruth(C.Structure):
    _fields_ = [(n, t) for n, t in [
        ("rows", C.c_uint64), ("data_rows", C.c_uint64), ("lat_fixed6_sum", C.c_int64),
        ("lon_fixed6_sum", C.c_int64), ("lat_checksum", C.c_uint64), ("lon_checksum", C.c_uint64),
        ("city_bytes", C.c_uint64), ("city_checksum", C.c_uint64), ("quoted_cities", C.c_uint64),
        ("id_sum", C.c_int64)]]


class HashTruth(C.Structure):
    _fields_ = [("lines", C.c_uint64), ("comment_lines", C.c_uint64), ("records", C.c_uint64),
                ("hazard_fields", C.c_uint64), ("i64_sum", C.c_int64)]


class NdjsonTruth(C.Structure):
    _fields_ = [("records", C.c_uint64), ("id_sum", C.c_int64), ("score_fixed6_sum", C.c_int64),
                ("escaped_names", C.c_uint64)]


class TsvTruth(C.Structure):
    _fields_ = [("rows", C.c_uint64), ("id_sum", C.c_int64)]


class LogfmtTruth(C.Structure):
    _fields_ = [("lines", C.c_uint64), ("pairs", C.c_uint64), ("key_bytes", C.c_uint64),
                ("value_bytes", C.c_uint64), ("checksum", C.c_uint64)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _build.datagen_lib_path()
        if not os.path.exists(path):
            _build.build_datagen()
        L = C.CDLL(path)
        for name, truth in [("datagen_geo_csv", GeoTruth), ("datagen_hash_csv", HashTruth),
                            ("datagen_ndjson", NdjsonTruth), ("datagen_tsv", TsvTruth),
                            ("datagen_logfmt", LogfmtTruth)]:
            f = getattr(L, name)
            f.restype = C.c_uint64
            f.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_uint64, C.POINTER(truth)]
        _lib = L
    return _lib


def _gen(fn: str, truth_t, target: int, seed: int, out: np.ndarray | None = None):
    cap = target + 4096
    buf = out if out is not None else np.empty(cap, dtype=np.uint8)
    if buf.size < cap:
        raise ValueError("output buffer too small")
    t = truth_t()
    n = getattr(lib(), fn)(buf.ctypes.data, target, cap, seed & (2**64 - 1), C.byref(t))
    if n == 0 and target > 0:
        raise RuntimeError(f"{fn}: buffer overflow")
    truth = {name: getattr(t, name) for name, _ in truth_t._fields_}
    return buf[:n], truth


def geo_csv(target: int, seed: int = SEED_CSV_GEO, out=None):
    """`id,City,Latitude,Longitude` CSV of at least `target` bytes."""
    return _gen("datagen_geo_csv", GeoTruth, target, seed, out)


def hash_csv(target: int, seed: int = SEED_CSV_HASH, out=None):
    """csv_hash-dialect CSV (`amount:i64,value:f64,label:string` + comments)."""
    return _gen("datagen_hash_csv", HashTruth, target, seed, out)


def ndjson(target: int, seed: int = SEED_NDJSON, out=None):
    """NDJSON records `{id, name, nested: {score, tags}}`."""
    return _gen("datagen_ndjson", NdjsonTruth, target, seed, out)


def tsv(target: int, seed: int = SEED_TSV, out=None):
    return _gen("datagen_tsv", TsvTruth, target, seed, out)


def logfmt(target: int, seed: int = SEED_LOGFMT, out=None):
    return _gen("datagen_logfmt", LogfmtTruth, target, seed, out)
