"""Arrow export of column results (the reference's examples/arrow_interop.rs
wraps its columns' buffers into Arrow arrays; here the buffers are in HBM
and leave through the Arrow C Device Data Interface, falx_export_arrow in
include/falx_gpu.h).

``export_column`` returns the C structs of one column, zero-copy over its
buffers. ``to_pyarrow`` imports a column whose buffers are host memory;
``download_to_pyarrow`` copies a device column's buffers to the host and
builds the same array there (pyarrow has no ROCm memory support).
"""
from __future__ import annotations

import ctypes as C
from typing import Tuple

import numpy as np

from . import gpu
from . import spec as S

ARROW_DEVICE_CPU = 1
ARROW_DEVICE_ROCM = 10


class ArrowSchema(C.Structure):
    pass


ArrowSchema._fields_ = [("format", C.c_char_p), ("name", C.c_char_p), ("metadata", C.c_char_p),
                        ("flags", C.c_int64), ("n_children", C.c_int64),
                        ("children", C.POINTER(C.POINTER(ArrowSchema))), ("dictionary", C.POINTER(ArrowSchema)),
                        ("release", C.c_void_p), ("private_data", C.c_void_p)]


class ArrowArray(C.Structure):
    pass


ArrowArray._fields_ = [("length", C.c_int64), ("null_count", C.c_int64), ("offset", C.c_int64),
                       ("n_buffers", C.c_int64), ("n_children", C.c_int64), ("buffers", C.POINTER(C.c_void_p)),
                       ("children", C.POINTER(C.POINTER(ArrowArray))), ("dictionary", C.POINTER(ArrowArray)),
                       ("release", C.c_void_p), ("private_data", C.c_void_p)]


class ArrowDeviceArray(C.Structure):
    _fields_ = [("array", ArrowArray), ("device_id", C.c_int64), ("device_type", C.c_int32),
                ("sync_event", C.c_void_p), ("reserved", C.c_int64 * 3)]


def export_column(spec: S.Spec, cols: S.CColumns, k: int, device_type: int = ARROW_DEVICE_ROCM,
                  device_id: int = 0) -> Tuple[ArrowDeviceArray, ArrowSchema]:
    arr, sch = ArrowDeviceArray(), ArrowSchema()
    cs = spec.to_c()
    st = gpu.load().falx_export_arrow(C.byref(cs), C.byref(cols), k, device_type, device_id,
                                      C.byref(arr), C.byref(sch))
    if st != S.OK:
        raise gpu.FalxError("falx_export_arrow", st)
    return arr, sch


def to_pyarrow(arr: ArrowDeviceArray, sch: ArrowSchema):
    """Imports an exported column whose buffers are host memory (pyarrow
    takes over the structs and calls their release callbacks)."""
    import pyarrow as pa
    return pa.Array._import_from_c(C.addressof(arr.array), C.addressof(sch))


def download_to_pyarrow(ctx: gpu.Context, arr: ArrowDeviceArray, sch: ArrowSchema):
    """The same array with its device buffers copied to the host."""
    import pyarrow as pa

    def fetch(ptr, nbytes):
        out = np.empty(max(1, nbytes), dtype=np.uint8)
        if ptr and nbytes:
            gpu._check("falx_download", gpu.load().falx_download(ctx.h, out.ctypes.data, ptr, nbytes))
        return pa.py_buffer(out[:nbytes].tobytes())

    n = arr.array.length
    fmt = sch.format.decode()
    valid = fetch(arr.array.buffers[0], (n + 63) // 64 * 8)
    if fmt in ("l", "g"):
        t = pa.int64() if fmt == "l" else pa.float64()
        out = pa.Array.from_buffers(t, n, [valid, fetch(arr.array.buffers[1], 8 * n)])
    elif fmt == "z":
        offs = fetch(arr.array.buffers[1], 4 * (n + 1))
        end = int(np.frombuffer(offs, dtype=np.int32)[-1]) if n else 0
        out = pa.Array.from_buffers(pa.binary(), n, [valid, offs, fetch(arr.array.buffers[2], end)])
    else:
        child = arr.array.children[0].contents
        values = pa.Array.from_buffers(pa.uint32(), 2 * n, [None, fetch(child.buffers[1], 8 * n)])
        out = pa.Array.from_buffers(pa.list_(pa.uint32(), 2), n, [valid], children=[values])
    release = C.CFUNCTYPE(None, C.c_void_p)
    if arr.array.release:
        release(arr.array.release)(C.addressof(arr.array))
    if sch.release:
        release(sch.release)(C.addressof(sch))
    return out
