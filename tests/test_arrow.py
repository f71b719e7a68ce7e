"""Arrow export (falx_export_arrow, the Arrow C Device Data Interface):
the layouts of every column type checked by pyarrow against the known
answers of the reference's examples/arrow_interop.rs (columns in host memory
here; the device case is in test_gpu_arrow.py)."""
import ctypes as C

import numpy as np
import pytest

from falx_amd import arrow as A
from falx_amd import spec as S

pa = pytest.importorskip("pyarrow")

# examples/arrow_interop.rs:30-36
EXAMPLE = (b"id,title,value,notes,label\n1,hello,2.5,skip,x\n2,\"quoted, with \"\"q\"\"\",3.25,skip,y\n"
           b"3,,4.5,skip,z\n4,unparsed-value,not-a-number,skip,w\n5\n")


def host_columns(ora, spec, data):
    """Columns parsed into host buffers (the CColumns struct points at them)."""
    d = np.frombuffer(data, dtype=np.uint8)
    cols, keep = ora.alloc_columns(spec, d.size + 1, d.size + 16)
    st = ora.lib().ora_parse_columns(C.byref(spec.to_c()), C.c_void_p(d.ctypes.data), d.size, 1, C.byref(cols))
    assert st == 0
    return cols, keep


def test_arrow_interop_known_answers(ora):
    spec = S.CSV_TYPED   # id i64 @0, title string @1, value f64 @2, label bytes @4
    cols, keep = host_columns(ora, spec, EXAMPLE)
    assert cols.rows == 6
    arrays = []
    for k in range(len(spec.columns)):
        arr, sch = A.export_column(spec, cols, k, A.ARROW_DEVICE_CPU)
        assert arr.device_type == A.ARROW_DEVICE_CPU and arr.array.length == 6 and arr.array.null_count == -1
        arrays.append(A.to_pyarrow(arr, sch))
    ids, title, value, label = arrays
    assert ids.type == pa.int64() and value.type == pa.float64() and title.type == pa.binary()
    # arrow_interop.rs:104-113: header + not-a-number + one-field rows are
    # null values; the one-field row is the only null title.
    assert value.null_count == 3 and title.null_count == 1
    assert value.to_pylist() == [None, 2.5, 3.25, 4.5, None, None]
    assert title.to_pylist() == [b"title", b"hello", b'quoted, with "q"', b"", b"unparsed-value", None]
    assert ids.to_pylist() == [None, 1, 2, 3, 4, 5]
    # bytes column: (start, end) spans of `label` into the input
    spans = label.to_pylist()
    assert spans[5] is None and [EXAMPLE[a:b] for a, b in spans[:5]] == [b"label", b"x", b"y", b"z", b"w"]


def test_arrow_export_matches_columns_everywhere(ora):
    """Every column of a randomized typed document: the imported array holds
    exactly the column's values and validity."""
    from test_gpu_parity import typed_random_doc
    from helpers import Rng
    rng = Rng(0xA770)
    spec = S.CSV_TYPED
    for _ in range(40):
        doc = typed_random_doc(rng)
        cols, keep = host_columns(ora, spec, doc)
        want = ora.parse_columns(spec, doc)
        for k, c in enumerate(spec.columns):
            got = A.to_pyarrow(*A.export_column(spec, cols, k, A.ARROW_DEVICE_CPU))
            w = want["cols"][k]
            valid = np.unpackbits(w["valid"].view(np.uint8), bitorder="little")[:want["rows"]].astype(bool)
            assert got.is_valid().to_numpy(zero_copy_only=False).tolist() == valid.tolist()
            if c.type in (S.I64, S.F64):
                vals = got.fill_null(0).to_numpy()
                assert np.array_equal(vals.view(np.uint64), w["values"].view(np.uint64))
            elif c.type == S.STR:
                o, d = w["offsets"], w["data"].tobytes()
                assert got.fill_null(b"").to_pylist() == [d[o[r]:o[r + 1]] for r in range(want["rows"])]


def test_arrow_export_rejects_bad_arguments():
    from falx_amd import gpu
    cols = S.CColumns()
    cols.rows_cap, cols.rows = 4, 5   # more rows than capacity
    with pytest.raises(gpu.FalxError):
        A.export_column(S.CSV_TYPED, cols, 0, A.ARROW_DEVICE_CPU)
    cols.rows = 2
    with pytest.raises(gpu.FalxError):
        A.export_column(S.CSV_TYPED, cols, 7, A.ARROW_DEVICE_CPU)   # no such column
