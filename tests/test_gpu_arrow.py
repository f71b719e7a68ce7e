"""Arrow export of device columns: the exported arrays point at the
columns' own HBM buffers (zero-copy, device_type ROCm), and their contents,
copied to the host, equal the oracle's columns."""
import numpy as np
import pytest

from falx_amd import arrow as A
from falx_amd import datagen
from falx_amd import spec as S

pa = pytest.importorskip("pyarrow")
pytestmark = pytest.mark.gpu


def test_device_columns_export_zero_copy(ctx, ora):
    spec = S.GEO_ALL.with_columns([(0, S.I64), (1, S.STR), (2, S.F64), (3, S.BYTES)], "geo_arrow")
    data, truth = datagen.geo_csv(2 * 1024 * 1024 + 3, seed=17)
    buf = ctx.upload(data)
    cols, keep = ctx.alloc_columns(spec, truth["rows"], data.size)
    assert ctx.parse_columns_into(spec, buf, data.size, cols) == S.OK
    want = ora.parse_columns(spec, data)
    for k, c in enumerate(spec.columns):
        arr, sch = A.export_column(spec, cols, k, A.ARROW_DEVICE_ROCM, 0)
        assert arr.device_type == A.ARROW_DEVICE_ROCM and arr.device_id == 0 and arr.array.length == truth["rows"]
        vals, valid, dbuf = keep[k]
        assert arr.array.buffers[0] == valid.ptr
        if c.type == S.BYTES:
            assert arr.array.children[0].contents.buffers[1] == vals.ptr
        else:
            assert arr.array.buffers[1] == vals.ptr
        if c.type == S.STR:
            assert arr.array.buffers[2] == dbuf.ptr
        got = A.download_to_pyarrow(ctx, arr, sch)
        w = want["cols"][k]
        rows = want["rows"]
        valid_w = np.unpackbits(w["valid"].view(np.uint8), bitorder="little")[:rows].astype(bool)
        assert got.is_valid().to_numpy(zero_copy_only=False).tolist() == valid_w.tolist(), k
        if c.type in (S.I64, S.F64):
            assert np.array_equal(got.fill_null(0).to_numpy().view(np.uint64), w["values"].view(np.uint64)), k
        elif c.type == S.STR:
            o, d = w["offsets"], w["data"].tobytes()
            assert got.fill_null(b"").to_pylist() == [d[o[r]:o[r + 1]] for r in range(rows)], k
        else:
            flat = got.values.to_numpy(zero_copy_only=False)
            assert np.array_equal(flat.reshape(-1, 2)[valid_w], w["values"][valid_w]), k
