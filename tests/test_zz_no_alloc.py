"""falx_index_structurals_static (no allocation, scratch in a file-scope
device array) against falx_index_structurals, bit for bit, through the C ABI:
two calls in a row with the second on a smaller input (the shared scratch
keeps the first call's contents), a capacity shortfall, and the size limit."""
import ctypes as C

import numpy as np
import pytest

from falx_amd import datagen, gpu
from falx_amd import spec as S

pytestmark = pytest.mark.gpu

STATIC_MAX = 64 << 20


def _static(ctx, spec, buf, n, out, cap):
    cnt = C.c_uint64()
    cs = spec.to_c()
    st = gpu.load().falx_index_structurals_static(ctx.h, C.byref(cs), buf.ptr, n, out.ptr if out else None,
                                                  cap, C.byref(cnt))
    return st, cnt.value


def _plain(ctx, spec, buf, n):
    out = ctx.alloc(4 * max(1, n))
    cnt = C.c_uint64()
    cs = spec.to_c()
    st = gpu.load().falx_index_structurals(ctx.h, C.byref(cs), buf.ptr, n, out.ptr, max(1, n), C.byref(cnt))
    assert st == S.OK
    return out.download(np.uint32, cnt.value)


def test_static_index_equals_plain_index(ctx):
    big, _ = datagen.hash_csv(STATIC_MAX - 4096, seed=31)       # near the limit, comment+quote regions
    small, _ = datagen.geo_csv(3 * 1024 * 1024 + 5, seed=32)    # smaller, second
    assert big.size <= STATIC_MAX
    cases = [(S.HASH_TYPED, big), (S.GEO_LATLON, small), (S.LOGFMT, datagen.logfmt(200_000, seed=3)[0])]
    for spec, data in cases:
        print(f"static index: {spec.name} {data.size} B", flush=True)
        buf = ctx.upload(data)
        n = int(data.size)
        want = _plain(ctx, spec, buf, n)
        out = ctx.alloc(4 * (want.size + 64))
        st, cnt = _static(ctx, spec, buf, n, out, want.size + 64)
        assert st == S.OK and cnt == want.size, (spec.name, st, cnt, want.size)
        assert np.array_equal(out.download(np.uint32, cnt), want), spec.name


def test_static_index_capacity_and_limit(ctx):
    data, _ = datagen.geo_csv(1 << 20, seed=33)
    buf = ctx.upload(data)
    n = int(data.size)
    want = _plain(ctx, S.GEO_LATLON, buf, n)
    out = ctx.alloc(4 * 100)
    st, cnt = _static(ctx, S.GEO_LATLON, buf, n, out, 100)
    assert st == S.E_CAPACITY and cnt == want.size
    assert np.array_equal(out.download(np.uint32, 100), want[:100])
    # Over the limit: rejected before any device work (the length is only
    # checked, the buffer is never read).
    st, cnt = _static(ctx, S.GEO_LATLON, buf, STATIC_MAX + 16, out, 100)
    assert st == S.E_TOO_LARGE and cnt == 0
    st, cnt = _static(ctx, S.GEO_LATLON, buf, 0, out, 100)
    assert st == S.OK and cnt == 0
