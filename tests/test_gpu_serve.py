"""The serving loop (falx_amd/serve.py): requests batched by continuous
batching over overlapping slots, or sent down the large-request lane (its
thread, or the loop itself on an idle device), get exactly the results that
one plain entry-point call per request gives, FP64 sums bit for bit."""
import numpy as np
import pytest

from falx_amd import serve
from falx_amd import spec as S



def _same(a, b):
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, float):
        return np.float64(a).tobytes() == np.float64(b).tobytes()
    return a == b


@pytest.mark.gpu
@pytest.mark.parametrize("slots,large", [(2, 0), (3, 1 << 20)], ids=["batched", "large_lane"])
def test_served_results_equal_plain_results(ctx, slots, large):
    corpus = serve.Corpus(ctx, 6 << 20, seed=11)
    reqs = serve.request_stream(corpus, 80, rate=4000.0, seed=5, lo=4096, hi=6 << 20)
    bs = serve.BatchServer(corpus, slots=slots, hbm_budget=2 << 30, max_batch=12, large=large,
                           large_out_budget=256 << 20)
    try:
        a = serve.fresh(reqs)
        sa = bs.run(a)
    finally:
        bs.close()
    p = serve.fresh(reqs)
    serve.PlainServer(corpus, out_budget=512 << 20).run(p)
    assert sa["errors"] == 0 and sa["batches"] > 1
    for i, (x, y) in enumerate(zip(a, p)):
        assert x.status == y.status, f"request {i}: {x.status} != {y.status}"
        assert _same(x.result, y.result), f"request {i} ({serve.SERVED[x.kind][0]}, {x.length} B) differs"


def test_request_stream_shape():
    class FakeCorpus:
        def size(self, kind):
            return 300 << 20
    reqs = serve.request_stream(FakeCorpus(), 2000, rate=100.0, seed=1)
    sizes = np.array([r.length for r in reqs])
    assert sizes.min() >= 4096 and sizes.max() <= 256 << 20
    assert all(r.offset % 16 == 0 for r in reqs)
    # log-uniform: the median is near the geometric mean of the bounds (1 MiB)
    assert 0.5 * (1 << 20) < np.median(sizes) < 2 * (1 << 20)
    gaps = np.diff([r.arrival for r in reqs])
    assert abs(gaps.mean() - 0.01) < 0.002
