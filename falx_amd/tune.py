"""Autotuner for batched execution: picks the persistent grid of the batch
stages and, per served spec and size class, the chunk (tiles per queue
unit), by timing batches on the device (falx_batch_last_ms).

Size class k holds documents of [4^k, 4^(k+1)) tiles of 8 KiB, as in
include/falx_gpu.h falx_batch_tuning. For each (spec, class) a batch of
documents of that class (windows of a resident corpus, as the serving loop
uses) is timed at every candidate chunk; the grid is chosen first, on the
whole mix. The result is written as JSON:

    {"executor": {"tile_workgroups": ..., "unit_workgroups": ..., "chunk": [8]},
     "chunk": {"<spec>": [8 per class]}, "device": "...", "trials": [...]}

Usage: python -m falx_amd.tune [out.json]
"""
from __future__ import annotations

import json
import os
import sys
import time
from typing import Dict, List

import numpy as np

from . import gpu, serve
from . import spec as S

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")
TILE = 8192
CHUNKS = [1, 2, 4, 8, 16]
GRID_MULT = [2, 4, 8, 12, 16]
CLASS_TILES = [1, 4, 16, 64, 256, 1024, 4096, 16384]   # one size per class
BATCH_BYTES = 96 << 20                                   # per timed batch (at most)
MAX_DOCS = 128


def load(path: str = DEFAULT_PATH) -> Dict:
    if not os.path.exists(path):
        return {}
    with open(path) as f:
        return json.load(f)


def _batch_reqs(corpus: serve.Corpus, kind: int, tiles: int, rng) -> List[serve.Req]:
    size = min(tiles * TILE - TILE // 2 if tiles > 1 else 6000, corpus.size(kind))
    n = max(1, min(MAX_DOCS, BATCH_BYTES // max(size, 1)))
    out = []
    for _ in range(n):
        off = int(rng.integers(0, corpus.size(kind) - size + 1)) & ~15
        out.append(serve.Req(kind, off, size, 0.0))
    return out


def _time(b: gpu.Batch, corpus: serve.Corpus, arena: serve.Arena, reqs, chunk: int, reps: int = 3) -> float:
    best = float("inf")
    for _ in range(reps):
        arena.reset()
        b.submit([serve._request(corpus, r, arena, chunk) for r in reqs])
        b.wait()
        best = min(best, b.last_ms)
    return best


def tune(corpus_bytes: int = 160 << 20, seed: int = 5, log=print) -> Dict:
    ctx = gpu.Context(0)
    corpus = serve.Corpus(ctx, corpus_bytes, seed=seed)
    b = gpu.Batch(ctx, 4 << 30)
    arena = serve.Arena(ctx, 6 << 30)
    base = b.tuning()
    n_cu = base.unit_workgroups
    rng = np.random.default_rng(seed)
    trials = []
    # 1. persistent grid on a mixed batch of every spec and class
    mix = []
    for kind in range(len(serve.SERVED)):
        for tiles in CLASS_TILES[:7]:
            mix += _batch_reqs(corpus, kind, tiles, rng)[:4]
    best_grid, best_ms = None, float("inf")
    for m in GRID_MULT:
        t = b.tuning()
        t.tile_workgroups = m * n_cu
        b.set_tuning(t)
        ms = _time(b, corpus, arena, mix, 0)
        trials.append({"stage": "grid", "tile_workgroups": m * n_cu, "ms": ms})
        log(f"grid {m}x{n_cu}: {ms:.3f} ms")
        if ms < best_ms:
            best_grid, best_ms = m * n_cu, ms
    t = b.tuning()
    t.tile_workgroups = best_grid
    b.set_tuning(t)
    # 2. chunk per (spec, size class)
    table: Dict[str, List[int]] = {}
    for kind, (name, _, _, _) in enumerate(serve.SERVED):
        row = []
        for k, tiles in enumerate(CLASS_TILES):
            reqs = _batch_reqs(corpus, kind, tiles, rng)
            best_c, best = 1, float("inf")
            for c in CHUNKS:
                if c > max(1, tiles):
                    break
                ms = _time(b, corpus, arena, reqs, c)
                trials.append({"stage": "chunk", "spec": name, "class": k, "chunk": c, "ms": ms,
                               "docs": len(reqs), "bytes": int(sum(r.length for r in reqs))})
                if ms < best * 0.98:   # prefer the smaller chunk unless clearly faster
                    best_c, best = c, ms
            row.append(best_c)
            log(f"{name} class {k} ({tiles} tiles x {len(reqs)}): chunk {best_c} ({best:.3f} ms)")
        table[name] = row
    votes = np.array(list(table.values()))
    exec_chunk = [int(np.median(votes[:, k])) for k in range(len(CLASS_TILES))]
    out = {
        "executor": {"tile_workgroups": int(best_grid), "unit_workgroups": int(n_cu), "chunk": exec_chunk},
        "chunk": table,
        "device": "gfx950",
        "trials": trials,
        "when": time.strftime("%Y-%m-%d %H:%M:%S"),
    }
    b.close()
    ctx.close()
    return out


if __name__ == "__main__":
    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_PATH
    res = tune()
    full = dict(res)
    res["trials"] = len(res["trials"])   # the table keeps the count; the log has every trial
    with open(path, "w") as f:
        json.dump(res, f, indent=1)
    for t in full["trials"]:
        print(json.dumps(t))
    print(json.dumps(res))
