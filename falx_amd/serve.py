"""Serving loop: parse/aggregate requests admitted as they arrive, batched
into the batch executor (include/falx_gpu.h falx_batch_*), results left in
HBM, latency measured from arrival to completion.

The reference has no server: callers invoke the generated parsers per
document (src/kernels/<fmt>.rs). This module is the request loop that the
served form of those calls needs on a GPU:

* :class:`Corpus` keeps one document per served spec resident in HBM; a
  request names a window of it (16-byte aligned, as every entry point
  requires) and an op.
* :class:`BatchServer` runs continuous batching over ``slots`` executors,
  each with its own context (stream), scratch and output arena. A free slot
  takes queued requests in arrival order while their scratch and worst-case
  outputs fit its HBM budget, submits them as one batch and is polled; a
  request that arrives while batches are on the device waits for the next
  free slot, so batches of different slots overlap on the GPU.
* Requests of ``large`` bytes or more skip batching: they fill the device
  on their own, and the plain entry points run them with the fast column
  passes. On an idle device the loop makes the call itself; otherwise a
  host thread with a context of its own makes those calls in arrival order
  (ctypes releases the GIL for each), so batches of small requests keep
  running beside them.
* :class:`PlainServer` is the baseline: one plain entry-point call (one set
  of launches, one synchronisation) per request, in arrival order.

Both record per-request latency (completion minus arrival, host clock) and
the aggregate input throughput.
"""
from __future__ import annotations

import ctypes as C
import math
import queue as _queue
import sys
import threading
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np

from . import datagen, gpu
from . import spec as S

KiB, MiB, GiB = 1 << 10, 1 << 20, 1 << 30

# The served mix: (name, op, spec, generator).
SERVED = [
    ("geo_csv", S.OP_COLUMNS_STATS, S.GEO_LATLON, datagen.geo_csv),
    ("csv_hash", S.OP_COLUMNS_STATS, S.HASH_TYPED, datagen.hash_csv),
    ("tsv", S.OP_COLUMNS_STATS, S.TSV_TYPED, datagen.tsv),
    ("logfmt", S.OP_LOGFMT_PAIRS, None, datagen.logfmt),
    ("ndjson", S.OP_NDJSON_ID_SCORE, None, datagen.ndjson),
]
MIN_ROW = 8   # no record of the served generators is shorter: rows <= len / 8 + 2
GRAPH_MIN = 8  # batches of at least this many requests run as captured graphs


@dataclass
class Req:
    kind: int            # index into SERVED
    offset: int
    length: int
    arrival: float       # seconds from the start of the run
    done: float = -1.0
    status: int = -1
    result: Optional[Dict] = None


class Corpus:
    """One generated document per served spec, resident in HBM."""

    def __init__(self, ctx: gpu.Context, nbytes: int, seed: int = 1):
        self.ctx = ctx
        self.host, self.dev = [], []
        for k, (_, _, _, gen) in enumerate(SERVED):
            data, _ = gen(nbytes, seed=seed + 7919 * k)
            self.host.append(data)
            self.dev.append(ctx.upload(data))

    def size(self, kind: int) -> int:
        return int(self.host[kind].size)


def request_stream(corpus: Corpus, n: int, rate: float, seed: int = 7, lo: int = 4 * KiB,
                   hi: int = 256 * MiB) -> List[Req]:
    """n requests: spec uniform over the served mix, size log-uniform in
    [lo, hi] (capped by the corpus), Poisson arrivals at `rate` per second
    (rate <= 0: all at time 0)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for _ in range(n):
        kind = int(rng.integers(len(SERVED)))
        size = int(math.exp(rng.uniform(math.log(lo), math.log(hi))))
        size = min(size, corpus.size(kind))
        off = int(rng.integers(0, corpus.size(kind) - size + 1)) & ~15
        if rate > 0:
            t += float(rng.exponential(1.0 / rate))
        out.append(Req(kind, off, size, t))
    return out


def out_bytes(kind: int, length: int) -> int:
    """Worst-case device outputs of a request (columns ops)."""
    name, op, spec, _ = SERVED[kind]
    if op not in (S.OP_COLUMNS, S.OP_COLUMNS_STATS):
        return 0
    rows = length // MIN_ROW + 2
    b = 0
    for c in spec.columns:
        b += 8 * ((rows + 63) // 64) + 256
        b += 4 * (rows + 1) + length + 256 if c.type == S.STR else 8 * rows
    return (b + 255) & ~255


class Arena:
    """A device buffer handed out by bump allocation, reset per batch."""

    def __init__(self, ctx: gpu.Context, nbytes: int):
        self.buf = ctx.alloc(nbytes)
        self.cap, self.used = nbytes, 0

    def take(self, n: int) -> int:
        n = (n + 255) & ~255
        if self.used + n > self.cap:
            raise MemoryError("arena exhausted")
        p = self.buf.ptr + self.used
        self.used += n
        return p

    def reset(self):
        self.used = 0


def arena_columns(spec: S.Spec, length: int, arena: Arena) -> S.CColumns:
    rows = length // MIN_ROW + 2
    cols = S.CColumns()
    cols.rows_cap = rows
    for k, c in enumerate(spec.columns):
        cols.cols[k].valid = arena.take(8 * ((rows + 63) // 64))
        if c.type == S.STR:
            cols.cols[k].values = arena.take(4 * (rows + 1))
            cols.cols[k].data = arena.take(max(16, length))
            cols.cols[k].data_cap = max(16, length)
        else:
            cols.cols[k].values = arena.take(8 * rows)
    return cols


class _View:
    """A window of a resident document as a DeviceBuffer-like pointer."""

    def __init__(self, buf: gpu.DeviceBuffer, offset: int, length: int):
        self.ptr = buf.ptr + offset
        self.nbytes = length


def _request(corpus: Corpus, r: Req, arena: Arena, chunk: int = 0) -> gpu.Request:
    name, op, spec, _ = SERVED[r.kind]
    view = _View(corpus.dev[r.kind], r.offset, r.length)
    cols = arena_columns(spec, r.length, arena) if op in (S.OP_COLUMNS, S.OP_COLUMNS_STATS) else None
    return gpu.Request(op, view, r.length, spec=spec, columns=cols, chunk=chunk)


def plain_call(ctx: gpu.Context, arena: Arena, corpus: Corpus, r: Req) -> Dict:
    """One request through its plain entry point (outputs in `arena`)."""
    name, op, spec, _ = SERVED[r.kind]
    view = _View(corpus.dev[r.kind], r.offset, r.length)
    lib = gpu.load()
    if op == S.OP_COLUMNS_STATS:
        arena.reset()
        cols = arena_columns(spec, r.length, arena)
        st = S.CCsvStats()
        cs = spec.to_c()
        status = lib.falx_parse_columns_with_stats(ctx.h, C.byref(cs), view.ptr, r.length,
                                                   C.byref(cols), C.byref(st))
        return {"status": status, "stats": gpu._stats_dict(spec, st)}
    if op == S.OP_LOGFMT_PAIRS:
        st = S.CLogfmtStats()
        status = lib.falx_parse_logfmt_pairs(ctx.h, view.ptr, r.length, C.byref(st))
        return {"status": status, "logfmt": {"pairs": st.pairs, "key_bytes": st.key_bytes,
                                             "value_bytes": st.value_bytes, "checksum": st.checksum}}
    st = S.CNdjsonIdScore()
    status = lib.falx_parse_ndjson_id_score(ctx.h, view.ptr, r.length, C.byref(st))
    return {"status": status, "ndjson": {"records": st.records, "invalid": st.invalid, "id_sum": st.id_sum,
                                         "score_fixed6_sum": st.score_fixed6_sum,
                                         "score_fixed6_cells": st.score_fixed6_cells,
                                         "score_sum": st.score_sum}}


class _LargeLane:
    """Large requests as plain calls on a host thread and context of their own."""

    def __init__(self, corpus: Corpus, out_budget: int):
        self.corpus = corpus
        self.ctx = gpu.Context(0)
        self.arena = Arena(self.ctx, out_budget)
        self.todo: _queue.Queue = _queue.Queue()
        self.finished: _queue.Queue = _queue.Queue()
        self.busy = 0
        self.thread = threading.Thread(target=self._work, daemon=True)
        self.thread.start()

    def _work(self):
        while True:
            item = self.todo.get()
            if item is None:
                return
            r, t0 = item
            try:
                x = plain_call(self.ctx, self.arena, self.corpus, r)
            except Exception as e:  # reported as the request's status
                x = {"status": S.E_DEVICE, "error": repr(e)}
            r.done, r.status, r.result = time.perf_counter() - t0, x["status"], x
            self.finished.put(r)

    def submit(self, r: Req, t0: float):
        self.busy += 1
        self.todo.put((r, t0))

    def reap(self) -> int:
        n = 0
        while True:
            try:
                self.finished.get_nowait()
            except _queue.Empty:
                break
            n += 1
        self.busy -= n
        return n

    def close(self):
        self.todo.put(None)
        self.thread.join()
        self.arena.buf.free()
        self.ctx.close()


@dataclass
class _Slot:
    ctx: gpu.Context
    batch: gpu.Batch
    arena: Arena
    budget: int
    reqs: List[Req] = field(default_factory=list)


class BatchServer:
    """Continuous batching over `slots` executors (see module docstring)."""

    def __init__(self, corpus: Corpus, slots: int = 2, hbm_budget: int = 8 * GiB, max_batch: int = 256,
                 tuning: Optional[Dict] = None, large: int = 0, large_out_budget: int = 2 * GiB,
                 exclusive: bool = False, hold_below: int = 0):
        self.corpus = corpus
        self.large = large
        self.exclusive = exclusive
        self.hold_below = hold_below
        self.lane = _LargeLane(corpus, large_out_budget) if large else None
        self.max_batch = max_batch
        self.tuning = tuning or {}
        per = hbm_budget // slots
        self.slots = []
        for _ in range(slots):
            ctx = gpu.Context(0)
            scratch = per // 4
            b = gpu.Batch(ctx, scratch)
            t = self.tuning.get("executor")
            bt = b.tuning()
            if t:
                bt.tile_workgroups = t["tile_workgroups"]
                bt.unit_workgroups = t["unit_workgroups"]
                for k, v in enumerate(t["chunk"]):
                    bt.chunk[k] = v
            # Batches of GRAPH_MIN requests or more replay captured graphs;
            # the small batches of light load launch their stages directly.
            bt.use_graph = GRAPH_MIN
            b.set_tuning(bt)
            self.slots.append(_Slot(ctx, b, Arena(ctx, per - scratch), per))
        self._scratch = {}

    def close(self):
        if self.lane:
            self.lane.close()
        for s in self.slots:
            s.batch.close()
            s.ctx.close()

    def _chunk(self, r: Req) -> int:
        per_spec = self.tuning.get("chunk", {}).get(SERVED[r.kind][0])
        if not per_spec:
            return 0
        tiles = (r.length + 8191) // 8192
        k = 0
        while k + 1 < len(per_spec) and tiles >= 4 ** (k + 1):
            k += 1
        return per_spec[k]

    def _scratch_of(self, r: Req) -> int:
        # falx_request_scratch_bytes depends on the op, spec and tile count only
        key = (r.kind, (r.length + 8191) // 8192)
        v = self._scratch.get(key)
        if v is None:
            name, op, spec, _ = SERVED[r.kind]
            v = gpu.Request(op, None, r.length, spec=spec).scratch_bytes()
            self._scratch[key] = v
        return v

    def _fill(self, s: _Slot, queue: deque, now: float) -> bool:
        take, scratch, outs = [], 0, 0
        scratch_cap = s.budget // 4
        while queue and len(take) < self.max_batch:
            r = queue[0]
            need_s, need_o = self._scratch_of(r), out_bytes(r.kind, r.length)
            if scratch + need_s > scratch_cap or outs + need_o > s.arena.cap:
                if take:
                    break
                raise MemoryError(f"request of {r.length} B exceeds a slot's HBM budget")
            take.append(queue.popleft())
            scratch += need_s
            outs += need_o
        if not take:
            return False
        s.arena.reset()
        reqs = [_request(self.corpus, r, s.arena, self._chunk(r)) for r in take]
        s.batch.submit(reqs)
        s.reqs = take
        return True

    def run(self, reqs: List[Req]) -> Dict:
        # The lane's thread waits for the GIL when its call returns: a short
        # switch interval keeps that wait small while this loop polls.
        old_switch = sys.getswitchinterval()
        sys.setswitchinterval(1e-4)
        try:
            return self._run(reqs)
        finally:
            sys.setswitchinterval(old_switch)

    def _run(self, reqs: List[Req]) -> Dict:
        pending = deque(sorted(reqs, key=lambda r: r.arrival))
        queue: deque = deque()
        large_q: deque = deque()
        left = len(reqs)
        batches = 0
        max_inflight = 0
        t0 = time.perf_counter()
        while left:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                (large_q if self.lane and r.length >= self.large else queue).append(r)
            progressed = False
            if self.lane and self.lane.busy:
                k = self.lane.reap()
                left -= k
                progressed = progressed or k > 0
            if large_q:
                if not self.lane.busy and not any(s.reqs for s in self.slots):
                    # Idle device: the call is made here, as the plain path
                    # would, with no hand-over to the lane's thread.
                    r = large_q.popleft()
                    x = plain_call(self.lane.ctx, self.lane.arena, self.corpus, r)
                    r.done, r.status, r.result = time.perf_counter() - t0, x["status"], x
                    left -= 1
                    progressed = True
                    continue
                while large_q:
                    self.lane.submit(large_q.popleft(), t0)
            for s in self.slots:
                if s.reqs and s.batch.done():
                    res = s.batch.wait()
                    t = time.perf_counter() - t0
                    for r, x in zip(s.reqs, res):
                        r.done, r.status, r.result = t, x["status"], x
                    left -= len(s.reqs)
                    s.reqs = []
                    progressed = True
            # With `exclusive`, small requests wait while a large one runs
            # (as they would behind it in arrival order) instead of sharing
            # the device with it; with `hold_below`, they wait only while
            # fewer than that many are queued (light load keeps the large
            # request's latency, a backlog shares the device).
            lane_busy = self.lane is not None and self.lane.busy > 0
            hold = lane_busy and (self.exclusive or len(queue) < self.hold_below)
            # One queued request on an idle device: a batch of one would only
            # add the batch set-up to its latency, so it is made as a plain
            # call here (on the first slot's context and output arena).
            if (len(queue) == 1 and not hold and not any(s.reqs for s in self.slots)
                    and not (self.lane and self.lane.busy)):
                r = queue.popleft()
                s0 = self.slots[0]
                x = plain_call(s0.ctx, s0.arena, self.corpus, r)
                r.done, r.status, r.result = time.perf_counter() - t0, x["status"], x
                left -= 1
                continue
            for s in self.slots:
                if not s.reqs and queue and not hold:
                    if self._fill(s, queue, now):
                        batches += 1
                        progressed = True
            max_inflight = max(max_inflight, sum(1 for s in self.slots if s.reqs))
            # Idle device: sleep until the next arrival. A batch in flight is
            # polled instead, so its completion is seen when it happens.
            busy = any(s.reqs for s in self.slots) or bool(self.lane and self.lane.busy)
            if not progressed and not queue and pending and not busy:
                wait = pending[0].arrival - (time.perf_counter() - t0)
                if wait > 2e-4:
                    time.sleep(wait - 1e-4)
            elif not progressed and self.lane and self.lane.busy:
                time.sleep(0)   # lets the lane's thread take the GIL when its call returns
        return summarize(reqs, extra={"batches": batches, "slots": len(self.slots), "max_inflight": max_inflight})


class PlainServer:
    """One plain entry-point call per request, in arrival order."""

    def __init__(self, corpus: Corpus, out_budget: int = 4 * GiB):
        self.corpus = corpus
        self.ctx = corpus.ctx
        self.arena = Arena(self.ctx, out_budget)

    def call(self, r: Req) -> Dict:
        return plain_call(self.ctx, self.arena, self.corpus, r)

    def run(self, reqs: List[Req]) -> Dict:
        t0 = time.perf_counter()
        for r in sorted(reqs, key=lambda r: r.arrival):
            while True:
                now = time.perf_counter() - t0
                if now >= r.arrival:
                    break
                if r.arrival - now > 2e-4:
                    time.sleep(r.arrival - now - 1e-4)
            x = self.call(r)
            r.done, r.status, r.result = time.perf_counter() - t0, x["status"], x
        return summarize(reqs)


def summarize(reqs: List[Req], extra: Optional[Dict] = None) -> Dict:
    lat = np.array([r.done - r.arrival for r in reqs])
    nbytes = sum(r.length for r in reqs)
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    out = {
        "requests": len(reqs),
        "input_bytes": nbytes,
        "seconds": span,
        "gib_s": nbytes / span / GiB if span > 0 else 0.0,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "errors": sum(1 for r in reqs if r.status not in (S.OK, S.E_SCHEMA)),
    }
    if extra:
        out.update(extra)
    return out


def fresh(reqs: List[Req]) -> List[Req]:
    return [Req(r.kind, r.offset, r.length, r.arrival) for r in reqs]
