#!/usr/bin/env python3
"""Headline benchmark: 1 GiB geo CSV (`id,City,Latitude,Longitude`), lat/lon
f64 column materialisation plus the fused count/sum of the same columns in
one pass over the HBM-resident input, on one MI355X.

A step is one `falx_parse_columns_with_stats` call over the whole 1 GiB
(inputs already in HBM, outputs preallocated). Prints one JSON line with the
input GiB/s, the roofline of the scan kernel (HIP events on its stream) and
the CPU oracle's parallel baseline timed on this host. Extra lanes are
reported under "lanes", each with the oracle's parallel CPU path on the same
input beside it:
  geo CSV 1 GiB: lat/lon columns, fused count/sum, City + lat/lon columns,
    structural index, pinned-H2D end-to-end, streamed tape in 64 MiB feeds,
    BGZF inflation and BGZF end-to-end (compressed upload + inflate + columns);
  NDJSON 1 GiB: sum(id + nested.score), JSON nested tape;
  csv_hash 1 GiB: structural index, typed columns (i64, f64, string),
    field bytes.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from fractions import Fraction

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
GIB = 1 << 30


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_latlon(n_in: int, rows: int) -> int:
    # input + two f64 columns + two validity bitmaps
    return n_in + 16 * rows + 2 * 8 * ((rows + 63) // 64)


def load_traffic(name: str):
    path = os.path.join(ROOT, "profiles", name)
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def time_steps(fn, steps: int, warmup: int, ctx):
    for _ in range(warmup):
        fn()
    kms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
        kms.append(ctx.last_kernel_ms)
    t1 = time.perf_counter()
    return (t1 - t0) * 1000.0 / steps, sum(kms) / len(kms)


def cpu_time(fn, runs: int, warmups: int = 2):
    """Median seconds of `runs` calls of fn after `warmups` untimed calls
    (the reference's bench_columns protocol: 2 warm-ups, 9 runs, median)."""
    for _ in range(warmups):
        fn()
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    times.sort()
    return times[len(times) // 2]


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads(requested: int) -> int:
    """Host threads for the CPU legs: the box's CPU share for one GPU (16),
    or fewer when the host has fewer."""
    return requested or min(os.cpu_count() or 1, 16)


def cpu_columns(spec, data, rows: int, data_cap: int, threads: int):
    """Closure running the oracle's parse_columns_par into preallocated
    columns (allocation outside the timed call, as on the GPU)."""
    from oracle import oracle as O
    cols, keep = O.alloc_columns(spec, rows + 1, data_cap)
    cs = spec.to_c()
    lib = O.lib()

    def run():
        st = lib.ora_parse_columns(C.byref(cs), C.c_void_p(data.ctypes.data), data.size, threads, C.byref(cols))
        assert st == 0, st
        return keep
    return run


def cpu_baseline(data, rows: int, runs: int, threads: int, warmups: int = 2):
    """The oracle's restatement of parse_columns_par (csv_geo.rs:1167-1284)
    followed by the sequential sums of the two columns (bench_columns.rs:
    151-179), at `threads` host threads, on the same input: `warmups`
    untimed runs, then `runs` timed ones (median and best)."""
    import numpy as np

    from falx_amd import spec as S
    from oracle import oracle as O

    spec = S.GEO_LATLON
    lib = O.lib()
    times = []
    lat = lon = None
    for it in range(warmups + runs):
        cols, keep = O.alloc_columns(spec, rows, 16)
        cs = spec.to_c()
        t0 = time.perf_counter()
        st = lib.ora_parse_columns(C.byref(cs), C.c_void_p(data.ctypes.data), data.size, threads, C.byref(cols))
        lat, lon = keep[0][0][:rows], keep[1][0][:rows]
        s_lat, s_lon = float(np.cumsum(lat)[-1]), float(np.cumsum(lon)[-1])
        if it >= warmups:
            times.append(time.perf_counter() - t0)
        assert st == 0 and cols.rows == rows
    times.sort()
    med = times[len(times) // 2]
    return data.size / GIB / med, med, times, (s_lat, s_lon)


def sum_errors(name: str, gpu_tree: float, ref_fold: float, exact) -> dict:
    """FP64 sum of a column: the GPU's blocked tree and the reference's
    sequential fold, each against the exact rational sum of the decimal
    values written (fixed-6 mantissas / 10^6)."""
    return {f"{name}_sum_gpu_tree": gpu_tree, f"{name}_sum_ref_fold": ref_fold, f"{name}_exact": float(exact),
            f"{name}_err_gpu_vs_exact": float(abs(Fraction(gpu_tree) - exact)),
            f"{name}_err_fold_vs_exact": float(abs(Fraction(ref_fold) - exact)),
            f"{name}_err_gpu_vs_fold": abs(gpu_tree - ref_fold)}


def nested_lane(ctx, nd, nd_buf, args, lanes, thr) -> None:
    """The JSON nested tape (json::parse_nested) of the 1 GiB NDJSON document:
    GPU throughput, algorithmic bytes = input + 8 B per tape entry, checked
    at full size against the JSON structural index (positions), the pairing
    invariants (partners mutual, opens before closes, matching kinds) and,
    with the CPU leg, the oracle's tape word for word."""
    import numpy as np

    from falx_amd import spec as S
    from oracle import oracle as O

    nd_n = int(nd.size)
    probe = ctx.alloc(8)
    _, need, _, _ = ctx.parse_nested_into(nd_buf, nd_n, probe)
    probe.free()
    tape = ctx.alloc(8 * max(1, need))
    res = [None]

    def step():
        res[0] = ctx.parse_nested_into(nd_buf, nd_n, tape)
    log("timing lane json_nested_tape")
    lms, lkms = time_steps(step, max(3, args.steps // 2), 1, ctx)
    st, cnt, err, _ = res[0]
    ab = nd_n + 8 * cnt
    lane = {"gib_s": round(nd_n / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3),
            "kernel_ms": round(lkms, 3), "hbm_frac": round(ab / (lkms / 1000.0) / 1e9 / HBM_PEAK_GBS, 3),
            "bytes": nd_n, "entries": int(cnt), "algorithmic_bytes": int(ab), "error": int(err)}
    got = tape.download(np.uint64, cnt)
    tape.free()
    jspec = S.Spec("json", b"{}[],:", ord('"'), ord("\\"), None, ())
    idx = ctx.index_structurals(jspec, nd_buf, nd_n)
    pos = (got & np.uint64(0xFFFFFFFF)).astype(np.uint32)
    part = (got >> np.uint64(32)).astype(np.int64)
    del got
    kinds = nd[pos]
    opens = (kinds == ord("{")) | (kinds == ord("["))
    closes = (kinds == ord("}")) | (kinds == ord("]"))
    i = np.arange(cnt, dtype=np.int64)
    po = part[opens]
    ok = (st == S.OK and err == 0 and np.array_equal(pos, idx) and bool(np.all(part[~(opens | closes)] == 0))
          and bool(np.all((po > i[opens]) & (po < cnt))) and bool(np.all(part[po] == i[opens]))
          and bool(np.all(kinds[po] == kinds[opens] + 2)) and bool(np.all(part[closes] < i[closes])))
    lane["verified"] = bool(ok)
    del idx, kinds, opens, closes, po, i
    if not args.no_cpu:
        log("timing CPU oracle lane json_nested_tape")
        lib = O.lib()
        ct = np.empty(max(1, nd_n), dtype=np.uint64)
        e, p = C.c_int32(), C.c_uint64()

        def cpu():
            return lib.ora_parse_nested(C.c_void_p(nd.ctypes.data), nd_n, thr, C.c_void_p(ct.ctypes.data), ct.size,
                                        C.byref(e), C.byref(p))
        lane["cpu_gib_s"] = round(nd_n / GIB / cpu_time(cpu, args.cpu_runs), 3)
        lane["cpu_threads"] = thr
        n_c = cpu()
        lane["oracle_equal"] = bool(n_c == cnt and e.value == err and np.array_equal(
            (ct[:n_c] >> np.uint64(32)).astype(np.int64), part) and np.array_equal(
            (ct[:n_c] & np.uint64(0xFFFFFFFF)).astype(np.uint32), pos))
        del ct
    lanes["json_nested_tape"] = lane


def stream_lane(ctx, data, dbuf, n_in, rows, args, lanes) -> None:
    """Streaming ingestion (StreamParser) of the 1 GiB geo CSV in 64 MiB feeds
    from device memory: each feed's tape of completed records. Verified by
    the closed forms (one terminator per record, three separators each).
    The CPU leg is the oracle's serial tape (the reference's stream is
    serial) on a 256 MiB prefix."""
    import numpy as np

    from falx_amd import gpu
    from falx_amd import spec as S

    lib = gpu.load()
    spec = S.GEO_LATLON
    cs = spec.to_c()
    feed = 64 << 20
    h = C.c_void_p()
    gpu._check("falx_stream_create", lib.falx_stream_create(ctx.h, C.byref(cs), 2 * feed, C.byref(h)))
    d_seps, d_ends = ctx.alloc(4 * (feed // 4)), ctx.alloc(8 * (feed // 8))
    tot = [0, 0]

    def run():
        gpu._check("falx_stream_reset", lib.falx_stream_reset(h))
        ns = ne = 0
        out = S.CStreamOut()
        for k in range(0, n_in, feed):
            gpu._check("falx_stream_feed", lib.falx_stream_feed(h, dbuf.ptr + k, min(feed, n_in - k), d_seps.ptr,
                                                                d_seps.nbytes // 4, d_ends.ptr, d_ends.nbytes // 8,
                                                                C.byref(out)))
            ns, ne = ns + out.n_seps, ne + out.n_ends
        gpu._check("falx_stream_finish", lib.falx_stream_finish(h, d_seps.ptr, d_seps.nbytes // 4, d_ends.ptr,
                                                                d_ends.nbytes // 8, C.byref(out)))
        tot[0], tot[1] = ns + out.n_seps, ne + out.n_ends + (1 if out.data_end else 0)
    log("timing lane stream_tape_64mib_feeds")
    lms, _ = time_steps(run, max(3, args.steps // 2), 1, ctx)
    lib.falx_stream_destroy(h)
    lane = {"gib_s": round(n_in / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3), "feed_bytes": feed,
            "records": int(tot[1]), "verified": tot[1] == rows and tot[0] == 3 * rows}
    if not args.no_cpu:
        from oracle import oracle as O
        sample = data[: 256 << 20]
        log("timing CPU oracle lane stream_tape_64mib_feeds (serial tape, 256 MiB prefix)")
        lane["cpu_gib_s"] = round(sample.size / GIB / cpu_time(lambda: O.parse(spec, sample, 1), 3, warmups=1), 3)
        lane["cpu_threads"] = 1
        lane["cpu_sample"] = "first 268435456 B, serial tape (ora_parse at 1 thread), 1 warm-up then 3 runs, median"
    lanes["stream_tape_64mib_feeds"] = lane
    d_seps.free()
    d_ends.free()


def bgzf_lanes(ctx, data, truth, spec, cols, args, lanes, thr) -> None:
    """The 1 GiB geo CSV as BGZF (bgzip framing, zlib level 6): inflation of
    the HBM-resident compressed bytes (GB/s of output; algorithmic bytes =
    compressed + inflated), and end to end from pinned host memory: the
    compressed upload, inflation and the headline columns + sums. CPU legs:
    zlib (the reference's flate2 backend family) over the blocks on `thr`
    threads, and that plus the oracle's parse_columns_par."""
    import zlib
    from concurrent.futures import ThreadPoolExecutor

    import numpy as np

    from falx_amd import bgzf, gpu
    from falx_amd import spec as S

    n_in, rows = int(data.size), int(truth["rows"])
    log("bgzipping the geo CSV (zlib level 6)")
    comp = bgzf.bgzip(memoryview(data), 6, threads=thr)
    blocks, total = bgzf.scan(comp)
    assert total == n_in
    c_n = len(comp)
    d_comp = ctx.upload(np.frombuffer(comp, dtype=np.uint8))
    d_out = ctx.alloc(n_in + 16)
    lib = gpu.load()
    b = np.ascontiguousarray(blocks)
    bad = C.c_uint64()

    def inflate():
        gpu._check("falx_bgzf_inflate", lib.falx_bgzf_inflate(ctx.h, d_comp.ptr, c_n, b.ctypes.data, b.size,
                                                               d_out.ptr, d_out.nbytes, C.byref(bad)))
    log("timing lane bgzf_inflate")
    lms, lkms = time_steps(inflate, max(3, args.steps // 2), 1, ctx)
    ab = c_n + n_in
    same = bool(np.array_equal(d_out.download(np.uint8, n_in), data))
    lane = {"gib_s": round(n_in / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3), "kernel_ms": round(lkms, 3),
            "hbm_frac": round(ab / (lkms / 1000.0) / 1e9 / HBM_PEAK_GBS, 3), "compressed_bytes": c_n,
            "blocks": int(b.size), "verified": same}
    # End to end: pinned compressed bytes -> HBM -> inflate -> columns + sums.
    hp = C.c_void_p()
    gpu._check("falx_host_alloc_pinned", lib.falx_host_alloc_pinned(ctx.h, c_n, C.byref(hp)))
    C.memmove(hp, comp, c_n)
    res = [None]

    def e2e():
        gpu._check("falx_upload", lib.falx_upload(ctx.h, d_comp.ptr, hp, c_n))
        inflate()
        res[0] = ctx.parse_columns_with_stats_into(spec, d_out, n_in, cols)
    log("timing lane bgzf_e2e_columns_stats")
    ems, _ = time_steps(e2e, max(3, args.steps // 2), 1, ctx)
    e_lane = {"gib_s": round(n_in / GIB / (ems / 1000.0), 2), "ms_per_step": round(ems, 3),
              "verified": (res[0]["records"] == rows and res[0]["cols"][0]["fixed6_sum"] == truth["lat_fixed6_sum"]
                           and res[0]["cols"][1]["fixed6_sum"] == truth["lon_fixed6_sum"])}
    lib.falx_host_free_pinned(ctx.h, hp)
    if not args.no_cpu:
        payloads = [(comp[int(x["payload"]):int(x["payload"]) + int(x["payload_len"])], int(x["out"]),
                     int(x["isize"])) for x in blocks]
        out = bytearray(n_in)
        mv = memoryview(out)

        def cpu_inflate():
            def one(p):
                mv[p[1]:p[1] + p[2]] = zlib.decompress(p[0], -15)
            with ThreadPoolExecutor(thr) as ex:
                list(ex.map(one, payloads, chunksize=64))
        log("timing CPU lane bgzf_inflate (zlib)")
        lane["cpu_gib_s"] = round(n_in / GIB / cpu_time(cpu_inflate, 3, warmups=1), 3)
        lane["cpu_threads"] = thr
        lane["cpu_kind"] = f"zlib {zlib.ZLIB_RUNTIME_VERSION} raw inflate per block, {thr} threads"
        lane["cpu_equal"] = bytes(out) == data.tobytes()
        run_cols = cpu_columns(spec, np.frombuffer(out, dtype=np.uint8), rows, 16, thr)

        def cpu_e2e():
            cpu_inflate()
            run_cols()
        log("timing CPU lane bgzf_e2e_columns_stats (zlib + oracle parse_columns_par)")
        e_lane["cpu_gib_s"] = round(n_in / GIB / cpu_time(cpu_e2e, 3, warmups=1), 3)
        e_lane["cpu_threads"] = thr
    lanes["bgzf_inflate"] = lane
    lanes["bgzf_e2e_columns_stats"] = e_lane
    d_comp.free()
    d_out.free()


def served_stream(ctx) -> dict:
    """The mixed request stream (geo CSV, csv_hash, TSV, logfmt, NDJSON;
    sizes log-uniform 4 KiB-256 MiB; Poisson arrivals) through continuous
    batching and through one plain call per request, on the same resident
    inputs: throughput and p50/p99 latency per arrival rate."""
    from falx_amd import serve, tune
    log("served stream: building the resident corpus")
    corpus = serve.Corpus(ctx, 256 << 20)
    tuning = tune.load()
    bs = serve.BatchServer(corpus, slots=3, hbm_budget=24 << 30, max_batch=64, tuning=tuning, large=32 << 20,
                           hold_below=8)
    ps = serve.PlainServer(corpus, out_budget=2 << 30)
    warm = serve.request_stream(corpus, 60, rate=0, seed=99)
    bs.run(serve.fresh(warm))
    ps.run(serve.fresh(warm))
    runs = []
    for label, n, rate, hi in [("mix_saturated", 150, 0.0, 256 << 20), ("mix_1000_per_s", 150, 1000.0, 256 << 20),
                               ("mix_250_per_s", 150, 250.0, 256 << 20),
                               ("small_saturated", 2000, 0.0, 1 << 20)]:
        reqs = serve.request_stream(corpus, n, rate=rate, seed=3, hi=hi)
        # One untimed pass of the same stream first (both servers): the
        # batched executors capture the launch sequences this stream's batch
        # mixes need, as a server that has been running would have.
        bs.run(serve.fresh(reqs))
        ps.run(serve.fresh(reqs))
        a = bs.run(serve.fresh(reqs))
        p = ps.run(serve.fresh(reqs))
        keep = ("gib_s", "p50_ms", "p99_ms", "errors")
        row = {"run": label, "requests": n, "rate_per_s": rate, "max_bytes": hi,
               "batched": {k: round(a[k], 3) if isinstance(a[k], float) else a[k] for k in keep + ("batches",)},
               "plain": {k: round(p[k], 3) if isinstance(p[k], float) else p[k] for k in keep},
               "throughput_x": round(a["gib_s"] / p["gib_s"], 2) if p["gib_s"] else None}
        runs.append(row)
        log(f"served {label}: batched {a['gib_s']:.2f} GiB/s p99 {a['p99_ms']:.2f} ms | "
            f"plain {p['gib_s']:.2f} GiB/s p99 {p['p99_ms']:.2f} ms")
    bs.close()
    return {"workload": "served mix of 5 specs, resident inputs, 3 executor slots, requests of 32 MiB or more "
                        "as plain calls beside the batches (small requests wait for them while fewer than 8 are "
                        "queued), results left in HBM",
            "tuning": {k: v for k, v in tuning.items() if k in ("executor", "chunk")} or None,
            "runs": runs}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=GIB)
    ap.add_argument("--cpu-runs", type=int, default=9)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lanes", action="store_true")
    ap.add_argument("--lanes", default="geo,ndjson,csv_hash",
                    help="comma-separated lane groups to run: geo, ndjson, csv_hash")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="only run a plain device copy of this many bytes (rocprof counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {args.gpus} is not supported (this path shards by "
              f"input range; see DESIGN.md)", file=sys.stderr)
        return 2

    import numpy as np

    from falx_amd import datagen, gpu
    from falx_amd import spec as S

    ctx = gpu.Context(0)
    if args.calibrate_copy:
        hip = C.CDLL("libamdhip64.so")
        src, dst = ctx.alloc(args.calibrate_copy), ctx.alloc(args.calibrate_copy)
        for _ in range(3):
            assert hip.hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), C.c_size_t(args.calibrate_copy), 3) == 0
        assert hip.hipDeviceSynchronize() == 0
        log(f"copied {args.calibrate_copy} bytes device-to-device x3")
        return 0

    log(f"generating {args.size} B geo CSV")
    data, truth = datagen.geo_csv(args.size)
    n_in, rows = int(data.size), int(truth["rows"])
    log(f"uploading {n_in} B ({rows} rows)")
    dbuf = ctx.upload(data)
    spec = S.GEO_LATLON
    cols, keep = ctx.alloc_columns(spec, rows, 16)

    def step():
        return ctx.parse_columns_with_stats_into(spec, dbuf, n_in, cols)

    log("timing headline: lat/lon columns + fused sum")
    ms, kms = time_steps(step, args.steps, args.warmup, ctx)
    st = step()
    # verify against the generator's closed forms
    verified = (cols.rows == rows and st["records"] == rows
                and st["cols"][0]["fixed6_sum"] == truth["lat_fixed6_sum"]
                and st["cols"][1]["fixed6_sum"] == truth["lon_fixed6_sum"]
                and st["cols"][0]["checksum"] == truth["lat_checksum"]
                and st["cols"][1]["checksum"] == truth["lon_checksum"])
    lat_gpu = keep[0][0].download(np.float64, rows)
    algo = algorithmic_bytes_latlon(n_in, rows)
    achieved = algo / (kms / 1000.0) / 1e9
    traffic = load_traffic("traffic_latlon_cols_stats.json")
    value = n_in / GIB / (ms / 1000.0)

    lanes = {}
    want = set(args.lanes.split(","))
    if not args.no_lanes:
        thr = cpu_threads(args.cpu_threads)
        if not args.no_cpu:
            from oracle import oracle as O
        if "geo" in want:
            cols3, _k3 = ctx.alloc_columns(S.GEO_CITY_LATLON, rows, n_in)
            lane_specs = [
                ("latlon_columns", lambda: ctx.parse_columns_into(spec, dbuf, n_in, cols), algo),
                ("fused_count_sum", lambda: ctx.csv_stats(spec, dbuf, n_in), n_in),
                ("city_latlon_columns", lambda: ctx.parse_columns_into(S.GEO_CITY_LATLON, dbuf, n_in, cols3),
                 algo + 4 * (rows + 1) + truth["city_bytes"] + 8 * ((rows + 63) // 64)),
            ]
            idx_out = ctx.alloc(4 * (n_in // 8 + 64))
            n_struct = [0]

            def idx():
                _, n_struct[0] = ctx.index_structurals(spec, dbuf, n_in, out=idx_out, download=False)
            lane_specs.append(("index_structurals", idx, None))
            for name, fn, ab in lane_specs:
                log(f"timing lane {name}")
                lms, lkms = time_steps(fn, max(3, args.steps // 2), 1, ctx)
                if name == "index_structurals":
                    ab = n_in + 4 * n_struct[0]
                lanes[name] = {"gib_s": round(n_in / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3),
                               "kernel_ms": round(lkms, 3), "hbm_frac": round(ab / (lkms / 1000.0) / 1e9 / HBM_PEAK_GBS, 3)}
                if name == "index_structurals":
                    # Closed form: three separators and a newline per record.
                    lanes[name]["verified"] = n_struct[0] == 4 * rows
            # end-to-end from pinned host memory: H2D + the headline step
            hp = C.c_void_p()
            gpu._check("falx_host_alloc_pinned", gpu.load().falx_host_alloc_pinned(ctx.h, n_in, C.byref(hp)))
            C.memmove(hp, data.ctypes.data, n_in)

            def e2e():
                gpu._check("falx_upload", gpu.load().falx_upload(ctx.h, dbuf.ptr, hp, n_in))
                step()
            log("timing lane e2e_pinned_h2d")
            ems, _ = time_steps(e2e, max(3, args.steps // 2), 1, ctx)
            lanes["e2e_pinned_h2d"] = {"gib_s": round(n_in / GIB / (ems / 1000.0), 2), "ms_per_step": round(ems, 3)}
            gpu.load().falx_host_free_pinned(ctx.h, hp)
            if not args.no_cpu:
                cpu_lane = {
                    "latlon_columns": cpu_columns(spec, data, rows, 16, thr),
                    "fused_count_sum": lambda: O.csv_stats(spec, data, thr),
                    "city_latlon_columns": cpu_columns(S.GEO_CITY_LATLON, data, rows, n_in, thr),
                    "index_structurals": lambda: O.index_structurals(spec, data, thr),
                }
                for name, fn in cpu_lane.items():
                    log(f"timing CPU oracle lane {name}")
                    lanes[name]["cpu_gib_s"] = round(n_in / GIB / cpu_time(fn, args.cpu_runs), 3)
                    lanes[name]["cpu_threads"] = thr
            del cols3, _k3, idx_out
            stream_lane(ctx, data, dbuf, n_in, rows, args, lanes)
            bgzf_lanes(ctx, data, truth, spec, cols, args, lanes, thr)

        if "ndjson" in want:
            # NDJSON sum(id + nested.score), 1 GiB
            log(f"generating {args.size} B NDJSON")
            nd, nd_truth = datagen.ndjson(args.size)
            nd_n = int(nd.size)
            nd_buf = ctx.upload(nd)
            nd_res = [None]

            def nd_step():
                nd_res[0] = ctx.ndjson_id_score(nd_buf, nd_n)
            log("timing lane ndjson_id_score_sum")
            lms, lkms = time_steps(nd_step, max(3, args.steps // 2), 1, ctx)
            r = nd_res[0]
            lanes["ndjson_id_score_sum"] = {
                "gib_s": round(nd_n / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3),
                "kernel_ms": round(lkms, 3), "hbm_frac": round(nd_n / (lkms / 1000.0) / 1e9 / HBM_PEAK_GBS, 3),
                "bytes": nd_n, "records": r["records"],
                "verified": (r["records"] == nd_truth["records"] and r["id_sum"] == nd_truth["id_sum"]
                             and r["score_fixed6_sum"] == nd_truth["score_fixed6_sum"] and r["invalid"] == 0)}
            if not args.no_cpu:
                log("timing CPU oracle lane ndjson_id_score_sum")
                lanes["ndjson_id_score_sum"]["cpu_gib_s"] = round(
                    nd_n / GIB / cpu_time(lambda: O.ndjson_id_score(nd, thr), args.cpu_runs), 3)
                lanes["ndjson_id_score_sum"]["cpu_threads"] = thr
                # Score sum: the GPU tree and the reference's sequential fold
                # (the oracle at one thread) against the exact decimal sum.
                fold = O.ndjson_id_score(nd, 1)["score_sum"]
                lanes["ndjson_id_score_sum"]["sums"] = sum_errors(
                    "score", r["score_sum"], fold, Fraction(nd_truth["score_fixed6_sum"], 10**6))
            nested_lane(ctx, nd, nd_buf, args, lanes, thr)
            nd_buf.free()
            del nd

        if "csv_hash" in want:
            # csv_hash 1 GiB: index, typed columns, field bytes
            log(f"generating {args.size} B csv_hash")
            hs, hs_truth = datagen.hash_csv(args.size)
            hs_n = int(hs.size)
            hs_buf = ctx.upload(hs)
            hspec = S.HASH_TYPED
            hs_rows = int(hs_truth["records"])
            hcols, hkeep = ctx.alloc_columns(hspec, hs_rows, hs_n)
            hidx = ctx.alloc(4 * (hs_n // 4 + 64))
            hs_struct = [0]

            def hs_index():
                _, hs_struct[0] = ctx.index_structurals(hspec, hs_buf, hs_n, out=hidx, download=False)
            hash_lanes = [
                ("csv_hash_index_structurals", hs_index, None,
                 lambda: O.index_structurals(hspec, hs, thr)),
                ("csv_hash_typed_columns", lambda: ctx.parse_columns_into(hspec, hs_buf, hs_n, hcols), None,
                 (cpu_columns(hspec, hs, hs_rows, hs_n, thr) if not args.no_cpu else None)),
                ("csv_hash_field_bytes", lambda: ctx.field_bytes(hspec, hs_buf, hs_n), hs_n,
                 lambda: O.field_bytes(hspec, hs, thr)),
            ]
            for name, fn, ab, cfn in hash_lanes:
                log(f"timing lane {name}")
                lms, lkms = time_steps(fn, max(3, args.steps // 2), 1, ctx)
                if name == "csv_hash_index_structurals":
                    ab = hs_n + 4 * hs_struct[0]
                elif ab is None:
                    # input + i64 and f64 values + string offsets and arena +
                    # three validity bitmaps
                    hr = int(hcols.rows)
                    ab = (hs_n + 16 * hr + 4 * (hr + 1) + int(hcols.cols[2].data_len) + 3 * 8 * ((hr + 63) // 64))
                lanes[name] = {"gib_s": round(hs_n / GIB / (lms / 1000.0), 2), "ms_per_step": round(lms, 3),
                               "kernel_ms": round(lkms, 3),
                               "hbm_frac": round(ab / (lkms / 1000.0) / 1e9 / HBM_PEAK_GBS, 3), "bytes": hs_n}
                if not args.no_cpu:
                    log(f"timing CPU oracle lane {name}")
                    lanes[name]["cpu_gib_s"] = round(hs_n / GIB / cpu_time(cfn, args.cpu_runs), 3)
                    lanes[name]["cpu_threads"] = thr
            # Closed forms of the generator: the row count and the wrapping sum
            # of the i64 column; the index and field bytes equal the oracle's.
            i64_col = hkeep[0][0].download(np.int64, int(hcols.rows))
            lanes["csv_hash_typed_columns"]["rows"] = int(hcols.rows)
            lanes["csv_hash_typed_columns"]["verified"] = (
                int(hcols.rows) == hs_rows and int(i64_col.view(np.uint64).sum(dtype=np.uint64)) ==
                hs_truth["i64_sum"] % 2**64)
            del i64_col
            if not args.no_cpu:
                lanes["csv_hash_index_structurals"]["verified"] = hs_struct[0] == O.index_structurals(hspec, hs, thr).size
                lanes["csv_hash_field_bytes"]["verified"] = (ctx.field_bytes(hspec, hs_buf, hs_n) ==
                                                             O.field_bytes(hspec, hs, thr))

    served = None
    if not args.no_served:
        served = served_stream(ctx)

    cpu = None
    sums = {}
    if not args.no_cpu:
        threads = cpu_threads(args.cpu_threads)
        log(f"timing CPU oracle baseline at {threads} threads")
        gibs, med, times, (s_lat, s_lon) = cpu_baseline(data, rows, args.cpu_runs, threads)
        # one thread on a 256 MiB prefix of the same input (1 warm-up, 3 runs)
        pre = data[: min(n_in, 256 << 20)]
        from oracle import oracle as O1
        one = pre.size / GIB / cpu_time(cpu_columns(spec, pre, rows, 16, 1), 3, warmups=1)
        cpu = {"value": round(gibs, 3), "unit": "GiB/s", "cores": threads, "kind": "port",
               "sample": f"the same {n_in} B input, 2 warm-ups then {args.cpu_runs} runs of parse_columns_par "
                         f"(lat/lon) + sequential column sums: median {med * 1000:.0f} ms, "
                         f"best {times[0] * 1000:.0f} ms",
               "best_gib_s": round(n_in / GIB / times[0], 3),
               "one_thread_gib_s": round(one, 3),
               "one_thread_sample": f"first {pre.size} B of the same input, parse_columns at 1 thread, "
                                    f"1 warm-up then 3 runs, median",
               "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
               "threads_note": "the box's CPU share for one GPU is 16 threads",
               "isa": O1.isa()}
        sums = {}
        sums.update(sum_errors("lat", st["cols"][0]["sum"], s_lat, Fraction(truth["lat_fixed6_sum"], 10**6)))
        sums.update(sum_errors("lon", st["cols"][1]["sum"], s_lon, Fraction(truth["lon_fixed6_sum"], 10**6)))

    out = {
        "metric": "input GiB/s (1 GiB CSV cols + fused sum), % HBM peak; batched p99 latency",
        "value": round(value, 3),
        "unit": "GiB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic (seeded xorshift64* geo CSV, generator v%d)" % datagen.GENERATOR_VERSION,
        "config": {"workload": "geo CSV id,City,Latitude,Longitude; lat/lon f64 columns + fused "
                               "count/sum (exact int64 mantissas) in one pass",
                   "bytes": n_in, "rows": rows, "seed": hex(datagen.SEED_CSV_GEO)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (traffic["bytes_per_launch"] if traffic else None)},
        "cpu_baseline": cpu,
        "kernel_ms": round(kms, 4),
        "verified": bool(verified) and bool(np.isfinite(lat_gpu).all()),
        "lanes": lanes,
        "served": served,
        "sums": sums,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
