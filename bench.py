#!/usr/bin/env python3
"""Headline benchmark: 1 GiB geo CSV (`id,City,Latitude,Longitude`), lat/lon
f64 column materialisation plus the fused count/sum of the same columns in
one pass over the HBM-resident input, on one MI355X.

A step is one `falx_parse_columns_with_stats` call over the whole 1 GiB
(inputs already in HBM, outputs preallocated). Prints one JSON line with the
input GiB/s, the roofline of the scan kernel (HIP events on its stream) and
the CPU oracle's parallel baseline timed on this host. Extra lanes (City +
lat/lon columns, fused sum alone, structural index, pinned-H2D end-to-end)
are reported under "lanes".

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


def cpu_baseline(data, rows: int, runs: int, threads: int):
    """The oracle's restatement of parse_columns_par (csv_geo.rs:1167-1284)
    followed by the sequential sums of the two columns (bench_columns.rs:
    151-179), at `threads` host threads, on the same input."""
    import numpy as np

    from falx_amd import spec as S
    from oracle import oracle as O

    spec = S.GEO_LATLON
    lib = O.lib()
    times = []
    lat = lon = None
    for _ in range(runs):
        cols, keep = O.alloc_columns(spec, rows, 16)
        cs = spec.to_c()
        t0 = time.perf_counter()
        st = lib.ora_parse_columns(C.byref(cs), C.c_void_p(data.ctypes.data), data.size, threads, C.byref(cols))
        lat, lon = keep[0][0][:rows], keep[1][0][:rows]
        s_lat, s_lon = float(np.cumsum(lat)[-1]), float(np.cumsum(lon)[-1])
        times.append(time.perf_counter() - t0)
        assert st == 0 and cols.rows == rows
    times.sort()
    med = times[len(times) // 2]
    return data.size / GIB / med, med, times, (s_lat, s_lon)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--size", type=int, default=GIB)
    ap.add_argument("--cpu-runs", type=int, default=3)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-lanes", action="store_true")
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
    if not args.no_lanes:
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

    cpu = None
    sums = {}
    if not args.no_cpu:
        threads = args.cpu_threads or min(os.cpu_count() or 1, 16)
        log(f"timing CPU oracle baseline at {threads} threads")
        gibs, med, times, (s_lat, s_lon) = cpu_baseline(data, rows, args.cpu_runs, threads)
        cpu = {"value": round(gibs, 3), "unit": "GiB/s", "cores": threads, "kind": "port",
               "sample": f"the same {n_in} B input, {args.cpu_runs} runs of parse_columns_par "
                         f"(lat/lon) + sequential column sums, median {med * 1000:.0f} ms"}
        exact_lat = Fraction(truth["lat_fixed6_sum"], 10**6)
        sums = {
            "lat_sum_gpu_tree": st["cols"][0]["sum"], "lat_sum_ref_fold": s_lat,
            "lat_exact": float(exact_lat),
            "lat_err_gpu_vs_exact": float(abs(Fraction(st["cols"][0]["sum"]) - exact_lat)),
            "lat_err_fold_vs_exact": float(abs(Fraction(s_lat) - exact_lat)),
            "lat_err_gpu_vs_fold": abs(st["cols"][0]["sum"] - s_lat),
        }

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
        "sums": sums,
    }
    print(json.dumps(out), flush=True)
    ctx.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
