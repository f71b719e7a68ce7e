"""Build the native pieces in-tree: libfalx_gpu.so (hipcc, gfx950) and the
dataset generator libfalx_datagen.so (g++). Outputs land in falx_amd/_build/.
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "_build")
REPO = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# The host translation unit (C ABI, batches, small kernels) and the tile
# kernels of each dialect kind (kinds.hip, compiled once per kind), built in
# parallel and linked into one library.
HOST_TU = "csrc/falx_gpu.hip"
KIND_TU = "csrc/kinds.hip"
N_KINDS = 6
GPU_DEPS = ["csrc/falx_gpu.hip", "csrc/kinds.hip", "csrc/scan.hip", "csrc/batch.hip", "csrc/cells.h", "csrc/ndjson_walk.h",
            "csrc/num_swar.h", "csrc/regions.h", "csrc/pow5_table.h", "csrc/swar_tile.h"]
# Packed-FP32 instructions stay off in every device compile.
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def gpu_lib_path() -> str:
    return os.path.join(BUILD, "libfalx_gpu.so")


def datagen_lib_path() -> str:
    return os.path.join(BUILD, "libfalx_datagen.so")


def _run(cmd):
    r = subprocess.run(cmd, cwd=HERE, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"build failed: {' '.join(cmd)}")
    return r


def build_gpu() -> str:
    os.makedirs(BUILD, exist_ok=True)
    out = gpu_lib_path()
    inc = ["-I", os.path.join(REPO, "include")]
    jobs = [(os.path.join(BUILD, "falx_gpu_host.o"), [os.path.join(HERE, HOST_TU)])]
    for k in range(N_KINDS):
        jobs.append((os.path.join(BUILD, f"falx_gpu_kind{k}.o"), [f"-DFX_KIND={k}", os.path.join(HERE, KIND_TU)]))
    procs = [(o, subprocess.Popen([HIPCC, *HIP_FLAGS, *inc, "-c", "-o", o, *args], cwd=HERE,
                                  stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True))
             for o, args in jobs]
    failed = []
    for o, p in procs:
        text, _ = p.communicate()
        if p.returncode != 0:
            failed.append(o)
            sys.stderr.write(text)
    if failed:
        raise RuntimeError(f"build failed: {', '.join(failed)}")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *[o for o, _ in jobs]])
    os.replace(tmp, out)
    return out


def build_datagen() -> str:
    os.makedirs(BUILD, exist_ok=True)
    out = datagen_lib_path()
    tmp = out + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", tmp,
          os.path.join(HERE, "datagen", "datagen.cpp")])
    os.replace(tmp, out)
    return out


def build_all() -> None:
    build_datagen()
    build_gpu()


if __name__ == "__main__":
    build_all()
    print("built", gpu_lib_path(), datagen_lib_path())
