"""falx_amd: MI355X (gfx950) backend for falx's structural-indexing and
column-parsing path. The compute lives in libfalx_gpu.so (HIP kernels behind
the C ABI of include/falx_gpu.h); this package is its host-side mirror.
"""
from . import spec  # noqa: F401
