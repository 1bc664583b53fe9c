"""exagoop_amd — MI355X-native explicit MPM step (ExaGOOP hot path).

The compute path is the HIP library libexagoop_amd.so behind the C ABI in
include/exagoop_amd.h; this package is its Python binding (capi), scene
builders for the benchmark and test workloads (scenes), and the build recipe
(_build).
"""
from . import capi, scenes  # noqa: F401
from .capi import Context, Specs, default_specs, MPMError  # noqa: F401
