"""opendba_amd: MI355X-native DTW / DBA hot path of OpenDBA.

The compute path is libodba_hip.so (hand-written HIP kernels for gfx950)
behind the C ABI declared in include/odba/odba.h. This package holds the
Python binding over that ABI (`native`) and the synthetic workload
generators used by the benchmark and tests (`workloads`).
"""
from . import native  # noqa: F401

__all__ = ["native"]
