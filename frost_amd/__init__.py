"""frost_amd: MI355X-native per-gulp single-pulse search path (RFI -> FDMT ->
matched filter -> candidates) behind frost's prepare / dedisp / search
interfaces. The compute is hand-written HIP for gfx950 in frost_amd/csrc,
exported through the C ABI in include/frost_amd.h; this package is the thin
Python mirror used by tests and the benchmark.
"""
from . import abi  # noqa: F401

__all__ = ["abi"]
