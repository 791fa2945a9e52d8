"""pe_amd: MI355X-native game value-function solver.

Host-side mirror of the reference entry points for the value-function path:

    pe_amd.reversible.free_dp.solve(config, RelativeGrid(...), out, horizon=...)
    pe_amd.local_game.reference.solve(config, arena, out, ...)

Compute runs in libpe_amd.so (hand-written HIP kernels for gfx950) through
the C ABI declared in include/pe_amd.h.
"""

from ._native import LIB_PATH, NativeLibraryMissing, device_info  # noqa: F401

__all__ = ["LIB_PATH", "NativeLibraryMissing", "device_info"]
