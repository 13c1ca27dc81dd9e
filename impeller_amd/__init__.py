"""impeller_amd -- MI355X-native (gfx950) BN254 NTT / MSM / fused-gate path.

The compute lives in libimpeller_hip.so (hand-written HIP, C ABI in
include/impeller_hip.h).  This package is the host-side mirror of the
reference's operator interface (core/src/ntt.rs, core/src/msm.rs,
core/src/fused_kernels.rs) used by tests and benchmarks.
"""
from . import _lib  # noqa: F401

__all__ = ["ntt", "msm", "fused", "runtime"]


def version():
    return _lib.call("impeller_version").decode()
