"""MI355X (gfx950) S2 two-point-correlation engine for RustMSPT.

The compute path is libs2hip.so (hand-written HIP kernels behind the C ABI in
include/s2hip.h); the Python modules mirror the reference's S2 entry points,
pipelines and configuration on top of it.
"""
from ._lib import S2HipError  # noqa: F401

__all__ = ["S2HipError"]
