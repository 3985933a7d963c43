"""MI355X (gfx950) backend for the OpenKSpace ReconStrategy hot path.

The compute path is liboks_hip.so (hand-written HIP kernels behind the C ABI
in include/oks.h); this package holds the ctypes binding, the host-side
mirror of the reference strategy API and the seeded synthetic generator.
"""
from . import abi  # noqa: F401
