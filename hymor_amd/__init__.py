"""MI355X (gfx950) engine for HYMOR's exp(A*T)*v linear-stability hot path.

The compute path is libhymor_hip.so (hand-written HIP kernels behind a C ABI,
include/hymor.h); this package holds its ctypes binding (``_lib``, ``core``),
the host-side mirror of the reference's Julia stability entry points
(``stability``) and config-shaped synthetic operators (``operators``).
"""
from ._lib import HymorError, device_available  # noqa: F401
