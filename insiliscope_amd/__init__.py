"""inSiliScope's image-formation path on AMD MI355X (gfx950): hand-written HIP kernels behind a C ABI
(include/iscg.h, lib/libiscg.so), with a host-side mirror of the reference renderer's GPU seams (host.py).

The kernels are the only implementation: there is no CPU fallback, and every entry point fails loudly when
libiscg.so is missing or no gfx950 device is present.
"""
from ._lib import IscgError, library_path  # noqa: F401

__all__ = ["IscgError", "library_path"]
