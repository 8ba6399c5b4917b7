"""MI355X-native ReSTIR PT / Conditional ReSTIR frame loop (HIP, gfx950).

The compute path is the HIP library ``libcrisoptix_amd.so`` (C ABI in include/crisoptix.h),
loaded on first use by :func:`crisoptix_amd.runtime.lib`.  Loading fails loudly when the
library has not been built; there is no CPU fallback.
"""
from . import abi, scenes  # noqa: F401

__all__ = ["abi", "scenes"]
