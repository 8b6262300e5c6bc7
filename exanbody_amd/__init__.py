"""exanbody_amd -- MI355X (gfx950) short-range particle hot path of exaNBody.

The compute path is libxnb_amd.so (hand-written HIP kernels behind a C ABI, declared in
include/xnb_amd.h).  This package holds the ctypes binding and a thin Python handle per box.
"""
from ._lib import (PRECISION_FP64, PRECISION_MIXED, XnbError, load)  # noqa: F401
from .batch import Batch  # noqa: F401
from .box import Box, cubic_domain, device_count, generate_fcc  # noqa: F401

__all__ = ["Batch", "Box", "cubic_domain", "device_count", "generate_fcc", "load", "XnbError", "PRECISION_FP64",
           "PRECISION_MIXED"]
