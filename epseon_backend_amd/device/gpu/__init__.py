"""Subpackage for interacting with GPU compute capabilities.

The native module ``_libepseon_gpu`` links the HIP runtime.  torch ships a
HIP runtime under the same SONAME; importing torch first (when installed)
makes the process map a single runtime whichever of the two is used later.
"""
from __future__ import annotations

import os

if not os.environ.get("EPSEON_NO_TORCH_PRELOAD"):
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
