"""MI355X-native Numerov-shooting backend with epseon_backend's device API.

* ``epseon_backend_amd.device.gpu._libepseon_gpu`` -- the reference's GPU
  module surface (context, device interface, task configurator, task handle),
  backed by hand-written gfx950 HIP kernels;
* ``epseon_backend_amd.device.cpu._libepseon_cpu`` -- the reference's CPU
  module surface (``greet``);
* ``epseon_backend_amd.capi`` -- ctypes binding of the C ABI
  (``include/epseon_hip.h``) for direct, batch-level use.
"""
from __future__ import annotations

__version__ = "0.1.0"
