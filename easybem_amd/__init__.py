"""easybem-hip: MI355X-native matrix-free layer-potential apply.

The product path is the HIP library libeasybem_hip.so (C-ABI in
include/easybem_hip.h), reached from Python through ``easybem_amd._lib``
(ctypes) or the drop-in pybind11 module ``easybem``.
"""
from . import meshes  # noqa: F401

__all__ = ["meshes"]
