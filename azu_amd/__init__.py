"""MI355X-native KinectFusion frame loop (integrate + raycast + ICP) for gfx950.

The compute path is libazu_hip.so (hand-written HIP kernels behind the C ABI of
include/azu_hip.h); this package is its Python host mirror: ``native`` (ctypes
binding), ``fusion`` (TSDFVolume / ICPTracker shaped like the reference's C++
classes), ``scene`` (synthetic depth input) and ``pipeline`` (the per-frame loop).
"""
from . import native  # noqa: F401  (loads libazu_hip.so; raises if it is missing)
