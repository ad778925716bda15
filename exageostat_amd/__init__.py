"""MI355X-native exact-dense Gaussian likelihood path of ExaGeoStat.

The compute lives in libexageostat_gpu.so (HIP kernels for gfx950 behind
the reference's C entry points, include/exageostat_gpu.h); this package
loads it with ctypes and mirrors the driver-level interface in ``mle``.
"""
from ._lib import load, LIB_PATH  # noqa: F401
from .mle import ExactMLE, DeviceMatrix, generate_xy_loc, generate_xy_loc_st, device_count  # noqa: F401
