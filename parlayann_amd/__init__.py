"""parlayann_amd -- MI355X-native batched graph beam search (ParlayANN's query
hot path) behind the reference's C++/Python search surface.

The device path lives in libpann.so (HIP kernels for gfx950 + a C ABI,
include/pann.h); this package is its Python front end.
"""
from .index import (DeviceIndex, GraphIndex, load_index, device_info, device_alloc_count,
                    FloatEuclidianIndex,
                    FloatMipsIndex, UInt8EuclidianIndex, UInt8MipsIndex, Int8EuclidianIndex,
                    Int8MipsIndex)

__all__ = ["DeviceIndex", "GraphIndex", "load_index", "device_info", "device_alloc_count",
           "FloatEuclidianIndex",
           "FloatMipsIndex", "UInt8EuclidianIndex", "UInt8MipsIndex", "Int8EuclidianIndex",
           "Int8MipsIndex"]
