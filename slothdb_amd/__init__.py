"""MI355X-native GROUP BY hash-aggregate operator for SlothDB's hot path.

The operator lives in lib/libslothdb_gpu.so (HIP kernels + C ABI,
include/slothdb_gpu.h); `gbagg` binds it, `datagen` produces seeded inputs.
"""
from . import gbagg  # noqa: F401
