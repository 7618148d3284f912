"""voreen_amd -- MI355X-native direct-volume-rendering ray march for Voreen.

The product is libhipdvr.so (HIP kernels for gfx950 behind the C-ABI of
include/hipdvr.h) and the HIPRaycaster Voreen module under
voreen_amd/hipraycaster.  This package adds a ctypes binding (hipdvr) and
seeded workload descriptions (workloads).
"""
from . import hipdvr, workloads  # noqa: F401
