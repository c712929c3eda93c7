"""konglong_amd: MI355X-native RTM propagation-and-imaging hot path.

The compute path is libkonglong_rtm.so (hand-written HIP kernels for gfx950
behind the C ABI in include/konglong_rtm.h); the Python modules are thin
ctypes wrappers that mirror the reference's class names (KernelCPU,
Receiver, WaveFieldCompress, Propagator).
"""
from ._lib import (lib, check, DeviceArray, device_count, synchronize, KlrtmError,  # noqa: F401
                   ENGINE_FD, ENGINE_CFD, ISO, VTI, TTI, SNAP_FP32, SNAP_C16, SNAP_FP16, SNAP_BF16, LIB_PATH)

__all__ = ["lib", "check", "DeviceArray", "device_count", "synchronize", "KlrtmError", "ENGINE_FD", "ENGINE_CFD",
           "ISO", "VTI", "TTI", "SNAP_FP32", "SNAP_C16", "SNAP_FP16", "SNAP_BF16", "LIB_PATH"]
