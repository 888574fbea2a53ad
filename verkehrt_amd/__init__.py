"""verkehrt_amd: an MI355X-native (gfx950 HIP) kernel variant for verkehrt's whole-array
primitives, behind a C ABI (include/vk_hip.h) that mirrors verkehrt's kernel-table entries.

``Registry`` / ``Dispatch`` mirror the reference's kernel-table API over that ABI; arrays
stay resident in HBM between calls.
"""
from .kernels import (Array, BatchPlan, Builder, Context, Dispatch, Input, KernelError, Reducer,
                      Registry, Scalar, Scanner, Stmt, TYPES)

__all__ = ["Array", "BatchPlan", "Builder", "Context", "Dispatch", "Input", "KernelError", "Reducer",
           "Registry", "Scalar", "Scanner", "Stmt", "TYPES"]
