"""MI355X-native (gfx950 HIP) time-stepping hot path of NS3DPeriodic.

The compute lives in libns3d_hip.so (C ABI: include/ns3d.h); `NS3D` is the
host-side mirror of the reference operators over that ABI.
"""
from .solver import NS3D, StepGraph  # noqa: F401

__all__ = ["NS3D", "StepGraph"]
