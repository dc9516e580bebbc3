"""swspec_amd: MI355X-native compute core for the swspec software spectrometer.

TaskCoreHIP mirrors the reference's TaskCore plugin interface
(src/TaskCore.h:36-89); the hot path runs in hand-written HIP kernels for
gfx950 behind the C ABI of include/swspec_hip.h.
"""
from .core import (Buffer, Settings, TaskCoreHIP, autotune, autotune_clear, derive_settings, plan_info,  # noqa: F401
                   unpack_device)

__all__ = ["Buffer", "Settings", "TaskCoreHIP", "autotune", "autotune_clear", "derive_settings", "plan_info",
           "unpack_device"]
