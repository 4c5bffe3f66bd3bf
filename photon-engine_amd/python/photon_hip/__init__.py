"""Python handles over libphoton_hip.so (the MI355X wavefront path tracer) and the seeded workloads."""
from . import capi, configs, document, scene  # noqa: F401
from .device import (Device, GpuFilm, GpuScene, camera_orthographic, camera_perspective, camera_thin_lens,  # noqa: F401
                     camera_with_shift, camera_with_shutter_end, device_count, render, settings)
