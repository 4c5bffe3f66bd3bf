"""Python handles over libphoton_hip.so (the MI355X wavefront path tracer) and the seeded workloads."""
from . import capi, configs, document, scene  # noqa: F401
from .device import Device, GpuFilm, GpuScene, camera_perspective, device_count, render, settings  # noqa: F401
