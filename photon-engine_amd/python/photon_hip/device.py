"""Thin Python handles over the C-ABI (device, scene, film), used by tests and bench.py."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, load, ptr


def device_count() -> int:
    n = C.c_int()
    check("pe_device_count", load().pe_device_count(C.byref(n)))
    return n.value


def camera_perspective(position, target, up, vfov, aspect) -> capi.Camera:
    cam = capi.Camera()
    f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
    check("pe_camera_perspective", load().pe_camera_perspective(f3(position), f3(target), f3(up), float(vfov),
                                                                  float(aspect), C.byref(cam)))
    return cam


def camera_thin_lens(position, target, up, vfov, aspect, aperture, focus, blades=0, rotation_deg=0.0) -> capi.Camera:
    """ThinLensCamera(position, target, up, vfov, aspect, aperture, focus) with setApertureShape(blades, rotation)."""
    cam = capi.Camera()
    f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
    check("pe_camera_thin_lens", load().pe_camera_thin_lens(f3(position), f3(target), f3(up), float(vfov),
                                                              float(aspect), float(aperture), float(focus), int(blades),
                                                              float(rotation_deg), C.byref(cam)))
    return cam


def camera_orthographic(position, target, up, vertical_extent, aspect) -> capi.Camera:
    """OrthographicCamera(position, target, up, verticalExtent, aspect)."""
    cam = capi.Camera()
    f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
    check("pe_camera_orthographic", load().pe_camera_orthographic(f3(position), f3(target), f3(up),
                                                                    float(vertical_extent), float(aspect), C.byref(cam)))
    return cam


def camera_with_shift(camera: capi.Camera, shift_x: float, shift_y: float) -> capi.Camera:
    """A copy of a perspective or thin-lens `camera` with a lens shift (setShift)."""
    out = capi.Camera.from_buffer_copy(camera)
    check("pe_camera_set_shift", load().pe_camera_set_shift(C.byref(out), float(shift_x), float(shift_y)))
    return out


def camera_with_shutter_end(camera: capi.Camera, end: capi.Camera) -> capi.Camera:
    """A copy of `camera` with `end`'s pose at shutter close (Camera::setShutterEnd)."""
    out = capi.Camera.from_buffer_copy(camera)
    check("pe_camera_set_shutter_end", load().pe_camera_set_shutter_end(C.byref(out), C.byref(end)))
    return out


def settings(max_bounces=8, rr_depth=3, shadow_quality=1, filter_type=capi.PE_FILTER_BOX, filter_width=1.0,
             seed=12345, samples_per_launch=1, clamp_indirect=0.0, adaptive=False, noise_threshold=0.02,
             min_adaptive_samples=16, paths_in_flight=0, motion_blur=False, ao_strength=0.0) -> capi.Settings:
    """RenderSettings as pe_settings (the reference's defaults for the adaptive fields, render_settings.h:30-32)."""
    s = capi.Settings()
    s.motion_blur = int(bool(motion_blur))
    s.ao_strength = float(ao_strength)
    s.max_bounces, s.rr_depth, s.shadow_quality = max_bounces, rr_depth, shadow_quality
    s.filter_type, s.filter_width, s.clamp_indirect = filter_type, filter_width, clamp_indirect
    s.seed, s.samples_per_launch, s.paths_in_flight = seed, samples_per_launch, paths_in_flight
    s.adaptive, s.noise_threshold, s.min_adaptive_samples = int(bool(adaptive)), noise_threshold, min_adaptive_samples
    return s


class Device:
    def __init__(self, ordinal: int = 0):
        self._lib = load()
        h = C.c_void_p()
        check("pe_device_create", self._lib.pe_device_create(ordinal, C.byref(h)))
        self.h = h

    def close(self):
        """Destroys the device. Scenes and films hold a reference to it, so only call this once they are gone."""
        if getattr(self, "h", None):
            self._lib.pe_device_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self, ordinal: int = 0) -> capi.DeviceInfo:
        info = capi.DeviceInfo()
        check("pe_device_info_get", self._lib.pe_device_info_get(ordinal, C.byref(info)))
        return info

    def synchronize(self):
        check("pe_device_synchronize", self._lib.pe_device_synchronize(self.h))

    def sobol(self, seed, px, py, index, dim) -> np.ndarray:
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        index = np.ascontiguousarray(index, np.uint32)
        dim = np.ascontiguousarray(dim, np.uint32)
        out = np.zeros((px.shape[0], 2), np.uint32)
        check("pe_sobol_batch", self._lib.pe_sobol_batch(self.h, seed, px.shape[0], ptr(px, C.c_int32),
                                                         ptr(py, C.c_int32), ptr(index, C.c_uint32),
                                                         ptr(dim, C.c_uint32), ptr(out, C.c_uint32)))
        return out

    def trace_diag(self) -> tuple[int, int]:
        """(deepest traversal stack, stack overflows) of the device's last closest-hit / any-hit batch."""
        depth, over = C.c_uint32(), C.c_uint64()
        check("pe_trace_diag", self._lib.pe_trace_diag(self.h, C.byref(depth), C.byref(over)))
        return depth.value, over.value

    def upload(self, scene) -> "GpuScene":
        return GpuScene(self, scene)

    def film(self, width: int, height: int) -> "GpuFilm":
        return GpuFilm(self, width, height)


class GpuScene:
    def __init__(self, dev: Device, scene):
        self.dev = dev
        self._lib = dev._lib
        desc, keep = scene.pack()
        h = C.c_void_p()
        check("pe_scene_upload", self._lib.pe_scene_upload(dev.h, C.byref(desc), C.byref(h)))
        del keep
        self.h = h

    def update(self, scene, kind: int) -> int:
        """pe_scene_update with the full updated scene; returns the status code (no exception)."""
        desc, keep = scene.pack()
        rc = self._lib.pe_scene_update(self.h, int(kind), C.byref(desc))
        del keep
        return rc

    def close(self):
        # A scene outliving its device (interpreter shutdown destroys objects in any order) is left to the
        # process's end: its handle died with the device.
        if getattr(self, "h", None) and getattr(self.dev, "h", None):
            self._lib.pe_scene_release(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def stats(self):
        t, n, b = C.c_uint64(), C.c_uint64(), C.c_uint64()
        check("pe_scene_stats", self._lib.pe_scene_stats(self.h, C.byref(t), C.byref(n), C.byref(b)))
        return {"triangles": t.value, "bvh_nodes": n.value, "device_bytes": b.value}

    def intersect(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.int32)
        bary = np.zeros((n, 2), np.float32)
        check("pe_intersect_batch", self._lib.pe_intersect_batch(self.dev.h, self.h, n, ptr(rays, C.c_float),
                                                                 ptr(hits, C.c_int32), ptr(bary, C.c_float)))
        return hits, bary

    def occluded(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(rays.shape[0], np.uint8)
        check("pe_occluded_batch", self._lib.pe_occluded_batch(self.dev.h, self.h, rays.shape[0],
                                                               ptr(rays, C.c_float), ptr(occ, C.c_uint8)))
        return occ


def display_settings(**overrides) -> capi.Display:
    """pe_display with the reference's defaults (pe_display_defaults), fields overridden by name."""
    d = capi.Display()
    check("pe_display_defaults", load().pe_display_defaults(C.byref(d)))
    for k, v in overrides.items():
        setattr(d, k, v)
    return d


class GpuFilm:
    def __init__(self, dev: Device, width: int, height: int):
        self.dev = dev
        self._lib = dev._lib
        self.width, self.height = width, height
        h = C.c_void_p()
        check("pe_film_create", self._lib.pe_film_create(dev.h, width, height, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and getattr(self.dev, "h", None):  # as GpuScene.close
            self._lib.pe_film_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def clear(self):
        check("pe_film_clear", self._lib.pe_film_clear(self.h))

    def download(self):
        """Returns (sum_rgb (H, W, 3), count (H, W), sum_lum2 (H, W), rejected)."""
        s = np.zeros((self.height, self.width, 3), np.float32)
        c = np.zeros((self.height, self.width), np.uint32)
        l2 = np.zeros((self.height, self.width), np.float32)
        rej = C.c_uint64()
        check("pe_film_download", self._lib.pe_film_download(self.h, ptr(s, C.c_float), ptr(c, C.c_uint32),
                                                             ptr(l2, C.c_float), C.byref(rej)))
        return s, c, l2, rej.value

    def display(self, settings: capi.Display | None = None):
        """pe_film_display -> (display (H, W, 3) float32 in [0, 1], rgb8 (H, W, 3) uint8)."""
        d = settings if settings is not None else display_settings()
        out = np.zeros((self.height, self.width, 3), np.float32)
        rgb8 = np.zeros((self.height, self.width, 3), np.uint8)
        check("pe_film_display", self._lib.pe_film_display(self.h, C.byref(d), ptr(out, C.c_float), ptr(rgb8, C.c_uint8)))
        return out, rgb8

    def mean(self) -> np.ndarray:
        s, c, _, _ = self.download()
        return s / np.maximum(c, 1)[..., None]


class CryptoFilm:
    """pe_crypto_film: per pixel up to PE_CRYPTO_SLOTS (id, count) entries in first-seen order and a total."""

    def __init__(self, dev: Device, width: int, height: int):
        self.dev, self.width, self.height = dev, width, height
        self._lib = load()
        h = C.c_void_p()
        check("pe_crypto_film_create", self._lib.pe_crypto_film_create(dev.h, width, height, C.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None) and getattr(self.dev, "h", None):
            self._lib.pe_crypto_film_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()

    def clear(self):
        check("pe_crypto_film_clear", self._lib.pe_crypto_film_clear(self.h))

    def download(self):
        """Returns (ids (H, W, S), counts (H, W, S), totals (H, W), dropped)."""
        S = capi.PE_CRYPTO_SLOTS
        ids = np.zeros((self.height, self.width, S), np.float32)
        counts = np.zeros((self.height, self.width, S), np.uint32)
        totals = np.zeros((self.height, self.width), np.uint32)
        dropped = C.c_uint64()
        check("pe_crypto_film_download", self._lib.pe_crypto_film_download(
            self.h, ptr(ids, C.c_float), ptr(counts, C.c_uint32), ptr(totals, C.c_uint32), C.byref(dropped)))
        return ids, counts, totals, dropped.value

    def pixel_ranks(self, x: int, y: int, data=None):
        """CryptoFilm::pixelRanks (cryptomatte.cpp:195-205): (id, coverage) by decreasing count."""
        ids, counts, totals, _ = data if data is not None else self.download()
        c, t = counts[y, x], totals[y, x]
        if t == 0:
            return []
        order = sorted((j for j in range(len(c)) if c[j] > 0), key=lambda j: -int(c[j]))
        inv = np.float32(1.0) / np.float32(t)
        return [(float(ids[y, x, j]), float(np.float32(c[j]) * inv)) for j in order]


class PassFilms:
    """RenderPassFilms without cryptomatte: a film per light pass (LightPass order), per light group of the
    scene (up to 16) and per data pass, each of the beauty film's size."""

    def __init__(self, dev: Device, width: int, height: int, groups: int = 0, light=True, data=True,
                 crypto_tables=None):
        mk = lambda: GpuFilm(dev, width, height)  # noqa: E731
        self.light = [mk() for _ in range(capi.PE_LIGHT_PASSES)] if light else []
        self.groups = [mk() for _ in range(min(groups, capi.PE_MAX_LIGHT_GROUPS))]
        self.depth, self.position, self.object_id, self.material_id = (mk() for _ in range(4)) if data else (None,) * 4
        # Cryptomatte: (object ids per mesh, material ids per material), as Scene.crypto_tables gives them.
        self.crypto_object = self.crypto_material = None
        self._tables = None
        if crypto_tables is not None:
            self.crypto_object = CryptoFilm(dev, width, height)
            self.crypto_material = CryptoFilm(dev, width, height)
            self._tables = tuple(np.ascontiguousarray(t, np.float32) for t in crypto_tables[:2])

    def targets(self) -> capi.PassTargets:
        pt = capi.PassTargets()
        if self.crypto_object is not None:
            pt.crypto_object, pt.crypto_material = self.crypto_object.h, self.crypto_material.h
            pt.object_ids, pt.material_ids = (ptr(t, C.c_float) for t in self._tables)
            pt.num_object_ids, pt.num_material_ids = (int(t.size) for t in self._tables)
        for k, f in enumerate(self.light):
            pt.light[k] = f.h
        for k, f in enumerate(self.groups):
            pt.light_groups[k] = f.h
        pt.depth, pt.position, pt.object_id, pt.material_id = (
            f.h if f is not None else None for f in (self.depth, self.position, self.object_id, self.material_id))
        return pt

    def means(self) -> dict:
        """Film::resolve of every film: {"light": (8, H, W, 3), "groups": (G, H, W, 3), "depth", ...}."""
        out = {"light": np.stack([f.mean() for f in self.light]) if self.light else None,
               "groups": np.stack([f.mean() for f in self.groups]) if self.groups else None}
        for name in ("depth", "position", "object_id", "material_id"):
            f = getattr(self, name)
            out[name] = f.mean() if f is not None else None
        return out


def render_targets(dev: Device, scene: GpuScene, camera: capi.Camera, st: capi.Settings, film: GpuFilm, first: int,
                   count: int, albedo: GpuFilm | None = None, normal: GpuFilm | None = None,
                   alpha: GpuFilm | None = None, cancel=None, passes: PassFilms | None = None) -> capi.BatchResult:
    """pe_render_targets: the beauty film plus the AOV and pass films given (RenderTargets with AovFilms and
    RenderPassFilms)."""
    res = capi.BatchResult()
    t = capi.Targets()
    t.beauty = film.h
    t.albedo, t.normal, t.alpha = (f.h if f is not None else None for f in (albedo, normal, alpha))
    if passes is not None:
        pt = passes.targets()
        t.passes = C.pointer(pt)
    cptr = ptr(cancel, C.c_int32) if cancel is not None else C.cast(None, C.POINTER(C.c_int32))
    rc = load().pe_render_targets(dev.h, scene.h, C.byref(camera), C.byref(st), C.byref(t), first, count, cptr,
                                  C.byref(res))
    if rc not in (capi.PE_OK, capi.PE_ERR_CANCELLED):
        check("pe_render_targets", rc)
    return res


def render(dev: Device, scene: GpuScene, camera: capi.Camera, st: capi.Settings, film: GpuFilm, first: int,
           count: int, cancel=None) -> capi.BatchResult:
    res = capi.BatchResult()
    cptr = ptr(cancel, C.c_int32) if cancel is not None else C.cast(None, C.POINTER(C.c_int32))
    rc = load().pe_render(dev.h, scene.h, C.byref(camera), C.byref(st), film.h, first, count, cptr, C.byref(res))
    if rc not in (capi.PE_OK, capi.PE_ERR_CANCELLED):
        check("pe_render", rc)
    return res
