"""ctypes mirror of include/photon_hip.h and the loader of libphoton_hip.so.

The library is loaded from the package directory (photon-engine_amd/libphoton_hip.so). There is no
fallback: if the library is missing or a call fails, a PhotonHipError is raised with the C-ABI's message.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parents[2]
LIB_PATH = Path(os.environ.get("PHOTON_HIP_LIB", PKG_DIR / "libphoton_hip.so"))

PE_OK = 0
PE_LINK_ALL, PE_LINK_INCLUDE, PE_LINK_EXCLUDE = 0, 1, 2  # light-link modes (pe_scene_desc.link_modes)
PE_ERR_INVALID = -1
PE_ERR_NO_DEVICE = -2
PE_ERR_HIP = -3
PE_ERR_OOM = -4
PE_ERR_UNSUPPORTED = -5
PE_ERR_CANCELLED = -6
PE_ERR_TIMEOUT = -7
PE_JOB_QUEUED, PE_JOB_ADMITTED, PE_JOB_DONE, PE_JOB_CANCELLED, PE_JOB_FAILED = 0, 1, 2, 3, 4

PE_MAT_DISNEY, PE_MAT_LAMBERT, PE_MAT_DIELECTRIC, PE_MAT_MIRROR = 0, 1, 2, 3
PE_LIGHT_AREA, PE_LIGHT_TRIANGLE = 1, 2
PE_SHAPE_CAMERA_INVISIBLE = 1
PE_FILTER_BOX, PE_FILTER_TRIANGLE, PE_FILTER_GAUSSIAN, PE_FILTER_BLACKMAN_HARRIS = 0, 1, 2, 3
PE_UPDATE_FULL, PE_UPDATE_MATERIALS, PE_UPDATE_LIGHTS, PE_UPDATE_TRANSFORMS = 0, 1, 2, 3

F3 = C.c_float * 3


class DeviceInfo(C.Structure):
    _fields_ = [("id", C.c_char * 32), ("name", C.c_char * 128), ("arch", C.c_char * 32), ("kind", C.c_int32),
                ("compute_units", C.c_int32), ("memory_bytes", C.c_uint64), ("usable", C.c_int32),
                ("note", C.c_char * 128)]


class DeviceFeatures(C.Structure):
    _fields_ = [("aovs", C.c_int32), ("light_passes", C.c_int32), ("adaptive_sampling", C.c_int32)]


class MaterialDesc(C.Structure):
    _fields_ = [
        ("type", C.c_int32), ("base_color", F3), ("emission", F3),
        ("metallic", C.c_float), ("roughness", C.c_float), ("specular", C.c_float),
        ("clearcoat", C.c_float), ("clearcoat_roughness", C.c_float), ("coat_ior", C.c_float),
        ("coat_tint", F3),
        ("ior", C.c_float), ("specular_transmission", C.c_float), ("anisotropy", C.c_float),
        ("sheen", C.c_float), ("sheen_roughness", C.c_float), ("sheen_color", F3), ("has_sheen_color", C.c_int32),
        ("diffuse_roughness", C.c_float), ("diffuse_transmission", C.c_float),
        ("subsurface", C.c_float), ("subsurface_radius", F3),
        ("subsurface_scale", C.c_float), ("subsurface_anisotropy", C.c_float),
        ("edge_tint", F3), ("thin_film_thickness", C.c_float), ("thin_film_ior", C.c_float),
        ("abbe", C.c_float), ("absorption_color", F3), ("absorption_distance", C.c_float),
        ("thin_walled", C.c_int32), ("opacity", C.c_float), ("has_opacity", C.c_int32),
        ("scatter_density", C.c_float), ("scatter_color", C.c_float * 3), ("scatter_anisotropy", C.c_float),
        ("flake_density", C.c_float), ("flake_size", C.c_float), ("flake_roughness", C.c_float),
        ("flake_color", F3), ("reserved", C.c_float * 3),
    ]


class MeshDesc(C.Structure):
    _fields_ = [("positions", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
                ("uvs", C.POINTER(C.c_float)), ("indices", C.POINTER(C.c_uint32)),
                ("num_vertices", C.c_uint32), ("num_triangles", C.c_uint32), ("material", C.c_int32),
                ("emitter_base", C.c_int32), ("standalone", C.c_int32), ("flags", C.c_uint32)]


class LightDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("p0", F3), ("p1", F3), ("p2", F3), ("radiance", F3),
                ("visible_to_camera", C.c_int32)]


class EnvDesc(C.Structure):
    _fields_ = [("rgb", C.POINTER(C.c_float)), ("width", C.c_int32), ("height", C.c_int32),
                ("rotation", C.c_float), ("intensity", C.c_float), ("has_sun", C.c_int32),
                ("sun_direction", F3), ("sun_cos_max", C.c_float), ("sun_radiance", F3)]


class SceneDesc(C.Structure):
    _fields_ = [("meshes", C.POINTER(MeshDesc)), ("num_meshes", C.c_int32),
                ("materials", C.POINTER(MaterialDesc)), ("num_materials", C.c_int32),
                ("lights", C.POINTER(LightDesc)), ("num_lights", C.c_int32),
                ("environment", C.POINTER(EnvDesc)), ("light_groups", C.POINTER(C.c_int32)),
                ("env_light_group", C.c_int32), ("num_light_groups", C.c_int32),
                ("node_uids", C.POINTER(C.c_uint64)), ("background_mode", C.c_int32),
                ("background_color", C.c_float * 3), ("link_modes", C.POINTER(C.c_int32)),
                ("link_offsets", C.POINTER(C.c_uint32)), ("link_uids", C.POINTER(C.c_uint64))]


# Background::Mode (engine/scene.h:24-28)
PE_BACKGROUND_ENVIRONMENT, PE_BACKGROUND_COLOR, PE_BACKGROUND_TRANSPARENT = 0, 1, 2


PE_MAX_LIGHT_GROUPS = 16
# LightPass order (integrators/path_tracer.h:21-24)
PE_PASS_DIFFUSE_DIRECT, PE_PASS_DIFFUSE_INDIRECT, PE_PASS_GLOSSY_DIRECT, PE_PASS_GLOSSY_INDIRECT = 0, 1, 2, 3
PE_PASS_TRANSMISSION_DIRECT, PE_PASS_TRANSMISSION_INDIRECT, PE_PASS_EMISSION, PE_PASS_BACKGROUND = 4, 5, 6, 7
PE_LIGHT_PASSES = 8
LIGHT_PASS_NAMES = ("DiffuseDirect", "DiffuseIndirect", "GlossyDirect", "GlossyIndirect", "TransmissionDirect",
                    "TransmissionIndirect", "Emission", "Background")


class PassTargets(C.Structure):
    _fields_ = [("light", C.c_void_p * PE_LIGHT_PASSES), ("light_groups", C.c_void_p * PE_MAX_LIGHT_GROUPS),
                ("depth", C.c_void_p), ("position", C.c_void_p), ("object_id", C.c_void_p),
                ("material_id", C.c_void_p), ("crypto_object", C.c_void_p), ("crypto_material", C.c_void_p),
                ("object_ids", C.POINTER(C.c_float)), ("material_ids", C.POINTER(C.c_float)),
                ("num_object_ids", C.c_int32), ("num_material_ids", C.c_int32)]


PE_CRYPTO_SLOTS = 8


class Camera(C.Structure):
    _fields_ = [("position", F3), ("lower_left", F3), ("horizontal", F3), ("vertical", F3),
                ("lens_radius", C.c_float), ("u", F3), ("v", F3), ("forward", F3), ("blades", C.c_int32),
                ("blade_rotation", C.c_float), ("orthographic", C.c_int32), ("has_shutter_end", C.c_int32),
                ("end_position", F3), ("end_lower_left", F3), ("end_horizontal", F3), ("end_vertical", F3),
                ("end_u", F3), ("end_v", F3), ("end_forward", F3), ("end_lens_radius", C.c_float)]


class Settings(C.Structure):
    _fields_ = [("max_bounces", C.c_int32), ("rr_depth", C.c_int32), ("shadow_quality", C.c_int32),
                ("filter_type", C.c_int32), ("filter_width", C.c_float), ("clamp_indirect", C.c_float),
                ("seed", C.c_uint64), ("samples_per_launch", C.c_int32), ("count_visits", C.c_int32),
                ("adaptive", C.c_int32), ("noise_threshold", C.c_float), ("min_adaptive_samples", C.c_int32),
                ("paths_in_flight", C.c_int32), ("motion_blur", C.c_int32), ("ao_strength", C.c_float)]


class BatchResult(C.Structure):
    _fields_ = [("completed", C.c_int32), ("samples_traced", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_visits", C.c_uint64),
                ("shadow_node_visits", C.c_uint64), ("shadow_tri_visits", C.c_uint64),
                ("max_stack_depth", C.c_uint32), ("stack_overflows", C.c_uint32), ("closest_hits", C.c_uint64)]


CANCEL_POLL = C.CFUNCTYPE(C.c_int, C.c_void_p)


class Targets(C.Structure):
    _fields_ = [("beauty", C.c_void_p), ("albedo", C.c_void_p), ("normal", C.c_void_p), ("alpha", C.c_void_p),
                ("guide", C.c_void_p), ("cancel_poll", CANCEL_POLL), ("cancel_user", C.c_void_p),
                ("passes", C.POINTER(PassTargets))]


PE_TONEMAP_REINHARD, PE_TONEMAP_REINHARD_EXTENDED, PE_TONEMAP_ACES, PE_TONEMAP_FILMIC = 0, 1, 2, 3
PE_TONEMAP_AGX, PE_TONEMAP_PBR_NEUTRAL, PE_TONEMAP_LINEAR = 4, 5, 6


class Display(C.Structure):
    _fields_ = [("tonemap", C.c_int32), ("exposure_ev", C.c_float), ("dither", C.c_int32),
                ("bloom_intensity", C.c_float), ("bloom_radius", C.c_float), ("bloom_threshold", C.c_float),
                ("glare_intensity", C.c_float), ("glare_blades", C.c_int32), ("glare_angle_deg", C.c_float),
                ("glare_length", C.c_float), ("white_balance_k", C.c_float), ("tint", C.c_float),
                ("saturation", C.c_float), ("contrast", C.c_float), ("vignette", C.c_float),
                ("chromatic_aberration", C.c_float), ("denoise", C.c_int32), ("reserved", C.c_float * 3)]


class KernelStats(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class GuideConfig(C.Structure):
    _fields_ = [("learning_rate", C.c_float), ("guide_fraction", C.c_float), ("batch", C.c_int32),
                ("steps_per_launch", C.c_int32), ("train_launches", C.c_int32), ("seed", C.c_uint32),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("weight_clamp", C.c_float),
                ("guided_bounces", C.c_int32), ("reserved", C.c_int32 * 4)]


PE_GUIDE_INPUTS, PE_GUIDE_HIDDEN, PE_GUIDE_OUTPUTS, PE_GUIDE_LAYERS, PE_GUIDE_LOBES = 160, 64, 64, 5, 8
PE_GUIDE_PARAMS = 26944


class ServerConfig(C.Structure):
    _fields_ = [("slots", C.c_int32), ("chunk_spp", C.c_int32), ("max_chunk_paths", C.c_uint64),
                ("memory_budget", C.c_uint64), ("max_admitted", C.c_int32), ("lane_paths", C.c_int32),
                ("reserved", C.c_int32 * 4)]


class ServerStats(C.Structure):
    _fields_ = [("memory_budget", C.c_uint64), ("memory_used", C.c_uint64), ("lane_bytes", C.c_uint64),
                ("scenes_parked", C.c_uint64), ("scenes_restored", C.c_uint64), ("admitted", C.c_int32),
                ("queued", C.c_int32), ("scenes_tracked", C.c_int32), ("scenes_resident", C.c_int32)]


class JobDesc(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("camera", Camera), ("settings", Settings), ("width", C.c_int32),
                ("height", C.c_int32), ("spp", C.c_int32), ("preview_spp", C.c_int32), ("reserved", C.c_int32 * 4)]


class JobStatus(C.Structure):
    _fields_ = [("state", C.c_int32), ("samples_done", C.c_int32), ("error", C.c_int32), ("chunks", C.c_int32),
                ("t_submit", C.c_double), ("t_admit", C.c_double), ("t_preview", C.c_double),
                ("t_done", C.c_double), ("extension_rays", C.c_uint64), ("shadow_rays", C.c_uint64),
                ("message", C.c_char * 128)]


class PhotonHipError(RuntimeError):
    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} -> {code}: {msg}")
        self.code = code


_VP = C.c_void_p
_P = C.POINTER

# Every symbol include/photon_hip.h declares, with its ctypes signature.
SIGNATURES = {
    "pe_last_error": (C.c_char_p, []),
    "pe_abi_version": (C.c_int, []),
    "pe_device_count": (C.c_int, [_P(C.c_int)]),
    "pe_device_info_get": (C.c_int, [C.c_int, _P(DeviceInfo)]),
    "pe_device_create": (C.c_int, [C.c_int, _P(_VP)]),
    "pe_device_destroy": (C.c_int, [_VP]),
    "pe_device_features_get": (C.c_int, [_VP, _P(DeviceFeatures)]),
    "pe_device_synchronize": (C.c_int, [_VP]),
    "pe_scene_upload": (C.c_int, [_VP, _P(SceneDesc), _P(_VP)]),
    "pe_scene_update": (C.c_int, [_VP, C.c_int, _P(SceneDesc)]),
    "pe_scene_release": (C.c_int, [_VP]),
    "pe_scene_stats": (C.c_int, [_VP, _P(C.c_uint64), _P(C.c_uint64), _P(C.c_uint64)]),
    "pe_camera_perspective": (C.c_int, [_P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float, C.c_float, _P(Camera)]),
    "pe_camera_thin_lens": (C.c_int, [_P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float, C.c_float, C.c_float,
                                      C.c_float, C.c_int, C.c_float, _P(Camera)]),
    "pe_camera_orthographic": (C.c_int, [_P(C.c_float), _P(C.c_float), _P(C.c_float), C.c_float, C.c_float,
                                         _P(Camera)]),
    "pe_camera_set_shutter_end": (C.c_int, [_P(Camera), _P(Camera)]),
    "pe_camera_set_shift": (C.c_int, [_P(Camera), C.c_float, C.c_float]),
    "pe_film_create": (C.c_int, [_VP, C.c_int, C.c_int, _P(_VP)]),
    "pe_film_destroy": (C.c_int, [_VP]),
    "pe_film_clear": (C.c_int, [_VP]),
    "pe_film_download": (C.c_int, [_VP, _P(C.c_float), _P(C.c_uint32), _P(C.c_float), _P(C.c_uint64)]),
    "pe_crypto_film_create": (C.c_int, [_VP, C.c_int, C.c_int, _P(_VP)]),
    "pe_crypto_film_destroy": (C.c_int, [_VP]),
    "pe_crypto_film_clear": (C.c_int, [_VP]),
    "pe_crypto_film_download": (C.c_int, [_VP, _P(C.c_float), _P(C.c_uint32), _P(C.c_uint32), _P(C.c_uint64)]),
    "pe_display_defaults": (C.c_int, [_P(Display)]),
    "pe_film_display": (C.c_int, [_VP, _P(Display), _P(C.c_float), _P(C.c_uint8)]),
    "pe_job_display": (C.c_int, [_VP, C.c_uint64, _P(Display), _P(C.c_float), _P(C.c_uint8)]),
    "pe_render_targets": (C.c_int, [_VP, _VP, _P(Camera), _P(Settings), _P(Targets), C.c_int, C.c_int,
                                    _P(C.c_int32), _P(BatchResult)]),
    "pe_render": (C.c_int, [_VP, _VP, _P(Camera), _P(Settings), _VP, C.c_int, C.c_int, _P(C.c_int32),
                            _P(BatchResult)]),
    "pe_profile_enable": (C.c_int, [_VP, C.c_int]),
    "pe_profile_read": (C.c_int, [_VP, _P(KernelStats), C.c_int, _P(C.c_int)]),
    "pe_sobol_batch": (C.c_int, [_VP, C.c_uint64, C.c_size_t, _P(C.c_int32), _P(C.c_int32), _P(C.c_uint32),
                                 _P(C.c_uint32), _P(C.c_uint32)]),
    "pe_intersect_batch": (C.c_int, [_VP, _VP, C.c_size_t, _P(C.c_float), _P(C.c_int32), _P(C.c_float)]),
    "pe_occluded_batch": (C.c_int, [_VP, _VP, C.c_size_t, _P(C.c_float), _P(C.c_uint8)]),
    "pe_trace_diag": (C.c_int, [_VP, _P(C.c_uint32), _P(C.c_uint64)]),
    "pe_intersect_batch_noalloc": (C.c_int, [_VP, _VP, C.c_size_t, _VP, _VP, _VP]),
    "pe_guide_create": (C.c_int, [_VP, _VP, _P(GuideConfig), _P(_VP)]),
    "pe_guide_destroy": (C.c_int, [_VP]),
    "pe_guide_get_params": (C.c_int, [_VP, _P(C.c_float)]),
    "pe_guide_set_params": (C.c_int, [_VP, _P(C.c_float)]),
    "pe_guide_infer": (C.c_int, [_VP, C.c_size_t, _P(C.c_float), _P(C.c_float), _P(C.c_float)]),
    "pe_guide_train_step": (C.c_int, [_VP, C.c_size_t, _P(C.c_float), _P(C.c_float), _P(C.c_float)]),
    "pe_guide_stats": (C.c_int, [_VP, _P(C.c_uint64), _P(C.c_float), _P(C.c_uint64), _P(C.c_uint64)]),
    "pe_render_guided": (C.c_int, [_VP, _VP, _VP, _P(Camera), _P(Settings), _VP, C.c_int, C.c_int, _P(C.c_int32),
                                   _P(BatchResult)]),
    "pe_server_create": (C.c_int, [_VP, _P(ServerConfig), _P(_VP)]),
    "pe_server_destroy": (C.c_int, [_VP]),
    "pe_server_clock": (C.c_double, [_VP]),
    "pe_job_submit": (C.c_int, [_VP, _P(JobDesc), _P(C.c_uint64)]),
    "pe_job_poll": (C.c_int, [_VP, C.c_uint64, _P(JobStatus)]),
    "pe_job_wait": (C.c_int, [_VP, C.c_uint64, C.c_int32, _P(JobStatus)]),
    "pe_job_cancel": (C.c_int, [_VP, C.c_uint64]),
    "pe_job_download": (C.c_int, [_VP, C.c_uint64, _P(C.c_float), _P(C.c_uint32), _P(C.c_float)]),
    "pe_job_release": (C.c_int, [_VP, C.c_uint64]),
    "pe_server_stats_get": (C.c_int, [_VP, _VP]),
}

_lib = None


def load() -> C.CDLL:
    """Load libphoton_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise FileNotFoundError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950) first")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(func: str, rc: int) -> None:
    if rc != PE_OK:
        raise PhotonHipError(func, rc, load().pe_last_error().decode(errors="replace"))


def ptr(a: np.ndarray | None, ctype):
    if a is None:
        return C.cast(None, C.POINTER(ctype))
    assert a.flags["C_CONTIGUOUS"], "array must be C-contiguous"
    return a.ctypes.data_as(C.POINTER(ctype))
