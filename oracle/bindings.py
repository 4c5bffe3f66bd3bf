"""ctypes bindings of the oracle libraries (test infrastructure only).

liboracle.so               the C++ restatement of the reference hot path (oracle_*.cpp)
_ref/libphoton_ref.so      the reference's own sources built with its own flags (oracle/Makefile.ref)
_ref/libphoton_ref_strict  the same sources built without floating-point contraction

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import this module.
"""
from __future__ import annotations

import ctypes as C
import sys
from pathlib import Path

import numpy as np

ORACLE_DIR = Path(__file__).resolve().parent
ROOT = ORACLE_DIR.parent
sys.path.insert(0, str(ROOT / "photon-engine_amd" / "python"))
from photon_hip import capi  # noqa: E402  (ctypes structure definitions shared with the product)

_VP = C.c_void_p
_P = C.POINTER
_F = _P(C.c_float)
_I = _P(C.c_int32)
_U = _P(C.c_uint32)


def _fp(a):
    return capi.ptr(a, C.c_float)


class _Lib:
    def __init__(self, path: Path, prefix: str, sigs: dict):
        self.path = path
        self.lib = C.CDLL(str(path))
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, prefix + name)
            fn.restype = res
            fn.argtypes = args

    def __getattr__(self, item):
        return getattr(self.lib, item)


_ORACLE_SIGS = {
    "sobol_batch": (C.c_int, [C.c_uint64, C.c_size_t, _I, _I, _U, _U, _U]),
    "reverse_bits": (C.c_uint32, [C.c_uint32]),
    "sobol_dim1": (C.c_uint32, [C.c_uint32]),
    "scene_build": (_VP, [_P(capi.SceneDesc)]),
    "scene_free": (None, [_VP]),
    "scene_triangles": (C.c_int, [_VP]),
    "intersect": (C.c_int, [_VP, C.c_size_t, _F, _I, _F]),
    "intersect_brute": (C.c_int, [_VP, C.c_size_t, _F, _I]),
    "occluded": (C.c_int, [_VP, C.c_size_t, _F, _P(C.c_uint8)]),
    "trace_samples": (C.c_int, [_VP, _P(capi.Camera), _P(capi.Settings), C.c_int, C.c_int, C.c_size_t, _I, _I, _I,
                                _F, _P(C.c_uint64)]),
    "render": (C.c_int, [_VP, _P(capi.Camera), _P(capi.Settings), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                         C.c_int, C.c_int, _F, _U, _F, _P(C.c_uint64), _P(C.c_double)]),
    "bsdf_eval": (C.c_int, [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _I, _F, _F, _F, _F]),
    "bsdf_sample": (C.c_int, [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _I, _F, _F, _F]),
    "microfacet": (C.c_int, [C.c_int, C.c_size_t, _F, _F, _F, _F]),
    "light_select": (C.c_int, [_VP, C.c_size_t, _F, _F, _F, _F, _I, _F, _I]),
    "spectral": (C.c_int, [C.c_size_t, _F, _F, _F, C.c_float, C.c_float, C.c_size_t, _F, _F]),
    "thin_film": (C.c_int, [C.c_size_t, _F, C.c_float, C.c_float, C.c_float, C.c_float, _F, _F, _F]),
    "display": (C.c_int, [C.c_int, C.c_int, _F, _U, _P(capi.Display), _F, _P(C.c_uint8)]),
    "trace_aovs": (C.c_int, [_VP, _P(capi.Camera), _P(capi.Settings), C.c_int, C.c_int, C.c_size_t, _I, _I, _I, _F,
                             _F]),
    "trace_passes": (C.c_int, [_VP, _P(capi.Camera), _P(capi.Settings), C.c_int, C.c_int, C.c_size_t, _I, _I, _I,
                               _F, _F]),
    "camera_rays": (C.c_int, [_P(capi.Camera), C.c_size_t, _F, _F]),
}

_REF_SIGS = {
    "sobol_batch": (C.c_int, [C.c_uint64, C.c_size_t, _I, _I, _U, _U, _F]),
    "scene_build": (_VP, [_P(capi.SceneDesc)]),
    "scene_free": (None, [_VP]),
    "scene_counts": (C.c_int, [_VP, _I, _I]),
    "intersect": (C.c_int, [_VP, C.c_size_t, _F, _I, _F]),
    "occluded": (C.c_int, [_VP, C.c_size_t, _F, _P(C.c_uint8)]),
    "camera_describe": (C.c_int, [_F, _F, _F, C.c_float, C.c_float, _F]),
    "camera_select": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int, _F,
                                _F, C.c_float]),
    "camera_describe_full": (C.c_int, [_F, _F, _F, C.c_float, C.c_float, _F]),
    "camera_shift": (C.c_int, [C.c_float, C.c_float]),
    "ao_strength": (C.c_int, [C.c_float]),
    "camera_rays": (C.c_int, [_F, _F, _F, C.c_float, C.c_float, C.c_size_t, _F, _F]),
    "microfacet": (C.c_int, [C.c_int, C.c_size_t, _F, _F, _F, _F]),
    "bsdf_eval": (C.c_int, [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _I, _F, _F, _F, _F]),
    "bsdf_sample": (C.c_int, [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _I, _F, _F, _F]),
    "trace_samples": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_float, C.c_uint64, C.c_size_t, _I, _I, _I, _F]),
    "render": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                         C.c_float, C.c_int, C.c_int, C.c_int, _F, _U, _P(C.c_double)]),
    "thin_film": (C.c_int, [C.c_size_t, _F, C.c_float, C.c_float, C.c_float, C.c_float, _F, _F, _F]),
    "display": (C.c_int, [C.c_int, C.c_int, _F, _U, _P(capi.Display), _F, _P(C.c_uint8)]),
    "render_aovs": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                              _F, _F, _F, _F]),
    "render_adaptive": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, _I, _I,
                                  C.c_float, C.c_int, _F, _U, _F, _P(C.c_uint64)]),
    "render_passes": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                C.c_int, _F, _F, _F]),
    "scene_build_groups": (_VP, [_P(capi.SceneDesc), _P(C.c_char_p), C.c_char_p, _P(C.c_char_p), _P(C.c_char_p)]),
    "scene_build_instanced": (_VP, [_P(capi.SceneDesc), _P(C.c_char_p), C.c_char_p, _P(C.c_char_p),
                                    _P(C.c_char_p), _F, _P(C.c_uint8)]),
    "crypto_hash": (C.c_int, [C.c_size_t, _P(C.c_char_p), _U, _F]),
    "mesh_encode": (C.c_size_t, [_F, _F, _F, C.c_uint32, _U, C.c_uint32, _P(C.c_uint8), C.c_size_t]),
    "mesh_decode": (C.c_int, [_P(C.c_uint8), C.c_size_t, _F, _F, _U, _U]),
    "temperature_color": (C.c_int, [C.c_size_t, _F, _F]),
    "scene_light_groups": (C.c_int, [_VP, _I, _I, C.c_char_p, C.c_size_t]),
    "trace_passes": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                               C.c_float, C.c_uint64, C.c_size_t, _I, _I, _I, _F, _F]),
}

def _scene_names(scene):
    """The names the reference's Scene carries and resolves itself: light-group names of the lights and of the
    environment, shape names and material names (C arrays for ref_scene_build_groups)."""
    groups = (C.c_char_p * max(1, len(scene.lights)))(*[g.encode() for g in scene.light_group_of])
    env = scene.environment.light_group.encode() if scene.environment is not None else None
    shapes = (C.c_char_p * max(1, len(scene.shapes)))(*[sh.name.encode() for sh in scene.shapes])
    mats = (C.c_char_p * max(1, len(scene.materials)))(
        *[scene.material_names.get(i, "").encode() for i in range(len(scene.materials))])
    return groups, env, shapes, mats


def _instancing(scene, instances, visibility):
    """Per-mesh local-to-world matrices (16 floats row-major, zeros where a mesh is not instanced) and visibility
    masks for ref_scene_build_instanced, from {mesh index: 4x4 matrix} and {mesh index: mask}."""
    n = max(1, len(scene.shapes))
    xf = np.zeros((n, 16), np.float32)
    for i, m in (instances or {}).items():
        xf[i] = np.asarray(m, np.float32).reshape(16)
    vis = np.full(n, 0x1F, np.uint8)
    for i, v in (visibility or {}).items():
        vis[i] = v
    return xf, vis


def _build(lib, prefix, scene, instances, visibility, keep):
    desc, k = scene.pack()
    keep += k
    args = _scene_names(scene)
    keep += list(args)
    if instances is None and visibility is None:
        return getattr(lib, prefix + "scene_build_groups")(C.byref(desc), *args)
    xf, vis = _instancing(scene, instances, visibility)
    keep += [xf, vis]
    return getattr(lib, prefix + "scene_build_instanced")(C.byref(desc), *args, _fp(xf), capi.ptr(vis, C.c_uint8))


def crypto_hash_reference(names):
    """The reference's cryptoHashName / cryptoHashToFloat of each name -> (uint32 hashes, float32 ids)."""
    lib = reference(strict=True)
    arr = (C.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    h = np.zeros(len(names), np.uint32)
    f = np.zeros(len(names), np.float32)
    lib.ref_crypto_hash(len(names), arr, capi.ptr(h, C.c_uint32), _fp(f))
    return h, f


PASS_FLOATS = 84  # per sample: 8 light passes, 16 light groups (rgb each), depth, position, object id, material id,
#                   then the first surface's shape and material index (-1: miss)


def split_passes(out: np.ndarray) -> dict:
    """The per-sample pass record (n, 84) as named arrays."""
    return {"light": out[:, :24].reshape(-1, 8, 3), "groups": out[:, 24:72].reshape(-1, 16, 3),
            "depth": out[:, 72], "position": out[:, 73:76], "object_id": out[:, 76:79],
            "material_id": out[:, 79:82], "shape": out[:, 82].astype(np.int32),
            "material": out[:, 83].astype(np.int32)}

_cache: dict = {}


def oracle() -> _Lib:
    if "oracle" not in _cache:
        path = ORACLE_DIR / "liboracle.so"
        if not path.exists():
            raise FileNotFoundError(f"{path} missing: run __graft_entry__.build()")
        _cache["oracle"] = _Lib(path, "oracle_", _ORACLE_SIGS)
    return _cache["oracle"]


def reference(strict: bool = False) -> _Lib | None:
    """The reference's own code (None where oracle/_ref was not built)."""
    key = "ref_strict" if strict else "ref"
    if key not in _cache:
        path = ORACLE_DIR / "_ref" / ("libphoton_ref_strict.so" if strict else "libphoton_ref.so")
        _cache[key] = _Lib(path, "ref_", _REF_SIGS) if path.exists() else None
    return _cache[key]


class OracleScene:
    """The restated scene (binary BVH built as the reference builds it)."""

    def __init__(self, scene):
        desc, self._keep = scene.pack()
        self._lib = oracle()
        self.h = self._lib.oracle_scene_build(C.byref(desc))

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.oracle_scene_free(self.h)
            self.h = None

    def intersect(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.int32)
        bary = np.zeros((n, 2), np.float32)
        self._lib.oracle_intersect(self.h, n, _fp(rays), capi.ptr(hits, C.c_int32), _fp(bary))
        return hits, bary

    def intersect_brute(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        hits = np.zeros((rays.shape[0], 4), np.int32)
        self._lib.oracle_intersect_brute(self.h, rays.shape[0], _fp(rays), capi.ptr(hits, C.c_int32))
        return hits

    def occluded(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(rays.shape[0], np.uint8)
        self._lib.oracle_occluded(self.h, rays.shape[0], _fp(rays), capi.ptr(occ, C.c_uint8))
        return occ

    def light_select(self, p, n, u, with_pmf=True):
        """LightSampler::Context at each (p, n): (pmf (N, candidates) or None, sampled index (N,), its pmf (N,),
        whether the light BVH is used)."""
        p = np.ascontiguousarray(p, np.float32)
        n = np.ascontiguousarray(n, np.float32)
        u = np.ascontiguousarray(u, np.float32)
        N = p.shape[0]
        cnt = self._lib.oracle_light_select(self.h, 0, _fp(p), _fp(n), _fp(u), None, None, None, None)
        pmf = np.zeros((N, max(cnt, 1)), np.float32) if with_pmf else None
        idx = np.zeros(N, np.int32)
        spmf = np.zeros(N, np.float32)
        bvh = C.c_int32()
        rc = self._lib.oracle_light_select(self.h, N, _fp(p), _fp(n), _fp(u), _fp(pmf) if with_pmf else None,
                                           capi.ptr(idx, C.c_int32), _fp(spmf), C.byref(bvh))
        if rc < 0:
            raise RuntimeError("oracle_light_select failed")
        return pmf, idx, spmf, bool(bvh.value)

    def trace_samples(self, camera, settings, width, height, px, py, s):
        """Per-sample radiance (tracePixel + PathTracer::Li). Returns (L (n, 3), (extension, shadow) rays)."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        L = np.zeros((px.shape[0], 3), np.float32)
        rays = (C.c_uint64 * 2)()
        rc = self._lib.oracle_trace_samples(self.h, C.byref(camera), C.byref(settings), width, height, px.shape[0],
                                            capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32), capi.ptr(s, C.c_int32),
                                            _fp(L), rays)
        if rc != 0:
            raise RuntimeError("oracle_trace_samples failed (feature outside the restated scope)")
        return L, (rays[0], rays[1])

    def trace_aovs(self, camera, settings, width, height, px, py, s):
        """Per-sample radiance and AOVs -> (L (n, 3), albedo (n, 3), normal (n, 3), alpha (n,))."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        L = np.zeros((px.shape[0], 3), np.float32)
        aov = np.zeros((px.shape[0], 7), np.float32)
        rc = self._lib.oracle_trace_aovs(self.h, C.byref(camera), C.byref(settings), width, height, px.shape[0],
                                         capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32), capi.ptr(s, C.c_int32),
                                         _fp(L), _fp(aov))
        if rc != 0:
            raise RuntimeError("oracle_trace_aovs failed (feature outside the restated scope)")
        return L, aov[:, 0:3], aov[:, 3:6], aov[:, 6]

    def trace_passes(self, camera, settings, width, height, px, py, s):
        """Per-sample radiance and render passes -> (L (n, 3), split_passes dict)."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        L = np.zeros((px.shape[0], 3), np.float32)
        out = np.zeros((px.shape[0], PASS_FLOATS), np.float32)
        rc = self._lib.oracle_trace_passes(self.h, C.byref(camera), C.byref(settings), width, height, px.shape[0],
                                           capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32), capi.ptr(s, C.c_int32),
                                           _fp(L), _fp(out))
        if rc != 0:
            raise RuntimeError("oracle_trace_passes failed (feature outside the restated scope)")
        return L, split_passes(out)

    def render(self, camera, settings, width, height, first, count, threads=1, rows=None, film=None):
        """Accumulates samples [first, first+count) of every pixel (rows [r0, r1) if given) into film
        (sum (H,W,3), count (H,W), lum2 (H,W)); returns (film, seconds, (extension, shadow) rays)."""
        if film is None:
            film = (np.zeros((height, width, 3), np.float32), np.zeros((height, width), np.uint32),
                    np.zeros((height, width), np.float32))
        r0, r1 = rows if rows is not None else (0, height)
        rays = (C.c_uint64 * 2)()
        secs = C.c_double()
        rc = self._lib.oracle_render(self.h, C.byref(camera), C.byref(settings), width, height, first, count, threads,
                                     r0, r1, _fp(film[0]), capi.ptr(film[1], C.c_uint32), _fp(film[2]), rays,
                                     C.byref(secs))
        if rc != 0:
            raise RuntimeError("oracle_render failed (feature outside the restated scope)")
        return film, secs.value, (rays[0], rays[1])


CAMERA_KINDS = {"perspective": 0, "thin_lens": 1, "orthographic": 2}


def camera_rays_oracle(camera, q):
    """The oracle's camera ray for q (n, 5): u, v, lens x, lens y, time (< 0: no shutter time) -> (n, 6)."""
    q = np.ascontiguousarray(q, np.float32)
    out = np.zeros((q.shape[0], 6), np.float32)
    oracle().oracle_camera_rays(C.byref(camera), q.shape[0], _fp(q), _fp(out))
    return out


class reference_camera:
    """Context manager: the reference driver builds its cameras as ThinLensCamera (aperture, focus, blades,
    rotation_deg) or OrthographicCamera (extent) instead of PerspectiveCamera, optionally with a shutter-end pose of
    the same kind at end_position / end_target (end_aperture) and the motion-blur flag; back to the pinhole on exit."""

    def __init__(self, strict: bool = False, kind="perspective", aperture=0.0, focus=1.0, blades=0, rotation_deg=0.0,
                 extent=2.0, motion_blur=False, end_position=None, end_target=None, end_aperture=0.0,
                 shift=(0.0, 0.0), ao_strength=0.0, harness: bool = False):
        """harness: select the camera of the HipDevice harness (its own copy of the reference driver)."""
        self.lib = hip_device_harness() if harness else reference(strict)
        if self.lib is None:
            raise FileNotFoundError("oracle/_ref was not built (the reference tree is absent)")
        if harness:
            for name in ("camera_select", "camera_shift", "ao_strength"):
                getattr(self.lib, "ref_" + name).argtypes = _REF_SIGS[name][1]
                getattr(self.lib, "ref_" + name).restype = C.c_int
        self.shift = (float(shift[0]), float(shift[1]))
        self.ao = float(ao_strength)
        self.args = (CAMERA_KINDS[kind], float(aperture), float(focus), int(blades), float(rotation_deg),
                     float(extent), int(bool(motion_blur)), int(end_position is not None),
                     (C.c_float * 3)(*map(float, end_position or (0, 0, 0))),
                     (C.c_float * 3)(*map(float, end_target or (0, 0, -1))), float(end_aperture))

    def __enter__(self):
        self.lib.ref_camera_select(*self.args)
        self.lib.ref_camera_shift(*self.shift)
        self.lib.ref_ao_strength(self.ao)
        return self

    def __exit__(self, *exc):
        self.lib.ref_camera_select(0, 0.0, 1.0, 0, 0.0, 2.0, 0, 0, None, None, 0.0)
        return False

    def describe(self, view):
        """describeGpu() and describeGpuEnd() of the selected camera: (2, 22) floats, then blades, rotation, ortho."""
        out = np.zeros(47, np.float32)
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        self.lib.ref_camera_describe_full(f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect,
                                          _fp(out))
        return out

    def rays(self, view, q):
        """Camera::generateRay for q (n, 5): u, v, lens x, lens y, time (< 0: the three-argument form) -> (n, 6)."""
        q = np.ascontiguousarray(q, np.float32)
        out = np.zeros((q.shape[0], 6), np.float32)
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        self.lib.ref_camera_rays(f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect, q.shape[0],
                                 _fp(q), _fp(out))
        return out


class RefScene:
    """engine::Scene built by the reference's own code from the same descriptor."""

    def __init__(self, scene, strict: bool = False, instances=None, visibility=None):
        """instances: {mesh index: 4x4 local-to-world matrix} makes those meshes MeshInstances of the (local)
        mesh; visibility: {mesh index: Shape visibility mask}."""
        self._lib = reference(strict)
        if self._lib is None:
            raise FileNotFoundError("oracle/_ref was not built (the reference tree is absent)")
        self._keep = []
        self.h = _build(self._lib, "ref_", scene, instances, visibility, self._keep)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.ref_scene_free(self.h)
            self.h = None

    def light_groups(self):
        """Scene::lightGroupNames, lightGroupIndex per light, envLightGroupIndex."""
        nl = self.counts()[1]
        per = np.zeros(max(1, nl), np.int32)
        env = C.c_int32()
        buf = C.create_string_buffer(4096)
        n = self._lib.ref_scene_light_groups(self.h, capi.ptr(per, C.c_int32), C.byref(env), buf, 4096)
        names = buf.value.decode().split("\n")[:n]
        return names, per[:nl].tolist(), env.value

    def trace_passes(self, view, px, py, s, shadow_quality=1, filter_type=0, filter_width=1.0, seed=12345):
        """Per-sample radiance and render passes along tracePixel's path -> (L (n, 3), split_passes dict)."""
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        L = np.zeros((px.shape[0], 3), np.float32)
        out = np.zeros((px.shape[0], PASS_FLOATS), np.float32)
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        self._lib.ref_trace_passes(self.h, f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect,
                                   view.width, view.height, view.max_bounces, shadow_quality, filter_type,
                                   filter_width, seed, px.shape[0], capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32),
                                   capi.ptr(s, C.c_int32), _fp(L), _fp(out))
        return L, split_passes(out)

    def counts(self):
        s, l = C.c_int32(), C.c_int32()
        self._lib.ref_scene_counts(self.h, C.byref(s), C.byref(l))
        return s.value, l.value

    def intersect(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.int32)
        geom = np.zeros((n, 9), np.float32)
        self._lib.ref_intersect(self.h, n, _fp(rays), capi.ptr(hits, C.c_int32), _fp(geom))
        return hits, geom

    def occluded(self, rays: np.ndarray):
        rays = np.ascontiguousarray(rays, dtype=np.float32)
        occ = np.zeros(rays.shape[0], np.uint8)
        self._lib.ref_occluded(self.h, rays.shape[0], _fp(rays), capi.ptr(occ, C.c_uint8))
        return occ

    def trace_samples(self, view, px, py, s, shadow_quality=1, filter_type=0, filter_width=1.0, seed=12345):
        px = np.ascontiguousarray(px, np.int32)
        py = np.ascontiguousarray(py, np.int32)
        s = np.ascontiguousarray(s, np.int32)
        L = np.zeros((px.shape[0], 3), np.float32)
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        self._lib.ref_trace_samples(self.h, f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect,
                                    view.width, view.height, view.max_bounces, shadow_quality, filter_type,
                                    filter_width, seed, px.shape[0], capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32),
                                    capi.ptr(s, C.c_int32), _fp(L))
        return L

    def render_aovs(self, view, first, count):
        """Renderer::renderSamples with AovFilms -> resolved (beauty, albedo, normal, alpha), each (h, w, 3)."""
        w, h = view.width, view.height
        out = [np.zeros((h, w, 3), np.float32) for _ in range(4)]
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        rc = self._lib.ref_render_aovs(self.h, f3(view.position), f3(view.target), f3(view.up), view.vfov,
                                       view.aspect, w, h, view.max_bounces, first, count, *[_fp(a) for a in out])
        if rc != 0:
            raise RuntimeError("ref_render_aovs failed")
        return out

    def render(self, view, first, count, threads, shadow_quality=1, filter_type=0, filter_width=1.0):
        w, h = view.width, view.height
        mean = np.zeros((h, w, 3), np.float32)
        cnt = np.zeros((h, w), np.uint32)
        secs = C.c_double()
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        self._lib.ref_render(self.h, f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect, w, h,
                             view.max_bounces, shadow_quality, filter_type, filter_width, first, count, threads,
                             _fp(mean), capi.ptr(cnt, C.c_uint32), C.byref(secs))
        return mean, cnt, secs.value

    def render_passes(self, view, first, count, threads=4):
        """Renderer::renderSamples with RenderPassFilms -> (beauty (H,W,3), films (28, H,W,3), groups, crypto
        ranks (2, 3, H, W, 4): CryptoFilm::resolveRanks(3) of objects and of materials)."""
        w, h = view.width, view.height
        beauty = np.zeros((h, w, 3), np.float32)
        films = np.zeros((28, h, w, 3), np.float32)
        crypto = np.zeros((2, 3, h, w, 4), np.float32)
        f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
        n = self._lib.ref_render_passes(self.h, f3(view.position), f3(view.target), f3(view.up), view.vfov,
                                        view.aspect, w, h, view.max_bounces, first, count, threads, _fp(beauty),
                                        _fp(films), _fp(crypto))
        if n < 0:
            raise RuntimeError("ref_render_passes failed")
        return beauty, films, n, crypto


def display(lib_kind: str, sum_rgb, count, settings):
    """The display path of a film (sums (h, w, 3), counts (h, w)) by the oracle ('oracle') or the reference
    ('ref' / 'ref_strict') -> (display (h, w, 3) float32, rgb8 (h, w, 3) uint8)."""
    sum_rgb = np.ascontiguousarray(sum_rgb, np.float32)
    count = np.ascontiguousarray(count, np.uint32)
    h, w = count.shape
    out = np.zeros((h, w, 3), np.float32)
    rgb8 = np.zeros((h, w, 3), np.uint8)
    if lib_kind == "oracle":
        fn = oracle().oracle_display
    else:
        fn = reference(strict=(lib_kind == "ref_strict")).ref_display
    rc = fn(w, h, _fp(sum_rgb), capi.ptr(count, C.c_uint32), C.byref(settings), _fp(out), capi.ptr(rgb8, C.c_uint8))
    if rc != 0:
        raise RuntimeError("display failed")
    return out, rgb8


def ref_render_adaptive(scene_h, lib, view, batches, threshold, min_spp):
    """The reference's Renderer::renderSamples with adaptive sampling over `batches` [(first, count), ...] into
    one film -> (mean (h, w, 3), counts (h, w), relative errors (h, w), samples traced)."""
    w, h = view.width, view.height
    mean = np.zeros((h, w, 3), np.float32)
    cnt = np.zeros((h, w), np.uint32)
    rel = np.zeros((h, w), np.float32)
    traced = C.c_uint64()
    firsts = np.array([b[0] for b in batches], np.int32)
    counts = np.array([b[1] for b in batches], np.int32)
    f3 = lambda v: (C.c_float * 3)(*map(float, v))  # noqa: E731
    rc = lib.ref_render_adaptive(scene_h, f3(view.position), f3(view.target), f3(view.up), view.vfov, view.aspect,
                                 w, h, view.max_bounces, len(batches), capi.ptr(firsts, C.c_int32),
                                 capi.ptr(counts, C.c_int32), threshold, min_spp, _fp(mean), capi.ptr(cnt, C.c_uint32),
                                 _fp(rel), C.byref(traced))
    if rc != 0:
        raise RuntimeError("ref_render_adaptive failed")
    return mean, cnt, rel, traced.value


def bsdf_eval(lib_kind: str, material, normal, tangent, front, wo, wi, points=None):
    """lib_kind 'oracle' or 'ref': Material::evalPdf on n tuples -> (f (n,3), pdf (n,)); points: the shading
    points (n, 3), where flakes are resolved (default: the origin)."""
    n = normal.shape[0]
    args = [np.ascontiguousarray(a, np.float32) for a in (normal, tangent)]
    front = np.ascontiguousarray(front, np.int32)
    wo = np.ascontiguousarray(wo, np.float32)
    wi = np.ascontiguousarray(wi, np.float32)
    pts = np.ascontiguousarray(points, np.float32) if points is not None else None
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib = oracle().lib if lib_kind == "oracle" else reference(strict=(lib_kind == "ref_strict")).lib
    fn = getattr(lib, "oracle_bsdf_eval_at" if lib_kind == "oracle" else "ref_bsdf_eval_at")
    fn.restype = C.c_int
    fn.argtypes = [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _F, _I, _F, _F, _F, _F]
    rc = fn(C.byref(material), n, _fp(pts) if pts is not None else None, _fp(args[0]), _fp(args[1]),
            capi.ptr(front, C.c_int32), _fp(wo), _fp(wi), _fp(f), _fp(pdf))
    if rc != 0:
        raise RuntimeError("bsdf_eval failed")
    return f, pdf


def bsdf_sample(lib_kind: str, material, normal, tangent, front, wo, u3, points=None):
    """Material::sampleFull -> (n, 14): ok, wi, f, pdf, delta, subsurface, passThrough, weight."""
    n = normal.shape[0]
    normal = np.ascontiguousarray(normal, np.float32)
    tangent = np.ascontiguousarray(tangent, np.float32)
    front = np.ascontiguousarray(front, np.int32)
    wo = np.ascontiguousarray(wo, np.float32)
    u3 = np.ascontiguousarray(u3, np.float32)
    pts = np.ascontiguousarray(points, np.float32) if points is not None else None
    out = np.zeros((n, 14), np.float32)
    lib = oracle().lib if lib_kind == "oracle" else reference(strict=(lib_kind == "ref_strict")).lib
    fn = getattr(lib, "oracle_bsdf_sample_at" if lib_kind == "oracle" else "ref_bsdf_sample_at")
    fn.restype = C.c_int
    fn.argtypes = [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _F, _I, _F, _F, _F]
    rc = fn(C.byref(material), n, _fp(pts) if pts is not None else None, _fp(normal), _fp(tangent),
            capi.ptr(front, C.c_int32), _fp(wo), _fp(u3), _fp(out))
    if rc != 0:
        raise RuntimeError("bsdf_sample failed")
    return out


def flake_resolve(lib_kind: str, material, points, normal, tangent, front):
    """DisneyMaterial::resolve's flake state -> (n, 4): inFlake, flakeNormal."""
    n = normal.shape[0]
    arrs = [np.ascontiguousarray(a, np.float32) for a in (points, normal, tangent)]
    front = np.ascontiguousarray(front, np.int32)
    out = np.zeros((n, 4), np.float32)
    lib = oracle().lib if lib_kind == "oracle" else reference(strict=(lib_kind == "ref_strict")).lib
    fn = getattr(lib, "oracle_flake_resolve" if lib_kind == "oracle" else "ref_flake_resolve")
    fn.restype = C.c_int
    fn.argtypes = [_P(capi.MaterialDesc), C.c_size_t, _F, _F, _F, _I, _F]
    rc = fn(C.byref(material), n, *[_fp(a) for a in arrs], capi.ptr(front, C.c_int32), _fp(out))
    if rc != 0:
        raise RuntimeError("flake_resolve failed")
    return out


def microfacet(lib_kind: str, which: int, a, b, c=None):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    c = np.ascontiguousarray(c if c is not None else np.zeros_like(a), np.float32)
    out = np.zeros_like(a)
    fn = oracle().oracle_microfacet if lib_kind == "oracle" else reference(strict=(lib_kind == "ref_strict")).ref_microfacet
    fn(which, a.shape[0], _fp(a), _fp(b), _fp(c), _fp(out))
    return out


def sobol_oracle(seed, px, py, idx, dim):
    px, py = np.ascontiguousarray(px, np.int32), np.ascontiguousarray(py, np.int32)
    idx, dim = np.ascontiguousarray(idx, np.uint32), np.ascontiguousarray(dim, np.uint32)
    out = np.zeros((px.shape[0], 2), np.uint32)
    oracle().oracle_sobol_batch(seed, px.shape[0], capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32),
                                capi.ptr(idx, C.c_uint32), capi.ptr(dim, C.c_uint32), capi.ptr(out, C.c_uint32))
    return out


def sobol_reference(seed, px, py, idx, dim):
    lib = reference()
    if lib is None:
        return None
    px, py = np.ascontiguousarray(px, np.int32), np.ascontiguousarray(py, np.int32)
    idx, dim = np.ascontiguousarray(idx, np.uint32), np.ascontiguousarray(dim, np.uint32)
    out = np.zeros((px.shape[0], 2), np.float32)
    lib.ref_sobol_batch(seed, px.shape[0], capi.ptr(px, C.c_int32), capi.ptr(py, C.c_int32),
                        capi.ptr(idx, C.c_uint32), capi.ptr(dim, C.c_uint32), _fp(out))
    return out


_HIPDEV_SIGS = {
    "ref_scene_build": (_VP, [_P(capi.SceneDesc)]),
    "ref_scene_build_groups": (_VP, [_P(capi.SceneDesc), _P(C.c_char_p), C.c_char_p, _P(C.c_char_p),
                                     _P(C.c_char_p)]),
    "ref_scene_build_instanced": (_VP, [_P(capi.SceneDesc), _P(C.c_char_p), C.c_char_p, _P(C.c_char_p),
                                        _P(C.c_char_p), _F, _P(C.c_uint8)]),
    "hipdev_session_render_passes": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                               C.c_int, C.c_int, _F, _F, _U, _I, _F, C.c_char_p, C.c_int]),
    "ref_scene_free": (None, [_VP]),
    "hipdev_session_render": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int, C.c_int,
                                        C.c_int, C.c_float, C.c_int, C.c_int, C.c_float, C.c_int, _F, _U, _I,
                                        C.c_char_p, C.c_int]),
    "hipdev_enumerate": (C.c_int, [C.c_char_p, C.c_int]),
    "hipdev_session_render_aovs": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                             C.c_int, C.c_int, _F, _F, _F, _F, _U, C.c_char_p, C.c_int]),
    "hipdev_session_render_adaptive": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                                 C.c_int, C.c_int, C.c_float, C.c_int, _F, _U, _I, C.c_char_p,
                                                 C.c_int]),
    "hipdev_session_interactive": (C.c_int, [_VP, _F, _F, _F, C.c_float, C.c_float, C.c_int, C.c_int, C.c_int,
                                             C.c_int, _F, _U, C.POINTER(C.c_double), C.c_int, C.c_int,
                                             C.POINTER(C.c_double), _I, _I, C.c_char_p, C.c_int]),
}


def hip_device_harness() -> _Lib | None:
    """oracle/_ref/libphoton_hipdevice.so: the HipDevice adapter driven by the reference's RenderSession
    (None where oracle/_ref was not built)."""
    if "hipdev" not in _cache:
        path = ORACLE_DIR / "_ref" / "libphoton_hipdevice.so"
        _cache["hipdev"] = _Lib(path, "", _HIPDEV_SIGS) if path.exists() else None
    return _cache["hipdev"]


class HipDeviceSession:
    """A reference Scene built from the descriptor, rendered by HipDevice under RenderSession (Final mode)."""

    def __init__(self, scene, instances=None, visibility=None):
        """instances / visibility as RefScene takes them."""
        self._lib = hip_device_harness()
        if self._lib is None:
            raise FileNotFoundError("oracle/_ref/libphoton_hipdevice.so was not built (the reference tree is absent)")
        self._keep = []
        self.h = _build(self._lib, "ref_", scene, instances, visibility, self._keep)

    def __del__(self):
        if getattr(self, "h", None):
            self._lib.ref_scene_free(self.h)
            self.h = None

    def render(self, view, spp, max_batch=64, shadow_quality=1, filter_type=0, filter_width=1.0, ao_strength=0.0,
               path_guiding=False):
        """Returns (rc, mean (H,W,3), counts (H,W), batches, error text)."""
        f3 = C.c_float * 3
        mean = np.zeros((view.height, view.width, 3), np.float32)
        counts = np.zeros((view.height, view.width), np.uint32)
        batches = C.c_int32()
        err = C.create_string_buffer(512)
        rc = self._lib.hipdev_session_render(
            self.h, f3(*view.position), f3(*view.target), f3(*view.up), view.vfov, view.aspect, view.width,
            view.height, view.max_bounces, shadow_quality, filter_type, filter_width, spp, max_batch, ao_strength,
            1 if path_guiding else 0, _fp(mean),
            capi.ptr(counts, C.c_uint32), C.byref(batches), err, 512)
        return rc, mean, counts, batches.value, err.value.decode(errors="replace")

    def render_passes(self, view, spp, max_batch):
        """RenderSession with SessionConfig::passes -> (rc, beauty (H,W,3), films (28, H,W,3), counts, groups,
        error, crypto ranks (2, 3, H, W, 4))."""
        f3 = C.c_float * 3
        beauty = np.zeros((view.height, view.width, 3), np.float32)
        films = np.zeros((28, view.height, view.width, 3), np.float32)
        counts = np.zeros((view.height, view.width), np.uint32)
        crypto = np.zeros((2, 3, view.height, view.width, 4), np.float32)
        groups = C.c_int32()
        err = C.create_string_buffer(512)
        rc = self._lib.hipdev_session_render_passes(
            self.h, f3(*view.position), f3(*view.target), f3(*view.up), view.vfov, view.aspect, view.width,
            view.height, view.max_bounces, spp, max_batch, _fp(beauty), _fp(films), capi.ptr(counts, C.c_uint32),
            C.byref(groups), _fp(crypto), err, 512)
        return rc, beauty, films, counts, groups.value, err.value.decode(errors="replace"), crypto

    def render_aovs(self, view, spp, max_batch):
        """RenderSession with AOVs -> (rc, [beauty, albedo, normal, alpha] each (H,W,3), counts (H,W), error)."""
        f3 = C.c_float * 3
        imgs = [np.zeros((view.height, view.width, 3), np.float32) for _ in range(4)]
        counts = np.zeros((view.height, view.width), np.uint32)
        err = C.create_string_buffer(512)
        rc = self._lib.hipdev_session_render_aovs(
            self.h, f3(*view.position), f3(*view.target), f3(*view.up), view.vfov, view.aspect, view.width,
            view.height, view.max_bounces, spp, max_batch, *[_fp(a) for a in imgs], capi.ptr(counts, C.c_uint32),
            err, 512)
        return rc, imgs, counts, err.value.decode(errors="replace")

    def render_adaptive(self, view, spp, max_batch, threshold, min_spp):
        """RenderSession with adaptive sampling -> (rc, mean (H,W,3), counts (H,W), batches, error text)."""
        f3 = C.c_float * 3
        mean = np.zeros((view.height, view.width, 3), np.float32)
        counts = np.zeros((view.height, view.width), np.uint32)
        batches = C.c_int32()
        err = C.create_string_buffer(512)
        rc = self._lib.hipdev_session_render_adaptive(
            self.h, f3(*view.position), f3(*view.target), f3(*view.up), view.vfov, view.aspect, view.width,
            view.height, view.max_bounces, spp, max_batch, threshold, min_spp, _fp(mean), capi.ptr(counts, C.c_uint32),
            C.byref(batches), err, 512)
        return rc, mean, counts, batches.value, err.value.decode(errors="replace")

    def interactive(self, view, spp, cancel_spp, cancel_after_ms):
        """Interactive one-sample passes, then a cancelled Final batch -> dict(rc, mean, counts, pass_seconds,
        cancel_latency, cancel_completed, cancel_samples, error)."""
        f3 = C.c_float * 3
        mean = np.zeros((view.height, view.width, 3), np.float32)
        counts = np.zeros((view.height, view.width), np.uint32)
        passes = np.zeros(spp, np.float64)
        lat, done, samples = C.c_double(), C.c_int32(), C.c_int32()
        err = C.create_string_buffer(512)
        rc = self._lib.hipdev_session_interactive(
            self.h, f3(*view.position), f3(*view.target), f3(*view.up), view.vfov, view.aspect, view.width,
            view.height, view.max_bounces, spp, _fp(mean), capi.ptr(counts, C.c_uint32),
            passes.ctypes.data_as(C.POINTER(C.c_double)), cancel_spp, cancel_after_ms, C.byref(lat), C.byref(done),
            C.byref(samples), err, 512)
        return dict(rc=rc, mean=mean, counts=counts, pass_seconds=passes, cancel_latency=lat.value,
                    cancel_completed=done.value, cancel_samples=samples.value,
                    error=err.value.decode(errors="replace"))


def spectral(u, nd=1.5, abbe=30.0, wavelengths=()):
    """Oracle sampleWavelength(u) -> (lambda (n,), weight (n, 3)) and cauchyIor(nd, abbe, wl) -> (m,)."""
    u = np.ascontiguousarray(u, np.float32)
    wl = np.ascontiguousarray(wavelengths, np.float32)
    lam = np.zeros(u.shape[0], np.float32)
    w = np.zeros((u.shape[0], 3), np.float32)
    ior = np.zeros(wl.shape[0], np.float32)
    oracle().oracle_spectral(u.shape[0], _fp(u), _fp(lam), _fp(w), nd, abbe, wl.shape[0], _fp(wl), _fp(ior))
    return lam, w, ior


def thin_film(which: str, cos1, n1, n2, n3, thickness, lam):
    """thinFilmReflectance for arrays of cos, thickness, wavelength: which = "oracle" or "ref_strict"."""
    lib = oracle() if which == "oracle" else reference(strict=True)
    cos1 = np.ascontiguousarray(cos1, np.float32)
    t = np.ascontiguousarray(np.broadcast_to(thickness, cos1.shape), np.float32)
    la = np.ascontiguousarray(np.broadcast_to(lam, cos1.shape), np.float32)
    out = np.zeros(cos1.shape[0], np.float32)
    fn = lib.oracle_thin_film if which == "oracle" else lib.ref_thin_film
    fn(cos1.shape[0], _fp(cos1), n1, n2, float(np.real(n3)), float(np.imag(n3)), _fp(t), _fp(la), _fp(out))
    return out
