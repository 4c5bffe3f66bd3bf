"""Scene assembly with the reference's Scene semantics, flattened into a pe_scene_desc.

Scene::addShape / addLight / setEnvironment (engine/scene.cpp:18-104) are mirrored:
  * add_shape(mesh, material): a TriangleMesh; when the material emits, every triangle also becomes a
    TriangleLight appended to lights() (scene.cpp:48-62), so emitter index = base + primIndex;
  * add_light(AreaLight): appends the light, then two standalone black emissive triangles
    (p0, p1, p2) and (p0, p2, p3) with the light's normal as vertex normals and uv (0,0),(1,0),(1,1),(0,1)
    (scene.cpp:65-99);
  * set_environment(EnvironmentLight).
The result is consumed by libphoton_hip (pe_scene_upload) and by the test oracles, from the same bytes.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi


def disney(base_color, metallic=0.0, roughness=0.5, specular=0.5, **kw) -> capi.MaterialDesc:
    """DisneyMaterial(baseColor, metallic, roughness, specular) with the reference's defaults (disney.h:210-245)."""
    m = capi.MaterialDesc()
    m.type = capi.PE_MAT_DISNEY
    m.base_color[:] = list(map(float, base_color))
    m.metallic, m.roughness, m.specular = float(metallic), float(roughness), float(specular)
    m.clearcoat, m.clearcoat_roughness, m.coat_ior = 0.0, 0.03, 1.5
    m.coat_tint[:] = [1.0, 1.0, 1.0]
    m.ior, m.specular_transmission, m.anisotropy = 1.5, 0.0, 0.0
    m.sheen, m.sheen_roughness = 0.0, 0.3
    m.sheen_color[:] = [1.0, 1.0, 1.0]
    m.has_sheen_color = 0
    m.diffuse_roughness, m.diffuse_transmission, m.subsurface = 0.0, 0.0, 0.0
    m.subsurface_radius[:] = [1.0, 0.2, 0.1]
    m.subsurface_scale, m.subsurface_anisotropy = 0.05, 0.0
    m.edge_tint[:] = [1.0, 1.0, 1.0]
    m.thin_film_thickness, m.thin_film_ior = 0.0, 1.4
    m.absorption_color[:] = [1.0, 1.0, 1.0]
    m.absorption_distance = 0.0
    m.scatter_color[:] = [1.0, 1.0, 1.0]
    m.flake_density, m.flake_size, m.flake_roughness = 0.0, 0.004, 0.3   # setFlakes defaults (disney.h:237-240)
    m.flake_color[:] = [0.9, 0.9, 0.9]
    if "opacity" in kw:                      # setOpacity (disney.h:148)
        m.has_opacity = 1
    for k, v in kw.items():
        cur = getattr(m, k)
        if isinstance(cur, C.Array):
            cur[:] = list(map(float, v))
        else:
            setattr(m, k, v)
    return m


def dielectric(ior=1.5, tint=(1.0, 1.0, 1.0), roughness=0.0, abbe=0.0, absorption_color=None,
               absorption_distance=0.0, thin_walled=False, thin_film_thickness=0.0, scatter_density=0.0,
               scatter_color=(1.0, 1.0, 1.0), scatter_anisotropy=0.0) -> capi.MaterialDesc:
    """Dielectric(ior, tint, roughness) (dielectric.h:18-20) with setAbbe (:74), setAbsorption(color, distance)
    (:57) when absorption_color is given, setThinWalled (:43), setThinFilm(thickness nm) (:46) and
    setScattering(density, color, anisotropy) (:61)."""
    m = disney(tint)
    m.scatter_density = float(scatter_density)
    m.scatter_color[:] = list(map(float, scatter_color))
    m.scatter_anisotropy = float(scatter_anisotropy)
    m.type = capi.PE_MAT_DIELECTRIC
    m.ior, m.roughness, m.abbe = float(ior), float(roughness), float(abbe)
    if absorption_color is not None:
        m.absorption_color[:] = list(map(float, absorption_color))
        m.absorption_distance = float(absorption_distance)
    m.thin_walled = 1 if thin_walled else 0
    m.thin_film_thickness = float(thin_film_thickness)
    return m


def lambert(albedo, emission=(0.0, 0.0, 0.0)) -> capi.MaterialDesc:
    """Lambertian(albedo, emission) (lambertian.h:16)."""
    m = disney(albedo)
    m.type = capi.PE_MAT_LAMBERT
    m.emission[:] = list(map(float, emission))
    return m


def mirror(color=(0.95, 0.95, 0.95)) -> capi.MaterialDesc:
    """Mirror(reflectance) (mirror.h:14; the project format's default colour, project_io.cpp:167-168)."""
    m = disney(color)
    m.type = capi.PE_MAT_MIRROR
    return m


@dataclass
class Mesh:
    positions: np.ndarray                 # (V, 3) float32
    indices: np.ndarray                   # (T, 3) uint32
    normals: np.ndarray | None = None     # (V, 3) float32
    uvs: np.ndarray | None = None         # (V, 2) float32

    def __post_init__(self):
        self.positions = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1, 3)
        self.indices = np.ascontiguousarray(self.indices, dtype=np.uint32).reshape(-1, 3)
        if self.normals is not None:
            self.normals = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1, 3)
        if self.uvs is not None:
            self.uvs = np.ascontiguousarray(self.uvs, dtype=np.float32).reshape(-1, 2)

    @property
    def num_triangles(self) -> int:
        return int(self.indices.shape[0])


@dataclass
class AreaLight:
    """AreaLight(position, u, v, radiance) (area_light.h:17-26)."""
    position: tuple
    u: tuple
    v: tuple
    radiance: tuple
    visible_to_camera: bool = True
    light_group: str = ""      # Light::setLightGroup (light.h:85): the EXR group LG_<name>; "" = none
    # Light::setLightLink (light.h:23-32, :81): receiver and shadow-blocker node uids with capi.PE_LINK_* modes
    link_mode: int = 0
    link_uids: tuple = ()
    shadow_mode: int = 0
    shadow_uids: tuple = ()


@dataclass
class Environment:
    """EnvironmentLight(image, rotation, intensity) [+ setSun] (environment_light.h:52-62)."""
    rgb: np.ndarray | None               # (H, W, 3) float32, row 0 = zenith; None = flat grey
    rotation: float = 0.0
    intensity: float = 1.0
    sun_direction: tuple | None = None
    sun_cos_max: float = 1.0
    sun_radiance: tuple = (0.0, 0.0, 0.0)
    light_group: str = ""


def crypto_hash(name: str) -> int:
    """cryptoHashName: MurmurHash3_x86_32 of the UTF-8 name with seed 0 (Psyop Cryptomatte 1.2.0;
    core/image/cryptomatte.cpp:28-76)."""
    data = name.encode("utf-8")
    c1, c2, mask = 0xCC9E2D51, 0x1B873593, 0xFFFFFFFF
    rotl = lambda x, r: ((x << r) | (x >> (32 - r))) & mask  # noqa: E731
    h = 0
    n4 = len(data) // 4
    for i in range(n4):
        k = int.from_bytes(data[4 * i:4 * i + 4], "little")
        k = rotl((k * c1) & mask, 15)
        h ^= (k * c2) & mask
        h = (rotl(h, 13) * 5 + 0xE6546B64) & mask
    tail = data[4 * n4:]
    if tail:
        k = int.from_bytes(tail, "little")
        k = rotl((k * c1) & mask, 15)
        h ^= (k * c2) & mask
    h ^= len(data)
    h ^= h >> 16
    h = (h * 0x85EBCA6B) & mask
    h ^= h >> 13
    h = (h * 0xC2B2AE35) & mask
    return h ^ (h >> 16)


def crypto_hash_to_float(h: int) -> float:
    """cryptoHashToFloat (cryptomatte.cpp:99-109): an exponent of 0 or 255 has its lowest bit flipped, so every
    id is a finite, normal float32."""
    exp = (h >> 23) & 0xFF
    if exp in (0, 0xFF):
        h = (h & 0x80000000) | ((exp ^ 1) << 23) | (h & 0x7FFFFF)
    return float(np.array([h], np.uint32).view(np.float32)[0])


@dataclass
class _Shape:
    mesh: Mesh
    material: int
    emitter_base: int
    standalone: bool
    flags: int = 0
    node_uid: int = 0          # Scene::tagShape: the owning scene-graph node (object-ID pass), 0 = untagged
    name: str = ""             # Scene::setShapeName (cryptomatte object name), "" = unnamed


@dataclass
class Scene:
    materials: list = field(default_factory=list)
    shapes: list = field(default_factory=list)
    lights: list = field(default_factory=list)     # LightDesc in Scene::lights() order
    environment: Environment | None = None
    light_group_of: list = field(default_factory=list)  # per light: its group name ("" = none)
    link_of: dict = field(default_factory=dict)  # light index -> (mode, uids, shadow mode, shadow uids)
    material_names: dict = field(default_factory=dict)  # material index -> Scene::setMaterialName name
    # Scene::setBackground (engine/scene.h:24-28): capi.PE_BACKGROUND_* and the colour of the Color mode
    background_mode: int = 0
    background_color: tuple = (0.18, 0.18, 0.18)

    def add_material(self, m: capi.MaterialDesc, name: str = "") -> int:
        self.materials.append(m)
        if name:
            self.material_names[len(self.materials) - 1] = name
        return len(self.materials) - 1

    def add_shape(self, mesh: Mesh, material: int, node_uid: int = 0, name: str = "") -> None:
        mat = self.materials[material]
        emissive = any(float(x) != 0.0 for x in mat.emission)
        base = -1
        if emissive:
            base = len(self.lights)
            p = mesh.positions[mesh.indices]
            for t in range(mesh.num_triangles):
                L = capi.LightDesc()
                L.type = capi.PE_LIGHT_TRIANGLE
                L.p0[:], L.p1[:], L.p2[:] = p[t, 0].tolist(), p[t, 1].tolist(), p[t, 2].tolist()
                L.radiance[:] = list(mat.emission)
                L.visible_to_camera = 1
                self.lights.append(L)
                self.light_group_of.append("")
        self.shapes.append(_Shape(mesh, material, base, False, 0, int(node_uid), name))

    def add_light(self, light: AreaLight) -> None:
        index = len(self.lights)
        L = capi.LightDesc()
        L.type = capi.PE_LIGHT_AREA
        L.p0[:] = list(map(float, light.position))
        L.p1[:] = list(map(float, light.u))
        L.p2[:] = list(map(float, light.v))
        L.radiance[:] = list(map(float, light.radiance))
        L.visible_to_camera = 1 if light.visible_to_camera else 0
        self.lights.append(L)
        self.light_group_of.append(light.light_group)
        if light.link_mode or light.shadow_mode:
            self.link_of[index] = (int(light.link_mode), tuple(light.link_uids), int(light.shadow_mode),
                                   tuple(light.shadow_uids))
        # The quad inserted into the BVH as two black emissive Lambert triangles (scene.cpp:77-98).
        mat = self.add_material(lambert((0.0, 0.0, 0.0), light.radiance))
        p0 = np.asarray(light.position, dtype=np.float32)
        u = np.asarray(light.u, dtype=np.float32)
        v = np.asarray(light.v, dtype=np.float32)
        p1 = p0 + u
        p2 = p1 + v
        p3 = p0 + v
        # AreaLight ctor (area_light.h:21-24): u.cross(v).normalized() = n * (1 / sqrt(x*x + y*y + z*z)), each
        # operation one float32 rounding in this order.
        f = np.float32
        n = np.array([f(u[1] * v[2]) - f(u[2] * v[1]), f(u[2] * v[0]) - f(u[0] * v[2]),
                      f(u[0] * v[1]) - f(u[1] * v[0])], dtype=np.float32)
        inv = f(f(1.0) / np.sqrt(f(f(f(n[0] * n[0]) + f(n[1] * n[1])) + f(n[2] * n[2]))))
        n = (n * inv).astype(np.float32)
        flags = 0 if light.visible_to_camera else capi.PE_SHAPE_CAMERA_INVISIBLE
        for tri, uvs in (((p0, p1, p2), ((0, 0), (1, 0), (1, 1))), ((p0, p2, p3), ((0, 0), (1, 1), (0, 1)))):
            m = Mesh(np.stack(tri), np.array([[0, 1, 2]]), np.stack([n, n, n]), np.array(uvs, dtype=np.float32))
            self.shapes.append(_Shape(m, mat, index, True, flags))

    def set_environment(self, env: Environment) -> None:
        self.environment = env

    @property
    def num_triangles(self) -> int:
        return sum(s.mesh.num_triangles for s in self.shapes)

    def crypto_tables(self):
        """Renderer's cryptomatte lookup (buildCryptoLookup, engine/renderer.cpp:65-118): every named material
        registers its name; then each shape in order registers its name, or object_<node uid> when tagged, or
        shape_<k> (k counting the anonymous shapes), and its material, unless registered, as
        material_<k> (k counting the anonymous materials). Returns (object id per shape, material id per
        material (0: never registered), object manifest {name: hash}, material manifest)."""
        obj = np.zeros(len(self.shapes), np.float32)
        mat = np.zeros(len(self.materials), np.float32)
        obj_manifest, mat_manifest = {}, {}
        done = set()
        for m, name in self.material_names.items():
            if name:
                mat_manifest[name] = crypto_hash(name)
                mat[m] = crypto_hash_to_float(mat_manifest[name])
                done.add(m)
        anon_shape = anon_mat = 0
        for i, sh in enumerate(self.shapes):
            if sh.name:
                name = sh.name
            elif sh.node_uid != 0:
                name = f"object_{sh.node_uid}"
            else:
                name = f"shape_{anon_shape}"
                anon_shape += 1
            obj_manifest[name] = crypto_hash(name)
            obj[i] = crypto_hash_to_float(obj_manifest[name])
            if sh.material not in done:
                name = f"material_{anon_mat}"
                anon_mat += 1
                mat_manifest[name] = crypto_hash(name)
                mat[sh.material] = crypto_hash_to_float(mat_manifest[name])
                done.add(sh.material)
        return obj, mat, obj_manifest, mat_manifest

    def light_groups(self):
        """Scene::buildAccelerator's light-group tables (engine/scene.cpp:144-173): when any light or the
        environment names a group, every light belongs to its group or "Default" and the environment to its
        group or "Env", groups numbered in first-seen order (lights first) and capped at 16, a name past the cap
        falling into group 0. Returns (names, per-light index, environment index); no groups: ([], [-1, ...], -1)."""
        env = self.environment
        explicit = (env is not None and env.light_group != "") or any(g != "" for g in self.light_group_of)
        if not explicit:
            return [], [-1] * len(self.lights), -1
        names: list = []

        def group(name):
            if name in names:
                return names.index(name)
            if len(names) >= capi.PE_MAX_LIGHT_GROUPS:
                return 0
            names.append(name)
            return len(names) - 1

        per_light = [group(g if g else "Default") for g in self.light_group_of]
        env_index = group(env.light_group if env.light_group else "Env") if env is not None else -1
        return names, per_light, env_index

    def pack(self):
        """Returns (SceneDesc, keepalive). keepalive owns every buffer the descriptor points into."""
        keep = []
        meshes = (capi.MeshDesc * max(1, len(self.shapes)))()
        for i, s in enumerate(self.shapes):
            md = meshes[i]
            md.positions = capi.ptr(s.mesh.positions, C.c_float)
            md.normals = capi.ptr(s.mesh.normals, C.c_float)
            md.uvs = capi.ptr(s.mesh.uvs, C.c_float)
            md.indices = capi.ptr(s.mesh.indices, C.c_uint32)
            md.num_vertices = s.mesh.positions.shape[0]
            md.num_triangles = s.mesh.num_triangles
            md.material = s.material
            md.emitter_base = s.emitter_base
            md.standalone = 1 if s.standalone else 0
            md.flags = s.flags
            keep.append(s.mesh)
        mats = (capi.MaterialDesc * max(1, len(self.materials)))(*self.materials)
        lights = (capi.LightDesc * max(1, len(self.lights)))(*self.lights)
        desc = capi.SceneDesc()
        desc.meshes = C.cast(meshes, C.POINTER(capi.MeshDesc))
        desc.num_meshes = len(self.shapes)
        desc.materials = C.cast(mats, C.POINTER(capi.MaterialDesc))
        desc.num_materials = len(self.materials)
        desc.lights = C.cast(lights, C.POINTER(capi.LightDesc))
        desc.num_lights = len(self.lights)
        keep += [meshes, mats, lights]
        names, per_light, env_index = self.light_groups()
        if names:
            groups = np.ascontiguousarray(per_light + [0], dtype=np.int32)
            keep.append(groups)
            desc.light_groups = capi.ptr(groups, C.c_int32)
            desc.env_light_group = env_index
            desc.num_light_groups = len(names)
        else:
            desc.light_groups = C.cast(None, C.POINTER(C.c_int32))
            desc.env_light_group = -1
            desc.num_light_groups = 0
        uids = np.array([s.node_uid for s in self.shapes] + [0], dtype=np.uint64)
        keep.append(uids)
        desc.node_uids = capi.ptr(uids, C.c_uint64)
        desc.background_mode = int(self.background_mode)
        desc.background_color[:] = [float(x) for x in self.background_color]
        if self.link_of:
            modes, offs, luids = [], [], []
            for i in range(len(self.lights)):
                mode, rec, smode, sh = self.link_of.get(i, (0, (), 0, ()))
                modes += [mode, smode]
                offs += [len(luids), len(rec)]
                luids += list(rec)
                offs += [len(luids), len(sh)]
                luids += list(sh)
            arrs = (np.ascontiguousarray(modes + [0], np.int32), np.ascontiguousarray(offs + [0], np.uint32),
                    np.ascontiguousarray(luids + [0], np.uint64))
            keep.extend(arrs)
            desc.link_modes = capi.ptr(arrs[0], C.c_int32)
            desc.link_offsets = capi.ptr(arrs[1], C.c_uint32)
            desc.link_uids = capi.ptr(arrs[2], C.c_uint64)
        if self.environment is not None:
            e = self.environment
            env = capi.EnvDesc()
            if e.rgb is not None:
                rgb = np.ascontiguousarray(e.rgb, dtype=np.float32)
                keep.append(rgb)
                env.rgb = capi.ptr(rgb, C.c_float)
                env.height, env.width = int(rgb.shape[0]), int(rgb.shape[1])
            env.rotation, env.intensity = float(e.rotation), float(e.intensity)
            if e.sun_direction is not None:
                env.has_sun = 1
                env.sun_direction[:] = list(map(float, e.sun_direction))
                env.sun_cos_max = float(e.sun_cos_max)
                env.sun_radiance[:] = list(map(float, e.sun_radiance))
            keep.append(env)
            desc.environment = C.pointer(env)
        else:
            desc.environment = C.cast(None, C.POINTER(capi.EnvDesc))
        return desc, keep
