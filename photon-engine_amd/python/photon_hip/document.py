"""Scene documents for the HIP path: the reference's .photon projects and the scene they compile to, OBJ models,
.mesh binaries and Radiance HDR environment maps.

What each part follows:
  * load_project / load_project_json: loadProjectJson (scene/project_io.cpp:640-680) with materialFromJson
    (:148-186, the advanced keys of presetAdvancedFromJson, scene/material_library.cpp:219-260), nodeFromJson
    and applyCommon (:268-291, :388-410), meshFromJson (:293-320), the re-import of a node's source model
    (:324-386, scene/model_import.cpp:255-292 for OBJ), lightFromJson (:473-512) and environmentFromJson
    (:530-571);
  * build_render_scene: buildRenderScene (scene/document.cpp:315-352): SceneGraph::compileNode
    (scene/scene_graph.cpp:170-216) with every mesh node baked to world space as the HIP device bakes a
    MeshInstance, the ground quad under the graph's bounds (placeGroundUnder, scene_graph.cpp:151-166),
    buildLights (document.cpp:66-115), buildEnvironment (:215-230) with the procedural studio sky (:117-140)
    and the background;
  * load_obj: ObjLoader::load (io/obj_loader.cpp:73-225) over tinyobjloader's reading of the format (v / vt /
    vn / f with 1-based or negative indices, o / g starting shapes, usemtl / mtllib with Kd);
  * read_mesh_binary / write_mesh_binary: scene/mesh_binary.cpp (PHMESH01);
  * read_hdr: the Radiance RGBE format (flat and run-length scanlines) decoded as stb_image decodes it.

What the HIP path does not render is refused with DocumentError, not approximated: texture maps and
rounded edges; point, spot and directional lights and light linking;
the physical sky; shadow catchers; OpenEXR, glTF, FBX, Alembic and USD files. Analytic sphere nodes are refused
unless the caller asks for them tessellated (sphere_segments), as the reference's USD arm of its golden test
renders them.
"""
from __future__ import annotations

import json
import math
import os
import re
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import capi
from .configs import View
from .scene import AreaLight, Environment, Mesh, Scene, dielectric, disney, lambert, mirror


class DocumentError(ValueError):
    """A document the HIP path cannot load or render as written."""


VIS_ALL, VIS_CAMERA = 0x1F, 0x01
GROUND_NODE_UID = (1 << 64) - 1          # kGroundNodeUid (scene/document.h:154)
MAX_PROJECT_BYTES = 512 << 20            # kMaxProjectBytes (project_io.cpp:40)
MAX_OBJ_BYTES = 512 << 20                # obj_loader.cpp:23-26
MAX_OBJ_VERTICES = 50_000_000
MAX_OBJ_TRIANGLES = 50_000_000


# ------------------------------------------------------------------------------------------------ OBJ

@dataclass
class ObjModel:
    """ObjLoadResult: one mesh per (shape, material) bucket and, parallel to it, the bucket's Lambertian of
    the MTL's Kd or None (the caller's default material)."""
    meshes: list = field(default_factory=list)
    materials: list = field(default_factory=list)


def _file_within(path: Path, cap: int) -> bool:
    try:
        return path.is_file() and path.stat().st_size <= cap
    except OSError:
        return False


def _read_mtl(path: Path, names: dict, kd: list) -> None:
    """newmtl / Kd of an MTL file into the material table (first definition of a name wins)."""
    current = None
    for raw in path.read_text(errors="replace").splitlines():
        tok = raw.split()
        if not tok or tok[0].startswith("#"):
            continue
        if tok[0] == "newmtl" and len(tok) > 1:
            name = " ".join(tok[1:])
            if name in names:
                current = None
                continue
            names[name] = len(kd)
            kd.append([1.0, 1.0, 1.0])
            current = names[name]
        elif tok[0] == "Kd" and current is not None and len(tok) >= 2:
            vals = [float(x) for x in tok[1:4]]
            while len(vals) < 3:
                vals.append(vals[-1])
            kd[current] = vals


def load_obj(path) -> ObjModel:
    """The OBJ file as ObjLoader::load reads it: n-gons fanned from their first corner, corners not welded, v
    flipped to image rows, one mesh per material bucket of each shape, a Lambertian of Kd per MTL material.
    A missing or oversized file, an index out of range, a non-finite position / normal / uv, or counts past the
    limits give an empty result."""
    path = Path(path)
    empty = ObjModel()
    if not _file_within(path, MAX_OBJ_BYTES):
        return empty
    text = path.read_text(errors="replace")
    # The referenced MTL files must be regular files within the size limit too (objInputsOk, :36-58).
    for line in text.splitlines():
        tok = line.split()
        if tok and tok[0] == "mtllib":
            for name in tok[1:]:
                mtl = Path(name) if os.path.isabs(name) else path.parent / name
                if mtl.exists() and not _file_within(mtl, MAX_OBJ_BYTES):
                    return empty
    v, vn, vt = [], [], []
    mat_names: dict = {}
    kd: list = []
    shapes = []                 # per shape: list of (material id, [corners of the face])
    faces = []
    current_mat = -1
    for raw in text.splitlines():
        tok = raw.split()
        if not tok or tok[0].startswith("#"):
            continue
        key = tok[0]
        try:
            if key == "v":
                xyz = [float(x) for x in tok[1:4]] + [0.0] * max(0, 4 - len(tok))
                v.append(xyz[:3])
            elif key == "vn":
                xyz = [float(x) for x in tok[1:4]] + [0.0] * max(0, 4 - len(tok))
                vn.append(xyz[:3])
            elif key == "vt":
                uv = [float(x) for x in tok[1:3]] + [0.0] * max(0, 3 - len(tok))
                vt.append(uv[:2])
            elif key == "f":
                corners = []
                for c in tok[1:]:
                    parts = c.split("/")
                    idx = []
                    for k, count in enumerate((len(v), len(vt), len(vn))):
                        s = parts[k] if k < len(parts) else ""
                        if s == "":
                            idx.append(-1)
                            continue
                        i = int(s)
                        # 1-based, or negative counting back from the elements read so far
                        idx.append(i - 1 if i > 0 else (count + i if i < 0 else -2))
                    corners.append(tuple(idx))
                faces.append((current_mat, corners))
            elif key in ("o", "g"):
                if faces:
                    shapes.append(faces)
                    faces = []
            elif key == "usemtl":
                current_mat = mat_names.get(" ".join(tok[1:]), -1)
            elif key == "mtllib":
                for name in tok[1:]:
                    mtl = Path(name) if os.path.isabs(name) else path.parent / name
                    if mtl.is_file():
                        _read_mtl(mtl, mat_names, kd)
        except ValueError:
            return empty
    if faces:
        shapes.append(faces)
    if len(v) > MAX_OBJ_VERTICES or len(vn) > MAX_OBJ_VERTICES or len(vt) > MAX_OBJ_VERTICES:
        return empty
    V = np.asarray(v, np.float32).reshape(-1, 3)
    VN = np.asarray(vn, np.float32).reshape(-1, 3)
    VT = np.asarray(vt, np.float32).reshape(-1, 2)
    out = ObjModel()
    tris = 0
    for shape in shapes:
        buckets: dict = {}      # material id -> (positions, normals, uvs), first-seen order
        for mat, corners in shape:
            fv = len(corners)
            if fv < 3:
                continue
            tris += fv - 2
            if tris > MAX_OBJ_TRIANGLES:
                return empty
            b = buckets.setdefault(mat, ([], [], []))
            for i in range(1, fv - 1):
                for c in (corners[0], corners[i], corners[i + 1]):
                    pi, ti, ni = c
                    if not 0 <= pi < len(V):
                        return empty
                    p = V[pi]
                    if not np.isfinite(p).all():
                        return empty
                    b[0].append(p)
                    if ni >= 0 or ni == -2:
                        if not 0 <= ni < len(VN) or not np.isfinite(VN[ni]).all():
                            return empty
                        b[1].append(VN[ni])
                    if ti >= 0 or ti == -2:
                        if not 0 <= ti < len(VT) or not np.isfinite(VT[ti]).all():
                            return empty
                        b[2].append((VT[ti][0], 1.0 - VT[ti][1]))
        for mat, (pos, nrm, uv) in buckets.items():
            if not pos:
                continue
            n = len(pos)
            mesh = Mesh(np.asarray(pos, np.float32), np.arange(n, dtype=np.uint32).reshape(-1, 3),
                        np.asarray(nrm, np.float32) if len(nrm) == n else None,
                        np.asarray(uv, np.float32) if len(uv) == n else None)
            out.meshes.append(mesh)
            out.materials.append(lambert(kd[mat]) if 0 <= mat < len(kd) else None)
    return out


# ------------------------------------------------------------------------------------------- .mesh

_MESH_MAGIC = b"PHMESH01"
_HAS_NORMALS, _HAS_UVS = 1, 2


def write_mesh_binary(mesh: Mesh) -> bytes:
    """encodeMeshBinary: magic, vertex count, index count, flags, 0, positions, normals, uvs, indices (LE)."""
    n = mesh.positions.shape[0]
    flags = (_HAS_NORMALS if mesh.normals is not None and len(mesh.normals) == n and n else 0) | \
            (_HAS_UVS if mesh.uvs is not None and len(mesh.uvs) == n and n else 0)
    parts = [_MESH_MAGIC, struct.pack("<4I", n, mesh.indices.size, flags, 0),
             np.ascontiguousarray(mesh.positions, "<f4").tobytes()]
    if flags & _HAS_NORMALS:
        parts.append(np.ascontiguousarray(mesh.normals, "<f4").tobytes())
    if flags & _HAS_UVS:
        parts.append(np.ascontiguousarray(mesh.uvs, "<f4").tobytes())
    parts.append(np.ascontiguousarray(mesh.indices, "<u4").tobytes())
    return b"".join(parts)


def read_mesh_binary(data: bytes) -> Mesh:
    """decodeMeshBinary: the size must be exactly what the header implies, positions finite, indices in
    range."""
    if len(data) < 24 or data[:8] != _MESH_MAGIC:
        raise DocumentError("not a .mesh file")
    vc, ic, flags, _ = struct.unpack_from("<4I", data, 8)
    expect = 24 + vc * 12 + (vc * 12 if flags & _HAS_NORMALS else 0) + (vc * 8 if flags & _HAS_UVS else 0) + ic * 4
    if vc == 0 or ic == 0 or ic % 3 or expect != len(data):
        raise DocumentError("corrupt .mesh file (size)")
    q = 24
    pos = np.frombuffer(data, "<f4", vc * 3, q).reshape(-1, 3)
    q += vc * 12
    if not np.isfinite(pos).all():
        raise DocumentError("corrupt .mesh file (non-finite position)")
    nrm = uv = None
    if flags & _HAS_NORMALS:
        nrm = np.frombuffer(data, "<f4", vc * 3, q).reshape(-1, 3)
        q += vc * 12
    if flags & _HAS_UVS:
        uv = np.frombuffer(data, "<f4", vc * 2, q).reshape(-1, 2)
        q += vc * 8
    idx = np.frombuffer(data, "<u4", ic, q)
    if (idx >= vc).any():
        raise DocumentError("corrupt .mesh file (index out of range)")
    return Mesh(pos.copy(), idx.reshape(-1, 3).copy(), None if nrm is None else nrm.copy(),
                None if uv is None else uv.copy())


# --------------------------------------------------------------------------------------------- HDR

def _rgbe_to_float(rgbe: np.ndarray) -> np.ndarray:
    e = rgbe[..., 3].astype(np.int32)
    scale = np.where(e > 0, np.ldexp(np.float32(1.0), e - 136), 0.0).astype(np.float32)
    return (rgbe[..., :3].astype(np.float32) * scale[..., None]).astype(np.float32)


def read_hdr(path) -> np.ndarray:
    """A Radiance .hdr file (32-bit_rle_rgbe, rows top to bottom, "-Y h +X w") as (h, w, 3) float32: each
    byte m with exponent byte e gives m * 2^(e - 136), e = 0 black (stb_image's decoding)."""
    data = Path(path).read_bytes()
    if not (data.startswith(b"#?RADIANCE") or data.startswith(b"#?RGBE")):
        raise DocumentError(f"{path}: not a Radiance HDR file")
    p = 0
    fmt_ok = False
    while True:
        e = data.index(b"\n", p)
        line = data[p:e].strip()
        p = e + 1
        if not line:
            break
        if line.startswith(b"FORMAT="):
            fmt_ok = line == b"FORMAT=32-bit_rle_rgbe"
            if not fmt_ok:
                raise DocumentError(f"{path}: unsupported HDR format {line!r}")
    e = data.index(b"\n", p)
    m = re.fullmatch(rb"-Y (\d+) \+X (\d+)", data[p:e].strip())
    if not m:
        raise DocumentError(f"{path}: unsupported HDR orientation {data[p:e]!r}")
    h, w = int(m.group(1)), int(m.group(2))
    p = e + 1
    buf = np.frombuffer(data, np.uint8, offset=p)
    out = np.zeros((h, w, 4), np.uint8)
    q = 0
    for y in range(h):
        rle = (8 <= w < 32768 and q + 4 <= buf.size and buf[q] == 2 and buf[q + 1] == 2 and
               (int(buf[q + 2]) << 8 | int(buf[q + 3])) == w and not buf[q + 2] & 0x80)
        if not rle:
            n = w * 4
            if q + n > buf.size:
                raise DocumentError(f"{path}: truncated HDR data")
            out[y] = buf[q:q + n].reshape(w, 4)
            q += n
            continue
        q += 4
        for c in range(4):
            x = 0
            while x < w:
                if q >= buf.size:
                    raise DocumentError(f"{path}: truncated HDR data")
                count = int(buf[q])
                q += 1
                if count > 128:
                    count -= 128
                    if x + count > w or q >= buf.size:
                        raise DocumentError(f"{path}: bad HDR run")
                    out[y, x:x + count, c] = buf[q]
                    q += 1
                else:
                    if count == 0 or x + count > w or q + count > buf.size:
                        raise DocumentError(f"{path}: bad HDR run")
                    out[y, x:x + count, c] = buf[q:q + count]
                    q += count
                x += count
    return _rgbe_to_float(out)


def write_hdr(path, rgb: np.ndarray) -> None:
    """(h, w, 3) float radiance as a flat (uncompressed) Radiance file, each pixel rounded to RGBE."""
    rgb = np.asarray(rgb, np.float32)
    h, w, _ = rgb.shape
    m = rgb.max(-1)
    mant, ex = np.frexp(m)
    ok = m >= 1e-32
    scale = np.where(ok, mant * 256.0 / np.where(ok, m, 1.0), 0.0)
    rgbe = np.zeros((h, w, 4), np.uint8)
    rgbe[..., :3] = np.clip(rgb * scale[..., None], 0, 255).astype(np.uint8)
    rgbe[..., 3] = np.where(ok, ex + 128, 0).astype(np.uint8)
    header = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode()
    Path(path).write_bytes(header + rgbe.tobytes())


# ------------------------------------------------------------------------------------------ project

@dataclass
class Node:
    """A SceneNode: a group, a mesh (in its local coordinates) or an analytic sphere."""
    name: str
    kind: str                                       # "group" | "mesh" | "sphere"
    xf: np.ndarray = field(default_factory=lambda: np.eye(4, dtype=np.float32))  # local to parent, row-major
    visible: bool = True
    visibility: int = VIS_ALL
    material: int = -1                              # Project.materials index, -1 = none
    mesh: Mesh | None = None
    radius: float = 0.0
    children: list = field(default_factory=list)
    uid: int = 0
    source_index: int = -1


@dataclass
class LightSpec:
    """LightDesc (scene/document.h:26-51) as a project file gives it."""
    type: str = "area"
    name: str = "Light"
    enabled: bool = True
    color: tuple = (1.0, 1.0, 1.0)
    temperature: float = 0.0
    intensity: float = 10.0
    position: tuple = (0.0, 3.0, 0.0)
    target: tuple = (0.0, 0.0, 0.0)
    width: float = 1.0
    height: float = 1.0
    azimuth: float = 45.0
    elevation: float = 45.0
    visible_to_camera: bool = False
    cone_angle: float = 45.0
    cone_blend: float = 0.15
    ies: str = ""
    light_group: str = ""
    light_link: dict | None = None


@dataclass
class EnvironmentSpec:
    """EnvironmentDesc (scene/document.h:53-81) as a project file gives it."""
    kind: str = "image"                              # "image" | "sky"
    hdr: str = ""
    zenith: tuple = (0.75, 0.8, 0.95)
    horizon: tuple = (0.55, 0.55, 0.58)
    rotation_deg: float = 0.0
    intensity: float = 1.0
    background_mode: str = "environment"             # "environment" | "color" | "transparent"
    background_color: tuple = (0.18, 0.18, 0.18)
    ground_enabled: bool = True
    ground_color: tuple = (0.42, 0.42, 0.42)
    ground_roughness: float = 0.55
    ground_shadow_catcher: bool = False
    sky: dict = field(default_factory=dict)


@dataclass
class Project:
    materials: list = field(default_factory=list)           # capi.MaterialDesc
    unsupported: list = field(default_factory=list)         # per material: why the HIP path refuses it, or None
    nodes: list = field(default_factory=list)               # the root's children
    lights: list = field(default_factory=list)              # LightSpec
    environment: EnvironmentSpec = field(default_factory=EnvironmentSpec)
    camera: dict = field(default_factory=dict)
    render: dict = field(default_factory=dict)
    base_dir: Path = Path(".")
    warnings: list = field(default_factory=list)


def _vec3(j: dict, key: str, default) -> tuple:
    a = j.get(key)
    if not isinstance(a, list) or len(a) < 3:
        return tuple(float(x) for x in default)
    return tuple(float(np.float32(x)) for x in a[:3])


def _f(j: dict, key: str, default: float) -> float:
    return float(np.float32(j.get(key, default)))


def _resolve(p: str, base: Path) -> str:
    if not p:
        return ""
    q = Path(p)
    return str(q if q.is_absolute() else base / q)


def _material(j: dict, base: Path):
    """materialFromJson -> (MaterialDesc, reason the HIP path refuses it or None)."""
    t = j.get("type", "disney")
    why = []
    if _f(j, "roundEdges", 0.0) > 0.0:
        why.append("rounded edges")
    if t == "glass":
        m = dielectric(_f(j, "ior", 1.5), _vec3(j, "tint", (1, 1, 1)), _f(j, "roughness", 0.0))
        thin = bool(j.get("thinWalled", False))
        m.thin_walled = 1 if thin else 0
        film = j.get("thinFilm") if isinstance(j.get("thinFilm"), dict) else {}
        m.thin_film_thickness = _f(film, "thickness", 0.0) if thin else 0.0
        ab = j.get("absorption")
        if isinstance(ab, dict):
            m.absorption_color[:] = list(_vec3(ab, "color", (1, 1, 1)))
            m.absorption_distance = _f(ab, "distance", 1.0)
        sc = j.get("scattering")
        if isinstance(sc, dict):
            m.scatter_density = _f(sc, "density", 0.0)
            m.scatter_color[:] = list(_vec3(sc, "color", (1, 1, 1)))
            m.scatter_anisotropy = _f(sc, "anisotropy", 0.0)
        m.abbe = _f(j, "abbe", 0.0)
        return m, ", ".join(why) or None
    if t == "lambert":
        return lambert(_vec3(j, "albedo", (0.8, 0.8, 0.8)), _vec3(j, "emission", (0, 0, 0))), None
    if t == "mirror":
        return mirror(_vec3(j, "color", (0.95, 0.95, 0.95))), None
    m = disney(_vec3(j, "baseColor", (0.8, 0.8, 0.8)), _f(j, "metallic", 0.0), _f(j, "roughness", 0.5),
               _f(j, "specular", 0.5))
    m.clearcoat = _f(j, "clearCoat", 0.0)
    m.clearcoat_roughness = _f(j, "clearCoatRoughness", 0.03)
    m.anisotropy = _f(j, "anisotropy", 0.0)
    m.sheen = _f(j, "sheen", 0.0)
    m.diffuse_transmission = _f(j, "diffuseTransmission", 0.0)
    m.emission[:] = list(_vec3(j, "emission", (0, 0, 0)))
    if isinstance(j.get("maps"), dict) and j["maps"]:
        why.append("texture maps")
    if "opacity" in j:                                   # setOpacity, clamped by the material (disney.h:148)
        m.opacity = _f(j, "opacity", 1.0)
        m.has_opacity = 1
    m.ior = _f(j, "ior", 1.5)
    m.specular_transmission = _f(j, "specularTransmission", 0.0)
    m.coat_ior = _f(j, "coatIor", 1.5)
    m.coat_tint[:] = list(_vec3(j, "coatTint", (1, 1, 1)))
    if "sheenColor" in j:
        m.sheen_color[:] = list(_vec3(j, "sheenColor", (1, 1, 1)))
        m.has_sheen_color = 1
    ss = j.get("subsurface")
    if isinstance(ss, dict):
        m.subsurface = _f(ss, "weight", 0.0)
        m.subsurface_radius[:] = list(_vec3(ss, "radius", (1.0, 0.2, 0.1)))
        m.subsurface_scale = _f(ss, "scale", 0.05)
        m.subsurface_anisotropy = _f(ss, "anisotropy", 0.0)
    m.edge_tint[:] = list(_vec3(j, "edgeTint", (1, 1, 1)))
    m.diffuse_roughness = _f(j, "diffuseRoughness", 0.0)
    m.sheen_roughness = _f(j, "sheenRoughness", 0.3)
    film = j.get("thinFilm")
    if isinstance(film, dict):
        m.thin_film_thickness = _f(film, "thickness", 0.0)
        m.thin_film_ior = _f(film, "ior", 1.4)
    fl = j.get("flakes")
    if isinstance(fl, dict):  # material_library.cpp:254-259 (MaterialPreset defaults, material_library.h:73-76)
        m.flake_density = _f(fl, "density", 0.0)
        m.flake_size = _f(fl, "size", 0.004)
        m.flake_roughness = _f(fl, "roughness", 0.3)
        m.flake_color[:] = list(_vec3(fl, "color", (0.9, 0.9, 0.9)))
    return m, ", ".join(why) or None


def _transform(j: dict) -> np.ndarray:
    a = j.get("xf")
    if not isinstance(a, list) or len(a) != 16:
        return np.eye(4, dtype=np.float32)
    return np.asarray(a, np.float32).reshape(4, 4)


class _Loader:
    def __init__(self, base: Path, project: Project):
        self.base = base
        self.project = project
        self.next_uid = 2            # a fresh reference process gives the scene graph's root uid 1

    def uid(self) -> int:
        u = self.next_uid
        self.next_uid += 1
        return u

    def common(self, n: Node, j: dict) -> None:
        """applyCommon: name, visibility, transform and material (kept where the index is out of range)."""
        n.name = j.get("name", n.name)
        n.visible = bool(j.get("visible", True))
        n.visibility = int(j.get("visibility", VIS_ALL)) & VIS_ALL
        n.xf = _transform(j)
        mi = int(j.get("material", -1))
        if 0 <= mi < len(self.project.materials):
            n.material = mi

    def mesh(self, g: dict) -> Mesh | None:
        if isinstance(g.get("bin"), str):
            path = _resolve(g["bin"], self.base)
            try:
                return read_mesh_binary(Path(path).read_bytes())
            except (OSError, DocumentError) as e:
                self.project.warnings.append(f"{e}: {g['bin']}")
                return None
        p = np.asarray(g.get("p", []), np.float32)
        p = p[:p.size // 3 * 3].reshape(-1, 3)
        idx = np.asarray(g.get("i", []), np.int64)
        if p.shape[0] == 0 or idx.size % 3 or (idx.size and (idx.min() < 0 or idx.max() >= p.shape[0])):
            return None
        nrm = np.asarray(g.get("n", []), np.float32)
        nrm = nrm[:nrm.size // 3 * 3].reshape(-1, 3)
        uv = np.asarray(g.get("uv", []), np.float32)
        uv = uv[:uv.size // 2 * 2].reshape(-1, 2)
        return Mesh(p, idx.astype(np.uint32).reshape(-1, 3), nrm if len(nrm) == len(p) else None,
                    uv if len(uv) == len(p) else None)

    def source(self, j: dict) -> Node | None:
        """An imported model's group: the file is read again and the saved tree rebuilt over its parts."""
        src = _resolve(j.get("source", ""), self.base)
        ext = Path(src).suffix.lower()
        if ext != ".obj":
            raise DocumentError(f"{src}: only OBJ models are read here ({ext or 'no extension'})")
        if not Path(src).exists():
            self.project.warnings.append(f"file not found: {src}")
            return None
        model = load_obj(src)
        if not model.meshes:
            self.project.warnings.append(f"no triangles in {src}")
            return None
        group = Node(Path(src).stem, "group", uid=self.uid())
        default = None
        pool = {}
        for i, (mesh, mat) in enumerate(zip(model.meshes, model.materials)):
            if mat is None:
                if default is None:       # importModelFile's default: DisneyMaterial(0.72, 0, 0.4, 0.5)
                    default = self.add_material(disney((0.72, 0.72, 0.72), 0.0, 0.4, 0.5))
                mi = default
            else:
                mi = self.add_material(mat)
            pool[i] = Node(f"Parça {i + 1}", "mesh", material=mi, mesh=mesh, uid=self.uid(), source_index=i)
        self.common(group, j)
        self.rebuild(j.get("children", []), group, pool)
        return group

    def add_material(self, m) -> int:
        self.project.materials.append(m)
        self.project.unsupported.append(None)
        return len(self.project.materials) - 1

    def rebuild(self, children: list, parent: Node, pool: dict) -> None:
        for cj in children:
            k = int(cj.get("sourceIndex", -1))
            if k >= 0:
                node = pool.pop(k, None)
                if node is None:                 # the part is gone from the file: its embedded copy, if any
                    extra = self.node(cj)
                    if extra is not None:
                        parent.children.append(extra)
                    continue
                self.common(node, cj)
                parent.children.append(node)
                self.rebuild(cj.get("children", []), node, pool)
            elif "source" not in cj and cj.get("type", "group") == "group":
                g = Node("Grup", "group", uid=self.uid())
                self.common(g, cj)
                parent.children.append(g)
                self.rebuild(cj.get("children", []), g, pool)
            else:
                extra = self.node(cj)
                if extra is not None:
                    parent.children.append(extra)

    def node(self, j: dict) -> Node | None:
        if "source" in j:
            return self.source(j)
        t = j.get("type", "group")
        if t == "mesh":
            n = Node("Mesh", "mesh", uid=self.uid())
            self.common(n, j)
            if isinstance(j.get("geometry"), dict):
                n.mesh = self.mesh(j["geometry"])
            if n.mesh is None:
                return None
        elif t == "sphere":
            n = Node("Küre", "sphere", uid=self.uid())
            self.common(n, j)
            n.radius = _f(j, "radius", 1.0)
        else:
            n = Node("Grup", "group", uid=self.uid())
            self.common(n, j)
        for c in j.get("children", []):
            child = self.node(c)
            if child is not None:
                n.children.append(child)
        return n


def _light(j: dict, base: Path) -> LightSpec:
    d = LightSpec()
    t = j.get("type", "area")
    d.type = t if t in ("directional", "point", "spot") else "area"
    d.name = j.get("name", d.name)
    d.enabled = bool(j.get("enabled", True))
    d.color = _vec3(j, "color", d.color)
    d.temperature = float(np.clip(_f(j, "temperature", 0.0), 0.0, 25000.0))
    d.intensity = _f(j, "intensity", d.intensity)
    d.position = _vec3(j, "position", d.position)
    d.target = _vec3(j, "target", d.target)
    d.width = _f(j, "width", d.width)
    d.height = _f(j, "height", d.height)
    d.azimuth = _f(j, "azimuth", d.azimuth)
    d.elevation = _f(j, "elevation", d.elevation)
    d.visible_to_camera = bool(j.get("visibleToCamera", True))   # projects before 0.2 had visible area lights
    d.cone_angle = _f(j, "coneAngle", d.cone_angle)
    d.cone_blend = _f(j, "coneBlend", d.cone_blend)
    d.ies = _resolve(j.get("ies", ""), base)
    d.light_group = j.get("lightGroup", "")
    ll = j.get("lightLink")
    if isinstance(ll, dict):
        d.light_link = ll
    return d


def _environment(j: dict, base: Path) -> EnvironmentSpec:
    e = EnvironmentSpec()
    e.hdr = _resolve(j.get("hdr", ""), base)
    e.zenith = _vec3(j, "zenith", e.zenith)
    e.horizon = _vec3(j, "horizon", e.horizon)
    e.rotation_deg = _f(j, "rotation", 0.0)
    e.intensity = _f(j, "intensity", 1.0)
    b = j.get("background")
    if isinstance(b, dict):
        mode = b.get("mode", "environment")
        e.background_mode = mode if mode in ("color", "transparent") else "environment"
        e.background_color = _vec3(b, "color", e.background_color)
    g = j.get("ground")
    if isinstance(g, dict):
        e.ground_enabled = bool(g.get("enabled", True))
        e.ground_color = _vec3(g, "color", e.ground_color)
        e.ground_roughness = _f(g, "roughness", e.ground_roughness)
        e.ground_shadow_catcher = bool(g.get("shadowCatcher", False))
    if j.get("kind", "image") == "sky":
        e.kind = "sky"
        e.sky = dict(j.get("sky", {}))
    return e


def load_project_json(text: str, base_dir) -> Project:
    """loadProjectJson: version 2 or later; materials, then the node tree, camera, render settings,
    environment and lights."""
    try:
        root = json.loads(text)
    except json.JSONDecodeError as e:
        raise DocumentError(f"invalid project file (JSON: {e})") from None
    if not isinstance(root, dict):
        raise DocumentError("invalid project file (not a JSON object)")
    if int(root.get("version", 0)) < 2:
        raise DocumentError("old project format (version 1) is not supported")
    base = Path(base_dir)
    pr = Project(base_dir=base)
    for mj in root.get("materials", []):
        m, why = _material(mj, base)
        pr.materials.append(m)
        pr.unsupported.append(why)
    loader = _Loader(base, pr)
    for nj in root.get("nodes", []):
        n = loader.node(nj)
        if n is not None:
            pr.nodes.append(n)
    pr.camera = dict(root.get("camera", {}))
    pr.render = dict(root.get("render", {}))
    if "environment" in root:
        pr.environment = _environment(root["environment"], base)
    pr.lights = [_light(lj, base) for lj in root.get("lights", [])]
    return pr


def load_project(path) -> Project:
    """loadProject: the file (not empty, at most 512 MiB) read from its folder."""
    path = Path(path)
    try:
        size = path.stat().st_size
    except OSError:
        size = 0
    if size == 0 or size > MAX_PROJECT_BYTES:
        raise DocumentError(f"project file empty, missing or too large: {path}")
    return load_project_json(path.read_text(encoding="utf-8"), path.resolve().parent)


# ----------------------------------------------------------------------------------------- compile

def _smooth_normals(p: np.ndarray, idx: np.ndarray) -> np.ndarray:
    """TriangleMesh::calculateNormals (geometry/mesh.cpp:34-55): face cross products summed per vertex, then
    normalised (float32, triangle order)."""
    acc = np.zeros_like(p, dtype=np.float32)
    t = p[idx]
    fn = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]).astype(np.float32)
    for k in range(3):
        np.add.at(acc, idx[:, k], fn)
    ln = np.linalg.norm(acc, axis=1, keepdims=True).astype(np.float32)
    return np.where(ln > 1e-20, acc / np.where(ln > 1e-20, ln, 1.0), 0.0).astype(np.float32)


def bake(mesh: Mesh, world: np.ndarray) -> Mesh:
    """A mesh node under its world matrix, as the HIP device bakes a MeshInstance: points through the matrix,
    the local (or computed) normals through its inverse transpose and normalised, the winding turned under a
    negative determinant."""
    M = np.asarray(world, np.float64)
    p = mesh.positions.astype(np.float64) @ M[:3, :3].T + M[:3, 3]
    n_local = mesh.normals if mesh.normals is not None else _smooth_normals(mesh.positions, mesh.indices)
    n = n_local.astype(np.float64) @ np.linalg.inv(M[:3, :3])
    ln = np.linalg.norm(n, axis=1, keepdims=True)
    n = np.where(ln > 0, n / np.where(ln > 0, ln, 1.0), n)
    idx = mesh.indices if np.linalg.det(M[:3, :3]) >= 0 else mesh.indices[:, [0, 2, 1]]
    return Mesh(p.astype(np.float32), idx.copy(), n.astype(np.float32),
                None if mesh.uvs is None else mesh.uvs.copy())


def uv_sphere(center, radius: float, lon: int = 128, lat: int = 64) -> Mesh:
    """A UV sphere of lon x lat segments with smooth normals (the tessellation the reference's USD importer
    gives a UsdGeomSphere, 128 segments); triangles counter-clockwise seen from outside, row by row, each
    quad's upper triangle before its lower one, the poles' degenerate halves left out."""
    th = np.linspace(0.0, np.pi, lat + 1)
    ph = np.linspace(0.0, 2.0 * np.pi, lon + 1)
    T, P = np.meshgrid(th, ph, indexing="ij")
    n = np.stack([np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)], -1).reshape(-1, 3)
    pos = (np.asarray(center, np.float64) + radius * n).astype(np.float32)
    uv = np.stack([P / (2 * np.pi), T / np.pi], -1).reshape(-1, 2).astype(np.float32)
    k = np.arange(lon)
    rows = []
    for r in range(lat):
        a = r * (lon + 1) + k
        b = a + lon + 1
        quads = []
        if r > 0:
            quads.append(np.stack([a, a + 1, b], -1))
        if r < lat - 1:
            quads.append(np.stack([a + 1, b + 1, b], -1))
        rows.append(np.stack(quads, 1).reshape(-1, 3))
    return Mesh(pos, np.concatenate(rows).astype(np.uint32), n.astype(np.float32), uv)


def _world_bounds(nodes: list, parent: np.ndarray):
    lo, hi = np.full(3, np.inf), np.full(3, -np.inf)
    for n in nodes:
        if not n.visible:
            continue
        W = parent @ n.xf.astype(np.float64)
        if n.mesh is not None:
            p = n.mesh.positions.astype(np.float64) @ W[:3, :3].T + W[:3, 3]
            lo, hi = np.minimum(lo, p.min(0)), np.maximum(hi, p.max(0))
        elif n.kind == "sphere" and n.radius > 0:
            c = W[:3, 3]
            lo, hi = np.minimum(lo, c - n.radius), np.maximum(hi, c + n.radius)
        clo, chi = _world_bounds(n.children, W)
        lo, hi = np.minimum(lo, clo), np.maximum(hi, chi)
    return lo, hi


def ground_quad(lo, hi) -> tuple:
    """placeGroundUnder: a square 200 times the scene's largest extent, centred under it, a hair below its
    lowest point -> (corner, edge u, edge v)."""
    lo = np.asarray(lo, np.float32)
    hi = np.asarray(hi, np.float32)
    d = hi - lo
    span = np.float32(max(d[0], d[2], d[1], 1e-3))
    pad = span * np.float32(100.0)
    cx = (lo[0] + hi[0]) * np.float32(0.5)
    cz = (lo[2] + hi[2]) * np.float32(0.5)
    corner = np.array([cx - pad, lo[1] - span * np.float32(0.0005), cz - pad], np.float32)
    return corner, np.array([pad * 2, 0, 0], np.float32), np.array([0, 0, pad * 2], np.float32)


def temperature_color(kelvin: float) -> tuple:
    """temperatureColor (core/image/imager.cpp:204-213): the illuminant's xy (Kim et al. 2002 cubic fits of
    the Planckian locus below 3500 K, the CIE daylight locus above 4500 K, blended linearly between), to
    linear sRGB through the D65 XYZ matrix, negatives clipped, scaled to luminance 1."""
    T = float(np.clip(kelvin, 1000.0, 25000.0))

    def planck(t):
        t = min(max(t, 1667.0), 25000.0)
        x = (-0.2661239e9 / t ** 3 - 0.2343589e6 / t ** 2 + 0.8776956e3 / t + 0.179910 if t <= 4000.0 else
             -3.0258469e9 / t ** 3 + 2.1070379e6 / t ** 2 + 0.2226347e3 / t + 0.240390)
        if t <= 2222.0:
            y = -1.1063814 * x ** 3 - 1.34811020 * x ** 2 + 2.18555832 * x - 0.20219683
        elif t <= 4000.0:
            y = -0.9549476 * x ** 3 - 1.37418593 * x ** 2 + 2.09137015 * x - 0.16748867
        else:
            y = 3.0817580 * x ** 3 - 5.87338670 * x ** 2 + 3.75112997 * x - 0.37001483
        return np.float32(x), np.float32(y)

    def daylight(t):
        t = min(max(t, 4000.0), 25000.0)
        x = (-4.6070e9 / t ** 3 + 2.9678e6 / t ** 2 + 0.09911e3 / t + 0.244063 if t <= 7000.0 else
             -2.0064e9 / t ** 3 + 1.9018e6 / t ** 2 + 0.24748e3 / t + 0.237040)
        return np.float32(x), np.float32(-3.0 * x * x + 2.870 * x - 0.275)

    (xp, yp), (xd, yd) = planck(T), daylight(T)
    k = np.float32(min(max((T - 3500.0) / 1000.0, 0.0), 1.0))
    x, y = xp + (xd - xp) * k, yp + (yd - yp) * k
    X, Y, Z = x / y, np.float32(1.0), (np.float32(1.0) - x - y) / y
    M = np.array([[3.2404542, -1.5371385, -0.4985314], [-0.9692660, 1.8760108, 0.0415560],
                  [0.0556434, -0.2040259, 1.0572252]], np.float32)
    rgb = np.maximum(np.float32(0.0), M @ np.array([X, Y, Z], np.float32))
    lum = np.float32(0.212671) * rgb[0] + np.float32(0.715160) * rgb[1] + np.float32(0.072169) * rgb[2]
    return tuple(float(c) for c in (rgb / lum if lum > 0 else np.ones(3, np.float32)))


def area_light(d: LightSpec) -> AreaLight:
    """buildLights' area arm: a width x height quad centred on the light's position, facing its target."""
    pos = np.asarray(d.position, np.float32)
    n = np.asarray(d.target, np.float32) - pos
    if float(n @ n) < 1e-12:
        n = np.array([0.0, -1.0, 0.0], np.float32)
    n = n / np.float32(np.linalg.norm(n))
    up = np.array([1.0, 0.0, 0.0] if abs(n[1]) > 0.95 else [0.0, 1.0, 0.0], np.float32)
    u = np.cross(up, n)
    u = u / np.float32(np.linalg.norm(u))
    v = np.cross(n, u)
    v = v / np.float32(np.linalg.norm(v))
    u = u * np.float32(max(1e-4, d.width))
    v = v * np.float32(max(1e-4, d.height))
    corner = pos - u * np.float32(0.5) - v * np.float32(0.5)
    color = temperature_color(d.temperature) if d.temperature > 0.0 else d.color
    L = tuple(float(np.float32(c) * np.float32(d.intensity)) for c in color)
    return AreaLight(tuple(map(float, corner)), tuple(map(float, u)), tuple(map(float, v)), L,
                     visible_to_camera=d.visible_to_camera, light_group=d.light_group)


def procedural_sky(zenith, horizon) -> np.ndarray:
    """makeProceduralSky: a 512 x 256 vertical gradient (zenith to horizon, then darkening to 35%) with two
    bright windows."""
    W, H = 512, 256
    t = (np.arange(H, dtype=np.float32) / np.float32(H - 1))[:, None]
    z, h = np.asarray(zenith, np.float32), np.asarray(horizon, np.float32)
    k1 = t * 2
    k2 = (t - 0.5) * 2
    col = np.where(t < 0.5, z * (1 - k1) + h * k1, h * (1 - 0.65 * k2)).astype(np.float32)
    img = np.repeat(col[:, None, :], W, axis=1)
    img[H * 14 // 100:H * 34 // 100, W * 40 // 100:W * 56 // 100] = (26.0, 25.0, 23.0)
    img[H * 22 // 100:H * 40 // 100, W * 6 // 100:W * 14 // 100] = (9.0, 10.0, 12.0)
    return img.astype(np.float32)


def build_render_scene(pr: Project, clay: bool = False, sphere_segments: tuple | None = None) -> Scene:
    """buildRenderScene: the visible nodes compiled in tree order (mesh nodes baked to world space, named
    after their node, their material named <node>_Mat by the first node that uses it), the ground quad, the
    enabled lights of positive intensity, the environment of positive intensity and the background."""
    sc = Scene()
    mat_index: dict = {}

    def material(mi: int, node: Node) -> int:
        if mi in mat_index:
            return mat_index[mi]
        if mi >= 0 and pr.unsupported[mi]:
            raise DocumentError(f"material {mi} of node {node.name!r}: {pr.unsupported[mi]} "
                                "not implemented on the HIP path")
        m = clay_mat if mi == -2 else pr.materials[mi]
        mat_index[mi] = sc.add_material(m, node.name + "_Mat")
        return mat_index[mi]

    clay_mat = disney((0.6, 0.6, 0.6), 0.0, 0.6, 0.5) if clay else None

    def compile_node(n: Node, parent: np.ndarray) -> None:
        if not n.visible:
            return
        W = parent @ n.xf.astype(np.float64)
        mi = -2 if clay else n.material
        if n.kind == "mesh" and n.mesh is not None:
            if mi == -1:
                raise DocumentError(f"mesh node {n.name!r} has no material")
            if n.visibility not in (VIS_ALL, VIS_ALL & ~VIS_CAMERA):
                raise DocumentError(f"node {n.name!r}: visibility mask {n.visibility:#x} not implemented on the "
                                    "HIP path")
            sc.add_shape(bake(n.mesh, W), material(mi, n), node_uid=n.uid, name=n.name)
            if n.visibility != VIS_ALL:
                sc.shapes[-1].flags |= capi.PE_SHAPE_CAMERA_INVISIBLE
        elif n.kind == "sphere" and mi != -1 and n.radius > 0:
            if sphere_segments is None:
                raise DocumentError(f"sphere node {n.name!r}: analytic spheres are not implemented on the HIP "
                                    "path (ask for them tessellated with sphere_segments)")
            s = max(np.linalg.norm(W[:3, k]) for k in range(3))
            sc.add_shape(uv_sphere(W[:3, 3], n.radius * s, *sphere_segments), material(mi, n), node_uid=n.uid,
                         name=n.name)
        for c in n.children:
            compile_node(c, W)

    for n in pr.nodes:
        compile_node(n, np.eye(4))
    env = pr.environment
    if env.ground_enabled and pr.nodes:
        lo, hi = _world_bounds(pr.nodes, np.eye(4))
        if lo[0] <= hi[0]:
            if env.ground_shadow_catcher:
                raise DocumentError("shadow catchers are not implemented on the HIP path")
            corner, eu, ev = ground_quad(lo, hi)
            p1 = corner + eu
            quad = np.array([corner, p1, p1 + ev, corner + ev], np.float32)
            gm = sc.add_material(disney(env.ground_color, 0.0, env.ground_roughness, 0.5))
            sc.add_shape(Mesh(quad, np.array([[0, 2, 1], [0, 3, 2]], np.uint32)), gm, node_uid=GROUND_NODE_UID,
                         name="Zemin")
    for d in pr.lights:
        if not d.enabled or d.intensity <= 0.0:
            continue
        if d.type != "area":
            raise DocumentError(f"light {d.name!r}: {d.type} lights are not implemented on the HIP path")
        if d.light_link and (d.light_link.get("mode", "all") != "all" or
                             d.light_link.get("shadowMode", "all") != "all"):
            raise DocumentError(f"light {d.name!r}: light linking is not implemented on the HIP path")
        sc.add_light(area_light(d))
    if env.intensity > 0.0:
        if env.kind == "sky":
            raise DocumentError("the physical sky environment is not implemented on the HIP path")
        img = None
        if env.hdr:
            if Path(env.hdr).suffix.lower() == ".exr":
                raise DocumentError(f"{env.hdr}: OpenEXR environments are not read here")
            try:
                img = read_hdr(env.hdr)
            except (OSError, DocumentError) as e:
                pr.warnings.append(str(e))
        if img is None:
            img = procedural_sky(env.zenith, env.horizon)
        sc.set_environment(Environment(img, float(np.float32(env.rotation_deg) * np.float32(math.pi / 180.0)),
                                       env.intensity))
    sc.background_mode = {"color": capi.PE_BACKGROUND_COLOR,
                          "transparent": capi.PE_BACKGROUND_TRANSPARENT}.get(env.background_mode,
                                                                             capi.PE_BACKGROUND_ENVIRONMENT)
    sc.background_color = env.background_color
    return sc


def orbit_view(pr: Project, width: int, height: int, spp: int = 16, max_bounces: int | None = None) -> View:
    """The project's orbit camera (photon_render's cameraFromJson, the golden test's orbitCamera,
    test_golden_reference.cpp:44-51): eye = target + r (sin th cos ph, cos th, sin th sin ph), up +y; the
    pinhole's view (orbit_camera gives the thin-lens or orthographic camera the block may ask for)."""
    c = pr.camera
    f32 = np.float32
    t = np.asarray(c.get("target", (0.0, 0.0, 0.0)), f32)
    r, th, ph = f32(c.get("radius", 5.0)), f32(c.get("theta", 1.15)), f32(c.get("phi", 0.45))
    eye = t + np.array([math.sin(th) * math.cos(ph), math.cos(th), math.sin(th) * math.sin(ph)], f32) * r
    mb = int(pr.render.get("maxBounces", 12)) if max_bounces is None else max_bounces
    return View(tuple(map(float, eye)), tuple(map(float, t)), (0.0, 1.0, 0.0), float(c.get("fov", 35.0)), width,
                height, spp, mb)


def orbit_camera(pr: Project, width: int, height: int):
    """The camera cameraFromJson builds from the project's camera block (main.cpp:89-111): orthographic with
    vertical extent 2 tan(fov / 2) r, a ThinLensCamera when aperture > 0 (focus = focusDistance when
    focusExplicit, else r; apertureBlades, bladeRotation, shiftX / shiftY) or a PerspectiveCamera with its
    lens shift; -> pe_camera."""
    from .device import camera_orthographic, camera_perspective, camera_thin_lens, camera_with_shift

    c = pr.camera
    view = orbit_view(pr, width, height)
    f32 = np.float32
    aspect = f32(width) / f32(height)
    r = f32(c.get("radius", 5.0))
    fov = f32(c.get("fov", 35.0))
    if c.get("orthographic", False):
        h = f32(2.0) * f32(math.tan(fov * f32(math.pi / 180.0) * f32(0.5))) * r
        return camera_orthographic(view.position, view.target, view.up, float(h), float(aspect))
    aperture = float(c.get("aperture", 0.0))
    if aperture > 0.0:
        focus = float(c.get("focusDistance", r)) if c.get("focusExplicit", False) else float(r)
        cam = camera_thin_lens(view.position, view.target, view.up, float(fov), float(aspect), aperture, focus,
                               int(c.get("apertureBlades", 0)), float(c.get("bladeRotation", 0.0)))
    else:
        cam = camera_perspective(view.position, view.target, view.up, float(fov), float(aspect))
    sx, sy = float(c.get("shiftX", 0.0)), float(c.get("shiftY", 0.0))
    return camera_with_shift(cam, sx, sy) if (sx or sy) else cam


