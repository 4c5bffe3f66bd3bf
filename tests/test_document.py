"""The document front end (photon_hip.document): .photon projects, OBJ models, .mesh binaries and Radiance
environment maps into a Scene for the HIP path.

Pinned by:
  * the reference's own OBJ tests: an MTL's Kd becomes the Lambertian of the mesh (test_import.cpp:63-89);
    empty, truncated, out-of-range and non-finite files load as nothing (test_loader_limits.cpp:108-130);
  * the reference's .mesh codec, built here from scene/mesh_binary.cpp: its bytes decode here and these bytes
    decode there, corrupt sizes and indices refused by both;
  * the reference's temperatureColor (core/image/imager.cpp:204-213) for the light colour of a temperature;
  * the Mitsuba golden images: the three golden documents go through this loader in test_golden_mitsuba.py;
  * for the rest, what the formats and formulas fix: fan triangulation and negative indices, flipped v,
    material buckets; an RLE Radiance file decodes to the flat one's pixels; the ground quad lies under the
    scene's bounds at 200 times its size; an area light's quad is centred on its position, faces its target and
    has its width and height; the tree's transforms compose; what the HIP path does not render is refused.
GPU: a project built by the loader renders on the device as the oracle renders it, sample for sample.
"""
from __future__ import annotations

import ctypes as C
import json

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi
from photon_hip import document as D
from photon_hip.scene import Mesh


def _write(p, text):
    p.write_text(text)
    return p


# ------------------------------------------------------------------------------------------------ OBJ

def test_obj_mtl_diffuse_is_assigned(tmp_path):
    """ObjLoader.MtlDiffuseIsAssigned (test_import.cpp:63-89)."""
    _write(tmp_path / "tri.mtl", "newmtl red\nKd 0.2 0.8 0.1\n")
    _write(tmp_path / "tri.obj", "mtllib tri.mtl\nusemtl red\nv 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    m = D.load_obj(tmp_path / "tri.obj")
    assert len(m.meshes) == 1 and len(m.materials) == 1
    assert m.materials[0].type == capi.PE_MAT_LAMBERT
    np.testing.assert_allclose(list(m.materials[0].base_color), [0.2, 0.8, 0.1], atol=1e-6)
    assert m.meshes[0].num_triangles == 1


def test_obj_broken_files_load_as_nothing(tmp_path):
    """LoaderLimits (test_loader_limits.cpp:108-130): empty, truncated, out-of-range and infinite."""
    cases = {"empty.obj": "", "trunc.obj": "v 0 0\n", "huge.obj": "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 999999999\n",
             "inf.obj": "v 1e999 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"}
    for name, text in cases.items():
        assert not D.load_obj(_write(tmp_path / name, text)).meshes, name
    assert len(D.load_obj(_write(tmp_path / "tri.obj", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")).meshes) == 1
    assert not D.load_obj(tmp_path / "missing.obj").meshes


def test_obj_fans_flips_v_and_buckets_by_material(tmp_path):
    _write(tmp_path / "m.mtl", "newmtl a\nKd 1 0 0\nnewmtl b\nKd 0 0 1\n")
    _write(tmp_path / "m.obj", "\n".join([
        "mtllib m.mtl", "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 0.5 1.5 0",
        "vt 0 0", "vt 1 0", "vt 1 1", "vt 0 1", "vn 0 0 1",
        "o first", "usemtl a", "f 1/1/1 2/2/1 3/3/1 4/4/1",          # a quad: two triangles fanned from 1
        "usemtl b", "f -2 -3 -1",                                      # negative indices: 4 3 5
        "usemtl a", "f 1 3 4",
        "g second", "f 2 3 5",
    ]) + "\n")
    m = D.load_obj(tmp_path / "m.obj")
    assert len(m.meshes) == 3                      # shape 1: buckets a, b; shape 2: bucket a
    a, b, second = m.meshes
    assert a.num_triangles == 3 and b.num_triangles == 1 and second.num_triangles == 1
    np.testing.assert_array_equal(a.positions[:6], [[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 0, 0], [1, 1, 0], [0, 1, 0]])
    np.testing.assert_array_equal(a.indices[:2], [[0, 1, 2], [3, 4, 5]])       # corners are not welded
    np.testing.assert_array_equal(b.positions, [[0, 1, 0], [1, 1, 0], [0.5, 1.5, 0]])
    assert a.uvs is None and a.normals is None     # only the quad's corners carry them: dropped for the bucket
    assert list(m.materials[0].base_color) == [1, 0, 0] and list(m.materials[1].base_color) == [0, 0, 1]
    q = D.load_obj(_write(tmp_path / "q.obj", "v 0 0 0\nv 1 0 0\nv 1 1 0\nvt 0 0.25\nvt 1 0\nvt 1 1\n"
                                                "vn 0 0 1\nf 1/1/1 2/2/1 3/3/1\n"))
    np.testing.assert_array_equal(q.meshes[0].uvs, [[0, 0.75], [1, 1], [1, 0]])   # v -> 1 - v
    np.testing.assert_array_equal(q.meshes[0].normals, [[0, 0, 1]] * 3)
    assert q.materials == [None]


# ------------------------------------------------------------------------------------------- .mesh

def _ref_or_skip():
    ref = bindings.reference(strict=True)
    if ref is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    return ref


def _ref_encode(ref, mesh):
    pos = np.ascontiguousarray(mesh.positions, np.float32)
    nrm = None if mesh.normals is None else np.ascontiguousarray(mesh.normals, np.float32)
    uv = None if mesh.uvs is None else np.ascontiguousarray(mesh.uvs, np.float32)
    idx = np.ascontiguousarray(mesh.indices.ravel(), np.uint32)
    fp = lambda a: None if a is None else a.ctypes.data_as(C.POINTER(C.c_float))  # noqa: E731
    n = ref.ref_mesh_encode(fp(pos), fp(nrm), fp(uv), len(pos), idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                            idx.size, None, 0)
    out = (C.c_uint8 * n)()
    ref.ref_mesh_encode(fp(pos), fp(nrm), fp(uv), len(pos), idx.ctypes.data_as(C.POINTER(C.c_uint32)), idx.size,
                        out, n)
    return bytes(out)


def _ref_decode(ref, data):
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    counts = np.array([1 << 20, 1 << 20, 0], np.uint32)
    pos = np.zeros((1 << 20) * 3, np.float32)
    uv = np.zeros((1 << 20) * 2, np.float32)
    idx = np.zeros(1 << 20, np.uint32)
    rc = ref.ref_mesh_decode(buf, len(data), pos.ctypes.data_as(C.POINTER(C.c_float)),
                             uv.ctypes.data_as(C.POINTER(C.c_float)), idx.ctypes.data_as(C.POINTER(C.c_uint32)),
                             counts.ctypes.data_as(C.POINTER(C.c_uint32)))
    if rc:
        return None
    nv, ni, has_uv = (int(x) for x in counts)
    return pos[:3 * nv].reshape(-1, 3), (uv[:2 * nv].reshape(-1, 2) if has_uv else None), idx[:ni].reshape(-1, 3)


def test_mesh_binary_matches_the_reference_codec():
    ref = _ref_or_skip()
    rng = np.random.default_rng(1)
    p = rng.normal(size=(40, 3)).astype(np.float32)
    idx = rng.integers(0, 40, (30, 3)).astype(np.uint32)
    n = rng.normal(size=(40, 3)).astype(np.float32)
    uv = rng.random((40, 2)).astype(np.float32)
    for mesh in (Mesh(p, idx), Mesh(p, idx, n, uv), Mesh(p, idx, n), Mesh(p, idx, None, uv)):
        ours = D.write_mesh_binary(mesh)
        if mesh.normals is not None:      # a reference TriangleMesh always has normals (computed when absent)
            assert ours == _ref_encode(ref, mesh)              # the same bytes
        back = D.read_mesh_binary(ours)
        np.testing.assert_array_equal(back.positions, mesh.positions)
        np.testing.assert_array_equal(back.indices, mesh.indices)
        assert (back.uvs is None) == (mesh.uvs is None) and (back.normals is None) == (mesh.normals is None)
        rp, ruv, ri = _ref_decode(ref, ours)
        np.testing.assert_array_equal(rp, back.positions)
        np.testing.assert_array_equal(ri, back.indices)
        if mesh.uvs is not None:
            np.testing.assert_array_equal(ruv, back.uvs)
    good = D.write_mesh_binary(Mesh(p, idx))
    bad_index = bytearray(good)
    bad_index[-4:] = (1000).to_bytes(4, "little")
    nan_pos = bytearray(good)
    nan_pos[24:28] = np.float32(np.nan).tobytes()
    for data in (good[:-4], bytes(bad_index), bytes(nan_pos), b"PHMESH02" + good[8:], good[:20]):
        with pytest.raises(D.DocumentError):
            D.read_mesh_binary(data)
        assert _ref_decode(ref, data) is None


# --------------------------------------------------------------------------------------------- HDR

def _rle_hdr(rgbe: np.ndarray) -> bytes:
    """Run-length Radiance scanlines (the 2 2 w marker, per channel: runs of a repeated byte, count | 128,
    and literal stretches) for the reader's other branch."""
    h, w, _ = rgbe.shape
    out = bytearray(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n" + f"-Y {h} +X {w}\n".encode())
    for y in range(h):
        out += bytes([2, 2, w >> 8, w & 255])
        for c in range(4):
            row = rgbe[y, :, c]
            x = 0
            while x < w:
                run = 1
                while x + run < w and run < 127 and row[x + run] == row[x]:
                    run += 1
                if run > 2:
                    out += bytes([128 + run, row[x]])
                    x += run
                else:
                    n = 1
                    while x + n < w and n < 128 and not (x + n + 2 < w and row[x + n] == row[x + n + 1] == row[x + n + 2]):
                        n += 1
                    out += bytes([n]) + bytes(row[x:x + n])
                    x += n
    return bytes(out)


def test_hdr_flat_and_run_length_files_decode_alike(tmp_path):
    rng = np.random.default_rng(2)
    rgbe = rng.integers(0, 256, (6, 40, 4)).astype(np.uint8)
    rgbe[2, 5:30] = rgbe[2, 5]                      # runs to encode
    rgbe[4, :, 3] = 0                               # exponent 0: black
    flat = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 6 +X 40\n" + rgbe.tobytes()
    (tmp_path / "flat.hdr").write_bytes(flat)
    (tmp_path / "rle.hdr").write_bytes(_rle_hdr(rgbe))
    a, b = D.read_hdr(tmp_path / "flat.hdr"), D.read_hdr(tmp_path / "rle.hdr")
    np.testing.assert_array_equal(a, b)
    # stb_image's rule: m * 2^(e - 136), e = 0 black
    e = rgbe[..., 3].astype(np.int64)
    want = np.where(e[..., None] > 0, rgbe[..., :3] * np.ldexp(1.0, e[..., None] - 136), 0.0)
    np.testing.assert_array_equal(a, want.astype(np.float32))
    img = rng.random((3, 9, 3)).astype(np.float32) * 20
    D.write_hdr(tmp_path / "w.hdr", img)
    back = D.read_hdr(tmp_path / "w.hdr")        # RGBE keeps 8 bits under the pixel's largest channel
    assert (np.abs(back - img) <= img.max(-1, keepdims=True) / 128).all()
    (tmp_path / "bad.hdr").write_bytes(b"#?RADIANCE\nFORMAT=32-bit_rle_xyze\n\n-Y 1 +X 1\n\0\0\0\0")
    with pytest.raises(D.DocumentError):
        D.read_hdr(tmp_path / "bad.hdr")


# ----------------------------------------------------------------------------------------- project

def test_temperature_color_matches_the_reference():
    ref = _ref_or_skip()
    ks = np.array([1000, 1500, 1900, 2222, 2700, 3200, 3500, 4000, 4300, 5000, 6504, 7000, 9000, 15000, 25000,
                   40000], np.float32)
    want = np.zeros((ks.size, 3), np.float32)
    ref.ref_temperature_color(ks.size, ks.ctypes.data_as(C.POINTER(C.c_float)),
                              want.ctypes.data_as(C.POINTER(C.c_float)))
    got = np.array([D.temperature_color(k) for k in ks], np.float32)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=1e-7)


def _project(tmp_path, **over):
    """A project with every kind of node: an inline mesh, a .mesh file, an imported OBJ (with a part dropped
    from the saved tree and one regrouped), a group with a transform, a sphere, lights and the procedural sky."""
    quad = Mesh(np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32),
                np.array([[0, 1, 2], [0, 2, 3]], np.uint32))
    (tmp_path / "part.mesh").write_bytes(D.write_mesh_binary(quad))
    _write(tmp_path / "model.mtl", "newmtl red\nKd 0.8 0.1 0.1\n")
    _write(tmp_path / "model.obj", "mtllib model.mtl\nv 0 0 0\nv 1 0 0\nv 0 0 1\nv 1 0 1\n"
                                   "o a\nusemtl red\nf 1 2 3\no b\nf 2 4 3\no c\nf 1 3 4\n")
    xf = lambda t, s=1.0: [s, 0, 0, t[0], 0, s, 0, t[1], 0, 0, s, t[2], 0, 0, 0, 1]  # noqa: E731
    doc = {
        "format": "photon-project", "version": 2,
        "camera": {"target": [0, 0.5, 0], "radius": 4.0, "theta": 1.2, "phi": 0.7, "fov": 40},
        "render": {"maxBounces": 6},
        "materials": [{"type": "disney", "baseColor": [0.7, 0.6, 0.5], "roughness": 0.3},
                      {"type": "glass", "ior": 1.45, "roughness": 0.1, "thinWalled": False,
                       "absorption": {"color": [0.9, 0.8, 0.7], "distance": 0.5}},
                      {"type": "lambert", "albedo": [0.2, 0.3, 0.4]},
                      {"type": "disney", "baseColor": [1, 1, 1], "subsurface": {"weight": 1.0, "scale": 0.01},
                       "sheenColor": [0.5, 0.5, 1.0], "thinFilm": {"thickness": 300, "ior": 1.3}}],
        "nodes": [
            {"type": "mesh", "name": "Plate", "material": 0, "xf": xf((0, 0.2, 0)),
             "geometry": {"p": [0, 0, 0, 1, 0, 0, 1, 0, 1], "i": [0, 2, 1]}},
            {"type": "group", "name": "Rig", "xf": xf((1, 0, 0), 2.0), "children": [
                {"type": "mesh", "name": "Panel", "material": 1, "geometry": {"bin": "part.mesh"}},
                {"type": "mesh", "name": "Hidden", "material": 2, "visible": False,
                 "geometry": {"p": [0, 0, 0, 1, 0, 0, 0, 1, 0], "i": [0, 1, 2]}}]},
            {"source": "model.obj", "name": "Model", "xf": xf((0, 0, -2)), "children": [
                {"sourceIndex": 0, "name": "Red part"},
                {"type": "group", "name": "Sub", "children": [{"sourceIndex": 2, "material": 3}]}]},
        ],
        "lights": [{"type": "area", "name": "Key", "position": [0, 3, 2], "target": [0, 0, 0], "width": 1.5,
                    "height": 0.5, "intensity": 5.0, "color": [1, 0.9, 0.8], "visibleToCamera": False,
                    "lightGroup": "key"},
                   {"type": "area", "name": "Warm", "position": [2, 2, 0], "target": [0, 0.5, 0],
                    "temperature": 3200, "intensity": 2.0},
                   {"type": "area", "name": "Off", "enabled": False}],
        "environment": {"hdr": "", "zenith": [0.3, 0.4, 0.6], "horizon": [0.5, 0.5, 0.5], "intensity": 0.5,
                        "rotation": 30, "background": {"mode": "color", "color": [1, 1, 1]},
                        "ground": {"enabled": True, "color": [0.3, 0.3, 0.3], "roughness": 0.7}},
    }
    for k, v in over.items():
        doc[k] = v
    p = tmp_path / "scene.photon"
    p.write_text(json.dumps(doc))
    return p, doc


def test_project_compiles_as_build_render_scene(tmp_path):
    path, doc = _project(tmp_path)
    pr = D.load_project(path)
    assert not pr.warnings, pr.warnings
    sc = D.build_render_scene(pr)
    names = [s.name for s in sc.shapes if not s.standalone]
    # tree order; the invisible node left out; the imported model's saved parts only (part "b" was dropped),
    # the regrouped part under its group; the ground last
    assert names == ["Plate", "Panel", "Red part", "Parça 3", "Zemin"], names
    uids = [s.node_uid for s in sc.shapes if not s.standalone]
    assert len(set(uids)) == len(uids) and uids[-1] == D.GROUND_NODE_UID
    # transforms compose: the panel's quad under Rig (translate (1, 0, 0), scale 2)
    panel = sc.shapes[1].mesh
    np.testing.assert_allclose(panel.positions, [[1, 0, 0], [3, 0, 0], [3, 2, 0], [1, 2, 0]], atol=1e-6)
    plate = sc.shapes[0].mesh
    np.testing.assert_allclose(plate.positions[:, 1], 0.2, atol=1e-7)
    # materials: named <node>_Mat by their first node; the OBJ part keeps its Kd, the regrouped part takes 3
    mats = {sc.material_names[i]: m for i, m in enumerate(sc.materials) if i in sc.material_names}
    assert set(mats) >= {"Plate_Mat", "Panel_Mat", "Red part_Mat", "Parça 3_Mat"}
    np.testing.assert_allclose(list(mats["Red part_Mat"].base_color), [0.8, 0.1, 0.1], atol=1e-6)
    assert mats["Panel_Mat"].type == capi.PE_MAT_DIELECTRIC
    np.testing.assert_allclose(list(mats["Panel_Mat"].absorption_color), [0.9, 0.8, 0.7], atol=1e-6)
    assert mats["Parça 3_Mat"].subsurface == 1.0 and mats["Parça 3_Mat"].has_sheen_color == 1
    assert abs(mats["Parça 3_Mat"].thin_film_thickness - 300) < 1e-4
    # the ground: under the lowest point, centred, 200 times the largest extent
    ground = sc.shapes[4].mesh.positions
    pts = np.concatenate([s.mesh.positions for s in sc.shapes[:4]])
    lo, hi = pts.min(0), pts.max(0)
    span = max(hi - lo)
    assert np.allclose(ground[:, 1], lo[1] - span * 0.0005, rtol=1e-5)
    np.testing.assert_allclose(ground.max(0)[[0, 2]] - ground.min(0)[[0, 2]], 200 * span, rtol=1e-5)
    np.testing.assert_allclose((ground.max(0) + ground.min(0))[[0, 2]] / 2, ((lo + hi) / 2)[[0, 2]], atol=1e-4)
    # lights: the enabled ones; each quad centred on its position, facing its target, width x height
    assert len(sc.lights) == 2 and sc.light_group_of == ["key", ""]
    for L, d in zip(sc.lights, doc["lights"][:2]):
        p0, u, v = (np.asarray(list(x), np.float64) for x in (L.p0, L.p1, L.p2))
        np.testing.assert_allclose(p0 + 0.5 * (u + v), d["position"], atol=1e-5)
        n = np.cross(u, v)
        to_target = np.asarray(d["target"], np.float64) - d["position"]
        assert np.dot(n, to_target) / np.linalg.norm(n) / np.linalg.norm(to_target) > 0.9999
        assert abs(np.dot(u, v)) < 1e-6
        np.testing.assert_allclose([np.linalg.norm(u), np.linalg.norm(v)], [d.get("width", 1.0), d.get("height", 1.0)],
                                   rtol=1e-6)
    np.testing.assert_allclose(list(sc.lights[0].radiance), [5.0, 4.5, 4.0], rtol=1e-6)
    warm = np.asarray(list(sc.lights[1].radiance))
    assert warm[0] > warm[1] > warm[2]                        # 3200 K is warm
    # environment: the procedural sky, its rotation in radians; the Color background
    assert sc.environment.rgb.shape == (256, 512, 3)
    assert abs(sc.environment.rotation - np.radians(30)) < 1e-6 and sc.environment.intensity == 0.5
    assert sc.background_mode == capi.PE_BACKGROUND_COLOR and tuple(sc.background_color) == (1.0, 1.0, 1.0)
    v = D.orbit_view(pr, 64, 48)
    assert v.max_bounces == 6
    np.testing.assert_allclose(np.linalg.norm(np.subtract(v.position, v.target)), 4.0, rtol=1e-6)


def test_project_flakes_load_as_material_presets(tmp_path):
    """A material's "flakes" block (material_library.cpp:254-259) becomes setFlakes(density, size, roughness,
    color) with the preset's defaults for missing keys (material_library.h:73-76)."""
    mats = [{"type": "disney", "baseColor": [0.5, 0.1, 0.1], "flakes": {"density": 0.4, "color": [1, 0.8, 0.5]}},
            {"type": "disney", "flakes": {"density": 0.9, "size": 0.002, "roughness": 0.7}}]
    path, _ = _project(tmp_path, materials=mats)
    pr = D.load_project(path)
    sc = D.build_render_scene(pr)
    m0, m1 = sc.materials[0], sc.materials[1]
    assert (m0.flake_density, m0.flake_size, m0.flake_roughness) == (np.float32(0.4), np.float32(0.004), np.float32(0.3))
    assert tuple(m0.flake_color) == (1.0, np.float32(0.8), 0.5)
    assert (m1.flake_density, m1.flake_size, m1.flake_roughness) == (np.float32(0.9), np.float32(0.002), np.float32(0.7))
    assert tuple(m1.flake_color) == (np.float32(0.9),) * 3


def test_project_cameras_as_camera_from_json(tmp_path):
    """cameraFromJson (main.cpp:89-111): the camera block's aperture makes a ThinLensCamera focused at r (or the
    explicit focus distance) with its blades, rotation and lens shift; orthographic makes an OrthographicCamera
    2 tan(fov / 2) r high; the rest a shifted pinhole. Each equals the reference's own camera of those
    parameters, description for description."""
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    base = {"target": [0, 0.5, 0], "radius": 4.0, "theta": 1.2, "phi": 0.7, "fov": 40}
    blocks = [
        (dict(base, aperture=0.08, apertureBlades=7, bladeRotation=12.0, shiftX=0.05, shiftY=-0.02),
         dict(kind="thin_lens", aperture=0.08, focus=4.0, blades=7, rotation_deg=12.0, shift=(0.05, -0.02))),
        (dict(base, aperture=0.05, focusExplicit=True, focusDistance=2.5),
         dict(kind="thin_lens", aperture=0.05, focus=2.5)),
        (dict(base, shiftY=0.1), dict(kind="perspective", shift=(0.0, 0.1))),
        (dict(base, orthographic=True), None),
    ]
    for block, spec in blocks:
        path, _ = _project(tmp_path, camera=block)
        pr = D.load_project(path)
        cam = D.orbit_camera(pr, 64, 48)
        view = D.orbit_view(pr, 64, 48)
        if spec is None:
            h = np.float32(2.0) * np.float32(np.tan(np.float32(40.0) * np.float32(np.pi / 180.0) * np.float32(0.5)))
            spec = dict(kind="orthographic", extent=float(h * np.float32(4.0)))
            assert cam.orthographic == 1
        with bindings.reference_camera(strict=True, **spec) as rc:
            want = rc.describe(view)
        got = np.array(list(cam.position) + list(cam.lower_left) + list(cam.horizontal) + list(cam.vertical)
                       + list(cam.u) + list(cam.v) + list(cam.forward) + [cam.lens_radius], np.float32)
        np.testing.assert_allclose(got, want[:22], rtol=2e-6, atol=2e-7)
        assert (cam.blades, cam.orthographic) == (int(want[44]), int(want[46]))
        assert abs(cam.blade_rotation - want[45]) <= 1e-7


def test_project_refuses_what_the_hip_path_does_not_render(tmp_path):
    cases = [
        ("materials", [{"type": "disney", "maps": {"albedo": "a.png"}}], "texture maps"),
        ("lights", [{"type": "point", "intensity": 1.0}], "point lights"),
        ("environment", {"kind": "sky", "intensity": 1.0}, "physical sky"),
        ("environment", {"intensity": 0.0, "ground": {"enabled": True, "shadowCatcher": True}}, "shadow catcher"),
    ]
    for key, value, what in cases:
        path, _ = _project(tmp_path, **{key: value})
        pr = D.load_project(path)
        with pytest.raises(D.DocumentError, match=what):
            D.build_render_scene(pr)
    sphere = {"type": "sphere", "name": "Ball", "material": 0, "radius": 0.5}
    path, doc = _project(tmp_path)
    doc["nodes"].append(sphere)
    path.write_text(json.dumps(doc))
    with pytest.raises(D.DocumentError, match="analytic spheres"):
        D.build_render_scene(D.load_project(path))
    sc = D.build_render_scene(D.load_project(path), sphere_segments=(32, 16))
    assert "Ball" in [s.name for s in sc.shapes]
    with pytest.raises(D.DocumentError, match="version"):
        D.load_project_json(json.dumps({"version": 1}), tmp_path)


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_renders_a_loaded_project_as_the_oracle(dev, tmp_path):
    from photon_hip.device import camera_perspective, render, settings

    path, _ = _project(tmp_path)
    pr = D.load_project(path)
    sc = D.build_render_scene(pr)
    view = D.orbit_view(pr, 48, 32)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    res = render(dev, gs, cam, settings(max_bounces=view.max_bounces, seed=12345), film, 3, 1)
    assert res.completed == 1
    s, c, _, _ = film.download()
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    Lo, _ = bindings.OracleScene(sc).trace_samples(cam, settings(max_bounces=view.max_bounces, seed=12345),
                                                   view.width, view.height, xx.ravel(), yy.ravel(),
                                                   np.full(xx.size, 3))
    Lg = s.reshape(-1, 3)
    close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
    print(f"loaded project: per-sample agreement with the oracle {close.mean():.5f} of {close.size}", flush=True)
    assert close.mean() >= 0.998
    assert Lo.mean() > 0.05
