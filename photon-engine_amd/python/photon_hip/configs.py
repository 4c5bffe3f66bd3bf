"""Seeded synthetic workloads.

box_icosphere   closed box of 6 inward quads (12 triangles) around an icosphere (level 5: 20,480 triangles),
                Disney principled everywhere (seeded baseColor U[0.2, 0.9], roughness 0.3, metallic 0,
                specular 0.5), one 0.5 x 0.5 area light of radiance 10 under the ceiling; 512 x 512, 64 spp,
                max depth 8, Russian roulette from depth 3.
product_sphere  a UV sphere of 1024 x 1000 segments (2,045,952 triangles) displaced by 0.05 * Perlin(4p),
                split into 16 longitude wedges, each its own Disney material (metallic in {0, 1},
                roughness U[0.05, 0.8], clearcoat U[0, 1]); a 2048 x 1024 sky-gradient environment with a
                0.5-degree sun disk at 1e4 x baked in; 1920 x 1080, 256 spp, max depth 12.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .scene import AreaLight, Environment, Mesh, Scene, dielectric, disney, lambert


@dataclass
class View:
    position: tuple
    target: tuple
    up: tuple
    vfov: float
    width: int
    height: int
    spp: int
    max_bounces: int

    @property
    def aspect(self) -> float:
        return self.width / self.height


def icosphere(level: int, radius: float, center) -> Mesh:
    t = (1.0 + 5.0 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    v = np.array(verts, dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array(faces, dtype=np.int64)
    for _ in range(level):
        edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.sort(edges, axis=1)
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = inv.reshape(-1)
        mid = v[uniq[:, 0]] + v[uniq[:, 1]]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        base = v.shape[0]
        v = np.concatenate([v, mid])
        nf = f.shape[0]
        a = base + inv[0:nf]
        b = base + inv[nf:2 * nf]
        c = base + inv[2 * nf:3 * nf]
        f = np.concatenate([np.stack([f[:, 0], a, c], 1), np.stack([f[:, 1], b, a], 1),
                            np.stack([f[:, 2], c, b], 1), np.stack([a, b, c], 1)])
    pos = (v * radius + np.asarray(center, dtype=np.float64)).astype(np.float32)
    return Mesh(pos, f.astype(np.uint32), v.astype(np.float32))


def _quad(p0, p1, p2, p3) -> Mesh:
    return Mesh(np.array([p0, p1, p2, p3], dtype=np.float32), np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32))


def box_icosphere(seed: int = 0, level: int = 5, width: int = 512, height: int = 512, spp: int = 64,
                  max_bounces: int = 8) -> tuple[Scene, View]:
    rng = np.random.default_rng(seed)
    sc = Scene()
    # Six inward-facing walls of [-1, 1]^3 (counter-clockwise seen from inside).
    walls = [
        ((-1, -1, -1), (1, -1, -1), (1, -1, 1), (-1, -1, 1)),   # floor, normal +y
        ((-1, 1, -1), (-1, 1, 1), (1, 1, 1), (1, 1, -1)),       # ceiling, normal -y
        ((-1, -1, -1), (-1, 1, -1), (1, 1, -1), (1, -1, -1)),   # back, normal +z
        ((-1, -1, 1), (1, -1, 1), (1, 1, 1), (-1, 1, 1)),       # front, normal -z
        ((-1, -1, -1), (-1, -1, 1), (-1, 1, 1), (-1, 1, -1)),   # left, normal +x
        ((1, -1, -1), (1, 1, -1), (1, 1, 1), (1, -1, 1)),       # right, normal -x
    ]
    for w in walls:
        m = sc.add_material(disney(rng.uniform(0.2, 0.9, 3), 0.0, 0.3, 0.5))
        sc.add_shape(_quad(*w), m)
    m = sc.add_material(disney(rng.uniform(0.2, 0.9, 3), 0.0, 0.3, 0.5))
    sc.add_shape(icosphere(level, 0.6, (0.0, -0.4, 0.0)), m)
    sc.add_light(AreaLight((-0.25, 0.999, -0.25), (0.5, 0.0, 0.0), (0.0, 0.0, 0.5), (10.0, 10.0, 10.0)))
    view = View((0.0, 0.0, 0.95), (0.0, -0.2, 0.0), (0.0, 1.0, 0.0), 60.0, width, height, spp, max_bounces)
    return sc, view


# ---- LED strip (many emitters) --------------------------------------------------------------------------

def led_strip(quads: int = 1000, width: int = 48, height: int = 48, spp: int = 8, max_bounces: int = 1,
              radiance: float = 4.0) -> tuple[Scene, View]:
    """The reference's many-emitter scene (tests/test_light_sampler.cpp:104-127): a 12 x 12 Lambert floor under a
    spiral of `quads` small emissive quads (2 * quads emissive triangles, Lambert(0, radiance)), so light
    selection runs on the light BVH (more than 33 candidates)."""
    sc = Scene()
    floor = sc.add_material(lambert((0.7, 0.7, 0.7)))
    up = np.array([0.0, 1.0, 0.0], dtype=np.float32)
    sc.add_shape(Mesh(np.array([[-6, 0, -6], [6, 0, -6], [6, 0, 6], [-6, 0, 6]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)), floor)
    f = np.float32
    t = np.arange(quads, dtype=np.float32) / f(quads)
    a, r = t * f(25.0), f(0.5) + f(4.5) * t
    c = np.stack([r * np.cos(a), np.full_like(t, 0.6), r * np.sin(a)], 1).astype(np.float32)
    off = np.array([[-0.03, 0, -0.03], [0.03, 0, -0.03], [0.03, 0, 0.03], [-0.03, 0, 0.03]], np.float32)
    pos = (c[:, None, :] + off[None, :, :]).reshape(-1, 3)
    b = (np.arange(quads, dtype=np.uint32) * 4)[:, None]
    idx = np.concatenate([b + np.uint32(0), b + np.uint32(1), b + np.uint32(2),
                          b + np.uint32(0), b + np.uint32(2), b + np.uint32(3)], 1).reshape(-1, 3)
    led = sc.add_material(lambert((0.0, 0.0, 0.0), (radiance, radiance, radiance)))
    sc.add_shape(Mesh(pos, idx), led)
    view = View((0.0, 7.0, 7.0), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 50.0, width, height, spp, max_bounces)
    return sc, view


# ---- transmissive studio (glass, dispersion, absorbing liquid, subsurface) -------------------------------

def box_mesh(lo, hi) -> Mesh:
    """Closed box with outward-facing triangles (a volume boundary for glass, liquid and subsurface walks)."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    p = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [3, 7, 6], [3, 6, 2], [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)
    return Mesh(p, f)


def prism_mesh(center, size, height, subdiv: int = 1) -> Mesh:
    """Closed triangular prism (outward normals) standing on y = center.y, its sides split into subdiv^2 quads."""
    cx, cy, cz = center
    a = [(cx + size * np.cos(t), cz + size * np.sin(t)) for t in (np.pi / 2, np.pi / 2 + 2 * np.pi / 3,
                                                                    np.pi / 2 + 4 * np.pi / 3)]
    pos, tris = [], []

    def quad_grid(p00, p10, p01, p11):
        base = len(pos)
        for j in range(subdiv + 1):
            for i in range(subdiv + 1):
                u, v = i / subdiv, j / subdiv
                q = (1 - u) * (1 - v) * np.array(p00) + u * (1 - v) * np.array(p10) + (1 - u) * v * np.array(p01) + \
                    u * v * np.array(p11)
                pos.append(q)
        for j in range(subdiv):
            for i in range(subdiv):
                k = base + j * (subdiv + 1) + i
                tris.append((k, k + 1, k + subdiv + 2))
                tris.append((k, k + subdiv + 2, k + subdiv + 1))

    for k in range(3):
        (x0, z0), (x1, z1) = a[k], a[(k + 1) % 3]
        quad_grid((x0, cy, z0), (x0, cy + height, z0), (x1, cy, z1), (x1, cy + height, z1))
    b = len(pos)
    pos += [(x, cy, z) for x, z in a] + [(x, cy + height, z) for x, z in a]
    tris += [(b, b + 1, b + 2), (b + 3, b + 5, b + 4)]
    return Mesh(np.array(pos, np.float32), np.array(tris, np.uint32))


def glass_studio(width: int = 64, height: int = 48, spp: int = 16, max_bounces: int = 16, led_quads: int = 0,
                 prism_subdiv: int = 2) -> tuple[Scene, View]:
    """Transmissive studio after the spectral/volume workload: a dispersive prism (Dielectric 1.5, Abbe 30), an
    absorbing liquid block (Dielectric 1.33, sigma_a U[0.1, 2] / cm), a subsurface block (Disney subsurface 1,
    mean free path 1-5 mm), a thin-walled pane, a rough glass sphere and a Disney material with specular and
    diffuse transmission, on a Lambert floor under an area light (plus an LED spiral of `led_quads` emissive
    quads when asked)."""
    rng = np.random.default_rng(2)
    sc = Scene()
    floor = sc.add_material(lambert((0.6, 0.6, 0.6)))
    up = np.array([0.0, 1.0, 0.0], np.float32)
    sc.add_shape(Mesh(np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)), floor)
    back = sc.add_material(disney((0.3, 0.35, 0.5), 0.0, 0.6, 0.5))
    sc.add_shape(_quad((-4, 0, -1.5), (4, 0, -1.5), (4, 3, -1.5), (-4, 3, -1.5)), back)
    prism = sc.add_material(dielectric(1.5, abbe=30.0))
    sc.add_shape(prism_mesh((-1.1, 0.002, 0.0), 0.45, 0.9, prism_subdiv), prism)
    # Tinted liquid: sigma_a ~ U[0.1, 2] per cm per channel (U[10, 200] per m; scene units are metres), given as
    # Dielectric::setAbsorption(exp(-sigma_a * 1 cm), 1 cm) (dielectric.cpp:123-144 recovers sigma_a).
    sigma_a = rng.uniform(10.0, 200.0, 3)
    liquid = sc.add_material(dielectric(1.33, absorption_color=np.exp(-sigma_a * 0.01), absorption_distance=0.01))
    sc.add_shape(box_mesh((-0.45, 0.002, -0.2), (0.05, 0.6, 0.3)), liquid)
    sss = sc.add_material(disney((0.85, 0.55, 0.4), 0.0, 0.4, 0.5, subsurface=1.0,
                                 subsurface_radius=rng.uniform(1.0, 5.0, 3), subsurface_scale=1e-3))
    sc.add_shape(box_mesh((0.3, 0.002, -0.25), (0.8, 0.5, 0.25)), sss)
    pane = sc.add_material(dielectric(1.5, tint=(0.9, 0.95, 1.0), thin_walled=True))
    sc.add_shape(_quad((0.9, 0.0, 0.5), (1.7, 0.0, 0.2), (1.7, 0.9, 0.2), (0.9, 0.9, 0.5)), pane)
    rough = sc.add_material(dielectric(1.45, tint=(1.0, 0.97, 0.9), roughness=0.25))
    sc.add_shape(icosphere(2, 0.25, (1.15, 0.25, -0.55)), rough)
    trans = sc.add_material(disney((0.5, 0.8, 0.4), 0.0, 0.35, 0.5, specular_transmission=0.6,
                                   diffuse_transmission=0.3, ior=1.4))
    sc.add_shape(icosphere(2, 0.22, (-0.2, 0.85, 0.6)), trans)
    sc.add_light(AreaLight((-0.6, 2.5, -0.2), (1.2, 0.0, 0.0), (0.0, 0.0, 0.8), (12.0, 12.0, 12.0)))
    if led_quads > 0:
        led = sc.add_material(lambert((0.0, 0.0, 0.0), (3.0, 2.5, 2.0)))
        t = np.arange(led_quads, dtype=np.float32) / np.float32(led_quads)
        a, r = t * np.float32(12.0), np.float32(0.6) + np.float32(2.2) * t
        c = np.stack([r * np.cos(a), np.full_like(t, 1.6), r * np.sin(a) - np.float32(0.3)], 1).astype(np.float32)
        off = np.array([[-0.03, 0, -0.03], [-0.03, 0, 0.03], [0.03, 0, 0.03], [0.03, 0, -0.03]], np.float32)
        pos = (c[:, None, :] + off[None, :, :]).reshape(-1, 3)
        b = (np.arange(led_quads, dtype=np.uint32) * 4)[:, None]
        idx = np.concatenate([b, b + 1, b + 2, b, b + 2, b + 3], 1).reshape(-1, 3).astype(np.uint32)
        sc.add_shape(Mesh(pos, idx), led)
    view = View((0.2, 1.3, 3.4), (0.0, 0.45, 0.0), (0.0, 1.0, 0.0), 45.0, width, height, spp, max_bounces)
    return sc, view


def lights_media_studio(width: int = 1280, height: int = 720, spp: int = 512, max_bounces: int = 16) -> tuple[Scene, View]:
    """The many-lights + media workload at full size: the glass studio with 1000 LED quads (2000 emissive
    triangles, light BVH) and a 200,934-triangle dispersive prism (sides split 183 x 183), 1280 x 720,
    512 spp, depth 16."""
    return glass_studio(width, height, spp, max_bounces, led_quads=1000, prism_subdiv=183)


# ---- slit room (learned guiding) -------------------------------------------------------------------------

def grid_quad(p0, u, v, nu: int, nv: int) -> Mesh:
    """The parallelogram p0 + s u + t v (s, t in [0, 1]) split into nu x nv quads (2 nu nv triangles); the
    triangles face u x v."""
    s = np.linspace(0.0, 1.0, nu + 1, dtype=np.float64)
    t = np.linspace(0.0, 1.0, nv + 1, dtype=np.float64)
    ss, tt = np.meshgrid(s, t, indexing="xy")
    p = (np.asarray(p0, np.float64)[None, None, :] + ss[..., None] * np.asarray(u, np.float64) +
         tt[..., None] * np.asarray(v, np.float64)).reshape(-1, 3).astype(np.float32)
    j, i = np.meshgrid(np.arange(nv), np.arange(nu), indexing="ij")
    k = (j * (nu + 1) + i).ravel()
    idx = np.concatenate([np.stack([k, k + 1, k + nu + 2], 1), np.stack([k, k + nu + 2, k + nu + 1], 1)])
    return Mesh(p, idx.astype(np.uint32))


def slit_room(width: int = 1920, height: int = 1080, spp: int = 64, max_bounces: int = 10, seed: int = 4,
              cell: float = 0.018, albedo: float = 1.0) -> tuple[Scene, View]:
    """The learned-guiding workload: a closed 4 x 4 x 3 m room (walls, floor and ceiling tessellated into
    `cell`-sized quads: about 500k triangles at the default), diffuse and glossy Disney surfaces and a few
    boxes, lit only through a 5 cm gap — a door left ajar in the +x wall, 2.1 m high — by an area light outside
    it. Everything inside is lit by the thin wedge of light the gap lets in, mostly after one or more diffuse
    bounces."""
    rng = np.random.default_rng(seed)
    sc = Scene()
    a = float(albedo)
    wall = sc.add_material(disney((0.72 * a, 0.7 * a, 0.66 * a), 0.0, 0.85, 0.5))
    floor = sc.add_material(disney((0.45 * a, 0.38 * a, 0.3 * a), 0.0, 0.3, 0.5))
    ceil = sc.add_material(disney((0.8 * a, 0.8 * a, 0.8 * a), 0.0, 0.9, 0.5))
    X, Y, Z = 2.0, 3.0, 2.0

    def face(p0, u, v, mat):
        nu = max(1, int(round(np.linalg.norm(u) / cell)))
        nv = max(1, int(round(np.linalg.norm(v) / cell)))
        sc.add_shape(grid_quad(p0, u, v, nu, nv), mat)

    face((-X, 0, -Z), (0, 0, 2 * Z), (2 * X, 0, 0), floor)          # floor, +y
    face((-X, Y, -Z), (2 * X, 0, 0), (0, 0, 2 * Z), ceil)           # ceiling, -y
    face((-X, 0, -Z), (2 * X, 0, 0), (0, Y, 0), wall)               # z = -2, +z
    face((-X, 0, Z), (0, Y, 0), (2 * X, 0, 0), wall)                # z = +2, -z
    face((-X, 0, -Z), (0, Y, 0), (0, 0, 2 * Z), wall)               # x = -2, +x
    g0, g1, gh = 0.40, 0.45, 2.1                                     # the gap: z in [g0, g1], y in [0, gh]
    face((X, 0, -Z), (0, 0, g0 + Z), (0, Y, 0), wall)               # x = +2, left of the gap
    face((X, 0, g1), (0, 0, Z - g1), (0, Y, 0), wall)               # right of it
    face((X, gh, g0), (0, 0, g1 - g0), (0, Y - gh, 0), wall)        # above it
    for k in range(5):                                               # furniture
        c = rng.uniform([-1.5, -1.5], [1.0, 1.5])
        sz = rng.uniform([0.3, 0.3, 0.3], [0.8, 1.1, 0.8])
        glossy = bool(rng.integers(2))
        m = sc.add_material(disney(rng.uniform(0.2, 0.8, 3), 0.0, 0.15 if glossy else 0.7, 0.5))
        sc.add_shape(box_mesh((c[0], 0.0, c[1]), (c[0] + sz[0], sz[1], c[1] + sz[2])), m)
    # Outside: a light 1 m beyond the gap, facing it (-x), wider and taller than the gap.
    sc.add_light(AreaLight((3.0, 0.2, -0.8), (0.0, 0.0, 2.4), (0.0, 2.4, 0.0), (120.0, 114.0, 100.0)))
    view = View((-1.7, 1.5, -1.6), (0.9, 0.6, 0.9), (0.0, 1.0, 0.0), 70.0, width, height, spp, max_bounces)
    return sc, view


# ---- product_sphere --------------------------------------------------------------------------------------

def _perlin3(p: np.ndarray, perm: np.ndarray) -> np.ndarray:
    """Improved Perlin noise (Perlin 2002) on points p (N, 3), permutation perm (512,)."""
    pi = np.floor(p).astype(np.int64)
    f = p - pi
    pi &= 255
    u = f * f * f * (f * (f * 6.0 - 15.0) + 10.0)

    def grad(h, x, y, z):
        h = h & 15
        a = np.where(h < 8, x, y)
        b = np.where(h < 4, y, np.where((h == 12) | (h == 14), x, z))
        return np.where(h & 1, -a, a) + np.where(h & 2, -b, b)

    X, Y, Z = pi[:, 0], pi[:, 1], pi[:, 2]
    x, y, z = f[:, 0], f[:, 1], f[:, 2]
    A = perm[X] + Y
    AA = perm[A] + Z
    AB = perm[A + 1] + Z
    B = perm[X + 1] + Y
    BA = perm[B] + Z
    BB = perm[B + 1] + Z

    def lerp(t, a, b):
        return a + t * (b - a)

    return lerp(u[:, 2],
                lerp(u[:, 1], lerp(u[:, 0], grad(perm[AA], x, y, z), grad(perm[BA], x - 1, y, z)),
                     lerp(u[:, 0], grad(perm[AB], x, y - 1, z), grad(perm[BB], x - 1, y - 1, z))),
                lerp(u[:, 1], lerp(u[:, 0], grad(perm[AA + 1], x, y, z - 1), grad(perm[BA + 1], x - 1, y, z - 1)),
                     lerp(u[:, 0], grad(perm[AB + 1], x, y - 1, z - 1), grad(perm[BB + 1], x - 1, y - 1, z - 1))))


def displaced_sphere(lon: int = 1024, lat: int = 1000, parts: int = 16, amplitude: float = 0.05,
                     seed: int = 1) -> list[Mesh]:
    """UV sphere with 2*lon*(lat-1) triangles, radius 1 + amplitude * Perlin(4p), as `parts` longitude wedges."""
    rng = np.random.default_rng(seed)
    perm = rng.permutation(256)
    perm = np.concatenate([perm, perm]).astype(np.int64)
    theta = np.linspace(0.0, np.pi, lat + 1)[1:-1]                     # interior rings
    phi = np.linspace(0.0, 2.0 * np.pi, lon + 1)                       # lon + 1 columns (seam duplicated)
    st, ct = np.sin(theta)[:, None], np.cos(theta)[:, None]
    ring = np.stack([st * np.cos(phi)[None, :], np.broadcast_to(ct, (lat - 1, lon + 1)), st * np.sin(phi)[None, :]], -1)
    dirs = np.concatenate([ring.reshape(-1, 3), [[0.0, 1.0, 0.0], [0.0, -1.0, 0.0]]])
    r = 1.0 + amplitude * _perlin3(4.0 * dirs, perm)
    pos_all = dirs * r[:, None]
    rows, cols = lat - 1, lon + 1
    north, south = rows * cols, rows * cols + 1

    def vid(i, j):
        return i * cols + j

    tris = []
    for j in range(lon):
        tris.append((north, vid(0, j + 1), vid(0, j)))
        for i in range(rows - 1):
            a, b, c, d = vid(i, j), vid(i, j + 1), vid(i + 1, j), vid(i + 1, j + 1)
            tris.append((a, b, d))
            tris.append((a, d, c))
        tris.append((south, vid(rows - 1, j), vid(rows - 1, j + 1)))
    tris = np.array(tris, dtype=np.int64).reshape(lon, -1, 3)
    # Smooth vertex normals: area-weighted face normals, accumulated per vertex.
    t = tris.reshape(-1, 3)
    fn = np.cross(pos_all[t[:, 1]] - pos_all[t[:, 0]], pos_all[t[:, 2]] - pos_all[t[:, 0]])
    nrm = np.zeros_like(pos_all)
    for k in range(3):
        np.add.at(nrm, t[:, k], fn)
    # Seam columns share the normal of their twin.
    seam = nrm[vid(np.arange(rows), 0)] + nrm[vid(np.arange(rows), lon)]
    nrm[vid(np.arange(rows), 0)] = seam
    nrm[vid(np.arange(rows), lon)] = seam
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    meshes = []
    per = lon // parts
    for p in range(parts):
        pt = tris[p * per:(p + 1) * per].reshape(-1, 3)
        used, local = np.unique(pt, return_inverse=True)
        meshes.append(Mesh(pos_all[used].astype(np.float32), local.reshape(-1, 3).astype(np.uint32),
                           nrm[used].astype(np.float32)))
    return meshes


def sky_with_sun(width: int = 2048, height: int = 1024, sun_elevation_deg: float = 45.0,
                 sun_azimuth_deg: float = 30.0, sun_diameter_deg: float = 0.5, sun_scale: float = 1e4) -> np.ndarray:
    """Sky gradient (zenith (0.3, 0.5, 1.0) to horizon 1.0, ground 0.2) with a sun disk baked in at 1e4 x."""
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    theta = v * np.pi
    elev = np.pi / 2 - theta
    zen = np.array([0.3, 0.5, 1.0])
    hor = np.array([1.0, 1.0, 1.0])
    s = np.clip(elev / (np.pi / 2), 0.0, 1.0)[:, None]
    sky = hor[None, :] * (1 - s) + zen[None, :] * s
    sky = np.where((elev < 0)[:, None], np.array([0.2, 0.2, 0.2])[None, :], sky)
    img = np.repeat(sky[:, None, :], width, axis=1)
    # Direction of every texel (environment_light.h:33-39 equirectDirection with rotation 0).
    a = u[None, :] * 2 * np.pi - np.pi
    st = np.sin(theta)[:, None]
    stw = np.broadcast_to(st, (height, width))
    ct = np.broadcast_to(np.cos(theta)[:, None], (height, width))
    d = np.stack([stw * np.cos(a), ct, -stw * np.sin(a)], -1)
    se, sa = np.radians(sun_elevation_deg), np.radians(sun_azimuth_deg)
    sun = np.array([np.cos(se) * np.cos(sa), np.sin(se), np.cos(se) * np.sin(sa)])
    cosmax = np.cos(np.radians(sun_diameter_deg / 2))
    disk = (d @ sun) >= cosmax
    img[disk] *= sun_scale
    return img.astype(np.float32)


def product_sphere(seed: int = 1, width: int = 1920, height: int = 1080, spp: int = 256, max_bounces: int = 12,
                   lon: int = 1024, lat: int = 1000, env_width: int = 2048, env_height: int = 1024) -> tuple[Scene, View]:
    rng = np.random.default_rng(seed + 100)
    sc = Scene()
    for mesh in displaced_sphere(lon, lat, 16, 0.05, seed):
        m = sc.add_material(disney(rng.uniform(0.2, 0.9, 3), float(rng.random() < 0.5), rng.uniform(0.05, 0.8), 0.5,
                                   clearcoat=float(rng.uniform(0.0, 1.0))))
        sc.add_shape(mesh, m)
    sc.set_environment(Environment(sky_with_sun(env_width, env_height)))
    view = View((0.0, 0.35, 3.2), (0.0, 0.0, 0.0), (0.0, 1.0, 0.0), 40.0, width, height, spp, max_bounces)
    return sc, view


def pass_studio(width: int = 64, height: int = 48, spp: int = 16, max_bounces: int = 6, level: int = 3,
                groups: bool = True) -> tuple[Scene, View]:
    """The render-pass test scene (tests/test_render_passes.cpp:33-53): a Lambert floor, a rough gold metal
    sphere, a smooth glass sphere, an emissive quad seen by the camera, an area light and a flat grey
    environment; spheres as icospheres. With groups the area light is in light group "Key" (the quad's mesh
    lights fall into "Default", the environment into "Env"), every shape carries a node uid, and the floor and
    the metal sphere and their materials carry names (cryptomatte: named, uid-named and anonymous objects)."""
    sc = Scene()
    up = np.array([[0.0, 1.0, 0.0]] * 4, dtype=np.float32)
    floor = sc.add_material(lambert((0.6, 0.6, 0.6)), name="FloorMat" if groups else "")
    metal = sc.add_material(disney((0.95, 0.8, 0.5), 1.0, 0.25, 0.5), name="Gold" if groups else "")
    glass = sc.add_material(dielectric(1.5, (1.0, 1.0, 1.0), 0.0))
    emitter = sc.add_material(lambert((0.0, 0.0, 0.0), (6.0, 6.0, 6.0)))
    fp = np.array([(-4, 0, -4), (4, 0, -4), (4, 0, 4), (-4, 0, 4)], dtype=np.float32)
    sc.add_shape(Mesh(fp, np.array([[0, 2, 1], [0, 3, 2]], dtype=np.uint32), up), floor, node_uid=11 if groups else 0,
                 name="Floor" if groups else "")
    sc.add_shape(icosphere(level, 0.6, (-0.8, 0.6, 0.0)), metal, node_uid=12 if groups else 0,
                 name="Hero_Metal" if groups else "")
    sc.add_shape(icosphere(level, 0.6, (0.8, 0.6, 0.0)), glass, node_uid=13 if groups else 0)
    ep = np.array([(-0.4, 1.8, -1.5), (0.4, 1.8, -1.5), (0.4, 2.4, -1.5), (-0.4, 2.4, -1.5)], dtype=np.float32)
    sc.add_shape(Mesh(ep, np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)), emitter, node_uid=14 if groups else 0)
    sc.add_light(AreaLight((-2.0, 3.0, 1.0), (1.0, 0.0, 0.0), (0.0, 0.0, 1.0), (8.0, 8.0, 8.0),
                           light_group="Key" if groups else ""))
    sc.set_environment(Environment(None, 0.0, 0.4))
    view = View((0.0, 1.4, 4.5), (0.0, 0.8, 0.0), (0.0, 1.0, 0.0), 45.0, width, height, spp, max_bounces)
    return sc, view


def tiny_scene(seed: int = 7, n_tris: int = 64) -> Scene:
    """A small random triangle soup (for hit parity against brute force)."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n_tris, 1, 3))
    tri = c + rng.normal(0, 0.25, (n_tris, 3, 3))
    sc = Scene()
    m = sc.add_material(disney((0.5, 0.5, 0.5)))
    sc.add_shape(Mesh(tri.reshape(-1, 3).astype(np.float32), np.arange(3 * n_tris, dtype=np.uint32).reshape(-1, 3)), m)
    return sc
