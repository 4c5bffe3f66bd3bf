"""Closest hit and any hit at the workload's size, and traversal stacks deep enough to spill out of LDS.

* The product-sphere mesh itself (2,045,952 triangles in 16 meshes, the 1080p workload's scene): camera rays
  over the whole frame plus cosine-distributed bounce rays leaving the surface, GPU against the oracle
  (geometry/bvh.cpp:322-446, triangle.cpp:11-50). Hits must agree bit for bit in t and the primitive id,
  except genuine ties (equal t, another triangle: the reference keeps the last one tested, an order that
  depends on the tree).
* A nested soup whose BVH boxes all overlap, so that a ray enters every child of every node: traversal pushes
  three entries per four-wide level and its stack passes the 16 entries kept in LDS into the global spill
  column. The device's deepest-stack statistic proves the spill was used; the hits still match the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import configs
from photon_hip.scene import Mesh, Scene, disney

pytestmark = pytest.mark.gpu

LDS_STACK_ENTRIES = 16  # kLdsDepth in csrc/pe_traverse.h


def _same_hits(a, b, allow_ties=False):
    eq = (a[:, 0] == b[:, 0]) & (a[:, 1] == b[:, 1]) & (a[:, 2] == b[:, 2])
    if allow_ties:
        eq |= (a[:, 0] == b[:, 0]) & (a[:, 1] >= 0) & (b[:, 1] >= 0)
    return eq


def _camera_rays(view, n, rng):
    """Pinhole rays through uniformly drawn film positions of the view."""
    pos, tgt, up = (np.asarray(v, np.float64) for v in (view.position, view.target, view.up))
    w = pos - tgt
    w /= np.linalg.norm(w)
    u = np.cross(up, w)
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    h = np.tan(np.radians(view.vfov) / 2)
    sx, sy = rng.uniform(-1, 1, n) * h * view.aspect, rng.uniform(-1, 1, n) * h
    d = sx[:, None] * u + sy[:, None] * v - w
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, :3], r[:, 3:6], r[:, 7] = pos, d, 1e30
    return r


def _bounce_rays(scene, rays, hits, rng):
    """Cosine-distributed rays leaving each hit point on the side the ray came from (geometric normal)."""
    keep = np.flatnonzero(hits[:, 1] >= 0)
    t = hits[keep, 0].view(np.float32).astype(np.float64)
    o = rays[keep, :3].astype(np.float64) + rays[keep, 3:6] * t[:, None]
    ng = np.zeros((keep.size, 3))
    for s, shape in enumerate(scene.shapes):
        sel = hits[keep, 1] == s
        if not sel.any():
            continue
        tri = shape.mesh.indices[hits[keep[sel], 2]]
        p = shape.mesh.positions.astype(np.float64)
        ng[sel] = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    ng /= np.linalg.norm(ng, axis=1, keepdims=True)
    ng *= -np.sign(np.einsum("ij,ij->i", ng, rays[keep, 3:6]))[:, None]
    a = np.where(np.abs(ng[:, :1]) > 0.9, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    tx = np.cross(ng, a)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(ng, tx)
    u1, u2 = rng.random(keep.size), rng.random(keep.size)
    rr, phi = np.sqrt(u1), 2 * np.pi * u2
    d = tx * (rr * np.cos(phi))[:, None] + ty * (rr * np.sin(phi))[:, None] + ng * np.sqrt(1 - u1)[:, None]
    out = np.zeros((keep.size, 8), np.float32)
    out[:, :3] = o + ng * 1e-4
    out[:, 3:6] = d
    out[:, 7] = 1e30
    return out


def test_gpu_hits_match_oracle_on_product_sphere_mesh(dev):
    sc, view = configs.product_sphere()
    assert sc.num_triangles == 2_045_952
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    rng = np.random.default_rng(31)
    cam = _camera_rays(view, 250_000, rng)
    cam_o, _ = o.intersect(cam)
    rays = np.concatenate([cam, _bounce_rays(sc, cam, cam_o, rng)])
    hg, bg = gs.intersect(rays)
    depth, overflows = dev.trace_diag()
    ho, bo = o.intersect(rays)
    ok = _same_hits(hg[:, :3], ho[:, :3], allow_ties=True)
    same = _same_hits(hg[:, :3], ho[:, :3])
    print(f"{rays.shape[0]} rays ({cam.shape[0]} camera), {np.mean(ho[:, 1] >= 0):.3f} hit; identical "
          f"{same.mean():.6f}, ties {np.sum(ok & ~same)}; deepest stack {depth}", flush=True)
    assert overflows == 0
    assert ok.all(), np.flatnonzero(~ok)[:10]
    assert same.mean() > 0.999
    np.testing.assert_array_equal(bg[same].view(np.uint32), bo[same].view(np.uint32))
    # Any hit: the shadow rays of the same set toward a point 0.7 of the way to the closest hit (occluded only
    # by something nearer) and unbounded.
    sh = rays.copy()
    t = ho[:, 0].view(np.float32)
    sh[:, 7] = np.where(ho[:, 1] >= 0, 0.7 * t, 1e30)
    np.testing.assert_array_equal(gs.occluded(sh), o.occluded(sh))
    np.testing.assert_array_equal(gs.occluded(rays), o.occluded(rays))


def _nested_soup(n, rng):
    """n large triangles all spanning the unit cube, so every BVH box overlaps every other."""
    c = rng.uniform(-0.05, 0.05, (n, 1, 3))
    tri = c + rng.uniform(-1.0, 1.0, (n, 3, 3))
    sc = Scene()
    m = sc.add_material(disney((0.5, 0.5, 0.5)))
    sc.add_shape(Mesh(tri.reshape(-1, 3).astype(np.float32), np.arange(3 * n, dtype=np.uint32).reshape(-1, 3)), m)
    return sc


def test_gpu_deep_stacks_spill_to_global_memory_and_match_oracle(dev):
    rng = np.random.default_rng(8)
    sc = _nested_soup(12_000, rng)
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    n = 20_000
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3] = rng.uniform(-1.5, 1.5, (n, 3))
    d = rng.normal(size=(n, 3))
    rays[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    rays[:, 7] = 1e30
    hg, _ = gs.intersect(rays)
    depth, overflows = dev.trace_diag()
    ho, _ = o.intersect(rays)
    ok = _same_hits(hg[:, :3], ho[:, :3], allow_ties=True)
    print(f"nested soup: deepest traversal stack {depth} entries (LDS holds {LDS_STACK_ENTRIES}); "
          f"identical {_same_hits(hg[:, :3], ho[:, :3]).mean():.6f}", flush=True)
    assert overflows == 0
    assert depth > LDS_STACK_ENTRIES, "the spill column was not exercised"
    assert ok.all(), np.flatnonzero(~ok)[:10]
    np.testing.assert_array_equal(gs.occluded(rays), o.occluded(rays))


def test_render_reports_the_deepest_stack_of_the_workload(dev):
    """The counting render variant reports the deepest stack over a whole frame of the workload (1 spp)."""
    from photon_hip import camera_perspective, render, settings

    sc, view = configs.product_sphere()
    gs = dev.upload(sc)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=1)
    st.count_visits = 1
    film = dev.film(view.width, view.height)
    r = render(dev, gs, cam, st, film, 0, 1)
    print(f"1080p frame: deepest traversal stack {r.max_stack_depth}, overflows {r.stack_overflows}", flush=True)
    assert r.stack_overflows == 0
    assert 0 < r.max_stack_depth < 128
