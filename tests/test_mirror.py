"""Mirror (materials/mirror.h, mirror.cpp): a delta reflection with weight R, on the integrator path as a delta
vertex (no light sample, emission seen after it counted in full, path_tracer.cpp:624-698).

CPU: the oracle's Mirror equals the reference's compiled Mirror bit for bit (sample and eval, 3000 random
frames); closed forms (wi = 2 (wo.n) n - wo, f |cos| / pdf = R, no sample from below); the oracle's per-sample
radiance on a scene with a flat and a curved mirror equals the reference's renderer (IEEE build).
GPU: the kernels against the oracle per sample and on a converged image.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi
from photon_hip.configs import View, _quad, icosphere
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import AreaLight, Mesh, Scene, disney, lambert, mirror

VIEW = View((0.3, 1.2, 3.2), (0.0, 0.5, 0.0), (0.0, 1.0, 0.0), 45.0, 48, 36, 16, 16)


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def mirror_scene() -> Scene:
    sc = Scene()
    up = np.array([0.0, 1.0, 0.0], np.float32)
    sc.add_shape(Mesh(np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)),
                 sc.add_material(lambert((0.6, 0.6, 0.6))))
    # A flat mirror facing the camera, a curved chrome ball and a diffuse ball that both reflect.
    sc.add_shape(_quad((-1.4, 0.0, -0.8), (1.4, 0.0, -0.8), (1.4, 1.6, -0.8), (-1.4, 1.6, -0.8)),
                 sc.add_material(mirror((0.9, 0.85, 0.8))))
    sc.add_shape(icosphere(3, 0.35, (-0.5, 0.35, 0.3)), sc.add_material(mirror()))
    sc.add_shape(icosphere(3, 0.3, (0.55, 0.3, 0.1)), sc.add_material(disney((0.8, 0.3, 0.2), 0.0, 0.4, 0.5)))
    sc.add_light(AreaLight((-0.5, 2.4, -0.2), (1.0, 0.0, 0.0), (0.0, 0.0, 0.7), (14.0, 14.0, 14.0)))
    return sc


def _random_frames(rng, n):
    d = rng.normal(size=(4, n, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    normal, tv, wo, wi = (x.astype(np.float32) for x in d)
    tangent = np.cross(normal, tv).astype(np.float32)
    front = (rng.random(n) < 0.7).astype(np.int32)
    return normal, tangent, front, wo, wi


def test_mirror_bsdf_equals_reference():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    rng = np.random.default_rng(5)
    n = 3000
    normal, tangent, front, wo, wi = _random_frames(rng, n)
    u3 = rng.random((n, 3)).astype(np.float32)
    m = mirror((0.9, 0.7, 0.5))
    so = bindings.bsdf_sample("oracle", m, normal, tangent, front, wo, u3)
    sr = bindings.bsdf_sample("ref_strict", m, normal, tangent, front, wo, u3)
    np.testing.assert_array_equal(so, sr)
    assert 0.3 < (sr[:, 0] > 0).mean() < 0.7     # both sides of the surface
    fo, po = bindings.bsdf_eval("oracle", m, normal, tangent, front, wo, wi)
    fr, pr = bindings.bsdf_eval("ref_strict", m, normal, tangent, front, wo, wi)
    np.testing.assert_array_equal(fo, fr)
    np.testing.assert_array_equal(po, pr)
    assert (fo == 0).all() and (po == 0).all()    # a Dirac delta has no finite value


def test_mirror_closed_form():
    rng = np.random.default_rng(6)
    n = 2000
    normal, tangent, front, wo, _ = _random_frames(rng, n)
    R = np.float32([0.9, 0.7, 0.5])
    s = bindings.bsdf_sample("oracle", mirror(tuple(R)), normal, tangent, front, wo, rng.random((n, 3)).astype(np.float32))
    cos_o = np.sum(wo * normal, 1)
    ok = s[:, 0] > 0
    np.testing.assert_array_equal(ok, cos_o > 0)   # from below the surface there is no sample
    wi = s[:, 1:4]
    refl = 2.0 * cos_o[:, None] * normal - wo
    np.testing.assert_allclose(wi[ok], refl[ok], atol=2e-6)
    # f |cos| / pdf = R (the cosine recomputed in world space: a few ulp apart from the local one at grazing angles)
    g = ok & (cos_o > 0.02)
    w = s[g, 4:7] * np.abs(np.sum(wi[g] * normal[g], 1))[:, None] / s[g, 7:8]
    np.testing.assert_allclose(w, np.broadcast_to(R, w.shape), rtol=1e-4)
    np.testing.assert_array_equal(s[ok, 7], 1.0)


def test_oracle_mirror_scene_equals_reference_per_sample():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc = mirror_scene()
    rng = np.random.default_rng(8)
    n = 3000
    px, py, s = rng.integers(0, VIEW.width, n), rng.integers(0, VIEW.height, n), rng.integers(0, 256, n)
    st = settings(max_bounces=VIEW.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(VIEW), st, VIEW.width, VIEW.height, px, py, s)
    Lr = bindings.RefScene(sc, strict=True).trace_samples(VIEW, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    assert Lo.max() > 0


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_mirror_scene_matches_oracle(dev):
    sc = mirror_scene()
    gs = dev.upload(sc)
    yy, xx = np.mgrid[0:VIEW.height, 0:VIEW.width]
    o = bindings.OracleScene(sc)
    for sample in (0, 7):
        film = dev.film(VIEW.width, VIEW.height)
        res = render(dev, gs, _cam(VIEW), settings(max_bounces=VIEW.max_bounces, seed=12345), film, sample, 1)
        s, c, _, rej = film.download()
        assert res.completed == 1 and (c == 1).all() and rej == 0
        Lo, _ = o.trace_samples(_cam(VIEW), settings(max_bounces=VIEW.max_bounces, seed=12345), VIEW.width,
                                VIEW.height, xx.ravel(), yy.ravel(), np.full(xx.size, sample))
        Lg = s.reshape(-1, 3)
        close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
        assert close.mean() >= 0.98, close.mean()
    spp = 256
    film = dev.film(VIEW.width, VIEW.height)
    render(dev, gs, _cam(VIEW), settings(max_bounces=VIEW.max_bounces, seed=12345, samples_per_launch=64), film, 0, spp)
    s, c, _, _ = film.download()
    gpu = s / c[..., None]
    (so, co, _), _, _ = o.render(_cam(VIEW), settings(max_bounces=VIEW.max_bounces, seed=12345), VIEW.width,
                                 VIEW.height, 0, spp, threads=8)
    ora = so / co[..., None]
    rel = float(np.mean((gpu - ora) ** 2 / (ora ** 2 + 0.01)))
    assert rel <= 1e-3, rel
