"""Thin-film interference (materials/fresnel.h:74-107): the Airy sum of a film between air and a dielectric or
metal base, averaged over three wavelengths per channel, in the Disney base specular Fresnel
(materials/disney.cpp:79-96) and in the soap-bubble wall of a thin-walled Dielectric (dielectric.cpp:150-158).

CPU: the oracle's restatement against the reference's own functions bit for bit; closed forms of the Airy sum
(zero thickness is the bare interface, a half-wave film is absent, a quarter-wave film of index sqrt(n3)
cancels normal reflection); Disney and Dielectric BSDFs with films and per-sample radiance against the
reference's strict build. GPU: the kernels against the oracle.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi
from photon_hip.configs import View, _quad, icosphere
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import AreaLight, Mesh, Scene, dielectric, disney, lambert


def _ref():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


def fresnel_unpolarised(cos1, n1, n3):
    """Bare-interface Fresnel reflectance in float64 (complex n3 allowed)."""
    cos1 = np.asarray(cos1, np.float64)
    sin1sq = 1 - cos1 ** 2
    cos3 = np.sqrt(1 - (n1 * n1 * sin1sq) / (n3 * n3) + 0j)
    rs = (n1 * cos1 - n3 * cos3) / (n1 * cos1 + n3 * cos3)
    rp = (n3 * cos1 - n1 * cos3) / (n3 * cos1 + n1 * cos3)
    return 0.5 * (np.abs(rs) ** 2 + np.abs(rp) ** 2)


@pytest.mark.parametrize("base", [1.0, 1.52, 2.4, complex(0.2, 3.1), complex(1.1, 2.4)])
def test_thin_film_reflectance_equals_reference(base):
    _ref()
    rng = np.random.default_rng(5)
    n = 5000
    cos1 = rng.uniform(0, 1, n).astype(np.float32)
    t = rng.uniform(0, 2000, n).astype(np.float32)
    lam = rng.uniform(380, 730, n).astype(np.float32)
    for n2 in (1.0001, 1.33, 1.8):
        a = bindings.thin_film("oracle", cos1, 1.0, n2, base, t, lam)
        b = bindings.thin_film("ref_strict", cos1, 1.0, n2, base, t, lam)
        np.testing.assert_array_equal(a, b)


def test_thin_film_closed_forms():
    cos1 = np.linspace(0.05, 1.0, 64).astype(np.float32)
    for n3 in (1.5, complex(0.3, 3.0)):
        # zero thickness: the film vanishes, leaving the bare air -> base interface
        r0 = bindings.thin_film("oracle", cos1, 1.0, 1.38, n3, 0.0, 550.0)
        np.testing.assert_allclose(r0, fresnel_unpolarised(cos1, 1.0, n3), atol=2e-6)
    # a half-wave film (n2 d = lambda / 2) is absent at normal incidence
    n2, lam = 1.38, 550.0
    r = bindings.thin_film("oracle", np.ones(1, np.float32), 1.0, n2, 1.52, lam / (2 * n2), lam)
    np.testing.assert_allclose(r, fresnel_unpolarised(1.0, 1.0, 1.52), atol=2e-6)
    # a quarter-wave film of index sqrt(n3) cancels normal reflection
    n3 = 1.69
    n2 = np.sqrt(n3)
    r = bindings.thin_film("oracle", np.ones(1, np.float32), 1.0, n2, n3, lam / (4 * n2), lam)
    assert r[0] < 1e-6, r
    # total internal reflection in the film
    r = bindings.thin_film("oracle", np.full(4, 0.1, np.float32), 1.5, 1.0, 1.5, 100.0, 500.0)
    np.testing.assert_array_equal(r, 1.0)


FILMS = {
    "iridescent_plastic": disney((0.2, 0.25, 0.3), 0.0, 0.2, 0.5, thin_film_thickness=420.0, thin_film_ior=1.33),
    "iridescent_metal": disney((0.9, 0.7, 0.4), 1.0, 0.25, 0.5, thin_film_thickness=310.0, thin_film_ior=1.5,
                               edge_tint=(0.8, 0.6, 0.5)),
    "bubble_wall": dielectric(1.33, thin_walled=True, thin_film_thickness=500.0),
}


@pytest.mark.parametrize("name", list(FILMS))
def test_film_bsdf_equals_reference(name):
    _ref()
    m = FILMS[name]
    rng = np.random.default_rng(11)
    n = 3000
    d = rng.normal(size=(4, n, 3))
    d /= np.linalg.norm(d, axis=1 if d.ndim == 2 else 2, keepdims=True)
    normal, tv, wo, wi = (x.astype(np.float32) for x in d)
    tangent = np.cross(normal, tv).astype(np.float32)
    front = (rng.random(n) < 0.7).astype(np.int32)
    u3 = rng.random((n, 3)).astype(np.float32)
    np.testing.assert_array_equal(bindings.bsdf_sample("oracle", m, normal, tangent, front, wo, u3),
                                  bindings.bsdf_sample("ref_strict", m, normal, tangent, front, wo, u3))
    fo, po = bindings.bsdf_eval("oracle", m, normal, tangent, front, wo, wi)
    fr, pr = bindings.bsdf_eval("ref_strict", m, normal, tangent, front, wo, wi)
    np.testing.assert_array_equal(fo, fr)
    np.testing.assert_array_equal(po, pr)


def film_scene() -> tuple[Scene, View]:
    sc = Scene()
    floor = sc.add_material(lambert((0.6, 0.6, 0.6)))
    up = np.array([0.0, 1.0, 0.0], np.float32)
    sc.add_shape(Mesh(np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)), floor)
    sc.add_shape(icosphere(3, 0.45, (-0.6, 0.45, 0.0)), sc.add_material(FILMS["iridescent_plastic"]))
    sc.add_shape(icosphere(3, 0.35, (0.5, 0.35, -0.3)), sc.add_material(FILMS["iridescent_metal"]))
    sc.add_shape(_quad((0.2, 0.0, 0.6), (1.2, 0.0, 0.4), (1.2, 0.9, 0.4), (0.2, 0.9, 0.6)),
                 sc.add_material(FILMS["bubble_wall"]))
    sc.add_light(AreaLight((-0.6, 2.5, -0.2), (1.2, 0.0, 0.0), (0.0, 0.0, 0.8), (12.0, 12.0, 12.0)))
    return sc, View((0.2, 1.3, 3.4), (0.0, 0.45, 0.0), (0.0, 1.0, 0.0), 45.0, 64, 48, 16, 12)


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def test_film_scene_radiance_equals_reference_per_sample():
    _ref()
    sc, view = film_scene()
    rng = np.random.default_rng(3)
    n = 4000
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 256, n)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = bindings.RefScene(sc, strict=True).trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    assert Lo.max() > 0


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


@pytest.mark.gpu
def test_gpu_film_scene_matches_oracle(dev):
    sc, view = film_scene()
    gs = dev.upload(sc)
    sample = 2
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=1)
    render(dev, gs, _cam(view), st, film, sample, 1)
    s, c, _, rej = film.download()
    assert (c == 1).all() and rej == 0
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), settings(max_bounces=view.max_bounces, seed=12345),
                                                   view.width, view.height, xx.ravel(), yy.ravel(),
                                                   np.full(xx.size, sample))
    Lg = s.reshape(-1, 3)
    close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
    assert close.mean() >= 0.97, close.mean()
    spp = 256
    film = dev.film(view.width, view.height)
    render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=64), film, 0, spp)
    s, c, _, _ = film.download()
    (so, co, _), _, _ = bindings.OracleScene(sc).render(_cam(view), settings(max_bounces=view.max_bounces, seed=12345),
                                                        view.width, view.height, 0, spp, threads=8)
    assert relmse(s / c[..., None], so / co[..., None]) <= 1e-3
