"""Disney metallic flakes: DisneyMaterial::setFlakes(density, size, roughness, color) (materials/disney.h:120-125).
Space is cut into cells of the flake size with one hashed seed per cell (3-D Voronoi); a point whose nearest seed
hashes below the density lies on a flake whose normal is tilted about the shading normal by a GGX-like slope
distribution of the flake roughness (resolveFlake, disney.cpp:407-460). On a flake the base layer is replaced by
a narrow GGX lobe (alpha 0.06) about the flake normal with Schlick reflectance of the flake colour, under the
clear coat; there is no diffuse, sheen or transmission (evalFlake / pdfFlake :104-126, evalLobes :223-311,
lobeWeights :313-338, sampleSpecularLobe :380-396).

CPU: the oracle's flake state, BSDF value, pdf and samples at shading points equal the reference's (IEEE build)
bit for bit; closed forms that do not need the reference (the flaked fraction of random points is the density,
flake tilts follow P(tan^2 theta <= x) = x / (r^2 + x) up to the 0.999 cap, roughness 0 keeps the shading
normal, a flake lobe of white colour conserves energy below 1); per-sample radiance of a flaked car-paint scene
equals the reference's tracePixel.
GPU: pe_render per sample against the oracle, and RenderSession through HipDevice (which packs the flakes)
against a direct render.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip.configs import View, grid_quad, icosphere
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import AreaLight, Environment, Scene, disney

MATS = {
    "paint": dict(base=(0.5, 0.05, 0.05), metallic=0.0, roughness=0.4, kw=dict(
        clearcoat=1.0, clearcoat_roughness=0.05, flake_density=0.5, flake_size=0.01, flake_roughness=0.35,
        flake_color=(0.95, 0.85, 0.6))),
    "dense_rough": dict(base=(0.2, 0.3, 0.6), metallic=0.7, roughness=0.3, kw=dict(
        flake_density=1.0, flake_size=0.003, flake_roughness=1.0, flake_color=(0.9, 0.9, 0.9))),
    "flat_flakes": dict(base=(0.8, 0.8, 0.8), metallic=0.0, roughness=0.6, kw=dict(
        clearcoat=0.3, flake_density=0.25, flake_size=0.02, flake_roughness=0.0, flake_color=(1.0, 0.4, 0.1))),
}


def _mat(name):
    d = MATS[name]
    return disney(d["base"], d["metallic"], d["roughness"], **d["kw"])


def _frames(n, seed):
    rng = np.random.default_rng(seed)
    pts = rng.uniform(-2.0, 2.0, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    tan = np.cross(nrm, [0.3, 0.9, 0.1])
    tan /= np.linalg.norm(tan, axis=1, keepdims=True)
    wo = nrm + rng.normal(size=(n, 3)) * 0.6
    wo /= np.linalg.norm(wo, axis=1, keepdims=True)
    wi = rng.normal(size=(n, 3))
    wi /= np.linalg.norm(wi, axis=1, keepdims=True)
    wi[: n // 2] = np.where((wi[: n // 2] * nrm[: n // 2]).sum(1, keepdims=True) < 0, -wi[: n // 2], wi[: n // 2])
    front = (rng.random(n) < 0.9).astype(np.int32)
    u = rng.random((n, 3)).astype(np.float32)
    return pts, nrm.astype(np.float32), tan.astype(np.float32), front, wo.astype(np.float32), wi.astype(np.float32), u


def _ref_or_skip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


@pytest.mark.parametrize("name", sorted(MATS))
def test_oracle_flakes_equal_reference(name):
    _ref_or_skip()
    m = _mat(name)
    pts, nrm, tan, front, wo, wi, u = _frames(6000, 5)
    a = bindings.flake_resolve("oracle", m, pts, nrm, tan, front)
    b = bindings.flake_resolve("ref_strict", m, pts, nrm, tan, front)
    np.testing.assert_array_equal(a, b)
    fo, po = bindings.bsdf_eval("oracle", m, nrm, tan, front, wo, wi, pts)
    fr, pr = bindings.bsdf_eval("ref_strict", m, nrm, tan, front, wo, wi, pts)
    np.testing.assert_array_equal(fo, fr)
    np.testing.assert_array_equal(po, pr)
    so = bindings.bsdf_sample("oracle", m, nrm, tan, front, wo, u, pts)
    sr = bindings.bsdf_sample("ref_strict", m, nrm, tan, front, wo, u, pts)
    np.testing.assert_array_equal(so, sr)
    # The flaked points do see the flake lobe: their values differ from the same material without flakes.
    plain = _mat(name)
    plain.flake_density = 0.0
    fp, _ = bindings.bsdf_eval("oracle", plain, nrm, tan, front, wo, wi, pts)
    on = a[:, 0] == 1.0
    assert on.any() and np.any(fo[on] != fp[on]) and np.array_equal(fo[~on], fp[~on])


def test_flake_closed_forms():
    pts, nrm, tan, front, wo, wi, _ = _frames(40000, 8)
    for density in (0.2, 0.7):
        m = disney((0.5, 0.5, 0.5), flake_density=density, flake_size=0.01, flake_roughness=0.5)
        st = bindings.flake_resolve("oracle", m, pts, nrm, tan, front)
        on = st[:, 0] == 1.0
        assert abs(on.mean() - density) < 0.015, (density, on.mean())
        # tilt distribution: tan^2 theta = r^2 u / (1 - u), u uniform below 0.999
        cos_t = np.clip((st[on, 1:] * nrm[on]).sum(1), -1.0, 1.0)
        tan2 = (1.0 - cos_t ** 2) / np.maximum(cos_t ** 2, 1e-12)
        r2 = 0.25
        for x in (0.05, 0.25, 1.0):
            assert abs((tan2 <= x).mean() - x / (r2 + x)) < 0.02
        assert np.abs(np.linalg.norm(st[on, 1:], axis=1) - 1.0).max() < 1e-6
    flat = disney((0.5, 0.5, 0.5), flake_density=1.0, flake_size=0.01, flake_roughness=0.0)
    st = bindings.flake_resolve("oracle", flat, pts, nrm, tan, front)
    assert (st[:, 0] == 1.0).all()
    np.testing.assert_allclose(st[:, 1:], nrm, atol=2e-7)
    # A white flake lobe alone (no coat) is a GGX mirror about the flake normal: measured against the flake's own
    # cosine, every sample's weight f cos_flake / pdf is G / G1(wo) <= 1; seen head-on along the flake normal
    # (no masking of wo) it is G1(wi), 1 for all but the rare grazing reflections of the narrow lobe.
    white = disney((0.5, 0.5, 0.5), 1.0, flake_density=1.0, flake_size=0.01, flake_roughness=0.3,
                   flake_color=(1.0, 1.0, 1.0))
    pts, nrm, tan, front, wo, _, u = _frames(20000, 9)
    front[:] = 1
    st = bindings.flake_resolve("oracle", white, pts, nrm, tan, front)
    s = bindings.bsdf_sample("oracle", white, nrm, tan, front, wo, u, pts)
    ok = s[:, 0] == 1.0
    cos_f = (s[ok, 1:4] * st[ok, 1:]).sum(1)
    w = s[ok, 4:7] * cos_f[:, None] / s[ok, 7:8]
    assert ok.mean() > 0.5 and w.max() <= 1.0 + 1e-4 and np.median(w) > 0.9
    s = bindings.bsdf_sample("oracle", white, nrm, tan, front, st[:, 1:].copy(), u, pts)
    ok = s[:, 0] == 1.0
    w = s[ok, 4:7] * (s[ok, 1:4] * st[ok, 1:]).sum(1)[:, None] / s[ok, 7:8]
    assert w.max() <= 1.0 + 1e-4 and (w > 0.99).mean() > 0.99


def _scene():
    """Three car-paint spheres (clear-coated flaked red, dense rough flakes, flat orange flakes) on a floor,
    under an area light and a dim sky; the flakes are a few millimetres across on 0.45 m spheres."""
    sc = Scene()
    sc.add_shape(grid_quad((-3, 0, -3), (0, 0, 6), (6, 0, 0), 2, 2), sc.add_material(disney((0.5, 0.5, 0.5), 0.0, 0.5)))
    for k, name in enumerate(sorted(MATS)):
        sc.add_shape(icosphere(4, 0.45, (-1.0 + 1.0 * k, 0.45, 0.0)), sc.add_material(_mat(name)))
    sc.add_light(AreaLight((-0.4, 2.5, 0.8), (0.8, 0.0, 0.0), (0.0, 0.0, 0.6), (12.0, 12.0, 12.0)))
    sc.set_environment(Environment(np.full((16, 32, 3), 0.3, np.float32)))
    return sc, View((0.0, 1.3, 3.0), (0.0, 0.4, 0.0), (0, 1, 0), 40.0, 40, 24, 16, 8)


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def test_oracle_flaked_scene_equals_reference_per_sample():
    _ref_or_skip()
    sc, view = _scene()
    rng = np.random.default_rng(3)
    n = 3000
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 512, n)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = bindings.RefScene(sc, strict=True).trace_samples(view, px, py, s, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_flakes_match_oracle_per_sample(dev):
    sc, view = _scene()
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    close_n = total = 0
    for sample in (0, 7, 21):
        film = dev.film(view.width, view.height)
        res = render(dev, gs, _cam(view), st, film, sample, 1)
        assert res.completed == 1
        sums, _, _, _ = film.download()
        L, _ = o.trace_samples(_cam(view), st, view.width, view.height, xx, yy, np.full(xx.size, sample))
        Lg = sums.reshape(-1, 3)
        close = np.all(np.abs(Lg - L) <= 1e-4 * np.maximum(1.0, np.abs(L)), axis=1)
        close_n += int(close.sum())
        total += close.size
        film.close()
    print(f"flakes: per-sample agreement {close_n / total:.5f} of {total}", flush=True)
    assert close_n >= 0.999 * total


@pytest.mark.gpu
def test_gpu_flakes_through_hip_device(dev):
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    sc, view = _scene()
    spp = 8
    rc, mean, counts, _, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=4)
    assert rc == 0, err
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4), film, 0,
           spp)
    s, c, _, _ = film.download()
    assert (counts == spp).all()
    np.testing.assert_allclose(mean, s / c[..., None], rtol=2e-5, atol=1e-6)
