"""Scene backgrounds: Scene::setBackground (engine/scene.h:24-28) with PathTracer::Li's miss branch
(integrators/path_tracer.cpp:483-495) and tracePixel's AOVs (engine/renderer.cpp:158-205). A camera path that
leaves the scene before its first vertex (straight, or past thin walls and camera-invisible lights) sees a Color
background's colour, credited to the Background pass and the environment's light group, or, for a Transparent
background, nothing: a camera ray that passed thin walls gives the alpha AOV their coverage,
1 - luminance(throughput). Every later miss sees the environment as before.

CPU: the oracle's per-sample radiance, AOVs and pass records equal the reference's (IEEE build) on an
environment-lit sphere with a Color background, the glass studio with a Color background and a thin-pane scene
with a Transparent one.
GPU: pe_render per sample against the oracle (radiance and AOVs, the coverage included), and RenderSession
through HipDevice (which packs Scene::background()) against a direct render.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.configs import View, box_mesh, grid_quad
from photon_hip.device import camera_perspective, render, render_targets, settings
from photon_hip.scene import AreaLight, Environment, Scene, dielectric, disney, lambert


def _panes():
    """Two tinted thin-walled panes between the camera and the sky, a block behind one of them, a floor below
    and a camera-invisible area light in the view: camera rays pass zero, one or two walls (and the light)
    before they leave."""
    sc = Scene()
    sc.add_shape(grid_quad((-3, 0, -3), (0, 0, 6), (6, 0, 0), 2, 2), sc.add_material(lambert((0.6, 0.6, 0.6))))
    thin = sc.add_material(dielectric(1.5, tint=(0.7, 0.85, 1.0), thin_walled=True))
    sc.add_shape(grid_quad((-1.2, 0.2, 0.8), (0.0, 1.6, 0.0), (1.4, 0.0, 0.0), 1, 1), thin)
    sc.add_shape(grid_quad((-0.6, 0.5, 0.3), (0.0, 1.4, 0.0), (1.6, 0.0, -0.2), 1, 1),
                 sc.add_material(dielectric(1.5, tint=(1.0, 0.8, 0.6), thin_walled=True)))
    sc.add_shape(box_mesh((0.6, 0.0, -0.6), (1.1, 0.7, -0.2)), sc.add_material(disney((0.8, 0.3, 0.2), roughness=0.4)))
    sc.add_light(AreaLight((-0.4, 1.3, 0.0), (0.5, 0.0, 0.0), (0.0, 0.3, 0.2), (6.0, 6.0, 6.0),
                           visible_to_camera=False))
    sky = np.ones((32, 64, 3), np.float32) * np.linspace(1.5, 0.3, 32, dtype=np.float32)[:, None, None]
    sc.set_environment(Environment(sky, 0.4, 1.0))
    sc.background_mode = capi.PE_BACKGROUND_TRANSPARENT
    return sc, View((0.1, 1.0, 3.2), (0.0, 0.9, 0.0), (0, 1, 0), 50.0, 32, 24, 16, 8)


def _scenes():
    sph, vs = configs.product_sphere(width=24, height=16, lon=64, lat=48, env_width=256, env_height=128)
    sph.background_mode = capi.PE_BACKGROUND_COLOR
    sph.background_color = (0.9, 0.45, 0.2)
    glass, vg = configs.glass_studio(width=24, height=16)
    glass.background_mode = capi.PE_BACKGROUND_COLOR
    glass.background_color = (0.2, 0.6, 0.3)
    panes, vp = _panes()
    return {"sphere_color": (sph, vs), "glass_color": (glass, vg), "panes_transparent": (panes, vp)}


NAMES = ["sphere_color", "glass_color", "panes_transparent"]


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _pixels(view):
    py, px = np.divmod(np.arange(view.width * view.height), view.width)
    return px, py


def _ref_or_skip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


@pytest.mark.parametrize("name", NAMES)
def test_oracle_radiance_equals_reference_per_sample(name):
    _ref_or_skip()
    sc, view = _scenes()[name]
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    rng = np.random.default_rng(3)
    n = 2000
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 256, n)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = ref.trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    if name == "sphere_color":      # the colour is seen: samples of exactly it
        assert (np.abs(Lo - np.float32(sc.background_color)).max(1) == 0).any()


@pytest.mark.parametrize("name", NAMES)
def test_oracle_aovs_equal_reference(name):
    """One sample per pixel through Renderer::renderSamples with AovFilms: albedo (the Color background's colour
    on a miss), normal and coverage (the thin walls' on the Transparent background)."""
    _ref_or_skip()
    sc, view = _scenes()[name]
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    px, py = _pixels(view)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    for s in (0, 7):
        L, alb, nrm, alpha = o.trace_aovs(_cam(view), st, view.width, view.height, px, py, np.full(px.size, s))
        b, a, n, c = (x.reshape(-1, 3) for x in ref.render_aovs(view, s, 1))
        np.testing.assert_array_equal(L, b)
        np.testing.assert_array_equal(alb, a)
        np.testing.assert_array_equal(nrm, n)
        np.testing.assert_array_equal(np.repeat(alpha[:, None], 3, 1), c)
    if name == "panes_transparent":       # coverage strictly between 0 and 1 behind the panes
        assert ((alpha > 0.0) & (alpha < 1.0)).any()


def test_oracle_background_pass_equals_reference():
    """The Color background is credited to the Background pass (and to the environment's light group)."""
    _ref_or_skip()
    sc, view = _scenes()["sphere_color"]
    sc.environment.light_group = "Sky"
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    rng = np.random.default_rng(5)
    n = 1500
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 64, n)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, po = o.trace_passes(_cam(view), st, view.width, view.height, px, py, s)
    Lr, pr = ref.trace_passes(view, px, py, s)
    np.testing.assert_array_equal(Lo, Lr)
    for k in po:
        np.testing.assert_array_equal(po[k], pr[k], err_msg=k)
    bg = po["light"][:, capi.PE_PASS_BACKGROUND]
    assert (np.abs(bg - np.float32(sc.background_color)).max(1) == 0).any()


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("name", NAMES)
def test_gpu_radiance_and_aovs_match_oracle_per_sample(dev, name):
    sc, view = _scenes()[name]
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    px, py = _pixels(view)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    close_n = total = agree = 0
    for sample in (2, 9):
        films = [dev.film(view.width, view.height) for _ in range(4)]
        res = render_targets(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345,
                                                           samples_per_launch=1),
                             films[0], sample, 1, albedo=films[1], normal=films[2], alpha=films[3])
        assert res.completed == 1
        beauty, alb, _, cov = (f.download() for f in films)
        L, alb_o, _, alpha_o = o.trace_aovs(_cam(view), st, view.width, view.height, px, py,
                                            np.full(px.size, sample))
        Lg = beauty[0].reshape(-1, 3)
        close = np.all(np.abs(Lg - L) <= 1e-4 * np.maximum(1.0, np.abs(L)), axis=1)
        close_n += int(close.sum())
        total += close.size
        cov_g = cov[0].reshape(-1, 3)[:, 0]
        same = np.abs(cov_g - alpha_o) <= 1e-6
        agree += int(same.sum())
        miss = same & (alpha_o == 0.0)
        np.testing.assert_allclose(alb[0].reshape(-1, 3)[miss], alb_o[miss], rtol=1e-4, atol=1e-6)
        if name == "panes_transparent":
            assert ((cov_g > 0.0) & (cov_g < 1.0)).any()
    print(f"{name}: per-sample agreement {close_n / total:.5f}, coverage agreement {agree / total:.5f} "
          f"of {total}", flush=True)
    assert close_n >= 0.998 * total
    assert agree >= 0.995 * total


@pytest.mark.gpu
def test_gpu_unknown_background_mode_is_refused(dev):
    sc, _ = _panes()
    sc.background_mode = 3
    with pytest.raises(Exception, match="background"):
        dev.upload(sc)


@pytest.mark.gpu
def test_gpu_background_through_hip_device(dev):
    """HipDevice packs Scene::background(): RenderSession's batches give the direct render's image."""
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    for name in ("sphere_color", "panes_transparent"):
        sc, view = _scenes()[name]
        spp = 8
        rc, mean, counts, _, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=4)
        assert rc == 0, err
        gs = dev.upload(sc)
        film = dev.film(view.width, view.height)
        render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4),
               film, 0, spp)
        s, c, _, _ = film.download()
        np.testing.assert_allclose(mean, s / c[..., None], rtol=2e-5, atol=1e-6)
