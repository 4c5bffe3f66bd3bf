"""PathTracer::Li driven by tracePixel / Renderer::renderSamples (integrators/path_tracer.cpp:321-890,
engine/renderer.cpp:139-531), on the box + icosphere and the environment-lit product sphere.

CPU: the oracle's restated integrator equals the reference's own strict build sample for sample (same Sobol
dimensions, same arithmetic); RenderDevice semantics (batch == single passes, accumulate into the film,
NaN guard) hold.
GPU: the wavefront kernels against the oracle on the same (pixel, sample) paths, per sample and as images
(relMSE = mean((x - ref)^2 / (ref^2 + 0.01)), scripts/compare_exr.py:30), and the device semantics.
Per-sample tolerance: the kernels contract a*b+c to FMA and use the device's sin/cos/acos/pow, so a path may
take another branch where a decision sits within an ulp of its threshold (a roulette or lobe draw); those
paths differ, the rest agree to float rounding. The image bar is the one the task states: relMSE <= 1e-3.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, settings


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def _scenes():
    box, vb = configs.box_icosphere(level=3, width=32, height=32)
    sph, vs = configs.product_sphere(width=48, height=27, lon=64, lat=48, env_width=256, env_height=128)
    return [("box_icosphere", box, vb), ("product_sphere", sph, vs)]


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere"])
def test_oracle_radiance_equals_reference_per_sample(name):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = {n: (s, v) for n, s, v in _scenes()}[name]
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    rng = np.random.default_rng(12)
    n = 3000
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 256, n)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = ref.trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)


def test_oracle_blackman_harris_filter_equals_reference():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = configs.box_icosphere(level=2, width=24, height=24)
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    rng = np.random.default_rng(3)
    px, py, s = rng.integers(0, 24, 800), rng.integers(0, 24, 800), rng.integers(0, 64, 800)
    st = settings(max_bounces=8, filter_type=capi.PE_FILTER_BLACKMAN_HARRIS, filter_width=1.5, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, 24, 24, px, py, s)
    Lr = ref.trace_samples(view, px, py, s, filter_type=3, filter_width=1.5, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)


def test_oracle_film_equals_reference_render_samples():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = configs.box_icosphere(level=3, width=16, height=16)
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    st = settings(max_bounces=8, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    (sum_, cnt, _), _, _ = o.render(_cam(view), st, 16, 16, 0, 6, threads=2)
    mean_r, cnt_r, _ = ref.render(view, 0, 6, threads=2, filter_type=0, filter_width=1.0)
    np.testing.assert_array_equal(cnt, cnt_r)
    inv = (1.0 / cnt.astype(np.float32)).astype(np.float32)
    np.testing.assert_array_equal(sum_ * inv[..., None], mean_r)


def test_oracle_batch_equals_single_passes():
    # test_render_session.cpp:96-115: one batch of 8 == eight passes of 1, bit for bit.
    sc, view = configs.box_icosphere(level=2, width=12, height=12)
    o = bindings.OracleScene(sc)
    st = settings(max_bounces=6, seed=12345)
    cam = _cam(view)
    (a, ca, la), _, _ = o.render(cam, st, 12, 12, 0, 8)
    film = (np.zeros((12, 12, 3), np.float32), np.zeros((12, 12), np.uint32), np.zeros((12, 12), np.float32))
    for k in range(8):
        o.render(cam, st, 12, 12, k, 1, film=film)
    np.testing.assert_array_equal(a, film[0])
    np.testing.assert_array_equal(ca, film[1])
    np.testing.assert_array_equal(la, film[2])


# ------------------------------------------------------------------------------------------------- GPU

def _gpu_image(dev, sc, view, spp, first=0, spl=1, max_bounces=None):
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=max_bounces or view.max_bounces, seed=12345, samples_per_launch=spl)
    res = render(dev, gs, _cam(view), st, film, first, spp)
    s, c, l2, rej = film.download()
    return s, c, l2, rej, res


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere"])
def test_gpu_radiance_matches_oracle_per_sample(dev, name):
    sc, view = {n: (s, v) for n, s, v in _scenes()}[name]
    sample = 5
    s, c, _, rej, res = _gpu_image(dev, sc, view, 1, first=sample)
    assert res.completed == 1 and res.samples_traced == view.width * view.height
    assert (c == 1).all() and rej == 0
    o = bindings.OracleScene(sc)
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, xx.ravel(), yy.ravel(),
                            np.full(xx.size, sample))
    Lg = s.reshape(-1, 3)
    close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
    print(f"{name}: per-sample agreement with the oracle {close.mean():.5f} of {close.size} samples", flush=True)
    assert close.mean() >= 0.999, close.mean()
    assert np.abs(Lg.mean(0) - Lo.mean(0)).max() <= 0.02 * max(1e-3, Lo.mean())


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere"])
def test_gpu_image_matches_oracle_relmse(dev, name):
    sc, view = {n: (s, v) for n, s, v in _scenes()}[name]
    spp = 64
    s, c, _, rej, res = _gpu_image(dev, sc, view, spp, spl=4)
    assert (c == spp).all()
    gpu = s / c[..., None]
    o = bindings.OracleScene(sc)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    (so, co, _), _, _ = o.render(_cam(view), st, view.width, view.height, 0, spp, threads=8)
    ora = so / co[..., None]
    assert relmse(gpu, ora) <= 1e-3, relmse(gpu, ora)
    assert res.extension_rays >= view.width * view.height * spp


@pytest.mark.gpu
def test_gpu_batch_equals_single_passes(dev):
    sc, view = configs.box_icosphere(level=3, width=40, height=24)
    gs = dev.upload(sc)
    cam = _cam(view)
    st = settings(max_bounces=8, seed=12345, samples_per_launch=8)
    f1 = dev.film(40, 24)
    render(dev, gs, cam, st, f1, 0, 8)
    f2 = dev.film(40, 24)
    st1 = settings(max_bounces=8, seed=12345, samples_per_launch=1)
    for k in range(8):
        render(dev, gs, cam, st1, f2, k, 1)
    a, b = f1.download(), f2.download()
    for x, y in zip(a[:3], b[:3]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.gpu
def test_gpu_cancel_and_errors(dev):
    sc, view = configs.box_icosphere(level=2, width=16, height=16)
    gs = dev.upload(sc)
    film = dev.film(16, 16)
    cancel = np.ones(1, np.int32)
    res = render(dev, gs, _cam(view), settings(), film, 0, 4, cancel=cancel)
    assert res.completed == 0 and res.samples_traced == 0
    lib = capi.load()
    bad = capi.BatchResult()
    cam = _cam(view)
    st = settings()
    assert lib.pe_render(dev.h, gs.h, C.byref(cam), C.byref(st), film.h, -1, 1, None, C.byref(bad)) == capi.PE_ERR_INVALID
    cam.lens_radius = -0.1
    assert lib.pe_render(dev.h, gs.h, C.byref(cam), C.byref(st), film.h, 0, 1, None, C.byref(bad)) == capi.PE_ERR_INVALID


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere"])
def test_gpu_converged_image_matches_reference_renderer(dev, name):
    """The workload's criterion: a converged 4096-spp GPU image against the reference's own CPU renderer
    (Renderer::renderSamples with its -O3 -ffast-math build, oracle/_ref) on the same frame, relMSE <= 1e-3."""
    ref_lib = bindings.reference(strict=False)
    if ref_lib is None:
        pytest.skip("oracle/_ref not built (reference tree absent when this tree was built)")
    if name == "box_icosphere":
        sc, view = configs.box_icosphere(level=3, width=20, height=20)
    else:
        sc, view = configs.product_sphere(width=48, height=27, lon=64, lat=48, env_width=256, env_height=128)
    spp = 4096
    s, c, _, rej, _ = _gpu_image(dev, sc, view, spp, spl=512)
    assert (c == spp).all() and rej == 0
    gpu = s / c[..., None]
    mean_r, cnt_r, _ = bindings.RefScene(sc, strict=False).render(view, 0, spp, threads=16)
    assert (cnt_r == spp).all()
    e = relmse(gpu, mean_r)
    assert e <= 1e-3, e


@pytest.mark.gpu
def test_gpu_full_size_box_icosphere_matches_reference_renderer(dev):
    """The small plain workload at its full size: 512 x 512, 64 spp, depth 8, the level-5 icosphere (20,480
    triangles) in the area-lit box, against the reference's own renderer on the same samples (seed 12345,
    Renderer::renderSamples with the reference's -O3 -ffast-math build): relMSE <= 1e-3."""
    if bindings.reference(strict=False) is None:
        pytest.skip("oracle/_ref not built (reference tree absent when this tree was built)")
    sc, view = configs.box_icosphere(level=5)
    assert (view.width, view.height, view.spp, view.max_bounces) == (512, 512, 64, 8)
    s, c, _, rej, res = _gpu_image(dev, sc, view, view.spp, spl=64)
    assert (c == view.spp).all() and rej == 0
    gpu = s / c[..., None]
    mean_r, cnt_r, _ = bindings.RefScene(sc, strict=False).render(view, 0, view.spp, threads=16)
    assert (cnt_r == view.spp).all()
    e = relmse(gpu, mean_r)
    print(f"512x512x64 box + icosphere: relMSE vs the reference renderer {e:.3e}", flush=True)
    assert e <= 1e-3, e


def _env_only():
    """An environment and no geometry: BVH::build of no primitives, every camera ray sees the environment."""
    from photon_hip.scene import Scene
    sph, view = configs.product_sphere(width=9, height=5, lon=32, lat=16, env_width=128, env_height=64)
    sc = Scene()
    sc.environment = sph.environment
    return sc, view


def test_oracle_empty_scene_equals_reference():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = _env_only()
    py, px = np.divmod(np.arange(view.width * view.height), view.width)
    for s in (0, 3):
        st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
        Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), st, view.width, view.height, px, py,
                                                       np.full(px.size, s))
        Lr = bindings.RefScene(sc, strict=True).trace_samples(view, px, py, np.full(px.size, s))
        np.testing.assert_array_equal(Lo, Lr)
        assert Lo.min() > 0


@pytest.mark.gpu
def test_gpu_edge_cases(dev):
    """An environment with no geometry, a 1x1 film, a zero-sample call and depth 0 (emission only), each against
    the oracle."""
    sc, view = _env_only()
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    res = render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345), film, 2, 1)
    assert res.completed == 1 and res.shadow_rays == 0
    s, c, _, _ = film.download()
    py, px = np.divmod(np.arange(view.width * view.height), view.width)
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), settings(max_bounces=view.max_bounces, seed=12345),
                                                   view.width, view.height, px, py, np.full(px.size, 2))
    np.testing.assert_allclose(s.reshape(-1, 3), Lo, rtol=1e-4)
    assert (c == 1).all()

    box, bview = configs.box_icosphere(level=2, width=1, height=1)
    gb = dev.upload(box)
    one = dev.film(1, 1)
    res = render(dev, gb, _cam(bview), settings(max_bounces=bview.max_bounces, seed=12345, samples_per_launch=16),
                 one, 0, 16)
    assert res.completed == 1 and res.samples_traced == 16
    (so, co, _), _, _ = bindings.OracleScene(box).render(_cam(bview), settings(max_bounces=bview.max_bounces), 1, 1, 0,
                                                          16)
    s1, c1, _, _ = one.download()
    assert (c1 == 16).all()
    np.testing.assert_allclose(s1, so, rtol=1e-3, atol=1e-5)
    res = render(dev, gb, _cam(bview), settings(), one, 16, 0)    # no samples: a no-op
    assert res.samples_traced == 0
    assert (one.download()[1] == 16).all()

    small, sview = configs.box_icosphere(level=2, width=12, height=10)
    gsm = dev.upload(small)
    f0 = dev.film(12, 10)
    render(dev, gsm, _cam(sview), settings(max_bounces=0, seed=12345), f0, 0, 1)
    py, px = np.divmod(np.arange(120), 12)
    L0, _ = bindings.OracleScene(small).trace_samples(_cam(sview), settings(max_bounces=0, seed=12345), 12, 10, px, py,
                                                      np.zeros(120, int))
    np.testing.assert_allclose(f0.download()[0].reshape(-1, 3), L0, rtol=1e-4, atol=1e-6)
