"""Light selection on the light BVH (LightSampler::build / Context::sample / pmf, lights/light_sampler.cpp:42-293;
LightBounds, lights/light_bounds.cpp:39-109) for scenes with more than 33 light candidates: the LED-strip scene
of tests/test_light_sampler.cpp:104-127 (2000 emissive triangles over a floor).

CPU: the oracle's light BVH is pinned by what the mathematics fixes (the pmf of every candidate sums to one at any
shading point; a sampled index reports the pmf that pmf(index) gives; sample frequencies follow the pmf) and by
the reference's own strict build, sample for sample (same topology, same pmf arithmetic).
GPU: the kernels against the oracle on the same paths and as images (tolerances as in test_integrator.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, settings


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _points(n, seed=3):
    rng = np.random.default_rng(seed)
    p = np.stack([rng.uniform(-6, 6, n), rng.uniform(0.0, 1.2, n), rng.uniform(-6, 6, n)], 1).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm[: n // 2] = [0.0, 1.0, 0.0]
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return p, nrm.astype(np.float32), rng.random(n).astype(np.float32)


def test_led_strip_uses_the_light_bvh_and_pmf_sums_to_one():
    sc, _ = configs.led_strip()
    o = bindings.OracleScene(sc)
    p, n, u = _points(256)
    pmf, idx, spmf, bvh = o.light_select(p, n, u)
    assert bvh
    assert pmf.shape[1] == 2000
    np.testing.assert_allclose(pmf.sum(1), 1.0, rtol=2e-4)
    assert (pmf >= 0).all()
    # A sampled candidate reports the pmf that pmf(index) computes along its bit trail.
    ok = idx >= 0
    assert ok.mean() > 0.99
    np.testing.assert_allclose(spmf[ok], pmf[np.arange(len(idx))[ok], idx[ok]], rtol=1e-5, atol=1e-12)


def test_light_bvh_sample_frequencies_follow_pmf():
    sc, _ = configs.led_strip(quads=40)   # 80 emitters: still the BVH, small enough to count
    o = bindings.OracleScene(sc)
    N = 200000
    p = np.tile(np.float32([[0.3, 0.1, -0.2]]), (N, 1))
    n = np.tile(np.float32([[0.0, 1.0, 0.0]]), (N, 1))
    u = ((np.arange(N) + 0.5) / N).astype(np.float32)          # stratified u: frequencies converge fast
    pmf, idx, _, bvh = o.light_select(p[:1], n[:1], u[:1])
    assert bvh
    _, idx, _, _ = o.light_select(p, n, u, with_pmf=False)
    freq = np.bincount(idx[idx >= 0], minlength=pmf.shape[1]) / N
    np.testing.assert_allclose(freq, pmf[0], atol=2e-4)


def test_oracle_radiance_equals_reference_per_sample_led_strip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = configs.led_strip()
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    rng = np.random.default_rng(5)
    k = 2000
    px, py, s = rng.integers(0, view.width, k), rng.integers(0, view.height, k), rng.integers(0, 64, k)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = ref.trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    assert Lo.max() > 0


@pytest.mark.gpu
def test_gpu_led_strip_matches_oracle(dev):
    sc, view = configs.led_strip(width=40, height=40, max_bounces=3)
    o = bindings.OracleScene(sc)
    gs = dev.upload(sc)
    cam = _cam(view)
    # One sample per pixel: per-sample agreement.
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=1)
    res = render(dev, gs, cam, st, film, 3, 1)
    assert res.completed == 1
    s, c, _, rej = film.download()
    assert (c == 1).all() and rej == 0
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    Lo, _ = o.trace_samples(cam, st, view.width, view.height, xx.ravel(), yy.ravel(), np.full(xx.size, 3))
    Lg = s.reshape(-1, 3)
    close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
    assert close.mean() >= 0.97, close.mean()
    # Images at 16 spp.
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=8)
    render(dev, gs, cam, st, film, 0, 16)
    s, c, _, _ = film.download()
    (so, co, _), _, _ = o.render(cam, st, view.width, view.height, 0, 16, threads=4)
    assert relmse(s / c[..., None], so / co[..., None]) <= 1e-3
