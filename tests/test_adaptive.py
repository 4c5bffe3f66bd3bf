"""Adaptive sampling (RenderSettings::adaptiveSampling, engine/renderer.cpp:438-505; Film::relativeError,
core/image/film.cpp:100-109).

A batch skips every pixel whose 3x3 neighbourhood is already below the noise threshold, and within a batch a
pixel stops after sample s when s + 1 is a multiple of 8, at least minAdaptiveSamples and not the batch's end,
and its own relative error is below the threshold.

CPU: the oracle's numpy restatement (oracle/adaptive.py) against closed forms of the relative error and, driven
by the reference's own per-sample radiance, against the reference's own Renderer::renderSamples (IEEE build)
over two batches into one film: sample counts, resolved means, relative errors and samples traced equal.
GPU: pe_render with adaptive sampling against the oracle driven by the GPU's own per-sample radiance (each
sample rendered alone into a fresh film): the film (sums, counts, luminance-square sums) equal bit for bit, and
the samples traced.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import adaptive, bindings
from photon_hip import configs
from photon_hip.device import camera_perspective, render, settings

F = np.float32


def _scene():
    sc, view = configs.box_icosphere(level=3, width=24, height=24)
    return sc, view


def test_relative_error_closed_forms():
    film = adaptive.Film(4, 1)
    # pixel 0: one sample (below two: 1e30); pixel 1: two equal samples (zero variance); pixels 2, 3: grey
    # samples 1 and 3 (luminance 1 and 3: mean 2, unbiased variance 2, error sqrt(2 / 2) / 2.01).
    film.add_samples(np.array([0, 1, 2, 3]), np.array([[1, 1, 1], [0.5, 0.5, 0.5], [1, 1, 1], [3, 3, 3]], F))
    film.add_samples(np.array([1, 2, 3]), np.array([[0.5, 0.5, 0.5], [3, 3, 3], [1, 1, 1]], F))
    err = film.relative_error()
    assert err[0] == F(1e30)
    assert err[1] == 0.0
    np.testing.assert_allclose(err[2:], np.sqrt(2.0 / 2.0) / 2.01, rtol=2e-6)
    # A rejected sample (NaN, negative) counts but adds nothing (film.cpp:56-71).
    film.add_samples(np.array([0, 1]), np.array([[np.nan, 0, 0], [-1, 1, 1]], F))
    assert film.count.tolist() == [2, 3, 2, 2] and film.rejected == 2
    assert film.sum[1].tolist() == [1.0, 1.0, 1.0]


@pytest.mark.parametrize("threshold", [0.1, 0.2])
def test_oracle_adaptive_equals_reference_renderer(threshold):
    lib = bindings.reference(strict=True)
    if lib is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = _scene()
    ref = bindings.RefScene(sc, strict=True)
    W, H = view.width, view.height
    py, px = np.divmod(np.arange(W * H), W)
    cache = {}

    def sample(s):
        if s not in cache:
            cache[s] = ref.trace_samples(view, px, py, np.full(W * H, s), filter_type=0, filter_width=1.0, seed=12345)
        return cache[s]

    film = adaptive.Film(W, H)
    batches = [(0, 16), (16, 32)]
    traced = sum(adaptive.render_batch(film, sample, f, c, threshold, 8) for f, c in batches)
    mean_r, cnt_r, rel_r, traced_r = bindings.ref_render_adaptive(ref.h, lib, view, batches, threshold, 8)
    np.testing.assert_array_equal(film.count.reshape(H, W), cnt_r)
    np.testing.assert_array_equal(film.resolved(), mean_r)
    np.testing.assert_array_equal(film.relative_error().reshape(H, W), rel_r)
    assert traced == traced_r
    # The case exercises both rules: pixels stopped early inside a batch and pixels skipped by the mask.
    assert film.count.min() < 32 and film.count.max() == 48


# ------------------------------------------------------------------------------------------------- GPU

def _gpu_samples(dev, gs, view, cam, n, max_bounces):
    """The GPU's radiance of samples 0..n-1 of every pixel, each rendered alone into a fresh film."""
    out = []
    for s in range(n):
        film = dev.film(view.width, view.height)
        render(dev, gs, cam, settings(max_bounces=max_bounces, seed=12345), film, s, 1)
        sm, c, _, _ = film.download()
        assert (c == 1).all()
        out.append(sm.reshape(-1, 3).astype(F))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("threshold,min_spp,batches", [(0.2, 8, [(0, 16), (16, 32)]),
                                                       (0.1, 2, [(0, 5), (5, 19), (24, 12)])])
def test_gpu_adaptive_matches_oracle(dev, threshold, min_spp, batches):
    sc, view = _scene()
    gs = dev.upload(sc)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    n = max(f + c for f, c in batches)
    Ls = _gpu_samples(dev, gs, view, cam, n, view.max_bounces)
    film_o = adaptive.Film(view.width, view.height)
    traced_o = [adaptive.render_batch(film_o, lambda s: Ls[s], f, c, threshold, min_spp) for f, c in batches]
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=8, adaptive=True,
                  noise_threshold=threshold, min_adaptive_samples=min_spp)
    traced = [render(dev, gs, cam, st, film, f, c).samples_traced for f, c in batches]
    sm, cnt, l2, _ = film.download()
    np.testing.assert_array_equal(cnt.reshape(-1), film_o.count)
    np.testing.assert_array_equal(sm.reshape(-1, 3), film_o.sum)
    np.testing.assert_array_equal(l2.reshape(-1), film_o.lum2)
    assert traced == traced_o
    assert film_o.count.min() < film_o.count.max()  # some pixels stopped


@pytest.mark.gpu
def test_gpu_adaptive_everything_clean(dev):
    """A threshold no pixel misses: every pixel stops at its 8th sample and the next batch traces nothing."""
    sc, view = _scene()
    gs = dev.upload(sc)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, adaptive=True, noise_threshold=1e30,
                  min_adaptive_samples=8)
    r1 = render(dev, gs, cam, st, film, 0, 16)
    r2 = render(dev, gs, cam, st, film, 16, 16)
    _, cnt, _, _ = film.download()
    assert (cnt == 8).all()
    assert r1.samples_traced == 8 * view.width * view.height and r2.samples_traced == 0
    assert r1.completed and r2.completed


@pytest.mark.gpu
def test_gpu_session_adaptive_through_hip_device(dev):
    """The reference's RenderSession on the HipDevice adapter with adaptive sampling (one-sample batches, so the
    batch boundaries do not depend on timing): per-pixel sample counts as the oracle predicts from the GPU's own
    per-sample radiance, and the session ends early once a batch traces nothing (render_session.cpp:99)."""
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    sc, view = _scene()
    gs = dev.upload(sc)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    spp, threshold, min_spp = 40, 0.5, 8
    Ls = _gpu_samples(dev, gs, view, cam, spp, view.max_bounces)
    film_o = adaptive.Film(view.width, view.height)
    traced = [adaptive.render_batch(film_o, lambda s: Ls[s], s, 1, threshold, min_spp) for s in range(spp)]
    rc, _, counts, batches, err = bindings.HipDeviceSession(sc).render_adaptive(view, spp, 1, threshold, min_spp)
    assert rc == 0, err
    np.testing.assert_array_equal(counts.reshape(-1), film_o.count)
    expect_batches = next((k + 1 for k, t in enumerate(traced) if t == 0), spp)
    assert batches == expect_batches
