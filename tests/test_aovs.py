"""AOVs: RenderTargets with AovFilms (engine/render_device.h:58-62, engine/renderer.h:42-53), recorded per sample
as tracePixel records them for an environment background (engine/renderer.cpp:175-205, primaryAlbedo :47-58,
addAovs :222-227).

CPU: the oracle's per-sample AOVs (oracle_trace_aovs) equal the reference's own Renderer::renderSamples with
AovFilms (IEEE build), one sample at a time so that each resolved pixel is that sample: albedo, normal, coverage
and beauty, on an area-lit box, an environment-lit sphere (camera rays that miss take the clamped environment
radiance as albedo) and the glass studio.
GPU: pe_render_targets against the oracle per sample (coverage and hit albedo exact where the traversal agrees,
environment albedo and shading normals to float tolerance: the kernels contract a*b+c and use the device's
trigonometry), a multi-sample launch equals the per-sample films added in sample order bit for bit, and asking
for AOVs leaves the beauty film unchanged.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, render_targets, settings


def _scenes():
    box, vb = configs.box_icosphere(level=3, width=20, height=16)
    sph, vs = configs.product_sphere(width=24, height=16, lon=64, lat=48, env_width=256, env_height=128)
    glass, vg = configs.glass_studio(width=24, height=16)
    return {"box_icosphere": (box, vb), "product_sphere": (sph, vs), "glass_studio": (glass, vg)}


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _pixels(view):
    py, px = np.divmod(np.arange(view.width * view.height), view.width)
    return px, py


@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere", "glass_studio"])
def test_oracle_aovs_equal_reference(name):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = _scenes()[name]
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    px, py = _pixels(view)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    for s in (0, 5):
        L, alb, nrm, alpha = o.trace_aovs(_cam(view), st, view.width, view.height, px, py, np.full(px.size, s))
        b, a, n, c = (x.reshape(-1, 3) for x in ref.render_aovs(view, s, 1))
        np.testing.assert_array_equal(L, b)
        np.testing.assert_array_equal(alb, a)
        np.testing.assert_array_equal(nrm, n)
        np.testing.assert_array_equal(np.repeat(alpha[:, None], 3, 1), c)
    if name == "product_sphere":
        assert 0.0 < alpha.mean() < 1.0  # both branches: surface and background


# ------------------------------------------------------------------------------------------------- GPU

def _gpu_aovs(dev, gs, view, first, count, spl=1):
    films = [dev.film(view.width, view.height) for _ in range(4)]
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl)
    res = render_targets(dev, gs, _cam(view), st, films[0], first, count, albedo=films[1], normal=films[2],
                         alpha=films[3])
    assert res.completed == 1 and res.samples_traced == count * view.width * view.height
    return [f.download() for f in films]


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere", "glass_studio"])
def test_gpu_aovs_match_oracle_per_sample(dev, name):
    sc, view = _scenes()[name]
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    px, py = _pixels(view)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    agree = total = 0
    for s in (0, 3, 11):
        beauty, alb, nrm, cov = _gpu_aovs(dev, gs, view, s, 1)
        assert (alb[1] == 1).all() and (nrm[1] == 1).all() and (cov[1] == 1).all()
        _, alb_o, nrm_o, alpha_o = o.trace_aovs(_cam(view), st, view.width, view.height, px, py, np.full(px.size, s))
        cov_g = cov[0].reshape(-1, 3)
        assert (cov_g[:, 0] == cov_g[:, 1]).all() and (cov_g[:, 0] == cov_g[:, 2]).all()
        same = cov_g[:, 0] == alpha_o  # the same hit/miss outcome (a traversal tie may differ)
        agree += int(same.sum())
        total += same.size
        hit = same & (alpha_o == 1.0)
        miss = same & (alpha_o == 0.0)
        alb_g, nrm_g = alb[0].reshape(-1, 3), nrm[0].reshape(-1, 3)
        np.testing.assert_array_equal(alb_g[hit], alb_o[hit])
        np.testing.assert_allclose(alb_g[miss], alb_o[miss], rtol=1e-4, atol=1e-6)
        np.testing.assert_allclose(nrm_g[hit], nrm_o[hit], atol=2e-5)
        assert (nrm_g[miss] == 0).all()
    assert agree >= 0.995 * total, (agree, total)


@pytest.mark.gpu
def test_gpu_aov_films_accumulate_in_sample_order(dev):
    """One launch of 6 samples equals the six single-sample films added in sample order, and the beauty film is
    the one a render without AOVs gives."""
    sc, view = _scenes()["product_sphere"]
    gs = dev.upload(sc)
    multi = _gpu_aovs(dev, gs, view, 0, 6, spl=6)
    singles = [_gpu_aovs(dev, gs, view, s, 1) for s in range(6)]
    for k in range(1, 4):
        acc = np.zeros_like(singles[0][k][0])
        for s in range(6):
            acc = (acc + singles[s][k][0]).astype(np.float32)
        np.testing.assert_array_equal(multi[k][0], acc)
        assert (multi[k][1] == 6).all()
    film = dev.film(view.width, view.height)
    render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=6), film, 0, 6)
    plain = film.download()
    for a, b in zip(plain, multi[0]):
        np.testing.assert_array_equal(a, b)


@pytest.mark.gpu
def test_gpu_session_aovs_through_hip_device(dev):
    """The reference's RenderSession with SessionConfig::aovs on the HipDevice adapter: the AovFilms it keeps
    resolve to the AOV images a direct pe_render_targets of the same samples gives (up to the float association
    of Film::addBatch across the session's batches)."""
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    sc, view = _scenes()["product_sphere"]
    spp = 12
    rc, imgs, counts, err = bindings.HipDeviceSession(sc).render_aovs(view, spp, 4)
    assert rc == 0, err
    assert (counts == spp).all()
    gs = dev.upload(sc)
    direct = _gpu_aovs(dev, gs, view, 0, spp, spl=spp)
    for k in range(4):
        s, c = direct[k][0], direct[k][1]
        np.testing.assert_allclose(imgs[k], s / c[..., None].astype(np.float32), rtol=1e-5, atol=1e-6)
