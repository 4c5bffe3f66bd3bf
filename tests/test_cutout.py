"""Disney opacity: DisneyMaterial::setOpacity (materials/disney.h:147-158) as a stochastic cutout. A ray meeting
a surface of opacity a passes it with probability 1 - a, a sampler draw (one dimension) deciding when 0 < a < 1
(passesCutout, integrators/path_tracer.cpp:159-167, in the skip loop :423-441 that also passes camera-invisible
lights); a shadow ray is attenuated by 1 - a at each such surface (shadowTransmittance :172-202,
disney.cpp:511-513); a cutout vertex is never guided (:624-627).

CPU: the oracle's per-sample radiance and AOVs equal the reference's (IEEE build) on a studio with half-opaque,
fully transparent and opaque-flagged surfaces; an object of opacity 0 leaves the samples as they were without it
(but for the few paths its restarted rays turn), and opacity 1 changes nothing.
GPU: pe_render per sample against the oracle (radiance and AOVs), and RenderSession through HipDevice (which
packs the material's opacity) against a direct render.
"""
from __future__ import annotations

import copy

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi
from photon_hip.configs import View, box_mesh, grid_quad, icosphere
from photon_hip.device import camera_perspective, render, render_targets, settings
from photon_hip.scene import AreaLight, Environment, Scene, disney, lambert


def _studio(ghost_opacity=0.0, with_ghost=True):
    """A floor, a half-opaque screen between the camera and a block, a quarter-opaque sphere casting a soft
    shadow, an opaque box flagged with opacity 1, and (optionally) a ghost box of the given opacity. The boxes
    stand 2 mm above the floor: a face coplanar with the floor would make every ray through a cutout box there
    a tie between the two surfaces, which the traversal order settles."""
    sc = Scene()
    sc.add_shape(grid_quad((-3, 0, -3), (0, 0, 6), (6, 0, 0), 2, 2), sc.add_material(lambert((0.6, 0.6, 0.6))))
    sc.add_shape(grid_quad((-1.0, 0.1, 0.9), (0.0, 1.2, 0.0), (1.6, 0.0, 0.0), 1, 1),
                 sc.add_material(disney((0.2, 0.5, 0.9), roughness=0.4, opacity=0.5)))
    sc.add_shape(box_mesh((-0.5, 0.002, -0.5), (0.1, 0.6, 0.0)), sc.add_material(disney((0.8, 0.3, 0.2), roughness=0.3)))
    sc.add_shape(icosphere(2, 0.3, (0.7, 0.9, 0.0)), sc.add_material(disney((0.9, 0.9, 0.2), 0.5, 0.3, opacity=0.25)))
    sc.add_shape(box_mesh((0.4, 0.002, -0.8), (0.9, 0.3, -0.3)), sc.add_material(disney((0.3, 0.8, 0.3), opacity=1.0)))
    if with_ghost:
        sc.add_shape(box_mesh((-1.2, 0.002, -0.2), (-0.7, 0.8, 0.3)),
                     sc.add_material(disney((1.0, 0.0, 1.0), roughness=0.2, opacity=ghost_opacity)))
    sc.add_light(AreaLight((0.3, 2.2, -0.2), (0.6, 0.0, 0.0), (0.0, 0.0, 0.6), (9.0, 9.0, 9.0)))
    sc.set_environment(Environment(np.full((16, 32, 3), 0.3, np.float32)))
    return sc, View((0.2, 1.1, 3.0), (0.0, 0.4, 0.0), (0, 1, 0), 45.0, 32, 24, 16, 8)


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _samples(view, n=2500, seed=4):
    rng = np.random.default_rng(seed)
    return rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 512, n)


def _ref_or_skip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


def test_oracle_cutouts_equal_reference_per_sample():
    _ref_or_skip()
    sc, view = _studio(ghost_opacity=0.6)
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    px, py, s = _samples(view)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    Lr = ref.trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    for k in (0, 5):
        L, alb, nrm, alpha = o.trace_aovs(_cam(view), st, view.width, view.height, xx, yy, np.full(xx.size, k))
        b, a, n, c = (x.reshape(-1, 3) for x in ref.render_aovs(view, k, 1))
        np.testing.assert_array_equal(L, b)
        np.testing.assert_array_equal(alb, a)
        np.testing.assert_array_equal(nrm, n)
        np.testing.assert_array_equal(np.repeat(alpha[:, None], 3, 1), c)


def test_oracle_invisible_and_opaque_cutouts():
    """Opacity 0 passes every ray and every shadow ray at full strength without a draw: with the ghost in the
    scene the samples are those of the scene without it, up to the ray restarting a few ulps off its line past
    each face (offsetRayOrigin), which turns a path here and there; opacity 1 flagged or not is the same
    material, sample for sample."""
    sc0, view = _studio(with_ghost=False)
    sc1, _ = _studio(ghost_opacity=0.0)
    px, py, s = _samples(view, 1500, 9)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    L0, _ = bindings.OracleScene(sc0).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    L1, _ = bindings.OracleScene(sc1).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    close = np.all(np.abs(L0 - L1) <= 1e-4 * np.maximum(1.0, np.abs(L0)), axis=1)
    assert close.mean() >= 0.995, close.mean()
    assert np.abs(L0.mean(0) - L1.mean(0)).max() <= 0.01 * L0.mean()
    sc2 = copy.deepcopy(sc0)
    sc2.materials[4].has_opacity = 0
    L2, _ = bindings.OracleScene(sc2).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    np.testing.assert_array_equal(L0, L2)


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
def test_gpu_cutouts_match_oracle_per_sample(dev):
    sc, view = _studio(ghost_opacity=0.6)
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    close_n = total = agree = 0
    for sample in (1, 6, 13):
        films = [dev.film(view.width, view.height) for _ in range(4)]
        res = render_targets(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345),
                             films[0], sample, 1, albedo=films[1], normal=films[2], alpha=films[3])
        assert res.completed == 1
        beauty, alb, _, cov = (f.download() for f in films)
        L, alb_o, _, alpha_o = o.trace_aovs(_cam(view), st, view.width, view.height, xx, yy, np.full(xx.size, sample))
        Lg = beauty[0].reshape(-1, 3)
        close = np.all(np.abs(Lg - L) <= 1e-4 * np.maximum(1.0, np.abs(L)), axis=1)
        close_n += int(close.sum())
        total += close.size
        hit = alpha_o == 1.0
        same_alb = np.all(alb[0].reshape(-1, 3) == alb_o, axis=1)
        agree += int((same_alb | ~hit).sum())
    print(f"cutouts: per-sample agreement {close_n / total:.5f}, primary albedo {agree / total:.5f} of {total}",
          flush=True)
    assert close_n >= 0.998 * total
    assert agree >= 0.995 * total


@pytest.mark.gpu
def test_gpu_cutouts_through_hip_device(dev):
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    sc, view = _studio(ghost_opacity=0.6)
    spp = 8
    rc, mean, counts, _, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=4)
    assert rc == 0, err
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4), film, 0,
           spp)
    s, c, _, _ = film.download()
    np.testing.assert_allclose(mean, s / c[..., None], rtol=2e-5, atol=1e-6)
