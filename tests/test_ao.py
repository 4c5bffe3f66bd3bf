"""Contact ambient occlusion: RenderSettings::aoStrength -> PathTracer(maxDepth, rrDepth, aoStrength, shadowQuality)
(renderer.cpp:432). At the first vertex of a camera path, shadowQuality cosine-distributed any-hit rays of 0.35
scene units (approxAo, path_tracer.cpp:27-44) scale the path's throughput by lerp(1, unoccluded fraction,
strength) before its NEE (:617-621); emission seen directly is not scaled.

CPU: the oracle's per-sample radiance with AO equals the reference's (IEEE build) bit for bit, for one and for
two shadow samples; closed forms (strength 0 is the plain render sample for sample; an open scene with nothing
within 0.35 of its surfaces is unchanged by AO of any strength; full strength in the box corners darkens).
GPU: pe_render per sample against the oracle (every AO ray a shadow-queue entry, its factor applied when the next
vertex or the path's finalisation folds the first vertex's NEE), in a scene with thin walls too.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import configs
from photon_hip.configs import View, box_mesh, grid_quad, icosphere
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import AreaLight, Environment, Scene, dielectric, disney, lambert


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _samples(view, n, seed):
    rng = np.random.default_rng(seed)
    return rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 512, n)


def _ref_or_skip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


def _walls_scene():
    """A floor, a box against a wall, a thin glass pane and a sphere: corners and contact shadows within 0.35."""
    sc = Scene()
    sc.add_shape(grid_quad((-2, 0, -2), (0, 0, 4), (4, 0, 0), 2, 2), sc.add_material(lambert((0.7, 0.7, 0.7))))
    sc.add_shape(grid_quad((-2, 0, -1), (4, 0, 0), (0, 2, 0), 2, 2), sc.add_material(disney((0.6, 0.5, 0.4), 0.0, 0.5)))
    sc.add_shape(box_mesh((-0.6, 0.0, -1.0), (0.0, 0.5, -0.5)), sc.add_material(disney((0.3, 0.5, 0.8), 0.0, 0.3)))
    sc.add_shape(icosphere(2, 0.3, (0.6, 0.3, -0.4)), sc.add_material(disney((0.8, 0.3, 0.2), 1.0, 0.25)))
    sc.add_shape(grid_quad((-0.4, 0.0, 0.4), (0.8, 0, 0), (0, 0.6, 0), 1, 1),
                 sc.add_material(dielectric(1.5, thin_walled=True)))
    sc.add_light(AreaLight((-0.3, 1.9, -0.3), (0.6, 0, 0), (0, 0, 0.6), (8.0, 8.0, 8.0)))
    sc.set_environment(Environment(np.full((16, 32, 3), 0.2, np.float32)))
    return sc, View((0.3, 1.0, 2.2), (0.0, 0.3, -0.5), (0, 1, 0), 50.0, 32, 24, 16, 6)


@pytest.mark.parametrize("shadow_quality", [1, 2])
@pytest.mark.parametrize("strength", [0.6, 1.0])
def test_oracle_ao_equals_reference_per_sample(strength, shadow_quality):
    _ref_or_skip()
    sc, view = configs.box_icosphere(level=2, width=24, height=24)
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    px, py, s = _samples(view, 2500, 3)
    st = settings(max_bounces=view.max_bounces, shadow_quality=shadow_quality, seed=12345, ao_strength=strength)
    Lo, _ = o.trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    with bindings.reference_camera(strict=True, ao_strength=strength):
        Lr = ref.trace_samples(view, px, py, s, shadow_quality=shadow_quality, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)


def test_ao_closed_forms():
    sc, view = configs.box_icosphere(level=2, width=24, height=24)
    o = bindings.OracleScene(sc)
    px, py, s = _samples(view, 3000, 5)
    plain, _ = o.trace_samples(_cam(view), settings(max_bounces=view.max_bounces), view.width, view.height, px, py, s)
    full, _ = o.trace_samples(_cam(view), settings(max_bounces=view.max_bounces, ao_strength=1.0), view.width,
                              view.height, px, py, s)
    assert full.mean() < 0.97 * plain.mean()
    # an isolated sphere under the sky: nothing within 0.35 of it, so AO changes no sample
    sc2 = Scene()
    sc2.add_shape(icosphere(2, 0.5, (0.0, 0.0, 0.0)), sc2.add_material(disney((0.6, 0.6, 0.6), 0.0, 0.4)))
    sc2.set_environment(Environment(np.full((16, 32, 3), 0.5, np.float32)))
    v2 = View((0.0, 0.3, 2.5), (0.0, 0.0, 0.0), (0, 1, 0), 40.0, 16, 16, 16, 4)
    o2 = bindings.OracleScene(sc2)
    px, py, s = _samples(v2, 1500, 6)
    a, _ = o2.trace_samples(_cam(v2), settings(max_bounces=4), 16, 16, px, py, s)
    b, _ = o2.trace_samples(_cam(v2), settings(max_bounces=4, ao_strength=1.0), 16, 16, px, py, s)
    hit = np.abs(a - b).max(axis=1) > 0
    # the draws AO consumes shift the later dimensions: compare the images' means, not samples
    assert abs(a.mean() - b.mean()) < 0.03 * a.mean(), (a.mean(), b.mean(), hit.mean())


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("case", ["box_icosphere", "walls"])
def test_gpu_ao_matches_oracle_per_sample(dev, case):
    if case == "walls":
        sc, view = _walls_scene()
    else:
        sc, view = configs.box_icosphere(level=2, width=24, height=24)
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    close_n = total = 0
    for sq, strength in ((1, 0.8), (2, 1.0)):
        st = settings(max_bounces=view.max_bounces, shadow_quality=sq, seed=12345, ao_strength=strength)
        for sample in (0, 9):
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
    print(f"ao {case}: per-sample agreement {close_n / total:.5f} of {total}", flush=True)
    # the bar of the other scenes with glass (a path through a pane may branch on a last-ulp difference)
    assert close_n >= (0.998 if case == "walls" else 0.999) * total
