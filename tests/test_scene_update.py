"""pe_scene_update: RenderDevice::updateScene (engine/render_device.h:77-79) with the Vulkan device's partial
rebuilds as the model (VulkanDevice::updateScene, engine/vulkan_device.cpp:1223-1256): Materials rebuilds the
material table, Lights the lights, environment and light selection, both with the BVH and triangles kept
resident; Full and Transforms rebuild everything.

What is checked: after each kind of update the scene renders exactly what a fresh upload of the updated
description renders (same samples, same paths, bit for bit); a refused update leaves the scene rendering what it
rendered before.
"""
from __future__ import annotations

import numpy as np
import pytest

from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import Environment, dielectric


def _image(dev, gs, view, spp=4):
    film = dev.film(view.width, view.height)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spp)
    res = render(dev, gs, cam, st, film, 0, spp)
    assert res.completed == 1
    s, c, _, rej = film.download()
    assert (c == spp).all() and rej == 0
    return s


def _small():
    return configs.box_icosphere(level=2, width=32, height=32)


@pytest.mark.gpu
def test_materials_update_equals_fresh_upload(dev):
    sc, view = _small()
    gs = dev.upload(sc)
    before = _image(dev, gs, view)
    for m in sc.materials[:3]:
        m.base_color[:] = [0.9, 0.15, 0.1]
    sc.materials[6].roughness = 0.05
    assert gs.update(sc, capi.PE_UPDATE_MATERIALS) == capi.PE_OK
    after = _image(dev, gs, view)
    np.testing.assert_array_equal(after, _image(dev, dev.upload(sc), view))
    assert not np.array_equal(after, before)


@pytest.mark.gpu
def test_materials_update_to_glass_switches_transport(dev):
    sc, view = _small()
    gs = dev.upload(sc)
    sc.materials[6] = dielectric(1.5, abbe=30.0)     # the icosphere becomes dispersive glass
    assert gs.update(sc, capi.PE_UPDATE_MATERIALS) == capi.PE_OK
    np.testing.assert_array_equal(_image(dev, gs, view), _image(dev, dev.upload(sc), view))


@pytest.mark.gpu
def test_lights_update_equals_fresh_upload(dev):
    sc, view = _small()
    gs = dev.upload(sc)
    before = _image(dev, gs, view)
    sc.lights[0].radiance[:] = [3.0, 6.0, 12.0]
    sc.set_environment(Environment(np.full((16, 32, 3), 0.25, np.float32)))
    assert gs.update(sc, capi.PE_UPDATE_LIGHTS) == capi.PE_OK
    after = _image(dev, gs, view)
    np.testing.assert_array_equal(after, _image(dev, dev.upload(sc), view))
    assert not np.array_equal(after, before)
    sc.environment = None                            # and away again
    assert gs.update(sc, capi.PE_UPDATE_LIGHTS) == capi.PE_OK
    np.testing.assert_array_equal(_image(dev, gs, view), _image(dev, dev.upload(sc), view))


@pytest.mark.gpu
def test_refused_update_leaves_scene_unchanged(dev):
    sc, view = _small()
    gs = dev.upload(sc)
    before = _image(dev, gs, view)
    short = configs.box_icosphere(level=2, width=32, height=32)[0]
    short.materials = short.materials[:3]            # meshes still use materials 3..7
    assert gs.update(short, capi.PE_UPDATE_MATERIALS) == capi.PE_ERR_INVALID
    assert gs.update(sc, 7) == capi.PE_ERR_INVALID   # unknown kind
    np.testing.assert_array_equal(_image(dev, gs, view), before)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", [capi.PE_UPDATE_FULL, capi.PE_UPDATE_TRANSFORMS])
def test_full_and_transform_updates_rebuild(dev, kind):
    sc, view = _small()
    gs = dev.upload(sc)
    moved, _ = _small()
    moved.shapes[6].mesh.positions = moved.shapes[6].mesh.positions + np.float32([0.2, 0.1, 0.0])
    assert gs.update(moved, kind) == capi.PE_OK
    np.testing.assert_array_equal(_image(dev, gs, view), _image(dev, dev.upload(moved), view))
    other, oview = configs.product_sphere(width=32, height=32, lon=48, lat=32, env_width=64, env_height=32)
    assert gs.update(other, kind) == capi.PE_OK
    assert gs.stats()["triangles"] == other.num_triangles
    np.testing.assert_array_equal(_image(dev, gs, oview), _image(dev, dev.upload(other), oview))
