"""The HipDevice adapter (photon-engine_amd/host/hip_device.cpp): the RenderDevice a maintainer adds to the
reference, compiled against the reference's own headers and driven by the reference's RenderSession
(engine/render_session.cpp:1-200) through oracle/hipdev_driver.cpp.

What is checked: the adapter packs a genuine reference Scene (built by the reference's code from the same
descriptor) into the C-ABI, and RenderSession's batches through HipDevice::render give the image that pe_render
gives directly (up to the float association of Film::addBatch across batches) and that the reference CPU renderer
gives (relMSE); features the HIP path lacks are refused with a message, not approximated.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import camera_perspective, configs, render, settings
from photon_hip import capi
from photon_hip.scene import Mesh, Scene, disney, lambert


def _harness_or_skip():
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def test_adapter_library_loads_and_exports():
    _harness_or_skip()
    lib = bindings.hip_device_harness()
    for name in ("hipdev_session_render", "hipdev_session_interactive", "hipdev_enumerate", "ref_scene_build",
                 "ref_scene_free"):
        assert hasattr(lib.lib, name), name


def _direct_mean(dev, sc, view, spp, spl, first=0):
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl)
    res = render(dev, gs, cam, st, film, first, spp)
    assert res.completed == 1
    s, c, _, rej = film.download()
    assert rej == 0
    return s / c[..., None].astype(np.float32), c


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["box_icosphere", "product_sphere"])
def test_session_through_hip_device_matches_direct_render(dev, name):
    _harness_or_skip()
    if name == "box_icosphere":
        sc, view = configs.box_icosphere(level=3, width=32, height=24)
    else:
        sc, view = configs.product_sphere(width=40, height=24, lon=64, lat=48, env_width=256, env_height=128)
    spp = 12
    rc, mean, counts, batches, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=4)
    assert rc == 0, err
    assert (counts == spp).all()
    assert batches >= 1
    direct, c = _direct_mean(dev, sc, view, spp, spl=4)
    assert (c == spp).all()
    # Same samples, same paths: the session's batches only change how the per-pixel sums are associated.
    np.testing.assert_allclose(mean, direct, rtol=2e-5, atol=1e-6)
    ref = bindings.reference(strict=True)
    if ref is not None:
        mean_r, cnt_r, _ = bindings.RefScene(sc, strict=True).render(view, 0, spp, threads=4)
        assert (cnt_r == spp).all()
        assert relmse(mean, mean_r) <= 1e-3, relmse(mean, mean_r)


@pytest.mark.gpu
def test_hip_device_refuses_what_it_does_not_implement(dev):
    """Settings the HIP path does not implement (here ambient occlusion) make the batch fail with a message
    instead of an approximation; the session reports it."""
    _harness_or_skip()
    sc, view = configs.box_icosphere(level=1, width=8, height=8)
    rc, _, _, _, err = bindings.HipDeviceSession(sc).render(view, 1, ao_strength=0.5)
    assert rc == 1
    assert "not implemented" in err, err


@pytest.mark.gpu
def test_interactive_passes_and_cancel_through_hip_device(dev):
    """RenderSession's Interactive mode (one-sample passes, render_session.cpp:69-71) through HipDevice gives the
    image of a direct render of the same samples, pass by pass at viewport latency; a cancelled Final batch
    returns soon after the flag (the library polls it between wavefront iterations), reports the batch
    incomplete, and the session keeps only its completed samples."""
    _harness_or_skip()
    sc, view = configs.product_sphere(width=320, height=180, lon=128, lat=96, env_width=512, env_height=256)
    spp = 8
    r = bindings.HipDeviceSession(sc).interactive(view, spp, cancel_spp=100000, cancel_after_ms=40)
    assert r["rc"] == 0, r["error"]
    assert (r["counts"] == spp).all()
    direct, c = _direct_mean(dev, sc, view, spp, spl=1)
    np.testing.assert_allclose(r["mean"], direct, rtol=2e-5, atol=1e-6)
    # The first pass carries the scene upload's first launches; the steady passes are viewport-fast.
    assert np.median(r["pass_seconds"][1:]) < 0.05, r["pass_seconds"]
    assert r["cancel_completed"] == 0
    assert r["cancel_samples"] == 1
    assert r["cancel_latency"] < 0.25, r["cancel_latency"]
    print(f"interactive pass median {np.median(r['pass_seconds'][1:]) * 1e3:.2f} ms, cancel latency "
          f"{r['cancel_latency'] * 1e3:.2f} ms")


@pytest.mark.gpu
def test_path_guiding_through_hip_device_converges_to_plain(dev):
    """RenderSettings::pathGuiding through HipDevice (renderer.cpp:413-429): RenderSession's batches train the
    learned guide on the first launches and then sample it with one-sample MIS; the estimate stays unbiased, so
    the guided image of the door-gap room converges to the plain one (relMSE against a plain 1024-spp image of
    the same frame no worse than the plain render's own)."""
    _harness_or_skip()
    sc, view = configs.slit_room(width=96, height=54, max_bounces=6)
    spp = 256
    rc, guided, counts, batches, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=32,
                                                                         path_guiding=True)
    assert rc == 0, err
    assert (counts == spp).all() and batches > 1
    rc, plain, _, _, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=32)
    assert rc == 0, err
    ref, _ = _direct_mean(dev, sc, view, 1024, spl=64, first=1 << 20)  # independent samples
    e_g, e_p = relmse(guided, ref), relmse(plain, ref)
    print(f"guided relMSE {e_g:.4f}, plain {e_p:.4f} against plain 1024 spp", flush=True)
    assert np.isfinite(guided).all()
    assert abs(guided.mean() / plain.mean() - 1.0) < 0.05
    assert e_g <= 1.5 * e_p
