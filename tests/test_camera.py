"""Cameras: PerspectiveCamera, ThinLensCamera (disk and polygonal apertures) and OrthographicCamera, with the
shutter-end pose of camera motion blur (camera/camera.h:19-91, perspective_camera.cpp, thin_lens_camera.cpp,
orthographic_camera.cpp). tracePixel draws the pixel-filter offset, the lens sample and, with
RenderSettings::motionBlur, a shutter time before Li (renderer.cpp:143-153).

CPU: pe_camera_* equal the reference's describeGpu / describeGpuEnd bit for bit; the oracle's camera rays equal
the reference's Camera::generateRay (IEEE build) bit for bit for every kind, with and without a shutter time;
closed forms that do not need the reference (every lens ray of a screen point passes through its point on the
focus plane, disk samples fill the aperture disk, polygon samples stay inside the blades-gon and cover every
slice equally, orthographic rays are parallel); the oracle's per-sample radiance and depth pass equal the
reference's tracePixel.
GPU: pe_render per sample against the oracle for every kind and for motion blur, the depth pass of a thin lens
(measured from the lens point), and RenderSession through HipDevice with a ThinLensCamera under motion blur
against a direct render.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip.configs import View, box_mesh, grid_quad, icosphere
from photon_hip.device import (PassFilms, camera_orthographic, camera_perspective, camera_thin_lens,
                               camera_with_shift, camera_with_shutter_end, render, render_targets, settings)
from photon_hip.scene import AreaLight, Environment, Scene, disney, lambert

FOCUS = 2.6
APERTURE = 0.12
END_POS = (0.5, 1.3, 2.9)
END_TARGET = (0.1, 0.3, 0.0)

# kind -> (reference_camera keywords, pe_camera builder, motion blur)
CASES = {
    "thin_lens_disk": (dict(kind="thin_lens", aperture=APERTURE, focus=FOCUS), False),
    "thin_lens_hexagon": (dict(kind="thin_lens", aperture=APERTURE, focus=FOCUS, blades=6, rotation_deg=15.0), False),
    "orthographic": (dict(kind="orthographic", extent=2.4), False),
    "perspective_moving": (dict(kind="perspective", motion_blur=True, end_position=END_POS, end_target=END_TARGET),
                           True),
    "thin_lens_moving": (dict(kind="thin_lens", aperture=APERTURE, focus=FOCUS, blades=5, motion_blur=True,
                              end_position=END_POS, end_target=END_TARGET, end_aperture=0.05), True),
    "orthographic_moving": (dict(kind="orthographic", extent=2.4, motion_blur=True, end_position=END_POS,
                                 end_target=END_TARGET), True),
    "perspective_shift": (dict(kind="perspective", shift=(0.12, -0.07)), False),
    "thin_lens_shift_moving": (dict(kind="thin_lens", aperture=APERTURE, focus=FOCUS, shift=(-0.05, 0.1),
                                    motion_blur=True, end_position=END_POS, end_target=END_TARGET,
                                    end_aperture=0.2), True),
}


def _studio():
    """Objects from 1 to 4.5 m away (a thin lens focused at 2.6 m blurs the near and far ones), a glossy floor,
    an area light and a dim environment."""
    sc = Scene()
    sc.add_shape(grid_quad((-4, 0, -4), (0, 0, 8), (8, 0, 0), 2, 2), sc.add_material(disney((0.6, 0.6, 0.6), 0.0, 0.35)))
    sc.add_shape(icosphere(2, 0.35, (-0.6, 0.35, 1.4)), sc.add_material(disney((0.9, 0.3, 0.2), roughness=0.3)))
    sc.add_shape(box_mesh((-0.2, 0.002, -0.4), (0.5, 0.9, 0.3)), sc.add_material(disney((0.2, 0.7, 0.3), 1.0, 0.2)))
    sc.add_shape(icosphere(2, 0.5, (1.0, 0.5, -1.6)), sc.add_material(lambert((0.3, 0.4, 0.9))))
    sc.add_light(AreaLight((0.2, 2.4, -0.3), (0.6, 0.0, 0.0), (0.0, 0.0, 0.6), (10.0, 10.0, 10.0)))
    sc.set_environment(Environment(np.full((16, 32, 3), 0.25, np.float32)))
    return sc, View((0.2, 1.2, 3.2), (0.0, 0.4, 0.0), (0, 1, 0), 40.0, 32, 24, 16, 8)


def _pe_camera(view, kw):
    """The pe_camera of what reference_camera(**kw) builds for `view`."""
    def at(pos, target, aperture):
        if kw["kind"] == "thin_lens":
            cam = camera_thin_lens(pos, target, view.up, view.vfov, view.aspect, aperture, kw["focus"],
                                   kw.get("blades", 0), kw.get("rotation_deg", 0.0))
        elif kw["kind"] == "orthographic":
            return camera_orthographic(pos, target, view.up, kw["extent"], view.aspect)
        else:
            cam = camera_perspective(pos, target, view.up, view.vfov, view.aspect)
        return camera_with_shift(cam, *kw["shift"]) if "shift" in kw else cam

    cam = at(view.position, view.target, kw.get("aperture", 0.0))
    if kw.get("end_position") is not None:
        cam = camera_with_shutter_end(cam, at(kw["end_position"], kw["end_target"], kw.get("end_aperture", 0.0)))
    return cam


def _desc_floats(cam):
    f = lambda a: list(a)  # noqa: E731
    open_ = (f(cam.position) + f(cam.lower_left) + f(cam.horizontal) + f(cam.vertical) + f(cam.u) + f(cam.v)
             + f(cam.forward) + [cam.lens_radius])
    if cam.has_shutter_end:
        end = (f(cam.end_position) + f(cam.end_lower_left) + f(cam.end_horizontal) + f(cam.end_vertical)
               + f(cam.end_u) + f(cam.end_v) + f(cam.end_forward) + [cam.end_lens_radius])
    else:
        end = open_
    return np.array(open_ + end + [float(cam.blades), cam.blade_rotation, float(cam.orthographic)], np.float32)


def _ref_or_skip():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


def _queries(n, seed, with_time):
    rng = np.random.default_rng(seed)
    q = rng.random((n, 5), dtype=np.float32)
    q[:8, 2:4] = [[0.5, 0.5], [0.0, 0.0], [0.999, 0.5], [0.5, 0.0], [0.25, 0.75], [1e-4, 0.999], [0.75, 0.75],
                  [0.5, 0.999]]
    if with_time:
        q[::7, 4] = 0.0
    else:
        q[:, 4] = -1.0
    return q


@pytest.mark.parametrize("case", sorted(CASES))
def test_camera_descriptions_equal_reference(case):
    _ref_or_skip()
    kw, _ = CASES[case]
    _, view = _studio()
    with bindings.reference_camera(strict=True, **kw) as rc:
        want = rc.describe(view)
    got = _desc_floats(_pe_camera(view, kw))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_camera_rays_equal_reference(case):
    _ref_or_skip()
    kw, moving = CASES[case]
    _, view = _studio()
    q = _queries(4000, 11, moving)
    with bindings.reference_camera(strict=True, **kw) as rc:
        want = rc.rays(view, q)
    got = bindings.camera_rays_oracle(_pe_camera(view, kw), q)
    np.testing.assert_array_equal(got, want)


def test_thin_lens_closed_forms():
    """Every lens ray of one screen point meets the focus plane at that point's image-plane position; disk
    apertures fill the disk of the lens radius, a hexagonal one stays inside the hexagon and gives each of its
    six slices a sixth of the samples; orthographic rays are parallel to the view direction."""
    _, view = _studio()
    rng = np.random.default_rng(3)
    for blades in (0, 6):
        cam = camera_thin_lens(view.position, view.target, view.up, view.vfov, view.aspect, APERTURE, FOCUS, blades,
                               30.0)
        pos, ll = np.array(cam.position), np.array(cam.lower_left)
        hor, ver, u, v = (np.array(a) for a in (cam.horizontal, cam.vertical, cam.u, cam.v))
        fwd = np.array(cam.forward)
        n = 20000
        q = np.zeros((n, 5), np.float32)
        q[:, :2] = rng.random(2)  # one screen point
        q[:, 2:4] = rng.random((n, 2))
        q[:, 4] = -1.0
        r = bindings.camera_rays_oracle(cam, q).astype(np.float64)
        target = ll + q[0, 0] * hor + q[0, 1] * ver
        depth = (target - pos) @ fwd
        t = depth / (r[:, 3:] @ fwd)
        hit = r[:, :3] + t[:, None] * r[:, 3:]
        assert np.abs(hit - target).max() < 2e-5 * FOCUS
        off = r[:, :3] - pos
        lx, ly = off @ u, off @ v
        rad = np.hypot(lx, ly) / APERTURE
        assert np.abs(off @ fwd).max() < 1e-6
        assert rad.max() <= 1.0 + 1e-5
        if blades == 0:
            # uniform on the disk: P(r < 0.5) = 1/4, P(r < 0.9) = 0.81
            assert abs((rad < 0.5).mean() - 0.25) < 0.01 and abs((rad < 0.9).mean() - 0.81) < 0.01
        else:
            ang = np.mod(np.arctan2(ly, lx) - np.deg2rad(30.0), 2 * np.pi)
            slice_ = np.floor(ang / (2 * np.pi / 6)).astype(int)
            assert np.abs(np.bincount(slice_, minlength=6) / n - 1 / 6).max() < 0.012
            # inside the hexagon: the distance along each slice's apothem direction is at most cos(30 deg)
            mid = np.deg2rad(30.0) + (slice_ + 0.5) * (2 * np.pi / 6)
            assert (rad * np.cos(ang + np.deg2rad(30.0) - mid)).max() <= np.cos(np.pi / 6) + 1e-5
    cam = camera_orthographic(view.position, view.target, view.up, 2.4, view.aspect)
    q = rng.random((1000, 5), dtype=np.float32)
    q[:, 4] = -1.0
    r = bindings.camera_rays_oracle(cam, q)
    assert np.abs(r[:, 3:] - np.array(cam.forward)).max() == 0.0
    span = (r[:, :3] - np.array(cam.lower_left)) @ np.array(cam.u)
    assert span.min() >= -1e-5 and span.max() <= np.linalg.norm(cam.horizontal) + 1e-5


def _samples(view, n=2000, seed=6):
    rng = np.random.default_rng(seed)
    return rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 512, n)


@pytest.mark.parametrize("case", sorted(CASES))
def test_oracle_per_sample_equals_reference(case):
    _ref_or_skip()
    kw, moving = CASES[case]
    sc, view = _studio()
    o, ref = bindings.OracleScene(sc), bindings.RefScene(sc, strict=True)
    px, py, s = _samples(view)
    st = settings(max_bounces=view.max_bounces, seed=12345, motion_blur=moving)
    Lo, _ = o.trace_samples(_pe_camera(view, kw), st, view.width, view.height, px, py, s)
    with bindings.reference_camera(strict=True, **kw):
        Lr = ref.trace_samples(view, px, py, s, seed=12345)
        Lp, passes = ref.trace_passes(view, px[:600], py[:600], s[:600], seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    Lo2, po = o.trace_passes(_pe_camera(view, kw), st, view.width, view.height, px[:600], py[:600], s[:600])
    np.testing.assert_array_equal(Lo2, Lp)
    np.testing.assert_array_equal(po["depth"], passes["depth"])
    np.testing.assert_array_equal(po["position"], passes["position"])


def test_lens_sample_is_drawn_by_every_camera():
    """The pinhole draws the lens sample too (renderer.cpp:151), so a thin lens of radius 0 renders the pinhole's
    samples exactly, and motion blur without a shutter end changes only the later dimensions."""
    sc, view = _studio()
    o = bindings.OracleScene(sc)
    px, py, s = _samples(view, 800, 2)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    pin = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    zero = camera_thin_lens(view.position, view.target, view.up, view.vfov, view.aspect, 0.0, 1.0)
    L0, _ = o.trace_samples(pin, st, view.width, view.height, px, py, s)
    L1, _ = o.trace_samples(zero, st, view.width, view.height, px, py, s)
    np.testing.assert_array_equal(L0, L1)


# ------------------------------------------------------------------------------------------------- GPU

@pytest.mark.gpu
@pytest.mark.parametrize("case", sorted(CASES))
def test_gpu_cameras_match_oracle_per_sample(dev, case):
    kw, moving = CASES[case]
    sc, view = _studio()
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    cam = _pe_camera(view, kw)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    st = settings(max_bounces=view.max_bounces, seed=12345, motion_blur=moving)
    close_n = total = 0
    for sample in (0, 5, 17):
        film = dev.film(view.width, view.height)
        res = render(dev, gs, cam, st, film, sample, 1)
        assert res.completed == 1
        sums, cnt, _, _ = film.download()
        L, _ = o.trace_samples(cam, st, view.width, view.height, xx, yy, np.full(xx.size, sample))
        Lg = sums.reshape(-1, 3)
        close = np.all(np.abs(Lg - L) <= 1e-4 * np.maximum(1.0, np.abs(L)), axis=1)
        close_n += int(close.sum())
        total += close.size
        film.close()
    print(f"camera {case}: per-sample agreement {close_n / total:.5f} of {total}", flush=True)
    assert close_n >= 0.999 * total


@pytest.mark.gpu
def test_gpu_thin_lens_depth_pass(dev):
    """The depth pass is measured from the camera ray's origin, a point on the lens (renderer.cpp:166)."""
    kw, _ = CASES["thin_lens_hexagon"]
    sc, view = _studio()
    gs = dev.upload(sc)
    o = bindings.OracleScene(sc)
    cam = _pe_camera(view, kw)
    st = settings(max_bounces=view.max_bounces, seed=12345)
    yy, xx = np.divmod(np.arange(view.width * view.height), view.width)
    film = dev.film(view.width, view.height)
    pf = PassFilms(dev, view.width, view.height, light=False)
    res = render_targets(dev, gs, cam, st, film, 3, 1, passes=pf)
    assert res.completed == 1
    _, po = o.trace_passes(cam, st, view.width, view.height, xx, yy, np.full(xx.size, 3))
    d = pf.depth.download()[0].reshape(-1, 3)[:, 0]
    hit = po["depth"] > 0
    same = np.abs(d - po["depth"]) <= 1e-5 * np.maximum(1.0, po["depth"])
    print(f"thin lens depth pass: {same.mean():.5f} of {same.size} equal, {hit.mean():.3f} hit", flush=True)
    assert same.mean() >= 0.999


@pytest.mark.gpu
def test_gpu_thin_lens_motion_blur_through_hip_device(dev):
    if bindings.hip_device_harness() is None:
        pytest.skip("oracle/_ref/libphoton_hipdevice.so not built here (reference tree absent)")
    kw, moving = CASES["thin_lens_moving"]
    sc, view = _studio()
    spp = 8
    with bindings.reference_camera(harness=True, **kw):
        rc, mean, counts, _, err = bindings.HipDeviceSession(sc).render(view, spp, max_batch=4)
    assert rc == 0, err
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    render(dev, gs, _pe_camera(view, kw),
           settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4, motion_blur=moving), film, 0, spp)
    s, c, _, _ = film.download()
    assert (counts == spp).all()
    np.testing.assert_allclose(mean, s / c[..., None], rtol=2e-5, atol=1e-6)
