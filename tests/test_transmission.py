"""Transmission, volumes and subsurface on the integrator path: Dielectric (smooth, rough, thin-walled,
dispersive; materials/dielectric.cpp), wavelength sampling and Cauchy dispersion (core/color/spectral.cpp),
the medium stack with Beer-Lambert absorption and scattering media (free flights with spectral MIS,
sampleFreeFlight :97-129), pass-through and shadow transmittance, and the subsurface random walk
(integrators/path_tracer.cpp:63-292, :457-479, :602-607, :744-786), plus Disney specular/diffuse transmission
(materials/disney.cpp:183-220).

CPU: the oracle against the reference's own strict build, sample for sample, on a studio that holds each of
these (and separately per feature, so a regression names its feature); the dielectric BSDF bit for bit;
closed forms of the spectral pieces (Cauchy through n_d and the Abbe number, unit-mean wavelength weights).
GPU: the kernels against the oracle (tolerances as in test_integrator.py).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.configs import View, _quad, box_mesh, icosphere, prism_mesh
from photon_hip.device import camera_perspective, render, settings
from photon_hip.scene import AreaLight, Mesh, Scene, dielectric, disney, lambert


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


VIEW = View((0.2, 1.3, 3.4), (0.0, 0.45, 0.0), (0.0, 1.0, 0.0), 45.0, 64, 48, 16, 16)

FEATURES = {
    "dispersive_prism": lambda sc: sc.add_shape(prism_mesh((-1.1, 0.002, 0.0), 0.45, 0.9, 2),
                                                sc.add_material(dielectric(1.5, abbe=30.0))),
    # sigma_a = (0.1, 0.45, 1.2) per cm, inside the lights + media workload's U[0.1, 2] / cm (scene units: metres),
    # as setAbsorption(exp(-sigma_a * 1 cm), 1 cm).
    "absorbing_liquid": lambda sc: sc.add_shape(box_mesh((-0.45, 0.002, -0.2), (0.05, 0.6, 0.3)),
                                                sc.add_material(dielectric(1.33, absorption_color=np.exp(
                                                    -np.array([10.0, 45.0, 120.0]) * 0.01), absorption_distance=0.01))),
    "subsurface_block": lambda sc: sc.add_shape(box_mesh((0.3, 0.002, -0.25), (0.8, 0.5, 0.25)),
                                                sc.add_material(disney((0.85, 0.55, 0.4), 0.0, 0.4, 0.5, subsurface=1.0,
                                                                       subsurface_radius=(2, 3, 4),
                                                                       subsurface_scale=1e-3))),
    "thin_pane": lambda sc: sc.add_shape(_quad((0.9, 0.0, 0.5), (1.7, 0.0, 0.2), (1.7, 0.9, 0.2), (0.9, 0.9, 0.5)),
                                         sc.add_material(dielectric(1.5, tint=(0.9, 0.95, 1.0), thin_walled=True))),
    "rough_glass": lambda sc: sc.add_shape(icosphere(2, 0.25, (1.15, 0.25, -0.55)),
                                           sc.add_material(dielectric(1.45, tint=(1.0, 0.97, 0.9), roughness=0.25))),
    # A milky glass block (setScattering: density 12 per metre, so a mean free path of 8 cm in a 50 cm block,
    # forward scattering) and a scattering, absorbing liquid: free flights with spectral MIS, Henyey-Greenstein
    # directions, roulette past 8 events.
    "scattering_block": lambda sc: sc.add_shape(box_mesh((0.3, 0.002, -0.25), (0.8, 0.5, 0.25)),
                                                sc.add_material(dielectric(1.45, roughness=0.15, scatter_density=12.0,
                                                                           scatter_color=(0.95, 0.8, 0.6),
                                                                           scatter_anisotropy=0.4))),
    "scattering_liquid": lambda sc: sc.add_shape(box_mesh((-0.45, 0.002, -0.2), (0.05, 0.6, 0.3)),
                                                 sc.add_material(dielectric(1.33, absorption_color=(0.8, 0.9, 0.6),
                                                                            absorption_distance=0.2,
                                                                            scatter_density=30.0,
                                                                            scatter_color=(0.9, 0.9, 0.95),
                                                                            scatter_anisotropy=-0.3))),
    "disney_transmission": lambda sc: sc.add_shape(icosphere(2, 0.22, (-0.2, 0.85, 0.6)),
                                                   sc.add_material(disney((0.5, 0.8, 0.4), 0.0, 0.35, 0.5,
                                                                          specular_transmission=0.6,
                                                                          diffuse_transmission=0.3, ior=1.4))),
}


def feature_scene(name: str) -> Scene:
    sc = Scene()
    floor = sc.add_material(lambert((0.6, 0.6, 0.6)))
    up = np.array([0.0, 1.0, 0.0], np.float32)
    sc.add_shape(Mesh(np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)), floor)
    FEATURES[name](sc)
    sc.add_light(AreaLight((-0.6, 2.5, -0.2), (1.2, 0.0, 0.0), (0.0, 0.0, 0.8), (12.0, 12.0, 12.0)))
    return sc


def _per_sample(sc, view, n=3000, seed=7):
    rng = np.random.default_rng(seed)
    px, py, s = rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, 256, n)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, rays = bindings.OracleScene(sc).trace_samples(_cam(view), st, view.width, view.height, px, py, s)
    return px, py, s, Lo, rays


@pytest.mark.parametrize("name", list(FEATURES))
def test_oracle_equals_reference_per_sample(name):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc = feature_scene(name)
    px, py, s, Lo, _ = _per_sample(sc, VIEW)
    Lr = bindings.RefScene(sc, strict=True).trace_samples(VIEW, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    assert Lo.max() > 0


def test_oracle_equals_reference_per_sample_glass_studio():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = configs.glass_studio(led_quads=100)    # 200 emissive triangles: the light BVH as well
    px, py, s, Lo, _ = _per_sample(sc, view, n=4000)
    Lr = bindings.RefScene(sc, strict=True).trace_samples(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)


@pytest.mark.parametrize("kind", ["smooth", "rough", "thin", "dispersive"])
def test_dielectric_bsdf_equals_reference(kind):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    m = {"smooth": dielectric(1.5), "rough": dielectric(1.45, tint=(0.9, 1.0, 0.8), roughness=0.25),
         "thin": dielectric(1.5, tint=(0.9, 0.95, 1.0), thin_walled=True), "dispersive": dielectric(1.5, abbe=30.0)}[kind]
    rng = np.random.default_rng(11)
    n = 3000
    d = rng.normal(size=(4, n, 3))
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    normal, tv, wo, wi = (x.astype(np.float32) for x in d)
    tangent = np.cross(normal, tv).astype(np.float32)
    front = (rng.random(n) < 0.7).astype(np.int32)
    u3 = rng.random((n, 3)).astype(np.float32)
    so = bindings.bsdf_sample("oracle", m, normal, tangent, front, wo, u3)
    sr = bindings.bsdf_sample("ref_strict", m, normal, tangent, front, wo, u3)
    np.testing.assert_array_equal(so, sr)
    fo, po = bindings.bsdf_eval("oracle", m, normal, tangent, front, wo, wi)
    fr, pr = bindings.bsdf_eval("ref_strict", m, normal, tangent, front, wo, wi)
    np.testing.assert_array_equal(fo, fr)
    np.testing.assert_array_equal(po, pr)


def test_smooth_glass_energy_split_closed_form():
    """Smooth glass at normal incidence: reflect with probability F = ((n-1)/(n+1))^2 and weight 1, refract with
    weight eta^2 = 1/n^2 (radiance scaling at the boundary, Veach 1997 5.2)."""
    n_ior = 1.5
    m = dielectric(n_ior)
    N = 4000
    normal = np.tile(np.float32([[0, 0, 1]]), (N, 1))
    tangent = np.tile(np.float32([[1, 0, 0]]), (N, 1))
    front = np.ones(N, np.int32)
    wo = normal.copy()
    u3 = np.zeros((N, 3), np.float32)
    u3[:, 0] = (np.arange(N) + 0.5) / N
    s = bindings.bsdf_sample("oracle", m, normal, tangent, front, wo, u3)
    F = ((n_ior - 1) / (n_ior + 1)) ** 2
    refl = s[:, 3] > 0
    assert abs(refl.mean() - F) < 2.0 / N
    w = s[:, 4] * np.abs(s[:, 3]) / s[:, 7]          # f |cos| / pdf
    np.testing.assert_allclose(w[refl], 1.0, rtol=1e-5)
    np.testing.assert_allclose(w[~refl], 1.0 / n_ior ** 2, rtol=1e-5)


def test_cauchy_dispersion_closed_form():
    """cauchyIor passes through n_d at the helium d line and its F - C spread is (n_d - 1) / V (the Abbe number's
    definition), in the Cauchy form n = A + B / lambda^2."""
    nd, abbe = 1.5, 30.0
    _, _, n = bindings.spectral(np.zeros(1), nd, abbe, [587.56, 486.13, 656.27, 400.0, 700.0])
    assert abs(n[0] - nd) < 2e-6
    assert abs((n[1] - n[2]) - (nd - 1) / abbe) < 2e-6
    assert n[3] > n[0] > n[4]       # normal dispersion: blue bends more
    _, _, same = bindings.spectral(np.zeros(1), nd, 0.0, [400.0, 700.0])
    np.testing.assert_array_equal(same, [nd, nd])


def test_wavelength_weights_have_unit_mean():
    """sampleWavelength importance-samples lambda and returns w(lambda) / (range * pdf), where w is the
    normalised RGB response: each channel's weight averages to one over u (a white path stays white on average),
    and the sampled wavelengths cover the [380, 730] nm range in order of u."""
    N = 200000
    u = ((np.arange(N) + 0.5) / N).astype(np.float32)
    lam, w, _ = bindings.spectral(u)
    assert lam.min() >= 380.0 and lam.max() <= 730.0
    assert np.all(np.diff(lam) >= 0)
    np.testing.assert_allclose(w.mean(0), 1.0, rtol=2e-3)
    assert (w >= 0).all()


# ------------------------------------------------------------------------------------------------- GPU

def _gpu_image(dev, sc, view, spp, first=0, spl=1):
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl)
    res = render(dev, gs, _cam(view), st, film, first, spp)
    s, c, _, rej = film.download()
    return s, c, rej, res


def _gpu_per_sample(dev, sc, view, sample=3):
    s, c, rej, res = _gpu_image(dev, sc, view, 1, first=sample)
    assert res.completed == 1 and (c == 1).all() and rej == 0
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    Lo, _ = bindings.OracleScene(sc).trace_samples(_cam(view), st, view.width, view.height, xx.ravel(), yy.ravel(),
                                                   np.full(xx.size, sample))
    Lg = s.reshape(-1, 3)
    # Same Sobol samples, same paths: a sample differs only where a device transcendental (exp, log, sin) one
    # ulp away from the host's flips a later decision.
    close = np.all(np.abs(Lg - Lo) <= 1e-4 * np.maximum(1.0, np.abs(Lo)), axis=1)
    print(f"per-sample agreement with the oracle: {close.mean():.5f} of {close.size} samples", flush=True)
    return close.mean(), Lg, Lo


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(FEATURES))
def test_gpu_radiance_matches_oracle_per_sample(dev, name):
    frac, Lg, Lo = _gpu_per_sample(dev, feature_scene(name), VIEW)
    assert frac >= 0.998, frac
    assert Lo.max() > 0


@pytest.mark.gpu
def test_gpu_glass_studio_matches_oracle_per_sample(dev):
    sc, view = configs.glass_studio(led_quads=100)
    frac, _, _ = _gpu_per_sample(dev, sc, view)
    assert frac >= 0.998, frac


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["glass_studio", "subsurface_block", "thin_pane", "absorbing_liquid",
                                  "scattering_block", "scattering_liquid"])
def test_gpu_image_matches_oracle_relmse(dev, name):
    if name == "glass_studio":
        sc, view = configs.glass_studio(led_quads=100)
    else:
        sc, view = feature_scene(name), VIEW
    # About 0.1% of the samples of these scenes take another path on the device (a caustic through glass after
    # a one-ulp difference in exp/log/sin); the reference's own -ffast-math build departs from its IEEE build
    # on the same samples at the same rate and by the same fireflies (relMSE 7.7e-3 at 32 spp, 1.6e-3 at 128 on
    # the glass studio). The image check therefore runs where those flips have averaged out.
    spp = 512
    s, c, rej, res = _gpu_image(dev, sc, view, spp, spl=64)
    assert (c == spp).all() and rej == 0
    gpu = s / c[..., None]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    (so, co, _), _, _ = bindings.OracleScene(sc).render(_cam(view), st, view.width, view.height, 0, spp, threads=8)
    ora = so / co[..., None]
    print(f"{name}: relMSE vs oracle at {spp} spp = {relmse(gpu, ora):.3e}", flush=True)
    assert relmse(gpu, ora) <= 1e-3, relmse(gpu, ora)
