"""The display path: Film::resolve -> applyImager -> toneMap -> quantizeUnorm8 with TPDF dither
(core/image/film.cpp:87-147, core/image/imager.cpp, core/image/tone_mapping.cpp, core/color/transfer.cpp,
the 8-bit loop of saveImagePNG, core/image/image_io.cpp:61-80).

CPU: the oracle's restatement (oracle/oracle_display.cpp) equals the reference's own functions (IEEE build)
float for float and byte for byte: all seven tone-map operators, exposure, and each imager stage alone and all
together (chromatic aberration, bloom, glare, white balance with tint, contrast, saturation, vignette), and the
denoiser of a build without OIDN (denoiseImage = softBlurDenoise, engine/denoiser.cpp:17-43, before the imager as
main.cpp:277-289 orders them), on a film with empty, dim and very bright pixels; closed forms for the operators
that have one, and for the denoiser: a flat image stays flat, an isolated bright pixel is pulled down.
GPU: pe_film_display on a rendered film against the oracle on the same sums (float tolerance: the device's pow,
log2 and exp differ from the host's in the last bits; 8-bit values equal or one apart).
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.device import camera_perspective, display_settings, render, settings

OPS = [capi.PE_TONEMAP_REINHARD, capi.PE_TONEMAP_REINHARD_EXTENDED, capi.PE_TONEMAP_ACES, capi.PE_TONEMAP_FILMIC,
       capi.PE_TONEMAP_AGX, capi.PE_TONEMAP_PBR_NEUTRAL, capi.PE_TONEMAP_LINEAR]

IMAGERS = {
    "identity": {},
    "denoise": {"denoise": 1},
    "denoise_all": {"denoise": 1, "bloom_intensity": 0.5, "glare_intensity": 0.3, "white_balance_k": 9000.0},
    "chromatic": {"chromatic_aberration": 3.0},
    "bloom": {"bloom_intensity": 0.6, "bloom_radius": 0.08, "bloom_threshold": 0.8},
    "glare": {"glare_intensity": 0.4, "glare_blades": 6, "glare_angle_deg": 20.0, "glare_length": 0.15},
    "colour": {"white_balance_k": 3200.0, "tint": 0.3, "contrast": 1.3, "saturation": 0.7, "vignette": 0.5},
    "all": {"chromatic_aberration": 2.0, "bloom_intensity": 0.5, "glare_intensity": 0.3, "white_balance_k": 9000.0,
            "tint": -0.2, "contrast": 0.8, "saturation": 1.4, "vignette": 0.3},
}


def _film(w=61, h=37, seed=0):
    rng = np.random.default_rng(seed)
    cnt = rng.integers(0, 6, (h, w)).astype(np.uint32)
    mean = rng.exponential(0.6, (h, w, 3)).astype(np.float32)
    mean[rng.random((h, w)) < 0.03] *= 200.0  # highlights above every threshold
    return (mean * cnt[..., None]).astype(np.float32), cnt


@pytest.mark.parametrize("op", OPS)
def test_oracle_tonemap_equals_reference(op):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    s, c = _film()
    for ev in (0.0, -1.5, 2.0):
        for dither in (1, 0):
            d = display_settings(tonemap=op, exposure_ev=ev, dither=dither)
            a, a8 = bindings.display("oracle", s, c, d)
            b, b8 = bindings.display("ref_strict", s, c, d)
            np.testing.assert_array_equal(a, b)
            np.testing.assert_array_equal(a8, b8)


@pytest.mark.parametrize("name", list(IMAGERS))
def test_oracle_imager_equals_reference(name):
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    s, c = _film(seed=1)
    d = display_settings(tonemap=capi.PE_TONEMAP_AGX, exposure_ev=0.5, **IMAGERS[name])
    a, a8 = bindings.display("oracle", s, c, d)
    b, b8 = bindings.display("ref_strict", s, c, d)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a8, b8)
    if name != "identity":  # the stage changes the picture
        base, _ = bindings.display("oracle", s, c, display_settings(tonemap=capi.PE_TONEMAP_AGX, exposure_ev=0.5))
        assert np.abs(a - base).max() > 1e-3


def test_oracle_wide_bloom_equals_reference():
    """A bloom wide enough that its octaves blur downsampled levels (sigma above 6 px, imager.cpp:98-103)."""
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    s, c = _film(w=160, h=90, seed=2)
    d = display_settings(tonemap=capi.PE_TONEMAP_FILMIC, bloom_intensity=0.7, bloom_radius=0.2, bloom_threshold=0.5)
    a, a8 = bindings.display("oracle", s, c, d)
    b, b8 = bindings.display("ref_strict", s, c, d)
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(a8, b8)


def test_tonemap_closed_forms():
    # One pixel of one sample per value: Reinhard x / (1 + x), Linear clamp, sRGB encode (transfer.cpp:18-27),
    # an empty pixel is black, and the 8-bit value without dither is floor(255 d + 0.5).
    vals = np.array([0.0, 0.002, 0.18, 1.0, 3.0], np.float32)
    s = np.repeat(vals[None, :, None], 3, 2)
    c = np.ones((1, vals.size), np.uint32)
    enc = lambda x: np.where(x <= 0.0031308, 12.92 * x, 1.055 * np.power(x, 1 / 2.4) - 0.055)  # noqa: E731
    r, r8 = bindings.display("oracle", s, c, display_settings(tonemap=capi.PE_TONEMAP_REINHARD, dither=0))
    np.testing.assert_allclose(r[0, :, 0], enc(vals / (1 + vals)), rtol=2e-6, atol=1e-7)
    np.testing.assert_array_equal(r8[0, :, 0], np.floor(r[0, :, 0].astype(np.float64) * 255 + 0.5).astype(np.uint8))
    l, _ = bindings.display("oracle", s, c, display_settings(tonemap=capi.PE_TONEMAP_LINEAR))
    np.testing.assert_allclose(l[0, :, 0], enc(np.minimum(vals, 1.0)), rtol=2e-6, atol=1e-7)
    e, e8 = bindings.display("oracle", np.zeros((1, 1, 3), np.float32), np.zeros((1, 1), np.uint32),
                             display_settings(dither=0))
    assert (e == 0).all() and (e8 == 0).all()


# ------------------------------------------------------------------------------------------------- GPU

def test_denoise_closed_forms():
    """softBlurDenoise: a flat image is a fixed point (every weight equal); a lone grey pixel c of luminance L on
    black has weight k = 1 / (1 + L^2) against its eight black neighbours' weights of 1, so it becomes
    c k / (8 + k)."""
    d = display_settings(tonemap=capi.PE_TONEMAP_LINEAR, denoise=1, dither=0)
    flat = np.full((5, 7, 3), 0.3, np.float32)
    cnt = np.ones((5, 7), np.uint32)
    a, _ = bindings.display("oracle", flat, cnt, d)
    b, _ = bindings.display("oracle", flat, cnt, display_settings(tonemap=capi.PE_TONEMAP_LINEAR, dither=0))
    np.testing.assert_array_equal(a, b)
    spike = np.zeros((5, 5, 3), np.float32)
    spike[2, 2] = 0.9                                   # luminance 0.9 (grey), within the linear operator's range
    raw = np.zeros((5, 5, 3), np.float32)
    k = 1.0 / (1.0 + 0.9 * 0.9)
    raw[2, 2] = 0.9 * k / (8.0 + k)                     # the centre: its own term over the nine weights
    assert raw[2, 2, 0] < 0.9 / 5                       # pulled far down
    out, _ = bindings.display("oracle", spike, np.ones((5, 5), np.uint32),
                              display_settings(tonemap=capi.PE_TONEMAP_LINEAR, denoise=1, dither=0))
    lin, _ = bindings.display("oracle", raw, np.ones((5, 5), np.uint32),
                              display_settings(tonemap=capi.PE_TONEMAP_LINEAR, dither=0))
    np.testing.assert_allclose(out[2, 2], lin[2, 2], rtol=1e-6)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["identity", "bloom", "glare", "all", "denoise", "denoise_all"])
def test_gpu_display_matches_oracle(dev, name):
    sc, view = configs.product_sphere(width=96, height=54, lon=96, lat=64, env_width=256, env_height=128)
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    render(dev, gs, cam, settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4), film, 0, 4)
    s, c, _, _ = film.download()
    for op in (capi.PE_TONEMAP_ACES, capi.PE_TONEMAP_AGX, capi.PE_TONEMAP_FILMIC, capi.PE_TONEMAP_PBR_NEUTRAL):
        d = display_settings(tonemap=op, exposure_ev=0.3, **IMAGERS[name])
        g, g8 = film.display(d)
        o, o8 = bindings.display("oracle", s, c, d)
        np.testing.assert_allclose(g, o, atol=2e-5)
        diff = np.abs(g8.astype(np.int32) - o8.astype(np.int32))
        assert diff.max() <= 1 and (diff == 0).mean() >= 0.995, (diff.max(), (diff == 0).mean())
