"""Render passes: the eight light passes, light groups and the data passes (RenderPassFilms, engine/renderer.h:
59-100; PathTracer::Li's crediting, integrators/path_tracer.cpp:372-395, :482-698, :790-812; tracePixel /
addPasses, engine/renderer.cpp:162-171, :207-219).

CPU: the oracle's restated crediting equals the reference's own strict build sample for sample (every pass,
every group, the data passes) on the scene of the reference's own pass test (tests/test_render_passes.cpp:33-53)
with light groups; Scene::buildAccelerator's group tables equal the scene description's; the passes sum to the
beauty value of every sample and each lobe lands where the reference's test expects it.
GPU: the wavefront kernels' pass films against the oracle per sample (one sample per pixel, so a film pixel is
one sample), the passes summing to the beauty film, and light groups summing to it.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import bindings
from photon_hip import capi, configs
from photon_hip.device import camera_perspective, settings
from photon_hip.scene import AreaLight, Environment, Scene, lambert


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _need_ref():
    if bindings.reference(strict=True) is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")


def _samples(view, n, seed, max_sample=64):
    rng = np.random.default_rng(seed)
    return (rng.integers(0, view.width, n), rng.integers(0, view.height, n), rng.integers(0, max_sample, n))


def _group_scene(env_group="", emitter_rows=2):
    """Area lights in named groups, an emissive mesh (no group) and an environment."""
    sc = Scene()
    m = sc.add_material(lambert((0.5, 0.5, 0.5)))
    e = sc.add_material(lambert((0.0, 0.0, 0.0), (3.0, 3.0, 3.0)))
    sc.add_shape(configs.icosphere(1, 0.5, (0.0, 0.0, 0.0)), m)
    sc.add_shape(configs.grid_quad((-1, 2, -1), (2, 0, 0), (0, 0, 2), emitter_rows, 1), e)
    sc.add_light(AreaLight((1, 1, 1), (0.3, 0, 0), (0, 0, 0.3), (5, 5, 5), light_group="Key"))
    sc.add_light(AreaLight((-1, 1, 1), (0.3, 0, 0), (0, 0, 0.3), (5, 5, 5)))
    sc.add_light(AreaLight((0, 1, -1), (0.3, 0, 0), (0, 0, 0.3), (5, 5, 5), light_group="Rim"))
    sc.add_light(AreaLight((0, -1, 1), (0.3, 0, 0), (0, 0, 0.3), (5, 5, 5), light_group="Key"))
    sc.set_environment(Environment(None, 0.0, 0.5, light_group=env_group))
    return sc


@pytest.mark.parametrize("env_group", ["", "Sky", "Key"])
def test_light_group_tables_equal_reference(env_group):
    _need_ref()
    sc = _group_scene(env_group)
    names, per, env = sc.light_groups()
    rn, rp, re = bindings.RefScene(sc, strict=True).light_groups()
    assert (names, per, env) == (rn, rp, re)
    assert names[:2] == ["Default", "Key"]


def test_light_group_tables_without_explicit_groups_and_past_the_cap():
    plain = Scene()
    plain.add_material(lambert((0.5, 0.5, 0.5)))
    plain.add_light(AreaLight((0, 1, 0), (1, 0, 0), (0, 0, 1), (1, 1, 1)))
    plain.set_environment(Environment(None))
    assert plain.light_groups() == ([], [-1], -1)
    # 20 named groups: the first 16 names get indices, later ones fall into group 0 (scene.cpp:160).
    many = Scene()
    many.add_material(lambert((0.5, 0.5, 0.5)))
    for k in range(20):
        many.add_light(AreaLight((k, 1, 0), (1, 0, 0), (0, 0, 1), (1, 1, 1), light_group=f"G{k}"))
    names, per, env = many.light_groups()
    assert len(names) == capi.PE_MAX_LIGHT_GROUPS and per[:16] == list(range(16)) and per[16:] == [0] * 4
    assert env == -1
    if bindings.reference(strict=True) is not None:
        assert bindings.RefScene(many, strict=True).light_groups() == (names, per, env)


@pytest.mark.parametrize("groups", [True, False])
def test_oracle_passes_equal_reference_per_sample(groups):
    _need_ref()
    sc, view = configs.pass_studio(groups=groups)
    o = bindings.OracleScene(sc)
    ref = bindings.RefScene(sc, strict=True)
    px, py, s = _samples(view, 2500, 5)
    st = settings(max_bounces=view.max_bounces, filter_type=capi.PE_FILTER_BOX, filter_width=1.0, seed=12345)
    Lo, po = o.trace_passes(_cam(view), st, view.width, view.height, px, py, s)
    Lr, pr = ref.trace_passes(view, px, py, s, filter_type=0, filter_width=1.0, seed=12345)
    np.testing.assert_array_equal(Lo, Lr)
    for key in po:
        np.testing.assert_array_equal(po[key], pr[key], err_msg=key)
    # Every pass is exercised by these samples (the reference test's expectations, test_render_passes.cpp:88-93).
    tot = po["light"].sum(axis=(0, 2))
    for k in (capi.PE_PASS_DIFFUSE_DIRECT, capi.PE_PASS_DIFFUSE_INDIRECT, capi.PE_PASS_GLOSSY_DIRECT,
              capi.PE_PASS_GLOSSY_INDIRECT, capi.PE_PASS_TRANSMISSION_INDIRECT, capi.PE_PASS_EMISSION,
              capi.PE_PASS_BACKGROUND):
        assert tot[k] > 0.0, capi.LIGHT_PASS_NAMES[k]
    if groups:
        assert (po["groups"][:, :3].sum(axis=(0, 2)) > 0).all() and not po["groups"][:, 3:].any()
        assert (po["object_id"].max(axis=1) > 0).mean() > 0.5
    else:
        assert not po["groups"].any() and not po["object_id"].any()


def test_oracle_passes_sum_to_beauty_and_land_in_place():
    """Light passes and light groups each sum to the beauty value (test_render_passes.cpp:71-86: relative
    1e-4); the metal sphere's centre is glossy, the floor diffuse, the glass transmission (:95-103)."""
    sc, view = configs.pass_studio(width=64, height=48)
    o = bindings.OracleScene(sc)
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    for sample in (0, 7):
        L, p = o.trace_passes(_cam(view), st, view.width, view.height, xx.ravel(), yy.ravel(),
                              np.full(xx.size, sample))
        for part in (p["light"].sum(axis=1), p["groups"].sum(axis=1)):
            assert (np.abs(part - L) <= 1e-4 * np.maximum(1.0, np.abs(L))).all()
    light = p["light"].reshape(view.height, view.width, 8, 3)
    mx, my = int(0.37 * view.width), int(0.52 * view.height)
    lum = light[my, mx] @ np.array([0.212671, 0.715160, 0.072169])
    assert lum[2] + lum[3] > 10 * (lum[0] + lum[1]) + 1e-4
    depth = p["depth"].reshape(view.height, view.width)
    assert abs(depth[my, mx] - 4.0) < 0.6
    floor_px = light[view.height - 2, 4]
    assert floor_px[0].sum() > 0 and floor_px[2:6].sum() == 0


# ------------------------------------------------------------------------------------------------- GPU

def _gpu_passes(dev, sc, view, first, count, spl=1, groups=0):
    from photon_hip.device import PassFilms, render_targets

    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    pf = PassFilms(dev, view.width, view.height, groups=groups)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spl)
    res = render_targets(dev, gs, _cam(view), st, film, first, count, passes=pf)
    assert res.completed == 1
    return film, pf, res


def _agree(a, b, rtol=1e-4):
    """Per-sample agreement of (n, ...) arrays: every component within rtol of max(1, |b|)."""
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    return np.all(np.abs(a - b) <= rtol * np.maximum(1.0, np.abs(b)), axis=1)


@pytest.mark.gpu
@pytest.mark.parametrize("groups", [True, False])
def test_gpu_passes_match_oracle_per_sample(dev, groups):
    """One sample per pixel, so every film pixel is one sample: each pass, group and data pass of the GPU against
    the oracle's for the same (pixel, sample) path. A sample agrees when all its passes do (the beauty bar: paths
    whose roulette or lobe draw sits within an ulp of its threshold may diverge)."""
    sc, view = configs.pass_studio(groups=groups)
    ng = len(sc.light_groups()[0])
    sample = 3
    film, pf, _ = _gpu_passes(dev, sc, view, sample, 1, groups=ng)
    m = pf.means()
    o = bindings.OracleScene(sc)
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    Lo, po = o.trace_passes(_cam(view), st, view.width, view.height, xx.ravel(), yy.ravel(), np.full(xx.size, sample))
    n = xx.size
    light = m["light"].reshape(8, n, 3).transpose(1, 0, 2)
    checks = {"beauty": _agree(film.mean().reshape(n, 3), Lo), "light": _agree(light, po["light"]),
              "depth": _agree(m["depth"].reshape(n, 3), np.repeat(po["depth"][:, None], 3, 1)),
              "position": _agree(m["position"].reshape(n, 3), po["position"]),
              "object_id": _agree(m["object_id"].reshape(n, 3), po["object_id"]),
              "material_id": _agree(m["material_id"].reshape(n, 3), po["material_id"])}
    if ng:
        checks["groups"] = _agree(m["groups"].reshape(ng, n, 3).transpose(1, 0, 2), po["groups"][:, :ng])
    every = np.logical_and.reduce(list(checks.values()))
    print("per-sample agreement with the oracle: " + ", ".join(f"{k} {v.mean():.4f}" for k, v in checks.items()) +
          f"; all {every.mean():.4f} of {n}", flush=True)
    assert every.mean() >= 0.995, {k: float(v.mean()) for k, v in checks.items()}
    tot = light.sum(axis=(0, 2))
    for k in (capi.PE_PASS_DIFFUSE_DIRECT, capi.PE_PASS_GLOSSY_DIRECT, capi.PE_PASS_TRANSMISSION_INDIRECT,
              capi.PE_PASS_EMISSION, capi.PE_PASS_BACKGROUND):
        assert tot[k] > 0.0, capi.LIGHT_PASS_NAMES[k]


@pytest.mark.gpu
def test_gpu_passes_sum_to_beauty_and_batch_equals_single(dev):
    """The light passes and the light groups each sum to the beauty film (test_render_passes.cpp:71-86, relative
    1e-4), and a batch of 8 samples in one window equals eight single-sample calls bit for bit."""
    sc, view = configs.pass_studio()
    ng = len(sc.light_groups()[0])
    film, pf, _ = _gpu_passes(dev, sc, view, 0, 8, spl=8, groups=ng)
    m = pf.means()
    beauty = film.mean()
    for part in (m["light"].sum(axis=0), m["groups"].sum(axis=0)):
        worst = float(np.max(np.abs(part - beauty) / np.maximum(1.0, beauty)))
        assert worst < 1e-4, worst
    from photon_hip.device import PassFilms, render_targets

    gs = dev.upload(sc)
    film1 = dev.film(view.width, view.height)
    pf1 = PassFilms(dev, view.width, view.height, groups=ng)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=1)
    for k in range(8):
        render_targets(dev, gs, _cam(view), st, film1, k, 1, passes=pf1)
    for a, b in zip(pf.light + pf.groups + [pf.depth, pf.position, pf.object_id, pf.material_id],
                    pf1.light + pf1.groups + [pf1.depth, pf1.position, pf1.object_id, pf1.material_id]):
        sa, ca, la, ra = a.download()
        sb, cb, lb, rb = b.download()
        np.testing.assert_array_equal(sa, sb)
        np.testing.assert_array_equal(ca, cb)
        assert ra == rb


@pytest.mark.gpu
def test_gpu_pass_film_errors(dev):
    from photon_hip.device import PassFilms, render_targets

    sc, view = configs.pass_studio(width=16, height=12)
    gs = dev.upload(sc)
    film = dev.film(16, 12)
    bad = PassFilms(dev, 8, 8, light=False, data=False, groups=1)
    with pytest.raises(RuntimeError):
        render_targets(dev, gs, _cam(view), settings(max_bounces=4, seed=1), film, 0, 1, passes=bad)


def _films_of(pf, ng):
    """The device pass films in pe_pass_targets order as (28, H, W, 3) means (absent groups zero)."""
    m = pf.means()
    h, w = m["depth"].shape[:2]
    out = np.zeros((28, h, w, 3), np.float32)
    out[:8] = m["light"]
    if ng:
        out[8:8 + ng] = m["groups"]
    for k, name in zip(range(24, 28), ("depth", "position", "object_id", "material_id")):
        out[k] = m[name]
    return out


@pytest.mark.gpu
def test_gpu_session_passes_through_hip_device(dev):
    """The reference's RenderSession with SessionConfig::passes on the HipDevice adapter (features().lightPasses):
    its RenderPassFilms, light groups named as the scene's, resolve to what a direct pe_render_targets of the same
    samples gives (up to Film::addBatch's association across batches), and to what the reference's own
    Renderer::renderSamples with RenderPassFilms gives for the same samples on the CPU (the same paths; the
    material-id colours differ by design: the reference hashes material addresses)."""
    if bindings.hip_device_harness() is None or bindings.reference() is None:
        pytest.skip("oracle/_ref not built here (reference tree absent)")
    sc, view = configs.pass_studio()
    names, _, _ = sc.light_groups()
    spp = 16
    rc, beauty, films, counts, groups, err, _ = bindings.HipDeviceSession(sc).render_passes(view, spp, 4)
    assert rc == 0, err
    assert (counts == spp).all() and groups == len(names)
    film, pf, _ = _gpu_passes(dev, sc, view, 0, spp, spl=spp, groups=len(names))
    direct = _films_of(pf, len(names))
    np.testing.assert_allclose(beauty, film.mean(), rtol=1e-5, atol=1e-6)
    np.testing.assert_allclose(films, direct, rtol=1e-5, atol=1e-6)
    rb, rfilms, rgroups, _ = bindings.RefScene(sc).render_passes(view, 0, spp, threads=8)
    assert rgroups == len(names)
    for k in list(range(8 + rgroups)) + [24, 25, 26]:
        err_k = float(np.mean((films[k] - rfilms[k]) ** 2 / (rfilms[k] ** 2 + 0.01)))
        assert err_k <= 1e-3, (k, err_k)
    assert float(np.mean((beauty - rb) ** 2 / (rb ** 2 + 0.01))) <= 1e-3


# ------------------------------------------------------------------------------------------------- cryptomatte

def _crypto_counts(ids_per_sample, width, height, px, py):
    """CryptoFilm::addSample over samples in order: per pixel {id: count} and the total (cryptomatte.cpp:121-133)."""
    buckets = {}
    for x, y, i in zip(px, py, ids_per_sample):
        b = buckets.setdefault((int(x), int(y)), [0, {}])
        b[0] += 1
        if i != 0.0:
            b[1][float(i)] = b[1].get(float(i), 0) + 1
    return buckets


def _pairs(bucket):
    """A bucket's (id, coverage) pairs as pixelRanks computes coverage (count * (1 / total), float32)."""
    total, entries = bucket
    inv = np.float32(1.0) / np.float32(total)
    return sorted((i, float(np.float32(c) * inv)) for i, c in entries.items())


def _rank_pairs(ranks, x, y):
    """(id, coverage) pairs of a pixel from resolveRanks(3) layers (3, H, W, 4); zero coverage = no entry."""
    out = []
    for layer in ranks:
        r, g, b, a = layer[y, x]
        out += [(float(r), float(g))] if g > 0 else []
        out += [(float(b), float(a))] if a > 0 else []
    return sorted(out)


def test_crypto_hash_equals_reference_and_known_answers():
    """cryptoHashName (MurmurHash3_x86_32, seed 0) against the published vectors and the reference's own build;
    cryptoHashToFloat keeps every id finite and normal (test_v04_production.cpp:146-153)."""
    from photon_hip.scene import crypto_hash, crypto_hash_to_float

    assert crypto_hash("") == 0 and crypto_hash("hello") == 0x248BFA47 and crypto_hash("test") == 0xBA6BD213
    tricky = [0x00000000, 0x00000001, 0x007FFFFF, 0x7F800000, 0x7FC00000, 0x7FFFFFFF, 0x80000000, 0xFF800000,
              0xFFFFFFFF]
    for h in tricky:
        f = np.float32(crypto_hash_to_float(h))
        assert np.isfinite(f) and f != 0 and abs(f) >= np.finfo(np.float32).tiny
    _need_ref()
    names = ["", "a", "ab", "abc", "abcd", "Hero_A", "Hero_B", "object_12", "shape_0", "material_3",
             "çok katmanlı", "x" * 37]
    h, f = bindings.crypto_hash_reference(names)
    assert [crypto_hash(n) for n in names] == h.tolist()
    np.testing.assert_array_equal(np.array([crypto_hash_to_float(int(v)) for v in h], np.float32), f)


def test_crypto_lookup_and_counts_equal_reference_renderer():
    """The scene description's cryptomatte lookup (named, uid-named and anonymous objects and materials) and
    CryptoFilm's counting, applied to the oracle's per-sample first surfaces, give the (id, coverage) ranks the
    reference's Renderer::renderSamples with RenderPassFilms writes for the same samples."""
    _need_ref()
    sc, view = configs.pass_studio(width=24, height=18)
    obj_ids, mat_ids, _, _ = sc.crypto_tables()
    spp = 4
    _, _, _, ranks = bindings.RefScene(sc, strict=True).render_passes(view, 0, spp, threads=4)
    o = bindings.OracleScene(sc)
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    st = settings(max_bounces=view.max_bounces, seed=12345)
    px = np.tile(xx.ravel(), spp)
    py = np.tile(yy.ravel(), spp)
    s = np.repeat(np.arange(spp), xx.size)
    _, p = o.trace_passes(_cam(view), st, view.width, view.height, px, py, s)
    order = np.lexsort((s, px, py))  # per pixel, in sample order
    for k, (key, table) in enumerate((("shape", obj_ids), ("material", mat_ids))):
        idx = p[key][order]
        ids = np.where(idx >= 0, table[np.clip(idx, 0, None)], 0.0)
        buckets = _crypto_counts(ids, view.width, view.height, px[order], py[order])
        for (x, y), b in buckets.items():
            assert _pairs(b) == _rank_pairs(ranks[k], x, y), (key, x, y)
    names = set(sc.crypto_tables()[2])
    assert {"Floor", "Hero_Metal", "object_13", "object_14", "shape_0", "shape_1"} <= names


@pytest.mark.gpu
def test_gpu_cryptomatte_matches_oracle_and_reference(dev):
    """Device CryptoFilms against the oracle's per-sample first surfaces counted as CryptoFilm counts them, and
    against the reference renderer's ranks for the same samples; then through HipDevice (RenderSession with
    passes), whose merged CryptoFilms give the same ranks."""
    from photon_hip.device import PassFilms, render_targets

    sc, view = configs.pass_studio(width=32, height=24)
    tables = sc.crypto_tables()
    spp = 8
    gs = dev.upload(sc)
    film = dev.film(view.width, view.height)
    pf = PassFilms(dev, view.width, view.height, light=False, data=False, crypto_tables=tables)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=spp)
    assert render_targets(dev, gs, _cam(view), st, film, 0, spp, passes=pf).completed == 1
    o = bindings.OracleScene(sc)
    yy, xx = np.mgrid[0:view.height, 0:view.width]
    px, py = np.tile(xx.ravel(), spp), np.tile(yy.ravel(), spp)
    s = np.repeat(np.arange(spp), xx.size)
    _, p = o.trace_passes(_cam(view), settings(max_bounces=view.max_bounces, seed=12345), view.width, view.height,
                          px, py, s)
    order = np.lexsort((s, px, py))
    _, _, _, ranks = bindings.RefScene(sc, strict=True).render_passes(view, 0, spp, threads=8)
    agree = []
    for k, (key, table, cf) in enumerate((("shape", tables[0], pf.crypto_object),
                                          ("material", tables[1], pf.crypto_material))):
        data = cf.download()
        assert data[3] == 0  # no pixel saw more than PE_CRYPTO_SLOTS ids
        assert (data[2] == spp).all()
        idx = p[key][order]
        ids = np.where(idx >= 0, table[np.clip(idx, 0, None)], 0.0)
        buckets = _crypto_counts(ids, view.width, view.height, px[order], py[order])
        for (x, y), b in buckets.items():
            mine = sorted(cf.pixel_ranks(x, y, data))
            agree.append(mine == _pairs(b) and mine == _rank_pairs(ranks[k], x, y))
    print(f"cryptomatte: pixels agreeing with the oracle and the reference renderer {np.mean(agree):.4f}", flush=True)
    assert np.mean(agree) >= 0.995
    if bindings.hip_device_harness() is not None:
        rc, _, _, _, _, err, hranks = bindings.HipDeviceSession(sc).render_passes(view, spp, 4)
        assert rc == 0, err
        same = [_rank_pairs(hranks[k], x, y) == _rank_pairs(ranks[k], x, y)
                for k in range(2) for y in range(view.height) for x in range(view.width)]
        print(f"cryptomatte through HipDevice: pixels agreeing with the reference renderer {np.mean(same):.4f}",
              flush=True)
        assert np.mean(same) >= 0.995
