"""Learned path guiding: the 8-lobe vMF mixture MLP (pe_guide_*) and guided rendering (pe_render_guided).

The network has no reference counterpart (the reference guides with an SD-tree, integrators/path_guiding.cpp),
so its kernels are checked against a plain PyTorch fp32 restatement of the same network: the encoding, the
forward pass with bf16-rounded weights and activations (what the MFMA kernels consume), the mixture, the loss
-sum w log q / sum w and its gradient by autograd, and the Adam update. Guided rendering is checked the way
the reference's own guiding is judged, against plain path tracing: the guided image converges to the plain
image (MIS with the BSDF keeps it unbiased), and on the slit room it has less error at equal sample count.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, settings

LAYERS = [(160, 64), (64, 64), (64, 64), (64, 64), (64, 64)]   # (in, out)
BOUNDS = ((-2.0, 0.0, -2.0), (2.0, 3.0, 2.0))


def _offsets():
    off, out = 0, []
    for i, o in LAYERS:
        out.append((off, off + o * i))
        off += o * i + o
    assert off == capi.PE_GUIDE_PARAMS
    return out


def features(pos, dirs, bounds=BOUNDS):
    """The encoding in float64: six coordinates in [0, 1], 12 frequencies of sin/cos of 2^k pi v (from the exact
    fractional part of 2^(k-1) v, in revolutions), the raw coordinates, zero padding to 160."""
    lo, hi = np.asarray(bounds[0], np.float32), np.asarray(bounds[1], np.float32)
    inv = (np.float32(1.0) / np.maximum(hi - lo, np.float32(1e-6))).astype(np.float32)
    v = np.concatenate([((pos.astype(np.float32) - lo) * inv), np.float32(0.5) * (dirs.astype(np.float32) + 1)], 1)
    v = v.astype(np.float32)
    n = v.shape[0]
    x = np.zeros((n, 160), np.float64)
    for c in range(6):
        for k in range(12):
            r = (v[:, c] * np.float32(2.0 ** (k - 1))).astype(np.float32)
            fr = (r - np.floor(r)).astype(np.float64)
            x[:, c * 24 + 2 * k] = np.sin(2 * np.pi * fr)
            x[:, c * 24 + 2 * k + 1] = np.cos(2 * np.pi * fr)
    x[:, 144:150] = v
    return x


def torch_forward(params, x):
    import torch

    bf = lambda t: t.to(torch.bfloat16).to(torch.float32)  # noqa: E731
    h = bf(torch.as_tensor(x, dtype=torch.float32))
    offs = _offsets()
    for li, ((i, o), (w0, b0)) in enumerate(zip(LAYERS, offs)):
        W = bf(params[w0:w0 + o * i].reshape(o, i))
        b = params[b0:b0 + o]
        h = h @ W.T + b
        if li < len(LAYERS) - 1:
            h = bf(torch.relu(h))
    return h


def torch_mixture_nll(o, d):
    """-log q(d) per row for raw outputs o (n, 64) and unit directions d (n, 3)."""
    import torch

    lobes = o[:, :40].reshape(-1, 8, 5)
    u = lobes[..., :3]
    ulen = u.norm(dim=-1, keepdim=True)
    mu = u / ulen
    k = torch.exp(torch.clamp(lobes[..., 3], -4.0, 7.0))
    logw = torch.log_softmax(lobes[..., 4], dim=-1)
    cosd = (mu * d[:, None, :]).sum(-1)
    logc = torch.log(k) - torch.log(2 * math.pi * (-torch.expm1(-2 * k)))
    return -torch.logsumexp(logw + logc + k * (cosd - 1), dim=-1)


def torch_lobes(o):
    import torch

    lobes = o[:, :40].reshape(-1, 8, 5)
    u = lobes[..., :3]
    mu = u / u.norm(dim=-1, keepdim=True)
    k = torch.exp(torch.clamp(lobes[..., 3], -4.0, 7.0))
    w = torch.softmax(lobes[..., 4], dim=-1)
    return torch.cat([mu, k[..., None], w[..., None]], -1)


def _unit(rng, n):
    d = rng.normal(size=(n, 3))
    return (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)


def _guide(dev, gs, **kw):
    from photon_hip.guide import Guide

    return Guide(dev, gs, bounds=BOUNDS, **kw)


@pytest.fixture(scope="module")
def small_room(dev):
    sc, view = configs.slit_room(width=96, height=54, cell=0.25, max_bounces=6)
    return sc, view, dev.upload(sc)


def _sharpen(g, rng, scale=3.0):
    """Larger random weights than the initialisation, so that lobes differ and gradients are not tiny."""
    p = g.params()
    p = (p * scale + rng.normal(0, 0.02, p.size).astype(np.float32) * (p != 0)).astype(np.float32)
    g.set_params(p)
    return p


@pytest.mark.gpu
def test_guide_inference_matches_torch(dev, small_room):
    import torch

    _, _, gs = small_room
    rng = np.random.default_rng(1)
    g = _guide(dev, gs, seed=3)
    p = _sharpen(g, rng)
    n = 3000
    pos = rng.uniform(BOUNDS[0], BOUNDS[1], (n, 3)).astype(np.float32)
    dirs = _unit(rng, n)
    got = g.infer(pos, dirs)
    ref = torch_lobes(torch_forward(torch.as_tensor(p), features(pos, dirs))).numpy()
    np.testing.assert_allclose(got[..., :3], ref[..., :3], atol=3e-2)
    np.testing.assert_allclose(got[..., 3], ref[..., 3], rtol=3e-2, atol=1e-3)
    np.testing.assert_allclose(got[..., 4], ref[..., 4], atol=1e-2)
    np.testing.assert_allclose(got[..., 4].sum(1), 1.0, atol=1e-5)


@pytest.mark.gpu
def test_guide_gradient_and_adam_match_torch(dev, small_room):
    import torch

    _, _, gs = small_room
    rng = np.random.default_rng(2)
    g = _guide(dev, gs, seed=5, batch=4096, learning_rate=1e-3)
    p0 = _sharpen(g, rng)
    n = 3000
    pos = rng.uniform(BOUNDS[0], BOUNDS[1], (n, 3)).astype(np.float32)
    wo, wi = _unit(rng, n), _unit(rng, n)
    w = rng.uniform(0.1, 2.0, n).astype(np.float32)
    recs = np.concatenate([pos, wo, wi, w[:, None]], 1).astype(np.float32)
    loss, grad = g.train_step(recs, want_grad=True)
    pt = torch.tensor(p0, requires_grad=True)
    o = torch_forward(pt, features(pos, wo))
    wt = torch.as_tensor(w)
    ref_loss = (wt * torch_mixture_nll(o, torch.as_tensor(wi))).sum() / wt.sum()
    ref_loss.backward()
    ref = pt.grad.numpy()
    assert abs(loss - ref_loss.item()) <= 2e-3 * abs(ref_loss.item()) + 1e-4, (loss, ref_loss.item())
    for (i, o_), (w0, b0) in zip(LAYERS, _offsets()):
        for a, b in ((w0, b0), (b0, b0 + o_)):
            gl, rl = grad[a:b], ref[a:b]
            err = np.linalg.norm(gl - rl) / max(np.linalg.norm(rl), 1e-12)
            assert err <= 6e-2, (a, b, err)
    cos = float(grad @ ref / (np.linalg.norm(grad) * np.linalg.norm(ref)))
    assert cos >= 0.998, cos
    # Adam's first step with bias correction: p - lr * g / (|g| + eps).
    p1 = g.params()
    expect = p0 - np.float32(1e-3) * grad / (np.sqrt(grad.astype(np.float64) ** 2) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-7)


@pytest.mark.gpu
def test_guide_learns_a_direction_field(dev, small_room):
    """Records whose directions follow a position-dependent vMF: the loss falls and the learned mixture
    puts most of its mass near the field's direction."""
    _, _, gs = small_room
    rng = np.random.default_rng(3)
    g = _guide(dev, gs, seed=7, batch=8192, learning_rate=3e-3)

    def batch(n):
        pos = rng.uniform(BOUNDS[0], BOUNDS[1], (n, 3)).astype(np.float32)
        mu = np.stack([np.sign(pos[:, 0]), np.full(n, 0.3), np.zeros(n)], 1)
        mu /= np.linalg.norm(mu, axis=1, keepdims=True)
        d = mu + rng.normal(0, 0.15, (n, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return pos, mu, np.concatenate([pos, _unit(rng, n), d, np.ones((n, 1))], 1).astype(np.float32)

    losses = []
    for _ in range(300):
        losses.append(g.train_step(batch(8192)[2])[0])
    assert losses[-1] < losses[0] - 1.5, (losses[0], losses[-1])
    pos, mu, recs = batch(512)
    lobes = g.infer(pos, recs[:, 3:6])
    mean_dir = (lobes[..., :3] * lobes[..., 4:5]).sum(1)
    assert float(np.mean(np.sum(mean_dir * mu, 1))) > 0.7


def _render(dev, gs, view, spp, guide=None, first=0):
    from photon_hip.guide import render_guided

    film = dev.film(view.width, view.height)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=8)
    if guide is None:
        render(dev, gs, cam, st, film, first, spp)
    else:
        render_guided(dev, gs, guide, cam, st, film, first, spp)
    s, c, _, rej = film.download()
    assert (c == spp).all() and rej == 0
    return s / c[..., None]


def relmse(x, ref):
    return float(np.mean((x - ref) ** 2 / (ref ** 2 + 0.01)))


@pytest.mark.gpu
def test_guided_render_converges_to_plain(dev, small_room):
    sc, view, gs = small_room
    ref = _render(dev, gs, view, 2048, first=4096)
    g = _guide(dev, gs, seed=11, batch=16384, steps_per_launch=8)
    guided = _render(dev, gs, view, 1024, guide=g)
    plain = _render(dev, gs, view, 1024)
    st = g.stats()
    assert st["steps"] > 0 and st["records"] > 0
    # Unbiased: the guided image agrees with plain path tracing about as well as another plain image does.
    e_g, e_p = relmse(guided, ref), relmse(plain, ref)
    assert e_g <= 2.0 * e_p + 1e-4, (e_g, e_p)
    assert abs(guided.mean() / ref.mean() - 1.0) < 0.02


@pytest.mark.gpu
def test_guided_mean_within_sigma_of_plain_over_independent_ranges(dev, small_room):
    """Unbiasedness as a statistical test: a guide trained for 4 launches and then frozen renders 12 disjoint
    64-spp sample ranges, plain path tracing 12 others; for the image mean and each quadrant's mean the guided and
    plain estimates differ by less than 4 standard errors of their difference. The test's reach is printed: the
    smallest relative bias it would flag (4 sigma of the difference over the mean)."""
    sc, view, gs = small_room
    g = _guide(dev, gs, seed=17, batch=16384, steps_per_launch=8, train_launches=4)
    _render(dev, gs, view, 4 * 8, guide=g, first=1 << 16)  # the training launches
    K, spp = 12, 64
    guided = np.stack([_render(dev, gs, view, spp, guide=g, first=k * spp) for k in range(K)])
    plain = np.stack([_render(dev, gs, view, spp, first=(K + k) * spp) for k in range(K)])
    h, w = view.height // 2, view.width // 2
    regions = {"image": (slice(None), slice(None))}
    for qy in range(2):
        for qx in range(2):
            regions[f"q{qy}{qx}"] = (slice(qy * h, (qy + 1) * h), slice(qx * w, (qx + 1) * w))
    reach = []
    for name, (ys, xs) in regions.items():
        a = guided[:, ys, xs].mean(axis=(1, 2, 3))
        b = plain[:, ys, xs].mean(axis=(1, 2, 3))
        se = np.sqrt(a.var(ddof=1) / K + b.var(ddof=1) / K)
        diff = a.mean() - b.mean()
        reach.append(4.0 * se / b.mean())
        assert abs(diff) <= 4.0 * se, (name, diff, se)
    print(f"guided vs plain over {K} independent ranges: would flag a relative bias above {max(reach):.4f} "
          f"(image {reach[0]:.4f})", flush=True)
    assert reach[0] < 0.05


@pytest.mark.gpu
def test_guiding_reduces_error_on_the_slit_room(dev):
    """At equal sample count, after 24 launches x 32 Adam steps of online training, the guided image of the
    door-gap room has less error than the plain one (measured: about 0.83 x at 240 x 135)."""
    sc, view = configs.slit_room(width=240, height=135, cell=0.1, max_bounces=8)
    gs = dev.upload(sc)
    ref = _render(dev, gs, view, 8192, first=1 << 20)
    g = _guide(dev, gs, seed=13, learning_rate=3e-3, steps_per_launch=32, train_launches=24)
    _render(dev, gs, view, 24 * 8, guide=g, first=20000)   # the training launches, 8 spp each
    guided = _render(dev, gs, view, 128, guide=g)
    plain = _render(dev, gs, view, 128)
    e_g, e_p = relmse(guided, ref), relmse(plain, ref)
    assert e_g < 0.95 * e_p, (e_g, e_p)


@pytest.mark.gpu
def test_guided_render_is_unbiased_through_subsurface(dev):
    """A guided vertex takes the BSDF branch with probability 1 - alpha: a subsurface entry, which the mixture
    cannot produce, must carry weight / (1 - alpha) (path_tracer.cpp:721-726). Without it the light through the
    block would come out scaled by 1 - alpha (half). Guided and plain renders of a close view of a subsurface
    block (the rough glass sphere beside it is a volume boundary, which is not guided) must agree in their mean
    to a few per cent."""
    from photon_hip.guide import render_guided
    from photon_hip.scene import AreaLight, Mesh, Scene, dielectric, disney, lambert

    sc = Scene()
    up = np.array([0.0, 1.0, 0.0], np.float32)
    sc.add_shape(Mesh(np.array([[-4, 0, -4], [4, 0, -4], [4, 0, 4], [-4, 0, 4]], np.float32),
                      np.array([[0, 2, 1], [0, 3, 2]], np.uint32), np.stack([up] * 4)),
                 sc.add_material(lambert((0.6, 0.6, 0.6))))
    sc.add_shape(configs.box_mesh((0.3, 0.002, -0.25), (0.8, 0.5, 0.25)),
                 sc.add_material(disney((0.85, 0.55, 0.4), 0.0, 0.4, 0.5, subsurface=1.0, subsurface_radius=(2, 3, 4),
                                        subsurface_scale=1e-3)))
    sc.add_shape(configs.icosphere(2, 0.15, (0.15, 0.15, 0.2)),
                 sc.add_material(dielectric(1.45, tint=(1.0, 0.97, 0.9), roughness=0.25)))
    sc.add_light(AreaLight((-0.6, 2.5, -0.2), (1.2, 0.0, 0.0), (0.0, 0.0, 0.8), (12.0, 12.0, 12.0)))
    view = configs.View((0.55, 0.6, 1.6), (0.55, 0.25, 0.0), (0.0, 1.0, 0.0), 40.0, 64, 48, 16, 8)
    gs = dev.upload(sc)
    cam = camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=16)
    g = _guide(dev, gs, seed=5, batch=16384, steps_per_launch=4, train_launches=1, guided_bounces=3)
    guided, plain = dev.film(view.width, view.height), dev.film(view.width, view.height)
    render_guided(dev, gs, g, cam, st, dev.film(view.width, view.height), 1 << 20, 16)   # trains: guide on
    assert g.stats()["steps"] > 0
    render_guided(dev, gs, g, cam, st, guided, 0, 512)
    render(dev, gs, cam, st, plain, 1 << 21, 512)
    mg, mp = guided.mean(), plain.mean()
    shift = float(mg.mean() / mp.mean() - 1.0)
    print(f"guided / plain mean through subsurface: {1 + shift:.4f}", flush=True)
    assert abs(shift) < 0.03, shift
