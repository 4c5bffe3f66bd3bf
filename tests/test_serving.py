"""The render-job server (pe_server_* / pe_job_*): continuous batching of render jobs of different scenes,
film sizes and sample counts on one device. Its model is RenderSession's batch loop (engine/
render_session.cpp:69-101) run for many sessions at once.

What is checked: a job's film is bit for bit the same alone and in a mix of concurrent jobs, and the same as
pe_render over the job's chunk boundaries; admission waits for device memory (films and resident scenes) and parks idle scenes in host memory to make
room, restoring them bit for bit; a short job's preview overtakes a long job's later chunks; a scene cannot be
released while jobs reference it, and scene updates wait for the chunks rendering it; cancel, release and the
error paths.
"""
from __future__ import annotations

import numpy as np
import pytest

from photon_hip import capi, configs
from photon_hip.device import camera_perspective, render, settings
from photon_hip.serve import Server

CHUNK_PATHS = 1 << 16


def _cam(view):
    return camera_perspective(view.position, view.target, view.up, view.vfov, view.aspect)


def _scenes():
    a = configs.box_icosphere(level=3, width=48, height=40, max_bounces=6)
    b = configs.product_sphere(width=64, height=36, lon=64, lat=48, env_width=128, env_height=64)
    c = configs.glass_studio(width=40, height=30)
    d = configs.led_strip(quads=60, width=32, height=32, max_bounces=2)
    return [a, b, c, d]


def _submit(srv, gs, view, spp, preview=8):
    st = settings(max_bounces=view.max_bounces, seed=12345)
    return srv.submit(gs, _cam(view), st, view.width, view.height, spp, preview)


@pytest.mark.gpu
def test_job_alone_equals_job_in_mix_and_chunked_render(dev):
    scenes = _scenes()
    gscenes = [dev.upload(sc) for sc, _ in scenes]
    (sc_a, view_a), gs_a = scenes[0], gscenes[0]
    srv = Server(dev, slots=4, chunk_spp=4, max_chunk_paths=CHUNK_PATHS)
    try:
        jid = _submit(srv, gs_a, view_a, spp=24)
        st = srv.wait(jid, 60000)
        assert st.state == capi.PE_JOB_DONE and st.samples_done == 24, st
        alone = srv.download(jid)
        srv.release(jid)
        ids = []
        for k in range(3):                       # the mix: every scene, some twice, all at once
            for (sc, view), gs in zip(scenes, gscenes):
                ids.append((gs is gs_a and k == 1, _submit(srv, gs, view, spp=12 + 4 * k)))
        target = [j for is_a, j in ids if is_a][0]
        for _, j in ids:
            s = srv.wait(j, 120000)
            assert s.state == capi.PE_JOB_DONE, s
        mixed = srv.download(target)            # 16 spp of scene a, rendered among eleven other jobs
        np.testing.assert_array_equal(mixed[1], np.full((view_a.height, view_a.width), 16, np.uint32))
        for _, j in ids:
            srv.release(j)
        jid2 = _submit(srv, gs_a, view_a, spp=16)  # the same job alone
        srv.wait(jid2, 60000)
        alone16 = srv.download(jid2)
        np.testing.assert_array_equal(mixed[0], alone16[0])
        np.testing.assert_array_equal(mixed[2], alone16[2])
    finally:
        srv.close()
    # pe_render over the same chunk boundaries (preview 8, then chunks of 4) gives the same film.
    film = dev.film(view_a.width, view_a.height)
    first = 0
    for n in (8, 4, 4, 4, 4):
        st = settings(max_bounces=view_a.max_bounces, seed=12345, samples_per_launch=n)
        render(dev, gs_a, _cam(view_a), st, film, first, n)
        first += n
    s, c, l2, _ = film.download()
    np.testing.assert_array_equal(s, alone[0])
    np.testing.assert_array_equal(l2, alone[2])


@pytest.mark.gpu
def test_admission_waits_for_film_memory(dev):
    (sc, view) = _scenes()[0]
    gs = dev.upload(sc)
    one_film = view.width * view.height * 20 + 8
    # The budget covers the (resident) scene and one film.
    srv = Server(dev, slots=2, chunk_spp=4, max_chunk_paths=CHUNK_PATHS,
                 memory_budget=gs.stats()["device_bytes"] + one_film)
    try:
        j1 = _submit(srv, gs, view, spp=8)
        j2 = _submit(srv, gs, view, spp=8)
        assert srv.wait(j1, 60000).state == capi.PE_JOB_DONE
        assert srv.poll(j2).state == capi.PE_JOB_QUEUED      # j1 still holds the only film
        srv.release(j1)
        s2 = srv.wait(j2, 60000)
        assert s2.state == capi.PE_JOB_DONE and s2.t_admit >= 0
        big = srv.try_submit(gs, _cam(view), settings(), view.width * 2, view.height, 4)
        assert big[0] == capi.PE_ERR_OOM                      # can never fit the budget
        srv.release(j2)
    finally:
        srv.close()


@pytest.mark.gpu
def test_short_job_preview_overtakes_long_job(dev):
    scenes = _scenes()
    (sc_l, view_l), (sc_s, view_s) = scenes[1], scenes[0]
    gl, gsh = dev.upload(sc_l), dev.upload(sc_s)
    srv = Server(dev, slots=1, chunk_spp=2, max_chunk_paths=CHUNK_PATHS)
    try:
        long_job = _submit(srv, gl, view_l, spp=200, preview=2)
        short_job = _submit(srv, gsh, view_s, spp=64, preview=4)
        s_short = srv.wait(short_job, 120000)
        s_long = srv.wait(long_job, 120000)
        assert s_short.state == s_long.state == capi.PE_JOB_DONE
        assert s_short.t_preview < s_long.t_done
        assert s_long.t_preview <= s_short.t_preview          # the long job's own preview went first
    finally:
        srv.close()


@pytest.mark.gpu
def test_cancel_release_and_errors(dev):
    (sc, view) = _scenes()[1]
    gs = dev.upload(sc)
    srv = Server(dev, slots=1, chunk_spp=1, max_chunk_paths=CHUNK_PATHS)
    try:
        j = _submit(srv, gs, view, spp=100000, preview=1)
        srv.cancel(j)
        s = srv.wait(j, 60000)
        assert s.state == capi.PE_JOB_CANCELLED and s.samples_done < 100000
        _, cnt, _ = srv.download(j)
        assert (cnt == s.samples_done).all()                  # whole chunks only
        srv.release(j)
        with pytest.raises(capi.PhotonHipError):
            srv.poll(j)
        assert srv.try_submit(gs, _cam(view), settings(), 0, 8, 4)[0] == capi.PE_ERR_INVALID
        assert srv.try_submit(gs, _cam(view), settings(), 8, 8, 0)[0] == capi.PE_ERR_INVALID
        j2 = _submit(srv, gs, view, spp=3, preview=1)
        assert srv.wait(j2, 60000).samples_done == 3
    finally:
        srv.close()


@pytest.mark.gpu
def test_job_display_is_the_display_of_its_film(dev):
    """pe_job_display of a finished job equals pe_film_display of the same samples rendered directly (the job's
    film equals pe_render over its chunk boundaries), with an imager and tone map that exercise every stage."""
    from photon_hip.device import display_settings

    (sc, view) = configs.product_sphere(width=64, height=36, lon=64, lat=48, env_width=128, env_height=64)
    gs = dev.upload(sc)
    d = display_settings(tonemap=capi.PE_TONEMAP_AGX, exposure_ev=0.5, bloom_intensity=0.4, glare_intensity=0.2,
                         white_balance_k=5000.0, contrast=1.2, vignette=0.3)
    srv = Server(dev, slots=2, chunk_spp=4, max_chunk_paths=CHUNK_PATHS)
    try:
        jid = _submit(srv, gs, view, spp=12)
        assert srv.wait(jid, 60000).state == capi.PE_JOB_DONE
        job_disp, job8 = srv.display(jid, d)
    finally:
        srv.close()
    film = dev.film(view.width, view.height)
    first = 0
    for n in (8, 4):
        render(dev, gs, _cam(view), settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=n), film,
               first, n)
        first += n
    disp, rgb8 = film.display(d)
    np.testing.assert_array_equal(job_disp, disp)
    np.testing.assert_array_equal(job8, rgb8)
    assert 0.05 < disp.mean() < 0.95


@pytest.mark.gpu
def test_idle_scene_is_parked_for_room_and_restored_bit_for_bit(dev):
    """A budget that holds one scene besides a film: a job on the second scene parks the first (idle), a later
    job on the first restores it, and its film equals the first job's bit for bit."""
    (sc_a, view_a), (sc_b, view_b) = _scenes()[1], _scenes()[0]
    ga, gb = dev.upload(sc_a), dev.upload(sc_b)
    bytes_a, bytes_b = ga.stats()["device_bytes"], gb.stats()["device_bytes"]
    film = max(view_a.width * view_a.height, view_b.width * view_b.height) * 20 + 8
    srv = Server(dev, slots=2, chunk_spp=4, max_chunk_paths=CHUNK_PATHS,
                 memory_budget=max(bytes_a, bytes_b) + min(bytes_a, bytes_b) // 2 + film)
    try:
        j1 = _submit(srv, ga, view_a, spp=8)
        assert srv.wait(j1, 60000).state == capi.PE_JOB_DONE
        first = srv.download(j1)
        srv.release(j1)
        j2 = _submit(srv, gb, view_b, spp=8)                  # room for scene b: scene a (idle) is parked
        assert srv.wait(j2, 60000).state == capi.PE_JOB_DONE
        st = srv.stats()
        assert st["scenes_parked"] == 1 and st["scenes_resident"] == 1, st
        assert st["memory_used"] <= st["memory_budget"], st
        srv.release(j2)
        j3 = _submit(srv, ga, view_a, spp=8)                  # scene a back: b is parked in turn
        assert srv.wait(j3, 60000).state == capi.PE_JOB_DONE
        again = srv.download(j3)
        st = srv.stats()
        assert st["scenes_restored"] == 1 and st["scenes_parked"] == 2, st
        for x, y in zip(first, again):
            np.testing.assert_array_equal(x, y)
        # A scene referenced by a job that is not released cannot be released; pe_render restores a parked scene.
        assert capi.load().pe_scene_release(ga.h) == capi.PE_ERR_INVALID
        srv.release(j3)
        film_b = dev.film(view_b.width, view_b.height)
        render(dev, gb, _cam(view_b), settings(max_bounces=view_b.max_bounces, seed=12345, samples_per_launch=8),
               film_b, 0, 8)
        assert (film_b.download()[1] == 8).all()
    finally:
        srv.close()


@pytest.mark.gpu
def test_scene_updates_wait_for_running_chunks(dev):
    """pe_scene_update may be called while server lanes render the scene: it takes the scene's lock, which every
    chunk holds shared, so it waits for the chunks in progress and no chunk starts on half-rebuilt buffers (the
    Full rebuild frees the old ones). The job still finishes every sample with finite values, and afterwards the
    scene renders bit for bit what a fresh upload of the last description renders."""
    import threading

    sc, view = configs.box_icosphere(level=3, width=48, height=40, max_bounces=6)
    gs = dev.upload(sc)
    srv = Server(dev, slots=2, chunk_spp=2, max_chunk_paths=CHUNK_PATHS)
    try:
        jobs = [_submit(srv, gs, view, spp=96), _submit(srv, gs, view, spp=96)]
        while srv.poll(jobs[0]).samples_done == 0:          # a lane is rendering the scene
            threading.Event().wait(0.001)
        for m in sc.materials[:3]:
            m.base_color[:] = [0.8, 0.2, 0.1]
        assert gs.update(sc, capi.PE_UPDATE_MATERIALS) == capi.PE_OK
        sc.lights[0].radiance[:] = [4.0, 4.0, 9.0]
        assert gs.update(sc, capi.PE_UPDATE_LIGHTS) == capi.PE_OK
        moved, _ = configs.box_icosphere(level=3, width=48, height=40, max_bounces=6)
        moved.shapes[6].mesh.positions = moved.shapes[6].mesh.positions + np.float32([0.15, 0.1, 0.0])
        assert gs.update(moved, capi.PE_UPDATE_FULL) == capi.PE_OK
        assert capi.load().pe_scene_release(gs.h) == capi.PE_ERR_INVALID   # jobs still reference it
        for j in jobs:
            st = srv.wait(j, 120000)
            assert st.state == capi.PE_JOB_DONE and st.samples_done == 96, st
            s, c, *_ = srv.download(j)
            assert (c == 96).all() and np.isfinite(s).all()
            srv.release(j)
    finally:
        srv.close()
    film = dev.film(view.width, view.height)
    st = settings(max_bounces=view.max_bounces, seed=12345, samples_per_launch=4)
    render(dev, gs, _cam(view), st, film, 0, 4)
    fresh = dev.film(view.width, view.height)
    render(dev, dev.upload(moved), _cam(view), st, fresh, 0, 4)
    np.testing.assert_array_equal(film.download()[0], fresh.download()[0])


@pytest.mark.gpu
def test_serving_stream_at_workload_sizes(dev):
    """The serving workload's shapes on the device: a seeded Poisson stream that holds a 2048 x 2048 film, a
    512-spp job and the 5.1M-triangle displaced sphere beside the small plain scene. Every job completes, every
    preview lands before its job completes, and each of three jobs re-rendered alone afterwards (the 5M-sphere
    job among them) gives the same film bit for bit as in the mix."""
    rng = np.random.default_rng(3)
    big_sc, big_view = configs.product_sphere(lon=1600, lat=1600, env_width=512, env_height=256)
    assert big_sc.num_triangles > 5_000_000
    small_sc, small_view = configs.box_icosphere(level=5)
    g_big, g_small = dev.upload(big_sc), dev.upload(small_sc)
    jobs = [(g_big, big_view, 2048, 2048, 24), (g_small, small_view, 512, 512, 512), (g_big, big_view, 256, 256, 64),
            (g_small, small_view, 256, 256, 16), (g_big, big_view, 1920, 1080, 32), (g_small, small_view, 512, 512, 40)]
    arrivals = np.cumsum(rng.exponential(1.0 / 20.0, len(jobs)))
    srv = Server(dev, slots=4, chunk_spp=16)
    try:
        import time

        t0 = time.perf_counter()
        ids = []
        for (gs, view, w, h, spp), at in zip(jobs, arrivals):
            while time.perf_counter() - t0 < at:
                time.sleep(0.001)
            st = settings(max_bounces=view.max_bounces, seed=12345)
            cam = camera_perspective(view.position, view.target, view.up, view.vfov, w / h)
            ids.append(srv.submit(gs, cam, st, w, h, spp, 16))
        stats = [srv.wait(j, 300000) for j in ids]
        for s, (_, _, w, h, spp) in zip(stats, jobs):
            assert s.state == capi.PE_JOB_DONE and s.samples_done == spp, s
            assert 0 <= s.t_preview <= s.t_done
        films = {k: srv.download(ids[k]) for k in (0, 1, 4)}
        for j in ids:
            srv.release(j)
        for k in (0, 1, 4):                                   # the same jobs alone
            gs, view, w, h, spp = jobs[k]
            cam = camera_perspective(view.position, view.target, view.up, view.vfov, w / h)
            j = srv.submit(gs, cam, settings(max_bounces=view.max_bounces, seed=12345), w, h, spp, 16)
            assert srv.wait(j, 300000).state == capi.PE_JOB_DONE
            alone = srv.download(j)
            srv.release(j)
            for x, y in zip(films[k], alone):
                np.testing.assert_array_equal(x, y)
            print(f"job {k}: {w}x{h} {spp} spp alone == in the mix", flush=True)
    finally:
        srv.close()
