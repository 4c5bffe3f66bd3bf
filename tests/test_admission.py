"""Device-memory admission of the render-job server (csrc/pe_admission.h, driven by serve.hip), on the CPU
against a capped allocator. The model generalised is RenderSession's one-film-per-session loop
(engine/render_session.cpp:69-101): a server admits many jobs, each holding a film, and keeps their scenes'
BVHs, triangles, materials and environments resident across jobs, parking idle scenes in host memory (least
recently used first) when the head of the queue needs the room.

Checked: the allocator never exceeds its cap over a seeded stream of admissions and releases; a scene an
admitted job renders is never parked; eviction order is least recently used; a job whose film and scene
exceed the budget is refused outright; a parked scene comes back when a job needs it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.fixture(scope="module")
def planner_lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("admission") / "libadmission.so"
    subprocess.run(["g++", "-O1", "-std=c++20", "-shared", "-fPIC", f"-I{ROOT / 'photon-engine_amd' / 'csrc'}",
                    str(ROOT / "tests" / "harness" / "admission_harness.cpp"), "-o", str(out)], check=True)
    lib = C.CDLL(str(out))
    lib.planner_new.restype = C.c_void_p
    lib.planner_new.argtypes = [C.c_uint64]
    lib.planner_delete.argtypes = [C.c_void_p]
    lib.planner_track.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.c_int]
    lib.planner_forget.argtypes = [C.c_void_p, C.c_uint64]
    lib.planner_plan.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.c_int,
                                 C.POINTER(C.c_int)]
    lib.planner_commit.argtypes = [C.c_void_p, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64), C.c_int, C.c_int]
    lib.planner_end_use.argtypes = [C.c_void_p, C.c_uint64]
    lib.planner_free_film.argtypes = [C.c_void_p, C.c_uint64]
    lib.planner_used.restype = C.c_uint64
    lib.planner_used.argtypes = [C.c_void_p]
    lib.planner_resident.argtypes = [C.c_void_p, C.c_uint64]
    return lib


class CappedDevice:
    """A device allocator with a hard cap: what the server allocates when it carries out a plan."""

    def __init__(self, cap):
        self.cap, self.used, self.peak = cap, 0, 0

    def alloc(self, n):
        assert self.used + n <= self.cap, f"allocation of {n} over the cap ({self.used}/{self.cap})"
        self.used += n
        self.peak = max(self.peak, self.used)

    def free(self, n):
        self.used -= n
        assert self.used >= 0


class Server:
    """The server's side of admission (serve.hip admit / release), with the planner deciding."""

    def __init__(self, lib, budget, scenes):
        self.lib, self.p = lib, lib.planner_new(budget)
        self.dev = CappedDevice(budget)
        self.scenes = dict(scenes)            # key -> bytes
        self.resident = {k: True for k in scenes}
        for k, b in scenes.items():           # scenes uploaded by the caller before the server saw them
            self.dev.alloc(b)
            lib.planner_track(self.p, k, b, 1)
        self.users = {k: 0 for k in scenes}
        self.parks = []

    def try_admit(self, film, key):
        ev = (C.c_uint64 * 64)()
        n = C.c_int()
        r = self.lib.planner_plan(self.p, film, key, ev, 64, C.byref(n))
        if r & 2:
            return "never"
        if not r & 1:
            return "wait"
        for k in ev[:n.value]:
            assert self.users[k] == 0, "a scene in use was chosen for parking"
            self.dev.free(self.scenes[k])
            self.resident[k] = False
            self.parks.append(k)
        if r & 4:
            self.dev.alloc(self.scenes[key])
            self.resident[key] = True
        self.dev.alloc(film)
        self.lib.planner_commit(self.p, film, key, ev, n.value, 1 if r & 4 else 0)
        self.users[key] += 1
        return "admit"

    def end(self, film, key):
        self.lib.planner_end_use(self.p, key)
        self.users[key] -= 1
        self.lib.planner_free_film(self.p, film)
        self.dev.free(film)


def test_idle_scenes_are_parked_least_recently_used_first(planner_lib):
    MB = 1 << 20
    srv = Server(planner_lib, 1000 * MB, {1: 300 * MB, 2: 300 * MB, 3: 300 * MB})  # 900 of 1000 resident
    assert srv.try_admit(50 * MB, 1) == "admit"
    srv.end(50 * MB, 1)                                   # scene 1 was used last; 2 and 3 are older
    assert srv.try_admit(200 * MB, 1) == "admit"          # needs 150 MB more than is free: park one idle scene
    assert srv.parks == [2]                               # the least recently used idle one
    assert srv.try_admit(450 * MB, 3) == "wait"           # scene 1 + its film (500) + 450 + scene 3 > budget now
    srv.end(200 * MB, 1)
    assert srv.try_admit(450 * MB, 3) == "admit"
    assert srv.parks == [2, 1]                            # scene 1, idle again, made room
    assert srv.try_admit(100 * MB, 2) == "wait"           # scene 2 must come back: 300 + 100 more than is free
    srv.end(450 * MB, 3)
    assert srv.try_admit(100 * MB, 2) == "admit"          # restored (scene 3 stays: it fits beside)
    assert srv.resident[2] and srv.resident[3]
    assert planner_lib.planner_used(srv.p) == srv.dev.used


def test_a_job_that_can_never_fit_is_refused(planner_lib):
    MB = 1 << 20
    srv = Server(planner_lib, 1000 * MB, {1: 600 * MB})
    assert srv.try_admit(401 * MB, 1) == "never"
    assert srv.try_admit(400 * MB, 1) == "admit"


def test_seeded_stream_never_exceeds_the_cap(planner_lib):
    rng = np.random.default_rng(3)
    MB = 1 << 20
    scenes = {k: int(rng.integers(20, 400)) * MB for k in range(1, 9)}
    budget = 1500 * MB
    srv = Server(planner_lib, budget, {k: b for k, b in list(scenes.items())[:3]})
    for k in list(scenes)[3:]:          # scenes the caller uploaded but the server may park: some start parked
        srv.scenes[k] = scenes[k]
        srv.resident[k] = False
        srv.users[k] = 0
        planner_lib.planner_track(srv.p, k, scenes[k], 0)
    queue, running, refused = [], [], 0
    for step in range(4000):
        if rng.random() < 0.55:
            film = int(rng.choice([256 * 256, 512 * 512, 1920 * 1080, 2048 * 2048])) * 20
            queue.append((film, int(rng.integers(1, 9))))
        # Admission in order; the head is never overtaken.
        while queue:
            r = srv.try_admit(*queue[0])
            if r == "wait":
                break
            if r == "never":
                refused += 1
            else:
                running.append(queue[0])
            queue.pop(0)
        if running and rng.random() < 0.5:
            srv.end(*running.pop(int(rng.integers(len(running)))))
        assert srv.dev.used <= budget
        assert planner_lib.planner_used(srv.p) == srv.dev.used
        for film, key in running:
            assert srv.resident[key], "a running job's scene is not resident"
    assert srv.parks, "the stream never needed to park a scene"
    assert refused == 0
    print(f"peak {srv.dev.peak / MB:.0f} of {budget / MB:.0f} MB; {len(srv.parks)} scenes parked")
