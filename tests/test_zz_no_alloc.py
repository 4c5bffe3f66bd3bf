"""The allocation-free closest-hit entry point (pe_intersect_batch_noalloc, no counterpart in the reference):
on caller-owned device buffers it returns, bit for bit, what pe_intersect_batch returns, on two calls in a row
(the second smaller, so stale scratch contents from the first would show)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from photon_hip import capi, configs

pytestmark = pytest.mark.gpu


def _rays(rng, n):
    r = np.zeros((n, 8), np.float32)
    r[:, :3] = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    r[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
    r[:, 7] = 1e30
    return r


class _Hip:
    """The caller's side: device buffers from the HIP runtime the library itself links (hipMalloc / hipMemcpy)."""

    def __init__(self):
        capi.load()
        # The runtime libphoton_hip.so is linked against, as already mapped into this process (torch, when
        # another test has imported it, maps a second copy of its own under the same soname).
        path = "libamdhip64.so.7"
        with open("/proc/self/maps") as maps:
            for line in maps:
                f = line.split()[-1]
                if "libamdhip64.so" in f and "/torch/" not in f:
                    path = f
                    break
        self.lib = C.CDLL(path)
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def h2d(self, dst, a):
        assert self.lib.hipMemcpy(dst, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0

    def d2h(self, a, src):
        assert self.lib.hipMemcpy(a.ctypes.data_as(C.c_void_p), src, a.nbytes, 2) == 0

    def free(self, p):
        assert self.lib.hipFree(p) == 0


def test_noalloc_intersect_equals_batch_entry_point(dev):
    hip = _Hip()
    sc, _ = configs.box_icosphere(level=5, width=32, height=32)
    gs = dev.upload(sc)
    lib = capi.load()
    rng = np.random.default_rng(11)
    for n in (300_000, 70_001):
        rays = _rays(rng, n)
        want_hits, want_bary = gs.intersect(rays)
        got_hits = np.full((n, 4), -7, np.int32)
        got_bary = np.full((n, 2), -7.0, np.float32)
        d_rays, d_hits, d_bary = hip.alloc(rays.nbytes), hip.alloc(got_hits.nbytes), hip.alloc(got_bary.nbytes)
        hip.h2d(d_rays, rays)
        hip.h2d(d_hits, got_hits)
        hip.h2d(d_bary, got_bary)
        capi.check("pe_intersect_batch_noalloc", lib.pe_intersect_batch_noalloc(dev.h, gs.h, n, d_rays, d_hits, d_bary))
        hip.d2h(got_hits, d_hits)
        hip.d2h(got_bary, d_bary)
        for p in (d_rays, d_hits, d_bary):
            hip.free(p)
        print(f"n={n}: hits {np.mean(want_hits[:, 1] >= 0):.3f} of rays", flush=True)
        assert np.array_equal(got_hits, want_hits)
        assert np.array_equal(got_bary.view(np.uint32), want_bary.view(np.uint32))
    gs.close()
