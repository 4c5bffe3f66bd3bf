"""Python side of the render-job server (pe_server_* / pe_job_* in include/photon_hip.h)."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import capi
from .capi import check, ptr


@dataclass
class JobStatus:
    state: int
    samples_done: int
    error: int
    chunks: int
    t_submit: float
    t_admit: float
    t_preview: float
    t_done: float
    extension_rays: int
    shadow_rays: int
    message: str

    @property
    def ended(self) -> bool:
        return self.state in (capi.PE_JOB_DONE, capi.PE_JOB_CANCELLED, capi.PE_JOB_FAILED)


def _status(st: capi.JobStatus) -> JobStatus:
    return JobStatus(st.state, st.samples_done, st.error, st.chunks, st.t_submit, st.t_admit, st.t_preview,
                     st.t_done, st.extension_rays, st.shadow_rays, st.message.decode(errors="replace"))


class Server:
    """Admits render jobs by device memory (films and resident scenes) and interleaves their chunks on `slots`
    render lanes."""

    def __init__(self, dev, slots=0, chunk_spp=0, max_chunk_paths=0, memory_budget=0, max_admitted=0, lane_paths=0):
        self.dev = dev
        self._lib = dev._lib
        cfg = capi.ServerConfig()
        cfg.slots, cfg.chunk_spp, cfg.max_chunk_paths = slots, chunk_spp, max_chunk_paths
        cfg.memory_budget, cfg.max_admitted, cfg.lane_paths = memory_budget, max_admitted, lane_paths
        h = C.c_void_p()
        check("pe_server_create", self._lib.pe_server_create(dev.h, C.byref(cfg), C.byref(h)))
        self.h = h
        self._sizes = {}

    def clock(self) -> float:
        return self._lib.pe_server_clock(self.h)

    def stats(self) -> dict:
        st = capi.ServerStats()
        check("pe_server_stats_get", self._lib.pe_server_stats_get(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in capi.ServerStats._fields_}

    def submit(self, gscene, camera: capi.Camera, st: capi.Settings, width: int, height: int, spp: int,
               preview_spp: int = 0) -> int:
        d = capi.JobDesc()
        d.scene = gscene.h
        d.camera = camera
        d.settings = st
        d.width, d.height, d.spp, d.preview_spp = width, height, spp, preview_spp
        jid = C.c_uint64()
        check("pe_job_submit", self._lib.pe_job_submit(self.h, C.byref(d), C.byref(jid)))
        self._sizes[jid.value] = (width, height)
        return jid.value

    def try_submit(self, *args, **kw) -> tuple[int, int]:
        """submit() returning (status, id) instead of raising."""
        try:
            return capi.PE_OK, self.submit(*args, **kw)
        except capi.PhotonHipError as e:
            return e.code, 0

    def poll(self, jid: int) -> JobStatus:
        st = capi.JobStatus()
        check("pe_job_poll", self._lib.pe_job_poll(self.h, jid, C.byref(st)))
        return _status(st)

    def wait(self, jid: int, timeout_ms: int = -1) -> JobStatus:
        st = capi.JobStatus()
        rc = self._lib.pe_job_wait(self.h, jid, timeout_ms, C.byref(st))
        if rc not in (capi.PE_OK, capi.PE_ERR_TIMEOUT):
            check("pe_job_wait", rc)
        return _status(st)

    def cancel(self, jid: int) -> None:
        check("pe_job_cancel", self._lib.pe_job_cancel(self.h, jid))

    def download(self, jid: int):
        w, h = self._sizes[jid]
        s = np.zeros((h, w, 3), np.float32)
        c = np.zeros((h, w), np.uint32)
        l2 = np.zeros((h, w), np.float32)
        check("pe_job_download", self._lib.pe_job_download(self.h, jid, ptr(s, C.c_float), ptr(c, C.c_uint32),
                                                           ptr(l2, C.c_float)))
        return s, c, l2

    def display(self, jid: int, settings=None):
        """pe_job_display -> (display (H, W, 3) float32, rgb8 (H, W, 3) uint8)."""
        from .device import display_settings

        w, h = self._sizes[jid]
        d = settings if settings is not None else display_settings()
        out = np.zeros((h, w, 3), np.float32)
        rgb8 = np.zeros((h, w, 3), np.uint8)
        check("pe_job_display", self._lib.pe_job_display(self.h, jid, C.byref(d), ptr(out, C.c_float),
                                                         ptr(rgb8, C.c_uint8)))
        return out, rgb8

    def release(self, jid: int) -> None:
        check("pe_job_release", self._lib.pe_job_release(self.h, jid))
        self._sizes.pop(jid, None)

    def close(self) -> None:
        dev = getattr(self, "dev", None)
        if getattr(self, "h", None) and (dev is None or getattr(dev, "h", None)):  # as GpuScene.close
            self._lib.pe_server_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()
