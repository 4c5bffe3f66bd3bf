"""Python side of the guiding network (pe_guide_* in include/photon_hip.h)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, ptr


class Guide:
    """The 8-lobe vMF mixture MLP of one scene: inference, Adam steps, parameters, guided rendering."""

    def __init__(self, dev, gscene, learning_rate=0.0, guide_fraction=0.0, batch=0, steps_per_launch=0,
                 train_launches=0, seed=0, bounds=None, weight_clamp=0.0, guided_bounces=0):
        self.dev = dev
        self._lib = dev._lib
        cfg = capi.GuideConfig()
        cfg.learning_rate, cfg.guide_fraction, cfg.batch = learning_rate, guide_fraction, batch
        cfg.steps_per_launch, cfg.train_launches, cfg.seed = steps_per_launch, train_launches, seed
        cfg.weight_clamp = weight_clamp
        cfg.guided_bounces = guided_bounces
        if bounds is not None:
            cfg.bounds_min[:] = list(map(float, bounds[0]))
            cfg.bounds_max[:] = list(map(float, bounds[1]))
        h = C.c_void_p()
        check("pe_guide_create", self._lib.pe_guide_create(dev.h, gscene.h, C.byref(cfg), C.byref(h)))
        self.h = h

    def params(self) -> np.ndarray:
        p = np.zeros(capi.PE_GUIDE_PARAMS, np.float32)
        check("pe_guide_get_params", self._lib.pe_guide_get_params(self.h, ptr(p, C.c_float)))
        return p

    def set_params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == capi.PE_GUIDE_PARAMS
        check("pe_guide_set_params", self._lib.pe_guide_set_params(self.h, ptr(p, C.c_float)))

    def infer(self, pos: np.ndarray, dirs: np.ndarray) -> np.ndarray:
        """(n, 8, 5): per lobe mean xyz, concentration, weight."""
        pos = np.ascontiguousarray(pos, dtype=np.float32)
        dirs = np.ascontiguousarray(dirs, dtype=np.float32)
        n = pos.shape[0]
        out = np.zeros((n, 8, 5), np.float32)
        check("pe_guide_infer", self._lib.pe_guide_infer(self.h, n, ptr(pos, C.c_float), ptr(dirs, C.c_float),
                                                         ptr(out, C.c_float)))
        return out

    def train_step(self, records: np.ndarray, want_grad: bool = False):
        """records (n, 10): pos xyz, outgoing dir xyz, sampled dir xyz, weight. Returns (loss, grad or None)."""
        records = np.ascontiguousarray(records, dtype=np.float32)
        loss = C.c_float()
        grad = np.zeros(capi.PE_GUIDE_PARAMS, np.float32) if want_grad else None
        check("pe_guide_train_step", self._lib.pe_guide_train_step(self.h, records.shape[0], ptr(records, C.c_float),
                                                                   C.byref(loss), ptr(grad, C.c_float)))
        return loss.value, grad

    def stats(self) -> dict:
        s, r, q = C.c_uint64(), C.c_uint64(), C.c_uint64()
        l = C.c_float()
        check("pe_guide_stats", self._lib.pe_guide_stats(self.h, C.byref(s), C.byref(l), C.byref(r), C.byref(q)))
        return {"steps": s.value, "last_loss": l.value, "records": r.value, "inferred": q.value}

    def close(self) -> None:
        dev = getattr(self, "dev", None)
        if getattr(self, "h", None) and (dev is None or getattr(dev, "h", None)):  # as GpuScene.close
            self._lib.pe_guide_destroy(self.h)
        self.h = None

    def __del__(self):
        self.close()


def render_guided(dev, gscene, guide, camera, st, film, first: int, count: int) -> capi.BatchResult:
    res = capi.BatchResult()
    gh = guide.h if guide is not None else None
    check("pe_render_guided", dev._lib.pe_render_guided(dev.h, gscene.h, gh, C.byref(camera), C.byref(st), film.h,
                                                        first, count, None, C.byref(res)))
    return res
