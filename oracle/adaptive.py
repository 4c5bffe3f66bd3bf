"""Adaptive sampling bookkeeping restated in numpy float32 (test infrastructure only: the oracle).

Restates Film::addSample (src/core/image/film.cpp:56-71), Film::relativeError (film.cpp:100-109),
Film::resolvedPixel (film.cpp:87-94) and the adaptive rules of Renderer::renderSamples
(src/engine/renderer.cpp:438-467, the batch's 3x3 skip mask; :495-502, a pixel's stop within a batch), on a film
of float32 sums, uint32 counts and float32 luminance-square sums. Every operation is a single float32 operation
in the reference's order (numpy does not contract a * b + c), which is the reference's IEEE build.
"""
from __future__ import annotations

import numpy as np

F = np.float32


class Film:
    def __init__(self, w: int, h: int):
        self.w, self.h = w, h
        self.sum = np.zeros((h * w, 3), F)
        self.count = np.zeros(h * w, np.uint32)
        self.lum2 = np.zeros(h * w, F)
        self.rejected = 0

    def add_samples(self, idx: np.ndarray, L: np.ndarray) -> None:
        """Film::addSample of one sample L[k] for each pixel idx[k] (film.cpp:56-71)."""
        L = L.astype(F)
        self.count[idx] += 1
        ok = np.isfinite(L).all(1) & (L >= 0).all(1)
        self.rejected += int((~ok).sum())
        i, Lk = idx[ok], L[ok]
        self.sum[i] = self.sum[i] + Lk
        lum = (F(0.212671) * Lk[:, 0] + F(0.715160) * Lk[:, 1]) + F(0.072169) * Lk[:, 2]
        self.lum2[i] = self.lum2[i] + lum * lum

    def relative_error(self, idx=None) -> np.ndarray:
        """Film::relativeError (film.cpp:100-109): 1e30 below two samples."""
        idx = np.arange(self.w * self.h) if idx is None else idx
        n = self.count[idx]
        s = self.sum[idx]
        with np.errstate(divide="ignore", invalid="ignore"):
            fn = n.astype(F)
            mean = ((F(0.212671) * s[:, 0] + F(0.715160) * s[:, 1]) + F(0.072169) * s[:, 2]) / fn
            mean_sq = self.lum2[idx] / fn
            var = np.maximum(F(0.0), mean_sq - mean * mean) * fn / (n - 1).astype(F)
            err = np.sqrt(var / fn) / (np.maximum(mean, F(0.0)) + F(0.01))
        return np.where(n < 2, F(1e30), err).astype(F)

    def resolved(self) -> np.ndarray:
        """Film::resolvedPixel (film.cpp:87-94): sum * (1 / n), black without samples."""
        with np.errstate(divide="ignore"):
            inv = F(1.0) / self.count.astype(F)
        out = self.sum * inv[:, None]
        return np.where(self.count[:, None] == 0, F(0.0), out).astype(F).reshape(self.h, self.w, 3)


def render_batch(film: Film, sample_fn, first: int, count: int, threshold: float, min_spp: int) -> int:
    """One Renderer::renderSamples batch [first, first + count) with adaptive sampling on. sample_fn(s) gives the
    radiance of sample s of every pixel, (h * w, 3). Returns the samples traced."""
    threshold = max(F(1e-4), F(threshold))
    min_spp = max(2, int(min_spp))
    w, h = film.w, film.h
    active = np.ones(w * h, bool)
    if first >= min_spp:  # renderer.cpp:442-467: skip a pixel whose 3x3 neighbourhood is clean
        clean = (film.relative_error() < threshold).reshape(h, w)
        pad = np.pad(clean, 1, mode="edge")
        allc = np.ones((h, w), bool)
        for dy in range(3):
            for dx in range(3):
                allc &= pad[dy:dy + h, dx:dx + w]
        active = ~allc.reshape(-1)
    end = first + count
    traced = 0
    for s in range(first, end):
        idx = np.nonzero(active)[0]
        if idx.size == 0:
            break
        film.add_samples(idx, sample_fn(s)[idx])
        traced += idx.size
        if s + 1 < end and s + 1 >= min_spp and (s + 1) % 8 == 0:  # renderer.cpp:500-502
            stop = film.relative_error(idx) < threshold
            active[idx[stop]] = False
    return traced
