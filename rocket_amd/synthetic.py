"""Seeded synthetic camera images for the PRNU workloads.

Every image is an 8-bit grayscale frame of a smooth random scene carrying the
multiplicative sensor fingerprint of its camera:

    I = clamp(round(S * (1 + 0.02 * K_c) + N(0, 2^2)), 0, 255)

* ``K_c ~ N(0, 1)`` per pixel, one field per camera ``c``;
* ``S`` is a sum of 8 random-phase 2-D cosines with integer frequencies
  ``(fx, fy)`` in ``[0..4]^2`` (not both 0), rescaled to ``[40, 215]``;
* image ``i`` belongs to camera ``i % n_cameras`` (round-robin in index order).

Randomness comes from ``numpy.random.SeedSequence(20261202)``: camera ``c``
draws from spawn key ``(0, c)`` and image ``i`` from spawn key ``(1, i)``, so
any image can be regenerated alone and in any process.

The reference has no generator (its tests use unseeded ``java.util.Random``
pixels, ``src/test/java/nl/esciencecenter/common_source_identification/
kernels/filter/AbstractFilterTest.java:28-43``); this one is the workload
definition for the benchmarks and tests of this package.
"""
from __future__ import annotations

import os
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass

import numpy as np

ROOT_SEED = 20261202
N_SINUSOIDS = 8
MAX_FREQ = 4
SCENE_LO, SCENE_HI = 40.0, 215.0
PRNU_GAIN = 0.02
NOISE_SIGMA = 2.0


@dataclass(frozen=True)
class Workload:
    """A named image set: ``count`` images of ``height`` x ``width`` from ``cameras``."""
    name: str
    count: int
    height: int
    width: int
    cameras: int

    @property
    def pixels(self) -> int:
        return self.height * self.width

    @property
    def pairs(self) -> int:
        return self.count * (self.count - 1) // 2


# Image sizes are quoted W x H; arrays are (height, width), row-major.
WORKLOADS = {
    "prnu64_512x512": Workload("prnu64_512x512", 64, 512, 512, 4),
    "ncc1024_2048x1536": Workload("ncc1024_2048x1536", 1024, 1536, 2048, 16),
    "pce256_1024x1024": Workload("pce256_1024x1024", 256, 1024, 1024, 8),
}


def _camera_rng(cam: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence(ROOT_SEED, spawn_key=(0, cam)))


def _image_rng(index: int) -> np.random.Generator:
    return np.random.default_rng(np.random.SeedSequence(ROOT_SEED, spawn_key=(1, index)))


def camera_fingerprint(cam: int, height: int, width: int) -> np.ndarray:
    """K_c ~ N(0,1), float64 (height, width)."""
    return _camera_rng(cam).standard_normal((height, width))


def scene(rng: np.random.Generator, height: int, width: int) -> np.ndarray:
    """Sum of 8 random-phase cosines, rescaled to [40, 215] (float64)."""
    freqs = []
    while len(freqs) < N_SINUSOIDS:
        fx, fy = rng.integers(0, MAX_FREQ + 1, size=2)
        if fx == 0 and fy == 0:
            continue
        freqs.append((int(fx), int(fy)))
    phases = rng.uniform(0.0, 2.0 * np.pi, size=N_SINUSOIDS)
    amps = rng.uniform(0.5, 1.0, size=N_SINUSOIDS)
    x = np.arange(width, dtype=np.float64)
    y = np.arange(height, dtype=np.float64)
    # cos(ax + by + p) = cos(ax) cos(by + p) - sin(ax) sin(by + p): rank-2 per cosine
    left = np.empty((height, 2 * N_SINUSOIDS))
    right = np.empty((2 * N_SINUSOIDS, width))
    for s, ((fx, fy), p, a) in enumerate(zip(freqs, phases, amps)):
        ax = 2.0 * np.pi * fx * x / width
        by = 2.0 * np.pi * fy * y / height + p
        left[:, 2 * s] = a * np.cos(by)
        left[:, 2 * s + 1] = -a * np.sin(by)
        right[2 * s] = np.cos(ax)
        right[2 * s + 1] = np.sin(ax)
    s_img = left @ right
    lo, hi = s_img.min(), s_img.max()
    return SCENE_LO + (SCENE_HI - SCENE_LO) * (s_img - lo) / (hi - lo)


def make_image(index: int, height: int, width: int, n_cameras: int,
               fingerprint: np.ndarray | None = None) -> np.ndarray:
    """Image ``index`` as uint8 (height, width)."""
    cam = index % n_cameras
    k = camera_fingerprint(cam, height, width) if fingerprint is None else fingerprint
    rng = _image_rng(index)
    s_img = scene(rng, height, width)
    noise = rng.standard_normal((height, width)) * NOISE_SIGMA
    img = np.rint(s_img * (1.0 + PRNU_GAIN * k) + noise)
    return np.clip(img, 0, 255).astype(np.uint8)


def _camera_batch(args):
    cam, indices, height, width, n_cameras = args
    k = camera_fingerprint(cam, height, width)
    return [(i, make_image(i, height, width, n_cameras, k)) for i in indices]


def make_images(count: int, height: int, width: int, n_cameras: int,
                workers: int | None = None, out: np.ndarray | None = None) -> np.ndarray:
    """All ``count`` images as uint8 (count, height, width); parallel over cameras."""
    if out is None:
        out = np.empty((count, height, width), dtype=np.uint8)
    jobs = [(c, list(range(c, count, n_cameras)), height, width, n_cameras)
            for c in range(min(n_cameras, count))]
    if workers is None:
        workers = min(16, os.cpu_count() or 1, len(jobs))
    if workers <= 1 or len(jobs) == 1:
        for job in jobs:
            for i, img in _camera_batch(job):
                out[i] = img
        return out
    with ProcessPoolExecutor(max_workers=workers) as pool:
        for batch in pool.map(_camera_batch, jobs):
            for i, img in batch:
                out[i] = img
    return out


def camera_of(index: int, n_cameras: int) -> int:
    return index % n_cameras
