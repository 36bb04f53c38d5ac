"""The benchmark workloads, built from seeds (see synth.py).

* ``detect_640``   -- 640x480, 20 blobs of 24-64 px, 3-stage 1/10/25 cascade.
* ``detect_1080``  -- 1920x1080, 200 blobs of 24-400 px, 38 stages / 6061 stumps
  (1, 10, 25, 25, 50, the rest spread evenly).
* ``training``     -- 10000 + 10000 24x24 patches, Basic pool (162,336 features).
* ``batch64``      -- 64 images, log-uniform area 320x240..3840x2160, aspect 4:3..21:9.

Generated cascades are cached as .bin model files under ``.vjcache/`` at the
repository root (or ``$VJ_CACHE_DIR``); the cache is an accelerator only, a
missing file is regenerated from the seed.
"""
import functools
import os
import zlib

import numpy as np

from . import synth
from .engine import loadCascadeModel, saveCascadeModel

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CACHE = os.environ.get("VJ_CACHE_DIR", os.path.join(ROOT, ".vjcache"))


def cascade(stage_sizes, seed=synth.DEFAULT_SEED):
    crc = zlib.crc32(np.asarray(stage_sizes, np.int64).tobytes())
    key = f"cascade_{len(stage_sizes)}x{sum(stage_sizes)}_s{seed}_{crc:08x}.bin"
    path = os.path.join(CACHE, key)
    if os.path.exists(path):
        _, f, s, g = loadCascadeModel(path)
        return f, s, g
    f, s, g = synth.blob_cascade(tuple(stage_sizes), seed)
    os.makedirs(CACHE, exist_ok=True)
    tmp = path + f".tmp{os.getpid()}"
    saveCascadeModel(tmp, (24, 24), f, s, g)
    os.replace(tmp, path)
    return f, s, g


def detect_640():
    img = synth.detect_image(640, 480, 20, 24, 64)
    return img, cascade((1, 10, 25))


def cascade_38():
    return cascade(synth.spread_stage_sizes(38, 6061))


def detect_1080():
    img = synth.detect_image(1920, 1080, 200, 24, 400)
    return img, cascade_38()


def training(n_pos=10000, n_neg=10000, seed=synth.DEFAULT_SEED):
    return synth.training_set(n_pos, n_neg, seed)


@functools.lru_cache(maxsize=2)
def batch64(seed=synth.DEFAULT_SEED):
    """The 64 batch images, generated once per process (callers copy from
    the arrays, never into them)."""
    sizes = synth.batch_sizes(64, seed=seed)
    return sizes, synth.batch_images(sizes, seed=seed)


def served(n=240, qps=100.0, seed=synth.DEFAULT_SEED):
    """Served detect: Poisson arrivals at `qps`, sizes log-uniform 240p..4K
    (16:9), each request a crop of one seeded 3840x2160 scene at a seeded
    offset.  Returns (arrival seconds, list of u8 image views)."""
    t, sizes = synth.request_stream(n, qps, seed=seed)
    base = synth.detect_image(3840, 2160, 400, 24, 400, seed=seed + 1)
    rng = np.random.default_rng(seed + 2)
    imgs = []
    for w, h in sizes:
        w, h = min(w, 3840), min(h, 2160)
        x0 = int(rng.integers(0, 3840 - w + 1))
        y0 = int(rng.integers(0, 2160 - h + 1))
        imgs.append(base[y0:y0 + h, x0:x0 + w])
    return t, imgs


def train_cli_inputs(n_pos=400, stages=3, seed=synth.DEFAULT_SEED):
    """Inputs of a `cuda_hello train` run, all from seeds: a positive pool of
    n_pos face patches, the stage negative files (10000 patches for stage 0,
    30000 for each later stage -- enough for any top-up), and four mining
    images (blurred noise with eye-banded blobs, 320x240 to 640x480)."""
    rng = np.random.default_rng(seed + 11)
    pos = synth.face_patches(n_pos, rng)
    negs = {0: synth.nonface_patches(10000, rng)}
    for s in range(1, stages):
        negs[s] = synth.nonface_patches(30000, rng)
    sizes = [(640, 480), (480, 360), (400, 300), (320, 240)]
    images = [synth.detect_image(w, h, 12, 24, 90, seed=seed + 20 + k) for k, (w, h) in enumerate(sizes)]
    return pos, negs, images
