"""Served detect: continuous batching of detect requests over device slots,
admitted by an HBM budget (the ``vj_server_*`` entry points of libvj_hip.so).

No reference counterpart: the reference detects one image per call
(src/face_vision_engine.cu:1541-1747).  Every request's result equals a plain
``FaceVisionEngine.detectMultiScale`` of its image with the server's
parameters.  ``submit`` and ``wait`` release the GIL inside the library, so
one thread can submit at a fixed rate while others wait.
"""
import ctypes as C

import numpy as np

from ._native import ServerConfig, ServerStats, check, lib
from .vjtypes import DETECTION, FEATURE, STAGE, STUMP


class DetectServer:
    def __init__(self, features, stumps, stages, win=(24, 24), max_image=(3840, 2160), slots=2, max_batch=16,
                 hbm_budget=8 << 30, scale_factor=1.25, min_neighbors=4, min_object_size=24,
                 max_detections=200000, apply_grouping=True):
        f = np.ascontiguousarray(features, FEATURE)
        s = np.ascontiguousarray(stumps, STUMP)
        g = np.ascontiguousarray(stages, STAGE)
        self.cfg = ServerConfig(win[0], win[1], max_image[0], max_image[1], slots, max_batch, int(hbm_budget),
                                scale_factor, min_neighbors, min_object_size, max_detections,
                                1 if apply_grouping else 0)
        self.max_detections = max_detections
        h = C.c_void_p()
        check(lib().vj_server_create(C.byref(self.cfg), f.ctypes.data, len(f), s.ctypes.data, len(s),
                                     g.ctypes.data, len(g), C.byref(h)), "vj_server_create")
        self._h = h

    def submit(self, image):
        """Queues a u8 image (H x W, any row stride); returns its ticket."""
        img = np.asarray(image, np.uint8)
        assert img.ndim == 2 and img.strides[1] == 1
        t = C.c_int64()
        check(lib().vj_server_submit(self._h, img.ctypes.data, img.shape[1], img.shape[0], img.strides[0],
                                     C.byref(t)), "vj_server_submit")
        return t.value

    def wait(self, ticket, cap=None):
        """Blocks until the request is done: (detections, latency in ms)."""
        cap = self.max_detections if cap is None else cap
        out = np.zeros(max(1, cap), DETECTION)
        n = C.c_int()
        lat = C.c_double()
        check(lib().vj_server_wait(self._h, ticket, out.ctypes.data, cap, C.byref(n), C.byref(lat)),
              "vj_server_wait")
        return out[:min(n.value, cap)].copy(), lat.value

    def stats(self):
        st = ServerStats()
        check(lib().vj_server_stats_get(self._h, C.byref(st)), "vj_server_stats_get")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        if self._h:
            lib().vj_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
