"""FaceVisionEngine and ModelIO over the C ABI.

Mirrors include/vj/face_vision_engine.cuh:19-196 and include/vj/model_io.hpp
(same method names and argument meaning).  Methods that return vj::Status in
the reference raise ``VJError`` carrying that status here; the constructor
raises where the reference throws (src/face_vision_engine.cu:659-725).
"""
import ctypes as C

import numpy as np

from ._native import DetectStats, IntegralView, check, lib
from .device import DeviceArray, _s
from .vjtypes import DETECTION, FEATURE, STAGE, STUMP

BASIC, CORE, ALL = 0, 1, 2


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return a.ptr
    if isinstance(a, int):
        return a
    return np.ascontiguousarray(a).ctypes.data


class FaceVisionEngine:
    Basic, Core, All = BASIC, CORE, ALL

    def __init__(self, win=(24, 24), maxSamples=1, maxImageW=1920, maxImageH=1080, enableTilted=False,
                 maxDetections=200000):
        h = C.c_void_p()
        check(lib().vj_engine_create(int(win[0]), int(win[1]), int(maxSamples), int(maxImageW), int(maxImageH),
                                     int(bool(enableTilted)), int(maxDetections), C.byref(h)),
              "FaceVisionEngine")
        self.handle = h.value
        self.win = (int(win[0]), int(win[1]))
        self.maxDetections = int(maxDetections)
        self._out = np.empty(0, DETECTION)  # detection buffer reused across calls

    def _out_buffer(self, cap):
        # the library writes the first n entries and only those are read, so
        # the buffer is neither zeroed nor reallocated while it is big enough
        if len(self._out) < cap:
            self._out = np.empty(cap, DETECTION)
        return self._out

    def close(self):
        if getattr(self, "handle", None):
            lib().vj_engine_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- feature pool (src/face_vision_engine.cu:877-1055) -------------------
    @staticmethod
    def countHaarFeatures(win=(24, 24), mode=CORE):
        n = C.c_int()
        check(lib().vj_haar_pool(int(win[0]), int(win[1]), int(mode), None, 0, C.byref(n)), "countHaarFeatures")
        return n.value

    @staticmethod
    def buildHaarFeaturePool(win=(24, 24), mode=CORE):
        n = FaceVisionEngine.countHaarFeatures(win, mode)
        out = np.zeros(n, FEATURE)
        m = C.c_int()
        check(lib().vj_haar_pool(int(win[0]), int(win[1]), int(mode), out.ctypes.data, n, C.byref(m)),
              "buildHaarFeaturePool")
        return out

    def uploadFeaturePool(self, features, stream=None):
        f = np.ascontiguousarray(features, FEATURE)
        check(lib().vj_engine_upload_feature_pool(self.handle, f.ctypes.data, len(f), _s(stream)),
              "uploadFeaturePool")

    def deviceFeaturePool(self):
        p, n = C.c_void_p(), C.c_int()
        check(lib().vj_engine_feature_pool(self.handle, C.byref(p), C.byref(n)), "deviceFeaturePool")
        return p.value, n.value

    # -- integrals (src/face_vision_engine.cu:1057-1162) ----------------------
    def computeIntegralBatchTransposed(self, dGray, imageW, imageH, imageStride, nSamples, computeInvNorm=True,
                                       stream=None):
        check(lib().vj_engine_integral_batch(self.handle, _ptr(dGray), int(imageW), int(imageH), int(imageStride),
                                             int(nSamples), int(bool(computeInvNorm)), _s(stream)),
              "computeIntegralBatchTransposed")

    def integralSetView(self):
        v = IntegralView()
        check(lib().vj_engine_integral_view(self.handle, C.byref(v)), "integralSetView")
        return v

    def integral_to_host(self):
        """(sumT u32 [iiArea][n], sqsumT u64 [iiArea][n], invNorm f32 [n]) of the current set."""
        v = self.integralSetView()
        area, n = v.iiArea, v.nSamples
        s = np.empty((area, n), np.uint32)
        q = np.empty((area, n), np.uint64)
        inv = np.empty(n, np.float32)
        check(lib().vj_memcpy_dtoh(s.ctypes.data, v.sumT, s.nbytes, None), "copy sumT")
        check(lib().vj_memcpy_dtoh(q.ctypes.data, v.sqsumT, q.nbytes, None), "copy sqsumT")
        check(lib().vj_memcpy_dtoh(inv.ctypes.data, v.invNorm, inv.nbytes, None), "copy invNorm")
        return s, q, inv

    # -- cascade (src/face_vision_engine.cu:1176-1354) ------------------------
    def setCascadeModel(self, dFeatures, featureCount, dStumps, stumpCount, stages, stream=None):
        st = np.ascontiguousarray(stages, STAGE)
        check(lib().vj_engine_set_cascade(self.handle, _ptr(dFeatures), int(featureCount), _ptr(dStumps),
                                          int(stumpCount), st.ctypes.data, len(st), _s(stream)),
              "setCascadeModel")

    def clearCascadeModel(self, stream=None):
        check(lib().vj_engine_clear_cascade(self.handle, _s(stream)), "clearCascadeModel")

    def set_cascade_host(self, features, stumps, stages):
        """Uploads host arrays and sets them as the cascade (copies are kept by the engine)."""
        df = DeviceArray.from_host(np.ascontiguousarray(features, FEATURE))
        ds = DeviceArray.from_host(np.ascontiguousarray(stumps, STUMP))
        self.setCascadeModel(df, len(features), ds, len(stumps), stages)
        df.close()
        ds.close()

    # -- detection (src/face_vision_engine.cu:1541-1747) ----------------------
    def detectMultiScale(self, dImage, imageW, imageH, imageStride, scaleFactor=1.25, minNeighbors=4,
                         minObjectSize=24, maxDetections=200000, applyGrouping=True, stream=None):
        cap = max(1, min(int(maxDetections), self.maxDetections))
        out = self._out_buffer(cap)
        n = C.c_int()
        check(lib().vj_engine_detect(self.handle, _ptr(dImage), int(imageW), int(imageH), int(imageStride),
                                     float(scaleFactor), int(minNeighbors), int(minObjectSize), int(maxDetections),
                                     int(bool(applyGrouping)), out.ctypes.data, cap, C.byref(n), _s(stream)),
              "detectMultiScale")
        return out[:min(n.value, cap)].copy()

    def detect_batch(self, dImages, sizes, scaleFactor=1.25, minNeighbors=4, minObjectSize=24,
                     maxDetections=200000, applyGrouping=True, cap=None, stream=None):
        """Detections of several device images in one pipeline; list of arrays."""
        k = len(dImages)
        ptrs = (C.c_void_p * k)(*[_ptr(d) for d in dImages])
        ws = (C.c_int * k)(*[int(s[0]) for s in sizes])
        hs = (C.c_int * k)(*[int(s[1]) for s in sizes])
        strides = (C.c_int * k)(*[int(s[2]) if len(s) > 2 else int(s[0]) for s in sizes])
        counts = (C.c_int * k)()
        if cap is None:
            cap = k * max(1, min(int(maxDetections), self.maxDetections))
        out = self._out_buffer(max(1, cap))
        check(lib().vj_engine_detect_batch(self.handle, k, ptrs, ws, hs, strides, float(scaleFactor),
                                           int(minNeighbors), int(minObjectSize), int(maxDetections),
                                           int(bool(applyGrouping)), out.ctypes.data, cap, counts, _s(stream)),
              "detect_batch")
        res, pos = [], 0
        for i in range(k):
            res.append(out[pos:pos + counts[i]].copy())
            pos += counts[i]
        return res

    def set_timing(self, on=True):
        check(lib().vj_engine_set_timing(self.handle, int(bool(on))), "set_timing")

    def set_graphs(self, on=True):
        """Capture the detect launch sequence in a HIP graph and replay it."""
        check(lib().vj_engine_set_graphs(self.handle, int(bool(on))), "set_graphs")

    def set_autotune(self, on=True):
        """Pick the cascade tile shape and the tail-pass split per batch shape by
        timing candidates; off restores the defaults."""
        check(lib().vj_engine_set_autotune(self.handle, int(bool(on))), "set_autotune")

    def tile_rows(self):
        """Cascade tile rows in use for window steps 1, 2 and 3."""
        a = (C.c_int * 3)()
        check(lib().vj_engine_tile_rows(self.handle, a), "tile_rows")
        return list(a)

    def cascade_split(self):
        """(first deep-pass stage, first tail-pass stage, stages) in use; the tail
        pass is off when the second equals the third."""
        a = (C.c_int * 3)()
        check(lib().vj_engine_cascade_split(self.handle, a), "cascade_split")
        return tuple(a)

    def phase_ms(self):
        """Device ms of the last detect: pyramid, integrals, cascade, compaction."""
        a = (C.c_float * 4)()
        check(lib().vj_engine_phase_ms(self.handle, a), "phase_ms")
        return list(a)

    def count_stages(self, on=True):
        """Counting mode: detect calls also count the windows entering each stage."""
        check(lib().vj_engine_count_stages(self.handle, int(bool(on))), "count_stages")

    def stage_windows(self):
        """Windows that entered each cascade stage in the last counted detect."""
        n = C.c_int()
        check(lib().vj_engine_stage_windows(self.handle, None, 0, C.byref(n)), "stage_windows")
        a = (C.c_int64 * max(1, n.value))()
        check(lib().vj_engine_stage_windows(self.handle, a, n.value, C.byref(n)), "stage_windows")
        return list(a)[:n.value]

    def last_stats(self):
        s = DetectStats()
        check(lib().vj_engine_last_stats(self.handle, C.byref(s)), "last_stats")
        return {f: getattr(s, f) for f, _ in DetectStats._fields_}


def crop_resize(gray, box, outW=24, outH=24):
    """cropResizeTo24 (src/main.cu:245-284): box (x, y, w, h) of a host gray
    image resampled to outW x outH, host C++."""
    g = np.ascontiguousarray(gray, np.uint8)
    h, w = g.shape
    b = np.zeros(1, DETECTION)
    b["x"], b["y"], b["w"], b["h"] = box[:4]
    out = np.zeros((outH, outW), np.uint8)
    check(lib().vj_crop_resize(g.ctypes.data, w, h, w, b.ctypes.data, int(outW), int(outH), out.ctypes.data),
          "cropResize")
    return out


def group_rectangles(dets, minNeighbors=4, eps=0.2):
    """groupRectanglesCpu (src/face_vision_engine.cu:1401-1539), host C++."""
    d = np.ascontiguousarray(dets, DETECTION)
    out = np.zeros(max(1, len(d)), DETECTION)
    n = C.c_int()
    check(lib().vj_group_rectangles(d.ctypes.data if len(d) else None, len(d), int(minNeighbors), float(eps),
                                    out.ctypes.data, C.byref(n)), "groupRectangles")
    return out[:n.value].copy()


def detect_multiscale_cpu(features, stumps, stages, image, win=(24, 24), scaleFactor=1.25, minObjectSize=24,
                          maxDetections=200000, maxScales=0):
    """detectMultiScaleCpu (src/main.cu:684-827): the host leg of
    detect_compare, single threaded.  Raw detections in (level, y, x) order,
    stopping at maxDetections."""
    f = np.ascontiguousarray(features, FEATURE)
    s = np.ascontiguousarray(stumps, STUMP)
    g = np.ascontiguousarray(stages, STAGE)
    img = np.ascontiguousarray(image, np.uint8)
    h, w = img.shape
    out = np.zeros(max(1, int(maxDetections)), DETECTION)
    n = C.c_int()
    check(lib().vj_cpu_detect_raw(f.ctypes.data, len(f), s.ctypes.data, len(s), g.ctypes.data, len(g), int(win[0]),
                                  int(win[1]), img.ctypes.data, w, h, w, float(scaleFactor), int(minObjectSize),
                                  int(maxDetections), int(maxScales), out.ctypes.data, len(out), C.byref(n)),
          "detectMultiScaleCpu")
    return out[:n.value].copy()


# -- ModelIO (src/model_io.cpp) ---------------------------------------------

def saveCascadeModel(path, win, features, stumps, stages):
    f = np.ascontiguousarray(features, FEATURE)
    s = np.ascontiguousarray(stumps, STUMP)
    g = np.ascontiguousarray(stages, STAGE)
    check(lib().vj_model_save(str(path).encode(), int(win[0]), int(win[1]), f.ctypes.data, len(f), s.ctypes.data,
                              len(s), g.ctypes.data, len(g)), "saveCascadeModel")


def loadCascadeModel(path):
    """Returns (win, features, stumps, stages)."""
    ww, wh = C.c_int(), C.c_int()
    nf, ns, ng = C.c_int64(), C.c_int64(), C.c_int64()
    p = str(path).encode()
    check(lib().vj_model_peek(p, C.byref(ww), C.byref(wh), C.byref(nf), C.byref(ns), C.byref(ng)),
          "loadCascadeModel")
    f = np.zeros(nf.value, FEATURE)
    s = np.zeros(ns.value, STUMP)
    g = np.zeros(ng.value, STAGE)
    check(lib().vj_model_load(p, f.ctypes.data, nf.value, s.ctypes.data, ns.value, g.ctypes.data, ng.value),
          "loadCascadeModel")
    return (ww.value, wh.value), f, s, g
