"""AdaBoostTrainer over the C ABI: the reference's per-op interface
(include/vj/boosting_core.cuh:20-161) and the fused stage/round form of the
training loop (src/main.cu:1936-2348)."""
import ctypes as C

import numpy as np

from ._native import RoundConfig, Weak, check, lib
from .device import DeviceArray, _s

FEATURE_BEST_SPLIT = np.dtype([("bestErr", "<f4"), ("bestTheta", "<f4"), ("bestSplitPos", "<i4"),
                               ("bestParity", "i1"), ("pad", "i1", (3,))])
TILE_BEST = np.dtype([("bestErr", "<f4"), ("bestTheta", "<f4"), ("featureIdx", "<i4"),
                      ("bestParity", "i1"), ("pad", "i1", (3,))])
assert FEATURE_BEST_SPLIT.itemsize == 16 and TILE_BEST.itemsize == 16


def _p(a):
    if a is None:
        return None
    return a.ptr if isinstance(a, DeviceArray) else int(a)


class AdaBoostTrainer:
    def __init__(self, win=(24, 24), maxSamples=20000, maxFeaturesPerTile=1024):
        h = C.c_void_p()
        check(lib().vj_trainer_create(int(win[0]), int(win[1]), int(maxSamples), int(maxFeaturesPerTile),
                                      C.byref(h)), "AdaBoostTrainer")
        self.handle = h.value
        self.n = 0
        self.nfeat = 0

    def close(self):
        if getattr(self, "handle", None):
            lib().vj_trainer_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- per-op interface ------------------------------------------------------
    def bindTrainingSet(self, iiView, dLabel, dActive, dWeight):
        check(lib().vj_trainer_bind(self.handle, C.byref(iiView), _p(dLabel), _p(dActive), _p(dWeight),
                                    iiView.nSamples), "bindTrainingSet")
        self.n = iiView.nSamples

    def evaluateFeatureResponses(self, dFeatures, featureBegin, featureCount, dOutResp, stream=None):
        check(lib().vj_trainer_evaluate_responses(self.handle, _p(dFeatures), int(featureBegin), int(featureCount),
                                                  _p(dOutResp), _s(stream)), "evaluateFeatureResponses")

    def sortSamplesPerFeature(self, dResp, featureCount, dOutSortedIdx, stream=None):
        check(lib().vj_trainer_sort_per_feature(self.handle, _p(dResp), int(featureCount), _p(dOutSortedIdx),
                                                _s(stream)), "sortSamplesPerFeature")

    def evaluateAndFindThreshold(self, dResp, dSortedIdx, featureCount, dOutBest, stream=None):
        check(lib().vj_trainer_find_threshold(self.handle, _p(dResp), _p(dSortedIdx), int(featureCount),
                                              _p(dOutBest), _s(stream)), "evaluateAndFindThreshold")

    def selectBestInTile(self, dBest, featureBegin, featureCount, dUsedMask, dOut, stream=None):
        check(lib().vj_trainer_select_best(self.handle, _p(dBest), int(featureBegin), int(featureCount),
                                           _p(dUsedMask), _p(dOut), _s(stream)), "selectBestInTile")

    # -- fused stage / round ---------------------------------------------------
    def stage_begin(self, dPatches, labels, numPos, dFeatures, nFeatures, active=None, stream=None):
        labels = np.ascontiguousarray(labels, np.int8)
        act = None if active is None else np.ascontiguousarray(active, np.uint8)
        check(lib().vj_trainer_stage_begin(self.handle, _p(dPatches), labels.ctypes.data,
                                           None if act is None else act.ctypes.data, len(labels), int(numPos),
                                           _p(dFeatures), int(nFeatures), _s(stream)), "stage_begin")
        self.n = len(labels)
        self.nfeat = int(nFeatures)

    def round(self, minHitRate=0.99, maxFalseAlarm=0.40, stream=None):
        w = Weak()
        cfg = RoundConfig(float(minHitRate), float(maxFalseAlarm))
        check(lib().vj_trainer_round(self.handle, C.byref(cfg), C.byref(w), _s(stream)), "round")
        return {f: getattr(w, f) for f, _ in Weak._fields_}

    def stage_end(self, threshold, stream=None):
        n = C.c_int()
        check(lib().vj_trainer_stage_end(self.handle, float(threshold), C.byref(n), _s(stream)), "stage_end")
        return n.value

    # -- sort and scan schedule (no reference counterpart) -------------------
    def set_autotune(self, on=True):
        check(lib().vj_trainer_set_autotune(self.handle, 1 if on else 0), "set_autotune")

    def set_schedule(self, sched=None):
        """Force (wide, ahead, blocks, packed) for the next stages; None: the default."""
        arr = None if sched is None else np.ascontiguousarray(sched, np.int32)
        check(lib().vj_trainer_set_schedule(self.handle, None if arr is None else arr.ctypes.data), "set_schedule")

    def schedule(self):
        out = np.zeros(4, np.int32)
        check(lib().vj_trainer_schedule(self.handle, out.ctypes.data), "schedule")
        return tuple(int(v) for v in out)

    # -- exact pruning of the round's search (no reference counterpart) ------
    def set_pruning(self, on=True):
        check(lib().vj_trainer_set_pruning(self.handle, 1 if on else 0), "set_pruning")

    def pruning_stats(self):
        """Features scanned by the pruned search since the stage began."""
        v = C.c_int64()
        check(lib().vj_trainer_pruning_stats(self.handle, C.byref(v)), "pruning_stats")
        return v.value

    def read_state(self):
        w = np.empty(self.n, np.float64)
        s = np.empty(self.n, np.float32)
        a = np.empty(self.n, np.uint8)
        inv = np.empty(self.n, np.float32)
        check(lib().vj_trainer_read_state(self.handle, w.ctypes.data, s.ctypes.data, a.ctypes.data,
                                          inv.ctypes.data), "read_state")
        return w, s, a, inv

    def read_orders(self, begin, count):
        out = np.empty((count, self.n), np.uint16)
        check(lib().vj_trainer_read_orders(self.handle, int(begin), int(count), out.ctypes.data), "read_orders")
        return out

    def read_splits(self, begin, count):
        """Valid-split flags (0/1) of the presorted orders, per position."""
        out = np.empty((count, self.n), np.uint8)
        check(lib().vj_trainer_read_splits(self.handle, int(begin), int(count), out.ctypes.data), "read_splits")
        return out


def cpu_best_weak(patches, labels, weights, features, active=None, win=(24, 24)):
    """The host leg of train_compare (runCpuWeakSearch, src/main.cu:587-682),
    single threaded, with the search semantics of the GPU per-op path
    (evaluateAndFindThreshold + selectBestInTile): returns a TILE_BEST
    record for features [0, len(features))."""
    p = np.ascontiguousarray(patches, np.uint8)
    lab = np.ascontiguousarray(labels, np.int8)
    w = np.ascontiguousarray(weights, np.float32)
    f = np.ascontiguousarray(features)
    act = None if active is None else np.ascontiguousarray(active, np.uint8)
    out = np.zeros(1, TILE_BEST)
    check(lib().vj_cpu_best_weak(p.ctypes.data, len(lab), int(win[0]), int(win[1]), lab.ctypes.data,
                                 None if act is None else act.ctypes.data, w.ctypes.data, f.ctypes.data, len(f),
                                 out.ctypes.data), "cpuBestWeak")
    return out[0]
