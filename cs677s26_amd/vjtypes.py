"""numpy views of the reference's plain-old-data records.

Byte layouts follow include/vj/common_types.cuh:26-143 of the reference and
the C-ABI header include/vj_api.h of this package:

* FastRect     20 B  {p0, p1, p2, p3 (1-D indices into a (winW+1)-wide
  integral), w}                                (common_types.cuh:26-37)
* HaarFeature  64 B  {r0, r1, r2, rectCount u8, tilted u8, reserved u16}
                                               (common_types.cuh:39-46)
* GpuStump4    16 B  {featureIdx as int32 bits, theta, leftVal, rightVal}
                                               (common_types.cuh:137-143)
* CascadeStage 12 B  {first, ntrees, threshold} (common_types.cuh:117-123)
* Detection    24 B  {x, y, w, h, scaleIdx, score} (common_types.cuh:126-135)
"""
import numpy as np

RECT = np.dtype([("p0", "<i4"), ("p1", "<i4"), ("p2", "<i4"), ("p3", "<i4"), ("w", "<f4")])
FEATURE = np.dtype([("r", RECT, (3,)), ("rectCount", "u1"), ("tilted", "u1"), ("reserved", "<u2")])
STUMP = np.dtype([("feature", "<i4"), ("theta", "<f4"), ("left", "<f4"), ("right", "<f4")], align=True)
STAGE = np.dtype([("first", "<i4"), ("ntrees", "<i4"), ("threshold", "<f4")])
DETECTION = np.dtype([("x", "<i4"), ("y", "<i4"), ("w", "<i4"), ("h", "<i4"),
                      ("scaleIdx", "<i4"), ("score", "<f4")])

assert RECT.itemsize == 20 and FEATURE.itemsize == 64
assert STUMP.itemsize == 16 and STAGE.itemsize == 12 and DETECTION.itemsize == 24

# Feature-pool modes (include/vj/face_vision_engine.cuh:25, FeatureMode).
POOL_BASIC, POOL_CORE, POOL_ALL = 0, 1, 2


def det_tuples(dets):
    """Detections as a list of plain tuples (x, y, w, h, scaleIdx, score-bits)."""
    d = np.asarray(dets, dtype=DETECTION)
    sb = d["score"].view("<u4")
    return [(int(a), int(b), int(c), int(e), int(f), int(g))
            for a, b, c, e, f, g in zip(d["x"], d["y"], d["w"], d["h"], d["scaleIdx"], sb)]
