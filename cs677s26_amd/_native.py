"""ctypes binding of libvj_hip.so (the C ABI declared in include/vj_api.h).

The library is built in-tree by ``__graft_entry__.build()`` (or ``make -C
cs677s26_amd/csrc``).  There is no fallback: if the shared library is missing
or does not load, every entry point raises ``NativeUnavailable``.
"""
import ctypes as C
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
# $VJ_LIB names another build of the same library (the sanitizer build)
LIB_PATH = os.environ.get("VJ_LIB") or os.path.join(_HERE, "libvj_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(_HERE), "include", "vj_api.h")

VJ_OK, VJ_INVALID_ARG, VJ_DEVICE_ERROR, VJ_OUT_OF_MEMORY, VJ_RUNTIME_ERROR = range(5)
STATUS_NAMES = {0: "kOk", 1: "kInvalidArg", 2: "kCudaError", 3: "kOutOfMemory", 4: "kRuntimeError"}


class NativeUnavailable(RuntimeError):
    """libvj_hip.so is not built or cannot be loaded."""


class VJError(RuntimeError):
    def __init__(self, status, where, message):
        super().__init__(f"{where} failed: {STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


P, I, F, D, I64, SZ = C.c_void_p, C.c_int, C.c_float, C.c_double, C.c_int64, C.c_size_t
PP = C.POINTER(C.c_void_p)


class IntegralView(C.Structure):
    _fields_ = [("nSamples", C.c_int32), ("iiWidth", C.c_int32), ("iiHeight", C.c_int32),
                ("iiArea", C.c_int32), ("sumT", P), ("sqsumT", P), ("invNorm", P)]


class DetectStats(C.Structure):
    _fields_ = [("levels", C.c_int32), ("windows", C.c_int64), ("rawDetections", C.c_int64),
                ("pyramidPixels", C.c_int64), ("integralBytes", C.c_int64), ("cascadeBytes", C.c_int64),
                ("imagesGroupedOnHost", C.c_int32)]


class Weak(C.Structure):
    _fields_ = [("feature", C.c_int32), ("theta", C.c_float), ("parity", C.c_int32), ("split", C.c_int32),
                ("err", C.c_double), ("alpha", C.c_double), ("left", C.c_float), ("right", C.c_float),
                ("stageThreshold", C.c_float), ("hitRate", C.c_float), ("falseAlarm", C.c_float),
                ("posPass", C.c_int32), ("negSurvive", C.c_int32), ("stageSatisfied", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("winW", C.c_int32), ("winH", C.c_int32), ("maxImageW", C.c_int32), ("maxImageH", C.c_int32),
                ("slots", C.c_int32), ("maxBatch", C.c_int32), ("hbmBudget", C.c_int64),
                ("scaleFactor", C.c_float), ("minNeighbors", C.c_int32), ("minObjectSize", C.c_int32),
                ("maxDetections", C.c_int32), ("applyGrouping", C.c_int32)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("rejected", C.c_int64),
                ("batches", C.c_int64), ("peakBytesInFlight", C.c_int64), ("maxSlotsBusy", C.c_int32),
                ("maxBatchSeen", C.c_int32)]


class RoundConfig(C.Structure):
    _fields_ = [("minHitRate", C.c_float), ("maxFalseAlarm", C.c_float)]


_PROTOS = {
    "vj_last_error": (C.c_char_p, []),
    "vj_device_count": (I, [C.POINTER(I)]),
    "vj_device_name": (I, [C.c_char_p, I]),
    "vj_malloc": (I, [PP, SZ]),
    "vj_free": (I, [P]),
    "vj_host_alloc": (I, [PP, SZ]),
    "vj_host_free": (I, [P]),
    "vj_memcpy_htod": (I, [P, P, SZ, P]),
    "vj_memcpy_dtoh": (I, [P, P, SZ, P]),
    "vj_memcpy_dtod": (I, [P, P, SZ, P]),
    "vj_memset": (I, [P, I, SZ, P]),
    "vj_stream_create": (I, [PP]),
    "vj_stream_destroy": (I, [P]),
    "vj_stream_synchronize": (I, [P]),
    "vj_device_synchronize": (I, []),
    "vj_event_create": (I, [PP]),
    "vj_event_destroy": (I, [P]),
    "vj_event_record": (I, [P, P]),
    "vj_event_synchronize": (I, [P]),
    "vj_event_elapsed_ms": (I, [P, P, C.POINTER(F)]),
    "vj_haar_pool": (I, [I, I, I, P, I, C.POINTER(I)]),
    "vj_model_save": (I, [C.c_char_p, I, I, P, I64, P, I64, P, I64]),
    "vj_model_peek": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I), C.POINTER(I64), C.POINTER(I64),
                          C.POINTER(I64)]),
    "vj_model_load": (I, [C.c_char_p, P, I64, P, I64, P, I64]),
    "vj_group_rectangles": (I, [P, I, I, F, P, C.POINTER(I)]),
    "vj_group_rectangles_many": (I, [P, P, I, I, F, P, P]),
    "vj_pgm_peek": (I, [C.c_char_p, C.POINTER(I), C.POINTER(I)]),
    "vj_pgm_load": (I, [C.c_char_p, P, SZ]),
    "vj_samples_count": (I, [C.c_char_p, I, C.POINTER(I)]),
    "vj_samples_load": (I, [C.c_char_p, I, P, SZ, C.POINTER(I)]),
    "vj_pgm_list": (I, [C.c_char_p, C.c_char_p, SZ, C.POINTER(SZ), C.POINTER(I)]),
    "vj_ppm_save_gray_boxes": (I, [C.c_char_p, P, I, I, P, I]),
    "vj_crop_resize": (I, [P, I, I, I, P, I, I, P]),
    "vj_engine_create": (I, [I, I, I, I, I, I, I, PP]),
    "vj_engine_destroy": (None, [P]),
    "vj_engine_upload_feature_pool": (I, [P, P, I, P]),
    "vj_engine_feature_pool": (I, [P, PP, C.POINTER(I)]),
    "vj_engine_integral_batch": (I, [P, P, I, I, I, I, I, P]),
    "vj_engine_integral_view": (I, [P, C.POINTER(IntegralView)]),
    "vj_integral_image_noalloc": (I, [P, I, I, I, P, P, P]),
    "vj_engine_set_cascade": (I, [P, P, I, P, I, P, I, P]),
    "vj_engine_clear_cascade": (I, [P, P]),
    "vj_engine_detect": (I, [P, P, I, I, I, F, I, I, I, I, P, I, C.POINTER(I), P]),
    "vj_engine_last_stats": (I, [P, C.POINTER(DetectStats)]),
    "vj_engine_set_timing": (I, [P, I]),
    "vj_engine_phase_ms": (I, [P, C.POINTER(F)]),
    "vj_engine_count_stages": (I, [P, I]),
    "vj_engine_stage_windows": (I, [P, P, I, C.POINTER(I)]),
    "vj_engine_detect_batch": (I, [P, I, P, P, P, P, F, I, I, I, I, P, I, P, P]),
    "vj_engine_detect_footprint": (I, [P, I, I, F, I, I, C.POINTER(I64)]),
    "vj_engine_set_graphs": (I, [P, I]),
    "vj_engine_set_autotune": (I, [P, I]),
    "vj_engine_tile_rows": (I, [P, C.POINTER(I)]),
    "vj_engine_cascade_split": (I, [P, C.POINTER(I)]),
    "vj_server_create": (I, [C.POINTER(ServerConfig), P, I, P, I, P, I, PP]),
    "vj_server_submit": (I, [P, P, I, I, I, C.POINTER(I64)]),
    "vj_server_wait": (I, [P, I64, P, I, C.POINTER(I), C.POINTER(D)]),
    "vj_server_stats_get": (I, [P, C.POINTER(ServerStats)]),
    "vj_server_destroy": (None, [P]),
    "vj_trainer_create": (I, [I, I, I, I, PP]),
    "vj_trainer_destroy": (None, [P]),
    "vj_trainer_bind": (I, [P, C.POINTER(IntegralView), P, P, P, I]),
    "vj_trainer_evaluate_responses": (I, [P, P, I, I, P, P]),
    "vj_trainer_sort_per_feature": (I, [P, P, I, P, P]),
    "vj_trainer_find_threshold": (I, [P, P, P, I, P, P]),
    "vj_trainer_select_best": (I, [P, P, I, I, P, P, P]),
    "vj_trainer_stage_begin": (I, [P, P, P, P, I, I, P, I, P]),
    "vj_trainer_round": (I, [P, C.POINTER(RoundConfig), C.POINTER(Weak), P]),
    "vj_trainer_stage_end": (I, [P, F, C.POINTER(I), P]),
    "vj_trainer_read_state": (I, [P, P, P, P, P]),
    "vj_trainer_read_orders": (I, [P, I, I, P]),
    "vj_trainer_read_splits": (I, [P, I, I, P]),
    "vj_trainer_set_autotune": (I, [P, I]),
    "vj_trainer_set_schedule": (I, [P, P]),
    "vj_trainer_schedule": (I, [P, P]),
    "vj_trainer_set_pruning": (I, [P, I]),
    "vj_trainer_pruning_stats": (I, [P, P]),
    "vj_cpu_detect_raw": (I, [P, I, P, I, P, I, I, I, P, I, I, I, F, I, I, I, P, I, C.POINTER(I)]),
    "vj_cpu_best_weak": (I, [P, I, I, I, P, P, P, P, I, P]),
}

_lib = None


def header_symbols():
    """Function names declared in include/vj_api.h."""
    text = open(HEADER_PATH).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vj_[a-z0-9_]+)\s*\(", text)))


def lib():
    """The loaded library; raises NativeUnavailable when it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise NativeUnavailable(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeUnavailable(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in _PROTOS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status, where):
    if status != VJ_OK:
        msg = lib().vj_last_error()
        raise VJError(status, where, msg.decode(errors="replace") if msg else "")


def device_count():
    n = I(0)
    check(lib().vj_device_count(C.byref(n)), "vj_device_count")
    return n.value
