"""ctypes wrapper of oracle/build/libvjoracle.so -- TEST INFRASTRUCTURE.

Imported only by tests/, the smoke entry point and bench.py's cpu_baseline
leg, where it is the checker.  The product (cs677s26_amd) never imports it.
Builds the library on first use when it is missing and gcc is present.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

from cs677s26_amd.vjtypes import DETECTION, FEATURE, STAGE, STUMP

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "libvjoracle.so")
_lib = None

LEVEL = np.dtype([("scaleIdx", "<i4"), ("w", "<i4"), ("h", "<i4"), ("step", "<i4"), ("scale", "<f4")])
FEATURE_BEST = np.dtype([("errQ", "<i8"), ("split", "<i4"), ("parity", "<i4")])


def build(target="all"):
    subprocess.check_call(["make", "-s", "-C", _HERE, target])


def _bind(L):
    P = C.c_void_p
    i, f = C.c_int, C.c_float
    L.or_haar_pool.argtypes = [i, i, i, P, i]
    L.or_integral.argtypes = [P, i, i, i, P, P]
    L.or_train_invnorm.argtypes = [P, P, i, i]
    L.or_train_invnorm.restype = f
    L.or_resize.argtypes = [P, i, i, i, P, i, i]
    L.or_crop_resize.argtypes = [P, i, i, i, i, i, i, i, i, i, P]
    L.or_levels.argtypes = [i, i, i, i, f, i, P, i]
    L.or_detect_raw.argtypes = [P, i, P, i, P, i, i, i, P, i, i, i, f, i, i, P, i, P]
    L.or_group.argtypes = [P, i, i, f, P]
    L.or_train_integrals.argtypes = [P, i, i, i, P, P, P]
    L.or_train_response.argtypes = [P, P, P, i, i]
    L.or_train_response.restype = f
    L.or_train_responses.argtypes = [P, P, P, i, P]
    L.or_presort.argtypes = [P, i, i, P, P, i, P, P]
    L.or_fixed_weight.argtypes = [f]
    L.or_fixed_weight.restype = C.c_int64
    L.or_search.argtypes = [P, P, i, i, P, P]
    L.or_select.argtypes = [P, i, P, P]
    L.or_exact_sum.argtypes = [P, i]
    L.or_exact_sum.restype = C.c_double
    return L


def lib():
    """The oracle library: build/libvjoracle.so, or the file $VJ_ORACLE_LIB
    names (the sanitizer build, for instance)."""
    global _lib
    if _lib is None:
        path = os.environ.get("VJ_ORACLE_LIB") or _LIB_PATH
        if not os.path.exists(path):
            build()
        _lib = _bind(C.CDLL(path))
    return _lib


@contextlib.contextmanager
def use_library(path):
    """Runs the block with every wrapper of this module bound to the oracle
    build at `path` (the planted-mistake builds of `make mutants`)."""
    global _lib
    saved = lib()
    _lib = _bind(C.CDLL(path))
    try:
        yield
    finally:
        _lib = saved


def mutant_path(k):
    """Path of the build with planted mistake k (built on first use)."""
    path = os.path.join(_HERE, "build", f"libvjoracle_m{k}.so")
    if not os.path.exists(path):
        build("mutants")
    return path


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n):
    lib().or_set_threads(int(n))


def haar_pool(win=24, mode=0):
    n = lib().or_haar_pool(win, win, mode, None, 0)
    out = np.zeros(n, FEATURE)
    lib().or_haar_pool(win, win, mode, _p(out), n)
    return out


def integral(img):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    s = np.zeros((h + 1, w + 1), np.uint32)
    q = np.zeros((h + 1, w + 1), np.uint64)
    lib().or_integral(_p(img), w, h, w, _p(s), _p(q))
    return s, q


def resize(img, dw, dh):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    out = np.zeros((dh, dw), np.uint8)
    lib().or_resize(_p(img), w, h, w, _p(out), dw, dh)
    return out


def crop_resize(img, box, out_w=24, out_h=24):
    """cropResizeTo24 (src/main.cu:245-284); box = (x, y, w, h)."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    out = np.zeros((out_h, out_w), np.uint8)
    lib().or_crop_resize(_p(img), w, h, w, int(box[0]), int(box[1]), int(box[2]), int(box[3]), out_w, out_h, _p(out))
    return out


def levels(w, h, sf=1.25, min_obj=24, win=24):
    n = lib().or_levels(w, h, win, win, C.c_float(sf), min_obj, None, 0)
    out = np.zeros(n, LEVEL)
    lib().or_levels(w, h, win, win, C.c_float(sf), min_obj, _p(out), n)
    return out


def detect_raw(features, stumps, stages, img, sf=1.25, min_obj=24, max_det=200000, win=24):
    """Raw detections in CPU order (src/main.cu:684-827) and windows visited."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    features = np.ascontiguousarray(features, FEATURE)
    stumps = np.ascontiguousarray(stumps, STUMP)
    stages = np.ascontiguousarray(stages, STAGE)
    cap = max_det
    out = np.zeros(cap, DETECTION)
    vis = C.c_int64(0)
    n = lib().or_detect_raw(_p(features), len(features), _p(stumps), len(stumps), _p(stages),
                            len(stages), win, win, _p(img), w, h, w, C.c_float(sf), min_obj,
                            max_det, _p(out), cap, C.byref(vis))
    return out[:n].copy(), vis.value


def group(dets, min_neighbors=4, eps=0.2):
    dets = np.ascontiguousarray(dets, DETECTION)
    out = np.zeros(max(1, len(dets)), DETECTION)
    n = lib().or_group(_p(dets), len(dets), min_neighbors, C.c_float(eps), _p(out))
    return out[:n].copy()


def train_integrals(patches):
    patches = np.ascontiguousarray(patches, np.uint8)
    n, h, w = patches.shape
    area = (w + 1) * (h + 1)
    s = np.zeros((area, n), np.uint32)
    q = np.zeros((area, n), np.uint64)
    inv = np.zeros(n, np.float32)
    lib().or_train_integrals(_p(patches), w, h, n, _p(s), _p(q), _p(inv))
    return s, q, inv


def presort(features, sumT, inv, fbegin=0, fcount=None, wide=False):
    features = np.ascontiguousarray(features, FEATURE)
    n = sumT.shape[1]
    if fcount is None:
        fcount = len(features) - fbegin
    if wide:
        out = np.zeros((fcount, n), np.uint32)
        lib().or_presort(_p(features), fbegin, fcount, _p(sumT), _p(inv), n, None, _p(out))
    else:
        out = np.zeros((fcount, n), np.uint16)
        lib().or_presort(_p(features), fbegin, fcount, _p(sumT), _p(inv), n, _p(out), None)
    return out


def responses(features, sumT, inv):
    features = np.ascontiguousarray(features, FEATURE)
    n = sumT.shape[1]
    out = np.zeros((len(features), n), np.float32)
    L = lib()
    for fi in range(len(features)):
        fp = features[fi:fi + 1]
        for s in range(n):
            out[fi, s] = L.or_train_response(_p(fp), _p(sumT), _p(inv), n, s)
    return out


def search(ord_, sw32):
    ord_ = np.ascontiguousarray(ord_)
    sw32 = np.ascontiguousarray(sw32, np.float32)
    fcount, n = ord_.shape
    out = np.zeros(fcount, FEATURE_BEST)
    if ord_.dtype == np.uint16:
        lib().or_search(_p(ord_), None, fcount, n, _p(sw32), _p(out))
    else:
        lib().or_search(None, _p(ord_), fcount, n, _p(sw32), _p(out))
    return out


def select(best, used=None):
    best = np.ascontiguousarray(best, FEATURE_BEST)
    e = C.c_int64(0)
    u = None if used is None else np.ascontiguousarray(used, np.uint8)
    idx = lib().or_select(_p(best), len(best), _p(u), C.byref(e))
    return idx, e.value


def fixed_weight(w):
    return lib().or_fixed_weight(C.c_float(w))


def exact_sum(w):
    w = np.ascontiguousarray(w, np.float64)
    return lib().or_exact_sum(_p(w), len(w))


def feature_responses(feature, sumT, inv):
    """Responses of one feature on every sample (src/main.cu:476-496)."""
    f = np.ascontiguousarray(np.asarray(feature, FEATURE).reshape(1))
    n = sumT.shape[1]
    out = np.zeros(n, np.float32)
    lib().or_train_responses(_p(f), _p(sumT), _p(inv), n, _p(out))
    return out


def _signed_search_weights(w, labels, active):
    f = w.astype(np.float32)
    return np.where(active != 0, np.where(labels > 0, f, -f), np.float32(0)).astype(np.float32)


def _normalise(w, active):
    """ApplyActiveMask + exact sum + divide (src/main.cu:885-906, 1725-1740)."""
    w = np.where(active != 0, w, 0.0)
    s = exact_sum(w)
    if s > 1e-30:
        w = w / s
    return w


def train_stage(patches, labels, num_pos, features, rounds, min_hit=0.99, max_fa=0.40, active=None,
                stop=False):
    """One cascade stage restated from src/main.cu:1936-2348.

    Stage start: integrals, invNorm, weights 0.5/numPos and 0.5/activeNeg
    normalised, strong = 0.  Each round: exact fixed-point search over the
    presorted orders (or_search / or_select), theta = 0.5f*(v_k + v_k+1),
    C = log((1-e')/e') in f64 with e' = clamp(e, 1e-6, 1-1e-6), leaves
    (float)(-+C) by parity, strong += (resp < theta) ? L : R in f32, active
    misclassified w *= exp(C), exact-sum normalisation, calibration by the
    cutIdx-th smallest positive score.  Returns the per-round records and the
    final (w, strong, active).
    """
    import math

    labels = np.asarray(labels, np.int8)
    n = len(labels)
    act = np.ones(n, np.uint8) if active is None else np.array(active, np.uint8)
    act[:num_pos] = 1
    sumT, _sq, inv = train_integrals(patches)
    # above 32,767 samples a 16-bit entry has no room for the split flag: the
    # 32-bit layout (index in bits 0-30, flag in bit 31)
    wide = n > 32767
    idx_mask = 0x7FFFFFFF if wide else 0x7FFF
    ords = presort(features, sumT, inv, wide=wide)
    active_neg = int(act[num_pos:].sum())
    w = np.where(np.arange(n) < num_pos, 0.5 / num_pos, np.where(act != 0, 0.5 / active_neg, 0.0))
    w = _normalise(w, act)
    strong = np.zeros(n, np.float32)
    used = np.zeros(len(features), np.uint8)
    recs = []
    min_pass = max(1, int(math.ceil(np.float32(min_hit) * np.float32(num_pos))))
    cut_idx = num_pos - min_pass
    for _ in range(rounds):
        sw = _signed_search_weights(w, labels, act)
        best = search(ords, sw)
        fi, eq = select(best, used)
        if fi < 0:
            break
        k, parity = int(best[fi]["split"]), int(best[fi]["parity"])
        resp = feature_responses(features[fi], sumT, inv)
        s0, s1 = int(ords[fi, k]) & idx_mask, int(ords[fi, k + 1]) & idx_mask
        theta = np.float32(0.5) * (resp[s0] + resp[s1])
        err = math.ldexp(float(eq), -62)
        ec = min(max(err, 1e-6), 1.0 - 1e-6)
        c = math.log((1.0 - ec) / ec)
        factor = math.exp(c)
        fc = np.float32(c)
        left, right = (-fc, fc) if parity > 0 else (fc, -fc)
        used[fi] = 1
        lt = resp < theta
        strong = (strong + np.where(lt, left, right)).astype(np.float32)
        pred = np.where(lt, -1, 1) if parity > 0 else np.where(lt, 1, -1)
        mis = (act != 0) & (pred != labels)
        w = np.where(mis, w * factor, w)
        w = _normalise(w, act)
        pos = np.sort(strong[:num_pos])
        cut = pos[cut_idx]
        pos_ge = int((strong[:num_pos] >= cut).sum())
        pos_gt = int((strong[:num_pos] > cut).sum())
        neg = strong[num_pos:][act[num_pos:] != 0]
        neg_ge = int((neg >= cut).sum())
        neg_gt = int((neg > cut).sum())
        strict = pos_gt >= min_pass
        thr = np.nextafter(cut, np.float32(np.inf)) if strict else cut
        pos_pass = pos_gt if strict else pos_ge
        neg_surv = neg_gt if strict else neg_ge
        hit = np.float32(pos_pass) / np.float32(num_pos)
        fa = np.float32(neg_surv) / np.float32(active_neg)
        sat = (hit >= np.float32(min_hit) and fa <= np.float32(max_fa)) or neg_surv == 0
        recs.append({"feature": int(fi), "split": k, "parity": parity, "errQ": int(eq), "err": err,
                     "theta": float(theta), "theta_bits": int(np.float32(theta).view(np.uint32)),
                     "alpha": c, "left": float(left), "right": float(right),
                     "threshold_bits": int(np.float32(thr).view(np.uint32)), "posPass": pos_pass,
                     "negSurvive": neg_surv, "satisfied": bool(sat)})
        if stop and sat:
            break
    return recs, w, strong, act
