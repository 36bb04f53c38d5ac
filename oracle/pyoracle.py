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
    L.or_detect_raw.argtypes = [P, i, P, i, P, i, i, i, P, i, i, i, f, i, i, P, i, P, P]
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


_std = None


def std_lib():
    """The standard-library steps of the reference's training loop
    (build/libvjoracle_std.so, from vj_stdlib.cpp)."""
    global _std
    if _std is None:
        path = os.path.join(_HERE, "build", "libvjoracle_std.so")
        if not os.path.exists(path):
            build()
        L = C.CDLL(path)
        L.or_std_shuffle.argtypes = [C.c_int, C.c_uint32, C.c_void_p]
        L.or_std_uniform_ints.argtypes = [C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_void_p]
        L.or_std_partial_sort_desc.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        _std = L
    return _std


def std_shuffle(n, seed):
    out = np.zeros(max(n, 1), np.int32)
    std_lib().or_std_shuffle(int(n), C.c_uint32(seed & 0xFFFFFFFF), out.ctypes.data)
    return out[:n]


def std_uniform_ints(count, lo, hi, seed):
    out = np.zeros(max(count, 1), np.int32)
    std_lib().or_std_uniform_ints(int(count), int(lo), int(hi), C.c_uint32(seed & 0xFFFFFFFF), out.ctypes.data)
    return out[:count]


def std_partial_sort_desc(scores, keep):
    scores = np.ascontiguousarray(scores, np.float32)
    out = np.zeros(max(len(scores), 1), np.int32)
    std_lib().or_std_partial_sort_desc(scores.ctypes.data, len(scores), int(keep), out.ctypes.data)
    return out[:len(scores)]


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


def detect_raw(features, stumps, stages, img, sf=1.25, min_obj=24, max_det=200000, win=24, stage_enter=None):
    """Raw detections in CPU order (src/main.cu:684-827) and windows visited;
    stage_enter (an int64 array of len(stages), or None) receives the windows
    that enter each stage."""
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape
    features = np.ascontiguousarray(features, FEATURE)
    stumps = np.ascontiguousarray(stumps, STUMP)
    stages = np.ascontiguousarray(stages, STAGE)
    cap = max_det
    out = np.zeros(cap, DETECTION)
    vis = C.c_int64(0)
    if stage_enter is not None:
        assert stage_enter.dtype == np.int64 and len(stage_enter) == len(stages)
        stage_enter[:] = 0
    n = lib().or_detect_raw(_p(features), len(features), _p(stumps), len(stumps), _p(stages),
                            len(stages), win, win, _p(img), w, h, w, C.c_float(sf), min_obj,
                            max_det, _p(out), cap, C.byref(vis),
                            None if stage_enter is None else _p(stage_enter))
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


def mine_hard_negatives(feats, stumps, stages, mine_images, stage_seed, neg, scale_factor=1.25, keep_per_image=256):
    """The mining step of a stage (src/main.cu:1800-1898): the images in
    std::shuffle order (std::mt19937(stage_seed)), each from 24x24 to
    2048x2048 scanned by a raw detect with the cascade so far (min size 24,
    at most 20000 boxes), the keep_per_image best boxes by std::partial_sort
    on score cropped and resized to the window (cropResizeTo24) into neg in
    that order, until neg is full.  Returns (count, [(image, box), ...])."""
    k_neg, win = len(neg), 24
    hard, used = 0, []
    for i in std_shuffle(len(mine_images), stage_seed):
        if hard >= k_neg:
            break
        img = np.ascontiguousarray(mine_images[i], np.uint8)
        h, w = img.shape
        if w < win or h < win or w > 2048 or h > 2048:
            continue
        raw, _ = detect_raw(feats, stumps, stages, img, sf=scale_factor, min_obj=win, max_det=max(20000, 2 * k_neg))
        keep = min(len(raw), keep_per_image)
        order = std_partial_sort_desc(raw["score"], keep) if len(raw) else []
        for j in range(keep):
            if hard >= k_neg:
                break
            d = raw[order[j]]
            neg[hard] = crop_resize(img, (d["x"], d["y"], d["w"], d["h"]), win, win)
            used.append((int(i), d.copy()))
            hard += 1
    return hard, used


def train_cascade(positives, stage_negatives, mine_images, max_stages=38, max_weak=250, seed=677,
                  min_hit=0.99, max_fa=0.40, scale_factor=1.25, cand_mul=3, pool_mode=1):
    """The cascade training loop of `cuda_hello train`, restated from
    src/main.cu:1523-1626 and 1760-2392, on data held in memory.

    positives:       the [N][24][24] u8 positive pool (faces_u8.bin).
    stage_negatives: {stage: [M][24][24] u8} -- the stage's negative file
                     (<stageNegCacheDir>/stage_negatives/non_faces_stageNN.bin); a
                     stage that needs negatives must have a file holding enough
                     of them (the sampler that writes a missing file is not
                     restated: it stands in for the reference's Python script).
    mine_images:     the mining images as u8 arrays, in sorted path order.

    Per stage: seed + stage * 1000003; the positives (all, or the first 10000
    of a std::shuffle of the pool when it is larger); stage 0 takes the first
    10000 negatives of its file; later stages mine -- the images in
    std::shuffle order, each no larger than 2048 x 2048 scanned by a raw
    detect with the cascade so far (min size 24, at most 20000 boxes), the
    256 best by std::partial_sort on score cropped and resized to the window
    -- and top up from the first max(10000, missing * cand_mul) patches of
    the stage's file by std::uniform_int_distribution picks; then boosting
    rounds (at most max_weak, within the 6061-stump budget) until the stage
    is satisfied, and the stage is kept with threshold - 1e-5f unless it
    failed with negatives left.  Returns (stumps, stages, per-stage info)."""
    from cs677s26_amd.vjtypes import STAGE, STUMP
    k_pos, k_neg, k_total, k_keep, win = 10000, 10000, 6061, 256, 24
    pos = np.ascontiguousarray(positives, np.uint8)
    pool_n = len(pos)
    num_pos = min(pool_n, k_pos)
    n = num_pos + k_neg
    feats = haar_pool(win, pool_mode)
    labels = np.full(n, -1, np.int8)
    labels[:num_pos] = 1
    stumps, stages, info = [], [], []
    for stage in range(max_stages):
        if len(stumps) >= k_total:
            break
        stage_seed = seed + stage * 1000003
        if pool_n <= num_pos:
            p = pos[:num_pos]
        else:
            p = pos[std_shuffle(pool_n, stage_seed ^ 0x7F4A7C15)[:num_pos]]
        neg = np.zeros((k_neg, win, win), np.uint8)
        hard = 0
        if stage == 0:
            src = stage_negatives[0]
            assert len(src) >= k_neg, "stage 0 needs a negative file of at least 10000 patches"
            neg[:] = src[:k_neg]
            hard = k_neg
        else:
            hard, _ = mine_hard_negatives(feats, np.array(stumps, STUMP), np.array(stages, STAGE), mine_images,
                                          stage_seed, neg, scale_factor)
        mined = hard if stage > 0 else 0
        if hard < k_neg:
            fallback = max(k_neg, (k_neg - hard) * cand_mul)
            src = stage_negatives[stage]
            assert len(src) >= fallback, f"stage {stage} needs a negative file of at least {fallback} patches"
            picks = std_uniform_ints(k_neg - hard, 0, fallback - 1, (stage_seed & 0xFFFFFFFF) ^ 0x5BD1E995)
            neg[hard:] = src[picks]
            hard = k_neg
        patches = np.concatenate([p, neg])
        first = len(stumps)
        limit = min(max_weak, k_total - first)
        recs, _w, strong, _act = train_stage(patches, labels, num_pos, feats, limit, min_hit, max_fa, stop=True)
        for r in recs:
            stumps.append((r["feature"], np.float32(r["theta"]), np.float32(r["left"]), np.float32(r["right"])))
        trees = len(recs)
        if trees == 0:
            break
        thr = np.uint32(recs[-1]["threshold_bits"]).view(np.float32)
        remain = int((strong[num_pos:] >= thr).sum())
        satisfied = bool(recs[-1]["satisfied"])
        info.append({"stage": stage, "mined": mined, "rounds": trees, "remain": remain, "satisfied": satisfied})
        if not satisfied and remain > 0:
            break
        stages.append((first, trees, np.float32(thr - np.float32(1e-5))))
        if remain == 0:
            break
    return np.array(stumps, STUMP), np.array(stages, STAGE), info
