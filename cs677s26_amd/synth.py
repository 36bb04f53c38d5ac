"""Seeded synthetic inputs for the detect and training workloads.

No network and no datasets: every image, patch set and cascade is generated
from a seed with numpy's PCG64 ``default_rng``, so the same seed yields the
same bytes on any machine with this numpy.  These build *inputs* only; every
result is computed by the HIP library and checked against the oracle.

Workloads (sizes as named by their functions' defaults):

* ``detect_image``   -- uniform noise 0-255, 5x5 box blur, then darker
  elliptical blobs with two dark eye bands.
* ``blob_cascade``   -- a cascade of Haar stumps on the 24x24 window with the
  stage sizes given, thresholds at random quantiles, stage thresholds set to
  pass 99% of synthetic blob windows.
* ``training_set``   -- face-like 24x24 positives (eye, nose and mouth bands,
  +-2 px jitter, Gaussian noise sigma 12) and negatives (blurred noise and
  random gradients).
* ``batch_sizes`` / ``request_stream`` -- uneven image sizes, Poisson arrivals.

The reference's default seed 677 (src/main.cu:1412) is used by default.
"""
import math

import numpy as np

from .vjtypes import FEATURE, STAGE, STUMP

DEFAULT_SEED = 677
WIN = 24


# ----------------------------------------------------------------------------
# images
# ----------------------------------------------------------------------------

def box_blur(img, k):
    """Integer k x k box mean with edge replication (uint8 in, uint8 out)."""
    a = np.asarray(img, dtype=np.int64)
    r = k // 2
    p = np.pad(a, ((r, k - 1 - r), (r, k - 1 - r)), mode="edge")
    c = np.zeros((p.shape[0] + 1, p.shape[1] + 1), dtype=np.int64)
    c[1:, 1:] = p.cumsum(0).cumsum(1)
    h, w = a.shape
    s = c[k:k + h, k:k + w] - c[0:h, k:k + w] - c[k:k + h, 0:w] + c[0:h, 0:w]
    return ((s + (k * k) // 2) // (k * k)).astype(np.uint8)


def render_blob(img, cx, cy, size, aspect, dark, eye_dark):
    """Darken an ellipse of width ``size`` centred at (cx, cy) and two eye bands.

    Works in place on an int16 canvas; coordinates may fall partly outside.
    """
    h, w = img.shape
    a = 0.5 * size
    b = 0.5 * size * aspect
    x0, x1 = max(0, int(cx - a) - 1), min(w, int(cx + a) + 2)
    y0, y1 = max(0, int(cy - b) - 1), min(h, int(cy + b) + 2)
    if x0 >= x1 or y0 >= y1:
        return
    yy, xx = np.mgrid[y0:y1, x0:x1]
    inside = ((xx + 0.5 - cx) / a) ** 2 + ((yy + 0.5 - cy) / b) ** 2 <= 1.0
    img[y0:y1, x0:x1][inside] -= dark
    # eye bands: rows around cy - 0.18*size, two horizontal bars
    ey = cy - 0.18 * size * aspect
    eh = max(1.0, 0.12 * size * aspect)
    for sgn in (-1.0, 1.0):
        ex0 = cx + sgn * 0.08 * size
        ex1 = cx + sgn * 0.36 * size
        lo, hi = min(ex0, ex1), max(ex0, ex1)
        bx0, bx1 = max(0, int(round(lo))), min(w, int(round(hi)))
        by0, by1 = max(0, int(round(ey - 0.5 * eh))), min(h, int(round(ey + 0.5 * eh)))
        if bx0 < bx1 and by0 < by1:
            img[by0:by1, bx0:bx1] -= eye_dark


def detect_image(w=640, h=480, n_blobs=20, blob_min=24, blob_max=64, seed=DEFAULT_SEED):
    """Blurred uniform noise with darker eye-banded elliptical blobs (uint8 h x w)."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    img = box_blur(base, 5).astype(np.int16)
    for _ in range(n_blobs):
        size = math.exp(rng.uniform(math.log(blob_min), math.log(blob_max)))
        cx = rng.uniform(0, w)
        cy = rng.uniform(0, h)
        aspect = rng.uniform(1.0, 1.25)
        dark = int(rng.integers(30, 80))
        eye = int(rng.integers(25, 60))
        render_blob(img, cx, cy, size, aspect, dark, eye)
    return np.clip(img, 0, 255).astype(np.uint8)


def box_blur_batch(imgs, k):
    """box_blur over a stack of equal-size images (n, h, w)."""
    a = np.asarray(imgs, dtype=np.int64)
    r = k // 2
    p = np.pad(a, ((0, 0), (r, k - 1 - r), (r, k - 1 - r)), mode="edge")
    c = np.zeros((p.shape[0], p.shape[1] + 1, p.shape[2] + 1), dtype=np.int64)
    c[:, 1:, 1:] = p.cumsum(1).cumsum(2)
    _, h, w = a.shape
    s = c[:, k:k + h, k:k + w] - c[:, 0:h, k:k + w] - c[:, k:k + h, 0:w] + c[:, 0:h, 0:w]
    return ((s + (k * k) // 2) // (k * k)).astype(np.uint8)


def _noise_patches(n, rng, k=5):
    return box_blur_batch(rng.integers(0, 256, size=(n, WIN, WIN), dtype=np.uint8), k)


def render_blobs_batch(imgs, cx, cy, size, aspect, dark, eye_dark):
    """render_blob applied to each image of an int16 stack (n, h, w) with
    per-image parameters (arrays of length n)."""
    n, h, w = imgs.shape
    yy, xx = np.mgrid[0:h, 0:w].astype(np.float64)
    a = (0.5 * size)[:, None, None]
    b = (0.5 * size * aspect)[:, None, None]
    inside = ((xx + 0.5 - cx[:, None, None]) / a) ** 2 + ((yy + 0.5 - cy[:, None, None]) / b) ** 2 <= 1.0
    imgs -= (inside * dark[:, None, None]).astype(np.int16)
    ey = cy - 0.18 * size * aspect
    eh = np.maximum(1.0, 0.12 * size * aspect)
    by0 = np.clip(np.round(ey - 0.5 * eh), 0, h)[:, None, None]
    by1 = np.clip(np.round(ey + 0.5 * eh), 0, h)[:, None, None]
    rows = (yy >= by0) & (yy < by1)
    for sgn in (-1.0, 1.0):
        e0, e1 = cx + sgn * 0.08 * size, cx + sgn * 0.36 * size
        bx0 = np.clip(np.round(np.minimum(e0, e1)), 0, w)[:, None, None]
        bx1 = np.clip(np.round(np.maximum(e0, e1)), 0, w)[:, None, None]
        band = rows & (xx >= bx0) & (xx < bx1)
        imgs -= (band * eye_dark[:, None, None]).astype(np.int16)


def blob_patches(n, rng):
    """24x24 windows holding one blob that fills the window, on blurred noise."""
    img = _noise_patches(n, rng).astype(np.int16)
    size = rng.uniform(20.0, 25.0, n)
    aspect = rng.uniform(1.0, 1.25, n)
    cx = 12.0 + rng.uniform(-1.0, 1.0, n)
    cy = 12.0 + rng.uniform(-1.0, 1.0, n)
    render_blobs_batch(img, cx, cy, size, aspect, rng.integers(30, 80, n), rng.integers(25, 60, n))
    return np.clip(img, 0, 255).astype(np.uint8)


# ----------------------------------------------------------------------------
# Haar features and responses (generator side; float64 is enough to pick data)
# ----------------------------------------------------------------------------

def make_feature(kind, x, y, dx, dy, win=WIN):
    """One upright feature of the reference's five Basic types (or the three
    Core extras), as a FEATURE record (src/face_vision_engine.cu:618-635,
    933-980 for the rectangles and weights)."""
    iiw = win + 1

    def rect(rx, ry, rw, rh, wt):
        return (rx + iiw * ry, (rx + rw) + iiw * ry, rx + iiw * (ry + rh),
                (rx + rw) + iiw * (ry + rh), wt)

    z = (0, 0, 0, 0, 0.0)
    if kind == 0:    # x2
        rr, n = [rect(x, y, 2 * dx, dy, -1.0), rect(x + dx, y, dx, dy, 2.0), z], 2
    elif kind == 1:  # y2
        rr, n = [rect(x, y, dx, 2 * dy, -1.0), rect(x, y + dy, dx, dy, 2.0), z], 2
    elif kind == 2:  # x3
        rr, n = [rect(x, y, 3 * dx, dy, -1.0), rect(x + dx, y, dx, dy, 2.0), z], 2
    elif kind == 3:  # y3
        rr, n = [rect(x, y, dx, 3 * dy, -1.0), rect(x, y + dy, dx, dy, 2.0), z], 2
    elif kind == 4:  # x2y2
        rr, n = [rect(x, y, 2 * dx, 2 * dy, -1.0), rect(x, y, dx, dy, 2.0),
                 rect(x + dx, y + dy, dx, dy, 2.0)], 3
    elif kind == 5:  # x4
        rr, n = [rect(x, y, 4 * dx, dy, -1.0), rect(x + dx, y, 2 * dx, dy, 2.0), z], 2
    elif kind == 6:  # y4
        rr, n = [rect(x, y, dx, 4 * dy, -1.0), rect(x, y + dy, dx, 2 * dy, 2.0), z], 2
    else:            # centre
        rr, n = [rect(x, y, 3 * dx, 3 * dy, -1.0), rect(x + dx, y + dy, dx, dy, 9.0), z], 2
    f = np.zeros((), dtype=FEATURE)
    for i, r in enumerate(rr):
        f["r"][i] = r
    f["rectCount"] = n
    return f


_EXTENT = {0: (2, 1), 1: (1, 2), 2: (3, 1), 3: (1, 3), 4: (2, 2), 5: (4, 1), 6: (1, 4), 7: (3, 3)}


def random_feature(rng, kinds=5, win=WIN):
    kind = int(rng.integers(0, kinds))
    mx, my = _EXTENT[kind]
    dx = int(rng.integers(1, win // mx + 1))
    dy = int(rng.integers(1, win // my + 1))
    x = int(rng.integers(0, win - mx * dx + 1))
    y = int(rng.integers(0, win - my * dy + 1))
    return make_feature(kind, x, y, dx, dy, win)


def patch_integrals(patches):
    """int64 (n, 25, 25) sum and square-sum tables of uint8 (n, 24, 24) patches."""
    p = patches.astype(np.int64)
    n, h, w = p.shape
    s = np.zeros((n, h + 1, w + 1), np.int64)
    q = np.zeros((n, h + 1, w + 1), np.int64)
    s[:, 1:, 1:] = p.cumsum(1).cumsum(2)
    q[:, 1:, 1:] = (p * p).cumsum(1).cumsum(2)
    return s, q


def patch_invnorm(s, q):
    """Window normaliser on the inset 22x22 rect, as the detector computes it."""
    n = s.shape[0]
    sf = s.reshape(n, -1)
    qf = q.reshape(n, -1)
    iw = s.shape[2]
    p0, p1, p2, p3 = 1 + iw, 23 + iw, 1 + 23 * iw, 23 + 23 * iw
    ss = (sf[:, p0] - sf[:, p1] - sf[:, p2] + sf[:, p3]).astype(np.float64)
    qq = (qf[:, p0] - qf[:, p1] - qf[:, p2] + qf[:, p3]).astype(np.float64)
    nf2 = 484.0 * qq - ss * ss
    return np.where(nf2 > 0, 1.0 / np.sqrt(np.maximum(nf2, 1e-30)), 0.0)


def feature_response(f, s, inv):
    return feature_responses(np.asarray(f, FEATURE).reshape(1), s, inv)[:, 0]


def feature_responses(feats, s, inv):
    """(n, T) normalised responses of T features on n patches (float64)."""
    sf = s.reshape(s.shape[0], -1)
    raw = np.zeros((s.shape[0], len(feats)), np.float64)
    for i in range(3):
        r = feats["r"][:, i]
        v = sf[:, r["p0"]] - sf[:, r["p1"]] - sf[:, r["p2"]] + sf[:, r["p3"]]
        raw += v * r["w"].astype(np.float64)
    return raw * inv[:, None]


def random_features(rng, count, kinds=5, win=WIN):
    out = np.zeros(count, FEATURE)
    for i in range(count):
        out[i] = random_feature(rng, kinds, win)
    return out


# ----------------------------------------------------------------------------
# cascades
# ----------------------------------------------------------------------------

def spread_stage_sizes(n_stages=38, total=6061, head=(1, 10, 25, 25, 50)):
    """Stage sizes: the given head, then the remainder spread evenly."""
    head = list(head)[:n_stages]
    rest = n_stages - len(head)
    left = total - sum(head)
    sizes = head + [left // rest + (1 if i < left % rest else 0) for i in range(rest)]
    assert sum(sizes) == total and len(sizes) == n_stages
    return sizes


def near_miss_patches(n, rng):
    """Background windows a detector must reject: blurred noise, and blobs that
    are off-centre or at the wrong scale for the 24x24 window."""
    img = _noise_patches(n, rng).astype(np.int16)
    kind = np.arange(n) % 3
    size = np.where(kind == 1, rng.uniform(18.0, 28.0, n),
                    np.where(rng.integers(0, 2, n) == 1, rng.uniform(8.0, 15.0, n),
                             rng.uniform(34.0, 60.0, n)))
    off_far = rng.uniform(5.0, 12.0, (n, 2)) * rng.choice([-1.0, 1.0], size=(n, 2))
    off_near = rng.uniform(-3.0, 3.0, (n, 2))
    off = np.where((kind == 1)[:, None], off_far, off_near)
    aspect = rng.uniform(1.0, 1.25, n)
    dark = np.where(kind == 0, 0, rng.integers(30, 80, n))
    eye = np.where(kind == 0, 0, rng.integers(25, 60, n))
    render_blobs_batch(img, 12.0 + off[:, 0], 12.0 + off[:, 1], size, aspect, dark, eye)
    return np.clip(img, 0, 255).astype(np.uint8)


def blob_cascade(stage_sizes=(1, 10, 25), seed=DEFAULT_SEED, n_sample=1500, kinds=5,
                 pass_rate=0.99, candidates=24):
    """Seeded cascade of Haar stumps that prefers blob windows.

    Built stage by stage the way a cascade is trained, on synthetic data:
    the negatives of stage k are near-miss windows (noise, off-centre or
    mis-scaled blobs) that stages 0..k-1 accept.  Each stump draws
    ``candidates`` random features (type, position, size), each with a
    threshold at a random (log-uniform, often extreme) quantile of the blob
    responses, and keeps the one separating blobs from the current negatives
    best.  Leaf values are +-C with C from that stump's balanced error; the
    leaf with more blob mass gets +C.  Each stage threshold passes
    ``pass_rate`` of the blob windows.  Returns (features, stumps, stages)
    record arrays; stump i uses feature i.
    """
    rng = np.random.default_rng(seed)
    pos = blob_patches(n_sample, rng)
    sp, qp = patch_integrals(pos)
    ip = patch_invnorm(sp, qp)
    total = int(sum(stage_sizes))
    feats = np.zeros(total, FEATURE)
    stumps = np.zeros(total, STUMP)
    stages = np.zeros(len(stage_sizes), STAGE)

    def stage_accepts(sj, s, inv):
        a, b = int(stages[sj]["first"]), int(stages[sj]["ntrees"])
        r = feature_responses(feats[a:a + b], s, inv).astype(np.float32)
        st = stumps[a:a + b]
        votes = np.where(r < st["theta"], st["left"], st["right"]).astype(np.float32)
        acc = np.zeros(len(inv), np.float32)
        for t in range(b):             # stump order, float32, as the detector adds
            acc += votes[:, t]
        return acc >= stages[sj]["threshold"]

    sn = np.zeros((0, WIN + 1, WIN + 1), np.int64)
    ineg = np.zeros(0)
    k = 0
    for si, ns in enumerate(stage_sizes):
        if si > 0 and len(ineg):
            keep = stage_accepts(si - 1, sn, ineg)
            sn, ineg = sn[keep], ineg[keep]
        for _try in range(4):          # top up the hard negatives
            if len(ineg) >= n_sample // 4:
                break
            s_new, q_new = patch_integrals(near_miss_patches(n_sample, rng))
            i_new = patch_invnorm(s_new, q_new)
            for sj in range(si):
                keep = stage_accepts(sj, s_new, i_new)
                s_new, i_new = s_new[keep], i_new[keep]
                if len(i_new) == 0:
                    break
            sn, ineg = np.concatenate([sn, s_new]), np.concatenate([ineg, i_new])
        stage_sum = np.zeros(n_sample, np.float32)
        for _ in range(int(ns)):
            cand = random_features(rng, candidates, kinds)
            rp_all = feature_responses(cand, sp, ip)
            rn_all = feature_responses(cand, sn, ineg) if len(ineg) else np.zeros((1, candidates))
            qs = np.exp(rng.uniform(math.log(0.003), math.log(0.5), size=candidates))
            flip = rng.integers(0, 2, size=candidates).astype(bool)
            qs = np.where(flip, 1.0 - qs, qs)
            srt = np.sort(rp_all, axis=0)
            rank = np.minimum((qs * n_sample).astype(np.int64), n_sample - 1)
            best = None
            for c_i in range(candidates):
                rp = rp_all[:, c_i]
                theta = np.float32(srt[rank[c_i], c_i])
                pl, nl = np.mean(rp < theta), np.mean(rn_all[:, c_i] < theta)
                gap = pl - nl
                if best is None or abs(gap) > abs(best[0]):
                    best = (gap, cand[c_i], theta, rp, pl, nl)
            gap, f, theta, rp, pl, nl = best
            err = 0.5 * ((1.0 - pl) + nl) if gap >= 0 else 0.5 * (pl + (1.0 - nl))
            err = min(max(err, 1e-3), 0.499)
            c = np.float32(0.5 * math.log((1.0 - err) / err) * rng.uniform(0.8, 1.2))
            left, right = (c, -c) if gap >= 0 else (-c, c)
            feats[k] = f
            stumps[k] = (k, theta, left, right)
            stage_sum += np.where(rp.astype(np.float32) < theta, np.float32(left), np.float32(right))
            k += 1
        thr = np.sort(stage_sum)[int((1.0 - pass_rate) * n_sample)] - np.float32(1e-4)
        stages[si] = (k - int(ns), int(ns), thr)
    return feats, stumps, stages


# ----------------------------------------------------------------------------
# training patches
# ----------------------------------------------------------------------------

def face_patches(n, rng):
    out = np.empty((n, WIN, WIN), dtype=np.uint8)
    for i in range(n):
        t = np.full((WIN + 4, WIN + 4), rng.uniform(120, 180), dtype=np.float64)
        o = 2
        eye = rng.uniform(50, 80)
        t[o + 7:o + 10, o + 4:o + 10] -= eye
        t[o + 7:o + 10, o + 14:o + 20] -= eye
        t[o + 10:o + 16, o + 10:o + 14] += rng.uniform(20, 40)
        t[o + 17:o + 20, o + 7:o + 17] -= rng.uniform(30, 60)
        jx, jy = rng.integers(-2, 3, size=2)
        win = t[o + jy:o + jy + WIN, o + jx:o + jx + WIN] + rng.normal(0.0, 12.0, size=(WIN, WIN))
        out[i] = np.clip(np.rint(win), 0, 255).astype(np.uint8)
    return out


def nonface_patches(n, rng):
    out = np.empty((n, WIN, WIN), dtype=np.uint8)
    yy, xx = np.mgrid[0:WIN, 0:WIN].astype(np.float64) - 11.5
    for i in range(n):
        if i % 2 == 0:
            out[i] = box_blur(rng.integers(0, 256, size=(WIN, WIN), dtype=np.uint8), 3)
        else:
            a = rng.uniform(40, 215)
            gx, gy = rng.uniform(-6, 6, size=2)
            g = a + gx * xx + gy * yy + rng.normal(0.0, 4.0, size=(WIN, WIN))
            out[i] = np.clip(np.rint(g), 0, 255).astype(np.uint8)
    return out


def training_set(n_pos=10000, n_neg=10000, seed=DEFAULT_SEED):
    """(patches uint8 [n_pos+n_neg, 24, 24], labels int8 +1/-1), positives first."""
    rng = np.random.default_rng(seed)
    pos = face_patches(n_pos, rng)
    neg = nonface_patches(n_neg, rng)
    labels = np.concatenate([np.ones(n_pos, np.int8), -np.ones(n_neg, np.int8)])
    return np.ascontiguousarray(np.concatenate([pos, neg])), labels


# ----------------------------------------------------------------------------
# batches and request streams
# ----------------------------------------------------------------------------

def batch_sizes(n=64, amin=320 * 240, amax=3840 * 2160, rmin=4 / 3, rmax=21 / 9,
                seed=DEFAULT_SEED):
    """(w, h) pairs with log-uniform area and uniform aspect ratio."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        area = math.exp(rng.uniform(math.log(amin), math.log(amax)))
        asp = rng.uniform(rmin, rmax)
        w = int(round(math.sqrt(area * asp)))
        h = int(round(area / w))
        out.append((max(w, WIN), max(h, WIN)))
    return out


def batch_images(sizes, seed=DEFAULT_SEED, blobs_per_mpix=40):
    out = []
    for i, (w, h) in enumerate(sizes):
        nb = max(1, int(round(blobs_per_mpix * w * h / 1e6)))
        out.append(detect_image(w, h, nb, 24, max(25, min(400, min(w, h) // 2)), seed + 7919 * (i + 1)))
    return out


def request_stream(n, qps, amin=426 * 240, amax=3840 * 2160, seed=DEFAULT_SEED):
    """Poisson arrivals (seconds from t=0) and log-uniform 240p..4K sizes."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / qps, size=n)
    t = np.cumsum(gaps) - gaps[0]
    sizes = []
    for _ in range(n):
        area = math.exp(rng.uniform(math.log(amin), math.log(amax)))
        asp = 16 / 9
        w = int(round(math.sqrt(area * asp)))
        sizes.append((w, max(WIN, int(round(area / w)))))
    return t, sizes
