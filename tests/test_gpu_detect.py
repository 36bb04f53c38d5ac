"""GPU parity of the detect path against the oracle (bit-exact).

Integrals (u32 sum, u64 square sum), the pyramid, the cascade and the
order-preserving compaction are checked through the C ABI against the C
restatement of the reference CPU path (src/main.cu:428-827) and the host
grouping (src/face_vision_engine.cu:1401-1539).
"""
import numpy as np
import pytest

from cs677s26_amd import synth
from cs677s26_amd.device import DeviceArray
from cs677s26_amd.engine import FaceVisionEngine
from cs677s26_amd.vjtypes import STAGE, det_tuples
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def cascade3():
    return synth.blob_cascade((1, 10, 25))


def _engine(w, h, cascade, max_det=200000):
    eng = FaceVisionEngine((24, 24), 1, w, h, False, max_det)
    eng.set_cascade_host(*cascade)
    return eng


@pytest.mark.parametrize("w,h", [(24, 24), (65, 70), (641, 479), (1920, 1080), (257, 1031), (16500, 130),
                                 (130, 9000)])
def test_integral_single_image(w, h):
    say(f"integral single {w}x{h}")
    rng = np.random.default_rng(w * 7 + h)
    img = rng.integers(0, 256, size=(h, w), dtype=np.uint8)
    eng = FaceVisionEngine((24, 24), 1, w, h)
    d = DeviceArray.from_host(img)
    eng.computeIntegralBatchTransposed(d, w, h, w, 1, True)
    s, q, inv = eng.integral_to_host()
    es, eq = O.integral(img)
    assert np.array_equal(s.reshape(h + 1, w + 1), es)
    assert np.array_equal(q.reshape(h + 1, w + 1), eq)


def test_integral_all_255_4k():
    w, h = 3840, 2160
    say("integral all-255 4K")
    img = np.full((h, w), 255, np.uint8)
    eng = FaceVisionEngine((24, 24), 1, w, h)
    d = DeviceArray.from_host(img)
    eng.computeIntegralBatchTransposed(d, w, h, w, 1, False)
    s, q, _ = eng.integral_to_host()
    s = s.reshape(h + 1, w + 1)
    q = q.reshape(h + 1, w + 1)
    yy, xx = np.mgrid[0:h + 1, 0:w + 1]
    assert np.array_equal(s.astype(np.int64), 255 * yy.astype(np.int64) * xx)
    assert np.array_equal(q.astype(np.int64), 65025 * yy.astype(np.int64) * xx)


def test_integral_batch_transposed():
    say("integral batch 24x24 x 3000")
    patches, _ = synth.training_set(1500, 1500, seed=3)
    n = len(patches)
    eng = FaceVisionEngine((24, 24), n, 24, 24)
    d = DeviceArray.from_host(patches)
    eng.computeIntegralBatchTransposed(d, 24, 24, 24, n, True)
    s, q, inv = eng.integral_to_host()
    es, eq, einv = O.train_integrals(patches)
    assert np.array_equal(s, es) and np.array_equal(q, eq)
    assert np.array_equal(inv.view(np.uint32), einv.view(np.uint32))


def test_detect_config1_raw_and_grouped(cascade3):
    say("detect 640x480 3-stage raw+grouped")
    img = synth.detect_image(640, 480, 20, 24, 64)
    eng = _engine(640, 480, cascade3)
    d = DeviceArray.from_host(img)
    raw = eng.detectMultiScale(d, 640, 480, 640, 1.25, 4, 24, 200000, applyGrouping=False)
    want, _ = O.detect_raw(*cascade3, img)
    assert len(raw) == len(want) > 0
    assert det_tuples(raw) == det_tuples(want)
    grouped = eng.detectMultiScale(d, 640, 480, 640, 1.25, 4, 24, 200000, applyGrouping=True)
    assert det_tuples(grouped) == det_tuples(O.group(want, 4))
    assert eng.last_stats()["imagesGroupedOnHost"] == 0  # grouped on the device
    st = eng.last_stats()
    assert st["windows"] == 653485 and st["levels"] == 14


@pytest.mark.parametrize("w,h", [(20000, 200), (200, 12000)])
def test_detect_sides_above_8192(cascade3, w, h):
    """Images wider or taller than 8192 (the integral's chunk carries and
    band prefixes in blocks; the sweep's x bins cover 4 columns each at
    20,000 wide): raw and grouped boxes equal the oracle's."""
    say(f"detect {w}x{h} 3-stage raw+grouped")
    img = synth.detect_image(w, h, 60, 24, 96, seed=w + h)
    eng = _engine(w, h, cascade3)
    d = DeviceArray.from_host(img)
    raw = eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, applyGrouping=False)
    want, _ = O.detect_raw(*cascade3, img)
    assert len(want) > 100
    assert det_tuples(raw) == det_tuples(want)
    grouped = eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, applyGrouping=True)
    assert det_tuples(grouped) == det_tuples(O.group(want, 4))


def test_detect_refuses_inexact_integral_sizes(cascade3):
    """w * h * 255 must stay below 2^31 (exact 32-bit integral sums)."""
    from cs677s26_amd._native import VJError
    w, h = 30000, 300  # 30000 * 300 * 255 > 2^31
    eng = _engine(w, h, cascade3)
    d = DeviceArray.from_host(np.zeros((h, w), np.uint8))
    with pytest.raises(VJError):
        eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, applyGrouping=False)


def test_detect_truncates_in_scan_order(cascade3):
    say("detect maxDetections truncation")
    img = synth.detect_image(640, 480, 20, 24, 64)
    feats, stumps, stages = cascade3
    eng = _engine(640, 480, cascade3, max_det=1000)
    d = DeviceArray.from_host(img)
    raw = eng.detectMultiScale(d, 640, 480, 640, 1.25, 4, 24, 1000, applyGrouping=False)
    want, _ = O.detect_raw(feats, stumps, stages, img, max_det=1000)
    assert len(want) == 1000
    assert det_tuples(raw) == det_tuples(want)


def test_detect_edge_cases(cascade3):
    say("detect edge cases")
    feats, stumps, stages = cascade3
    eng = _engine(700, 500, cascade3)
    # constant image: every window fails the variance gate
    flat = np.full((200, 300), 128, np.uint8)
    d = DeviceArray.from_host(flat)
    assert len(eng.detectMultiScale(d, 300, 200, 300, 1.25, 4, 24, 200000, False)) == 0
    # image smaller than the window: no levels
    tiny = np.zeros((20, 30), np.uint8)
    d = DeviceArray.from_host(tiny)
    assert len(eng.detectMultiScale(d, 30, 20, 30, 1.25, 4, 24, 200000, False)) == 0
    # row stride wider than the image, and a large minObjectSize (level 0 skipped)
    img = synth.detect_image(333, 250, 8, 24, 60, seed=5)
    padded = np.zeros((250, 341), np.uint8)
    padded[:, :333] = img
    d = DeviceArray.from_host(padded)
    for min_obj, sf in ((24, 1.25), (40, 1.1), (31, 1.5)):
        raw = eng.detectMultiScale(d, 333, 250, 341, sf, 4, min_obj, 200000, False)
        want, _ = O.detect_raw(feats, stumps, stages, img, sf=sf, min_obj=min_obj)
        assert det_tuples(raw) == det_tuples(want), (min_obj, sf)
    # threshold -inf everywhere: exactly the windows passing the variance gate
    st_open = np.array(stages, STAGE)
    st_open["threshold"] = -np.inf
    eng2 = _engine(333, 250, (feats, stumps, st_open))
    raw = eng2.detectMultiScale(d, 333, 250, 341, 1.25, 4, 24, 10 ** 6, False)
    want, vis = O.detect_raw(feats, stumps, st_open, img, max_det=10 ** 6)
    assert det_tuples(raw) == det_tuples(want)
    st_shut = np.array(stages, STAGE)
    st_shut["threshold"] = np.inf
    eng3 = _engine(333, 250, (feats, stumps, st_shut))
    assert len(eng3.detectMultiScale(d, 333, 250, 341, 1.25, 4, 24, 200000, False)) == 0


def test_detect_large_levels_with_unaligned_rows(cascade3):
    """Levels large enough for the pyramid's per-depth launches (at least 512 K
    pixels at a depth), read from rows that start at odd addresses: an odd
    width, then the same image in a buffer with an odd wider stride, then a
    minimum object size that skips level 0 (the first level resized straight
    from the input at a larger ratio).  Raw boxes equal the oracle's."""
    say("detect 1283x777, odd strides, per-depth resize launches")
    feats, stumps, stages = cascade3
    w, h = 1283, 777
    img = synth.detect_image(w, h, 60, 24, 90, seed=17)
    eng = _engine(w, h, cascade3)
    padded = np.zeros((h, w + 8), np.uint8)
    padded[:, :w] = img
    for buf, stride in ((img, w), (padded, w + 8)):
        d = DeviceArray.from_host(buf)
        for min_obj, sf in ((24, 1.25), (31, 1.5)):
            raw = eng.detectMultiScale(d, w, h, stride, sf, 4, min_obj, 200000, False)
            want, _ = O.detect_raw(feats, stumps, stages, img, sf=sf, min_obj=min_obj)
            assert det_tuples(raw) == det_tuples(want), (stride, min_obj, sf)


def test_detect_batch_equals_single(cascade3):
    say("detect batch of 6 uneven images")
    sizes = [(640, 480), (333, 250), (1000, 431), (24, 24), (800, 600), (97, 201)]
    imgs = [synth.detect_image(w, h, max(1, w * h // 15000), 24, 64, seed=100 + i) for i, (w, h) in enumerate(sizes)]
    eng = _engine(1000, 600, cascade3)
    dev = [DeviceArray.from_host(im) for im in imgs]
    for grouping in (False, True):
        got = eng.detect_batch(dev, [(w, h, w) for w, h in sizes], 1.25, 4, 24, 200000, grouping)
        for im, g, d, (w, h) in zip(imgs, got, dev, sizes):
            single = eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, grouping)
            assert det_tuples(g) == det_tuples(single)
            want, _ = O.detect_raw(*cascade3, im)
            if grouping:
                want = O.group(want, 4)
            assert det_tuples(g) == det_tuples(want)


def test_detect_config3_golden():
    """1920x1080, 38-stage / 6061-stump cascade: raw and grouped boxes equal
    the oracle's (golden fixture written by tests/golden/make_golden.py)."""
    import os
    say("detect 1080p 38-stage vs golden")
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "detect_c3.npz"))
    from cs677s26_amd import workloads
    img, casc = workloads.detect_1080()
    eng = _engine(1920, 1080, casc)
    d = DeviceArray.from_host(img)
    raw = eng.detectMultiScale(d, 1920, 1080, 1920, 1.25, 4, 24, 200000, applyGrouping=False)
    assert det_tuples(raw) == det_tuples(g["raw"])
    assert eng.last_stats()["windows"] == int(g["windows"]) == 4729998
    grouped = eng.detectMultiScale(d, 1920, 1080, 1920, 1.25, 4, 24, 200000, applyGrouping=True)
    assert det_tuples(grouped) == det_tuples(g["grouped"])
    assert eng.last_stats()["imagesGroupedOnHost"] == 0  # grouped on the device


def test_stage_windows_equal_oracle():
    """Counting mode counts the windows entering each stage (first, deep and
    tail passes of the 38-stage cascade): equal to the oracle's counts for
    one image and, summed, for a batch of two."""
    from cs677s26_amd import workloads
    say("stage window counts, 38-stage cascade, 640x480 and 800x600 crops")
    img, casc = workloads.detect_1080()
    crops = [np.ascontiguousarray(img[300:780, 600:1240]), np.ascontiguousarray(img[100:700, 1000:1800])]
    eng = _engine(800, 600, casc)
    eng.count_stages(True)
    want = []
    for im in crops:
        cnt = np.zeros(len(casc[2]), np.int64)
        O.detect_raw(*casc, im, stage_enter=cnt)
        want.append(cnt)
    assert want[0][-1] > 0 and want[0][14] > 0  # windows reach the tail pass
    d = DeviceArray.from_host(crops[0])
    eng.detectMultiScale(d, 640, 480, 640, 1.25, 4, 24, 200000, applyGrouping=False)
    assert eng.stage_windows() == want[0].tolist()
    dev = [DeviceArray.from_host(im) for im in crops]
    eng.detect_batch(dev, [(640, 480, 640), (800, 600, 800)], 1.25, 4, 24, 200000, False)
    assert eng.stage_windows() == (want[0] + want[1]).tolist()


def test_graph_replays_and_autotune_match_eager(cascade3):
    """The captured launch graph, replayed several times with eager detects of
    other shapes in between, and the autotuned tile shapes both give exactly
    the eager detections."""
    say("graph replays and autotune vs eager")
    sizes = [(640, 480), (333, 250), (97, 201), (512, 300)]
    imgs = [synth.detect_image(w, h, max(1, w * h // 15000), 24, 64, seed=300 + i) for i, (w, h) in enumerate(sizes)]
    dev = [DeviceArray.from_host(im) for im in imgs]
    spec = [(w, h, w) for w, h in sizes]
    eager = _engine(640, 480, cascade3)
    eager.set_graphs(False)
    want = [det_tuples(x) for x in eager.detect_batch(dev, spec, 1.25, 4, 24, 200000, False)]
    other = synth.detect_image(300, 200, 5, 24, 60, seed=77)
    d_other = DeviceArray.from_host(other)
    want_other = det_tuples(eager.detectMultiScale(d_other, 300, 200, 300, 1.25, 4, 24, 200000, False))
    eng = _engine(640, 480, cascade3)
    for rep in range(4):
        got = eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, False)
        assert [det_tuples(x) for x in got] == want, rep
        # eager work of another shape on the same engine between replays
        eng.set_graphs(False)
        assert det_tuples(eng.detectMultiScale(d_other, 300, 200, 300, 1.25, 4, 24, 200000, False)) == want_other
        eng.set_graphs(True)
    tuned = _engine(640, 480, cascade3)
    tuned.set_autotune(True)
    for rep in range(2):
        got = tuned.detect_batch(dev, spec, 1.25, 4, 24, 200000, False)
        assert [det_tuples(x) for x in got] == want, rep
    say(f"  tuned tile rows {tuned.tile_rows()}")
    assert tuned.tile_rows() in ([8, 4, 4], [4, 4, 4], [8, 8, 8], [8, 8, 4], [4, 8, 8])


@pytest.fixture(scope="module")
def cascade7():
    return synth.blob_cascade((1, 10, 25, 25, 50, 40, 40), seed=11)


def test_two_pass_cascade_matches_oracle(cascade7):
    """A 7-stage / 191-stump cascade: the first pass runs the leading stages
    (111 stumps), the deep pass the last two over slices of the survivors;
    raw and grouped boxes equal the oracle's, per image and in a batch."""
    say("two-pass cascade 7 stages vs oracle")
    sizes = [(640, 480), (333, 250), (1000, 431)]
    imgs = [synth.detect_image(w, h, max(2, w * h // 12000), 24, 96, seed=500 + i) for i, (w, h) in enumerate(sizes)]
    eng = _engine(1000, 480, cascade7)
    dev = [DeviceArray.from_host(im) for im in imgs]
    wants = []
    for im, d, (w, h) in zip(imgs, dev, sizes):
        want, _ = O.detect_raw(*cascade7, im)
        assert len(want) > 0
        wants.append(want)
        raw = eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, applyGrouping=False)
        assert det_tuples(raw) == det_tuples(want)
        grouped = eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, applyGrouping=True)
        assert det_tuples(grouped) == det_tuples(O.group(want, 4))
    got = eng.detect_batch(dev, [(w, h, w) for w, h in sizes], 1.25, 4, 24, 200000, False)
    for g, want in zip(got, wants):
        assert det_tuples(g) == det_tuples(want)


@pytest.mark.parametrize("offset", [0.0, 3.5, -17.25])
def test_pair_mode_margins_near_threshold(cascade7, offset):
    """Leaves of +-(1e6 + a few units) make every stage sum's rounding depend
    on the order of its additions, and thresholds near zero put the windows
    with as many positive as negative leaves within the pair mode's margin of
    the threshold, so its sequential re-summation decides them (and computes
    every last-stage score).  Raw boxes and scores equal the oracle's."""
    say(f"pair-mode margins, threshold offset {offset}")
    feats, stumps, _ = cascade7
    rng = np.random.default_rng(23)
    st = stumps.copy()
    n = len(st)
    st["left"] = np.where(stumps["left"] >= 0, 1, -1) * (1e6 + rng.uniform(-30, 30, n)).astype(np.float32)
    st["right"] = np.where(stumps["right"] >= 0, 1, -1) * (1e6 + rng.uniform(-30, 30, n)).astype(np.float32)
    sizes = (1, 10, 24, 26, 50, 40, 40)
    assert sum(sizes) == n
    stages = np.zeros(len(sizes), STAGE)
    stages["first"] = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    stages["ntrees"] = sizes
    stages["threshold"] = np.float32(offset)
    img = synth.detect_image(640, 480, 20, 24, 64, seed=77)
    want, _ = O.detect_raw(feats, st, stages, img)
    say(f"  oracle raw detections {len(want)}")
    assert len(want) > 100
    eng = _engine(640, 480, (feats, st, stages))
    raw = eng.detectMultiScale(DeviceArray.from_host(img), 640, 480, 640, 1.25, 4, 24, 200000, applyGrouping=False)
    assert det_tuples(raw) == det_tuples(want)


def test_grouping_device_pairs_capacity_and_host_images():
    """Grouping of dense images: a 16,000-box image (about 40 pairs per box)
    overflows the first call's pair capacity, so that call groups on the host
    and the capacity grows; the next calls, and a 20,000-box image, group on
    the device unless a component passes the device limits (then on the
    host).  Every result equals the oracle's grouping of the oracle's raw
    boxes."""
    feats, stumps, stages = synth.blob_cascade((1,), seed=5, n_sample=200)
    stages = stages.copy()
    stages["threshold"] = np.float32(-1e30)
    img = synth.detect_image(640, 480, 20, 24, 64, seed=41)
    d = DeviceArray.from_host(img)
    eng = _engine(640, 480, (feats, stumps, stages))
    for max_det in (16000, 20000):
        want_raw, _ = O.detect_raw(feats, stumps, stages, img, max_det=max_det)
        assert len(want_raw) == max_det
        want = det_tuples(O.group(want_raw, 4))
        say(f"  maxDetections {max_det}: {len(want)} groups")
        for rep in range(2):
            got = eng.detectMultiScale(d, 640, 480, 640, 1.25, 4, 24, max_det, applyGrouping=True)
            assert det_tuples(got) == want, (max_det, rep)
            say(f"    call {rep}: images grouped on the host {eng.last_stats()['imagesGroupedOnHost']}")


def test_grouping_widest_images_and_empty_ones():
    """Device grouping help at the widest supported image (8192 columns: the
    sweep kernel's x bins take just over 64 KB of LDS) in a batch that also
    holds a flat image (no detections) and a narrow one: every image's
    grouped boxes equal the oracle's grouping of the oracle's raw boxes."""
    say("grouping: 8192-wide image, flat image, narrow image in one batch")
    feats, stumps, stages = synth.blob_cascade((1,), seed=5, n_sample=200)
    stages = stages.copy()
    stages["threshold"] = np.float32(-1e30)  # every window that passes the variance gate
    wide = synth.detect_image(8192, 40, 12, 24, 36, seed=61)
    flat = np.full((120, 200), 77, np.uint8)  # zero variance: no window passes the gate
    narrow = synth.detect_image(60, 300, 3, 24, 40, seed=62)
    imgs = [wide, flat, narrow]
    eng = _engine(8192, 300, (feats, stumps, stages))
    dev = [DeviceArray.from_host(im) for im in imgs]
    sizes = [(im.shape[1], im.shape[0], im.shape[1]) for im in imgs]
    got = eng.detect_batch(dev, sizes, 1.25, 3, 24, 200000, True)
    for im, g in zip(imgs, got):
        raw, _ = O.detect_raw(feats, stumps, stages, im)
        assert det_tuples(g) == det_tuples(O.group(raw, 3)), im.shape
    assert len(got[1]) == 0
    # the plain path on the wide image alone
    single = eng.detectMultiScale(dev[0], 8192, 40, 8192, 1.25, 3, 24, 200000, True)
    assert det_tuples(single) == det_tuples(got[0])


def test_batch64_workload_equals_plain_per_image():
    """The batched form at its full benchmark size (64 seeded images, sizes
    log-uniform from 320x240 to 4K, the 38-stage / 6,061-stump cascade, graph
    replay and the autotuner on): every image's detections, grouped and raw,
    equal the plain per-image path's on that image."""
    from cs677s26_amd import workloads
    say("batch of 64 (benchmark workload) vs plain per image")
    sizes, imgs = workloads.batch64()
    feats, stumps, stages = workloads.cascade_38()
    mw, mh = max(s[0] for s in sizes), max(s[1] for s in sizes)
    eng = _engine(mw, mh, (feats, stumps, stages))
    dev = [DeviceArray.from_host(im) for im in imgs]
    spec = [(w, h, w) for w, h in sizes]
    eng.set_autotune(True)
    for grouping in (True, False):
        for rep in range(3):  # direct launch, capture, replay
            got = eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, grouping)
            if rep == 0:
                first = [det_tuples(x) for x in got]
            else:
                assert [det_tuples(x) for x in got] == first, (grouping, rep)
        plain = FaceVisionEngine((24, 24), 1, mw, mh, False, 200000)
        plain.set_cascade_host(feats, stumps, stages)
        for i, (d, (w, h)) in enumerate(zip(dev, sizes)):
            assert det_tuples(plain.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, grouping)) == first[i], (grouping, i)
    say(f"  raw boxes {sum(len(x) for x in first)}, tuned tile rows {eng.tile_rows()}, split {eng.cascade_split()}")
    s0, s1, ns = eng.cascade_split()
    assert 0 < s0 <= s1 <= ns == len(stages)
