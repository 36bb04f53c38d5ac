"""Served detect (vj_server_*): every request's result equals the plain
per-image detect of its image, alone and in a mix of concurrent requests;
admission by HBM budget refuses requests that cannot fit."""
import threading

import numpy as np
import pytest

from cs677s26_amd import _native, synth
from cs677s26_amd._native import VJError
from cs677s26_amd.device import DeviceArray
from cs677s26_amd.engine import FaceVisionEngine
from cs677s26_amd.server import DetectServer
from cs677s26_amd.vjtypes import det_tuples
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def cascade3():
    return synth.blob_cascade((1, 10, 25))


@pytest.fixture(scope="module")
def mix():
    """Crops of one seeded scene at mixed sizes, some with a row stride."""
    base = synth.detect_image(1200, 900, 60, 24, 120, seed=41)
    rng = np.random.default_rng(5)
    out = []
    for i in range(14):
        w = int(rng.integers(40, 1200))
        h = int(rng.integers(30, 900))
        x0 = int(rng.integers(0, 1200 - w + 1))
        y0 = int(rng.integers(0, 900 - h + 1))
        out.append(base[y0:y0 + h, x0:x0 + w])
    return out


def plain(cascade, img, grouping, max_det=200000):
    eng = FaceVisionEngine((24, 24), 1, img.shape[1], img.shape[0], False, max_det)
    eng.set_cascade_host(*cascade)
    d = DeviceArray.from_host(np.ascontiguousarray(img))
    return eng.detectMultiScale(d, img.shape[1], img.shape[0], img.shape[1], 1.25, 4, 24, max_det, grouping)


@pytest.mark.parametrize("grouping", [True, False])
def test_mix_equals_alone_and_plain(cascade3, mix, grouping):
    say(f"serve mix of {len(mix)} (grouping={grouping})")
    want = [det_tuples(plain(cascade3, im, grouping)) for im in mix]
    assert sum(len(x) for x in want) > 0
    with DetectServer(*cascade3, max_image=(1200, 900), slots=2, max_batch=4, hbm_budget=1 << 30,
                      apply_grouping=grouping) as srv:
        # the mix: every request queued at once from two threads
        tickets = [None] * len(mix)

        def submit(idx):
            for i in idx:
                tickets[i] = srv.submit(mix[i])

        th = [threading.Thread(target=submit, args=(range(k, len(mix), 2),)) for k in range(2)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        got = [srv.wait(t) for t in tickets]
        st = srv.stats()
        say(f"  stats {st}")
        assert st["completed"] == st["submitted"] == len(mix) and st["rejected"] == 0
        for i, (dets, lat) in enumerate(got):
            assert det_tuples(dets) == want[i], i
            assert lat >= 0.0
        # alone: the same requests one at a time through the same server
        for i in (0, 5, 11):
            dets, _ = srv.wait(srv.submit(mix[i]))
            assert det_tuples(dets) == want[i], i


def test_raw_results_equal_oracle(cascade3, mix):
    say("serve raw vs oracle")
    with DetectServer(*cascade3, max_image=(1200, 900), slots=3, max_batch=8, hbm_budget=1 << 30,
                      apply_grouping=False) as srv:
        tickets = [srv.submit(im) for im in mix[:6]]
        for im, t in zip(mix[:6], tickets):
            dets, _ = srv.wait(t)
            want, _ = O.detect_raw(*cascade3, np.ascontiguousarray(im))
            assert det_tuples(dets) == det_tuples(want)


def test_admission_by_budget(cascade3):
    say("serve admission")
    big = synth.detect_image(1000, 800, 10, 24, 64, seed=3)
    small = synth.detect_image(120, 90, 2, 24, 40, seed=4)
    with DetectServer(*cascade3, max_image=(1000, 800), slots=2, max_batch=4, hbm_budget=24 << 20,
                      max_detections=1000) as srv:
        with pytest.raises(VJError) as ei:
            srv.submit(big)
        assert ei.value.status == _native.VJ_OUT_OF_MEMORY
        dets, _ = srv.wait(srv.submit(small))
        assert det_tuples(dets) == det_tuples(plain(cascade3, small, True, 1000))
        assert srv.stats()["rejected"] == 1


def test_served_workload_poisson_stream():
    """The served workload itself: 240 requests, 240p..4K crops of a seeded
    scene, in their Poisson arrival order at 100 QPS (arrivals replayed in
    real time), through the 38-stage / 6,061-stump cascade with the bench's
    server shape (3 slots, batches of up to 8, 16 GiB budget).  Every result
    equals the plain per-image detect of its image; the three smallest
    requests also equal the oracle (raw detect, then grouping)."""
    import time

    from cs677s26_amd import workloads
    arrivals, imgs = workloads.served()
    casc = workloads.cascade_38()
    say(f"served workload: {len(imgs)} requests, {sum(im.size for im in imgs) / 1e6:.1f} Mpix")
    eng = FaceVisionEngine((24, 24), 1, 3840, 2160, False, 200000)
    eng.set_cascade_host(*casc)
    want = []
    for im in imgs:
        d = DeviceArray.from_host(np.ascontiguousarray(im))
        want.append(det_tuples(eng.detectMultiScale(d, im.shape[1], im.shape[0], im.shape[1], 1.25, 4, 24, 200000,
                                                    True)))
    say("  plain per-image results done")
    with DetectServer(*casc, max_image=(3840, 2160), slots=3, max_batch=8, hbm_budget=16 << 30) as srv:
        t0 = time.perf_counter()
        tickets = []
        for a, im in zip(arrivals, imgs):
            delay = t0 + a - time.perf_counter()
            if delay > 0:
                time.sleep(delay)
            tickets.append(srv.submit(im))
        got = [srv.wait(t)[0] for t in tickets]
        st = srv.stats()
    say(f"  stats {st}")
    assert st["completed"] == len(imgs) and st["rejected"] == 0 and st["batches"] > 1
    for i, dets in enumerate(got):
        assert det_tuples(dets) == want[i], (i, imgs[i].shape)
    assert sum(len(w) for w in want) > 0
    for i in sorted(range(len(imgs)), key=lambda k: imgs[k].size)[:3]:
        raw, _ = O.detect_raw(*casc, np.ascontiguousarray(imgs[i]))
        assert want[i] == det_tuples(O.group(raw, 4)), (i, imgs[i].shape)
    say("  served results equal plain and oracle")
