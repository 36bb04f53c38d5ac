"""The drop-in command line on the GPU: `cuda_hello detect` and the compare
verbs (`detect_compare`, `train_compare`) against the
oracle, and `cuda_hello train` end to end (seeded stage-0 negatives, hard
negatives mined with the cascade trained so far, the top-up sample, and a
model file the detector loads)."""
import os
import re
import subprocess

import numpy as np
import pytest

from cs677s26_amd import synth
from cs677s26_amd.engine import loadCascadeModel, saveCascadeModel
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "cs677s26_amd", "bin", "cuda_hello")


def say(msg):
    print(msg, flush=True)


def write_pgm(path, img):
    h, w = img.shape
    with open(path, "wb") as f:
        f.write(b"P5\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(img, np.uint8).tobytes())


def test_cli_detect_matches_oracle(tmp_path):
    say("cli detect vs oracle")
    feats, stumps, stages = synth.blob_cascade((1, 10, 25))
    model = tmp_path / "m.bin"
    saveCascadeModel(model, (24, 24), feats, stumps, stages)
    img = synth.detect_image(640, 480, 20, 24, 64, seed=31)
    write_pgm(tmp_path / "a.pgm", img)
    r = subprocess.run([EXE, "detect", str(model), str(tmp_path / "a.pgm"), str(tmp_path / "o.ppm")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    raw, _ = O.detect_raw(feats, stumps, stages, img)
    want = len(O.group(raw, 4))
    m = re.search(r"\[DETECT\] detections=(\d+)", r.stdout)
    assert m and int(m.group(1)) == want
    ppm = (tmp_path / "o.ppm").read_bytes()
    assert ppm.startswith(b"P6\n640 480\n255\n") and len(ppm) == len(b"P6\n640 480\n255\n") + 640 * 480 * 3


def test_cli_train_mines_hard_negatives(tmp_path):
    """Three stages on 2,000 synthetic faces: stage 0 samples its negatives
    from the PGM images (and caches them), stages 1-2 mine hard negatives
    with the cascade so far and top them up from a seeded sample; the saved
    model loads and its stages index its stumps."""
    say("cli train with hard-negative mining")
    rng = np.random.default_rng(8)
    faces = synth.face_patches(2000, rng)
    faces.astype(np.uint8).tofile(tmp_path / "faces_u8.bin")
    imgs = tmp_path / "neg"
    imgs.mkdir()
    for i, (w, h) in enumerate([(320, 240), (400, 300), (256, 256), (500, 200), (300, 420), (640, 480)]):
        write_pgm(imgs / f"n{i}.pgm", synth.detect_image(w, h, 6, 24, 80, seed=900 + i))
    cache = tmp_path / "cache"
    out = tmp_path / "model.bin"
    args = [EXE, "train", str(tmp_path / "faces_u8.bin"), "10000", "12", "200000", "4", "1.25", "3", "0.99", "0.4",
            "24", str(cache), str(imgs), "677", "3", str(out)]
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    say(r.stdout[-2000:])
    assert r.returncode == 0, r.stderr
    assert (cache / "stage_negatives" / "non_faces_stage00.bin").stat().st_size == 10000 * 576
    assert re.search(r"\[TRAIN\] stage=1 hardNegSelected=(\d+)/10000", r.stdout)
    win, f, s, g = loadCascadeModel(out)
    assert win == (24, 24) and len(g) >= 2
    assert int(g["first"][-1]) + int(g["ntrees"][-1]) == len(s)
    assert (s["feature"] >= 0).all() and (s["feature"] < len(f)).all()
    # the same seed gives the same model
    out2 = tmp_path / "model2.bin"
    r2 = subprocess.run(args[:-1] + [str(out2)], capture_output=True, text=True, timeout=300)
    assert r2.returncode == 0, r2.stderr
    assert out.read_bytes() == out2.read_bytes()


def test_cli_train_equals_oracle_driver(tmp_path):
    """`cuda_hello train` for three stages against the oracle's restatement of
    the reference's training loop (oracle/pyoracle.py train_cascade,
    src/main.cu:1523-2392), fixed in tests/golden/train_cli.json: the same
    stumps (feature, theta, leaves, bit for bit) and stage records
    (first, ntrees, threshold - 1e-5f), so the same stage-0 negatives, the
    same hard negatives mined in the same order (std::shuffle of the images,
    the 256 best of each by std::partial_sort, cropResizeTo24), the same
    top-up picks and the same stage ends.  The model holds the Core pool."""
    _train_cli_vs_golden(tmp_path, None)


def _train_cli_vs_golden(tmp_path, debug_every):
    import json
    from cs677s26_amd import workloads
    say(f"cli train vs the oracle's training driver (golden), VJ_DEBUG_EVERY={debug_every}")
    gold = json.load(open(os.path.join(ROOT, "tests", "golden", "train_cli.json")))
    pos, negs, images = workloads.train_cli_inputs()
    pos.tofile(tmp_path / "faces_u8.bin")
    cache = tmp_path / "cache"
    (cache / "stage_negatives").mkdir(parents=True)
    for s, arr in negs.items():
        arr.tofile(cache / "stage_negatives" / f"non_faces_stage{s:02d}.bin")
    imgs = tmp_path / "neg"
    imgs.mkdir()
    for i, im in enumerate(images):
        write_pgm(imgs / f"img{i}.pgm", im)
    out = tmp_path / "model.bin"
    a = gold["args"]
    args = [EXE, "train", str(tmp_path / "faces_u8.bin"), "10000", str(a["maxWeakPerStage"]), "200000", "4",
            str(a["scaleFactor"]), str(a["maxStages"]), str(a["minHitRate"]), str(a["maxFalseAlarm"]), "24",
            str(cache), str(imgs), str(a["seed"]), str(a["hardNegCandidateMultiplier"]), str(out)]
    env = dict(os.environ)
    env.pop("VJ_DEBUG_EVERY", None)
    if debug_every is not None:
        env["VJ_DEBUG_EVERY"] = str(debug_every)
    r = subprocess.run(args, capture_output=True, text=True, timeout=300, env=env)
    say(r.stdout[-1500:])
    assert r.returncode == 0, r.stderr
    for inf in gold["stage_info"][1:]:
        assert f"[TRAIN] stage={inf['stage']} hardNegSelected={inf['mined']}/10000" in r.stdout
    win, f, s, g = loadCascadeModel(out)
    assert win == (24, 24)
    got_stumps = [[int(x["feature"]), int(np.float32(x["theta"]).view(np.uint32)),
                   int(np.float32(x["left"]).view(np.uint32)), int(np.float32(x["right"]).view(np.uint32))] for x in s]
    got_stages = [[int(x["first"]), int(x["ntrees"]), int(np.float32(x["threshold"]).view(np.uint32))] for x in g]
    assert got_stumps == gold["stumps"]
    assert got_stages == gold["stages"]
    assert np.array_equal(np.asarray(f).view(np.uint8), O.haar_pool(24, 1).view(np.uint8))
    return r.stdout


def test_cli_train_debug_lines(tmp_path):
    """VJ_DEBUG_EVERY=k (src/main.cu:1708-1720, 1967-1983, 2105-2162,
    2227-2283): the same model as without it, plus the [DBG] lines.  At
    stage start the normalised weights split 0.5 / 0.5.  On every k-th
    round (and a stage's last allowed round) the chosen stump's errors are
    summed over the updated weights before the normalisation, where
    AdaBoost's update leaves the stump at exactly one half of the total: both
    parities' errors come to 1 - err (the misclassified weight e times
    (1-e)/e, and the rest, 1 - e).  The score line's weights sum to one and
    its threshold is the round's calibrated stage threshold."""
    import re
    out = _train_cli_vs_golden(tmp_path, 2)
    assert "[INFO] Debug logging enabled: VJ_DEBUG_EVERY=2" in out
    init = re.findall(r"\[DBG\] stage=(\d+) initWeightSum pos=(\S+) neg=(\S+) activeNeg=(\d+)", out)
    assert len(init) >= 2
    for _st, pw, nw, an in init:
        assert abs(float(pw) - 0.5) < 1e-3 and abs(float(nw) - 0.5) < 1e-3 and int(an) == 10000
    train = {(int(a), int(b)): float(e) for a, b, e in re.findall(r"\[TRAIN\] stage=(\d+) weak=(\d+) feature=\d+ err=(\S+)", out)}
    errs = re.findall(r"\[DBG\] stage=(\d+) weak=(\d+) bestParityRaw=(\S+) parityUsed=(\S+) errP=(\S+) errN=(\S+) theta=(\S+)", out)
    assert errs
    for st, wk, praw, pused, ep, en, _th in errs:
        assert int(wk) % 2 == 0 or (int(st), int(wk) + 1) not in train
        assert praw == pused
        e = train[(int(st), int(wk))]
        assert abs(float(ep) - (1 - e)) < 2e-3 and abs(float(en) - (1 - e)) < 2e-3
    scores = re.findall(r"\[DBG\] stage=\d+ weak=\d+ score_pos\[min,max,mean,p50\]=\[(\S+)\] score_neg\[min,max,mean\]=\[(\S+)\] threshold=(\S+) weightSum\[pos,neg\]=\[(\S+)\]", out)
    assert len(scores) == len(errs)
    for pos, neg, _thr, ws in scores:
        pmin, pmax, pmean, p50 = map(float, pos.split(","))
        assert pmin <= p50 <= pmax and pmin <= pmean <= pmax
        nmin, nmax, nmean = map(float, neg.split(","))
        assert nmin <= nmean <= nmax
        assert abs(sum(map(float, ws.split(","))) - 1.0) < 1e-3


def test_cli_detect_compare_legs_agree(tmp_path):
    """detect_compare (src/main.cu:1216-1368): the host leg and the GPU leg
    report the oracle's raw detection count and agree box for box; the PPM
    is written; maxScales other than 0 is refused as in the reference."""
    say("cli detect_compare")
    feats, stumps, stages = synth.blob_cascade((1, 10, 25))
    model = tmp_path / "m.bin"
    saveCascadeModel(model, (24, 24), feats, stumps, stages)
    img = synth.detect_image(480, 360, 12, 24, 64, seed=33)
    write_pgm(tmp_path / "a.pgm", img)
    raw, _ = O.detect_raw(feats, stumps, stages, img)
    r = subprocess.run([EXE, "detect_compare", str(model), str(tmp_path / "a.pgm"), str(tmp_path / "o.ppm"),
                        "1.25", "24", "200000", "0", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    m = re.search(r"\[DETECT_COMPARE\] cpuDetections=(\d+) gpuDetections=(\d+) runs=2 maxScales=0", r.stdout)
    assert m and int(m.group(1)) == len(raw) and int(m.group(2)) == len(raw), r.stdout
    assert "[DETECT_COMPARE] match=yes" in r.stdout
    assert re.search(r"\[DETECT_COMPARE\] speedup=[0-9.e+]+x", r.stdout)
    assert (tmp_path / "o.ppm").read_bytes().startswith(b"P6\n480 360\n255\n")
    # the cap: both legs stop at the first 100 in scan order
    r = subprocess.run([EXE, "detect_compare", str(model), str(tmp_path / "a.pgm"), str(tmp_path / "o.ppm"),
                        "1.25", "24", "100", "0", "1"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "[DETECT_COMPARE] match=yes" in r.stdout, r.stdout + r.stderr
    assert f"cpuDetections={min(100, len(raw))} gpuDetections={min(100, len(raw))}" in r.stdout
    r = subprocess.run([EXE, "detect_compare", str(model), str(tmp_path / "a.pgm"), str(tmp_path / "o.ppm"),
                        "1.25", "24", "200000", "3"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "requires maxScales=0" in r.stderr


def test_cli_train_compare_legs_agree(tmp_path):
    """train_compare (src/main.cu:1032-1214): both legs report the same stump
    over the first 1,500 Core features, and it is the oracle's (per-feature
    exact search, then the smallest float error with the smaller index)."""
    say("cli train_compare")
    patches, labels = synth.training_set(300, 300, seed=21)
    patches[:300].tofile(tmp_path / "pos.bin")
    patches[300:].tofile(tmp_path / "neg.bin")
    r = subprocess.run([EXE, "train_compare", str(tmp_path / "pos.bin"), str(tmp_path / "neg.bin"), "256", "200",
                        "1500", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "[TRAIN_COMPARE] samples=456 numPos=256 numNeg=200 features=1500 runs=2" in r.stdout, r.stdout
    assert "[TRAIN_COMPARE] match=yes" in r.stdout, r.stdout
    pool = O.haar_pool(24, 1)[:1500]
    sel = np.concatenate([patches[:256], patches[300:500]])
    sumT, _, inv = O.train_integrals(sel)
    w = np.concatenate([np.full(256, np.float32(0.5) / np.float32(256), np.float32),
                        np.full(200, np.float32(0.5) / np.float32(200), np.float32)])
    sw = np.concatenate([w[:256], -w[256:]]).astype(np.float32)
    best = O.search(O.presort(pool, sumT, inv), sw)
    errs = [np.float32(np.ldexp(float(b["errQ"]), -62)) for b in best]
    fi = min(range(len(best)), key=lambda i: (errs[i], i))
    m = re.search(r"\[TRAIN_COMPARE\] GPU best feature=(-?\d+) theta=\S+ parity=(-?\d+)", r.stdout)
    assert m and int(m.group(1)) == fi and int(m.group(2)) == int(best[fi]["parity"]), r.stdout
