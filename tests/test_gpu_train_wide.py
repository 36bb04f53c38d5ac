"""The trainer above the LDS sort limit, up to the reference's 65,535 samples
(bindTrainingSet, src/boosting_core.cu:704-706): the wide layout (presort
through global ping-pong buffers, plain 16-bit indices with the split flags
in separate mask rows, weights gathered from global memory) against the
oracle's 32-bit layout, bit for bit."""
import math

import numpy as np
import pytest

from cs677s26_amd import synth
from cs677s26_amd.device import DeviceArray
from cs677s26_amd.engine import FaceVisionEngine
from cs677s26_amd.trainer import FEATURE_BEST_SPLIT, AdaBoostTrainer
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)


def _pool(step):
    basic = O.haar_pool(24, 0)
    core = O.haar_pool(24, 1)
    centre = core[core["r"][:, 1]["w"] == 9.0]
    return np.concatenate([basic[::step], centre[::7 * step]])


def _check_stage(n_pos, n_neg, pool, rounds, seed):
    patches, labels = synth.training_set(n_pos, n_neg, seed=seed)
    n = len(labels)
    tr = AdaBoostTrainer((24, 24), n, 1024)
    d_p, d_f = DeviceArray.from_host(patches), DeviceArray.from_host(pool)
    tr.stage_begin(d_p, labels, n_pos, d_f, len(pool))
    sumT, _, inv = O.train_integrals(patches)
    want = O.presort(pool, sumT, inv, wide=True)
    assert np.array_equal(tr.read_orders(0, len(pool)), (want & 0x7FFFFFFF).astype(np.uint16))
    assert np.array_equal(tr.read_splits(0, len(pool)), (want >> 31).astype(np.uint8))
    say(f"  presort of {len(pool)} features x {n} samples equals the oracle")
    recs, w_or, strong_or, _ = O.train_stage(patches, labels, n_pos, pool, rounds)
    w_prev, _, _, _ = tr.read_state()
    for i, rec in enumerate(recs):
        r = tr.round(0.99, 0.40)
        assert (r["feature"], r["split"], r["parity"]) == (rec["feature"], rec["split"], rec["parity"]), i
        assert r["err"] == rec["err"], i
        assert np.float32(r["theta"]).view(np.uint32) == rec["theta_bits"], i
        assert np.float32(r["stageThreshold"]).view(np.uint32) == rec["threshold_bits"], i
        assert (r["posPass"], r["negSurvive"]) == (rec["posPass"], rec["negSurvive"]), i
        resp = O.feature_responses(pool[r["feature"]], sumT, inv)
        lt = resp < np.float32(r["theta"])
        pred = np.where(lt, -1, 1) if r["parity"] > 0 else np.where(lt, 1, -1)
        e64 = math.fsum(w_prev[pred != labels])
        assert abs(r["err"] - e64) <= 1e-6 * max(e64, 1e-30), (i, r["err"], e64)
        w_prev, _, _, _ = tr.read_state()
    w, strong, _, _ = tr.read_state()
    assert np.array_equal(w.view(np.uint64), w_or.view(np.uint64))
    assert np.array_equal(strong.view(np.uint32), strong_or.view(np.uint32))


def test_stage_40000_samples():
    say("wide stage: 20000 + 20000 samples, 8 rounds")
    _check_stage(20000, 20000, _pool(97), 8, seed=31)


def test_stage_at_the_sample_limit():
    say("wide stage: 32768 + 32767 = 65535 samples, 4 rounds")
    _check_stage(32768, 32767, _pool(601), 4, seed=33)


def test_stage_just_above_the_lds_limit():
    say("wide stage: 10241 + 10240 = 20481 samples, 5 rounds")
    _check_stage(10241, 10240, _pool(211), 5, seed=35)


def test_per_op_interface_40000():
    """sortSamplesPerFeature and evaluateAndFindThreshold at 40,000 samples."""
    say("per-op trainer interface at 40000 samples")
    patches, labels = synth.training_set(18000, 22000, seed=37)
    n = len(labels)
    pool = _pool(181)[:700]
    eng = FaceVisionEngine((24, 24), n, 24, 24)
    d_p = DeviceArray.from_host(patches)
    eng.computeIntegralBatchTransposed(d_p, 24, 24, 24, n, True)
    view = eng.integralSetView()
    rng = np.random.default_rng(8)
    weights = rng.random(n).astype(np.float32)
    weights /= weights.sum()
    active = (rng.random(n) > 0.1).astype(np.uint8)
    d_lab, d_act, d_w = DeviceArray.from_host(labels), DeviceArray.from_host(active), DeviceArray.from_host(weights)
    tr = AdaBoostTrainer((24, 24), n, 1024)
    tr.bindTrainingSet(view, d_lab, d_act, d_w)
    d_f = DeviceArray.from_host(pool)
    d_resp = DeviceArray((len(pool), n), np.float32)
    tr.evaluateFeatureResponses(d_f, 0, len(pool), d_resp)
    d_sorted = DeviceArray((len(pool), n), np.uint16)
    tr.sortSamplesPerFeature(d_resp, len(pool), d_sorted)
    srt = d_sorted.to_host()
    sumT, _, inv = O.train_integrals(patches)
    ords = O.presort(pool, sumT, inv, wide=True)
    assert np.array_equal(srt, (ords & 0x7FFFFFFF).astype(np.uint16))
    d_best = DeviceArray(len(pool), FEATURE_BEST_SPLIT)
    tr.evaluateAndFindThreshold(d_resp, d_sorted, len(pool), d_best)
    best = d_best.to_host()
    sw = np.where(active != 0, np.where(labels > 0, weights, -weights), 0).astype(np.float32)
    ob = O.search(ords, sw)
    assert np.array_equal(best["bestSplitPos"], ob["split"])
    assert np.array_equal(best["bestParity"].astype(np.int32), np.where(ob["split"] >= 0, ob["parity"], 1))
    valid = ob["split"] >= 0
    assert np.array_equal(best["bestErr"][valid],
                          np.ldexp(ob["errQ"][valid].astype(np.float64), -62).astype(np.float32))
    resp = d_resp.to_host()
    for i in np.nonzero(valid)[0][::37]:
        k = int(ob["split"][i])
        s0, s1 = int(ords[i, k]) & 0x7FFFFFFF, int(ords[i, k + 1]) & 0x7FFFFFFF
        assert best["bestTheta"][i] == np.float32(0.5) * (resp[i, s0] + resp[i, s1])
