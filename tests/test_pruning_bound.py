"""The stump search's pruning bound, checked on the CPU with the oracle.

A fused round skips a feature when lb[f] -- its last exact error scaled by
every round's rho since, rho being the smallest sample-by-sample ratio of the
search weights' fixed-point magnitudes to the previous round's -- exceeds an
error already found in the round.  That is exact only if lb never exceeds the
feature's true error.  Here every feature's exact error is computed every
round (oracle or_search over all features) and the bound is carried the way
the device carries it: lb stays a lower bound of every feature's error, round
after round, whether the feature was scanned or skipped, so no skipped
feature could have won or tied.
"""
import math

import numpy as np

from cs677s26_amd import synth
from oracle import pyoracle as O


def test_lower_bound_holds_every_round():
    patches, labels = synth.training_set(800, 800, seed=5)
    labels = np.asarray(labels, np.int8)
    n = len(labels)
    num_pos = 800
    pool = O.haar_pool(24, 0)[::211]
    act = np.ones(n, np.uint8)
    sumT, _sq, inv = O.train_integrals(patches)
    ords = O.presort(pool, sumT, inv)
    w = O._normalise(np.where(np.arange(n) < num_pos, 0.5 / num_pos, 0.5 / (n - num_pos)), act)
    used = np.zeros(len(pool), np.uint8)
    lb = np.zeros(len(pool))
    q_prev = None
    skipped_total = 0
    for t in range(12):
        sw = O._signed_search_weights(w, labels, act)
        q = np.array([float(abs(O.fixed_weight(float(v)))) for v in sw])
        best = O.search(ords, sw)
        fi, eq = O.select(best, used)
        errs = best["errQ"].astype(np.float64)
        if q_prev is not None:
            m = q_prev > 0
            rho = float(np.min(q[m] / q_prev[m])) * (1.0 - 2.0 ** -40)
            lb = lb * rho * (1.0 - 2.0 ** -50)
        valid = best["split"] >= 0
        # the bound: never above a feature's exact error this round
        assert np.all(lb[valid] <= errs[valid]), t
        # what the device skips: bound above the round's best error
        skip = (lb > float(eq) * (1.0 + 2.0 ** -40)) & (used == 0)
        assert not skip[fi]
        skipped_total += int(skip.sum())
        # scanned features take their exact error as the bound
        lb = np.where(skip, lb, np.where(valid, errs * (1.0 - 2.0 ** -50), lb))
        q_prev = q
        # the round update (oracle train_stage, src/main.cu:2082-2103)
        k, parity = int(best[fi]["split"]), int(best[fi]["parity"])
        resp = O.feature_responses(pool[fi], sumT, inv)
        theta = np.float32(0.5) * (resp[int(ords[fi, k]) & 0x7FFF] + resp[int(ords[fi, k + 1]) & 0x7FFF])
        e = min(max(math.ldexp(float(eq), -62), 1e-6), 1.0 - 1e-6)
        used[fi] = 1
        lt = resp < theta
        pred = np.where(lt, -1, 1) if parity > 0 else np.where(lt, 1, -1)
        w = O._normalise(np.where(pred != labels, w * math.exp(math.log((1.0 - e) / e)), w), act)
    assert skipped_total > 0  # the bound does prune
