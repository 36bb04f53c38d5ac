"""GPU parity of every sort and scan schedule of a training stage.

A schedule picks the presort layout (LDS radix sort with split flags in the
entries, or the global-memory sort with split-mask rows), how many 16-byte
order loads each lane of the threshold scan keeps in flight, the scan's
workgroups per CU and the form of the scan's weights in LDS.  The sums are
exact integers and the sort is a stable total order, so every schedule must
give the oracle's rounds bit for bit, ties included; the autotuner only picks
among them.
"""
import numpy as np
import pytest

from cs677s26_amd import synth
from cs677s26_amd.device import DeviceArray
from cs677s26_amd.trainer import AdaBoostTrainer
from oracle import pyoracle as O

pytestmark = pytest.mark.gpu

# (wide, ahead, blocks per CU, packed): every value of every field appears
SCHEDULES = [(0, 2, 1, 0), (0, 8, 4, 1), (0, 4, 2, 1), (1, 4, 2, 0), (1, 2, 4, 0), (1, 8, 1, 0)]


def say(msg):
    print(msg, flush=True)


def _pool():
    basic = O.haar_pool(24, 0)
    core = O.haar_pool(24, 1)
    centre = core[core["r"][:, 1]["w"] == 9.0]
    return np.concatenate([basic[::53], centre[::97]])


def _check_rounds(tr, recs, w_or, what):
    for i, rec in enumerate(recs):
        r = tr.round(0.99, 0.40)
        got = (r["feature"], r["split"], r["parity"], r["err"], int(np.float32(r["theta"]).view(np.uint32)))
        want = (rec["feature"], rec["split"], rec["parity"], rec["err"], int(rec["theta_bits"]))
        assert got == want, (what, i, got, want)
        assert np.float32(r["stageThreshold"]).view(np.uint32) == rec["threshold_bits"], (what, i)
    w, _, _, _ = tr.read_state()
    assert np.array_equal(w.view(np.uint64), w_or.view(np.uint64)), what


@pytest.fixture(scope="module")
def case():
    patches, labels = synth.training_set(1000, 1000, seed=9)
    pool = _pool()
    recs, w_or, _, _ = O.train_stage(patches, labels, 1000, pool, 8)
    sumT, _, inv = O.train_integrals(patches)
    orders = O.presort(pool[:64], sumT, inv)
    # duplicated features: every round two features tie on the error
    sub = pool[:200]
    dup = np.concatenate([sub, sub])
    recs_dup, w_dup, _, _ = O.train_stage(patches, labels, 1000, dup, 5)
    return patches, labels, pool, recs, w_or, orders, dup, recs_dup, w_dup


@pytest.mark.parametrize("sched", SCHEDULES)
def test_every_schedule_gives_the_oracle_rounds(case, sched):
    patches, labels, pool, recs, w_or, orders, dup, recs_dup, w_dup = case
    say(f"schedule {sched}: 8 rounds on 2000 samples, 5 rounds on duplicated features")
    d_p = DeviceArray.from_host(patches)
    tr = AdaBoostTrainer((24, 24), len(labels), 1024)
    tr.set_schedule(sched)
    d_f = DeviceArray.from_host(pool)
    tr.stage_begin(d_p, labels, 1000, d_f, len(pool))
    assert tr.schedule() == sched
    # the presort in the schedule's layout: plain indices and split flags
    got = tr.read_orders(0, 64)
    flags = tr.read_splits(0, 64)
    idx = got if sched[0] else (got & 0x7FFF)
    assert np.array_equal(idx, orders & 0x7FFF)
    assert np.array_equal(flags, (orders >> 15).astype(np.uint8))
    _check_rounds(tr, recs, w_or, sched)
    d_dup = DeviceArray.from_host(dup)
    tr.stage_begin(d_p, labels, 1000, d_dup, len(dup))
    _check_rounds(tr, recs_dup, w_dup, ("duplicates",) + tuple(sched))


def test_autotune_picks_a_candidate_and_reuses_it(case):
    patches, labels, pool, recs, w_or, _, _, _, _ = case
    say("autotuned schedule: 8 rounds on 2000 samples, then the bucket's cached pick")
    d_p, d_f = DeviceArray.from_host(patches), DeviceArray.from_host(pool)
    tr = AdaBoostTrainer((24, 24), len(labels), 1024)
    tr.set_autotune(True)
    tr.stage_begin(d_p, labels, 1000, d_f, len(pool))
    sched = tr.schedule()
    assert sched[0] in (0, 1) and sched[1] in (2, 4, 8) and sched[2] in (1, 2, 4) and sched[3] in (0, 1)
    _check_rounds(tr, recs, w_or, ("autotuned",) + sched)
    tr2 = AdaBoostTrainer((24, 24), len(labels), 1024)
    tr2.set_autotune(True)
    tr2.stage_begin(d_p, labels, 1000, d_f, len(pool))
    assert tr2.schedule() == sched
    # forcing a schedule turns the tuner off; None goes back to the default
    tr2.set_schedule((1, 2, 1, 0))
    tr2.stage_begin(d_p, labels, 1000, d_f, len(pool))
    assert tr2.schedule() == (1, 2, 1, 0)
    tr2.set_schedule(None)
    tr2.stage_begin(d_p, labels, 1000, d_f, len(pool))
    assert tr2.schedule() == (0, 4, 2, 0)


@pytest.mark.parametrize("prune", [False, True])
def test_pruning_gives_the_oracle_rounds(case, prune):
    """The round's search with and without its exact pruning (a feature is
    skipped only when a lower bound on its error exceeds an error already
    found, so it can neither win nor tie): the oracle's rounds bit for bit,
    on the pool and on duplicated features that tie every round.  With the
    pruning on, fewer features than a full scan's are scanned."""
    patches, labels, pool, recs, w_or, orders, dup, recs_dup, w_dup = case
    say(f"pruning {prune}: 8 rounds on 2000 samples, 5 rounds on duplicated features")
    d_p = DeviceArray.from_host(patches)
    tr = AdaBoostTrainer((24, 24), len(labels), 1024)
    tr.set_schedule((0, 4, 2, 0))
    tr.set_pruning(prune)
    d_f = DeviceArray.from_host(pool)
    tr.stage_begin(d_p, labels, 1000, d_f, len(pool))
    _check_rounds(tr, recs, w_or, ("pruning", prune))
    scanned = tr.pruning_stats()
    if prune:
        # the first round scans everything; later ones skip some features
        assert len(pool) <= scanned < len(recs) * len(pool), scanned
    else:
        assert scanned == 0
    d_dup = DeviceArray.from_host(dup)
    tr.stage_begin(d_p, labels, 1000, d_dup, len(dup))
    _check_rounds(tr, recs_dup, w_dup, ("pruning duplicates", prune))
