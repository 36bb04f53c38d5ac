"""Writes the golden fixtures under tests/golden/ from the oracle.

    python tests/golden/make_golden.py [detect] [stage]

* detect_c1.npz -- raw and grouped detections of the 640x480 image with the
  3-stage cascade (sf 1.25, minNeighbors 4, minObjectSize 24).
* detect_c3.npz -- the same for 1920x1080 with the 38-stage cascade.
* stage_c2.json -- 50 boosting rounds of one stage: 10000 + 10000 patches,
  Basic pool (162,336 features), per round the feature, split, parity, exact
  error, theta bits and calibrated threshold bits.

Inputs come from the seeded generators; outputs from the oracle (the C
restatement of the reference CPU path), which tests/test_oracle.py pins
against closed forms, brute force and invariants.
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import numpy as np  # noqa: E402

from cs677s26_amd import workloads  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def detect():
    for name, (img, casc) in (("detect_c1", workloads.detect_640()), ("detect_c3", workloads.detect_1080())):
        t = time.time()
        raw, vis = O.detect_raw(*casc, img)
        grp = O.group(raw, 4)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), raw=raw, grouped=grp, windows=np.int64(vis))
        print(name, len(raw), len(grp), vis, f"{time.time() - t:.1f}s", flush=True)


def stage():
    patches, labels = workloads.training()
    pool = O.haar_pool(24, 0)
    t = time.time()
    recs, w, strong, act = O.train_stage(patches, labels, 10000, pool, 50)
    out = {"samples": int(len(labels)), "features": int(len(pool)), "rounds": recs,
           "final_weight_sum": float(w.sum()), "final_strong_sum": float(strong.astype(np.float64).sum())}
    with open(os.path.join(HERE, "stage_c2.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("stage_c2", len(recs), f"{time.time() - t:.1f}s", flush=True)


if __name__ == "__main__":
    which = sys.argv[1:] or ["detect", "stage"]
    if "detect" in which:
        detect()
    if "stage" in which:
        stage()
