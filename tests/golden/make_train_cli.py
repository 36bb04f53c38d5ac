"""Writes tests/golden/train_cli.json: the cascade the oracle's training
driver (oracle/pyoracle.py train_cascade, restating src/main.cu:1523-2392)
trains from the seeded inputs of cs677s26_amd.workloads.train_cli_inputs
with the arguments of tests/test_gpu_cli.py::test_cli_train_equals_oracle_driver
(3 stages, at most 20 rounds each).  Takes about 90 s on 8 cores."""
import json
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from cs677s26_amd import workloads  # noqa: E402
from oracle import pyoracle as O  # noqa: E402


def main():
    pos, negs, images = workloads.train_cli_inputs()
    stumps, stages, info = O.train_cascade(pos, negs, images, max_stages=3, max_weak=20)
    out = {
        "args": {"maxWeakPerStage": 20, "maxStages": 3, "minHitRate": 0.99, "maxFalseAlarm": 0.40,
                 "scaleFactor": 1.25, "seed": 677, "hardNegCandidateMultiplier": 3},
        "stumps": [[int(s["feature"]), int(np.float32(s["theta"]).view(np.uint32)),
                    int(np.float32(s["left"]).view(np.uint32)), int(np.float32(s["right"]).view(np.uint32))]
                   for s in stumps],
        "stages": [[int(g["first"]), int(g["ntrees"]), int(np.float32(g["threshold"]).view(np.uint32))]
                   for g in stages],
        "stage_info": info,
    }
    path = os.path.join(ROOT, "tests", "golden", "train_cli.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {path}: {len(stumps)} stumps, {len(stages)} stages")


if __name__ == "__main__":
    main()
