"""Benchmark: detect Mpix/s on 1080p with the 38-stage cascade (one MI355X).

A step is one plain detect call (pyramid, integrals, cascade, compaction,
device-to-host copy and host grouping) on a 1920x1080 u8 frame that is
already resident in HBM.  Prints ONE JSON line (see README of this file's
fields below).  Extra legs, each optional:

* ``batched``  -- 64 uneven images per detect_batch call (Mpix/s);
* ``stump``    -- one AdaBoost stage: presort of 162,336 features x 20,000
  samples, then 50 boosting rounds (seconds per stage);
* ``cpu``      -- the oracle (C restatement of the reference CPU path) on the
  same 1080p frame, one thread (the reference CPU path is single-threaded)
  and every allowed core; one warm-up, then the median of three runs.

``roofline`` is for the cascade phase (the dominant one: ``cascade_kernel``,
the first pass over every tile, then the deep work-list kernels, the persistent
``cascade_deep_kernel`` over the survivors and ``cascade_tail_kernel`` over the
last stages' few windows): algorithmic bytes per frame
(12*(W+1)*(H+1) summed over levels, plus model and output bytes) over the
phase's average duration from HIP events on the detect stream.
``lds_roofline`` is the same phase against the LDS, where its stump
evaluations gather their corners: four f32 reads per rectangle per (window,
stump) evaluation, with the windows entering each stage counted by the engine
in its counting mode (one extra call), over 256 CUs x 128 B per clock x 2.4 GHz.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
LDS_PEAK_GBS = 256 * 128 * 2.4
METRIC = "detect Mpix/s (1080p, 38-stage); stump-search s/stage; p99 latency at QPS"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def detect_leg(args):
    import numpy as np

    from cs677s26_amd import workloads
    from cs677s26_amd.device import DeviceArray, synchronize
    from cs677s26_amd.engine import FaceVisionEngine

    img, (feats, stumps, stages) = workloads.detect_1080()
    h, w = img.shape
    eng = FaceVisionEngine((24, 24), 1, w, h, False, 200000)
    eng.set_cascade_host(feats, stumps, stages)
    d_img = DeviceArray.from_host(img)
    for _ in range(args.warmup):
        eng.detectMultiScale(d_img, w, h, w, 1.25, 4, 24, 200000, True)
    # timed region: K whole detect calls, device synchronised on both sides
    synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dets = eng.detectMultiScale(d_img, w, h, w, 1.25, 4, 24, 200000, True)
    synchronize()
    wall = time.perf_counter() - t0
    stats = eng.last_stats()
    # per-kernel device time from HIP events on the detect stream
    eng.set_timing(True)
    phases = np.zeros(4)
    for _ in range(args.steps):
        eng.detectMultiScale(d_img, w, h, w, 1.25, 4, 24, 200000, True)
        phases += np.array(eng.phase_ms())
    eng.set_timing(False)
    # one more call in counting mode: the windows entering each stage
    eng.count_stages(True)
    eng.detectMultiScale(d_img, w, h, w, 1.25, 4, 24, 200000, True)
    enter = np.array(eng.stage_windows(), np.int64)
    eng.count_stages(False)
    phases /= args.steps
    # LDS gather bytes of the cascade: every (window, stump) evaluation reads
    # four f32 corners per rectangle of its feature from the staged footprint
    rects = feats["rectCount"][stumps["feature"]].astype(np.int64)
    stage_bytes = np.array([16 * int(rects[g["first"]:g["first"] + g["ntrees"]].sum()) for g in stages], np.int64)
    return {
        "mpix": w * h / 1e6,
        "wall_s": wall,
        "ms_per_step": 1e3 * wall / args.steps,
        "detections": int(len(dets)),
        "stats": stats,
        "phase_ms": phases.tolist(),
        "stage_windows": enter.tolist(),
        "evals": int((enter * stages["ntrees"].astype(np.int64)).sum()),
        "gather_bytes": int((enter * stage_bytes).sum()),
    }


def batched_leg(args):
    from cs677s26_amd import workloads
    from cs677s26_amd.device import DeviceArray, synchronize
    from cs677s26_amd.engine import FaceVisionEngine

    sizes, imgs = workloads.batch64()
    mw = max(s[0] for s in sizes)
    mh = max(s[1] for s in sizes)
    feats, stumps, stages = workloads.cascade_38()
    eng = FaceVisionEngine((24, 24), 1, mw, mh, False, 200000)
    eng.set_cascade_host(feats, stumps, stages)
    dev = [DeviceArray.from_host(im) for im in imgs]
    spec = [(w, h, w) for (w, h) in sizes]
    # warm-up: the autotuner picks the cascade tile shape for this batch
    # shape, then the launch sequence is captured once and replayed
    eng.set_autotune(True)
    eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, True)
    eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, True)
    tile_rows = eng.tile_rows()
    split = eng.cascade_split()
    synchronize()
    reps = max(1, min(args.steps, 3))
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, True)
    synchronize()
    wall = time.perf_counter() - t0
    # integral phase at scale: pyramid + integral kernels of one batch, HIP
    # events on the detect stream, against the batch's algorithmic bytes
    eng.set_timing(True)
    eng.detect_batch(dev, spec, 1.25, 4, 24, 200000, True)
    ph = eng.phase_ms()
    st = eng.last_stats()
    eng.set_timing(False)
    int_ms = ph[0] + ph[1]
    int_roof = {"achieved": st["integralBytes"] / (int_ms * 1e-3) / 1e9 if int_ms > 0 else 0.0,
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": st["integralBytes"], "avg_ms": int_ms,
                "cascade_ms": ph[2]}
    int_roof["frac"] = int_roof["achieved"] / HBM_PEAK_GBS
    # the plain per-image path on the same 64 images: one detect call each
    eng.set_autotune(False)
    t1 = time.perf_counter()
    for d, (w, h) in zip(dev, sizes):
        eng.detectMultiScale(d, w, h, w, 1.25, 4, 24, 200000, True)
    synchronize()
    wall_plain = time.perf_counter() - t1
    mpix = sum(w * h for w, h in sizes) / 1e6
    return {"mpix_per_s": mpix * reps / wall, "mpix_per_batch": mpix, "ms_per_batch": 1e3 * wall / reps,
            "plain_per_image_mpix_per_s": mpix / wall_plain, "tile_rows": tile_rows,
            "tail_pass": split[1] < split[2],
            "speedup_vs_plain": (mpix * reps / wall) / (mpix / wall_plain), "integral_roofline": int_roof}


def served_leg(args):
    """Poisson stream at a fixed QPS through the served form.  Latency is
    scheduled arrival -> result ready: the submitter's lag behind the
    schedule plus the server's submit -> done time (so the order in which
    this harness's threads collect results does not count)."""
    import queue
    import threading

    import numpy as np

    from cs677s26_amd import workloads
    from cs677s26_amd.server import DetectServer

    arrivals, imgs = workloads.served(args.requests, args.qps)
    feats, stumps, stages = workloads.cascade_38()
    srv = DetectServer(feats, stumps, stages, max_image=(3840, 2160), slots=args.slots, max_batch=args.max_batch,
                       hbm_budget=16 << 30)
    # warm-up: every slot sees a batch of the largest requests
    big = sorted(range(len(imgs)), key=lambda i: -imgs[i].size)[:args.max_batch * args.slots]
    for t in [srv.submit(imgs[i]) for i in big]:
        srv.wait(t)
    n = len(imgs)
    done = np.zeros(n)
    sub = np.zeros(n)
    srv_lat = np.zeros(n)
    served_dets = [None] * n
    q = queue.Queue()

    def waiter():
        while True:
            item = q.get()
            if item is None:
                return
            i, tk = item
            dets, lat = srv.wait(tk)
            done[i] = time.perf_counter()
            srv_lat[i] = lat
            served_dets[i] = dets

    ws = [threading.Thread(target=waiter) for _ in range(8)]
    for w in ws:
        w.start()
    t0 = time.perf_counter()
    for i in range(n):
        delay = t0 + arrivals[i] - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        sub[i] = time.perf_counter()
        q.put((i, srv.submit(imgs[i])))
    for _ in ws:
        q.put(None)
    for w in ws:
        w.join()
    lat = 1e3 * (sub - (t0 + arrivals)) + srv_lat
    mpix = sum(im.size for im in imgs) / 1e6
    span = max(done.max(), (sub + srv_lat / 1e3).max()) - t0
    # saturation: the same requests submitted at once
    t1 = time.perf_counter()
    for t in [srv.submit(im) for im in imgs]:
        srv.wait(t)
    sat = time.perf_counter() - t1
    st = srv.stats()
    srv.close()
    # the plain per-image path on the same requests (one engine, image by image)
    from cs677s26_amd.device import DeviceArray, synchronize
    from cs677s26_amd.engine import FaceVisionEngine
    eng = FaceVisionEngine((24, 24), 1, 3840, 2160, False, 200000)
    eng.set_cascade_host(feats, stumps, stages)
    dev = [DeviceArray.from_host(im) for im in imgs]
    eng.detectMultiScale(dev[0], imgs[0].shape[1], imgs[0].shape[0], imgs[0].shape[1], 1.25, 4, 24, 200000, True)
    synchronize()
    t2 = time.perf_counter()
    plain_dets = [eng.detectMultiScale(d, im.shape[1], im.shape[0], im.shape[1], 1.25, 4, 24, 200000, True)
                  for d, im in zip(dev, imgs)]
    synchronize()
    plain = time.perf_counter() - t2
    # outside the timed regions: every result of the Poisson run equals the
    # plain per-image detect of its request
    from cs677s26_amd.vjtypes import det_tuples
    mismatched = [i for i in range(n) if det_tuples(served_dets[i]) != det_tuples(plain_dets[i])]
    return {"qps": args.qps, "requests": n, "mpix": mpix, "p50_ms": float(np.percentile(lat, 50)),
            "p99_ms": float(np.percentile(lat, 99)), "server_p50_ms": float(np.percentile(srv_lat, 50)),
            "server_p99_ms": float(np.percentile(srv_lat, 99)), "achieved_qps": n / span,
            "mpix_per_s": mpix / span, "saturated_mpix_per_s": mpix / sat,
            "plain_per_image_mpix_per_s": mpix / plain, "saturated_speedup_vs_plain": plain / sat, "slots": args.slots,
            "max_slots_busy": st["maxSlotsBusy"], "max_batch": st["maxBatchSeen"], "batches": st["batches"],
            "results_checked": n, "results_equal_plain": not mismatched, "mismatched_requests": mismatched[:10]}


def stump_leg(args):
    from cs677s26_amd import workloads
    from cs677s26_amd.device import DeviceArray, synchronize
    from cs677s26_amd.engine import FaceVisionEngine
    from cs677s26_amd.trainer import AdaBoostTrainer

    patches, labels = workloads.training()
    pool = FaceVisionEngine.buildHaarFeaturePool((24, 24), FaceVisionEngine.Basic)
    n = len(labels)
    num_pos = int((labels > 0).sum())
    tr = AdaBoostTrainer((24, 24), n, 1024)
    d_p = DeviceArray.from_host(patches)
    d_f = DeviceArray.from_host(pool)

    def stage():
        synchronize()
        t0 = time.perf_counter()
        tr.stage_begin(d_p, labels, num_pos, d_f, len(pool))
        synchronize()
        t1 = time.perf_counter()
        rounds = [tr.round(0.99, 0.40) for _ in range(args.rounds)]
        synchronize()
        return t0, t1, time.perf_counter(), rounds

    # the autotuned schedule: the first stage_begin of the bucket times the
    # candidates (reported as tune_s, outside the timed stages), then one
    # untimed stage and the timed one reuse the pick; the default schedule's
    # stage is timed after it for comparison
    tr.set_autotune(True)
    synchronize()
    tt = time.perf_counter()
    tr.stage_begin(d_p, labels, num_pos, d_f, len(pool))
    synchronize()
    tune_s = time.perf_counter() - tt
    stage()
    t0, t1, t2, rounds = stage()
    scanned = tr.pruning_stats()  # features scanned in the timed stage's rounds (seed passes included)
    sched = tr.schedule()
    # the same stage with every feature scanned every round (no pruning)
    tr.set_pruning(False)
    f0, f1, f2, full_rounds = stage()
    tr.set_pruning(True)
    tr.set_schedule(None)
    d0, d1, d2, _ = stage()
    same = [(r["feature"], r["split"], r["parity"], r["err"]) for r in rounds] == \
        [(r["feature"], r["split"], r["parity"], r["err"]) for r in full_rounds]
    round_ms = 1e3 * (t2 - t1) / max(1, args.rounds)
    full_ms = 1e3 * (f2 - f1) / max(1, args.rounds)
    # the threshold scan streams each scanned feature's u16 order with its
    # split bit: N*2 + N/8 bytes per feature, over the whole round's wall time
    per_feature = n * 2 + n // 8
    scan_bytes = scanned * per_feature // max(1, args.rounds)
    full_bytes = len(pool) * per_feature
    scan_gbs = scan_bytes / (round_ms * 1e-3) / 1e9 if round_ms > 0 else 0.0
    per_feature_traffic = measured_traffic_search()
    full_gbs = full_bytes / (full_ms * 1e-3) / 1e9 if full_ms > 0 else 0.0
    return {"s_per_stage": t2 - t0, "presort_s": t1 - t0, "round_ms": round_ms,
            "pruning": {"features_scanned_per_round": scanned / max(1, args.rounds),
                        "full_scan_s_per_stage": f2 - f0, "full_scan_round_ms": full_ms,
                        "rounds_equal_full_scan": same},
            "schedule": {"wide": sched[0], "loads_ahead": sched[1], "blocks_per_cu": sched[2], "packed": sched[3],
                         "tune_s": tune_s, "default_s_per_stage": d2 - d0, "default_presort_s": d1 - d0},
            "rounds": args.rounds, "features": len(pool), "samples": n,
            "first_feature": rounds[0]["feature"] if rounds else -1,
            "search_roofline": {"bound": "hbm", "kernel": "search_kernel (whole round, pruned)", "achieved": scan_gbs,
                                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": scan_gbs / HBM_PEAK_GBS,
                                "algorithmic_bytes": scan_bytes, "avg_ms": round_ms,
                                "traffic": None if per_feature_traffic is None
                                else per_feature_traffic * scanned / max(1, args.rounds),
                                "full_scan": {"achieved": full_gbs, "frac": full_gbs / HBM_PEAK_GBS,
                                              "algorithmic_bytes": full_bytes, "avg_ms": full_ms}}}


def cpu_info():
    """The host CPU model, the cores this process may run on, and the thread
    count of the all-core runs (the job's OpenMP share when it sets one)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    nproc = len(os.sched_getaffinity(0))
    threads = nproc
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        threads = min(nproc, int(env))
    return {"cpu_model": model, "nproc": nproc, "all_core_threads": threads}


def timed_median(fn, runs=3):
    """One untimed warm-up call, then the median of ``runs`` timed calls
    (monotonic clock), as the reference's compare modes repeat their runs."""
    import statistics
    fn()
    ts = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        ts.append(time.perf_counter() - t0)
    return statistics.median(ts), ts


def cpu_leg(args):
    """The oracle's detect + grouping of the 1080p frame: one thread (the
    reference CPU path is single-threaded), then every allowed core."""
    from cs677s26_amd import workloads
    from oracle import pyoracle as O

    img, (feats, stumps, stages) = workloads.detect_1080()
    h, w = img.shape
    info = cpu_info()
    mpix = w * h / 1e6

    def run():
        raw, _ = O.detect_raw(feats, stumps, stages, img)
        O.group(raw, 4)

    O.set_threads(1)
    t1, runs1 = timed_median(run)
    O.set_threads(info["all_core_threads"])
    tn, runsn = timed_median(run)
    O.set_threads(1)
    return {"value": mpix / t1, "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": "one 1920x1080 frame, 38-stage cascade, detect + grouping, oracle C restatement; "
                      "1 warm-up, median of 3",
            "seconds": t1, "runs_s": runs1, **info,
            "all_cores": {"value": mpix / tn, "unit": "Mpix/s", "cores": info["all_core_threads"], "seconds": tn,
                          "runs_s": runsn}}


def cpu_batched_leg(args):
    """The oracle's detect + grouping over the batch's images in order,
    skipping any that would take the sample past 1.6 Mpix, until 1.2 Mpix are
    covered: the CPU path's Mpix/s on the batched workload's image mix, one
    thread and every allowed core."""
    from cs677s26_amd import workloads
    from oracle import pyoracle as O

    sizes, imgs = workloads.batch64()
    feats, stumps, stages = workloads.cascade_38()
    info = cpu_info()
    pick, mpix = [], 0.0
    for (w, h), im in zip(sizes, imgs):
        if mpix + w * h / 1e6 > 1.6:
            continue
        pick.append(im)
        mpix += w * h / 1e6
        if mpix >= 1.2:
            break

    def run():
        for im in pick:
            raw, _ = O.detect_raw(feats, stumps, stages, im)
            O.group(raw, 4)

    O.set_threads(1)
    t1, runs1 = timed_median(run)
    O.set_threads(info["all_core_threads"])
    tn, runsn = timed_median(run)
    O.set_threads(1)
    return {"value": mpix / t1, "unit": "Mpix/s", "cores": 1, "kind": "port",
            "sample": f"{len(pick)} of the 64 batch images ({mpix:.2f} Mpix, in order, none past a 1.6 Mpix "
                      "total), 38-stage cascade, detect + grouping per image, oracle C restatement; 1 warm-up, "
                      "median of 3", "seconds": t1, "runs_s": runs1, **info,
            "all_cores": {"value": mpix / tn, "unit": "Mpix/s", "cores": info["all_core_threads"], "seconds": tn,
                          "runs_s": runsn}}


def measured_traffic(name):
    """HBM bytes per dispatch of a kernel on a workload: the committed rocprofv3 --pmc
    summary's FETCH_SIZE + WRITE_SIZE, each scaled by the same counter's
    reading for a plain copy kernel of known size (None if not measured)."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_dispatch:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def measured_traffic_search():
    """HBM bytes per scanned feature of search_kernel from the committed
    rocprofv3 --pmc summary of the stump search (None if not measured)."""
    path = os.path.join(ROOT, "profiles", "stump_search_162336x20000_kernels.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_feature (search_kernel):"):
                    return float(line.split(":", 1)[1])
    except OSError:
        return None
    return None


def cpu_stump_leg(args):
    """The oracle's stump search (presort + exact sorted sweep, the
    reference's runCpuWeakSearch semantics) over the first features of the
    162,336-feature pool and all 20,000 samples, timed for one presort and
    one round and scaled to the full pool and 50 rounds (extrapolated, as the
    reference's own compare mode samples features): 2,048 features on one
    thread, 16,384 on every allowed core."""
    import numpy as np

    from cs677s26_amd import workloads
    from oracle import pyoracle as O

    patches, labels = workloads.training()
    info = cpu_info()
    sumT, _, inv = O.train_integrals(patches)
    pool = O.haar_pool(24, 0)
    n = len(labels)
    w = np.full(n, 1.0 / n)
    sw = np.where(labels > 0, w, -w).astype(np.float32)

    def measure(nfeat, threads):
        O.set_threads(threads)
        sub = pool[:nfeat]
        box = {}

        def presort():
            box["ords"] = O.presort(sub, sumT, inv)

        def round_():
            O.select(O.search(box["ords"], sw))

        tp, runs_p = timed_median(presort)
        tr, runs_r = timed_median(round_)
        scale = len(pool) / nfeat
        return tp * scale + args.rounds * tr * scale, tp, tr, runs_p, runs_r

    v1, tp1, tr1, rp1, rr1 = measure(2048, 1)
    vn, tpn, trn, rpn, rrn = measure(16384, info["all_core_threads"])
    O.set_threads(1)
    return {"value": v1, "unit": "s/stage (extrapolated)", "cores": 1, "kind": "port",
            "sample": f"2048 of {len(pool)} features x {n} samples: presort {tp1:.3f} s and one round "
                      f"{tr1:.4f} s (each 1 warm-up, median of 3), scaled to the full pool and {args.rounds} "
                      "rounds",
            "higher_is_better": False, "presort_runs_s": rp1, "round_runs_s": rr1, **info,
            "all_cores": {"value": vn, "unit": "s/stage (extrapolated)", "cores": info["all_core_threads"],
                          "sample": f"16384 of {len(pool)} features: presort {tpn:.3f} s, one round {trn:.4f} s",
                          "presort_runs_s": rpn, "round_runs_s": rrn}}


def detect_fields(args):
    d = detect_leg(args)
    out = {
        "value": d["mpix"] * args.steps / d["wall_s"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": d["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "detect 1920x1080 u8, seeded 200 blobs 24-400 px, 38-stage 6061-stump cascade, "
                               "sf 1.25, minNeighbors 4, minObjectSize 24, maxDetections 200000, plain path",
                   "levels": d["stats"]["levels"], "windows": d["stats"]["windows"],
                   "raw_detections": d["stats"]["rawDetections"], "detections": d["detections"]},
    }
    casc_ms = d["phase_ms"][2]
    casc_bytes = d["stats"]["cascadeBytes"]
    achieved = casc_bytes / (casc_ms * 1e-3) / 1e9 if casc_ms > 0 else 0.0
    out["roofline"] = {"bound": "hbm", "kernel": "cascade_kernel+deep_list_kernel+cascade_deep_kernel+cascade_tail_kernel",
                       "achieved": achieved, "peak": HBM_PEAK_GBS,
                       "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                       "traffic": measured_traffic("cascade_kernel_detect1080"),
                       "algorithmic_bytes": casc_bytes, "avg_ms": casc_ms}
    # the same phase against the LDS: its algorithmic gather bytes over the
    # chip's LDS bandwidth (256 CUs x 128 B per clock x 2.4 GHz)
    lds_gbs = d["gather_bytes"] / (casc_ms * 1e-3) / 1e9 if casc_ms > 0 else 0.0
    out["lds_roofline"] = {"bound": "lds", "achieved": lds_gbs, "peak": LDS_PEAK_GBS, "unit": "GB/s",
                           "frac": lds_gbs / LDS_PEAK_GBS, "gather_bytes": d["gather_bytes"],
                           "evals": d["evals"], "stage_windows": d["stage_windows"], "avg_ms": casc_ms}
    int_ms = d["phase_ms"][1] + d["phase_ms"][0]
    int_bytes = d["stats"]["integralBytes"]
    out["phases_ms"] = {"pyramid": d["phase_ms"][0], "integral": d["phase_ms"][1], "cascade": d["phase_ms"][2],
                        "compaction": d["phase_ms"][3]}
    out["integral_roofline"] = {"achieved": int_bytes / (int_ms * 1e-3) / 1e9 if int_ms > 0 else 0.0,
                                "peak": HBM_PEAK_GBS, "unit": "GB/s", "algorithmic_bytes": int_bytes,
                                "avg_ms": int_ms}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rounds", type=int, default=50)
    ap.add_argument("--requests", type=int, default=240, help="served leg: requests in the Poisson stream")
    ap.add_argument("--qps", type=float, default=100.0, help="served leg: arrival rate")
    ap.add_argument("--slots", type=int, default=3, help="served leg: device slots")
    ap.add_argument("--max-batch", type=int, default=8, help="served leg: requests per slot batch")
    ap.add_argument("--skip", default="", help="comma list of legs to skip: detect,batched,served,stump,cpu")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X; "
              "run without --gpus or with --gpus 1", file=sys.stderr)
        sys.exit(2)
    skip = set(x for x in args.skip.split(",") if x)

    out = {"metric": METRIC, "value": None}
    if "detect" not in skip:
        log("detect 1080p / 38-stage")
        out.update(detect_fields(args))
    if "batched" not in skip:
        log("batched 64 images")
        b = batched_leg(args)
        out["batched_mpix_per_s"] = b["mpix_per_s"]
        out["batched"] = b
    if "stump" not in skip:
        log("stump search stage (162,336 x 20,000, 50 rounds)")
        s = stump_leg(args)
        out["stump_search_s_per_stage"] = s["s_per_stage"]
        out["stump"] = s
    out["p99_latency_ms"] = None
    if "served" not in skip:
        log(f"served: Poisson {args.requests} requests at {args.qps} QPS, 240p..4K")
        v = served_leg(args)
        out["served"] = v
        out["p99_latency_ms"] = v["p99_ms"]
    if "cpu" not in skip:
        log("cpu baseline (oracle, 1 thread and all cores, warm-up + median of 3)")
        out["cpu_baseline"] = cpu_leg(args)
        log("cpu baseline: stump search (oracle, 2,048 features on 1 thread, 16,384 on all cores)")
        out["cpu_baseline_stump"] = cpu_stump_leg(args)
        log("cpu baseline: batched workload images (oracle, about 1.2 Mpix, 1 thread and all cores)")
        out["cpu_baseline_batched"] = cpu_batched_leg(args)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
