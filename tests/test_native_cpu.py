"""Host-side checks of libvj_hip.so that need no GPU: exports, the Haar pool,
the .bin model format, grouping, PGM/PPM I/O and CLI error handling."""
import os
import struct
import subprocess

import numpy as np
import pytest

from cs677s26_amd import _native, synth
from cs677s26_amd._native import VJError
from cs677s26_amd.engine import FaceVisionEngine, group_rectangles, loadCascadeModel, saveCascadeModel
from cs677s26_amd.vjtypes import DETECTION, det_tuples
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    names = _native.header_symbols()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_pool_equals_oracle(mode):
    got = FaceVisionEngine.buildHaarFeaturePool((24, 24), mode)
    want = O.haar_pool(24, mode)
    assert got.tobytes() == want.tobytes()


def test_model_roundtrip_and_header_bytes(tmp_path):
    feats, stumps, stages = synth.blob_cascade((1, 3), seed=9, n_sample=200)
    p = tmp_path / "m.bin"
    saveCascadeModel(p, (24, 24), feats, stumps, stages)
    raw = p.read_bytes()
    # header: magic, u32 version 1, i32 winW, i32 winH, 4 zero bytes, u64 x 3
    assert raw[:8] == b"VJCBIN1\0"
    assert struct.unpack("<IiiI", raw[8:24]) == (1, 24, 24, 0)
    assert struct.unpack("<QQQ", raw[24:48]) == (len(feats), len(stumps), len(stages))
    assert len(raw) == 48 + 64 * len(feats) + 16 * len(stumps) + 12 * len(stages)
    assert raw[48:48 + 64 * len(feats)] == feats.tobytes()
    win, f2, s2, g2 = loadCascadeModel(p)
    assert win == (24, 24) and f2.tobytes() == feats.tobytes()
    assert s2.tobytes() == stumps.tobytes() and g2.tobytes() == stages.tobytes()


def test_model_loader_rejects_bad_files(tmp_path):
    feats, stumps, stages = synth.blob_cascade((1,), seed=9, n_sample=100)
    good = tmp_path / "good.bin"
    saveCascadeModel(good, (24, 24), feats, stumps, stages)
    raw = bytearray(good.read_bytes())
    cases = {
        "magic": raw[:0] + b"VJCBIN2\0" + raw[8:],
        "version": raw[:8] + struct.pack("<I", 2) + raw[12:],
        "window": raw[:12] + struct.pack("<i", 0) + raw[16:],
        "zero_count": raw[:24] + struct.pack("<Q", 0) + raw[32:],
        "truncated": raw[:-5],
        "short_header": raw[:20],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.bin"
        p.write_bytes(bytes(data))
        with pytest.raises(VJError) as ei:
            loadCascadeModel(p)
        assert ei.value.status == _native.VJ_INVALID_ARG, name
    with pytest.raises(VJError):
        saveCascadeModel(tmp_path / "empty.bin", (24, 24), feats[:0], stumps, stages)


def test_grouping_equals_oracle_on_random_boxes():
    rng = np.random.default_rng(3)
    for trial in range(40):
        n = int(rng.integers(0, 400))
        d = np.zeros(n, DETECTION)
        centres = rng.integers(0, 600, size=(max(1, n // 8 + 1), 2))
        for i in range(n):
            cx, cy = centres[rng.integers(0, len(centres))]
            s = int(rng.integers(24, 90))
            d[i] = (cx + rng.integers(-6, 7), cy + rng.integers(-6, 7), s, s, int(rng.integers(0, 9)),
                    np.float32(rng.normal()))
        mn = int(rng.integers(0, 6))
        assert det_tuples(group_rectangles(d, mn)) == det_tuples(O.group(d, mn)), trial


def test_grouping_sweep_order_negative_and_wide_x():
    """The sweep order is a counting sort over x for moderate x ranges and a
    key sort otherwise: negative x, and x spread over 10^7 (the key sort),
    group as the oracle groups them."""
    rng = np.random.default_rng(5)
    for trial, (lo, span) in enumerate([(-500, 1200), (-3000000, 10000000), (0, 1), (2**30, 5000)]):
        n = 600
        d = np.zeros(n, DETECTION)
        centres = np.stack([rng.integers(lo, lo + span, 60), rng.integers(0, 800, 60)], 1)
        for i in range(n):
            cx, cy = centres[rng.integers(0, len(centres))]
            s = int(rng.integers(24, 90))
            d[i] = (cx + rng.integers(-5, 6), cy + rng.integers(-5, 6), s, s, 0, np.float32(rng.normal()))
        assert det_tuples(group_rectangles(d, 2)) == det_tuples(O.group(d, 2)), trial


def test_grouping_many_lists_equals_one_by_one():
    """vj_group_rectangles_many (every list's steps on the host pool) returns
    for each list what vj_group_rectangles returns for it alone: lists of 0,
    1, a few hundred and several thousand boxes."""
    import ctypes as C
    from cs677s26_amd._native import lib
    rng = np.random.default_rng(12)
    g = np.load(os.path.join(ROOT, "tests", "golden", "detect_c3.npz"))
    lists = [np.zeros(0, DETECTION), g["raw"][:1].copy(), g["raw"], g["raw"][::-1].copy()]
    for n in (300, 2500, 7000):
        d = np.zeros(n, DETECTION)
        centres = rng.integers(0, 3000, size=(n // 20 + 1, 2))
        for i in range(n):
            cx, cy = centres[rng.integers(0, len(centres))]
            s = int(rng.integers(24, 120))
            d[i] = (cx + rng.integers(-8, 9), cy + rng.integers(-8, 9), s, s, 0, np.float32(rng.normal()))
        lists.append(d)
    k = len(lists)
    outs = [np.zeros(max(1, len(x)), DETECTION) for x in lists]
    ins = (C.c_void_p * k)(*[x.ctypes.data if len(x) else None for x in lists])
    ons = (C.c_void_p * k)(*[o.ctypes.data for o in outs])
    ns = (C.c_int * k)(*[len(x) for x in lists])
    cnt = (C.c_int * k)()
    assert lib().vj_group_rectangles_many(ins, ns, k, 3, C.c_float(0.2), ons, cnt) == 0
    for x, o, c in zip(lists, outs, cnt):
        assert det_tuples(o[:c]) == det_tuples(group_rectangles(x, 3))


def test_crop_resize_equals_oracle():
    """Mined-box crops (src/main.cu:245-284) equal the oracle's, byte for
    byte, for boxes inside, across and beyond the image border."""
    from cs677s26_amd.engine import crop_resize
    rng = np.random.default_rng(6)
    img = rng.integers(0, 256, size=(120, 161), dtype=np.uint8)
    for _ in range(200):
        s = int(rng.integers(1, 140))
        box = (int(rng.integers(-20, 160)), int(rng.integers(-20, 120)), s, int(rng.integers(1, 140)))
        assert np.array_equal(crop_resize(img, box), O.crop_resize(img, box)), box


def test_pgm_reader(tmp_path):
    from cs677s26_amd._native import lib
    import ctypes as C
    img = np.arange(35 * 20, dtype=np.uint32).astype(np.uint8).reshape(20, 35)
    p = tmp_path / "a.pgm"
    p.write_bytes(b"P5\n# a comment\n35 20\n# another\n255\n" + img.tobytes())
    w, h = C.c_int(), C.c_int()
    assert lib().vj_pgm_peek(str(p).encode(), C.byref(w), C.byref(h)) == 0 and (w.value, h.value) == (35, 20)
    out = np.zeros(700, np.uint8)
    assert lib().vj_pgm_load(str(p).encode(), out.ctypes.data, out.nbytes) == 0
    assert np.array_equal(out.reshape(20, 35), img)
    for bad in (b"P2\n35 20\n255\n", b"P5\n35 20\n65535\n", b"P5\n0 20\n255\n"):
        q = tmp_path / "b.pgm"
        q.write_bytes(bad + bytes(700))
        assert lib().vj_pgm_peek(str(q).encode(), C.byref(w), C.byref(h)) == _native.VJ_INVALID_ARG


def test_sample_files_and_pgm_list(tmp_path):
    """loadRawSampleBin (src/main.cu:122-162) and collectPgmPaths
    (src/main.cu:226-244) through the library: patch counts, the reference's
    error messages, and every *.pgm under a tree in sorted order."""
    from cs677s26_amd._native import lib
    import ctypes as C
    px = 24 * 24
    data = np.random.default_rng(3).integers(0, 256, 3 * px, dtype=np.uint8)
    good = tmp_path / "s.bin"
    good.write_bytes(data.tobytes())
    n = C.c_int()
    assert lib().vj_samples_count(str(good).encode(), px, C.byref(n)) == 0 and n.value == 3
    out = np.zeros(3 * px, np.uint8)
    assert lib().vj_samples_load(str(good).encode(), px, out.ctypes.data, out.nbytes, C.byref(n)) == 0
    assert n.value == 3 and np.array_equal(out, data)
    assert lib().vj_samples_load(str(good).encode(), px, out.ctypes.data, out.nbytes - 1,
                                 C.byref(n)) == _native.VJ_INVALID_ARG
    (tmp_path / "empty.bin").write_bytes(b"")
    (tmp_path / "ragged.bin").write_bytes(bytes(px + 1))
    for name, msg in (("missing.bin", "Cannot open: "), ("empty.bin", "Empty file: "),
                      ("ragged.bin", "File size is not divisible by samplePixels: ")):
        path = str(tmp_path / name)
        assert lib().vj_samples_count(path.encode(), px, C.byref(n)) == _native.VJ_INVALID_ARG
        assert lib().vj_last_error().decode() == msg + path

    root = tmp_path / "imgs"
    for rel in ("b.pgm", "a/x.PGM", "a/y.txt", "a/deep/z.pgm", "c.pgm.bak", "a/deep/.pgm"):
        (root / rel).parent.mkdir(parents=True, exist_ok=True)
        (root / rel).write_bytes(b"P5\n1 1\n255\n\0")
    need, cnt = C.c_size_t(), C.c_int()
    assert lib().vj_pgm_list(str(root).encode(), None, 0, C.byref(need), C.byref(cnt)) == 0
    buf = C.create_string_buffer(need.value)
    assert lib().vj_pgm_list(str(root).encode(), buf, need.value, C.byref(need), C.byref(cnt)) == 0
    got = buf.raw[:need.value].split(b"\0")[:-1]
    want = sorted(str(root / r).encode() for r in ("b.pgm", "a/x.PGM", "a/deep/z.pgm"))
    assert cnt.value == 3 and got == want
    assert lib().vj_pgm_list(str(tmp_path / "nope").encode(), None, 0, C.byref(need), C.byref(cnt)) == 0
    assert cnt.value == 0 and need.value == 0


def test_no_device_is_reported_not_faked():
    if _native.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(VJError) as ei:
        FaceVisionEngine((24, 24), 1, 64, 64)
    assert ei.value.status == _native.VJ_DEVICE_ERROR


def test_facade_compiles_and_constructors_throw_without_a_device():
    """include/vj_facade.hpp (the reference's FaceVisionEngine / AdaBoostTrainer
    over the C ABI) compiles with -Wall -Wextra -Werror in reference-style call
    sites (tests/facade/reference_callers.cpp) and links with libvj_hip.so;
    without a device both constructors throw std::runtime_error with the
    library's message, as the reference's constructors throw."""
    subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "tests", "facade")])
    exe = os.path.join(ROOT, "tests", "facade", "build", "reference_callers")
    out = subprocess.run([exe, "nodevice"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    assert "pool 162336" in out.stdout
    if _native.device_count() == 0:
        assert "FaceVisionEngine threw: FaceVisionEngine: vj_engine_create: no HIP device" in out.stdout
        assert "AdaBoostTrainer threw: AdaBoostTrainer: vj_trainer_create: no HIP device" in out.stdout
    usage = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert usage.returncode == 2 and "usage" in usage.stderr


def test_pair_capacity_rule(tmp_path):
    """The device grouping's pair-buffer rule (csrc/vj_pair_plan.hpp): the
    device result is used only when every pair was stored; an overflow sends
    the call to the host and grows the buffer for the next call by a quarter
    plus 64K entries, never past the 2^30-entry cap (beyond it such calls
    stay on the host) and never shrinking; a pair count of 2^31 or more
    (which 32-bit offsets could not address) is an overflow like any other."""
    src = tmp_path / "plan.cpp"
    src.write_text(r"""
#include <cstdio>
#include "vj_pair_plan.hpp"
int main() {
    using vj::pair_plan;
    int64_t next = 0;
    const struct { int64_t n, cap; int use; int64_t next; } cases[] = {
        {0, 100, 1, 100}, {100, 100, 1, 100}, {101, 100, 0, 101 + 25 + 65536},
        {1000000, 2000000, 1, 2000000}, {int64_t{1} << 31, 1000, 0, vj::kPairCapMax},
        {int64_t{3} << 30, vj::kPairCapMax, 0, vj::kPairCapMax}, {int64_t{1} << 29, int64_t{1} << 29, 1, int64_t{1} << 29},
    };
    for (const auto& c : cases) {
        const int use = pair_plan(c.n, c.cap, &next) ? 1 : 0;
        if (use != c.use || next != c.next) {
            std::printf("fail n=%lld cap=%lld use=%d next=%lld\n", (long long)c.n, (long long)c.cap, use, (long long)next);
            return 1;
        }
    }
    std::printf("ok\n");
    return 0;
}
""")
    exe = tmp_path / "plan"
    subprocess.check_call(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "cs677s26_amd", "csrc"), str(src), "-o",
                           str(exe)])
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and out.stdout.strip() == "ok", out.stdout


def test_cli_usage_and_errors(tmp_path):
    exe = os.path.join(ROOT, "cs677s26_amd", "bin", "cuda_hello")
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage" in r.stderr
    r = subprocess.run([exe, "bogus"], capture_output=True, text=True)
    assert r.returncode == 1 and "[ERR] Unknown mode" in r.stderr
    r = subprocess.run([exe, "detect", str(tmp_path / "missing.bin"), "x.pgm", "y.ppm"], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "[ERR] loadCascadeModel failed" in r.stderr
    r = subprocess.run([exe, "train", str(tmp_path / "m.bin")], capture_output=True, text=True)
    assert r.returncode == 1 and "stageNegCacheDir is required" in r.stderr


def test_grouping_large_input_equals_golden():
    """4,488 raw boxes of the 1080p workload (the parallel pair search runs
    above 1,024 boxes): grouped boxes equal the oracle's golden fixture."""
    g = np.load(os.path.join(ROOT, "tests", "golden", "detect_c3.npz"))
    assert det_tuples(group_rectangles(g["raw"], 4)) == det_tuples(g["grouped"])
    assert det_tuples(group_rectangles(g["raw"], 4)) == det_tuples(O.group(g["raw"], 4))
    # a shuffled copy: the same set of boxes in another input order groups
    # exactly as the oracle groups that order
    rng = np.random.default_rng(4)
    sh = g["raw"][rng.permutation(len(g["raw"]))]
    assert det_tuples(group_rectangles(sh, 3)) == det_tuples(O.group(sh, 3))


# ---- CPU legs of the compare verbs (src/main.cu:1032-1368) -----------------

def test_cpu_detect_leg_equals_oracle_and_golden():
    """detect_compare's host leg (detectMultiScaleCpu, src/main.cu:684-827)
    gives the oracle's raw detections, in order, on the 640x480 workload and
    under a small cap (the early return at maxDetections)."""
    from cs677s26_amd.engine import detect_multiscale_cpu
    g = np.load(os.path.join(ROOT, "tests", "golden", "detect_c1.npz"))
    feats, stumps, stages = synth.blob_cascade((1, 10, 25))
    img = synth.detect_image(640, 480, 20, 24, 64)
    got = detect_multiscale_cpu(feats, stumps, stages, img)
    assert det_tuples(got) == det_tuples(g["raw"])
    want, _ = O.detect_raw(feats, stumps, stages, img)
    assert det_tuples(got) == det_tuples(want)
    for cap in (100, 257):
        got = detect_multiscale_cpu(feats, stumps, stages, img, maxDetections=cap)
        assert det_tuples(got) == det_tuples(want[:cap])


def test_cpu_detect_leg_max_scales_and_parameters():
    """maxScales stops the scale loop at that index (skipped levels count);
    a larger minimum object size skips leading levels; scale 1.1 and a
    non-multiple-of-4 image follow the oracle too."""
    from cs677s26_amd.engine import detect_multiscale_cpu
    feats, stumps, stages = synth.blob_cascade((1, 10, 25))
    img = synth.detect_image(333, 251, 8, 24, 64, seed=3)
    for sf, mo in ((1.25, 24), (1.1, 40), (1.5, 30)):
        want, _ = O.detect_raw(feats, stumps, stages, img, sf=sf, min_obj=mo)
        got = detect_multiscale_cpu(feats, stumps, stages, img, scaleFactor=sf, minObjectSize=mo)
        assert det_tuples(got) == det_tuples(want), (sf, mo)
        for ms in (1, 3):
            got = detect_multiscale_cpu(feats, stumps, stages, img, scaleFactor=sf, minObjectSize=mo, maxScales=ms)
            assert det_tuples(got) == det_tuples(want[want["scaleIdx"] < ms]), (sf, mo, ms)
    # a stump that names a feature outside the pool is refused
    bad = stumps.copy()
    bad["feature"][0] = len(feats)
    with pytest.raises(VJError):
        detect_multiscale_cpu(feats, bad, stages, img)


@pytest.mark.parametrize("weights", ["uniform", "random"])
def test_cpu_search_leg_equals_oracle(weights):
    """train_compare's host leg: the best stump of features [0, F) equals the
    oracle's per-feature exact search followed by the per-op selection
    (float errors, smaller index on ties), with theta from the oracle's
    responses.  Uniform weights make exact ties between features common."""
    from cs677s26_amd.trainer import cpu_best_weak
    patches, labels = synth.training_set(150, 150, seed=5)
    n = len(labels)
    pool = FaceVisionEngine.buildHaarFeaturePool((24, 24), FaceVisionEngine.Core)[:1500]
    rng = np.random.default_rng(2)
    if weights == "uniform":
        w = np.full(n, 0.5 / 150, np.float32)
    else:
        w = rng.random(n).astype(np.float32)
        w = (w / w.sum()).astype(np.float32)
    active = (rng.random(n) > 0.1).astype(np.uint8)
    for act in (None, active):
        got = cpu_best_weak(patches, labels, w, pool, active=act)
        sumT, _, inv = O.train_integrals(patches)
        ords = O.presort(pool, sumT, inv)
        sw = np.where(labels > 0, w, -w).astype(np.float32)
        if act is not None:
            sw = np.where(act != 0, sw, np.float32(0))
        best = O.search(ords, sw)
        errs = [np.float32(np.ldexp(float(b["errQ"]), -62)) if b["split"] >= 0 else np.float32(np.finfo(np.float32).max)
                for b in best]
        fi = min(range(len(best)), key=lambda i: (errs[i], i))
        assert got["featureIdx"] == fi
        assert got["bestErr"] == errs[fi]
        assert got["bestParity"] == best[fi]["parity"]
        v = O.feature_responses(pool[fi], sumT, inv)
        k = int(best[fi]["split"])
        s0, s1 = int(ords[fi, k]) & 0x7FFF, int(ords[fi, k + 1]) & 0x7FFF
        assert got["bestTheta"] == np.float32(np.float32(0.5) * (v[s0] + v[s1]))


def test_cli_compare_verbs_usage_and_errors(tmp_path):
    exe = os.path.join(ROOT, "cs677s26_amd", "bin", "cuda_hello")
    r = subprocess.run([exe, "train_compare", "a.bin"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: cuda_hello train_compare" in r.stderr
    r = subprocess.run([exe, "train_compare", str(tmp_path / "p.bin"), str(tmp_path / "n.bin")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "[ERR] Cannot open" in r.stderr
    (tmp_path / "odd.bin").write_bytes(bytes(577))
    r = subprocess.run([exe, "train_compare", str(tmp_path / "odd.bin"), str(tmp_path / "odd.bin")],
                       capture_output=True, text=True)
    assert r.returncode == 1 and "not divisible by samplePixels" in r.stderr
    r = subprocess.run([exe, "detect_compare", "m.bin", "a.pgm"], capture_output=True, text=True)
    assert r.returncode == 1 and "Usage: cuda_hello detect_compare" in r.stderr
    r = subprocess.run([exe, "detect_compare", str(tmp_path / "m.bin"), "a.pgm", "o.ppm"], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "[ERR] loadCascadeModel failed" in r.stderr
