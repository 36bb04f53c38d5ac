// cli.cpp -- cuda_hello drop-in command line (src/main.cu:2581-2615).
//
//   cuda_hello detect <in_model.bin> <image.pgm> <out.ppm>
//              [scaleFactor=1.25] [minNeighbors=4] [minObjectSize=24]
//              [maxDetections=200000] [benchmarkRuns=0]
//   cuda_hello train  [faces_u8.bin] [numNeg] [maxWeakPerStage] [maxDetections]
//              [minNeighbors] [scaleFactor] [maxStages] [minHitRate]
//              [maxFalseAlarm] [minObjectSize] [stageNegCacheDir]
//              [stageNegImagesDir] [stageNegSeed] [hardNegCandidateMultiplier]
//              <out_model.bin>
//   cuda_hello train_compare <faces_u8.bin> <negatives_u8.bin> [numPos=256]
//              [numNeg=256] [featureLimit=1024] [runs=3]
//   cuda_hello detect_compare <in_model.bin> <image.pgm> <out.ppm>
//              [scaleFactor=1.25] [minObjectSize=24] [maxDetections=200000]
//              [maxScales=0] [runs=3]
//
// Same verbs, positional order, defaults, clamps and stdout tags as the
// reference (src/main.cu:1032-1368, 1370-1531, 2394-2579); exit status 0 or 1.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <limits>
#include <numeric>
#include <random>
#include <string>
#include <vector>

#include "vj_api.h"

namespace {

int fail(const char* what) {
    std::cerr << "[ERR] " << what << ": " << vj_last_error() << "\n";
    return 1;
}

struct DevMem {
    void* p = nullptr;
    ~DevMem() {
        if (p) vj_free(p);
    }
};

struct Model {
    int winW = 0, winH = 0;
    std::vector<vj_feature> features;
    std::vector<vj_stump> stumps;
    std::vector<vj_stage> stages;
};

bool load_model(const std::string& path, Model& m) {
    int64_t nF = 0, nS = 0, nSt = 0;
    if (vj_model_peek(path.c_str(), &m.winW, &m.winH, &nF, &nS, &nSt) != VJ_OK) return false;
    m.features.resize(static_cast<size_t>(nF));
    m.stumps.resize(static_cast<size_t>(nS));
    m.stages.resize(static_cast<size_t>(nSt));
    return vj_model_load(path.c_str(), m.features.data(), nF, m.stumps.data(), nS, m.stages.data(), nSt) == VJ_OK;
}

// src/main.cu:2394-2579
int run_detect(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "Usage: cuda_hello detect <in_model.bin> <image.pgm> <out.ppm>"
                  << " [scaleFactor=1.25] [minNeighbors=4] [minObjectSize=24] [maxDetections=200000] [benchmarkRuns=0]\n";
        return 1;
    }
    const std::string modelPath = argv[1], imagePath = argv[2], outPath = argv[3];
    float scaleFactor = 1.25f;
    int minNeighbors = 4, minObjectSize = 24, maxDetections = 200000, benchmarkRuns = 0;
    try {
        if (argc > 4) scaleFactor = std::stof(argv[4]);
        if (argc > 5) minNeighbors = std::max(1, std::stoi(argv[5]));
        if (argc > 6) minObjectSize = std::max(1, std::stoi(argv[6]));
        if (argc > 7) maxDetections = std::max(100, std::stoi(argv[7]));
        if (argc > 8) benchmarkRuns = std::max(0, std::stoi(argv[8]));
    } catch (...) {
        std::cerr << "[ERR] invalid numeric argument\n";
        return 1;
    }
    Model m;
    if (!load_model(modelPath, m)) {
        std::cerr << "[ERR] loadCascadeModel failed: " << modelPath << "\n";
        return 1;
    }
    int w = 0, h = 0;
    if (vj_pgm_peek(imagePath.c_str(), &w, &h) != VJ_OK) return fail("loadPGM");
    std::vector<uint8_t> img(static_cast<size_t>(w) * h);
    if (vj_pgm_load(imagePath.c_str(), img.data(), img.size()) != VJ_OK) return fail("loadPGM");

    vj_engine* eng = nullptr;
    if (vj_engine_create(m.winW, m.winH, 1, std::max(m.winW, w), std::max(m.winH, h), 0, maxDetections, &eng) != VJ_OK)
        return fail("FaceVisionEngine");
    struct Guard {
        vj_engine* e;
        ~Guard() { vj_engine_destroy(e); }
    } guard{eng};
    if (vj_engine_upload_feature_pool(eng, m.features.data(), static_cast<int>(m.features.size()), nullptr) != VJ_OK)
        return fail("uploadFeaturePool");
    DevMem dStumps, dImage;
    if (vj_malloc(&dStumps.p, m.stumps.size() * sizeof(vj_stump)) != VJ_OK ||
        vj_malloc(&dImage.p, img.size()) != VJ_OK ||
        vj_memcpy_htod(dStumps.p, m.stumps.data(), m.stumps.size() * sizeof(vj_stump), nullptr) != VJ_OK ||
        vj_memcpy_htod(dImage.p, img.data(), img.size(), nullptr) != VJ_OK)
        return fail("cudaMemcpy to device");
    const vj_feature* dPool = nullptr;
    int poolN = 0;
    vj_engine_feature_pool(eng, &dPool, &poolN);
    if (vj_engine_set_cascade(eng, dPool, poolN, static_cast<const vj_stump*>(dStumps.p),
                              static_cast<int>(m.stumps.size()), m.stages.data(), static_cast<int>(m.stages.size()),
                              nullptr) != VJ_OK)
        return fail("setCascadeModel");

    std::vector<vj_detection> dets(static_cast<size_t>(maxDetections));
    int count = 0;
    auto detect_once = [&]() {
        return vj_engine_detect(eng, static_cast<const uint8_t*>(dImage.p), w, h, w, scaleFactor, minNeighbors,
                                minObjectSize, maxDetections, 1, dets.data(), static_cast<int>(dets.size()), &count,
                                nullptr);
    };
    if (benchmarkRuns > 0) {
        for (int i = 0; i < 3; ++i)
            if (detect_once() != VJ_OK) return fail("warmup detectMultiScale");
        vj_event a = nullptr, b = nullptr;
        vj_event_create(&a);
        vj_event_create(&b);
        float sum = 0.0f, mn = std::numeric_limits<float>::infinity(), mx = 0.0f;
        for (int i = 0; i < benchmarkRuns; ++i) {
            vj_event_record(a, nullptr);
            if (detect_once() != VJ_OK) return fail("detectMultiScale");
            vj_event_record(b, nullptr);
            vj_event_synchronize(b);
            float ms = 0.0f;
            vj_event_elapsed_ms(a, b, &ms);
            sum += ms;
            mn = std::min(mn, ms);
            mx = std::max(mx, ms);
        }
        vj_event_destroy(a);
        vj_event_destroy(b);
        const float avg = sum / static_cast<float>(benchmarkRuns);
        std::cout << "[DETECT_BENCH] runs=" << benchmarkRuns << " detections(last)=" << count << " avg(ms): " << avg
                  << " min(ms): " << mn << " max(ms): " << mx << " avg(fps): " << (avg > 0.0f ? 1000.0f / avg : 0.0f)
                  << "\n";
    } else {
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (detect_once() != VJ_OK) return fail("detectMultiScale");
        const auto t1 = std::chrono::high_resolution_clock::now();
        std::cout << "[DETECT] detections=" << count << " elapsed(ms): "
                  << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count() << "\n";
    }
    const int shown = std::min(count, static_cast<int>(dets.size()));
    if (vj_ppm_save_gray_boxes(outPath.c_str(), img.data(), w, h, dets.data(), shown) != VJ_OK) {
        std::cerr << "[ERR] savePPM failed\n";
        return 1;
    }
    std::cout << "[OK] Saved result to: " << outPath << "\n";
    return 0;
}

void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " train ... <out_model.bin>\n";
    std::cerr << "   or: " << prog << " detect <in_model.bin> <image.pgm> <out.ppm>\n";
    std::cerr << "   or: " << prog << " train_compare <faces_u8.bin> <negatives_u8.bin>\n";
    std::cerr << "   or: " << prog << " detect_compare <in_model.bin> <image.pgm> <out.ppm>\n";
}

}  // namespace

int run_train(int argc, char** argv);
int run_train_compare(int argc, char** argv);
int run_detect_compare(int argc, char** argv);

int main(int argc, char** argv) {
    if (argc < 2) {
        usage(argv[0]);
        return 1;
    }
    const std::string mode = argv[1];
    if (mode == "detect") return run_detect(argc - 1, argv + 1);
    if (mode == "train") return run_train(argc - 1, argv + 1);
    if (mode == "train_compare") return run_train_compare(argc - 1, argv + 1);
    if (mode == "detect_compare") return run_detect_compare(argc - 1, argv + 1);
    std::cerr << "[ERR] Unknown mode: " << mode << "\n";
    usage(argv[0]);
    return 1;
}

// ---------------------------------------------------------------------------
// train (src/main.cu:1370-2392).  Positives from faces_u8.bin ([N][24][24]
// u8), a seeded subset per stage.  Negatives per stage as the reference
// gathers them (src/main.cu:1765-1925):
//   * stage 0: <stageNegCacheDir>/stage_negatives/non_faces_stage00.bin when
//     it holds enough patches, else random square crops of the PGM images in
//     <stageNegImagesDir>, written to that file (the reference runs its
//     python sampler for this, scripts/stage_negatives.py:149-228);
//   * later stages: hard negatives mined with the cascade trained so far --
//     raw detections on the shuffled PGM images, the 256 best-scoring boxes
//     of each image cropped and resized to the window (cropResizeTo24) --
//     topped up, when too few, by random picks from a fresh sample of
//     max(numNeg, missing * hardNegCandidateMultiplier) crops.
// The boosting rounds run on the fused trainer (presort once per stage).
// ---------------------------------------------------------------------------
namespace {

// A raw sample file through the library (count patches of px bytes); the
// library's message names the failure.
bool load_bin(const std::string& path, int px, std::vector<uint8_t>& out, int& count, bool quiet = false) {
    int n = 0;
    if (vj_samples_count(path.c_str(), px, &n) == VJ_OK) {
        out.resize(static_cast<size_t>(n) * static_cast<size_t>(px));
        if (vj_samples_load(path.c_str(), px, out.data(), out.size(), &count) == VJ_OK) return true;
    }
    if (!quiet) std::cerr << "[ERR] " << vj_last_error() << "\n";
    return false;
}

std::vector<std::string> pgm_paths(const std::string& dir) {
    size_t bytes = 0;
    int n = 0;
    std::vector<std::string> out;
    if (vj_pgm_list(dir.c_str(), nullptr, 0, &bytes, &n) != VJ_OK || n == 0) return out;
    std::vector<char> buf(bytes);
    if (vj_pgm_list(dir.c_str(), buf.data(), buf.size(), &bytes, &n) != VJ_OK) return out;
    for (size_t at = 0; at < bytes; at += out.back().size() + 1) out.emplace_back(buf.data() + at);
    return out;
}

struct Gray {
    int w = 0, h = 0;
    std::vector<uint8_t> px;
};

bool load_gray(const std::string& path, Gray& g) {
    if (vj_pgm_peek(path.c_str(), &g.w, &g.h) != VJ_OK) return false;
    g.px.resize(static_cast<size_t>(g.w) * g.h);
    return vj_pgm_load(path.c_str(), g.px.data(), g.px.size()) == VJ_OK;
}

// count random square crops (side uniform in [win, min(w, h)], position
// uniform), each resized to win x win, from images drawn uniformly from the
// shuffled usable files, at most 6 crops per draw.
bool sample_negatives(const std::vector<std::string>& files, int count, int win, uint32_t seed,
                      std::vector<uint8_t>& out) {
    std::mt19937 rng(seed);
    std::vector<std::string> order = files;
    std::shuffle(order.begin(), order.end(), rng);
    std::vector<std::string> usable;
    for (const std::string& f : order) {
        int w = 0, h = 0;
        if (vj_pgm_peek(f.c_str(), &w, &h) == VJ_OK && std::min(w, h) >= win) usable.push_back(f);
    }
    if (usable.empty()) return false;
    const size_t px = static_cast<size_t>(win) * win;
    out.assign(static_cast<size_t>(count) * px, 0);
    std::uniform_int_distribution<size_t> pickImage(0, usable.size() - 1);
    Gray g;
    std::string loaded;
    int n = 0;
    while (n < count) {
        const std::string& f = usable[pickImage(rng)];
        if (f != loaded) {
            if (!load_gray(f, g)) return false;
            loaded = f;
        }
        const int k = std::min(count - n, 6);
        for (int i = 0; i < k; ++i) {
            const int side = std::uniform_int_distribution<int>(win, std::min(g.w, g.h))(rng);
            const int x0 = std::uniform_int_distribution<int>(0, g.w - side)(rng);
            const int y0 = std::uniform_int_distribution<int>(0, g.h - side)(rng);
            const vj_detection box{x0, y0, side, side, 0, 0.0f};
            if (vj_crop_resize(g.px.data(), g.w, g.h, g.w, &box, win, win, out.data() + static_cast<size_t>(n) * px) !=
                VJ_OK)
                return false;
            ++n;
        }
    }
    return true;
}

// count negative patches for a stage: the cached stage file when it holds at
// least count patches, else a fresh sample (written to the cache file).
bool stage_negatives(const std::string& cacheDir, const std::vector<std::string>& images, int stage, int count,
                     int win, int seed, std::vector<uint8_t>& out) {
    char name[64];
    std::snprintf(name, sizeof name, "/stage_negatives/non_faces_stage%02d.bin", stage);
    const std::string path = cacheDir + name;
    const int px = win * win;
    int have = 0;
    if (load_bin(path, px, out, have, true) && have >= count) {
        out.resize(static_cast<size_t>(count) * px);
        return true;
    }
    if (!sample_negatives(images, count, win, static_cast<uint32_t>(seed), out)) return false;
    std::error_code ec;
    std::filesystem::create_directories(cacheDir + "/stage_negatives", ec);
    std::ofstream ofs(path, std::ios::binary);
    if (ofs) ofs.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size()));
    return true;
}

}  // namespace

// VJ_DEBUG_EVERY diagnostics of the train verb (src/main.cu:1708-1720,
// 1967-1983, 2105-2162, 2227-2283): host copies of the trainer's state, and
// the chosen feature's responses recomputed on the host from the stage's
// patches (src/main.cu:476-496: w0*S0 + w1*S1 (+ w2*S2), then * invNorm, each
// S the float of an integer rectangle sum of the patch integral).
struct TrainDebug {
    int every = 0;
    int n = 0, numPos = 0, win = 24;
    std::vector<uint32_t> ii;  // [n][(win+1)^2] patch integrals
    std::vector<double> w;
    std::vector<float> strong, inv;
    std::vector<uint8_t> active;

    void integrals(const uint8_t* patches) {
        const int iw = win + 1, area = iw * iw;
        ii.assign(static_cast<size_t>(n) * area, 0u);
        for (int s = 0; s < n; ++s) {
            const uint8_t* p = patches + static_cast<size_t>(s) * win * win;
            uint32_t* t = ii.data() + static_cast<size_t>(s) * area;
            for (int y = 0; y < win; ++y) {
                uint32_t row = 0;
                for (int x = 0; x < win; ++x) {
                    row += p[y * win + x];
                    t[(y + 1) * iw + x + 1] = t[y * iw + x + 1] + row;
                }
            }
        }
    }
    float response(const vj_feature& f, int s) const {
        const uint32_t* t = ii.data() + static_cast<size_t>(s) * (win + 1) * (win + 1);
        auto sum = [t](const vj_rect& r) {
            return static_cast<float>(static_cast<int32_t>(t[r.p0] - t[r.p1] - t[r.p2] + t[r.p3]));
        };
        float resp = f.r[0].w * sum(f.r[0]) + f.r[1].w * sum(f.r[1]);
        if (f.rectCount == 3) resp += f.r[2].w * sum(f.r[2]);
        return resp * inv[static_cast<size_t>(s)];
    }
    bool read(vj_trainer* t) {
        w.resize(static_cast<size_t>(n));
        strong.resize(static_cast<size_t>(n));
        inv.resize(static_cast<size_t>(n));
        active.resize(static_cast<size_t>(n));
        return vj_trainer_read_state(t, w.data(), strong.data(), active.data(), inv.data()) == VJ_OK;
    }
    void weight_sums(float& pos, float& neg) const {
        pos = 0.0f;
        neg = 0.0f;
        for (int i = 0; i < numPos; ++i) pos += static_cast<float>(w[static_cast<size_t>(i)]);
        for (int i = numPos; i < n; ++i) neg += static_cast<float>(w[static_cast<size_t>(i)]);
    }
};

int run_train(int argc, char** argv) {
    constexpr int kStages = 38, kTotal = 6061, kPos = 10000, kNeg = 10000, kWin = 24, kPerImageKeep = 256;
    const int modelArg = argc - 1;
    if (modelArg < 1) {
        std::cerr << "Usage: cuda_hello train [faces_u8.bin] [numNeg] [maxWeakPerStage] [maxDetections] [minNeighbors]"
                  << " [scaleFactor] [maxStages] [minHitRate] [maxFalseAlarm] [minObjectSize] [stageNegCacheDir]"
                  << " [stageNegImagesDir] [stageNegSeed] [hardNegCandidateMultiplier] <out_model.bin>\n";
        return 1;
    }
    const std::string outPath = argv[modelArg];
    const std::string posPath = modelArg > 1 ? argv[1] : "faces_u8.bin";
    int numNeg = kNeg, maxWeak = 250, maxStages = kStages, seed = 677, maxDetections = 200000, minNeighbors = 4;
    int minObjectSize = 24, candMul = 3;
    float minHit = 0.99f, maxFa = 0.40f, scaleFactor = 1.25f;
    std::string cacheDir, imagesDir = "data_cache/coco_val2017_nonperson_pgm";
    try {
        if (modelArg > 2) numNeg = std::max(1, std::stoi(argv[2]));
        if (modelArg > 3) maxWeak = std::max(1, std::stoi(argv[3]));
        if (modelArg > 4) maxDetections = std::max(100, std::stoi(argv[4]));
        if (modelArg > 5) minNeighbors = std::max(1, std::stoi(argv[5]));
        if (modelArg > 6) {
            scaleFactor = std::stof(argv[6]);
            if (scaleFactor <= 1.0f) {
                std::cerr << "[ERR] scaleFactor must be > 1.0\n";
                return 1;
            }
        }
        if (modelArg > 7) maxStages = std::max(1, std::stoi(argv[7]));
        if (modelArg > 8) {
            minHit = std::stof(argv[8]);
            if (minHit <= 0.0f || minHit > 1.0f) {
                std::cerr << "[ERR] minHitRate must be in (0, 1]\n";
                return 1;
            }
        }
        if (modelArg > 9) {
            maxFa = std::stof(argv[9]);
            if (maxFa < 0.0f || maxFa > 1.0f) {
                std::cerr << "[ERR] maxFalseAlarm must be in [0, 1]\n";
                return 1;
            }
        }
        if (modelArg > 10) minObjectSize = std::max(1, std::stoi(argv[10]));
        if (modelArg > 11) cacheDir = argv[11];
        if (modelArg > 12) imagesDir = argv[12];
        if (modelArg > 13) seed = std::stoi(argv[13]);
        if (modelArg > 14) candMul = std::max(1, std::stoi(argv[14]));
    } catch (...) {
        std::cerr << "[ERR] invalid numeric argument\n";
        return 1;
    }
    if (numNeg != kNeg) {
        std::cout << "[INFO] numNeg is fixed to paper-style per-stage setting: " << kNeg << "\n";
        numNeg = kNeg;
    }
    if (maxStages > kStages) {
        std::cout << "[INFO] maxStages capped to " << kStages << "\n";
        maxStages = kStages;
    }
    std::cout << "[INFO] Config: maxWeakPerStage=" << maxWeak << " maxStages=" << maxStages
              << " scaleFactor=" << scaleFactor << " minNeighbors=" << minNeighbors
              << " minObjectSize=" << minObjectSize << " maxDetections=" << maxDetections
              << " minHitRate=" << minHit << " maxFalseAlarm=" << maxFa << " numNeg=" << numNeg
              << " hardNegCandMul=" << candMul << "\n";
    if (cacheDir.empty()) {
        std::cerr << "[ERR] stageNegCacheDir is required. Dynamic negatives are mandatory.\n";
        return 1;
    }
    const int px = kWin * kWin;
    std::vector<uint8_t> posRaw;
    int posPool = 0;
    if (!load_bin(posPath, px, posRaw, posPool)) return 1;
    const int numPos = std::min(posPool, kPos);
    const int n = numPos + numNeg;
    int poolN = 0;
    vj_haar_pool(kWin, kWin, 1, nullptr, 0, &poolN);
    std::vector<vj_feature> pool(static_cast<size_t>(poolN));
    vj_haar_pool(kWin, kWin, 1, pool.data(), poolN, &poolN);
    std::cout << "[INFO] Feature count: " << poolN << "\n";
    const std::vector<std::string> images = pgm_paths(imagesDir);

    vj_trainer* tr = nullptr;
    if (vj_trainer_create(kWin, kWin, n, 1024, &tr) != VJ_OK) return fail("AdaBoostTrainer");
    struct TGuard {
        vj_trainer* t;
        ~TGuard() { vj_trainer_destroy(t); }
    } tguard{tr};
    // the mining detector: one image of up to 2048 x 2048 at a time
    // (src/main.cu:1614-1626)
    const int mineMaxW = std::max(2048, kWin), mineMaxH = std::max(2048, kWin);
    vj_engine* miner = nullptr;
    if (vj_engine_create(kWin, kWin, 1, mineMaxW, mineMaxH, 0, maxDetections, &miner) != VJ_OK)
        return fail("FaceVisionEngine");
    struct EGuard {
        vj_engine* e;
        ~EGuard() { vj_engine_destroy(e); }
    } eguard{miner};
    DevMem dPool, dTrain, dStumps, dMineImage;
    if (vj_malloc(&dPool.p, pool.size() * sizeof(vj_feature)) != VJ_OK ||
        vj_memcpy_htod(dPool.p, pool.data(), pool.size() * sizeof(vj_feature), nullptr) != VJ_OK ||
        vj_malloc(&dTrain.p, static_cast<size_t>(n) * px) != VJ_OK ||
        vj_malloc(&dStumps.p, static_cast<size_t>(kTotal) * sizeof(vj_stump)) != VJ_OK ||
        vj_malloc(&dMineImage.p, static_cast<size_t>(mineMaxW) * mineMaxH) != VJ_OK)
        return fail("device buffers");
    std::vector<int8_t> labels(static_cast<size_t>(n), -1);
    for (int i = 0; i < numPos; ++i) labels[static_cast<size_t>(i)] = 1;
    std::vector<uint8_t> trainRaw(static_cast<size_t>(n) * px);
    std::vector<vj_stump> stumps;
    std::vector<vj_stage> stages;
    const int mineCap = std::max(20000, numNeg * 2);
    std::vector<vj_detection> mined(static_cast<size_t>(mineCap));
    TrainDebug dbg;
    if (const char* env = std::getenv("VJ_DEBUG_EVERY")) {
        try {
            dbg.every = std::max(0, std::stoi(env));
        } catch (...) {
            dbg.every = 0;
        }
    }
    if (dbg.every > 0) {
        dbg.n = n;
        dbg.numPos = numPos;
        dbg.win = kWin;
        std::cout << "[INFO] Debug logging enabled: VJ_DEBUG_EVERY=" << dbg.every << "\n";
    }
    const auto t0 = std::chrono::high_resolution_clock::now();
    std::cout << "[TRAIN] Start cascade training: maxStages=" << maxStages << " maxWeakPerStage=" << maxWeak
              << " totalFeatureBudget=" << kTotal << "\n";
    for (int stage = 0; stage < maxStages; ++stage) {
        if (static_cast<int>(stumps.size()) >= kTotal) {
            std::cout << "[TRAIN] Reached total feature budget " << kTotal << ", stop at stage " << stage << "\n";
            break;
        }
        const int stageSeed = seed + stage * 1000003;
        // positives: all, or a seeded subset of the pool (src/main.cu:1582-1598)
        if (posPool <= numPos) {
            std::memcpy(trainRaw.data(), posRaw.data(), static_cast<size_t>(numPos) * px);
        } else {
            std::vector<int> idx(static_cast<size_t>(posPool));
            std::iota(idx.begin(), idx.end(), 0);
            std::mt19937 rng(static_cast<uint32_t>(stageSeed ^ 0x7f4a7c15));
            std::shuffle(idx.begin(), idx.end(), rng);
            for (int i = 0; i < numPos; ++i)
                std::memcpy(trainRaw.data() + static_cast<size_t>(i) * px,
                            posRaw.data() + static_cast<size_t>(idx[static_cast<size_t>(i)]) * px, px);
        }
        uint8_t* neg = trainRaw.data() + static_cast<size_t>(numPos) * px;
        int hard = 0;
        std::vector<uint8_t> sample;
        if (stage == 0) {
            if (!stage_negatives(cacheDir, images, stage, numNeg, kWin, stageSeed, sample)) {
                std::cerr << "[ERR] stage negative sampling failed at stage " << stage << "\n";
                return 1;
            }
            std::memcpy(neg, sample.data(), static_cast<size_t>(numNeg) * px);
            hard = numNeg;
        } else {
            // hard negatives: what the cascade so far still accepts
            if (vj_memcpy_htod(dStumps.p, stumps.data(), stumps.size() * sizeof(vj_stump), nullptr) != VJ_OK ||
                vj_engine_set_cascade(miner, static_cast<const vj_feature*>(dPool.p), poolN,
                                      static_cast<const vj_stump*>(dStumps.p), static_cast<int>(stumps.size()),
                                      stages.data(), static_cast<int>(stages.size()), nullptr) != VJ_OK)
                return fail("set mining cascade model");
            if (images.empty()) {
                std::cerr << "[ERR] No PGM mining images found in " << imagesDir << "\n";
                return 1;
            }
            std::vector<std::string> order = images;
            std::mt19937 mineRng(static_cast<uint32_t>(stageSeed));
            std::shuffle(order.begin(), order.end(), mineRng);
            Gray g;
            int scanned = 0;
            for (const std::string& path : order) {
                if (hard >= numNeg) break;
                if (!load_gray(path, g)) continue;
                if (g.w < kWin || g.h < kWin || g.w > mineMaxW || g.h > mineMaxH) continue;
                if (vj_memcpy_htod(dMineImage.p, g.px.data(), g.px.size(), nullptr) != VJ_OK) continue;
                int found = 0;
                if (vj_engine_detect(miner, static_cast<const uint8_t*>(dMineImage.p), g.w, g.h, g.w, scaleFactor, 0,
                                     kWin, mineCap, 0, mined.data(), mineCap, &found, nullptr) != VJ_OK)
                    continue;
                found = std::min(found, mineCap);
                // the best-scoring boxes of each image only, so that a few
                // images cannot fill the stage (src/main.cu:1866-1878)
                const int keep = std::min(found, kPerImageKeep);
                std::partial_sort(mined.begin(), mined.begin() + keep, mined.begin() + found,
                                  [](const vj_detection& a, const vj_detection& b) { return a.score > b.score; });
                for (int i = 0; i < keep && hard < numNeg; ++i, ++hard)
                    vj_crop_resize(g.px.data(), g.w, g.h, g.w, &mined[static_cast<size_t>(i)], kWin, kWin,
                                   neg + static_cast<size_t>(hard) * px);
                ++scanned;
                if (scanned % 20 == 0)
                    std::cout << "[MINE] stage=" << stage << " scanned=" << scanned << " hard=" << hard << "/"
                              << numNeg << "\n";
            }
            std::cout << "[TRAIN] stage=" << stage << " hardNegSelected=" << hard << "/" << numNeg << "\n";
        }
        if (hard < numNeg) {
            // too few hard negatives: random picks from a fresh sample
            // (src/main.cu:1900-1925)
            const int fallback = std::max(numNeg, (numNeg - hard) * candMul);
            if (!stage_negatives(cacheDir, images, stage, fallback, kWin, stageSeed ^ 0x9e3779b9, sample)) {
                std::cerr << "[ERR] fallback stage negative sampling failed at stage " << stage << "\n";
                return 1;
            }
            std::mt19937 rng(static_cast<uint32_t>(stageSeed) ^ 0x5bd1e995u);
            std::uniform_int_distribution<int> pick(0, fallback - 1);
            for (; hard < numNeg; ++hard)
                std::memcpy(neg + static_cast<size_t>(hard) * px, sample.data() + static_cast<size_t>(pick(rng)) * px,
                            px);
        }
        if (vj_memcpy_htod(dTrain.p, trainRaw.data(), trainRaw.size(), nullptr) != VJ_OK) return fail("copy stage set");
        if (vj_trainer_stage_begin(tr, static_cast<const uint8_t*>(dTrain.p), labels.data(), nullptr, n, numPos,
                                   static_cast<const vj_feature*>(dPool.p), poolN, nullptr) != VJ_OK)
            return fail("stage begin");
        if (dbg.every > 0) {
            if (!dbg.read(tr)) return fail("read trainer state");
            dbg.integrals(trainRaw.data());
            float pw, nw;
            dbg.weight_sums(pw, nw);
            std::cout << "[DBG] stage=" << stage << " initWeightSum pos=" << pw << " neg=" << nw
                      << " activeNeg=" << numNeg << "\n";
        }
        const int first = static_cast<int>(stumps.size());
        const int limit = std::min(maxWeak, kTotal - first);
        vj_round_config cfg{minHit, maxFa};
        float stageThreshold = 0.0f;
        bool satisfied = false;
        for (int r = 0; r < limit; ++r) {
            const bool emitDebug = dbg.every > 0 && (r % dbg.every == 0 || r == limit - 1);
            // the weights before the round: the update multiplies the
            // misclassified ones by exp(C) before the normalisation
            if (emitDebug && !dbg.read(tr)) return fail("read trainer state");
            vj_weak wk;
            if (vj_trainer_round(tr, &cfg, &wk, nullptr) != VJ_OK) return fail("round");
            if (wk.feature < 0) {
                std::cerr << "[WARN] No valid weak found at stage " << stage << " weakRound " << r << "\n";
                break;
            }
            stumps.push_back({wk.feature, wk.theta, wk.left, wk.right});
            stageThreshold = wk.stageThreshold;
            if (emitDebug) {
                const vj_feature& f = pool[static_cast<size_t>(wk.feature)];
                const double grow = std::exp(wk.alpha);
                double errP = 0.0, errN = 0.0;
                for (int i = 0; i < n; ++i) {
                    if (!dbg.active[static_cast<size_t>(i)]) continue;
                    const float resp = dbg.response(f, i);
                    const int y = labels[static_cast<size_t>(i)];
                    const bool lt = resp < wk.theta;
                    const int predP = lt ? -1 : 1, predN = lt ? 1 : -1;
                    const int predUsed = wk.parity > 0 ? predP : predN;
                    const double wi = dbg.w[static_cast<size_t>(i)] * (predUsed != y ? grow : 1.0);
                    const float wf = static_cast<float>(wi);
                    if (predP != y) errP += wf;
                    if (predN != y) errN += wf;
                }
                std::cout << "[DBG] stage=" << stage << " weak=" << r << " bestParityRaw=" << wk.parity
                          << " parityUsed=" << wk.parity << " errP=" << errP << " errN=" << errN
                          << " theta=" << wk.theta << "\n";
                if (!dbg.read(tr)) return fail("read trainer state");
                float posMin = std::numeric_limits<float>::infinity(), posMax = -posMin;
                double posSum = 0.0;
                std::vector<float> posScores(dbg.strong.begin(), dbg.strong.begin() + numPos);
                for (float sc : posScores) {
                    posMin = std::min(posMin, sc);
                    posMax = std::max(posMax, sc);
                    posSum += sc;
                }
                float posMedian = 0.0f;
                if (!posScores.empty()) {
                    auto mid = posScores.begin() + static_cast<std::ptrdiff_t>(posScores.size() / 2);
                    std::nth_element(posScores.begin(), mid, posScores.end());
                    posMedian = *mid;
                }
                float negMin = std::numeric_limits<float>::infinity(), negMax = -negMin;
                double negSum = 0.0;
                for (int i = numPos; i < n; ++i) {
                    const float sc = dbg.strong[static_cast<size_t>(i)];
                    negMin = std::min(negMin, sc);
                    negMax = std::max(negMax, sc);
                    negSum += sc;
                }
                const float negMean = n == numPos ? 0.0f : static_cast<float>(negSum / static_cast<double>(n - numPos));
                float pw, nw;
                dbg.weight_sums(pw, nw);
                std::cout << "[DBG] stage=" << stage << " weak=" << r << " score_pos[min,max,mean,p50]=[" << posMin
                          << "," << posMax << "," << static_cast<float>(posSum / static_cast<double>(numPos)) << ","
                          << posMedian << "]" << " score_neg[min,max,mean]=[" << negMin << "," << negMax << ","
                          << negMean << "]" << " threshold=" << wk.stageThreshold << " weightSum[pos,neg]=[" << pw
                          << "," << nw << "]\n";
            }
            std::cout << "[TRAIN] stage=" << stage << " weak=" << r << " feature=" << wk.feature << " err="
                      << static_cast<float>(std::min(std::max(wk.err, 1e-6), 1.0 - 1e-6)) << " alpha(C)=" << wk.alpha
                      << " parity=" << wk.parity << " hit=" << wk.hitRate << " fa=" << wk.falseAlarm
                      << " negSurvive=" << wk.negSurvive << "/" << numNeg << "\n";
            if (wk.stageSatisfied) {
                satisfied = true;
                break;
            }
        }
        const int trees = static_cast<int>(stumps.size()) - first;
        if (trees <= 0) break;
        int remain = 0;
        if (vj_trainer_stage_end(tr, stageThreshold, &remain, nullptr) != VJ_OK) return fail("stage end");
        if (!satisfied && remain > 0) {
            std::cerr << "[WARN] Stage " << stage << " did not meet target. Increase maxWeakPerStage.\n";
            break;
        }
        stages.push_back({first, trees, stageThreshold - 1e-5f});
        std::cout << "[TRAIN] Stage " << stage << " done: ntrees=" << trees << " threshold=" << stageThreshold
                  << " remainNeg=" << remain << "/" << numNeg << "\n";
        if (remain == 0) break;
    }
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::cout << "[TRAIN] elapsed(ms): " << std::chrono::duration_cast<std::chrono::milliseconds>(t1 - t0).count()
              << "\n";
    if (stumps.empty() || stages.empty()) {
        std::cerr << "[ERR] no trained cascade stages\n";
        return 1;
    }
    if (vj_model_save(outPath.c_str(), 24, 24, pool.data(), poolN, stumps.data(), static_cast<int64_t>(stumps.size()),
                      stages.data(), static_cast<int64_t>(stages.size())) != VJ_OK) {
        std::cerr << "[ERR] saveCascadeModel failed: " << outPath << "\n";
        return 1;
    }
    std::cout << "[OK] Saved model to: " << outPath << " (features=" << poolN << ", stumps=" << stumps.size()
              << ", stages=" << stages.size() << ")\n";
    return 0;
}

// ---------------------------------------------------------------------------
// train_compare / detect_compare (src/main.cu:1032-1368): the same work on
// the host (vj_cpu_*) and on the GPU, averaged over `runs`, with both
// results and the speed-up printed under the reference's tags.  Besides the
// reference's lines each verb prints one "match=" line saying whether the
// two legs agree (the same stump; the same raw detections in order).
// ---------------------------------------------------------------------------
namespace {

struct Timer {
    vj_event a = nullptr, b = nullptr;
    Timer() {
        vj_event_create(&a);
        vj_event_create(&b);
    }
    ~Timer() {
        if (a) vj_event_destroy(a);
        if (b) vj_event_destroy(b);
    }
};

double ms_since(std::chrono::high_resolution_clock::time_point t0) {
    const auto t1 = std::chrono::high_resolution_clock::now();
    return static_cast<double>(std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count()) / 1000.0;
}

}  // namespace

int run_train_compare(int argc, char** argv) {
    if (argc < 3) {
        std::cerr << "Usage: cuda_hello train_compare <faces_u8.bin> <negatives_u8.bin>"
                  << " [numPos=256] [numNeg=256] [featureLimit=1024] [runs=3]\n";
        return 1;
    }
    constexpr int kWin = 24, kPx = kWin * kWin;
    int reqPos = 256, reqNeg = 256, featureLimit = 1024, runs = 3;
    try {
        if (argc > 3) reqPos = std::max(1, std::stoi(argv[3]));
        if (argc > 4) reqNeg = std::max(1, std::stoi(argv[4]));
        if (argc > 5) featureLimit = std::max(1, std::stoi(argv[5]));
        if (argc > 6) runs = std::max(1, std::stoi(argv[6]));
    } catch (...) {
        std::cerr << "[ERR] invalid numeric argument\n";
        return 1;
    }
    std::vector<uint8_t> posRaw, negRaw;
    int posCount = 0, negCount = 0;
    if (!load_bin(argv[1], kPx, posRaw, posCount) || !load_bin(argv[2], kPx, negRaw, negCount)) return 1;
    const int numPos = std::min(reqPos, posCount), numNeg = std::min(reqNeg, negCount);
    const int n = numPos + numNeg;
    if (n <= 1) {
        std::cerr << "[ERR] Need at least 2 samples for train_compare\n";
        return 1;
    }
    // positives first, then negatives; weights 0.5/numPos and 0.5/numNeg
    std::vector<uint8_t> train(static_cast<size_t>(n) * kPx);
    std::memcpy(train.data(), posRaw.data(), static_cast<size_t>(numPos) * kPx);
    std::memcpy(train.data() + static_cast<size_t>(numPos) * kPx, negRaw.data(), static_cast<size_t>(numNeg) * kPx);
    std::vector<int8_t> labels(static_cast<size_t>(n), -1);
    std::vector<uint8_t> active(static_cast<size_t>(n), 1);
    std::vector<float> weights(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i) {
        if (i < numPos) labels[static_cast<size_t>(i)] = +1;
        weights[static_cast<size_t>(i)] = i < numPos ? 0.5f / static_cast<float>(numPos) : 0.5f / static_cast<float>(numNeg);
    }
    int poolN = 0;
    if (vj_haar_pool(kWin, kWin, 1, nullptr, 0, &poolN) != VJ_OK) return fail("buildHaarFeaturePool");
    std::vector<vj_feature> pool(static_cast<size_t>(poolN));
    if (vj_haar_pool(kWin, kWin, 1, pool.data(), poolN, &poolN) != VJ_OK) return fail("buildHaarFeaturePool");
    featureLimit = std::min(featureLimit, poolN);

    vj_tile_best cpu{};
    double cpuMs = 0.0;
    for (int r = 0; r < runs; ++r) {
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (vj_cpu_best_weak(train.data(), n, kWin, kWin, labels.data(), active.data(), weights.data(), pool.data(),
                             featureLimit, &cpu) != VJ_OK)
            return fail("CPU weak search");
        cpuMs += ms_since(t0);
    }
    cpuMs /= runs;

    vj_engine* eng = nullptr;
    if (vj_engine_create(kWin, kWin, n, kWin, kWin, 0, 1, &eng) != VJ_OK) return fail("FaceVisionEngine");
    struct EngineGuard {
        vj_engine* e;
        ~EngineGuard() { vj_engine_destroy(e); }
    } eg{eng};
    vj_trainer* tr = nullptr;
    if (vj_trainer_create(kWin, kWin, n, featureLimit, &tr) != VJ_OK) return fail("AdaBoostTrainer");
    struct TrainerGuard {
        vj_trainer* t;
        ~TrainerGuard() { vj_trainer_destroy(t); }
    } tg{tr};
    if (vj_engine_upload_feature_pool(eng, pool.data(), poolN, nullptr) != VJ_OK) return fail("uploadFeaturePool");
    const size_t fn = static_cast<size_t>(featureLimit) * n;
    DevMem dTrain, dLabel, dActive, dWeight, dResp, dSorted, dBest, dTile;
    if (vj_malloc(&dTrain.p, train.size()) != VJ_OK || vj_malloc(&dLabel.p, labels.size()) != VJ_OK ||
        vj_malloc(&dActive.p, active.size()) != VJ_OK || vj_malloc(&dWeight.p, weights.size() * 4) != VJ_OK ||
        vj_malloc(&dResp.p, fn * 4) != VJ_OK || vj_malloc(&dSorted.p, fn * 2) != VJ_OK ||
        vj_malloc(&dBest.p, static_cast<size_t>(featureLimit) * 16) != VJ_OK || vj_malloc(&dTile.p, 16) != VJ_OK)
        return fail("device allocation");
    if (vj_memcpy_htod(dTrain.p, train.data(), train.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dLabel.p, labels.data(), labels.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dActive.p, active.data(), active.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dWeight.p, weights.data(), weights.size() * 4, nullptr) != VJ_OK)
        return fail("cudaMemcpy failed");
    if (vj_engine_integral_batch(eng, static_cast<const uint8_t*>(dTrain.p), kWin, kWin, kWin, n, 1, nullptr) != VJ_OK)
        return fail("computeIntegralBatchTransposed failed");
    vj_integral_view view{};
    if (vj_engine_integral_view(eng, &view) != VJ_OK ||
        vj_trainer_bind(tr, &view, static_cast<const int8_t*>(dLabel.p), static_cast<const uint8_t*>(dActive.p),
                        static_cast<const float*>(dWeight.p), n) != VJ_OK)
        return fail("bindTrainingSet failed");
    const vj_feature* dPool = nullptr;
    int dPoolN = 0;
    vj_engine_feature_pool(eng, &dPool, &dPoolN);
    Timer tm;
    float gpuMs = 0.0f;
    for (int r = 0; r < runs; ++r) {
        vj_event_record(tm.a, nullptr);
        int st = vj_trainer_evaluate_responses(tr, dPool, 0, featureLimit, static_cast<float*>(dResp.p), nullptr);
        if (st == VJ_OK)
            st = vj_trainer_sort_per_feature(tr, static_cast<const float*>(dResp.p), featureLimit,
                                             static_cast<uint16_t*>(dSorted.p), nullptr);
        if (st == VJ_OK)
            st = vj_trainer_find_threshold(tr, static_cast<const float*>(dResp.p),
                                           static_cast<const uint16_t*>(dSorted.p), featureLimit, dBest.p, nullptr);
        if (st == VJ_OK) st = vj_trainer_select_best(tr, dBest.p, 0, featureLimit, nullptr, dTile.p, nullptr);
        vj_event_record(tm.b, nullptr);
        vj_event_synchronize(tm.b);
        if (st != VJ_OK) return fail("GPU benchmark failed");
        float ms = 0.0f;
        vj_event_elapsed_ms(tm.a, tm.b, &ms);
        gpuMs += ms;
    }
    gpuMs /= static_cast<float>(runs);
    vj_tile_best gpu{};
    if (vj_memcpy_dtoh(&gpu, dTile.p, sizeof gpu, nullptr) != VJ_OK) return fail("cudaMemcpy failed");

    std::cout << "[TRAIN_COMPARE] samples=" << n << " numPos=" << numPos << " numNeg=" << numNeg
              << " features=" << featureLimit << " runs=" << runs << "\n";
    std::cout << "[TRAIN_COMPARE] CPU best feature=" << cpu.feature << " theta=" << cpu.theta
              << " parity=" << static_cast<int>(cpu.parity) << " err=" << cpu.err << " avg(ms)=" << cpuMs << "\n";
    std::cout << "[TRAIN_COMPARE] GPU best feature=" << gpu.feature << " theta=" << gpu.theta
              << " parity=" << static_cast<int>(gpu.parity) << " err=" << gpu.err << " avg(ms)=" << gpuMs << "\n";
    if (gpuMs > 0.0f) std::cout << "[TRAIN_COMPARE] speedup=" << (cpuMs / static_cast<double>(gpuMs)) << "x\n";
    const bool same = cpu.feature == gpu.feature && cpu.parity == gpu.parity &&
                      std::memcmp(&cpu.theta, &gpu.theta, 4) == 0 && std::memcmp(&cpu.err, &gpu.err, 4) == 0;
    std::cout << "[TRAIN_COMPARE] match=" << (same ? "yes" : "no") << "\n";
    return 0;
}

int run_detect_compare(int argc, char** argv) {
    if (argc < 4) {
        std::cerr << "Usage: cuda_hello detect_compare <in_model.bin> <image.pgm> <out.ppm>"
                  << " [scaleFactor=1.25] [minObjectSize=24] [maxDetections=200000] [maxScales=0] [runs=3]\n";
        return 1;
    }
    const std::string modelPath = argv[1], imagePath = argv[2], outPath = argv[3];
    float scaleFactor = 1.25f;
    int minObjectSize = 24, maxDetections = 200000, maxScales = 0, runs = 3;
    try {
        if (argc > 4) scaleFactor = std::stof(argv[4]);
        if (argc > 5) minObjectSize = std::max(1, std::stoi(argv[5]));
        if (argc > 6) maxDetections = std::max(100, std::stoi(argv[6]));
        if (argc > 7) maxScales = std::max(0, std::stoi(argv[7]));
        if (argc > 8) runs = std::max(1, std::stoi(argv[8]));
    } catch (...) {
        std::cerr << "[ERR] invalid numeric argument\n";
        return 1;
    }
    if (maxScales != 0) {
        std::cerr << "[ERR] detect_compare currently requires maxScales=0 so CPU/GPU use the same scale schedule\n";
        return 1;
    }
    Model m;
    if (!load_model(modelPath, m)) {
        std::cerr << "[ERR] loadCascadeModel failed: " << modelPath << "\n";
        return 1;
    }
    Gray img;
    if (!load_gray(imagePath, img)) return fail("loadPGM");
    const int w = img.w, h = img.h;

    std::vector<vj_detection> cpuDets(static_cast<size_t>(maxDetections));
    int cpuCount = 0;
    double cpuMs = 0.0;
    for (int r = 0; r < runs; ++r) {
        const auto t0 = std::chrono::high_resolution_clock::now();
        if (vj_cpu_detect_raw(m.features.data(), static_cast<int>(m.features.size()), m.stumps.data(),
                              static_cast<int>(m.stumps.size()), m.stages.data(), static_cast<int>(m.stages.size()),
                              m.winW, m.winH, img.px.data(), w, h, w, scaleFactor, minObjectSize, maxDetections,
                              maxScales, cpuDets.data(), static_cast<int>(cpuDets.size()), &cpuCount) != VJ_OK)
            return fail("detectMultiScaleCpu");
        cpuMs += ms_since(t0);
    }
    cpuMs /= runs;

    vj_engine* eng = nullptr;
    if (vj_engine_create(m.winW, m.winH, 1, std::max(m.winW, w), std::max(m.winH, h), 0, maxDetections, &eng) != VJ_OK)
        return fail("FaceVisionEngine");
    struct Guard {
        vj_engine* e;
        ~Guard() { vj_engine_destroy(e); }
    } guard{eng};
    if (vj_engine_upload_feature_pool(eng, m.features.data(), static_cast<int>(m.features.size()), nullptr) != VJ_OK)
        return fail("uploadFeaturePool");
    DevMem dStumps, dImage;
    if (vj_malloc(&dStumps.p, m.stumps.size() * sizeof(vj_stump)) != VJ_OK ||
        vj_malloc(&dImage.p, img.px.size()) != VJ_OK ||
        vj_memcpy_htod(dStumps.p, m.stumps.data(), m.stumps.size() * sizeof(vj_stump), nullptr) != VJ_OK ||
        vj_memcpy_htod(dImage.p, img.px.data(), img.px.size(), nullptr) != VJ_OK)
        return fail("cudaMemcpy failed");
    const vj_feature* dPool = nullptr;
    int poolN = 0;
    vj_engine_feature_pool(eng, &dPool, &poolN);
    if (vj_engine_set_cascade(eng, dPool, poolN, static_cast<const vj_stump*>(dStumps.p),
                              static_cast<int>(m.stumps.size()), m.stages.data(), static_cast<int>(m.stages.size()),
                              nullptr) != VJ_OK)
        return fail("setCascadeModel");
    std::vector<vj_detection> gpuDets(static_cast<size_t>(maxDetections));
    int gpuCount = 0;
    Timer tm;
    float gpuMs = 0.0f;
    for (int r = 0; r < runs; ++r) {
        vj_event_record(tm.a, nullptr);
        const int st = vj_engine_detect(eng, static_cast<const uint8_t*>(dImage.p), w, h, w, scaleFactor, 0,
                                        minObjectSize, maxDetections, 0, gpuDets.data(),
                                        static_cast<int>(gpuDets.size()), &gpuCount, nullptr);
        vj_event_record(tm.b, nullptr);
        vj_event_synchronize(tm.b);
        if (st != VJ_OK) return fail("detectMultiScale failed");
        float ms = 0.0f;
        vj_event_elapsed_ms(tm.a, tm.b, &ms);
        gpuMs += ms;
    }
    gpuMs /= static_cast<float>(runs);
    const int shown = std::min(gpuCount, static_cast<int>(gpuDets.size()));

    std::cout << "[DETECT_COMPARE] cpuDetections=" << cpuCount << " gpuDetections=" << gpuCount << " runs=" << runs
              << " maxScales=" << maxScales << "\n";
    std::cout << "[DETECT_COMPARE] CPU avg(ms)=" << cpuMs << "\n";
    std::cout << "[DETECT_COMPARE] GPU avg(ms)=" << gpuMs << "\n";
    if (gpuMs > 0.0f) std::cout << "[DETECT_COMPARE] speedup=" << (cpuMs / static_cast<double>(gpuMs)) << "x\n";
    const bool same = cpuCount == gpuCount &&
                      std::memcmp(cpuDets.data(), gpuDets.data(), sizeof(vj_detection) * static_cast<size_t>(shown)) == 0;
    std::cout << "[DETECT_COMPARE] match=" << (same ? "yes" : "no") << "\n";
    if (vj_ppm_save_gray_boxes(outPath.c_str(), img.px.data(), w, h, gpuDets.data(), shown) != VJ_OK) {
        std::cerr << "[ERR] savePPM failed\n";
        return 1;
    }
    std::cout << "[OK] Saved GPU detection visualization to: " << outPath << "\n";
    return 0;
}
