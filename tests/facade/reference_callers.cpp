// Call sites in the shape of the reference's runDetect (src/main.cu:2402-2470)
// and runTrainCompare GPU leg (src/main.cu:1119-1180), written against
// include/vj_facade.hpp.  Test program: the CPU suite builds it and runs the
// `nodevice` verb; the GPU suite runs `detect` and `train` and checks the
// printed results against the oracle.
//
//   reference_callers nodevice
//   reference_callers detect <model.bin> <image.pgm> <minNeighbors> <grouping 0|1>
//   reference_callers train <patches.u8> <numPos> <numNeg> <featureLimit>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "vj_facade.hpp"

using vj::AdaBoostTrainer;
using vj::CascadeStage;
using vj::Detection;
using vj::FaceVisionEngine;
using vj::FeatureTileView;
using vj::GpuStump4;
using vj::HaarFeature;
using vj::LabelsView;
using vj::Status;
using vj::TileBestCandidate;
using vj::WeightsView;
using vj::WindowSpec;

namespace {

template <typename T>
struct DeviceBuffer {  // the role of the reference's CudaBuffer (src/main.cu:52-101)
    T* p = nullptr;
    explicit DeviceBuffer(size_t n) {
        if (vj_malloc(reinterpret_cast<void**>(&p), std::max<size_t>(n, 1) * sizeof(T)) != VJ_OK)
            throw std::runtime_error("device allocation failed");
    }
    ~DeviceBuffer() { vj_free(p); }
    T* data() { return p; }
};

uint32_t bits(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    return u;
}

int run_detect(const char* modelPath, const char* imagePath, int minNeighbors, bool grouping) {
    int winW = 0, winH = 0;
    int64_t nF = 0, nS = 0, nSt = 0;
    if (vj_model_peek(modelPath, &winW, &winH, &nF, &nS, &nSt) != VJ_OK) {
        std::cerr << "[ERR] loadCascadeModel failed: " << modelPath << "\n";
        return 1;
    }
    const WindowSpec win{winW, winH};
    std::vector<HaarFeature> hFeatures(static_cast<size_t>(nF));
    std::vector<GpuStump4> hStumps(static_cast<size_t>(nS));
    std::vector<CascadeStage> hStages(static_cast<size_t>(nSt));
    if (vj_model_load(modelPath, reinterpret_cast<vj_feature*>(hFeatures.data()), nF,
                      reinterpret_cast<vj_stump*>(hStumps.data()), nS, reinterpret_cast<vj_stage*>(hStages.data()),
                      nSt) != VJ_OK)
        return 1;
    int w = 0, h = 0;
    if (vj_pgm_peek(imagePath, &w, &h) != VJ_OK) return 1;
    std::vector<uint8_t> img(static_cast<size_t>(w) * h);
    if (vj_pgm_load(imagePath, img.data(), img.size()) != VJ_OK) return 1;

    FaceVisionEngine detectEngine(win, 1, std::max(win.winW, w), std::max(win.winH, h), false, 200000);
    Status st = detectEngine.uploadFeaturePool(hFeatures);
    if (st != Status::kOk) return 1;
    DeviceBuffer<GpuStump4> dStumps(hStumps.size());
    DeviceBuffer<uint8_t> dImage(img.size());
    if (vj_memcpy_htod(dStumps.data(), hStumps.data(), hStumps.size() * sizeof(GpuStump4), nullptr) != VJ_OK ||
        vj_memcpy_htod(dImage.data(), img.data(), img.size(), nullptr) != VJ_OK)
        return 1;
    st = detectEngine.setCascadeModel(detectEngine.deviceFeaturePool(), detectEngine.featurePoolSize(),
                                      dStumps.data(), static_cast<int>(hStumps.size()), hStages);
    if (st != Status::kOk) return 1;
    std::vector<Detection> detections;
    st = detectEngine.detectMultiScale(dImage.data(), w, h, w, 1.25f, minNeighbors, 24, 200000, detections, grouping);
    if (st != Status::kOk) {
        std::cerr << "[ERR] detectMultiScale failed: " << vj_last_error() << "\n";
        return 1;
    }
    for (const Detection& d : detections)
        std::printf("%d %d %d %d %d %u\n", d.x, d.y, d.w, d.h, d.scaleIdx, bits(d.score));
    return 0;
}

int run_train(const char* patchPath, int numPos, int numNeg, int featureLimit) {
    const WindowSpec win{24, 24};
    const int nSamples = numPos + numNeg;
    std::vector<uint8_t> trainRaw(static_cast<size_t>(nSamples) * 24 * 24);
    std::ifstream f(patchPath, std::ios::binary);
    if (!f.read(reinterpret_cast<char*>(trainRaw.data()), static_cast<std::streamsize>(trainRaw.size()))) return 1;
    std::vector<int8_t> labels(static_cast<size_t>(nSamples), -1);
    std::vector<uint8_t> active(static_cast<size_t>(nSamples), 1);
    std::vector<float> weights(static_cast<size_t>(nSamples), 1.0f / static_cast<float>(nSamples));
    for (int i = 0; i < numPos; ++i) labels[static_cast<size_t>(i)] = 1;
    std::vector<HaarFeature> features = FaceVisionEngine::buildHaarFeaturePool(win, FaceVisionEngine::FeatureMode::Basic);
    features.resize(static_cast<size_t>(featureLimit));

    FaceVisionEngine engine(win, nSamples, win.winW, win.winH, false, 1);
    AdaBoostTrainer trainer(win, nSamples, featureLimit);
    Status st = engine.uploadFeaturePool(features);
    if (st != Status::kOk) return 1;
    DeviceBuffer<uint8_t> dTrain(trainRaw.size());
    DeviceBuffer<int8_t> dLabel(static_cast<size_t>(nSamples));
    DeviceBuffer<uint8_t> dActive(static_cast<size_t>(nSamples));
    DeviceBuffer<float> dWeight(static_cast<size_t>(nSamples));
    DeviceBuffer<TileBestCandidate> dTileBest(1);
    if (vj_memcpy_htod(dTrain.data(), trainRaw.data(), trainRaw.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dLabel.data(), labels.data(), labels.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dActive.data(), active.data(), active.size(), nullptr) != VJ_OK ||
        vj_memcpy_htod(dWeight.data(), weights.data(), weights.size() * sizeof(float), nullptr) != VJ_OK)
        return 1;
    st = engine.computeIntegralBatchTransposed(dTrain.data(), win.winW, win.winH, win.winW, nSamples);
    if (st != Status::kOk) return 1;
    st = trainer.bindTrainingSet(engine.integralSetView(), LabelsView{nSamples, dLabel.data(), dActive.data()},
                                 WeightsView{nSamples, dWeight.data()});
    if (st != Status::kOk) return 1;
    st = trainer.evaluateFeatureResponses(engine.deviceFeaturePool(), 0, featureLimit);
    if (st == Status::kOk) st = trainer.sortSamplesPerFeature(trainer.responseBuffer(), featureLimit);
    if (st == Status::kOk) {
        FeatureTileView tile{};
        tile.featureBegin = 0;
        tile.featureCount = featureLimit;
        tile.d_features = engine.deviceFeaturePool();
        tile.d_resp = trainer.responseBuffer();
        tile.d_sortedIdx = trainer.sortedIndexBuffer();
        st = trainer.evaluateAndFindThreshold(tile);
    }
    if (st == Status::kOk) st = trainer.selectBestInTile(0, featureLimit, nullptr, dTileBest.data());
    if (st != Status::kOk) {
        std::cerr << "[ERR] GPU weak search failed: " << vj_last_error() << "\n";
        return 1;
    }
    TileBestCandidate best{};
    if (vj_memcpy_dtoh(&best, dTileBest.data(), sizeof best, nullptr) != VJ_OK) return 1;
    std::printf("%d %u %d %u\n", best.featureIdx, bits(best.bestTheta), static_cast<int>(best.bestParity),
                bits(best.bestErr));
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    const std::string verb = argc > 1 ? argv[1] : "";
    try {
        if (verb == "nodevice") {
            // without a device the constructors throw, as the reference's do
            try {
                FaceVisionEngine e(WindowSpec{24, 24}, 1, 64, 64);
                std::printf("engine constructed\n");
            } catch (const std::runtime_error& ex) {
                std::printf("FaceVisionEngine threw: %s\n", ex.what());
            }
            try {
                AdaBoostTrainer t(WindowSpec{24, 24}, 100, 10);
                std::printf("trainer constructed\n");
            } catch (const std::runtime_error& ex) {
                std::printf("AdaBoostTrainer threw: %s\n", ex.what());
            }
            std::printf("pool %d\n", FaceVisionEngine::countHaarFeatures(WindowSpec{24, 24},
                                                                         FaceVisionEngine::FeatureMode::Basic));
            return 0;
        }
        if (verb == "detect" && argc == 6) return run_detect(argv[2], argv[3], std::atoi(argv[4]), std::atoi(argv[5]) != 0);
        if (verb == "train" && argc == 6)
            return run_train(argv[2], std::atoi(argv[3]), std::atoi(argv[4]), std::atoi(argv[5]));
    } catch (const std::exception& ex) {
        std::cerr << "[ERR] " << ex.what() << "\n";
        return 1;
    }
    std::cerr << "usage: reference_callers nodevice | detect <model> <pgm> <mn> <grouping> | "
                 "train <patches> <numPos> <numNeg> <features>\n";
    return 2;
}
