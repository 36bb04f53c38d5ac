/*
 * vj_facade.hpp -- the reference's C++ classes over libvj_hip.so.
 *
 * A caller written against vj::FaceVisionEngine and vj::AdaBoostTrainer
 * (include/vj/face_vision_engine.cuh:19-196, include/vj/boosting_core.cuh:20-161
 * in the reference) includes this header instead of those two and of
 * include/vj/common_types.cuh, and links -lvj_hip.  Class names, method names,
 * argument order, defaults and the public record types are the reference's;
 * streams are the library's vj_stream (a hipStream_t converts to it), where the
 * reference takes a cudaStream_t.  Constructors throw std::runtime_error on
 * failure, as the reference's do; methods return vj::Status, whose numbers are
 * the VJ_* codes.
 *
 * The record types restate include/vj/common_types.cuh:11-143 field for field
 * (the static_asserts below tie each to its vj_api.h twin), with two
 * differences a caller may notice:
 *  - IntegralImageSetT::d_sqsumT points at 64-bit square sums (the low word of
 *    each element is the reference's int32 value);
 *  - GpuStump4::st is a 16-byte record {x, y, z, w} in place of CUDA's float4.
 */
#ifndef VJ_FACADE_HPP
#define VJ_FACADE_HPP

#include <algorithm>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "vj_api.h"

namespace vj {

// include/vj/common_types.cuh:11-17
enum class Status : int {
    kOk = VJ_OK,
    kInvalidArg = VJ_INVALID_ARG,
    kCudaError = VJ_DEVICE_ERROR,
    kOutOfMemory = VJ_OUT_OF_MEMORY,
    kRuntimeError = VJ_RUNTIME_ERROR
};

// include/vj/common_types.cuh:19-143
struct WindowSpec {
    int winW;
    int winH;
};
struct FastRect {
    int p0, p1, p2, p3;
    float w;
};
struct HaarFeature {
    FastRect r0;
    FastRect r1;
    FastRect r2;
    uint8_t rectCount;
    uint8_t tilted;
    uint16_t reserved;
};
struct IntegralImageSetT {
    int nSamples;
    int iiWidth;
    int iiHeight;
    int iiArea;
    const int32_t* d_sumT;      // [iiArea][N]
    const int32_t* d_sqsumT;    // [iiArea][N] 64-bit elements (see above)
    const int32_t* d_tiltedT;   // never filled, as in the reference
    const float* d_invNorm;     // [N]
};
struct LabelsView {
    int nSamples;
    const int8_t* d_label;
    const uint8_t* d_active;
};
struct WeightsView {
    int nSamples;
    float* d_weight;
};
struct FeatureTileView {
    int featureBegin;
    int featureCount;
    const HaarFeature* d_features;
    const float* d_resp;
    const uint16_t* d_sortedIdx;
};
struct FeatureBestSplit {
    float bestErr;
    float bestTheta;
    int32_t bestSplitPos;
    int8_t bestParity;
    int8_t pad0, pad1, pad2;
};
struct TileBestCandidate {
    float bestErr;
    float bestTheta;
    int32_t featureIdx;
    int8_t bestParity;
    int8_t pad0, pad1, pad2;
};
struct WeakClassifier {
    int featureIdx;
    float theta;
    int8_t parity;
    int8_t pad0, pad1, pad2;
    float err;
    float alpha;
    float leftVal;
    float rightVal;
};
struct CascadeStage {
    int first;
    int ntrees;
    float threshold;
};
struct Detection {
    int x;
    int y;
    int w;
    int h;
    int scaleIdx;
    float score;
};
struct alignas(16) Stump4 {
    float x, y, z, w;  // {featureIdx as int bits, theta, leftVal, rightVal}
};
struct alignas(16) GpuStump4 {
    Stump4 st;
};

static_assert(sizeof(HaarFeature) == sizeof(vj_feature) && offsetof(HaarFeature, rectCount) ==
                  offsetof(vj_feature, rectCount), "HaarFeature layout");
static_assert(sizeof(GpuStump4) == sizeof(vj_stump), "GpuStump4 layout");
static_assert(sizeof(CascadeStage) == sizeof(vj_stage), "CascadeStage layout");
static_assert(sizeof(Detection) == sizeof(vj_detection), "Detection layout");
static_assert(sizeof(TileBestCandidate) == sizeof(vj_tile_best), "TileBestCandidate layout");
static_assert(sizeof(FeatureBestSplit) == 16, "FeatureBestSplit layout");

inline std::runtime_error facade_error(const char* what) {
    const char* msg = vj_last_error();
    return std::runtime_error(std::string(what) + ": " + (msg ? msg : "unknown error"));
}

// include/vj/face_vision_engine.cuh:19-196, src/face_vision_engine.cu:644-1747
class FaceVisionEngine {
public:
    enum class FeatureMode : int { Basic = 0, Core = 1, All = 2 };

    FaceVisionEngine(const WindowSpec& win, int maxSamples, int maxImageW, int maxImageH,
                     bool enableTilted = true, int maxDetections = 200000)
        : win_(win) {
        if (vj_engine_create(win.winW, win.winH, maxSamples, maxImageW, maxImageH, enableTilted ? 1 : 0,
                             maxDetections, &e_) != VJ_OK)
            throw facade_error("FaceVisionEngine");
    }
    ~FaceVisionEngine() { vj_engine_destroy(e_); }
    FaceVisionEngine(const FaceVisionEngine&) = delete;
    FaceVisionEngine& operator=(const FaceVisionEngine&) = delete;
    FaceVisionEngine(FaceVisionEngine&& o) noexcept : win_(o.win_), e_(std::exchange(o.e_, nullptr)) {}
    FaceVisionEngine& operator=(FaceVisionEngine&& o) noexcept {
        if (this != &o) {
            vj_engine_destroy(e_);
            win_ = o.win_;
            e_ = std::exchange(o.e_, nullptr);
        }
        return *this;
    }

    static int countHaarFeatures(const WindowSpec& win, FeatureMode mode) {
        int n = 0;
        return vj_haar_pool(win.winW, win.winH, static_cast<int>(mode), nullptr, 0, &n) == VJ_OK ? n : 0;
    }
    static std::vector<HaarFeature> buildHaarFeaturePool(const WindowSpec& win, FeatureMode mode) {
        std::vector<HaarFeature> v(static_cast<size_t>(countHaarFeatures(win, mode)));
        int n = 0;
        if (vj_haar_pool(win.winW, win.winH, static_cast<int>(mode), reinterpret_cast<vj_feature*>(v.data()),
                         static_cast<int>(v.size()), &n) != VJ_OK)
            v.clear();
        return v;
    }
    Status uploadFeaturePool(const std::vector<HaarFeature>& hFeatures, vj_stream stream = nullptr) {
        return Status(vj_engine_upload_feature_pool(e_, reinterpret_cast<const vj_feature*>(hFeatures.data()),
                                                    static_cast<int>(hFeatures.size()), stream));
    }
    const HaarFeature* deviceFeaturePool() const noexcept {
        const vj_feature* d = nullptr;
        int n = 0;
        vj_engine_feature_pool(e_, &d, &n);
        return reinterpret_cast<const HaarFeature*>(d);
    }
    int featurePoolSize() const noexcept {
        const vj_feature* d = nullptr;
        int n = 0;
        vj_engine_feature_pool(e_, &d, &n);
        return n;
    }
    Status computeIntegralBatchTransposed(const uint8_t* dGrayBatch, int imageW, int imageH, int imageStride,
                                          int nSamples, bool computeInvNorm = true, vj_stream stream = nullptr) {
        return Status(vj_engine_integral_batch(e_, dGrayBatch, imageW, imageH, imageStride, nSamples,
                                               computeInvNorm ? 1 : 0, stream));
    }
    IntegralImageSetT integralSetView() const noexcept {
        vj_integral_view v{};
        vj_engine_integral_view(e_, &v);
        IntegralImageSetT o{};
        o.nSamples = v.nSamples;
        o.iiWidth = v.iiWidth;
        o.iiHeight = v.iiHeight;
        o.iiArea = v.iiArea;
        o.d_sumT = reinterpret_cast<const int32_t*>(v.sumT);
        o.d_sqsumT = reinterpret_cast<const int32_t*>(v.sqsumT);
        o.d_tiltedT = nullptr;
        o.d_invNorm = v.invNorm;
        return o;
    }
    Status setCascadeModel(const HaarFeature* dUsedFeatures, int usedFeatureCount, const GpuStump4* dStumps,
                           int stumpCount, const std::vector<CascadeStage>& hStages, vj_stream stream = nullptr) {
        return Status(vj_engine_set_cascade(e_, reinterpret_cast<const vj_feature*>(dUsedFeatures), usedFeatureCount,
                                            reinterpret_cast<const vj_stump*>(dStumps), stumpCount,
                                            reinterpret_cast<const vj_stage*>(hStages.data()),
                                            static_cast<int>(hStages.size()), stream));
    }
    Status clearCascadeModel(vj_stream stream = nullptr) { return Status(vj_engine_clear_cascade(e_, stream)); }
    Status detectMultiScale(const uint8_t* dImageGray, int imageW, int imageH, int imageStride, float scaleFactor,
                            int minNeighbors, int minObjectSize, int maxDetections,
                            std::vector<Detection>& outDetections, bool applyGrouping = true,
                            vj_stream stream = nullptr) {
        outDetections.resize(static_cast<size_t>(std::max(maxDetections, 1)));
        int n = 0;
        const int st = vj_engine_detect(e_, dImageGray, imageW, imageH, imageStride, scaleFactor, minNeighbors,
                                        minObjectSize, maxDetections, applyGrouping ? 1 : 0,
                                        reinterpret_cast<vj_detection*>(outDetections.data()),
                                        static_cast<int>(outDetections.size()), &n, stream);
        outDetections.resize(st == VJ_OK ? static_cast<size_t>(std::min<int>(n, static_cast<int>(outDetections.size())))
                                         : 0);
        return Status(st);
    }

private:
    WindowSpec win_{};
    vj_engine* e_ = nullptr;
};

// include/vj/boosting_core.cuh:20-161, src/boosting_core.cu:466-925.  The
// trainer owns [maxFeaturesPerTile][maxSamples] response and sorted-index
// buffers and the per-feature and tile bests, as the reference's does.
class AdaBoostTrainer {
public:
    AdaBoostTrainer(const WindowSpec& win, int maxSamples, int maxFeaturesPerTile)
        : win_(win), maxSamples_(maxSamples), maxFeaturesPerTile_(maxFeaturesPerTile) {
        if (vj_trainer_create(win.winW, win.winH, maxSamples, maxFeaturesPerTile, &t_) != VJ_OK)
            throw facade_error("AdaBoostTrainer");
        const size_t cells = static_cast<size_t>(maxSamples) * static_cast<size_t>(maxFeaturesPerTile);
        if (vj_malloc(reinterpret_cast<void**>(&dResp_), cells * sizeof(float)) != VJ_OK ||
            vj_malloc(reinterpret_cast<void**>(&dSorted_), cells * sizeof(uint16_t)) != VJ_OK ||
            vj_malloc(reinterpret_cast<void**>(&dBest_), static_cast<size_t>(maxFeaturesPerTile) *
                                                             sizeof(FeatureBestSplit)) != VJ_OK ||
            vj_malloc(reinterpret_cast<void**>(&dTileBest_), sizeof(TileBestCandidate)) != VJ_OK) {
            std::runtime_error err = facade_error("AdaBoostTrainer buffers");
            release();
            throw err;
        }
    }
    ~AdaBoostTrainer() { release(); }
    AdaBoostTrainer(const AdaBoostTrainer&) = delete;
    AdaBoostTrainer& operator=(const AdaBoostTrainer&) = delete;
    AdaBoostTrainer(AdaBoostTrainer&& o) noexcept { steal(o); }
    AdaBoostTrainer& operator=(AdaBoostTrainer&& o) noexcept {
        if (this != &o) {
            release();
            steal(o);
        }
        return *this;
    }

    Status bindTrainingSet(const IntegralImageSetT& iiSet, const LabelsView& labels, const WeightsView& weights) {
        vj_integral_view v{};
        v.nSamples = iiSet.nSamples;
        v.iiWidth = iiSet.iiWidth;
        v.iiHeight = iiSet.iiHeight;
        v.iiArea = iiSet.iiArea;
        v.sumT = reinterpret_cast<const uint32_t*>(iiSet.d_sumT);
        v.sqsumT = reinterpret_cast<const uint64_t*>(iiSet.d_sqsumT);
        v.invNorm = iiSet.d_invNorm;
        if (labels.nSamples != iiSet.nSamples || weights.nSamples != iiSet.nSamples) return Status::kInvalidArg;
        const int st = vj_trainer_bind(t_, &v, labels.d_label, labels.d_active, weights.d_weight, iiSet.nSamples);
        if (st == VJ_OK) n_ = iiSet.nSamples;
        return Status(st);
    }
    Status evaluateFeatureResponses(const HaarFeature* dFeatures, int featureBegin, int featureCount,
                                    float* dOutResp = nullptr, vj_stream stream = nullptr) const {
        return Status(vj_trainer_evaluate_responses(t_, reinterpret_cast<const vj_feature*>(dFeatures), featureBegin,
                                                    featureCount, dOutResp ? dOutResp : dResp_, stream));
    }
    Status evaluateAndFindThreshold(const FeatureTileView& tile, FeatureBestSplit* dOutBestPerFeature = nullptr,
                                    vj_stream stream = nullptr) const {
        return Status(vj_trainer_find_threshold(t_, tile.d_resp, tile.d_sortedIdx, tile.featureCount,
                                                dOutBestPerFeature ? dOutBestPerFeature : dBest_, stream));
    }
    Status selectBestInTile(int featureBegin, int featureCount, const uint8_t* dUsedFeatureMask,
                            TileBestCandidate* dOut = nullptr, vj_stream stream = nullptr) const {
        return Status(vj_trainer_select_best(t_, dBest_, featureBegin, featureCount, dUsedFeatureMask,
                                             dOut ? dOut : dTileBest_, stream));
    }
    Status sortSamplesPerFeature(const float* dResp, int featureCount, uint16_t* dOutSortedIdx = nullptr,
                                 vj_stream stream = nullptr) {
        return Status(vj_trainer_sort_per_feature(t_, dResp, featureCount, dOutSortedIdx ? dOutSortedIdx : dSorted_,
                                                  stream));
    }

    float* responseBuffer() noexcept { return dResp_; }
    const float* responseBuffer() const noexcept { return dResp_; }
    FeatureBestSplit* bestBuffer() noexcept { return dBest_; }
    const FeatureBestSplit* bestBuffer() const noexcept { return dBest_; }
    TileBestCandidate* tileBestBuffer() noexcept { return dTileBest_; }
    const TileBestCandidate* tileBestBuffer() const noexcept { return dTileBest_; }
    const uint16_t* sortedIndexBuffer() const noexcept { return dSorted_; }
    int maxSamples() const noexcept { return maxSamples_; }
    int maxFeaturesPerTile() const noexcept { return maxFeaturesPerTile_; }

private:
    void release() noexcept {
        vj_free(dResp_);
        vj_free(dSorted_);
        vj_free(dBest_);
        vj_free(dTileBest_);
        vj_trainer_destroy(t_);
        dResp_ = nullptr;
        dSorted_ = nullptr;
        dBest_ = nullptr;
        dTileBest_ = nullptr;
        t_ = nullptr;
    }
    void steal(AdaBoostTrainer& o) noexcept {
        win_ = o.win_;
        maxSamples_ = o.maxSamples_;
        maxFeaturesPerTile_ = o.maxFeaturesPerTile_;
        n_ = o.n_;
        t_ = std::exchange(o.t_, nullptr);
        dResp_ = std::exchange(o.dResp_, nullptr);
        dSorted_ = std::exchange(o.dSorted_, nullptr);
        dBest_ = std::exchange(o.dBest_, nullptr);
        dTileBest_ = std::exchange(o.dTileBest_, nullptr);
    }

    WindowSpec win_{};
    int maxSamples_ = 0;
    int maxFeaturesPerTile_ = 0;
    int n_ = 0;
    vj_trainer* t_ = nullptr;
    float* dResp_ = nullptr;
    uint16_t* dSorted_ = nullptr;
    FeatureBestSplit* dBest_ = nullptr;
    TileBestCandidate* dTileBest_ = nullptr;
};

}  // namespace vj

#endif  // VJ_FACADE_HPP
