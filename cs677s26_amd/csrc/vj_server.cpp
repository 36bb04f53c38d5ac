// vj_server.cpp -- served detect: a request loop with continuous batching over
// a few device slots, admitted by an HBM budget.  No reference counterpart:
// the reference detects one image per call (src/face_vision_engine.cu:1541-1747).
//
// Each slot owns one engine (its work buffers and a resident copy of the
// cascade) and one stream, and runs on its own host thread.  A free slot takes
// the oldest queued requests that fit its share of the budget and runs them as
// one batched detect; meanwhile the other slots keep their batches on the GPU,
// so new arrivals are admitted as soon as any slot frees up.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "vj_common.hpp"

using Clock = std::chrono::steady_clock;

struct vj_server {
    struct Request {
        std::vector<uint8_t> img;  // dense w*h copy
        int w = 0, h = 0;
        int64_t bytes = 0;  // admitted footprint
        std::vector<vj_detection> dets;
        int status = VJ_OK;
        std::string err;
        bool done = false;
        Clock::time_point t0, t1;
    };
    struct Slot {
        vj_engine* e = nullptr;
        hipStream_t s = nullptr;
        vj::DevBuf<uint8_t> dImg;
        vj::PinnedBuf<uint8_t> hImg;
        std::unique_ptr<vj_detection[]> out;
        size_t outCap = 0;
        std::thread th;
    };

    vj_server_config cfg{};
    int device = 0;
    int64_t slotCap = 0;  // budget share of one slot
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<int64_t, std::shared_ptr<Request>> byId;
    int64_t nextId = 1;
    bool stopping = false;
    int busy = 0;
    int64_t inFlight = 0;
    vj_server_stats st{};
    std::vector<std::unique_ptr<Slot>> slots;
};

namespace {

// Runs one batch on a slot: pack the images into pinned memory, one upload,
// one batched detect on the slot's stream, results back to the requests.
int run_batch(vj_server* S, vj_server::Slot* sl, std::vector<std::shared_ptr<vj_server::Request>>& batch) {
    const int n = static_cast<int>(batch.size());
    size_t total = 0;
    for (auto& r : batch) total += r->img.size();
    VJ_TRY(sl->hImg.ensure(total));
    VJ_TRY(sl->dImg.ensure(total));
    std::vector<const uint8_t*> ptr(static_cast<size_t>(n));
    std::vector<int> w(static_cast<size_t>(n)), h(static_cast<size_t>(n)), stride(static_cast<size_t>(n)),
        counts(static_cast<size_t>(n));
    size_t off = 0;
    for (int i = 0; i < n; ++i) {
        auto& r = batch[static_cast<size_t>(i)];
        std::memcpy(sl->hImg.get() + off, r->img.data(), r->img.size());
        ptr[static_cast<size_t>(i)] = sl->dImg.get() + off;
        w[static_cast<size_t>(i)] = r->w;
        h[static_cast<size_t>(i)] = r->h;
        stride[static_cast<size_t>(i)] = r->w;
        off += r->img.size();
    }
    VJ_HIP(hipMemcpyAsync(sl->dImg.get(), sl->hImg.get(), total, hipMemcpyHostToDevice, sl->s));
    const size_t cap = static_cast<size_t>(n) * static_cast<size_t>(S->cfg.maxDetections);
    if (sl->outCap < cap) {
        sl->out.reset(new vj_detection[cap]);
        sl->outCap = cap;
    }
    VJ_TRY(vj_engine_detect_batch(sl->e, n, ptr.data(), w.data(), h.data(), stride.data(), S->cfg.scaleFactor,
                                  S->cfg.minNeighbors, S->cfg.minObjectSize, S->cfg.maxDetections,
                                  S->cfg.applyGrouping, sl->out.get(), static_cast<int>(cap), counts.data(), sl->s));
    size_t pos = 0;
    for (int i = 0; i < n; ++i) {
        const int c = counts[static_cast<size_t>(i)];
        batch[static_cast<size_t>(i)]->dets.assign(sl->out.get() + pos, sl->out.get() + pos + c);
        pos += static_cast<size_t>(c);
    }
    return VJ_OK;
}

void slot_loop(vj_server* S, vj_server::Slot* sl) {
    (void)hipSetDevice(S->device);
    for (;;) {
        std::vector<std::shared_ptr<vj_server::Request>> batch;
        int64_t bytes = 0;
        {
            std::unique_lock<std::mutex> lk(S->mu);
            S->cvWork.wait(lk, [&] { return S->stopping || !S->pending.empty(); });
            if (S->pending.empty()) return;  // stopping, queue drained
            while (!S->pending.empty() && static_cast<int>(batch.size()) < S->cfg.maxBatch) {
                const auto& r = S->pending.front();
                if (!batch.empty() && bytes + r->bytes > S->slotCap) break;
                bytes += r->bytes;
                batch.push_back(r);
                S->pending.pop_front();
            }
            ++S->busy;
            S->inFlight += bytes;
            S->st.maxSlotsBusy = std::max(S->st.maxSlotsBusy, S->busy);
            S->st.maxBatchSeen = std::max(S->st.maxBatchSeen, static_cast<int32_t>(batch.size()));
            S->st.peakBytesInFlight = std::max(S->st.peakBytesInFlight, S->inFlight);
        }
        // more queued work: let another free slot start on it now
        S->cvWork.notify_one();
        int status = VJ_RUNTIME_ERROR;
        std::string err;
        try {
            status = run_batch(S, sl, batch);
            if (status != VJ_OK) err = vj_last_error();
        } catch (const std::exception& ex) {
            err = ex.what();
        }
        {
            std::lock_guard<std::mutex> lk(S->mu);
            --S->busy;
            S->inFlight -= bytes;
            ++S->st.batches;
            const Clock::time_point t1 = Clock::now();
            for (auto& r : batch) {
                r->status = status;
                r->err = err;
                r->t1 = t1;
                r->done = true;
                ++S->st.completed;
            }
        }
        S->cvDone.notify_all();
    }
}

}  // namespace

extern "C" int vj_server_create(const vj_server_config* cfg, const vj_feature* hF, int nF, const vj_stump* hS,
                                int nS, const vj_stage* hSt, int nSt, vj_server** out) {
    VJ_CHECK(out && cfg && hF && hS && hSt && nF > 0 && nS > 0 && nSt > 0, "vj_server_create: bad arguments");
    *out = nullptr;
    VJ_CHECK(cfg->slots >= 1 && cfg->slots <= 8 && cfg->maxBatch >= 1 && cfg->hbmBudget > 0,
             "vj_server_create: slots must be 1..8, maxBatch >= 1, hbmBudget > 0");
    VJ_CHECK(cfg->maxImageW >= cfg->winW && cfg->maxImageH >= cfg->winH, "vj_server_create: bad image limits");
    VJ_CHECK(cfg->scaleFactor > 1.0f && cfg->minObjectSize >= 1 && cfg->minNeighbors >= 0 && cfg->maxDetections >= 1,
             "vj_server_create: bad detect parameters");
    std::unique_ptr<vj_server> S(new vj_server());
    S->cfg = *cfg;
    VJ_HIP(hipGetDevice(&S->device));
    S->slotCap = cfg->hbmBudget / cfg->slots;
    // the cascade, uploaded once and copied into every slot's engine
    void *dF = nullptr, *dS = nullptr;
    VJ_TRY(vj_malloc(&dF, sizeof(vj_feature) * static_cast<size_t>(nF)));
    int status = vj_malloc(&dS, sizeof(vj_stump) * static_cast<size_t>(nS));
    if (status == VJ_OK) status = vj_memcpy_htod(dF, hF, sizeof(vj_feature) * static_cast<size_t>(nF), nullptr);
    if (status == VJ_OK) status = vj_memcpy_htod(dS, hS, sizeof(vj_stump) * static_cast<size_t>(nS), nullptr);
    for (int k = 0; k < cfg->slots && status == VJ_OK; ++k) {
        auto sl = std::make_unique<vj_server::Slot>();
        status = vj_engine_create(cfg->winW, cfg->winH, 1, cfg->maxImageW, cfg->maxImageH, 0, cfg->maxDetections,
                                  &sl->e);
        if (status != VJ_OK) break;
        if (hipStreamCreateWithFlags(&sl->s, hipStreamNonBlocking) != hipSuccess) {
            status = ::vj::hip_fail(hipGetLastError(), "hipStreamCreateWithFlags", __FILE__, __LINE__);
            vj_engine_destroy(sl->e);
            break;
        }
        status = vj_engine_set_cascade(sl->e, static_cast<const vj_feature*>(dF), nF,
                                       static_cast<const vj_stump*>(dS), nS, hSt, nSt, sl->s);
        S->slots.push_back(std::move(sl));
    }
    if (dF) (void)vj_free(dF);
    if (dS) (void)vj_free(dS);
    if (status != VJ_OK) {
        for (auto& sl : S->slots) {
            vj_engine_destroy(sl->e);
            (void)hipStreamDestroy(sl->s);
        }
        return status;
    }
    for (auto& sl : S->slots) sl->th = std::thread(slot_loop, S.get(), sl.get());
    *out = S.release();
    return VJ_OK;
}

extern "C" int vj_server_submit(vj_server* S, const uint8_t* hImage, int w, int h, int stride, int64_t* ticket) {
    VJ_CHECK(S && hImage && ticket && w > 0 && h > 0 && stride >= w, "vj_server_submit: bad arguments");
    VJ_CHECK(w <= S->cfg.maxImageW && h <= S->cfg.maxImageH, "vj_server_submit: image larger than maxImageW x maxImageH");
    int64_t bytes = 0;
    VJ_TRY(vj_engine_detect_footprint(S->slots[0]->e, w, h, S->cfg.scaleFactor, S->cfg.minObjectSize,
                                      S->cfg.maxDetections, &bytes));
    if (bytes > S->slotCap) {
        std::lock_guard<std::mutex> lk(S->mu);
        ++S->st.rejected;
        ::vj::set_error("vj_server_submit: request needs " + std::to_string(bytes) +
                        " device bytes, above one slot's budget of " + std::to_string(S->slotCap));
        return VJ_OUT_OF_MEMORY;
    }
    auto r = std::make_shared<vj_server::Request>();
    r->w = w;
    r->h = h;
    r->bytes = bytes;
    r->img.resize(static_cast<size_t>(w) * static_cast<size_t>(h));
    for (int y = 0; y < h; ++y)
        std::memcpy(r->img.data() + static_cast<size_t>(y) * w, hImage + static_cast<size_t>(y) * stride,
                    static_cast<size_t>(w));
    r->t0 = Clock::now();
    {
        std::lock_guard<std::mutex> lk(S->mu);
        *ticket = S->nextId++;
        S->byId[*ticket] = r;
        S->pending.push_back(r);
        ++S->st.submitted;
    }
    S->cvWork.notify_one();
    return VJ_OK;
}

extern "C" int vj_server_wait(vj_server* S, int64_t ticket, vj_detection* hOut, int cap, int* count,
                              double* latencyMs) {
    VJ_CHECK(S && count && cap >= 0 && (cap == 0 || hOut), "vj_server_wait: bad arguments");
    *count = 0;
    std::shared_ptr<vj_server::Request> r;
    {
        std::unique_lock<std::mutex> lk(S->mu);
        auto it = S->byId.find(ticket);
        VJ_CHECK(it != S->byId.end(), "vj_server_wait: unknown or already waited ticket");
        r = it->second;
        S->cvDone.wait(lk, [&] { return r->done; });
        S->byId.erase(ticket);
    }
    if (latencyMs) *latencyMs = std::chrono::duration<double, std::milli>(r->t1 - r->t0).count();
    if (r->status != VJ_OK) {
        ::vj::set_error(r->err);
        return r->status;
    }
    *count = static_cast<int>(r->dets.size());
    const size_t m = std::min(r->dets.size(), static_cast<size_t>(cap));
    if (m) std::memcpy(hOut, r->dets.data(), m * sizeof(vj_detection));
    return VJ_OK;
}

extern "C" int vj_server_stats_get(vj_server* S, vj_server_stats* out) {
    VJ_CHECK(S && out, "vj_server_stats_get: bad arguments");
    std::lock_guard<std::mutex> lk(S->mu);
    *out = S->st;
    return VJ_OK;
}

extern "C" void vj_server_destroy(vj_server* S) {
    if (!S) return;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stopping = true;
    }
    S->cvWork.notify_all();
    for (auto& sl : S->slots)
        if (sl->th.joinable()) sl->th.join();
    for (auto& sl : S->slots) {
        vj_engine_destroy(sl->e);
        (void)hipStreamDestroy(sl->s);
    }
    delete S;
}
