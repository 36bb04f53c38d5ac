// vj_host.cpp -- host-side parts of the drop-in boundary: error reporting,
// the Haar feature pool, the .bin model format, rectangle grouping and
// PGM/PPM I/O.  None of this runs on the GPU in the reference either.
#include <algorithm>
#include <atomic>
#include <cctype>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <dirent.h>
#include <sys/stat.h>
#include <fstream>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "vj_common.hpp"

namespace vj {

namespace {
thread_local std::string g_error;
}

void set_error(const std::string& msg) { g_error = msg; }

int hip_fail(hipError_t e, const char* what, const char* file, int line) {
    g_error = std::string(hipGetErrorName(e)) + " (" + hipGetErrorString(e) + ") in " + what +
              " at " + file + ":" + std::to_string(line);
    (void)hipGetLastError();
    return VJ_DEVICE_ERROR;
}

// ---------------------------------------------------------------------------
// Haar pool.  Enumeration order and rectangles follow
// src/face_vision_engine.cu:882-1020: x, y, dx, dy loops; within each, the
// five Basic templates (x2, y2, x3, y3, x2y2), then for Core x4, y4 and the
// centre template, then for All the six tilted ones.  Weights -1/+2 (and +9
// for the centre, +3 for tilted x3/y3) are the reference's.
// ---------------------------------------------------------------------------
namespace {

struct PoolWriter {
    vj_feature* out;
    int cap;
    int n = 0;
    int iiW;

    vj_rect rect(int x, int y, int w, int h, float wt, bool tilted) const {
        vj_rect r;
        if (!tilted) {
            r.p0 = x + iiW * y;
            r.p1 = (x + w) + iiW * y;
            r.p2 = x + iiW * (y + h);
            r.p3 = (x + w) + iiW * (y + h);
        } else {  // src/face_vision_engine.cu:626-631
            r.p0 = x + iiW * y;
            r.p1 = (x - h) + iiW * (y + h);
            r.p2 = (x + w) + iiW * (y + w);
            r.p3 = (x + w - h) + iiW * (y + w + h);
        }
        r.w = wt;
        return r;
    }
    void emit(bool tilted, std::initializer_list<vj_rect> rs) {
        if (out && n < cap) {
            vj_feature f;
            std::memset(&f, 0, sizeof f);
            int i = 0;
            for (const vj_rect& r : rs) f.r[i++] = r;
            f.rectCount = static_cast<uint8_t>(i);
            f.tilted = tilted ? 1 : 0;
            out[n] = f;
        }
        ++n;
    }
};

}  // namespace

}  // namespace vj

using namespace vj;

extern "C" const char* vj_last_error(void) { return g_error.c_str(); }

extern "C" int vj_haar_pool(int W, int H, int mode, vj_feature* out, int cap, int* count) {
    VJ_CHECK(W > 0 && H > 0 && mode >= 0 && mode <= 2 && count, "vj_haar_pool: bad arguments");
    PoolWriter pw{out, cap, 0, W + 1};
    const bool core = mode != 0, tilt = mode == 2;
    for (int x = 0; x < W; ++x) {
        for (int y = 0; y < H; ++y) {
            for (int dx = 1; dx <= W; ++dx) {
                for (int dy = 1; dy <= H; ++dy) {
                    if (x + 2 * dx <= W && y + dy <= H)
                        pw.emit(false, {pw.rect(x, y, 2 * dx, dy, -1.f, false), pw.rect(x + dx, y, dx, dy, 2.f, false)});
                    if (x + dx <= W && y + 2 * dy <= H)
                        pw.emit(false, {pw.rect(x, y, dx, 2 * dy, -1.f, false), pw.rect(x, y + dy, dx, dy, 2.f, false)});
                    if (x + 3 * dx <= W && y + dy <= H)
                        pw.emit(false, {pw.rect(x, y, 3 * dx, dy, -1.f, false), pw.rect(x + dx, y, dx, dy, 2.f, false)});
                    if (x + dx <= W && y + 3 * dy <= H)
                        pw.emit(false, {pw.rect(x, y, dx, 3 * dy, -1.f, false), pw.rect(x, y + dy, dx, dy, 2.f, false)});
                    if (x + 2 * dx <= W && y + 2 * dy <= H)
                        pw.emit(false, {pw.rect(x, y, 2 * dx, 2 * dy, -1.f, false), pw.rect(x, y, dx, dy, 2.f, false),
                                        pw.rect(x + dx, y + dy, dx, dy, 2.f, false)});
                    if (core) {
                        if (x + 4 * dx <= W && y + dy <= H)
                            pw.emit(false, {pw.rect(x, y, 4 * dx, dy, -1.f, false), pw.rect(x + dx, y, 2 * dx, dy, 2.f, false)});
                        if (x + dx <= W && y + 4 * dy <= H)
                            pw.emit(false, {pw.rect(x, y, dx, 4 * dy, -1.f, false), pw.rect(x, y + dy, dx, 2 * dy, 2.f, false)});
                        if (x + 3 * dx <= W && y + 3 * dy <= H)
                            pw.emit(false, {pw.rect(x, y, 3 * dx, 3 * dy, -1.f, false), pw.rect(x + dx, y + dy, dx, dy, 9.f, false)});
                    }
                    if (tilt) {
                        if (x + 2 * dx <= W && y + 2 * dx + dy <= H && x - dy >= 0)
                            pw.emit(true, {pw.rect(x, y, 2 * dx, dy, -1.f, true), pw.rect(x, y, dx, dy, 2.f, true)});
                        if (x + dx <= W && y + dx + 2 * dy <= H && x - 2 * dy >= 0)
                            pw.emit(true, {pw.rect(x, y, dx, 2 * dy, -1.f, true), pw.rect(x, y, dx, dy, 2.f, true)});
                        if (x + 3 * dx <= W && y + 3 * dx + dy <= H && x - dy >= 0)
                            pw.emit(true, {pw.rect(x, y, 3 * dx, dy, -1.f, true), pw.rect(x + dx, y + dx, dx, dy, 3.f, true)});
                        if (x + dx <= W && y + dx + 3 * dy <= H && x - 3 * dy >= 0)
                            pw.emit(true, {pw.rect(x, y, dx, 3 * dy, -1.f, true), pw.rect(x - dy, y + dy, dx, dy, 3.f, true)});
                        if (x + 4 * dx <= W && y + 4 * dx + dy <= H && x - dy >= 0)
                            pw.emit(true, {pw.rect(x, y, 4 * dx, dy, -1.f, true), pw.rect(x + dx, y + dx, 2 * dx, dy, 2.f, true)});
                        if (x + dx <= W && y + dx + 4 * dy <= H && x - 4 * dy >= 0)
                            pw.emit(true, {pw.rect(x, y, dx, 4 * dy, -1.f, true), pw.rect(x - dy, y + dy, dx, 2 * dy, 2.f, true)});
                    }
                }
            }
        }
    }
    *count = pw.n;
    return VJ_OK;
}

// ---------------------------------------------------------------------------
// .bin model: [header 48 B][HaarFeature x nF][GpuStump4 x nS][CascadeStage x nSt]
// Header: magic "VJCBIN1\0", u32 version 1, i32 winW, i32 winH, 4 zero bytes
// of padding, u64 featureCount, u64 stumpCount, u64 stageCount; native
// little-endian (src/model_io.cpp:15-28, 66-167).
// ---------------------------------------------------------------------------
namespace {

struct ModelHeader {
    char magic[8];
    uint32_t version;
    int32_t winW;
    int32_t winH;
    uint32_t pad;
    uint64_t featureCount;
    uint64_t stumpCount;
    uint64_t stageCount;
};
static_assert(sizeof(ModelHeader) == 48, "the .bin header is 48 bytes");
const char kMagic[8] = {'V', 'J', 'C', 'B', 'I', 'N', '1', '\0'};

struct File {
    FILE* f;
    explicit File(FILE* p) : f(p) {}
    ~File() {
        if (f) std::fclose(f);
    }
};

int read_header(FILE* f, ModelHeader& h, const char* path) {
    if (std::fread(&h, sizeof h, 1, f) != 1) {
        set_error(std::string("model file too short: ") + path);
        return VJ_INVALID_ARG;
    }
    // src/model_io.cpp:148-152: magic, version, positive window, non-zero counts.
    if (std::memcmp(h.magic, kMagic, 8) != 0 || h.version != 1u || h.winW <= 0 || h.winH <= 0 ||
        h.featureCount == 0 || h.stumpCount == 0 || h.stageCount == 0) {
        set_error(std::string("not a VJCBIN1 v1 model (bad magic, version, window or counts): ") + path);
        return VJ_INVALID_ARG;
    }
    return VJ_OK;
}

}  // namespace

extern "C" int vj_model_save(const char* path, int winW, int winH, const vj_feature* feats,
                             int64_t nF, const vj_stump* stumps, int64_t nS, const vj_stage* stages,
                             int64_t nSt) {
    // src/model_io.cpp:84-86: every array must be non-empty.
    VJ_CHECK(path && path[0] && winW > 0 && winH > 0 && feats && nF > 0 && stumps && nS > 0 &&
                 stages && nSt > 0,
             "vj_model_save: empty model or bad arguments");
    File fp(std::fopen(path, "wb"));
    VJ_CHECK(fp.f, std::string("cannot open for writing: ") + path);
    ModelHeader h;
    std::memset(&h, 0, sizeof h);
    std::memcpy(h.magic, kMagic, 8);
    h.version = 1;
    h.winW = winW;
    h.winH = winH;
    h.featureCount = static_cast<uint64_t>(nF);
    h.stumpCount = static_cast<uint64_t>(nS);
    h.stageCount = static_cast<uint64_t>(nSt);
    bool ok = std::fwrite(&h, sizeof h, 1, fp.f) == 1 &&
              std::fwrite(feats, sizeof(vj_feature), static_cast<size_t>(nF), fp.f) == static_cast<size_t>(nF) &&
              std::fwrite(stumps, sizeof(vj_stump), static_cast<size_t>(nS), fp.f) == static_cast<size_t>(nS) &&
              std::fwrite(stages, sizeof(vj_stage), static_cast<size_t>(nSt), fp.f) == static_cast<size_t>(nSt);
    if (!ok) {
        set_error(std::string("short write: ") + path);
        return VJ_RUNTIME_ERROR;
    }
    return VJ_OK;
}

extern "C" int vj_model_peek(const char* path, int* winW, int* winH, int64_t* nF, int64_t* nS,
                             int64_t* nSt) {
    VJ_CHECK(path && path[0], "vj_model_peek: empty path");
    File fp(std::fopen(path, "rb"));
    VJ_CHECK(fp.f, std::string("cannot open model: ") + path);
    ModelHeader h;
    VJ_TRY(read_header(fp.f, h, path));
    if (winW) *winW = h.winW;
    if (winH) *winH = h.winH;
    if (nF) *nF = static_cast<int64_t>(h.featureCount);
    if (nS) *nS = static_cast<int64_t>(h.stumpCount);
    if (nSt) *nSt = static_cast<int64_t>(h.stageCount);
    return VJ_OK;
}

extern "C" int vj_model_load(const char* path, vj_feature* feats, int64_t capF, vj_stump* stumps,
                             int64_t capS, vj_stage* stages, int64_t capSt) {
    VJ_CHECK(path && path[0], "vj_model_load: empty path");
    File fp(std::fopen(path, "rb"));
    VJ_CHECK(fp.f, std::string("cannot open model: ") + path);
    ModelHeader h;
    VJ_TRY(read_header(fp.f, h, path));
    VJ_CHECK(feats && stumps && stages && static_cast<uint64_t>(capF) >= h.featureCount &&
                 static_cast<uint64_t>(capS) >= h.stumpCount && static_cast<uint64_t>(capSt) >= h.stageCount,
             "vj_model_load: output arrays too small");
    bool ok = std::fread(feats, sizeof(vj_feature), h.featureCount, fp.f) == h.featureCount &&
              std::fread(stumps, sizeof(vj_stump), h.stumpCount, fp.f) == h.stumpCount &&
              std::fread(stages, sizeof(vj_stage), h.stageCount, fp.f) == h.stageCount;
    VJ_CHECK(ok, std::string("model file truncated: ") + path);
    return VJ_OK;
}

// ---------------------------------------------------------------------------
// Rectangle grouping (host), src/face_vision_engine.cu:585-616, 1401-1539.
// Sweep over boxes sorted by (x, index) joining similar boxes in a union-find
// (union by rank, path compression); clusters with more than
// max(1, minNeighbors) members become the integer-mean box with the first
// maximal score; a second sweep drops clusters that sit inside a stronger one.
// ---------------------------------------------------------------------------
namespace vj {

namespace {

// Union-find over caller-owned arrays (union by rank, path compression).
struct UnionFind {
    int* parent;
    int* rank;
    int root(int x) {
        int r = x;
        while (parent[r] != r) r = parent[r];
        while (parent[x] != r) {
            int nx = parent[x];
            parent[x] = r;
            x = nx;
        }
        return r;
    }
    // Joins the set whose root is ra with b's set (union by rank) and returns
    // the root of the union.
    int join_root(int ra, int b) {
        int rb = root(b);
        if (ra == rb) return ra;
        if (rank[ra] < rank[rb]) std::swap(ra, rb);
        parent[rb] = ra;
        if (rank[ra] == rank[rb]) ++rank[ra];
        return ra;
    }
};

// A few persistent host threads for data-parallel host loops.  run() hands
// tasks 0..n-1 out dynamically to the workers and the caller and returns
// when all are done; a caller that finds the pool busy (another thread is
// using it) runs its tasks itself.
class HostPool {
public:
    static HostPool& get() {
        static HostPool pool;
        return pool;
    }
    void run(int nTasks, const std::function<void(int)>& fn) {
        std::unique_lock<std::mutex> use(useMu_, std::try_to_lock);
        if (!use.owns_lock() || workers_.empty() || nTasks <= 1) {
            for (int t = 0; t < nTasks; ++t) fn(t);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &fn;
            total_ = nTasks;
            next_.store(0);
            active_ = static_cast<int>(workers_.size());
            ++gen_;
        }
        cv_.notify_all();
        drain(fn);
        std::unique_lock<std::mutex> lk(mu_);
        done_.wait(lk, [&] { return active_ == 0; });
        job_ = nullptr;
    }
    ~HostPool() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
            ++gen_;
        }
        cv_.notify_all();
        for (std::thread& t : workers_) t.join();
    }

private:
    HostPool() {
        const unsigned hw = std::thread::hardware_concurrency();
        const int n = static_cast<int>(std::min(hw > 1 ? hw - 1 : 0u, 15u));
        for (int k = 0; k < n; ++k) workers_.emplace_back([this] { loop(); });
    }
    void drain(const std::function<void(int)>& fn) {
        for (int t = next_.fetch_add(1); t < total_; t = next_.fetch_add(1)) fn(t);
    }
    void loop() {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return gen_ != seen; });
                seen = gen_;
                if (stop_) return;
                job = job_;
            }
            if (job) drain(*job);
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--active_ == 0) done_.notify_all();
            }
        }
    }
    std::vector<std::thread> workers_;
    std::mutex useMu_, mu_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    std::atomic<int> next_{0};
    int total_ = 0, active_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

struct ClusterAcc {
    int64_t x, y, w, h;
    int count;
    float best;
    int scaleIdx;
};

// Working memory of one grouping, kept between calls so that a call
// allocates only when its input outgrows every earlier one.
struct GroupScratch {
    const vj_detection* in = nullptr;
    int n = 0, nBlocks = 0;
    float eps = 0.0f;
    std::vector<int> parent, rank, idx, X, Y, W, H, cnt, members, corder;
    std::vector<uint64_t> key;
    std::vector<std::vector<int>> edges;
    std::vector<ClusterAcc> acc;
    std::vector<vj_detection> clusters;
};

constexpr int kGroupBlock = 32;  // sweep positions per pair-search task

// Positions 0..n-1 of boxes in (x, index) order into idx: a stable counting
// sort over x when the x range is moderate, else a sort of (x, index) keys.
void sweep_order(const vj_detection* d, int n, GroupScratch& g) {
    g.idx.resize(static_cast<size_t>(n));
    int lo = d[0].x, hi = d[0].x;
    for (int i = 1; i < n; ++i) {
        lo = std::min(lo, d[i].x);
        hi = std::max(hi, d[i].x);
    }
    const int64_t range = static_cast<int64_t>(hi) - lo + 1;
    if (range <= 4ll * n + 65536) {
        g.cnt.assign(static_cast<size_t>(range) + 1, 0);
        for (int i = 0; i < n; ++i) ++g.cnt[static_cast<size_t>(d[i].x - lo) + 1];
        for (int64_t v = 1; v <= range; ++v) g.cnt[static_cast<size_t>(v)] += g.cnt[static_cast<size_t>(v - 1)];
        for (int i = 0; i < n; ++i) g.idx[static_cast<size_t>(g.cnt[static_cast<size_t>(d[i].x - lo)]++)] = i;
        return;
    }
    g.key.resize(static_cast<size_t>(n));
    for (int i = 0; i < n; ++i)
        g.key[static_cast<size_t>(i)] =
            (static_cast<uint64_t>(static_cast<uint32_t>(d[i].x) ^ 0x80000000u) << 32) | static_cast<uint32_t>(i);
    std::sort(g.key.begin(), g.key.end());
    for (int k = 0; k < n; ++k) g.idx[static_cast<size_t>(k)] = static_cast<int>(g.key[static_cast<size_t>(k)] & 0xffffffffu);
}

// Grouping in three steps: prepare (sweep order, the boxes' edges in that
// order as flat arrays, union-find reset), the pair search in blocks of
// sweep positions (independent, any thread), and finish (the unions, the
// clusters, the containment pass).  n >= 1.
void group_prepare(GroupScratch& g, const vj_detection* in, int n, float eps) {
    g.in = in;
    g.n = n;
    g.eps = eps;
    const size_t un = static_cast<size_t>(n);
    g.parent.resize(un);
    g.rank.assign(un, 0);
    std::iota(g.parent.begin(), g.parent.end(), 0);
    sweep_order(in, n, g);
    g.X.resize(un);
    g.Y.resize(un);
    g.W.resize(un);
    g.H.resize(un);
    for (int k = 0; k < n; ++k) {
        const vj_detection& b = in[g.idx[static_cast<size_t>(k)]];
        g.X[static_cast<size_t>(k)] = b.x;
        g.Y[static_cast<size_t>(k)] = b.y;
        g.W[static_cast<size_t>(k)] = b.w;
        g.H[static_cast<size_t>(k)] = b.h;
    }
    g.nBlocks = (n + kGroupBlock - 1) / kGroupBlock;
    if (g.edges.size() < static_cast<size_t>(g.nBlocks)) g.edges.resize(static_cast<size_t>(g.nBlocks));
}

// The similar pairs (a, b), a before b in sweep order, for the a's of one
// block: for each a the candidates are the b's whose x is within a's reach;
// their similarRects test (the tolerance (int)(eps * (min w + min h) *
// 0.5f), every edge within it) is evaluated branch-free over runs of 64
// candidates, then the hits are collected in order.  The block's list holds,
// for each a with hits, a marker -1 - a followed by its b's.
void group_find_block(GroupScratch& g, int blk) {
    const int n = g.n;
    const float eps = g.eps;
    const int* X = g.X.data();
    const int* Y = g.Y.data();
    const int* W = g.W.data();
    const int* H = g.H.data();
    std::vector<int>& e = g.edges[static_cast<size_t>(blk)];
    e.clear();
    uint8_t flag[64];
    for (int a = blk * kGroupBlock; a < std::min(n, (blk + 1) * kGroupBlock); ++a) {
        const size_t mark = e.size();
        e.push_back(-1 - a);
        const int ax = X[a], ay = Y[a], aw = W[a], ah = H[a];
        const int ar = ax + aw, ab = ay + ah;
        const int reach = static_cast<int>(eps * static_cast<float>(aw + ah) * 0.5f);
        int end = a + 1;
        while (end < n && X[end] - ax <= reach) ++end;
        for (int b0 = a + 1; b0 < end; b0 += 64) {
            const int m = std::min(64, end - b0);
            const int* xb = X + b0;
            const int* yb = Y + b0;
            const int* wb = W + b0;
            const int* hb = H + b0;
            for (int k = 0; k < m; ++k) {
                const int d = static_cast<int>(eps * static_cast<float>(std::min(aw, wb[k]) + std::min(ah, hb[k])) * 0.5f);
                flag[k] = static_cast<uint8_t>((std::abs(ax - xb[k]) <= d) & (std::abs(ay - yb[k]) <= d) &
                                               (std::abs(ar - (xb[k] + wb[k])) <= d) &
                                               (std::abs(ab - (yb[k] + hb[k])) <= d));
            }
            for (int k = 0; k < m; ++k)
                if (flag[k]) e.push_back(b0 + k);
        }
        if (e.size() == mark + 1) e.pop_back();
    }
}

// The clusters of the unions made so far (integer-mean boxes with the first
// best score, more than max(1, minNeighbors) members), then the containment
// pass, src/face_vision_engine.cu:1447-1537.
int group_clusters(GroupScratch& g, int minNeighbors, vj_detection* out, int* count) {
    const int n = g.n;
    const float eps = g.eps;
    const vj_detection* in = g.in;
    UnionFind uf{g.parent.data(), g.rank.data()};
    g.acc.resize(static_cast<size_t>(n));
    ClusterAcc* acc = g.acc.data();
    for (int i = 0; i < n; ++i) acc[i] = ClusterAcc{0, 0, 0, 0, 0, -std::numeric_limits<float>::infinity(), 0};
    for (int i = 0; i < n; ++i) {
        ClusterAcc& c = acc[uf.root(i)];
        c.x += in[i].x;
        c.y += in[i].y;
        c.w += in[i].w;
        c.h += in[i].h;
        c.count += 1;
        if (in[i].score > c.best) {
            c.best = in[i].score;
            c.scaleIdx = in[i].scaleIdx;
        }
    }
    const int keep = std::max(1, minNeighbors);
    g.clusters.clear();
    g.members.clear();
    for (int i = 0; i < n; ++i) {
        const ClusterAcc& c = acc[i];
        if (g.parent[static_cast<size_t>(i)] != i || c.count <= keep) continue;
        vj_detection d;
        d.x = static_cast<int>(c.x / c.count);
        d.y = static_cast<int>(c.y / c.count);
        d.w = static_cast<int>(c.w / c.count);
        d.h = static_cast<int>(c.h / c.count);
        d.scaleIdx = c.scaleIdx;
        d.score = c.best;
        g.clusters.push_back(d);
        g.members.push_back(c.count);
    }
    const int nc = static_cast<int>(g.clusters.size());
    const vj_detection* clusters = g.clusters.data();
    int maxW = 0;
    for (int i = 0; i < nc; ++i) maxW = std::max(maxW, clusters[i].w);
    g.corder.resize(static_cast<size_t>(nc));
    std::iota(g.corder.begin(), g.corder.end(), 0);
    std::sort(g.corder.begin(), g.corder.end(),
              [clusters](int a, int b) { return clusters[a].x != clusters[b].x ? clusters[a].x < clusters[b].x : a < b; });
    int no = 0;
    for (int i = 0; i < nc; ++i) {
        const vj_detection& r1 = clusters[i];
        const int n1 = g.members[static_cast<size_t>(i)];
        const float lo = static_cast<float>(r1.x + r1.w) - (1.0f + eps) * static_cast<float>(maxW);
        const float hi = static_cast<float>(r1.x) + eps * static_cast<float>(maxW);
        bool inside = false;
        // the clusters left of lo are skipped by the scan; start past them
        const int loI = static_cast<int>(lo);
        auto it = std::lower_bound(g.corder.begin(), g.corder.end(), loI,
                                   [clusters](int o, int v) { return clusters[o].x < v; });
        for (; it != g.corder.end(); ++it) {
            const int o = *it;
            const vj_detection& r2 = clusters[o];
            if (r2.x > static_cast<int>(hi)) break;
            if (o == i) continue;
            const int n2 = g.members[static_cast<size_t>(o)];
            const int dx = static_cast<int>(static_cast<float>(r2.w) * eps);
            const int dy = static_cast<int>(static_cast<float>(r2.h) * eps);
            if (r1.x >= r2.x - dx && r1.y >= r2.y - dy && r1.x + r1.w <= r2.x + r2.w + dx &&
                r1.y + r1.h <= r2.y + r2.h + dy && (n2 > std::max(3, n1) || n1 < 3)) {
                inside = true;
                break;
            }
        }
        if (!inside) out[no++] = r1;
    }
    *count = no;
    return VJ_OK;
}

// The unions in exactly the sequential sweep's order (a ascending, then b
// ascending) from the pair-search blocks, then the clusters.
int group_finish(GroupScratch& g, int minNeighbors, vj_detection* out, int* count) {
    const int* idx = g.idx.data();
    UnionFind uf{g.parent.data(), g.rank.data()};
    for (int blk = 0; blk < g.nBlocks; ++blk) {
        int ra = -1, a = -1;  // ra: root of a's set once a has been joined
        for (int v : g.edges[static_cast<size_t>(blk)]) {
            if (v < 0) {
                a = -1 - v;
                ra = -1;
            } else {
                ra = uf.join_root(ra < 0 ? uf.root(idx[a]) : ra, idx[v]);
            }
        }
    }
    return group_clusters(g, minNeighbors, out, count);
}

}  // namespace

int group_rectangles(const vj_detection* in, int n, int minNeighbors, float eps, vj_detection* out,
                     int* count) {
    *count = 0;
    if (n <= 0) return VJ_OK;
    thread_local GroupScratch scratch;
    GroupScratch& g = scratch;
    group_prepare(g, in, n, eps);
    if (n >= 1024) {
        // (an allocation failure on a pool thread is reported, not thrown there)
        std::vector<char> failed(static_cast<size_t>(g.nBlocks), 0);
        HostPool::get().run(g.nBlocks, [&](int blk) {
            try {
                group_find_block(g, blk);
            } catch (const std::exception&) {
                failed[static_cast<size_t>(blk)] = 1;
            }
        });
        for (char f : failed)
            if (f) return VJ_OUT_OF_MEMORY;
    } else {
        for (int blk = 0; blk < g.nBlocks; ++blk) group_find_block(g, blk);
    }
    return group_finish(g, minNeighbors, out, count);
}

int group_rectangles_many(const vj_detection* const* in, const int* n, int nImages, int minNeighbors, float eps,
                          vj_detection* const* out, int* count) {
    // every image's steps on the host pool: prepares, then all pair-search
    // blocks of all images, then finishes, largest images first
    // (a named reference, so that the pool's workers reach this thread's
    // scratch: a thread_local named inside a lambda is the running thread's)
    thread_local std::vector<GroupScratch> scratchTls;
    std::vector<GroupScratch>& scratch = scratchTls;
    if (scratch.size() < static_cast<size_t>(nImages)) scratch.resize(static_cast<size_t>(nImages));
    std::vector<int> order;
    for (int i = 0; i < nImages; ++i) {
        count[i] = 0;
        if (n[i] > 0) order.push_back(i);
    }
    std::sort(order.begin(), order.end(), [n](int a, int b) { return n[a] != n[b] ? n[a] > n[b] : a < b; });
    const int m = static_cast<int>(order.size());
    HostPool& pool = HostPool::get();
    // (an allocation failure on a pool thread is recorded, not thrown there)
    std::vector<char> failed(static_cast<size_t>(m), 0);
    pool.run(m, [&](int k) {
        const int i = order[static_cast<size_t>(k)];
        try {
            group_prepare(scratch[static_cast<size_t>(i)], in[i], n[i], eps);
        } catch (const std::exception&) {
            failed[static_cast<size_t>(k)] = 1;
        }
    });
    for (char f : failed)
        if (f) return VJ_OUT_OF_MEMORY;
    std::vector<std::pair<int, int>> tasks;
    for (int i : order)
        for (int blk = 0; blk < scratch[static_cast<size_t>(i)].nBlocks; ++blk) tasks.emplace_back(i, blk);
    std::vector<char> taskFailed(tasks.size(), 0);
    pool.run(static_cast<int>(tasks.size()), [&](int t) {
        try {
            group_find_block(scratch[static_cast<size_t>(tasks[static_cast<size_t>(t)].first)],
                             tasks[static_cast<size_t>(t)].second);
        } catch (const std::exception&) {
            taskFailed[static_cast<size_t>(t)] = 1;
        }
    });
    for (char f : taskFailed)
        if (f) return VJ_OUT_OF_MEMORY;
    std::vector<int> status(static_cast<size_t>(nImages), VJ_OK);
    pool.run(m, [&](int k) {
        const int i = order[static_cast<size_t>(k)];
        try {
            GroupScratch& g = scratch[static_cast<size_t>(i)];
            status[static_cast<size_t>(i)] = group_finish(g, minNeighbors, out[i], &count[i]);
        } catch (const std::exception&) {
            status[static_cast<size_t>(i)] = VJ_OUT_OF_MEMORY;
        }
    });
    for (int i = 0; i < nImages; ++i)
        if (status[static_cast<size_t>(i)] != VJ_OK) return status[static_cast<size_t>(i)];
    return VJ_OK;
}

}  // namespace vj

extern "C" int vj_group_rectangles_many(const vj_detection* const* in, const int* n, int nImages, int minNeighbors,
                                        float eps, vj_detection* const* out, int* count) {
    VJ_CHECK(nImages >= 0 && (nImages == 0 || (in && n && out && count)), "vj_group_rectangles_many: bad arguments");
    for (int i = 0; i < nImages; ++i)
        VJ_CHECK(n[i] >= 0 && (n[i] == 0 || (in[i] && out[i])), "vj_group_rectangles_many: bad list");
    try {
        return vj::group_rectangles_many(in, n, nImages, minNeighbors, eps, out, count);
    } catch (const std::exception& ex) {
        vj::set_error(std::string("vj_group_rectangles_many: ") + ex.what());
        return VJ_OUT_OF_MEMORY;
    }
}

extern "C" int vj_group_rectangles(const vj_detection* in, int n, int minNeighbors, float eps,
                                   vj_detection* out, int* count) {
    VJ_CHECK(count && n >= 0 && (n == 0 || (in && out)), "vj_group_rectangles: bad arguments");
    try {
        return group_rectangles(in, n, minNeighbors, eps, out, count);
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return VJ_RUNTIME_ERROR;
    }
}

// ---------------------------------------------------------------------------
// PGM (P5, maxVal <= 255, '#' comments) and PPM (P6) -- src/main.cu:286-410.
// ---------------------------------------------------------------------------
namespace {

bool next_token(std::istream& is, std::string& tok) {
    tok.clear();
    while (is.good()) {
        const int c = is.peek();
        if (c == '#') {
            std::string line;
            std::getline(is, line);
            continue;
        }
        if (c != EOF && std::isspace(static_cast<unsigned char>(c))) {
            is.get();
            continue;
        }
        break;
    }
    if (!is.good()) return false;
    is >> tok;
    return !tok.empty();
}

int pgm_open(const char* path, std::ifstream& ifs, int& w, int& h) {
    ifs.open(path, std::ios::binary);
    VJ_CHECK(ifs, std::string("cannot open PGM: ") + path);
    std::string tok;
    VJ_CHECK(next_token(ifs, tok) && tok == "P5", std::string("unsupported PGM format (need P5): ") + path);
    int maxVal = 0;
    try {
        VJ_CHECK(next_token(ifs, tok), "PGM header truncated");
        w = std::stoi(tok);
        VJ_CHECK(next_token(ifs, tok), "PGM header truncated");
        h = std::stoi(tok);
        VJ_CHECK(next_token(ifs, tok), "PGM header truncated");
        maxVal = std::stoi(tok);
    } catch (...) {
        set_error("invalid PGM header");
        return VJ_INVALID_ARG;
    }
    VJ_CHECK(w > 0 && h > 0 && maxVal > 0 && maxVal <= 255, "invalid PGM metadata");
    ifs.get();  // the single whitespace byte after maxVal
    return VJ_OK;
}

}  // namespace

extern "C" int vj_pgm_peek(const char* path, int* w, int* h) {
    VJ_CHECK(path && w && h, "vj_pgm_peek: bad arguments");
    std::ifstream ifs;
    return pgm_open(path, ifs, *w, *h);
}

extern "C" int vj_pgm_load(const char* path, uint8_t* out, size_t cap) {
    VJ_CHECK(path && out, "vj_pgm_load: bad arguments");
    std::ifstream ifs;
    int w = 0, h = 0;
    VJ_TRY(pgm_open(path, ifs, w, h));
    const size_t n = static_cast<size_t>(w) * static_cast<size_t>(h);
    VJ_CHECK(cap >= n, "vj_pgm_load: output buffer too small");
    VJ_CHECK(ifs.read(reinterpret_cast<char*>(out), static_cast<std::streamsize>(n)),
             "failed to read PGM pixel payload");
    return VJ_OK;
}

namespace {

// Size of a raw sample file in whole patches (src/main.cu:122-162).
int samples_open(const char* path, int px, FILE** f, long* bytes) {
    VJ_CHECK(path && px > 0, "vj_samples: bad arguments");
    *f = std::fopen(path, "rb");
    VJ_CHECK(*f, std::string("Cannot open: ") + path);
    long sz = -1;
    if (std::fseek(*f, 0, SEEK_END) == 0) sz = std::ftell(*f);
    std::rewind(*f);
    if (sz <= 0 || sz % px != 0) {
        std::fclose(*f);
        *f = nullptr;
        VJ_CHECK(sz > 0, std::string("Empty file: ") + path);
        VJ_CHECK(false, std::string("File size is not divisible by samplePixels: ") + path);
    }
    *bytes = sz;
    return VJ_OK;
}

// Regular *.pgm files under dir, depth first (order fixed by the caller).
void pgm_walk(const std::string& dir, std::vector<std::string>& out) {
    DIR* d = opendir(dir.c_str());
    if (!d) return;
    while (const dirent* ent = readdir(d)) {
        const std::string name = ent->d_name;
        if (name == "." || name == "..") continue;
        const std::string path = dir + "/" + name;
        struct stat st;
        if (stat(path.c_str(), &st) != 0) continue;
        if (S_ISDIR(st.st_mode)) {
            pgm_walk(path, out);
        } else if (S_ISREG(st.st_mode) && name.size() > 4) {
            std::string ext = name.substr(name.size() - 4);
            for (char& c : ext) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
            if (ext == ".pgm") out.push_back(path);
        }
    }
    closedir(d);
}

}  // namespace

extern "C" int vj_samples_count(const char* path, int samplePixels, int* count) {
    VJ_CHECK(count, "vj_samples_count: bad arguments");
    FILE* f = nullptr;
    long sz = 0;
    VJ_TRY(samples_open(path, samplePixels, &f, &sz));
    std::fclose(f);
    *count = static_cast<int>(sz / samplePixels);
    return VJ_OK;
}

extern "C" int vj_samples_load(const char* path, int samplePixels, uint8_t* hOut, size_t cap, int* count) {
    VJ_CHECK(hOut && count, "vj_samples_load: bad arguments");
    FILE* f = nullptr;
    long sz = 0;
    VJ_TRY(samples_open(path, samplePixels, &f, &sz));
    const size_t n = static_cast<size_t>(sz);
    const bool fits = cap >= n;
    const bool ok = fits && std::fread(hOut, 1, n, f) == n;
    std::fclose(f);
    VJ_CHECK(fits, "vj_samples_load: output buffer too small");
    VJ_CHECK(ok, std::string("Failed to read: ") + path);
    *count = static_cast<int>(sz / samplePixels);
    return VJ_OK;
}

extern "C" int vj_pgm_list(const char* dir, char* buf, size_t cap, size_t* bytes, int* count) {
    VJ_CHECK(dir && bytes && count, "vj_pgm_list: bad arguments");
    std::vector<std::string> paths;
    std::string root = dir;
    while (root.size() > 1 && root.back() == '/') root.pop_back();
    if (!root.empty()) pgm_walk(root, paths);
    std::sort(paths.begin(), paths.end());
    size_t need = 0;
    for (const std::string& q : paths) need += q.size() + 1;
    *bytes = need;
    *count = static_cast<int>(paths.size());
    if (buf && cap >= need) {
        size_t at = 0;
        for (const std::string& q : paths) {
            std::memcpy(buf + at, q.c_str(), q.size() + 1);
            at += q.size() + 1;
        }
    }
    return VJ_OK;
}

// Crop and resample of a mined box, src/main.cu:245-284: source position
// (x + (o + 0.5) * (w / outW) - 0.5) per axis in float, floor clamped to the
// image (the fractional weight is taken from the clamped tap, as there),
// float lerp across then down, round half up, clamp to 0..255.
extern "C" int vj_crop_resize(const uint8_t* gray, int w, int h, int stride, const vj_detection* box, int outW,
                              int outH, uint8_t* out) {
    VJ_CHECK(gray && box && out && w > 0 && h > 0 && stride >= w && outW > 0 && outH > 0,
             "vj_crop_resize: bad arguments");
    const float sx = static_cast<float>(std::max(1, box->w));
    const float sy = static_cast<float>(std::max(1, box->h));
    for (int oy = 0; oy < outH; ++oy) {
        for (int ox = 0; ox < outW; ++ox) {
            const float fx = static_cast<float>(box->x) + (static_cast<float>(ox) + 0.5f) * (sx / outW) - 0.5f;
            const float fy = static_cast<float>(box->y) + (static_cast<float>(oy) + 0.5f) * (sy / outH) - 0.5f;
            const int x0 = std::clamp(static_cast<int>(std::floor(fx)), 0, w - 1);
            const int y0 = std::clamp(static_cast<int>(std::floor(fy)), 0, h - 1);
            const int x1 = std::min(x0 + 1, w - 1), y1 = std::min(y0 + 1, h - 1);
            const float ax = fx - static_cast<float>(x0), ay = fy - static_cast<float>(y0);
            const uint8_t* r0 = gray + static_cast<size_t>(y0) * stride;
            const uint8_t* r1 = gray + static_cast<size_t>(y1) * stride;
            const float p00 = r0[x0], p01 = r0[x1], p10 = r1[x0], p11 = r1[x1];
            const float top = p00 + (p01 - p00) * ax;
            const float bot = p10 + (p11 - p10) * ax;
            const float v = top + (bot - top) * ay;
            out[static_cast<size_t>(oy) * outW + ox] = static_cast<uint8_t>(std::clamp(static_cast<int>(v + 0.5f), 0, 255));
        }
    }
    return VJ_OK;
}

extern "C" int vj_ppm_save_gray_boxes(const char* path, const uint8_t* gray, int w, int h,
                                      const vj_detection* boxes, int n) {
    VJ_CHECK(path && gray && w > 0 && h > 0 && n >= 0, "vj_ppm_save_gray_boxes: bad arguments");
    std::vector<uint8_t> rgb(static_cast<size_t>(w) * h * 3);
    for (size_t i = 0; i < static_cast<size_t>(w) * h; ++i) rgb[3 * i] = rgb[3 * i + 1] = rgb[3 * i + 2] = gray[i];
    auto put = [&](int x, int y) {
        uint8_t* p = &rgb[(static_cast<size_t>(y) * w + x) * 3];
        p[0] = p[1] = p[2] = 255;
    };
    // 2-pixel white frame per box, src/main.cu:373-410.
    for (int b = 0; b < n; ++b) {
        const vj_detection& d = boxes[b];
        const int x0 = std::max(0, d.x), y0 = std::max(0, d.y);
        const int x1 = std::min(w - 1, d.x + d.w - 1), y1 = std::min(h - 1, d.y + d.h - 1);
        if (x0 > x1 || y0 > y1) continue;
        for (int t = 0; t < 2; ++t) {
            const int yt = std::min(h - 1, y0 + t), yb = std::max(0, y1 - t);
            for (int x = x0; x <= x1; ++x) {
                put(x, yt);
                put(x, yb);
            }
            const int xl = std::min(w - 1, x0 + t), xr = std::max(0, x1 - t);
            for (int y = y0; y <= y1; ++y) {
                put(xl, y);
                put(xr, y);
            }
        }
    }
    std::ofstream ofs(path, std::ios::binary);
    VJ_CHECK(ofs, std::string("cannot write PPM: ") + path);
    ofs << "P6\n" << w << " " << h << "\n255\n";
    ofs.write(reinterpret_cast<const char*>(rgb.data()), static_cast<std::streamsize>(rgb.size()));
    VJ_CHECK(static_cast<bool>(ofs), "short write of PPM");
    return VJ_OK;
}
