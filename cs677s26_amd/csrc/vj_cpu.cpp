// vj_cpu.cpp -- host legs of the compare verbs (cuda_hello detect_compare /
// train_compare, src/main.cu:1032-1368).
//
// The reference's compare modes run the same work on the CPU and on the GPU
// and print both results with their times.  These are the CPU halves, single
// threaded as the reference's are.  No GPU entry point calls them: every
// detect and training call of the library runs its HIP kernels.
//
//   vj_cpu_detect_raw  -- detectMultiScaleCpu (src/main.cu:684-827): scale
//                         schedule, cumulative bilinear pyramid, int32
//                         integral, exact window normalisation, stage loop in
//                         stump order, raw boxes in (level, y, x) order with
//                         the early return at maxDetections.
//   vj_cpu_best_weak   -- one best-stump search over features [0, count):
//                         responses, stable sort, threshold sweep and
//                         arg-min, with the GPU path's semantics (exact
//                         fixed-point sums, ties to the smaller split then
//                         parity +1 within a feature, float errors with the
//                         smaller index across features), so that both
//                         halves of train_compare print the same stump.  The
//                         reference's own CPU leg (runCpuWeakSearch,
//                         src/main.cu:587-682) sums in float and breaks exact
//                         ties the other way; the library's trainer follows
//                         the reference GPU comparator
//                         (src/boosting_core.cu:59-99).
#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <numeric>
#include <vector>

#include "vj_common.hpp"

namespace {

// resizeGrayBilinearHost, src/main.cu:545-585: pixel-centre mapping, clamped
// taps, bilinear with the border extrapolation of a negative fraction.
void resize_bilinear(const uint8_t* src, int sw, int sh, int sstride, uint8_t* dst, int dw, int dh) {
    const float kx = static_cast<float>(sw) / static_cast<float>(dw);
    const float ky = static_cast<float>(sh) / static_cast<float>(dh);
    for (int y = 0; y < dh; ++y) {
        const float fy = (static_cast<float>(y) + 0.5f) * ky - 0.5f;
        const int y0 = std::clamp(static_cast<int>(std::floor(fy)), 0, sh - 1);
        const int y1 = std::min(y0 + 1, sh - 1);
        const float ay = fy - static_cast<float>(y0);
        const uint8_t* r0 = src + static_cast<size_t>(y0) * sstride;
        const uint8_t* r1 = src + static_cast<size_t>(y1) * sstride;
        for (int x = 0; x < dw; ++x) {
            const float fx = (static_cast<float>(x) + 0.5f) * kx - 0.5f;
            const int x0 = std::clamp(static_cast<int>(std::floor(fx)), 0, sw - 1);
            const int x1 = std::min(x0 + 1, sw - 1);
            const float ax = fx - static_cast<float>(x0);
            const float a = r0[x0], b = r0[x1], c = r1[x0], d = r1[x1];
            const float top = a + (b - a) * ax;
            const float bot = c + (d - c) * ax;
            const float v = top + (bot - top) * ay;
            dst[static_cast<size_t>(y) * dw + x] = static_cast<uint8_t>(std::clamp(static_cast<int>(v + 0.5f), 0, 255));
        }
    }
}

// Integral and square integral with a zero first row and column, both mod
// 2^32 (the reference keeps int32; the sum of a 4K frame stays below 2^31,
// and the square sum's window deltas are read mod 2^32).
void integral_u32(const uint8_t* img, int w, int h, int stride, std::vector<uint32_t>& s, std::vector<uint32_t>& q) {
    const int iw = w + 1;
    s.assign(static_cast<size_t>(iw) * (h + 1), 0u);
    q.assign(s.size(), 0u);
    for (int y = 1; y <= h; ++y) {
        const uint8_t* row = img + static_cast<size_t>(y - 1) * stride;
        uint32_t rs = 0, rq = 0;
        uint32_t* so = s.data() + static_cast<size_t>(y) * iw;
        uint32_t* qo = q.data() + static_cast<size_t>(y) * iw;
        for (int x = 1; x <= w; ++x) {
            const uint32_t p = row[x - 1];
            rs += p;
            rq += p * p;
            so[x] = so[x - iw] + rs;
            qo[x] = qo[x - iw] + rq;
        }
    }
}

// A stump's feature with its corners turned into offsets from a window's
// top-left integral element (decodeModelPoint, src/main.cu:498-502).
struct CpuFeature {
    int n;
    int off[3][4];
    float w[3];
};

}  // namespace

extern "C" int vj_cpu_detect_raw(const vj_feature* feats, int nFeatures, const vj_stump* stumps, int nStumps,
                                 const vj_stage* stages, int nStages, int winW, int winH, const uint8_t* img,
                                 int w, int h, int stride, float scaleFactor, int minObjectSize,
                                 int maxDetections, int maxScales, vj_detection* out, int cap, int* count) {
    try {
        VJ_CHECK(feats && stumps && stages && img && count, "vj_cpu_detect_raw: null argument");
        VJ_CHECK(nFeatures > 0 && nStumps > 0 && nStages > 0, "vj_cpu_detect_raw: empty cascade");
        VJ_CHECK(winW >= 3 && winH >= 3 && w > 0 && h > 0 && stride >= w, "vj_cpu_detect_raw: bad image or window");
        VJ_CHECK(scaleFactor > 1.0f && maxDetections > 0 && cap >= 0 && (cap == 0 || out),
                 "vj_cpu_detect_raw: bad parameters");
        *count = 0;
        const int mW = winW + 1, mArea = (winW + 1) * (winH + 1);
        for (int i = 0; i < nStages; ++i)
            VJ_CHECK(stages[i].first >= 0 && stages[i].ntrees >= 0 &&
                         static_cast<int64_t>(stages[i].first) + stages[i].ntrees <= nStumps,
                     "vj_cpu_detect_raw: stage range outside the stumps");
        for (int i = 0; i < nStumps; ++i)
            VJ_CHECK(stumps[i].feature >= 0 && stumps[i].feature < nFeatures,
                     "vj_cpu_detect_raw: stump feature index out of range");
        auto corners_ok = [&](const vj_feature& f) {
            if (f.rectCount != 2 && f.rectCount != 3) return false;
            for (int r = 0; r < f.rectCount; ++r)
                for (int p : {f.r[r].p0, f.r[r].p1, f.r[r].p2, f.r[r].p3})
                    if (p < 0 || p >= mArea) return false;
            return true;
        };
        for (int i = 0; i < nStumps; ++i)
            VJ_CHECK(corners_ok(feats[stumps[i].feature]), "vj_cpu_detect_raw: feature corner outside the window");

        std::vector<uint8_t> prev, cur;
        const uint8_t* prevPtr = img;
        int prevW = w, prevH = h, prevStride = stride;
        std::vector<uint32_t> S, Q;
        std::vector<CpuFeature> cf(static_cast<size_t>(nStumps));
        int produced = 0;
        float scale = 1.0f;
        for (int scaleIdx = 0;; ++scaleIdx, scale *= scaleFactor) {
            if (maxScales > 0 && scaleIdx >= maxScales) break;
            const int lw = static_cast<int>(std::round(static_cast<float>(w) / scale));
            const int lh = static_cast<int>(std::round(static_cast<float>(h) / scale));
            const int ow = static_cast<int>(std::round(static_cast<float>(winW) * scale));
            const int oh = static_cast<int>(std::round(static_cast<float>(winH) * scale));
            if (lw < winW || lh < winH) break;
            if (ow < minObjectSize || oh < minObjectSize) continue;
            const uint8_t* lp = img;
            int ls = stride;
            if (scaleIdx != 0) {
                // each level is resized from the previous processed one
                cur.resize(static_cast<size_t>(lw) * lh);
                resize_bilinear(prevPtr, prevW, prevH, prevStride, cur.data(), lw, lh);
                prev.swap(cur);
                prevPtr = prev.data();
                prevW = lw;
                prevH = lh;
                prevStride = lw;
                lp = prev.data();
                ls = lw;
            }
            integral_u32(lp, lw, lh, ls, S, Q);
            const int iw = lw + 1;
            for (int i = 0; i < nStumps; ++i) {
                const vj_feature& f = feats[stumps[i].feature];
                CpuFeature& c = cf[static_cast<size_t>(i)];
                c.n = f.rectCount;
                for (int r = 0; r < f.rectCount; ++r) {
                    const int32_t p[4] = {f.r[r].p0, f.r[r].p1, f.r[r].p2, f.r[r].p3};
                    for (int k = 0; k < 4; ++k) c.off[r][k] = (p[k] / mW) * iw + (p[k] % mW);
                    c.w[r] = f.r[r].w;
                }
            }
            const int step = (scale >= 3.0f) ? 3 : ((scale >= 2.0f) ? 2 : 1);
            const int narea = (winW - 2) * (winH - 2);
            const int n0 = 1 + iw, n1 = (winW - 1) + iw, n2 = 1 + iw * (winH - 1), n3 = (winW - 1) + iw * (winH - 1);
            for (int y = 0; y + winH <= lh; y += step) {
                for (int x = 0; x + winW <= lw; x += step) {
                    const size_t b = static_cast<size_t>(y) * iw + x;
                    const int32_t ns = static_cast<int32_t>(S[b + n0] - S[b + n1] - S[b + n2] + S[b + n3]);
                    const uint32_t nq = Q[b + n0] - Q[b + n1] - Q[b + n2] + Q[b + n3];
                    const double sd = static_cast<double>(ns);
                    const double nf2 = static_cast<double>(narea) * static_cast<double>(nq) - sd * sd;
                    if (!(nf2 > 0.0)) continue;
                    const float inv = static_cast<float>(1.0 / std::sqrt(nf2));
                    if (!(static_cast<double>(narea) * static_cast<double>(inv) < 1e-1)) continue;
                    const uint32_t* sw = S.data() + b;
                    bool pass = true;
                    float last = 0.0f;
                    for (int st = 0; st < nStages && pass; ++st) {
                        float acc = 0.0f;
                        const int e = stages[st].first + stages[st].ntrees;
                        for (int i = stages[st].first; i < e; ++i) {
                            const CpuFeature& c = cf[static_cast<size_t>(i)];
                            float raw = 0.0f;
                            for (int r = 0; r < c.n; ++r) {
                                const float v = ((static_cast<float>(static_cast<int32_t>(sw[c.off[r][0]])) -
                                                  static_cast<float>(static_cast<int32_t>(sw[c.off[r][1]]))) -
                                                 static_cast<float>(static_cast<int32_t>(sw[c.off[r][2]]))) +
                                                static_cast<float>(static_cast<int32_t>(sw[c.off[r][3]]));
                                raw = (r == 0) ? v * c.w[0] : raw + v * c.w[r];
                            }
                            acc += (raw * inv < stumps[i].theta) ? stumps[i].left : stumps[i].right;
                        }
                        if (acc < stages[st].threshold) pass = false;
                        else last = acc;
                    }
                    if (!pass) continue;
                    if (produced < cap) {
                        vj_detection& d = out[produced];
                        d.x = static_cast<int>(static_cast<float>(x) * scale + 0.5f);
                        d.y = static_cast<int>(static_cast<float>(y) * scale + 0.5f);
                        d.w = static_cast<int>(static_cast<float>(winW) * scale + 0.5f);
                        d.h = static_cast<int>(static_cast<float>(winH) * scale + 0.5f);
                        d.scaleIdx = scaleIdx;
                        d.score = last;
                    }
                    if (++produced >= maxDetections) {
                        *count = produced;
                        return VJ_OK;
                    }
                }
            }
        }
        *count = produced;
        return VJ_OK;
    } catch (const std::bad_alloc&) {
        vj::set_error("vj_cpu_detect_raw: out of host memory");
        return VJ_OUT_OF_MEMORY;
    }
}

extern "C" int vj_cpu_best_weak(const uint8_t* patches, int n, int winW, int winH, const int8_t* labels,
                                const uint8_t* active, const float* weights, const vj_feature* feats, int count,
                                vj_tile_best* out) {
    try {
        VJ_CHECK(patches && labels && weights && feats && out, "vj_cpu_best_weak: null argument");
        VJ_CHECK(n > 1 && n <= 65535 && count > 0 && winW >= 3 && winH >= 3, "vj_cpu_best_weak: bad sizes");
        const int iw = winW + 1, ih = winH + 1, area = iw * ih;
        for (int i = 0; i < count; ++i) {
            const vj_feature& f = feats[i];
            VJ_CHECK(f.rectCount == 2 || f.rectCount == 3, "vj_cpu_best_weak: feature rectCount must be 2 or 3");
            for (int r = 0; r < f.rectCount; ++r)
                for (int p : {f.r[r].p0, f.r[r].p1, f.r[r].p2, f.r[r].p3})
                    VJ_CHECK(p >= 0 && p < area, "vj_cpu_best_weak: feature corner outside the window integral");
        }
        // exact weights: each signed weight once to an int64 count of 2^-62
        double tot = 0.0;
        std::vector<int64_t> q(static_cast<size_t>(n));
        int64_t totPos = 0, totNeg = 0;
        for (int s = 0; s < n; ++s) {
            VJ_CHECK(std::isfinite(weights[s]), "vj_cpu_best_weak: non-finite weight");
            tot += std::fabs(static_cast<double>(weights[s]));
            const float sw = (active && active[s] == 0) ? 0.0f : (labels[s] > 0 ? weights[s] : -weights[s]);
            const int64_t m = static_cast<int64_t>(std::ldexp(std::fabs(static_cast<double>(sw)), 62));
            q[static_cast<size_t>(s)] = (sw < 0.0f) ? -m : m;
            if (sw < 0.0f) totNeg += m;
            else totPos += m;
        }
        VJ_CHECK(tot <= 1.5, "vj_cpu_best_weak: weights must be normalised (sum of |w| <= 1.5)");

        // per-sample integrals (row-major) and normalisers of the inset
        // (winW-2) x (winH-2) rectangle (src/main.cu:457-472)
        std::vector<uint32_t> S, Q, all(static_cast<size_t>(n) * area);
        std::vector<float> inv(static_cast<size_t>(n));
        const int a0 = 1 + iw, a1 = (iw - 2) + iw, a2 = 1 + iw * (ih - 2), a3 = (iw - 2) + iw * (ih - 2);
        const double narea = static_cast<double>(iw - 3) * (ih - 3);
        for (int s = 0; s < n; ++s) {
            integral_u32(patches + static_cast<size_t>(s) * winW * winH, winW, winH, winW, S, Q);
            std::copy(S.begin(), S.end(), all.begin() + static_cast<size_t>(s) * area);
            const double sm = static_cast<double>(static_cast<int32_t>(S[a0] - S[a1] - S[a2] + S[a3]));
            const double sq = static_cast<double>(static_cast<int32_t>(Q[a0] - Q[a1] - Q[a2] + Q[a3]));
            const double nf2 = narea * sq - sm * sm;
            inv[static_cast<size_t>(s)] = (nf2 > 1e-12) ? static_cast<float>(1.0 / std::sqrt(nf2)) : 0.0f;
        }

        std::vector<float> v(static_cast<size_t>(n));
        std::vector<uint64_t> key(static_cast<size_t>(n));
        vj_tile_best best{FLT_MAX, 0.0f, INT_MAX, +1, 0, 0, 0};
        for (int fi = 0; fi < count; ++fi) {
            const vj_feature& f = feats[fi];
            for (int s = 0; s < n; ++s) {
                const uint32_t* si = all.data() + static_cast<size_t>(s) * area;
                float sums[3];
                for (int r = 0; r < f.rectCount; ++r)
                    sums[r] = ((static_cast<float>(static_cast<int32_t>(si[f.r[r].p0])) -
                                static_cast<float>(static_cast<int32_t>(si[f.r[r].p1]))) -
                               static_cast<float>(static_cast<int32_t>(si[f.r[r].p2]))) +
                              static_cast<float>(static_cast<int32_t>(si[f.r[r].p3]));
                float resp = f.r[0].w * sums[0] + f.r[1].w * sums[1];
                if (f.rectCount == 3) resp += f.r[2].w * sums[2];
                resp *= inv[static_cast<size_t>(s)];
                v[static_cast<size_t>(s)] = resp;
                uint32_t bits;
                std::memcpy(&bits, &resp, 4);
                // order-preserving key: negative floats reversed below the positives
                bits = (bits & 0x80000000u) ? ~bits : (bits | 0x80000000u);
                key[static_cast<size_t>(s)] = (static_cast<uint64_t>(bits) << 32) | static_cast<uint32_t>(s);
            }
            std::sort(key.begin(), key.end());
            int64_t D = 0, bE = INT64_MAX;
            int bK = -1, bP = +1;
            for (int k = 0; k + 1 < n; ++k) {
                const uint32_t s0 = static_cast<uint32_t>(key[static_cast<size_t>(k)]);
                const uint32_t s1 = static_cast<uint32_t>(key[static_cast<size_t>(k) + 1]);
                D += q[s0];
                if (!(v[s0] + 1e-7f < v[s1])) continue;
                const int64_t eP = totPos - D;  // parity -1
                const int64_t eN = totNeg + D;  // parity +1
                if (eP < bE) {
                    bE = eP;
                    bK = k;
                    bP = -1;
                }
                if (eN < bE || (eN == bE && bK == k)) {
                    bE = eN;
                    bK = k;
                    bP = +1;
                }
            }
            float err = FLT_MAX, theta = 0.0f;
            int par = +1;
            if (bK >= 0) {
                err = static_cast<float>(std::ldexp(static_cast<double>(bE), -62));
                theta = 0.5f * (v[static_cast<uint32_t>(key[static_cast<size_t>(bK)])] +
                                v[static_cast<uint32_t>(key[static_cast<size_t>(bK) + 1])]);
                par = bP;
            }
            if (err < best.err || (err == best.err && fi < best.feature)) {
                best.err = err;
                best.theta = theta;
                best.feature = fi;
                best.parity = static_cast<int8_t>(par);
            }
        }
        *out = best;
        return VJ_OK;
    } catch (const std::bad_alloc&) {
        vj::set_error("vj_cpu_best_weak: out of host memory");
        return VJ_OUT_OF_MEMORY;
    }
}
