// vj_engine.hip -- FaceVisionEngine on gfx950: pyramid, integral images,
// cascade evaluation and order-preserving compaction of detections.
//
// Reference: src/face_vision_engine.cu (FaceVisionEngine) and the CPU path it
// must equal, src/main.cu:428-827.  Arithmetic follows the CPU path exactly:
// exact u32/u64 integrals, int window sums, f64 normaliser, f32 Haar sums from
// absolute integral values, f32 stage sums in stump order, CPU box rounding.
// Build with -ffp-contract=off (no fused multiply-add anywhere).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <type_traits>
#include <vector>

#include "vj_common.hpp"
#include "vj_kernels.hpp"
#include "vj_integral_dev.hpp"
#include "vj_pair_plan.hpp"

namespace vj {

// ===========================================================================
// Kernels
// ===========================================================================

// Vectors with element alignment only, for multi-word global stores at
// element-aligned addresses.
typedef uint32_t u32x4_a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef uint64_t u64x2_a8 __attribute__((ext_vector_type(2), aligned(8)));

// Pixel-centre bilinear resize, the CPU path's float sequence
// (src/main.cu:565-581): fx = (x + 0.5) * sx - 0.5, clamped floor, lerp in x
// then y, (u8) clamp((int)(v + 0.5), 0, 255).  The taps of an output
// coordinate c along one axis (source length n): i0 and i1 = min(i0 + 1,
// n - 1) with the weight of i1.
struct Taps {
    int i0, i1;
    float a;
};
__host__ __device__ inline Taps bilinear_taps(int c, float scale, int n) {
    const float f = (static_cast<float>(c) + 0.5f) * scale - 0.5f;
    int i0 = static_cast<int>(floorf(f));
    i0 = i0 < 0 ? 0 : (i0 > n - 1 ? n - 1 : i0);
    return Taps{i0, i0 + 1 < n - 1 ? i0 + 1 : n - 1, f - static_cast<float>(i0)};
}
__device__ __forceinline__ uint8_t bilinear_pixel(float p00, float p01, float p10, float p11, float ax, float ay) {
    const float top = p00 + (p01 - p00) * ax;
    const float bot = p10 + (p11 - p10) * ax;
    const float val = top + (bot - top) * ay;
    const int v = static_cast<int>(val + 0.5f);
    return static_cast<uint8_t>(v < 0 ? 0 : (v > 255 ? 255 : v));
}

// One resize depth in one launch: every level of that depth (one per image)
// from its source level, a job per level.  Used for the large levels, where a
// launch per depth costs little next to the work and no pixel is computed
// twice.  A workgroup makes `rows` output rows (up to kResizeRows; fewer when
// the depth has too few rows to fill the GPU otherwise) of a segment of up to
// 1024 columns.  It first copies the source rectangle those pixels tap into LDS
// with word loads (byte gathers from global memory would each hold the
// texture addresser for a whole instruction), then each thread makes four
// adjacent pixels of every row from LDS, keeping their column taps, and
// stores them as one word.  A rectangle larger than kResizeLds (source levels
// far larger than the output) is read from global memory directly.
struct LevelJob {
    int32_t level;    // destination level
    int32_t itemOff;  // first work item of this job
};
constexpr int kResizeSeg = 1024;
constexpr int kResizeRows = 8;
constexpr int kResizeLds = 24 * 1024;

__global__ __launch_bounds__(256) void resize_level_kernel(const LevelDesc* __restrict__ L,
                                                           const LevelJob* __restrict__ jobs, int nJobs,
                                                           const float* __restrict__ rs,
                                                           const int32_t* __restrict__ src, int rows) {
    __shared__ __align__(16) uint8_t rect[kResizeLds];
    const int item = blockIdx.x;
    int lo = 0, hi = nJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].itemOff <= item) lo = mid;
        else hi = mid - 1;
    }
    const int li = jobs[lo].level;
    const LevelDesc dst = L[li];
    const LevelDesc sl = L[src[li]];
    const int segs = (dst.w + kResizeSeg - 1) / kResizeSeg;
    const int local = item - jobs[lo].itemOff;
    const int band = local / segs;
    const int xs = (local - band * segs) * kResizeSeg;
    const int xe = min(xs + kResizeSeg, dst.w);
    const int y0 = band * rows, y1 = min(y0 + rows, dst.h);
    const float sx = rs[2 * li], sy = rs[2 * li + 1];
    // the source rectangle: rows [sr0, sr1], columns [c0, c1] (taps are
    // monotone), each LDS row starting at the word boundary below column c0
    // of its source row (the pitch keeps every row's words aligned)
    const int c0 = bilinear_taps(xs, sx, sl.w).i0, c1 = bilinear_taps(xe - 1, sx, sl.w).i1;
    const int sr0 = bilinear_taps(y0, sy, sl.h).i0, sr1 = bilinear_taps(y1 - 1, sy, sl.h).i1;
    const gpix_t pix = global_pix(sl.pix);
    const int pitch = (c1 - c0 + 1 + 3 + 3) & ~3;  // room for up to 3 bytes before c0
    const bool staged = static_cast<int64_t>(pitch) * (sr1 - sr0 + 1) <= kResizeLds;
    if (staged) {
        const int tid = static_cast<int>(threadIdx.x), lane = tid & (kWave - 1), wave = tid / kWave;
        for (int r = wave; r <= sr1 - sr0; r += 256 / kWave) {
            const gpix_t row = pix + static_cast<size_t>(sr0 + r) * sl.stride;
            const uintptr_t lo8 = reinterpret_cast<uintptr_t>(row + c0), hi8 = reinterpret_cast<uintptr_t>(row + c1) + 1;
            const uintptr_t alo = lo8 & ~uintptr_t{3};
            const uintptr_t a0 = (lo8 + 3) & ~uintptr_t{3}, a1 = hi8 & ~uintptr_t{3};
            uint8_t* dstRow = rect + r * pitch;
            // whole words inside [c0, c1], then the up to 3 + 3 bytes around them
            for (uintptr_t a = a0 + 4 * static_cast<uintptr_t>(lane); a < a1; a += 4 * kWave)
                *reinterpret_cast<uint32_t*>(dstRow + (a - alo)) =
                    *(const uint32_t __attribute__((address_space(1)))*)(a);
            const int head = static_cast<int>(a0 - lo8), tail = static_cast<int>(hi8 - a1);
            if (a0 > a1) {  // no whole word: every byte of the span
                if (lane <= c1 - c0) dstRow[lo8 - alo + lane] = row[c0 + lane];
            } else if (lane < head) {
                dstRow[lo8 - alo + lane] = row[c0 + lane];
            } else if (lane >= 8 && lane < 8 + tail) {
                dstRow[a1 - alo + (lane - 8)] = row[c1 + 1 - tail + (lane - 8)];
            }
        }
        __syncthreads();
    }
    const int x0 = xs + static_cast<int>(threadIdx.x) * 4;
    if (x0 >= xe) return;
    const int nx = min(4, xe - x0);
    const int shift = static_cast<int>(reinterpret_cast<uintptr_t>(pix + c0) & 3u);  // LDS offset of column c0
    Taps tx[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) tx[k] = bilinear_taps(min(x0 + k, xe - 1), sx, sl.w);
    for (int y = y0; y < y1; ++y) {
        const Taps ty = bilinear_taps(y, sy, sl.h);
        float p[4][4];
        if (staged) {
            // LDS row of source row s: (s - sr0) * pitch, shifted by the
            // row's own misalignment (the stride may not be a multiple of 4)
            auto at = [&](int sr, int c) {
                const uintptr_t rb = reinterpret_cast<uintptr_t>(pix + static_cast<size_t>(sr) * sl.stride + c0);
                return static_cast<float>(rect[(sr - sr0) * pitch + static_cast<int>(rb & 3u) + (c - c0)]);
            };
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                p[k][0] = at(ty.i0, tx[k].i0);
                p[k][1] = at(ty.i0, tx[k].i1);
                p[k][2] = at(ty.i1, tx[k].i0);
                p[k][3] = at(ty.i1, tx[k].i1);
            }
        } else {
            const gpix_t r0 = pix + static_cast<size_t>(ty.i0) * sl.stride;
            const gpix_t r1 = pix + static_cast<size_t>(ty.i1) * sl.stride;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                p[k][0] = static_cast<float>(r0[tx[k].i0]);
                p[k][1] = static_cast<float>(r0[tx[k].i1]);
                p[k][2] = static_cast<float>(r1[tx[k].i0]);
                p[k][3] = static_cast<float>(r1[tx[k].i1]);
            }
        }
        (void)shift;
        uint32_t packed = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k)
            packed |= static_cast<uint32_t>(bilinear_pixel(p[k][0], p[k][1], p[k][2], p[k][3], tx[k].a, ty.a))
                      << (8 * k);
        const gpix_out_t out = global_pix_out(dst.pix) + static_cast<size_t>(y) * dst.stride;
        // pyramid rows start 64-byte aligned, so four pixels from x0 (a
        // multiple of four) form one aligned word
        if (nx == 4) {
            *(uint32_t __attribute__((address_space(1)))*)(out + x0) = packed;
        } else {
            for (int k = 0; k < nx; ++k) out[x0 + k] = static_cast<uint8_t>(packed >> (8 * k));
        }
    }
}

// The pyramid in a few launches: a resize group is up to kResizeDepth
// consecutive levels of one image's resize chain (each level resized from
// the one before, as the CPU path does), computed from the level before the
// group.  The group's last level is cut into a grid of gx x gy tiles, and
// every level of the group into the same grid of proportional pieces
// [i * W / gx, (i + 1) * W / gx), so each workgroup owns one piece of every
// level.  It computes, level by level into LDS, the region of each level
// that its own pieces need (its piece plus the taps of the next level's
// region), so no workgroup waits for another; it writes only its own pieces.
// Every pixel is the same function of the same source pixels however often it
// is recomputed, so the levels equal a level-by-level resize.
constexpr int kResizeDepth = 6;
constexpr int kResizeCap = 9216;  // pixels of one level's region in LDS
constexpr int kResizeThreads = 256;

struct ResizeJob {
    int32_t src;                  // level the group resizes from
    int32_t lv[kResizeDepth];     // the group's levels in chain order
    int32_t n;                    // levels in the group
    int32_t gx, gy;               // tile grid
    int32_t itemOff;              // first work item (tile) of this job
};

// Half-open pixel ranges of one axis of every level of a group for tile i:
// own[j] the tile's piece of level j (1-based; 0 is the source), need[j] what
// the tile computes of level j, need[0] the source range it reads.
struct AxisRanges {
    int own0[kResizeDepth + 1], own1[kResizeDepth + 1];
    int need0[kResizeDepth + 1], need1[kResizeDepth + 1];
};
// len[j] is level j's length on the axis, sc[j] the scale level j is resized
// with (source over destination length), g the tile count, i the tile.
// (Loops run over the fixed depth so that the arrays stay in registers.)
__host__ __device__ inline void axis_ranges(int n, const int* len, const float* sc, int g, int i, AxisRanges& r) {
#pragma unroll
    for (int j = 1; j <= kResizeDepth; ++j) {
        r.own0[j] = j <= n ? static_cast<int>(static_cast<int64_t>(i) * len[j] / g) : 0;
        r.own1[j] = j <= n ? static_cast<int>(static_cast<int64_t>(i + 1) * len[j] / g) : 0;
        r.need0[j] = r.own0[j];
        r.need1[j] = r.own1[j];
    }
#pragma unroll
    for (int j = kResizeDepth; j >= 1; --j) {
        if (j > n) continue;
        // taps of need[j] in level j - 1 (monotone in the coordinate)
        const int lo = bilinear_taps(r.need0[j], sc[j], len[j - 1]).i0;
        const int hi = bilinear_taps(r.need1[j] - 1, sc[j], len[j - 1]).i1 + 1;
        r.need0[j - 1] = (j - 1 >= 1 && r.own0[j - 1] < lo) ? r.own0[j - 1] : lo;
        r.need1[j - 1] = (j - 1 >= 1 && r.own1[j - 1] > hi) ? r.own1[j - 1] : hi;
    }
}

__global__ __launch_bounds__(kResizeThreads) void resize_group_kernel(const LevelDesc* __restrict__ L,
                                                                      const ResizeJob* __restrict__ jobs,
                                                                      int nJobs, const float* __restrict__ rs) {
    __shared__ uint8_t buf[2][kResizeCap];
    const int item = blockIdx.x;
    int lo = 0, hi = nJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].itemOff <= item) lo = mid;
        else hi = mid - 1;
    }
    const ResizeJob jb = jobs[lo];
    const int local = item - jb.itemOff;
    const int ty = local / jb.gx, tx = local - ty * jb.gx;
    const int n = jb.n;
    int W[kResizeDepth + 1], H[kResizeDepth + 1];
    float sx[kResizeDepth + 1], sy[kResizeDepth + 1];
    const LevelDesc s0 = L[jb.src];
    W[0] = s0.w;
    H[0] = s0.h;
    sx[0] = sy[0] = 1.0f;
#pragma unroll
    for (int j = 1; j <= kResizeDepth; ++j) {
        const int li = jb.lv[j <= n ? j - 1 : 0];
        W[j] = L[li].w;
        H[j] = L[li].h;
        sx[j] = rs[2 * li];
        sy[j] = rs[2 * li + 1];
    }
    AxisRanges rx, ry;
    axis_ranges(n, W, sx, jb.gx, tx, rx);
    axis_ranges(n, H, sy, jb.gy, ty, ry);
    const int lane = lane_id(), wave = static_cast<int>(threadIdx.x) / kWave;
    constexpr int kWaves = kResizeThreads / kWave;
#pragma unroll
    for (int j = 1; j <= kResizeDepth; ++j) {
        if (j > n) break;
        const int x0 = rx.need0[j], nx = rx.need1[j] - x0;
        const int y0 = ry.need0[j], ny = ry.need1[j] - y0;
        if (nx * ny > kResizeCap) return;  // the plan sizes tiles so this never happens
        const LevelDesc d = L[jb.lv[j - 1]];
        const gpix_out_t out = global_pix_out(d.pix);
        uint8_t* cur = buf[j & 1];
        const uint8_t* prev = buf[(j - 1) & 1];
        const int px0 = rx.need0[j - 1], pitch = rx.need1[j - 1] - px0, py0 = ry.need0[j - 1];
        for (int yy = wave; yy < ny; yy += kWaves) {
            const int y = y0 + yy;
            const Taps ty2 = bilinear_taps(y, sy[j], H[j - 1]);
            const bool ownRow = y >= ry.own0[j] && y < ry.own1[j];
            for (int xx = lane; xx < nx; xx += kWave) {
                const int x = x0 + xx;
                const Taps tx2 = bilinear_taps(x, sx[j], W[j - 1]);
                float p00, p01, p10, p11;
                if (j == 1) {
                    const gpix_t r0 = global_pix(s0.pix) + static_cast<size_t>(ty2.i0) * s0.stride;
                    const gpix_t r1 = global_pix(s0.pix) + static_cast<size_t>(ty2.i1) * s0.stride;
                    p00 = static_cast<float>(r0[tx2.i0]);
                    p01 = static_cast<float>(r0[tx2.i1]);
                    p10 = static_cast<float>(r1[tx2.i0]);
                    p11 = static_cast<float>(r1[tx2.i1]);
                } else {
                    const uint8_t* r0 = prev + (ty2.i0 - py0) * pitch - px0;
                    const uint8_t* r1 = prev + (ty2.i1 - py0) * pitch - px0;
                    p00 = static_cast<float>(r0[tx2.i0]);
                    p01 = static_cast<float>(r0[tx2.i1]);
                    p10 = static_cast<float>(r1[tx2.i0]);
                    p11 = static_cast<float>(r1[tx2.i1]);
                }
                const uint8_t v = bilinear_pixel(p00, p01, p10, p11, tx2.a, ty2.a);
                if (j < n) cur[yy * nx + xx] = v;
                if (ownRow && x >= rx.own0[j] && x < rx.own1[j]) out[static_cast<size_t>(y) * d.stride + x] = v;
            }
        }
        __syncthreads();  // level j's region is complete before level j + 1 reads it
    }
}

// Pixels x..x+3 of a row packed little-endian into one word (zero beyond w).
__device__ __forceinline__ uint32_t load4_packed(gpix_t row, int x, int w) {
    if (x + 3 < w && ((reinterpret_cast<uintptr_t>(row + x) & 3u) == 0))
        return *(const uint32_t __attribute__((address_space(1)))*)(row + x);
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (x + k < w) v |= static_cast<uint32_t>(row[x + k]) << (8 * k);
    return v;
}

// Every row of a band, four pixels per lane, loaded before any is used.
__device__ __forceinline__ void load_band(const LevelDesc& lv, int x, int y0, int y1, uint32_t pv[kBandRows]) {
#pragma unroll
    for (int r = 0; r < kBandRows; ++r)
        pv[r] = (y0 + r < y1) ? load4_packed(global_pix(lv.pix) + static_cast<size_t>(y0 + r) * lv.stride, x, lv.w) : 0u;
}

// Integral phase 1, one wavefront per (level, band of 32 rows, chunk of 256
// columns): column sums over the band (pixels and squares) and each row's
// total over the chunk.
__global__ __launch_bounds__(64) void ii_colsum_kernel(const LevelDesc* __restrict__ L, int nItems,
                                                       const int32_t* __restrict__ itemOffs, int nLevels,
                                                       uint32_t* __restrict__ colS, uint32_t* __restrict__ colQ,
                                                       uint32_t* __restrict__ rowS, uint32_t* __restrict__ rowQ,
                                                       uint32_t* __restrict__ totS, uint64_t* __restrict__ totQ) {
    const int item = blockIdx.x;
    if (item >= nItems) return;
    const int li = find_segment(itemOffs, nLevels, item);
    const LevelDesc lv = L[li];
    const int local = item - lv.itemOff;
    const int b = local / lv.nChunks;
    const int c = local - b * lv.nChunks;
    const int lane = lane_id();
    const int x = c * kChunkCols + lane * 4;
    const int y0 = b * kBandRows;
    const int y1 = min(lv.h, y0 + kBandRows);
    uint32_t cs[4] = {0, 0, 0, 0}, cq[4] = {0, 0, 0, 0};
    uint32_t bandS = 0;
    uint64_t bandQ = 0;
    uint32_t pv[kBandRows];
    load_band(lv, x, y0, y1, pv);
#pragma unroll
    for (int r = 0; r < kBandRows; ++r) {
        const int y = y0 + r;
        if (y >= y1) break;
        const uint32_t p[4] = {pv[r] & 0xFFu, (pv[r] >> 8) & 0xFFu, (pv[r] >> 16) & 0xFFu, pv[r] >> 24};
        uint32_t rs = 0, rq = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            cs[k] += p[k];
            cq[k] += p[k] * p[k];
            rs += p[k];
            rq += p[k] * p[k];
        }
        rs = wave_sum_dpp(rs);
        rq = wave_sum_dpp(rq);
        bandS += rs;
        bandQ += rq;
        if (lane == 0) {
            rowS[lv.rowOff + static_cast<int64_t>(y) * lv.nChunks + c] = rs;
            rowQ[lv.rowOff + static_cast<int64_t>(y) * lv.nChunks + c] = rq;
        }
    }
    // the band's total over the chunk (turned into a prefix over bands by
    // the vertical scan)
    if (lane == 0) {
        totS[lv.totOff + static_cast<int64_t>(b) * lv.nChunks + c] = bandS;
        totQ[lv.totOff + static_cast<int64_t>(b) * lv.nChunks + c] = bandQ;
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        if (x + k < lv.w) {
            colS[lv.colOff + static_cast<int64_t>(b) * lv.w + x + k] = cs[k];
            colQ[lv.colOff + static_cast<int64_t>(b) * lv.w + x + k] = cq[k];
        }
    }
}

// Integral phase 2, one workgroup per (level, chunk of 256 columns): turns
// the band column sums into exclusive vertical prefixes (column sums above
// each band, in place), and the band totals of the chunk (from phase 1) into
// their exclusive prefix over the bands: the sum of all pixels above the band
// in the chunk's columns.
__global__ __launch_bounds__(256) void ii_vscan_kernel(const LevelDesc* __restrict__ L,
                                                       const int32_t* __restrict__ chunkOffs, int nLevels,
                                                       uint32_t* __restrict__ colS, uint32_t* __restrict__ colQ,
                                                       uint32_t* __restrict__ totS, uint64_t* __restrict__ totQ,
                                                       int32_t* __restrict__ zeroCtl, int nZero) {
    // the cascade's work-list counters (detect only), zeroed here to save a launch
    if (zeroCtl && blockIdx.x == 0)
        for (int i = static_cast<int>(threadIdx.x); i < nZero; i += blockDim.x) zeroCtl[i] = 0;
    const int item = blockIdx.x;
    const int li = find_segment(chunkOffs, nLevels, item);
    const LevelDesc lv = L[li];
    const int c = item - lv.chunkItemOff;
    const int x = c * kChunkCols + threadIdx.x;
    if (x < lv.w) {
        uint32_t s = 0, q = 0;
        constexpr int kU = 4;  // bands loaded ahead of their use
        int b0 = 0;
        for (; b0 + kU <= lv.nBands; b0 += kU) {
            uint32_t a[kU], aq[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = lv.colOff + static_cast<int64_t>(b0 + u) * lv.w + x;
                a[u] = colS[i];
                aq[u] = colQ[i];
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int64_t i = lv.colOff + static_cast<int64_t>(b0 + u) * lv.w + x;
                colS[i] = s;
                colQ[i] = q;
                s += a[u];
                q += aq[u];
            }
        }
        for (int b = b0; b < lv.nBands; ++b) {
            const int64_t i = lv.colOff + static_cast<int64_t>(b) * lv.w + x;
            const uint32_t a = colS[i], aq = colQ[i];
            colS[i] = s;
            colQ[i] = q;
            s += a;
            q += aq;
        }
    }
    if (threadIdx.x == 0) {
        uint32_t s = 0;
        uint64_t q = 0;
        for (int b = 0; b < lv.nBands; ++b) {
            const int64_t i = lv.totOff + static_cast<int64_t>(b) * lv.nChunks + c;
            const uint32_t a = totS[i];
            const uint64_t aq = totQ[i];
            totS[i] = s;
            totQ[i] = q;
            s += a;
            q += aq;
        }
    }
}

// Integral phase 3, one wavefront per (level, band, chunk): the integral at
// the band's top edge comes from the vertical prefixes (scanned across the
// chunk and the chunks to its left), each row adds its row prefix; writes
// ii[y+1][x+1] as u32 sum and QT square sum: u64 for the integral views, u32
// (the square sum modulo 2^32) for detect, whose window square sums are below
// 2^32 and so exact as u32 differences.
template <typename QT>
__global__ __launch_bounds__(64) void ii_write_kernel(const LevelDesc* __restrict__ L, int nItems,
                                                      const int32_t* __restrict__ itemOffs, int nLevels,
                                                      const uint32_t* __restrict__ colS,
                                                      const uint32_t* __restrict__ colQ,
                                                      const uint32_t* __restrict__ totS,
                                                      const uint64_t* __restrict__ totQ,
                                                      const uint32_t* __restrict__ rowS,
                                                      const uint32_t* __restrict__ rowQ,
                                                      uint32_t* __restrict__ iiS, QT* __restrict__ iiQ) {
    const int item = blockIdx.x;
    if (item >= nItems) return;
    const int li = find_segment(itemOffs, nLevels, item);
    const LevelDesc lv = L[li];
    const int local = item - lv.itemOff;
    const int b = local / lv.nChunks;
    const int c = local - b * lv.nChunks;
    const int lane = lane_id();
    const int x = c * kChunkCols + lane * 4;
    const int y0 = b * kBandRows;
    const int y1 = min(lv.h, y0 + kBandRows);
    const int64_t pitch = lv.iiPitch;
    uint32_t* outS = iiS + lv.iiOff;
    QT* outQ = iiQ + lv.iiOff;

    // carry from the chunks to the left: band totals, one chunk per lane (64
    // chunks at a time), and per-row totals, loaded four chunks at a time
    const int64_t tb = lv.totOff + static_cast<int64_t>(b) * lv.nChunks;
    uint32_t ccS = 0;
    QT ccQ = 0;
    for (int k0 = 0; k0 < c; k0 += kWave) {
        ccS += wave_sum(k0 + lane < c ? totS[tb + k0 + lane] : 0u);
        ccQ += wave_sum(k0 + lane < c ? static_cast<QT>(totQ[tb + k0 + lane]) : QT{0});
    }
    uint32_t rcS = 0, rcQ = 0;
    if (lane < kBandRows && y0 + lane < lv.h) {
        const int64_t r = lv.rowOff + static_cast<int64_t>(y0 + lane) * lv.nChunks;
        for (int k0 = 0; k0 < c; k0 += 4) {
            uint32_t vs[4], vq[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                vs[u] = k0 + u < c ? rowS[r + k0 + u] : 0u;
                vq[u] = k0 + u < c ? rowQ[r + k0 + u] : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                rcS += vs[u];
                rcQ += vq[u];
            }
        }
    }

    // integral along the band's top edge for this lane's four columns
    uint32_t vS[4], vQ[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const bool in = x + k < lv.w;
        vS[k] = in ? colS[lv.colOff + static_cast<int64_t>(b) * lv.w + x + k] : 0u;
        vQ[k] = in ? colQ[lv.colOff + static_cast<int64_t>(b) * lv.w + x + k] : 0u;
    }
    const uint32_t tS = vS[0] + vS[1] + vS[2] + vS[3];
    const QT tQ = static_cast<QT>(vQ[0]) + vQ[1] + vQ[2] + vQ[3];
    const uint32_t eS = wave_incl_scan(tS) - tS;
    const QT eQ = wave_incl_scan(tQ) - tQ;
    uint32_t aS[4];
    QT aQ[4];
    {
        uint32_t rs = ccS + eS;
        QT rq = ccQ + eQ;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rs += vS[k];
            rq += vQ[k];
            aS[k] = rs;
            aQ[k] = rq;
        }
    }
    // A lane writes integral elements x .. x+3 of each row: ii at pixel
    // columns x-1 .. x+2, its left neighbour's last value and its first three
    // (so that with a row pitch of whole 128-byte lines every store is one
    // aligned 16-byte piece); lane 0 takes the value left of the chunk, and
    // the lane whose last pixel is the row's last (x + 4 == w) at the end of
    // the last chunk also writes element w.
    const int w = lv.w;
    const bool tailEl = x + 4 == w && lane == kWave - 1;
    auto put_row = [&](int64_t row, uint32_t s0, const uint32_t (&sv)[4], QT q0, const QT (&qv)[4]) {
        uint32_t* rS = outS + row * pitch;
        QT* rQ = outQ + row * pitch;
        if (x + 3 <= w) {
            *reinterpret_cast<u32x4_a4*>(rS + x) = u32x4_a4{s0, sv[0], sv[1], sv[2]};
            if constexpr (sizeof(QT) == 4) {
                *reinterpret_cast<u32x4_a4*>(rQ + x) = u32x4_a4{q0, qv[0], qv[1], qv[2]};
            } else {
                *reinterpret_cast<u64x2_a8*>(rQ + x) = u64x2_a8{q0, qv[0]};
                *reinterpret_cast<u64x2_a8*>(rQ + x + 2) = u64x2_a8{qv[1], qv[2]};
            }
        } else if (x <= w) {
            rS[x] = s0;
            rQ[x] = q0;
#pragma unroll
            for (int k = 1; k < 4; ++k)
                if (x + k <= w) {
                    rS[x + k] = sv[k - 1];
                    rQ[x + k] = qv[k - 1];
                }
        }
        if (tailEl) {
            rS[w] = sv[3];
            rQ[w] = qv[3];
        }
    };
    if (b == 0) {  // row 0 of the integral is zero
        const uint32_t zs[4] = {0u, 0u, 0u, 0u};
        const QT zq[4] = {0u, 0u, 0u, 0u};
        put_row(0, 0u, zs, QT{0}, zq);
    }
    // the value left of the chunk (element c * 256) on the current row
    uint32_t leftS = ccS;
    QT leftQ = ccQ;
    uint32_t pv[kBandRows];
    load_band(lv, x, y0, y1, pv);
#pragma unroll
    for (int r = 0; r < kBandRows; ++r) {
        const int y = y0 + r;
        if (y >= y1) break;
        const uint32_t p[4] = {pv[r] & 0xFFu, (pv[r] >> 8) & 0xFFu, (pv[r] >> 16) & 0xFFu, pv[r] >> 24};
        const uint32_t ls = p[0] + p[1] + p[2] + p[3];
        const uint32_t lq = p[0] * p[0] + p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
        const uint32_t rowCarryS = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rcS), r));
        const uint32_t rowCarryQ = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(rcQ), r));
        uint32_t rs = wave_incl_scan_dpp(ls) - ls + rowCarryS;
        uint32_t rq = wave_incl_scan_dpp(lq) - lq + rowCarryQ;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            rs += p[k];
            rq += p[k] * p[k];
            aS[k] += rs;
            aQ[k] += rq;
        }
        leftS += rowCarryS;
        leftQ += rowCarryQ;
        uint32_t prevS = static_cast<uint32_t>(__shfl_up(static_cast<int>(aS[3]), 1, kWave));
        QT prevQ;
        if constexpr (sizeof(QT) == 4) prevQ = static_cast<QT>(__shfl_up(static_cast<int>(aQ[3]), 1, kWave));
        else prevQ = static_cast<QT>(__shfl_up(static_cast<unsigned long long>(aQ[3]), 1, kWave));
        if (lane == 0) {
            prevS = leftS;
            prevQ = leftQ;
        }
        put_row(y + 1, prevS, aS, prevQ, aQ);
    }
}

__global__ __launch_bounds__(256) void batch_integral_kernel(const uint8_t* __restrict__ gray, int w, int h,
                                                             int stride, int n, uint32_t* __restrict__ sumT,
                                                             uint64_t* __restrict__ sqT) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) batch_integral_one(gray, w, h, stride, n, s, sumT, sqT);
}

__global__ __launch_bounds__(256) void invnorm_kernel(const uint32_t* __restrict__ sumT,
                                                      const uint64_t* __restrict__ sqT, int iw, int ih, int n,
                                                      float* __restrict__ inv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) inv[s] = train_invnorm(sumT, sqT, iw, ih, n, s);
}

// Weighted rectangle sum of one stump's feature at a window origin in LDS
// (a byte pointer; the corner offsets are in bytes),
// in the CPU path's float order: ((p0 - p1) - p2) + p3 per rectangle, times
// its weight, summed r0 + r1 (+ r2) (src/main.cu:522-542).
__device__ __forceinline__ float haar_raw(const char* ob, const KStump& sp) {
    auto at = [ob](uint32_t off) { return *reinterpret_cast<const float*>(ob + off); };
    float raw = (((at(sp.off[0]) - at(sp.off[1])) - at(sp.off[2])) + at(sp.off[3])) * sp.w[0] +
                (((at(sp.off[4]) - at(sp.off[5])) - at(sp.off[6])) + at(sp.off[7])) * sp.w[1];
    if (sp.rectCount == 3) raw += (((at(sp.off[8]) - at(sp.off[9])) - at(sp.off[10])) + at(sp.off[11])) * sp.w[2];
    return raw;
}

constexpr int kCascadeBlock = 256;
// Survivors of the first pass are handed to the deep pass in slices of up to
// kDeepSlice windows (at most kDeepSlices = kTileMaxWin / kDeepSlice slices
// per tile; a work item is tile * kDeepSlices + slice).
constexpr int kDeepSlice = 64;
constexpr int kDeepSlices = kTileMaxWin / kDeepSlice;
// Layout of the deep and tail passes' counters (see deep_list_kernel).
constexpr int kDeepHist = 4, kDeepCursor = kDeepHist + kTileMaxWin + 1, kTailCtl = kDeepCursor + kTileMaxWin + 1;
constexpr int kDeepCtlInts = kTailCtl + 2;
// Stumps of the leading stages the first pass runs, and the deep pass's
// persistent grid (eight workgroups on each of 256 CUs).
constexpr int kSplitStumps = 128;
constexpr int kDeepGrid = 2048;
// The deep pass stops once it has run about kDeepStumps more stumps; its
// survivors go on to a tail pass in slices of kTailSlice windows, so a stage
// of the tail runs 256 / kTailSlice stumps of each window in parallel.  By
// then few windows are left, and slices of kDeepSlice would leave a few
// workgroups with long sequential stage sums after all others are done.
constexpr int kDeepStumps = 1600;
constexpr int kTailSlice = 8;
constexpr int kTailPerDeep = kDeepSlice / kTailSlice;
// Floats of LDS the deep and tail passes keep for pair-mode leaf values
// (after the staged footprint): eight windows of a 181-stump stage fit.
constexpr int kLeafCap = 2048;
// The deep pass when it runs every remaining stage (no tail pass): a 512-float
// leaf cache and at most 64 VGPRs, so that eight workgroups fit on a CU.
constexpr int kLeafCapWide = 512;

// LDS of one cascade workgroup besides the staged footprint.
struct CascadeSmem {
    uint16_t listIdx[kTileMaxWin];               // live windows (tile-local index)
    float invTab[kTileMaxWin];                   // normaliser by tile-local index
    float red[kCascadeBlock];                    // pair mode: partial stage sums
    int listCnt[2];
    int item;                                    // deep pass: current work item
    int tailBase;                                // deep pass: first tail item of the slice
    int box[4];                                  // deep / tail item: window columns and rows it covers
    unsigned long long maskWords[32];            // final bits per tile row
};

// A tile of windows: level, position and shape.
struct TileGeo {
    LevelDesc lv;
    int tx, ty, TH, lgc, TW, step, wx0, wy0;
};

__device__ __forceinline__ TileGeo tile_geo(const LevelDesc* __restrict__ L, const int32_t* __restrict__ lvlIdx,
                                            const int32_t* __restrict__ tileOffs, int nLv, int t) {
    TileGeo g;
    const int j = find_segment(tileOffs, nLv, t);
    g.lv = L[lvlIdx[j]];
    const int local = t - tileOffs[j];
    g.ty = local / g.lv.tilesX;
    g.tx = local - g.ty * g.lv.tilesX;
    g.TH = g.lv.tileRows;
    g.lgc = g.lv.tileLgCols;
    g.TW = 1 << g.lgc;
    g.step = g.lv.step;
    g.wx0 = g.tx * g.TW;
    g.wy0 = g.ty * g.TH;
    return g;
}

// Stages rows [ra, rb) x columns [ca, cb) of the tile's integral footprint
// (FH x FW entries, pitch FW) into LDS as f32, row by row: wave k takes rows
// ra + k, ra + k + 4, ..., four rows per step, each lane columns ca + lane
// and ca + lane + 64 of a row (a footprint row is at most 128 entries: the
// window is at most 64 wide), so eight loads are in flight per lane before
// their stores.  The first pass stages the whole footprint; a deep or tail
// item only the box its windows cover.
__device__ __forceinline__ void stage_footprint(const TileGeo& g, const uint32_t* __restrict__ iiS, float* lds,
                                                int winW, int ra, int rb, int ca, int cb) {
    constexpr int kWaves = kCascadeBlock / kWave;
    const int FW = (g.TW - 1) * g.step + winW + 1;
    const int ix0 = g.wx0 * g.step, iy0 = g.wy0 * g.step;
    const int64_t iw = g.lv.iiPitch;
    const uint32_t* sumL = iiS + g.lv.iiOff;
    const int lane = lane_id(), wave = static_cast<int>(threadIdx.x) / kWave;
    const int cx0 = ca + lane, cx1 = cx0 + kWave;
    const bool c0 = cx0 < cb && ix0 + cx0 <= g.lv.w;
    const bool c1 = cx1 < cb && ix0 + cx1 <= g.lv.w;
    for (int r0 = ra + wave; r0 < rb; r0 += 4 * kWaves) {
        float v[4][2];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = r0 + u * kWaves, gy = iy0 + r;
            const bool rowOk = r < rb && gy <= g.lv.h;
            const uint32_t* row = sumL + static_cast<int64_t>(gy) * iw + ix0;
            v[u][0] = (rowOk && c0) ? static_cast<float>(static_cast<int32_t>(row[cx0])) : 0.0f;
            v[u][1] = (rowOk && c1) ? static_cast<float>(static_cast<int32_t>(row[cx1])) : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = r0 + u * kWaves;
            if (r < rb) {
                if (cx0 < cb) lds[r * FW + cx0] = v[u][0];
                if (cx1 < cb) lds[r * FW + cx1] = v[u][1];
            }
        }
    }
}

// The footprint box of the windows in sm.box (tile-local window columns
// [box[0], box[1]] and rows [box[2], box[3]]).
__device__ __forceinline__ void stage_window_box(const TileGeo& g, const uint32_t* __restrict__ iiS, float* lds,
                                                 const int* box, int winW, int winH) {
    if (box[1] < box[0]) return;  // no windows
    stage_footprint(g, iiS, lds, winW, box[2] * g.step, box[3] * g.step + winH + 1, box[0] * g.step,
                    box[1] * g.step + winW + 1);
}

// Variance gate (exact int sum, u32 square-sum delta, f64 normaliser) into
// the live list; a tile has at most two windows per thread, both gated
// before either is appended so their global loads overlap.
__device__ __forceinline__ void gate_windows(const TileGeo& g, const uint32_t* __restrict__ iiS,
                                             const uint32_t* __restrict__ iiQ, CascadeSmem& sm, int winW,
                                             int winH) {
    const int64_t iw = g.lv.iiPitch;
    const uint32_t* sumL = iiS + g.lv.iiOff;
    const uint32_t* sqL = iiQ + g.lv.iiOff;
    const int nw = winW - 2, nh = winH - 2, narea = nw * nh;
    const int64_t np0 = 1 + iw, np1 = (1 + nw) + iw;
    const int64_t np2 = 1 + static_cast<int64_t>(iw) * (1 + nh), np3 = (1 + nw) + static_cast<int64_t>(iw) * (1 + nh);
    const int lane = lane_id();
    const int nWin = g.TW * g.TH;
    static_assert(kTileMaxWin <= 2 * kCascadeBlock, "two gate slots per thread");
    bool live[2];
    float inv[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int widx = static_cast<int>(threadIdx.x) + k * kCascadeBlock;
        const int wx = g.wx0 + (widx & (g.TW - 1));
        const int wy = g.wy0 + (widx >> g.lgc);
        live[k] = false;
        inv[k] = 0.0f;
        if (widx < nWin && wx < g.lv.scanW && wy < g.lv.scanH) {
            const int64_t o = static_cast<int64_t>(wy * g.step) * iw + wx * g.step;
            const int32_t ns = static_cast<int32_t>(sumL[o + np0] - sumL[o + np1] - sumL[o + np2] + sumL[o + np3]);
            const uint32_t nq = static_cast<uint32_t>(sqL[o + np0] - sqL[o + np1] - sqL[o + np2] + sqL[o + np3]);
            const double normSum = static_cast<double>(ns);
            const double nf2 = static_cast<double>(narea) * static_cast<double>(nq) - normSum * normSum;
            if (nf2 > 0.0) {
                inv[k] = static_cast<float>(1.0 / sqrt(nf2));
                live[k] = static_cast<double>(narea) * static_cast<double>(inv[k]) < 1e-1;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (k * kCascadeBlock >= nWin) break;
        const int widx = static_cast<int>(threadIdx.x) + k * kCascadeBlock;
        const uint64_t m = __ballot(live[k]);
        int wbase = 0;
        if (lane == 0 && m) wbase = atomicAdd(&sm.listCnt[0], __popcll(m));
        wbase = __shfl(wbase, 0, kWave);
        if (live[k]) {
            sm.listIdx[wbase + mask_rank(m)] = static_cast<uint16_t>(widx);
            sm.invTab[widx] = inv[k];
        }
    }
}

// Runs stages [si0, si1) over the live list sm.listIdx[0 .. sm.listCnt[0]).
// A window that passes the cascade's last stage sets its bit in
// sm.maskWords and stores its last stage sum; one that passes an earlier
// stage joins the next list.  The list is rewritten in place: every stage
// first reads the entries it needs into registers, then (after a barrier)
// appends its survivors; the two counters alternate between "read" and
// "append".  Returns the counter slot holding the survivors of stage si1-1.
// leaf (LDS, leafCap floats, or null) holds a pair-mode stage's leaf values
// when they fit, so a sequential stage sum adds them up again instead of
// re-evaluating the stumps.  With kCount, enter counts the windows entering
// each stage (the engine's counting mode; the kernels of ordinary calls are
// built without it).
// Stage rule and window normaliser: src/main.cu:752-804.
template <bool kCount>
__device__ int run_stages(const TileGeo& g, CascadeSmem& sm, const float* lds, const KStump* __restrict__ stumps,
                          const KStage* __restrict__ stages, int si0, int si1, int nStages, int winW,
                          float* __restrict__ scores, unsigned long long* __restrict__ enter, float* leaf = nullptr,
                          int leafCap = 0) {
    const int lane = lane_id();
    const int FW = (g.TW - 1) * g.step + winW + 1;
    auto settle = [&](bool pass, bool last, uint32_t widx, float stageSum, int nxt) {
        if (last) {
            if (pass) {
                const int wx = g.wx0 + static_cast<int>(widx & (g.TW - 1));
                const int wy = g.wy0 + static_cast<int>(widx >> g.lgc);
                const int64_t word = g.lv.wordOff + static_cast<int64_t>(wy) * g.lv.rowWords + (wx >> 6);
                scores[word * 64 + (wx & 63)] = stageSum;
                atomicOr(&sm.maskWords[widx >> g.lgc], 1ull << (widx & (g.TW - 1)));
            }
        } else {
            const uint64_t m = __ballot(pass);
            int wbase = 0;
            if (lane == 0 && m) wbase = atomicAdd(&sm.listCnt[nxt], __popcll(m));
            wbase = __shfl(wbase, 0, kWave);
            if (pass) sm.listIdx[wbase + mask_rank(m)] = static_cast<uint16_t>(widx);
        }
    };
    auto origin = [&](uint32_t widx) {
        return reinterpret_cast<const char*>(lds + static_cast<int>(widx >> g.lgc) * g.step * FW +
                                             static_cast<int>(widx & (g.TW - 1)) * g.step);
    };

    int cur = 0;
    for (int si = si0; si < si1; ++si) {
        const int n = sm.listCnt[cur];
        if (n == 0) break;
        if constexpr (kCount) {
            if (threadIdx.x == 0) atomicAdd(&enter[si], static_cast<unsigned long long>(n));
        }
        const int nxt = cur ^ 1;
        const KStage st = stages[si];
        const bool last = (si == nStages - 1);
        const int tid = threadIdx.x;
        const KStump* __restrict__ sg = stumps + st.first;
        const int nt = st.ntrees;
        if (n >= kCascadeBlock) {
            // window mode: one live window per lane (two rounds above 256
            // windows), stumps in order (the stump is uniform across the
            // wave, so its fields are scalar loads); two stumps per step so
            // their loads overlap, the two leaf values still added in order.
            const uint32_t e0 = sm.listIdx[tid];
            const uint32_t e1 = (tid + kCascadeBlock < n) ? sm.listIdx[tid + kCascadeBlock] : 0u;
            if (tid == 0) sm.listCnt[nxt] = 0;
            __syncthreads();
            for (int r = 0; r < 2; ++r) {
                if (r * kCascadeBlock >= n) break;
                const int i = r * kCascadeBlock + tid;
                const uint32_t widx = r ? e1 : e0;
                bool pass = false;
                float stageSum = 0.0f;
                if (i < n) {
                    const float invNorm = sm.invTab[widx];
                    const char* o = origin(widx);
                    int k = 0;
                    for (; k + 1 < nt; k += 2) {
                        const KStump s0 = sg[k];
                        const KStump s1 = sg[k + 1];
                        const float r0 = haar_raw(o, s0), r1 = haar_raw(o, s1);
                        stageSum += (r0 * invNorm < s0.theta) ? s0.left : s0.right;
                        stageSum += (r1 * invNorm < s1.theta) ? s1.left : s1.right;
                    }
                    if (k < nt) {
                        const KStump s0 = sg[k];
                        stageSum += (haar_raw(o, s0) * invNorm < s0.theta) ? s0.left : s0.right;
                    }
                    pass = !(stageSum < st.threshold);
                }
                settle(pass, last, widx, stageSum, nxt);
            }
        } else {
            // pair mode (fewer live windows than threads): the windows are
            // padded to nwp = 2^lgw slots and each pass evaluates spp =
            // 256 / nwp stumps on every slot, windows fastest, so a lane
            // keeps one window for the whole stage and a wave shares one
            // stump (scalar loads) whenever nwp >= 64; below that, each
            // lane's next stump record is fetched while the current one is
            // evaluated.  Every lane adds its stumps' leaf values into a
            // partial sum; the partial sums of a window are then added up
            // (within the wave, then across the waves).  That sum S differs
            // from the sequential stage sum (the reference's float order) by
            // at most the stage's bound B (both are sums of the same leaf
            // values, each within gamma_(n+16) * sum|leaf| of the exact
            // sum), so S < lo = thr - B fails and S > hi = thr + B passes
            // exactly as the sequential sum would.  Windows in between, and
            // every window the cascade's last stage may pass (its score is
            // the sequential sum), are summed again sequentially: from the
            // leaf values kept in LDS (row w, ntp entries per row) when they
            // fit, else by evaluating the stumps again.
            int lgw = 0;
            while ((1 << lgw) < n) ++lgw;
            const int nwp = 1 << lgw;
            const int spp = kCascadeBlock >> lgw;
            const int w = tid & (nwp - 1);
            const int js0 = tid >> lgw;
            const bool evalLive = w < n;
            const int ntp = (nt + 3) & ~3;
            float* const leafRow = (leaf && n * ntp <= leafCap) ? leaf + w * ntp : nullptr;
            const uint32_t ew = evalLive ? sm.listIdx[w] : 0u;
            const bool mine = tid < n;
            const uint32_t myIdx = mine ? sm.listIdx[tid] : 0u;
            if (tid == 0) sm.listCnt[nxt] = 0;
            const char* o = origin(ew);
            const float inv = sm.invTab[ew];
            float part = 0.0f;
            if (nwp >= kWave) {
                for (int jb = 0; jb < nt; jb += spp) {
                    // js is uniform across the wave; the record is a scalar
                    // load used at once (a scalar load still in flight would
                    // make every wait on the LDS gathers wait for it too)
                    const int js = __builtin_amdgcn_readfirstlane(jb + js0);
                    if (js < nt && evalLive) {
                        const KStump sp = sg[js];
                        const float v = (haar_raw(o, sp) * inv < sp.theta) ? sp.left : sp.right;
                        part += v;
                        if (leafRow) leafRow[js] = v;
                    }
                }
            } else if (nt > 0) {
                // each lane's records come one pass ahead, from indices clamped
                // into the stage (always valid; a clamped record is loaded but
                // not used), in two registers that alternate, so no record is
                // zeroed or copied
                const int last = nt - 1;
                KStump recA = sg[min(js0, last)];
                for (int jb = 0; jb < nt; jb += 2 * spp) {
                    const int jsA = jb + js0, jsB = jsA + spp;
                    const KStump recB = sg[min(jsB, last)];
                    if (evalLive && jsA < nt) {
                        const float v = (haar_raw(o, recA) * inv < recA.theta) ? recA.left : recA.right;
                        part += v;
                        if (leafRow) leafRow[jsA] = v;
                    }
                    if (jb + spp >= nt) break;
                    recA = sg[min(jsB + spp, last)];
                    if (evalLive && jsB < nt) {
                        const float v = (haar_raw(o, recB) * inv < recB.theta) ? recB.left : recB.right;
                        part += v;
                        if (leafRow) leafRow[jsB] = v;
                    }
                }
            }
            // lanes l and l ^ d (d >= nwp) of a wave hold the same window
            for (int d = nwp; d < kWave; d <<= 1) part += __shfl_xor(part, d, kWave);
            sm.red[tid] = part;
            __syncthreads();
            float stageSum = 0.0f;
            bool pass = false;
            if (mine) {
                // window tid's partial sums: lanes tid + k * max(nwp, 64)
                const int per = max(nwp, kWave);
                float S = sm.red[tid];
                for (int k = per; k < kCascadeBlock; k += per) S += sm.red[tid + k];
                if (S < st.lo) {
                    pass = false;
                } else if (S > st.hi && !last) {
                    pass = true;
                } else if (leafRow) {
                    // leaf values of window tid, in stump order
                    const float4* lr = reinterpret_cast<const float4*>(leaf + tid * ntp);
                    int k = 0;
                    for (; k + 4 <= nt; k += 4) {
                        const float4 v = lr[k >> 2];
                        stageSum += v.x;
                        stageSum += v.y;
                        stageSum += v.z;
                        stageSum += v.w;
                    }
                    for (; k < nt; ++k) stageSum += leaf[tid * ntp + k];
                    pass = !(stageSum < st.threshold);
                } else {
                    const char* om = origin(myIdx);
                    const float invm = sm.invTab[myIdx];
                    for (int k = 0; k < nt; ++k) {
                        const KStump sp = sg[k];
                        stageSum += (haar_raw(om, sp) * invm < sp.theta) ? sp.left : sp.right;
                    }
                    pass = !(stageSum < st.threshold);
                }
            }
            if (tid < kWave * ((n + kWave - 1) / kWave)) settle(pass, last, myIdx, stageSum, nxt);
        }
        __syncthreads();
        cur = nxt;
    }
    return cur;
}

// Cascade evaluation, pass 1: one workgroup of kCascadeBlock threads per tile
// of 2^lgc x TH windows of one level.  The tile's integral footprint is
// staged in LDS as f32 (each absolute u32 value converted once: the CPU path
// converts the same absolute values, so every Haar sum rounds identically).
// Windows failing the variance gate never enter the live list; each stage
// then runs over the live list only and the survivors are compacted by
// wavefront ballot, so lanes stay busy however unevenly windows reject.
// Stages [0, stageSplit) run here.  When that is the whole cascade, the
// survivors' mask bits are written here; otherwise the tile's mask pieces
// are written as zero and its survivors (tile-local index and normaliser)
// are appended to a global list for the deep pass.
template <bool kCount>
__global__ __launch_bounds__(kCascadeBlock) void cascade_kernel(
    const LevelDesc* __restrict__ L, const int32_t* __restrict__ lvlIdx, const int32_t* __restrict__ tileOffs,
    int nLv, int nTiles, const uint32_t* __restrict__ iiS, const uint32_t* __restrict__ iiQ,
    const KStump* __restrict__ stumpsAll, int nStumps, const KStage* __restrict__ stages, int nStages,
    int stageSplit, int winW, int winH, uint64_t* __restrict__ words, float* __restrict__ scores,
    int32_t* __restrict__ survTotal, uint16_t* __restrict__ survIdx, float* __restrict__ survInv,
    int32_t* __restrict__ tileSurvOff, int32_t* __restrict__ tileCount, unsigned long long* __restrict__ enter) {
    extern __shared__ float lds[];
    __shared__ CascadeSmem sm;
    __shared__ int survBase;

    // XCD-aware remap: consecutive tiles land on the same XCD (shared L2).
    int t;
    {
        const int i = blockIdx.x, q = nTiles >> 3, r = nTiles & 7;
        const int xcd = i & 7, pos = i >> 3;
        t = (xcd < r) ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
    }
    const TileGeo g = tile_geo(L, lvlIdx, tileOffs, nLv, t);
    const bool deep = stageSplit < nStages;

    // Tile row r is the TW-bit piece of mask word (wy0 + r, wx0 / 64) that
    // starts at bit wx0 % 64 (little-endian: a 32- or 16-bit store).  Every
    // piece of every word is written by exactly one tile, so the mask needs
    // no clearing.
    auto write_mask = [&](int r, uint64_t bits) {
        const int64_t word = g.lv.wordOff + static_cast<int64_t>(g.wy0 + r) * g.lv.rowWords + (g.wx0 >> 6);
        if (g.lgc == 6) {
            words[word] = bits;
        } else if (g.lgc == 5) {
            reinterpret_cast<uint32_t*>(words)[word * 2 + ((g.wx0 & 63) >> 5)] = static_cast<uint32_t>(bits);
        } else {
            reinterpret_cast<uint16_t*>(words)[word * 4 + ((g.wx0 & 63) >> 4)] = static_cast<uint16_t>(bits);
        }
    };
    if (g.wx0 >= g.lv.scanW) {  // padding tile past the scan width: zero pieces
        for (int r = threadIdx.x; r < g.TH; r += kCascadeBlock)
            if (g.wy0 + r < g.lv.scanH) write_mask(r, 0ull);
        if (deep && threadIdx.x == 0) tileCount[t] = 0;
        return;
    }

    stage_footprint(g, iiS, lds, winW, 0, (g.TH - 1) * g.step + winH + 1, 0, (g.TW - 1) * g.step + winW + 1);
    if (threadIdx.x < 32) sm.maskWords[threadIdx.x] = 0ull;
    if (threadIdx.x == 0) {
        sm.listCnt[0] = 0;
        sm.listCnt[1] = 0;
    }
    __syncthreads();
    gate_windows(g, iiS, iiQ, sm, winW, winH);
    __syncthreads();
    const KStump* __restrict__ stumps = stumpsAll + static_cast<size_t>(g.step - 1) * nStumps;
    const int slot =
        run_stages<kCount>(g, sm, lds, stumps, stages, 0, min(stageSplit, nStages), nStages, winW, scores, enter);

    for (int r = threadIdx.x; r < g.TH; r += kCascadeBlock)
        if (g.wy0 + r < g.lv.scanH) write_mask(r, deep ? 0ull : sm.maskWords[r]);
    if (deep) {
        const int n = sm.listCnt[slot];
        if (threadIdx.x == 0) {
            survBase = n ? atomicAdd(survTotal, n) : 0;
            tileSurvOff[t] = survBase;
            tileCount[t] = n;
            if (n > 0) atomicAdd(&survTotal[kDeepHist + n], 1);  // the work list's histogram
        }
        __syncthreads();
        for (int i = threadIdx.x; i < n; i += kCascadeBlock) {
            const uint32_t widx = sm.listIdx[i];
            survIdx[survBase + i] = static_cast<uint16_t>(widx);
            survInv[survBase + i] = sm.invTab[widx];
        }
    }
}

// Deep-pass work list: every tile with survivors after the first pass
// contributes ceil(count / kDeepSlice) items (tile * kDeepSlices + slice),
// ordered by survivor count, largest first, so the longest items start
// first.  The counters are zeroed by the integral pass (ii_vscan_kernel), the
// first pass adds each tile to the histogram of survivor counts, and one
// launch then turns the histogram into each count's first item (counts in
// descending order, recomputed by every workgroup) and places every tile's
// items at its count's offset (tiles of one count in any order).  ctl layout:
// [0] survivor total, [1] items, [2] work counter, then per count its tile
// histogram and its cursor, then the tail pass's item total and work counter.

__global__ __launch_bounds__(256) void zero_u64_kernel(unsigned long long* __restrict__ a, int n) {
    for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0ull;
}

constexpr int kListThreads = 256;
static_assert(kTileMaxWin == 2 * kListThreads, "the count scan takes two counts per thread");

__global__ __launch_bounds__(kListThreads) void deep_list_kernel(const int32_t* __restrict__ tileCount, int nTiles,
                                                                 int32_t* __restrict__ ctl,
                                                                 int32_t* __restrict__ items) {
    __shared__ int32_t base[kTileMaxWin + 1];
    __shared__ int32_t cnt[kTileMaxWin + 1];
    __shared__ int32_t off[kTileMaxWin + 1];
    __shared__ int32_t waveTot[kListThreads / kWave];
    const int tid = static_cast<int>(threadIdx.x);
    // each count's first item: thread i takes counts kTileMaxWin - 2i and
    // kTileMaxWin - 2i - 1 (descending counts)
    {
        const int c0 = kTileMaxWin - 2 * tid, c1 = c0 - 1;
        const int32_t v0 = ctl[kDeepHist + c0] * ((c0 + kDeepSlice - 1) / kDeepSlice);
        const int32_t v1 = ctl[kDeepHist + c1] * ((c1 + kDeepSlice - 1) / kDeepSlice);
        const int32_t t = v0 + v1;
        const int32_t inc = wave_incl_scan(t);
        if (lane_id() == kWave - 1) waveTot[tid / kWave] = inc;
        __syncthreads();
        int32_t b = inc - t, total = 0;
        for (int w = 0; w < kListThreads / kWave; ++w) {
            if (w < tid / kWave) b += waveTot[w];
            total += waveTot[w];
        }
        base[c0] = b;
        base[c1] = b + v0;
        if (blockIdx.x == 0 && tid == 0) ctl[1] = total;
    }
    for (int t0 = blockIdx.x * kListThreads; t0 < nTiles; t0 += gridDim.x * kListThreads) {
        for (int c = tid; c <= kTileMaxWin; c += kListThreads) cnt[c] = 0;
        __syncthreads();
        const int t = t0 + tid;
        const int c = t < nTiles ? tileCount[t] : 0;
        const int local = c > 0 ? atomicAdd(&cnt[c], 1) : 0;
        __syncthreads();
        for (int b = tid; b <= kTileMaxWin; b += kListThreads)
            if (cnt[b]) off[b] = atomicAdd(&ctl[kDeepCursor + b], cnt[b]);
        __syncthreads();
        if (c > 0) {
            const int ns = (c + kDeepSlice - 1) / kDeepSlice;
            const int pos = base[c] + (off[c] + local) * ns;
            for (int k = 0; k < ns; ++k) items[pos + k] = t * kDeepSlices + k;
        }
        __syncthreads();  // cnt and off are rewritten by the next chunk
    }
}

// Cascade evaluation, deep pass: persistent workgroups take work items
// (a slice of up to kDeepSlice first-pass survivors of one tile) from a
// counter until none are left, restage the tile's footprint and run stages
// [stageSplit, stageTail) over the slice.  When stageTail is the cascade's
// end, the final bits are ORed into the mask words (several slices of a tile
// share words).  Otherwise the slice's survivors are written back over its
// own entries of survIdx / survInv (no other item reads them) and appended
// to the tail pass's work list in items of up to kTailSlice windows.
template <bool kCount, int kWaves = 6>
__global__ __launch_bounds__(kCascadeBlock) __attribute__((amdgpu_waves_per_eu(kWaves == 8 ? 8 : 1, 8))) void cascade_deep_kernel(
    const LevelDesc* __restrict__ L, const int32_t* __restrict__ lvlIdx, const int32_t* __restrict__ tileOffs,
    int nLv, const uint32_t* __restrict__ iiS, const KStump* __restrict__ stumpsAll, int nStumps,
    const KStage* __restrict__ stages, int nStages, int stageSplit, int stageTail, int winW, int winH, int leafOff,
    unsigned long long* __restrict__ words, float* __restrict__ scores, uint16_t* __restrict__ survIdx,
    float* __restrict__ survInv, const int32_t* __restrict__ tileSurvOff, const int32_t* __restrict__ tileCount,
    const int32_t* __restrict__ items, const int32_t* __restrict__ nItems, int32_t* __restrict__ workCounter,
    int4* __restrict__ tailItems, int32_t* __restrict__ tailTotal, unsigned long long* __restrict__ enter) {
    extern __shared__ float lds[];
    __shared__ CascadeSmem sm;
    const int total = *nItems;
    const bool tail = stageTail < nStages;
    for (;;) {
        if (threadIdx.x == 0) {
            sm.item = atomicAdd(workCounter, 1);
            sm.box[0] = sm.box[2] = INT_MAX;
            sm.box[1] = sm.box[3] = -1;
        }
        __syncthreads();
        const int item = sm.item;
        if (item >= total) break;
        const int t = items[item] / kDeepSlices, slice = items[item] % kDeepSlices;
        const TileGeo g = tile_geo(L, lvlIdx, tileOffs, nLv, t);
        const int first = slice * kDeepSlice;
        const int m = min(kDeepSlice, tileCount[t] - first);
        const int off = tileSurvOff[t] + first;
        if (threadIdx.x < 32) sm.maskWords[threadIdx.x] = 0ull;
        if (static_cast<int>(threadIdx.x) < m) {
            const uint32_t widx = survIdx[off + threadIdx.x];
            sm.listIdx[threadIdx.x] = static_cast<uint16_t>(widx);
            sm.invTab[widx] = survInv[off + threadIdx.x];
            const int wx = static_cast<int>(widx & static_cast<uint32_t>(g.TW - 1)), wy = static_cast<int>(widx >> g.lgc);
            atomicMin(&sm.box[0], wx);
            atomicMax(&sm.box[1], wx);
            atomicMin(&sm.box[2], wy);
            atomicMax(&sm.box[3], wy);
        }
        if (threadIdx.x == 0) {
            sm.listCnt[0] = m;
            sm.listCnt[1] = 0;
        }
        __syncthreads();
        stage_window_box(g, iiS, lds, sm.box, winW, winH);
        __syncthreads();
        const KStump* __restrict__ stumps = stumpsAll + static_cast<size_t>(g.step - 1) * nStumps;
        const int slot = run_stages<kCount>(g, sm, lds, stumps, stages, stageSplit, stageTail, nStages, winW,
                                            scores, enter, lds + leafOff, kWaves == 8 ? kLeafCapWide : kLeafCap);
        if (tail) {
            const int n = sm.listCnt[slot];
            const int nt = (n + kTailSlice - 1) / kTailSlice;
            if (static_cast<int>(threadIdx.x) < n) {
                const uint32_t widx = sm.listIdx[threadIdx.x];
                survIdx[off + threadIdx.x] = static_cast<uint16_t>(widx);
                survInv[off + threadIdx.x] = sm.invTab[widx];
            }
            if (threadIdx.x == 0) sm.tailBase = nt ? atomicAdd(tailTotal, nt) : 0;
            __syncthreads();
            if (static_cast<int>(threadIdx.x) < nt) {
                const int a = static_cast<int>(threadIdx.x) * kTailSlice;
                tailItems[sm.tailBase + threadIdx.x] = make_int4(t, off + a, min(kTailSlice, n - a), 0);
            }
        } else {
            for (int r = threadIdx.x; r < g.TH; r += kCascadeBlock) {
                const unsigned long long bits = sm.maskWords[r];
                if (bits) {
                    const int64_t word =
                        g.lv.wordOff + static_cast<int64_t>(g.wy0 + r) * g.lv.rowWords + (g.wx0 >> 6);
                    atomicOr(&words[word], bits << (g.wx0 & 63));
                }
            }
        }
        __syncthreads();
    }
}

// Cascade evaluation, tail pass: persistent workgroups take the deep pass's
// tail items (up to kTailSlice windows of one tile: tile, first survIdx
// entry, count) from a counter, restage the tile's footprint, run stages
// [stageTail, nStages) and OR the final bits into the mask words.
template <bool kCount>
__global__ __launch_bounds__(kCascadeBlock) void cascade_tail_kernel(
    const LevelDesc* __restrict__ L, const int32_t* __restrict__ lvlIdx, const int32_t* __restrict__ tileOffs,
    int nLv, const uint32_t* __restrict__ iiS, const KStump* __restrict__ stumpsAll, int nStumps,
    const KStage* __restrict__ stages, int nStages, int stageTail, int winW, int winH, int leafOff,
    unsigned long long* __restrict__ words, float* __restrict__ scores, const uint16_t* __restrict__ survIdx,
    const float* __restrict__ survInv, const int4* __restrict__ tailItems, const int32_t* __restrict__ tailTotal,
    int32_t* __restrict__ workCounter, unsigned long long* __restrict__ enter) {
    extern __shared__ float lds[];
    __shared__ CascadeSmem sm;
    const int total = *tailTotal;
    for (;;) {
        if (threadIdx.x == 0) {
            sm.item = atomicAdd(workCounter, 1);
            sm.box[0] = sm.box[2] = INT_MAX;
            sm.box[1] = sm.box[3] = -1;
        }
        __syncthreads();
        const int item = sm.item;
        if (item >= total) break;
        const int4 it = tailItems[item];
        const TileGeo g = tile_geo(L, lvlIdx, tileOffs, nLv, it.x);
        if (threadIdx.x < 32) sm.maskWords[threadIdx.x] = 0ull;
        if (static_cast<int>(threadIdx.x) < it.z) {
            const uint32_t widx = survIdx[it.y + threadIdx.x];
            sm.listIdx[threadIdx.x] = static_cast<uint16_t>(widx);
            sm.invTab[widx] = survInv[it.y + threadIdx.x];
            const int wx = static_cast<int>(widx & static_cast<uint32_t>(g.TW - 1)), wy = static_cast<int>(widx >> g.lgc);
            atomicMin(&sm.box[0], wx);
            atomicMax(&sm.box[1], wx);
            atomicMin(&sm.box[2], wy);
            atomicMax(&sm.box[3], wy);
        }
        if (threadIdx.x == 0) {
            sm.listCnt[0] = it.z;
            sm.listCnt[1] = 0;
        }
        __syncthreads();
        stage_window_box(g, iiS, lds, sm.box, winW, winH);
        __syncthreads();
        const KStump* __restrict__ stumps = stumpsAll + static_cast<size_t>(g.step - 1) * nStumps;
        run_stages<kCount>(g, sm, lds, stumps, stages, stageTail, nStages, nStages, winW, scores, enter,
                           lds + leafOff, kLeafCap);
        for (int r = threadIdx.x; r < g.TH; r += kCascadeBlock) {
            const unsigned long long bits = sm.maskWords[r];
            if (bits) {
                const int64_t word = g.lv.wordOff + static_cast<int64_t>(g.wy0 + r) * g.lv.rowWords + (g.wx0 >> 6);
                atomicOr(&words[word], bits << (g.wx0 & 63));
            }
        }
        __syncthreads();
    }
}

// Output compaction in three steps over blocks of kWordBlock mask words
// (blocks never straddle images): per-block survivor counts, one scan that
// turns them into per-image output offsets, and a block-parallel emission.
constexpr int kWordBlock = 1024;

__global__ __launch_bounds__(256) void block_count_kernel(const int64_t* __restrict__ blockWord,
                                                          const uint64_t* __restrict__ words,
                                                          int32_t* __restrict__ blockCount) {
    __shared__ int32_t part[4];
    const int b = blockIdx.x;
    const int64_t a = blockWord[b], e = blockWord[b + 1];
    int32_t c = 0;
    for (int64_t i = a + threadIdx.x; i < e; i += blockDim.x) c += __popcll(words[i]);
    c = wave_sum(c);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) blockCount[b] = part[0] + part[1] + part[2] + part[3];
}

// One workgroup: image totals (raw counts), each image's output base
// sum_{j<i} min(count_j, cap), and each block's offset within its image.
__global__ __launch_bounds__(1024) void block_scan_kernel(const int32_t* __restrict__ imgBlock, int nImages,
                                                          const int32_t* __restrict__ blockCount,
                                                          int32_t* __restrict__ blockOff,
                                                          int32_t* __restrict__ counts,
                                                          int32_t* __restrict__ imgBase, int cap) {
    __shared__ int32_t waveTot[16];
    __shared__ int32_t carry;
    int32_t base = 0;
    for (int img = 0; img < nImages; ++img) {
        const int a = imgBlock[img], e = imgBlock[img + 1];
        if (threadIdx.x == 0) carry = 0;
        __syncthreads();
        for (int c0 = a; c0 < e; c0 += blockDim.x) {
            const int i = c0 + threadIdx.x;
            const int32_t v = i < e ? blockCount[i] : 0;
            const int32_t inc = wave_incl_scan(v);
            if ((threadIdx.x & 63) == 63) waveTot[threadIdx.x >> 6] = inc;
            __syncthreads();
            int32_t off = carry + inc - v;
            int32_t tot = 0;
            for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) {
                if (k < static_cast<int>(threadIdx.x >> 6)) off += waveTot[k];
                tot += waveTot[k];
            }
            if (i < e) blockOff[i] = off;
            __syncthreads();
            if (threadIdx.x == 0) carry += tot;
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            counts[img] = carry;
            imgBase[img] = base;
        }
        base += min(carry, cap);
        __syncthreads();
    }
}

// Emission: detections of block b go to out[imgBase + blockOff[b] + rank]
// while the rank within the image stays below cap; word order is (level, y,
// x) scan order.  Box rounding: src/main.cu:806-813.
__global__ __launch_bounds__(256) void emit_kernel(const LevelDesc* __restrict__ L,
                                                   const int64_t* __restrict__ lvWordOff, int nLevels,
                                                   const int64_t* __restrict__ blockWord,
                                                   const int32_t* __restrict__ blockImg,
                                                   const uint64_t* __restrict__ words,
                                                   const float* __restrict__ scores,
                                                   const int32_t* __restrict__ blockOff,
                                                   const int32_t* __restrict__ imgBase, int cap, int winW,
                                                   int winH, vj_detection* __restrict__ out) {
    __shared__ int32_t wavePart[4];
    const int b = blockIdx.x;
    const int64_t a = blockWord[b], e = blockWord[b + 1];
    const int32_t off0 = blockOff[b];
    if (off0 >= cap) return;
    const int64_t nw = e - a;
    const int64_t per = (nw + blockDim.x - 1) / blockDim.x;
    const int64_t my0 = a + min(nw, per * threadIdx.x);
    const int64_t my1 = a + min(nw, per * (threadIdx.x + 1));
    int32_t c = 0;
    for (int64_t i = my0; i < my1; ++i) c += __popcll(words[i]);
    const int32_t inc = wave_incl_scan(c);
    if ((threadIdx.x & 63) == 63) wavePart[threadIdx.x >> 6] = inc;
    __syncthreads();
    int32_t rank = off0 + inc - c;
    for (int k = 0; k < static_cast<int>(threadIdx.x >> 6); ++k) rank += wavePart[k];
    if (c == 0 || rank >= cap) return;
    vj_detection* dst = out + imgBase[blockImg[b]];
    int li = -1;
    LevelDesc lv{};
    for (int64_t i = my0; i < my1 && rank < cap; ++i) {
        uint64_t m = words[i];
        if (!m) continue;
        if (li < 0 || i < lv.wordOff || i >= lv.wordOff + static_cast<int64_t>(lv.scanH) * lv.rowWords) {
            li = find_segment64(lvWordOff, nLevels, i);
            lv = L[li];
        }
        const int64_t rel = i - lv.wordOff;
        const int wy = static_cast<int>(rel / lv.rowWords);
        const int wxw = static_cast<int>(rel - static_cast<int64_t>(wy) * lv.rowWords);
        while (m && rank < cap) {
            const int bit = __builtin_ctzll(m);
            m &= m - 1;
            const int wx = wxw * 64 + bit;
            vj_detection d;
            d.x = static_cast<int>(static_cast<float>(wx * lv.step) * lv.scale + 0.5f);
            d.y = static_cast<int>(static_cast<float>(wy * lv.step) * lv.scale + 0.5f);
            d.w = static_cast<int>(static_cast<float>(winW) * lv.scale + 0.5f);
            d.h = static_cast<int>(static_cast<float>(winH) * lv.scale + 0.5f);
            d.scaleIdx = lv.scaleIdx;
            d.score = scores[i * 64 + bit];
            dst[rank] = d;
            ++rank;
        }
    }
}

// ===========================================================================
// Host side
// ===========================================================================

// Offsets of a plan's tables: elements in the 32- and 64-bit metadata arrays
// and bytes in the engine's plan buffer.
struct PlanTables {
    size_t oItem = 0, oChunk = 0, oSrc = 0, oLvl = 0, oTile = 0, oLvWord = 0, oBlockWord = 0, oBlockImg = 0,
           oImgBlock = 0;
    size_t aLv = 0, aRs = 0, a32 = 0, a64 = 0, aLJobs = 0, aGJobs = 0;
    int nBlocks = 0;
};

struct Plan {
    std::vector<LevelDesc> levels;
    std::vector<float> rs;          // per level: resize scale x, y
    std::vector<int32_t> src;       // per level: source level index (itself for level 0)
    std::vector<int32_t> itemOffs, chunkOffs;
    std::vector<int64_t> lvWordOff;
    std::vector<int64_t> imgWordOff;
    // the pyramid's launches in order: one per large resize depth (a job per
    // level), one per group of small depths (a job per image)
    struct ResizeStep {
        bool group;
        int32_t first, count;  // jobs in levelJobs / groupJobs
        int32_t items;         // workgroups
        int32_t rows;          // a depth's launch: output rows per workgroup
    };
    std::vector<ResizeStep> resizeSteps;
    std::vector<LevelJob> levelJobs;
    std::vector<ResizeJob> groupJobs;
    PlanTables tab;  // where run_pipeline placed the tables in the plan buffer
    std::vector<int32_t> tileLvl, tileOff;  // scanned levels and their first cascade tile
    int32_t nTiles = 0;
    int32_t maxStep = 0;
    int32_t nItems = 0, nChunkItems = 0;
    int64_t pyrBytes = 0, iiElems = 0, colElems = 0, totElems = 0, rowElems = 0, nWords = 0;
    int64_t windows = 0, pyramidPixels = 0, integralBytes = 0, cascadeBytes = 0;
    std::vector<int32_t> imgLevels;  // levels per image
    std::vector<uint8_t> inPyr;      // per level: pixels live in the pyramid buffer
    std::vector<int64_t> pyrOffs;    // pyramid byte offsets of the inPyr levels, in order
};

}  // namespace vj

using namespace vj;

// ---------------------------------------------------------------------------
// Grouping help on the device: the sweep order and the similar pairs of each
// image's raw boxes (src/face_vision_engine.cu:1417-1445), so that the host
// only replays the unions and builds the clusters.  Box a's sweep rank is the
// number of boxes before it in (x, index) order; its pairs are the boxes b
// after it in that order with similarRects(a, b) (the tolerance (int)(eps *
// (min w + min h) * 0.5f), every edge within it), which all lie within a's
// x reach, eps * (w + h) / 2.  Images with more than kGroupDevMax raw boxes
// are left to the host.  Every kernel derives the images' box counts from
// the device counts.
// ---------------------------------------------------------------------------
constexpr int kGroupDevMax = 1 << 24;

// Grouping on the device.  The components of the similarity graph do not
// depend on the order of the unions, so the pair search unions every similar
// pair at once in a lock-free union-find over sweep positions (the larger
// root is linked under the smaller, so a component's representative is its
// smallest sweep position).  Per component: box sums, count, last member and
// the best score (the smallest raw index among the members of the highest
// score, as the reference's strict > in input order keeps the first).  Only
// the reference's cluster order needs its exact union-by-rank replay: the
// output lists clusters by the raw index of the DSU root that
// groupRectanglesCpu's unions leave, so each kept component's own pairs are
// replayed in the reference's order (src/face_vision_engine.cu:1430-1445).
struct GrpCount {
    int32_t cnt, last;  // members; largest member position
};
// A kept component's cluster: integer-mean box, member count, the raw index
// of the reference's DSU root, the best score and its scale index.
struct GrpCluster {
    int32_t x, y, w, h, n, rootRaw, scaleIdx;
    float score;
};

// Union-find over sweep positions.  A label changes once, from itself to a
// smaller position (a root linked under a smaller one), so the labels never
// form a cycle and a find over values that are not yet the latest still ends
// at a node of the same component.
__device__ __forceinline__ int32_t uf_find(const int32_t* label, int32_t x) {
    for (;;) {
        const int32_t px = label[x];
        if (px == x) return x;
        x = px;
    }
}

// Links the components of a and b, the larger root under the smaller.  A
// failed compare-and-swap returns b's new parent, the next candidate root.
__device__ __forceinline__ void uf_union(int32_t* label, int32_t a, int32_t b) {
    a = uf_find(label, a);
    b = uf_find(label, b);
    while (a != b) {
        if (a > b) {
            const int32_t t = a;
            a = b;
            b = t;
        }
        const int32_t old = atomicCAS(label + b, b, a);
        if (old == b) return;
        b = uf_find(label, old);
        a = uf_find(label, a);
    }
}

// Ordered key of a cluster score candidate: 0 for NaN and -inf (they never
// beat the initial -inf), and -0 keyed as +0 (the two compare equal).
__device__ __forceinline__ uint32_t best_key(float s) {
    if (!(s > -INFINITY)) return 0u;
    const uint32_t b = __float_as_uint(s + 0.0f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ int grouped_boxes(const int32_t* __restrict__ counts, int i, int cap) {
    const int n = min(counts[i], cap);
    return n <= kGroupDevMax ? n : 0;
}

__device__ __forceinline__ int64_t all_boxes(const int32_t* __restrict__ counts, int nImages, int cap) {
    int64_t total = 0;
    for (int i = 0; i < nImages; ++i) total += min(counts[i], cap);
    return total;
}

// The image of box g (boxes of the images concatenated, base imgBase[i]).
__device__ __forceinline__ int image_of(const int32_t* __restrict__ imgBase, int nImages, int64_t g) {
    int lo = 0, hi = nImages - 1;  // the last image whose base is at most g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (imgBase[mid] <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Sweep order without a sort, one workgroup per image.  A box's rank in
// its image's (x, index) order is the number of the image's boxes with a
// smaller x plus the number with the same x and a smaller index.  An x
// histogram in LDS (one bin per column 0 .. w, or per 2^s columns for an
// image wider than kSweepMaxBins - 1) is scanned into each bin's first rank;
// each bin's boxes are listed (in any order) in a bucket at that rank, and a
// box's ties are the bucket entries before it in (x, index) order.  The
// boxes are then placed in sweep order: sorted[base + rank] = the box's
// edges, order[base + rank] = its index in its image.
constexpr int kSweepThreads = 1024;
constexpr int kSweepMaxBins = 8193;

// Column shift of the x bins of an image w wide.
__host__ __device__ inline int sweep_shift(int w) {
    int sh = 0;
    while ((w >> sh) + 1 > kSweepMaxBins) ++sh;
    return sh;
}

// An image with at most kSweepLdsBoxes boxes keeps its buckets in LDS
// (after the bins): the tie counts then read LDS, not global memory.
constexpr int kSweepLdsBoxes = 16384;

inline size_t sweep_lds_bytes(int maxW, int ldsBoxes) {
    const size_t nb = static_cast<size_t>(maxW >> sweep_shift(maxW)) + 1;
    return sizeof(int32_t) * (2 * nb + 1 + static_cast<size_t>(ldsBoxes));
}

__global__ __launch_bounds__(kSweepThreads) void pair_sweep_kernel(const vj_detection* __restrict__ dets,
                                                                  const int32_t* __restrict__ counts,
                                                                  const int32_t* __restrict__ imgBase,
                                                                  const int32_t* __restrict__ imgW, int cap,
                                                                  int ldsBoxes, int32_t* __restrict__ bucket,
                                                                  int4* __restrict__ sorted,
                                                                  int32_t* __restrict__ order,
                                                                  int32_t* __restrict__ grpImg,
                                                                  int64_t* __restrict__ ctl) {
    extern __shared__ int32_t xs[];  // cnt[nb] (then the fill cursor), first[nb + 1], LDS buckets
    __shared__ int32_t waveTot[kSweepThreads / kWave];
    const int i = blockIdx.x;
    if (threadIdx.x == 0) {
        // this call's grouping counters: kept clusters, output count, host flag
        grpImg[3 * i] = 0;
        grpImg[3 * i + 1] = 0;
        grpImg[3 * i + 2] = 0;
        if (i == 0) ctl[1] = 0;
    }
    const int n = grouped_boxes(counts, i, cap);
    if (n == 0) return;
    const int W = imgW[i], sh = sweep_shift(W);
    const int nb = (W >> sh) + 1;
    int32_t* cnt = xs;
    int32_t* first = xs + nb;
    const vj_detection* D = dets + imgBase[i];
    auto bin = [W, sh](int x) { return min(max(x, 0), W) >> sh; };
    for (int k = threadIdx.x; k < nb; k += kSweepThreads) cnt[k] = 0;
    __syncthreads();
    for (int g = threadIdx.x; g < n; g += kSweepThreads) atomicAdd(&cnt[bin(D[g].x)], 1);
    __syncthreads();
    // exclusive scan of cnt into first: each thread a run of consecutive bins
    const int per = (nb + kSweepThreads - 1) / kSweepThreads;
    const int k0 = min(static_cast<int>(threadIdx.x) * per, nb), k1 = min(k0 + per, nb);
    int32_t run = 0;
    for (int k = k0; k < k1; ++k) run += cnt[k];
    const int32_t inc = wave_incl_scan(run);
    if (lane_id() == kWave - 1) waveTot[threadIdx.x / kWave] = inc;
    __syncthreads();
    int32_t off = inc - run;
    for (int w = 0; w < static_cast<int>(threadIdx.x / kWave); ++w) off += waveTot[w];
    for (int k = k0; k < k1; ++k) {
        first[k] = off;
        off += cnt[k];
    }
    if (threadIdx.x == kSweepThreads - 1) first[nb] = n;
    __syncthreads();
    for (int k = threadIdx.x; k < nb; k += kSweepThreads) cnt[k] = 0;
    __syncthreads();
    int32_t* B = n <= ldsBoxes ? xs + 2 * nb + 1 : bucket + imgBase[i];
    for (int g = threadIdx.x; g < n; g += kSweepThreads) {
        const int x = bin(D[g].x);
        B[first[x] + atomicAdd(&cnt[x], 1)] = g;
    }
    __syncthreads();  // the bucket (LDS, or global and this workgroup's own) is complete
    for (int g = threadIdx.x; g < n; g += kSweepThreads) {
        const vj_detection d = D[g];
        const int x = bin(d.x);
        const int b0 = first[x], b1 = first[x + 1];
        int32_t ties = 0;
        if (sh == 0) {
            for (int j = b0; j < b1; ++j) ties += B[j] < g;
        } else {
            for (int j = b0; j < b1; ++j) {
                const int o = B[j];
                const int ox = D[o].x;
                ties += ox < d.x || (ox == d.x && o < g);
            }
        }
        const int64_t p = imgBase[i] + b0 + ties;
        sorted[p] = make_int4(d.x, d.y, d.w, d.h);
        order[p] = g;
    }
}

// Per sweep position p, one wavefront: the similar boxes after it, 64
// candidates at a time in sweep order until the x reach is passed (kWrite
// false: their count into pairCnt[p]; true: their sweep ranks, ascending,
// from pairOff[p], placed by ballot rank so that the order is kept).
template <bool kWrite>
__global__ __launch_bounds__(256) void pair_find_kernel(const int32_t* __restrict__ counts,
                                                       const int32_t* __restrict__ imgBase, int nImages, int cap,
                                                       float eps, const int4* __restrict__ sorted,
                                                       int32_t* __restrict__ pairCnt,
                                                       const int64_t* __restrict__ pairOff,
                                                       int32_t* __restrict__ pairs, int64_t pairCap,
                                                       int32_t* __restrict__ label, GrpCount* __restrict__ gc) {
    const int64_t total = all_boxes(counts, nImages, cap);
    const int lane = lane_id();
    const int64_t wavesTotal = static_cast<int64_t>(gridDim.x) * (blockDim.x / kWave);
    for (int64_t p = static_cast<int64_t>(blockIdx.x) * (blockDim.x / kWave) + threadIdx.x / kWave; p < total;
         p += wavesTotal) {
        const int i = image_of(imgBase, nImages, p);
        const int n = grouped_boxes(counts, i, cap);
        if constexpr (!kWrite) {
            // every position starts as its own component, with empty sums
            if (lane == 0) {
                label[p] = static_cast<int32_t>(p);
                gc[p] = GrpCount{0, 0};
            }
        }
        if (n == 0) {  // an image left to the host: no pairs here
            if constexpr (!kWrite) {
                if (lane == 0) pairCnt[p] = 0;
            }
            continue;
        }
        const int64_t base = imgBase[i], end = base + n;
        const int4 A = sorted[p];
        const int ar = A.x + A.z, ab = A.y + A.w;
        const int reach = static_cast<int>(eps * static_cast<float>(A.z + A.w) * 0.5f);
        int64_t o = 0;
        bool store = true;
        if constexpr (kWrite) {
            o = pairOff[p];
            store = o + pairCnt[p] <= pairCap;
        }
        int found = 0;
        for (int64_t q0 = p + 1; q0 < end; q0 += kWave) {
            const int64_t q = q0 + lane;
            const bool valid = q < end;
            const int4 B = valid ? sorted[q] : A;
            const bool inReach = valid && B.x - A.x <= reach;
            bool sim = false;
            if (inReach) {
                const int d = static_cast<int>(eps * static_cast<float>(min(A.z, B.z) + min(A.w, B.w)) * 0.5f);
                sim = abs(A.x - B.x) <= d && abs(A.y - B.y) <= d && abs(ar - (B.x + B.z)) <= d &&
                      abs(ab - (B.y + B.w)) <= d;
            }
            const uint64_t m = __ballot(sim);
            if constexpr (kWrite) {
                if (sim) {
                    if (store) pairs[o + found + mask_rank(m)] = static_cast<int32_t>(q - base);
                    uf_union(label, static_cast<int32_t>(p), static_cast<int32_t>(q));
                }
            }
            found += __popcll(m);
            if (__ballot(valid && !inReach)) break;  // sweep order: the rest is out of reach too
        }
        if constexpr (!kWrite) {
            if (lane == 0) pairCnt[p] = found;
        }
    }
}

// Exclusive scan of the pair counts of all sweep positions, in two launches
// over blocks of kScanItems counts (each thread kScanPer consecutive ones):
// pair_scan_local_kernel writes each block's own exclusive offsets and its
// total; pair_scan_add_kernel adds the totals of the blocks before (64-bit
// offsets) and block 0 of the last block writes the grand total into ctl[0].
constexpr int kScanThreads = 256, kScanPer = 16, kScanItems = kScanThreads * kScanPer;

__global__ __launch_bounds__(kScanThreads) void pair_scan_local_kernel(const int32_t* __restrict__ counts,
                                                                      int nImages, int cap,
                                                                      const int32_t* __restrict__ pairCnt,
                                                                      int64_t* __restrict__ pairOff,
                                                                      int64_t* __restrict__ blockTot) {
    __shared__ int64_t waveTot[kScanThreads / kWave];
    const int64_t total = all_boxes(counts, nImages, cap);
    const int64_t nb = (total + kScanItems - 1) / kScanItems;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const int64_t g0 = b * kScanItems + static_cast<int64_t>(threadIdx.x) * kScanPer;
        int32_t v[kScanPer];
        int64_t t = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            v[k] = g0 + k < total ? pairCnt[g0 + k] : 0;
            t += v[k];
        }
        const int64_t inc = static_cast<int64_t>(wave_incl_scan(static_cast<uint64_t>(t)));
        if (lane == kWave - 1) waveTot[wave] = inc;
        __syncthreads();
        int64_t off = inc - t, blk = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) {
            if (w < wave) off += waveTot[w];
            blk += waveTot[w];
        }
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            if (g0 + k < total) pairOff[g0 + k] = off;
            off += v[k];
        }
        if (threadIdx.x == 0) blockTot[b] = blk;
        __syncthreads();  // waveTot is rewritten by the next block
    }
}

__global__ __launch_bounds__(kScanThreads) void pair_scan_add_kernel(const int32_t* __restrict__ counts,
                                                                    int nImages, int cap,
                                                                    const int64_t* __restrict__ blockTot,
                                                                    int64_t* __restrict__ pairOff,
                                                                    int64_t* __restrict__ ctl) {
    __shared__ int64_t part[kScanThreads / kWave];
    const int64_t total = all_boxes(counts, nImages, cap);
    const int64_t nb = (total + kScanItems - 1) / kScanItems;
    const int lane = lane_id(), wave = threadIdx.x / kWave;
    if (nb == 0 && blockIdx.x == 0 && threadIdx.x == 0) ctl[0] = 0;
    for (int64_t b = blockIdx.x; b < nb; b += gridDim.x) {
        int64_t pre = 0;
        for (int64_t k = threadIdx.x; k < b; k += kScanThreads) pre += blockTot[k];
        pre = wave_sum(pre);
        if (lane == 0) part[wave] = pre;
        __syncthreads();
        pre = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) pre += part[w];
        const int64_t g0 = b * kScanItems + static_cast<int64_t>(threadIdx.x) * kScanPer;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k)
            if (g0 + k < total) pairOff[g0 + k] += pre;
        if (b == nb - 1 && threadIdx.x == 0) ctl[0] = pre + blockTot[b];
        __syncthreads();  // part is rewritten by the next block
    }
}

// Per sweep position: its component (the root of the union-find, written
// back into label, which from here on names each position's component), and
// the component's member count and last member.  Lanes of a wave that share
// a component add up before one atomic each.  Nothing runs when the pair
// buffer overflowed (the call is then grouped on the host).
__global__ __launch_bounds__(256) void grp_comp_kernel(const int32_t* __restrict__ counts,
                                                       const int32_t* __restrict__ imgBase, int nImages, int cap,
                                                       int32_t* __restrict__ label, GrpCount* __restrict__ gc,
                                                       const int64_t* __restrict__ ctl, int64_t pairCap) {
    if (ctl[0] > pairCap) return;
    const int64_t total = all_boxes(counts, nImages, cap);
    const int lane = lane_id();
    for (int64_t g0 = (static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x) & ~int64_t{kWave - 1}; g0 < total;
         g0 += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        const int64_t g = g0 + lane;
        bool valid = g < total;
        int32_t r = -1;
        if (valid) {
            valid = grouped_boxes(counts, image_of(imgBase, nImages, g), cap) > 0;
            if (valid) {
                r = uf_find(label, static_cast<int32_t>(g));
                label[g] = r;
            }
        }
        uint64_t act = __ballot(valid);
        while (act) {
            const int leader = __ffsll(static_cast<long long>(act)) - 1;
            const int32_t rl = __builtin_amdgcn_readlane(r, leader);
            const uint64_t m = __ballot(valid && r == rl);
            if (lane == leader) {
                atomicAdd(&gc[rl].cnt, __popcll(m));
                atomicMax(&gc[rl].last, static_cast<int32_t>(g0 + 63 - __clzll(static_cast<long long>(m))));
            }
            act &= ~m;
        }
    }
}

// Each component root whose count passes the minNeighbors rule (count >
// max(1, minNeighbors)) gets the next cluster slot of its image (grpImg[3 i]
// counts them) and a place in the list of all kept components (ctl[1]).
__global__ __launch_bounds__(256) void grp_list_kernel(const int32_t* __restrict__ counts,
                                                       const int32_t* __restrict__ imgBase, int nImages, int cap,
                                                       int minNeighbors, const int32_t* __restrict__ label,
                                                       const GrpCount* __restrict__ gc, int32_t* __restrict__ keptAll,
                                                       int32_t* __restrict__ keptSlot, int32_t* __restrict__ grpImg,
                                                       int64_t* __restrict__ ctl, int64_t pairCap) {
    if (ctl[0] > pairCap) return;
    const int64_t total = all_boxes(counts, nImages, cap);
    const int mn = max(1, minNeighbors);
    for (int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; g < total;
         g += static_cast<int64_t>(gridDim.x) * blockDim.x) {
        if (label[g] != g || gc[g].cnt <= mn) continue;
        const int i = image_of(imgBase, nImages, g);
        if (grouped_boxes(counts, i, cap) == 0) continue;
        const int64_t k = static_cast<int64_t>(atomicAdd(reinterpret_cast<unsigned long long*>(&ctl[1]), 1ull));
        keptAll[k] = static_cast<int32_t>(g);
        keptSlot[k] = atomicAdd(&grpImg[3 * i], 1);
    }
}

// One kept component per wavefront.  First the span between its root (its
// smallest sweep position) and its last member is staged in LDS, all loads
// in parallel: the member bits and each position's pair-list offset; the
// cluster sums and the best score are reduced on the way.  Then the
// reference's union by rank (ties to the first argument's root) is replayed
// over the members' pairs in the reference's order -- members by ascending
// sweep position, each one's partners in ascending sweep order
// (src/face_vision_engine.cu:1430-1445).  Pair lists are stored in sweep
// order, so the span's lists form one range of pair entries, and that range
// in order is exactly this order (entries of other components are masked
// out).  It is read 64 entries at a time through an LDS window of
// kReplayEdges entries (each entry with its owner position), reloaded by all
// lanes together.  Per step of kReplayU chunks the lanes find the roots and
// ranks of both ends of their entries at once; the unions the chunks need
// are then applied one by one in order from registers (every slot whose
// root lost takes the winner and its new rank), and each union writes only
// the loser's parent and the winner's rank to LDS for the steps after.  The replay stops once
// the members form one set.  The root's raw index orders the clusters.  A
// span of more than kReplaySpan positions leaves its image to the host.
constexpr int kReplaySpan = 4096;
constexpr int kReplayEdges = 8192;
constexpr int kReplayU = 4;  // chunks of 64 pair entries per replay step
struct ReplayLds {
    int32_t off[kReplaySpan + 1];                 // pair-list offsets, relative to the span's first
    unsigned long long member[kReplaySpan / 64];  // member bits
    int16_t par[kReplaySpan];                     // union-find parents (span-local)
    int16_t edge[kReplayEdges];                   // window: partner (span-local; -1 outside)
    int16_t owner[kReplayEdges];                  // window: the position whose list holds the entry
    uint8_t rnk[kReplaySpan];                     // union-by-rank ranks
};

__global__ __launch_bounds__(64) void grp_replay_kernel(
    const int32_t* __restrict__ imgBase, int nImages, const vj_detection* __restrict__ dets,
    const int4* __restrict__ sorted, const int32_t* __restrict__ order, const int32_t* __restrict__ label,
    const GrpCount* __restrict__ gc, const int32_t* __restrict__ keptAll, const int32_t* __restrict__ keptSlot,
    const int64_t* __restrict__ ctl, int64_t pairCap, const int64_t* __restrict__ pairOff,
    const int32_t* __restrict__ pairCnt, const int32_t* __restrict__ pairs, int32_t* __restrict__ grpImg,
    GrpCluster* __restrict__ clusters) {
    __shared__ ReplayLds S;
    if (ctl[0] > pairCap) return;
    const int64_t nKept = ctl[1];
    const int lane = lane_id();
    auto find = [&](int x) {
        for (;;) {
            const int p = S.par[x];
            if (p == x) return x;
            const int gp = S.par[p];
            if (gp == p) return p;
            S.par[x] = static_cast<int16_t>(gp);  // path halving: the root is unchanged
            x = gp;
        }
    };
    for (int64_t k = blockIdx.x; k < nKept; k += gridDim.x) {
        const int32_t r = keptAll[k];
        const int i = image_of(imgBase, nImages, r);
        const int32_t base = imgBase[i];
        const int32_t last = gc[r].last;
        const int span = last - r + 1;
        if (span > kReplaySpan) {
            if (lane == 0) grpImg[3 * i + 2] = 1;
            continue;
        }
        const int64_t e0 = pairOff[r];
        const int64_t E = pairOff[last] + pairCnt[last] - e0;
        unsigned long long sx = 0, sy = 0, sw = 0, sh = 0;
        int n = 0;
        uint32_t bestKey = 0;
        int32_t bestRaw = INT_MAX;
        // four positions per lane per step: every independent load of the
        // step is issued before the first is used
        constexpr int kU = 4;
        for (int j0 = 0; j0 < span; j0 += kU * kWave) {
            int32_t lab[kU], raw[kU];
            int64_t po[kU];
            int4 B[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int j = j0 + u * kWave + lane;
                const int32_t p = r + min(j, span - 1);
                lab[u] = label[p];
                po[u] = pairOff[p];
                B[u] = sorted[p];
                raw[u] = order[p];
            }
            float sc[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int j = j0 + u * kWave + lane;
                sc[u] = (j < span && lab[u] == r) ? dets[base + raw[u]].score : 0.0f;
            }
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const int j = j0 + u * kWave + lane;
                const bool in = j < span;
                const bool mem = in && lab[u] == r;
                const uint64_t mb = __ballot(mem);
                if (lane == 0 && j0 + u * kWave < span) S.member[(j0 + u * kWave) >> 6] = mb;
                if (in) {
                    S.off[j] = static_cast<int32_t>(po[u] - e0);
                    S.par[j] = static_cast<int16_t>(j);
                    S.rnk[j] = 0;
                }
                if (mem) {
                    sx += static_cast<unsigned long long>(B[u].x);
                    sy += static_cast<unsigned long long>(B[u].y);
                    sw += static_cast<unsigned long long>(B[u].z);
                    sh += static_cast<unsigned long long>(B[u].w);
                    ++n;
                    const uint32_t kk = best_key(sc[u]);
                    if (kk && (kk > bestKey || (kk == bestKey && raw[u] < bestRaw))) {
                        bestKey = kk;
                        bestRaw = raw[u];
                    }
                }
            }
        }
        if (lane == 0) S.off[span] = static_cast<int32_t>(E);
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sw = wave_sum(sw);
        sh = wave_sum(sh);
        n = wave_sum(n);
        const uint32_t topKey = wave_max(bestKey);
        int32_t topRaw = (bestKey == topKey && topKey) ? bestRaw : INT_MAX;
#pragma unroll
        for (int d = kWave / 2; d > 0; d >>= 1) topRaw = min(topRaw, __shfl_xor(topRaw, d, kWave));

        // the window of pair entries [winLo, winHi): partners and owners
        int64_t winLo = 0, winHi = 0;
        auto load_window = [&](int64_t from) {
            __syncthreads();  // the staged span is complete; the old window is no longer read
            winLo = from;
            winHi = min(E, from + kReplayEdges);
            // kWinLoads loads in flight per lane before the first is used
            constexpr int kWinLoads = 32;
            for (int64_t t0 = from; t0 < winHi; t0 += kWinLoads * kWave) {
                int32_t pv[kWinLoads];
#pragma unroll
                for (int u = 0; u < kWinLoads; ++u) {
                    const int64_t t = t0 + u * kWave + lane;
                    pv[u] = t < winHi ? pairs[e0 + t] : 0;
                }
#pragma unroll
                for (int u = 0; u < kWinLoads; ++u) {
                    const int64_t t = t0 + u * kWave + lane;
                    const int32_t q = base + pv[u] - r;
                    if (t < winHi) S.edge[t - from] = static_cast<int16_t>(q >= 0 && q < span ? q : -1);
                }
            }
            int lo = 0, hi = span - 1;  // the last position whose list starts at or before `from`
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (S.off[mid] <= from) lo = mid;
                else hi = mid - 1;
            }
            for (int j = lo + lane; j < span && S.off[j] < winHi; j += kWave) {
                const int64_t a0 = max(static_cast<int64_t>(S.off[j]), from);
                const int64_t a1 = min(static_cast<int64_t>(S.off[j + 1]), winHi);
                for (int64_t t = a0; t < a1; ++t) S.owner[t - from] = static_cast<int16_t>(j);
            }
            __syncthreads();
        };
        load_window(0);
        // once the members form one set no later pair changes a root.  Each
        // step takes kReplayU chunks of 64 entries (entry t0 + 64 u + lane in
        // slot u): the roots of all their ends are found together, walking
        // the parent links in lockstep (every link of a round is read before
        // any is used), then the chunks' unions are applied in entry order,
        // each one relabelling the roots held by the slots not yet done.
        int sets = n;
        for (int64_t t0 = 0; t0 < E && sets > 1; t0 += kReplayU * kWave) {
            if (t0 + kReplayU * kWave > winHi && winHi < E) load_window(t0);
            bool valid[kReplayU];
            int x[2 * kReplayU];
#pragma unroll
            for (int u = 0; u < kReplayU; ++u) {
                const int64_t t = t0 + u * kWave + lane;
                int src = 0, dst = -1;
                if (t < winHi) {
                    src = S.owner[t - winLo];
                    dst = S.edge[t - winLo];
                }
                // (a member's partners are members, so inside the span)
                valid[u] = dst >= 0 && ((S.member[src >> 6] >> (src & 63)) & 1ull);
                x[2 * u] = valid[u] ? src : 0;
                x[2 * u + 1] = valid[u] ? dst : 0;
            }
            int start[2 * kReplayU];
#pragma unroll
            for (int k = 0; k < 2 * kReplayU; ++k) start[k] = x[k];
            for (;;) {
                int p[2 * kReplayU];
#pragma unroll
                for (int k = 0; k < 2 * kReplayU; ++k) p[k] = S.par[x[k]];
                bool moved = false;
#pragma unroll
                for (int k = 0; k < 2 * kReplayU; ++k) {
                    moved |= p[k] != x[k];
                    x[k] = p[k];
                }
                if (!__ballot(moved)) break;
            }
            int rs[kReplayU], rq[kReplayU], ks[kReplayU], kq[kReplayU];
#pragma unroll
            for (int u = 0; u < kReplayU; ++u) {
                rs[u] = x[2 * u];
                rq[u] = x[2 * u + 1];
                ks[u] = S.rnk[rs[u]];
                kq[u] = S.rnk[rq[u]];
            }
            // path compression of the ends (every lane writing a node writes
            // the root all lanes found for it; no root changes)
#pragma unroll
            for (int u = 0; u < kReplayU; ++u) {
                if (valid[u]) {
                    S.par[start[2 * u]] = static_cast<int16_t>(rs[u]);
                    S.par[start[2 * u + 1]] = static_cast<int16_t>(rq[u]);
                }
            }
#pragma unroll
            for (int u = 0; u < kReplayU; ++u) {
                // the roots of this slot were relabelled by every earlier
                // union; their ranks may have grown since, so they are read
                // again (the unions wrote them to LDS)
                if (u > 0) {
                    ks[u] = S.rnk[rs[u]];
                    kq[u] = S.rnk[rq[u]];
                }
                const uint64_t vm = __ballot(valid[u]);
                uint64_t nm = vm & __builtin_amdgcn_sicmp(rs[u], rq[u], 33 /* ne */);
                while (nm) {
                    const int f = __ffsll(static_cast<long long>(nm)) - 1;  // uniform: read the lane directly
                    const int a = __builtin_amdgcn_readlane(rs[u], f), b = __builtin_amdgcn_readlane(rq[u], f);
                    const int ra = __builtin_amdgcn_readlane(ks[u], f), rb = __builtin_amdgcn_readlane(kq[u], f);
                    const int win = ra < rb ? b : a, lose = ra < rb ? a : b;
                    const int rw = ra == rb ? ra + 1 : max(ra, rb);
                    if (lane == 0) {
                        S.par[lose] = static_cast<int16_t>(win);
                        S.rnk[win] = static_cast<uint8_t>(rw);
                    }
#pragma unroll
                    for (int v = u; v < kReplayU; ++v) {
                        rs[v] = rs[v] == lose ? win : rs[v];
                        rq[v] = rq[v] == lose ? win : rq[v];
                    }
                    ks[u] = rs[u] == win ? rw : ks[u];
                    kq[u] = rq[u] == win ? rw : kq[u];
                    nm = vm & __builtin_amdgcn_sicmp(rs[u], rq[u], 33 /* ne */);
                    --sets;
                }
            }
        }
        const int root = find(0);
        if (lane == 0) {
            GrpCluster c;
            c.x = static_cast<int32_t>(sx / static_cast<unsigned long long>(n));
            c.y = static_cast<int32_t>(sy / static_cast<unsigned long long>(n));
            c.w = static_cast<int32_t>(sw / static_cast<unsigned long long>(n));
            c.h = static_cast<int32_t>(sh / static_cast<unsigned long long>(n));
            c.n = n;
            c.rootRaw = order[r + root];
            if (topRaw != INT_MAX) {
                const vj_detection& d = dets[base + topRaw];
                c.scaleIdx = d.scaleIdx;
                c.score = d.score;
            } else {
                c.scaleIdx = 0;
                c.score = -INFINITY;
            }
            clusters[base + keptSlot[k]] = c;
        }
        __syncthreads();  // the LDS serves the next component
    }
}

// Clusters of one image (one workgroup per image): the kept components in
// ascending order of their DSU root's raw index, minus those inside a
// stronger cluster (src/face_vision_engine.cu:1483-1536: r1 is dropped when
// some other r2 with (int)leftBound <= r2.x <= (int)rightBound contains it
// within (int)(r2.w * eps), (int)(r2.h * eps) and n2 > max(3, n1) || n1 < 3).
// One wavefront per cluster compares it with 64 others at a time.  More
// than kFinishMax clusters leave the image to the host.  The clusters of
// image i are written from position sum_{j<i} (kept components of image j)
// of the output, so one copy of that prefix brings every image's.
constexpr int kFinishMax = 2048;
constexpr int kFinishThreads = 1024;

__global__ __launch_bounds__(kFinishThreads) void grp_finish_kernel(
    const int32_t* __restrict__ counts, const int32_t* __restrict__ imgBase, int cap, float eps,
    const GrpCluster* __restrict__ clusters, const int64_t* __restrict__ ctl, int64_t pairCap,
    int32_t* __restrict__ grpImg, vj_detection* __restrict__ grouped) {
    __shared__ int32_t X[kFinishMax], Y[kFinishMax], W[kFinishMax], H[kFinishMax], N[kFinishMax], R[kFinishMax];
    __shared__ int32_t byRank[kFinishMax];
    __shared__ uint8_t keepFlag[kFinishMax];
    __shared__ int32_t waveTot[kFinishThreads / kWave];
    __shared__ int32_t sMaxW;
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = lane_id(), wave = tid / kWave;
    if (ctl[0] > pairCap || grouped_boxes(counts, i, cap) == 0 || grpImg[3 * i + 2]) return;
    const int K = grpImg[3 * i];
    if (K > kFinishMax) {
        if (tid == 0) grpImg[3 * i + 2] = 1;
        return;
    }
    const int32_t base = imgBase[i];
    int64_t outBase = 0;
    for (int j = 0; j < i; ++j) outBase += grpImg[3 * j];
    if (tid == 0) sMaxW = 0;
    __syncthreads();
    int mw = 0;
    for (int c = tid; c < K; c += kFinishThreads) {
        const GrpCluster g = clusters[base + c];
        X[c] = g.x;
        Y[c] = g.y;
        W[c] = g.w;
        H[c] = g.h;
        N[c] = g.n;
        R[c] = g.rootRaw;
        mw = max(mw, g.w);
    }
    mw = wave_max(mw);
    if (lane == 0) atomicMax(&sMaxW, mw);
    __syncthreads();
    const float maxW = static_cast<float>(sMaxW);
    for (int c = wave; c < K; c += kFinishThreads / kWave) {
        const int x1 = X[c], y1 = Y[c], w1 = W[c], h1 = H[c], n1 = N[c], root1 = R[c];
        const int lo = static_cast<int>(static_cast<float>(x1 + w1) - (1.0f + eps) * maxW);
        const int hi = static_cast<int>(static_cast<float>(x1) + eps * maxW);
        int rank = 0;
        bool inside = false;
        for (int o = lane; o < K; o += kWave) {
            rank += R[o] < root1;
            const int x2 = X[o];
            if (o == c || x2 < lo || x2 > hi) continue;
            const int w2 = W[o], h2 = H[o], y2 = Y[o], n2 = N[o];
            const int dx = static_cast<int>(static_cast<float>(w2) * eps);
            const int dy = static_cast<int>(static_cast<float>(h2) * eps);
            inside = inside || (x1 >= x2 - dx && y1 >= y2 - dy && x1 + w1 <= x2 + w2 + dx &&
                                y1 + h1 <= y2 + h2 + dy && (n2 > max(3, n1) || n1 < 3));
        }
        rank = wave_sum(rank);
        const bool drop = __ballot(inside) != 0;
        if (lane == 0) {
            byRank[rank] = c;
            keepFlag[rank] = drop ? 0 : 1;
        }
    }
    __syncthreads();
    // survivors in rank order: an exclusive scan of the flags, two per thread
    const int r0 = 2 * tid, r1 = r0 + 1;
    const int f0 = r0 < K ? keepFlag[r0] : 0, f1 = r1 < K ? keepFlag[r1] : 0;
    const int t = f0 + f1;
    const int inc = wave_incl_scan(t);
    if (lane == kWave - 1) waveTot[wave] = inc;
    __syncthreads();
    int off = inc - t;
    for (int w = 0; w < wave; ++w) off += waveTot[w];
    auto put = [&](int rk, int pos) {
        const int c = byRank[rk];
        const GrpCluster& g = clusters[base + c];
        grouped[outBase + pos] = vj_detection{X[c], Y[c], W[c], H[c], g.scaleIdx, g.score};
    };
    if (f0) put(r0, off);
    if (f1) put(r1, off + f0);
    if (tid == kFinishThreads - 1) grpImg[3 * i + 1] = off + t;
}

struct vj_engine {
    int winW = 24, winH = 24, maxSamples = 1, maxW = 1, maxH = 1, maxDet = 1;
    bool tilted = false;

    DevBuf<vj_feature> pool;
    int poolN = 0;

    // integral batch (API view)
    DevBuf<uint32_t> bSum;
    DevBuf<uint64_t> bSq;
    DevBuf<float> bInv;
    int curN = 0, curW = 0, curH = 0;

    // cascade model
    DevBuf<vj_feature> modelFeat;
    DevBuf<KStump> kst;
    DevBuf<KStage> kstage;
    int nModelFeat = 0, nStumps = 0, nStages = 0;

    // detect work buffers
    DevBuf<uint8_t> pyr;
    DevBuf<uint32_t> iiS, colS, colQ, rowS, rowQ, totS;
    DevBuf<uint32_t> iiQ;  // detect's square-sum integrals modulo 2^32
    DevBuf<uint64_t> totQ;
    DevBuf<uint64_t> words;
    DevBuf<float> scores;
    DevBuf<LevelDesc> dLevels;
    DevBuf<float> dRs;
    DevBuf<int32_t> dMeta32;
    DevBuf<int64_t> dMeta64;
    DevBuf<uint8_t> dPlan;                     // a detect call's tables (levels, scales, offsets, jobs)
    DevBuf<int32_t> dCounts, dBlockCount, dBlockOff, dImgBase;
    // two-pass cascade: first-pass survivors and the deep pass's work list
    int stageSplit = 0;                        // stages run by the first pass
    int stageTail = 0;                         // first stage of the tail pass (nStages: none)
    int stageTailDefault = 0;                  // stageTail from the split rule (autotune off)
    DevBuf<uint16_t> survIdx;
    DevBuf<float> survInv;
    DevBuf<int32_t> tileSurvOff, tileCount, deepItems, deepCtl;  // deepCtl: total, items, counter, list scratch
    DevBuf<int4> tailItems;                    // tail pass: tile, first survivor entry, count
    bool countStages = false;                  // counting mode (vj_engine_count_stages)
    DevBuf<unsigned long long> stageEnter;     // counting mode: windows entering each stage
    std::vector<int64_t> stageWindows;         // the same, of the last counted call
    DevBuf<vj_detection> dOut;
    PinnedBuf<vj_detection> hOut;
    PinnedBuf<int32_t> hCounts;
    // device grouping: per raw box its pair count and pair-list offset, the
    // boxes in sweep order and their raw indices, the pair lists, the
    // union-find labels (then components) and per-component sums; the kept
    // components per image and over all images; per image its kept count,
    // output count and host flag; the clusters of all images
    DevBuf<int32_t> gPairCnt, gOrder, gPairs, gLabel, gKeptAll, gKeptSlot, gGrpImg;
    DevBuf<int64_t> gPairOff;
    DevBuf<GrpCount> gCount;
    DevBuf<GrpCluster> gClusters;
    DevBuf<vj_detection> gOutG;
    PinnedBuf<int32_t> hGrpImg;
    PinnedBuf<vj_detection> hOutG;
    // sweep order: per image its width (its x bins), the boxes listed by x
    DevBuf<int32_t> gImgW, gBucket;
    std::vector<int32_t> imgWUp;  // the image widths gImgW holds
    PinnedBuf<int32_t> hImgW;
    DevBuf<int4> gSorted;
    DevBuf<int64_t> gCtl;
    DevBuf<int64_t> gScanTot;  // pair-count scan: per-block totals
    PinnedBuf<int64_t> hCtl;
    int64_t pairCap = 0;  // capacity of gPairs (grows after an overflow)
    PinnedBuf<uint8_t> hStage;  // staging for plan uploads

    vj_detect_stats stats{};

    // cascade tile rows per window step (index 1..3), chosen per batch shape
    // by the autotuner when it is on
    int tileRows[4] = {0, tile_rows_for_step(1), tile_rows_for_step(2), tile_rows_for_step(3)};
    bool autotune = false;
    std::map<std::vector<int64_t>, std::array<int, 5>> tuned;  // tile rows 1..3, tail start

    // the detect launch sequence captured as a HIP graph, replayed while the
    // plan shape, the tile rows and every buffer address stay the same.  A
    // shape is captured on its second consecutive call (lastKey): a caller
    // whose image sizes change every call launches directly and pays no
    // capture.
    // the last detect call's plan, reused by a call with the same images
    Plan plan;
    std::vector<int64_t> planKey;
    const uint8_t* planPyr = nullptr;
    bool planValid = false;

    bool graphs = true;
    std::vector<int64_t> graphKey, lastKey;
    hipGraphExec_t graphExec = nullptr;
    hipStream_t captureStream = nullptr;  // captures run here; replays on the caller's stream

    // optional per-phase timing: events before resize, integral, cascade,
    // compaction and after compaction
    bool timing = false;
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    float phaseMs[4] = {0, 0, 0, 0};

    ~vj_engine() {
        for (hipEvent_t& x : ev)
            if (x) (void)hipEventDestroy(x);
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        if (captureStream) (void)hipStreamDestroy(captureStream);
    }
};

namespace vj {

namespace {

inline int64_t align_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }

// LDS row pitch of a cascade tile's integral footprint at a window step, and
// the footprint's size in elements.
inline int cascade_pitch(int step, int winW) { return ((1 << tile_lgcols_for_step(step)) - 1) * step + winW + 1; }
inline size_t cascade_footprint(int step, int rows, int winW, int winH) {
    return static_cast<size_t>(cascade_pitch(step, winW)) * static_cast<size_t>((rows - 1) * step + winH + 1);
}

// Pyramid schedule of one image, src/face_vision_engine.cu:1596-1714 (and the
// CPU loop src/main.cu:717-747): float scale, round(W/scale), skip levels
// whose object size is under minObjectSize, stop below the window.
struct LevelSpec {
    int scaleIdx, w, h, step;
    float scale;
};

std::vector<LevelSpec> schedule(int W, int H, int winW, int winH, float sf, int minObj) {
    std::vector<LevelSpec> out;
    float scale = 1.0f;
    int scaleIdx = 0;
    for (;;) {
        const int sw = static_cast<int>(std::round(static_cast<float>(W) / scale));
        const int sh = static_cast<int>(std::round(static_cast<float>(H) / scale));
        const int ow = static_cast<int>(std::round(static_cast<float>(winW) * scale));
        const int oh = static_cast<int>(std::round(static_cast<float>(winH) * scale));
        if (sw < winW || sh < winH) break;
        if (ow < minObj || oh < minObj) {
            scale *= sf;
            ++scaleIdx;
            continue;
        }
        out.push_back({scaleIdx, sw, sh, (scale >= 3.0f) ? 3 : ((scale >= 2.0f) ? 2 : 1), scale});
        scale *= sf;
        ++scaleIdx;
        if (scaleIdx > 100000) break;
    }
    return out;
}

struct ImageIn {
    const uint8_t* pix;
    int w, h, stride;
};

// Tile grid of a resize group: the largest tile side (on the group's last
// level) whose every tile's level regions fit in LDS and that gives at least
// kResizeTiles tiles, else the smallest side that fits (most tiles: the
// small trailing levels have little work per launch, so parallelism counts
// more than the recomputed margins).  The regions are computed with the
// kernel's own arithmetic.
constexpr int kResizeTiles = 256;
constexpr int64_t kResizeBigPixels = int64_t{1} << 19;
void resize_grid(const Plan& P, ResizeJob& jb) {
    int W[kResizeDepth + 1], H[kResizeDepth + 1];
    float sx[kResizeDepth + 1], sy[kResizeDepth + 1];
    const LevelDesc& s0 = P.levels[static_cast<size_t>(jb.src)];
    W[0] = s0.w;
    H[0] = s0.h;
    sx[0] = sy[0] = 1.0f;
    for (int j = 1; j <= kResizeDepth; ++j) {
        const int li = jb.lv[j <= jb.n ? j - 1 : 0];
        W[j] = P.levels[static_cast<size_t>(li)].w;
        H[j] = P.levels[static_cast<size_t>(li)].h;
        sx[j] = P.rs[2 * static_cast<size_t>(li)];
        sy[j] = P.rs[2 * static_cast<size_t>(li) + 1];
    }
    // widest region of every level over the tiles of one axis
    auto widest = [&](const int* len, const float* sc, int g, int* out) {
        for (int j = 0; j <= kResizeDepth; ++j) out[j] = 0;
        AxisRanges r;
        for (int i = 0; i < g; ++i) {
            axis_ranges(jb.n, len, sc, g, i, r);
            for (int j = 1; j <= jb.n; ++j) out[j] = std::max(out[j], r.need1[j] - r.need0[j]);
        }
    };
    static constexpr int kSides[] = {64, 48, 32, 24, 16, 12, 8, 6, 4, 3, 2, 1};
    const int Wn = W[jb.n], Hn = H[jb.n];
    int pick = -1;
    for (int side : kSides) {
        const int gx = (Wn + side - 1) / side, gy = (Hn + side - 1) / side;
        int wx[kResizeDepth + 1], wy[kResizeDepth + 1];
        widest(W, sx, gx, wx);
        widest(H, sy, gy, wy);
        bool fits = true;
        for (int j = 1; j <= jb.n; ++j) fits = fits && static_cast<int64_t>(wx[j]) * wy[j] <= kResizeCap;
        if (!fits) continue;
        pick = side;
        if (side <= 4 || static_cast<int64_t>(gx) * gy >= kResizeTiles) break;
    }
    // a one-pixel tile needs at most 2^n source pixels per axis, so one fits
    jb.gx = (Wn + pick - 1) / pick;
    jb.gy = (Hn + pick - 1) / pick;
}

// Builds the level table for a set of images.  Level 0 of an image is the
// input itself (stride as given); each further level is resized from the
// previous scanned level (or from the input when level 0 is skipped).
// The first image whose level 0 is skipped still needs its input as a
// source, so such images get an unscanned "source" level entry.
int build_plan(const vj_engine* e, const std::vector<ImageIn>& imgs, float sf, int minObj, Plan& P) {
    P = Plan();
    const int winW = e->winW, winH = e->winH;
    int64_t pyrOff = 0;
    P.imgWordOff.push_back(0);
    int maxDepth = 0;
    struct Pending {
        int level, depth;
    };
    std::vector<Pending> resized;
    auto push_level = [&P](const LevelDesc& d, int src, float rsx, float rsy, bool inPyr) {
        P.levels.push_back(d);
        P.rs.push_back(rsx);
        P.rs.push_back(rsy);
        const int idx = static_cast<int>(P.levels.size()) - 1;
        P.src.push_back(src < 0 ? idx : src);
        P.inPyr.push_back(inPyr ? 1 : 0);
        return idx;
    };
    for (size_t im = 0; im < imgs.size(); ++im) {
        const ImageIn& in = imgs[im];
        const std::vector<LevelSpec> sch = schedule(in.w, in.h, winW, winH, sf, minObj);
        P.imgLevels.push_back(static_cast<int32_t>(sch.size()));
        int prev = -1;  // previous scanned level of this image (its resize source)
        int depth = 0;
        for (const LevelSpec& ls : sch) {
            LevelDesc d{};
            d.w = ls.w;
            d.h = ls.h;
            d.step = ls.step;
            d.scale = ls.scale;
            d.scaleIdx = ls.scaleIdx;
            d.image = static_cast<int32_t>(im);
            if (ls.scaleIdx == 0) {
                d.pix = in.pix;
                d.stride = in.stride;
                prev = push_level(d, -1, 1.0f, 1.0f, false);
                continue;
            }
            if (prev < 0) {
                // level 0 was skipped: the input itself is the first source,
                // as an entry that is resized from but never scanned (step 0)
                LevelDesc s0{};
                s0.pix = in.pix;
                s0.w = in.w;
                s0.h = in.h;
                s0.stride = in.stride;
                s0.image = static_cast<int32_t>(im);
                prev = push_level(s0, -1, 1.0f, 1.0f, false);
            }
            const LevelDesc& sl = P.levels[static_cast<size_t>(prev)];
            const float rsx = static_cast<float>(sl.w) / static_cast<float>(ls.w);
            const float rsy = static_cast<float>(sl.h) / static_cast<float>(ls.h);
            d.stride = static_cast<int32_t>(align_up(ls.w, 64));
            d.pix = nullptr;
            P.pyrOffs.push_back(pyrOff);
            pyrOff += align_up(static_cast<int64_t>(d.stride) * ls.h, 256);
            ++depth;
            const int idx = push_level(d, prev, rsx, rsy, true);
            resized.push_back({idx, depth});
            maxDepth = std::max(maxDepth, depth);
            prev = idx;
        }
    }
    // per-level work layout (scanned levels only; source-only entries get empty ranges)
    int64_t iiOff = 0, colOff = 0, totOff = 0, rowOff = 0, wordOff = 0;
    int32_t items = 0, chunkItems = 0;
    int curImg = 0;
    for (size_t li = 0; li < P.levels.size(); ++li) {
        LevelDesc& d = P.levels[li];
        while (curImg < d.image) {
            P.imgWordOff.push_back(wordOff);
            ++curImg;
        }
        d.itemOff = items;
        d.chunkItemOff = chunkItems;
        d.iiOff = iiOff;
        d.colOff = colOff;
        d.totOff = totOff;
        d.rowOff = rowOff;
        d.wordOff = wordOff;
        d.tileOff = 0;
        if (d.step == 0) {  // source-only entry
            d.nBands = d.nChunks = 0;
            d.scanW = d.scanH = d.rowWords = 0;
            d.tilesX = d.tilesY = 0;
            d.tileRows = 1;
            d.tileLgCols = 6;
            P.itemOffs.push_back(items);
            P.chunkOffs.push_back(chunkItems);
            P.lvWordOff.push_back(wordOff);
            continue;
        }
        d.nBands = (d.h + kBandRows - 1) / kBandRows;
        d.nChunks = (d.w + kChunkCols - 1) / kChunkCols;
        d.iiPitch = static_cast<int32_t>(align_up(d.w + 1, kIIPitchAlign));
        d.scanW = (d.w - e->winW + 1 + d.step - 1) / d.step;
        d.scanH = (d.h - e->winH + 1 + d.step - 1) / d.step;
        d.rowWords = (d.scanW + 63) / 64;
        d.tileRows = e->tileRows[d.step];
        d.tileLgCols = tile_lgcols_for_step(d.step);
        d.tilesX = d.rowWords << (6 - d.tileLgCols);  // whole mask words
        d.tilesY = (d.scanH + d.tileRows - 1) / d.tileRows;
        P.itemOffs.push_back(items);
        P.chunkOffs.push_back(chunkItems);
        P.lvWordOff.push_back(wordOff);
        items += d.nBands * d.nChunks;
        chunkItems += d.nChunks;
        iiOff += align_up(static_cast<int64_t>(d.iiPitch) * (d.h + 1), 64);
        colOff += static_cast<int64_t>(d.nBands) * d.w;
        totOff += static_cast<int64_t>(d.nBands) * d.nChunks;
        rowOff += static_cast<int64_t>(d.h) * d.nChunks;
        wordOff += static_cast<int64_t>(d.scanH) * d.rowWords;
        P.tileLvl.push_back(static_cast<int32_t>(li));
        P.tileOff.push_back(P.nTiles);
        d.tileOff = P.nTiles;
        P.nTiles += d.tilesX * d.tilesY;
        P.maxStep = std::max(P.maxStep, d.step);
        P.windows += static_cast<int64_t>(d.scanW) * d.scanH;
        P.pyramidPixels += static_cast<int64_t>(d.w) * d.h;
        P.integralBytes += static_cast<int64_t>(d.w) * d.h + 8ll * (d.w + 1) * (d.h + 1);  // u32 sum + u32 square sum
        if (d.scaleIdx != 0) P.integralBytes += static_cast<int64_t>(d.w) * d.h;  // resize write
        P.cascadeBytes += 8ll * (d.w + 1) * (d.h + 1);
    }
    while (curImg < static_cast<int>(imgs.size())) {
        P.imgWordOff.push_back(wordOff);
        ++curImg;
    }
    // resize reads: each resized level reads its source once
    for (const Pending& r : resized) {
        const LevelDesc& s = P.levels[static_cast<size_t>(P.src[static_cast<size_t>(r.level)])];
        P.integralBytes += static_cast<int64_t>(s.w) * s.h;
    }
    P.nItems = items;
    P.nChunkItems = chunkItems;
    P.pyrBytes = pyrOff;
    P.iiElems = iiOff;
    P.colElems = colOff;
    P.totElems = totOff;
    P.rowElems = rowOff;
    P.nWords = wordOff;
    // the pyramid's launches: a depth whose levels (over the call's images)
    // hold at least kResizeBigPixels pixels gets a launch of its own; runs of
    // smaller depths go into groups of up to kResizeDepth, a job per image
    std::vector<int64_t> depthPix(static_cast<size_t>(maxDepth) + 1, 0);
    std::vector<std::vector<int32_t>> atDepth(static_cast<size_t>(maxDepth) + 1);
    for (const Pending& r : resized) {  // listed image by image, depth by depth
        const LevelDesc& d = P.levels[static_cast<size_t>(r.level)];
        depthPix[static_cast<size_t>(r.depth)] += static_cast<int64_t>(d.w) * d.h;
        atDepth[static_cast<size_t>(r.depth)].push_back(r.level);
    }
    for (int dpt = 1; dpt <= maxDepth;) {
        Plan::ResizeStep st{};
        if (depthPix[static_cast<size_t>(dpt)] >= kResizeBigPixels) {
            st.group = false;
            st.first = static_cast<int32_t>(P.levelJobs.size());
            // rows per workgroup: the most (up to kResizeRows) that still
            // give kResizeTiles * 4 workgroups
            int64_t segRows = 0;
            for (int32_t li : atDepth[static_cast<size_t>(dpt)]) {
                const LevelDesc& d = P.levels[static_cast<size_t>(li)];
                segRows += static_cast<int64_t>((d.w + kResizeSeg - 1) / kResizeSeg) * d.h;
            }
            st.rows = kResizeRows;
            while (st.rows > 1 && segRows / st.rows < 4 * kResizeTiles) st.rows >>= 1;
            for (int32_t li : atDepth[static_cast<size_t>(dpt)]) {
                const LevelDesc& d = P.levels[static_cast<size_t>(li)];
                P.levelJobs.push_back({li, st.items});
                st.items += (d.w + kResizeSeg - 1) / kResizeSeg * ((d.h + st.rows - 1) / st.rows);
            }
            st.count = static_cast<int32_t>(P.levelJobs.size()) - st.first;
            ++dpt;
        } else {
            int end = dpt;  // depths [dpt, end)
            while (end <= maxDepth && end - dpt < kResizeDepth && depthPix[static_cast<size_t>(end)] < kResizeBigPixels)
                ++end;
            st.group = true;
            st.first = static_cast<int32_t>(P.groupJobs.size());
            // each image's levels at these depths, in chain order
            std::map<int32_t, ResizeJob> byImage;
            for (int g = dpt; g < end; ++g)
                for (int32_t li : atDepth[static_cast<size_t>(g)]) {
                    ResizeJob& jb = byImage[P.levels[static_cast<size_t>(li)].image];
                    if (jb.n == 0) jb.src = P.src[static_cast<size_t>(li)];
                    jb.lv[jb.n++] = li;
                }
            for (auto& kv : byImage) {
                ResizeJob& jb = kv.second;
                resize_grid(P, jb);
                jb.itemOff = st.items;
                st.items += jb.gx * jb.gy;
                P.groupJobs.push_back(jb);
            }
            st.count = static_cast<int32_t>(P.groupJobs.size()) - st.first;
            dpt = end;
        }
        P.resizeSteps.push_back(st);
    }
    return VJ_OK;
}

int check_image_limits(const vj_engine* e, int w, int h) {
    VJ_CHECK(w <= e->maxW && h <= e->maxH, "image larger than the engine's maxImageW x maxImageH");
    VJ_CHECK(static_cast<int64_t>(w) * h * 255 < (int64_t{1} << 31),
             "image too large for exact 32-bit integrals (w*h*255 must be < 2^31)");
    return VJ_OK;
}

template <typename T>
int upload(DevBuf<T>& dst, const std::vector<T>& v, PinnedBuf<uint8_t>& staging, size_t& stOff,
           hipStream_t s) {
    VJ_TRY(dst.ensure(v.size()));
    if (v.empty()) return VJ_OK;
    const size_t bytes = v.size() * sizeof(T);
    std::memcpy(staging.get() + stOff, v.data(), bytes);
    VJ_HIP(hipMemcpyAsync(dst.get(), staging.get() + stOff, bytes, hipMemcpyHostToDevice, s));
    stOff += align_up(static_cast<int64_t>(bytes), 256);
    return VJ_OK;
}

// Runs the whole detect pipeline for the images of a plan; leaves per-image
// counts in e->dCounts and detections (capped per image) in e->dOut.
int run_pipeline(vj_engine* e, Plan& P, int cap, hipStream_t s, bool tablesCurrent) {
    // buffers
    VJ_TRY(e->pyr.ensure(static_cast<size_t>(std::max<int64_t>(P.pyrBytes, 1))));
    {
        size_t k = 0;
        for (size_t li = 0; li < P.levels.size(); ++li)
            if (P.inPyr[li]) P.levels[li].pix = e->pyr.get() + P.pyrOffs[k++];
    }
    VJ_TRY(e->iiS.ensure(static_cast<size_t>(P.iiElems)));
    VJ_TRY(e->iiQ.ensure(static_cast<size_t>(P.iiElems)));
    VJ_TRY(e->colS.ensure(static_cast<size_t>(P.colElems)));
    VJ_TRY(e->colQ.ensure(static_cast<size_t>(P.colElems)));
    VJ_TRY(e->totS.ensure(static_cast<size_t>(P.totElems)));
    VJ_TRY(e->totQ.ensure(static_cast<size_t>(P.totElems)));
    VJ_TRY(e->rowS.ensure(static_cast<size_t>(P.rowElems)));
    VJ_TRY(e->rowQ.ensure(static_cast<size_t>(P.rowElems)));
    VJ_TRY(e->words.ensure(static_cast<size_t>(P.nWords)));
    VJ_TRY(e->scores.ensure(static_cast<size_t>(P.nWords) * 64));
    const bool deep = e->stageSplit < e->nStages && P.nTiles > 0;
    if (deep) {
        VJ_TRY(e->survIdx.ensure(static_cast<size_t>(std::max<int64_t>(P.windows, 1))));
        VJ_TRY(e->survInv.ensure(static_cast<size_t>(std::max<int64_t>(P.windows, 1))));
        VJ_TRY(e->tileSurvOff.ensure(static_cast<size_t>(P.nTiles)));
        VJ_TRY(e->tileCount.ensure(static_cast<size_t>(P.nTiles)));
        VJ_TRY(e->deepItems.ensure(static_cast<size_t>(P.nTiles) * kDeepSlices));
        VJ_TRY(e->deepCtl.ensure(kDeepCtlInts));
        if (e->stageTail < e->nStages)
            VJ_TRY(e->tailItems.ensure(static_cast<size_t>(P.nTiles) * kDeepSlices * kTailPerDeep));
    }
    const int nImages = static_cast<int>(P.imgWordOff.size()) - 1;
    VJ_TRY(e->dCounts.ensure(static_cast<size_t>(nImages)));
    VJ_TRY(e->dOut.ensure(static_cast<size_t>(cap) * nImages));

    // The call's tables (levels, scales, offsets, jobs, compaction blocks) in
    // one staging buffer and one copy, unless the engine's plan buffer already
    // holds this very plan (the previous call's, with the same images).
    if (!tablesCurrent) {
        PlanTables& T = P.tab;
        std::vector<int32_t> meta32;
        T.oItem = meta32.size();
        meta32.insert(meta32.end(), P.itemOffs.begin(), P.itemOffs.end());
        T.oChunk = meta32.size();
        meta32.insert(meta32.end(), P.chunkOffs.begin(), P.chunkOffs.end());
        T.oSrc = meta32.size();
        meta32.insert(meta32.end(), P.src.begin(), P.src.end());
        T.oLvl = meta32.size();
        meta32.insert(meta32.end(), P.tileLvl.begin(), P.tileLvl.end());
        T.oTile = meta32.size();
        meta32.insert(meta32.end(), P.tileOff.begin(), P.tileOff.end());
        std::vector<int64_t> meta64;
        T.oLvWord = meta64.size();
        meta64.insert(meta64.end(), P.lvWordOff.begin(), P.lvWordOff.end());
        // compaction blocks: each image's words cut into runs of kWordBlock
        std::vector<int64_t> blockWord;
        std::vector<int32_t> blockImg, imgBlock;
        for (int im = 0; im < nImages; ++im) {
            imgBlock.push_back(static_cast<int32_t>(blockImg.size()));
            for (int64_t a = P.imgWordOff[static_cast<size_t>(im)]; a < P.imgWordOff[static_cast<size_t>(im) + 1];
                 a += kWordBlock) {
                blockWord.push_back(a);
                blockImg.push_back(im);
            }
        }
        imgBlock.push_back(static_cast<int32_t>(blockImg.size()));
        blockWord.push_back(P.imgWordOff.back());
        T.oBlockWord = meta64.size();
        meta64.insert(meta64.end(), blockWord.begin(), blockWord.end());
        T.oBlockImg = meta32.size();
        meta32.insert(meta32.end(), blockImg.begin(), blockImg.end());
        T.oImgBlock = meta32.size();
        meta32.insert(meta32.end(), imgBlock.begin(), imgBlock.end());
        T.nBlocks = static_cast<int>(blockImg.size());
        const size_t stageBytes = P.levels.size() * sizeof(LevelDesc) + P.rs.size() * 4 + meta32.size() * 4 +
                                  meta64.size() * 8 + P.levelJobs.size() * sizeof(LevelJob) +
                                  P.groupJobs.size() * sizeof(ResizeJob) + 7 * 256;
        VJ_TRY(e->hStage.ensure(stageBytes));
        VJ_HIP(hipStreamSynchronize(s));  // staging buffer reuse across calls
        size_t stOff = 0;
        auto place = [&](const void* src, size_t bytes) {
            const size_t at = stOff;
            if (bytes) std::memcpy(e->hStage.get() + at, src, bytes);
            stOff += static_cast<size_t>(align_up(static_cast<int64_t>(bytes), 256));
            return at;
        };
        T.aLv = place(P.levels.data(), P.levels.size() * sizeof(LevelDesc));
        T.aRs = place(P.rs.data(), P.rs.size() * sizeof(float));
        T.a32 = place(meta32.data(), meta32.size() * sizeof(int32_t));
        T.a64 = place(meta64.data(), meta64.size() * sizeof(int64_t));
        T.aLJobs = place(P.levelJobs.data(), P.levelJobs.size() * sizeof(LevelJob));
        T.aGJobs = place(P.groupJobs.data(), P.groupJobs.size() * sizeof(ResizeJob));
        VJ_TRY(e->dPlan.ensure(std::max<size_t>(stOff, 256)));
        VJ_HIP(hipMemcpyAsync(e->dPlan.get(), e->hStage.get(), stOff, hipMemcpyHostToDevice, s));
    }
    VJ_TRY(e->dBlockCount.ensure(static_cast<size_t>(P.tab.nBlocks)));
    VJ_TRY(e->dBlockOff.ensure(static_cast<size_t>(P.tab.nBlocks)));
    VJ_TRY(e->dImgBase.ensure(static_cast<size_t>(nImages)));
    const size_t oItem = P.tab.oItem, oChunk = P.tab.oChunk, oSrc = P.tab.oSrc, oLvl = P.tab.oLvl,
                 oTile = P.tab.oTile, oLvWord = P.tab.oLvWord, oBlockWord = P.tab.oBlockWord,
                 oBlockImg = P.tab.oBlockImg, oImgBlock = P.tab.oImgBlock;
    const size_t aLv = P.tab.aLv, aRs = P.tab.aRs, a32 = P.tab.a32, a64 = P.tab.a64, aLJobs = P.tab.aLJobs,
                 aGJobs = P.tab.aGJobs;

    const LevelDesc* dL = reinterpret_cast<const LevelDesc*>(e->dPlan.get() + aLv);
    const float* dRs = reinterpret_cast<const float*>(e->dPlan.get() + aRs);
    const int32_t* m32 = reinterpret_cast<const int32_t*>(e->dPlan.get() + a32);
    const int64_t* m64 = reinterpret_cast<const int64_t*>(e->dPlan.get() + a64);
    const LevelJob* dLJobs = reinterpret_cast<const LevelJob*>(e->dPlan.get() + aLJobs);
    const ResizeJob* dGJobs = reinterpret_cast<const ResizeJob*>(e->dPlan.get() + aGJobs);
    const int nLevels = static_cast<int>(P.levels.size());

    size_t ldsBytes = 0;
    for (int st = 1; st <= P.maxStep; ++st)
        ldsBytes = std::max(ldsBytes, cascade_footprint(st, e->tileRows[st], e->winW, e->winH) * sizeof(float));
    // deep and tail passes: the leaf values after the footprint, 16-byte aligned
    const int leafOff = static_cast<int>((ldsBytes / sizeof(float) + 3) & ~size_t(3));
    const size_t deepLds = (static_cast<size_t>(leafOff) + kLeafCap) * sizeof(float);
    // without a tail pass the deep pass runs every remaining stage at eight
    // workgroups per CU (smaller leaf cache, at most 64 VGPRs)
    const size_t deepLdsWide = (static_cast<size_t>(leafOff) + kLeafCapWide) * sizeof(float);
    const int nBlocks = P.tab.nBlocks;
    const int deepGrid = deep ? std::min(P.nTiles * kDeepSlices, kDeepGrid) : 0;
    const bool tail = deep && e->stageTail < e->nStages;
    const int stageTail = deep ? e->stageTail : e->nStages;

    // counting mode: the windows entering each stage
    unsigned long long* enter = nullptr;
    if (e->countStages) {
        VJ_TRY(e->stageEnter.ensure(static_cast<size_t>(e->nStages)));
        enter = e->stageEnter.get();
    }
    auto launch_all = [&](hipStream_t ls) -> int {
        if (enter) {
            hipLaunchKernelGGL(zero_u64_kernel, dim3(1), dim3(256), 0, ls, enter, e->nStages);
            VJ_HIP(hipGetLastError());
        }
        if (e->timing) VJ_HIP(hipEventRecord(e->ev[0], ls));
        // 1. pyramid: one launch per large depth or group of small depths
        //    (each needs the one before)
        for (const Plan::ResizeStep& st : P.resizeSteps) {
            if (!st.count || !st.items) continue;
            if (st.group)
                hipLaunchKernelGGL(resize_group_kernel, dim3(st.items), dim3(kResizeThreads), 0, ls, dL,
                                   dGJobs + st.first, st.count, dRs);
            else
                hipLaunchKernelGGL(resize_level_kernel, dim3(st.items), dim3(256), 0, ls, dL, dLJobs + st.first,
                                   st.count, dRs, m32 + oSrc, st.rows);
            VJ_HIP(hipGetLastError());
        }
        if (e->timing) VJ_HIP(hipEventRecord(e->ev[1], ls));
        // 2. integrals of every level
        if (P.nItems > 0) {
            hipLaunchKernelGGL(ii_colsum_kernel, dim3(P.nItems), dim3(64), 0, ls, dL, P.nItems, m32 + oItem, nLevels,
                               e->colS.get(), e->colQ.get(), e->rowS.get(), e->rowQ.get(), e->totS.get(),
                               e->totQ.get());
            VJ_HIP(hipGetLastError());
            hipLaunchKernelGGL(ii_vscan_kernel, dim3(P.nChunkItems), dim3(256), 0, ls, dL, m32 + oChunk, nLevels,
                               e->colS.get(), e->colQ.get(), e->totS.get(), e->totQ.get(),
                               deep ? e->deepCtl.get() : nullptr, kDeepCtlInts);
            VJ_HIP(hipGetLastError());
            hipLaunchKernelGGL(ii_write_kernel<uint32_t>, dim3(P.nItems), dim3(64), 0, ls, dL, P.nItems, m32 + oItem, nLevels,
                               e->colS.get(), e->colQ.get(), e->totS.get(), e->totQ.get(), e->rowS.get(),
                               e->rowQ.get(), e->iiS.get(), e->iiQ.get());
            VJ_HIP(hipGetLastError());
        }
        // 3. cascade: every tile of every level in one launch; LDS sized for the
        //    largest window step present (its tile footprint is the widest)
        if (e->timing) VJ_HIP(hipEventRecord(e->ev[2], ls));
        if (P.nTiles > 0) {
            hipLaunchKernelGGL(enter ? cascade_kernel<true> : cascade_kernel<false>, dim3(P.nTiles), dim3(kCascadeBlock), ldsBytes, ls, dL, m32 + oLvl,
                               m32 + oTile, static_cast<int>(P.tileLvl.size()), P.nTiles, e->iiS.get(), e->iiQ.get(),
                               e->kst.get(), e->nStumps, e->kstage.get(), e->nStages, e->stageSplit, e->winW, e->winH,
                               e->words.get(), e->scores.get(), e->deepCtl.get(), e->survIdx.get(), e->survInv.get(),
                               e->tileSurvOff.get(), e->tileCount.get(), enter);
            VJ_HIP(hipGetLastError());
            if (deep) {
                const int listGrid = std::min((P.nTiles + kListThreads - 1) / kListThreads, 1024);
                hipLaunchKernelGGL(deep_list_kernel, dim3(listGrid), dim3(kListThreads), 0, ls, e->tileCount.get(),
                                   P.nTiles, e->deepCtl.get(), e->deepItems.get());
                VJ_HIP(hipGetLastError());
                hipLaunchKernelGGL(enter ? cascade_deep_kernel<true> : tail ? cascade_deep_kernel<false> : (cascade_deep_kernel<false, 8>),
                                   dim3(deepGrid), dim3(kCascadeBlock), enter || tail ? deepLds : deepLdsWide, ls, dL,
                                   m32 + oLvl, m32 + oTile, static_cast<int>(P.tileLvl.size()), e->iiS.get(),
                                   e->kst.get(), e->nStumps, e->kstage.get(), e->nStages, e->stageSplit, stageTail,
                                   e->winW, e->winH, leafOff, reinterpret_cast<unsigned long long*>(e->words.get()),
                                   e->scores.get(), e->survIdx.get(), e->survInv.get(), e->tileSurvOff.get(),
                                   e->tileCount.get(), e->deepItems.get(), e->deepCtl.get() + 1,
                                   e->deepCtl.get() + 2, e->tailItems.get(), e->deepCtl.get() + kTailCtl, enter);
                VJ_HIP(hipGetLastError());
            }
            if (tail) {
                hipLaunchKernelGGL(enter ? cascade_tail_kernel<true> : cascade_tail_kernel<false>, dim3(deepGrid), dim3(kCascadeBlock), deepLds, ls, dL,
                                   m32 + oLvl, m32 + oTile, static_cast<int>(P.tileLvl.size()), e->iiS.get(),
                                   e->kst.get(), e->nStumps, e->kstage.get(), e->nStages, stageTail, e->winW,
                                   e->winH, leafOff, reinterpret_cast<unsigned long long*>(e->words.get()), e->scores.get(),
                                   e->survIdx.get(), e->survInv.get(), e->tailItems.get(),
                                   e->deepCtl.get() + kTailCtl, e->deepCtl.get() + kTailCtl + 1, enter);
                VJ_HIP(hipGetLastError());
            }
        }
        if (e->timing) VJ_HIP(hipEventRecord(e->ev[3], ls));
        // 4. counts and order-preserving emission
        if (nBlocks > 0) {
            hipLaunchKernelGGL(block_count_kernel, dim3(nBlocks), dim3(256), 0, ls, m64 + oBlockWord, e->words.get(),
                               e->dBlockCount.get());
            VJ_HIP(hipGetLastError());
        }
        hipLaunchKernelGGL(block_scan_kernel, dim3(1), dim3(1024), 0, ls, m32 + oImgBlock, nImages, e->dBlockCount.get(),
                           e->dBlockOff.get(), e->dCounts.get(), e->dImgBase.get(), cap);
        VJ_HIP(hipGetLastError());
        if (nBlocks > 0) {
            hipLaunchKernelGGL(emit_kernel, dim3(nBlocks), dim3(256), 0, ls, dL, m64 + oLvWord, nLevels, m64 + oBlockWord,
                               m32 + oBlockImg, e->words.get(), e->scores.get(), e->dBlockOff.get(), e->dImgBase.get(),
                               cap, e->winW, e->winH, e->dOut.get());
            VJ_HIP(hipGetLastError());
        }
        if (e->timing) VJ_HIP(hipEventRecord(e->ev[4], ls));
        return VJ_OK;
    };
    if (e->timing || e->countStages || !e->graphs) return launch_all(s);

    // Graph replay: the captured kernels read their per-call tables (levels,
    // jobs, offsets) from the device buffers uploaded above, so the graph is
    // valid while every grid size, scalar argument and buffer address matches.
    std::vector<int64_t> key;
    auto k = [&key](int64_t v) { key.push_back(v); };
    for (const Plan::ResizeStep& st : P.resizeSteps) {
        k(st.group ? 1 : 0);
        k(st.first);
        k(st.count);
        k(st.items);
        k(st.rows);
    }
    for (int64_t v : {int64_t(e->stageSplit), int64_t(stageTail), int64_t(deepGrid), int64_t(nLevels), int64_t(P.nItems), int64_t(P.nChunkItems), int64_t(P.nTiles),
                      int64_t(P.tileLvl.size()), int64_t(ldsBytes), int64_t(nBlocks), int64_t(nImages), int64_t(cap),
                      int64_t(e->nStumps), int64_t(e->nStages), int64_t(e->winW), int64_t(e->winH),
                      int64_t(oSrc), int64_t(oItem), int64_t(oChunk), int64_t(oLvl), int64_t(oTile), int64_t(oLvWord),
                      int64_t(oBlockWord), int64_t(oBlockImg), int64_t(oImgBlock)})
        k(v);
    for (const void* ptr : {static_cast<const void*>(dL), static_cast<const void*>(m32), static_cast<const void*>(m64),
                            static_cast<const void*>(dLJobs), static_cast<const void*>(dGJobs),
                            static_cast<const void*>(dRs),
                            static_cast<const void*>(e->colS.get()), static_cast<const void*>(e->colQ.get()),
                            static_cast<const void*>(e->rowS.get()), static_cast<const void*>(e->rowQ.get()),
                            static_cast<const void*>(e->totS.get()), static_cast<const void*>(e->totQ.get()),
                            static_cast<const void*>(e->iiS.get()), static_cast<const void*>(e->iiQ.get()),
                            static_cast<const void*>(e->kst.get()), static_cast<const void*>(e->kstage.get()),
                            static_cast<const void*>(e->words.get()), static_cast<const void*>(e->scores.get()),
                            static_cast<const void*>(e->dBlockCount.get()), static_cast<const void*>(e->dBlockOff.get()),
                            static_cast<const void*>(e->dCounts.get()), static_cast<const void*>(e->dImgBase.get()),
                            static_cast<const void*>(e->dOut.get()), static_cast<const void*>(e->survIdx.get()),
                            static_cast<const void*>(e->survInv.get()), static_cast<const void*>(e->tileSurvOff.get()),
                            static_cast<const void*>(e->tileCount.get()), static_cast<const void*>(e->deepItems.get()),
                            static_cast<const void*>(e->deepCtl.get()), static_cast<const void*>(e->tailItems.get())})
        k(static_cast<int64_t>(reinterpret_cast<uintptr_t>(ptr)));
    if (e->graphExec && key == e->graphKey) {
        VJ_HIP(hipGraphLaunch(e->graphExec, s));
        return VJ_OK;
    }
    if (key != e->lastKey) {
        // first call of this shape: launch directly, capture if it repeats
        e->lastKey = std::move(key);
        return launch_all(s);
    }
    if (e->graphExec) {
        (void)hipGraphExecDestroy(e->graphExec);
        e->graphExec = nullptr;
        e->graphKey.clear();
    }
    // capture on the engine's own stream (the null stream cannot capture);
    // capturing runs nothing, so only the replay is ordered on s
    if (!e->captureStream) VJ_HIP(hipStreamCreateWithFlags(&e->captureStream, hipStreamNonBlocking));
    hipGraph_t g = nullptr;
    VJ_HIP(hipStreamBeginCapture(e->captureStream, hipStreamCaptureModeThreadLocal));
    const int st = launch_all(e->captureStream);
    const hipError_t ce = hipStreamEndCapture(e->captureStream, &g);
    if (st != VJ_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ce != hipSuccess) return hip_fail(ce, "hipStreamEndCapture", __FILE__, __LINE__);
    hipGraphExec_t ex = nullptr;
    const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) return hip_fail(ie, "hipGraphInstantiate", __FILE__, __LINE__);
    e->graphExec = ex;
    e->graphKey = std::move(key);
    VJ_HIP(hipGraphLaunch(ex, s));
    return VJ_OK;
}

}  // namespace
}  // namespace vj

// ===========================================================================
// C ABI
// ===========================================================================

extern "C" int vj_engine_create(int winW, int winH, int maxSamples, int maxImageW, int maxImageH,
                                int enableTilted, int maxDetections, vj_engine** out) {
    // src/face_vision_engine.cu:659-661: every size must be positive.
    VJ_CHECK(out, "vj_engine_create: null output");
    *out = nullptr;
    VJ_CHECK(winW >= 3 && winH >= 3 && winW <= 64 && winH <= 64, "vj_engine_create: window must be 3..64");
    VJ_CHECK(maxSamples > 0 && maxImageW > 0 && maxImageH > 0 && maxDetections > 0,
             "vj_engine_create: invalid constructor args");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("vj_engine_create: no HIP device");
        return VJ_DEVICE_ERROR;
    }
    {
        // the sweep kernel's x bins (at most kSweepMaxBins) and LDS buckets
        // (kSweepLdsBoxes) take about 128 KB of LDS
        static const hipError_t ldsAttr = hipFuncSetAttribute(
            reinterpret_cast<const void*>(pair_sweep_kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
            static_cast<int>(sweep_lds_bytes(kSweepMaxBins - 1, kSweepLdsBoxes)));
        if (ldsAttr != hipSuccess) {
            (void)hipGetLastError();
            set_error("vj_engine_create: cannot raise the grouping kernel's LDS limit");
            return VJ_DEVICE_ERROR;
        }
    }
    vj_engine* e = new (std::nothrow) vj_engine();
    if (!e) {
        set_error("vj_engine_create: out of host memory");
        return VJ_OUT_OF_MEMORY;
    }
    e->winW = winW;
    e->winH = winH;
    e->maxSamples = maxSamples;
    e->maxW = maxImageW;
    e->maxH = maxImageH;
    e->tilted = enableTilted != 0;
    e->maxDet = maxDetections;
    int st = e->hCounts.ensure(1);
    if (st == VJ_OK) st = e->hOut.ensure(static_cast<size_t>(std::min(maxDetections, 1 << 16)));
    if (st != VJ_OK) {
        delete e;
        return st;
    }
    *out = e;
    return VJ_OK;
}

extern "C" void vj_engine_destroy(vj_engine* e) { delete e; }

extern "C" int vj_engine_upload_feature_pool(vj_engine* e, const vj_feature* hFeatures, int n, vj_stream s) {
    VJ_CHECK(e && hFeatures && n > 0, "uploadFeaturePool: empty pool");
    VJ_TRY(e->pool.ensure(static_cast<size_t>(n)));
    VJ_HIP(hipMemcpyAsync(e->pool.get(), hFeatures, sizeof(vj_feature) * static_cast<size_t>(n),
                          hipMemcpyHostToDevice, static_cast<hipStream_t>(s)));
    VJ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(s)));
    e->poolN = n;
    return VJ_OK;
}

extern "C" int vj_engine_feature_pool(vj_engine* e, const vj_feature** d, int* n) {
    VJ_CHECK(e && d && n, "vj_engine_feature_pool: null");
    *d = e->pool.get();
    *n = e->poolN;
    return VJ_OK;
}

extern "C" int vj_engine_integral_batch(vj_engine* e, const uint8_t* dGray, int w, int h, int stride,
                                        int n, int computeInvNorm, vj_stream sv) {
    // src/face_vision_engine.cu:1064-1069
    VJ_CHECK(e && dGray && w > 0 && h > 0 && stride >= w && n > 0, "computeIntegralBatch: invalid arguments");
    VJ_CHECK(w <= e->maxW && h <= e->maxH && n <= e->maxSamples, "computeIntegralBatch: exceeds engine limits");
    hipStream_t s = static_cast<hipStream_t>(sv);
    const size_t area = static_cast<size_t>(w + 1) * (h + 1);
    VJ_TRY(e->bSum.ensure(area * n));
    VJ_TRY(e->bSq.ensure(area * n));
    VJ_TRY(e->bInv.ensure(static_cast<size_t>(n)));
    if (n == 1 && (w > 64 || h > 64)) {
        // a single image goes through the detect integral kernels (bands of
        // rows in parallel) straight into the view buffers
        VJ_TRY(check_image_limits(e, w, h));
        LevelDesc d{};
        d.pix = dGray;
        d.w = w;
        d.h = h;
        d.stride = stride;
        d.step = 1;
        d.nBands = (h + kBandRows - 1) / kBandRows;
        d.nChunks = (w + kChunkCols - 1) / kChunkCols;
        d.iiPitch = w + 1;  // the view's dense layout
        VJ_TRY(e->colS.ensure(static_cast<size_t>(d.nBands) * w));
        VJ_TRY(e->colQ.ensure(static_cast<size_t>(d.nBands) * w));
        VJ_TRY(e->totS.ensure(static_cast<size_t>(d.nBands) * d.nChunks));
        VJ_TRY(e->totQ.ensure(static_cast<size_t>(d.nBands) * d.nChunks));
        VJ_TRY(e->rowS.ensure(static_cast<size_t>(h) * d.nChunks));
        VJ_TRY(e->rowQ.ensure(static_cast<size_t>(h) * d.nChunks));
        VJ_TRY(e->hStage.ensure(sizeof(LevelDesc) + 4 * 256));
        VJ_HIP(hipStreamSynchronize(s));
        size_t stOff = 0;
        std::vector<LevelDesc> lv{d};
        std::vector<int32_t> offs{0, 0};
        VJ_TRY(upload(e->dLevels, lv, e->hStage, stOff, s));
        VJ_TRY(upload(e->dMeta32, offs, e->hStage, stOff, s));
        const int items = d.nBands * d.nChunks;
        hipLaunchKernelGGL(ii_colsum_kernel, dim3(items), dim3(64), 0, s, e->dLevels.get(), items, e->dMeta32.get(), 1,
                           e->colS.get(), e->colQ.get(), e->rowS.get(), e->rowQ.get(), e->totS.get(), e->totQ.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(ii_vscan_kernel, dim3(d.nChunks), dim3(256), 0, s, e->dLevels.get(),
                           e->dMeta32.get() + 1, 1, e->colS.get(), e->colQ.get(), e->totS.get(), e->totQ.get(),
                           nullptr, 0);
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(ii_write_kernel<uint64_t>, dim3(items), dim3(64), 0, s, e->dLevels.get(), items, e->dMeta32.get(), 1,
                           e->colS.get(), e->colQ.get(), e->totS.get(), e->totQ.get(), e->rowS.get(),
                           e->rowQ.get(), e->bSum.get(), e->bSq.get());
        VJ_HIP(hipGetLastError());
    } else {
        hipLaunchKernelGGL(batch_integral_kernel, dim3((n + 255) / 256), dim3(256), 0, s, dGray, w, h, stride, n,
                           e->bSum.get(), e->bSq.get());
        VJ_HIP(hipGetLastError());
    }
    if (computeInvNorm) {
        hipLaunchKernelGGL(invnorm_kernel, dim3((n + 255) / 256), dim3(256), 0, s, e->bSum.get(), e->bSq.get(),
                           w + 1, h + 1, n, e->bInv.get());
        VJ_HIP(hipGetLastError());
    }
    e->curN = n;
    e->curW = w;
    e->curH = h;
    return VJ_OK;
}

extern "C" int vj_engine_integral_view(vj_engine* e, vj_integral_view* out) {
    VJ_CHECK(e && out, "integralSetView: null");
    out->nSamples = e->curN;
    out->iiWidth = e->curW + 1;
    out->iiHeight = e->curH + 1;
    out->iiArea = out->iiWidth * out->iiHeight;
    out->sumT = e->bSum.get();
    out->sqsumT = e->bSq.get();
    out->invNorm = e->bInv.get();
    return VJ_OK;
}

// The call's small results in one launch instead of four copies: the
// grouping's control words, each image's cluster counts and flags, the raw
// counts and the first grouped boxes (as many as the images' kept clusters,
// at most nOut), stored straight into the caller-visible pinned buffers
// (device addresses of host memory; visible to the host once the stream is
// synchronised).
__global__ __launch_bounds__(256) void publish_kernel(const int64_t* __restrict__ ctl,
                                                      const int32_t* __restrict__ grpImg,
                                                      const int32_t* __restrict__ counts,
                                                      const vj_detection* __restrict__ outG, int nImages, int nOut,
                                                      int64_t* __restrict__ hCtl, int32_t* __restrict__ hGrpImg,
                                                      int32_t* __restrict__ hCounts, vj_detection* __restrict__ hOutG) {
    __shared__ int32_t part[4];
    const int tid = static_cast<int>(threadIdx.x);
    if (tid < 2) hCtl[tid] = ctl[tid];
    int32_t kept = 0;
    for (int i = tid; i < nImages; i += 256) {
        hGrpImg[3 * i] = grpImg[3 * i];
        hGrpImg[3 * i + 1] = grpImg[3 * i + 1];
        hGrpImg[3 * i + 2] = grpImg[3 * i + 2];
        hCounts[i] = counts[i];
        kept += max(grpImg[3 * i], 0);
    }
    kept = wave_sum(kept);
    if ((tid & (kWave - 1)) == 0) part[tid / kWave] = kept;
    __syncthreads();
    const int64_t total = static_cast<int64_t>(part[0]) + part[1] + part[2] + part[3];
    const int m = static_cast<int>(total < nOut ? total : static_cast<int64_t>(nOut));
    for (int k = tid; k < m; k += 256) hOutG[k] = outG[k];
}

// ---- one image's integral without allocating ------------------------------
// Scratch of vj_integral_image_noalloc: the level descriptor and item table
// of the integral kernels, then the column band sums, row-chunk totals and
// band totals of one image.  Each array is sized for the largest image the
// integral kernels accept (sides up to 8192, w * h * 255 < 2^31): at most
// 270,878 column sums, 40,960 row-chunk totals and 1,280 band totals.  Every
// element the kernels read was written earlier in the same call, so nothing
// needs clearing.  One copy per GPU, held until the process ends.
constexpr int64_t kNaCols = 270878, kNaRows = 40960, kNaTots = 1280;
constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t kNaDesc = 0;
constexpr size_t kNaOffs = na_align(sizeof(LevelDesc));
constexpr size_t kNaColS = kNaOffs + 256;
constexpr size_t kNaColQ = kNaColS + na_align(4 * kNaCols);
constexpr size_t kNaRowS = kNaColQ + na_align(4 * kNaCols);
constexpr size_t kNaRowQ = kNaRowS + na_align(4 * kNaRows);
constexpr size_t kNaTotS = kNaRowQ + na_align(4 * kNaRows);
constexpr size_t kNaTotQ = kNaTotS + na_align(4 * kNaTots);
constexpr size_t kNaBytes = kNaTotQ + na_align(8 * kNaTots);
static_assert(kNaBytes < (size_t{3} << 20), "integral scratch stays under 3 MiB");

static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Writes the level descriptor and the one-level item table into the scratch
// (the kernels read them from device memory; no host copy is recorded).
__global__ __launch_bounds__(64) void na_setup_kernel(LevelDesc d, unsigned char* base) {
    if (threadIdx.x == 0) {
        *reinterpret_cast<LevelDesc*>(base + kNaDesc) = d;
        int32_t* offs = reinterpret_cast<int32_t*>(base + kNaOffs);
        offs[0] = 0;
        offs[1] = 0;
    }
}

namespace {
// The scratch's device address on each GPU, looked up once per device with
// that device current (host code never passes the array's own name).
constexpr int kNaMaxDevices = 64;
std::mutex gNaMutex;
std::atomic<unsigned char*> gNaPtr[kNaMaxDevices];

int na_scratch(unsigned char** out) {
    int dev = 0;
    VJ_HIP(hipGetDevice(&dev));
    VJ_CHECK(dev >= 0 && dev < kNaMaxDevices, "integral_image_noalloc: device index above 63");
    unsigned char* p = gNaPtr[dev].load(std::memory_order_acquire);
    if (!p) {
        std::lock_guard<std::mutex> lock(gNaMutex);
        p = gNaPtr[dev].load(std::memory_order_relaxed);
        if (!p) {
            void* a = nullptr;
            VJ_HIP(hipGetSymbolAddress(&a, HIP_SYMBOL(scratch)));
            VJ_CHECK(a != nullptr, "integral_image_noalloc: no device address for the scratch");
            p = static_cast<unsigned char*>(a);
            gNaPtr[dev].store(p, std::memory_order_release);
        }
    }
    *out = p;
    return VJ_OK;
}
}  // namespace

extern "C" int vj_integral_image_noalloc(const uint8_t* dGray, int w, int h, int stride, uint32_t* dSum,
                                         uint64_t* dSq, vj_stream sv) {
    VJ_CHECK(dGray && dSum && dSq && w > 0 && h > 0 && stride >= w, "integral_image_noalloc: invalid arguments");
    VJ_CHECK(w <= 8192 && h <= 8192, "integral_image_noalloc: sides above 8192 exceed the scratch");
    VJ_CHECK(static_cast<int64_t>(w) * h * 255 < (int64_t{1} << 31),
             "integral_image_noalloc: image too large for exact 32-bit integrals (w*h*255 must be < 2^31)");
    LevelDesc d{};
    d.pix = dGray;
    d.w = w;
    d.h = h;
    d.stride = stride;
    d.step = 1;
    d.nBands = (h + kBandRows - 1) / kBandRows;
    d.nChunks = (w + kChunkCols - 1) / kChunkCols;
    d.iiPitch = w + 1;  // the caller's dense (h+1) x (w+1) layout
    VJ_CHECK(static_cast<int64_t>(d.nBands) * w <= kNaCols && static_cast<int64_t>(h) * d.nChunks <= kNaRows &&
                 static_cast<int64_t>(d.nBands) * d.nChunks <= kNaTots,
             "integral_image_noalloc: image exceeds the scratch");
    unsigned char* base = nullptr;
    VJ_TRY(na_scratch(&base));
    hipStream_t s = static_cast<hipStream_t>(sv);
    const LevelDesc* L = reinterpret_cast<const LevelDesc*>(base + kNaDesc);
    const int32_t* offs = reinterpret_cast<const int32_t*>(base + kNaOffs);
    uint32_t* colS = reinterpret_cast<uint32_t*>(base + kNaColS);
    uint32_t* colQ = reinterpret_cast<uint32_t*>(base + kNaColQ);
    uint32_t* rowS = reinterpret_cast<uint32_t*>(base + kNaRowS);
    uint32_t* rowQ = reinterpret_cast<uint32_t*>(base + kNaRowQ);
    uint32_t* totS = reinterpret_cast<uint32_t*>(base + kNaTotS);
    uint64_t* totQ = reinterpret_cast<uint64_t*>(base + kNaTotQ);
    const int items = d.nBands * d.nChunks;
    hipLaunchKernelGGL(na_setup_kernel, dim3(1), dim3(64), 0, s, d, base);
    VJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(ii_colsum_kernel, dim3(items), dim3(64), 0, s, L, items, offs, 1, colS, colQ, rowS, rowQ,
                       totS, totQ);
    VJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(ii_vscan_kernel, dim3(d.nChunks), dim3(256), 0, s, L, offs + 1, 1, colS, colQ, totS, totQ,
                       nullptr, 0);
    VJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(ii_write_kernel<uint64_t>, dim3(items), dim3(64), 0, s, L, items, offs, 1, colS, colQ, totS, totQ, rowS,
                       rowQ, dSum, dSq);
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

// The pair mode's decision margins of one stage.  Any float summation of n
// terms x_i, in any order, is within gamma_m * sum|x_i| of the exact sum
// (gamma_m = m u / (1 - m u), u = 2^-24, m >= the number of roundings any
// term goes through); the pair mode's sum and the sequential one are two such
// summations of the same chosen leaf values, |x_i| <= max(|left_i|,
// |right_i|).  So with B = 2 gamma_m sum_i max(|left_i|, |right_i|) and m =
// n + 16 (room for the additions of zero partial sums), lo is a float at or
// below threshold - B and hi a float at or above threshold + B.
void stage_margins(const vj_stump* st, int n, float threshold, float& lo, float& hi) {
    double M = 0.0;
    for (int k = 0; k < n; ++k) M += std::max(std::fabs(static_cast<double>(st[k].left)), std::fabs(static_cast<double>(st[k].right)));
    const double u = std::ldexp(1.0, -24);
    const double m = static_cast<double>(n) + 16.0;
    const double gamma = m * u / (1.0 - m * u);
    const double B = n > 1 ? 2.0 * gamma * M * (1.0 + 1e-6) + 1e-30 : 0.0;
    const double dlo = static_cast<double>(threshold) - B, dhi = static_cast<double>(threshold) + B;
    lo = static_cast<float>(dlo);
    if (static_cast<double>(lo) > dlo) lo = std::nextafter(lo, -std::numeric_limits<float>::infinity());
    hi = static_cast<float>(dhi);
    if (static_cast<double>(hi) < dhi) hi = std::nextafter(hi, std::numeric_limits<float>::infinity());
}

extern "C" int vj_engine_set_cascade(vj_engine* e, const vj_feature* dFeatures, int nF, const vj_stump* dStumps,
                                     int nS, const vj_stage* hStages, int nSt, vj_stream sv) {
    // src/face_vision_engine.cu:1182-1185
    VJ_CHECK(e && dFeatures && dStumps && hStages && nF > 0 && nS > 0 && nSt > 0,
             "setCascadeModel: invalid arguments");
    hipStream_t s = static_cast<hipStream_t>(sv);
    std::vector<vj_feature> hf(static_cast<size_t>(nF));
    std::vector<vj_stump> hs(static_cast<size_t>(nS));
    VJ_HIP(hipMemcpyAsync(hf.data(), dFeatures, sizeof(vj_feature) * hf.size(), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipMemcpyAsync(hs.data(), dStumps, sizeof(vj_stump) * hs.size(), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < nSt; ++k) {
        VJ_CHECK(hStages[k].first >= 0 && hStages[k].ntrees >= 0 &&
                     static_cast<int64_t>(hStages[k].first) + hStages[k].ntrees <= nS,
                 "setCascadeModel: stage range outside the stump array");
    }
    const int mW = e->winW + 1;
    const int area = mW * (e->winH + 1);
    // LUT of corner offsets for each step's LDS pitch (the detector decodes
    // p into (p / (winW+1), p % (winW+1)) as src/face_vision_engine.cu:360-365)
    std::vector<KStump> ks(static_cast<size_t>(nS) * 3);
    for (int i = 0; i < nS; ++i) {
        const vj_stump& st = hs[static_cast<size_t>(i)];
        VJ_CHECK(st.feature >= 0 && st.feature < nF, "setCascadeModel: stump feature index out of range");
        const vj_feature& f = hf[static_cast<size_t>(st.feature)];
        VJ_CHECK(f.rectCount == 2 || f.rectCount == 3, "setCascadeModel: feature rectCount must be 2 or 3");
        for (int k = 0; k < 3; ++k) {
            const int pitch = cascade_pitch(k + 1, e->winW);
            KStump& o = ks[static_cast<size_t>(k) * nS + i];
            std::memset(&o, 0, sizeof o);
            for (int r = 0; r < 3; ++r) {
                const int32_t p[4] = {f.r[r].p0, f.r[r].p1, f.r[r].p2, f.r[r].p3};
                for (int c = 0; c < 4; ++c) {
                    int v = p[c];
                    if (r >= f.rectCount) v = 0;
                    VJ_CHECK(v >= 0 && v < area, "setCascadeModel: feature corner outside the window");
                    const int dy = v / mW, dx = v - dy * mW;
                    o.off[r * 4 + c] = static_cast<uint16_t>((dy * pitch + dx) * 4);  // LDS byte offset
                }
                o.w[r] = (r < f.rectCount) ? f.r[r].w : 0.0f;
            }
            o.theta = st.theta;
            o.left = st.left;
            o.right = st.right;
            o.rectCount = f.rectCount;
        }
    }
    std::vector<KStage> kg(static_cast<size_t>(nSt));
    for (int k = 0; k < nSt; ++k) {
        KStage& g = kg[static_cast<size_t>(k)];
        g = KStage{hStages[k].first, hStages[k].ntrees, hStages[k].threshold, 0.0f, 0.0f, {0, 0, 0}};
        stage_margins(hs.data() + hStages[k].first, hStages[k].ntrees, hStages[k].threshold, g.lo, g.hi);
    }
    VJ_TRY(e->modelFeat.ensure(static_cast<size_t>(nF)));
    VJ_TRY(e->kst.ensure(ks.size()));
    VJ_TRY(e->kstage.ensure(kg.size()));
    VJ_HIP(hipMemcpyAsync(e->modelFeat.get(), dFeatures, sizeof(vj_feature) * nF, hipMemcpyDeviceToDevice, s));
    VJ_HIP(hipMemcpyAsync(e->kst.get(), ks.data(), sizeof(KStump) * ks.size(), hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(e->kstage.get(), kg.data(), sizeof(KStage) * kg.size(), hipMemcpyHostToDevice, s));
    VJ_HIP(hipStreamSynchronize(s));
    e->nModelFeat = nF;
    e->nStumps = nS;
    e->nStages = nSt;
    e->planValid = false;
    // The first pass runs the leading stages whose stumps add up to at most
    // kSplitStumps (at least one stage); the rest go to the deep pass.
    {
        int64_t cum = 0;
        int k = 0;
        while (k < nSt && (k == 0 || cum + hStages[k].ntrees <= kSplitStumps)) cum += hStages[k++].ntrees;
        e->stageSplit = k;
        // the deep pass runs stages until it has about kDeepStumps stumps
        // (at least one stage); the tail pass runs the rest
        cum = 0;
        while (k < nSt && cum < kDeepStumps) cum += hStages[k++].ntrees;
        e->stageTail = e->stageTailDefault = k;
    }
    return VJ_OK;
}

extern "C" int vj_engine_clear_cascade(vj_engine* e, vj_stream) {
    VJ_CHECK(e, "clearCascadeModel: null");
    e->modelFeat.release();
    e->kst.release();
    e->kstage.release();
    e->nModelFeat = e->nStumps = e->nStages = 0;
    return VJ_OK;
}

namespace {

// Cascade tile rows per window step (1, 2, 3) the autotuner tries; each keeps
// a tile within kTileMaxWin windows and 32 rows.
constexpr std::array<int, 4> kTileCandidates[] = {
    {0, 8, 8, 8}, {0, 8, 4, 4}, {0, 4, 4, 4}, {0, 8, 8, 4}, {0, 4, 8, 8}};

// Autotuner: the first call of a batch shape (image sizes, scale factor,
// minimum object size, cascade size) times the cascade kernels for each
// candidate tile shape, with and without the tail pass (the deep pass then
// runs every remaining stage: slower on one frame, whose few late windows
// the tail spreads over more lanes, faster on large batches), best of two
// runs, HIP events on the call's stream, and keeps the fastest for that
// shape; later calls reuse it.  Every choice gives the same detections;
// only the speed differs.
int autotune(vj_engine* e, const std::vector<ImageIn>& imgs, float sf, int minObj, int capPer, hipStream_t s) {
    std::vector<int64_t> key;
    uint32_t sfBits;
    std::memcpy(&sfBits, &sf, 4);
    for (int64_t v : {int64_t(sfBits), int64_t(minObj), int64_t(e->nStages), int64_t(e->nStumps)}) key.push_back(v);
    for (const ImageIn& in : imgs) {
        key.push_back(in.w);
        key.push_back(in.h);
    }
    auto it = e->tuned.find(key);
    if (it == e->tuned.end()) {
        if (!e->ev[0])
            for (hipEvent_t& x : e->ev) VJ_HIP(hipEventCreate(&x));
        const bool timing = e->timing;
        e->timing = true;
        float best = 0.0f;
        std::array<int, 5> bestRows{};
        int status = VJ_OK;
        std::vector<std::array<int, 5>> cands;
        for (const std::array<int, 4>& c : kTileCandidates) {
            cands.push_back({c[0], c[1], c[2], c[3], e->stageTailDefault});
            if (e->stageTailDefault < e->nStages) cands.push_back({c[0], c[1], c[2], c[3], e->nStages});
        }
        for (const std::array<int, 5>& c : cands) {
            std::copy(c.begin(), c.begin() + 4, e->tileRows);
            e->stageTail = c[4];
            float t = -1.0f;
            for (int rep = 0; rep < 2 && status == VJ_OK; ++rep) {
                Plan P;
                status = build_plan(e, imgs, sf, minObj, P);
                e->planValid = false;  // the tuning plans overwrite the plan buffer
                if (status == VJ_OK) status = run_pipeline(e, P, capPer, s, false);
                float ms = 0.0f;
                if (status == VJ_OK && hipStreamSynchronize(s) == hipSuccess &&
                    hipEventElapsedTime(&ms, e->ev[2], e->ev[3]) == hipSuccess) {
                    if (t < 0.0f || ms < t) t = ms;
                } else if (status == VJ_OK) {
                    status = hip_fail(hipGetLastError(), "autotune timing", __FILE__, __LINE__);
                }
            }
            if (status != VJ_OK) break;
            if (best == 0.0f || t < best) {
                best = t;
                bestRows = c;
            }
        }
        e->timing = timing;
        if (status != VJ_OK) return status;
        it = e->tuned.emplace(key, bestRows).first;
    }
    std::copy(it->second.begin(), it->second.begin() + 4, e->tileRows);
    e->stageTail = it->second[4];
    return VJ_OK;
}

// Grouped boxes copied back together with the per-image counts.
constexpr size_t kOutPrefetch = 4096;

int detect_images(vj_engine* e, const std::vector<ImageIn>& imgs, float sf, int minNeighbors, int minObj,
                  int maxDet, int applyGrouping, vj_detection* hOut, int cap, int* hCounts, bool batch,
                  hipStream_t s) {
    VJ_CHECK(sf > 1.0f && minObj >= 1, "detectMultiScale: scaleFactor must be > 1 and minObjectSize >= 1");
    VJ_CHECK(e->nStages > 0 && e->kst.get(), "detectMultiScale: no cascade model set");
    for (const ImageIn& in : imgs) {
        VJ_CHECK(in.pix && in.w > 0 && in.h > 0 && in.stride >= in.w, "detectMultiScale: invalid image");
        VJ_TRY(check_image_limits(e, in.w, in.h));
    }
    const int capPer = std::max(1, std::min(maxDet, e->maxDet));
    if (e->autotune) VJ_TRY(autotune(e, imgs, sf, minObj, capPer, s));
    // The plan depends only on the images (addresses, sizes, strides), the
    // scale factor, the minimum object size and the tile rows; a call that
    // repeats the previous call's reuses its plan and the tables it left in
    // the plan buffer (while the pyramid buffer has not moved).
    std::vector<int64_t> key;
    {
        uint32_t sfBits;
        std::memcpy(&sfBits, &sf, 4);
        for (int64_t v : {int64_t(sfBits), int64_t(minObj), int64_t(capPer), int64_t(e->tileRows[1]),
                          int64_t(e->tileRows[2]), int64_t(e->tileRows[3]), int64_t(e->nStages)})
            key.push_back(v);
        for (const ImageIn& in : imgs)
            for (int64_t v : {int64_t(reinterpret_cast<uintptr_t>(in.pix)), int64_t(in.w), int64_t(in.h),
                              int64_t(in.stride)})
                key.push_back(v);
    }
    const bool reuse = e->planValid && key == e->planKey && e->pyr.get() == e->planPyr;
    Plan& P = e->plan;
    if (!reuse) {
        e->planValid = false;
        VJ_TRY(build_plan(e, imgs, sf, minObj, P));
    }
    const int nImages = static_cast<int>(imgs.size());
    VJ_TRY(run_pipeline(e, P, capPer, s, reuse));
    if (!reuse) {
        e->planKey = std::move(key);
        e->planPyr = e->pyr.get();
        e->planValid = true;
    }
    // grouping on the device, every image of the call in the same launches:
    // sweep order, similar pairs (and their unions), component statistics,
    // the reference's DSU roots of the kept components, the clusters
    const size_t capTotal = static_cast<size_t>(capPer) * static_cast<size_t>(nImages);
    const bool devGroup = applyGrouping != 0 && capTotal < (size_t{1} << 31);
    size_t outPrefetch = 0;
    if (devGroup) {
        if (e->pairCap == 0) e->pairCap = static_cast<int64_t>(std::min<size_t>(capTotal, size_t{1} << 20) * 16 + 65536);
        // image widths: one x bin per column 0 .. w of an image (LDS of the sweep)
        int maxW = 1;
        VJ_TRY(e->hImgW.ensure(static_cast<size_t>(nImages)));
        for (int i = 0; i < nImages; ++i) {
            e->hImgW.get()[i] = imgs[static_cast<size_t>(i)].w;
            maxW = std::max(maxW, imgs[static_cast<size_t>(i)].w);
        }
        // the widths go up only when they differ from those already there (a
        // buffer that grows is a new allocation and gets them again)
        const bool fresh = !e->gImgW.get() || static_cast<size_t>(nImages) > e->gImgW.capacity();
        VJ_TRY(e->gImgW.ensure(static_cast<size_t>(nImages)));
        const std::vector<int32_t> widths(e->hImgW.get(), e->hImgW.get() + nImages);
        if (fresh || widths != e->imgWUp) {
            VJ_HIP(hipMemcpyAsync(e->gImgW.get(), e->hImgW.get(), sizeof(int32_t) * nImages, hipMemcpyHostToDevice, s));
            e->imgWUp = widths;
        }
        VJ_TRY(e->gBucket.ensure(capTotal));
        VJ_TRY(e->gPairCnt.ensure(capTotal));
        VJ_TRY(e->gPairOff.ensure(capTotal));
        VJ_TRY(e->gSorted.ensure(capTotal));
        VJ_TRY(e->gOrder.ensure(capTotal));
        VJ_TRY(e->gPairs.ensure(static_cast<size_t>(e->pairCap)));
        VJ_TRY(e->gLabel.ensure(capTotal));
        VJ_TRY(e->gCount.ensure(capTotal));
        VJ_TRY(e->gClusters.ensure(capTotal));
        VJ_TRY(e->gKeptAll.ensure(capTotal));
        VJ_TRY(e->gKeptSlot.ensure(capTotal));
        VJ_TRY(e->gOutG.ensure(capTotal));
        VJ_TRY(e->gGrpImg.ensure(3 * static_cast<size_t>(nImages)));
        VJ_TRY(e->gCtl.ensure(2));
        VJ_TRY(e->gScanTot.ensure((capTotal + kScanItems - 1) / kScanItems));
        const int sweepLdsBoxes = std::min(capPer, kSweepLdsBoxes);
        hipLaunchKernelGGL(pair_sweep_kernel, dim3(nImages), dim3(kSweepThreads), sweep_lds_bytes(maxW, sweepLdsBoxes),
                           s, e->dOut.get(), e->dCounts.get(), e->dImgBase.get(), e->gImgW.get(), capPer,
                           sweepLdsBoxes, e->gBucket.get(), e->gSorted.get(), e->gOrder.get(), e->gGrpImg.get(), e->gCtl.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(pair_find_kernel<false>, dim3(2048), dim3(256), 0, s, e->dCounts.get(), e->dImgBase.get(),
                           nImages, capPer, 0.2f, e->gSorted.get(), e->gPairCnt.get(), nullptr, nullptr, int64_t{0},
                           e->gLabel.get(), e->gCount.get());
        VJ_HIP(hipGetLastError());
        const int scanGrid = static_cast<int>(std::min<size_t>((capTotal + kScanItems - 1) / kScanItems, 1024));
        hipLaunchKernelGGL(pair_scan_local_kernel, dim3(scanGrid), dim3(kScanThreads), 0, s, e->dCounts.get(),
                           nImages, capPer, e->gPairCnt.get(), e->gPairOff.get(), e->gScanTot.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(pair_scan_add_kernel, dim3(scanGrid), dim3(kScanThreads), 0, s, e->dCounts.get(),
                           nImages, capPer, e->gScanTot.get(), e->gPairOff.get(), e->gCtl.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(pair_find_kernel<true>, dim3(2048), dim3(256), 0, s, e->dCounts.get(), e->dImgBase.get(),
                           nImages, capPer, 0.2f, e->gSorted.get(), e->gPairCnt.get(), e->gPairOff.get(),
                           e->gPairs.get(), e->pairCap, e->gLabel.get(), nullptr);
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(grp_comp_kernel, dim3(1024), dim3(256), 0, s, e->dCounts.get(), e->dImgBase.get(), nImages,
                           capPer, e->gLabel.get(), e->gCount.get(), e->gCtl.get(), e->pairCap);
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(grp_list_kernel, dim3(1024), dim3(256), 0, s, e->dCounts.get(), e->dImgBase.get(), nImages,
                           capPer, minNeighbors, e->gLabel.get(), e->gCount.get(), e->gKeptAll.get(),
                           e->gKeptSlot.get(), e->gGrpImg.get(), e->gCtl.get(), e->pairCap);
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(grp_replay_kernel, dim3(2048), dim3(kWave), 0, s, e->dImgBase.get(), nImages, e->dOut.get(),
                           e->gSorted.get(), e->gOrder.get(), e->gLabel.get(), e->gCount.get(), e->gKeptAll.get(),
                           e->gKeptSlot.get(), e->gCtl.get(), e->pairCap, e->gPairOff.get(), e->gPairCnt.get(),
                           e->gPairs.get(), e->gGrpImg.get(), e->gClusters.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(grp_finish_kernel, dim3(nImages), dim3(kFinishThreads), 0, s, e->dCounts.get(),
                           e->dImgBase.get(), capPer, 0.2f, e->gClusters.get(), e->gCtl.get(), e->pairCap,
                           e->gGrpImg.get(), e->gOutG.get());
        VJ_HIP(hipGetLastError());
        VJ_TRY(e->hCtl.ensure(2));
        VJ_TRY(e->hGrpImg.ensure(3 * static_cast<size_t>(nImages)));
        // the first grouped boxes come back with the counts, so a call with
        // few of them needs no second round trip
        outPrefetch = std::min<size_t>(capTotal, kOutPrefetch);
        VJ_TRY(e->hOutG.ensure(outPrefetch));
    }
    VJ_TRY(e->hCounts.ensure(static_cast<size_t>(nImages)));
    if (devGroup) {
        // one launch stores the control words, per-image counts and the first
        // grouped boxes straight into the pinned buffers
        void *dCtl, *dGrpImg, *dCounts, *dOutG;
        VJ_HIP(hipHostGetDevicePointer(&dCtl, e->hCtl.get(), 0));
        VJ_HIP(hipHostGetDevicePointer(&dGrpImg, e->hGrpImg.get(), 0));
        VJ_HIP(hipHostGetDevicePointer(&dCounts, e->hCounts.get(), 0));
        VJ_HIP(hipHostGetDevicePointer(&dOutG, e->hOutG.get(), 0));
        hipLaunchKernelGGL(publish_kernel, dim3(1), dim3(256), 0, s, e->gCtl.get(), e->gGrpImg.get(),
                           e->dCounts.get(), e->gOutG.get(), nImages, static_cast<int>(outPrefetch),
                           static_cast<int64_t*>(dCtl), static_cast<int32_t*>(dGrpImg),
                           static_cast<int32_t*>(dCounts), static_cast<vj_detection*>(dOutG));
        VJ_HIP(hipGetLastError());
    } else {
        VJ_HIP(hipMemcpyAsync(e->hCounts.get(), e->dCounts.get(), sizeof(int32_t) * nImages, hipMemcpyDeviceToHost,
                              s));
    }
    VJ_HIP(hipStreamSynchronize(s));
    if (e->timing)
        for (int k = 0; k < 4; ++k) VJ_HIP(hipEventElapsedTime(&e->phaseMs[k], e->ev[k], e->ev[k + 1]));
    if (e->countStages) {
        std::vector<unsigned long long> cnt(static_cast<size_t>(e->nStages), 0ull);
        if (e->stageEnter.get())
            VJ_HIP(hipMemcpy(cnt.data(), e->stageEnter.get(), sizeof(unsigned long long) * cnt.size(),
                             hipMemcpyDeviceToHost));
        e->stageWindows.assign(cnt.begin(), cnt.end());
    }
    int64_t total = 0;
    int64_t raw = 0;
    for (int i = 0; i < nImages; ++i) {
        raw += e->hCounts.get()[i];
        total += std::min(e->hCounts.get()[i], capPer);
    }
    e->stats.levels = 0;
    for (int32_t n : P.imgLevels) e->stats.levels += n;
    e->stats.windows = P.windows;
    e->stats.rawDetections = raw;
    e->stats.pyramidPixels = P.pyramidPixels;
    e->stats.integralBytes = P.integralBytes;
    e->stats.cascadeBytes = P.cascadeBytes + static_cast<int64_t>(e->nStumps) * 64 + 24 * total;
    std::vector<int64_t> rawOff(static_cast<size_t>(nImages) + 1, 0);
    for (int i = 0; i < nImages; ++i)
        rawOff[static_cast<size_t>(i) + 1] = rawOff[static_cast<size_t>(i)] + std::min(e->hCounts.get()[i], capPer);
    auto rawCount = [&](int i) { return static_cast<int>(rawOff[static_cast<size_t>(i) + 1] - rawOff[static_cast<size_t>(i)]); };
    VJ_TRY(e->hOut.ensure(static_cast<size_t>(std::max<int64_t>(total, 1))));

    // per image: where its result comes from.  Grouped on the device unless
    // the pair buffer overflowed (it grows for the next call), the image has
    // more boxes than the device sweep takes, or a component or the cluster
    // list passed the device limits; those images are grouped on the host
    // (every image's steps spread over the host thread pool).
    std::vector<char> onDev(static_cast<size_t>(nImages), 0);
    std::vector<int64_t> devOff(static_cast<size_t>(nImages), 0);
    std::vector<int> hostImgs;
    int64_t devTotal = 0;
    if (applyGrouping) {
        bool usable = false;
        if (devGroup) {
            int64_t nextCap = e->pairCap;
            usable = pair_plan(e->hCtl.get()[0], e->pairCap, &nextCap);
            e->pairCap = nextCap;
        }
        for (int i = 0; i < nImages; ++i) {
            const int n = rawCount(i);
            devOff[static_cast<size_t>(i)] = devTotal;
            if (devGroup) devTotal += e->hGrpImg.get()[3 * i];
            if (n == 0) continue;
            if (usable && n <= kGroupDevMax && !e->hGrpImg.get()[3 * i + 2]) onDev[static_cast<size_t>(i)] = 1;
            else hostImgs.push_back(i);
        }
        bool copied = false;
        if (!hostImgs.empty()) {
            copied = true;
            if (static_cast<int>(hostImgs.size()) == nImages) {
                VJ_HIP(hipMemcpyAsync(e->hOut.get(), e->dOut.get(), sizeof(vj_detection) * total, hipMemcpyDeviceToHost, s));
            } else {
                for (int i : hostImgs)
                    VJ_HIP(hipMemcpyAsync(e->hOut.get() + rawOff[static_cast<size_t>(i)],
                                          e->dOut.get() + rawOff[static_cast<size_t>(i)],
                                          sizeof(vj_detection) * rawCount(i), hipMemcpyDeviceToHost, s));
            }
        }
        const bool anyDev = std::find(onDev.begin(), onDev.end(), 1) != onDev.end();
        if (anyDev && devTotal > static_cast<int64_t>(outPrefetch)) {
            copied = true;
            VJ_TRY(e->hOutG.ensure(static_cast<size_t>(devTotal)));
            VJ_HIP(hipMemcpyAsync(e->hOutG.get(), e->gOutG.get(), sizeof(vj_detection) * devTotal,
                                  hipMemcpyDeviceToHost, s));
        }
        if (copied) VJ_HIP(hipStreamSynchronize(s));
    } else if (total > 0) {
        VJ_HIP(hipMemcpyAsync(e->hOut.get(), e->dOut.get(), sizeof(vj_detection) * total, hipMemcpyDeviceToHost, s));
        VJ_HIP(hipStreamSynchronize(s));
    }
    e->stats.imagesGroupedOnHost = static_cast<int32_t>(hostImgs.size());
    std::vector<std::vector<vj_detection>> grouped(static_cast<size_t>(nImages));
    if (!hostImgs.empty()) {
        const size_t m = hostImgs.size();
        std::vector<const vj_detection*> ins(m);
        std::vector<vj_detection*> outs(m);
        std::vector<int> ns(m), cnt(m);
        for (size_t k = 0; k < m; ++k) {
            const int i = hostImgs[k];
            ns[k] = rawCount(i);
            grouped[static_cast<size_t>(i)].resize(static_cast<size_t>(ns[k]));
            ins[k] = e->hOut.get() + rawOff[static_cast<size_t>(i)];
            outs[k] = grouped[static_cast<size_t>(i)].data();
        }
        VJ_CHECK(group_rectangles_many(ins.data(), ns.data(), static_cast<int>(m), minNeighbors, 0.2f, outs.data(),
                                       cnt.data()) == VJ_OK,
                 "groupRectangles failed (host memory)");
        for (size_t k = 0; k < m; ++k) grouped[static_cast<size_t>(hostImgs[k])].resize(static_cast<size_t>(cnt[k]));
    }
    int64_t outPos = 0;
    for (int i = 0; i < nImages; ++i) {
        const vj_detection* src = e->hOut.get() + rawOff[static_cast<size_t>(i)];
        int m = rawCount(i);
        if (applyGrouping) {
            if (onDev[static_cast<size_t>(i)]) {
                m = e->hGrpImg.get()[3 * i + 1];
                src = m > 0 ? e->hOutG.get() + devOff[static_cast<size_t>(i)] : nullptr;
            } else {
                src = grouped[static_cast<size_t>(i)].data();
                m = static_cast<int>(grouped[static_cast<size_t>(i)].size());
            }
        }
        if (hCounts) hCounts[i] = m;
        if (batch) {
            VJ_CHECK(outPos + m <= cap, "detect_batch: output capacity too small");
        }
        const int64_t room = std::max<int64_t>(0, std::min<int64_t>(m, cap - outPos));
        if (hOut && room > 0) std::memcpy(hOut + outPos, src, sizeof(vj_detection) * static_cast<size_t>(room));
        outPos += m;
    }
    return VJ_OK;
}

}  // namespace

extern "C" int vj_engine_detect(vj_engine* e, const uint8_t* dImage, int w, int h, int stride, float sf,
                                int minNeighbors, int minObj, int maxDet, int applyGrouping, vj_detection* hOut,
                                int cap, int* count, vj_stream sv) {
    VJ_CHECK(e && count && cap >= 0, "detectMultiScale: bad arguments");
    *count = 0;
    try {
        std::vector<ImageIn> imgs{{dImage, w, h, stride}};
        int c = 0;
        VJ_TRY(detect_images(e, imgs, sf, minNeighbors, minObj, maxDet, applyGrouping, hOut, cap, &c, false,
                             static_cast<hipStream_t>(sv)));
        *count = c;
        return VJ_OK;
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return VJ_RUNTIME_ERROR;
    }
}

extern "C" int vj_engine_detect_batch(vj_engine* e, int nImages, const uint8_t* const* dImages, const int* w,
                                      const int* h, const int* stride, float sf, int minNeighbors, int minObj,
                                      int maxDet, int applyGrouping, vj_detection* hOut, int cap, int* hCounts,
                                      vj_stream sv) {
    VJ_CHECK(e && nImages > 0 && dImages && w && h && stride && hCounts && cap >= 0,
             "detect_batch: bad arguments");
    try {
        std::vector<ImageIn> imgs;
        for (int i = 0; i < nImages; ++i) imgs.push_back({dImages[i], w[i], h[i], stride[i]});
        return detect_images(e, imgs, sf, minNeighbors, minObj, maxDet, applyGrouping, hOut, cap, hCounts, true,
                             static_cast<hipStream_t>(sv));
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return VJ_RUNTIME_ERROR;
    }
}

extern "C" int vj_engine_detect_footprint(vj_engine* e, int w, int h, float sf, int minObj, int maxDet,
                                          int64_t* bytes) {
    VJ_CHECK(e && bytes && w > 0 && h > 0 && sf > 1.0f && minObj >= 1 && maxDet >= 1,
             "detect_footprint: bad arguments");
    *bytes = 0;
    try {
        Plan P;
        std::vector<ImageIn> imgs{{nullptr, w, h, w}};
        VJ_TRY(build_plan(e, imgs, sf, minObj, P));
        const int64_t cap = std::max(1, std::min(maxDet, e->maxDet));
        *bytes = static_cast<int64_t>(w) * h + P.pyrBytes + 8 * P.iiElems + 8 * P.colElems + 12 * P.totElems +
                 8 * P.rowElems + 8 * P.nWords + 4 * 64 * P.nWords + 24 * cap + 6 * P.windows +
                 44 * static_cast<int64_t>(P.nTiles) +
                 static_cast<int64_t>(P.levels.size()) * static_cast<int64_t>(sizeof(LevelDesc) + 64);
        return VJ_OK;
    } catch (const std::exception& ex) {
        set_error(ex.what());
        return VJ_RUNTIME_ERROR;
    }
}

extern "C" int vj_engine_set_graphs(vj_engine* e, int on) {
    VJ_CHECK(e, "vj_engine_set_graphs: null");
    e->graphs = on != 0;
    return VJ_OK;
}

extern "C" int vj_engine_set_autotune(vj_engine* e, int on) {
    VJ_CHECK(e, "vj_engine_set_autotune: null");
    e->autotune = on != 0;
    if (!e->autotune) {
        for (int st = 1; st <= 3; ++st) e->tileRows[st] = tile_rows_for_step(st);
        e->stageTail = e->stageTailDefault;
    }
    return VJ_OK;
}

extern "C" int vj_engine_cascade_split(vj_engine* e, int* split3) {
    VJ_CHECK(e && split3, "vj_engine_cascade_split: null");
    split3[0] = e->stageSplit;
    split3[1] = e->stageTail;
    split3[2] = e->nStages;
    return VJ_OK;
}

extern "C" int vj_engine_tile_rows(vj_engine* e, int* rows) {
    VJ_CHECK(e && rows, "vj_engine_tile_rows: null");
    for (int st = 1; st <= 3; ++st) rows[st - 1] = e->tileRows[st];
    return VJ_OK;
}

extern "C" int vj_engine_set_timing(vj_engine* e, int on) {
    VJ_CHECK(e, "vj_engine_set_timing: null");
    if (on && !e->ev[0])
        for (hipEvent_t& x : e->ev) VJ_HIP(hipEventCreate(&x));
    e->timing = on != 0;
    return VJ_OK;
}

extern "C" int vj_engine_phase_ms(vj_engine* e, float* ms) {
    VJ_CHECK(e && ms, "vj_engine_phase_ms: null");
    for (int k = 0; k < 4; ++k) ms[k] = e->phaseMs[k];
    return VJ_OK;
}

extern "C" int vj_engine_count_stages(vj_engine* e, int on) {
    VJ_CHECK(e, "vj_engine_count_stages: null");
    e->countStages = on != 0;
    return VJ_OK;
}

extern "C" int vj_engine_stage_windows(vj_engine* e, int64_t* out, int cap, int* nStages) {
    VJ_CHECK(e && nStages && (out || cap == 0), "vj_engine_stage_windows: null");
    *nStages = static_cast<int>(e->stageWindows.size());
    for (int k = 0; k < std::min(cap, *nStages); ++k) out[k] = e->stageWindows[static_cast<size_t>(k)];
    return VJ_OK;
}

extern "C" int vj_engine_last_stats(vj_engine* e, vj_detect_stats* out) {
    VJ_CHECK(e && out, "vj_engine_last_stats: null");
    *out = e->stats;
    return VJ_OK;
}
