/*
 * vj_oracle.c -- CPU restatement of the reference hot path (TEST INFRASTRUCTURE).
 *
 * Only tests/, the smoke entry point and bench.py's cpu_baseline leg load this
 * library, as the checker.  It restates the reference's single-threaded CPU
 * routines (src/main.cu:412-827) and the host-side engine code
 * (src/face_vision_engine.cu:585-635, 882-1020, 1401-1714) line for line in
 * their arithmetic, so that every rounding step matches.  Where the reference
 * GPU path is followed instead of its CPU path (stump-search tie-breaking and
 * the -0/+0 key order of its radix sort) the function says so.
 *
 * Compile: gcc -O2 -ffp-contract=off -fno-fast-math (see oracle/Makefile).
 */
#include "vj_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

/* OR_MUTANT (0 in every normal build) plants one known mistake, so the tests
 * can show that their brute-force checks catch it (oracle/Makefile target
 * `mutants`; tests/test_oracle.py::test_mutants_fail_the_brute_force_checks):
 *   1  detect: the third corner of a rectangle is added, not subtracted
 *   2  detect: a stage's stumps are summed in reverse order
 *   3  training: the second rectangle takes the first one's weight
 *   4  training: the last corner of the first rectangle is read one column left
 *   5  detect: the window normaliser uses the inset rect's p1 corner twice */
#ifndef OR_MUTANT
#define OR_MUTANT 0
#endif

#define MAXI(a, b) ((a) > (b) ? (a) : (b))
#define MINI(a, b) ((a) < (b) ? (a) : (b))

/* ---- Haar pool: src/face_vision_engine.cu:618-635, 882-1020 -------------- */

static or_rect make_rect(int x, int y, int w, int h, int iiW, float weight, int tilted) {
    or_rect r;
    if (!tilted) {
        r.p0 = x + iiW * y;
        r.p1 = (x + w) + iiW * y;
        r.p2 = x + iiW * (y + h);
        r.p3 = (x + w) + iiW * (y + h);
    } else {
        r.p0 = x + iiW * y;
        r.p1 = (x - h) + iiW * (y + h);
        r.p2 = (x + w) + iiW * (y + w);
        r.p3 = (x + w - h) + iiW * (y + w + h);
    }
    r.w = weight;
    return r;
}

typedef struct { or_feature* out; int cap; int n; int iiW; } pool_ctx;

static void push2(pool_ctx* c, int tilted, int x0, int y0, int w0, int h0, float t0,
                  int x1, int y1, int w1, int h1, float t1) {
    if (c->out && c->n < c->cap) {
        or_feature f;
        memset(&f, 0, sizeof f);
        f.r[0] = make_rect(x0, y0, w0, h0, c->iiW, t0, tilted);
        f.r[1] = make_rect(x1, y1, w1, h1, c->iiW, t1, tilted);
        f.rectCount = 2;
        f.tilted = (uint8_t)tilted;
        c->out[c->n] = f;
    }
    c->n++;
}

static void push3(pool_ctx* c, int x0, int y0, int w0, int h0, float t0,
                  int x1, int y1, int w1, int h1, float t1,
                  int x2, int y2, int w2, int h2, float t2) {
    if (c->out && c->n < c->cap) {
        or_feature f;
        memset(&f, 0, sizeof f);
        f.r[0] = make_rect(x0, y0, w0, h0, c->iiW, t0, 0);
        f.r[1] = make_rect(x1, y1, w1, h1, c->iiW, t1, 0);
        f.r[2] = make_rect(x2, y2, w2, h2, c->iiW, t2, 0);
        f.rectCount = 3;
        c->out[c->n] = f;
    }
    c->n++;
}

int or_haar_pool(int W, int H, int mode, or_feature* out, int cap) {
    pool_ctx c = {out, cap, 0, W + 1};
    const int core = mode != 0, tilt = mode == 2;
    for (int x = 0; x < W; ++x)
        for (int y = 0; y < H; ++y)
            for (int dx = 1; dx <= W; ++dx)
                for (int dy = 1; dy <= H; ++dy) {
                    if (x + 2 * dx <= W && y + dy <= H)
                        push2(&c, 0, x, y, 2 * dx, dy, -1.0f, x + dx, y, dx, dy, 2.0f);
                    if (x + dx <= W && y + 2 * dy <= H)
                        push2(&c, 0, x, y, dx, 2 * dy, -1.0f, x, y + dy, dx, dy, 2.0f);
                    if (x + 3 * dx <= W && y + dy <= H)
                        push2(&c, 0, x, y, 3 * dx, dy, -1.0f, x + dx, y, dx, dy, 2.0f);
                    if (x + dx <= W && y + 3 * dy <= H)
                        push2(&c, 0, x, y, dx, 3 * dy, -1.0f, x, y + dy, dx, dy, 2.0f);
                    if (x + 2 * dx <= W && y + 2 * dy <= H)
                        push3(&c, x, y, 2 * dx, 2 * dy, -1.0f, x, y, dx, dy, 2.0f,
                              x + dx, y + dy, dx, dy, 2.0f);
                    if (core) {
                        if (x + 4 * dx <= W && y + dy <= H)
                            push2(&c, 0, x, y, 4 * dx, dy, -1.0f, x + dx, y, 2 * dx, dy, 2.0f);
                        if (x + dx <= W && y + 4 * dy <= H)
                            push2(&c, 0, x, y, dx, 4 * dy, -1.0f, x, y + dy, dx, 2 * dy, 2.0f);
                        if (x + 3 * dx <= W && y + 3 * dy <= H)
                            push2(&c, 0, x, y, 3 * dx, 3 * dy, -1.0f, x + dx, y + dy, dx, dy, 9.0f);
                    }
                    if (tilt) {
                        if (x + 2 * dx <= W && y + 2 * dx + dy <= H && x - dy >= 0)
                            push2(&c, 1, x, y, 2 * dx, dy, -1.0f, x, y, dx, dy, 2.0f);
                        if (x + dx <= W && y + dx + 2 * dy <= H && x - 2 * dy >= 0)
                            push2(&c, 1, x, y, dx, 2 * dy, -1.0f, x, y, dx, dy, 2.0f);
                        if (x + 3 * dx <= W && y + 3 * dx + dy <= H && x - dy >= 0)
                            push2(&c, 1, x, y, 3 * dx, dy, -1.0f, x + dx, y + dx, dx, dy, 3.0f);
                        if (x + dx <= W && y + dx + 3 * dy <= H && x - 3 * dy >= 0)
                            push2(&c, 1, x, y, dx, 3 * dy, -1.0f, x - dy, y + dy, dx, dy, 3.0f);
                        if (x + 4 * dx <= W && y + 4 * dx + dy <= H && x - dy >= 0)
                            push2(&c, 1, x, y, 4 * dx, dy, -1.0f, x + dx, y + dx, 2 * dx, dy, 2.0f);
                        if (x + dx <= W && y + dx + 4 * dy <= H && x - 4 * dy >= 0)
                            push2(&c, 1, x, y, dx, 4 * dy, -1.0f, x - dy, y + dy, dx, 2 * dy, 2.0f);
                    }
                }
    return c.n;
}

/* ---- integral images: src/main.cu:428-455 ------------------------------- */

void or_integral(const uint8_t* img, int w, int h, int stride, uint32_t* sum, uint64_t* sq) {
    const int iiW = w + 1;
    memset(sum, 0, sizeof(uint32_t) * (size_t)iiW * (h + 1));
    memset(sq, 0, sizeof(uint64_t) * (size_t)iiW * (h + 1));
    for (int y = 1; y <= h; ++y) {
        uint32_t rowSum = 0;
        uint64_t rowSq = 0;
        for (int x = 1; x <= w; ++x) {
            const uint32_t pix = img[(size_t)(y - 1) * stride + (x - 1)];
            rowSum += pix;
            rowSq += (uint64_t)(pix * pix);
            const size_t idx = (size_t)y * iiW + x;
            sum[idx] = sum[idx - iiW] + rowSum;
            sq[idx] = sq[idx - iiW] + rowSq;
        }
    }
}

/* src/main.cu:457-472: rect (1,1,iiW-3,iiH-3); deltas in the reference's int32
 * arithmetic (exact for training patches), nf2 and 1/sqrt in double. */
float or_train_invnorm(const uint32_t* sum, const uint64_t* sq, int iiW, int iiH) {
    if (iiW < 3 || iiH < 3) return 1.0f;
    const int x = 1, y = 1, ww = iiW - 3, hh = iiH - 3;
    const int area = ww * hh;
    const int p0 = x + iiW * y, p1 = (x + ww) + iiW * y;
    const int p2 = x + iiW * (y + hh), p3 = (x + ww) + iiW * (y + hh);
    const double s = (double)(int32_t)(sum[p0] - sum[p1] - sum[p2] + sum[p3]);
    const double q = (double)(int64_t)(sq[p0] - sq[p1] - sq[p2] + sq[p3]);
    const double nf2 = (double)area * q - s * s;
    return (nf2 > 1e-12) ? (float)(1.0 / sqrt(nf2)) : 0.0f;
}

/* ---- resize: src/main.cu:545-585 ---------------------------------------- */

void or_resize(const uint8_t* src, int sw, int sh, int sstride, uint8_t* dst, int dw, int dh) {
    const float scaleX = (float)sw / (float)dw;
    const float scaleY = (float)sh / (float)dh;
    for (int y = 0; y < dh; ++y) {
        for (int x = 0; x < dw; ++x) {
            const float fx = ((float)x + 0.5f) * scaleX - 0.5f;
            const float fy = ((float)y + 0.5f) * scaleY - 0.5f;
            const int sx0 = MAXI(0, MINI((int)floorf(fx), sw - 1));
            const int sy0 = MAXI(0, MINI((int)floorf(fy), sh - 1));
            const int sx1 = MINI(sx0 + 1, sw - 1);
            const int sy1 = MINI(sy0 + 1, sh - 1);
            const float ax = fx - (float)sx0;
            const float ay = fy - (float)sy0;
            const float p00 = (float)src[(size_t)sy0 * sstride + sx0];
            const float p01 = (float)src[(size_t)sy0 * sstride + sx1];
            const float p10 = (float)src[(size_t)sy1 * sstride + sx0];
            const float p11 = (float)src[(size_t)sy1 * sstride + sx1];
            const float top = p00 + (p01 - p00) * ax;
            const float bot = p10 + (p11 - p10) * ax;
            const float val = top + (bot - top) * ay;
            int v = (int)(val + 0.5f);
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            dst[(size_t)y * dw + x] = (uint8_t)v;
        }
    }
}

/* ---- mined-box crop: cropResizeTo24, src/main.cu:245-284 ----------------- */

void or_crop_resize(const uint8_t* img, int w, int h, int stride, int bx, int by, int bw, int bh,
                    int outW, int outH, uint8_t* out) {
    const float sx = (float)MAXI(1, bw);
    const float sy = (float)MAXI(1, bh);
    for (int oy = 0; oy < outH; ++oy) {
        for (int ox = 0; ox < outW; ++ox) {
            const float srcFx = (float)bx + ((float)ox + 0.5f) * (sx / (float)outW) - 0.5f;
            const float srcFy = (float)by + ((float)oy + 0.5f) * (sy / (float)outH) - 0.5f;
            const int x0 = MAXI(0, MINI((int)floorf(srcFx), w - 1));
            const int y0 = MAXI(0, MINI((int)floorf(srcFy), h - 1));
            const int x1 = MINI(x0 + 1, w - 1);
            const int y1 = MINI(y0 + 1, h - 1);
            const float ax = srcFx - (float)x0;
            const float ay = srcFy - (float)y0;
            const float p00 = (float)img[(size_t)y0 * stride + x0];
            const float p01 = (float)img[(size_t)y0 * stride + x1];
            const float p10 = (float)img[(size_t)y1 * stride + x0];
            const float p11 = (float)img[(size_t)y1 * stride + x1];
            const float top = p00 + (p01 - p00) * ax;
            const float bot = p10 + (p11 - p10) * ax;
            const float val = top + (bot - top) * ay;
            int v = (int)(val + 0.5f);
            v = v < 0 ? 0 : (v > 255 ? 255 : v);
            out[(size_t)oy * outW + ox] = (uint8_t)v;
        }
    }
}

/* ---- pyramid schedule: src/main.cu:717-747 ------------------------------ */

int or_levels(int imageW, int imageH, int winW, int winH, float sf, int minObj,
              or_level* out, int cap) {
    float scale = 1.0f;
    int scaleIdx = 0, n = 0;
    if (!(sf > 1.0f)) return 0;
    for (;;) {
        const int sw = (int)roundf((float)imageW / scale);
        const int sh = (int)roundf((float)imageH / scale);
        const int ow = (int)roundf((float)winW * scale);
        const int oh = (int)roundf((float)winH * scale);
        if (sw < winW || sh < winH) break;
        if (ow < minObj || oh < minObj) {
            scale *= sf;
            ++scaleIdx;
            continue;
        }
        if (out && n < cap) {
            out[n].scaleIdx = scaleIdx;
            out[n].w = sw;
            out[n].h = sh;
            out[n].step = (scale >= 3.0f) ? 3 : ((scale >= 2.0f) ? 2 : 1);
            out[n].scale = scale;
        }
        ++n;
        scale *= sf;
        ++scaleIdx;
    }
    return n;
}

/* ---- detection: src/main.cu:504-543, 684-827 ---------------------------- */

static inline float ffi(const uint32_t* s, int p) { return (float)(int32_t)s[p]; }

static float det_rect(const or_rect* r, const uint32_t* sum, int iiW, int base, int mW) {
    const int dy0 = r->p0 / mW, dx0 = r->p0 - dy0 * mW;
    const int dy1 = r->p1 / mW, dx1 = r->p1 - dy1 * mW;
    const int dy2 = r->p2 / mW, dx2 = r->p2 - dy2 * mW;
    const int dy3 = r->p3 / mW, dx3 = r->p3 - dy3 * mW;
    const int p0 = base + dy0 * iiW + dx0, p1 = base + dy1 * iiW + dx1;
    const int p2 = base + dy2 * iiW + dx2, p3 = base + dy3 * iiW + dx3;
#if OR_MUTANT == 1
    return (ffi(sum, p0) - ffi(sum, p1) + ffi(sum, p2) + ffi(sum, p3)) * r->w;
#else
    return (ffi(sum, p0) - ffi(sum, p1) - ffi(sum, p2) + ffi(sum, p3)) * r->w;
#endif
}

/* One window at integral offset base: the variance gate, then every stage
 * in order, each stage sum in stump order; returns whether it passes all of
 * them, with the last stage sum in *last (src/main.cu:752-804).  enter (or
 * NULL) counts the stages the window enters. */
static int det_window(const or_feature* features, const or_stump* stumps, const or_stage* stages, int nStages,
                      const uint32_t* sum, const uint64_t* sq, int iiW, int base, int mW, int narea,
                      int np0, int np1, int np2, int np3, float* last, int64_t* enter) {
#if OR_MUTANT == 5
    const double normSum = (double)(int32_t)(sum[base + np0] - sum[base + np1] -
                                             sum[base + np1] + sum[base + np3]);
#else
    const double normSum = (double)(int32_t)(sum[base + np0] - sum[base + np1] -
                                             sum[base + np2] + sum[base + np3]);
#endif
    const uint32_t normSqU = (uint32_t)sq[base + np0] - (uint32_t)sq[base + np1] -
                             (uint32_t)sq[base + np2] + (uint32_t)sq[base + np3];
    const double nf2 = (double)narea * (double)normSqU - normSum * normSum;
    if (!(nf2 > 0.0)) return 0;
    const float invNorm = (float)(1.0 / sqrt(nf2));
    if (!((double)narea * (double)invNorm < 1e-1)) return 0;
    float lastSum = 0.0f;
    for (int si = 0; si < nStages; ++si) {
        if (enter) {
#pragma omp atomic
            enter[si] += 1;
        }
        float stageSum = 0.0f;
        for (int wj = 0; wj < stages[si].ntrees; ++wj) {
#if OR_MUTANT == 2
            const int wi = stages[si].ntrees - 1 - wj;
#else
            const int wi = wj;
#endif
            const or_stump* st = &stumps[stages[si].first + wi];
            const or_feature* f = &features[st->feature];
            float raw = det_rect(&f->r[0], sum, iiW, base, mW) + det_rect(&f->r[1], sum, iiW, base, mW);
            if (f->rectCount == 3) raw += det_rect(&f->r[2], sum, iiW, base, mW);
            const float resp = raw * invNorm;
            stageSum += (resp < st->theta) ? st->left : st->right;
        }
        if (stageSum < stages[si].threshold) return 0;
        lastSum = stageSum;
    }
    *last = lastSum;
    return 1;
}

int or_detect_raw(const or_feature* features, int nFeatures,
                  const or_stump* stumps, int nStumps,
                  const or_stage* stages, int nStages,
                  int winW, int winH,
                  const uint8_t* img, int w, int h, int stride,
                  float sf, int minObj, int maxDet,
                  or_det* out, int cap, int64_t* visited, int64_t* stageEnter) {
    (void)nFeatures;
    (void)nStumps;
    int count = 0;
    int64_t nvis = 0;
    const int mW = winW + 1;
    float scale = 1.0f;
    int scaleIdx = 0;
    const uint8_t* prev = img;
    int prevW = w, prevH = h, prevStride = stride;
    uint8_t* bufA = NULL;
    uint8_t* bufB = NULL;
    uint32_t* sum = NULL;
    uint64_t* sq = NULL;
    if (!(sf > 1.0f)) return 0;
    bufA = (uint8_t*)malloc((size_t)w * h);
    bufB = (uint8_t*)malloc((size_t)w * h);
    sum = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(w + 1) * (h + 1));
    sq = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(w + 1) * (h + 1));
    for (;;) {
        const int sw = (int)roundf((float)w / scale);
        const int sh = (int)roundf((float)h / scale);
        const int ow = (int)roundf((float)winW * scale);
        const int oh = (int)roundf((float)winH * scale);
        if (sw < winW || sh < winH) break;
        if (ow < minObj || oh < minObj) {
            scale *= sf;
            ++scaleIdx;
            continue;
        }
        const uint8_t* lvl;
        int lstride;
        if (scaleIdx == 0) {
            lvl = img;
            lstride = stride;
        } else {
            uint8_t* dst = (prev == bufA) ? bufB : bufA;
            or_resize(prev, prevW, prevH, prevStride, dst, sw, sh);
            lvl = dst;
            lstride = sw;
            prev = dst;
            prevW = sw;
            prevH = sh;
            prevStride = sw;
        }
        or_integral(lvl, sw, sh, lstride, sum, sq);
        const int iiW = sw + 1;
        const int workW = sw - winW + 1, workH = sh - winH + 1;
        const int step = (scale >= 3.0f) ? 3 : ((scale >= 2.0f) ? 2 : 1);
        const int nw = winW - 2, nh = winH - 2, narea = nw * nh;
        const int np0 = 1 + iiW * 1, np1 = (1 + nw) + iiW * 1;
        const int np2 = 1 + iiW * (1 + nh), np3 = (1 + nw) + iiW * (1 + nh);
        /* every window of the level (rows spread over the OpenMP threads),
         * then the passing ones emitted in (y, x) order */
        const int nRy = (workH + step - 1) / step, nRx = (workW + step - 1) / step;
        uint8_t* pf = (uint8_t*)malloc((size_t)nRy * nRx + 1);
        float* lastSum = (float*)malloc(sizeof(float) * ((size_t)nRy * nRx + 1));
#pragma omp parallel for schedule(dynamic, 4)
        for (int ry = 0; ry < nRy; ++ry) {
            for (int rx = 0; rx < nRx; ++rx) {
                const int base = ry * step * iiW + rx * step;
                pf[(size_t)ry * nRx + rx] = (uint8_t)det_window(features, stumps, stages, nStages, sum, sq, iiW,
                                                                base, mW, narea, np0, np1, np2, np3,
                                                                &lastSum[(size_t)ry * nRx + rx], stageEnter);
            }
        }
        for (int ry = 0; ry < nRy; ++ry) {
            for (int rx = 0; rx < nRx; ++rx) {
                ++nvis;
                if (!pf[(size_t)ry * nRx + rx]) continue;
                or_det d;
                d.x = (int)((float)(rx * step) * scale + 0.5f);
                d.y = (int)((float)(ry * step) * scale + 0.5f);
                d.w = (int)((float)winW * scale + 0.5f);
                d.h = (int)((float)winH * scale + 0.5f);
                d.scaleIdx = scaleIdx;
                d.score = lastSum[(size_t)ry * nRx + rx];
                if (count < cap) out[count] = d;
                ++count;
                if (count >= maxDet) {
                    free(pf);
                    free(lastSum);
                    goto done;
                }
            }
        }
        free(pf);
        free(lastSum);
        scale *= sf;
        ++scaleIdx;
    }
done:
    free(bufA);
    free(bufB);
    free(sum);
    free(sq);
    if (visited) *visited = nvis;
    return count < cap ? count : cap;
}

/* ---- grouping: src/face_vision_engine.cu:585-616, 1401-1539 ------------- */

static int dsu_find(int* p, int x) {
    int r = x;
    while (p[r] != r) r = p[r];
    while (p[x] != r) {
        const int nx = p[x];
        p[x] = r;
        x = nx;
    }
    return r;
}

static void dsu_unite(int* p, int* rk, int a, int b) {
    a = dsu_find(p, a);
    b = dsu_find(p, b);
    if (a == b) return;
    if (rk[a] < rk[b]) {
        const int t = a;
        a = b;
        b = t;
    }
    p[b] = a;
    if (rk[a] == rk[b]) ++rk[a];
}

static int similar(const or_det* a, const or_det* b, float eps) {
    const int delta = (int)(eps * (float)(MINI(a->w, b->w) + MINI(a->h, b->h)) * 0.5f);
    return abs(a->x - b->x) <= delta && abs(a->y - b->y) <= delta &&
           abs((a->x + a->w) - (b->x + b->w)) <= delta &&
           abs((a->y + a->h) - (b->y + b->h)) <= delta;
}

static const or_det* g_sort_dets;
static int cmp_by_x(const void* pa, const void* pb) {
    const int a = *(const int*)pa, b = *(const int*)pb;
    if (g_sort_dets[a].x != g_sort_dets[b].x) return g_sort_dets[a].x < g_sort_dets[b].x ? -1 : 1;
    return (a > b) - (a < b);
}

int or_group(const or_det* in, int n, int minNeighbors, float eps, or_det* out) {
    if (n <= 0) return 0;
    int* p = (int*)malloc(sizeof(int) * n);
    int* rk = (int*)calloc(n, sizeof(int));
    int* order = (int*)malloc(sizeof(int) * n);
    for (int i = 0; i < n; ++i) p[i] = order[i] = i;
    g_sort_dets = in;
    qsort(order, n, sizeof(int), cmp_by_x);
    for (int oi = 0; oi < n; ++oi) {
        const int i = order[oi];
        const or_det* a = &in[i];
        const int maxDx = (int)(eps * (float)(a->w + a->h) * 0.5f);
        for (int oj = oi + 1; oj < n; ++oj) {
            const int j = order[oj];
            const or_det* b = &in[j];
            if (b->x - a->x > maxDx) break;
            if (similar(a, b, eps)) dsu_unite(p, rk, i, j);
        }
    }
    int64_t* sx = (int64_t*)calloc(n, sizeof(int64_t));
    int64_t* sy = (int64_t*)calloc(n, sizeof(int64_t));
    int64_t* sw = (int64_t*)calloc(n, sizeof(int64_t));
    int64_t* sh = (int64_t*)calloc(n, sizeof(int64_t));
    int* cnt = (int*)calloc(n, sizeof(int));
    float* best = (float*)malloc(sizeof(float) * n);
    int* bidx = (int*)calloc(n, sizeof(int));
    for (int i = 0; i < n; ++i) best[i] = -INFINITY;
    for (int i = 0; i < n; ++i) {
        const int r = dsu_find(p, i);
        sx[r] += in[i].x;
        sy[r] += in[i].y;
        sw[r] += in[i].w;
        sh[r] += in[i].h;
        cnt[r] += 1;
        if (in[i].score > best[r]) {
            best[r] = in[i].score;
            bidx[r] = in[i].scaleIdx;
        }
    }
    or_det* cl = (or_det*)malloc(sizeof(or_det) * n);
    int* cc = (int*)malloc(sizeof(int) * n);
    int nc = 0;
    const int mn = MAXI(1, minNeighbors);
    for (int i = 0; i < n; ++i) {
        if (dsu_find(p, i) != i || cnt[i] <= 0) continue;
        if (cnt[i] <= mn) continue;
        or_det d;
        d.x = (int)(sx[i] / cnt[i]);
        d.y = (int)(sy[i] / cnt[i]);
        d.w = (int)(sw[i] / cnt[i]);
        d.h = (int)(sh[i] / cnt[i]);
        d.scaleIdx = bidx[i];
        d.score = best[i];
        cl[nc] = d;
        cc[nc] = cnt[i];
        ++nc;
    }
    int maxW = 0;
    for (int i = 0; i < nc; ++i) maxW = MAXI(maxW, cl[i].w);
    int* corder = (int*)malloc(sizeof(int) * (nc > 0 ? nc : 1));
    for (int i = 0; i < nc; ++i) corder[i] = i;
    g_sort_dets = cl;
    qsort(corder, nc, sizeof(int), cmp_by_x);
    int no = 0;
    for (int i = 0; i < nc; ++i) {
        const or_det* r1 = &cl[i];
        const int n1 = cc[i];
        int discard = 0;
        const float leftBound = (float)(r1->x + r1->w) - (1.0f + eps) * (float)maxW;
        const float rightBound = (float)r1->x + eps * (float)maxW;
        for (int k = 0; k < nc; ++k) {
            const int o = corder[k];
            const or_det* r2 = &cl[o];
            if (r2->x < (int)leftBound) continue;
            if (r2->x > (int)rightBound) break;
            if (o == i) continue;
            const int n2 = cc[o];
            const int dx = (int)((float)r2->w * eps);
            const int dy = (int)((float)r2->h * eps);
            if (r1->x >= r2->x - dx && r1->y >= r2->y - dy && r1->x + r1->w <= r2->x + r2->w + dx &&
                r1->y + r1->h <= r2->y + r2->h + dy && (n2 > MAXI(3, n1) || n1 < 3)) {
                discard = 1;
                break;
            }
        }
        if (!discard) out[no++] = *r1;
    }
    free(p);
    free(rk);
    free(order);
    free(sx);
    free(sy);
    free(sw);
    free(sh);
    free(cnt);
    free(best);
    free(bidx);
    free(cl);
    free(cc);
    free(corder);
    return no;
}

/* ---- training ------------------------------------------------------------ */

void or_train_integrals(const uint8_t* patches, int w, int h, int n,
                        uint32_t* sumT, uint64_t* sqT, float* invNorm) {
    const int iiW = w + 1, iiH = h + 1, area = iiW * iiH;
    uint32_t* s = (uint32_t*)malloc(sizeof(uint32_t) * area);
    uint64_t* q = (uint64_t*)malloc(sizeof(uint64_t) * area);
    for (int i = 0; i < n; ++i) {
        or_integral(patches + (size_t)i * w * h, w, h, w, s, q);
        for (int p = 0; p < area; ++p) {
            sumT[(size_t)p * n + i] = s[p];
            if (sqT) sqT[(size_t)p * n + i] = q[p];
        }
        invNorm[i] = or_train_invnorm(s, q, iiW, iiH);
    }
    free(s);
    free(q);
}

/* src/main.cu:476-496: corners to float, ((a-b)-c)+d, w0*s0 + w1*s1 (+ w2*s2),
 * times invNorm.  Every value before the last product is an exact integer. */
float or_train_response(const or_feature* f, const uint32_t* sumT, const float* invNorm,
                        int n, int s) {
#define RS(P) ((float)(int32_t)sumT[(size_t)(P) * n + s])
#if OR_MUTANT == 4
    const float s0 = RS(f->r[0].p0) - RS(f->r[0].p1) - RS(f->r[0].p2) + RS(f->r[0].p3 - 1);
#else
    const float s0 = RS(f->r[0].p0) - RS(f->r[0].p1) - RS(f->r[0].p2) + RS(f->r[0].p3);
#endif
    const float s1 = RS(f->r[1].p0) - RS(f->r[1].p1) - RS(f->r[1].p2) + RS(f->r[1].p3);
#if OR_MUTANT == 3
    float resp = f->r[0].w * s0 + f->r[0].w * s1;
#else
    float resp = f->r[0].w * s0 + f->r[1].w * s1;
#endif
    if (f->rectCount == 3) {
        const float s2 = RS(f->r[2].p0) - RS(f->r[2].p1) - RS(f->r[2].p2) + RS(f->r[2].p3);
        resp += f->r[2].w * s2;
    }
#undef RS
    return resp * invNorm[s];
}

void or_train_responses(const or_feature* f, const uint32_t* sumT, const float* invNorm, int n, float* out) {
    for (int s = 0; s < n; ++s) out[s] = or_train_response(f, sumT, invNorm, n, s);
}

static inline uint32_t float_key(float v) {
    uint32_t b;
    memcpy(&b, &v, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

static int cmp_u64(const void* a, const void* b) {
    const uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

void or_presort(const or_feature* features, int fBegin, int fCount,
                const uint32_t* sumT, const float* invNorm, int n,
                uint16_t* out16, uint32_t* out32) {
#pragma omp parallel
    {
        float* v = (float*)malloc(sizeof(float) * n);
        uint64_t* kk = (uint64_t*)malloc(sizeof(uint64_t) * n);
#pragma omp for schedule(dynamic, 16)
        for (int fi = 0; fi < fCount; ++fi) {
            const or_feature* f = &features[fBegin + fi];
            for (int s = 0; s < n; ++s) {
                v[s] = or_train_response(f, sumT, invNorm, n, s);
                kk[s] = ((uint64_t)float_key(v[s]) << 32) | (uint32_t)s;
            }
            qsort(kk, n, sizeof(uint64_t), cmp_u64);
            for (int k = 0; k < n; ++k) {
                const uint32_t s = (uint32_t)(kk[k] & 0xFFFFFFFFu);
                int valid = 0;
                if (k + 1 < n) {
                    const uint32_t s1 = (uint32_t)(kk[k + 1] & 0xFFFFFFFFu);
                    valid = (v[s] + 1e-7f < v[s1]);
                }
                if (out16) out16[(size_t)fi * n + k] = (uint16_t)(s | (valid ? 0x8000u : 0u));
                if (out32) out32[(size_t)fi * n + k] = s | (valid ? 0x80000000u : 0u);
            }
        }
        free(v);
        free(kk);
    }
}

int64_t or_fixed_weight(float w) {
    const double a = ldexp(fabs((double)w), 62);
    const int64_t q = (int64_t)a;
    return (w < 0.0f) ? -q : q;
}

/* Threshold sweep of src/boosting_core.cu:289-456 with exact sums:
 * D_k = sum_{j<=k} q(sw[ord[j]]), errN (parity +1) = totNeg + D_k,
 * errP (parity -1) = totPos - D_k; candidates where the valid bit is set;
 * per-feature order: smaller err, then smaller split, then parity +1. */
void or_search(const uint16_t* ord16, const uint32_t* ord32, int fCount, int n,
               const float* sw32, or_feature_best* out) {
    int64_t* q = (int64_t*)malloc(sizeof(int64_t) * n);
    int64_t totPos = 0, totNeg = 0;
    for (int s = 0; s < n; ++s) {
        q[s] = or_fixed_weight(sw32[s]);
        if (q[s] > 0) totPos += q[s];
        else totNeg -= q[s];
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int fi = 0; fi < fCount; ++fi) {
        or_feature_best b = {INT64_MAX, -1, +1};
        int64_t D = 0;
        for (int k = 0; k < n; ++k) {
            uint32_t e, s;
            int valid;
            if (ord16) {
                e = ord16[(size_t)fi * n + k];
                s = e & 0x7FFFu;
                valid = (e >> 15) & 1;
            } else {
                e = ord32[(size_t)fi * n + k];
                s = e & 0x7FFFFFFFu;
                valid = (int)(e >> 31);
            }
            D += q[s];
            if (!valid) continue;
            const int64_t errP = totPos - D;
            const int64_t errN = totNeg + D;
            if (errP < b.errQ) {
                b.errQ = errP;
                b.split = k;
                b.parity = -1;
            }
            if (errN < b.errQ || (errN == b.errQ && b.split == k)) {
                b.errQ = errN;
                b.split = k;
                b.parity = +1;
            }
        }
        out[fi] = b;
    }
    free(q);
}

int or_select(const or_feature_best* best, int fCount, const uint8_t* used, int64_t* errQ) {
    int bi = -1;
    int64_t be = INT64_MAX;
    for (int f = 0; f < fCount; ++f) {
        if (used && used[f]) continue;
        if (best[f].split < 0) continue;
        if (best[f].errQ < be) {
            be = best[f].errQ;
            bi = f;
        }
    }
    if (errQ) *errQ = be;
    return bi;
}

double or_exact_sum(const double* w, int n) {
    __int128 acc = 0;
    for (int i = 0; i < n; ++i) acc += (__int128)ldexp(w[i], 90);
    return ldexp((double)acc, -90);
}

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
