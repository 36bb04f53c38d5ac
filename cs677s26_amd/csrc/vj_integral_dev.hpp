// vj_integral_dev.hpp -- device routines for the transposed integrals of a
// batch of small images (training patches), shared by the engine's
// computeIntegralBatchTransposed and the trainer's stage start.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vj {

// Integral and square integral of sample s, layout [iiArea][n] (n == 1 is a
// row-major image).  One thread walks its sample; the row above is re-read
// from the output (same thread, program order).
__device__ inline void batch_integral_one(const uint8_t* __restrict__ gray, int w, int h, int stride, int n, int s,
                                          uint32_t* __restrict__ sumT, uint64_t* __restrict__ sqT) {
    const int iw = w + 1;
    const uint8_t* img = gray + static_cast<size_t>(s) * h * stride;
    for (int x = 0; x <= w; ++x) {
        sumT[static_cast<size_t>(x) * n + s] = 0u;
        sqT[static_cast<size_t>(x) * n + s] = 0u;
    }
    for (int y = 1; y <= h; ++y) {
        uint32_t rs = 0;
        uint64_t rq = 0;
        const size_t rowBase = static_cast<size_t>(y) * iw;
        sumT[rowBase * n + s] = 0u;
        sqT[rowBase * n + s] = 0u;
        for (int x = 1; x <= w; ++x) {
            const uint32_t p = img[static_cast<size_t>(y - 1) * stride + (x - 1)];
            rs += p;
            rq += p * p;
            const size_t i = (rowBase + x) * n + s;
            sumT[i] = sumT[i - static_cast<size_t>(iw) * n] + rs;
            sqT[i] = sqT[i - static_cast<size_t>(iw) * n] + rq;
        }
    }
}

// Training normaliser of sample s on the inset rect (1,1,iiW-3,iiH-3): int
// deltas, nf2 = area*sq - s*s in f64, (float)(1/sqrt(nf2)) or 0 when
// nf2 <= 1e-12 (src/face_vision_engine.cu:216-259, src/main.cu:457-472).
__device__ inline float train_invnorm(const uint32_t* __restrict__ sumT, const uint64_t* __restrict__ sqT, int iw,
                                      int ih, int n, int s) {
    if (iw < 3 || ih < 3) return 1.0f;
    const int ww = iw - 3, hh = ih - 3;
    const int area = ww * hh;
    const size_t p0 = 1 + static_cast<size_t>(iw), p1 = (1 + ww) + static_cast<size_t>(iw);
    const size_t p2 = 1 + static_cast<size_t>(iw) * (1 + hh), p3 = (1 + ww) + static_cast<size_t>(iw) * (1 + hh);
    const double ss = static_cast<double>(static_cast<int32_t>(sumT[p0 * n + s] - sumT[p1 * n + s] -
                                                               sumT[p2 * n + s] + sumT[p3 * n + s]));
    const double qq = static_cast<double>(static_cast<int64_t>(sqT[p0 * n + s] - sqT[p1 * n + s] -
                                                               sqT[p2 * n + s] + sqT[p3 * n + s]));
    const double nf2 = static_cast<double>(area) * qq - ss * ss;
    return (nf2 > 1e-12) ? static_cast<float>(1.0 / sqrt(nf2)) : 0.0f;
}

}  // namespace vj
