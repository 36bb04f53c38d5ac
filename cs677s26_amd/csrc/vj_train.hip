// vj_train.hip -- AdaBoostTrainer on gfx950: Haar responses over all
// samples, per-feature presort, exact weighted threshold scan, arg-min,
// weight update, normalisation and stage calibration.
//
// Reference: src/boosting_core.cu (AdaBoostTrainer, kernels K8-K12 and the
// CUB segmented sort), the round kernels src/main.cu:829-1028 and the round
// logic src/main.cu:2005-2348.
//
// Exactness.  Sums of weights are order-independent: a search weight is the
// float32 rounding of the f64 weight, turned into an int64 count of 2^-62
// units (exact for weights >= 2^-39), and prefix sums are exact integer sums.
// Normalisation divides by the exact sum of the f64 weights (int128 in 2^-90
// units, rounded to nearest once).  Ties therefore resolve identically on any
// schedule: per feature smaller error, then smaller split, then parity +1
// (src/boosting_core.cu:59-79); across features smaller error, then smaller
// index (src/boosting_core.cu:87-99).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <array>
#include <chrono>
#include <cstring>
#include <iterator>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "vj_common.hpp"
#include "vj_kernels.hpp"
#include "vj_integral_dev.hpp"

namespace vj {

constexpr int kSortThreads = 1024;
constexpr int kSortMaxPerThread = 20;              // N <= 20480
constexpr int kSortMaxN = kSortThreads * kSortMaxPerThread;
constexpr int kSearchThreads = 1024;               // 16 waves, one feature each at a time
constexpr size_t kSearchWideLds = 160 * 1024 - 2048; // dynamic LDS limit of the search kernels
constexpr int kPackedMaxN = 32767;                 // bit 15 of an order entry is its split flag

struct RoundBest {
    int64_t errQ;
    int32_t feature;
    int32_t split;
    int32_t parity;
    float theta;
    int32_t pad[2];
};

struct Calib {
    float cut;
    uint32_t posGe, posGt, negGe, negGt;
    int32_t pad[3];
};

struct WaveBest {
    int64_t errQ;
    int32_t split;
    int32_t parity;
    int32_t feature;
    int32_t pad;
};

__device__ __forceinline__ uint32_t float_key(float v) {
    const uint32_t b = __float_as_uint(v);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__device__ __forceinline__ float key_float(uint32_t k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

// trunc(|w| * 2^62) with the sign of w, exactly (w finite, |w| <= 1).
__device__ __forceinline__ int64_t fixed_weight(float w) {
    const uint32_t b = __float_as_uint(w);
    int e = static_cast<int>((b >> 23) & 0xFFu);
    uint64_t m = b & 0x7FFFFFu;
    if (e == 0) e = 1;
    else m |= 0x800000u;
    const int sh = min(e - 88, 39);
    const uint64_t q = sh >= 0 ? (m << sh) : (sh > -64 ? (m >> (-sh)) : 0ull);
    return (b >> 31) ? -static_cast<int64_t>(q) : static_cast<int64_t>(q);
}

// Integer sum of one rectangle of the transposed training integral.
__device__ __forceinline__ int32_t rect_sum(const uint32_t* __restrict__ sumT, const vj_rect& r, int n, int s) {
    return static_cast<int32_t>(sumT[static_cast<size_t>(r.p0) * n + s] - sumT[static_cast<size_t>(r.p1) * n + s] -
                                sumT[static_cast<size_t>(r.p2) * n + s] + sumT[static_cast<size_t>(r.p3) * n + s]);
}

// Normalised response of feature f on sample s: w0*S0 + w1*S1 (+ w2*S2) in
// float, times invNorm (src/main.cu:476-496).  Each rectangle sum is an exact
// integer below 2^24, so converting it once equals the reference's
// corner-by-corner float arithmetic.
__device__ __forceinline__ float response(const vj_feature& f, const uint32_t* __restrict__ sumT,
                                          const float* __restrict__ inv, int n, int s) {
    const float s0 = static_cast<float>(rect_sum(sumT, f.r[0], n, s));
    const float s1 = static_cast<float>(rect_sum(sumT, f.r[1], n, s));
    float resp = f.r[0].w * s0 + f.r[1].w * s1;
    if (f.rectCount == 3) resp += f.r[2].w * static_cast<float>(rect_sum(sumT, f.r[2], n, s));
    return resp * inv[s];
}

// ---------------------------------------------------------------------------
// Presort: one 1024-thread workgroup per feature, a stable LSD radix sort in
// LDS.  Keys are the order-preserving bit patterns of the responses (-0
// before +0, as the reference's radix sort orders them), rebased to the
// feature's smallest key so that only the bits in which the feature's keys
// differ are sorted, 8 per pass.  The sorted positions are padded to whole
// waves (nPad = NB * 1024) with the largest key, which sorts them after every
// real sample; stability keeps equal keys in sample order.  Each wave owns
// NB * 64 consecutive positions and ranks them 64 at a time: eight ballots
// give every lane the lanes holding the same digit, so a lane's rank is its
// count of same-digit lanes below it plus the wave's running count of that
// digit; one exclusive scan over (digit, wave) then places every element.
// Each lane computes the responses of the positions it ranks in the first
// pass (coalesced gathers of 64 consecutive samples).  Output per position
// k: the sample index, with bit 15 set when fl(v_k + 1e-7f) < v_{k+1}
// (src/boosting_core.cu:395-405).  mode 0 computes the responses from the
// integrals; mode 1 reads them from resp (row f of [count][n]) and writes
// plain indices (sortSamplesPerFeature semantics).  Workgroup i takes
// feature xcd_spread(i, count): the workgroups that one XCD runs together
// take consecutive features, whose corner rows mostly coincide, so the
// gathers hit that XCD's L2.
// ---------------------------------------------------------------------------
constexpr int kSortWaves = kSortThreads / kWave;

// The lanes of a full wavefront whose 8-bit digit equals this lane's: a lane
// differs from this one where some bit differs, so OR over the bits of
// (lanes with the bit set) XOR (all ones if this lane has it); the last bit's
// step takes the complement.
__device__ __forceinline__ uint64_t same_digit_lanes(uint32_t dg) {
    uint32_t xlo = 0u, xhi = 0u;
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        const uint32_t ext = static_cast<uint32_t>(__builtin_amdgcn_sbfe(static_cast<int>(dg), b, 1));
        // the ballot of ext != 0, compared on ext itself
        uint64_t bb;
        asm("v_cmp_ne_u32_e64 %0, 0, %1" : "=s"(bb) : "v"(ext));
        if (b < 7) {
            xlo = __builtin_amdgcn_bitop3_b32(ext, xlo, static_cast<uint32_t>(bb), 0xDE);   // xlo | (bb ^ ext)
            xhi = __builtin_amdgcn_bitop3_b32(ext, xhi, static_cast<uint32_t>(bb >> 32), 0xDE);
        } else {
            xlo = __builtin_amdgcn_bitop3_b32(ext, xlo, static_cast<uint32_t>(bb), 0x21);   // ~(xlo | (bb ^ ext))
            xhi = __builtin_amdgcn_bitop3_b32(ext, xhi, static_cast<uint32_t>(bb >> 32), 0x21);
        }
    }
    return (static_cast<uint64_t>(xhi) << 32) | xlo;
}
// Digit counters: one row of 256 per wave, rows kCntPitch words apart, so a
// wave's lanes reading the counters of distinct digits hit distinct banks.
constexpr int kCntPitch = 257;
static_assert(kSortWaves == 16 && 256 * kSortWaves == 4 * kSortThreads, "the counter scan takes 4 entries per thread");

// Block i of a count-block grid -> item: the hardware deals blocks to the 8
// XCDs in turn, so XCD x runs blocks x, x + 8, ...; they take the x-th
// contiguous eighth of [0, count).
__device__ __forceinline__ int xcd_spread(int i, int count) {
    const int q = count >> 3, r = count & 7, xcd = i & 7, pos = i >> 3;
    return xcd < r ? xcd * (q + 1) + pos : r * (q + 1) + (xcd - r) * q + pos;
}

// Keys of positions p0 + 64 i (i < NB) of feature ft with R rectangles
// (samples past n read sample n - 1; the caller replaces their keys).  The
// corner rows are uniform base pointers and a lane's sample a 32-bit byte
// offset from them; four positions' corners are loaded before any is used.
template <int NB, int R>
__device__ __forceinline__ void feature_keys(const vj_feature& ft, const uint32_t* __restrict__ sumT,
                                             const float* __restrict__ inv, int n, int p0, uint32_t (&kr)[NB]) {
    const uint8_t* c[R][4];
    float w[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        c[r][0] = reinterpret_cast<const uint8_t*>(sumT + static_cast<size_t>(ft.r[r].p0) * n);
        c[r][1] = reinterpret_cast<const uint8_t*>(sumT + static_cast<size_t>(ft.r[r].p1) * n);
        c[r][2] = reinterpret_cast<const uint8_t*>(sumT + static_cast<size_t>(ft.r[r].p2) * n);
        c[r][3] = reinterpret_cast<const uint8_t*>(sumT + static_cast<size_t>(ft.r[r].p3) * n);
        w[r] = ft.r[r].w;
    }
    const uint8_t* iv = reinterpret_cast<const uint8_t*>(inv);
    constexpr int G = 4;
#pragma unroll
    for (int i0 = 0; i0 < NB; i0 += G) {
        uint32_t v[G][R][4];
        float sc[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (i0 + g < NB) {
                const uint32_t off = static_cast<uint32_t>(min(p0 + (i0 + g) * kWave, n - 1)) * 4u;
#pragma unroll
                for (int r = 0; r < R; ++r)
#pragma unroll
                    for (int k = 0; k < 4; ++k) v[g][r][k] = *reinterpret_cast<const uint32_t*>(c[r][k] + off);
                sc[g] = *reinterpret_cast<const float*>(iv + off);
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            if (i0 + g < NB) {
                float s[R];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    s[r] = static_cast<float>(static_cast<int32_t>(v[g][r][0] - v[g][r][1] - v[g][r][2] + v[g][r][3]));
                // src/main.cu:476-496: w0*S0 + w1*S1 (+ w2*S2), then * invNorm
                float resp = w[0] * s[0] + w[1] * s[1];
                if (R == 3) resp += w[2] * s[2];
                kr[i0 + g] = float_key(resp * sc[g]);
            }
        }
    }
}

template <int NB>
__global__ __launch_bounds__(kSortThreads) void presort_kernel(const vj_feature* __restrict__ feats, int fBegin,
                                                               int count, const uint32_t* __restrict__ sumT,
                                                               const float* __restrict__ inv,
                                                               const float* __restrict__ resp, int n,
                                                               uint16_t* __restrict__ out, int64_t outPitch,
                                                               int mode) {
    extern __shared__ uint8_t smem[];
    constexpr int nPad = NB * kSortThreads;
    uint32_t* cnt = reinterpret_cast<uint32_t*>(smem);                   // [kSortWaves][kCntPitch]
    uint32_t* keys = cnt + kCntPitch * kSortWaves;                       // [nPad] rebased keys
    uint16_t* idx = reinterpret_cast<uint16_t*>(keys + nPad);            // [nPad]
    __shared__ uint32_t redMin[kSortWaves], redMax[kSortWaves], waveTot[kSortWaves];

    const int f = xcd_spread(blockIdx.x, count);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int base = wave * NB * kWave;           // this wave's positions
    // per element: the rebased key, and (sample index << 16) | rank
    uint32_t kr[NB], pk[NB];
    uint32_t kmin = 0xFFFFFFFFu, kmax = 0u;
    if (mode == 0) {
        const vj_feature ft = feats[fBegin + f];
        if (ft.rectCount == 3) feature_keys<NB, 3>(ft, sumT, inv, n, base + lane, kr);
        else feature_keys<NB, 2>(ft, sumT, inv, n, base + lane, kr);
    } else {
        const float* row = resp + static_cast<size_t>(f) * n;
#pragma unroll
        for (int i = 0; i < NB; ++i) kr[i] = float_key(row[min(base + i * kWave + lane, n - 1)]);
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        pk[i] = static_cast<uint32_t>(base + i * kWave + lane) << 16;
        if ((pk[i] >> 16) < static_cast<uint32_t>(n)) {
            kmin = min(kmin, kr[i]);
            kmax = max(kmax, kr[i]);
        }
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        kmin = min(kmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(kmin), d, kWave)));
        kmax = max(kmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(kmax), d, kWave)));
    }
    if (lane == 0) {
        redMin[wave] = kmin;
        redMax[wave] = kmax;
    }
    __syncthreads();
    kmin = redMin[0];
    kmax = redMax[0];
#pragma unroll
    for (int w = 1; w < kSortWaves; ++w) {
        kmin = min(kmin, redMin[w]);
        kmax = max(kmax, redMax[w]);
    }
    const uint32_t span = kmax - kmin;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        keys[base + i * kWave + lane] = (pk[i] >> 16) < static_cast<uint32_t>(n) ? kr[i] - kmin : span;
        idx[base + i * kWave + lane] = static_cast<uint16_t>(pk[i] >> 16);
    }
    __syncthreads();

    const int bits = span ? 32 - __clz(static_cast<int>(span)) : 0;
    uint32_t* wcnt = cnt + wave * kCntPitch;
    for (int shift = 0; shift < bits; shift += 8) {
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            kr[i] = keys[base + i * kWave + lane];
            pk[i] = static_cast<uint32_t>(idx[base + i * kWave + lane]) << 16;
        }
#pragma unroll
        for (int d = lane; d < 256; d += kWave) wcnt[d] = 0u;
        // rank: same-digit lanes below this one plus the wave's earlier
        // elements of the digit; every lane of a digit writes the same new
        // count
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const uint32_t dg = __builtin_amdgcn_ubfe(kr[i], static_cast<uint32_t>(shift), 8u);
            const uint64_t m = same_digit_lanes(dg);
            const uint32_t c = wcnt[dg];
            pk[i] |= c + static_cast<uint32_t>(mask_rank(m));
            wcnt[dg] = c + static_cast<uint32_t>(__popcll(m));
            // the rank is formed here: sunk to the scatter, every element's
            // mask and count would stay live across the scan
            asm volatile("" : "+v"(pk[i]));
        }
        __syncthreads();
        // exclusive scan over (digit, wave) in that order: 4 consecutive
        // entries per thread
        {
            uint32_t c4[4];
            uint32_t t = 0;
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = tid * 4 + k;
                c4[k] = cnt[(e & (kSortWaves - 1)) * kCntPitch + (e >> 4)];
                t += c4[k];
            }
            const uint32_t inc = wave_incl_scan_dpp(t);
            if (lane == 63) waveTot[wave] = inc;
            __syncthreads();
            uint32_t off = inc - t;
            for (int w = 0; w < wave; ++w) off += waveTot[w];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const int e = tid * 4 + k;
                cnt[(e & (kSortWaves - 1)) * kCntPitch + (e >> 4)] = off;
                off += c4[k];
            }
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const uint32_t dg = __builtin_amdgcn_ubfe(kr[i], static_cast<uint32_t>(shift), 8u);
            const uint32_t pos = wcnt[dg] + (pk[i] & 0xFFFFu);
            keys[pos] = kr[i];
            idx[pos] = static_cast<uint16_t>(pk[i] >> 16);
        }
        __syncthreads();
    }
    uint16_t* o = out + static_cast<int64_t>(f) * outPitch;
    if (mode != 0) {
        for (int k = tid; k < n; k += kSortThreads) o[k] = idx[k];
        return;
    }
    // four positions per lane: one 8-byte store of the entries (mode 0 rows
    // are padded to 256 entries, so o + k is 8-byte aligned), the split flags
    // from the keys of positions k .. k + 4
    for (int k = tid * 4; k < n; k += kSortThreads * 4) {
        const uint2 iv = *reinterpret_cast<const uint2*>(idx + k);
        const uint4 kv = *reinterpret_cast<const uint4*>(keys + k);
        const uint32_t k4 = k + 4 < n ? keys[k + 4] : 0u;
        const float v0 = key_float(kv.x + kmin), v1 = key_float(kv.y + kmin), v2 = key_float(kv.z + kmin),
                    v3 = key_float(kv.w + kmin), v4 = key_float(k4 + kmin);
        uint32_t lo = iv.x, hi = iv.y;
        if (k + 1 < n && v0 + 1e-7f < v1) lo |= 0x8000u;
        if (k + 2 < n && v1 + 1e-7f < v2) lo |= 0x80000000u;
        if (k + 3 < n && v2 + 1e-7f < v3) hi |= 0x8000u;
        if (k + 4 < n && v3 + 1e-7f < v4) hi |= 0x80000000u;
        if (k + 4 <= n) {
            *reinterpret_cast<uint2*>(o + k) = make_uint2(lo, hi);
        } else {
            o[k] = static_cast<uint16_t>(lo);      // k + 3 >= n here
            if (k + 1 < n) o[k + 1] = static_cast<uint16_t>(lo >> 16);
            if (k + 2 < n) o[k + 2] = static_cast<uint16_t>(hi);
        }
    }
}

// ---------------------------------------------------------------------------
// Presort above the LDS limit (kSortMaxN < n <= 65,535): the same stable LSD
// radix sort of rebased keys, 8 bits per pass, with the keys and indices in
// two global ping-pong buffers of the workgroup's slot, streamed through in
// tiles of kWideTile positions.  Each pass counts the digit histogram (the
// counts do not depend on the order), scans it into each digit's first
// position, then tile by tile every wave ranks its 256 consecutive positions
// with the same ballot ranks as the LDS sort, a scan over the waves of each
// digit places the tile's elements after the earlier tiles' elements of that
// digit, and the elements are written to the other buffer.  Persistent
// workgroups take features blockIdx.x, blockIdx.x + gridDim.x, ...  Output:
// mode 0 writes, per position k < outPitch, the sample index (0 from n on,
// so that every entry is a valid sample) and bit k of the feature's split
// mask row, set when fl(v_k + 1e-7f) < v_{k+1}; mode 1 writes plain indices
// for k < n (sortSamplesPerFeature).
// ---------------------------------------------------------------------------
constexpr int kWidePer = 4;                           // 64-position batches per wave per tile
constexpr int kWideTile = kSortThreads * kWidePer;    // 4096 positions
constexpr int kTrainMaxN = 65535;                     // u16 sample indices (src/boosting_core.cu:704-706)

__global__ __launch_bounds__(kSortThreads) void presort_wide_kernel(
        const vj_feature* __restrict__ feats, int fBegin, const uint32_t* __restrict__ sumT,
        const float* __restrict__ inv, const float* __restrict__ resp, int n, int nFeatures,
        uint16_t* __restrict__ out, int64_t outPitch, uint64_t* __restrict__ mask, int64_t maskPitch, int mode,
        uint32_t* __restrict__ scrKeys, uint16_t* __restrict__ scrIdx, int64_t scrPitch) {
    __shared__ uint32_t cnt[kSortWaves * kCntPitch];
    __shared__ uint32_t run[256];
    __shared__ uint32_t grp[4][256];
    __shared__ uint32_t redMin[kSortWaves], redMax[kSortWaves], wtot[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int nPad = (n + kWideTile - 1) / kWideTile * kWideTile;
    uint32_t* const kA = scrKeys + static_cast<int64_t>(blockIdx.x) * 2 * scrPitch;
    uint32_t* const kB = kA + scrPitch;
    uint16_t* const iA = scrIdx + static_cast<int64_t>(blockIdx.x) * 2 * scrPitch;
    uint16_t* const iB = iA + scrPitch;
    uint32_t* const wcnt = cnt + wave * kCntPitch;
    for (int f = blockIdx.x; f < nFeatures; f += gridDim.x) {
        uint32_t kmin = 0xFFFFFFFFu, kmax = 0u;
        if (mode == 0) {
            const vj_feature ft = feats[fBegin + f];
            for (int s = tid; s < n; s += kSortThreads) {
                const uint32_t k = float_key(response(ft, sumT, inv, n, s));
                kA[s] = k;
                iA[s] = static_cast<uint16_t>(s);
                kmin = min(kmin, k);
                kmax = max(kmax, k);
            }
        } else {
            const float* row = resp + static_cast<size_t>(f) * n;
            for (int s = tid; s < n; s += kSortThreads) {
                const uint32_t k = float_key(row[s]);
                kA[s] = k;
                iA[s] = static_cast<uint16_t>(s);
                kmin = min(kmin, k);
                kmax = max(kmax, k);
            }
        }
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            kmin = min(kmin, static_cast<uint32_t>(__shfl_xor(static_cast<int>(kmin), d, kWave)));
            kmax = max(kmax, static_cast<uint32_t>(__shfl_xor(static_cast<int>(kmax), d, kWave)));
        }
        if (lane == 0) {
            redMin[wave] = kmin;
            redMax[wave] = kmax;
        }
        __syncthreads();
        kmin = redMin[0];
        kmax = redMax[0];
        for (int w = 1; w < kSortWaves; ++w) {
            kmin = min(kmin, redMin[w]);
            kmax = max(kmax, redMax[w]);
        }
        // pads take the largest key: stability keeps them after every sample
        for (int s = n + tid; s < nPad; s += kSortThreads) {
            kA[s] = kmax;
            iA[s] = 0xFFFFu;
        }
        __syncthreads();
        const uint32_t span = kmax - kmin;
        const int bits = span ? 32 - __clz(static_cast<int>(span)) : 0;
        uint32_t* sk = kA;
        uint32_t* dk = kB;
        uint16_t* si = iA;
        uint16_t* di = iB;
        for (int shift = 0; shift < bits; shift += 8) {
            // digit histogram, each wave into its own counter row
            for (int d = lane; d < 256; d += kWave) wcnt[d] = 0u;
            for (int p = wave * kWave; p < nPad; p += kSortThreads) {
                const uint32_t dg = ((sk[p + lane] - kmin) >> shift) & 0xFFu;
                const uint64_t m = same_digit_lanes(dg);
                const uint32_t c = wcnt[dg];
                if (mask_rank(m) == 0) wcnt[dg] = c + static_cast<uint32_t>(__popcll(m));
            }
            __syncthreads();
            // each digit's first position: an exclusive scan of the totals
            uint32_t t = 0, inc = 0;
            if (tid < 256) {
                for (int w = 0; w < kSortWaves; ++w) t += cnt[w * kCntPitch + tid];
                inc = wave_incl_scan_dpp(t);
                if (lane == 63) wtot[wave] = inc;
            }
            __syncthreads();
            if (tid < 256) {
                uint32_t off = inc - t;
                for (int w = 0; w < wave; ++w) off += wtot[w];
                run[tid] = off;
            }
            __syncthreads();
            for (int t0 = 0; t0 < nPad; t0 += kWideTile) {
                for (int d = lane; d < 256; d += kWave) wcnt[d] = 0u;
                const int base = t0 + wave * (kWidePer * kWave);
                uint32_t kr[kWidePer], rk[kWidePer], dgs[kWidePer];
                uint16_t ir[kWidePer];
#pragma unroll
                for (int i = 0; i < kWidePer; ++i) {
                    kr[i] = sk[base + i * kWave + lane];
                    ir[i] = si[base + i * kWave + lane];
                }
#pragma unroll
                for (int i = 0; i < kWidePer; ++i) {
                    dgs[i] = ((kr[i] - kmin) >> shift) & 0xFFu;
                    const uint64_t m = same_digit_lanes(dgs[i]);
                    const uint32_t below = static_cast<uint32_t>(mask_rank(m));
                    const uint32_t c = wcnt[dgs[i]];
                    rk[i] = c + below;
                    if (below == 0) wcnt[dgs[i]] = c + static_cast<uint32_t>(__popcll(m));
                }
                __syncthreads();
                // per digit, four threads of four waves each: the tile's
                // elements of the earlier waves, after the earlier tiles'
                {
                    const int d = tid & 255, g = tid >> 8;
                    uint32_t v[4], gt = 0;
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        v[k] = cnt[(4 * g + k) * kCntPitch + d];
                        gt += v[k];
                    }
                    grp[g][d] = gt;
                    __syncthreads();
                    uint32_t off = run[d];
                    for (int gg = 0; gg < g; ++gg) off += grp[gg][d];
#pragma unroll
                    for (int k = 0; k < 4; ++k) {
                        cnt[(4 * g + k) * kCntPitch + d] = off;
                        off += v[k];
                    }
                    __syncthreads();
                    if (g == 3) run[d] = off;
                }
#pragma unroll
                for (int i = 0; i < kWidePer; ++i) {
                    const uint32_t pos = wcnt[dgs[i]] + rk[i];
                    dk[pos] = kr[i];
                    di[pos] = ir[i];
                }
                __syncthreads();  // counters, run and the written tile are complete
            }
            uint32_t* tk = sk;
            sk = dk;
            dk = tk;
            uint16_t* ti = si;
            si = di;
            di = ti;
        }
        uint16_t* o = out + static_cast<int64_t>(f) * outPitch;
        if (mode == 0) {
            uint64_t* mrow = mask + static_cast<int64_t>(f) * maskPitch;
            for (int k0 = wave * kWave; k0 < outPitch; k0 += kSortThreads) {
                const int k = k0 + lane;
                bool valid = false;
                uint16_t e = 0;
                if (k < n) {
                    e = si[k];
                    if (k + 1 < n) valid = key_float(sk[k]) + 1e-7f < key_float(sk[k + 1]);
                }
                o[k] = e;
                const uint64_t word = __ballot(valid);
                if (lane == 0) mrow[k0 >> 6] = word;
            }
        } else {
            for (int k = tid; k < n; k += kSortThreads) o[k] = si[k];
        }
        __syncthreads();  // the slot's buffers and the LDS serve the next feature
    }
}

// ---------------------------------------------------------------------------
// Threshold scan: every wave takes one feature at a time and walks its
// sorted order, 8 entries per lane per step; D_k = sum_{j<=k} q(order j)
// exactly in int64 (signed: +w positives, -w negatives, 0 inactive);
// errN (parity +1) = totNeg + D_k, errP (parity -1) = totPos - D_k at every
// position whose valid bit is set (src/boosting_core.cu:289-456).  So the
// best parity -1 split is the first valid position of largest D and the best
// parity +1 split the first valid position of smallest D; the two are
// compared at the end with the reference's tie rules.  The search weights
// sit in LDS packed one word per sample; each WG keeps the best of its
// features.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool better(int64_t e, int32_t k, int32_t p, int64_t be, int32_t bk, int32_t bp) {
    if (e != be) return e < be;
    if (k != bk) return k < bk;
    return p > bp;
}

// A search weight packed in one word: bits 0-24 hold the signed significand
// sm (|sm| < 2^24) and bits 25-31 a shift sh <= 39, with sm << sh equal to
// fixed_weight(w) exactly (significands of weights below 2^-39 are truncated
// first, as fixed_weight does).
__device__ __forceinline__ uint32_t pack_weight(float w) {
    const uint32_t b = __float_as_uint(w);
    int e = static_cast<int>((b >> 23) & 0xFFu);
    uint32_t m = b & 0x7FFFFFu;
    if (e == 0) e = 1;
    else m |= 0x800000u;
    int sh = min(e - 88, 39);
    if (sh < 0) {
        m = sh > -32 ? (m >> (-sh)) : 0u;
        sh = 0;
    }
    const int32_t sm = (b >> 31) ? -static_cast<int32_t>(m) : static_cast<int32_t>(m);
    return (static_cast<uint32_t>(sh) << 25) | (static_cast<uint32_t>(sm) & 0x1FFFFFFu);
}

__device__ __forceinline__ int64_t unpack_weight(uint32_t p) {
    const int32_t sm = static_cast<int32_t>(p << 7) >> 7;
    return static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(sm)) << (p >> 25));
}

// Inclusive wave64 prefix sum of an int64 with DPP moves of both halves
// (the same row shifts and row broadcasts as wave_incl_scan_dpp).
__device__ __forceinline__ int64_t wave_incl_scan_dpp64(int64_t v) {
#define VJ_DPP64_STEP(ctrl, rmask, bc)                                                                          \
    {                                                                                                           \
        const uint32_t lo = static_cast<uint32_t>(                                                              \
            __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint64_t>(v)), ctrl, rmask, 0xf, bc));  \
        const uint32_t hi = static_cast<uint32_t>(                                                              \
            __builtin_amdgcn_update_dpp(0, static_cast<int>(static_cast<uint64_t>(v) >> 32), ctrl, rmask, 0xf, bc)); \
        v += static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);                                      \
    }
    VJ_DPP64_STEP(0x111, 0xf, true)
    VJ_DPP64_STEP(0x112, 0xf, true)
    VJ_DPP64_STEP(0x114, 0xf, true)
    VJ_DPP64_STEP(0x118, 0xf, true)
    VJ_DPP64_STEP(0x142, 0xa, false)
    VJ_DPP64_STEP(0x143, 0xc, false)
#undef VJ_DPP64_STEP
    return v;
}

constexpr int kScanPer = 8;  // sorted entries per lane per step (one 16-byte load)

// A search weight as the scan reads it from LDS: a packed word, or (when the
// samples' int64 weights fit the workgroup's LDS) the int64 value itself.
__device__ __forceinline__ int64_t weight_value(uint32_t w) { return unpack_weight(w); }
__device__ __forceinline__ int64_t weight_value(int64_t w) { return w; }
__device__ __forceinline__ void store_weight(uint32_t* lsw, int s, uint32_t packed) { lsw[s] = packed; }
__device__ __forceinline__ void store_weight(int64_t* lsw, int s, uint32_t packed) { lsw[s] = unpack_weight(packed); }

// Exact threshold scan of one feature's sorted order by one wavefront; returns
// the feature's best (error, split, parity) in every lane (split -1: none).
// With kSplit false only the error is exact: the scan keeps the largest and
// smallest prefix but not where they occur, and fK is 0 when the feature has
// a split at all (the error does not depend on which split or parity wins a
// tie, so the search runs this form and rescans only its winners).
// v ? b : a for each lane v of the wave's lane mask (v_cndmask_b32 on both
// halves; VALU only).
__device__ __forceinline__ int64_t select_lanes(int64_t a, int64_t b, uint64_t mask) {
    uint32_t lo, hi;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(lo) : "v"(static_cast<uint32_t>(a)), "v"(static_cast<uint32_t>(b)), "s"(mask));
    asm("v_cndmask_b32 %0, %1, %2, %3"
        : "=v"(hi)
        : "v"(static_cast<uint32_t>(static_cast<uint64_t>(a) >> 32)), "v"(static_cast<uint32_t>(static_cast<uint64_t>(b) >> 32)),
          "s"(mask));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// kWide: entries are plain 16-bit sample indices and the split flags are the
// bits of the feature's mask row (byte k/8, bit k%8 for position k), read one
// byte per lane per step; otherwise bit 15 of an entry is its flag.
template <bool kSplit, bool kWide, int kAhead = 4, typename W>
__device__ __forceinline__ void scan_feature(const uint16_t* __restrict__ row, const uint8_t* __restrict__ mrow,
                                             int n, const W* lsw, int64_t totPos, int64_t totNeg, int lane,
                                             int64_t& fE, int32_t& fK, int32_t& fP) {
    constexpr uint32_t kIdxMask = kWide ? 0xFFFFu : 0x7FFFu;
    int64_t carry = 0;
    int64_t hiD = INT64_MIN, loD = INT64_MAX;
    int32_t hiK = INT_MAX, loK = INT_MAX;
    // one step of 64 * kScanPer sorted entries; kTail adds the bounds checks
    // that only the last, partial step needs (without them the eight LDS
    // reads of a lane issue together)
    auto step = [&](int base, const uint4 e8, const uint32_t mb, auto tailTag) {
        constexpr bool kTail = decltype(tailTag)::value;
        const int k0 = base + lane * kScanPer;
        const uint32_t ew[4] = {e8.x, e8.y, e8.z, e8.w};
        // all eight weight words are read before any is used
        W pw[kScanPer];
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) pw[j] = lsw[(ew[j >> 1] >> (16 * (j & 1))) & kIdxMask];
        int64_t q[kScanPer];
        int64_t c = 0;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            q[j] = (!kTail || k0 + j < n) ? weight_value(pw[j]) : 0;
            c += q[j];
        }
        const int64_t inc = wave_incl_scan_dpp64(c);
        int64_t D = carry + inc - c;
#pragma unroll
        for (int j = 0; j < kScanPer; ++j) {
            D += q[j];
            const uint32_t e = (ew[j >> 1] >> (16 * (j & 1))) & 0xFFFFu;
            // branch-free: the two selects chains are independent
            if constexpr (kSplit) {
                const bool flag = kWide ? ((mb >> j) & 1u) != 0u : (e & 0x8000u) != 0u;
                const bool valid = flag && (!kTail || k0 + j < n);
                const bool up = valid && D > hiD;
                const bool down = valid && D < loD;
                hiD = up ? D : hiD;
                loD = down ? D : loD;
                hiK = up ? k0 + j : hiK;
                loK = down ? k0 + j : loK;
            } else {
                // two selects per chain under lane masks formed straight from
                // the compares (written as plain selects, the compiler turns
                // them into a 64-bit max followed by a select on valid: four
                // per chain)
                uint64_t vm;
                if constexpr (kWide) {
                    vm = __builtin_amdgcn_sicmp(static_cast<int32_t>((mb >> j) & 1u), 0, 33 /* ne */);
                } else {
                    const int32_t ev = (j & 1) ? static_cast<int32_t>(ew[j >> 1])
                                               : static_cast<int32_t>(static_cast<int16_t>(ew[j >> 1] & 0xFFFFu));
                    vm = __builtin_amdgcn_sicmp(ev, 0, 40 /* slt: bit 15 of the entry */);
                }
                if constexpr (kTail) vm &= __builtin_amdgcn_sicmp(k0 + j, n, 40 /* slt */);
                hiD = select_lanes(hiD, D, vm & __builtin_amdgcn_sicmpl(D, hiD, 38 /* sgt */));
                loD = select_lanes(loD, D, vm & __builtin_amdgcn_sicmpl(D, loD, 40 /* slt */));
            }
        }
        const uint64_t last = static_cast<uint64_t>(inc);
        carry += static_cast<int64_t>(
            (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(last >> 32), 63)))
             << 32) |
            static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(last), 63)));
    };
    // whole steps, each step's entries loaded kAhead steps before it is
    // scanned (with one workgroup per CU the scan waits on HBM latency unless
    // several 1 KB loads per wave are in flight), then the partial last step
    // (rows are padded to 256 entries, so a lane's 16-byte load starting
    // below n stays inside its row)
    constexpr int kStep = kWave * kScanPer;
    auto load8 = [&](int base) { return *reinterpret_cast<const uint4*>(row + base + lane * kScanPer); };
    auto loadm = [&](int base) -> uint32_t {
        if constexpr (kWide) return mrow[(base >> 3) + lane];
        else return 0u;
    };
    const int nFull = n / kStep;
    uint4 buf[kAhead];
    uint32_t mbuf[kAhead];
#pragma unroll
    for (int u = 0; u < kAhead; ++u) {
        buf[u] = (u < nFull) ? load8(u * kStep) : make_uint4(0u, 0u, 0u, 0u);
        mbuf[u] = (u < nFull) ? loadm(u * kStep) : 0u;
    }
    int s0 = 0;
    for (; s0 + kAhead <= nFull; s0 += kAhead) {
#pragma unroll
        for (int u = 0; u < kAhead; ++u) {
            const uint4 cur = buf[u];
            const uint32_t curm = mbuf[u];
            const int nx = s0 + u + kAhead;
            buf[u] = (nx < nFull) ? load8(nx * kStep) : make_uint4(0u, 0u, 0u, 0u);
            mbuf[u] = (nx < nFull) ? loadm(nx * kStep) : 0u;
            step((s0 + u) * kStep, cur, curm, std::false_type{});
        }
    }
#pragma unroll
    for (int u = 0; u < kAhead; ++u)
        if (s0 + u < nFull) step((s0 + u) * kStep, buf[u], mbuf[u], std::false_type{});
    const int base = nFull * kStep;
    if (base < n) {
        const bool in = base + lane * kScanPer < n;
        const uint4 e8 = in ? load8(base) : make_uint4(0u, 0u, 0u, 0u);
        const uint32_t m8 = in ? loadm(base) : 0u;
        step(base, e8, m8, std::true_type{});
    }
    if constexpr (!kSplit) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) {
            const int64_t oH = __shfl_xor(hiD, d, kWave);
            const int64_t oL = __shfl_xor(loD, d, kWave);
            hiD = oH > hiD ? oH : hiD;
            loD = oL < loD ? oL : loD;
        }
        fE = INT64_MAX;
        fK = -1;
        fP = +1;
        if (loD != INT64_MAX) {
            const int64_t eN = totNeg + loD, eP = totPos - hiD;
            fE = eP < eN ? eP : eN;
            fK = 0;
        }
        return;
    }
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const int64_t oH = __shfl_xor(hiD, d, kWave);
        const int32_t oHK = __shfl_xor(hiK, d, kWave);
        if (oH > hiD || (oH == hiD && oHK < hiK)) {
            hiD = oH;
            hiK = oHK;
        }
        const int64_t oL = __shfl_xor(loD, d, kWave);
        const int32_t oLK = __shfl_xor(loK, d, kWave);
        if (oL < loD || (oL == loD && oLK < loK)) {
            loD = oL;
            loK = oLK;
        }
    }
    fE = INT64_MAX;
    fK = -1;
    fP = +1;
    if (loK != INT_MAX) {
        fE = totNeg + loD;
        fK = loK;
        const int64_t eP = totPos - hiD;
        if (better(eP, hiK, -1, fE, fK, fP)) {
            fE = eP;
            fK = hiK;
            fP = -1;
        }
    }
}

// Packs the search weights into LDS and returns the exact totals.
template <typename W>
__device__ __forceinline__ void load_weights(const float* __restrict__ sw32, int n, W* lsw, int64_t* totP,
                                             int64_t* totN, int64_t& totPos, int64_t& totNeg) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    int64_t tp = 0, tn = 0;
    for (int s = tid; s < n; s += blockDim.x) {
        const uint32_t pw = pack_weight(sw32[s]);
        store_weight(lsw, s, pw);
        const int64_t q = unpack_weight(pw);
        if (q > 0) tp += q;
        else tn -= q;
    }
    tp = wave_sum(tp);
    tn = wave_sum(tn);
    if (lane == 0) {
        totP[wave] = tp;
        totN[wave] = tn;
    }
    __syncthreads();
    totPos = 0;
    totNeg = 0;
    for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) {
        totPos += totP[k];
        totNeg += totN[k];
    }
}

// Exact pruning of the search (no counterpart in the reference, which scans
// every feature every round).  A feature's error is the sum of the search
// weights' magnitudes q over the samples its best stump misclassifies, and
// every round's weights are at least rho times the previous round's, sample
// by sample (rho: the smallest ratio, rounded down; normalize_kernel).  So a
// feature's error is at least rho times its error of the round before, and
// lb[f] (its last exact error, scaled down by every rho since and rounded
// down) bounds it from below.  A feature whose bound exceeds the smallest
// error found so far this round (gBest, which only falls to the round's
// minimum) can neither win nor tie, and is skipped; the chosen stump is the
// one the full scan finds.  A seed pass first scans the previous round's
// per-workgroup winners, so gBest starts near the round's minimum.
template <typename W, int kAhead>
__global__ __launch_bounds__(kSearchThreads) __attribute__((amdgpu_waves_per_eu(1, 4))) void search_kernel(const uint16_t* __restrict__ ord, int64_t pitch,
                                                               int n, int nFeatures,
                                                               const float* __restrict__ sw32,
                                                               const uint8_t* __restrict__ used,
                                                               WaveBest* __restrict__ blockBest,
                                                               double* __restrict__ lb, const double* __restrict__ rho,
                                                               unsigned long long* __restrict__ gBest,
                                                               const WaveBest* __restrict__ seeds, int nSeeds,
                                                               unsigned long long* __restrict__ scanned) {
    extern __shared__ int64_t lsmem[];
    W* lsw = reinterpret_cast<W*>(lsmem);  // [n] search weights
    __shared__ int64_t totP[16], totN[16];
    __shared__ WaveBest wb[16];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    int64_t totPos, totNeg;
    load_weights(sw32, n, lsw, totP, totN, totPos, totNeg);
    int64_t bE = INT64_MAX;
    int32_t bK = -1, bP = +1, bF = INT_MAX;
    const int nWavesTotal = gridDim.x * (blockDim.x >> 6);
    const bool prune = lb != nullptr;
    const double grow = prune && !seeds ? rho[0] : 0.0;
    const int count = seeds ? nSeeds : nFeatures;
    unsigned long long done = 0;
    unsigned long long g = prune ? __hip_atomic_load(gBest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0ull;
    // this wave's features k0, k0 + T, k0 + 2T, ... (T waves in the grid),
    // 64 at a time: lane j loads the j-th one's index, used flag and bound
    for (int k0 = blockIdx.x * (blockDim.x >> 6) + wave; k0 < count; k0 += kWave * nWavesTotal) {
        const int kk = k0 + lane * nWavesTotal;
        int fl = -1;
        if (kk < count) fl = seeds ? seeds[kk].feature : kk;
        if (fl >= nFeatures) fl = -1;
        const bool take = fl >= 0 && !(used && used[fl]);
        double bl = 0.0;
        if (take && prune && !seeds) bl = lb[fl] * grow * (1.0 - 0x1p-50);
        uint64_t todo = __ballot(take);
        while (todo) {
            const int j = __ffsll(static_cast<long long>(todo)) - 1;
            todo &= todo - 1;
            const int f = __builtin_amdgcn_readlane(fl, j);
            if (prune && !seeds) {
                const uint64_t bb = static_cast<uint64_t>(__double_as_longlong(bl));
                const double b = __longlong_as_double(static_cast<long long>(
                    (static_cast<uint64_t>(static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bb >> 32), j)))
                     << 32) |
                    static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(bb), j))));
                if (b > static_cast<double>(g) * (1.0 + 0x1p-40)) {
                    if (lane == 0) lb[f] = b;  // skipped: the bound, one round on
                    continue;
                }
            }
            // the round's smallest error so far, read again while this
            // feature is scanned (for the next one's test)
            const unsigned long long gNext =
                prune ? __hip_atomic_load(gBest, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ~0ull;
            int64_t fE;
            int32_t fK, fP;
            scan_feature<false, false, kAhead>(ord + static_cast<int64_t>(f) * pitch, nullptr, n, lsw, totPos, totNeg,
                                               lane, fE, fK, fP);
            ++done;
            g = gNext < g ? gNext : g;
            if (prune) {
                if (!seeds && lane == 0) lb[f] = static_cast<double>(fE) * (1.0 - 0x1p-50);
                if (fK >= 0 && static_cast<unsigned long long>(fE) < g) {
                    g = static_cast<unsigned long long>(fE);
                    if (lane == 0) atomicMin(gBest, g);
                }
            }
            if (!seeds && fK >= 0 && (fE < bE || (fE == bE && f < bF))) {
                bE = fE;
                bF = f;
            }
        }
    }
    if (scanned && lane == 0 && done) atomicAdd(scanned, done);
    if (seeds) return;  // the seed pass only lowers gBest (uniform over the workgroup)
    // the wave's winner again, now with its split and parity (same error)
    if (bF != INT_MAX) {
        int64_t fE;
        scan_feature<true, false, kAhead>(ord + static_cast<int64_t>(bF) * pitch, nullptr, n, lsw, totPos, totNeg, lane,
                                          fE, bK, bP);
    }
    if (lane == 0) wb[wave] = {bE, bK, bP, bF, 0};
    __syncthreads();
    if (tid == 0) {
        WaveBest r = wb[0];
        for (int k = 1; k < 16; ++k) {
            const WaveBest& o = wb[k];
            if (o.split >= 0 && (r.split < 0 || o.errQ < r.errQ || (o.errQ == r.errQ && o.feature < r.feature))) r = o;
        }
        blockBest[blockIdx.x] = r;
    }
}

// Search weights of the wide layout: each sample's exact fixed-point weight
// (the value of its packed word) in global memory, where the scan gathers it
// (the array stays in L2), and the exact positive and negative totals.  One
// workgroup.
__global__ __launch_bounds__(1024) void wide_weights_kernel(const float* __restrict__ sw32, int n,
                                                            int64_t* __restrict__ q, int64_t* __restrict__ tot) {
    __shared__ int64_t tp[16], tn[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t a = 0, b = 0;
    for (int s = threadIdx.x; s < n; s += blockDim.x) {
        const int64_t v = unpack_weight(pack_weight(sw32[s]));
        q[s] = v;
        if (v > 0) a += v;
        else b -= v;
    }
    a = wave_sum(a);
    b = wave_sum(b);
    if (lane == 0) {
        tp[wave] = a;
        tn[wave] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t A = 0, B = 0;
        for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) {
            A += tp[k];
            B += tn[k];
        }
        tot[0] = A;
        tot[1] = B;
    }
}

// search_kernel over the wide layout (plain indices, split mask rows, weights
// gathered from global memory).
template <int kAhead>
__global__ __launch_bounds__(kSearchThreads) void search_wide_kernel(const uint16_t* __restrict__ ord,
                                                                    const uint64_t* __restrict__ mask, int64_t pitch,
                                                                    int n, int nFeatures,
                                                                    const int64_t* __restrict__ q,
                                                                    const int64_t* __restrict__ tot,
                                                                    const uint8_t* __restrict__ used,
                                                                    WaveBest* __restrict__ blockBest) {
    __shared__ WaveBest wb[16];
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    const int64_t totPos = tot[0], totNeg = tot[1];
    const int64_t mPitch = pitch / 64;
    int64_t bE = INT64_MAX;
    int32_t bK = -1, bP = +1, bF = INT_MAX;
    const int nWavesTotal = gridDim.x * (blockDim.x >> 6);
    for (int f = blockIdx.x * (blockDim.x >> 6) + wave; f < nFeatures; f += nWavesTotal) {
        if (used && used[f]) continue;
        int64_t fE;
        int32_t fK, fP;
        scan_feature<false, true, kAhead>(ord + static_cast<int64_t>(f) * pitch,
                                  reinterpret_cast<const uint8_t*>(mask + static_cast<int64_t>(f) * mPitch), n, q,
                                  totPos, totNeg, lane, fE, fK, fP);
        if (fK >= 0 && (fE < bE || (fE == bE && f < bF))) {
            bE = fE;
            bF = f;
        }
    }
    if (bF != INT_MAX) {
        int64_t fE;
        scan_feature<true, true, kAhead>(ord + static_cast<int64_t>(bF) * pitch,
                                 reinterpret_cast<const uint8_t*>(mask + static_cast<int64_t>(bF) * mPitch), n, q,
                                 totPos, totNeg, lane, fE, bK, bP);
    }
    if (lane == 0) wb[wave] = {bE, bK, bP, bF, 0};
    __syncthreads();
    if (tid == 0) {
        WaveBest r = wb[0];
        for (int k = 1; k < 16; ++k) {
            const WaveBest& o = wb[k];
            if (o.split >= 0 && (r.split < 0 || o.errQ < r.errQ || (o.errQ == r.errQ && o.feature < r.feature))) r = o;
        }
        blockBest[blockIdx.x] = r;
    }
}

// Round arg-min over the workgroup bests; then the threshold of the winner
// from the two samples around its split, theta = 0.5f*(v_k + v_k+1).
__global__ __launch_bounds__(256) void reduce_best_kernel(const WaveBest* __restrict__ bb, int nb,
                                                          const uint16_t* __restrict__ ord, int64_t pitch,
                                                          const vj_feature* __restrict__ feats,
                                                          const uint32_t* __restrict__ sumT,
                                                          const float* __restrict__ inv, int n,
                                                          uint32_t idxMask, RoundBest* __restrict__ out) {
    __shared__ WaveBest sb[256];
    WaveBest r{INT64_MAX, -1, +1, INT_MAX, 0};
    for (int i = threadIdx.x; i < nb; i += blockDim.x) {
        const WaveBest o = bb[i];
        if (o.split >= 0 && (r.split < 0 || o.errQ < r.errQ || (o.errQ == r.errQ && o.feature < r.feature))) r = o;
    }
    sb[threadIdx.x] = r;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            const WaveBest o = sb[threadIdx.x + st];
            WaveBest& a = sb[threadIdx.x];
            if (o.split >= 0 && (a.split < 0 || o.errQ < a.errQ || (o.errQ == a.errQ && o.feature < a.feature))) a = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const WaveBest b = sb[0];
        RoundBest rb{};
        rb.errQ = b.errQ;
        rb.feature = b.split >= 0 ? b.feature : -1;
        rb.split = b.split;
        rb.parity = b.parity;
        rb.theta = 0.0f;
        if (b.split >= 0 && b.split + 1 < n) {
            const uint16_t* row = ord + static_cast<int64_t>(b.feature) * pitch;
            const int s0 = row[b.split] & idxMask, s1 = row[b.split + 1] & idxMask;
            const vj_feature f = feats[b.feature];
            const float v0 = response(f, sumT, inv, n, s0);
            const float v1 = response(f, sumT, inv, n, s1);
            rb.theta = 0.5f * (v0 + v1);
        }
        *out = rb;
    }
}

// Stage-start integrals and normalisers of the n 24x24 patches.
__global__ __launch_bounds__(256) void stage_integral_kernel(const uint8_t* __restrict__ patches, int w, int h, int n,
                                                             uint32_t* __restrict__ sumT, uint64_t* __restrict__ sqT,
                                                             float* __restrict__ inv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    batch_integral_one(patches, w, h, w, n, s, sumT, sqT);
    inv[s] = train_invnorm(sumT, sqT, w + 1, h + 1, n, s);
}

// Weak update (src/main.cu:829-883): strong += (resp < theta) ? L : R for
// every sample; active misclassified samples get w *= factor, factor =
// exp(C) computed on the host in f64.
__global__ __launch_bounds__(256) void apply_kernel(const vj_feature* __restrict__ feats, int feature,
                                                    const uint32_t* __restrict__ sumT,
                                                    const float* __restrict__ inv, int n,
                                                    const int8_t* __restrict__ label,
                                                    const uint8_t* __restrict__ active,
                                                    double* __restrict__ w, float* __restrict__ strong,
                                                    float theta, int parity, float left, float right,
                                                    double factor, uint8_t* __restrict__ used) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0) used[feature] = 1;
    if (s >= n) return;
    const vj_feature f = feats[feature];
    const float r = response(f, sumT, inv, n, s);
    const bool lt = r < theta;
    strong[s] += lt ? left : right;
    if (!active[s]) return;
    const int pred = (parity > 0) ? (lt ? -1 : +1) : (lt ? +1 : -1);
    if (pred != static_cast<int>(label[s])) w[s] *= factor;
}

// Exact fixed-point value of a non-negative double in 2^-90 units (int128 as
// hi:lo), truncated like a C cast of ldexp(w, 90).
__device__ __forceinline__ void fixed90(double w, uint64_t& lo, uint64_t& hi) {
    lo = hi = 0;
    const uint64_t b = static_cast<uint64_t>(__double_as_longlong(w));
    int e = static_cast<int>((b >> 52) & 0x7FFu);
    uint64_t m = b & 0xFFFFFFFFFFFFFull;
    if (e == 0) e = 1;
    else m |= 0x10000000000000ull;
    const int sh = e - 985;  // value = m * 2^(e-1075) * 2^90
    if (sh >= 0) {
        if (sh < 64) {
            lo = m << sh;
            hi = sh ? (m >> (64 - sh)) : 0;
        } else if (sh < 128) {
            hi = m << (sh - 64);
        }
    } else if (sh > -64) {
        lo = m >> (-sh);
    }
}

// Nearest double (ties to even) to the non-negative int128 hi:lo.
__device__ __forceinline__ double u128_to_double(uint64_t lo, uint64_t hi) {
    if (!hi && !lo) return 0.0;
    int p = hi ? 127 - __clzll(hi) : 63 - __clzll(lo);  // msb position
    if (p <= 52) return static_cast<double>(lo);
    const int sh = p - 52;
    uint64_t mant, rem, half;
    // mant = v >> sh, rem = v & ((1<<sh)-1)
    if (sh < 64) {
        mant = (lo >> sh) | (sh ? (hi << (64 - sh)) : 0);
        rem = sh ? (lo & ((1ull << sh) - 1)) : 0;
        // bits of hi below sh-64 do not exist when sh < 64
        half = 1ull << (sh - 1);
        const bool up = rem > half || (rem == half && (mant & 1ull));
        if (up) ++mant;
    } else {
        const int s2 = sh - 64;
        mant = hi >> s2;
        const uint64_t remHi = s2 ? (hi & ((1ull << s2) - 1)) : 0;
        // compare remHi:lo with the half point (1 << (sh-1))
        bool up;
        if (s2 == 0) {
            up = lo > (1ull << 63) || (lo == (1ull << 63) && (mant & 1ull));
        } else {
            const uint64_t halfHi = 1ull << (s2 - 1);
            up = remHi > halfHi || (remHi == halfHi && (lo > 0 || (mant & 1ull)));
        }
        if (up) ++mant;
    }
    return ldexp(static_cast<double>(mant), sh);
}

// Weight normalisation (src/main.cu:1725-1740, kernels K14/K15): inactive
// weights to 0, exact sum, divide when the sum exceeds 1e-30; also writes
// the next round's float search weights (+w positive, -w negative, 0
// inactive).  One workgroup.
__global__ __launch_bounds__(1024) void normalize_kernel(double* __restrict__ w, const int8_t* __restrict__ label,
                                                         const uint8_t* __restrict__ active, int n,
                                                         float* __restrict__ sw32, double* __restrict__ rho,
                                                         unsigned long long* __restrict__ gBest) {
    __shared__ uint64_t sLo[16], sHi[16];
    __shared__ double sS;
    __shared__ double sRho[16];
    // kNormU samples per thread per step, all loaded before any is used (one
    // workgroup: without loads in flight the loop waits on each one)
    constexpr int kNormU = 8;
    const int bd = static_cast<int>(blockDim.x);
    uint64_t lo = 0, hi = 0;
    for (int s0 = threadIdx.x; s0 < n; s0 += bd * kNormU) {
        double wv[kNormU];
        uint8_t av[kNormU];
#pragma unroll
        for (int u = 0; u < kNormU; ++u) {
            const int s = s0 + u * bd;
            wv[u] = s < n ? w[s] : 0.0;
            av[u] = s < n ? active[s] : 1;
        }
#pragma unroll
        for (int u = 0; u < kNormU; ++u) {
            const int s = s0 + u * bd;
            if (!av[u]) {
                w[s] = 0.0;
                wv[u] = 0.0;
            }
            uint64_t a, b;
            fixed90(wv[u], a, b);
            const uint64_t nl = lo + a;
            hi += b + (nl < lo ? 1ull : 0ull);
            lo = nl;
        }
    }
    // wave reduction of the 128-bit partial sums
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) {
        const uint64_t ol = __shfl_xor(lo, d, kWave), oh = __shfl_xor(hi, d, kWave);
        const uint64_t nl = lo + ol;
        hi += oh + (nl < lo ? 1ull : 0ull);
        lo = nl;
    }
    if ((threadIdx.x & 63) == 0) {
        sLo[threadIdx.x >> 6] = lo;
        sHi[threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t L = 0, H = 0;
        for (int k = 0; k < static_cast<int>(blockDim.x >> 6); ++k) {
            const uint64_t nl = L + sLo[k];
            H += sHi[k] + (nl < L ? 1ull : 0ull);
            L = nl;
        }
        sS = ldexp(u128_to_double(L, H), -90);
    }
    __syncthreads();
    const double S = sS;
    // the smallest ratio of a sample's new search-weight magnitude (as the
    // scan adds it) to its previous one: the search's pruning factor
    double rmin = INFINITY;
    for (int s0 = threadIdx.x; s0 < n; s0 += bd * kNormU) {
        double wv[kNormU];
        uint8_t av[kNormU];
        int8_t lv[kNormU];
        float ov[kNormU];
#pragma unroll
        for (int u = 0; u < kNormU; ++u) {
            const int s = s0 + u * bd;
            wv[u] = s < n ? w[s] : 0.0;
            av[u] = s < n ? active[s] : 0;
            lv[u] = s < n ? label[s] : 0;
            ov[u] = s < n && rho ? sw32[s] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < kNormU; ++u) {
            const int s = s0 + u * bd;
            if (s < n) {
                double x = wv[u];
                if (S > 1e-30) x = x / S;
                w[s] = x;
                const float f = static_cast<float>(x);
                const float nw = av[u] ? (lv[u] > 0 ? f : -f) : 0.0f;
                sw32[s] = nw;
                if (rho) {
                    const int64_t qo = unpack_weight(pack_weight(ov[u])), qn = unpack_weight(pack_weight(nw));
                    const double ao = static_cast<double>(qo < 0 ? -qo : qo), an = static_cast<double>(qn < 0 ? -qn : qn);
                    if (ao > 0.0) rmin = fmin(rmin, an / ao);
                }
            }
        }
    }
    if (rho) {
#pragma unroll
        for (int d = 32; d > 0; d >>= 1) rmin = fmin(rmin, __shfl_xor(rmin, d, kWave));
        if ((threadIdx.x & 63) == 0) sRho[threadIdx.x >> 6] = rmin;
        __syncthreads();
        if (threadIdx.x == 0) {
            double r = sRho[0];
            for (int k = 1; k < static_cast<int>(blockDim.x >> 6); ++k) r = fmin(r, sRho[k]);
            // each ratio is within 3 units of 2^-53 of the exact one: the
            // margin keeps rho at or below every exact ratio
            rho[0] = r < INFINITY ? r * (1.0 - 0x1p-40) : 0.0;
            gBest[0] = ~0ull;
        }
    }
}

// Fills a[0..n) with v (the search's error bounds at a stage start).
__global__ __launch_bounds__(256) void fill_f64_kernel(double* __restrict__ a, int n, double v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Stage weight reset (src/main.cu:908-924): positives 0.5/numPos, active
// negatives 0.5/activeNeg, others 0; strong scores 0.
__global__ __launch_bounds__(256) void reset_kernel(double* __restrict__ w, float* __restrict__ strong,
                                                    const uint8_t* __restrict__ active, int n, int numPos,
                                                    double wPos, double wNeg) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    strong[s] = 0.0f;
    w[s] = (s < numPos) ? wPos : (active[s] ? wNeg : 0.0);
}

// Stage-threshold calibration (src/main.cu:2172-2226): the cutIdx-th
// smallest positive strong score by radix select over the order-preserving
// keys (the reference sorts them with a radix sort), then the four counts
// pos >= cut, pos > cut, active neg >= cut, active neg > cut.
__global__ __launch_bounds__(1024) void calibrate_kernel(const float* __restrict__ strong,
                                                         const uint8_t* __restrict__ active, int n,
                                                         int numPos, int cutIdx, Calib* __restrict__ out) {
    __shared__ uint32_t hist[256];
    __shared__ uint32_t sPrefix, sMask, sK;
    __shared__ uint32_t red[4][16];
    __shared__ uint32_t binTot[4];
    if (threadIdx.x == 0) {
        sPrefix = 0;
        sMask = 0;
        sK = static_cast<uint32_t>(cutIdx);
    }
    constexpr int kCalU = 8;  // scores per thread per step, all loaded before any is used
    const int bd = static_cast<int>(blockDim.x);
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63;
    for (int pass = 3; pass >= 0; --pass) {
        const int shift = pass * 8;
        for (int b = tid; b < 256; b += bd) hist[b] = 0;
        __syncthreads();
        const uint32_t prefix = sPrefix, mask = sMask;
        for (int i0 = tid; i0 < numPos; i0 += bd * kCalU) {
            uint32_t kv[kCalU];
#pragma unroll
            for (int u = 0; u < kCalU; ++u) {
                const int i = i0 + u * bd;
                kv[u] = i < numPos ? float_key(strong[i]) : 0u;
            }
#pragma unroll
            for (int u = 0; u < kCalU; ++u) {
                const int i = i0 + u * bd;
                const bool in = i < numPos && (kv[u] & mask) == prefix;
                const uint32_t dg = (kv[u] >> shift) & 255u;
                // the lanes sharing the first participating lane's digit add
                // with one atomic (the leading digits of the scores mostly
                // agree); the others add one by one
                const uint64_t part = __ballot(in);
                if (!part) continue;
                const int lead = __ffsll(static_cast<long long>(part)) - 1;
                const uint32_t d0 = static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(dg), lead));
                const uint64_t same = __ballot(in && dg == d0);
                if (lane == lead) atomicAdd(&hist[d0], static_cast<uint32_t>(__popcll(same)));
                if (in && dg != d0) atomicAdd(&hist[dg], 1u);
            }
        }
        __syncthreads();
        // the digit holding the k-th remaining score: an exclusive scan over
        // the 256 bins (one per thread of the first four waves)
        uint32_t h = 0, inc = 0;
        if (tid < 256) {  // whole waves
            h = hist[tid];
            inc = wave_incl_scan_dpp(h);
            if (lane == 63) binTot[tid >> 6] = inc;
        }
        const uint32_t kk = sK;
        __syncthreads();
        if (tid < 256) {
            uint32_t before = inc - h;
            for (int wv = 0; wv < (tid >> 6); ++wv) before += binTot[wv];
            // the bin with before <= kk < before + h; bin 255 takes the rest
            // (the intervals of the bins are disjoint, so one thread writes)
            if (kk >= before && (kk - before < h || tid == 255)) {
                sK = kk - before;
                sPrefix = prefix | (static_cast<uint32_t>(tid) << shift);
                sMask = mask | (255u << shift);
            }
        }
        __syncthreads();
    }
    const float cut = key_float(sPrefix);
    uint32_t c[4] = {0, 0, 0, 0};
    for (int i0 = tid; i0 < n; i0 += bd * kCalU) {
        float sv[kCalU];
        uint8_t av[kCalU];
#pragma unroll
        for (int u = 0; u < kCalU; ++u) {
            const int i = i0 + u * bd;
            sv[u] = i < n ? strong[i] : 0.0f;
            av[u] = (i < n && i >= numPos) ? active[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < kCalU; ++u) {
            const int i = i0 + u * bd;
            const uint32_t ge = sv[u] >= cut, gt = sv[u] > cut;
            if (i < numPos) {
                c[0] += ge;
                c[1] += gt;
            } else if (i < n && av[u]) {
                c[2] += ge;
                c[3] += gt;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        c[k] = wave_sum(c[k]);
        if ((threadIdx.x & 63) == 0) red[k][threadIdx.x >> 6] = c[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Calib r{};
        r.cut = cut;
        uint32_t t[4] = {0, 0, 0, 0};
        for (int k = 0; k < 4; ++k)
            for (int wv = 0; wv < static_cast<int>(blockDim.x >> 6); ++wv) t[k] += red[k][wv];
        r.posGe = t[0];
        r.posGt = t[1];
        r.negGe = t[2];
        r.negGt = t[3];
        *out = r;
    }
}

// Stage end (src/main.cu:960-983): negatives stay active iff strong >= threshold.
__global__ __launch_bounds__(256) void stage_end_kernel(const float* __restrict__ strong, uint8_t* __restrict__ active,
                                                        int n, int numPos, float threshold,
                                                        uint32_t* __restrict__ count) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t keep = 0;
    if (s < n && s >= numPos && active[s]) {
        keep = strong[s] >= threshold;
        active[s] = static_cast<uint8_t>(keep);
    }
    keep = wave_sum(keep);
    if ((threadIdx.x & 63) == 0 && keep) atomicAdd(count, keep);
}

// Per-op API kernels --------------------------------------------------------

// evaluateFeatureResponses (src/boosting_core.cu:221-261): out[f][s].
__global__ __launch_bounds__(256) void responses_kernel(const vj_feature* __restrict__ feats, int begin,
                                                        const uint32_t* __restrict__ sumT,
                                                        const float* __restrict__ inv, int n,
                                                        float* __restrict__ out) {
    const int f = blockIdx.y;
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const vj_feature ft = feats[begin + f];
    out[static_cast<size_t>(f) * n + s] = response(ft, sumT, inv, n, s);
}

// Search weights from the reference's float weights, labels and active mask.
__global__ __launch_bounds__(256) void search_weights_kernel(const float* __restrict__ w,
                                                             const int8_t* __restrict__ label,
                                                             const uint8_t* __restrict__ active, int n,
                                                             float* __restrict__ sw32) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const bool a = active == nullptr || active[s] != 0;
    sw32[s] = a ? (label[s] > 0 ? w[s] : -w[s]) : 0.0f;
}

// Marks valid splits of a plain sorted-index array from the responses.
__global__ __launch_bounds__(256) void mark_valid_kernel(const float* __restrict__ resp,
                                                         const uint16_t* __restrict__ sorted, int n,
                                                         uint16_t* __restrict__ out, int64_t pitch) {
    const int f = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const float* row = resp + static_cast<size_t>(f) * n;
    const uint16_t* sr = sorted + static_cast<size_t>(f) * n;
    uint32_t e = sr[k];
    if (k + 1 < n && row[sr[k]] + 1e-7f < row[sr[k + 1]]) e |= 0x8000u;
    out[static_cast<int64_t>(f) * pitch + k] = static_cast<uint16_t>(e);
}

// FeatureBestSplit per feature from per-feature scan results.
struct FeatureBestSplit {
    float bestErr;
    float bestTheta;
    int32_t bestSplitPos;
    int8_t bestParity, pad0, pad1, pad2;
};
struct TileBestCandidate {
    float bestErr;
    float bestTheta;
    int32_t featureIdx;
    int8_t bestParity, pad0, pad1, pad2;
};

__global__ __launch_bounds__(kSearchThreads) void search_each_kernel(const uint16_t* __restrict__ ord, int64_t pitch,
                                                                    int n, int nFeatures,
                                                                    const float* __restrict__ sw32,
                                                                    const float* __restrict__ resp,
                                                                    FeatureBestSplit* __restrict__ out) {
    extern __shared__ uint32_t lsw[];  // [n] packed search weights
    __shared__ int64_t totP[16], totN[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t totPos, totNeg;
    load_weights(sw32, n, lsw, totP, totN, totPos, totNeg);
    const int nWavesTotal = gridDim.x * (blockDim.x >> 6);
    for (int f = blockIdx.x * (blockDim.x >> 6) + wave; f < nFeatures; f += nWavesTotal) {
        const uint16_t* row = ord + static_cast<int64_t>(f) * pitch;
        int64_t fE;
        int32_t fK, fP;
        scan_feature<true, false>(row, nullptr, n, lsw, totPos, totNeg, lane, fE, fK, fP);
        if (lane == 0) {
            FeatureBestSplit r{};
            if (fK >= 0 && fK + 1 < n) {
                r.bestErr = static_cast<float>(ldexp(static_cast<double>(fE), -62));
                r.bestSplitPos = fK;
                r.bestParity = static_cast<int8_t>(fP);
                const float* rr = resp + static_cast<size_t>(f) * n;
                r.bestTheta = 0.5f * (rr[row[fK] & 0x7FFF] + rr[row[fK + 1] & 0x7FFF]);
            } else {
                r.bestErr = FLT_MAX;
                r.bestTheta = 0.0f;
                r.bestSplitPos = -1;
                r.bestParity = +1;
            }
            out[f] = r;
        }
    }
}

// mark_valid_kernel for the wide layout: rows padded to pitch (entries 0 past
// n), split flags as mask bits.  The grid covers pitch positions per row.
__global__ __launch_bounds__(256) void mark_valid_wide_kernel(const float* __restrict__ resp,
                                                              const uint16_t* __restrict__ sorted, int n,
                                                              uint16_t* __restrict__ out, uint64_t* __restrict__ mask,
                                                              int64_t pitch) {
    const int f = blockIdx.y;
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const float* row = resp + static_cast<size_t>(f) * n;
    const uint16_t* sr = sorted + static_cast<size_t>(f) * n;
    uint16_t e = 0;
    bool valid = false;
    if (k < n) {
        e = sr[k];
        if (k + 1 < n) valid = row[sr[k]] + 1e-7f < row[sr[k + 1]];
    }
    out[static_cast<int64_t>(f) * pitch + k] = e;
    const uint64_t word = __ballot(valid);
    if ((threadIdx.x & 63) == 0) mask[static_cast<int64_t>(f) * (pitch / 64) + k / 64] = word;
}

// search_each_kernel over the wide layout.
__global__ __launch_bounds__(kSearchThreads) void search_each_wide_kernel(
        const uint16_t* __restrict__ ord, const uint64_t* __restrict__ mask, int64_t pitch, int n, int nFeatures,
        const int64_t* __restrict__ q, const int64_t* __restrict__ tot, const float* __restrict__ resp,
        FeatureBestSplit* __restrict__ out) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t totPos = tot[0], totNeg = tot[1];
    const int nWavesTotal = gridDim.x * (blockDim.x >> 6);
    for (int f = blockIdx.x * (blockDim.x >> 6) + wave; f < nFeatures; f += nWavesTotal) {
        const uint16_t* row = ord + static_cast<int64_t>(f) * pitch;
        int64_t fE;
        int32_t fK, fP;
        scan_feature<true, true>(row, reinterpret_cast<const uint8_t*>(mask + static_cast<int64_t>(f) * (pitch / 64)),
                                 n, q, totPos, totNeg, lane, fE, fK, fP);
        if (lane == 0) {
            FeatureBestSplit r{};
            if (fK >= 0 && fK + 1 < n) {
                r.bestErr = static_cast<float>(ldexp(static_cast<double>(fE), -62));
                r.bestSplitPos = fK;
                r.bestParity = static_cast<int8_t>(fP);
                const float* rr = resp + static_cast<size_t>(f) * n;
                r.bestTheta = 0.5f * (rr[row[fK]] + rr[row[fK + 1]]);
            } else {
                r.bestErr = FLT_MAX;
                r.bestTheta = 0.0f;
                r.bestSplitPos = -1;
                r.bestParity = +1;
            }
            out[f] = r;
        }
    }
}

// selectBestInTile (src/boosting_core.cu:155-192): masked arg-min, smaller
// index on ties.
__global__ __launch_bounds__(256) void select_kernel(const FeatureBestSplit* __restrict__ best, int begin,
                                                     int count, const uint8_t* __restrict__ used,
                                                     TileBestCandidate* __restrict__ out) {
    __shared__ float sE[256];
    __shared__ int sF[256];
    __shared__ float sT[256];
    __shared__ int sP[256];
    float e = FLT_MAX, th = 0.0f;
    int fi = INT_MAX, par = +1;
    for (int i = threadIdx.x; i < count; i += blockDim.x) {
        const int f = begin + i;
        if (used && used[f]) continue;
        const FeatureBestSplit b = best[i];
        if (b.bestErr < e || (b.bestErr == e && f < fi)) {
            e = b.bestErr;
            th = b.bestTheta;
            fi = f;
            par = b.bestParity;
        }
    }
    sE[threadIdx.x] = e;
    sF[threadIdx.x] = fi;
    sT[threadIdx.x] = th;
    sP[threadIdx.x] = par;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            const int o = threadIdx.x + st;
            if (sE[o] < sE[threadIdx.x] || (sE[o] == sE[threadIdx.x] && sF[o] < sF[threadIdx.x])) {
                sE[threadIdx.x] = sE[o];
                sF[threadIdx.x] = sF[o];
                sT[threadIdx.x] = sT[o];
                sP[threadIdx.x] = sP[o];
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        TileBestCandidate r{};
        r.bestErr = sE[0];
        r.bestTheta = sT[0];
        r.featureIdx = sF[0];
        r.bestParity = static_cast<int8_t>(sP[0]);
        *out = r;
    }
}

}  // namespace vj

using namespace vj;

// How a stage sorts and scans; every schedule gives the same rounds (the
// sums are exact integers and the sort is a stable total order), only the
// speed differs.
struct TrainSchedule {
    int wide = 0;    // global-memory sort and split-mask layout (always above the LDS sort limit)
    int ahead = 4;   // search: steps of 16-byte order loads per lane in flight (2, 4 or 8)
    int blocks = 2;  // search: workgroups per CU (1, 2 or 4)
    int packed = 0;  // search: packed 32-bit weights in LDS even where int64 weights fit
};

struct vj_trainer {
    int winW = 24, winH = 24, maxSamples = 0, maxTile = 0;
    // per-op binding
    vj_integral_view ii{};
    const int8_t* bLabel = nullptr;
    const uint8_t* bActive = nullptr;
    const float* bWeight = nullptr;
    int bN = 0;
    DevBuf<float> opSw;
    DevBuf<uint16_t> opOrd;
    DevBuf<uint64_t> opMask;

    // fused stage state
    int n = 0, numPos = 0, nFeat = 0, stageNegBase = 0;
    int64_t pitch = 0;
    const vj_feature* feats = nullptr;
    DevBuf<uint32_t> sumT;
    DevBuf<uint64_t> sqT;
    DevBuf<float> inv;
    DevBuf<uint16_t> ord;
    // wide layout (n above the LDS sort limit): plain indices in ord, split
    // flags in mask rows of pitch / 64 words, exact weights and totals for the
    // scan, and the presort's per-workgroup ping-pong buffers
    bool wide = false;
    DevBuf<uint64_t> mask;
    DevBuf<int64_t> wq, wtot;
    DevBuf<uint32_t> scrKeys;
    DevBuf<uint16_t> scrIdx;
    DevBuf<int8_t> label;
    DevBuf<uint8_t> active;
    DevBuf<double> w;
    DevBuf<float> strong;
    DevBuf<float> sw32;
    DevBuf<uint8_t> used;
    DevBuf<WaveBest> blockBest;
    DevBuf<RoundBest> roundBest;
    DevBuf<Calib> calib;
    DevBuf<uint32_t> counter;
    PinnedBuf<RoundBest> hRound;
    PinnedBuf<Calib> hCalib;
    PinnedBuf<uint32_t> hCounter;
    int searchBlocks = 0;
    int cus = 256;
    // exact pruning of the search (see search_kernel): per-feature error
    // bounds, the weights' growth factor, the round's smallest error so far,
    // features scanned since the stage began, and the last round's grid
    bool prune = true;
    DevBuf<double> lb, rho;
    DevBuf<unsigned long long> gBest, scanCnt;
    int lastBlocks = 0;
    // the stage's sort and scan schedule (see TrainSchedule)
    TrainSchedule sched;
    bool autotune = false;   // pick the schedule per (samples, features) bucket by timing
    bool forced = false;     // a schedule set by the caller
};

namespace {

int device_cus() {
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) == hipSuccess) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) == hipSuccess && p.multiProcessorCount > 0) cus = p.multiProcessorCount;
    }
    (void)hipGetLastError();
    return cus;
}

constexpr int kAheadChoices[] = {2, 4, 8};
constexpr int kBlockChoices[] = {1, 2, 4};
constexpr int kMaxBlocksPerCu = 4;

// The wide presort over count features into out (rows outPitch apart) and
// mask (rows outPitch / 64 words apart; mode 0 only).
int launch_presort_wide(vj_trainer* t, const vj_feature* feats, const uint32_t* sumT, const float* inv,
                        const float* resp, int n, int count, uint16_t* out, int64_t outPitch, uint64_t* mask, int mode,
                        hipStream_t s) {
    const int grid = std::max(1, std::min(count, t->searchBlocks));
    const int64_t nPad = (n + kWideTile - 1) / kWideTile * kWideTile;
    VJ_TRY(t->scrKeys.ensure(static_cast<size_t>(grid) * 2 * static_cast<size_t>(nPad)));
    VJ_TRY(t->scrIdx.ensure(static_cast<size_t>(grid) * 2 * static_cast<size_t>(nPad)));
    hipLaunchKernelGGL(presort_wide_kernel, dim3(grid), dim3(kSortThreads), 0, s, feats, 0, sumT, inv, resp, n, count,
                       out, outPitch, mask, outPitch / 64, mode, t->scrKeys.get(), t->scrIdx.get(), nPad);
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

size_t presort_lds(int n) {
    const size_t nPad = static_cast<size_t>((n + kSortThreads - 1) / kSortThreads * kSortThreads);
    return static_cast<size_t>(kCntPitch) * kSortWaves * 4 + nPad * 4 + nPad * 2;
}

// presort_kernel<NB> for NB = ceil(n / 1024) positions per thread (1..20).
template <int NB = 1>
void launch_presort(int count, hipStream_t s, const vj_feature* feats, const uint32_t* sumT, const float* inv,
                    const float* resp, int n, uint16_t* out, int64_t outPitch, int mode) {
    if constexpr (NB <= kSortMaxPerThread) {
        if ((n + kSortThreads - 1) / kSortThreads != NB)
            return launch_presort<NB + 1>(count, s, feats, sumT, inv, resp, n, out, outPitch, mode);
        hipLaunchKernelGGL(presort_kernel<NB>, dim3(count), dim3(kSortThreads), presort_lds(n), s, feats, 0, count,
                           sumT, inv, resp, n, out, outPitch, mode);
    }
}

template <int NB = 1>
int presort_lds_limits() {
    if constexpr (NB <= kSortMaxPerThread) {
        VJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(presort_kernel<NB>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 1024));
        return presort_lds_limits<NB + 1>();
    }
    return VJ_OK;
}

// Every corner of every feature must index the (winW+1) x (winH+1) window
// integral; checked on the host before any kernel reads through them.
int validate_features(const vj_feature* dFeatures, int begin, int count, int winW, int winH, hipStream_t s) {
    std::vector<vj_feature> hf(static_cast<size_t>(count));
    VJ_HIP(hipMemcpyAsync(hf.data(), dFeatures + begin, sizeof(vj_feature) * hf.size(), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    const int area = (winW + 1) * (winH + 1);
    for (const vj_feature& f : hf) {
        VJ_CHECK(f.rectCount == 2 || f.rectCount == 3, "feature rectCount must be 2 or 3");
        for (int r = 0; r < f.rectCount; ++r) {
            const int32_t p[4] = {f.r[r].p0, f.r[r].p1, f.r[r].p2, f.r[r].p3};
            for (int c = 0; c < 4; ++c) VJ_CHECK(p[c] >= 0 && p[c] < area, "feature corner outside the window integral");
        }
    }
    return VJ_OK;
}

int set_lds_limits() {
    static bool done = false;
    if (done) return VJ_OK;
    VJ_TRY(presort_lds_limits());
    for (const void* k : {reinterpret_cast<const void*>(search_kernel<uint32_t, 2>),
                          reinterpret_cast<const void*>(search_kernel<uint32_t, 4>),
                          reinterpret_cast<const void*>(search_kernel<uint32_t, 8>),
                          reinterpret_cast<const void*>(search_kernel<int64_t, 2>),
                          reinterpret_cast<const void*>(search_kernel<int64_t, 4>),
                          reinterpret_cast<const void*>(search_kernel<int64_t, 8>)})
        VJ_HIP(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 2048));
    VJ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(search_each_kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024 - 2048));
    done = true;
    return VJ_OK;
}

}  // namespace

extern "C" int vj_trainer_create(int winW, int winH, int maxSamples, int maxTile, vj_trainer** out) {
    // src/boosting_core.cu:468-476
    VJ_CHECK(out, "vj_trainer_create: null output");
    *out = nullptr;
    VJ_CHECK(winW > 0 && winH > 0 && maxSamples > 0 && maxTile > 0, "AdaBoostTrainer: invalid maxSamples/maxFeaturesPerTile");
    VJ_CHECK(static_cast<int64_t>(maxSamples) * maxTile <= INT_MAX,
             "AdaBoostTrainer: maxSamples*maxFeaturesPerTile exceeds int range");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
        (void)hipGetLastError();
        set_error("vj_trainer_create: no HIP device");
        return VJ_DEVICE_ERROR;
    }
    VJ_TRY(set_lds_limits());
    vj_trainer* t = new (std::nothrow) vj_trainer();
    VJ_CHECK(t, "vj_trainer_create: out of host memory");
    t->winW = winW;
    t->winH = winH;
    t->maxSamples = maxSamples;
    t->maxTile = maxTile;
    t->cus = device_cus();
    t->searchBlocks = 2 * t->cus;
    int st = t->hRound.ensure(1);
    if (st == VJ_OK) st = t->hCalib.ensure(1);
    if (st == VJ_OK) st = t->hCounter.ensure(1);
    if (st != VJ_OK) {
        delete t;
        return st;
    }
    *out = t;
    return VJ_OK;
}

extern "C" void vj_trainer_destroy(vj_trainer* t) { delete t; }

extern "C" int vj_trainer_bind(vj_trainer* t, const vj_integral_view* ii, const int8_t* dLabel,
                               const uint8_t* dActive, const float* dWeight, int n) {
    // src/boosting_core.cu:688-702
    VJ_CHECK(t && ii && ii->sumT && ii->invNorm && dLabel && dWeight, "bindTrainingSet: null view");
    VJ_CHECK(ii->nSamples > 0 && ii->nSamples <= t->maxSamples && ii->nSamples <= 65535,
             "bindTrainingSet: sample count out of range");
    VJ_CHECK(n == ii->nSamples, "bindTrainingSet: label/weight count mismatch");
    t->ii = *ii;
    t->bLabel = dLabel;
    t->bActive = dActive;
    t->bWeight = dWeight;
    t->bN = n;
    return VJ_OK;
}

extern "C" int vj_trainer_evaluate_responses(vj_trainer* t, const vj_feature* dFeatures, int begin, int count,
                                             float* dOut, vj_stream sv) {
    VJ_CHECK(t && dFeatures && t->ii.sumT && t->ii.invNorm, "evaluateFeatureResponses: not bound");
    VJ_CHECK(begin >= 0 && count > 0 && count <= t->maxTile && dOut, "evaluateFeatureResponses: bad range");
    VJ_CHECK(t->ii.iiWidth == t->winW + 1 && t->ii.iiHeight == t->winH + 1, "evaluateFeatureResponses: window mismatch");
    VJ_TRY(validate_features(dFeatures, begin, count, t->winW, t->winH, static_cast<hipStream_t>(sv)));
    const int n = t->bN;
    hipLaunchKernelGGL(responses_kernel, dim3((n + 255) / 256, count), dim3(256), 0, static_cast<hipStream_t>(sv),
                       dFeatures, begin, t->ii.sumT, t->ii.invNorm, n, dOut);
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

extern "C" int vj_trainer_sort_per_feature(vj_trainer* t, const float* dResp, int count, uint16_t* dOut,
                                           vj_stream sv) {
    VJ_CHECK(t && dResp && dOut && count > 0 && count <= t->maxTile, "sortSamplesPerFeature: bad arguments");
    const int n = t->bN;
    VJ_CHECK(n > 0 && n <= kTrainMaxN, "sortSamplesPerFeature: sample count must be 1..65535");
    if (n > kSortMaxN)
        return launch_presort_wide(t, nullptr, nullptr, nullptr, dResp, n, count, dOut, static_cast<int64_t>(n),
                                   nullptr, 1, static_cast<hipStream_t>(sv));
    launch_presort(count, static_cast<hipStream_t>(sv), nullptr, nullptr, nullptr, dResp, n, dOut,
                   static_cast<int64_t>(n), 1);
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

extern "C" int vj_trainer_find_threshold(vj_trainer* t, const float* dResp, const uint16_t* dSorted, int count,
                                         void* dOutBest, vj_stream sv) {
    VJ_CHECK(t && dResp && dSorted && dOutBest && t->bLabel && t->bWeight, "evaluateAndFindThreshold: not bound");
    VJ_CHECK(count > 0 && count <= t->maxTile, "evaluateAndFindThreshold: bad count");
    const int n = t->bN;
    VJ_CHECK(n > 1 && n <= kTrainMaxN, "evaluateAndFindThreshold: sample count must be 2..65535");
    hipStream_t s = static_cast<hipStream_t>(sv);
    // exact sums count 2^-62 units in int64: the weights must be normalised
    // (every boosting round leaves them summing to 1)
    {
        std::vector<float> hw(static_cast<size_t>(n));
        VJ_HIP(hipMemcpyAsync(hw.data(), t->bWeight, sizeof(float) * hw.size(), hipMemcpyDeviceToHost, s));
        VJ_HIP(hipStreamSynchronize(s));
        double tot = 0.0;
        for (float x : hw) {
            VJ_CHECK(std::isfinite(x), "evaluateAndFindThreshold: non-finite weight");
            tot += std::fabs(static_cast<double>(x));
        }
        VJ_CHECK(tot <= 1.5, "evaluateAndFindThreshold: weights must be normalised (sum of |w| <= 1.5)");
    }
    const int64_t pitch = (n + 255) / 256 * 256;
    VJ_TRY(t->opSw.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->opOrd.ensure(static_cast<size_t>(pitch) * count));
    hipLaunchKernelGGL(search_weights_kernel, dim3((n + 255) / 256), dim3(256), 0, s, t->bWeight, t->bLabel,
                       t->bActive, n, t->opSw.get());
    VJ_HIP(hipGetLastError());
    if (n > kPackedMaxN) {
        // indices need all 16 bits: the wide layout
        VJ_TRY(t->opMask.ensure(static_cast<size_t>(pitch / 64) * count));
        VJ_TRY(t->wq.ensure(static_cast<size_t>(n)));
        VJ_TRY(t->wtot.ensure(2));
        hipLaunchKernelGGL(wide_weights_kernel, dim3(1), dim3(1024), 0, s, t->opSw.get(), n, t->wq.get(),
                           t->wtot.get());
        VJ_HIP(hipGetLastError());
        hipLaunchKernelGGL(mark_valid_wide_kernel, dim3(static_cast<unsigned>(pitch / 256), count), dim3(256), 0, s,
                           dResp, dSorted, n, t->opOrd.get(), t->opMask.get(), pitch);
        VJ_HIP(hipGetLastError());
        const int blocks = std::min((count + 15) / 16, t->searchBlocks);
        hipLaunchKernelGGL(search_each_wide_kernel, dim3(blocks), dim3(kSearchThreads), 0, s, t->opOrd.get(),
                           t->opMask.get(), pitch, n, count, t->wq.get(), t->wtot.get(), dResp,
                           static_cast<FeatureBestSplit*>(dOutBest));
        VJ_HIP(hipGetLastError());
        return VJ_OK;
    }
    hipLaunchKernelGGL(mark_valid_kernel, dim3((n + 255) / 256, count), dim3(256), 0, s, dResp, dSorted, n,
                       t->opOrd.get(), pitch);
    VJ_HIP(hipGetLastError());
    const int blocks = std::min((count + 15) / 16, t->searchBlocks);
    hipLaunchKernelGGL(search_each_kernel, dim3(blocks), dim3(kSearchThreads), static_cast<size_t>(n) * 4, s,
                       t->opOrd.get(), pitch, n, count, t->opSw.get(), dResp,
                       static_cast<FeatureBestSplit*>(dOutBest));
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

extern "C" int vj_trainer_select_best(vj_trainer* t, const void* dBest, int begin, int count, const uint8_t* dUsed,
                                      void* dOut, vj_stream sv) {
    VJ_CHECK(t && dBest && dOut && begin >= 0 && count > 0 && count <= t->maxTile, "selectBestInTile: bad arguments");
    hipLaunchKernelGGL(select_kernel, dim3(1), dim3(256), 0, static_cast<hipStream_t>(sv),
                       static_cast<const FeatureBestSplit*>(dBest), begin, count, dUsed,
                       static_cast<TileBestCandidate*>(dOut));
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

namespace {

// The sample order of features [0, count) of the stage, in the layout of
// t->sched (LDS radix sort with flags in the entries, or the global sort with
// split-mask rows).
int presort_stage(vj_trainer* t, int count, hipStream_t s) {
    const int n = t->n;
    t->wide = t->sched.wide != 0;
    if (t->wide) {
        VJ_TRY(t->mask.ensure(static_cast<size_t>(t->pitch / 64) * t->nFeat));
        VJ_TRY(t->wq.ensure(static_cast<size_t>(n)));
        VJ_TRY(t->wtot.ensure(2));
        return launch_presort_wide(t, t->feats, t->sumT.get(), t->inv.get(), nullptr, n, count, t->ord.get(),
                                   t->pitch, t->mask.get(), 0, s);
    }
    launch_presort(count, s, t->feats, t->sumT.get(), t->inv.get(), nullptr, n, t->ord.get(), t->pitch, 0);
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

// One round's threshold scan over features [0, count) with schedule sc (the
// layout of the presorted orders, t->wide, is the one presort_stage wrote);
// blocks receives the grid (the entries of t->blockBest written).
int launch_search(vj_trainer* t, const TrainSchedule& sc, int count, int& blocks, hipStream_t s, bool prune = false) {
    const int n = t->n;
    blocks = std::max(1, std::min((count + 15) / 16, sc.blocks * t->cus));
    const dim3 g(static_cast<unsigned>(blocks)), b(kSearchThreads);
    if (t->wide) {
        hipLaunchKernelGGL(wide_weights_kernel, dim3(1), dim3(1024), 0, s, t->sw32.get(), n, t->wq.get(),
                           t->wtot.get());
        VJ_HIP(hipGetLastError());
        auto go = [&](auto kern) {
            hipLaunchKernelGGL(kern, g, b, 0, s, t->ord.get(), t->mask.get(), t->pitch, n, count, t->wq.get(),
                               t->wtot.get(), t->used.get(), t->blockBest.get());
        };
        if (sc.ahead == 2) go(search_wide_kernel<2>);
        else if (sc.ahead == 8) go(search_wide_kernel<8>);
        else go(search_wide_kernel<4>);
    } else {
        // int64 weights in LDS when they fit beside the kernel's static LDS
        // and the schedule does not ask for packed words
        const bool i64 = !sc.packed && static_cast<size_t>(n) * 8 <= kSearchWideLds;
        const size_t lds = static_cast<size_t>(n) * (i64 ? 8 : 4);
        double* lb = prune ? t->lb.get() : nullptr;
        // the seed pass (the previous round's per-workgroup winners, one
        // feature per wave, so it keeps eight loads in flight per lane)
        // then the round's scan
        auto go = [&](auto seedKern, auto kern) {
            if (prune && t->lastBlocks > 0)
                hipLaunchKernelGGL(seedKern, dim3(static_cast<unsigned>((t->lastBlocks + 15) / 16)), b, lds, s,
                                   t->ord.get(), t->pitch, n, count, t->sw32.get(), t->used.get(), t->blockBest.get(),
                                   lb, t->rho.get(), t->gBest.get(), t->blockBest.get(), t->lastBlocks,
                                   t->scanCnt.get());
            hipLaunchKernelGGL(kern, g, b, lds, s, t->ord.get(), t->pitch, n, count, t->sw32.get(), t->used.get(),
                               t->blockBest.get(), lb, t->rho.get(), t->gBest.get(), nullptr, 0,
                               prune ? t->scanCnt.get() : nullptr);
        };
        if (i64) {
            if (sc.ahead == 2) go(search_kernel<int64_t, 8>, search_kernel<int64_t, 2>);
            else if (sc.ahead == 8) go(search_kernel<int64_t, 8>, search_kernel<int64_t, 8>);
            else go(search_kernel<int64_t, 8>, search_kernel<int64_t, 4>);
        } else {
            if (sc.ahead == 2) go(search_kernel<uint32_t, 8>, search_kernel<uint32_t, 2>);
            else if (sc.ahead == 8) go(search_kernel<uint32_t, 8>, search_kernel<uint32_t, 8>);
            else go(search_kernel<uint32_t, 8>, search_kernel<uint32_t, 4>);
        }
        if (prune) t->lastBlocks = blocks;
    }
    VJ_HIP(hipGetLastError());
    return VJ_OK;
}

// A stage start (or the tuner's restored state): every feature's bound 0,
// so the next round scans every feature; no seeds.
int reset_pruning(vj_trainer* t, int nFeatures, hipStream_t s) {
    hipLaunchKernelGGL(fill_f64_kernel, dim3((nFeatures + 255) / 256), dim3(256), 0, s, t->lb.get(), nFeatures, 0.0);
    VJ_HIP(hipGetLastError());
    t->lastBlocks = 0;
    return VJ_OK;
}

// Schedules tuned per (sample count / 1024, log2 feature count, window).
// The first stage of a bucket picks the presort layout from a sample of its
// features (the presort and one default scan of the sample in each layout,
// scaled to one presort plus kTuneRounds rounds of the whole stage), presorts
// every feature in that layout, then times every scan candidate on the whole
// stage and keeps the fastest.  Later stages of the bucket reuse the pick.
constexpr int kTuneSample = 16384;
constexpr int kTuneRounds = 50;
std::mutex gTuneMu;
std::map<std::array<int, 5>, TrainSchedule> gTuned;

std::array<int, 5> tune_key(const vj_trainer* t) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) (void)hipGetLastError();
    int lgF = 0;
    while ((1 << lgF) < t->nFeat) ++lgF;
    return {dev, (t->n + 1023) / 1024, lgF, t->winW, t->winH};
}

// The best of three timed runs of launch on stream s, in ms.
template <typename F>
int time_best(hipStream_t s, F&& launch, float& ms) {
    hipEvent_t e0 = nullptr, e1 = nullptr;
    VJ_HIP(hipEventCreate(&e0));
    const hipError_t ce = hipEventCreate(&e1);
    if (ce != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(ce, "hipEventCreate", __FILE__, __LINE__);
    }
    int st = VJ_OK;
    float best = 1e30f;
    for (int rep = 0; rep < 3 && st == VJ_OK; ++rep) {
        float m = 0.0f;
        hipError_t he = hipEventRecord(e0, s);
        if (he == hipSuccess) st = launch();
        if (he == hipSuccess && st == VJ_OK) he = hipEventRecord(e1, s);
        if (he == hipSuccess && st == VJ_OK) he = hipEventSynchronize(e1);
        if (he == hipSuccess && st == VJ_OK) he = hipEventElapsedTime(&m, e0, e1);
        if (he != hipSuccess && st == VJ_OK) st = hip_fail(he, "tuning timer", __FILE__, __LINE__);
        best = std::min(best, m);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ms = best;
    return st;
}

// The presort layout of the stage (t->sched.wide) from a sample of features.
int tune_layout(vj_trainer* t, hipStream_t s) {
    if (t->n > kSortMaxN) {
        t->sched.wide = 1;
        return VJ_OK;
    }
    const int sample = std::min(t->nFeat, kTuneSample);
    const double scale = static_cast<double>(t->nFeat) / sample;
    double bestCost = 1e300;
    int bestWide = 0;
    for (int wide = 0; wide <= 1; ++wide) {
        TrainSchedule sc;
        sc.wide = wide;
        t->sched = sc;
        float sortMs = 0.0f, scanMs = 0.0f;
        VJ_TRY(time_best(s, [&] { return presort_stage(t, sample, s); }, sortMs));
        int blocks = 0;
        VJ_TRY(time_best(s, [&] { return launch_search(t, sc, sample, blocks, s); }, scanMs));
        const double cost = scale * (sortMs + kTuneRounds * static_cast<double>(scanMs));
        if (cost < bestCost) {
            bestCost = cost;
            bestWide = wide;
        }
    }
    t->sched = TrainSchedule{};
    t->sched.wide = bestWide;
    return VJ_OK;
}

// The scan schedule of the stage, in the presorted layout, from timed rounds:
// each candidate runs the stage's first two rounds (search, arg-min, weight
// update, calibration) from the same saved state, on the host's clock, as the
// rounds of the stage will (a scan kernel timed alone ranked the workgroup
// counts differently from the rounds).  The candidates take turns, three
// times round, so that a drift of the clocks over the tuning favours none;
// each keeps its best time, and one replaces the default only when it is
// faster by more than kTuneMargin.  The state is restored afterwards.
constexpr double kTuneMargin = 0.03;
constexpr int kTuneRoundsTimed = 2;  // rounds per timing (the same rounds under every candidate)
int round_impl(vj_trainer* t, const vj_round_config* cfg, vj_weak* out, hipStream_t s);

int tune_scan(vj_trainer* t, hipStream_t s) {
    std::vector<TrainSchedule> cand;
    for (int ahead : kAheadChoices)
        for (int blocksPerCu : kBlockChoices)
            for (int packed = 0; packed <= (t->sched.wide ? 0 : 1); ++packed) {
                TrainSchedule sc = t->sched;
                sc.ahead = ahead;
                sc.blocks = blocksPerCu;
                sc.packed = packed;
                cand.push_back(sc);
            }
    const size_t n = static_cast<size_t>(t->n), nf = static_cast<size_t>(t->nFeat);
    DevBuf<double> w0;
    DevBuf<float> strong0, sw0;
    DevBuf<uint8_t> used0;
    VJ_TRY(w0.ensure(n));
    VJ_TRY(strong0.ensure(n));
    VJ_TRY(sw0.ensure(n));
    VJ_TRY(used0.ensure(nf));
    auto copy = [&](void* dst, const void* src, size_t bytes) {
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, s);
    };
    VJ_HIP(copy(w0.get(), t->w.get(), n * 8));
    VJ_HIP(copy(strong0.get(), t->strong.get(), n * 4));
    VJ_HIP(copy(sw0.get(), t->sw32.get(), n * 4));
    VJ_HIP(copy(used0.get(), t->used.get(), nf));
    auto restore = [&]() -> int {
        VJ_HIP(copy(t->w.get(), w0.get(), n * 8));
        VJ_HIP(copy(t->strong.get(), strong0.get(), n * 4));
        VJ_HIP(copy(t->sw32.get(), sw0.get(), n * 4));
        VJ_HIP(copy(t->used.get(), used0.get(), nf));
        VJ_TRY(reset_pruning(t, t->nFeat, s));
        VJ_HIP(hipStreamSynchronize(s));
        return VJ_OK;
    };
    std::vector<double> ms(cand.size(), 1e300);
    for (int rep = 0; rep < 3; ++rep)
        for (size_t c = 0; c < cand.size(); ++c) {
            VJ_TRY(restore());
            t->sched = cand[c];
            vj_weak out;
            const auto t0 = std::chrono::steady_clock::now();
            for (int r = 0; r < kTuneRoundsTimed; ++r) VJ_TRY(round_impl(t, nullptr, &out, s));
            ms[c] = std::min(ms[c], std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
        }
    VJ_TRY(restore());
    const TrainSchedule dflt{};
    size_t best = 0, def = 0;
    for (size_t c = 0; c < cand.size(); ++c) {
        if (ms[c] < ms[best]) best = c;
        if (cand[c].ahead == dflt.ahead && cand[c].blocks == dflt.blocks && cand[c].packed == dflt.packed) def = c;
    }
    t->sched = ms[def] <= ms[best] * (1.0 + kTuneMargin) ? cand[def] : cand[best];
    return VJ_OK;
}

}  // namespace

extern "C" int vj_trainer_stage_begin(vj_trainer* t, const uint8_t* dPatches, const int8_t* hLabels,
                                      const uint8_t* hActive, int n, int numPos, const vj_feature* dFeatures,
                                      int nFeatures, vj_stream sv) {
    VJ_CHECK(t && dPatches && hLabels && dFeatures && nFeatures > 0, "stage_begin: bad arguments");
    VJ_CHECK(n > 1 && n <= t->maxSamples && n <= kTrainMaxN,
             "stage_begin: sample count must be 2..min(maxSamples, 65535)");
    VJ_CHECK(numPos > 0 && numPos < n, "stage_begin: need positives and negatives");
    VJ_CHECK(static_cast<int64_t>(t->winW) * t->winH * 255 < (1 << 24),
             "stage_begin: window too large for exact float rectangle sums");
    hipStream_t s = static_cast<hipStream_t>(sv);
    VJ_TRY(validate_features(dFeatures, 0, nFeatures, t->winW, t->winH, s));
    const int iw = t->winW + 1, ih = t->winH + 1;
    const size_t area = static_cast<size_t>(iw) * ih;
    t->pitch = (n + 255) / 256 * 256;
    VJ_TRY(t->sumT.ensure(area * n));
    VJ_TRY(t->sqT.ensure(area * n));
    VJ_TRY(t->inv.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->ord.ensure(static_cast<size_t>(t->pitch) * nFeatures));
    VJ_TRY(t->label.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->lb.ensure(static_cast<size_t>(nFeatures)));
    VJ_TRY(t->rho.ensure(1));
    VJ_TRY(t->gBest.ensure(1));
    VJ_TRY(t->scanCnt.ensure(1));
    VJ_HIP(hipMemsetAsync(t->scanCnt.get(), 0, sizeof(unsigned long long), s));
    VJ_TRY(t->active.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->w.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->strong.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->sw32.ensure(static_cast<size_t>(n)));
    VJ_TRY(t->used.ensure(static_cast<size_t>(nFeatures)));
    VJ_TRY(t->blockBest.ensure(static_cast<size_t>(kMaxBlocksPerCu) * t->cus));
    VJ_TRY(t->roundBest.ensure(1));
    VJ_TRY(t->calib.ensure(1));
    VJ_TRY(t->counter.ensure(1));
    std::vector<uint8_t> act(static_cast<size_t>(n), 1);
    if (hActive) std::memcpy(act.data(), hActive, static_cast<size_t>(n));
    for (int i = 0; i < numPos; ++i) act[static_cast<size_t>(i)] = 1;
    int activeNeg = 0;
    for (int i = numPos; i < n; ++i) activeNeg += act[static_cast<size_t>(i)] ? 1 : 0;
    VJ_CHECK(activeNeg > 0, "stage_begin: no active negatives");
    VJ_HIP(hipMemcpyAsync(t->label.get(), hLabels, static_cast<size_t>(n), hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemcpyAsync(t->active.get(), act.data(), static_cast<size_t>(n), hipMemcpyHostToDevice, s));
    VJ_HIP(hipMemsetAsync(t->used.get(), 0, static_cast<size_t>(nFeatures), s));
    // integrals and normalisers (src/main.cu:1945)
    hipLaunchKernelGGL(stage_integral_kernel, dim3((n + 255) / 256), dim3(256), 0, s, dPatches, t->winW,
                       t->winH, n, t->sumT.get(), t->sqT.get(), t->inv.get());
    VJ_HIP(hipGetLastError());
    // weights 0.5/numPos and 0.5/activeNeg, normalised (src/main.cu:1955-1962)
    hipLaunchKernelGGL(reset_kernel, dim3((n + 255) / 256), dim3(256), 0, s, t->w.get(), t->strong.get(),
                       t->active.get(), n, numPos, 0.5 / static_cast<double>(numPos), 0.5 / static_cast<double>(activeNeg));
    VJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(normalize_kernel, dim3(1), dim3(1024), 0, s, t->w.get(), t->label.get(), t->active.get(), n,
                       t->sw32.get(), t->rho.get(), t->gBest.get());
    VJ_HIP(hipGetLastError());
    VJ_TRY(reset_pruning(t, nFeatures, s));
    VJ_HIP(hipGetLastError());
    t->n = n;
    t->numPos = numPos;
    t->nFeat = nFeatures;
    t->feats = dFeatures;
    t->stageNegBase = activeNeg;
    // the stage's schedule: the caller's, the tuned one of this bucket, or
    // the default; the wide layout is the only one above the LDS sort limit
    bool tuneScan = false;
    std::array<int, 5> key{};
    if (!t->forced) {
        t->sched = TrainSchedule{};
        if (t->autotune) {
            key = tune_key(t);
            std::lock_guard<std::mutex> g(gTuneMu);
            auto it = gTuned.find(key);
            if (it != gTuned.end()) t->sched = it->second;
            else tuneScan = true;
        }
        if (tuneScan) VJ_TRY(tune_layout(t, s));
    }
    if (n > kSortMaxN) t->sched.wide = 1;
    // sample order of every feature, once for the whole stage
    VJ_TRY(presort_stage(t, nFeatures, s));
    if (tuneScan) {
        VJ_TRY(tune_scan(t, s));
        std::lock_guard<std::mutex> g(gTuneMu);
        gTuned[key] = t->sched;
    }
    VJ_HIP(hipStreamSynchronize(s));  // hLabels / act staging must stay valid until here
    return VJ_OK;
}

extern "C" int vj_trainer_set_autotune(vj_trainer* t, int on) {
    VJ_CHECK(t, "set_autotune: null trainer");
    t->autotune = on != 0;
    t->forced = false;
    return VJ_OK;
}

extern "C" int vj_trainer_set_schedule(vj_trainer* t, const int* sched) {
    VJ_CHECK(t, "set_schedule: null trainer");
    if (!sched) {
        t->forced = false;
        t->autotune = false;
        return VJ_OK;
    }
    TrainSchedule sc;
    sc.wide = sched[0];
    sc.ahead = sched[1];
    sc.blocks = sched[2];
    sc.packed = sched[3];
    VJ_CHECK(sc.wide == 0 || sc.wide == 1, "set_schedule: wide must be 0 or 1");
    VJ_CHECK(std::find(std::begin(kAheadChoices), std::end(kAheadChoices), sc.ahead) != std::end(kAheadChoices),
             "set_schedule: loads ahead must be 2, 4 or 8");
    VJ_CHECK(std::find(std::begin(kBlockChoices), std::end(kBlockChoices), sc.blocks) != std::end(kBlockChoices),
             "set_schedule: workgroups per CU must be 1, 2 or 4");
    VJ_CHECK(sc.packed == 0 || sc.packed == 1, "set_schedule: packed must be 0 or 1");
    t->sched = sc;
    t->forced = true;
    t->autotune = false;
    return VJ_OK;
}

extern "C" int vj_trainer_set_pruning(vj_trainer* t, int on) {
    VJ_CHECK(t, "set_pruning: null trainer");
    t->prune = on != 0;
    t->lastBlocks = 0;
    return VJ_OK;
}

extern "C" int vj_trainer_pruning_stats(vj_trainer* t, int64_t* scanned) {
    VJ_CHECK(t && scanned, "pruning_stats: null argument");
    *scanned = 0;
    if (!t->scanCnt.get()) return VJ_OK;
    unsigned long long v = 0;
    VJ_HIP(hipMemcpy(&v, t->scanCnt.get(), sizeof v, hipMemcpyDeviceToHost));
    *scanned = static_cast<int64_t>(v);
    return VJ_OK;
}

extern "C" int vj_trainer_schedule(vj_trainer* t, int* sched) {
    VJ_CHECK(t && sched, "schedule: null argument");
    sched[0] = t->sched.wide;
    sched[1] = t->sched.ahead;
    sched[2] = t->sched.blocks;
    sched[3] = t->sched.packed;
    return VJ_OK;
}

extern "C" int vj_trainer_round(vj_trainer* t, const vj_round_config* cfg, vj_weak* out, vj_stream sv) {
    VJ_CHECK(t && out && t->n > 0, "round: no stage begun");
    return round_impl(t, cfg, out, static_cast<hipStream_t>(sv));
}

namespace {

int round_impl(vj_trainer* t, const vj_round_config* cfg, vj_weak* out, hipStream_t s) {
    const float minHit = cfg ? cfg->minHitRate : 0.99f;
    const float maxFa = cfg ? cfg->maxFalseAlarm : 0.40f;
    std::memset(out, 0, sizeof *out);
    out->feature = -1;
    int blocks = 0;
    VJ_TRY(launch_search(t, t->sched, t->nFeat, blocks, s, t->prune && !t->wide));
    hipLaunchKernelGGL(reduce_best_kernel, dim3(1), dim3(256), 0, s, t->blockBest.get(), blocks, t->ord.get(), t->pitch,
                       t->feats, t->sumT.get(), t->inv.get(), t->n, t->wide ? 0xFFFFu : 0x7FFFu,
                       t->roundBest.get());
    VJ_HIP(hipGetLastError());
    VJ_HIP(hipMemcpyAsync(t->hRound.get(), t->roundBest.get(), sizeof(RoundBest), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    const RoundBest rb = *t->hRound.get();
    if (rb.feature < 0) return VJ_OK;  // no valid split anywhere (src/main.cu:2065-2070)
    // host round logic (src/main.cu:2082-2103), in f64
    const double err = std::ldexp(static_cast<double>(rb.errQ), -62);
    const double ec = std::min(std::max(err, 1e-6), 1.0 - 1e-6);
    const double C = std::log((1.0 - ec) / ec);
    const double factor = std::exp(C);
    const float fc = static_cast<float>(C);
    out->feature = rb.feature;
    out->theta = rb.theta;
    out->parity = rb.parity;
    out->split = rb.split;
    out->err = err;
    out->alpha = C;
    out->left = rb.parity > 0 ? -fc : fc;
    out->right = rb.parity > 0 ? fc : -fc;
    hipLaunchKernelGGL(apply_kernel, dim3((t->n + 255) / 256), dim3(256), 0, s, t->feats, rb.feature, t->sumT.get(),
                       t->inv.get(), t->n, t->label.get(), t->active.get(), t->w.get(), t->strong.get(), rb.theta,
                       rb.parity, out->left, out->right, factor, t->used.get());
    VJ_HIP(hipGetLastError());
    hipLaunchKernelGGL(normalize_kernel, dim3(1), dim3(1024), 0, s, t->w.get(), t->label.get(), t->active.get(), t->n,
                       t->sw32.get(), t->rho.get(), t->gBest.get());
    VJ_HIP(hipGetLastError());
    // calibration (src/main.cu:2172-2226)
    const int minPass = std::max(1, static_cast<int>(std::ceil(minHit * static_cast<float>(t->numPos))));
    const int cutIdx = t->numPos - minPass;
    VJ_CHECK(cutIdx >= 0 && cutIdx < t->numPos, "round: minHitRate gives no valid cut index");
    hipLaunchKernelGGL(calibrate_kernel, dim3(1), dim3(1024), 0, s, t->strong.get(), t->active.get(), t->n, t->numPos,
                       cutIdx, t->calib.get());
    VJ_HIP(hipGetLastError());
    VJ_HIP(hipMemcpyAsync(t->hCalib.get(), t->calib.get(), sizeof(Calib), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    const Calib c = *t->hCalib.get();
    const bool strictGt = static_cast<int>(c.posGt) >= minPass;
    out->stageThreshold = strictGt ? std::nextafter(c.cut, INFINITY) : c.cut;
    out->posPass = static_cast<int32_t>(strictGt ? c.posGt : c.posGe);
    out->negSurvive = static_cast<int32_t>(strictGt ? c.negGt : c.negGe);
    out->hitRate = static_cast<float>(out->posPass) / static_cast<float>(t->numPos);
    out->falseAlarm = static_cast<float>(out->negSurvive) / static_cast<float>(t->stageNegBase);
    out->stageSatisfied = ((out->hitRate >= minHit && out->falseAlarm <= maxFa) || out->negSurvive == 0) ? 1 : 0;
    return VJ_OK;
}

}  // namespace

extern "C" int vj_trainer_stage_end(vj_trainer* t, float threshold, int* remainNeg, vj_stream sv) {
    VJ_CHECK(t && remainNeg && t->n > 0, "stage_end: no stage begun");
    hipStream_t s = static_cast<hipStream_t>(sv);
    *t->hCounter.get() = 0u;
    VJ_HIP(hipMemcpyAsync(t->counter.get(), t->hCounter.get(), sizeof(uint32_t), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(stage_end_kernel, dim3((t->n + 255) / 256), dim3(256), 0, s, t->strong.get(), t->active.get(),
                       t->n, t->numPos, threshold, t->counter.get());
    VJ_HIP(hipGetLastError());
    VJ_HIP(hipMemcpyAsync(t->hCounter.get(), t->counter.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    VJ_HIP(hipStreamSynchronize(s));
    *remainNeg = static_cast<int>(*t->hCounter.get());
    return VJ_OK;
}

extern "C" int vj_trainer_read_state(vj_trainer* t, double* hW, float* hStrong, uint8_t* hActive, float* hInv) {
    VJ_CHECK(t && t->n > 0, "read_state: no stage begun");
    const size_t n = static_cast<size_t>(t->n);
    if (hW) VJ_HIP(hipMemcpy(hW, t->w.get(), n * sizeof(double), hipMemcpyDeviceToHost));
    if (hStrong) VJ_HIP(hipMemcpy(hStrong, t->strong.get(), n * sizeof(float), hipMemcpyDeviceToHost));
    if (hActive) VJ_HIP(hipMemcpy(hActive, t->active.get(), n, hipMemcpyDeviceToHost));
    if (hInv) VJ_HIP(hipMemcpy(hInv, t->inv.get(), n * sizeof(float), hipMemcpyDeviceToHost));
    return VJ_OK;
}

extern "C" int vj_trainer_read_orders(vj_trainer* t, int begin, int count, uint16_t* hOut) {
    VJ_CHECK(t && hOut && begin >= 0 && count > 0 && begin + count <= t->nFeat, "read_orders: bad range");
    VJ_HIP(hipMemcpy2D(hOut, static_cast<size_t>(t->n) * 2, t->ord.get() + static_cast<size_t>(begin) * t->pitch,
                       static_cast<size_t>(t->pitch) * 2, static_cast<size_t>(t->n) * 2, static_cast<size_t>(count),
                       hipMemcpyDeviceToHost));
    return VJ_OK;
}

extern "C" int vj_trainer_read_splits(vj_trainer* t, int begin, int count, uint8_t* hOut) {
    VJ_CHECK(t && hOut && begin >= 0 && count > 0 && begin + count <= t->nFeat, "read_splits: bad range");
    const size_t n = static_cast<size_t>(t->n);
    if (t->wide) {
        const size_t words = static_cast<size_t>(t->pitch / 64);
        std::vector<uint64_t> m(words * static_cast<size_t>(count));
        VJ_HIP(hipMemcpy(m.data(), t->mask.get() + static_cast<size_t>(begin) * words, m.size() * 8,
                         hipMemcpyDeviceToHost));
        for (size_t f = 0; f < static_cast<size_t>(count); ++f)
            for (size_t k = 0; k < n; ++k) hOut[f * n + k] = static_cast<uint8_t>((m[f * words + k / 64] >> (k % 64)) & 1u);
    } else {
        std::vector<uint16_t> o(n * static_cast<size_t>(count));
        VJ_TRY(vj_trainer_read_orders(t, begin, count, o.data()));
        for (size_t i = 0; i < o.size(); ++i) hOut[i] = static_cast<uint8_t>(o[i] >> 15);
    }
    return VJ_OK;
}
