// vj_kernels.hpp -- device-side records and wavefront helpers shared by the
// detect and training kernels (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace vj {

constexpr int kWave = 64;

// Integral-image tiling: bands of 32 rows, chunks of 256 columns (one
// wavefront, 4 columns per lane).
constexpr int kBandRows = 32;
constexpr int kChunkCols = 256;

// Cascade tiling: a tile is 2^lgcols x rows windows of one level, shaped per
// window step so that every step's integral footprint stays within 14 KB of
// LDS (64x8 windows at step 1, 32x8 at step 2, 16x8 at step 3: 88x32, 87x39
// and 70x46 integral entries for a 24x24 window) and eight 256-thread
// workgroups fit on a CU.  A tile row is a 64-, 32- or 16-bit piece of one
// 64-bit mask word.  (The autotuner may pick other row counts per batch.)
constexpr int kTileMaxWin = 512;

__host__ __device__ constexpr int tile_lgcols_for_step(int step) {
    return step == 1 ? 6 : (step == 2 ? 5 : 4);
}
__host__ __device__ constexpr int tile_rows_for_step(int) {
    return 8;
}

// One pyramid level of one image, as every level-parallel kernel sees it.
struct LevelDesc {
    const uint8_t* pix;   // level pixels
    int64_t iiOff;        // first element of the level's (h+1) x (w+1) integral
    int64_t colOff;       // first element of the level's [nBands][w] column sums
    int64_t totOff;       // first element of the level's [nBands][nChunks] chunk totals
    int64_t rowOff;       // first element of the level's [h][nChunks] row-chunk totals
    int64_t wordOff;      // first 64-bit word of the level's window mask
    int32_t w, h, stride; // level size and pixel row stride
    int32_t step;         // window step (1, 2 or 3)
    float scale;          // level scale
    int32_t scaleIdx;     // index in the scale sequence (counts skipped scales)
    int32_t image;        // image index within the call
    int32_t nBands, nChunks;
    int32_t scanW, scanH, rowWords;
    int32_t tileRows, tileLgCols, tilesX, tilesY;
    int32_t itemOff;      // first (band, chunk) item
    int32_t chunkItemOff; // first column-chunk item
    int32_t tileOff;      // first cascade tile
    int32_t srcLevel;     // level this one is resized from (-1: level 0 of its image)
    int32_t iiPitch;      // elements per integral row: w + 1, or padded to whole 128-byte lines
};

// Integral row pitch of the detect pyramid, in elements: whole 128-byte lines
// of u32 sums and of u32 square sums, so the integral kernels' 16-byte stores
// are aligned.
constexpr int kIIPitchAlign = 32;

// Cascade stump with its corner offsets decoded for one LDS tile pitch.
// 64 bytes so a wavefront reads it with one scalar load.
struct alignas(16) KStump {
    uint16_t off[12];     // r0.p0..p3, r1.p0..p3, r2.p0..p3 as LDS byte offsets 4 * (dy*pitch + dx)
    float w[3];
    float theta, left, right;
    int32_t rectCount;
    int32_t pad[3];
};
static_assert(sizeof(KStump) == 64, "KStump is one 64-byte scalar load");

// Cascade stage.  lo <= threshold - B and hi >= threshold + B, where B bounds
// how far any two float summations of the stage's leaf values can differ.
struct KStage {
    int32_t first, ntrees;
    float threshold, lo, hi;
    int32_t pad[3];
};

// Pixel pointers read from a LevelDesc in memory have no known address space,
// so the compiler would access them with flat instructions (which also count
// against the LDS wait counter); they always point to global memory.
typedef const uint8_t __attribute__((address_space(1)))* gpix_t;
typedef uint8_t __attribute__((address_space(1)))* gpix_out_t;
__device__ __forceinline__ gpix_t global_pix(const uint8_t* p) { return (gpix_t)p; }
__device__ __forceinline__ gpix_out_t global_pix_out(const uint8_t* p) { return (gpix_out_t)(const_cast<uint8_t*>(p)); }

__device__ __forceinline__ int lane_id() {
    return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}

// Number of set bits of a 64-bit wave mask strictly below this lane.
__device__ __forceinline__ int mask_rank(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                     __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
}

template <typename T>
__device__ __forceinline__ T wave_incl_scan(T v) {
    const int lane = lane_id();
#pragma unroll
    for (int d = 1; d < kWave; d <<= 1) {
        T u = __shfl_up(v, d, kWave);
        if (lane >= d) v += u;
    }
    return v;
}

// Inclusive wave64 prefix sum of a u32 with DPP moves: shifts of 1, 2, 4, 8
// within each 16-lane row, then row_bcast15 / row_bcast31 carry the row
// totals into the rows above (GFX9-family DPP, valid on gfx950).
__device__ __forceinline__ uint32_t wave_incl_scan_dpp(uint32_t v) {
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x111, 0xf, 0xf, true));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x112, 0xf, 0xf, true));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x114, 0xf, 0xf, true));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x118, 0xf, 0xf, true));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x142, 0xa, 0xf, false));
    v += static_cast<uint32_t>(__builtin_amdgcn_update_dpp(0, static_cast<int>(v), 0x143, 0xc, 0xf, false));
    return v;
}

// Sum over the wave of a u32, returned in every lane.
__device__ __forceinline__ uint32_t wave_sum_dpp(uint32_t v) {
    return static_cast<uint32_t>(__builtin_amdgcn_readlane(static_cast<int>(wave_incl_scan_dpp(v)), 63));
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) v += __shfl_xor(v, d, kWave);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
        const T o = __shfl_xor(v, d, kWave);
        v = o > v ? o : v;
    }
    return v;
}

// Index of the last element of sorted offsets[0..n) that is <= key.
__device__ __forceinline__ int find_segment(const int32_t* offsets, int n, int key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_segment64(const int64_t* offsets, int n, int64_t key) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (offsets[mid] <= key) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace vj
