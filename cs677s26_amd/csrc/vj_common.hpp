// vj_common.hpp -- internal helpers shared by the library's translation units:
// status/error reporting, grow-only device buffers, POD layout checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>

#include "vj_api.h"

namespace vj {

static_assert(sizeof(vj_rect) == 20, "FastRect is 20 bytes");
static_assert(sizeof(vj_feature) == 64, "HaarFeature is 64 bytes");
static_assert(sizeof(vj_stump) == 16, "GpuStump4 is 16 bytes");
static_assert(sizeof(vj_stage) == 12, "CascadeStage is 12 bytes");
static_assert(sizeof(vj_detection) == 24, "Detection is 24 bytes");

void set_error(const std::string& msg);

// Records a HIP failure (with the call site) and yields VJ_DEVICE_ERROR.
int hip_fail(hipError_t e, const char* what, const char* file, int line);

#define VJ_HIP(call)                                                         \
    do {                                                                     \
        hipError_t vj_e_ = (call);                                           \
        if (vj_e_ != hipSuccess) return ::vj::hip_fail(vj_e_, #call, __FILE__, __LINE__); \
    } while (0)

#define VJ_CHECK(cond, msg)                                                  \
    do {                                                                     \
        if (!(cond)) {                                                       \
            ::vj::set_error(msg);                                            \
            return VJ_INVALID_ARG;                                           \
        }                                                                    \
    } while (0)

#define VJ_TRY(expr)                                                         \
    do {                                                                     \
        int vj_s_ = (expr);                                                  \
        if (vj_s_ != VJ_OK) return vj_s_;                                    \
    } while (0)

// Device buffer that only grows; contents are not preserved across growth.
template <typename T>
class DevBuf {
public:
    DevBuf() = default;
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;

    int ensure(size_t count) {
        if (count <= cap_ && ptr_) return VJ_OK;
        release();
        size_t n = count ? count : 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr_), n * sizeof(T));
        if (e != hipSuccess) {
            ptr_ = nullptr;
            cap_ = 0;
            (void)hipGetLastError();
            set_error("device allocation of " + std::to_string(n * sizeof(T)) + " bytes failed");
            return VJ_OUT_OF_MEMORY;
        }
        cap_ = n;
        return VJ_OK;
    }
    void release() {
        if (ptr_) (void)hipFree(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
    }
    T* get() const { return ptr_; }
    size_t capacity() const { return cap_; }

private:
    T* ptr_ = nullptr;
    size_t cap_ = 0;
};

// Pinned host buffer that only grows.
template <typename T>
class PinnedBuf {
public:
    PinnedBuf() = default;
    ~PinnedBuf() { release(); }
    PinnedBuf(const PinnedBuf&) = delete;
    PinnedBuf& operator=(const PinnedBuf&) = delete;

    int ensure(size_t count) {
        if (count <= cap_ && ptr_) return VJ_OK;
        release();
        size_t n = count ? count : 1;
        hipError_t e = hipHostMalloc(reinterpret_cast<void**>(&ptr_), n * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr_ = nullptr;
            cap_ = 0;
            (void)hipGetLastError();
            set_error("pinned host allocation failed");
            return VJ_OUT_OF_MEMORY;
        }
        cap_ = n;
        return VJ_OK;
    }
    void release() {
        if (ptr_) (void)hipHostFree(ptr_);
        ptr_ = nullptr;
        cap_ = 0;
    }
    T* get() const { return ptr_; }

private:
    T* ptr_ = nullptr;
    size_t cap_ = 0;
};

// Host-side helpers implemented in vj_host.cpp and used by the engine.
int group_rectangles(const vj_detection* in, int n, int minNeighbors, float eps,
                     vj_detection* out, int* count);
// The same for many images at once (in[i], n[i] boxes -> out[i], count[i]),
// their steps spread over the host thread pool.
int group_rectangles_many(const vj_detection* const* in, const int* n, int nImages, int minNeighbors, float eps,
                          vj_detection* const* out, int* count);

}  // namespace vj
