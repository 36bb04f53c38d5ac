// vj_device.cpp -- device memory, streams and events behind the C ABI, so
// that host callers (Python via ctypes, the CLI) need no HIP headers.
#include <cstring>

#include "vj_common.hpp"

using namespace vj;

extern "C" {

int vj_device_count(int* n) {
    VJ_CHECK(n, "vj_device_count: null");
    *n = 0;
    hipError_t e = hipGetDeviceCount(n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *n = 0;
    }
    return VJ_OK;
}

int vj_device_name(char* buf, int cap) {
    VJ_CHECK(buf && cap > 0, "vj_device_name: bad buffer");
    hipDeviceProp_t p;
    VJ_HIP(hipGetDeviceProperties(&p, 0));
    std::strncpy(buf, p.gcnArchName, static_cast<size_t>(cap) - 1);
    buf[cap - 1] = '\0';
    return VJ_OK;
}

int vj_malloc(void** p, size_t bytes) {
    VJ_CHECK(p, "vj_malloc: null");
    hipError_t e = hipMalloc(p, bytes ? bytes : 1);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipMalloc of " + std::to_string(bytes) + " bytes failed");
        *p = nullptr;
        return VJ_OUT_OF_MEMORY;
    }
    return VJ_OK;
}

int vj_free(void* p) {
    if (p) VJ_HIP(hipFree(p));
    return VJ_OK;
}

int vj_host_alloc(void** p, size_t bytes) {
    VJ_CHECK(p, "vj_host_alloc: null");
    hipError_t e = hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        set_error("hipHostMalloc failed");
        *p = nullptr;
        return VJ_OUT_OF_MEMORY;
    }
    return VJ_OK;
}

int vj_host_free(void* p) {
    if (p) VJ_HIP(hipHostFree(p));
    return VJ_OK;
}

int vj_memcpy_htod(void* dst, const void* src, size_t bytes, vj_stream s) {
    if (!bytes) return VJ_OK;
    VJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, static_cast<hipStream_t>(s)));
    VJ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_memcpy_dtoh(void* dst, const void* src, size_t bytes, vj_stream s) {
    if (!bytes) return VJ_OK;
    VJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, static_cast<hipStream_t>(s)));
    VJ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_memcpy_dtod(void* dst, const void* src, size_t bytes, vj_stream s) {
    if (!bytes) return VJ_OK;
    VJ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_memset(void* dst, int v, size_t bytes, vj_stream s) {
    if (!bytes) return VJ_OK;
    VJ_HIP(hipMemsetAsync(dst, v, bytes, static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_stream_create(vj_stream* s) {
    VJ_CHECK(s, "vj_stream_create: null");
    hipStream_t h = nullptr;
    VJ_HIP(hipStreamCreateWithFlags(&h, hipStreamNonBlocking));
    *s = h;
    return VJ_OK;
}

int vj_stream_destroy(vj_stream s) {
    if (s) VJ_HIP(hipStreamDestroy(static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_stream_synchronize(vj_stream s) {
    VJ_HIP(hipStreamSynchronize(static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_device_synchronize(void) {
    VJ_HIP(hipDeviceSynchronize());
    return VJ_OK;
}

int vj_event_create(vj_event* e) {
    VJ_CHECK(e, "vj_event_create: null");
    hipEvent_t h = nullptr;
    VJ_HIP(hipEventCreate(&h));
    *e = h;
    return VJ_OK;
}

int vj_event_destroy(vj_event e) {
    if (e) VJ_HIP(hipEventDestroy(static_cast<hipEvent_t>(e)));
    return VJ_OK;
}

int vj_event_record(vj_event e, vj_stream s) {
    VJ_HIP(hipEventRecord(static_cast<hipEvent_t>(e), static_cast<hipStream_t>(s)));
    return VJ_OK;
}

int vj_event_synchronize(vj_event e) {
    VJ_HIP(hipEventSynchronize(static_cast<hipEvent_t>(e)));
    return VJ_OK;
}

int vj_event_elapsed_ms(vj_event a, vj_event b, float* ms) {
    VJ_CHECK(ms, "vj_event_elapsed_ms: null");
    VJ_HIP(hipEventElapsedTime(ms, static_cast<hipEvent_t>(a), static_cast<hipEvent_t>(b)));
    return VJ_OK;
}

}  // extern "C"
