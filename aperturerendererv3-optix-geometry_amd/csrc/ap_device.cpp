// Device plumbing entry points (include/aperture/ap_device.h).
#include <cstring>

#include "aperture/ap_device.h"
#include "ap_internal.h"

using namespace ap;

extern "C" {

int ap_device_count(int* count) {
    if (!count) return fail(AP_ERR_INVALID_ARG, "ap_device_count: NULL count");
    *count = 0;
    const hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) { *count = 0; return AP_OK; }
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    return AP_OK;
}

int ap_device_set(int device) {
    const hipError_t e = hipSetDevice(device);
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipSetDevice");
}

int ap_device_name(int device, char* buf, size_t len) {
    if (!buf || len == 0) return fail(AP_ERR_INVALID_ARG, "ap_device_name: NULL buffer");
    hipDeviceProp_t p;
    const hipError_t e = hipGetDeviceProperties(&p, device);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceProperties");
    std::strncpy(buf, p.gcnArchName, len - 1);
    buf[len - 1] = 0;
    return AP_OK;
}

int ap_device_malloc(size_t bytes, void** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_device_malloc: NULL out");
    *out = nullptr;
    if (bytes == 0) return AP_OK;
    const hipError_t e = hipMalloc(out, bytes);
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipMalloc");
}

int ap_device_free(void* ptr) {
    if (!ptr) return AP_OK;
    const hipError_t e = hipFree(ptr);
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipFree");
}

int ap_memcpy(void* dst, const void* src, size_t bytes, int kind) {
    if (bytes == 0) return AP_OK;
    if (!dst || !src) return fail(AP_ERR_INVALID_ARG, "ap_memcpy: NULL pointer");
    hipMemcpyKind k;
    switch (kind) {
        case AP_COPY_HOST_TO_DEVICE: k = hipMemcpyHostToDevice; break;
        case AP_COPY_DEVICE_TO_HOST: k = hipMemcpyDeviceToHost; break;
        case AP_COPY_DEVICE_TO_DEVICE: k = hipMemcpyDeviceToDevice; break;
        default: return fail(AP_ERR_INVALID_ARG, "ap_memcpy: unknown kind");
    }
    const hipError_t e = hipMemcpy(dst, src, bytes, k);
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipMemcpy");
}

int ap_memset_async(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return AP_OK;
    if (!dst) return fail(AP_ERR_INVALID_ARG, "ap_memset_async: NULL pointer");
    const hipError_t e = hipMemsetAsync(dst, value, bytes, static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipMemsetAsync");
}

int ap_stream_create(void** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_stream_create: NULL out");
    hipStream_t s = nullptr;
    const hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreateWithFlags");
    *out = s;
    return AP_OK;
}

int ap_stream_destroy(void* stream) {
    if (!stream) return AP_OK;
    const hipError_t e = hipStreamDestroy(static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipStreamDestroy");
}

int ap_stream_synchronize(void* stream) {
    const hipError_t e = hipStreamSynchronize(static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipStreamSynchronize");
}

int ap_device_synchronize(void) {
    const hipError_t e = hipDeviceSynchronize();
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipDeviceSynchronize");
}

int ap_event_create(void** out) {
    if (!out) return fail(AP_ERR_INVALID_ARG, "ap_event_create: NULL out");
    hipEvent_t ev = nullptr;
    const hipError_t e = hipEventCreate(&ev);
    if (e != hipSuccess) return hip_fail(e, "hipEventCreate");
    *out = ev;
    return AP_OK;
}

int ap_event_destroy(void* ev) {
    if (!ev) return AP_OK;
    const hipError_t e = hipEventDestroy(static_cast<hipEvent_t>(ev));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipEventDestroy");
}

int ap_event_record(void* ev, void* stream) {
    const hipError_t e = hipEventRecord(static_cast<hipEvent_t>(ev), static_cast<hipStream_t>(stream));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipEventRecord");
}

int ap_event_elapsed_ms(void* start, void* end, float* out_ms) {
    if (!out_ms) return fail(AP_ERR_INVALID_ARG, "ap_event_elapsed_ms: NULL out");
    hipError_t e = hipEventSynchronize(static_cast<hipEvent_t>(end));
    if (e != hipSuccess) return hip_fail(e, "hipEventSynchronize");
    e = hipEventElapsedTime(out_ms, static_cast<hipEvent_t>(start), static_cast<hipEvent_t>(end));
    return e == hipSuccess ? AP_OK : hip_fail(e, "hipEventElapsedTime");
}

}  // extern "C"
