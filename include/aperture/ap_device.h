/*
 * Device plumbing of libaperture_hip.so: memory, copies, streams and
 * timing events on the library's own HIP runtime.
 *
 * No reference counterpart: the reference host calls cudaMalloc/cudaMemcpy
 * directly (core/scene.cu). These entry points exist for hosts that have no
 * HIP runtime of their own (the Python harness, FFI bindings), so that one
 * process never mixes two HIP runtimes. Streams are created at the default
 * priority.
 */
#ifndef APERTURE_AP_DEVICE_H
#define APERTURE_AP_DEVICE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AP_COPY_HOST_TO_DEVICE 1
#define AP_COPY_DEVICE_TO_HOST 2
#define AP_COPY_DEVICE_TO_DEVICE 3

int ap_device_count(int* count);
int ap_device_set(int device);
int ap_device_name(int device, char* buf, size_t buf_len);
int ap_device_malloc(size_t bytes, void** out_ptr);
int ap_device_free(void* ptr);
/* Synchronous copy of `bytes` (kind = AP_COPY_*). */
int ap_memcpy(void* dst, const void* src, size_t bytes, int kind);
/* Stream-ordered device-to-device copy by a plain streaming kernel (16 B per
 * lane; pointers 16-byte aligned). Reads each source byte and writes each
 * destination byte once, so it is also the known-size reference for scaling
 * HBM traffic counters. */
int ap_device_copy(void* dst, const void* src, size_t bytes, void* stream);
/* Stream-ordered fill of `bytes` with the byte `value`. */
int ap_memset_async(void* dst, int value, size_t bytes, void* stream);
int ap_stream_create(void** out_stream);
int ap_stream_destroy(void* stream);
int ap_stream_synchronize(void* stream);
int ap_device_synchronize(void);
int ap_event_create(void** out_event);
int ap_event_destroy(void* event);
int ap_event_record(void* event, void* stream);
/* Milliseconds between two recorded events (synchronises on `end`). */
int ap_event_elapsed_ms(void* start, void* end, float* out_ms);

#ifdef __cplusplus
}
#endif
#endif /* APERTURE_AP_DEVICE_H */
