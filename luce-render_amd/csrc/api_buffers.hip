// Device buffers, stream synchronization and event timing: the slice of
// luce-gpu (gpu.Buffer.create/allocate/read/upload, Compute.wait, timestamps;
// used at src/render/device.lucb:137-156 and integrator.lucb:848-885) that
// callers of the tracing entry points need around them.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lr_internal.h"
#include "lr_scene.h"

using namespace lr;

extern "C" {

int lr_buffer_alloc(lr_device *device, size_t bytes, void **out) {
    if (!device || !out) return fail(LR_INVALID, "lr_buffer_alloc needs a device and an output");
    *out = nullptr;
    (void)hipSetDevice(device->ordinal);
    hipError_t e = hipMalloc(out, bytes ? bytes : 64);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc");
    return 0;
}

void lr_buffer_free(lr_device *device, void *buffer) {
    if (!device || !buffer) return;
    (void)hipSetDevice(device->ordinal);
    (void)hipStreamSynchronize(device->stream);
    (void)hipFree(buffer);
}

int lr_buffer_write(lr_device *device, void *dst, const void *src, size_t bytes) {
    if (!device || (bytes && (!dst || !src))) return fail(LR_INVALID, "lr_buffer_write needs its arguments");
    (void)hipSetDevice(device->ordinal);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(device->stream);
    if (e != hipSuccess) return hip_fail(e, "buffer write");
    return 0;
}

int lr_buffer_read(lr_device *device, void *dst, const void *src, size_t bytes) {
    if (!device || (bytes && (!dst || !src))) return fail(LR_INVALID, "lr_buffer_read needs its arguments");
    (void)hipSetDevice(device->ordinal);
    hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(device->stream);
    if (e != hipSuccess) return hip_fail(e, "buffer read");
    return 0;
}

void *lr_device_stream(lr_device *device) { return device ? (void *)device->stream : nullptr; }

int lr_device_synchronize(lr_device *device) {
    if (!device) return fail(LR_INVALID, "lr_device_synchronize needs a device");
    hipError_t e = hipStreamSynchronize(device->stream);
    if (e != hipSuccess) return hip_fail(e, "stream synchronize");
    return 0;
}

int lr_event_create(void **event) {
    if (!event) return fail(LR_INVALID, "lr_event_create needs an output");
    hipEvent_t e;
    hipError_t s = hipEventCreate(&e);
    if (s != hipSuccess) return hip_fail(s, "hipEventCreate");
    *event = (void *)e;
    return 0;
}

int lr_event_record(void *event, void *stream) {
    hipError_t s = hipEventRecord((hipEvent_t)event, (hipStream_t)stream);
    if (s != hipSuccess) return hip_fail(s, "hipEventRecord");
    return 0;
}

int lr_event_elapsed(void *start, void *stop, float *ms) {
    if (!ms) return fail(LR_INVALID, "lr_event_elapsed needs an output");
    hipError_t s = hipEventSynchronize((hipEvent_t)stop);
    if (s == hipSuccess) s = hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop);
    if (s != hipSuccess) return hip_fail(s, "hipEventElapsedTime");
    return 0;
}

void lr_event_destroy(void *event) {
    if (event) (void)hipEventDestroy((hipEvent_t)event);
}

}  // extern "C"

namespace {

// A plain device-to-device copy (16 bytes a lane, grid-stride): the reference
// kernel of known traffic against which hardware byte counters are scaled.
__global__ void copy_kernel(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) dst[i] = src[i];
}

}  // namespace

extern "C" int lr_buffer_copy(lr_device *device, void *dst, const void *src, size_t bytes) {
    if (!device || (bytes && (!dst || !src))) return fail(LR_INVALID, "lr_buffer_copy needs its arguments");
    if (bytes % 16 || ((uintptr_t)dst | (uintptr_t)src) % 16) return fail(LR_INVALID, "lr_buffer_copy copies whole 16-byte words");
    (void)hipSetDevice(device->ordinal);
    size_t n = bytes / 16;
    if (n == 0) return 0;
    unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 256 * 64);
    hipLaunchKernelGGL(copy_kernel, dim3(blocks), dim3(256), 0, device->stream, static_cast<const float4 *>(src),
                       static_cast<float4 *>(dst), n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "copy launch");
    return 0;
}
