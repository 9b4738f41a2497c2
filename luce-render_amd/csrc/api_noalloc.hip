// The allocation-free closest-hit entry point, for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). It traces as lr_trace_closest_device does,
// by the persistent refilled schedule, with the traversal's spill area and
// the launch's counter pair in a fixed file-scope array instead of buffers
// the scene allocates on demand.
#include <hip/hip_runtime.h>

#include "lr_internal.h"
#include "lr_scene.h"
#include "render_kernels.h"
#include "trace_kernels.h"

namespace {

// 32 MiB a device, held from the library's load until the process ends: the
// counter pair (padded to 256 bytes), then the spill columns of the grid.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr size_t kScratchWords = kScratchBytes / 4;
constexpr size_t kCounterWords = 64;
constexpr uint32_t kDefaultRefill = 32;

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchWords];

namespace lr {

void scratch_attach(lr_device *device) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    device->scratch = e == hipSuccess ? ptr : nullptr;
    device->scratch_error = (int)e;
}

}  // namespace lr

using namespace lr;

extern "C" int lr_trace_closest_device_noalloc(const lr_gpu_scene *scene, const float *rays, size_t n,
                                               uint32_t *hits, void *stream) {
    if (!scene || (n && (!rays || !hits))) return fail(LR_INVALID, "lr_trace_closest_device_noalloc needs its arguments");
    if (n >= 0x80000000ull) return fail(LR_LIMIT, "at most 2^31 - 1 rays a call without allocation");
    if (n == 0) return 0;
    const lr_device *dev = scene->device;
    if (!dev->scratch) return hip_fail(dev->scratch_error, "the scratch array's device address");
    (void)hipSetDevice(dev->ordinal);
    // Spill entries a lane needs beyond its LDS stack; the grid is cut to the
    // workgroups whose columns fit the array.
    const lr_trace_config &cfg = scene->config;
    uint32_t bound = scene->info.stack_bound[bound_index(cfg.width)];
    uint64_t depth = bound > cfg.stack ? bound - cfg.stack : 0;
    uint64_t room = kScratchWords - kCounterWords;
    uint64_t max_blocks = depth ? room / (depth * cfg.group) : 0xFFFFFFFFull;
    if (max_blocks == 0)
        return fail(LR_LIMIT, "the scene's traversal stack needs more than the 32 MiB scratch array holds for one "
                              "workgroup (use lr_trace_closest_device)");
    if (max_blocks > 0xFFFFFFFFull) max_blocks = 0xFFFFFFFFull;
    uint32_t *base = static_cast<uint32_t *>(dev->scratch);
    DeviceScene sc = scene->device_scene();
    sc.work = base;
    sc.spill = base + kCounterWords;
    hipStream_t s = (hipStream_t)stream;
    // The counter pair starts at zero (the last wave of a launch also clears
    // it): a fill kernel, not a memset, so the call can be captured.
    launch_fill(sc.work, 2, 0u, s);
    int e = scene->variant().closest_persistent(sc, reinterpret_cast<const float4 *>(rays), (uint32_t)n,
                                                reinterpret_cast<uint2 *>(hits), cfg.refill ? cfg.refill : kDefaultRefill,
                                                (uint32_t)max_blocks, s);
    if (e) return hip_fail(e, "closest launch without allocation");
    return 0;
}
