// Launchers of the tracing kernels, one set per (width, stack, group) variant.
#pragma once

#include "trace_device.h"

namespace lr {

struct TraceVariant {
    // Each sets sc.spill_stride to the threads it launches; the caller has
    // sized sc.spill for them (launch_threads below).
    // refill 0: one ray per thread; 1..64: persistent waves that take new
    // rays once that many lanes are idle (needs sc.work). The probe always
    // gives each thread one ray.
    int (*probe)(const DeviceScene &sc, const float *frame, uint32_t width, uint32_t height, uint2 *out,
                 uint32_t refill, hipStream_t stream);
    int (*closest)(const DeviceScene &sc, const float4 *rays, uint32_t n, uint2 *out, uint32_t refill,
                   hipStream_t stream);
    // A frame's primary rays (as probe) and a ray batch (as closest) in one
    // persistent launch when refill is above 0, else the two launches.
    int (*frame_closest)(const DeviceScene &sc, const float *frame, uint32_t width, uint32_t height, uint2 *frame_out,
                         const float4 *rays, uint32_t n, uint2 *out, uint32_t refill, hipStream_t stream);
    int (*any)(const DeviceScene &sc, const float4 *rays, uint32_t n, uint32_t *out, uint32_t refill,
               hipStream_t stream);
    // Workgroups a persistent (refilled) launch over `items` rays takes on
    // the current device: the closest-hit kernel, or the any-hit one.
    uint32_t (*resident)(uint64_t items, int any);
    // Closest hits by persistent waves (refill 1..64) on at most `max_blocks`
    // workgroups (fewer when fewer are resident or needed); sc.work and
    // sc.spill (for the grid launched) are the caller's.
    int (*closest_persistent)(const DeviceScene &sc, const float4 *rays, uint32_t n, uint2 *out, uint32_t refill,
                              uint32_t max_blocks, hipStream_t stream);
};

// 60 variants: index (layout 2/4/8/20/24, stack 16/24/32/48, group 64/128/256).
constexpr int kStackSizes = 4;
constexpr int kTraceVariants = 5 * kStackSizes * 3;
const TraceVariant *trace_variants();
int variant_index(uint32_t width, uint32_t stack, uint32_t group);

// Threads a launch over `items` work items covers at group size `group`.
inline uint64_t launch_threads(uint64_t items, uint32_t group) { return (items + group - 1) / group * group; }
__host__ __device__ inline uint64_t probe_items(uint32_t width, uint32_t height) {
    return (uint64_t)((width + 7) / 8) * ((height + 7) / 8) * 64;
}

}  // namespace lr
