// Host-side handle types behind the C ABI's opaque pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "lr_internal.h"
#include "trace_kernels.h"

struct lr_device {
    int ordinal = 0;
    hipStream_t stream = nullptr;
    int compute_units = 0;
};

struct lr_gpu_scene {
    lr_device *device = nullptr;
    void *wide[3] = {nullptr, nullptr, nullptr};  // widths 2, 4, 8
    void *tris = nullptr;                         // 48 B per leaf slot
    void *faces = nullptr;
    void *shading = nullptr;
    void *uvs = nullptr;
    void *spill = nullptr;
    uint64_t spill_entries = 0;
    uint32_t *work = nullptr;  // the persistent launches' two counters
    void *ranges = nullptr;    // width-2 nodes' triangle slot ranges (uint2 each)
    float root_lo[3] = {0, 0, 0};
    float root_hi[3] = {0, 0, 0};
    lr_scene_info info{};
    lr_trace_config config{4, 64, 24, 32};

    ~lr_gpu_scene();
    int ensure_spill(uint64_t threads, uint32_t stack);
    int config_width_index() const;
    lr::DeviceScene device_scene() const;
    const lr::TraceVariant &variant() const;
};

namespace lr {
using GpuScene = lr_gpu_scene;
}
