// Host side of the learned guide (guide_net.h): its device buffers.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "guide_net.h"

struct lr_device;

namespace lr {

struct GuideNet {
    lr_device *device = nullptr;
    net::NetBuffers b{};
    float *partials = nullptr;  // waves rows of kParams + 1 gradient sums
    uint32_t waves = 0;
    std::vector<void *> buffers;

    ~GuideNet();
    // Fresh network for points within [lo, hi]; `max_waves` bounds the
    // gradient rows a training batch may use.
    int create(lr_device *device, const float lo[3], const float hi[3], uint32_t seed, uint32_t max_waves);
    int set_params(const float *params);  // also clears the Adam moments
    void refresh(hipStream_t s);          // FP16 fragments from the master copy
};

}  // namespace lr

struct lr_guide_net {
    lr::GuideNet net;
};
