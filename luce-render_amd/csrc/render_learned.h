// Launchers of the learned guide inside the integrator (render_learned.hip).
#pragma once

#include <algorithm>

#include "render_common.h"

namespace lr {

// Waves of a training step's gradient kernel, and the partial rows they add
// into (wave w into row w % kNetRows).
constexpr uint32_t kNetWaves = 256u;
constexpr uint32_t kNetRows = 32u;

// Mixtures for this bounce's surface queue of the render at p[0] (`band`:
// its camera paths, an upper bound of the queue).
void launch_learned_infer(const rt::Params *p, uint32_t band, hipStream_t s);
// `steps` Adam steps on the render's path records; the kernels do nothing
// once the render is past its training samples.
void launch_learned_train(const rt::Params *p, bool spectral, uint32_t steps, hipStream_t s);

}  // namespace lr
