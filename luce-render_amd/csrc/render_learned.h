// Launchers of the learned guide inside the integrator (render_learned.hip).
#pragma once

#include <algorithm>

#include "render_common.h"

namespace lr {

// Records a training step draws, the workgroups of its gradient kernel (32
// records each, so every workgroup writes its partial row) and those rows.
constexpr uint32_t kTrainBatch = 32768u;
constexpr uint32_t kGradGroups = kTrainBatch / 32u;
constexpr uint32_t kNetRows = kGradGroups;

// Mixtures for this bounce's surface queue of the render at p[0] (`band`:
// its camera paths, an upper bound of the queue).
void launch_learned_infer(const rt::Params *p, uint32_t band, hipStream_t s);
// Mixtures at the first hits of the pass's first sample (`band`: its camera
// paths), which every sample's first vertex reads (without light paths).
void launch_learned_infer_first(const rt::Params *p, uint32_t band, hipStream_t s);
// `steps` Adam steps on the render's path records; the kernels do nothing
// once the render is past its training samples.
void launch_learned_train(const rt::Params *p, bool spectral, uint32_t steps, hipStream_t s);

}  // namespace lr
