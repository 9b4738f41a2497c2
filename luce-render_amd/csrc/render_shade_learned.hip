// Camera-path shading with the learned guide: the lobe-wise variants of
// shade_kernel, where the network's mixture stands in for part of the
// diffuse lobe's cosine sampling and the glossy lobes keep their own rate. A
// translation unit of their own, so that the plain and grid-guided variants
// in render_kernels.hip compile as they do without them.
#include "render_bsdf.h"
#include "render_guide.h"
#include "render_kernels.h"
#include "render_program.h"
#include "render_shade.h"

namespace lr {

void launch_shade_learned(bool spectral, bool programs, const rt::Params *p, uint32_t n, uint32_t current,
                          uint32_t blocks, hipStream_t s) {
    if (n == 0) return;
    dim3 grid(staged_blocks(128, blocks)), group(128);
    if (spectral) {
        if (programs) hipLaunchKernelGGL((rt::shade_kernel<true, false, true, true>), grid, group, 0, s, p, n, current);
        else hipLaunchKernelGGL((rt::shade_kernel<true, false, false, true>), grid, group, 0, s, p, n, current);
    } else {
        if (programs) hipLaunchKernelGGL((rt::shade_kernel<false, false, true, true>), grid, group, 0, s, p, n, current);
        else hipLaunchKernelGGL((rt::shade_kernel<false, false, false, true>), grid, group, 0, s, p, n, current);
    }
}

}  // namespace lr
