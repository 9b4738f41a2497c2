// Launchers of the wavefront integrator's kernels.
#pragma once

#include <algorithm>
#include <vector>

#include "render_common.h"

namespace lr {

// ShadeProgram / ShadeLightProgram: the shading stages that run materials'
// node-graph programs.
enum class Stage {
    BeginPass,
    Camera,
    LightEmit,
    Route,
    Shade,
    ShadeLight,
    ShadeProgram,
    ShadeLightProgram,
    Splat,
    SplatShadow,
    Miss,
    Finish,
    GuideFit,
    Adaptive,
    Convert,
    EndPass
};

// The tracing kernels of a pass for one (width, stack, group) variant.
struct RenderTraceVariant {
    // `p`: a device array of `n` requests' parameters; `blocks` workgroups
    // share their items.
    void (*closest)(const rt::Params *p, uint32_t n, uint32_t current, uint32_t blocks, hipStream_t s);
    // `cutouts`: some request has cut-out materials (the walk through them)
    void (*shadow)(const rt::Params *p, uint32_t n, uint32_t blocks, bool cutouts, hipStream_t s);
    uint32_t group;
};
const RenderTraceVariant *render_trace_variants();

// `p`: a device array of `n` (at most rt::kMaxRows) requests' parameters;
// `blocks` workgroups share every request's items (1-D grid).
// `arg`: the closest-hit queue being read (Route, Shade, ShadeLight), the
// phase (Adaptive) or "convert every pass" (Convert).
void launch_render_stage(Stage stage, bool spectral, const rt::Params *p, uint32_t n, uint32_t arg, uint32_t blocks,
                         hipStream_t s);
// Camera-path shading with the learned guide (render_shade_learned.hip): the
// Shade / ShadeProgram stage with the guide standing in for part of the
// diffuse lobe only.
void launch_shade_learned(bool spectral, bool programs, const rt::Params *p, uint32_t n, uint32_t current,
                          uint32_t blocks, hipStream_t s);
// Workgroups of a pushing stage asked for `wanted`: capped at about 32 waves a
// compute unit (the kernels loop).
uint32_t staged_blocks(uint32_t group, uint32_t wanted);
void launch_tables(float *tables, hipStream_t s);
void launch_fill(uint32_t *dst, size_t words, uint32_t value, hipStream_t s);
// Adds the render at p[0]'s converged pixels into *out.
void launch_converged(const rt::Params *p, uint32_t pixels, uint32_t *out, hipStream_t s);
// The culling tree of `n` analytic lights from their host records (28
// floats each): rt::LightNode breadth-first from index 1 (index 0 unused),
// light_tree_nodes(h) nodes for the h lights that can be hit (none for
// distant and dome lights).
std::vector<rt::LightNode> light_tree(const float *lights, uint32_t n);
// Nodes of the culling tree over h lights: twice the next power of two (0
// for none); for h at most n, at most light_tree_nodes(n).
uint32_t light_tree_nodes(uint32_t h);
// The distant and dome lights of `n` host records, in index order.
std::vector<uint32_t> far_lights(const float *lights, uint32_t n);
// Copies a tree to `out` (room for nodes.size()) and waits for the copy.
hipError_t upload_light_tree(const std::vector<rt::LightNode> &nodes, rt::LightNode *out, hipStream_t s);

}  // namespace lr
