// One render's device state (a request), and the launch sequence of a pass
// over any number of renders at once.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <vector>

#include "guide_net_host.h"
#include "lr_internal.h"
#include "lr_scene.h"
#include "render_kernels.h"

namespace lr {

// Everything a render keeps on the device: loaded arrays, path state, queues,
// film, images, its pass state and the kernel parameters (a host copy and a
// device copy, which a single render's launches read).
struct RenderJob {
    lr_device *device = nullptr;
    lr_gpu_scene *scene = nullptr;
    lr_render_settings settings{};
    rt::Params params{};
    rt::Params *params_device = nullptr;
    uint32_t rounds = 0;
    uint32_t slots = 0;
    uint32_t bands = 1;
    uint32_t pass_samples = 1;  // samples a pass may carry (Params::pass_samples)
    int trace_variant = -1;
    size_t bytes = 0;  // device memory held
    std::vector<void *> buffers;
    std::unique_ptr<GuideNet> net;  // the learned guide (guiding 2)
    uint32_t train_steps = 0;
    bool adaptive = false;  // a noise threshold is set: passes check convergence
    bool cutouts = false;   // some material has geometry opacity below one: shadow rays walk through
    bool programs = false;  // some material has a node-graph program: shading runs it
    bool splats = false;    // the scene has Gaussian splats: the splat stages run
    // The loaded arrays' sizes this job's buffers hold.
    size_t light_floats = 0, material_floats = 0, emitter_words = 0, program_words = 0, texture_floats = 0,
           texture_count = 0;
    uint32_t light_paths = 0, through = 0, analytic_lights = 0;
    // What the view's buffers hold room for (pixels; path slots).
    uint32_t cap_pixels = 0, cap_slots = 0;
    uint32_t cap_width = 0, cap_height = 0;  // the view the buffers were made for
    // The constants the device holds (a host copy): set_display writes its
    // row there too, so a restart of the same loaded arrays keeps it.
    float constants[128] = {};
    // The display row (exposure, view) set by set_display.
    void set_display_row(const float row[4]);
    // Whether two constant arrays differ in the view rows (12..15) and the
    // display row at most: a restart of the same loaded arrays.
    static bool same_but_view(const float a[128], const float b[128]);
    // A job made again for a restart takes the held job's display row.
    int adopt_display(const RenderJob &held);

    ~RenderJob();
    void *alloc(size_t bytes, int &rc);
    // Validates, allocates and uploads (integrator.lucb:343-438); the film,
    // queues and pass state start cleared.
    int create(lr_device *device, lr_gpu_scene *scene, const lr_loaded *loaded, const lr_render_settings *settings,
               const float frame[16], float *shared_tables);
    // A new view or new lights and materials over the same scene whose
    // image and path slots fit the buffers held (integrator.lucb:463-540:
    // the view's buffers are kept when they fit): the loaded arrays are
    // copied into the buffers already held, the film, queues and pass state
    // cleared; at the same size the image and display keep the last frame
    // until the first new pass replaces them. kNotInPlace when the request
    // needs other buffers (the caller makes a new job).
    static constexpr int kNotInPlace = -1;
    int restart_in_place(const lr_loaded *loaded, const lr_render_settings *settings, const float frame[16]);
    // Samples the next pass carries when `done` samples are in.
    uint32_t samples_in_pass(uint32_t done) const {
        if (bands != 1u) return 1u;
        uint32_t left = settings.samples > done ? settings.samples - done : 1u;
        uint32_t k = pass_samples < left ? pass_samples : left;
        // the learned guide trains after each pass of a quarter of its training samples
        if (params.learned && done < params.train_samples)
            k = std::min(k, std::min(params.train_samples - done, std::max(1u, params.train_samples / 4u)));
        return k;
    }
    // Frees everything but the finished image; returns the bytes still held.
    size_t keep_image_only();
    // Device bytes a render of these settings holds (its admission cost).
    static size_t footprint(const lr_gpu_scene *scene, const lr_loaded *loaded, const lr_render_settings *settings);
};

uint32_t grid_blocks(uint64_t items, uint32_t group);
uint64_t spill_columns(uint64_t slots, uint32_t group);

// Launches one pass of every render in `jobs` (sharing each launch: grid row
// y is render y), their parameters at `params` (a device array in job
// order). All jobs share the spectral mode and traversal variant. With
// `marks`, an event after each stage, tagged by kernel kind.
// `train`: include the learned guide's training steps (a single render still
// within its training samples; the kernels also check on the device).
void record_pass(const std::vector<RenderJob *> &jobs, const rt::Params *params, hipStream_t s,
                 std::vector<std::pair<int, hipEvent_t>> *marks, bool train = true);

enum KernelKind { KCamera, KClosest, KSurface, KMiss, KShadow, KLight, KFinish, KConvert, KAdaptive, KRest, KKinds };

}  // namespace lr
