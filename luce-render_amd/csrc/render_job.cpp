// A render's device state and the pass launch sequence
// (src/render/integrator.lucb:417-438 allocate_view, :788-828 record).
#include "render_job.h"

#include <algorithm>
#include <cstring>
#include <memory>

#include "render_bsdf.h"
#include "render_guide.h"
#include "render_learned.h"

namespace lr {

uint32_t grid_blocks(uint64_t items, uint32_t group) {
    uint64_t b = (items + group - 1) / group;
    if (b < 1) b = 1;
    if (b > 65535) b = 65535;
    return (uint32_t)b;
}


// Columns of a render's traversal spill area. The render's tracing kernels
// take an item's index in its render's queue (below the render's slots) as
// its spill column, whatever grid the launch takes, so there is a column for
// every slot: no clamp to a launch's grid.
uint64_t spill_columns(uint64_t slots, uint32_t group) {
    uint64_t g = group ? group : 1;
    return (slots + g - 1) / g * g;
}

namespace {

struct Shape {
    uint32_t pixels, band, light_paths, slots, bands;
    uint32_t pass_samples;  // samples a pass carries at most; slots hold that many
};

// Slots a pass of several samples aims at: about 2^26, 20-28 GB of path state
// and traversal spill, under a tenth of the 288 GB a GPU holds. Each pass
// pays the tail rounds' floor (the last incoherent rays' chains of node
// fetches) once, however many samples it carries: 1080p with light paths
// takes 16 samples a pass (30.4M-triangle scene, 16 samples: 19.2 ms a sample
// one sample a pass, 10.0 in passes of 10 and 6, 9.5 in one pass of 16).
constexpr uint64_t kPassSlots = 1ull << 26;
constexpr uint32_t kMostPassSamples = 16;

Shape shape_of(const lr_loaded *loaded, const lr_render_settings &st) {
    Shape s;
    s.pixels = st.width * st.height;
    uint32_t pool = st.pool ? st.pool : 2097152u;  // integrator.lucb:96-98
    s.band = pool < s.pixels ? pool : s.pixels;
    s.light_paths = loaded->light_paths < s.band ? loaded->light_paths : s.band;
    uint32_t one = s.band + s.light_paths;
    s.bands = (s.pixels + s.band - 1) / s.band;
    // Several samples a pass only where a sample is one band and adaptive
    // sampling (which checks between samples) is off. A guided render's
    // guide then learns after each pass of samples (the reference's after
    // each sample: pass_samples 1); the learned guide trains after each pass
    // of a quarter of its training samples.
    bool alone = s.bands == 1u && !(loaded->constants[44] > 0.0f);
    uint64_t k = 1;
    if (alone) {
        k = st.pass_samples ? st.pass_samples : std::max<uint64_t>(1, std::min<uint64_t>(kMostPassSamples, kPassSlots / one));
        k = std::min<uint64_t>(k, std::max<uint32_t>(st.samples, 1u));
        while (k > 1 && (uint64_t)one * k > 0xFFFFFF00ull) k--;
    }
    s.pass_samples = (uint32_t)k;
    s.slots = (uint32_t)((uint64_t)one * k);
    return s;
}

int width_index(uint32_t w) { return bound_index(w); }

}  // namespace

RenderJob::~RenderJob() {
    for (void *b : buffers) (void)hipFree(b);
}

void *RenderJob::alloc(size_t n, int &rc) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, n ? n : 64);
    if (e != hipSuccess) {
        rc = hip_fail(e, "hipMalloc(render state)");
        return nullptr;
    }
    buffers.push_back(p);
    bytes += n ? n : 64;
    return p;
}

static int check_request(const lr_gpu_scene *sc, const lr_loaded *loaded, const lr_render_settings &st,
                         const float frame[16], bool &with_programs);

int RenderJob::restart_in_place(const lr_loaded *loaded, const lr_render_settings *settings_in, const float frame[16]) {
    if (!loaded || !settings_in || !frame) return fail(LR_INVALID, "a render needs its arguments");
    const lr_render_settings &st = *settings_in;
    bool with_programs = false;
    if (int crc = check_request(scene, loaded, st, frame, with_programs)) return crc;
    uint32_t w = st.trace_width ? st.trace_width : scene->config.width;
    uint32_t stack = st.trace_stack ? st.trace_stack : scene->config.stack;
    uint32_t group = st.trace_group ? st.trace_group : scene->config.group;
    // Buffers that fit: the image and the path slots within what is held,
    // the loaded arrays of the same sizes, the same launches but for their
    // item counts.
    Shape sh = shape_of(loaded, st);
    bool with_splats = scene->splat_count > 0 && loaded->constants[106] > 0.0f;
    bool same = with_splats == splats && sh.pixels <= cap_pixels && sh.slots <= cap_slots && st.max_bounces == settings.max_bounces &&
                st.spectral == settings.spectral && st.pool == settings.pool && st.guiding == settings.guiding &&
                st.guiding != 2 && w == settings.trace_width && stack == settings.trace_stack &&
                group == settings.trace_group && loaded->light_floats == light_floats &&
                loaded->material_floats == material_floats && loaded->emitter_words == emitter_words &&
                loaded->through == through && with_programs == programs &&
                (!with_programs || (loaded->program_words == program_words && loaded->texture_floats == texture_floats &&
                                    loaded->texture_count == texture_count)) &&
                (uint32_t)loaded->constants[0] == analytic_lights;
    if (!same) return kNotInPlace;
    rt::Params &p = params;
    hipStream_t s = device->stream;
    hipError_t e = hipSuccess;
    auto put = [&](const void *dst, const void *src, size_t n) {
        if (e == hipSuccess && n) e = hipMemcpyAsync(const_cast<void *>(dst), src, n, hipMemcpyHostToDevice, s);
    };
    // A restart (integrator.lucb:463-466): the same loaded arrays but for the
    // view rows (12..15) keep the display row set_display wrote (the
    // reference uploads the view rows alone). New lights or materials (a
    // reload) take the display row they were packed with.
    float next[128];
    std::memcpy(next, loaded->constants, sizeof(next));
    if (same_but_view(next, constants))
        for (int k = 0; k < 4; k++) next[rt::K_DISPLAY * 4 + k] = constants[rt::K_DISPLAY * 4 + k];
    put(p.lights, loaded->lights, loaded->light_floats * 4);
    put(p.constants, next, 128 * 4);
    put(p.materials, loaded->materials, loaded->material_floats * 4);
    put(p.emitters, loaded->emitters, loaded->emitter_words * 4);
    if (with_programs) {
        put(p.programs, loaded->programs, loaded->program_words * 4);
        if (loaded->texture_count) {
            put(p.textures, loaded->textures, loaded->texture_floats / 4 * 16);
            put(p.texture_table, loaded->texture_table, loaded->texture_count * 16);
        }
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s);  // `next` is on this stack
    if (e != hipSuccess) return hip_fail(e, "restart upload");
    std::memcpy(constants, next, sizeof(constants));
    {
        // The same light count: a tree of at most the nodes held.
        std::vector<rt::LightNode> tree = light_tree(loaded->lights, analytic_lights);
        e = upload_light_tree(tree, const_cast<rt::LightNode *>(p.light_tree), s);
        if (e != hipSuccess) return hip_fail(e, "restart light tree");
        p.light_nodes = (uint32_t)tree.size();
        std::vector<uint32_t> far = far_lights(loaded->lights, analytic_lights);
        if (!far.empty()) e = hipMemcpyAsync(const_cast<uint32_t *>(p.far_lights), far.data(), far.size() * 4, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);  // `far` is on this stack
        if (e != hipSuccess) return hip_fail(e, "restart far lights");
        p.far_count = (uint32_t)far.size();
    }
    bool resized = st.width != settings.width || st.height != settings.height;
    settings.samples = st.samples;
    settings.seed = st.seed;
    settings.width = st.width;
    settings.height = st.height;
    std::memcpy(p.frame, frame, 64);
    p.seed = st.seed;
    slots = sh.slots;
    bands = sh.bands;
    pass_samples = sh.pass_samples;
    light_paths = loaded->light_paths;
    p.pool = slots;
    p.pass_samples = pass_samples;
    p.sample_target = st.samples;
    p.pixels = sh.pixels;
    p.width = st.width;
    p.height = st.height;
    p.band_size = sh.band;
    p.light_paths = sh.light_paths;
    p.scene.spill_stride = (uint32_t)spill_columns(slots, group);
    if (resized) {
        launch_fill((uint32_t *)p.image, (size_t)p.pixels * 4, 0u, s);
        launch_fill((uint32_t *)p.display, (size_t)p.pixels * 4, 0u, s);
    }
    adaptive = loaded->constants[44] > 0.0f;
    cutouts = loaded->constants[3] >= 0.5f;
    launch_fill((uint32_t *)p.film, (size_t)4 * p.pixels * 4, 0u, s);
    launch_fill(p.queues, rt::kQueueBase + (size_t)rt::kQueues * slots, 0u, s);
    launch_fill(p.state_u, (size_t)rt::kStateU * slots, 0u, s);
    if (p.guide) launch_fill((uint32_t *)p.guide, rt::kGuideBytes / 4, 0u, s);
    rt::PassState init{};
    init.bands = bands;
    e = hipMemcpyAsync(p.pass, &init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(params_device, &p, sizeof(p), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "restart");
    return 0;
}

bool RenderJob::same_but_view(const float a[128], const float b[128]) {
    for (int k = 0; k < 128; k++) {
        int row = k / 4;
        if (row == rt::K_DISPLAY || (row >= 12 && row <= 15)) continue;
        if (std::memcmp(&a[k], &b[k], 4) != 0) return false;
    }
    return true;
}

int RenderJob::adopt_display(const RenderJob &held) {
    if (!same_but_view(constants, held.constants)) return 0;
    const float *row = held.constants + rt::K_DISPLAY * 4;
    set_display_row(row);
    hipError_t e = hipMemcpyAsync((char *)params.constants + rt::K_DISPLAY * 16, row, 16, hipMemcpyHostToDevice,
                                  device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(device->stream);
    return e == hipSuccess ? 0 : hip_fail(e, "display settings");
}

void RenderJob::set_display_row(const float row[4]) {
    for (int k = 0; k < 4; k++) constants[rt::K_DISPLAY * 4 + k] = row[k];
}

size_t RenderJob::keep_image_only() {
    net.reset();
    std::vector<void *> kept;
    for (void *b : buffers) {
        if (b == (void *)params.image) kept.push_back(b);
        else (void)hipFree(b);
    }
    buffers.swap(kept);
    bytes = (size_t)params.pixels * 16;
    return bytes;
}

size_t RenderJob::footprint(const lr_gpu_scene *scene, const lr_loaded *loaded, const lr_render_settings *st) {
    Shape s = shape_of(loaded, *st);
    size_t b = 0;
    b += loaded->light_floats * 4 + 512 + loaded->material_floats * 4 + loaded->emitter_words * 4;
    b += loaded->program_words * 4 + loaded->texture_floats * 4 + loaded->texture_count * 16;
    uint32_t nlights = (uint32_t)loaded->constants[0];
    b += (size_t)std::max(1u, light_tree_nodes(nlights)) * sizeof(rt::LightNode);  // the lights' culling tree, as create allocates
    b += (size_t)std::max(1u, nlights) * 4;                             // and the distant lights' list (at most all)
    b += (size_t)rt::kStateV * s.slots * 16 + (size_t)rt::kStateU * s.slots * 4;
    b += ((size_t)rt::kQueueBase + (size_t)rt::kQueues * s.slots) * 4;
    b += (size_t)6 * s.pixels * 16;  // film (4 planes), image, display
    if (st->guiding == 1) b += rt::kGuideBytes;
    if (st->guiding == 2)
        b += (size_t)2 * s.slots * 16 + (size_t)kNetRows * (net::kParams + 1) * 4 + (size_t)net::kParams * 12 +
             net::kFragmentHalves * 4 + 128;
    uint32_t w = st->trace_width ? st->trace_width : scene->config.width;
    uint32_t stack = st->trace_stack ? st->trace_stack : scene->config.stack;
    uint32_t group = st->trace_group ? st->trace_group : scene->config.group;
    uint32_t bound = scene->info.stack_bound[width_index(w)];
    if (bound > stack) b += (size_t)(bound - stack) * spill_columns(s.slots, group) * 4;
    return b;
}

// A request's checks (integrator.lucb:343-438): sizes, the loaded arrays,
// the node-graph programs' words and textures. `with_programs`: some material
// has a program.
static int check_request(const lr_gpu_scene *sc, const lr_loaded *loaded, const lr_render_settings &st,
                         const float frame[16], bool &with_programs) {
    // integrator.lucb:344, 376-377
    if (st.width == 0 || st.height == 0 || st.width > 16384 || st.height > 16384)
        return fail(LR_INVALID, "an image is 1 to 16384 pixels on a side");
    if (st.samples == 0 || st.max_bounces == 0 || st.max_bounces > 1024)
        return fail(LR_INVALID, "a render needs samples and 1 to 1024 bounces");
    if (st.guiding > 2) return fail(LR_INVALID, "guiding is 0 (off), 1 (hash grid) or 2 (learned)");
    if ((uint32_t)frame[11] != st.width || (uint32_t)frame[15] != st.height)
        return fail(LR_INVALID, "the frame's size is not the settings' image size");
    if (!loaded->constants || !loaded->lights || !loaded->materials || !loaded->emitters || loaded->material_floats < 64)
        return fail(LR_INVALID, "a render needs the loaded lights, constants, materials (the default one at least) and emitters");
    if (!sc->shading) return fail(LR_INVALID, "rendering needs the scene's shading attributes");
    // Node-graph programs: every entry a material names lies in the words,
    // every texture's texels in the texel array (the kernels also stop at
    // the words' end).
    with_programs = false;
    for (size_t m = 0; (m + 1) * 64 <= loaded->material_floats; m++) {
        float entry = loaded->materials[m * 64 + 52];  // row 13 x
        if (!(entry >= 0.0f)) continue;
        if (!loaded->programs || (double)entry >= (double)loaded->program_words)
            return fail(LR_INVALID, "a material's node-graph program lies outside the program words");
        with_programs = true;
    }
    if (with_programs) {
        if (!sc->uvs) return fail(LR_INVALID, "node-graph programs need the scene's texture coordinates");
        if (loaded->program_words > 0xFFFFFFFFull || loaded->texture_count > 0xFFFFFFFFull)
            return fail(LR_INVALID, "too many program words or textures");
        if (loaded->texture_count && (!loaded->texture_table || !loaded->textures))
            return fail(LR_INVALID, "textures need their table and texels");
        size_t texels = loaded->texture_floats / 4;
        for (size_t t = 0; t < loaded->texture_count; t++) {
            const uint32_t *row = loaded->texture_table + t * 4;
            if ((uint64_t)row[0] + (uint64_t)row[1] * row[2] > texels)
                return fail(LR_INVALID, "a texture's texels lie outside the texel array");
        }
    }
    return 0;
}

int RenderJob::create(lr_device *dev, lr_gpu_scene *sc, const lr_loaded *loaded, const lr_render_settings *settings_in,
                      const float frame[16], float *shared_tables) {
    if (!dev || !sc || !loaded || !settings_in || !frame) return fail(LR_INVALID, "a render needs its arguments");
    const lr_render_settings &st = *settings_in;
    bool with_programs = false;
    if (int crc = check_request(sc, loaded, st, frame, with_programs)) return crc;
    device = dev;
    scene = sc;
    settings = st;
    uint32_t w = st.trace_width ? st.trace_width : sc->config.width;
    uint32_t stack = st.trace_stack ? st.trace_stack : sc->config.stack;
    uint32_t group = st.trace_group ? st.trace_group : sc->config.group;
    trace_variant = variant_index(w, stack, group);
    if (trace_variant < 0) return fail(LR_INVALID, "traversal layouts are 2, 4, 8, 20, 24; stacks 16, 24, 32, 48; groups 64, 128, 256");
    settings.trace_width = w;
    settings.trace_stack = stack;
    settings.trace_group = group;
    Shape sh = shape_of(loaded, st);
    slots = sh.slots;
    bands = sh.bands;
    pass_samples = sh.pass_samples;
    cap_pixels = sh.pixels;
    cap_slots = sh.slots;
    cap_width = st.width;
    cap_height = st.height;
    rounds = st.max_bounces + 1 + loaded->through;
    adaptive = loaded->constants[44] > 0.0f;  // the noise threshold (loading.lucb:267)
    cutouts = loaded->constants[3] >= 0.5f;   // K_INFO.w
    int rc = 0;
    rt::Params &p = params;
    std::memcpy(p.frame, frame, 64);
    p.pool = slots;
    p.pass_samples = pass_samples;
    p.sample_target = st.samples;
    p.seed = st.seed;
    p.pixels = sh.pixels;
    p.width = st.width;
    p.height = st.height;
    p.band_size = sh.band;
    p.light_paths = sh.light_paths;
    p.spectral = st.spectral;
    hipStream_t s = dev->stream;
    auto upload = [&](const void *src, size_t n) -> void * {
        void *d = alloc(n, rc);
        if (!d) return nullptr;
        if (n) {
            hipError_t e = hipMemcpyAsync(d, src, n, hipMemcpyHostToDevice, s);
            if (e != hipSuccess) {
                rc = hip_fail(e, "render upload");
                return nullptr;
            }
        }
        return d;
    };
    light_floats = loaded->light_floats;
    material_floats = loaded->material_floats;
    emitter_words = loaded->emitter_words;
    program_words = loaded->program_words;
    texture_floats = loaded->texture_floats;
    texture_count = loaded->texture_count;
    light_paths = loaded->light_paths;
    through = loaded->through;
    std::memcpy(constants, loaded->constants, sizeof(constants));
    p.lights = (const float4 *)upload(loaded->lights, loaded->light_floats * 4);
    p.constants = (const float4 *)upload(constants, 128 * 4);
    p.materials = (const float4 *)upload(loaded->materials, loaded->material_floats * 4);
    p.emitters = (const uint32_t *)upload(loaded->emitters, loaded->emitter_words * 4);
    if (rc) return rc;
    // The analytic lights (their count is the first constant): the culling
    // tree of those a ray can hit and the list of the distant ones.
    uint32_t nlights = (uint32_t)loaded->constants[0];
    if ((size_t)nlights * 28u > loaded->light_floats) return fail(LR_INVALID, "the light count exceeds the light records");
    analytic_lights = nlights;
    {
        std::vector<rt::LightNode> tree = light_tree(loaded->lights, nlights);
        // room for a tree of every light, so a reload keeps the buffer
        p.light_tree = (const rt::LightNode *)alloc((size_t)std::max(1u, light_tree_nodes(nlights)) * sizeof(rt::LightNode), rc);
        if (rc) return rc;
        if (hipError_t e = upload_light_tree(tree, const_cast<rt::LightNode *>(p.light_tree), s))
            return hip_fail(e, "light tree");
        p.light_nodes = (uint32_t)tree.size();
        std::vector<uint32_t> far = far_lights(loaded->lights, nlights);
        // room for every light, so a reload keeps the buffer
        p.far_lights = (const uint32_t *)alloc((size_t)std::max(1u, nlights) * 4, rc);
        if (rc) return rc;
        hipError_t e = far.empty() ? hipSuccess
                                   : hipMemcpyAsync(const_cast<uint32_t *>(p.far_lights), far.data(), far.size() * 4,
                                                    hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);  // `far` is on this stack
        if (e != hipSuccess) return hip_fail(e, "far lights");
        p.far_count = (uint32_t)far.size();
    }
    float *tables = shared_tables;
    if (!tables) {
        tables = (float *)alloc((size_t)rt::TABLE_SIZE * 4, rc);
        if (rc) return rc;
        launch_tables(tables, s);
    }
    p.tables = tables;
    p.shading = (const uint4 *)sc->shading;
    p.uvs = (const float *)sc->uvs;
    programs = with_programs;
    splats = sc->splat_count > 0 && loaded->constants[106] > 0.0f;  // K_SPLAT_TREE z
    if (splats) {
        p.splat_records = (const uint4 *)sc->splat_records;
        p.splat_words = (const uint32_t *)sc->splat_words;
        p.splat_nodes = (const float4 *)sc->splat_nodes;
        p.splat_count = (uint32_t)sc->splat_count;
    }
    if (with_programs) {
        p.programs = (const uint32_t *)upload(loaded->programs, loaded->program_words * 4);
        p.program_words = (uint32_t)loaded->program_words;
        p.texture_count = (uint32_t)loaded->texture_count;
        if (loaded->texture_count) {
            p.textures = (const float4 *)upload(loaded->textures, loaded->texture_floats / 4 * 16);
            p.texture_table = (const uint4 *)upload(loaded->texture_table, loaded->texture_count * 16);
        }
        if (rc) return rc;
    }
    p.state_v = (float4 *)alloc((size_t)rt::kStateV * slots * 16, rc);
    p.state_u = (uint32_t *)alloc((size_t)rt::kStateU * slots * 4, rc);
    size_t queue_words = rt::kQueueBase + (size_t)rt::kQueues * slots;
    p.queues = (uint32_t *)alloc(queue_words * 4, rc);
    p.film = (float4 *)alloc((size_t)4 * sh.pixels * 16, rc);
    p.image = (float4 *)alloc((size_t)sh.pixels * 16, rc);
    p.display = (float4 *)alloc((size_t)sh.pixels * 16, rc);
    p.pass = (rt::PassState *)alloc(sizeof(rt::PassState), rc);
    if (st.guiding == 1) p.guide = (uint8_t *)alloc(rt::kGuideBytes, rc);
    if (st.guiding == 2) {
        net.reset(new GuideNet());
        int nrc = net->create(dev, sc->root_lo, sc->root_hi, st.seed, kNetRows);
        if (nrc) return nrc;
        p.net = net->b;
        p.net_partials = net->partials;
        p.guide_mix = (uint4 *)alloc((size_t)2 * slots * 16, rc);
        p.learned = 1;
        p.train_samples = st.train_samples ? st.train_samples : 8u;
        p.train_batch = kTrainBatch;
        p.train_rate = 5e-3f;
        train_steps = st.train_steps ? st.train_steps : 8u;
    }
    params_device = (rt::Params *)alloc(sizeof(rt::Params), rc);
    if (rc) return rc;
    // The traversal's spill area for the largest grid a pass launches, and
    // the node layout of the chosen width.
    p.scene = sc->device_scene();
    p.scene.nodes = (const uint8_t *)sc->wide[layout_index(w)];
    uint32_t bound = sc->info.stack_bound[width_index(w)];
    uint64_t depth = bound > stack ? bound - stack : 0;
    uint64_t threads = spill_columns(slots, group);
    if (threads > 0xFFFFFFFFull || threads < slots) return fail(LR_LIMIT, "too many path slots for the spill area");
    p.scene.spill = depth ? (uint32_t *)alloc(depth * threads * 4, rc) : nullptr;
    p.scene.spill_stride = (uint32_t)threads;
    p.scene.spill_depth = (uint32_t)depth;
    p.scene.work = nullptr;  // the render's tracing launches give each thread one ray
    if (rc) return rc;
    launch_fill((uint32_t *)p.film, (size_t)4 * sh.pixels * 4, 0u, s);
    launch_fill((uint32_t *)p.image, (size_t)sh.pixels * 4, 0u, s);
    launch_fill(p.queues, queue_words, 0u, s);
    launch_fill(p.state_u, (size_t)rt::kStateU * slots, 0u, s);
    if (p.guide) launch_fill((uint32_t *)p.guide, rt::kGuideBytes / 4, 0u, s);
    if (p.guide_mix) launch_fill((uint32_t *)p.guide_mix, (size_t)2 * slots * 4, 0u, s);
    rt::PassState init{};
    init.bands = bands;
    hipError_t e = hipMemcpyAsync(p.pass, &init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(params_device, &p, sizeof(p), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "render set-up");
    return 0;
}

void record_pass(const std::vector<RenderJob *> &jobs, const rt::Params *params, hipStream_t s,
                 std::vector<std::pair<int, hipEvent_t>> *marks, bool train) {
    if (jobs.empty()) return;
    uint32_t n = (uint32_t)jobs.size();
    bool sp = jobs[0]->settings.spectral != 0;
    const RenderTraceVariant &tv = render_trace_variants()[jobs[0]->trace_variant];
    uint32_t rounds = 0;
    uint64_t lights = 0;
    bool guide = false;
    for (RenderJob *j : jobs) {
        rounds = std::max(rounds, j->rounds);
        lights += j->params.light_paths;
        guide = guide || j->params.guide;
    }
    // The learned guide runs for single renders (the server refuses it).
    bool learned = n == 1 && jobs[0]->params.learned;
    // Workgroups of a stage: enough for every request's largest item count
    // (the pushing stages' launchers cut this down; their kernels loop).
    auto blocks = [&](uint32_t group, auto items) {
        uint64_t b = 0;
        for (RenderJob *j : jobs) b += (items(j) + group - 1) / group;
        return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, 0x7fffffffu));
    };
    // camera paths and light paths a pass holds (its samples' together)
    auto band = [](RenderJob *j) -> uint64_t { return (uint64_t)j->params.band_size * j->pass_samples; };
    auto light_paths = [](RenderJob *j) -> uint64_t { return (uint64_t)j->params.light_paths * j->pass_samples; };
    auto band_items = [](RenderJob *j) -> uint64_t { return j->params.band_size; };
    auto slots = [](RenderJob *j) -> uint64_t { return j->slots; };
    auto pixels = [](RenderJob *j) -> uint64_t { return j->params.pixels; };
    bool adaptive = false, cutouts = false, programs = false, splats = false;
    for (RenderJob *j : jobs) adaptive = adaptive || j->adaptive;
    for (RenderJob *j : jobs) cutouts = cutouts || j->cutouts;
    for (RenderJob *j : jobs) programs = programs || j->programs;
    for (RenderJob *j : jobs) splats = splats || j->splats;
    auto mark = [&](int kind) {
        if (!marks) return;
        hipEvent_t e;
        (void)hipEventCreate(&e);
        (void)hipEventRecord(e, s);
        marks->push_back({kind, e});
    };
    launch_render_stage(Stage::BeginPass, sp, params, n, 0, 1, s);
    mark(KRest);
    launch_render_stage(Stage::Camera, sp, params, n, 0, blocks(256, band), s);
    mark(KCamera);
    if (lights) {
        launch_render_stage(Stage::LightEmit, sp, params, n, 0, blocks(256, light_paths), s);
        mark(KLight);
    }
    // One round more than bounces: the last BSDF rays still find lights
    // (integrator.lucb:806-814).
    for (uint32_t b = 0; b < rounds; b++) {
        uint32_t cur = b % 2u;
        tv.closest(params, n, cur, blocks(tv.group, slots), s);
        if (splats) launch_render_stage(Stage::Splat, sp, params, n, cur, blocks(128, slots), s);
        launch_render_stage(Stage::Route, sp, params, n, cur, blocks(256, slots), s);
        mark(KClosest);
        // Shading consults the guide only at the first GUIDE_RING vertices
        // below the bounce limit; a path's vertex count trails its round by
        // at most the cut-out pass-throughs, so later rounds infer nothing
        // anyone reads.
        // With light paths the guide is consulted at the second vertex alone,
        // so only the second round (plus the pass-throughs) infers. Without
        // them, the first round infers once a pixel: at the first sample's
        // hits, which every sample's first vertex reads.
        const uint32_t guided_rounds = lights > 0 ? 2u : rt::GUIDE_RING;
        if (learned && b < std::min(guided_rounds, jobs[0]->settings.max_bounces) + jobs[0]->through &&
            !(b == 0 && lights > 0)) {
            if (b == 0) launch_learned_infer_first(params, jobs[0]->params.band_size, s);
            else launch_learned_infer(params, jobs[0]->params.band_size * jobs[0]->pass_samples, s);
            mark(KRest);
        }
        if (learned) launch_shade_learned(sp, programs, params, n, cur, blocks(128, band), s);
        else launch_render_stage(programs ? Stage::ShadeProgram : Stage::Shade, sp, params, n, cur, blocks(128, band), s);
        if (lights)
            launch_render_stage(programs ? Stage::ShadeLightProgram : Stage::ShadeLight, sp, params, n, cur,
                                blocks(128, light_paths), s);
        mark(KSurface);
        launch_render_stage(Stage::Miss, sp, params, n, 0, blocks(256, slots), s);
        mark(KMiss);
        if (splats) launch_render_stage(Stage::SplatShadow, sp, params, n, 0, blocks(128, slots), s);
        tv.shadow(params, n, blocks(tv.group, slots), cutouts, s);
        mark(KShadow);
    }
    launch_render_stage(Stage::Finish, sp, params, n, 0, blocks(256, band_items), s);
    mark(KFinish);
    if (learned && train) {
        // train_steps for each training sample the pass carried
        RenderJob *j = jobs[0];
        uint32_t carried = std::min(j->pass_samples, std::max(1u, j->params.train_samples / 4u));
        launch_learned_train(params, sp, j->train_steps * carried, s);
        mark(KRest);
    }
    if (guide) {
        launch_render_stage(Stage::GuideFit, sp, params, n, 0, 2048, s);
        mark(KRest);
    }
    if (adaptive) {
        launch_render_stage(Stage::Adaptive, sp, params, n, 0, blocks(256, pixels), s);
        launch_render_stage(Stage::Adaptive, sp, params, n, 1, blocks(256, pixels), s);
        mark(KAdaptive);
    }
    launch_render_stage(Stage::Convert, sp, params, n, 0, blocks(256, pixels), s);
    mark(KConvert);
    launch_render_stage(Stage::EndPass, sp, params, n, 0, 1, s);
    mark(KRest);
}

}  // namespace lr
