// The C ABI's renderer (Renderer.create / advance / finish / samples_done /
// image / read / kernel_times / destroy, src/render/integrator.lucb:343-903).
//
// A pass is the reference's record() (integrator.lucb:788-828): begin, camera,
// light emission, max_bounces + 1 (+ cut-out) rounds of closest / shade / miss
// / shadow, finish, guide fit, film convert, end. Its launches never change
// (sample and band live on the device), so the pass is recorded once into a
// hipGraph and replayed; up to two passes are in flight, as advance keeps them
// (integrator.lucb:651-655).
#include <hip/hip_runtime.h>

#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "lr_internal.h"
#include "lr_scene.h"
#include "render_bsdf.h"
#include "render_guide.h"
#include "render_kernels.h"

using namespace lr;

namespace {

constexpr int kKinds = 10;  // integrator.lucb:115 kernel_names
enum Kind { KCamera, KClosest, KSurface, KMiss, KShadow, KLight, KFinish, KConvert, KAdaptive, KRest };

struct Pass {
    hipEvent_t done;
    uint32_t completes;  // samples every pixel has once it finishes
    std::vector<std::pair<int, hipEvent_t>> marks;
    hipEvent_t start = nullptr;
};

}  // namespace

struct lr_renderer {
    lr_device *device = nullptr;
    lr_gpu_scene *scene = nullptr;
    lr_render_settings settings{};
    rt::Params params{};
    uint32_t rounds = 0;
    uint32_t slots = 0;
    uint32_t bands = 1;
    // host mirror of the device pass state
    uint32_t sample = 0, band = 0;
    uint32_t completed = 0;
    std::deque<Pass> in_flight;
    double kernel_ns[kKinds] = {0};
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<void *> buffers;
    int trace_variant = -1;

    void *alloc(size_t bytes, int &rc) {
        void *p = nullptr;
        hipError_t e = hipMalloc(&p, bytes ? bytes : 64);
        if (e != hipSuccess) {
            rc = hip_fail(e, "hipMalloc(renderer)");
            return nullptr;
        }
        buffers.push_back(p);
        return p;
    }
    ~lr_renderer() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &ps : in_flight) {
            (void)hipEventDestroy(ps.done);
            if (ps.start) (void)hipEventDestroy(ps.start);
            for (auto &m : ps.marks) (void)hipEventDestroy(m.second);
        }
        for (void *b : buffers) (void)hipFree(b);
    }

    uint32_t blocks(uint64_t items, uint32_t group) const {
        uint64_t b = (items + group - 1) / group;
        if (b < 1) b = 1;
        if (b > 65535) b = 65535;
        return (uint32_t)b;
    }

    // Launches one pass; with `marks`, an event after each stage.
    void record(hipStream_t s, std::vector<std::pair<int, hipEvent_t>> *marks) {
        bool sp = settings.spectral != 0;
        const RenderTraceVariant &tv = render_trace_variants()[trace_variant];
        auto mark = [&](int kind) {
            if (!marks) return;
            hipEvent_t e;
            (void)hipEventCreate(&e);
            (void)hipEventRecord(e, s);
            marks->push_back({kind, e});
        };
        launch_render_stage(Stage::BeginPass, sp, params, 0, 1, s);
        mark(KRest);
        launch_render_stage(Stage::Camera, sp, params, 0, blocks(params.band_size, 256), s);
        mark(KCamera);
        if (params.light_paths) {
            launch_render_stage(Stage::LightEmit, sp, params, 0, blocks(params.light_paths, 256), s);
            mark(KLight);
        }
        for (uint32_t b = 0; b < rounds; b++) {
            uint32_t cur = b % 2u;
            tv.closest(params, cur, blocks(slots, tv.group), s);
            mark(KClosest);
            launch_render_stage(Stage::Shade, sp, params, cur, blocks(slots, 128), s);
            mark(KSurface);
            launch_render_stage(Stage::Miss, sp, params, 0, blocks(slots, 256), s);
            mark(KMiss);
            tv.shadow(params, blocks(slots, tv.group), s);
            mark(KShadow);
        }
        launch_render_stage(Stage::Finish, sp, params, 0, blocks(params.band_size, 256), s);
        mark(KFinish);
        if (params.guide) {
            launch_render_stage(Stage::GuideFit, sp, params, 0, 2048, s);
            mark(KRest);
        }
        launch_render_stage(Stage::Convert, sp, params, 0, blocks(params.pixels, 256), s);
        mark(KConvert);
        launch_render_stage(Stage::EndPass, sp, params, 0, 1, s);
        mark(KRest);
    }

    // Retire finished passes (integrator.lucb:622-638).
    int retire(bool wait_all) {
        while (!in_flight.empty()) {
            Pass &ps = in_flight.front();
            hipError_t q = wait_all ? hipEventSynchronize(ps.done) : hipEventQuery(ps.done);
            if (q == hipErrorNotReady) break;
            if (q != hipSuccess) return hip_fail(q, "pass");
            completed = ps.completes;
            if (!ps.marks.empty()) {
                hipEvent_t prev = ps.start;
                for (auto &m : ps.marks) {
                    float ms = 0;
                    if (prev && hipEventElapsedTime(&ms, prev, m.second) == hipSuccess) kernel_ns[m.first] += ms * 1e6;
                    prev = m.second;
                }
                for (auto &m : ps.marks) (void)hipEventDestroy(m.second);
                (void)hipEventDestroy(ps.start);
            }
            (void)hipEventDestroy(ps.done);
            in_flight.pop_front();
        }
        return 0;
    }

    int launch_pass() {
        hipStream_t s = device->stream;
        Pass ps;
        if (settings.profile) {
            (void)hipEventCreate(&ps.start);
            (void)hipEventRecord(ps.start, s);
            record(s, &ps.marks);
        } else if (settings.use_graph) {
            if (!exec) {
                hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
                record(s, nullptr);
                e = hipStreamEndCapture(s, &graph);
                if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
                e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            }
            hipError_t e = hipGraphLaunch(exec, s);
            if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        } else {
            record(s, nullptr);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "pass launch");
        band += 1;
        if (band == bands) {
            band = 0;
            sample += 1;
        }
        ps.completes = sample;
        (void)hipEventCreateWithFlags(&ps.done, hipEventDisableTiming);
        (void)hipEventRecord(ps.done, s);
        in_flight.push_back(ps);
        return 0;
    }
};

extern "C" {

int lr_renderer_create(lr_device *device, lr_gpu_scene *scene, const lr_loaded *loaded,
                       const lr_render_settings *settings, const float frame[16], lr_renderer **out) {
    if (!device || !scene || !loaded || !settings || !frame || !out)
        return fail(LR_INVALID, "lr_renderer_create needs its arguments");
    *out = nullptr;
    const lr_render_settings &st = *settings;
    // integrator.lucb:344, 376-377
    if (st.width == 0 || st.height == 0 || st.width > 16384 || st.height > 16384)
        return fail(LR_INVALID, "an image is 1 to 16384 pixels on a side");
    if (st.samples == 0 || st.max_bounces == 0 || st.max_bounces > 1024)
        return fail(LR_INVALID, "a render needs samples and 1 to 1024 bounces");
    if ((uint32_t)frame[11] != st.width || (uint32_t)frame[15] != st.height)
        return fail(LR_INVALID, "the frame's size is not the settings' image size");
    if (!loaded->constants || !loaded->lights || !loaded->materials || !loaded->emitters || loaded->material_floats < 64)
        return fail(LR_INVALID, "a renderer needs the loaded lights, constants, materials (the default one at least) and emitters");
    (void)hipSetDevice(device->ordinal);
    std::unique_ptr<lr_renderer> r(new lr_renderer());
    r->device = device;
    r->scene = scene;
    r->settings = st;
    int vi = variant_index(st.trace_width ? st.trace_width : scene->config.width,
                           st.trace_stack ? st.trace_stack : scene->config.stack,
                           st.trace_group ? st.trace_group : scene->config.group);
    if (vi < 0) return fail(LR_INVALID, "traversal widths are 2, 4, 8; stacks 24, 32, 48; groups 64, 128, 256");
    r->trace_variant = vi;
    uint32_t pixels = st.width * st.height;
    uint32_t pool = st.pool ? st.pool : 2097152u;  // integrator.lucb:96-98
    uint32_t band = pool < pixels ? pool : pixels;
    uint32_t light_paths = loaded->light_paths < band ? loaded->light_paths : band;
    r->slots = band + light_paths;
    r->bands = (pixels + band - 1) / band;
    r->rounds = st.max_bounces + 1 + loaded->through;
    int rc = 0;
    rt::Params &p = r->params;
    std::memcpy(p.frame, frame, 64);
    p.pool = r->slots;
    p.seed = st.seed;
    p.pixels = pixels;
    p.width = st.width;
    p.height = st.height;
    p.band_size = band;
    p.light_paths = light_paths;
    p.spectral = st.spectral;
    auto upload = [&](const void *src, size_t bytes) -> void * {
        void *d = r->alloc(bytes, rc);
        if (!d) return nullptr;
        if (bytes) {
            hipError_t e = hipMemcpy(d, src, bytes, hipMemcpyHostToDevice);
            if (e != hipSuccess) {
                rc = hip_fail(e, "renderer upload");
                return nullptr;
            }
        }
        return d;
    };
    p.lights = (const float4 *)upload(loaded->lights, loaded->light_floats * 4);
    p.constants = (const float4 *)upload(loaded->constants, 128 * 4);
    p.materials = (const float4 *)upload(loaded->materials, loaded->material_floats * 4);
    p.emitters = (const uint32_t *)upload(loaded->emitters, loaded->emitter_words * 4);
    if (rc) return rc;
    float *tables = (float *)r->alloc((size_t)rt::TABLE_SIZE * 4, rc);
    p.tables = tables;
    p.shading = (const uint4 *)scene->shading;
    p.uvs = (const float *)scene->uvs;
    if (!scene->shading) return fail(LR_INVALID, "rendering needs the scene's shading attributes");
    size_t slots = r->slots;
    p.state_v = (float4 *)r->alloc(rt::kStateV * slots * 16, rc);
    p.state_u = (uint32_t *)r->alloc(rt::kStateU * slots * 4, rc);
    size_t queue_words = rt::kQueueBase + rt::kQueues * slots;
    p.queues = (uint32_t *)r->alloc(queue_words * 4, rc);
    p.film = (float4 *)r->alloc((size_t)4 * pixels * 16, rc);
    p.image = (float4 *)r->alloc((size_t)pixels * 16, rc);
    p.display = (float4 *)r->alloc((size_t)pixels * 16, rc);
    p.pass = (rt::PassState *)r->alloc(sizeof(rt::PassState), rc);
    if (st.guiding) p.guide = (uint8_t *)r->alloc(rt::kGuideBytes, rc);
    if (rc) return rc;
    // The traversal's spill area for the largest grid a pass launches.
    p.scene = scene->device_scene();
    {
        const TraceVariant &tv0 = trace_variants()[vi];
        (void)tv0;
        uint32_t group = render_trace_variants()[vi].group;
        uint64_t threads = (uint64_t)r->blocks(slots, group) * group;
        uint32_t width = st.trace_width ? st.trace_width : scene->config.width;
        uint32_t stack = st.trace_stack ? st.trace_stack : scene->config.stack;
        uint32_t bound = scene->info.stack_bound[width == 2 ? 0 : width == 4 ? 1 : 2];
        uint64_t depth = bound > stack ? bound - stack : 0;
        p.scene.spill = depth ? (uint32_t *)r->alloc(depth * threads * 4, rc) : nullptr;
        if (rc) return rc;
        p.scene.nodes = (const uint8_t *)scene->wide[width == 2 ? 0 : width == 4 ? 1 : 2];
    }
    hipStream_t s = device->stream;
    launch_tables(tables, s);
    launch_fill((uint32_t *)p.film, (size_t)4 * pixels * 4, 0u, s);
    launch_fill((uint32_t *)p.image, (size_t)pixels * 4, 0u, s);
    launch_fill(p.queues, queue_words, 0u, s);
    launch_fill(p.state_u, rt::kStateU * slots, 0u, s);
    if (p.guide) launch_fill((uint32_t *)p.guide, rt::kGuideBytes / 4, 0u, s);
    rt::PassState init{};
    init.bands = r->bands;
    hipError_t e = hipMemcpyAsync(p.pass, &init, sizeof(init), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "renderer set-up");
    *out = r.release();
    return 0;
}

int lr_renderer_advance(lr_renderer *r, int *busy) {
    if (!r) return fail(LR_INVALID, "lr_renderer_advance needs a renderer");
    (void)hipSetDevice(r->device->ordinal);
    int rc = r->retire(false);
    if (rc) return rc;
    size_t depth = r->settings.profile ? 1 : 2;
    if (r->in_flight.size() >= depth) {
        if (busy) *busy = 1;
        return 0;
    }
    if (r->sample >= r->settings.samples) {
        if (busy) *busy = r->in_flight.empty() ? 0 : 1;
        return 0;
    }
    rc = r->launch_pass();
    if (busy) *busy = 1;
    return rc;
}

int lr_renderer_finish(lr_renderer *r) {
    if (!r) return fail(LR_INVALID, "lr_renderer_finish needs a renderer");
    for (;;) {
        int busy = 0;
        int rc = lr_renderer_advance(r, &busy);
        if (rc) return rc;
        if (!busy) break;
        if (r->in_flight.size() >= (r->settings.profile ? 1u : 2u) || r->sample >= r->settings.samples) {
            rc = r->retire(true);
            if (rc) return rc;
        }
    }
    return r->retire(true);
}

uint32_t lr_renderer_samples_done(lr_renderer *r) {
    if (!r) return 0;
    (void)r->retire(false);
    return r->completed;
}

int lr_renderer_image(lr_renderer *r, lr_image *out) {
    if (!r || !out) return fail(LR_INVALID, "lr_renderer_image needs a renderer and an output");
    out->rgba = r->params.image;
    out->display = r->params.display;
    out->width = r->settings.width;
    out->height = r->settings.height;
    return 0;
}

int lr_renderer_read(lr_renderer *r, float *rgba, size_t count) {
    if (!r || !rgba) return fail(LR_INVALID, "lr_renderer_read needs a renderer and an output");
    size_t pixels = (size_t)r->settings.width * r->settings.height;
    if (count != pixels * 4) return fail(LR_INVALID, "reading the image needs four floats a pixel");
    int rc = r->retire(true);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(rgba, r->params.image, pixels * 16, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "image read");
    return 0;
}

int lr_renderer_read_film(lr_renderer *r, float *planes, size_t count) {
    if (!r || !planes) return fail(LR_INVALID, "lr_renderer_read_film needs a renderer and an output");
    size_t pixels = (size_t)r->settings.width * r->settings.height;
    if (count != pixels * 16) return fail(LR_INVALID, "the film is four planes of four floats a pixel");
    int rc = r->retire(true);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(planes, r->params.film, pixels * 64, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "film read");
    return 0;
}

int lr_renderer_kernel_times(lr_renderer *r, double out[10]) {
    if (!r || !out) return fail(LR_INVALID, "kernel times need ten numbers");
    (void)r->retire(true);
    for (int k = 0; k < kKinds; k++) out[k] = r->kernel_ns[k] * 1e-6;
    return 0;
}

int lr_renderer_stats(lr_renderer *r, lr_render_stats *out) {
    if (!r || !out) return fail(LR_INVALID, "lr_renderer_stats needs a renderer and an output");
    int rc = r->retire(true);
    if (rc) return rc;
    rt::PassState ps;
    hipError_t e = hipMemcpyAsync(&ps, r->params.pass, sizeof(ps), hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "stats read");
    out->samples_done = ps.samples_done;
    out->light_samples = ps.light_samples;
    out->closest_rays = ps.closest_rays;
    out->shadow_rays = ps.shadow_rays;
    return 0;
}

void lr_renderer_destroy(lr_renderer *r) {
    if (!r) return;
    (void)hipSetDevice(r->device->ordinal);
    (void)r->retire(true);
    (void)hipStreamSynchronize(r->device->stream);
    delete r;
}

}  // extern "C"
