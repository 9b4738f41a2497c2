// The C ABI's renderer (Renderer.create / advance / finish / samples_done /
// image / read / kernel_times / destroy, src/render/integrator.lucb:343-903).
//
// A pass is the reference's record() (integrator.lucb:788-828): begin, camera,
// light emission, max_bounces + 1 (+ cut-out) rounds of closest / shade / miss
// / shadow, finish, guide fit, film convert, end. Its launches never change
// (sample and band live on the device), so the pass is recorded once into a
// hipGraph and replayed; up to two passes are in flight, as advance keeps them
// (integrator.lucb:651-655). With the learned guide, passes within its
// training samples replay a second graph that also holds the training steps.
#include <hip/hip_runtime.h>

#include <cstring>
#include <deque>
#include <memory>
#include <vector>

#include "lr_internal.h"
#include "lr_scene.h"
#include "render_guide.h"
#include "render_job.h"

using namespace lr;

namespace {

struct Pass {
    hipEvent_t done;
    uint32_t completes;  // samples every pixel has once it finishes
    std::vector<std::pair<int, hipEvent_t>> marks;
    hipEvent_t start = nullptr;
};

}  // namespace

struct lr_renderer {
    std::unique_ptr<RenderJob> job{new RenderJob()};
    lr_device *device = nullptr;
    lr_render_settings settings{};
    // host mirror of the device pass state
    uint32_t sample = 0, band = 0;
    uint32_t completed = 0;
    // The display settings changed since the display image was last made.
    bool display_stale = false;
    std::deque<Pass> in_flight;
    double kernel_ns[KKinds] = {0};
    hipGraph_t graph[2] = {nullptr, nullptr};  // without, with the training steps
    hipGraphExec_t exec[2] = {nullptr, nullptr};

    ~lr_renderer() {
        for (int k = 0; k < 2; k++) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
        }
        for (auto &ps : in_flight) {
            (void)hipEventDestroy(ps.done);
            if (ps.start) (void)hipEventDestroy(ps.start);
            for (auto &m : ps.marks) (void)hipEventDestroy(m.second);
        }
    }

    void drop_graphs() {
        for (int k = 0; k < 2; k++) {
            if (exec[k]) (void)hipGraphExecDestroy(exec[k]);
            if (graph[k]) (void)hipGraphDestroy(graph[k]);
            exec[k] = nullptr;
            graph[k] = nullptr;
        }
    }

    void record(hipStream_t s, std::vector<std::pair<int, hipEvent_t>> *marks, bool train) {
        std::vector<RenderJob *> one{job.get()};
        record_pass(one, job->params_device, s, marks, train);
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
        uint32_t carried = job->samples_in_pass(sample);  // as begin_pass_kernel counts
        bool train = job->params.learned && sample < job->params.train_samples;
        if (settings.profile) {
            (void)hipEventCreate(&ps.start);
            (void)hipEventRecord(ps.start, s);
            record(s, &ps.marks, train);
        } else if (settings.use_graph) {
            int k = train ? 1 : 0;
            if (!exec[k]) {
                hipError_t e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
                if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
                record(s, nullptr, train);
                e = hipStreamEndCapture(s, &graph[k]);
                if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
                e = hipGraphInstantiate(&exec[k], graph[k], nullptr, nullptr, 0);
                if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
            }
            hipError_t e = hipGraphLaunch(exec[k], s);
            if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
        } else {
            record(s, nullptr, train);
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "pass launch");
        band += 1;
        if (band == job->bands) {
            band = 0;
            sample += carried;
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
    (void)hipSetDevice(device->ordinal);
    std::unique_ptr<lr_renderer> r(new lr_renderer());
    int rc = r->job->create(device, scene, loaded, settings, frame, nullptr);
    if (rc) return rc;
    r->device = device;
    r->settings = r->job->settings;
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
        if (!r->display_stale) {
            if (busy) *busy = r->in_flight.empty() ? 0 : 1;
            return 0;
        }
        // Only the display changed: the film converted again, after the
        // passes in flight (integrator.lucb:656-669).
        r->display_stale = false;
        uint32_t pixels = r->settings.width * r->settings.height;
        launch_render_stage(Stage::Convert, r->settings.spectral != 0, r->job->params_device, 1, 1,
                            (pixels + 255u) / 256u, r->device->stream);
        hipError_t e = hipGetLastError();
        if (busy) *busy = 1;
        return e == hipSuccess ? 0 : hip_fail(e, "display conversion");
    }
    r->display_stale = false;  // the pass's last band converts with the new settings
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
    out->rgba = r->job->params.image;
    out->display = r->job->params.display;
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
    hipError_t e = hipMemcpyAsync(rgba, r->job->params.image, pixels * 16, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "image read");
    return 0;
}

int lr_renderer_read_display(lr_renderer *r, float *rgba, size_t count) {
    if (!r || !rgba) return fail(LR_INVALID, "lr_renderer_read_display needs a renderer and an output");
    size_t pixels = (size_t)r->settings.width * r->settings.height;
    if (count != pixels * 4) return fail(LR_INVALID, "reading the display image needs four floats a pixel");
    int rc = r->retire(true);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(rgba, r->job->params.display, pixels * 16, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "display read");
    return 0;
}

int lr_renderer_set_display(lr_renderer *r, float exposure, int view) {
    if (!r) return fail(LR_INVALID, "lr_renderer_set_display needs a renderer");
    if (view != 0 && view != 1) return fail(LR_INVALID, "the view is 0 (standard) or 1 (neutral tone mapping)");
    (void)hipSetDevice(r->device->ordinal);
    const float row[4] = {exposure, (float)view, 0.0f, 0.0f};
    hipStream_t s = r->device->stream;
    hipError_t e = hipMemcpyAsync((char *)r->job->params.constants + rt::K_DISPLAY * 16, row, sizeof(row),
                                  hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) return hip_fail(e, "display settings");
    r->job->set_display_row(row);
    r->display_stale = true;
    return 0;
}

int lr_renderer_read_film(lr_renderer *r, float *planes, size_t count) {
    if (!r || !planes) return fail(LR_INVALID, "lr_renderer_read_film needs a renderer and an output");
    size_t pixels = (size_t)r->settings.width * r->settings.height;
    if (count != pixels * 16) return fail(LR_INVALID, "the film is four planes of four floats a pixel");
    int rc = r->retire(true);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(planes, r->job->params.film, pixels * 64, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "film read");
    return 0;
}

int lr_renderer_read_guide(lr_renderer *r, void *out, size_t bytes) {
    if (!r || !out) return fail(LR_INVALID, "lr_renderer_read_guide needs a renderer and an output");
    if (!r->job->params.guide) return fail(LR_INVALID, "the render does not guide with the hash grid");
    if (bytes != rt::kGuideBytes) return fail(LR_INVALID, "the guide grid's size is lr_renderer_read_guide's bytes");
    int rc = r->retire(true);
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(out, r->job->params.guide, bytes, hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "guide read");
    return 0;
}

int lr_renderer_kernel_times(lr_renderer *r, double out[10]) {
    if (!r || !out) return fail(LR_INVALID, "kernel times need ten numbers");
    (void)r->retire(true);
    for (int k = 0; k < KKinds; k++) out[k] = r->kernel_ns[k] * 1e-6;
    return 0;
}

int lr_renderer_stats(lr_renderer *r, lr_render_stats *out) {
    if (!r || !out) return fail(LR_INVALID, "lr_renderer_stats needs a renderer and an output");
    int rc = r->retire(true);
    if (rc) return rc;
    rt::PassState ps;
    hipError_t e = hipMemcpyAsync(&ps, r->job->params.pass, sizeof(ps), hipMemcpyDeviceToHost, r->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(r->device->stream);
    if (e != hipSuccess) return hip_fail(e, "stats read");
    out->samples_done = ps.samples_done;
    out->light_samples = ps.light_samples;
    out->closest_rays = ps.closest_rays;
    out->shadow_rays = ps.shadow_rays;
    return 0;
}

// A side of a view's buffers that grew: a quarter more than wanted, rounded
// up to 64, never below what is held, at most 16384 (integrator.lucb:986-989).
static uint32_t roomy_side(uint32_t wanted, uint32_t held) {
    uint32_t side = (wanted + wanted / 4u + 63u) / 64u * 64u;
    if (side < held) side = held;
    return side <= 16384u ? side : 16384u;
}

int lr_renderer_reset(lr_renderer *r, lr_gpu_scene *scene, const lr_loaded *loaded, const lr_render_settings *settings,
                      const float frame[16]) {
    if (!r || !loaded || !settings || !frame) return fail(LR_INVALID, "lr_renderer_reset needs its arguments");
    if ((frame[3] > 0.5f) != (r->job->params.frame[3] > 0.5f)) return fail(LR_INVALID, "a restart keeps the view's projection");
    (void)hipSetDevice(r->device->ordinal);
    // Passes in flight finish and count for nothing.
    int rc = r->retire(true);
    if (rc) return rc;
    // Same scene and shape: the buffers are kept (the image and display show
    // the last frame until the first new pass).
    if (!scene || scene == r->job->scene) {
        rc = r->job->restart_in_place(loaded, settings, frame);
        if (rc != RenderJob::kNotInPlace) {
            if (rc) return rc;
            r->drop_graphs();
            r->settings = r->job->settings;
            r->sample = r->band = r->completed = 0;
            r->display_stale = false;
            for (double &k : r->kernel_ns) k = 0.0;
            return 0;
        }
    }
    // The new state is made before the old goes, so a failure leaves the
    // renderer as it was.
    std::unique_ptr<RenderJob> job(new RenderJob());
    rc = RenderJob::kNotInPlace;
    const RenderJob &held = *r->job;
    if ((!scene || scene == held.scene) && (uint64_t)settings->width * settings->height > held.cap_pixels &&
        settings->guiding != 2) {
        // A view that outgrows the buffers: made again with a quarter more
        // room each way (and light paths in proportion), then the view taken
        // in place, so a view growing a few pixels at a time (a window
        // resized) makes its buffers again seldom (integrator.lucb:474-478).
        lr_render_settings room = *settings;
        room.width = roomy_side(settings->width, held.cap_width);
        room.height = roomy_side(settings->height, held.cap_height);
        float room_frame[16];
        std::memcpy(room_frame, frame, sizeof(room_frame));
        room_frame[11] = (float)room.width;
        room_frame[15] = (float)room.height;
        lr_loaded room_loaded = *loaded;
        uint64_t pixels = (uint64_t)settings->width * settings->height, room_pixels = (uint64_t)room.width * room.height;
        room_loaded.light_paths = (uint32_t)(((uint64_t)loaded->light_paths * room_pixels + pixels - 1) / pixels);
        rc = job->create(r->device, held.scene, &room_loaded, &room, room_frame, nullptr);
        if (rc == 0) rc = job->restart_in_place(loaded, settings, frame);
        if (rc) {
            // No room to spare (or none needed): the exact size, as for a new
            // renderer.
            job.reset(new RenderJob());
            rc = RenderJob::kNotInPlace;
        }
    }
    if (rc == RenderJob::kNotInPlace) {
        rc = job->create(r->device, scene ? scene : r->job->scene, loaded, settings, frame, nullptr);
        if (rc) return rc;
    }
    if (job->scene == held.scene) {
        rc = job->adopt_display(held);
        if (rc) return rc;
    }
    r->drop_graphs();
    r->job.swap(job);
    r->settings = r->job->settings;
    r->sample = r->band = r->completed = 0;
    r->display_stale = false;
    for (double &k : r->kernel_ns) k = 0.0;
    return 0;
}

int lr_renderer_converged_pixels(lr_renderer *r, size_t *out) {
    if (!r || !out) return fail(LR_INVALID, "lr_renderer_converged_pixels needs a renderer and an output");
    int rc = r->retire(true);
    if (rc) return rc;
    hipStream_t s = r->device->stream;
    uint32_t *counter = nullptr;
    hipError_t e = hipMalloc(&counter, 64);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(converged count)");
    launch_fill(counter, 1, 0u, s);
    launch_converged(r->job->params_device, r->job->params.pixels, counter, s);
    uint32_t count = 0;
    e = hipMemcpyAsync(&count, counter, 4, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    (void)hipFree(counter);
    if (e != hipSuccess) return hip_fail(e, "converged count");
    *out = count;
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
