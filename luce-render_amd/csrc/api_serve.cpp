// The serving form: many renders (requests) of different scenes, resolutions
// and sample counts resident on one GPU, advanced together. Each step runs
// one pass of every active request in shared launches (each request's items
// laid end to end), so small requests fill the chip together instead of each
// leaving most of its 256 compute units idle; a small request's pass carries
// several samples, so it does not wait a step per sample behind large ones. Requests are admitted in
// arrival order as their device memory fits the budget; a finished request
// keeps its image until it is read and released. No reference counterpart
// (the reference renders one view at a time).
#include <hip/hip_runtime.h>

#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <vector>

#include "lr_internal.h"
#include "lr_scene.h"
#include "render_bsdf.h"
#include "render_job.h"

using namespace lr;

namespace {

struct Request {
    uint64_t id = 0;
    int state = 0;  // 0 pending, 1 active, 2 done
    lr_gpu_scene *scene = nullptr;
    lr_render_settings settings{};
    float frame[16];
    // the loaded arrays, copied at submission (the caller's are borrowed)
    std::vector<float> lights, constants, materials, textures;
    std::vector<uint32_t> emitters, programs, texture_table;
    lr_loaded loaded{};
    size_t footprint = 0;
    std::unique_ptr<RenderJob> job;
    uint32_t samples = 0;
    rt::PassState final_state{};  // its counters, kept once its state is freed
};

}  // namespace

struct lr_server {
    lr_device *device = nullptr;
    lr_server_settings settings{};
    uint64_t next_id = 1;
    size_t used = 0;
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    std::deque<uint64_t> pending;
    std::vector<Request *> active;
    rt::Params *params = nullptr;  // device array of the active requests' parameters
    size_t params_capacity = 0;
    bool dirty = true;
    float *tables = nullptr;

    ~lr_server() {
        if (params) (void)hipFree(params);
        if (tables) (void)hipFree(tables);
    }

    int admit() {
        while (!pending.empty()) {
            Request *r = requests[pending.front()].get();
            if (settings.max_active && active.size() >= settings.max_active) break;
            if (active.size() >= rt::kMaxRows) break;  // rows a shared launch can carry
            if (used + r->footprint > settings.budget_bytes) break;
            r->job.reset(new RenderJob());
            int rc = r->job->create(device, r->scene, &r->loaded, &r->settings, r->frame, tables);
            if (rc) return rc;
            used += r->footprint;
            r->state = 1;
            active.push_back(r);
            pending.pop_front();
            dirty = true;
        }
        return 0;
    }

    int upload_params() {
        if (!dirty) return 0;
        if (active.size() > params_capacity) {
            if (params) (void)hipFree(params);
            params = nullptr;
            params_capacity = 0;
            size_t cap = std::max<size_t>(16, active.size() * 2);
            hipError_t e = hipMalloc(&params, cap * sizeof(rt::Params));
            if (e != hipSuccess) return hip_fail(e, "hipMalloc(server params)");
            params_capacity = cap;
        }
        std::vector<rt::Params> host;
        for (Request *r : active) host.push_back(r->job->params);
        if (!host.empty()) {
            hipError_t e = hipMemcpyAsync(params, host.data(), host.size() * sizeof(rt::Params), hipMemcpyHostToDevice,
                                          device->stream);
            if (e != hipSuccess) return hip_fail(e, "server params upload");
        }
        dirty = false;
        return 0;
    }
};

extern "C" {

int lr_server_create(lr_device *device, const lr_server_settings *settings, lr_server **out) {
    if (!device || !settings || !out) return fail(LR_INVALID, "lr_server_create needs its arguments");
    *out = nullptr;
    if (variant_index(settings->trace_width, settings->trace_stack, settings->trace_group) < 0)
        return fail(LR_INVALID, "traversal layouts are 2, 4, 8, 20, 24; stacks 16, 24, 32, 48; groups 64, 128, 256");
    (void)hipSetDevice(device->ordinal);
    std::unique_ptr<lr_server> s(new lr_server());
    s->device = device;
    s->settings = *settings;
    hipError_t e = hipMalloc(&s->tables, (size_t)rt::TABLE_SIZE * 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(tables)");
    launch_tables(s->tables, device->stream);
    *out = s.release();
    return 0;
}

int lr_server_submit(lr_server *s, lr_gpu_scene *scene, const lr_loaded *loaded, const lr_render_settings *settings,
                     const float frame[16], uint64_t *id) {
    if (!s || !scene || !loaded || !settings || !frame || !id) return fail(LR_INVALID, "lr_server_submit needs its arguments");
    if (settings->width == 0 || settings->height == 0 || settings->width > 16384 || settings->height > 16384)
        return fail(LR_INVALID, "an image is 1 to 16384 pixels on a side");
    if (settings->samples == 0 || settings->max_bounces == 0 || settings->max_bounces > 1024)
        return fail(LR_INVALID, "a render needs samples and 1 to 1024 bounces");
    if ((settings->spectral != 0) != (s->settings.spectral != 0))
        return fail(LR_INVALID, "a server's requests share its spectral mode");
    if (settings->guiding > 1) return fail(LR_INVALID, "served requests guide with the hash grid at most");
    std::unique_ptr<Request> r(new Request());
    r->scene = scene;
    r->settings = *settings;
    r->settings.trace_width = s->settings.trace_width;
    r->settings.trace_group = s->settings.trace_group;
    r->settings.trace_stack = s->settings.trace_stack;
    r->settings.pool = settings->width * settings->height;  // one band: a step runs one pass, of one or more samples
    r->settings.profile = 0;
    r->settings.use_graph = 0;
    std::memcpy(r->frame, frame, 64);
    r->lights.assign(loaded->lights, loaded->lights + loaded->light_floats);
    r->constants.assign(loaded->constants, loaded->constants + 128);
    r->materials.assign(loaded->materials, loaded->materials + loaded->material_floats);
    r->emitters.assign(loaded->emitters, loaded->emitters + loaded->emitter_words);
    // The programs and textures too: the request uploads them at admission,
    // a later step, after the caller's arrays may be gone.
    if (loaded->programs) r->programs.assign(loaded->programs, loaded->programs + loaded->program_words);
    if (loaded->textures) r->textures.assign(loaded->textures, loaded->textures + loaded->texture_floats);
    if (loaded->texture_table)
        r->texture_table.assign(loaded->texture_table, loaded->texture_table + loaded->texture_count * 4);
    r->loaded = *loaded;
    r->loaded.lights = r->lights.data();
    r->loaded.constants = r->constants.data();
    r->loaded.materials = r->materials.data();
    r->loaded.emitters = r->emitters.data();
    r->loaded.programs = loaded->programs ? r->programs.data() : nullptr;
    r->loaded.textures = loaded->textures ? r->textures.data() : nullptr;
    r->loaded.texture_table = loaded->texture_table ? r->texture_table.data() : nullptr;
    r->footprint = RenderJob::footprint(scene, &r->loaded, &r->settings);
    if (r->footprint > s->settings.budget_bytes)
        return fail(LR_LIMIT, "the request alone needs more device memory than the server's budget");
    r->id = s->next_id++;
    *id = r->id;
    s->pending.push_back(r->id);
    s->requests[r->id] = std::move(r);
    return 0;
}

int lr_server_step(lr_server *s, uint32_t *active_out, uint32_t *pending_out) {
    if (!s) return fail(LR_INVALID, "lr_server_step needs a server");
    (void)hipSetDevice(s->device->ordinal);
    int rc = s->admit();
    if (rc) return rc;
    if (!s->active.empty()) {
        rc = s->upload_params();
        if (rc) return rc;
        std::vector<RenderJob *> jobs;
        std::vector<uint32_t> carried;
        for (Request *r : s->active) {
            jobs.push_back(r->job.get());
            carried.push_back(r->job->samples_in_pass(r->samples));
        }
        record_pass(jobs, s->params, s->device->stream, nullptr);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(s->device->stream);
        if (e != hipSuccess) return hip_fail(e, "server step");
        std::vector<Request *> still;
        for (size_t k = 0; k < s->active.size(); k++) {
            Request *r = s->active[k];
            r->samples += carried[k];
            if (r->samples >= r->settings.samples) {
                // Finished: its image stays until read and released; the
                // rest of its memory goes back to the budget.
                r->state = 2;
                (void)hipMemcpy(&r->final_state, r->job->params.pass, sizeof(rt::PassState), hipMemcpyDeviceToHost);
                s->used -= r->footprint;
                r->footprint = r->job->keep_image_only();
                s->used += r->footprint;
                s->dirty = true;
            } else {
                still.push_back(r);
            }
        }
        s->active.swap(still);
    }
    if (active_out) *active_out = (uint32_t)s->active.size();
    if (pending_out) *pending_out = (uint32_t)s->pending.size();
    return 0;
}

int lr_server_status(lr_server *s, uint64_t id, lr_request_status *out) {
    if (!s || !out) return fail(LR_INVALID, "lr_server_status needs a server and an output");
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(LR_INVALID, "no such request");
    Request *r = it->second.get();
    out->state = (uint32_t)r->state;
    out->samples_done = r->samples;
    out->samples = r->settings.samples;
    out->bytes = r->footprint;
    return 0;
}

int lr_server_read(lr_server *s, uint64_t id, float *rgba, size_t count) {
    if (!s || !rgba) return fail(LR_INVALID, "lr_server_read needs a server and an output");
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(LR_INVALID, "no such request");
    Request *r = it->second.get();
    if (r->state != 2) return fail(LR_INVALID, "the request has not finished");
    size_t pixels = (size_t)r->settings.width * r->settings.height;
    if (count != pixels * 4) return fail(LR_INVALID, "reading the image needs four floats a pixel");
    hipError_t e = hipMemcpyAsync(rgba, r->job->params.image, pixels * 16, hipMemcpyDeviceToHost, s->device->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->device->stream);
    if (e != hipSuccess) return hip_fail(e, "request read");
    return 0;
}

int lr_server_stats(lr_server *s, uint64_t id, lr_render_stats *out) {
    if (!s || !out) return fail(LR_INVALID, "lr_server_stats needs a server and an output");
    auto it = s->requests.find(id);
    if (it == s->requests.end() || !it->second->job) return fail(LR_INVALID, "no such admitted request");
    rt::PassState ps = it->second->final_state;
    if (it->second->state != 2) {
        hipError_t e = hipMemcpyAsync(&ps, it->second->job->params.pass, sizeof(ps), hipMemcpyDeviceToHost, s->device->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(s->device->stream);
        if (e != hipSuccess) return hip_fail(e, "request stats");
    }
    out->samples_done = ps.samples_done;
    out->light_samples = ps.light_samples;
    out->closest_rays = ps.closest_rays;
    out->shadow_rays = ps.shadow_rays;
    return 0;
}

int lr_server_release(lr_server *s, uint64_t id) {
    if (!s) return fail(LR_INVALID, "lr_server_release needs a server");
    auto it = s->requests.find(id);
    if (it == s->requests.end()) return fail(LR_INVALID, "no such request");
    Request *r = it->second.get();
    (void)hipStreamSynchronize(s->device->stream);
    if (r->state == 0) {
        for (auto p = s->pending.begin(); p != s->pending.end(); ++p)
            if (*p == id) {
                s->pending.erase(p);
                break;
            }
    } else {
        if (r->state == 1) {
            for (auto a = s->active.begin(); a != s->active.end(); ++a)
                if (*a == r) {
                    s->active.erase(a);
                    break;
                }
            s->dirty = true;
        }
        s->used -= r->footprint;
    }
    s->requests.erase(it);
    return 0;
}

void lr_server_destroy(lr_server *s) {
    if (!s) return;
    (void)hipSetDevice(s->device->ordinal);
    (void)hipStreamSynchronize(s->device->stream);
    delete s;
}

}  // extern "C"
