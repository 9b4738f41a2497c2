// The C ABI's device, scene and tracing entry points (include/luce_render_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>

#include "lr_internal.h"
#include "lr_scene.h"
#include "trace_kernels.h"

namespace lr {

namespace {
thread_local std::string g_error;
}

int fail(int code, const std::string &message) {
    g_error = message;
    return code;
}

int hip_fail(int status, const char *what) {
    return fail(LR_DEVICE_ERROR, std::string(what) + ": " + hipGetErrorString((hipError_t)status));
}

}  // namespace lr

using namespace lr;

int lr_gpu_scene::ensure_spill(uint64_t threads, uint32_t stack) {
    uint32_t bound = info.stack_bound[bound_index(config.width)];
    uint64_t depth = bound > stack ? bound - stack : 0;
    uint64_t need = depth * threads;
    if (need <= spill_entries) return 0;
    if (spill) (void)hipFree(spill);
    spill = nullptr;
    spill_entries = 0;
    hipError_t e = hipMalloc(&spill, need * 4);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(spill)");
    spill_entries = need;
    return 0;
}

int lr_gpu_scene::config_width_index() const { return layout_index(config.width); }

lr::DeviceScene lr_gpu_scene::device_scene() const {
    DeviceScene sc{};
    int wi = config_width_index();
    sc.nodes = static_cast<const uint8_t *>(wide[wi]);
    sc.binary = static_cast<const uint8_t *>(wide[0]);
    sc.ranges = static_cast<const uint2 *>(ranges);
    sc.tris = static_cast<const float *>(tris);
    sc.faces = static_cast<const uint32_t *>(faces);
    std::memcpy(sc.root_lo, root_lo, 12);
    std::memcpy(sc.root_hi, root_hi, 12);
    sc.empty = info.triangles == 0 ? 1u : 0u;
    sc.spill = static_cast<uint32_t *>(spill);
    sc.spill_stride = 0;
    sc.spill_depth = 0;
    sc.work = work;
    return sc;
}

const lr::TraceVariant &lr_gpu_scene::variant() const {
    return trace_variants()[variant_index(config.width, config.stack, config.group)];
}

lr_gpu_scene::~lr_gpu_scene() {
    for (void *p : {wide[0], wide[1], wide[2], wide[3], wide[4], tris, faces, shading, uvs, spill, (void *)work, ranges, (void *)binary,
                    (void *)indices, (void *)parent, (void *)arrive, (void *)slots[0], (void *)slots[1], (void *)slots[2],
                    splat_records, splat_words, splat_nodes})
        if (p) (void)hipFree(p);
}


namespace {

template <class T> int upload(void **dst, const T *src, size_t count, const char *what) {
    size_t bytes = std::max<size_t>(count * sizeof(T), 64);  // an empty array still gets a word (device.lucb:125-128)
    hipError_t e = hipMalloc(dst, bytes);
    if (e != hipSuccess) return hip_fail(e, what);
    if (count) {
        e = hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice);
        if (e != hipSuccess) return hip_fail(e, what);
    } else {
        e = hipMemset(*dst, 0, bytes);
        if (e != hipSuccess) return hip_fail(e, what);
    }
    return 0;
}

}  // namespace

extern "C" {

size_t lr_last_error(char *buffer, size_t capacity) {
    const std::string &m = g_error;
    if (buffer && capacity) {
        size_t n = std::min(capacity - 1, m.size());
        std::memcpy(buffer, m.data(), n);
        buffer[n] = 0;
    }
    return m.size();
}

int lr_device_open(int ordinal, lr_device **out) {
    if (!out) return fail(LR_INVALID, "lr_device_open needs an output handle");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return fail(LR_UNAVAILABLE, "no GPU");
    if (ordinal < 0 || ordinal >= count) return fail(LR_INVALID, "no GPU of that ordinal");
    hipError_t e = hipSetDevice(ordinal);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    auto *d = new lr_device();
    d->ordinal = ordinal;
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete d;
        return hip_fail(e, "hipStreamCreate");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, ordinal) == hipSuccess) d->compute_units = prop.multiProcessorCount;
    scratch_attach(d);
    *out = d;
    return 0;
}

void lr_device_close(lr_device *device) {
    if (!device) return;
    (void)hipSetDevice(device->ordinal);
    (void)hipStreamSynchronize(device->stream);
    (void)hipStreamDestroy(device->stream);
    delete device;
}

int lr_bvh_build(const float *positions, size_t points, const uint32_t *indices, size_t triangles, int threads,
                 lr_bvh_node *nodes, size_t *node_count, uint32_t *order) {
    if (!nodes || !node_count || (triangles && (!positions || !indices || !order)))
        return fail(LR_INVALID, "lr_bvh_build needs positions, indices, nodes and order");
    if (points > 4294967295ull) return fail(LR_LIMIT, "a scene holds at most 2^32 points");
    if (triangles >= (1ull << 27)) return fail(LR_LIMIT, "a scene holds at most 2^27 triangles");
    for (size_t i = 0; i < triangles * 3; i++)
        if (indices[i] >= points) return fail(LR_INVALID, "a triangle names a point the scene lacks");
    *node_count = build_binary_bvh(positions, indices, triangles, threads, nodes, order);
    return 0;
}

int lr_bvh_build_boxes(const float *boxes, size_t count, int threads, lr_bvh_node *nodes, size_t *node_count,
                       uint32_t *order) {
    if (!nodes || !node_count || (count && (!boxes || !order)))
        return fail(LR_INVALID, "lr_bvh_build_boxes needs boxes, nodes and order");
    if (count >= (1ull << 27)) return fail(LR_LIMIT, "a tree holds at most 2^27 boxes");
    for (size_t i = 0; i < count * 6; i++)
        if (!std::isfinite(boxes[i])) return fail(LR_INVALID, "a box has a coordinate that is not finite");
    *node_count = build_box_bvh(boxes, count, threads, nodes, order);
    return 0;
}

}  // extern "C"

namespace {

template <class T> struct Scratch {
    T *p = nullptr;
    ~Scratch() {
        if (p) (void)hipFree(p);
    }
    int alloc(size_t count, const char *what) {
        hipError_t e = hipMalloc((void **)&p, std::max<size_t>(count, 1) * sizeof(T));
        return e == hipSuccess ? 0 : hip_fail(e, what);
    }
    int put(const T *src, size_t count, hipStream_t s, const char *what) {
        int rc = alloc(count, what);
        if (rc || !count) return rc;
        hipError_t e = hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s);
        return e == hipSuccess ? 0 : hip_fail(e, what);
    }
};

// The scene arrays' common checks (device.lucb:41-89 rejects what the
// kernels could not index).
int check_arrays(const lr_scene_arrays *a, bool need_nodes) {
    if (a->triangles && (!a->positions || !a->indices || !a->faces))
        return fail(LR_INVALID, "a scene with triangles needs positions, indices and faces");
    if (need_nodes && (!a->nodes || a->node_count == 0)) return fail(LR_INVALID, "a scene needs its BVH (one node at least)");
    if (a->triangles >= (1ull << 27)) return fail(LR_LIMIT, "a scene holds at most 2^27 triangles");
    if (a->points > 4294967295ull) return fail(LR_LIMIT, "a scene holds at most 2^32 points");
    for (size_t i = 0; i < a->triangles * 3; i++)
        if (a->indices[i] >= a->points) return fail(LR_INVALID, "a triangle names a point the scene lacks");
    return 0;
}

// The splats as given (device.lucb:54-62): records, harmonics and their tree,
// checked so that every index the kernels follow stays inside the arrays.
int upload_splats(lr_gpu_scene *s, const lr_scene_arrays *a) {
    const size_t n = a->splats;
    if (n == 0) return 0;
    if (n > (1ull << 26)) return fail(LR_LIMIT, "a render holds at most 2^26 splats");
    if (!a->splat_records || !a->splat_nodes || a->splat_node_count == 0 || a->splat_node_count > 2 * n)
        return fail(LR_INVALID, "splats need their records and their BVH (at most 2 nodes a splat)");
    for (size_t i = 0; i < n; i++) {
        const uint32_t *r = a->splat_records + i * 16;
        uint32_t items = r[15] & 31u;
        if (items > 15u) return fail(LR_INVALID, "a splat has more than 15 SH items");
        if (items >= 3u) {
            uint64_t stride = 2u + (items * 3u + 1u) / 2u;
            if (!a->splat_harmonics || (uint64_t)r[11] + stride > a->splat_words)
                return fail(LR_INVALID, "a splat's harmonics lie outside the harmonics words");
        }
    }
    for (size_t i = 0; i < a->splat_node_count; i++) {
        const lr_bvh_node &nd = a->splat_nodes[i];
        if (nd.count == 0) {
            if ((size_t)nd.first + 1 >= a->splat_node_count || nd.first == 0)
                return fail(LR_INVALID, "a splat BVH node names a child past the array");
        } else if ((size_t)nd.first + nd.count > n || nd.count > 16) {
            return fail(LR_INVALID, "a splat BVH leaf names splats past the array or more than 16");
        }
    }
    int rc;
    if ((rc = upload(&s->splat_records, a->splat_records, n * 16, "upload splat records")) ||
        (rc = upload(&s->splat_nodes, a->splat_nodes, a->splat_node_count, "upload splat BVH")))
        return rc;
    if (a->splat_harmonics && a->splat_words &&
        (rc = upload(&s->splat_words, a->splat_harmonics, a->splat_words, "upload splat harmonics")))
        return rc;
    s->splat_count = n;
    return 0;
}

// The traversal layouts and triangles of a scene whose leaf-order indices and
// binary tree are in device memory: triangles packed, widths 2, 4, 8
// collapsed, all on the device.
int finish_scene(lr_gpu_scene *s, const float *positions, uint32_t *indices, lr_bvh_node *nodes, size_t node_count) {
    lr_device *device = s->device;
    hipStream_t st = device->stream;
    hipError_t e;
    const size_t n = s->info.triangles;
    // The tree and leaf-order indices stay for refit (the scene frees them).
    s->binary = nodes;
    s->indices = indices;
    if ((e = hipMalloc(&s->tris, std::max<size_t>(n, 1) * 36)) != hipSuccess) return hip_fail(e, "hipMalloc(triangles)");
    pack_triangles_device(device, positions, indices, n, static_cast<float *>(s->tris));
    lr_bvh_node root;
    if ((e = hipMemcpyAsync(&root, nodes, sizeof(root), hipMemcpyDeviceToHost, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(root)");
    std::memcpy(s->root_lo, root.lo, 12);
    std::memcpy(s->root_hi, root.hi, 12);
    const bool empty = n == 0 || root.lo[0] > root.hi[0];
    const int widths[3] = {2, 4, 8};
    for (int wi = 0; wi < 3; wi++) {
        DeviceLayout layout;
        int rc = collapse_wide_device(device, nodes, node_count, empty, widths[wi], layout);
        if (rc) return rc;
        s->wide[wi] = layout.data;
        if (wi == 0) s->ranges = layout.ranges;
        s->info.wide_nodes[wi] = layout.nodes;
        s->info.wide_bytes[wi] = layout.bytes;
        s->info.stack_bound[wi] = layout.stack_bound;
    }
    int qrc;
    if ((qrc = quantize_wide_device(device, s->wide[1], s->info.wide_nodes[1], 4, &s->wide[3])) ||
        (qrc = quantize_wide_device(device, s->wide[2], s->info.wide_nodes[2], 8, &s->wide[4])))
        return qrc;
    s->info.triangle_bytes = n * 36;
    if ((e = hipMalloc((void **)&s->work, 64)) != hipSuccess || (e = hipMemsetAsync(s->work, 0, 64, st)) != hipSuccess ||
        (e = hipStreamSynchronize(st)) != hipSuccess)
        return hip_fail(e, "ray counters");
    // Default traversal: the quantized width-4 layout, 64-wide groups,
    // 16-entry LDS stack, batches refilled at 32 idle lanes (the fastest on
    // the 30M-triangle scene).
    s->config = lr_trace_config{20, 64, 16, 32};
    return 0;
}

}  // namespace

extern "C" {

int lr_scene_upload(lr_device *device, const lr_scene_arrays *a, lr_gpu_scene **out) {
    if (!device || !a || !out) return fail(LR_INVALID, "lr_scene_upload needs a device, arrays and a handle");
    *out = nullptr;
    int rc;
    if ((rc = check_arrays(a, true))) return rc;
    if (a->triangles && a->node_count > 2 * a->triangles) return fail(LR_INVALID, "more BVH nodes than 2 per triangle");
    // The nodes must describe a tree over exactly these triangles.
    for (size_t i = 0; i < a->node_count; i++) {
        const lr_bvh_node &n = a->nodes[i];
        if (a->triangles == 0) break;
        if (n.count == 0) {
            if ((size_t)n.first + 1 >= a->node_count || n.first == 0) return fail(LR_INVALID, "a BVH node names a child past the array");
        } else if ((size_t)n.first + n.count > a->triangles || n.count > 16) {
            return fail(LR_INVALID, "a BVH leaf names triangles past the array or more than 16");
        }
    }
    (void)hipSetDevice(device->ordinal);
    hipStream_t st = device->stream;
    std::unique_ptr<lr_gpu_scene> s(new lr_gpu_scene());
    s->device = device;
    s->info.triangles = a->triangles;
    s->info.binary_nodes = a->node_count;
    Scratch<float> pos;
    Scratch<uint32_t> idx;
    Scratch<lr_bvh_node> nodes;
    if ((rc = pos.put(a->positions, a->triangles ? a->points * 3 : 0, st, "upload points")) ||
        (rc = idx.put(a->indices, a->triangles * 3, st, "upload indices")) ||
        (rc = nodes.put(a->nodes, a->node_count, st, "upload BVH")))
        return rc;
    if ((rc = upload(&s->faces, a->faces, a->triangles, "upload faces"))) return rc;
    if (a->shading && (rc = upload(&s->shading, a->shading, a->triangles * 4, "upload shading"))) return rc;
    if (a->uvs && (rc = upload(&s->uvs, a->uvs, a->triangles * 6, "upload uvs"))) return rc;
    if ((rc = upload_splats(s.get(), a))) return rc;
    s->points = a->triangles ? a->points : 0;
    rc = finish_scene(s.get(), pos.p, idx.p, nodes.p, a->node_count);
    idx.p = nullptr;  // owned by the scene from here
    nodes.p = nullptr;
    if (rc) return rc;
    *out = s.release();
    return 0;
}

int lr_scene_build(lr_device *device, const lr_scene_arrays *a, lr_gpu_scene **out, lr_bvh_node *nodes_out,
                   size_t *node_count, uint32_t *order_out) {
    if (!device || !a || !out) return fail(LR_INVALID, "lr_scene_build needs a device, arrays and a handle");
    *out = nullptr;
    int rc;
    if ((rc = check_arrays(a, false))) return rc;
    (void)hipSetDevice(device->ordinal);
    hipStream_t st = device->stream;
    hipError_t e;
    const size_t n = a->triangles;
    std::unique_ptr<lr_gpu_scene> s(new lr_gpu_scene());
    s->device = device;
    s->info.triangles = n;
    Scratch<float> pos;
    Scratch<uint32_t> idx, idx_leaf, order;
    Scratch<lr_bvh_node> nodes;
    if ((rc = pos.put(a->positions, n ? a->points * 3 : 0, st, "upload points")) ||
        (rc = idx.put(a->indices, n * 3, st, "upload indices")) || (rc = idx_leaf.alloc(n * 3, "hipMalloc(indices)")) ||
        (rc = order.alloc(n, "hipMalloc(order)")) || (rc = nodes.alloc(std::max<size_t>(1, 2 * n), "hipMalloc(BVH)")))
        return rc;
    size_t count = 0;
    if ((rc = build_bvh_device(device, pos.p, idx.p, n, nodes.p, order.p, &count))) return rc;
    s->info.binary_nodes = count;
    // Per-triangle arrays into leaf order (bvh.lucb:198-217), on the device.
    auto reorder = [&](const void *host, uint32_t words, void **dst, const char *what) -> int {
        Scratch<uint32_t> src;
        int r = src.put(static_cast<const uint32_t *>(host), n * words, st, what);
        if (r) return r;
        hipError_t e2 = hipMalloc(dst, std::max<size_t>(n * words * 4, 64));
        if (e2 != hipSuccess) return hip_fail(e2, what);
        gather_rows_device(device, src.p, order.p, n, words, static_cast<uint32_t *>(*dst));
        e2 = hipStreamSynchronize(st);  // the source scratch goes out of scope
        return e2 == hipSuccess ? 0 : hip_fail(e2, what);
    };
    gather_rows_device(device, idx.p, order.p, n, 3, idx_leaf.p);
    if ((rc = reorder(a->faces, 1, &s->faces, "upload faces"))) return rc;
    if (a->shading && (rc = reorder(a->shading, 4, &s->shading, "upload shading"))) return rc;
    if (a->uvs && (rc = reorder(a->uvs, 6, &s->uvs, "upload uvs"))) return rc;
    if ((rc = upload_splats(s.get(), a))) return rc;
    s->points = n ? a->points : 0;
    rc = finish_scene(s.get(), pos.p, idx_leaf.p, nodes.p, count);
    lr_bvh_node *tree = nodes.p;
    idx_leaf.p = nullptr;  // owned by the scene from here
    nodes.p = nullptr;
    if (rc) return rc;
    if (nodes_out && (e = hipMemcpyAsync(nodes_out, tree, count * sizeof(lr_bvh_node), hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(BVH)");
    if (order_out && n && (e = hipMemcpyAsync(order_out, order.p, n * 4, hipMemcpyDeviceToHost, st)) != hipSuccess)
        return hip_fail(e, "hipMemcpy(order)");
    if ((e = hipStreamSynchronize(st)) != hipSuccess) return hip_fail(e, "lr_scene_build");
    if (node_count) *node_count = count;
    *out = s.release();
    return 0;
}

int lr_scene_refit(lr_gpu_scene *scene, const float *positions, size_t points) {
    if (!scene || (points && !positions)) return fail(LR_INVALID, "lr_scene_refit needs a scene and its points");
    if (points != scene->points)
        return fail(LR_INVALID, "a refit moves the scene's points: the same number as it was built with");
    (void)hipSetDevice(scene->device->ordinal);
    if (scene->info.triangles == 0) return 0;
    Scratch<float> pos;
    int rc = pos.put(positions, points * 3, scene->device->stream, "upload points");
    if (rc) return rc;
    return refit_scene_device(scene, pos.p);
}

int lr_scene_read_layout(const lr_gpu_scene *scene, int width, void *out, size_t capacity, size_t *bytes) {
    if (!scene || !bytes) return fail(LR_INVALID, "lr_scene_read_layout needs a scene and a size");
    int wi = layout_index((uint32_t)width);
    if (wi < 0) return fail(LR_INVALID, "layouts are 2, 4, 8 and the quantized 20, 24");
    *bytes = wi < 3 ? scene->info.wide_bytes[wi]
                    : std::max<size_t>(scene->info.wide_nodes[wi - 2], 1) * (wi == 3 ? 64 : 128);
    if (!out) return 0;
    if (capacity < *bytes) return fail(LR_INVALID, "the buffer is smaller than the layout");
    (void)hipSetDevice(scene->device->ordinal);
    hipError_t e = hipMemcpy(out, scene->wide[wi], *bytes, hipMemcpyDeviceToHost);
    return e == hipSuccess ? 0 : hip_fail(e, "hipMemcpy(wide nodes)");
}

int lr_bvh_collapse(const lr_bvh_node *nodes, size_t node_count, int width, void *out, size_t capacity, size_t *bytes,
                    uint32_t *stack_bound) {
    if (!nodes || !bytes || (width != 2 && width != 4 && width != 8))
        return fail(LR_INVALID, "lr_bvh_collapse needs nodes, a size and a width of 2, 4 or 8");
    WideLayout layout = collapse_wide(nodes, node_count, width);
    *bytes = layout.bytes.size();
    if (stack_bound) *stack_bound = layout.stack_bound;
    if (!out) return 0;
    if (capacity < *bytes) return fail(LR_INVALID, "the buffer is smaller than the layout");
    std::memcpy(out, layout.bytes.data(), *bytes);
    return 0;
}

void lr_scene_destroy(lr_gpu_scene *scene) {
    if (!scene) return;
    (void)hipSetDevice(scene->device->ordinal);
    (void)hipStreamSynchronize(scene->device->stream);
    delete scene;
}

int lr_scene_set_trace_config(lr_gpu_scene *scene, const lr_trace_config *config) {
    if (!scene || !config) return fail(LR_INVALID, "lr_scene_set_trace_config needs a scene and a config");
    if (variant_index(config->width, config->stack, config->group) < 0)
        return fail(LR_INVALID, "traversal layouts are 2, 4, 8, 20, 24; stacks 16, 24, 32, 48; groups 64, 128, 256");
    if (config->refill > 64) return fail(LR_INVALID, "refill is 0 (one ray a thread) or 1..64 idle lanes");
    scene->config = *config;
    return 0;
}

int lr_scene_get_trace_config(const lr_gpu_scene *scene, lr_trace_config *config) {
    if (!scene || !config) return fail(LR_INVALID, "lr_scene_get_trace_config needs a scene and a config");
    *config = scene->config;
    return 0;
}

int lr_scene_info_get(const lr_gpu_scene *scene, lr_scene_info *info) {
    if (!scene || !info) return fail(LR_INVALID, "lr_scene_info_get needs a scene");
    *info = scene->info;
    return 0;
}

int lr_probe_device(const lr_gpu_scene *scene_c, const float frame[16], uint32_t width, uint32_t height,
                    uint32_t *hits_device, void *stream) {
    lr_gpu_scene *scene = const_cast<lr_gpu_scene *>(scene_c);
    if (!scene || !frame || !hits_device) return fail(LR_INVALID, "lr_probe_device needs a scene, a frame and output");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(LR_INVALID, "an image is 1..16384 pixels on a side");
    const TraceVariant &v = scene->variant();
    int rc = scene->ensure_spill(launch_threads(probe_items(width, height), scene->config.group), scene->config.stack);
    if (rc) return rc;
    DeviceScene sc = scene->device_scene();
    int e = v.probe(sc, frame, width, height, reinterpret_cast<uint2 *>(hits_device), scene->config.refill,
                    (hipStream_t)stream);
    if (e) return hip_fail(e, "probe launch");
    return 0;
}

int lr_probe(lr_device *device, const lr_gpu_scene *scene, const float frame[16], uint32_t width, uint32_t height,
             uint32_t *face_out, float *t_out, size_t count) {
    if (!device || !scene || !frame || !face_out || !t_out) return fail(LR_INVALID, "lr_probe needs its arguments");
    if ((size_t)width * height != count) return fail(LR_INVALID, "a probe needs one result per pixel");  // device.lucb:139-140
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(LR_INVALID, "an image is 1..16384 pixels on a side");
    (void)hipSetDevice(device->ordinal);
    void *hits = nullptr;
    hipError_t e = hipMalloc(&hits, count * 8);
    if (e != hipSuccess) return hip_fail(e, "hipMalloc(probe results)");
    int rc = lr_probe_device(scene, frame, width, height, static_cast<uint32_t *>(hits), device->stream);
    std::vector<uint32_t> host(count * 2);
    if (!rc) {
        e = hipMemcpyAsync(host.data(), hits, count * 8, hipMemcpyDeviceToHost, device->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(device->stream);
        if (e != hipSuccess) rc = hip_fail(e, "probe readback");
    }
    (void)hipFree(hits);
    if (rc) return rc;
    for (size_t i = 0; i < count; i++) {
        face_out[i] = host[i * 2];
        std::memcpy(&t_out[i], &host[i * 2 + 1], 4);
    }
    return 0;
}

int lr_trace_closest_device(const lr_gpu_scene *scene_c, const float *rays, size_t n, uint32_t *hits, void *stream) {
    lr_gpu_scene *scene = const_cast<lr_gpu_scene *>(scene_c);
    if (!scene || (n && (!rays || !hits))) return fail(LR_INVALID, "lr_trace_closest_device needs its arguments");
    if (n > 0xffffff00ull) return fail(LR_LIMIT, "at most 2^32 - 256 rays a call");
    int rc = scene->ensure_spill(launch_threads(n, scene->config.group), scene->config.stack);
    if (rc) return rc;
    int e = scene->variant().closest(scene->device_scene(), reinterpret_cast<const float4 *>(rays), (uint32_t)n,
                                     reinterpret_cast<uint2 *>(hits), scene->config.refill, (hipStream_t)stream);
    if (e) return hip_fail(e, "closest launch");
    return 0;
}

int lr_trace_frame_and_closest_device(const lr_gpu_scene *scene_c, const float frame[16], uint32_t width,
                                      uint32_t height, uint32_t *frame_hits, const float *rays, size_t n,
                                      uint32_t *hits, void *stream) {
    lr_gpu_scene *scene = const_cast<lr_gpu_scene *>(scene_c);
    if (!scene || !frame || !frame_hits || (n && (!rays || !hits)))
        return fail(LR_INVALID, "lr_trace_frame_and_closest_device needs its arguments");
    if (width == 0 || height == 0 || width > 16384 || height > 16384)
        return fail(LR_INVALID, "an image is 1..16384 pixels on a side");
    if (n > 0xffffff00ull) return fail(LR_LIMIT, "at most 2^32 - 256 rays a call");
    int rc = scene->ensure_spill(launch_threads(probe_items(width, height) + n, scene->config.group), scene->config.stack);
    if (rc) return rc;
    int e = scene->variant().frame_closest(scene->device_scene(), frame, width, height,
                                           reinterpret_cast<uint2 *>(frame_hits), reinterpret_cast<const float4 *>(rays),
                                           (uint32_t)n, reinterpret_cast<uint2 *>(hits), scene->config.refill,
                                           (hipStream_t)stream);
    if (e) return hip_fail(e, "frame and closest launch");
    return 0;
}

int lr_trace_resident_grid(const lr_gpu_scene *scene, size_t n, int any, uint32_t *blocks) {
    if (!scene || !blocks) return fail(LR_INVALID, "lr_trace_resident_grid needs a scene and an output");
    (void)hipSetDevice(scene->device->ordinal);
    *blocks = scene->variant().resident((uint64_t)n, any);
    return 0;
}

int lr_trace_any_device(const lr_gpu_scene *scene_c, const float *rays, size_t n, uint32_t *occluded, void *stream) {
    lr_gpu_scene *scene = const_cast<lr_gpu_scene *>(scene_c);
    if (!scene || (n && (!rays || !occluded))) return fail(LR_INVALID, "lr_trace_any_device needs its arguments");
    if (n > 0xffffff00ull) return fail(LR_LIMIT, "at most 2^32 - 256 rays a call");
    int rc = scene->ensure_spill(launch_threads(n, scene->config.group), scene->config.stack);
    if (rc) return rc;
    int e = scene->variant().any(scene->device_scene(), reinterpret_cast<const float4 *>(rays), (uint32_t)n, occluded,
                                 scene->config.refill, (hipStream_t)stream);
    if (e) return hip_fail(e, "any-hit launch");
    return 0;
}

}  // extern "C"
