/*
 * luce-render on MI355X: the C ABI that replaces luce-gpu plus the embedded
 * SPIR-V/MSL kernels under luce-render's API. Plain pointers and sizes only.
 *
 * Status codes follow src/render/module.lucb:60-63: 0 ok, 400 invalid,
 * 401 limit, 402 stopped; LR_UNAVAILABLE (503) when no GPU can be opened
 * (luce-gpu's `gpu.unavailable`, which tests/gpu/main.lucb:18-22 turns into a
 * skip). The message of the last failure on the calling thread is read with
 * lr_last_error.
 *
 * Ownership: host arrays are borrowed for the call (uploads copy them);
 * handles own their device memory; destroying a NULL handle is a no-op
 * (device.lucb:119-121 destroys default handles the same way). A scene must
 * outlive the renderers that use it. Calls on one handle are not re-entrant.
 */
#ifndef LUCE_RENDER_HIP_H
#define LUCE_RENDER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LR_OK 0
#define LR_INVALID 400
#define LR_LIMIT 401
#define LR_STOPPED 402
#define LR_UNAVAILABLE 503
#define LR_DEVICE_ERROR 500

typedef struct lr_device lr_device;
typedef struct lr_gpu_scene lr_gpu_scene;

/* src/render/bvh.lucb:23-27: the 32-byte std430 node { vec3 lo; u32 first;
 * vec3 hi; u32 count; }. Interior: count 0, children at first and first + 1;
 * leaf: `count` triangles from leaf-order slot `first`. */
typedef struct lr_bvh_node {
    float lo[3];
    uint32_t first;
    float hi[3];
    uint32_t count;
} lr_bvh_node;

/* The compiled scene's arrays (src/render/scene.lucb:31-99) as the host hands
 * them over after build_bvh: positions xyz per point, three point indices per
 * triangle in leaf order, each triangle's polygon, its shading words
 * (3 octahedral normals + material), its six uv floats, and the binary BVH. */
typedef struct lr_scene_arrays {
    const float *positions;
    size_t points;
    const uint32_t *indices;
    const uint32_t *faces;
    const uint32_t *shading; /* 4 per triangle; may be NULL */
    const float *uvs;        /* 6 per triangle; may be NULL */
    size_t triangles;
    const lr_bvh_node *nodes;
    size_t node_count;
    /* Gaussian splats (CompiledSplats, src/render/splats.lucb:45-70), already
     * in their tree's leaf order: 16 words a record, the harmonics words the
     * records name, and the splats' own binary BVH. May be empty; both upload
     * and build take them as given. */
    const uint32_t *splat_records;
    size_t splats;
    const uint32_t *splat_harmonics;
    size_t splat_words;
    const lr_bvh_node *splat_nodes;
    size_t splat_node_count;
} lr_scene_arrays;

/* Replaces gpu.Device.open (luce-gpu, used at tests/gpu/main.lucb:18). */
int lr_device_open(int ordinal, lr_device **out);
void lr_device_close(lr_device *device);

/* The message of the calling thread's last failure; returns its length. */
size_t lr_last_error(char *buffer, size_t capacity);

/* Replaces build_bvh (src/render/bvh.lucb:66-217): the binned-SAH binary BVH
 * over `triangles` triangles, the same tree and leaf order the reference
 * builds. `nodes` holds at least max(1, 2 * triangles) entries; `order`
 * receives each leaf slot's original triangle; `threads` 0 picks the host's
 * cores. The caller reorders its per-triangle arrays by `order`. */
int lr_bvh_build(const float *positions, size_t points, const uint32_t *indices, size_t triangles,
                 int threads, lr_bvh_node *nodes, size_t *node_count, uint32_t *order);

/* build_tree (src/render/bvh.lucb:85-133) over boxes, as build_splat_bvh
 * (src/render/splats.lucb:193-213) builds the splats' tree: `boxes` six floats
 * each (low xyz, high xyz); `nodes` room for max(1, 2 * count); the leaf order
 * in `order`. */
int lr_bvh_build_boxes(const float *boxes, size_t count, int threads, lr_bvh_node *nodes, size_t *node_count,
                       uint32_t *order);

/* References for a build over pieces of triangles (no reference counterpart:
 * for scenes whose triangles overlap each other's boxes by the hundreds,
 * where whole-triangle boxes make every ray test hundreds of triangles).
 * Each triangle whose box is longer than `min_extent` on an axis is cut at
 * that axis's middle, its pieces cut again while long, at most `max_depth`
 * (0..8) cuts deep; `splittable` (null: all) marks the triangles that may be
 * cut. Out: per reference six floats (low xyz, high xyz: the piece's box, a
 * few ulps wider, inside the triangle's box) and its triangle, in triangle
 * order. Null `boxes` or `tris` gives the count alone in `count`. The
 * references are built over with lr_bvh_build_boxes; a leaf slot then holds
 * the triangle of its reference (a triangle may sit in several leaves; hits
 * are the same). */
int lr_bvh_split_triangles(const float *positions, size_t points, const uint32_t *indices, size_t triangles,
                           const uint8_t *splittable, float min_extent, uint32_t max_depth, int threads,
                           float *boxes, uint32_t *tris, size_t capacity, size_t *count);

/* build_bvh (src/render/bvh.lucb:66-217) on the GPU: the same tree, node
 * numbering and leaf order as lr_bvh_build (node boxes equal as values; a zero
 * plane that occurs with both signs may keep either sign). Host arrays as in
 * lr_bvh_build; the device holds the work (about 120 bytes per triangle). */
int lr_bvh_build_device(lr_device *device, const float *positions, size_t points, const uint32_t *indices,
                        size_t triangles, lr_bvh_node *nodes, size_t *node_count, uint32_t *order);

/* Replaces GpuScene.upload (src/render/device.lucb:41-89). The binary BVH is
 * repacked on upload into the traversal layouts (width 2, 4 and 8 with float
 * child boxes, 64/128/256 byte nodes; widths 4 and 8 with child boxes
 * quantized to bytes, 64/128 byte nodes) and a leaf-order triangle array
 * (p0, e1, e2). */
int lr_scene_upload(lr_device *device, const lr_scene_arrays *arrays, lr_gpu_scene **out);
/* compile_scene's build_bvh and reorder (bvh.lucb:66-217) followed by
 * GpuScene.upload (device.lucb:41-89), all on the device: `arrays` in the
 * order the scene was flattened (nodes ignored); the BVH is built on the GPU
 * (the tree lr_bvh_build gives), the per-triangle arrays are put in leaf order
 * there, and the traversal layouts collapsed there. When given, `nodes` (room
 * for max(1, 2 * triangles)), `node_count` and `order` receive the tree and
 * leaf order for the host's copy of the scene. */
int lr_scene_build(lr_device *device, const lr_scene_arrays *arrays, lr_gpu_scene **out, lr_bvh_node *nodes,
                   size_t *node_count, uint32_t *order);

/* Moves the scene's points (the same number it was built with; triangles,
 * tree and leaf order unchanged) and refits on the device: every node's box
 * recomputed bottom-up from the new points, each wide node keeping the
 * children it was built with, the triangles repacked. No reference
 * counterpart: the reference recompiles and re-uploads a changed scene; this
 * serves animated geometry whose topology holds. Renderers on the scene trace
 * the new geometry from their next pass (restart them with lr_renderer_reset to
 * drop samples of the old); emitters' records come from the loaded arrays. */
int lr_scene_refit(lr_gpu_scene *scene, const float *positions, size_t points);

/* The width-`width` traversal nodes as uploaded (`bytes` receives their size;
 * with `out` NULL only the size), and the host collapse of a binary tree into
 * the same layout (the repacking of bvh_wide.cpp, for comparison). */
int lr_scene_read_layout(const lr_gpu_scene *scene, int width, void *out, size_t capacity, size_t *bytes);
int lr_bvh_collapse(const lr_bvh_node *nodes, size_t node_count, int width, void *out, size_t capacity, size_t *bytes,
                    uint32_t *stack_bound);

/* GpuScene.destroy (device.lucb:108). */
void lr_scene_destroy(lr_gpu_scene *scene);

/* Replaces probe (src/render/device.lucb:138-160): one pixel-centre ray per
 * pixel of the 16-float frame (camera.lucb:41-47), width x height results row
 * by row from the top: polygon (0xffffffff on a miss) and distance along the
 * frame's unnormalized direction. `count` must equal width * height. */
int lr_probe(lr_device *device, const lr_gpu_scene *scene, const float frame[16], uint32_t width,
             uint32_t height, uint32_t *face_out, float *t_out, size_t count);

/* Traversal configuration: the node layout (width 2, 4, 8 with float child
 * boxes; LR_QUANTIZED | 4 and LR_QUANTIZED | 8 with 8-bit child boxes, each
 * holding its float box), workgroup size (64, 128, 256),
 * LDS stack entries per lane (16, 24, 32, 48) and the ray schedule of the ray-batch
 * launches: refill 0 gives each thread one ray; refill k (1..64) runs
 * persistent waves that take new rays from a shared counter once k of their
 * lanes are idle (the probe keeps one ray per thread). Every configuration returns the same hits. No reference
 * counterpart: the reference fixes width 2, 64-wide groups, a 32-entry stack
 * (bvh.glsl:49) and one ray per invocation. Launches on one scene are ordered
 * (one stream at a time): they share its spill area and ray counter. */
#define LR_QUANTIZED 16
typedef struct lr_trace_config {
    uint32_t width;
    uint32_t group;
    uint32_t stack;
    uint32_t refill;
} lr_trace_config;
int lr_scene_set_trace_config(lr_gpu_scene *scene, const lr_trace_config *config);
int lr_scene_get_trace_config(const lr_gpu_scene *scene, lr_trace_config *config);

/* The probe with device-resident output (for timing and for callers that keep
 * results on the GPU): hits[i] = {polygon, t bits}, as probe.comp:33 writes
 * them. `stream` is a hipStream_t (NULL: the default stream); nothing waits. */
int lr_probe_device(const lr_gpu_scene *scene, const float frame[16], uint32_t width, uint32_t height,
                    uint32_t *hits_device, void *stream);

/* Closest hits of n device-resident rays, 8 floats each (origin xyz, t_max,
 * direction xyz, unused), into device hits[i] = {polygon, t bits}. The ray
 * semantics are trace_closest's (bvh.glsl:51-90). */
int lr_trace_closest_device(const lr_gpu_scene *scene, const float *rays_device, size_t n,
                            uint32_t *hits_device, void *stream);

/* lr_probe_device of a frame and lr_trace_closest_device of n rays in one
 * call: no reference counterpart. For callers that trace a frame's primary
 * rays and a batch together (a renderer's camera and bounce rays): with the
 * scene's refill above 0, one persistent launch takes the frame's 8x8 tiles
 * and then the batch's rays from one counter, so the two share one ramp-up
 * and one drain; with refill 0, the two launches. The same hits as the two
 * calls, bit for bit. */
int lr_trace_frame_and_closest_device(const lr_gpu_scene *scene, const float frame[16], uint32_t width,
                                      uint32_t height, uint32_t *frame_hits_device, const float *rays_device,
                                      size_t n, uint32_t *hits_device, void *stream);

/* Whether anything blocks each of n device-resident rays before its t_max
 * (trace_any, bvh.glsl:93-122): occluded[i] = 1 or 0. */
int lr_trace_any_device(const lr_gpu_scene *scene, const float *rays_device, size_t n,
                        uint32_t *occluded_device, void *stream);

/* lr_trace_closest_device without allocating: no reference counterpart. For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same hits, bit
 * for bit, by persistent refilled waves whose spill area and counters live
 * in a 32 MiB array the library holds on each device from load to process
 * end; the grid is cut to the workgroups whose spill columns fit it. Calls
 * that overlap on one device race on that array (serialize them on one
 * stream). Fails with LR_LIMIT for n >= 2^31 or when one workgroup's spill
 * columns exceed the array. */
int lr_trace_closest_device_noalloc(const lr_gpu_scene *scene, const float *rays_device, size_t n,
                                    uint32_t *hits_device, void *stream);

/* Workgroups a refilled (persistent) launch of the scene's traversal variant
 * takes over n rays on its device: the resident workgroups of that variant's
 * kernel (closest hit, or any hit when `any`), at most what n rays need. No
 * reference counterpart (the autotuner's grid sizing, reported). */
int lr_trace_resident_grid(const lr_gpu_scene *scene, size_t n, int any, uint32_t *blocks);

/* Scene statistics for roofline accounting: node counts per width, bytes of
 * each layout, triangle bytes, and the stack bound of each width. */
typedef struct lr_scene_info {
    size_t triangles;
    size_t binary_nodes;
    size_t wide_nodes[3];     /* widths 2, 4, 8 */
    size_t wide_bytes[3];
    size_t triangle_bytes;
    uint32_t stack_bound[3];  /* deepest stack any ray can need, per width */
} lr_scene_info;
int lr_scene_info_get(const lr_gpu_scene *scene, lr_scene_info *info);

/* ------------------------------------------------------------------ renderer */

typedef struct lr_renderer lr_renderer;

/* What Loaded.create packs beside the geometry (src/render/loading.lucb:51-161):
 * light records (7 float4 each) followed by the light tree's nodes (3 float4
 * each); the 32 constants rows (128 floats, loading.lucb:253-284 and the view
 * rows of :279-284); material records (16 float4 each, the default material
 * first); the emitter words (tree links, emitter records, triangle map); the
 * light paths a sample (Loaded.light_paths, :166-169) and the extra rounds
 * cut-outs need (8 or 0). */
typedef struct lr_loaded {
    const float *lights;
    size_t light_floats;
    const float *constants;
    const float *materials;
    size_t material_floats;
    const uint32_t *emitters;
    size_t emitter_words;
    uint32_t light_paths;
    uint32_t through;
    /* Materials' node-graph programs (shader_graph.lucb: u32 words, each
     * material's entry in its record's row 13 x, -1 none) and the textures
     * their image nodes read (RGBA f32 texels, rows from the top; per texture
     * slot the first texel, width, height and a zero). May be empty. */
    const uint32_t *programs;
    size_t program_words;
    const float *textures;
    size_t texture_floats;
    const uint32_t *texture_table;
    size_t texture_count;
} lr_loaded;

/* RenderSettings (src/render/integrator.lucb:42-85) as far as the device
 * needs them; the rest is in the loaded constants. pool 0: min(2^21, pixels).
 * use_graph: replay each pass from one recorded hipGraph. trace_*: the
 * traversal variant (0: the scene's configuration). guiding: 0 off, 1 the
 * reference's hash grid of vMF mixtures, 2 the learned guide (an MLP on the
 * matrix cores trained online during the first train_samples samples, 0: 8,
 * in passes of a quarter of them, each followed by train_steps Adam steps for
 * each sample it carried, 0: 8; single renders only). */
typedef struct lr_render_settings {
    uint32_t width;
    uint32_t height;
    uint32_t samples;
    uint32_t max_bounces;
    uint32_t spectral;
    uint32_t pool;
    uint32_t seed;
    uint32_t guiding;
    uint32_t profile;
    uint32_t use_graph;
    uint32_t trace_width;
    uint32_t trace_group;
    uint32_t trace_stack;
    uint32_t train_samples;
    uint32_t train_steps;
    /* Samples one pass may carry when a sample is one band (no reference
     * counterpart: the reference's passes carry one): 0 picks as many as fit
     * in about 2^26 path slots, at most 16; 1 is the reference's schedule.
     * Renders that stop pixels by a noise threshold take one; the learned
     * guide's training samples go in passes of a quarter of them; the hash grid
     * learns after each pass. An unguided image is the same either way
     * (camera paths bit for bit). */
    uint32_t pass_samples;
} lr_render_settings;

/* Renderer.image / display: device pointers to the rgba32f ACEScg image and
 * the exposed, clipped linear Rec.709 display image (rgba32f here). */
typedef struct lr_image {
    void *rgba;
    void *display;
    uint32_t width;
    uint32_t height;
} lr_image;

/* Renderer.create (integrator.lucb:343): `frame` is the 16-float RayFrame of
 * the settings' size. The scene must outlive the renderer. */
int lr_renderer_create(lr_device *device, lr_gpu_scene *scene, const lr_loaded *loaded,
                       const lr_render_settings *settings, const float frame[16], lr_renderer **out);
/* Renderer.advance (:649): launches the next pass unless two are in flight;
 * *busy = 0 once every sample is in the image. Never waits for the GPU. */
int lr_renderer_advance(lr_renderer *renderer, int *busy);
/* Renderer.finish (:724): every remaining pass, waited for. */
int lr_renderer_finish(lr_renderer *renderer);
/* Renderer.samples_done (:732): samples every pixel has in finished passes. */
uint32_t lr_renderer_samples_done(lr_renderer *renderer);
/* Renderer.image (:737). */
int lr_renderer_image(lr_renderer *renderer, lr_image *out);
/* Renderer.read (:778): width * height * 4 floats, row by row from the top. */
int lr_renderer_read(lr_renderer *renderer, float *rgba, size_t count);
/* The display image (Renderer.display, :743: exposed, linear Rec.709, through
 * the view, clamped to [0, 1]) read like lr_renderer_read. */
int lr_renderer_read_display(lr_renderer *renderer, float *rgba, size_t count);
/* Renderer.set_display (:750-756): show the image at `exposure` stops through
 * `view` (0 standard, 1 Khronos PBR Neutral tone mapping). The next pass's
 * conversion applies it, or, once every sample is in, a conversion of its own
 * that lr_renderer_advance / lr_renderer_finish launch (:656-669). A restart,
 * reload or scene replacement takes the settings of its loaded constants. */
int lr_renderer_set_display(lr_renderer *renderer, float exposure, int view);
/* The film's four planes (sum, half, scratch, splat; common.glsl:64-75),
 * width * height * 16 floats: no reference counterpart, for tests. */
int lr_renderer_read_film(lr_renderer *renderer, float *planes, size_t count);
/* The hash-grid guide's buffer (guiding 1; shaders/guide.glsl:12-31's regions:
 * 2^19 u32 keys, then per cell 4 lobes and their weights as float4, then per
 * cell 4 x 5 float statistics), after the passes submitted so far. `bytes`
 * must be its size, 524,288 x 164. */
int lr_renderer_read_guide(lr_renderer *renderer, void *out, size_t bytes);
/* Renderer.kernel_times (:873): milliseconds per kernel kind (profile on),
 * in the order camera, intersect closest, shade surface, shade miss,
 * intersect shadow, light emission, finish, film convert, adaptive, rest. */
int lr_renderer_kernel_times(lr_renderer *renderer, double out[10]);
/* Counters of the work done so far (waits for passes in flight): no
 * reference counterpart, for throughput reports. */
typedef struct lr_render_stats {
    uint32_t samples_done;
    uint32_t light_samples;
    uint64_t closest_rays;
    uint64_t shadow_rays;
} lr_render_stats;
int lr_renderer_stats(lr_renderer *renderer, lr_render_stats *out);
/* Renderer.destroy (:887): waits for passes in flight. */
/* Renderer.restart / reload / replace_scene (integrator.lucb:463-597): start
 * over with new loaded arrays, settings and frame, and with `scene` in place of
 * the renderer's when not NULL (which must then outlive the renderer). Passes
 * in flight finish and count for nothing; the film, sample count and kernel
 * times start over. The projection (perspective or orthographic) is kept.
 * With the same scene, a view (or lights and materials) whose image and path
 * slots fit the buffers held keeps them, and the image and display pointers
 * stay; a view with more pixels than they hold makes them again with a
 * quarter more room each way, rounded up to 64 (integrator.lucb:474-478,
 * 986-989); otherwise the device state is made anew and the pointers change.
 * On failure the renderer is left as it was. */
int lr_renderer_reset(lr_renderer *renderer, lr_gpu_scene *scene, const lr_loaded *loaded,
                      const lr_render_settings *settings, const float frame[16]);
/* Renderer.converged_pixels (integrator.lucb:759): pixels adaptive sampling
 * has stopped (0 without a noise threshold); waits for passes in flight. */
int lr_renderer_converged_pixels(lr_renderer *renderer, size_t *out);
void lr_renderer_destroy(lr_renderer *renderer);

/* ------------------------------------------------------------------- serving */

/* Many renders resident on one GPU, advanced together: each step runs one
 * sample of every active request in shared kernel launches. Requests are
 * admitted in arrival order once their device memory fits the budget. No
 * reference counterpart (the reference renders one view at a time). */
typedef struct lr_server lr_server;
typedef struct lr_server_settings {
    uint64_t budget_bytes;  /* device memory for requests' state, films and images */
    uint32_t spectral;      /* every request's build */
    uint32_t trace_width;   /* the traversal variant every request uses */
    uint32_t trace_group;
    uint32_t trace_stack;
    uint32_t max_active;    /* 0: no limit beyond the budget (and 128, the rows of a shared launch) */
} lr_server_settings;
typedef struct lr_request_status {
    uint32_t state;         /* 0 waiting for admission, 1 rendering, 2 finished */
    uint32_t samples_done;
    uint32_t samples;
    uint64_t bytes;         /* its device memory */
} lr_request_status;
int lr_server_create(lr_device *device, const lr_server_settings *settings, lr_server **out);
/* Queues a render of `scene` (which must outlive the request); the loaded
 * arrays are copied. 401 when the request alone exceeds the budget. */
int lr_server_submit(lr_server *server, lr_gpu_scene *scene, const lr_loaded *loaded,
                     const lr_render_settings *settings, const float frame[16], uint64_t *id);
/* Admits what fits, runs one sample of every active request and waits for
 * it; requests that reach their samples finish. */
int lr_server_step(lr_server *server, uint32_t *active, uint32_t *pending);
int lr_server_status(lr_server *server, uint64_t id, lr_request_status *out);
/* A finished request's image (rgba32f, row by row from the top). */
int lr_server_read(lr_server *server, uint64_t id, float *rgba, size_t count);
int lr_server_stats(lr_server *server, uint64_t id, lr_render_stats *out);
/* Frees a request: cancels it when waiting or rendering. */
int lr_server_release(lr_server *server, uint64_t id);
void lr_server_destroy(lr_server *server);

/* Device buffers and timing: the part of luce-gpu (gpu.Buffer.allocate /
 * create / read / upload, Compute.wait and timestamps; used at
 * src/render/device.lucb:137-156 and integrator.lucb:848-885) that callers of
 * the device-resident entry points need. Copies run on the device's stream
 * and complete before returning. */
int lr_buffer_alloc(lr_device *device, size_t bytes, void **out);
void lr_buffer_free(lr_device *device, void *buffer);
int lr_buffer_write(lr_device *device, void *dst_device, const void *src_host, size_t bytes);
int lr_buffer_read(lr_device *device, void *dst_host, const void *src_device, size_t bytes);
/* A plain device-to-device copy kernel on the device's stream (16-byte
 * aligned, whole 16-byte words; nothing waits): no reference counterpart, the
 * known-traffic kernel hardware byte counters are scaled against. */
int lr_buffer_copy(lr_device *device, void *dst_device, const void *src_device, size_t bytes);
/* The device's stream (a hipStream_t) and a wait for all work on it. */
void *lr_device_stream(lr_device *device);
int lr_device_synchronize(lr_device *device);
/* HIP events for timing work on a stream. */
int lr_event_create(void **event);
int lr_event_record(void *event, void *stream);
int lr_event_elapsed(void *start, void *stop, float *ms);
void lr_event_destroy(void *event);

/* ------------------------------------------------------------------------
 * The learned guide: an MLP (64 one-blob inputs of a surface point and its
 * normal, three hidden layers of 64, 32 outputs) giving a mixture of four von
 * Mises-Fisher lobes over incident directions, run on the matrix cores in
 * FP16 with FP32 accumulation and trained by Adam on the loss
 * -weight * log q(direction). No reference counterpart (the reference guides
 * with its hash grid only); these entry points expose the network on its own
 * (the renderer trains it internally when guiding is 2).
 *   samples: 10 floats each (position xyz, normal xyz, direction xyz, weight)
 *   points:  6 floats each (position xyz, normal xyz)
 *   mixtures: 20 floats a point (lobe k: direction xyz, concentration at
 *             4k..4k+3; the four lobe weights at 16..19) */
typedef struct lr_guide_net lr_guide_net;
int lr_guide_net_create(lr_device *device, const float lo[3], const float hi[3], uint32_t seed, lr_guide_net **out);
void lr_guide_net_destroy(lr_guide_net *net);
size_t lr_guide_net_param_count(void);
/* Parameters: W0, W1, W2 (64 x 64), W3 (64 x 32) row = input, then biases
 * b0, b1, b2 (64), b3 (32). Setting them clears the Adam moments. */
int lr_guide_net_set_params(lr_guide_net *net, const float *params, size_t count);
int lr_guide_net_get_params(lr_guide_net *net, float *params, size_t count);
/* The mean gradient of the batch loss sum_s (w_s / sum w) (-log q(d_s | x_s))
 * by every parameter, and the mean per-sample loss. */
int lr_guide_net_gradient(lr_guide_net *net, const float *samples, size_t n, float *grad, size_t count, float *loss);
/* One Adam step (beta 0.9 / 0.99) on the batch at learning rate `rate`. */
int lr_guide_net_step(lr_guide_net *net, const float *samples, size_t n, float rate, float *loss);
/* Mixtures of the points. After steps they come from the moving average of
 * the trained parameters (weight 0.8 on the average so far, bias-corrected as
 * Adam's moments are), which the renderer's guide uses too; after
 * lr_guide_net_set_params, from the parameters set. lr_guide_net_get_params
 * returns the trained parameters. */
int lr_guide_net_mixtures(lr_guide_net *net, const float *points, size_t n, float *out);

#ifdef __cplusplus
}
#endif

#endif
