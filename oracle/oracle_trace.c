/*
 * Oracle (test infrastructure only): a CPU restatement of luce-render's
 * closest-hit path, used by the tests and by bench.py's CPU-baseline leg as the
 * checker. Nothing in the product links or calls this file.
 *
 * Contents and the reference text each part follows:
 *   or_ray_frame / or_look_frame   src/render/camera.lucb:57-76, 84-95 (f64)
 *   or_pack_frame                  src/render/camera.lucb:41-47
 *   or_bvh_build                   src/render/bvh.lucb:29-217 (f32, single thread)
 *   or_trace_closest               shaders/bvh.glsl:11-90 (f32 mirror, same op order)
 *   or_trace_any                   shaders/bvh.glsl:93-122
 *   or_probe                       shaders/probe.comp:21-34 (pixel-centre raygen)
 *   or_brute_force                 tests/gpu/main.lucb:82-119 (f64 Moller-Trumbore,
 *                                  nearest polygon plus the margin to the next one)
 *
 * Pinning (what checks each function beyond this restatement):
 *   ray frame      the reference's closed forms (tests.lucb:63-91);
 *   bvh build      check_bvh invariants (tests.lucb:9-28), an independent f64
 *                  recompute of every split's SAH cost, and the grid/cube counts;
 *   trace mirror   the f64 brute force on every pixel of the probe scenes
 *                  (same polygon and |dt| <= 1e-4 t unless the margin < 1e-3);
 *   brute force    closed-form hits of hand-placed triangles.
 *
 * Build: gcc -O2 -ffp-contract=off (no fast-math), see oracle/Makefile.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>

/* ------------------------------------------------------------------ frames */

/* camera.lucb:57-76. m holds the placement's axes x, y, z (rows 0..2, not yet
 * normalized) and translation (row 3). out: position[3], corner[3], du[3],
 * dv[3], forward[3], lens_radius, focus, orthographic. */
int or_ray_frame(const double m[12], int orthographic, double horizontal_aperture,
                 double focal_length, double f_stop, double focus_distance,
                 uint32_t width, uint32_t height, double out[18]) {
    if (width == 0 || height == 0) return 400;
    double ax[3][3];
    for (int r = 0; r < 3; r++) {
        double l = sqrt(m[r * 3] * m[r * 3] + m[r * 3 + 1] * m[r * 3 + 1] + m[r * 3 + 2] * m[r * 3 + 2]);
        if (!(l > 0.0)) return 400;
        for (int k = 0; k < 3; k++) ax[r][k] = m[r * 3 + k] / l;
    }
    double forward[3] = {-ax[2][0], -ax[2][1], -ax[2][2]};
    double aspect = (double)height / (double)width;
    double half_width = orthographic ? 0.05 * horizontal_aperture : 0.5 * horizontal_aperture / focal_length;
    double half_height = half_width * aspect;
    for (int k = 0; k < 3; k++) {
        out[k] = m[9 + k];
        out[3 + k] = ax[0][k] * -half_width + ax[1][k] * half_height + (orthographic ? 0.0 : forward[k]);
        out[6 + k] = ax[0][k] * (2.0 * half_width / (double)width);
        out[9 + k] = ax[1][k] * (-2.0 * half_height / (double)height);
        out[12 + k] = forward[k];
    }
    out[15] = (f_stop > 0.0 && !orthographic) ? 0.0005 * focal_length / f_stop : 0.0;
    out[16] = focus_distance;
    out[17] = orthographic ? 1.0 : 0.0;
    return 0;
}

/* camera.lucb:84-95. */
int or_look_frame(const double position[3], const double forward[3], const double up[3],
                  double tan_half_height, uint32_t width, uint32_t height, int orthographic,
                  double out[18]) {
    if (width == 0 || height == 0) return 400;
    if (!(tan_half_height > 0.0 && tan_half_height < 1e6)) return 400;
    double l = sqrt(forward[0] * forward[0] + forward[1] * forward[1] + forward[2] * forward[2]);
    if (!(l > 0.0)) return 400;
    double a[3] = {forward[0] / l, forward[1] / l, forward[2] / l};
    double r[3] = {a[1] * up[2] - a[2] * up[1], a[2] * up[0] - a[0] * up[2], a[0] * up[1] - a[1] * up[0]};
    l = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (!(l > 0.0)) return 400;
    for (int k = 0; k < 3; k++) r[k] /= l;
    double above[3] = {r[1] * a[2] - r[2] * a[1], r[2] * a[0] - r[0] * a[2], r[0] * a[1] - r[1] * a[0]};
    double half_height = tan_half_height;
    double half_width = half_height * (double)width / (double)height;
    for (int k = 0; k < 3; k++) {
        out[k] = position[k];
        out[3 + k] = r[k] * -half_width + above[k] * half_height + (orthographic ? 0.0 : a[k]);
        out[6 + k] = r[k] * (2.0 * half_width / (double)width);
        out[9 + k] = above[k] * (-2.0 * half_height / (double)height);
        out[12 + k] = a[k];
    }
    out[15] = 0.0;
    out[16] = 1.0;
    out[17] = orthographic ? 1.0 : 0.0;
    return 0;
}

/* camera.lucb:41-47: the 16 floats the kernels read. */
void or_pack_frame(const double frame[18], uint32_t width, uint32_t height, float out[16]) {
    double v[16] = {frame[0], frame[1], frame[2], frame[17],
                    frame[3], frame[4], frame[5], 0.0,
                    frame[6], frame[7], frame[8], (double)width,
                    frame[9], frame[10], frame[11], (double)height};
    for (int k = 0; k < 16; k++) out[k] = (float)v[k];
}

/* --------------------------------------------------------------- BVH build */

typedef struct { float lo[3]; uint32_t first; float hi[3]; uint32_t count; } or_node;
typedef struct { float lo[3], hi[3]; } or_box;

#define OR_BINS 12
#define OR_LEAF_SIZE 4
#define OR_LEAF_LIMIT 16

static or_box box_empty(void) {
    or_box b = {{3.0e38f, 3.0e38f, 3.0e38f}, {-3.0e38f, -3.0e38f, -3.0e38f}};
    return b;
}
static float fmin32(float a, float b) { return a < b ? a : b; }
static float fmax32(float a, float b) { return a > b ? a : b; }
static or_box box_grow(or_box a, or_box b) {
    or_box r;
    for (int k = 0; k < 3; k++) { r.lo[k] = fmin32(a.lo[k], b.lo[k]); r.hi[k] = fmax32(a.hi[k], b.hi[k]); }
    return r;
}
static float box_area(const or_box *b) {
    float x = b->hi[0] - b->lo[0];
    float y = b->hi[1] - b->lo[1];
    float z = b->hi[2] - b->lo[2];
    if (x < 0.0f || y < 0.0f || z < 0.0f) return 0.0f;
    float xy = x * y, yz = y * z, zx = z * x;
    float s = xy + yz;
    s = s + zx;
    return 2.0f * s;
}

static size_t bin_of(const float c[3], const or_box *middle, int axis) {
    float lo = middle->lo[axis];
    float extent = middle->hi[axis] - lo;
    if (extent <= 0.0f) return 0;
    float scaled = (float)OR_BINS * (c[axis] - lo);
    float q = scaled / extent;
    int64_t at = (int64_t)q;
    if (at < 0) return 0;
    return at < OR_BINS ? (size_t)at : OR_BINS - 1;
}

typedef struct { int axis; size_t bin; float cost; } or_split;

static or_split best_split(const or_box *boxes, const float (*centers)[3], const uint32_t *order, size_t n,
                           const or_box *parent, const or_box *middle, double *f64_cost) {
    or_split best = {-1, 0, 3.0e38f};
    float area = box_area(parent);
    if (area <= 0.0f) return best;
    for (int axis = 0; axis < 3; axis++) {
        if (middle->hi[axis] - middle->lo[axis] <= 0.0f) continue;
        size_t counts[OR_BINS];
        or_box bounds[OR_BINS];
        for (int b = 0; b < OR_BINS; b++) { counts[b] = 0; bounds[b] = box_empty(); }
        for (size_t i = 0; i < n; i++) {
            size_t b = bin_of(centers[order[i]], middle, axis);
            counts[b]++;
            bounds[b] = box_grow(bounds[b], boxes[order[i]]);
        }
        float right_area[OR_BINS];
        size_t right_count[OR_BINS];
        or_box sweep = box_empty();
        size_t total = 0;
        for (int b = OR_BINS - 1; b > 0; b--) {
            sweep = box_grow(sweep, bounds[b]);
            total += counts[b];
            right_area[b] = box_area(&sweep);
            right_count[b] = total;
        }
        or_box left = box_empty();
        size_t left_count = 0;
        for (int s = 1; s < OR_BINS; s++) {
            left = box_grow(left, bounds[s - 1]);
            left_count += counts[s - 1];
            if (left_count == 0 || right_count[s] == 0) continue;
            float la = box_area(&left);
            float l = la * (float)left_count;
            float r = right_area[s] * (float)right_count[s];
            float sum = l + r;
            float cost = 1.0f + sum / area;
            if (cost < best.cost) {
                best.axis = axis; best.bin = (size_t)s; best.cost = cost;
                if (f64_cost) *f64_cost = 1.0 + ((double)la * (double)left_count + (double)right_area[s] * (double)right_count[s]) / (double)area;
            }
        }
    }
    return best;
}

static size_t partition(const float (*centers)[3], uint32_t *order, size_t n, const or_box *middle, int axis, size_t bin) {
    size_t low = 0, high = n;
    while (low < high) {
        if (bin_of(centers[order[low]], middle, axis) < bin) low++;
        else {
            high--;
            uint32_t held = order[low];
            order[low] = order[high];
            order[high] = held;
        }
    }
    return low;
}

typedef struct { size_t node, first, count; } or_pending;

/* bvh.lucb:66-129. Fills nodes (capacity >= 2*count, or 1 for an empty scene)
 * and order; returns the node count. max_cost_gap (optional) receives the
 * largest |f32 cost - f64 recompute| over the splits taken. */
size_t or_bvh_build(const float *positions, const uint32_t *indices, size_t count,
                    or_node *nodes, uint32_t *order, double *max_cost_gap) {
    if (count == 0) {
        or_node n = {{1.0f, 1.0f, 1.0f}, 0, {-1.0f, -1.0f, -1.0f}, 0};
        nodes[0] = n;
        return 1;
    }
    or_box *boxes = (or_box *)malloc(count * sizeof(or_box));
    float (*centers)[3] = (float (*)[3])malloc(count * sizeof(float[3]));
    for (size_t t = 0; t < count; t++) {
        or_box b = box_empty();
        for (int k = 0; k < 3; k++) {
            const float *p = positions + (size_t)indices[t * 3 + k] * 3;
            or_box q = {{p[0], p[1], p[2]}, {p[0], p[1], p[2]}};
            b = box_grow(b, q);
        }
        boxes[t] = b;
        for (int k = 0; k < 3; k++) centers[t][k] = 0.5f * (b.lo[k] + b.hi[k]);
        order[t] = (uint32_t)t;
    }
    size_t node_count = 1;
    memset(&nodes[0], 0, sizeof(or_node));
    size_t cap = 64, top = 0;
    or_pending *stack = (or_pending *)malloc(cap * sizeof(or_pending));
    stack[top++] = (or_pending){0, 0, count};
    double gap = 0.0;
    while (top > 0) {
        or_pending job = stack[--top];
        or_box bounds = box_empty(), middle = box_empty();
        for (size_t at = job.first; at < job.first + job.count; at++) {
            bounds = box_grow(bounds, boxes[order[at]]);
            or_box c = {{centers[order[at]][0], centers[order[at]][1], centers[order[at]][2]},
                        {centers[order[at]][0], centers[order[at]][1], centers[order[at]][2]}};
            middle = box_grow(middle, c);
        }
        float leaf_cost = (float)job.count;
        size_t left_count = 0;
        or_split split = {-1, 0, 3.0e38f};
        double c64 = 0.0;
        if (job.count > OR_LEAF_SIZE)
            split = best_split(boxes, (const float (*)[3])centers, order + job.first, job.count, &bounds, &middle, &c64);
        if (job.count > OR_LEAF_SIZE && (split.cost < leaf_cost || job.count > OR_LEAF_LIMIT)) {
            if (split.axis >= 0 && split.cost < 3.0e38f) {
                left_count = partition((const float (*)[3])centers, order + job.first, job.count, &middle, split.axis, split.bin);
                double d = fabs((double)split.cost - c64);
                if (d > gap) gap = d;
            }
            if (left_count == 0 || left_count == job.count) left_count = job.count / 2;
        }
        if (left_count == 0) {
            or_node *n = &nodes[job.node];
            memcpy(n->lo, bounds.lo, 12); memcpy(n->hi, bounds.hi, 12);
            n->first = (uint32_t)job.first; n->count = (uint32_t)job.count;
            continue;
        }
        size_t child = node_count;
        memset(&nodes[child], 0, 2 * sizeof(or_node));
        node_count += 2;
        or_node *n = &nodes[job.node];
        memcpy(n->lo, bounds.lo, 12); memcpy(n->hi, bounds.hi, 12);
        n->first = (uint32_t)child; n->count = 0;
        if (top + 2 > cap) { cap *= 2; stack = (or_pending *)realloc(stack, cap * sizeof(or_pending)); }
        stack[top++] = (or_pending){child + 1, job.first + left_count, job.count - left_count};
        stack[top++] = (or_pending){child, job.first, left_count};
    }
    free(stack);
    free(boxes);
    free(centers);
    if (max_cost_gap) *max_cost_gap = gap;
    return node_count;
}

/* bvh.lucb:198-217: triangle attributes into leaf order (stride words each). */
void or_reorder_u32(const uint32_t *src, uint32_t *dst, const uint32_t *order, size_t count, size_t stride) {
    for (size_t at = 0; at < count; at++)
        for (size_t k = 0; k < stride; k++) dst[at * stride + k] = src[(size_t)order[at] * stride + k];
}

/* tests.lucb:9-28 plus "children nested": 0 when every invariant holds, else a
 * code naming the first failure. */
int or_check_bvh(const or_node *nodes, size_t node_count, const float *positions, const uint32_t *indices, size_t count) {
    uint32_t *seen = (uint32_t *)calloc(count ? count : 1, 4);
    int bad = 0;
    for (size_t i = 0; i < node_count && !bad; i++) {
        const or_node *n = &nodes[i];
        if (n->count == 0) {
            if (count == 0) continue;
            for (int c = 0; c < 2; c++) {
                const or_node *in = &nodes[n->first + c];
                if (n->first + c >= node_count) { bad = 1; break; }
                for (int k = 0; k < 3; k++)
                    if (in->lo[k] < n->lo[k] || in->hi[k] > n->hi[k]) bad = 2;
            }
            continue;
        }
        for (uint32_t s = n->first; s < n->first + n->count; s++) {
            if (s >= count) { bad = 3; break; }
            seen[s]++;
            for (int k = 0; k < 3; k++) {
                const float *p = positions + (size_t)indices[s * 3 + k] * 3;
                for (int a = 0; a < 3; a++)
                    if (p[a] < n->lo[a] || p[a] > n->hi[a]) bad = 4;
            }
        }
    }
    for (size_t t = 0; t < count && !bad; t++) if (seen[t] != 1) bad = 5;
    free(seen);
    return bad;
}

/* --------------------------------------------------------- traversal (f32) */

typedef struct { float x, y, z; } v3;
static v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static v3 vmul(v3 a, v3 b) { v3 r = {a.x * b.x, a.y * b.y, a.z * b.z}; return r; }
static float vdot(v3 a, v3 b) { float s = a.x * b.x; s = s + a.y * b.y; s = s + a.z * b.z; return s; }
static v3 vcross(v3 a, v3 b) {
    v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
    return r;
}
static v3 point_at(const float *positions, uint32_t index) {
    v3 r = {positions[(size_t)index * 3], positions[(size_t)index * 3 + 1], positions[(size_t)index * 3 + 2]};
    return r;
}

/* bvh.glsl:17-25 */
static float box_entry(const float lo[3], const float hi[3], v3 o, v3 inv, float t_max) {
    v3 l = {lo[0], lo[1], lo[2]}, h = {hi[0], hi[1], hi[2]};
    v3 a = vmul(vsub(l, o), inv);
    v3 b = vmul(vsub(h, o), inv);
    float nx = fmin32(a.x, b.x), ny = fmin32(a.y, b.y), nz = fmin32(a.z, b.z);
    float fx = fmax32(a.x, b.x), fy = fmax32(a.y, b.y), fz = fmax32(a.z, b.z);
    float enter = fmax32(fmax32(nx, ny), fmax32(nz, 0.0f));
    float leave = fmin32(fmin32(fx, fy), fmin32(fz, t_max));
    return enter <= leave ? enter : -1.0f;
}

typedef struct { float t; uint32_t slot; float u, v; } or_hit;

/* bvh.glsl:27-47 */
static int triangle_hit(const float *positions, const uint32_t *indices, uint32_t tri, v3 o, v3 d, or_hit *hit) {
    v3 p0 = point_at(positions, indices[tri * 3]);
    v3 e1 = vsub(point_at(positions, indices[tri * 3 + 1]), p0);
    v3 e2 = vsub(point_at(positions, indices[tri * 3 + 2]), p0);
    v3 p = vcross(d, e2);
    float det = vdot(e1, p);
    if (fabsf(det) < 1e-20f) return 0;
    float inv = 1.0f / det;
    v3 s = vsub(o, p0);
    float u = vdot(s, p) * inv;
    if (u < 0.0f || u > 1.0f) return 0;
    v3 q = vcross(s, e1);
    float v = vdot(d, q) * inv;
    if (v < 0.0f || u + v > 1.0f) return 0;
    float t = vdot(e2, q) * inv;
    if (t <= 0.0f || t >= hit->t) return 0;
    hit->t = t; hit->slot = tri; hit->u = u; hit->v = v;
    return 1;
}

static float safe_inverse(float d) {
    float s = d;
    if (fabsf(d) < 1e-20f) {
        float sg = d > 0.0f ? 1.0f : (d < 0.0f ? -1.0f : 0.0f);
        s = sg * 1e-20f + (d == 0.0f ? 1.0f : 0.0f) * 1e-20f;
    }
    return 1.0f / s;
}

#define OR_STACK 32

/* bvh.glsl:51-90. counters (optional): [0] interior nodes visited, [1] leaf
 * nodes visited, [2] triangles tested, [3] stack entries dropped. */
or_hit or_trace_closest(const or_node *nodes, const float *positions, const uint32_t *indices,
                        const float origin[3], const float direction[3], float t_max, uint64_t *counters) {
    or_hit hit = {t_max, 0xffffffffu, 0.0f, 0.0f};
    v3 o = {origin[0], origin[1], origin[2]}, d = {direction[0], direction[1], direction[2]};
    v3 inv = {safe_inverse(d.x), safe_inverse(d.y), safe_inverse(d.z)};
    uint32_t stack[OR_STACK];
    uint32_t depth = 0, node = 0;
    if (nodes[0].lo[0] > nodes[0].hi[0] || box_entry(nodes[0].lo, nodes[0].hi, o, inv, hit.t) < 0.0f) return hit;
    for (;;) {
        const or_node *cur = &nodes[node];
        if (cur->count > 0) {
            if (counters) { counters[1]++; counters[2] += cur->count; }
            for (uint32_t i = 0; i < cur->count; i++) triangle_hit(positions, indices, cur->first + i, o, d, &hit);
        } else {
            if (counters) counters[0]++;
            uint32_t left = cur->first, right = left + 1;
            float a = box_entry(nodes[left].lo, nodes[left].hi, o, inv, hit.t);
            float b = box_entry(nodes[right].lo, nodes[right].hi, o, inv, hit.t);
            int tl = a >= 0.0f, tr = b >= 0.0f;
            if (tl && tr) {
                uint32_t near = a <= b ? left : right;
                uint32_t far = a <= b ? right : left;
                if (depth < OR_STACK) stack[depth++] = far;
                else if (counters) counters[3]++;
                node = near;
                continue;
            }
            if (tl) { node = left; continue; }
            if (tr) { node = right; continue; }
        }
        if (depth == 0) break;
        node = stack[--depth];
    }
    return hit;
}

/* bvh.glsl:93-122 */
int or_trace_any(const or_node *nodes, const float *positions, const uint32_t *indices,
                 const float origin[3], const float direction[3], float t_max) {
    or_hit hit = {t_max, 0xffffffffu, 0.0f, 0.0f};
    v3 o = {origin[0], origin[1], origin[2]}, d = {direction[0], direction[1], direction[2]};
    v3 inv = {safe_inverse(d.x), safe_inverse(d.y), safe_inverse(d.z)};
    uint32_t stack[OR_STACK];
    uint32_t depth = 0, node = 0;
    if (nodes[0].lo[0] > nodes[0].hi[0] || box_entry(nodes[0].lo, nodes[0].hi, o, inv, t_max) < 0.0f) return 0;
    for (;;) {
        const or_node *cur = &nodes[node];
        if (cur->count > 0) {
            for (uint32_t i = 0; i < cur->count; i++)
                if (triangle_hit(positions, indices, cur->first + i, o, d, &hit)) return 1;
        } else {
            uint32_t left = cur->first;
            int tl = box_entry(nodes[left].lo, nodes[left].hi, o, inv, t_max) >= 0.0f;
            int tr = box_entry(nodes[left + 1].lo, nodes[left + 1].hi, o, inv, t_max) >= 0.0f;
            if (tl && tr) {
                if (depth < OR_STACK) stack[depth++] = left + 1;
                node = left;
                continue;
            }
            if (tl) { node = left; continue; }
            if (tr) { node = left + 1; continue; }
        }
        if (depth == 0) break;
        node = stack[--depth];
    }
    return 0;
}

/* probe.comp:21-30: the pixel-centre ray of the 16-float frame. */
static void probe_ray(const float f[16], uint32_t x, uint32_t y, float o[3], float d[3]) {
    float fx = (float)x + 0.5f, fy = (float)y + 0.5f;
    v3 film;
    film.x = f[4] + fx * f[8]; film.x = film.x + fy * f[12];
    film.y = f[5] + fx * f[9]; film.y = film.y + fy * f[13];
    film.z = f[6] + fx * f[10]; film.z = film.z + fy * f[14];
    int ortho = f[3] > 0.5f;
    if (ortho) {
        o[0] = f[0] + film.x; o[1] = f[1] + film.y; o[2] = f[2] + film.z;
        v3 du = {f[8], f[9], f[10]}, dv = {f[12], f[13], f[14]};
        v3 fw = vcross(du, dv);
        float l = sqrtf(vdot(fw, fw));
        d[0] = fw.x / l; d[1] = fw.y / l; d[2] = fw.z / l;
    } else {
        o[0] = f[0]; o[1] = f[1]; o[2] = f[2];
        d[0] = film.x; d[1] = film.y; d[2] = film.z;
    }
}

/* probe.comp:21-34 over every pixel: face (0xffffffff on a miss) and t bits. */
void or_probe(const or_node *nodes, const float *positions, const uint32_t *indices, const uint32_t *faces,
              const float frame[16], uint32_t width, uint32_t height,
              uint32_t *face_out, float *t_out, uint32_t *slot_out, uint64_t counters[4]) {
    for (uint32_t y = 0; y < height; y++)
        for (uint32_t x = 0; x < width; x++) {
            float o[3], d[3];
            probe_ray(frame, x, y, o, d);
            or_hit h = or_trace_closest(nodes, positions, indices, o, d, 3.0e38f, counters);
            size_t p = (size_t)y * width + x;
            face_out[p] = h.slot == 0xffffffffu ? 0xffffffffu : faces[h.slot];
            t_out[p] = h.t;
            if (slot_out) slot_out[p] = h.slot;
        }
}

/* The f32 mirror over explicit rays (origin xyz, direction xyz per ray). */
void or_trace_rays(const or_node *nodes, const float *positions, const uint32_t *indices, const uint32_t *faces,
                   const float *rays, size_t n, float t_max, uint32_t *face_out, float *t_out, uint64_t counters[4]) {
    for (size_t i = 0; i < n; i++) {
        or_hit h = or_trace_closest(nodes, positions, indices, rays + i * 6, rays + i * 6 + 3, t_max, counters);
        face_out[i] = h.slot == 0xffffffffu ? 0xffffffffu : faces[h.slot];
        t_out[i] = h.t;
    }
}

/* ------------------------------------------------------ f64 brute force */

typedef struct { double x, y, z; } d3;
static d3 dsub(d3 a, d3 b) { d3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static double ddot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 dcross(d3 a, d3 b) { d3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; return r; }

/* tests/gpu/main.lucb:105-119 */
static double hit64(const float *p, const uint32_t *c, d3 o, d3 d) {
    d3 a = {p[(size_t)c[0] * 3], p[(size_t)c[0] * 3 + 1], p[(size_t)c[0] * 3 + 2]};
    d3 b = {p[(size_t)c[1] * 3], p[(size_t)c[1] * 3 + 1], p[(size_t)c[1] * 3 + 2]};
    d3 cc = {p[(size_t)c[2] * 3], p[(size_t)c[2] * 3 + 1], p[(size_t)c[2] * 3 + 2]};
    d3 e1 = dsub(b, a), e2 = dsub(cc, a);
    d3 q = dcross(d, e2);
    double det = ddot(e1, q);
    if (fabs(det) < 1e-30) return -1.0;
    d3 s = dsub(o, a);
    double u = ddot(s, q) / det;
    if (u < 0.0 || u > 1.0) return -1.0;
    d3 r = dcross(s, e1);
    double v = ddot(d, r) / det;
    if (v < 0.0 || u + v > 1.0) return -1.0;
    double t = ddot(e2, r) / det;
    return t > 0.0 ? t : -1.0;
}

/* tests/gpu/main.lucb:82-98: nearest polygon (-1 none), its distance, and the
 * relative margin to the next different polygon. */
static void nearest64(const float *positions, const uint32_t *indices, const uint32_t *faces, size_t count,
                      d3 o, d3 d, int64_t *face, double *dist, double *margin) {
    double best = 1.0e300, second = 1.0e300;
    int64_t f = -1;
    for (size_t at = 0; at < count; at++) {
        double t = hit64(positions, indices + at * 3, o, d);
        if (t < 0.0) continue;
        int64_t polygon = (int64_t)faces[at];
        if (t < best) {
            if (polygon != f) second = best;
            best = t;
            f = polygon;
        } else if (t < second && polygon != f) {
            second = t;
        }
    }
    *face = f;
    *dist = best;
    *margin = f >= 0 ? (second - best) / best : 1.0;
}

typedef struct {
    const float *positions; const uint32_t *indices; const uint32_t *faces; size_t count;
    const double *frame; uint32_t width, height;
    const double *rays; size_t n;
    int64_t *face; double *dist; double *margin;
    size_t begin, end;
} bf_job;

static void *bf_run(void *arg) {
    bf_job *j = (bf_job *)arg;
    for (size_t i = j->begin; i < j->end; i++) {
        d3 o, d;
        if (j->rays) {
            const double *r = j->rays + i * 6;
            o = (d3){r[0], r[1], r[2]}; d = (d3){r[3], r[4], r[5]};
        } else {
            /* camera.lucb:51-55 at the pixel centre, in f64 */
            const double *f = j->frame;
            double x = (double)(i % j->width) + 0.5, y = (double)(i / j->width) + 0.5;
            d3 film = {f[3] + f[6] * x + f[9] * y, f[4] + f[7] * x + f[10] * y, f[5] + f[8] * x + f[11] * y};
            if (f[17] > 0.5) { o = (d3){f[0] + film.x, f[1] + film.y, f[2] + film.z}; d = (d3){f[12], f[13], f[14]}; }
            else { o = (d3){f[0], f[1], f[2]}; d = film; }
        }
        nearest64(j->positions, j->indices, j->faces, j->count, o, d, &j->face[i], &j->dist[i], &j->margin[i]);
    }
    return NULL;
}

/* Brute force over every pixel of a frame (rays == NULL) or over n explicit
 * f64 rays (origin, direction), split over `threads` threads by rows. */
void or_brute_force(const float *positions, const uint32_t *indices, const uint32_t *faces, size_t count,
                    const double frame[18], uint32_t width, uint32_t height,
                    const double *rays, size_t n, int threads,
                    int64_t *face, double *dist, double *margin) {
    size_t total = rays ? n : (size_t)width * height;
    if (threads < 1) threads = 1;
    pthread_t th[256];
    bf_job jobs[256];
    if (threads > 256) threads = 256;
    for (int k = 0; k < threads; k++) {
        jobs[k] = (bf_job){positions, indices, faces, count, frame, width, height, rays, n, face, dist, margin,
                           total * (size_t)k / (size_t)threads, total * (size_t)(k + 1) / (size_t)threads};
        pthread_create(&th[k], NULL, bf_run, &jobs[k]);
    }
    for (int k = 0; k < threads; k++) pthread_join(th[k], NULL);
}
