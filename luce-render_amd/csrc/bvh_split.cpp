// Triangle references with tighter boxes, for scenes whose triangles overlap
// each other's boxes by the hundreds (dense random-triangle clouds): each
// triangle whose box is longer than `min_extent` on some axis is cut at the
// middle of that axis, the two pieces of the triangle (clipped polygons) cut
// again while they are long, up to `max_depth` cuts deep, after Ernst and
// Greiner's early split clipping (2007). Every piece keeps a box that holds
// it with a margin of a few ulps (so the slab test finds a hit on a cut in at
// least one piece) and never leaves the triangle's own box. The references
// are built over by their boxes (lr_bvh_build_boxes); each leaf slot then
// holds its triangle. No reference counterpart: the reference builds over
// whole triangles (bvh.lucb:66-81), which is what lr_bvh_build keeps.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

#include "lr_internal.h"

namespace {

struct Poly {
    double v[12][3];
    int n = 0;
};

struct Box {
    double lo[3], hi[3];
};

Box box_of(const Poly &p) {
    Box b;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = b.hi[a] = p.v[0][a];
        for (int k = 1; k < p.n; k++) {
            b.lo[a] = std::min(b.lo[a], p.v[k][a]);
            b.hi[a] = std::max(b.hi[a], p.v[k][a]);
        }
    }
    return b;
}

// The part of `p` on one side of the plane x[axis] = m (keep_low: x <= m).
Poly clip(const Poly &p, int axis, double m, bool keep_low) {
    Poly out;
    for (int k = 0; k < p.n; k++) {
        const double *a = p.v[k], *b = p.v[(k + 1) % p.n];
        bool ia = keep_low ? a[axis] <= m : a[axis] >= m;
        bool ib = keep_low ? b[axis] <= m : b[axis] >= m;
        if (ia && out.n < 12) {
            std::memcpy(out.v[out.n], a, sizeof(out.v[0]));
            out.n++;
        }
        if (ia != ib && out.n < 12) {
            double t = (m - a[axis]) / (b[axis] - a[axis]);
            for (int c = 0; c < 3; c++) out.v[out.n][c] = a[c] + t * (b[c] - a[c]);
            out.v[out.n][axis] = m;
            out.n++;
        }
    }
    return out;
}

// A float box around [lo, hi] grown by a few ulps of its magnitude, kept
// inside the triangle's own float box [tlo, thi].
void emit(const Box &b, const float tlo[3], const float thi[3], float out[6]) {
    for (int a = 0; a < 3; a++) {
        double mag = std::max(1e-30, std::max(std::fabs(b.lo[a]), std::fabs(b.hi[a])));
        double margin = mag * 1e-6;
        float lo = (float)(b.lo[a] - margin), hi = (float)(b.hi[a] + margin);
        if ((double)lo > b.lo[a] - margin) lo = std::nextafter(lo, -INFINITY);
        if ((double)hi < b.hi[a] + margin) hi = std::nextafter(hi, INFINITY);
        out[a] = std::max(lo, tlo[a]);
        out[3 + a] = std::min(hi, thi[a]);
    }
}

// The references of triangle t, appended in depth-first order (low side
// first).
void split_one(const float *positions, const uint32_t *indices, size_t t, bool splittable, double min_extent,
               uint32_t max_depth, std::vector<float> &boxes, std::vector<uint32_t> &tris) {
    float tlo[3], thi[3];
    Poly p;
    p.n = 3;
    for (int k = 0; k < 3; k++)
        for (int a = 0; a < 3; a++) p.v[k][a] = positions[(size_t)indices[t * 3 + k] * 3 + a];
    for (int a = 0; a < 3; a++) {
        float x0 = positions[(size_t)indices[t * 3] * 3 + a], x1 = positions[(size_t)indices[t * 3 + 1] * 3 + a],
              x2 = positions[(size_t)indices[t * 3 + 2] * 3 + a];
        tlo[a] = std::min(x0, std::min(x1, x2));
        thi[a] = std::max(x0, std::max(x1, x2));
    }
    struct Item {
        Poly p;
        uint32_t depth;
    };
    Item stack[40];
    int top = 0;
    stack[top++] = {p, 0u};
    while (top > 0) {
        Item it = stack[--top];
        Box b = box_of(it.p);
        int axis = 0;
        for (int a = 1; a < 3; a++)
            if (b.hi[a] - b.lo[a] > b.hi[axis] - b.lo[axis]) axis = a;
        double extent = b.hi[axis] - b.lo[axis];
        if (!splittable || it.depth >= max_depth || !(extent > min_extent) || top + 2 > 40) {
            float out[6];
            if (it.depth == 0) {
                std::memcpy(out, tlo, sizeof(tlo));
                std::memcpy(out + 3, thi, sizeof(thi));
            } else {
                emit(b, tlo, thi, out);
            }
            boxes.insert(boxes.end(), out, out + 6);
            tris.push_back((uint32_t)t);
            continue;
        }
        double m = 0.5 * (b.lo[axis] + b.hi[axis]);
        Poly hi = clip(it.p, axis, m, false), lo = clip(it.p, axis, m, true);
        if (hi.n >= 3) stack[top++] = {hi, it.depth + 1};
        if (lo.n >= 3) stack[top++] = {lo, it.depth + 1};
    }
}

}  // namespace

extern "C" {

int lr_bvh_split_triangles(const float *positions, size_t points, const uint32_t *indices, size_t triangles,
                           const uint8_t *splittable, float min_extent, uint32_t max_depth, int threads,
                           float *boxes, uint32_t *tris, size_t capacity, size_t *count) {
    if (!count || (triangles && (!positions || !indices)))
        return lr::fail(LR_INVALID, "lr_bvh_split_triangles needs its arguments");
    if (max_depth > 8) return lr::fail(LR_INVALID, "at most 8 cuts deep");
    for (size_t i = 0; i < triangles * 3; i++)
        if (indices[i] >= points) return lr::fail(LR_INVALID, "a triangle names a point past the positions");
    unsigned n_threads = threads > 0 ? (unsigned)threads : std::max(1u, std::thread::hardware_concurrency());
    n_threads = (unsigned)std::min<size_t>(n_threads, std::max<size_t>(1, triangles / 4096));
    std::vector<std::vector<float>> part_boxes(n_threads);
    std::vector<std::vector<uint32_t>> part_tris(n_threads);
    auto work = [&](unsigned w) {
        size_t a = triangles * w / n_threads, b = triangles * (w + 1) / n_threads;
        for (size_t t = a; t < b; t++)
            split_one(positions, indices, t, !splittable || splittable[t], (double)min_extent, max_depth,
                      part_boxes[w], part_tris[w]);
    };
    std::vector<std::thread> pool;
    for (unsigned w = 1; w < n_threads; w++) pool.emplace_back(work, w);
    work(0);
    for (auto &th : pool) th.join();
    size_t total = 0;
    for (auto &v : part_tris) total += v.size();
    *count = total;
    if (!boxes || !tris) return 0;  // the count alone
    if (capacity < total) return lr::fail(LR_LIMIT, "the reference arrays are shorter than the references");
    size_t at = 0;
    for (unsigned w = 0; w < n_threads; w++) {
        std::memcpy(boxes + at * 6, part_boxes[w].data(), part_boxes[w].size() * sizeof(float));
        std::memcpy(tris + at, part_tris[w].data(), part_tris[w].size() * sizeof(uint32_t));
        at += part_tris[w].size();
    }
    return 0;
}

}  // extern "C"
