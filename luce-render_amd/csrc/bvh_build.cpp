// Host BVH build: the binned-SAH binary tree of src/render/bvh.lucb:66-217,
// built in parallel.
//
// The reference builds single-threaded from an explicit stack, allocating each
// split node's two children as a pair at the end of the node array. Its tree
// depends only on the split rules and the in-place partition order, not on the
// order in which independent subtrees are processed, so this builder runs
// subtrees as tasks on a small thread pool (binning of very large nodes is
// also split across threads: bin counts and boxes merge exactly, min/max and
// integer sums being associative) and afterwards numbers the nodes in the
// reference's order: a pre-order walk in which a node's child pair takes the
// next two indices when the node is visited, left subtree before right.
//
// Arithmetic is f32 without contraction (build flags carry -ffp-contract=off),
// as the reference's math32 boxes: areas 2(xy + yz + zx), costs
// 1 + (A_L N_L + A_R N_R) / A, bins floor(12 (c - lo) / extent).
#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>

#include "lr_internal.h"

namespace lr {
namespace {

constexpr int kBins = 12;         // bvh.lucb:29
constexpr size_t kLeafSize = 4;   // bvh.lucb:30
constexpr size_t kLeafLimit = 16; // bvh.lucb:32
constexpr size_t kTaskGrain = 4096;      // subtrees at least this big become tasks
constexpr size_t kParallelBinning = 1u << 20;

struct Box {
    float lo[3], hi[3];
};

inline Box empty_box() {
    return Box{{3.0e38f, 3.0e38f, 3.0e38f}, {-3.0e38f, -3.0e38f, -3.0e38f}};
}
inline float mn(float a, float b) { return a < b ? a : b; }
inline float mx(float a, float b) { return a > b ? a : b; }
inline void grow(Box &a, const Box &b) {
    for (int k = 0; k < 3; k++) {
        a.lo[k] = mn(a.lo[k], b.lo[k]);
        a.hi[k] = mx(a.hi[k], b.hi[k]);
    }
}
inline void grow_point(Box &a, const float *p) {
    for (int k = 0; k < 3; k++) {
        a.lo[k] = mn(a.lo[k], p[k]);
        a.hi[k] = mx(a.hi[k], p[k]);
    }
}
inline float area(const Box &b) {
    float x = b.hi[0] - b.lo[0], y = b.hi[1] - b.lo[1], z = b.hi[2] - b.lo[2];
    if (x < 0.0f || y < 0.0f || z < 0.0f) return 0.0f;
    float s = x * y + y * z;
    s = s + z * x;
    return 2.0f * s;
}

struct Centers {
    std::vector<float> xyz;  // 3 per primitive
    const float *at(uint32_t i) const { return &xyz[(size_t)i * 3]; }
};

inline int bin_of(const float *c, const Box &middle, int axis) {
    float lo = middle.lo[axis];
    float extent = middle.hi[axis] - lo;
    if (extent <= 0.0f) return 0;
    float q = ((float)kBins * (c[axis] - lo)) / extent;
    int64_t at = (int64_t)q;
    if (at < 0) return 0;
    return at < kBins ? (int)at : kBins - 1;
}

struct Bins {
    size_t count[3][kBins];
    Box bounds[3][kBins];
    void clear() {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < kBins; b++) {
                count[a][b] = 0;
                bounds[a][b] = empty_box();
            }
    }
    void merge(const Bins &o) {
        for (int a = 0; a < 3; a++)
            for (int b = 0; b < kBins; b++) {
                count[a][b] += o.count[a][b];
                grow(bounds[a][b], o.bounds[a][b]);
            }
    }
};

struct TmpNode {
    Box box;
    uint32_t left = 0xffffffffu;  // tmp index of the left child (right = left + 1)
    uint32_t first = 0, count = 0;
};

class Builder {
  public:
    Builder(const Box *boxes, const Centers &centers, uint32_t *order, size_t n, int threads)
        : boxes_(boxes), centers_(centers), order_(order), threads_(threads) {
        tmp_.resize(std::max<size_t>(1, 2 * n));
        next_.store(1);
    }

    void run(size_t n) {
        push_task(0, 0, n);
        std::vector<std::thread> pool;
        for (int t = 0; t < threads_; t++) pool.emplace_back([this] { worker(); });
        for (auto &th : pool) th.join();
    }

    const std::vector<TmpNode> &nodes() const { return tmp_; }
    size_t used() const { return next_.load(); }

  private:
    struct Task {
        uint32_t node;
        size_t first, count;
    };

    void push_task(uint32_t node, size_t first, size_t count) {
        {
            std::lock_guard<std::mutex> lock(mutex_);
            tasks_.push_back(Task{node, first, count});
            outstanding_++;
        }
        cv_.notify_one();
    }

    void worker() {
        for (;;) {
            Task task;
            {
                std::unique_lock<std::mutex> lock(mutex_);
                cv_.wait(lock, [this] { return !tasks_.empty() || outstanding_ == 0; });
                if (tasks_.empty()) return;
                task = tasks_.back();
                tasks_.pop_back();
            }
            build(task.node, task.first, task.count);
            {
                std::lock_guard<std::mutex> lock(mutex_);
                outstanding_--;
                if (outstanding_ == 0) cv_.notify_all();
            }
        }
    }

    void bin_range(Bins &bins, const uint32_t *order, size_t n, const Box &middle) const {
        for (size_t i = 0; i < n; i++) {
            uint32_t p = order[i];
            const float *c = centers_.at(p);
            for (int axis = 0; axis < 3; axis++) {
                int b = bin_of(c, middle, axis);
                bins.count[axis][b]++;
                grow(bins.bounds[axis][b], boxes_[p]);
            }
        }
    }

    void bounds_range(Box &bounds, Box &middle, const uint32_t *order, size_t n) const {
        for (size_t i = 0; i < n; i++) {
            uint32_t p = order[i];
            grow(bounds, boxes_[p]);
            grow_point(middle, centers_.at(p));
        }
    }

    template <class F> void parallel_chunks(size_t n, F &&f) const {
        int k = threads_;
        std::vector<std::thread> th;
        for (int t = 0; t < k; t++) th.emplace_back([&, t] { f((size_t)t, n * t / k, n * (t + 1) / k); });
        for (auto &x : th) x.join();
    }

    // best_split (bvh.lucb:147-182) on gathered bins.
    void best_split(const Bins &bins, const Box &parent, const Box &middle, int &axis_out, int &bin_out,
                    float &cost_out) const {
        axis_out = -1;
        bin_out = 0;
        cost_out = 3.0e38f;
        float parent_area = area(parent);
        if (parent_area <= 0.0f) return;
        for (int axis = 0; axis < 3; axis++) {
            if (middle.hi[axis] - middle.lo[axis] <= 0.0f) continue;
            float right_area[kBins];
            size_t right_count[kBins];
            Box sweep = empty_box();
            size_t total = 0;
            for (int b = kBins - 1; b > 0; b--) {
                grow(sweep, bins.bounds[axis][b]);
                total += bins.count[axis][b];
                right_area[b] = area(sweep);
                right_count[b] = total;
            }
            Box left = empty_box();
            size_t left_count = 0;
            for (int s = 1; s < kBins; s++) {
                grow(left, bins.bounds[axis][s - 1]);
                left_count += bins.count[axis][s - 1];
                if (left_count == 0 || right_count[s] == 0) continue;
                float l = area(left) * (float)left_count;
                float r = right_area[s] * (float)right_count[s];
                float cost = 1.0f + (l + r) / parent_area;
                if (cost < cost_out) {
                    axis_out = axis;
                    bin_out = s;
                    cost_out = cost;
                }
            }
        }
    }

    // partition (bvh.lucb:185-196): the reference's in-place swap order.
    size_t partition(uint32_t *order, size_t n, const Box &middle, int axis, int bin) const {
        size_t low = 0, high = n;
        while (low < high) {
            if (bin_of(centers_.at(order[low]), middle, axis) < bin) {
                low++;
            } else {
                high--;
                std::swap(order[low], order[high]);
            }
        }
        return low;
    }

    void build(uint32_t node, size_t first, size_t count) {
        for (;;) {
            uint32_t *order = order_ + first;
            Box bounds = empty_box(), middle = empty_box();
            bool wide = count >= kParallelBinning && threads_ > 1;
            if (wide) {
                std::vector<Box> pb(threads_, empty_box()), pm(threads_, empty_box());
                parallel_chunks(count, [&](size_t t, size_t b, size_t e) { bounds_range(pb[t], pm[t], order + b, e - b); });
                for (int t = 0; t < threads_; t++) {
                    grow(bounds, pb[t]);
                    grow(middle, pm[t]);
                }
            } else {
                bounds_range(bounds, middle, order, count);
            }
            size_t left_count = 0;
            int axis = -1, bin = 0;
            float cost = 3.0e38f;
            if (count > kLeafSize) {
                Bins bins;
                bins.clear();
                if (wide) {
                    std::vector<Bins> pbins(threads_);
                    parallel_chunks(count, [&](size_t t, size_t b, size_t e) {
                        pbins[t].clear();
                        bin_range(pbins[t], order + b, e - b, middle);
                    });
                    for (int t = 0; t < threads_; t++) bins.merge(pbins[t]);
                } else {
                    bin_range(bins, order, count, middle);
                }
                best_split(bins, bounds, middle, axis, bin, cost);
            }
            if (count > kLeafSize && (cost < (float)count || count > kLeafLimit)) {
                if (axis >= 0 && cost < 3.0e38f) left_count = partition(order, count, middle, axis, bin);
                if (left_count == 0 || left_count == count) left_count = count / 2;
            }
            TmpNode &t = tmp_[node];
            t.box = bounds;
            if (left_count == 0) {
                t.first = (uint32_t)first;
                t.count = (uint32_t)count;
                return;
            }
            uint32_t child = (uint32_t)next_.fetch_add(2);
            t.left = child;
            t.count = 0;
            size_t right_first = first + left_count, right_count = count - left_count;
            // The bigger side continues here; a big other side becomes a task.
            if (right_count >= kTaskGrain && threads_ > 1) {
                push_task(child + 1, right_first, right_count);
                node = child;
                count = left_count;
                continue;
            }
            build(child + 1, right_first, right_count);
            node = child;
            count = left_count;
        }
    }

    const Box *boxes_;
    const Centers &centers_;
    uint32_t *order_;
    int threads_;
    std::vector<TmpNode> tmp_;
    std::atomic<size_t> next_{1};
    std::mutex mutex_;
    std::condition_variable cv_;
    std::deque<Task> tasks_;
    size_t outstanding_ = 0;
};

}  // namespace

namespace {

// build_tree (bvh.lucb:85-133) over primitive boxes: `box_of(t)` gives
// primitive t's box; centres are the boxes' midpoints.
template <class BoxOf>
size_t build_over_boxes(size_t n, int threads, lr_bvh_node *nodes, uint32_t *order, BoxOf box_of) {
    if (n == 0) {
        nodes[0] = lr_bvh_node{{1.0f, 1.0f, 1.0f}, 0, {-1.0f, -1.0f, -1.0f}, 0};
        return 0 + 1;
    }
    if (threads <= 0) threads = (int)std::max(1u, std::thread::hardware_concurrency());
    threads = std::min(threads, 64);
    std::vector<Box> boxes(n);
    Centers centers;
    centers.xyz.resize(n * 3);
    auto prep = [&](size_t b, size_t e) {
        for (size_t t = b; t < e; t++) {
            Box box = box_of(t);
            boxes[t] = box;
            for (int k = 0; k < 3; k++) centers.xyz[t * 3 + k] = 0.5f * (box.lo[k] + box.hi[k]);
            order[t] = (uint32_t)t;
        }
    };
    {
        std::vector<std::thread> th;
        for (int t = 0; t < threads; t++) th.emplace_back(prep, n * t / threads, n * (t + 1) / threads);
        for (auto &x : th) x.join();
    }
    Builder builder(boxes.data(), centers, order, n, threads);
    builder.run(n);
    const std::vector<TmpNode> &tmp = builder.nodes();
    // Number in the reference's order: visiting a split node gives its pair
    // the next two indices; the left subtree is finished before the right.
    size_t next = 1;
    struct Item {
        uint32_t tmp;
        size_t index;
    };
    std::vector<Item> stack;
    stack.push_back(Item{0, 0});
    while (!stack.empty()) {
        Item it = stack.back();
        stack.pop_back();
        const TmpNode &t = tmp[it.tmp];
        lr_bvh_node &out = nodes[it.index];
        std::memcpy(out.lo, t.box.lo, 12);
        std::memcpy(out.hi, t.box.hi, 12);
        if (t.left == 0xffffffffu) {
            out.first = t.first;
            out.count = t.count;
            continue;
        }
        size_t child = next;
        next += 2;
        out.first = (uint32_t)child;
        out.count = 0;
        stack.push_back(Item{t.left + 1, child + 1});
        stack.push_back(Item{t.left, child});
    }
    return next;
}

}  // namespace

size_t build_binary_bvh(const float *positions, const uint32_t *indices, size_t n, int threads, lr_bvh_node *nodes,
                        uint32_t *order) {
    return build_over_boxes(n, threads, nodes, order, [&](size_t t) {
        Box box = empty_box();
        for (int k = 0; k < 3; k++) grow_point(box, positions + (size_t)indices[t * 3 + k] * 3);
        return box;
    });
}

size_t build_box_bvh(const float *boxes, size_t n, int threads, lr_bvh_node *nodes, uint32_t *order) {
    return build_over_boxes(n, threads, nodes, order, [&](size_t t) {
        Box box;
        for (int k = 0; k < 3; k++) {
            box.lo[k] = boxes[t * 6 + k];
            box.hi[k] = boxes[t * 6 + 3 + k];
        }
        return box;
    });
}

}  // namespace lr
