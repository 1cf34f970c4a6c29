// bvh.cpp — see bvh.hpp.
#include "bvh.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>

namespace pt {

namespace {

struct Node2 {
    Aabb box;
    int left = -1, right = -1;  // children (binary), -1 for a leaf
    int first = 0, count = 0;   // leaf range in the permuted primitive order
};

float area(const Aabb& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
    return 2.0f * (dx * dy + dy * dz + dz * dx);
}
Aabb empty_box() {
    Aabb b;
    for (int k = 0; k < 3; ++k) {
        b.lo[k] = std::numeric_limits<float>::infinity();
        b.hi[k] = -std::numeric_limits<float>::infinity();
    }
    return b;
}
void grow(Aabb& a, const Aabb& b) {
    for (int k = 0; k < 3; ++k) {
        a.lo[k] = std::min(a.lo[k], b.lo[k]);
        a.hi[k] = std::max(a.hi[k], b.hi[k]);
    }
}

struct Builder {
    const std::vector<Aabb>& prims;
    std::vector<int> order;
    std::vector<Node2> nodes;
    int max_leaf;

    explicit Builder(const std::vector<Aabb>& p, int ml) : prims(p), max_leaf(ml) {
        order.resize(p.size());
        for (size_t i = 0; i < p.size(); ++i) order[i] = (int)i;
    }
    float centroid(int i, int axis) const { return 0.5f * (prims[(size_t)i].lo[axis] + prims[(size_t)i].hi[axis]); }
    // SAH bin of a centroid, clamped to [0, kBins-1]; a NaN offset (non-finite
    // boxes, which callers reject before building) lands in bin 0 rather than
    // in an out-of-range int conversion.
    template <int kBins>
    static int bin_of(float c, float lo, float scale) {
        const float f = (c - lo) * scale;
        if (!(f > 0.0f)) return 0;
        if (f >= (float)(kBins - 1)) return kBins - 1;
        return (int)f;
    }

    int build(int first, int count) {
        int id = (int)nodes.size();
        nodes.emplace_back();
        Aabb box = empty_box(), cbox = empty_box();
        for (int i = first; i < first + count; ++i) {
            grow(box, prims[(size_t)order[(size_t)i]]);
            Aabb c;
            for (int k = 0; k < 3; ++k) c.lo[k] = c.hi[k] = centroid(order[(size_t)i], k);
            grow(cbox, c);
        }
        nodes[(size_t)id].box = box;
        nodes[(size_t)id].first = first;
        nodes[(size_t)id].count = count;
        if (count <= 1) return id;

        // Binned SAH over the three axes.
        constexpr int kBins = 16;
        float best_cost = std::numeric_limits<float>::infinity();
        int best_axis = -1, best_split = 0;
        for (int axis = 0; axis < 3; ++axis) {
            float lo = cbox.lo[axis], hi = cbox.hi[axis];
            if (!(hi > lo)) continue;
            float scale = kBins / (hi - lo);
            Aabb bin_box[kBins];
            int bin_n[kBins] = {0};
            for (int b = 0; b < kBins; ++b) bin_box[b] = empty_box();
            for (int i = first; i < first + count; ++i) {
                int p = order[(size_t)i];
                int b = bin_of<kBins>(centroid(p, axis), lo, scale);
                bin_n[b]++;
                grow(bin_box[b], prims[(size_t)p]);
            }
            float right_area[kBins];
            int right_n[kBins];
            Aabb acc = empty_box();
            int n = 0;
            for (int b = kBins - 1; b > 0; --b) {
                grow(acc, bin_box[b]);
                n += bin_n[b];
                right_area[b] = area(acc);
                right_n[b] = n;
            }
            acc = empty_box();
            n = 0;
            for (int b = 0; b < kBins - 1; ++b) {
                grow(acc, bin_box[b]);
                n += bin_n[b];
                if (n == 0 || right_n[b + 1] == 0) continue;
                float cost = area(acc) * (float)n + right_area[b + 1] * (float)right_n[b + 1];
                if (cost < best_cost) {
                    best_cost = cost;
                    best_axis = axis;
                    best_split = b + 1;
                }
            }
        }
        const float leaf_cost = area(box) * (float)count;
        // SAH with unit traversal cost against the intersection cost of the leaf.
        bool make_leaf = count <= max_leaf && (best_axis < 0 || best_cost + area(box) >= leaf_cost);
        if (make_leaf) return id;

        int mid;
        if (best_axis < 0) {
            // All centroids coincide: split the list in half.
            mid = first + count / 2;
        } else {
            float lo = cbox.lo[best_axis], scale = kBins / (cbox.hi[best_axis] - lo);
            auto it = std::partition(order.begin() + first, order.begin() + first + count, [&](int p) {
                int b = bin_of<kBins>(centroid(p, best_axis), lo, scale);
                return b < best_split;
            });
            mid = (int)(it - order.begin());
            if (mid == first || mid == first + count) mid = first + count / 2;
        }
        int l = build(first, mid - first);
        int r = build(mid, first + count - mid);
        nodes[(size_t)id].left = l;
        nodes[(size_t)id].right = r;
        return id;
    }
};

}  // namespace

WideBvh build_wide_bvh(const std::vector<Aabb>& prims, const std::vector<uint32_t>& refs, int width, int max_leaf) {
    WideBvh out;
    out.width = width;
    const int W = width;
    const size_t node_words = (size_t)8 * W;
    auto new_node = [&]() {
        size_t base = out.words.size();
        out.words.resize(base + node_words, 0);
        float inf = std::numeric_limits<float>::infinity(), ninf = -inf;
        for (int j = 0; j < W; ++j) {
            for (int k = 0; k < 3; ++k) {
                std::memcpy(&out.words[base + (size_t)(k * W + j)], &inf, 4);
                std::memcpy(&out.words[base + (size_t)((3 + k) * W + j)], &ninf, 4);
            }
            int32_t minus1 = -1;
            std::memcpy(&out.words[base + (size_t)(7 * W + j)], &minus1, 4);
        }
        return (int)(base / node_words);
    };
    if (prims.empty()) {
        new_node();
        out.n_nodes = 1;
        out.depth = 1;
        out.max_stack = 1;
        return out;
    }
    Builder b(prims, max_leaf);
    int root = b.build(0, (int)prims.size());
    out.prim_refs.resize(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) out.prim_refs[i] = refs[(size_t)b.order[i]];

    // Collapse: each wide node gathers up to W binary descendants.
    struct Item { int bnode; int wnode; int level; int stack_before; };
    std::vector<Item> work;
    int wroot = new_node();
    work.push_back({root, wroot, 1, 0});
    while (!work.empty()) {
        Item it = work.back();
        work.pop_back();
        const Node2& bn = b.nodes[(size_t)it.bnode];
        std::vector<int> kids;
        if (bn.left < 0) {
            kids.push_back(it.bnode);  // a leaf root: one leaf child
        } else {
            kids = {bn.left, bn.right};
            while ((int)kids.size() < W) {
                int best = -1;
                float best_a = -1.0f;
                for (size_t k = 0; k < kids.size(); ++k) {
                    const Node2& c = b.nodes[(size_t)kids[k]];
                    if (c.left >= 0 && area(c.box) > best_a) {
                        best_a = area(c.box);
                        best = (int)k;
                    }
                }
                if (best < 0) break;
                int open = kids[(size_t)best];
                kids.erase(kids.begin() + best);
                kids.push_back(b.nodes[(size_t)open].left);
                kids.push_back(b.nodes[(size_t)open].right);
            }
        }
        out.depth = std::max(out.depth, it.level);
        // Traversal pushes every hit child but the nearest: the stack can hold
        // (#children - 1) more entries below this node.
        int stack_here = it.stack_before + (int)kids.size() - 1;
        out.max_stack = std::max(out.max_stack, stack_here + 1);
        for (size_t j = 0; j < kids.size(); ++j) {
            const Node2& c = b.nodes[(size_t)kids[j]];
            size_t base = (size_t)it.wnode * node_words;
            for (int k = 0; k < 3; ++k) {
                std::memcpy(&out.words[base + (size_t)(k * W) + j], &c.box.lo[k], 4);
                std::memcpy(&out.words[base + (size_t)((3 + k) * W) + j], &c.box.hi[k], 4);
            }
            int32_t child, count;
            if (c.left < 0) {
                child = c.first;
                count = c.count;
            } else {
                int w = new_node();
                child = w;
                count = 0;
                work.push_back({kids[j], w, it.level + 1, stack_here});
            }
            std::memcpy(&out.words[(size_t)it.wnode * node_words + (size_t)(6 * W) + j], &child, 4);
            std::memcpy(&out.words[(size_t)it.wnode * node_words + (size_t)(7 * W) + j], &count, 4);
        }
    }
    out.n_nodes = (int)(out.words.size() / node_words);
    return out;
}

}  // namespace pt
