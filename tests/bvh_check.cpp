// Checks the host BVH builder (pathtracer_amd/csrc/bvh.cpp) on the CPU, on
// spheres drawn like the benchmark's random-sphere scene or in degenerate
// layouts. What the traversal kernels rely on:
//   - the leaves' ranges cover prim_refs once, and every reference is one input;
//   - every slot's box holds its leaf's primitive boxes or its child node's
//     slots (exact float containment), and empty slots are inverted boxes;
//   - leaves hold at most max_leaf primitives (at least one);
//   - a traversal that pushes every hit child but the nearest never holds more
//     than max_stack entries;
//   - that traversal finds the same closest sphere as testing all of them.
//
// Usage: bvh_check n width max_leaf layout seed
//   layout: 0 random spheres (centres in the unit cube, radius exp(U[ln 0.002,
//   ln 0.05])), 1 all coincident, 2 collinear, 3 two tight clusters.
// Prints "ok nodes=N depth=D max_stack=S max_seen=M" or the first violation.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "bvh.hpp"

using pt::Aabb;

struct Sphere {
    float c[3], r;
};

static int fail(const char* what, long a, long b) {
    std::printf("FAIL %s (%ld, %ld)\n", what, a, b);
    return 1;
}

static float word_f(const pt::WideBvh& t, size_t i) {
    float f;
    std::memcpy(&f, &t.words[i], 4);
    return f;
}
static int32_t word_i(const pt::WideBvh& t, size_t i) {
    int32_t v;
    std::memcpy(&v, &t.words[i], 4);
    return v;
}

struct Slot {
    float lo[3], hi[3];
    int child, count;
};
static Slot slot(const pt::WideBvh& t, int node, int j) {
    const int W = t.width;
    const size_t base = (size_t)node * 8 * W;
    Slot s;
    for (int k = 0; k < 3; ++k) {
        s.lo[k] = word_f(t, base + (size_t)(k * W + j));
        s.hi[k] = word_f(t, base + (size_t)((3 + k) * W + j));
    }
    s.child = word_i(t, base + (size_t)(6 * W + j));
    s.count = word_i(t, base + (size_t)(7 * W + j));
    return s;
}

static bool contains(const Slot& outer, const float lo[3], const float hi[3]) {
    for (int k = 0; k < 3; ++k)
        if (!(outer.lo[k] <= lo[k] && outer.hi[k] >= hi[k])) return false;
    return true;
}

// Entry distance of a ray into a box, padded so that rounding in the slab test
// cannot drop a sphere whose hit lies on the box's face; -1 on a miss.
static float box_entry(const Slot& s, const float o[3], const float d[3], float tmax) {
    float t0 = 0.0f, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        const float pad = 1e-5f * (1.0f + std::fabs(s.lo[k]) + std::fabs(s.hi[k]));
        const float lo = s.lo[k] - pad, hi = s.hi[k] + pad;
        if (d[k] == 0.0f) {
            if (o[k] < lo || o[k] > hi) return -1.0f;
            continue;
        }
        float a = (lo - o[k]) / d[k], b = (hi - o[k]) / d[k];
        if (a > b) std::swap(a, b);
        t0 = std::max(t0, a);
        t1 = std::min(t1, b);
        if (t0 > t1) return -1.0f;
    }
    return t0;
}

static float hit_sphere(const Sphere& s, const float o[3], const float d[3]) {
    const float oc[3] = {o[0] - s.c[0], o[1] - s.c[1], o[2] - s.c[2]};
    const float b = oc[0] * d[0] + oc[1] * d[1] + oc[2] * d[2];
    const float c = oc[0] * oc[0] + oc[1] * oc[1] + oc[2] * oc[2] - s.r * s.r;
    const float disc = b * b - c;
    if (disc < 0.0f) return -1.0f;
    const float q = std::sqrt(disc);
    float t = -b - q;
    if (t <= 1e-6f) t = -b + q;
    return t > 1e-6f ? t : -1.0f;
}

constexpr uint32_t kRefMask = 0x5A5A0000u;  // references are index ^ mask

int main(int argc, char** argv) {
    if (argc != 6) {
        std::fprintf(stderr, "usage: %s n width max_leaf layout seed\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), W = std::atoi(argv[2]), max_leaf = std::atoi(argv[3]);
    const int layout = std::atoi(argv[4]);
    std::mt19937 rng((unsigned)std::atoi(argv[5]));
    std::uniform_real_distribution<float> u01(0.0f, 1.0f);
    std::vector<Sphere> sph((size_t)n);
    for (int i = 0; i < n; ++i) {
        Sphere& s = sph[(size_t)i];
        const float lr = std::log(0.002f) + u01(rng) * (std::log(0.05f) - std::log(0.002f));
        s.r = std::exp(lr);
        if (layout == 0) {
            for (int k = 0; k < 3; ++k) s.c[k] = u01(rng);
        } else if (layout == 1) {
            s.c[0] = s.c[1] = s.c[2] = 0.5f;
            s.r = 0.01f;
        } else if (layout == 2) {
            s.c[0] = u01(rng);
            s.c[1] = s.c[2] = 0.5f;
        } else if (layout == 4) {
            // random boxes plus one whose upper corner overflows to +inf
            for (int k = 0; k < 3; ++k) s.c[k] = u01(rng);
            if (i == n / 2) {
                for (int k = 0; k < 3; ++k) s.c[k] = 3e38f;
                s.r = 1e38f;
            }
        } else {
            const float base = (i % 2) ? 0.2f : 0.8f;
            for (int k = 0; k < 3; ++k) s.c[k] = base + 1e-4f * u01(rng);
            if (i % 97 == 0) s.c[0] = u01(rng);
        }
    }
    std::vector<Aabb> boxes((size_t)n);
    std::vector<uint32_t> refs((size_t)n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            boxes[(size_t)i].lo[k] = sph[(size_t)i].c[k] - sph[(size_t)i].r;
            boxes[(size_t)i].hi[k] = sph[(size_t)i].c[k] + sph[(size_t)i].r;
        }
        refs[(size_t)i] = (uint32_t)i ^ kRefMask;  // distinct, not the index itself
    }
    const pt::WideBvh t = pt::build_wide_bvh(boxes, refs, W, max_leaf);
    if (t.width != W) return fail("width", t.width, W);
    if (t.words.size() != (size_t)t.n_nodes * 8 * W) return fail("node count", (long)t.words.size(), t.n_nodes);
    if ((int)t.prim_refs.size() != n) return fail("prim_refs size", (long)t.prim_refs.size(), n);

    // reference -> input index
    std::vector<int> of_ref((size_t)n, -1);
    for (int j = 0; j < n; ++j) {
        const uint32_t r = t.prim_refs[(size_t)j];
        const uint32_t i = r ^ kRefMask;
        if (i >= (uint32_t)n) return fail("unknown reference", j, (long)r);
        if (i >= (uint32_t)n || of_ref[(size_t)j] != -1) return fail("bad reference", j, (long)r);
        of_ref[(size_t)j] = (int)i;
    }
    std::vector<int> seen((size_t)n, 0), covered((size_t)n, 0), node_seen((size_t)t.n_nodes, 0);
    for (int j = 0; j < n; ++j)
        if (seen[(size_t)of_ref[(size_t)j]]++) return fail("reference stored twice", j, of_ref[(size_t)j]);

    // structure: walk from the root
    std::vector<int> todo = {0};
    node_seen[0] = 1;
    while (!todo.empty()) {
        const int node = todo.back();
        todo.pop_back();
        int used = 0;
        for (int j = 0; j < W; ++j) {
            const Slot s = slot(t, node, j);
            if (s.count == -1) {
                for (int k = 0; k < 3; ++k)
                    if (!(s.lo[k] > s.hi[k])) return fail("empty slot with a box", node, j);
                continue;
            }
            ++used;
            if (s.count > 0) {
                if (s.count > std::max(1, max_leaf)) return fail("leaf too large", node, s.count);
                if (s.child < 0 || s.child + s.count > n) return fail("leaf range", node, s.child);
                for (int q = s.child; q < s.child + s.count; ++q) {
                    if (covered[(size_t)q]++) return fail("leaf ranges overlap", node, q);
                    const Aabb& b = boxes[(size_t)of_ref[(size_t)q]];
                    if (!contains(s, b.lo, b.hi)) return fail("leaf box misses a primitive", node, q);
                }
            } else if (s.count == 0) {
                if (s.child <= 0 || s.child >= t.n_nodes) return fail("child index", node, s.child);
                if (node_seen[(size_t)s.child]++) return fail("node reached twice", node, s.child);
                for (int jj = 0; jj < W; ++jj) {
                    const Slot c = slot(t, s.child, jj);
                    if (c.count != -1 && !contains(s, c.lo, c.hi)) return fail("slot misses a child slot", node, jj);
                }
                todo.push_back(s.child);
            } else {
                return fail("slot count", node, s.count);
            }
        }
        if (n > 0 && used == 0) return fail("node without children", node, 0);
    }
    for (int q = 0; q < n; ++q)
        if (!covered[(size_t)q]) return fail("primitive in no leaf", q, 0);
    for (int k = 0; k < t.n_nodes; ++k)
        if (!node_seen[(size_t)k]) return fail("unreachable node", k, 0);

    if (layout == 4) {  // the structure holds; ray tests need finite spheres
        std::printf("ok nodes=%d depth=%d (non-finite box)\n", t.n_nodes, t.depth);
        return 0;
    }
    // rays: closest hit by traversal against all spheres, and the stack bound
    std::uniform_real_distribution<float> um(-1.0f, 1.0f);
    int max_seen = 0;
    for (int ray = 0; ray < 4000; ++ray) {
        float o[3], d[3];
        for (int k = 0; k < 3; ++k) o[k] = 0.5f + 1.5f * um(rng);
        const float aim[3] = {u01(rng), u01(rng), u01(rng)};
        float len = 0.0f;
        for (int k = 0; k < 3; ++k) {
            d[k] = (ray % 5 == 0) ? um(rng) : aim[k] - o[k];
            len += d[k] * d[k];
        }
        len = std::sqrt(len);
        if (len == 0.0f) continue;
        for (int k = 0; k < 3; ++k) d[k] /= len;
        if (ray % 7 == 0) d[ray % 3] = 0.0f;  // axis-parallel components

        float best_t = 1e30f;
        int best = -1;
        for (int i = 0; i < n; ++i) {
            const float th = hit_sphere(sph[(size_t)i], o, d);
            if (th > 0.0f && (th < best_t || (th == best_t && i < best))) {
                best_t = th;
                best = i;
            }
        }
        float bvh_t = 1e30f;
        int bvh_best = -1;
        std::vector<int> stack;  // node indices, or ~leaf-slot encoded as -(2 + node * W + j)
        int cur = 0;
        while (true) {
            if (cur >= 0) {
                struct Hit { float t; int node, j; };
                std::vector<Hit> hits;
                for (int j = 0; j < W; ++j) {
                    const Slot s = slot(t, cur, j);
                    if (s.count == -1) continue;
                    const float te = box_entry(s, o, d, bvh_t);
                    if (te >= 0.0f) hits.push_back({te, cur, j});
                }
                std::sort(hits.begin(), hits.end(), [](const Hit& a, const Hit& b) { return a.t < b.t; });
                for (size_t h = hits.size(); h-- > 1;) stack.push_back(-(2 + hits[h].node * W + hits[h].j));
                max_seen = std::max(max_seen, (int)stack.size());
                cur = hits.empty() ? -1 : -(2 + hits[0].node * W + hits[0].j);
                if (cur != -1) continue;
            } else if (cur <= -2) {
                const int code = -cur - 2, node = code / W, j = code % W;
                const Slot s = slot(t, node, j);
                if (s.count == 0) {
                    cur = s.child;
                    continue;
                }
                for (int q = s.child; q < s.child + s.count; ++q) {
                    const int i = of_ref[(size_t)q];
                    const float th = hit_sphere(sph[(size_t)i], o, d);
                    if (th > 0.0f && (th < bvh_t || (th == bvh_t && i < bvh_best))) {
                        bvh_t = th;
                        bvh_best = i;
                    }
                }
            }
            if (stack.empty()) break;
            cur = stack.back();
            stack.pop_back();
        }
        if (bvh_best != best || (best >= 0 && bvh_t != best_t)) return fail("closest hit differs", ray, bvh_best);
    }
    if (max_seen > t.max_stack) return fail("stack above max_stack", max_seen, t.max_stack);
    std::printf("ok nodes=%d depth=%d max_stack=%d max_seen=%d\n", t.n_nodes, t.depth, t.max_stack, max_seen);
    return 0;
}
