// Runs the persistent kernels' queue-share code (pathtracer_amd/csrc/wave_queue.h)
// on the CPU, lane by lane, for a grid of waves interleaved in a random order,
// and checks what the kernels rely on:
//   - every queue entry is handed to exactly one lane, and none outside [0, n);
//   - every wave leaves its loop (each segment has a home wave that drains it);
//   - the wave-uniform state stays equal across the 64 lanes of a wave;
//   - a lane that is not idle is handed nothing.
// The loop is the persistent kernels' loop: refill when at least `refill` lanes
// are idle (or, with `all_idle`, when none is busy), step the busy lanes, leave
// once no lane is busy and the share is drained. A lane's work takes 0..max_work
// steps; 0 is a ray that the march kernel skips (its SDF-box mask bit is clear).
//
// Usage: wave_queue_sim blocks threads_per_block n refill all_idle max_work seed
// Prints "ok <iterations>" or the first violation, and exits 0 or 1.
#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

struct Dim {
    unsigned x = 0;
};
static Dim blockDim, gridDim, blockIdx, threadIdx;
static int g_lane = 0;
static int g_first = 0;

#define __device__
#define __forceinline__ inline
using std::max;
using std::min;
static inline unsigned __lane_id() { return (unsigned)g_lane; }
static inline int __popcll(uint64_t m) { return __builtin_popcountll(m); }
static inline uint32_t lanes_below(uint64_t m) {
    return (uint32_t)__builtin_popcountll(m & ((1ull << g_lane) - 1ull));
}
static inline int atomicAdd(int* p, int v) {
    const int old = *p;
    *p += v;
    return old;
}
// lanes are run in order 0..63, so lane 0's value is known to the others
static inline int sim_readfirstlane(int x) {
    if (g_lane == 0) g_first = x;
    return g_first;
}
#define __builtin_amdgcn_readfirstlane(x) sim_readfirstlane(x)

#include "wave_queue.h"

struct Wave {
    int block = 0, wib = 0;
    WaveChunk c[64];
    int entry[64];
    int left[64];
    bool done = false;
};

static bool same(const WaveChunk& a, const WaveChunk& b) {
    return a.lo == b.lo && a.hi == b.hi && a.seg_lo == b.seg_lo && a.seg_hi == b.seg_hi &&
           a.chunk == b.chunk && a.cur == b.cur && a.exhausted == b.exhausted;
}

static int fail(const char* what, int w, int lane, int v) {
    std::printf("FAIL %s wave=%d lane=%d value=%d\n", what, w, lane, v);
    return 1;
}

int main(int argc, char** argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s blocks threads n refill all_idle max_work seed\n", argv[0]);
        return 2;
    }
    const int blocks = std::atoi(argv[1]), threads = std::atoi(argv[2]), n = std::atoi(argv[3]);
    const int refill = std::atoi(argv[4]), all_idle = std::atoi(argv[5]), max_work = std::atoi(argv[6]);
    std::mt19937 rng((unsigned)std::atoi(argv[7]));
    if (blocks < 1 || threads < 64 || threads % 64 || n < 0 || refill < 1 || refill > 64 || max_work < 0) return 2;

    std::vector<int> cursors((size_t)kSeg * kCntStride, 0);
    blockDim.x = (unsigned)threads;
    gridDim.x = (unsigned)blocks;
    const int nw = threads / 64;
    std::vector<Wave> waves((size_t)blocks * nw);
    for (int b = 0; b < blocks; ++b)
        for (int k = 0; k < nw; ++k) {
            Wave& W = waves[(size_t)b * nw + k];
            W.block = b;
            W.wib = k;
            blockIdx.x = (unsigned)b;
            for (int l = 0; l < 64; ++l) {
                g_lane = l;
                threadIdx.x = (unsigned)(k * 64 + l);
                W.c[l] = wave_chunk_init(cursors.data(), n);
                W.entry[l] = -1;
                W.left[l] = 0;
            }
        }

    std::vector<int> seen((size_t)n, 0);
    std::uniform_int_distribution<int> work(0, max_work);
    const long long cap = 1000000LL + 8LL * ((long long)n + 64) * (max_work + 2);
    long long iters = 0;
    int live = (int)waves.size();
    std::vector<int> order(waves.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
    while (live > 0) {
        if (++iters > cap) return fail("no progress (a wave never leaves its loop)", -1, -1, live);
        // one loop iteration of every live wave, in a random order per sweep
        std::shuffle(order.begin(), order.end(), rng);
        for (int w : order) {
            Wave& W = waves[(size_t)w];
            if (W.done) continue;
            blockIdx.x = (unsigned)W.block;
            uint64_t idle = 0;
            for (int l = 0; l < 64; ++l)
                if (W.entry[l] < 0) idle |= 1ull << l;
            const bool want = __builtin_popcountll(idle) >= refill || (all_idle && ~idle == 0ull);
            if (!W.c[0].drained() && want) {
                for (int l = 0; l < 64; ++l) {
                    g_lane = l;
                    threadIdx.x = (unsigned)(W.wib * 64 + l);
                    const int e = wave_refill(idle, W.c[l]);
                    if (e < -1 || e >= n) return fail("entry out of range", w, l, e);
                    if (!((idle >> l) & 1ull)) {
                        if (e != -1) return fail("busy lane handed an entry", w, l, e);
                        continue;
                    }
                    if (e < 0) continue;
                    if (seen[(size_t)e]++) return fail("entry handed out twice", w, l, e);
                    const int steps = work(rng);
                    if (steps > 0) {
                        W.entry[l] = e;
                        W.left[l] = steps;
                    }
                }
                for (int l = 1; l < 64; ++l)
                    if (!same(W.c[0], W.c[l])) return fail("wave-uniform state differs across lanes", w, l, 0);
            }
            bool busy = false;
            for (int l = 0; l < 64; ++l) busy = busy || W.entry[l] >= 0;
            if (!busy) {
                if (W.c[0].drained()) {
                    W.done = true;
                    --live;
                }
                continue;
            }
            for (int l = 0; l < 64; ++l)
                if (W.entry[l] >= 0 && --W.left[l] == 0) W.entry[l] = -1;
        }
    }
    for (int e = 0; e < n; ++e)
        if (seen[(size_t)e] != 1) return fail("entry never handed out", -1, -1, e);
    std::printf("ok %lld\n", iters);
    return 0;
}
