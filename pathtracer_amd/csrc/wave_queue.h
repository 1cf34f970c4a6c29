// Wave-level shares of a segmented work queue, used by the persistent kernels:
// the queue's n entries are split into min(kSeg, waves in the grid) segments,
// each with a cursor on its own 256-byte line; a wave takes `chunk` entries per
// atomic from its home segment and hands them to its idle lanes from registers.
// Every entry is handed out exactly once, and every segment has a home wave.
//
// Include after the HIP runtime and lanes_below() (ptdev.h). The functions are
// wave-uniform: every lane of a wave calls them with the same arguments.
#pragma once

// Queue counters and cursors sit kCntStride ints (256 bytes) apart: device-scope
// atomics on one line serialise across the chip.
constexpr int kCntStride = 64;
// Segments per persistent queue.
constexpr int kSeg = 32;
// Most entries a wave takes per cursor atomic.
constexpr int kChunkMax = 64;

// A wave's share of a segmented queue; wave-uniform.
struct WaveChunk {
    int lo = 0, hi = 0;          // entries taken but not handed out yet
    int seg_lo = 0, seg_hi = 0;  // the home segment
    int chunk = 1;
    int* cur = nullptr;          // the home segment's cursor
    bool exhausted = false;      // the home segment has no entries left to take
    __device__ __forceinline__ bool drained() const { return exhausted && lo >= hi; }
};

__device__ __forceinline__ WaveChunk wave_chunk_init(int* cursors, int n) {
    WaveChunk c;
    const int nw = (int)(blockDim.x >> 6);
    const int waves = (int)gridDim.x * nw;
    // every segment needs at least one home wave: a grid of fewer than kSeg
    // waves uses as many segments as it has waves
    const int nseg = min(kSeg, waves);
    const int seg = ((int)blockIdx.x * nw + __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) % nseg;
    const int len = (n + nseg - 1) / nseg;
    c.seg_lo = min(seg * len, n);
    c.seg_hi = min(c.seg_lo + len, n);
    c.chunk = max(1, min(kChunkMax, n / max(1, waves)));
    c.cur = cursors + seg * kCntStride;
    c.exhausted = c.seg_lo >= c.seg_hi;
    return c;
}

// Gives the idle lanes the next entries of the wave's segment; returns this
// lane's entry or -1. Call from converged control flow.
__device__ __forceinline__ int wave_refill(uint64_t idle, WaveChunk& c) {
    const int lane = (int)__lane_id();
    const int k = (int)__popcll(idle);
    const int avail = c.hi - c.lo;
    int nlo = 0, nhi = 0;
    if (avail < k && !c.exhausted) {
        int base = 0;
        if (lane == 0) base = atomicAdd(c.cur, c.chunk);
        base = __builtin_amdgcn_readfirstlane(base);  // lane 0's value, in a scalar register
        nlo = min(c.seg_lo + base, c.seg_hi);
        nhi = min(nlo + c.chunk, c.seg_hi);
        if (nhi >= c.seg_hi) c.exhausted = true;
    }
    const int r = (int)lanes_below(idle);
    int e = -1;
    if ((idle >> lane) & 1ull) {
        if (r < avail) e = c.lo + r;
        else if (r - avail < nhi - nlo) e = nlo + (r - avail);
    }
    if (k <= avail) {
        c.lo += k;
    } else {
        c.lo = nlo + min(k - avail, nhi - nlo);
        c.hi = nhi;
    }
    return e;
}
