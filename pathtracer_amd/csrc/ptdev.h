// ptdev.h — host/device helpers of the gfx950 tracer: a small vector type with
// GLSL's component-wise semantics, the PCG hash, wave64 queue appends.
//
// Every float expression keeps the operand order GLSL evaluates
// (left-associative, no contraction; the build passes -ffp-contract=off), so
// the reference-mode kernels perform the same IEEE operations as the
// reference shader's expressions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "ptmath.h"

#define PT_HD __host__ __device__ __forceinline__

namespace ptd {

struct f3 {
    float x, y, z;
};
struct f4 {
    float x, y, z, w;
};

PT_HD f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
PT_HD f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
PT_HD f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
PT_HD f3 operator*(f3 a, f3 b) { return f3{a.x * b.x, a.y * b.y, a.z * b.z}; }
PT_HD f3 operator/(f3 a, f3 b) { return f3{a.x / b.x, a.y / b.y, a.z / b.z}; }
PT_HD f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
PT_HD f3 operator*(float s, f3 a) { return f3{s * a.x, s * a.y, s * a.z}; }
PT_HD f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
PT_HD f3 operator-(f3 a) { return f3{-a.x, -a.y, -a.z}; }
PT_HD f3 fma3(f3 a, f3 b, f3 c) { return f3{fmaf(a.x, b.x, c.x), fmaf(a.y, b.y, c.y), fmaf(a.z, b.z, c.z)}; }
PT_HD f3 fma3s(f3 a, float t, f3 c) { return f3{fmaf(a.x, t, c.x), fmaf(a.y, t, c.y), fmaf(a.z, t, c.z)}; }
PT_HD float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
PT_HD float length(f3 a) { return sqrtf(dot(a, a)); }
PT_HD f3 normalize(f3 a) { return a / length(a); }
PT_HD f3 abs3(f3 a) { return f3{fabsf(a.x), fabsf(a.y), fabsf(a.z)}; }
PT_HD f3 min3(f3 a, f3 b) { return f3{fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)}; }
PT_HD f3 max3(f3 a, f3 b) { return f3{fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)}; }
PT_HD f3 cross(f3 a, f3 b) { return f3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
PT_HD float comp(f3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }

PT_HD f4 mk4(float s) { return f4{s, s, s, s}; }
PT_HD f4 operator+(f4 a, f4 b) { return f4{a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w}; }
PT_HD f4 operator*(f4 a, f4 b) { return f4{a.x * b.x, a.y * b.y, a.z * b.z, a.w * b.w}; }
PT_HD f4 operator*(f4 a, float s) { return f4{a.x * s, a.y * s, a.z * s, a.w * s}; }

// GLSL mat3, stored as columns c0, c1, c2.
struct m3 {
    f3 c0, c1, c2;
};
// v * M  (= transpose(M) v)
PT_HD f3 vmul(f3 v, const m3& m) { return f3{dot(v, m.c0), dot(v, m.c1), dot(v, m.c2)}; }
// M * v
PT_HD f3 mvmul(const m3& m, f3 v) {
    return f3{m.c0.x * v.x + m.c1.x * v.y + m.c2.x * v.z, m.c0.y * v.x + m.c1.y * v.y + m.c2.y * v.z,
              m.c0.z * v.x + m.c1.z * v.y + m.c2.z * v.z};
}
// A * B with entry (col j, row r) = A[0][r] B[j][0] + A[1][r] B[j][1] + A[2][r] B[j][2]
PT_HD m3 mmul(const m3& a, const m3& b) {
    auto col = [&](f3 bj) {
        return f3{a.c0.x * bj.x + a.c1.x * bj.y + a.c2.x * bj.z, a.c0.y * bj.x + a.c1.y * bj.y + a.c2.y * bj.z,
                  a.c0.z * bj.x + a.c1.z * bj.y + a.c2.z * bj.z};
    };
    return m3{col(b.c0), col(b.c1), col(b.c2)};
}
// RotationMatrix(angle degrees) = mX * mY * mZ (src/shader.comp:142-152).
PT_HD m3 rotation_deg(f3 a) {
    a = a * 0.0174532925199f;
    float sx = ptm_sin(a.x), sy = ptm_sin(a.y), sz = ptm_sin(a.z);
    float cx = ptm_cos(a.x), cy = ptm_cos(a.y), cz = ptm_cos(a.z);
    m3 mx{f3{1.0f, 0.0f, 0.0f}, f3{0.0f, cx, -sx}, f3{0.0f, sx, cx}};
    m3 my{f3{cy, 0.0f, sy}, f3{0.0f, 1.0f, 0.0f}, f3{-sy, 0.0f, cy}};
    m3 mz{f3{cz, -sz, 0.0f}, f3{sz, cz, 0.0f}, f3{0.0f, 0.0f, 1.0f}};
    return mmul(mmul(mx, my), mz);
}

// GLSL refract(I, N, eta).
PT_HD f3 refract(f3 I, f3 N, float eta) {
    float d = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return f3{0.0f, 0.0f, 0.0f};
    return eta * I - (eta * d + sqrtf(k)) * N;
}

// pcg_hash (Jarzynski & Olano 2020), the reference's PCG32 (src/shader.comp:936-941).
PT_HD uint32_t pcg(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
// RandomFloatPCG32: advances the chain, returns float(seed) / 2^32 in [0, 1].
PT_HD float rand_chain(uint32_t& seed) {
    seed = pcg(seed);
    return (float)seed * 2.3283064365386963e-10f;  // exact: division by 2^32
}

// A load through the constant address space: it marks the memory read-only
// for the kernel's duration (the scene tables are), which a generic pointer
// into memory the kernel also writes cannot show the compiler, so a uniform
// address becomes a scalar load through the scalar cache instead of a vector
// load with its full memory latency. Divergent addresses still give vector
// loads.
template <typename T>
__device__ __forceinline__ T ld_ro(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// Appends `pred` lanes of the calling wave to a queue whose length lives at
// *counter: one atomic per wave (ballot + popcount), slot order = lane order.
// Returns the slot for a predicated lane, -1 otherwise. Every active lane of
// the wave must call it.
__device__ __forceinline__ int wave_append(int* counter, bool pred) {
    uint64_t mask = __ballot(pred);
    if (mask == 0) return -1;
    int lane = (int)__lane_id();
    int leader = __ffsll((unsigned long long)mask) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(counter, (int)__popcll(mask));
    base = __shfl(base, leader);
    uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
    return pred ? base + (int)below : -1;
}
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Appends the calling thread to each of NQ queues whose bit is set in `want`,
// with one device atomic per queue per block (per-wave ballots, block totals
// in LDS): device-scope atomics on one address serialise across the chip, so
// a block of BLK threads issues BLK/64 times fewer than per-wave appends.
// slot[q] receives the thread's slot in queue q (-1 if not appended); slots
// follow (wave, lane) order. Every thread of the block must call it, from
// block-uniform control flow (it synchronises the block).
template <int BLK, int NQ>
__device__ __forceinline__ void block_append(int* const (&ctr)[NQ], uint32_t want, int (&slot)[NQ]) {
    constexpr int NW = BLK / 64;
    __shared__ int s_n[NQ][NW];
    __shared__ int s_base[NQ];
    const int lane = (int)__lane_id(), wv = (int)threadIdx.x >> 6;
    uint64_t m[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        m[q] = __ballot((want >> q) & 1u);
        if (lane == 0) s_n[q][wv] = (int)__popcll(m[q]);
    }
    __syncthreads();
    if ((int)threadIdx.x < NQ) {
        const int q = (int)threadIdx.x;
        int tot = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) tot += s_n[q][w];
        s_base[q] = tot ? atomicAdd(ctr[q], tot) : 0;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        int off = s_base[q];
        for (int w = 0; w < wv; ++w) off += s_n[q][w];
        slot[q] = ((want >> q) & 1u) ? off + (int)lanes_below(m[q]) : -1;
    }
    __syncthreads();  // s_n and s_base are rewritten by the next call
}

// Device counters (pt_counters): kStatLines copies of kStatSlots 64-bit slots,
// one 256-byte line per copy; a block adds into copy blockIdx.x % kStatLines and
// the host sums the copies, so per-wave additions from a whole grid do not
// serialise on one address.
constexpr int kStatSlots = 20, kStatLines = 64, kStatStride = 32;
// Slots 14..17: FP32 operations of extension-ray sphere tracing by SDF kind
// (PT_SDF_MANDELBULB .. PT_SDF_TERRAIN); slot 18: those of shadow-ray marching;
// slot 19: hit records naming no primitive of their scene (always 0 unless a
// kernel left a hit record unwritten or corrupt).
constexpr int kStatMarchFlops = 14, kStatShadowMarchFlops = 18, kStatBadRefs = 19;
__device__ __forceinline__ unsigned long long* stat_slot(unsigned long long* stats, int slot) {
    return stats + (size_t)(blockIdx.x % kStatLines) * kStatStride + slot;
}

// Sum of `v` over the 64 lanes of the wave, added once to *dst. Call from
// converged control flow with the whole wave active (lanes without work pass 0).
__device__ __forceinline__ void wave_add_u64(unsigned long long* dst, unsigned long long v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
    if (__lane_id() == 0 && v != 0ull) atomicAdd(dst, v);
}

}  // namespace ptd
