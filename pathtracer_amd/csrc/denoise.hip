// denoise.hip — a small U-Net denoiser for the tracer's low-sample images,
// trained and run on gfx950 matrix cores (pt_dn_*). No counterpart in the
// reference.
//
// Network (NHWC BF16 activations, FP32 accumulation, FP32 master weights):
//   in  16 ch: log1p(XYZ) (3), albedo (3), normal (3), 1 / (1 + depth) (1), zeros (6)
//   e0: conv 16->32, conv 32->32                      (H x W)      skip0
//   e1: avgpool 2, conv 32->64, conv 64->64           (H/2)        skip1
//   b : avgpool 2, conv 64->128, conv 128->128        (H/4)
//   d1: upsample 2, conv [128 | skip1 64]->64, conv 64->64
//   d0: upsample 2, conv [64 | skip0 32]->32, conv 32->32
//   out: conv 32->16 (channels 0..2 used), no activation
// Every conv is 3x3, zero padding 1, bias, ReLU except the output. The output
// is a correction in log space: log1p(denoised XYZ) = log1p(input XYZ) + out.
// Loss: mean |log1p(denoised) - log1p(target)| over pixels and the 3 channels.
//
// Convolutions are implicit GEMMs on v_mfma_f32_32x32x16_bf16 (lane l holds
// A[row l % 32][k (l / 32) * 8 + i], B[k (l / 32) * 8 + i][col l % 32]; the
// 16 results of lane l are rows (v / 4) * 8 + (l / 32) * 4 + v % 4, column
// l % 32):
//   forward / backward-data: rows = pixels, columns = output channels,
//     K = 9 taps x input channels in steps of 16 (each lane reads 8
//     consecutive channels: 16 bytes of NHWC), both operands staged in LDS per
//     tile; backward-data runs the forward kernel on the output gradient with
//     the weights flipped and transposed;
//   backward-weights: rows = output channels, columns = input channels,
//     K = pixels; each wave stages 16 pixels x 32 channels of both operands in
//     LDS to read them transposed, and adds its partial sums into the FP32
//     gradient with atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../include/pathtracer.h"
#include "tracer.hpp"

namespace ptdn {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));
typedef unsigned short u16;

constexpr int kIn = 16;      // input channels (12 used)
constexpr int kFeat = 12;    // floats per pixel of a feature image
constexpr int kOutC = 16;    // output channels (3 used)
constexpr int kThreads = 256;

__device__ __forceinline__ u16 f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((uint32_t)h << 16); }

union Frag {
    uint4 u;
    bf8 v;
};
// A compile-time range of 32-channel output tiles [J0, J0 + N) (k_conv_t's epilogue).
template <int J0, int N>
struct SliceRange {
    static constexpr int j0 = J0, count = N;
};

// ---------------------------------------------------------------------------------
// Conv forward (also backward-data). X1 (C1 channels) and X2 (C2 channels, may
// be null) form the input [N, H, W, C1 + C2]; Wt is [9][Cout][C1 + C2] BF16;
// output channels [0, Co1) go to Y1 (Co1 channels) and [Co1, Cout) to Y2;
// bias and ReLU are applied at the store. `flip` reads tap 8 - t
// (backward-data).
// Pixel index of (n, y, x) of an H x W tensor (q), or with `up`, of the
// (H / 2) x (W / 2) tensor that a nearest-neighbour 2x upsample turns into it:
// the skip layers read their upsampled input this way instead of a stored copy.
__device__ __forceinline__ size_t up_index(int up, size_t q, int n, int y, int x, int H, int W) {
    return up ? ((size_t)n * (H >> 1) + (y >> 1)) * (W >> 1) + (x >> 1) : q;
}

struct ConvArgs {
    const u16* X1;
    const u16* X2;
    int C1, C2;
    const u16* Wt;
    const float* bias;  // null: none
    u16* Y1;
    u16* Y2;
    int Co1, Cout;
    int N, H, W;
    int relu;
    int flip;
    int tiles_per_block;  // consecutive pixel tiles a block computes in turn
    int up1;              // X1 is at half resolution, upsampled 2x (nearest) as it is read
};

// ---------------------------------------------------------------------------------
// Conv forward (also backward-data), tiled: a block computes tiles of TH rows x
// 32 columns of one image's pixels x 32 * NC output channels. Per chunk of 32
// input channels it stages in LDS the (TH + 2) x 34 input tile around its
// pixels (zero outside the image) and the chunk's weights for all 9 taps; each
// wave then runs its TH / 4 pixel rows x NC channel tiles of 32 x 32 MFMAs over
// 9 taps x 2 steps of 16 channels with both operands read from LDS. Every input
// pixel is read from memory by about (TH + 2)(34) / (32 TH) tiles instead of
// once per tap. Pixels per 16-byte vector in LDS are padded to 40 channels so
// that 32 lanes reading 32 neighbouring pixels spread over the banks.
//
// Staging goes through registers: each thread issues all of its 16-byte loads
// of a chunk back to back (one wait for the lot), and the next chunk's loads
// are issued before the current chunk's MFMAs, so they are in flight while the
// matrix cores work. The 9 taps are unrolled, which lets the compiler issue the
// LDS reads of the next MFMAs ahead of the current ones.
//  - Several input chunks: a block computes one tile; its output goes through
//    the weight tile's LDS to 16-byte stores.
//  - One input chunk (ONE_CHUNK, Cin <= 32: the full-resolution layers whose
//    load, MFMAs and store would otherwise follow one another): a block
//    computes `tiles_per_block` consecutive tiles (along x first) with the
//    weights staged once; the next tile's input loads are in flight under the
//    current tile's MFMAs and output stores, and the output goes through the
//    input tile's LDS, 32 channels at a time.
// Register budget: the LDS tiles allow 2 blocks per CU at NC = 2 (73 KB each)
// and 3 at NC = 1 (50 KB), so the launch bounds ask for that many waves per SIMD.
template <int NC, int TH, bool ONE_CHUNK>
__global__ void __launch_bounds__(kThreads, NC == 2 ? 2 : 3) k_conv_t(ConvArgs a) {
    constexpr int CK = 32, CKP = CK + 8;  // channels per chunk; LDS pitch per pixel / weight row
    constexpr int TW = 32, IW = TW + 2, IH = TH + 2;
    constexpr int R = TH / 4;             // pixel rows per wave
    constexpr int NXV = IH * IW * (CK / 8), NWV = 9 * 32 * NC * (CK / 8);  // 16-byte vectors per chunk
    constexpr int NXT = (NXV + kThreads - 1) / kThreads, NWT = (NWV + kThreads - 1) / kThreads;
    __shared__ __align__(16) u16 s_x[IH * IW * CKP];
    __shared__ __align__(16) u16 s_w[9 * 32 * NC * CKP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int tiles_x = (a.W + TW - 1) / TW, tiles_y = (a.H + TH - 1) / TH;
    const int co0 = blockIdx.y * 32 * NC;
    const int Cin = a.C1 + a.C2;
    const int kq = (lane >> 5) * 8;  // this lane's 8 channels within a 16-channel step
    const int col = lane & 31;
    f16v acc[R][NC];
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
        for (int j = 0; j < NC; ++j) acc[r][j] = f16v{};
    float bias[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
        const int co = co0 + j * 32 + col;
        bias[j] = a.bias && co < a.Cout ? a.bias[co] : 0.0f;
    }
    // tile -> image n and the tile's first pixel (x0, y0)
    auto origin = [&](int tile, int& n, int& x0, int& y0) {
        x0 = (tile % tiles_x) * TW;
        const int rest = tile / tiles_x;
        y0 = (rest % tiles_y) * TH;
        n = rest / tiles_y;
    };
    uint4 xr[NXT], wr[NWT];
    // this thread's share of a chunk: input vectors i = threadIdx.x + u * kThreads
    // (pixel i / 4 of the halo tile, channels 8 (i % 4) ..), weight vectors likewise
    // (tap, output row, 8 channels)
    auto fetch_x = [&](int n, int x0, int y0, int c0) {
#pragma unroll
        for (int u = 0; u < NXT; ++u) {
            const int i = threadIdx.x + u * kThreads;
            const int v = i % (CK / 8), pix = i / (CK / 8);
            const int y = y0 + pix / IW - 1, x = x0 + pix % IW - 1;
            const int c = c0 + v * 8;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (i < NXV && y >= 0 && y < a.H && x >= 0 && x < a.W && c < Cin) {
                const size_t q = ((size_t)n * a.H + y) * a.W + x;
                const u16* src = c < a.C1 ? a.X1 + up_index(a.up1, q, n, y, x, a.H, a.W) * a.C1 + c
                                          : a.X2 + q * a.C2 + (c - a.C1);
                val = *reinterpret_cast<const uint4*>(src);
            }
            xr[u] = val;
        }
    };
    auto fetch_w = [&](int c0) {
#pragma unroll
        for (int u = 0; u < NWT; ++u) {
            const int i = threadIdx.x + u * kThreads;
            const int v = i % (CK / 8), row = (i / (CK / 8)) % (32 * NC), t = i / ((CK / 8) * 32 * NC);
            const int tt = a.flip ? 8 - t : t;
            const int co = co0 + row, c = c0 + v * 8;
            uint4 val = make_uint4(0, 0, 0, 0);
            if (i < NWV && co < a.Cout && c < Cin)
                val = *reinterpret_cast<const uint4*>(a.Wt + ((size_t)tt * a.Cout + co) * Cin + c);
            wr[u] = val;
        }
    };
    auto put_x = [&]() {
#pragma unroll
        for (int u = 0; u < NXT; ++u) {
            const int i = threadIdx.x + u * kThreads;
            if (i < NXV) *reinterpret_cast<uint4*>(&s_x[(i / (CK / 8)) * CKP + (i % (CK / 8)) * 8]) = xr[u];
        }
    };
    auto put_w = [&]() {
#pragma unroll
        for (int u = 0; u < NWT; ++u) {
            const int i = threadIdx.x + u * kThreads;
            if (i < NWV) *reinterpret_cast<uint4*>(&s_w[(i / (CK / 8)) * CKP + (i % (CK / 8)) * 8]) = wr[u];
        }
    };
    // the MFMAs of the chunk at c0 staged in LDS
    auto mma = [&](int c0) {
#pragma unroll
        for (int t = 0; t < 9; ++t) {
            const int dy = t / 3, dx = t % 3;  // offsets into the halo-padded tile
#pragma unroll
            for (int k = 0; k < CK; k += 16) {
                if (c0 + k >= Cin) break;  // a 16-channel input's second step is all zeros
                Frag A[R];
#pragma unroll
                for (int r = 0; r < R; ++r)
                    A[r].u = *reinterpret_cast<const uint4*>(&s_x[((wave * R + r + dy) * IW + col + dx) * CKP + k + kq]);
#pragma unroll
                for (int j = 0; j < NC; ++j) {
                    Frag B;
                    B.u = *reinterpret_cast<const uint4*>(&s_w[(t * 32 * NC + j * 32 + col) * CKP + k + kq]);
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        acc[r][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A[r].v, B.v, acc[r][j], 0, 0, 0);
                }
            }
        }
    };
    // Output channels [j0, j0 + 32 J) of the tile at (n, x0, y0) staged in s_o
    // (pitch OP per pixel): bias, ReLU and BF16 rounding per accumulator into LDS
    // in pixel-major order, then 16-byte vectors of 8 channels to memory (the
    // tensors' channel counts and the Y1 / Y2 split are multiples of 8), one
    // bounds check per vector. The caller synchronises before (LDS free) and
    // after (LDS read) it.
    auto store_slices = [&](u16* s_o, auto jc, int n, int x0, int y0) {
        constexpr int J0 = decltype(jc)::j0, J = decltype(jc)::count, OP = 32 * J + 8;
#pragma unroll
        for (int jj = 0; jj < J; ++jj) {
            const int j = J0 + jj;
            const int co = co0 + j * 32 + col;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int row = wave * R + r;
#pragma unroll
                for (int v = 0; v < 16; ++v) {
                    const int px = (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3);
                    float o = acc[r][j][v] + bias[j];
                    if (a.relu) o = fmaxf(o, 0.0f);
                    s_o[(row * TW + px) * OP + jj * 32 + col] = co < a.Cout ? f2bf(o) : (u16)0;
                }
            }
        }
        __syncthreads();
        constexpr int VPP = 32 * J / 8;  // 16-byte vectors per pixel
#pragma unroll
        for (int u = 0; u < (TH * TW * VPP) / kThreads; ++u) {
            const int i = threadIdx.x + u * kThreads;
            const int g = i % VPP, pix = i / VPP;
            const int y = y0 + pix / TW, x = x0 + pix % TW, c = co0 + J0 * 32 + g * 8;
            if (y >= a.H || x >= a.W || c >= a.Cout) continue;
            const uint4 val = *reinterpret_cast<const uint4*>(&s_o[pix * OP + g * 8]);
            const size_t p = ((size_t)n * a.H + y) * a.W + x;
            u16* dst = c < a.Co1 ? a.Y1 + p * a.Co1 + c : a.Y2 + p * (a.Cout - a.Co1) + (c - a.Co1);
            *reinterpret_cast<uint4*>(dst) = val;
        }
    };
    if constexpr (!ONE_CHUNK) {
        int n, x0, y0;
        origin(blockIdx.x, n, x0, y0);
        fetch_x(n, x0, y0, 0);
        fetch_w(0);
#pragma unroll 1
        for (int c0 = 0; c0 < Cin; c0 += CK) {
            __syncthreads();  // every wave is done with the previous chunk's tiles
            put_x();
            put_w();
            __syncthreads();
            if (c0 + CK < Cin) {
                fetch_x(n, x0, y0, c0 + CK);
                fetch_w(c0 + CK);
            }
            mma(c0);
        }
        static_assert(TH * TW * (32 * NC + 8) <= 9 * 32 * NC * CKP, "the output tile must fit the weight tile's LDS");
        __syncthreads();  // every wave is done with the last chunk's tiles
        store_slices(s_w, SliceRange<0, NC>{}, n, x0, y0);
    } else {
        static_assert(TH * TW * (32 + 8) <= IH * IW * CKP, "an output slice must fit the input tile's LDS");
        const int ntiles = a.N * tiles_x * tiles_y;
        const int t_begin = blockIdx.x * a.tiles_per_block;
        const int t_count = min(a.tiles_per_block, ntiles - t_begin);
        int n, x0, y0;
        origin(t_begin, n, x0, y0);
        fetch_x(n, x0, y0, 0);
        fetch_w(0);
#pragma unroll 1
        for (int k = 0; k < t_count; ++k) {
            __syncthreads();  // every wave is done with the previous tile's output
            put_x();
            if (k == 0) put_w();
            __syncthreads();
            int n1 = n, x1 = x0, y1 = y0;
            if (k + 1 < t_count) {
                origin(t_begin + k + 1, n1, x1, y1);
                fetch_x(n1, x1, y1, 0);
            }
            mma(0);
            __syncthreads();  // every wave is done with the input tile
            store_slices(s_x, SliceRange<0, 1>{}, n, x0, y0);
            if constexpr (NC == 2) {
                __syncthreads();
                store_slices(s_x, SliceRange<1, 1>{}, n, x0, y0);
            }
#pragma unroll
            for (int r = 0; r < R; ++r)
#pragma unroll
                for (int j = 0; j < NC; ++j) acc[r][j] = f16v{};
            n = n1;
            x0 = x1;
            y0 = y1;
        }
    }
}

// ---------------------------------------------------------------------------------
// Conv backward-weights: dW[t][co][ci] += sum over pixels of dZ[p][co] * X[p + tap t][ci]
// (FP32 atomics), dbias[co] += sum of dZ[p][co] (tap 4 blocks). Grid: x = pixel
// chunks, y = co tiles of 32, z = 9 taps x ci tiles of 32.
struct WgradArgs {
    const u16* dZ;
    int Cout;
    const u16* X1;
    const u16* X2;
    int C1, C2;
    float* dW;   // [9][Cout][C1 + C2]
    float* dB;   // [Cout]
    int N, H, W;
    int chunk;   // pixels per block (multiple of 64)
    int up1;     // X1 at half resolution, as in ConvArgs
};

__global__ void __launch_bounds__(kThreads) k_conv_wgrad(WgradArgs a) {
    __shared__ u16 s_dz[4][16][32 + 8];
    __shared__ u16 s_x[4][16][32 + 8];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int P = a.N * a.H * a.W;
    const int Cin = a.C1 + a.C2;
    const int nci = (Cin + 31) / 32;
    const int t = blockIdx.z / nci, ci0 = (blockIdx.z % nci) * 32;
    const int co0 = blockIdx.y * 32;
    const int dy = t / 3 - 1, dx = t % 3 - 1;
    const int pbeg = blockIdx.x * a.chunk, pend = min(P, pbeg + a.chunk);
    f16v acc = f16v{};
    float bsum = 0.0f;  // lanes 0..31: this co's sum of dZ (tap 4 only)
    const int lr = lane >> 2, lc = (lane & 3) * 8;  // staging: pixel row, 8-channel group
    // each wave takes every 4th group of 16 pixels of the block's chunk
    for (int g = pbeg + wave * 16; g < pend; g += 64) {
        const int p = g + lr;
        Frag dz, xv;
        dz.u = make_uint4(0, 0, 0, 0);
        xv.u = make_uint4(0, 0, 0, 0);
        if (p < pend) {
            if (co0 + lc < a.Cout) dz.u = *reinterpret_cast<const uint4*>(a.dZ + (size_t)p * a.Cout + co0 + lc);
            const int n = p / (a.H * a.W), rem = p % (a.H * a.W);
            const int y = rem / a.W + dy, x = rem % a.W + dx;
            const int c = ci0 + lc;
            if (y >= 0 && y < a.H && x >= 0 && x < a.W && c < Cin) {
                const size_t q = ((size_t)n * a.H + y) * a.W + x;
                const u16* src = c < a.C1 ? a.X1 + up_index(a.up1, q, n, y, x, a.H, a.W) * a.C1 + c
                                          : a.X2 + q * a.C2 + (c - a.C1);
                xv.u = *reinterpret_cast<const uint4*>(src);
            }
        }
        *reinterpret_cast<uint4*>(&s_dz[wave][lr][lc]) = dz.u;
        *reinterpret_cast<uint4*>(&s_x[wave][lr][lc]) = xv.u;
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_s_waitcnt(0);
        __builtin_amdgcn_wave_barrier();
        Frag A, B;
        const int kq = (lane >> 5) * 8, col = lane & 31;
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            reinterpret_cast<u16*>(&A.u)[i] = s_dz[wave][kq + i][col];
            reinterpret_cast<u16*>(&B.u)[i] = s_x[wave][kq + i][col];
        }
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(A.v, B.v, acc, 0, 0, 0);
        if (t == 4 && blockIdx.z % nci == 0 && lane < 32)
            for (int i = 0; i < 16; ++i) bsum += bf2f(s_dz[wave][i][lane]);
        __builtin_amdgcn_wave_barrier();
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        const int co = co0 + (v >> 2) * 8 + (lane >> 5) * 4 + (v & 3), ci = ci0 + (lane & 31);
        if (co < a.Cout && ci < Cin && acc[v] != 0.0f) atomicAdd(&a.dW[((size_t)t * a.Cout + co) * Cin + ci], acc[v]);
    }
    if (t == 4 && blockIdx.z % nci == 0 && lane < 32 && co0 + lane < a.Cout && bsum != 0.0f)
        atomicAdd(&a.dB[co0 + lane], bsum);
}

// ---------------------------------------------------------------------------------
// Elementwise kernels (NHWC BF16). Pooling and upsampling move 8 channels (one
// 16-byte vector) per thread: C is a multiple of 8, and the tensors hold fewer
// than 2^31 such groups (the denoiser's size limits), so index arithmetic is
// 32-bit. Each channel's arithmetic is the one-channel formula, in the same order.
__device__ __forceinline__ float bf_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }
__device__ __forceinline__ uint32_t bf_pack(float lo, float hi) {
    return (uint32_t)f2bf(lo) | ((uint32_t)f2bf(hi) << 16);
}
// per channel: scale * (((a + b) + c) + d)
__device__ __forceinline__ uint32_t sum4(uint32_t a, uint32_t b, uint32_t c, uint32_t d, float scale) {
    return bf_pack(scale * (((bf_lo(a) + bf_lo(b)) + bf_lo(c)) + bf_lo(d)),
                   scale * (((bf_hi(a) + bf_hi(b)) + bf_hi(c)) + bf_hi(d)));
}
__device__ __forceinline__ uint4 sum4(uint4 a, uint4 b, uint4 c, uint4 d, float scale) {
    return make_uint4(sum4(a.x, b.x, c.x, d.x, scale), sum4(a.y, b.y, c.y, d.y, scale),
                      sum4(a.z, b.z, c.z, d.z, scale), sum4(a.w, b.w, c.w, d.w, scale));
}
__device__ __forceinline__ uint32_t scale2(uint32_t a, float scale) {
    return bf_pack(scale * bf_lo(a), scale * bf_hi(a));
}

__global__ void k_pool2(const u16* x, u16* y, int N, int H, int W, int C) {  // 2x2 average
    const uint32_t Ho = H / 2, Wo = W / 2, C8 = C >> 3;
    const uint32_t n = (uint32_t)N * Ho * Wo * C8;
    const uint4* xv = reinterpret_cast<const uint4*>(x);
    uint4* yv = reinterpret_cast<uint4*>(y);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = i % C8, r = i / C8;
        const uint32_t xo = r % Wo, r2 = r / Wo;
        const uint32_t yo = r2 % Ho, b = r2 / Ho;
        const uint32_t q = (b * (uint32_t)H + 2 * yo) * (uint32_t)W + 2 * xo;
        yv[i] = sum4(xv[q * C8 + g], xv[(q + 1) * C8 + g], xv[(q + W) * C8 + g], xv[(q + W + 1) * C8 + g], 0.25f);
    }
}
// gradient of the 2x2 average: dX[q] = 0.25 dY (dX written, not accumulated)
__global__ void k_pool2_bwd(const u16* dy, u16* dx, int N, int H, int W, int C) {
    const uint32_t C8 = C >> 3;
    const uint32_t n = (uint32_t)N * H * W * C8;
    const uint4* dyv = reinterpret_cast<const uint4*>(dy);
    uint4* dxv = reinterpret_cast<uint4*>(dx);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = i % C8, r = i / C8;
        const uint32_t x = r % W, r2 = r / W;
        const uint32_t y = r2 % H, b = r2 / H;
        const uint4 v = dyv[((b * (uint32_t)(H / 2) + y / 2) * (uint32_t)(W / 2) + x / 2) * C8 + g];
        dxv[i] = make_uint4(scale2(v.x, 0.25f), scale2(v.y, 0.25f), scale2(v.z, 0.25f), scale2(v.w, 0.25f));
    }
}
// gradient of the nearest 2x upsample the skip layers apply to their first
// input as they read it (up_index): dX = sum of the 4 dY it fed
__global__ void k_up2_bwd(const u16* dy, u16* dx, int N, int H, int W, int C) {
    const uint32_t C8 = C >> 3;
    const uint32_t n = (uint32_t)N * H * W * C8;
    const uint4* dyv = reinterpret_cast<const uint4*>(dy);
    uint4* dxv = reinterpret_cast<uint4*>(dx);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t g = i % C8, r = i / C8;
        const uint32_t x = r % W, r2 = r / W;
        const uint32_t y = r2 % H, b = r2 / H;
        const uint32_t q = (b * 2 * (uint32_t)H + 2 * y) * 2 * (uint32_t)W + 2 * x;
        dxv[i] = sum4(dyv[q * C8 + g], dyv[(q + 1) * C8 + g], dyv[(q + 2 * W) * C8 + g],
                      dyv[(q + 2 * W + 1) * C8 + g], 1.0f);
    }
}
// dZ = dY * (Y > 0): gradient through the ReLU that produced Y; a second
// gradient `add` (skip connection) is summed in when given
__global__ void k_relu_bwd(const u16* dy, const u16* add, const u16* y, u16* dz, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        float g = bf2f(dy[i]) + (add ? bf2f(add[i]) : 0.0f);
        dz[i] = f2bf(bf2f(y[i]) > 0.0f ? g : 0.0f);
    }
}
// features (N, H, W, 10) FP32 -> input tensor (16 channels BF16)
__global__ void k_prep(const float* f, u16* x, size_t P) {
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        const float* s = f + p * kFeat;
        u16* d = x + p * kIn;
        for (int c = 0; c < 3; ++c) d[c] = f2bf(log1pf(fmaxf(s[c], 0.0f)));
        for (int c = 3; c < 9; ++c) d[c] = f2bf(s[c]);
        d[9] = f2bf(1.0f / (1.0f + fmaxf(s[9], 0.0f)));
        d[10] = f2bf(log1pf(fmaxf(s[10], 0.0f)));
        d[11] = f2bf(rsqrtf(fmaxf(s[11], 1.0f)));
        for (int c = 12; c < kIn; ++c) d[c] = 0;
    }
}
// prediction z = log1p(denoised) = log1p(in) + out[0..2]; loss and its
// gradient with respect to z, mean over 3P values: |z - log1p(target)|
// (l2 == 0) or (expm1(z) - target)^2 (l2 == 1); also the denoised image
// max(expm1(z), 0).
__global__ void k_head(const float* f, const u16* out, const float* target, float* denoised, u16* dout,
                       float* loss, size_t P, int l2) {
    float l = 0.0f;
    const float inv = 1.0f / (3.0f * (float)P);
    for (size_t p = blockIdx.x * (size_t)blockDim.x + threadIdx.x; p < P; p += (size_t)gridDim.x * blockDim.x) {
        for (int c = 0; c < kOutC; ++c) {
            if (c >= 3) {
                if (dout) dout[p * kOutC + c] = 0;
                continue;
            }
            const float z = log1pf(fmaxf(f[p * kFeat + c], 0.0f)) + bf2f(out[p * kOutC + c]);
            if (denoised) denoised[p * 3 + c] = fmaxf(expm1f(z), 0.0f);
            if (target && l2) {
                const float y = expm1f(z), d = y - fmaxf(target[p * 3 + c], 0.0f);
                l += d * d;
                if (dout) dout[p * kOutC + c] = f2bf(2.0f * d * (y + 1.0f) * inv);
            } else if (target) {
                const float d = z - log1pf(fmaxf(target[p * 3 + c], 0.0f));
                l += fabsf(d);
                if (dout) dout[p * kOutC + c] = f2bf(d > 0.0f ? inv : (d < 0.0f ? -inv : 0.0f));
            }
        }
    }
    if (loss) {
        for (int off = 32; off > 0; off >>= 1) l += __shfl_down(l, off);
        if ((threadIdx.x & 63) == 0 && l != 0.0f) atomicAdd(loss, l * inv);
    }
}
// Adam on FP32 masters; then the BF16 forward copy Wf[t][co][ci] and the
// backward-data copy Wb[t][ci][co] (k_conv_t with `flip` reads tap 8 - t: the
// gradient of tap t's offset o is gathered from offset -o, which is tap 8 - t).
__global__ void k_adam(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    }
}
__global__ void k_wcopy(const float* w, u16* wf, u16* wb, int Cout, int Cin) {
    const int n = 9 * Cout * Cin;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int ci = i % Cin, co = (i / Cin) % Cout, t = i / (Cin * Cout);
        const u16 h = f2bf(w[i]);
        wf[i] = h;
        wb[((size_t)t * Cin + ci) * Cout + co] = h;
    }
}

// ---------------------------------------------------------------------------------
// Served requests: feature images assembled from a request's per-pixel moments
// and its first-hit features, and the error sums of the dual-buffer estimate.

// which: 0 all samples, 1 the even-index half, 2 the odd-index half. XYZ and
// the standard error of Y in texel units times inv_scale; the standard error
// of a half uses the sample variance of all samples.
__global__ void k_dn_assemble(ptk::DnRequest q, int which, float inv_scale, float* feat) {
    const int P = q.W * q.H;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const int n = q.nsamp[p], no = q.n_odd[p];
        const int cnt = which == 0 ? n : (which == 1 ? n - no : no);
        float* f = feat + (size_t)p * kFeat;
        const float k = cnt > 0 ? q.exposure * inv_scale / (float)cnt : 0.0f;
        for (int c = 0; c < 3; ++c) {
            const float s = q.sum[c * P + p], so = q.sum_odd[c * P + p];
            f[c] = (which == 0 ? s : (which == 1 ? s - so : so)) * k;
        }
        float var = 0.0f;
        if (n > 1) {
            const float s1 = q.sum[P + p], s2 = q.sumsq[P + p];
            var = fmaxf(s2 - s1 * (s1 / (float)n), 0.0f) / (float)(n - 1);
        }
        f[10] = cnt > 0 ? sqrtf(var / (float)cnt) * q.exposure * inv_scale : 0.0f;
        f[11] = (float)cnt;
        const int na = q.aov_n[p];
        const float ka = na > 0 ? 1.0f / (float)na : 0.0f;
        for (int c = 3; c < 10; ++c) f[c] = q.aov[(size_t)p * kFeat + c] * ka;
    }
}

// Against the other half (other: 1 even, 2 odd) of the same request, per block
// then one double atomic per block into line blockIdx.x % 64 (stride 32):
// slot 0 sum (den - other)^2, slot 1 sum of the other half's variance of the
// mean, slot 2 sum of the full mean^2, slot 3 sum (den - prev)^2 when prev
// (the other half's denoised image) is given; all over pixels and XYZ in
// texel units times inv_scale.
__global__ void k_dn_cross(ptk::DnRequest q, const float* den, const float* prev, int other, float inv_scale,
                           double* lines) {
    __shared__ double red[4][4];
    const int P = q.W * q.H;
    double e = 0.0, v = 0.0, m = 0.0, dd = 0.0;
    for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < P; p += gridDim.x * blockDim.x) {
        const int n = q.nsamp[p], no = q.n_odd[p];
        const int cnt = other == 1 ? n - no : no;
        if (cnt < 1 || n < 2) continue;
        const double k = (double)q.exposure * inv_scale;
        for (int c = 0; c < 3; ++c) {
            const double s = q.sum[c * P + p], so = q.sum_odd[c * P + p], s2 = q.sumsq[c * P + p];
            const double o = (other == 1 ? s - so : so) / cnt * k;
            const double d = (double)den[(size_t)p * 3 + c] - o;
            e += d * d;
            const double var = fmax(s2 - s * (s / n), 0.0) / (double)(n - 1);
            v += var / cnt * k * k;
            const double mean = s / n * k;
            m += mean * mean;
            if (prev) {
                const double g = (double)den[(size_t)p * 3 + c] - (double)prev[(size_t)p * 3 + c];
                dd += g * g;
            }
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        e += __shfl_down(e, off);
        v += __shfl_down(v, off);
        m += __shfl_down(m, off);
        dd += __shfl_down(dd, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][wv] = e; red[1][wv] = v; red[2][wv] = m; red[3][wv] = dd; }
    __syncthreads();
    if (threadIdx.x < 4) {
        double t = 0.0;
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) t += red[threadIdx.x][i];
        atomicAdd(lines + (size_t)(blockIdx.x % 64) * 32 + threadIdx.x, t);
    }
}

__global__ void k_dn_scale(const float* den, float scale, float* out, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out[i] = den[i] * scale;
}

// ---------------------------------------------------------------------------------
// Host side.
struct Layer {
    int cin1, cin2, cout;  // input split (concat), output channels
    bool relu;
    bool up1;              // the first input is the half-resolution decoder tensor, upsampled as it is read
    size_t w_off, b_off;   // into the FP32 parameter vector
    u16 *wf = nullptr, *wb = nullptr;
};

}  // namespace ptdn

using namespace ptdn;

struct pt_denoiser {
    pt_ctx* ctx = nullptr;
    int cus = 256;  // compute units of the context's device
    int maxN = 0, maxH = 0, maxW = 0;
    std::vector<Layer> L;
    size_t nparam = 0;
    int64_t step = 0;
    int loss_kind = PT_DN_LOSS_L1_LOG;
    DevAlloc mem;
    float *w = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
    u16* wbf = nullptr;  // BF16 copies, 2 per weight
    // activations (BF16) and gradients
    u16 *x0 = nullptr, *e0a = nullptr, *e0 = nullptr, *p1 = nullptr, *e1a = nullptr, *e1 = nullptr, *p2 = nullptr,
        *ba = nullptr, *bb = nullptr, *d1a = nullptr, *d1 = nullptr, *d0a = nullptr, *d0 = nullptr, *out = nullptr;
    u16 *ga = nullptr, *gb = nullptr, *gc = nullptr, *gs1 = nullptr, *gs0 = nullptr;  // gradient scratch
    float *feat = nullptr, *tgt = nullptr, *den = nullptr, *loss = nullptr;
    std::string err;
};

namespace {

int dn_fail(pt_denoiser* d, int st, const std::string& msg) {
    d->err = msg;
    return st;
}

void conv(pt_denoiser* d, const Layer& l, const u16* x1, const u16* x2, u16* y, int N, int H, int W, bool bwd_data,
          u16* y2 = nullptr, int co1 = 0) {
    ConvArgs a{};
    a.N = N;
    a.H = H;
    a.W = W;
    if (!bwd_data) {
        a.X1 = x1; a.X2 = x2; a.C1 = l.cin1; a.C2 = l.cin2;
        a.up1 = l.up1 ? 1 : 0;
        a.Wt = l.wf; a.bias = d->w + l.b_off;
        a.Y1 = y; a.Y2 = nullptr; a.Co1 = l.cout; a.Cout = l.cout;
        a.relu = l.relu; a.flip = 0;
    } else {  // input = dZ (cout channels), output = dX split (cin1 | cin2)
        a.X1 = x1; a.X2 = nullptr; a.C1 = l.cout; a.C2 = 0;
        a.Wt = l.wb; a.bias = nullptr;
        a.Y1 = y; a.Y2 = y2; a.Co1 = co1 ? co1 : l.cin1 + l.cin2; a.Cout = l.cin1 + l.cin2;
        a.relu = 0; a.flip = 1;
    }
    // tiles of 8 x 32 pixels x 64 output channels (32 below 64 channels); each
    // block takes consecutive tiles so that the grid is about two rounds of the
    // blocks the CUs hold at once (2 per CU at 64 channels, 3 at 32), at most 16
    constexpr int kTH = 8;
    const int nc = a.Cout >= 64 ? 2 : 1;
    const int ntiles = N * ((H + kTH - 1) / kTH) * ((W + 31) / 32);
    const int cotiles = (a.Cout + 32 * nc - 1) / (32 * nc);
    const int64_t resident = (int64_t)d->cus * (nc == 2 ? 2 : 3);
    const int64_t want = ((int64_t)ntiles * cotiles + 2 * resident - 1) / (2 * resident);
    const bool one = a.C1 + a.C2 <= 32;  // one input chunk: the weights stay in LDS for all the block's tiles
    a.tiles_per_block = one ? (int)std::max<int64_t>(1, std::min<int64_t>(16, want)) : 1;
    dim3 grid((unsigned)((ntiles + a.tiles_per_block - 1) / a.tiles_per_block), cotiles);
    if (nc == 2) {
        if (one) ptk::launch(d->ctx, "k_dn_conv", k_conv_t<2, kTH, true>, grid, dim3(kThreads), a);
        else ptk::launch(d->ctx, "k_dn_conv", k_conv_t<2, kTH, false>, grid, dim3(kThreads), a);
    } else {
        if (one) ptk::launch(d->ctx, "k_dn_conv", k_conv_t<1, kTH, true>, grid, dim3(kThreads), a);
        else ptk::launch(d->ctx, "k_dn_conv", k_conv_t<1, kTH, false>, grid, dim3(kThreads), a);
    }
}

void wgrad(pt_denoiser* d, const Layer& l, const u16* dz, const u16* x1, const u16* x2, int N, int H, int W) {
    WgradArgs a{};
    a.dZ = dz; a.Cout = l.cout; a.X1 = x1; a.X2 = x2; a.C1 = l.cin1; a.C2 = l.cin2;
    a.up1 = l.up1 ? 1 : 0;
    a.dW = d->g + l.w_off; a.dB = d->g + l.b_off;
    a.N = N; a.H = H; a.W = W;
    const int P = N * H * W;
    const int cin = l.cin1 + l.cin2;
    // about 2048 pixels per block, at least 16 blocks per (tap, tile)
    a.chunk = std::max(64, std::min(2048, (P / 16 + 63) / 64 * 64));
    dim3 grid((P + a.chunk - 1) / a.chunk, (l.cout + 31) / 32, 9 * ((cin + 31) / 32));
    ptk::launch(d->ctx, "k_dn_wgrad", k_conv_wgrad, grid, dim3(kThreads), a);
}

dim3 ew_grid(size_t n) { return dim3((unsigned)std::min<size_t>((n + 255) / 256, 8192)); }

// Forward pass of features already in d->feat (N images H x W).
void forward(pt_denoiser* d, int N, int H, int W) {
    pt_ctx* ctx = d->ctx;
    const int P = N * H * W, P1 = P / 4, P2 = P / 16;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    ptk::launch(ctx, "k_dn_prep", k_prep, ew_grid(P), dim3(256), (const float*)d->feat, d->x0, (size_t)P);
    conv(d, d->L[0], d->x0, nullptr, d->e0a, N, H, W, false);
    conv(d, d->L[1], d->e0a, nullptr, d->e0, N, H, W, false);
    ptk::launch(ctx, "k_dn_pool", k_pool2, ew_grid((size_t)P1 * 32 / 8), dim3(256), (const u16*)d->e0, d->p1, N, H, W, 32);
    conv(d, d->L[2], d->p1, nullptr, d->e1a, N, H1, W1, false);
    conv(d, d->L[3], d->e1a, nullptr, d->e1, N, H1, W1, false);
    ptk::launch(ctx, "k_dn_pool", k_pool2, ew_grid((size_t)P2 * 64 / 8), dim3(256), (const u16*)d->e1, d->p2, N, H1, W1, 64);
    conv(d, d->L[4], d->p2, nullptr, d->ba, N, H2, W2, false);
    conv(d, d->L[5], d->ba, nullptr, d->bb, N, H2, W2, false);
    conv(d, d->L[6], d->bb, d->e1, d->d1a, N, H1, W1, false);  // bb upsampled as it is read
    conv(d, d->L[7], d->d1a, nullptr, d->d1, N, H1, W1, false);
    conv(d, d->L[8], d->d1, d->e0, d->d0a, N, H, W, false);    // d1 upsampled as it is read
    conv(d, d->L[9], d->d0a, nullptr, d->d0, N, H, W, false);
    conv(d, d->L[10], d->d0, nullptr, d->out, N, H, W, false);
}

// Backward pass: d->gc holds dL/d(out) (kOutC channels); gradients summed into d->g.
void backward(pt_denoiser* d, int N, int H, int W) {
    pt_ctx* ctx = d->ctx;
    const int P = N * H * W, P1 = P / 4, P2 = P / 16;
    const int H1 = H / 2, W1 = W / 2, H2 = H / 4, W2 = W / 4;
    auto relu_bwd = [&](const u16* dy, const u16* add, const u16* y, u16* dz, size_t n) {
        ptk::launch(ctx, "k_dn_relu_bwd", k_relu_bwd, ew_grid(n), dim3(256), dy, add, y, dz, n);
    };
    // out = conv10(d0): no activation
    wgrad(d, d->L[10], d->gc, d->d0, nullptr, N, H, W);
    conv(d, d->L[10], d->gc, nullptr, d->ga, N, H, W, true);          // dL/d d0 (32)
    relu_bwd(d->ga, nullptr, d->d0, d->gb, (size_t)P * 32);            // dZ9
    wgrad(d, d->L[9], d->gb, d->d0a, nullptr, N, H, W);
    conv(d, d->L[9], d->gb, nullptr, d->ga, N, H, W, true);           // dL/d d0a (32)
    relu_bwd(d->ga, nullptr, d->d0a, d->gb, (size_t)P * 32);           // dZ8
    wgrad(d, d->L[8], d->gb, d->d1, d->e0, N, H, W);
    conv(d, d->L[8], d->gb, nullptr, d->gc, N, H, W, true, d->gs0, 64);  // dL/d up(d1) (64) | dL/d e0 via skip (32)
    ptk::launch(ctx, "k_dn_up_bwd", k_up2_bwd, ew_grid((size_t)P1 * 64 / 8), dim3(256), (const u16*)d->gc, d->ga, N, H1, W1,
                64);                                                   // dL/d d1 (64)
    relu_bwd(d->ga, nullptr, d->d1, d->gb, (size_t)P1 * 64);           // dZ7
    wgrad(d, d->L[7], d->gb, d->d1a, nullptr, N, H1, W1);
    conv(d, d->L[7], d->gb, nullptr, d->ga, N, H1, W1, true);         // dL/d d1a (64)
    relu_bwd(d->ga, nullptr, d->d1a, d->gb, (size_t)P1 * 64);          // dZ6
    wgrad(d, d->L[6], d->gb, d->bb, d->e1, N, H1, W1);
    conv(d, d->L[6], d->gb, nullptr, d->gc, N, H1, W1, true, d->gs1, 128);  // dL/d up(bb) (128) | dL/d e1 via skip (64)
    ptk::launch(ctx, "k_dn_up_bwd", k_up2_bwd, ew_grid((size_t)P2 * 128 / 8), dim3(256), (const u16*)d->gc, d->ga, N, H2,
                W2, 128);                                              // dL/d bb (128)
    relu_bwd(d->ga, nullptr, d->bb, d->gb, (size_t)P2 * 128);          // dZ5
    wgrad(d, d->L[5], d->gb, d->ba, nullptr, N, H2, W2);
    conv(d, d->L[5], d->gb, nullptr, d->ga, N, H2, W2, true);         // dL/d ba
    relu_bwd(d->ga, nullptr, d->ba, d->gb, (size_t)P2 * 128);          // dZ4
    wgrad(d, d->L[4], d->gb, d->p2, nullptr, N, H2, W2);
    conv(d, d->L[4], d->gb, nullptr, d->ga, N, H2, W2, true);         // dL/d p2 (64)
    ptk::launch(ctx, "k_dn_pool_bwd", k_pool2_bwd, ew_grid((size_t)P1 * 64 / 8), dim3(256), (const u16*)d->ga, d->gc, N, H1,
                W1, 64);                                               // dL/d e1 via pool (64)
    relu_bwd(d->gc, d->gs1, d->e1, d->gb, (size_t)P1 * 64);            // dZ3 (pool + skip)
    wgrad(d, d->L[3], d->gb, d->e1a, nullptr, N, H1, W1);
    conv(d, d->L[3], d->gb, nullptr, d->ga, N, H1, W1, true);         // dL/d e1a
    relu_bwd(d->ga, nullptr, d->e1a, d->gb, (size_t)P1 * 64);          // dZ2
    wgrad(d, d->L[2], d->gb, d->p1, nullptr, N, H1, W1);
    conv(d, d->L[2], d->gb, nullptr, d->ga, N, H1, W1, true);         // dL/d p1 (32)
    ptk::launch(ctx, "k_dn_pool_bwd", k_pool2_bwd, ew_grid((size_t)P * 32 / 8), dim3(256), (const u16*)d->ga, d->gc, N, H, W,
                32);                                                   // dL/d e0 via pool (32)
    relu_bwd(d->gc, d->gs0, d->e0, d->gb, (size_t)P * 32);             // dZ1 (pool + skip)
    wgrad(d, d->L[1], d->gb, d->e0a, nullptr, N, H, W);
    conv(d, d->L[1], d->gb, nullptr, d->ga, N, H, W, true);           // dL/d e0a
    relu_bwd(d->ga, nullptr, d->e0a, d->gb, (size_t)P * 32);           // dZ0
    wgrad(d, d->L[0], d->gb, d->x0, nullptr, N, H, W);
}

void refresh_bf16(pt_denoiser* d) {
    for (const Layer& l : d->L)
        ptk::launch(d->ctx, "k_dn_wcopy", k_wcopy, ew_grid((size_t)9 * l.cout * (l.cin1 + l.cin2)), dim3(256),
                    (const float*)(d->w + l.w_off), l.wf, l.wb, l.cout, l.cin1 + l.cin2);
}

int upload(pt_denoiser* d, float* dst, const float* src, size_t n) {
    hipError_t e = hipMemcpyAsync(dst, src, n * 4, hipMemcpyHostToDevice, d->ctx->stream);
    if (e != hipSuccess) return dn_fail(d, ptk::hip_fail(d->ctx, e, "denoiser upload"), d->ctx->err);
    return PT_OK;
}

int check_shape(pt_denoiser* d, int N, int H, int W) {
    if (N <= 0 || H <= 0 || W <= 0 || N > d->maxN || H > d->maxH || W > d->maxW || H % 4 || W % 4)
        return dn_fail(d, PT_ERR_INVALID_ARG, "denoiser: batch shape outside the network's limits (H, W multiples of 4)");
    return PT_OK;
}

}  // namespace

namespace ptk {

int dn_fits(const pt_denoiser* d, int W, int H) {
    return d && W > 0 && H > 0 && W <= d->maxW && H <= d->maxH && W % 4 == 0 && H % 4 == 0;
}

pt_ctx* dn_ctx(const pt_denoiser* d) { return d ? d->ctx : nullptr; }

namespace {
// Denoises half `which` of the request into d->den (texel units / scale).
void dn_half(pt_denoiser* d, const DnRequest& q, int which, float inv_scale) {
    const size_t P = (size_t)q.W * q.H;
    ptk::launch(d->ctx, "k_dn_assemble", k_dn_assemble, ew_grid(P), dim3(256), q, which, inv_scale, d->feat);
    forward(d, 1, q.H, q.W);
    ptk::launch(d->ctx, "k_dn_head", k_head, ew_grid(P), dim3(256), (const float*)d->feat, (const u16*)d->out,
                (const float*)nullptr, d->den, (u16*)nullptr, (float*)nullptr, P, 0);
}
// Both halves in one batch of 2: D(even half) into d->den, D(odd half) into
// d->den + 3 P. Each image's pixels are computed as in a batch of one, so the
// results equal two dn_half calls; a small image's two passes share the launches.
void dn_halves(pt_denoiser* d, const DnRequest& q, float inv_scale) {
    const size_t P = (size_t)q.W * q.H;
    ptk::launch(d->ctx, "k_dn_assemble", k_dn_assemble, ew_grid(P), dim3(256), q, 1, inv_scale, d->feat);
    ptk::launch(d->ctx, "k_dn_assemble", k_dn_assemble, ew_grid(P), dim3(256), q, 2, inv_scale, d->feat + P * kFeat);
    forward(d, 2, q.H, q.W);
    ptk::launch(d->ctx, "k_dn_head", k_head, ew_grid(2 * P), dim3(256), (const float*)d->feat, (const u16*)d->out,
                (const float*)nullptr, d->den, (u16*)nullptr, (float*)nullptr, 2 * P, 0);
}
}  // namespace

int dn_request_estimate(pt_denoiser* d, const DnRequest& q, float scale, int spp, double* d_lines, double* rel) {
    if (!dn_fits(d, q.W, q.H)) return dn_fail(d, PT_ERR_INVALID_ARG, "denoiser: request outside the network's limits");
    pt_ctx* ctx = d->ctx;
    const size_t P = (size_t)q.W * q.H;
    const float inv = 1.0f / scale;
    hipError_t e = hipMemsetAsync(d_lines, 0, 2 * 64 * 32 * sizeof(double), ctx->stream);
    if (e != hipSuccess) return dn_fail(d, ptk::hip_fail(ctx, e, "denoiser estimate"), ctx->err);
    const dim3 grid((unsigned)std::min<size_t>((P + 255) / 256, 2048));
    if (d->maxN >= 2) {
        dn_halves(d, q, inv);
        // D(even half) against the odd half; D(odd half) against the even half and D(even half)
        ptk::launch(ctx, "k_dn_cross", k_dn_cross, grid, dim3(256), q, (const float*)d->den, (const float*)nullptr, 2,
                    inv, d_lines);
        ptk::launch(ctx, "k_dn_cross", k_dn_cross, grid, dim3(256), q, (const float*)(d->den + P * 3),
                    (const float*)d->den, 1, inv, d_lines + 64 * 32);
    } else {
        dn_half(d, q, 1, inv);  // D(even half) against the odd half
        ptk::launch(ctx, "k_dn_cross", k_dn_cross, grid, dim3(256), q, (const float*)d->den, (const float*)nullptr, 2,
                    inv, d_lines);
        // keep D(A) (the training-target buffer is free during inference)
        e = hipMemcpyAsync(d->tgt, d->den, P * 3 * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream);
        if (e != hipSuccess) return dn_fail(d, ptk::hip_fail(ctx, e, "denoiser estimate"), ctx->err);
        dn_half(d, q, 2, inv);  // D(odd half) against the even half, and against D(even half)
        ptk::launch(ctx, "k_dn_cross", k_dn_cross, grid, dim3(256), q, (const float*)d->den, (const float*)d->tgt, 1,
                    inv, d_lines + 64 * 32);
    }
    std::vector<double> h((size_t)2 * 64 * 32);
    e = hipMemcpyAsync(h.data(), d_lines, h.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return dn_fail(d, ptk::hip_fail(ctx, e, "denoiser estimate"), ctx->err);
    double s[2][4] = {};
    for (int k = 0; k < 2; ++k)
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < 4; ++j) s[k][j] += h[(size_t)k * 64 * 32 + (size_t)l * 32 + j];
    // E|D(A) - B|^2 = |D(A) - mu|^2 + E|B - mu|^2 for independent halves A, B:
    // the squared error of a denoised half image, bias included
    const double err2_half = 0.5 * ((s[0][0] - s[0][1]) + (s[1][0] - s[1][1]));
    // E|D(A) - D(B)|^2 = 2 var(D(half)). With the bias independent of the
    // sample count and the variance falling as 1/n, the full image's squared
    // error is err2_half - var(D(half)) / 2. At few samples the network's
    // output variance falls more slowly than 1/n (it smooths low-sample
    // inputs harder), so that credit is phased in linearly up to 64 samples.
    const double credit = std::min(1.0, spp / 64.0);
    const double err2 = err2_half - credit * 0.25 * s[1][3];
    const double m = s[0][2];
    // Both terms are differences of sums. On heavy-tailed scenes at few samples
    // (fireflies in a half's variance) they can come out zero or negative: the
    // estimate is then inconclusive, never a zero error that would finish the
    // request denoised.
    *rel = (m > 0.0 && err2_half > 0.0 && err2 > 0.0) ? std::sqrt(err2 / m) : INFINITY;
    return PT_OK;
}

int dn_request_denoise(pt_denoiser* d, const DnRequest& q, float scale, float* d_out) {
    if (!dn_fits(d, q.W, q.H)) return dn_fail(d, PT_ERR_INVALID_ARG, "denoiser: request outside the network's limits");
    const size_t P = (size_t)q.W * q.H;
    dn_half(d, q, 0, 1.0f / scale);
    ptk::launch(d->ctx, "k_dn_scale", k_dn_scale, ew_grid(P * 3), dim3(256), (const float*)d->den, scale, d_out,
                (int)(P * 3));
    return PT_OK;
}

}  // namespace ptk

extern "C" {

pt_status pt_dn_create(pt_ctx* ctx, int max_batch, int max_height, int max_width, uint64_t seed, pt_denoiser** out) {
    if (!ctx || !out || max_batch <= 0 || max_height <= 0 || max_width <= 0 || max_height % 4 || max_width % 4)
        return PT_ERR_INVALID_ARG;
    *out = nullptr;
    pt_denoiser* d = new (std::nothrow) pt_denoiser();
    if (!d) return PT_ERR_INVALID_ARG;
    d->ctx = ctx;
    {
        int cu = 0;
        if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, ctx->device) == hipSuccess && cu > 0)
            d->cus = cu;
    }
    d->maxN = max_batch;
    d->maxH = max_height;
    d->maxW = max_width;
    const int spec[11][4] = {{kIn, 0, 32, 1}, {32, 0, 32, 1}, {32, 0, 64, 1}, {64, 0, 64, 1}, {64, 0, 128, 1},
                             {128, 0, 128, 1}, {128, 64, 64, 1}, {64, 0, 64, 1}, {64, 32, 32, 1}, {32, 0, 32, 1},
                             {32, 0, kOutC, 0}};
    size_t off = 0;
    for (auto& sp : spec) {
        Layer l;
        l.cin1 = sp[0]; l.cin2 = sp[1]; l.cout = sp[2]; l.relu = sp[3] != 0;
        l.up1 = l.cin2 > 0;  // the decoder layers that concatenate a skip connection
        l.w_off = off;
        off += (size_t)9 * l.cout * (l.cin1 + l.cin2);
        l.b_off = off;
        off += (size_t)l.cout;
        d->L.push_back(l);
        // k_conv_t reads and writes 8 channels per 16-byte vector, splits the
        // input at C1 and the output at Co1, and skips the second 16-channel
        // step of a 16-channel input: channel counts of both passes (forward,
        // backward-data) are multiples of 16
        const int cin = l.cin1 + l.cin2;
        if (cin % 16 || l.cout % 16 || l.cin1 % 16) {
            delete d;
            return PT_ERR_INVALID_ARG;
        }
    }
    d->nparam = off;
    const size_t P = (size_t)max_batch * max_height * max_width;
    // activations: BF16 element counts
    const size_t act[] = {P * kIn, P * 32, P * 32, P / 4 * 32, P / 4 * 64, P / 4 * 64, P / 16 * 64, P / 16 * 128,
                          P / 16 * 128, P / 4 * 64, P / 4 * 64, P * 32, P * 32, P * kOutC,
                          P * 64, P * 64, P * 64, P / 4 * 64, P * 32};
    size_t bytes = 4 * off * 4 + 2 * off * 2 + (P * kFeat + P * 3 + P * 3 + 64) * 4 + 64 * 256;
    for (size_t a : act) bytes += (a * 2 + 255) / 256 * 256;
    int st = ptk::dev_alloc(ctx, d->mem, bytes);
    if (st != PT_OK) {
        delete d;
        return st;
    }
    char* p = static_cast<char*>(d->mem.p);
    auto carve = [&](size_t b) {
        char* q = p;
        p += (b + 255) / 256 * 256;
        return q;
    };
    d->w = reinterpret_cast<float*>(carve(off * 4));
    d->g = reinterpret_cast<float*>(carve(off * 4));
    d->m = reinterpret_cast<float*>(carve(off * 4));
    d->v = reinterpret_cast<float*>(carve(off * 4));
    d->wbf = reinterpret_cast<u16*>(carve(off * 4));
    u16** acts[] = {&d->x0, &d->e0a, &d->e0, &d->p1, &d->e1a, &d->e1, &d->p2, &d->ba, &d->bb, &d->d1a, &d->d1,
                    &d->d0a, &d->d0, &d->out, &d->ga, &d->gb, &d->gc, &d->gs1, &d->gs0};
    static_assert(sizeof(act) / sizeof(act[0]) == sizeof(acts) / sizeof(acts[0]), "one size per activation");
    for (size_t i = 0; i < sizeof(act) / sizeof(act[0]); ++i) *acts[i] = reinterpret_cast<u16*>(carve(act[i] * 2));
    d->feat = reinterpret_cast<float*>(carve(P * kFeat * 4));
    d->tgt = reinterpret_cast<float*>(carve(P * 3 * 4));
    d->den = reinterpret_cast<float*>(carve(P * 3 * 4));
    d->loss = reinterpret_cast<float*>(carve(64 * 4));
    u16* wb = d->wbf;
    for (Layer& l : d->L) {
        const size_t n = (size_t)9 * l.cout * (l.cin1 + l.cin2);
        l.wf = wb;
        l.wb = wb + n;
        wb += 2 * n;
    }
    // He-normal weights, zero biases
    std::vector<float> h(off, 0.0f);
    std::mt19937_64 rng(seed);
    for (const Layer& l : d->L) {
        const int cin = l.cin1 + l.cin2;
        std::normal_distribution<float> nd(0.0f, std::sqrt(2.0f / (9.0f * (float)cin)));
        for (size_t i = 0; i < (size_t)9 * l.cout * cin; ++i) h[l.w_off + i] = nd(rng);
    }
    for (int t = 0; t < 9; ++t)  // the output layer starts small: the network starts near the identity
        for (int co = 0; co < kOutC; ++co)
            for (int ci = 0; ci < 32; ++ci) h[d->L[10].w_off + ((size_t)t * kOutC + co) * 32 + ci] *= 0.1f;
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemcpyAsync(d->w, h.data(), off * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d->m, 0, off * 8, ctx->stream);  // m and v are adjacent
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        ptk::dev_free(ctx, d->mem);
        delete d;
        return ptk::hip_fail(ctx, e, "pt_dn_create");
    }
    refresh_bf16(d);
    *out = d;
    return PT_OK;
}

void pt_dn_destroy(pt_denoiser* d) {
    if (!d) return;
    (void)hipSetDevice(d->ctx->device);
    (void)hipStreamSynchronize(d->ctx->stream);
    ptk::dev_free(d->ctx, d->mem);
    delete d;
}

const char* pt_dn_last_error(const pt_denoiser* d) { return d ? d->err.c_str() : ""; }

int64_t pt_dn_param_count(const pt_denoiser* d) { return d ? (int64_t)d->nparam : 0; }

pt_status pt_dn_set_loss(pt_denoiser* d, int kind) {
    if (!d || (kind != PT_DN_LOSS_L1_LOG && kind != PT_DN_LOSS_L2)) return PT_ERR_INVALID_ARG;
    d->loss_kind = kind;
    return PT_OK;
}

pt_status pt_dn_get_params(pt_denoiser* d, float* host, size_t n) {
    if (!d || !host || n != d->nparam) return PT_ERR_INVALID_ARG;
    hipError_t e = hipMemcpyAsync(host, d->w, n * 4, hipMemcpyDeviceToHost, d->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->ctx->stream);
    return e == hipSuccess ? PT_OK : dn_fail(d, ptk::hip_fail(d->ctx, e, "pt_dn_get_params"), d->ctx->err);
}

pt_status pt_dn_set_params(pt_denoiser* d, const float* host, size_t n) {
    if (!d || !host || n != d->nparam) return PT_ERR_INVALID_ARG;
    int st = upload(d, d->w, host, n);
    if (st != PT_OK) return st;
    refresh_bf16(d);
    hipError_t e = hipStreamSynchronize(d->ctx->stream);
    return e == hipSuccess ? PT_OK : dn_fail(d, ptk::hip_fail(d->ctx, e, "pt_dn_set_params"), d->ctx->err);
}

pt_status pt_dn_denoise(pt_denoiser* d, const float* features, int N, int H, int W, float* denoised) {
    if (!d || !features || !denoised) return PT_ERR_INVALID_ARG;
    int st = check_shape(d, N, H, W);
    if (st != PT_OK) return st;
    const size_t P = (size_t)N * H * W;
    (void)hipSetDevice(d->ctx->device);
    st = upload(d, d->feat, features, P * kFeat);
    if (st != PT_OK) return st;
    forward(d, N, H, W);
    ptk::launch(d->ctx, "k_dn_head", k_head, ew_grid(P), dim3(256), (const float*)d->feat, (const u16*)d->out,
                (const float*)nullptr, d->den, (u16*)nullptr, (float*)nullptr, P, 0);
    hipError_t e = hipMemcpyAsync(denoised, d->den, P * 3 * 4, hipMemcpyDeviceToHost, d->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(d->ctx->stream);
    return e == hipSuccess ? PT_OK : dn_fail(d, ptk::hip_fail(d->ctx, e, "pt_dn_denoise"), d->ctx->err);
}

pt_status pt_dn_gradients(pt_denoiser* d, const float* features, const float* target, int N, int H, int W,
                          float* loss, float* host_grads, size_t n_grads) {
    if (!d || !features || !target) return PT_ERR_INVALID_ARG;
    if (host_grads && n_grads != d->nparam) return PT_ERR_INVALID_ARG;
    int st = check_shape(d, N, H, W);
    if (st != PT_OK) return st;
    const size_t P = (size_t)N * H * W;
    pt_ctx* ctx = d->ctx;
    (void)hipSetDevice(ctx->device);
    st = upload(d, d->feat, features, P * kFeat);
    if (st == PT_OK) st = upload(d, d->tgt, target, P * 3);
    if (st != PT_OK) return st;
    hipError_t e = hipMemsetAsync(d->g, 0, d->nparam * 4, ctx->stream);
    if (e == hipSuccess) e = hipMemsetAsync(d->loss, 0, 4, ctx->stream);
    if (e != hipSuccess) return dn_fail(d, ptk::hip_fail(ctx, e, "pt_dn_gradients"), ctx->err);
    forward(d, N, H, W);
    ptk::launch(ctx, "k_dn_head", k_head, ew_grid(P), dim3(256), (const float*)d->feat, (const u16*)d->out,
                (const float*)d->tgt, (float*)nullptr, d->gc, d->loss, P, d->loss_kind);
    backward(d, N, H, W);
    if (loss) e = hipMemcpyAsync(loss, d->loss, 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess && host_grads) e = hipMemcpyAsync(host_grads, d->g, n_grads * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? PT_OK : dn_fail(d, ptk::hip_fail(ctx, e, "pt_dn_gradients"), ctx->err);
}

pt_status pt_dn_train_step(pt_denoiser* d, const float* features, const float* target, int N, int H, int W, float lr,
                           float* loss) {
    int st = pt_dn_gradients(d, features, target, N, H, W, loss, nullptr, 0);
    if (st != PT_OK) return st;
    d->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - std::pow(b1, (float)d->step), c2 = 1.0f - std::pow(b2, (float)d->step);
    ptk::launch(d->ctx, "k_dn_adam", k_adam, ew_grid(d->nparam), dim3(256), d->w, (const float*)d->g, d->m, d->v,
                d->nparam, lr, b1, b2, 1e-8f, c1, c2);
    refresh_bf16(d);
    hipError_t e = hipStreamSynchronize(d->ctx->stream);
    return e == hipSuccess ? PT_OK : dn_fail(d, ptk::hip_fail(d->ctx, e, "pt_dn_train_step"), d->ctx->err);
}

}  // extern "C"
