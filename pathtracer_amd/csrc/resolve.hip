// resolve.hip — display resolve of the XYZ accumulator and device-side
// diagnostics.
#include <algorithm>
#include <vector>

#include "tracer.hpp"

namespace ptk {

using namespace ptd;

// sRGBCompanding (src/pathtracer.cpp:1061-1070, src/shader.frag:46-55).
__device__ __forceinline__ float srgb_oetf(float x) {
    x = ptm_clamp(x, 0.0f, 1.0f);
    if (x <= 0.0031308f) return 12.92f * x;
    return 1.055f * ptm_pow(x, 1.0f / 2.4f) - 0.055f;
}
__device__ __forceinline__ float tonemap1(float x, int tm) {
    if (tm == 1) x = x / (1.0f + x);                                                     // Reinhard
    if (tm == 2) x = x * (2.51f * x + 0.03f) / (x * (2.43f * x + 0.59f) + 0.14f);       // ACES (Narkowicz 2016)
    if (tm == 3) x = ptm_exp(-0.25f / x);                                                // DEUCES
    return x;
}

// SaveRender's per-pixel conversion (src/pathtracer.cpp:3501-3511): Bradford
// E -> D65, XYZ -> linear sRGB (Lindbloom matrices), clamp at 0, tonemap,
// OETF, then truncation to 8 bits.
__global__ void k_resolve(const float4* texels, int n, int tm, uint8_t* rgb) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 t = texels[i];
    f3 c = mk3(t.x, t.y, t.z);
    f3 e = mk3(0.9531874f * c.x + -0.0265906f * c.y + 0.0238731f * c.z,
               -0.0382467f * c.x + 1.0288406f * c.y + 0.0094060f * c.z,
               0.0026068f * c.x + -0.0030332f * c.y + 1.0892565f * c.z);
    f3 r = mk3(3.2404542f * e.x + -1.5371385f * e.y + -0.4985314f * e.z,
               -0.9692660f * e.x + 1.8760108f * e.y + 0.0415560f * e.z,
               0.0556434f * e.x + -0.2040259f * e.y + 1.0572252f * e.z);
    r = max3(r, mk3(0.0f, 0.0f, 0.0f));
    float v[3] = {r.x, r.y, r.z};
    for (int k = 0; k < 3; ++k) {
        float x = srgb_oetf(tonemap1(v[k], tm));
        rgb[3 * i + k] = (uint8_t)(int)((double)x * 255.0);
    }
}

int resolve_srgb8(pt_ctx* ctx, pt_accum* acc, int tonemap, uint8_t* host_rgb) {
    int n = acc->w * acc->h;
    DevAlloc buf;
    int st = dev_alloc(ctx, buf, (size_t)n * 3);
    if (st != PT_OK) return st;
    launch(ctx, "k_resolve", k_resolve, dim3((n + 255) / 256), dim3(256), reinterpret_cast<const float4*>(acc->d_texels),
           n, tonemap, static_cast<uint8_t*>(buf.p));
    hipError_t e = hipMemcpyAsync(host_rgb, buf.p, (size_t)n * 3, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, buf);
    if (e != hipSuccess) return hip_fail(ctx, e, "pt_resolve_srgb8");
    return PT_OK;
}

__global__ void k_math(int fn, const float* x, const float* y, float* out, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a = x[i];
    float r;
    switch (fn) {
        case 0: r = ptm_sin(a); break;
        case 1: r = ptm_cos(a); break;
        case 2: r = ptm_exp(a); break;
        case 3: r = ptm_log(a); break;
        case 4: r = ptm_pow(a, y[i]); break;
        case 5: r = ptm_acos(a); break;
        case 6: r = ptm_sqrt(a); break;
        default: r = ptm_rsqrt(a); break;
    }
    out[i] = r;
}

int math_eval_device(pt_ctx* ctx, int fn, const float* x, const float* y, float* out, size_t n) {
    if (n == 0) return PT_OK;
    if (n > (size_t)1 << 26) {
        ctx->err = "pt_math_eval_device: at most 2^26 values per call";
        return PT_ERR_INVALID_ARG;
    }
    DevAlloc buf;
    int st = dev_alloc(ctx, buf, n * 12);
    if (st != PT_OK) return st;
    float* dx = static_cast<float*>(buf.p);
    float* dy = dx + n;
    float* dout = dy + n;
    hipError_t e = hipMemcpyAsync(dx, x, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && y) e = hipMemcpyAsync(dy, y, n * 4, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) {
        launch(ctx, "k_math", k_math, dim3((unsigned)((n + 255) / 256)), dim3(256), fn, (const float*)dx,
               (const float*)dy, dout, (int)n);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(out, dout, n * 4, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, buf);
    if (e != hipSuccess) return hip_fail(ctx, e, "pt_math_eval_device");
    return PT_OK;
}

// Measured device peaks for the rooflines. k_peak_copy streams n4 float4s from
// src to dst (grid-stride, 32 bytes of HBM traffic per element). k_peak_fma runs
// `iters` rounds of 8 independent fused multiply-adds per thread (explicit fmaf:
// this library is built without FP contraction, and without packed FP32, so the
// loop measures what unpacked FP32 code can issue) and stores one sum per thread
// so that nothing is dead.
__global__ void __launch_bounds__(256) k_peak_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) dst[i] = src[i];
}

__global__ void __launch_bounds__(256) k_peak_fma(float m, float c, int iters, float* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float a[8];
    for (int k = 0; k < 8; ++k) a[k] = (float)(t + k) * 1e-9f;
#pragma unroll 4
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) a[k] = __builtin_fmaf(a[k], m, c);
    }
    out[t] = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
}

int peak_probe(pt_ctx* ctx, int what, size_t n, int reps, double* units) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device) != hipSuccess || cus <= 0)
        cus = 256;
    DevAlloc buf;
    hipError_t e = hipSuccess;
    if (what == 0) {
        const size_t n4 = n / 4;
        int st = dev_alloc(ctx, buf, n * 8);
        if (st != PT_OK) return st;
        const float4* src = static_cast<const float4*>(buf.p);
        float4* dst = static_cast<float4*>(buf.p) + n4;
        e = hipMemsetAsync(buf.p, 0, n * 8, ctx->stream);
        const size_t want = (n4 + 255) / 256;
        const unsigned blocks = (unsigned)std::min<size_t>(want, (size_t)cus * 64);
        for (int r = 0; e == hipSuccess && r <= reps; ++r) {  // the first launch warms up
            launch(ctx, r == 0 ? "k_peak_copy_warmup" : "k_peak_copy", k_peak_copy, dim3(blocks), dim3(256), src, dst, n4);
            e = hipGetLastError();
        }
        *units = 8.0 * (double)n4 * 4.0;
    } else {
        const unsigned blocks = (unsigned)cus * 16;
        int st = dev_alloc(ctx, buf, (size_t)blocks * 256 * 4);
        if (st != PT_OK) return st;
        for (int r = 0; e == hipSuccess && r <= reps; ++r) {
            launch(ctx, r == 0 ? "k_peak_fma_warmup" : "k_peak_fma", k_peak_fma, dim3(blocks), dim3(256), 0.999999f,
                   1e-7f, (int)n, static_cast<float*>(buf.p));
            e = hipGetLastError();
        }
        *units = 16.0 * (double)n * (double)blocks * 256.0;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, buf);
    if (e != hipSuccess) return hip_fail(ctx, e, "pt_peak_probe");
    return PT_OK;
}

__global__ void k_bump_dispatch(unsigned long long* stats) { stats[5] += 1ull; }

int prof_kernel_id(pt_ctx* ctx, const char* name) {
    for (size_t i = 0; i < ctx->prof_totals.size(); ++i)
        if (ctx->prof_totals[i].name == name) return (int)i;
    ProfTotal t;
    t.name = name;
    ctx->prof_totals.push_back(t);
    return (int)ctx->prof_totals.size() - 1;
}

hipEvent_t prof_event(pt_ctx* ctx) {
    if (!ctx->prof_pool.empty()) {
        hipEvent_t e = ctx->prof_pool.back();
        ctx->prof_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

int count_dispatch(pt_ctx* ctx) {
    launch(ctx, "k_bump_dispatch", k_bump_dispatch, dim3(1), dim3(1), ctx->d_stats);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "dispatch counter");
    return PT_OK;
}

}  // namespace ptk
