// tracer.hpp — internal definitions shared by the C ABI (capi.cpp) and the
// kernel translation units: device-side scene records, wave buffers, and the
// opaque pt_ctx / pt_scene / pt_accum objects.
#pragma once

#include <hip/hip_runtime.h>

#include <memory>
#include <string>
#include <vector>

#include "../../include/pathtracer.h"
#include "ptdev.h"
#include "scene.hpp"

namespace ptk {

using ptd::f3;
using ptd::m3;

constexpr float kMaxDist = 1e5f;  // MAXDIST, src/shader.comp:8
constexpr int kUboFloats = 4097;
constexpr int kUboObj = 7, kUboSdf = 7 + 1024, kUboMat = kUboSdf + 768, kUboLight = kUboMat + 783,
              kUboLightId = kUboLight + 128, kUboCie = kUboLightId + 64;

// ---- reference mode: objects decoded once from the uniform block -----------------
// Values that src/shader.comp recomputes per call from the packed floats
// (rotation matrices, integer IDs, bounding radii, lens slice offsets) are
// computed once here with the same operations, so they carry the same bits.
struct RefSphere { float px, py, pz, r, mat, light; };
struct RefPlane { float px, py, pz, mat, light; };
struct RefBox { f3 pos, size; m3 M; float bsr2, mat, light; };
struct RefLens {
    f3 pos; m3 M;
    float radius, half, slicePos, br2, twoF, mat, light;
    int conv;
};
struct RefCyclide { f3 pos, scale; m3 M; float a, b, c, d, brad2, mat, light; };
struct RefSdf { f3 pos, size; pt_sdf_desc d; };

struct RefSceneDev {
    const float* ubo;  // the 4097-float uniform block
    const RefSphere* sph;
    const RefPlane* pln;
    const RefBox* box;
    const RefLens* lens;
    const RefCyclide* cyc;
    const RefSdf* sdf;
    int ns, np, nb, nl, nc, nsdf;
    float numLights;   // numObjects[6]
    int sdf_kind;      // PT_SDF_* when all SDFs share one kind, 0 when mixed or none
};

// Per-dispatch constants (push constants with the camera precomputed).
struct RefFrame {
    int W, H, frame, spf, pathLength;
    float camSize, apSize, apDist;
    f3 camPos, fwd;
    m3 camM;
    // camera lens of TracePathLens (src/shader.comp:1409-1444)
    f3 lensPos;
    m3 lensM;
    float lensRadius, lensHalf, lensSlicePos, lensTwoF;
    // the wave: pixels [p0, p0 + pc) x samples [k0, k0 + m)
    int p0, pc, k0, m;
};

// SoA path state of one wave (capacity N paths). Records are written once and
// read once per stage.
struct RefWave {
    float *ox, *oy, *oz, *dx, *dy, *dz;
    uint32_t* seed;
    float* lh;
    float4* T;  // throughput per wavelength
    float4* L;  // radiance per wavelength
    float* w;   // MIS weight of the next emitter hit
    float *ht, *hnx, *hny, *hnz, *hmat, *hlight;
    int* q[2];  // ping-pong path queues
    // shadow (NEE) records
    int* s_path;
    float *sox, *soy, *soz, *sdx, *sdy, *sdz;
    int* s_lobj;
    float4* s_c;
    float* s_t;
    int* s_id;
    float* sum;      // per-pixel running sums, 3 x pc
    int* counts;     // [0..1] queue lengths, [2] shadow length
    unsigned long long* stats;  // pt_counters fields
    int cap;
};

// ---- extended mode ----------------------------------------------------------------
struct ExtScene;   // defined in ext_tracer.hip
struct ExtWave;

}  // namespace ptk

// ---- opaque C-ABI objects ------------------------------------------------------------
struct DevAlloc {
    void* p = nullptr;
    size_t bytes = 0;
};

struct ProfPending {
    int kernel;
    hipEvent_t start, stop;
};
struct ProfTotal {
    std::string name;
    uint64_t launches = 0;
    double ms = 0.0;
};

struct pt_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    pt_config cfg{};
    std::string err;
    // pt_dispatch_noalloc: whether this device's static wave was found at pt_create
    int static_wave_status = PT_OK;
    std::string static_wave_err;
    int64_t allocated = 0;
    unsigned long long* d_stats = nullptr;  // pt_counters fields, 16 slots
    // kernel timing (pt_profile_enable)
    bool prof = false;
    std::vector<ProfPending> prof_pending;
    std::vector<hipEvent_t> prof_pool;
    std::vector<ProfTotal> prof_totals;
    // reference-mode wave
    DevAlloc ref_mem;
    ptk::RefWave ref_wave{};
    // extended-mode wave (owned by ext_tracer.hip)
    void* ext_state = nullptr;
};

struct pt_scene {
    pt_ctx* ctx = nullptr;
    pt::SceneModel model;
    bool has_model = false;
    bool ref_ok = false;
    std::string ref_err;
    pt_packed_scene ubo{};
    std::vector<pt_sdf_desc> sdfs;
    // reference-mode device data (one allocation)
    DevAlloc ref_mem;
    ptk::RefSceneDev ref_dev{};
    // extended-mode device data (built on first use)
    void* ext = nullptr;
    int64_t device_bytes = 0;
};

struct pt_accum {
    pt_ctx* ctx = nullptr;
    int w = 0, h = 0;
    float* d_texels = nullptr;
    // pt_accum_create_exportable: the texels are a virtual-memory mapping of a
    // physical allocation (vmm_handle, vmm_bytes long) that can be exported
    bool vmm = false;
    hipMemGenericAllocationHandle_t vmm_handle{};
    size_t vmm_bytes = 0;
};

namespace ptk {

// Device memory with the context's budget applied. Returns PT_OK,
// PT_ERR_BUDGET or PT_ERR_HIP (message in ctx->err).
int dev_alloc(pt_ctx* ctx, DevAlloc& a, size_t bytes);
void dev_free(pt_ctx* ctx, DevAlloc& a);
int hip_fail(pt_ctx* ctx, hipError_t e, const char* what);

// Reference mode (ref_tracer.hip).
int ref_scene_upload(pt_ctx* ctx, pt_scene* s);
int ref_dispatch(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc);
void ref_wave_release(pt_ctx* ctx);
// The allocation-free variant: the same kernels through a static device wave.
// ref_static_setup looks the wave's address up on the context's device (once
// per device; call with that device current).
int ref_static_setup(pt_ctx* ctx);
int ref_dispatch_static(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc);

// Extended mode (ext_tracer.hip).
int ext_scene_build(pt_ctx* ctx, pt_scene* s);
void ext_scene_release(pt_ctx* ctx, pt_scene* s);
int ext_dispatch(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc);
// The samples [frame - spf, frame) of every pixel added to a request's
// per-pixel moments (sum, sumsq: 3 x P floats, channel-major; nsamp: P ints).
int ext_dispatch_moments(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, float* sum, float* sumsq,
                         int* nsamp);
// The samples [frame - spf, frame) of every pixel: per-pixel sums of the
// per-sample XYZ, first-hit features (albedo, normal, distance) and Y^2 into
// feat (12 floats per pixel, the last unused), sample counts into nsamp.
int ext_dispatch_features(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, float* feat, int* nsamp);
// One request of a batched wave: its scene, the push constants of its quantum
// (camera, resolution, pathLength; samples [frame - samplesPerFrame, frame)),
// and its per-pixel moments.
struct ExtBatchItem {
    const pt_scene* scene = nullptr;
    pt_push_constants pc{};
    float* sum = nullptr;
    float* sumsq = nullptr;
    int* nsamp = nullptr;
    float* sum_odd = nullptr;  // optional: sums / counts of the odd-index samples
    int* n_odd = nullptr;
};
constexpr int kMaxBatch = 64;

// A served request's buffers as the denoiser reads them: per-pixel moments
// (sum, sumsq: 3P floats channel-major; nsamp: P), sums and counts of the
// odd-index samples, and the sums / counts of a first-hit features pass
// (PT_FEATURES floats per pixel as ext_dispatch_features writes them).
struct DnRequest {
    const float* sum;
    const float* sumsq;
    const int* nsamp;
    const float* sum_odd;
    const int* n_odd;
    const float* aov;
    const int* aov_n;
    int W, H;
    float exposure;  // apertureSize^2 * ISO
};
// Whether the network takes a W x H image (within its limits, multiples of 4).
int dn_fits(const pt_denoiser* d, int W, int H);
pt_ctx* dn_ctx(const pt_denoiser* d);
// Dual-buffer estimate of the relative RMSE of the request's denoised image:
// each half-sample image is denoised and compared with the other, independent
// half, whose own variance is subtracted (the squared error of a denoised
// half, bias included); half the variance of a denoised half, from the
// difference of the two denoised halves, is then taken off for the full
// sample count (phased in up to 64 samples per pixel); over sqrt(sum of
// mean^2). Inputs are divided by `scale` (an RMS of the image). d_lines:
// device scratch of 2 x 64 x 32 doubles. Synchronous.
int dn_request_estimate(pt_denoiser* d, const DnRequest& q, float scale, int spp, double* d_lines, double* rel);
// The denoised image of all samples into d_out (3 floats per pixel, texel units).
int dn_request_denoise(pt_denoiser* d, const DnRequest& q, float scale, float* d_out);
// All items' paths in one wave (their total must fit the context's wave
// capacity); each request's samples are added to its moments in sample order,
// so a request's result does not depend on what else is in the wave.
int ext_dispatch_batch(pt_ctx* ctx, const ExtBatchItem* items, int n);
// Adds (sum over pixels and channels of variance of the mean, sum of mean^2)
// into d_out: 64 lines of 32 doubles, slots 0 and 1 of each line.
int ext_rel_error(pt_ctx* ctx, int P, const float* sum, const float* sumsq, const int* nsamp, double* d_out);
void ext_wave_release(pt_ctx* ctx);

// Display resolve and diagnostics (resolve.hip).
int resolve_srgb8(pt_ctx* ctx, pt_accum* acc, int tonemap, uint8_t* host_rgb);
int math_eval_device(pt_ctx* ctx, int fn, const float* x, const float* y, float* out, size_t n);
int peak_probe(pt_ctx* ctx, int what, size_t n, int reps, double* units);
int count_dispatch(pt_ctx* ctx);

// Kernel launch on the context's stream; with profiling on, bracketed by two
// HIP events whose elapsed time is added to the kernel's total on read.
int prof_kernel_id(pt_ctx* ctx, const char* name);
hipEvent_t prof_event(pt_ctx* ctx);
template <typename Kernel, typename... Args>
inline void launch(pt_ctx* ctx, const char* name, Kernel kernel, dim3 grid, dim3 block, Args... args) {
    if (!ctx->prof) {
        hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
        return;
    }
    ProfPending pp{prof_kernel_id(ctx, name), prof_event(ctx), prof_event(ctx)};
    (void)hipEventRecord(pp.start, ctx->stream);
    hipLaunchKernelGGL(kernel, grid, block, 0, ctx->stream, args...);
    (void)hipEventRecord(pp.stop, ctx->stream);
    ctx->prof_pending.push_back(pp);
}

}  // namespace ptk
