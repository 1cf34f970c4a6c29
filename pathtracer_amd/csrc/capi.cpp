// capi.cpp — the C ABI declared in include/pathtracer.h.
//
// Errors never cross the boundary as exceptions: every entry point catches,
// stores a message (per context, or thread-local for context-free calls) and
// returns a PT_ERR_* status, as the reference's VkResult checks turn into one
// failure per call (src/pathtracer.cpp:3878-3880).
#include <hip/hip_runtime.h>

#include <cstring>
#include <exception>
#include <new>
#include <string>

#include "../../include/pathtracer.h"
#include "scene.hpp"
#include "tracer.hpp"

namespace {
thread_local std::string g_err;

int set_global(int st, const std::string& msg) {
    g_err = msg;
    return st;
}
int set_err(pt_ctx* ctx, int st, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return st;
}
}  // namespace

namespace ptk {

int hip_fail(pt_ctx* ctx, hipError_t e, const char* what) {
    ctx->err = std::string(what) + ": " + hipGetErrorString(e);
    return PT_ERR_HIP;
}

int dev_alloc(pt_ctx* ctx, DevAlloc& a, size_t bytes) {
    if (bytes == 0) bytes = 256;
    if (ctx->cfg.hbm_budget > 0 && ctx->allocated + (int64_t)bytes > ctx->cfg.hbm_budget) {
        ctx->err = "HBM budget exceeded: " + std::to_string(ctx->allocated + (int64_t)bytes) + " > " +
                   std::to_string(ctx->cfg.hbm_budget) + " bytes";
        return PT_ERR_BUDGET;
    }
    hipError_t e = hipMalloc(&a.p, bytes);
    if (e != hipSuccess) {
        a.p = nullptr;
        return hip_fail(ctx, e, "hipMalloc");
    }
    a.bytes = bytes;
    ctx->allocated += (int64_t)bytes;
    return PT_OK;
}

void dev_free(pt_ctx* ctx, DevAlloc& a) {
    if (a.p) {
        (void)hipFree(a.p);
        ctx->allocated -= (int64_t)a.bytes;
    }
    a.p = nullptr;
    a.bytes = 0;
}

}  // namespace ptk

extern "C" {

const char* pt_last_error(const pt_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }
const char* pt_last_error_global(void) { return g_err.c_str(); }

pt_status pt_create(pt_ctx** out, int device_ordinal, const pt_config* cfg) {
    if (!out) return set_global(PT_ERR_INVALID_ARG, "pt_create: out is NULL");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_global(PT_ERR_NO_DEVICE, "pt_create: no HIP device is visible");
    if (device_ordinal < 0 || device_ordinal >= ndev)
        return set_global(PT_ERR_INVALID_ARG, "pt_create: device ordinal out of range");
    pt_ctx* ctx = new (std::nothrow) pt_ctx();
    if (!ctx) return set_global(PT_ERR_INVALID_ARG, "pt_create: out of host memory");
    ctx->device = device_ordinal;
    if (cfg) ctx->cfg = *cfg;
    if (ctx->cfg.wave_paths < 0 || (ctx->cfg.block_size != 0 && ctx->cfg.block_size != 256) ||
        ctx->cfg.blocks_per_cu < 0 || ctx->cfg.blocks_per_cu > 64 || ctx->cfg.sdf_max_steps < 0 ||
        (ctx->cfg.bvh_width != 0 && ctx->cfg.bvh_width != 2 && ctx->cfg.bvh_width != 4 && ctx->cfg.bvh_width != 8) ||
        ctx->cfg.traversal < 0 || ctx->cfg.traversal > 2 || ctx->cfg.bvh_leaf < 0 || ctx->cfg.bvh_leaf > 16 ||
        ctx->cfg.graph_mode < 0 || ctx->cfg.graph_mode > 1 || ctx->cfg.march < 0 || ctx->cfg.march > 1 ||
        ctx->cfg.march_blocks_per_cu < 0 || ctx->cfg.march_blocks_per_cu > 64) {
        delete ctx;
        return set_global(PT_ERR_INVALID_ARG, "pt_create: invalid pt_config");
    }
    hipError_t e = hipSetDevice(device_ordinal);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    const size_t stat_bytes = (size_t)ptd::kStatLines * ptd::kStatStride * sizeof(unsigned long long);
    if (e == hipSuccess) e = hipMalloc(&ctx->d_stats, stat_bytes);
    if (e == hipSuccess) e = hipMemset(ctx->d_stats, 0, stat_bytes);
    if (e != hipSuccess) {
        std::string msg = std::string("pt_create: ") + hipGetErrorString(e);
        if (ctx->d_stats) (void)hipFree(ctx->d_stats);
        if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
        delete ctx;
        return set_global(PT_ERR_HIP, msg);
    }
    // The address of this device's static wave for pt_dispatch_noalloc; a
    // failure is reported by that entry point, not here.
    ctx->static_wave_status = ptk::ref_static_setup(ctx);
    ctx->static_wave_err = ctx->static_wave_status == PT_OK ? "" : ctx->err;
    ctx->err.clear();
    *out = ctx;
    return PT_OK;
}

void pt_destroy(pt_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ptk::ref_wave_release(ctx);
    ptk::ext_wave_release(ctx);
    for (const ProfPending& p : ctx->prof_pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (hipEvent_t ev : ctx->prof_pool) (void)hipEventDestroy(ev);
    if (ctx->d_stats) (void)hipFree(ctx->d_stats);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

static pt_status finish_scene(pt_ctx* ctx, pt_scene* s, pt_scene** out) {
    if (s->ref_ok) {
        int st = ptk::ref_scene_upload(ctx, s);
        if (st != PT_OK) {
            pt_scene_release(s);
            return st;
        }
    }
    *out = s;
    return PT_OK;
}

pt_status pt_scene_from_json(pt_ctx* ctx, const char* json, size_t len, pt_scene** out) {
    if (!ctx || !json || !out) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_scene_from_json: NULL argument");
    *out = nullptr;
    try {
        (void)hipSetDevice(ctx->device);
        pt_scene* s = new pt_scene();
        s->ctx = ctx;
        std::string err;
        int st = pt::load_scene_json(json, len, s->model, err);
        if (st != PT_OK) {
            delete s;
            return set_err(ctx, st, "scene JSON: " + err);
        }
        s->has_model = true;
        for (const pt::RSdf& d : s->model.sdfs) s->sdfs.push_back(d.desc);
        std::string perr;
        s->ref_ok = pt::pack_reference_ubo(s->model, s->ubo, perr) == PT_OK;
        if (!s->ref_ok) s->ref_err = perr;
        return finish_scene(ctx, s, out);
    } catch (const std::exception& e) {
        return set_err(ctx, PT_ERR_JSON, std::string("pt_scene_from_json: ") + e.what());
    }
}

pt_status pt_scene_from_packed(pt_ctx* ctx, const pt_packed_scene* ubo, const pt_sdf_desc* sdfs, int n_sdfs,
                               pt_scene** out) {
    if (!ctx || !ubo || !out || n_sdfs < 0 || (n_sdfs > 0 && !sdfs))
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_scene_from_packed: bad argument");
    *out = nullptr;
    try {
        (void)hipSetDevice(ctx->device);
        pt_scene* s = new pt_scene();
        s->ctx = ctx;
        s->ubo = *ubo;
        s->sdfs.assign(sdfs, sdfs + n_sdfs);
        s->ref_ok = true;
        return finish_scene(ctx, s, out);
    } catch (const std::exception& e) {
        return set_err(ctx, PT_ERR_INVALID_ARG, std::string("pt_scene_from_packed: ") + e.what());
    }
}

pt_status pt_scene_get_packed(const pt_scene* s, pt_packed_scene* out) {
    if (!s || !out) return set_global(PT_ERR_INVALID_ARG, "pt_scene_get_packed: NULL argument");
    if (!s->ref_ok) return set_global(PT_ERR_CAPACITY, "scene has no reference packing: " + s->ref_err);
    *out = s->ubo;
    return PT_OK;
}

pt_status pt_scene_get_sdfs(const pt_scene* s, pt_sdf_desc* out, int cap, int* n_out) {
    if (!s || !n_out || cap < 0 || (cap > 0 && !out)) return set_global(PT_ERR_INVALID_ARG, "pt_scene_get_sdfs: bad argument");
    *n_out = (int)s->sdfs.size();
    for (int i = 0; i < cap && i < (int)s->sdfs.size(); ++i) out[i] = s->sdfs[(size_t)i];
    return PT_OK;
}

pt_status pt_scene_camera(const pt_scene* s, int shot, pt_push_constants* pc) {
    if (!s || !pc) return set_global(PT_ERR_INVALID_ARG, "pt_scene_camera: NULL argument");
    if (!s->has_model) return set_global(PT_ERR_STATE, "pt_scene_camera: scene was not loaded from JSON");
    std::string err;
    int st = pt::camera_to_push(s->model, shot, *pc, err);
    if (st != PT_OK) return set_global(st, err);
    return PT_OK;
}

pt_status pt_scene_device_bytes(const pt_scene* s, int64_t* bytes) {
    if (!s || !bytes) return set_global(PT_ERR_INVALID_ARG, "pt_scene_device_bytes: NULL argument");
    *bytes = s->device_bytes;
    return PT_OK;
}

void pt_scene_release(pt_scene* s) {
    if (!s) return;
    pt_ctx* ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ptk::dev_free(ctx, s->ref_mem);
    ptk::ext_scene_release(ctx, s);
    delete s;
}

pt_status pt_sdf_recognize(const char* glsl, size_t len, pt_sdf_desc* out) {
    if (!glsl || !out) return set_global(PT_ERR_INVALID_ARG, "pt_sdf_recognize: NULL argument");
    try {
        std::string err;
        int st = pt::recognize_sdf(std::string(glsl, len), *out, err);
        if (st != PT_OK) return set_global(st, err);
        return PT_OK;
    } catch (const std::exception& e) {
        return set_global(PT_ERR_UNSUPPORTED_SDF, e.what());
    }
}

pt_status pt_pack_json(const char* json, size_t len, int shot, pt_packed_scene* ubo, pt_push_constants* pc,
                       pt_sdf_desc* sdfs, int cap, int* n_sdfs) {
    if (!json || cap < 0 || (cap > 0 && !sdfs)) return set_global(PT_ERR_INVALID_ARG, "pt_pack_json: bad argument");
    try {
        pt::SceneModel m;
        std::string err;
        int st = pt::load_scene_json(json, len, m, err);
        if (st != PT_OK) return set_global(st, "scene JSON: " + err);
        if (ubo) {
            st = pt::pack_reference_ubo(m, *ubo, err);
            if (st != PT_OK) return set_global(st, err);
        }
        if (pc) {
            st = pt::camera_to_push(m, shot, *pc, err);
            if (st != PT_OK) return set_global(st, err);
        }
        if (n_sdfs) *n_sdfs = (int)m.sdfs.size();
        for (int i = 0; i < cap && i < (int)m.sdfs.size(); ++i) sdfs[i] = m.sdfs[(size_t)i].desc;
        return PT_OK;
    } catch (const std::exception& e) {
        return set_global(PT_ERR_JSON, std::string("pt_pack_json: ") + e.what());
    }
}

static pt_status write_text(const std::string& text, char* out, size_t cap, size_t* len) {
    if (len) *len = text.size();
    if (!out) return PT_OK;
    if (cap < text.size() + 1) return set_global(PT_ERR_CAPACITY, "scene JSON needs " + std::to_string(text.size() + 1) + " bytes");
    std::memcpy(out, text.data(), text.size());
    out[text.size()] = '\0';
    return PT_OK;
}

pt_status pt_scene_write_json(const pt_scene* scene, char* out, size_t cap, size_t* len) {
    if (!scene) return set_global(PT_ERR_INVALID_ARG, "pt_scene_write_json: NULL scene");
    if (!scene->has_model)
        return set_global(PT_ERR_STATE, "pt_scene_write_json: the scene came from a packed buffer and has no JSON model");
    try {
        return write_text(pt::write_scene_json(scene->model), out, cap, len);
    } catch (const std::exception& e) {
        return set_global(PT_ERR_JSON, std::string("pt_scene_write_json: ") + e.what());
    }
}

pt_status pt_json_rewrite(const char* json, size_t json_len, char* out, size_t cap, size_t* len) {
    if (!json) return set_global(PT_ERR_INVALID_ARG, "pt_json_rewrite: NULL json");
    try {
        pt::SceneModel m;
        std::string err;
        int st = pt::load_scene_json(json, json_len, m, err);
        if (st != PT_OK) return set_global(st, "scene JSON: " + err);
        return write_text(pt::write_scene_json(m), out, cap, len);
    } catch (const std::exception& e) {
        return set_global(PT_ERR_JSON, std::string("pt_json_rewrite: ") + e.what());
    }
}

pt_status pt_accum_create(pt_ctx* ctx, int width, int height, pt_accum** out) {
    if (!ctx || !out) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create: NULL argument");
    *out = nullptr;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)1 << 28)
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create: resolution out of range");
    (void)hipSetDevice(ctx->device);
    pt_accum* a = new (std::nothrow) pt_accum();
    if (!a) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create: out of host memory");
    a->ctx = ctx;
    a->w = width;
    a->h = height;
    DevAlloc m;
    int st = ptk::dev_alloc(ctx, m, (size_t)width * height * 16);
    if (st != PT_OK) {
        delete a;
        return st;
    }
    a->d_texels = static_cast<float*>(m.p);
    hipError_t e = hipMemsetAsync(a->d_texels, 0, (size_t)width * height * 16, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        (void)hipFree(a->d_texels);
        ctx->allocated -= (int64_t)width * height * 16;
        delete a;
        return ptk::hip_fail(ctx, e, "pt_accum_create");
    }
    *out = a;
    return PT_OK;
}

pt_status pt_accum_reset(pt_accum* a) {
    if (!a) return set_global(PT_ERR_INVALID_ARG, "pt_accum_reset: NULL argument");
    (void)hipSetDevice(a->ctx->device);
    hipError_t e = hipMemsetAsync(a->d_texels, 0, (size_t)a->w * a->h * 16, a->ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(a->ctx, e, "pt_accum_reset");
    return PT_OK;
}

pt_status pt_accum_read_xyz(pt_accum* a, float* host, size_t n) {
    if (!a || !host) return set_global(PT_ERR_INVALID_ARG, "pt_accum_read_xyz: NULL argument");
    if (n != (size_t)a->w * a->h * 4) return set_err(a->ctx, PT_ERR_INVALID_ARG, "pt_accum_read_xyz: size mismatch");
    (void)hipSetDevice(a->ctx->device);
    hipError_t e = hipMemcpyAsync(host, a->d_texels, n * 4, hipMemcpyDeviceToHost, a->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(a->ctx, e, "pt_accum_read_xyz");
    return PT_OK;
}

pt_status pt_accum_write_xyz(pt_accum* a, const float* host, size_t n) {
    if (!a || !host) return set_global(PT_ERR_INVALID_ARG, "pt_accum_write_xyz: NULL argument");
    if (n != (size_t)a->w * a->h * 4) return set_err(a->ctx, PT_ERR_INVALID_ARG, "pt_accum_write_xyz: size mismatch");
    (void)hipSetDevice(a->ctx->device);
    hipError_t e = hipMemcpyAsync(a->d_texels, host, n * 4, hipMemcpyHostToDevice, a->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(a->ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(a->ctx, e, "pt_accum_write_xyz");
    return PT_OK;
}

namespace {
// Unmaps and frees an exportable accumulator's memory (each step only once the
// previous ones have been made).
void vmm_free(pt_accum* a, bool mapped, bool reserved, bool created) {
    if (mapped) (void)hipMemUnmap(a->d_texels, a->vmm_bytes);
    if (reserved) (void)hipMemAddressFree(a->d_texels, a->vmm_bytes);
    if (created) (void)hipMemRelease(a->vmm_handle);
}
}  // namespace

void pt_accum_release(pt_accum* a) {
    if (!a) return;
    (void)hipSetDevice(a->ctx->device);
    (void)hipStreamSynchronize(a->ctx->stream);
    if (a->vmm) {
        vmm_free(a, true, true, true);
        a->ctx->allocated -= (int64_t)a->vmm_bytes;
    } else {
        DevAlloc m;
        m.p = a->d_texels;
        m.bytes = (size_t)a->w * a->h * 16;
        ptk::dev_free(a->ctx, m);
    }
    delete a;
}

pt_status pt_accum_create_exportable(pt_ctx* ctx, int width, int height, pt_accum** out) {
    if (!ctx || !out) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create_exportable: NULL argument");
    *out = nullptr;
    if (width <= 0 || height <= 0 || (int64_t)width * height > (int64_t)1 << 28)
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create_exportable: resolution out of range");
    (void)hipSetDevice(ctx->device);
    hipMemAllocationProp prop{};
    prop.type = hipMemAllocationTypePinned;
    prop.requestedHandleType = hipMemHandleTypePosixFileDescriptor;
    prop.location.type = hipMemLocationTypeDevice;
    prop.location.id = ctx->device;
    size_t gran = 0;
    hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityMinimum);
    if (e != hipSuccess) return ptk::hip_fail(ctx, e, "pt_accum_create_exportable: hipMemGetAllocationGranularity");
    if (gran == 0) gran = 1;
    const size_t bytes = (size_t)width * height * 16;
    const size_t sz = (bytes + gran - 1) / gran * gran;
    if (ctx->cfg.hbm_budget > 0 && ctx->allocated + (int64_t)sz > ctx->cfg.hbm_budget)
        return set_err(ctx, PT_ERR_BUDGET, "HBM budget exceeded: " + std::to_string(ctx->allocated + (int64_t)sz) +
                                               " > " + std::to_string(ctx->cfg.hbm_budget) + " bytes");
    pt_accum* a = new (std::nothrow) pt_accum();
    if (!a) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_accum_create_exportable: out of host memory");
    a->ctx = ctx;
    a->w = width;
    a->h = height;
    a->vmm = true;
    a->vmm_bytes = sz;
    e = hipMemCreate(&a->vmm_handle, sz, &prop, 0);
    if (e != hipSuccess) {
        delete a;
        return ptk::hip_fail(ctx, e, "pt_accum_create_exportable: hipMemCreate");
    }
    void* ptr = nullptr;
    e = hipMemAddressReserve(&ptr, sz, gran, nullptr, 0);
    if (e != hipSuccess) {
        vmm_free(a, false, false, true);
        delete a;
        return ptk::hip_fail(ctx, e, "pt_accum_create_exportable: hipMemAddressReserve");
    }
    a->d_texels = static_cast<float*>(ptr);
    e = hipMemMap(ptr, sz, 0, a->vmm_handle, 0);
    if (e != hipSuccess) {
        vmm_free(a, false, true, true);
        delete a;
        return ptk::hip_fail(ctx, e, "pt_accum_create_exportable: hipMemMap");
    }
    hipMemAccessDesc access{};
    access.location = prop.location;
    access.flags = hipMemAccessFlagsProtReadWrite;
    e = hipMemSetAccess(ptr, sz, &access, 1);
    if (e == hipSuccess) e = hipMemsetAsync(ptr, 0, sz, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
        vmm_free(a, true, true, true);
        delete a;
        return ptk::hip_fail(ctx, e, "pt_accum_create_exportable");
    }
    ctx->allocated += (int64_t)sz;
    *out = a;
    return PT_OK;
}

pt_status pt_accum_export_fd(pt_accum* a, int* fd, uint64_t* bytes) {
    if (!a || !fd || !bytes) return set_global(PT_ERR_INVALID_ARG, "pt_accum_export_fd: NULL argument");
    *fd = -1;
    *bytes = 0;
    if (!a->vmm)
        return set_err(a->ctx, PT_ERR_STATE, "pt_accum_export_fd: the accumulator was not made by pt_accum_create_exportable");
    (void)hipSetDevice(a->ctx->device);
    int f = -1;
    hipError_t e = hipMemExportToShareableHandle(&f, a->vmm_handle, hipMemHandleTypePosixFileDescriptor, 0);
    if (e != hipSuccess) return ptk::hip_fail(a->ctx, e, "pt_accum_export_fd: hipMemExportToShareableHandle");
    *fd = f;
    *bytes = a->vmm_bytes;
    return PT_OK;
}

pt_status pt_dispatch(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc, pt_mode mode) {
    if (!ctx || !s || !acc || !pc) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch: NULL argument");
    if (s->ctx != ctx || acc->ctx != ctx) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch: objects of another context");
    if (pc->resolution[0] != acc->w || pc->resolution[1] != acc->h)
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch: push-constant resolution differs from the accumulator");
    if (pc->samplesPerFrame <= 0 || pc->pathLength < 0 || pc->pathLength > 1024 || pc->currentSamples < pc->samplesPerFrame)
        return set_err(ctx, PT_ERR_INVALID_ARG,
                       "pt_dispatch: need samplesPerFrame > 0, 0 <= pathLength <= 1024, currentSamples >= samplesPerFrame");
    (void)hipSetDevice(ctx->device);
    try {
        int st;
        if (mode == PT_MODE_REF) {
            if (!s->ref_ok) return set_err(ctx, PT_ERR_CAPACITY, "reference mode cannot hold this scene: " + s->ref_err);
            st = ptk::ref_dispatch(ctx, s, acc, pc);
        } else if (mode == PT_MODE_EXT) {
            if (!s->has_model) return set_err(ctx, PT_ERR_STATE, "extended mode needs a scene loaded from JSON");
            st = ptk::ext_dispatch(ctx, s, acc, pc);
        } else {
            return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch: unknown mode");
        }
        if (st == PT_OK) st = ptk::count_dispatch(ctx);
        return st;
    } catch (const std::exception& e) {
        return set_err(ctx, PT_ERR_HIP, std::string("pt_dispatch: ") + e.what());
    }
}

pt_status pt_dispatch_noalloc(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc) {
    if (!ctx || !s || !acc || !pc) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch_noalloc: NULL argument");
    if (s->ctx != ctx || acc->ctx != ctx)
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch_noalloc: objects of another context");
    if (pc->resolution[0] != acc->w || pc->resolution[1] != acc->h)
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_dispatch_noalloc: push-constant resolution differs from the accumulator");
    if (pc->samplesPerFrame <= 0 || pc->pathLength < 0 || pc->pathLength > 1024 || pc->currentSamples < pc->samplesPerFrame)
        return set_err(ctx, PT_ERR_INVALID_ARG,
                       "pt_dispatch_noalloc: need samplesPerFrame > 0, 0 <= pathLength <= 1024, "
                       "currentSamples >= samplesPerFrame");
    if (!s->ref_ok) return set_err(ctx, PT_ERR_CAPACITY, "reference mode cannot hold this scene: " + s->ref_err);
    if (ctx->static_wave_status != PT_OK)
        return set_err(ctx, ctx->static_wave_status, "pt_dispatch_noalloc: " + ctx->static_wave_err);
    // kernel timing records events, which this entry point does not create
    if (ctx->prof) return set_err(ctx, PT_ERR_STATE, "pt_dispatch_noalloc: kernel profiling is enabled");
    (void)hipSetDevice(ctx->device);
    int st = ptk::ref_dispatch_static(ctx, s, acc, pc);
    if (st == PT_OK) st = ptk::count_dispatch(ctx);
    return st;
}

pt_status pt_render_features(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, float* host, size_t n_floats) {
    if (!ctx || !s || !pc || !host) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_render_features: NULL argument");
    if (s->ctx != ctx) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_render_features: scene of another context");
    if (!s->has_model) return set_err(ctx, PT_ERR_STATE, "pt_render_features: needs a scene loaded from JSON");
    const int w = pc->resolution[0], h = pc->resolution[1];
    if (w <= 0 || h <= 0 || (int64_t)w * h > (int64_t)1 << 28 || pc->samplesPerFrame <= 0 || pc->pathLength < 0 ||
        pc->pathLength > 1024 || pc->currentSamples < pc->samplesPerFrame)
        return set_err(ctx, PT_ERR_INVALID_ARG,
                       "pt_render_features: need a resolution of 1 to 2^28 pixels, samplesPerFrame > 0, "
                       "0 <= pathLength <= 1024, currentSamples >= samplesPerFrame");
    const size_t P = (size_t)w * h;
    if (n_floats != P * 12) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_render_features: size mismatch");
    (void)hipSetDevice(ctx->device);
    DevAlloc m;
    const size_t feat_bytes = (P * 48 + 255) / 256 * 256;
    int st = ptk::dev_alloc(ctx, m, feat_bytes + P * 4);
    if (st != PT_OK) return st;
    float* d_feat = (float*)m.p;
    int* d_n = (int*)((char*)m.p + feat_bytes);
    std::vector<int> n(P);
    try {
        hipError_t e = hipMemsetAsync(m.p, 0, feat_bytes + P * 4, ctx->stream);
        if (e != hipSuccess) st = ptk::hip_fail(ctx, e, "pt_render_features");
        if (st == PT_OK) st = ptk::ext_dispatch_features(ctx, s, pc, d_feat, d_n);
        if (st == PT_OK) st = ptk::count_dispatch(ctx);
        if (st == PT_OK) {
            e = hipMemcpyAsync(host, d_feat, P * 48, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipMemcpyAsync(n.data(), d_n, P * 4, hipMemcpyDeviceToHost, ctx->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
            if (e != hipSuccess) st = ptk::hip_fail(ctx, e, "pt_render_features");
        }
    } catch (const std::exception& ex) {
        st = set_err(ctx, PT_ERR_HIP, std::string("pt_render_features: ") + ex.what());
    }
    (void)hipStreamSynchronize(ctx->stream);
    ptk::dev_free(ctx, m);
    if (st != PT_OK) return st;
    // XYZ in the accumulator's units: exposure apertureSize^2 * ISO per sample,
    // as the texel resolve applies it
    const float ex = pc->apertureSize * pc->apertureSize * (float)pc->ISO;
    for (size_t i = 0; i < P; ++i) {
        float* f = host + i * 12;
        const double sy = f[1], syy = f[10];
        const float inv = n[i] > 0 ? 1.0f / (float)n[i] : 0.0f;
        const float e = n[i] > 0 ? ex / (float)n[i] : 0.0f;
        for (int c = 0; c < 3; ++c) f[c] *= e;
        for (int c = 3; c < 10; ++c) f[c] *= inv;
        // standard error of the mean of Y: sqrt(sample variance / n)
        double var = 0.0;
        if (n[i] > 1) var = std::max(syy - sy * sy / n[i], 0.0) / ((double)n[i] - 1.0) / (double)n[i];
        f[10] = (float)(std::sqrt(var) * (double)ex);
        f[11] = (float)n[i];
    }
    return PT_OK;
}

pt_status pt_sync(pt_ctx* ctx) {
    if (!ctx) return set_global(PT_ERR_INVALID_ARG, "pt_sync: NULL context");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(ctx, e, "pt_sync");
    return PT_OK;
}

pt_status pt_counters_get(pt_ctx* ctx, pt_counters* out) {
    if (!ctx || !out) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_counters_get: NULL argument");
    (void)hipSetDevice(ctx->device);
    std::vector<unsigned long long> lines((size_t)ptd::kStatLines * ptd::kStatStride);
    hipError_t e = hipMemcpyAsync(lines.data(), ctx->d_stats, lines.size() * sizeof(unsigned long long),
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(ctx, e, "pt_counters_get");
    unsigned long long h[ptd::kStatSlots] = {};
    for (int l = 0; l < ptd::kStatLines; ++l)
        for (int k = 0; k < ptd::kStatSlots; ++k) h[k] += lines[(size_t)l * ptd::kStatStride + (size_t)k];
    out->segments = h[0];
    out->shadow_segments = h[1];
    out->sdf_steps = h[2];
    out->paths = h[3];
    out->nan_samples = h[4];
    out->dispatches = h[5];
    out->escaped = h[6];
    out->emitter_hits = h[7];
    out->lambert_hits = h[8];
    out->dielectric_hits = h[9];
    out->conductor_hits = h[10];
    out->bvh_nodes = h[11];
    out->shadow_bvh_nodes = h[12];
    out->sdf_exhausted = h[13];
    out->march_flops_mandelbulb = h[ptd::kStatMarchFlops + 0];
    out->march_flops_menger = h[ptd::kStatMarchFlops + 1];
    out->march_flops_blob = h[ptd::kStatMarchFlops + 2];
    out->march_flops_terrain = h[ptd::kStatMarchFlops + 3];
    out->shadow_march_flops = h[ptd::kStatShadowMarchFlops];
    out->bad_refs = h[ptd::kStatBadRefs];
    return PT_OK;
}

pt_status pt_counters_reset(pt_ctx* ctx) {
    if (!ctx) return set_global(PT_ERR_INVALID_ARG, "pt_counters_reset: NULL context");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipMemsetAsync(ctx->d_stats, 0, (size_t)ptd::kStatLines * ptd::kStatStride * sizeof(unsigned long long),
                                  ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(ctx, e, "pt_counters_reset");
    return PT_OK;
}

pt_status pt_profile_enable(pt_ctx* ctx, int enable) {
    if (!ctx) return set_global(PT_ERR_INVALID_ARG, "pt_profile_enable: NULL context");
    ctx->prof = enable != 0;
    return PT_OK;
}

pt_status pt_profile_read(pt_ctx* ctx, pt_kernel_stat* out, int cap, int* n_out, int reset) {
    if (!ctx || cap < 0 || (cap > 0 && !out)) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_profile_read: bad argument");
    (void)hipSetDevice(ctx->device);
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return ptk::hip_fail(ctx, e, "pt_profile_read");
    for (const ProfPending& p : ctx->prof_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
            ctx->prof_totals[(size_t)p.kernel].ms += ms;
            ctx->prof_totals[(size_t)p.kernel].launches += 1;
        }
        ctx->prof_pool.push_back(p.start);
        ctx->prof_pool.push_back(p.stop);
    }
    ctx->prof_pending.clear();
    int n = 0;
    for (const ProfTotal& t : ctx->prof_totals) {
        if (t.launches == 0) continue;
        if (n < cap) {
            std::memset(out[n].name, 0, sizeof out[n].name);
            std::strncpy(out[n].name, t.name.c_str(), sizeof out[n].name - 1);
            out[n].launches = t.launches;
            out[n].total_ms = t.ms;
        }
        ++n;
    }
    if (n_out) *n_out = n;
    if (reset)
        for (ProfTotal& t : ctx->prof_totals) { t.launches = 0; t.ms = 0.0; }
    return PT_OK;
}

pt_status pt_resolve_srgb8(pt_ctx* ctx, pt_accum* acc, int tonemap, uint8_t* host_rgb, size_t n_bytes) {
    if (!ctx || !acc || !host_rgb) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_resolve_srgb8: NULL argument");
    if (n_bytes != (size_t)acc->w * acc->h * 3) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_resolve_srgb8: size mismatch");
    if (tonemap < 0 || tonemap > 3) return set_err(ctx, PT_ERR_INVALID_ARG, "pt_resolve_srgb8: tonemap must be 0..3");
    (void)hipSetDevice(ctx->device);
    return ptk::resolve_srgb8(ctx, acc, tonemap, host_rgb);
}

pt_status pt_math_eval_host(int fn, const float* x, const float* y, float* out, size_t n) {
    if (!x || !out || fn < 0 || fn > 7 || (fn == 4 && !y)) return set_global(PT_ERR_INVALID_ARG, "pt_math_eval_host: bad argument");
    for (size_t i = 0; i < n; ++i) {
        float a = x[i];
        switch (fn) {
            case 0: out[i] = ptm_sin(a); break;
            case 1: out[i] = ptm_cos(a); break;
            case 2: out[i] = ptm_exp(a); break;
            case 3: out[i] = ptm_log(a); break;
            case 4: out[i] = ptm_pow(a, y[i]); break;
            case 5: out[i] = ptm_acos(a); break;
            case 6: out[i] = ptm_sqrt(a); break;
            default: out[i] = ptm_rsqrt(a); break;
        }
    }
    return PT_OK;
}

pt_status pt_math_eval_device(pt_ctx* ctx, int fn, const float* x, const float* y, float* out, size_t n) {
    if (!ctx || !x || !out || fn < 0 || fn > 7 || (fn == 4 && !y))
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_math_eval_device: bad argument");
    (void)hipSetDevice(ctx->device);
    return ptk::math_eval_device(ctx, fn, x, y, out, n);
}

pt_status pt_peak_probe(pt_ctx* ctx, int what, size_t n, int reps, double* units_out) {
    if (!ctx || !units_out || what < 0 || what > 1 || reps < 1 || reps > 1000 || n == 0 ||
        (what == 0 && (n % 4 != 0 || n > ((size_t)1 << 30))) || (what == 1 && n > ((size_t)1 << 20)))
        return set_err(ctx, PT_ERR_INVALID_ARG, "pt_peak_probe: bad argument");
    (void)hipSetDevice(ctx->device);
    return ptk::peak_probe(ctx, what, n, reps, units_out);
}

}  // extern "C"
