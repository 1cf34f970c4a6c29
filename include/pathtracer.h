/*
 * pathtracer.h — C ABI of the MI355X (gfx950) spectral path tracer.
 *
 * The reference renders through one Vulkan compute dispatch of
 * src/shader.comp. Its "operator API" is:
 *   - a 16,388-byte uniform buffer `UniformBufferObject`
 *     (src/pathtracer.cpp:187-195, packed by UpdateUniformBuffer :3642-3811),
 *   - 88 bytes of push constants `PushConstantValues`
 *     (src/pathtracer.cpp:197-218, filled by UpdatePushConstant :3813-3834),
 *   - an RGBA32F CIE-XYZ texel buffer of W*H*16 bytes that the kernel
 *     reads, blends and writes (src/pathtracer.cpp:2250-2255,
 *     src/shader.comp:1525-1533),
 *   - SDF programs given as GLSL text in the scene JSON and spliced into the
 *     shader (src/pathtracer.cpp:2004-2054),
 *   - vkCmdDispatch(ceil(W/16), ceil(H/16), 1) per frame
 *     (src/pathtracer.cpp:3603), after the caller advanced `frame` and
 *     `currentSamples` (src/pathtracer.cpp:4042-4048).
 * The entry points below replace exactly that contract. Plain C types only;
 * no exception crosses this boundary. Every call returns a pt_status; the
 * message for the last failure on a context is pt_last_error(ctx).
 *
 * Threading: a context is externally synchronised (one host thread at a
 * time), as the reference's single render thread is. Work is asynchronous on
 * the context's HIP stream until pt_sync() or a read-back call.
 */
#ifndef PATHTRACER_H
#define PATHTRACER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int pt_status;
#define PT_OK 0
#define PT_ERR_INVALID_ARG (-1)
#define PT_ERR_JSON (-2)
#define PT_ERR_CAPACITY (-3)
#define PT_ERR_UNSUPPORTED_SDF (-4)
#define PT_ERR_HIP (-5)
#define PT_ERR_NO_DEVICE (-6)
#define PT_ERR_BUDGET (-7)
#define PT_ERR_STATE (-8)

/* Capacities of the packed scene; same values as src/shader.comp:11-15 and
 * src/pathtracer.cpp:39-43. */
#define PT_MAX_OBJECTS_SIZE 1024
#define PT_MAX_SDFS_SIZE 768
#define PT_MAX_MATERIALS_SIZE 783
#define PT_MAX_LIGHTS_SIZE 128
#define PT_MAX_LIGHTIDS_SIZE 64
#define PT_CIE_SIZE 1323
/* SDF slots addressable by the reference's 32-bit set mask (src/shader.comp:753-770). */
#define PT_MAX_SDF_OBJECTS 32

/* Binary-identical to UniformBufferObject (src/pathtracer.cpp:187-195). */
typedef struct pt_packed_scene {
    float numObjects[7]; /* spheres, planes, boxes, lenses, cyclides, sdfs, lights */
    float objects[PT_MAX_OBJECTS_SIZE];
    float sdfs[PT_MAX_SDFS_SIZE];
    float materials[PT_MAX_MATERIALS_SIZE];
    float lights[PT_MAX_LIGHTS_SIZE];
    float lightIDs[PT_MAX_LIGHTIDS_SIZE];
    float cie[PT_CIE_SIZE];
} pt_packed_scene;

/* Binary-identical to PushConstantValues (src/pathtracer.cpp:197-218). */
typedef struct pt_push_constants {
    int32_t resolution[2];
    int32_t frame;           /* samples dispatched so far, including this dispatch */
    int32_t currentSamples;  /* samples in the running mean after this dispatch */
    int32_t samplesPerFrame;
    float FPS;
    float persistence;
    int32_t pathLength;      /* maximum segments per path */
    float cameraAngle[2];    /* (-pitch, yaw) degrees, src/pathtracer.cpp:3821 */
    float cameraPosX, cameraPosY, cameraPosZ;
    int32_t ISO;
    float cameraSize;
    float apertureSize;
    float apertureDist;
    float lensRadius;
    float lensFocalLength;
    float lensThickness;
    float lensDistance;
    int32_t tonemap;
} pt_push_constants;

/* Built-in SDF programs. The reference splices arbitrary GLSL; a HIP build
 * cannot compile GLSL, so pt_sdf_recognize maps the shipped programs
 * (sdfs/{blob,mandelbulb,menger,terrain}.glsl and the `glsl` strings of scenes/scene{3..10}.json) to these kinds. */
#define PT_SDF_NONE 0
#define PT_SDF_MANDELBULB 1 /* power-8 Chebyshev form, sdfs/mandelbulb.glsl */
#define PT_SDF_MENGER 2     /* sdfs/menger.glsl */
#define PT_SDF_BLOB 3       /* four smooth-min spheres, sdfs/blob.glsl */
#define PT_SDF_TERRAIN 4    /* sum of sines height field, sdfs/terrain.glsl */

#define PT_SDFMAT_CONSTANT 0 /* material = m[0] */
#define PT_SDFMAT_RADIAL 1   /* mix(m[0], m[1], |p|^2 / (m[2] + |p|^2)) */
#define PT_SDFMAT_BLOB 2     /* smooth-min of the four blob spheres carrying m[0..3] */

typedef struct pt_sdf_desc {
    int32_t kind;          /* PT_SDF_* */
    int32_t iterations;    /* mandelbulb iterations / menger levels */
    float bailout;         /* mandelbulb escape radius */
    int32_t material_kind; /* PT_SDFMAT_* */
    float material[4];
    float pos[3];          /* bounding-box centre; the SDF is evaluated at p - pos */
    float size[3];         /* bounding-box size */
} pt_sdf_desc;

/* Tracing semantics.
 * PT_MODE_REF: the reference's method bit for bit (4 rotated wavelengths in
 *   [390, 720) nm, sequential PCG chain keyed on (frame, pixel), brute-force
 *   closest hit, its MIS/RR rules and quirks).
 * PT_MODE_EXT: the extended physically based mode (quads, multi-Gaussian
 *   reflectance, dispersive dielectrics, conductors, environment light, BVH,
 *   hero wavelengths over [380, 780) nm, counter-based PCG keyed on
 *   (seed, pixel, sample, bounce, dimension)); it also takes the reference's
 *   lenses and cyclides, shaded with their materials (emitting ones are found
 *   by BSDF sampling only). */
typedef enum pt_mode { PT_MODE_REF = 0, PT_MODE_EXT = 1 } pt_mode;

/* Launch/tuning knobs (0 = library default). */
typedef struct pt_config {
    int32_t wave_paths;    /* paths in flight per wave (queue capacity); 0 = 2^27 for pt_dispatch (allocated as a
                              dispatch needs, about 330 B per path) and 2^22 for a server's batched waves */
    int32_t block_size;    /* threads per block of the queue kernels: 0 or 256 (they are compiled for 256
                              lanes, 4 wavefronts; other values are rejected) */
    int32_t bvh_width;     /* 2, 4 or 8 (0 = 2) */
    int32_t sdf_max_steps; /* sphere-tracing budget of the extended mode */
    int64_t hbm_budget;    /* bytes this context may allocate (0 = no limit) */
    int32_t traversal;     /* extended mode, single-scene dispatches: 0 = the default (1), 1 = persistent lanes
                              refilled from the queue, 2 = one ray per lane. Batched waves and scenes with a
                              parallelogram BVH take one ray per lane. */
    int32_t bvh_leaf;      /* most primitives per BVH leaf, 1..16 (0 = 1) */
    int32_t graph_mode;    /* batched serving waves: 0 = captured once per launch shape into a hipGraph and
                              replayed, 1 = kernels launched one by one */
    int32_t blocks_per_cu; /* extended mode: grid of the queue kernels, blocks per compute unit, 1..64 (0 = 8) */
    int32_t march;         /* extended mode, SDF sphere tracing: 0 = persistent lanes refilled from the queue,
                              1 = one ray per lane */
    int32_t march_blocks_per_cu; /* extended mode: grid of the persistent march kernels, blocks per compute unit,
                                    1..64 (0 = the default) */
} pt_config;

typedef struct pt_counters {
    uint64_t segments;        /* extension rays traced (one per path segment) */
    uint64_t shadow_segments; /* NEE visibility rays traced */
    uint64_t sdf_steps;       /* sphere-tracing SDF evaluations */
    uint64_t paths;           /* camera paths started */
    uint64_t nan_samples;     /* samples discarded as NaN (src/shader.comp:1479-1487) */
    uint64_t dispatches;
    /* extended mode: path vertices by class */
    uint64_t escaped;         /* rays leaving the scene */
    uint64_t emitter_hits;
    uint64_t lambert_hits;
    uint64_t dielectric_hits;
    uint64_t conductor_hits;
    uint64_t bvh_nodes;       /* BVH nodes visited by extension rays (extended mode) */
    uint64_t shadow_bvh_nodes;/* BVH nodes visited by shadow rays (extended mode) */
    uint64_t sdf_exhausted;   /* SDF marches that ran out of sdf_max_steps (taken as misses; extended mode) */
    /* extended mode: FP32 operations of sphere tracing (distance-estimator
     * evaluations plus the step update), extension rays by SDF kind, then shadow rays */
    uint64_t march_flops_mandelbulb;
    uint64_t march_flops_menger;
    uint64_t march_flops_blob;
    uint64_t march_flops_terrain;
    uint64_t shadow_march_flops;
    uint64_t bad_refs;        /* extended mode: hit records that name no primitive of their scene; classified as
                                 misses and counted here (0 in a correct run) */
} pt_counters;

typedef struct pt_ctx pt_ctx;
typedef struct pt_scene pt_scene;
typedef struct pt_accum pt_accum;

/* Context: one HIP device and one stream (the reference's device + compute queue,
 * src/pathtracer.cpp:1534-1536). */
pt_status pt_create(pt_ctx** out, int device_ordinal, const pt_config* cfg);
void pt_destroy(pt_ctx* ctx);
const char* pt_last_error(const pt_ctx* ctx);
/* Message of the last failure of a call that had no context (create, recognise). */
const char* pt_last_error_global(void);

/* Scene from the reference's JSON format (ReadJSON + UpdateFromJSON +
 * UpdateUniformBuffer + InsertSDF, src/pathtracer.cpp:893-908, 2576-2722,
 * 3642-3811, 2004-2054). Unknown SDF text fails with PT_ERR_UNSUPPORTED_SDF.
 * The extended keys (quad, material types, environment, render) are accepted
 * and used by PT_MODE_EXT. */
pt_status pt_scene_from_json(pt_ctx* ctx, const char* json, size_t len, pt_scene** out);
/* Scene from an already packed uniform buffer plus SDF descriptors
 * (UpdateUniformBuffer's output, src/pathtracer.cpp:3642-3811). */
pt_status pt_scene_from_packed(pt_ctx* ctx, const pt_packed_scene* ubo, const pt_sdf_desc* sdfs,
                               int n_sdfs, pt_scene** out);
/* Copies the packed uniform buffer of a scene (the bytes the reference uploads). */
pt_status pt_scene_get_packed(const pt_scene* scene, pt_packed_scene* out);
/* Copies up to `cap` SDF descriptors; *n_out receives the scene's count. */
pt_status pt_scene_get_sdfs(const pt_scene* scene, pt_sdf_desc* out, int cap, int* n_out);
/* Fills the camera fields of `pc` from camera shot `shot` (1-based) as
 * UpdateFromJSON + UpdatePushConstant do (src/pathtracer.cpp:2577-2606, 3813-3834). */
pt_status pt_scene_camera(const pt_scene* scene, int shot, pt_push_constants* pc);
/* Device bytes the scene holds (for admission by HBM footprint). */
pt_status pt_scene_device_bytes(const pt_scene* scene, int64_t* bytes);
void pt_scene_release(pt_scene* scene);

/* GLSL SDF text -> built-in kind and parameters. */
pt_status pt_sdf_recognize(const char* glsl, size_t len, pt_sdf_desc* out);

/* Host-only packing, no device needed: the scene JSON -> the uniform buffer
 * the reference uploads (UpdateFromJSON + UpdateUniformBuffer,
 * src/pathtracer.cpp:2576-2722, 3642-3811), the camera fields of the push
 * constants for camera shot `shot` (1-based, UpdatePushConstant :3813-3834),
 * and up to `cap` SDF descriptors (*n_sdfs receives the count). Any output
 * pointer may be NULL. */
pt_status pt_pack_json(const char* json, size_t len, int shot, pt_packed_scene* ubo, pt_push_constants* pc,
                       pt_sdf_desc* sdfs, int cap, int* n_sdfs);

/* Scene writer, UpdateToJSON + SaveScene (src/pathtracer.cpp:2724-2858,
 * 3465-3470): the reference's keys in its order with every float rounded to
 * the 1e-5 grid as the reference saves it (RoundDecimal, :935-943), 4-space
 * indentation; the extended keys follow with their floats exact. Writes at most
 * `cap` bytes including a terminating NUL into `out` (which may be NULL to ask
 * for the size); *len receives the text length without the NUL. Returns
 * PT_ERR_CAPACITY when `cap` is too small. */
pt_status pt_scene_write_json(const pt_scene* scene, char* out, size_t cap, size_t* len);
/* Host-only: parses `json` and writes it back as pt_scene_write_json does. */
pt_status pt_json_rewrite(const char* json, size_t json_len, char* out, size_t cap, size_t* len);

/* Accumulator: device RGBA32F, W*H texels, zeroed (the reference's texel buffer). */
pt_status pt_accum_create(pt_ctx* ctx, int width, int height, pt_accum** out);
pt_status pt_accum_reset(pt_accum* acc);
/* Host copies of the texels (4 floats per pixel, row gid.y = 0 first). */
pt_status pt_accum_read_xyz(pt_accum* acc, float* host_rgba, size_t n_floats);
pt_status pt_accum_write_xyz(pt_accum* acc, const float* host_rgba, size_t n_floats);
void pt_accum_release(pt_accum* acc);

/* Zero-copy display (replaces the texel buffer the reference's fragment pass
 * samples in place, src/shader.frag:81-99, bound at src/pathtracer.cpp:2294-2366).
 * pt_accum_create_exportable makes an accumulator like pt_accum_create whose
 * device memory can be shared with another API: pt_accum_export_fd returns a new
 * dma-buf file descriptor of it and the exported size in bytes (W*H*16 rounded
 * up to the allocation granularity; the texels start at offset 0). The caller
 * owns the descriptor: Vulkan imports it with VkImportMemoryFdInfoKHR
 * (VK_EXTERNAL_MEMORY_HANDLE_TYPE_DMA_BUF_BIT_EXT), or close() it. Dispatches
 * are asynchronous: pt_sync (or a semaphore of the caller's) before the other
 * API reads. Every other pt_accum_* call works on either kind; an accumulator
 * from pt_accum_create returns PT_ERR_STATE here. */
pt_status pt_accum_create_exportable(pt_ctx* ctx, int width, int height, pt_accum** out);
pt_status pt_accum_export_fd(pt_accum* acc, int* fd, uint64_t* bytes);

/* == vkCmdPushConstants + vkCmdDispatch + vkQueueSubmit: trace
 * pc->samplesPerFrame samples per pixel and fold them into `acc`
 * (src/shader.comp:1509-1533). Asynchronous on the context's stream. */
pt_status pt_dispatch(pt_ctx* ctx, const pt_scene* scene, pt_accum* acc, const pt_push_constants* pc,
                      pt_mode mode);
pt_status pt_sync(pt_ctx* ctx);
/* No counterpart in the reference: pt_dispatch(..., PT_MODE_REF) that allocates
 * no device memory, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). The path state
 * lives in a static device array of 65,536 paths (11,010,304 B), held by each
 * device the library runs on until the process ends; a dispatch of more paths
 * runs as several waves through it, so it has no size limit beyond
 * pt_dispatch's. Texels equal pt_dispatch's bit for bit. All contexts on one
 * device share that array: calls that overlap on one device race. Returns
 * PT_ERR_STATE while kernel profiling is enabled (it records events). */
pt_status pt_dispatch_noalloc(pt_ctx* ctx, const pt_scene* scene, pt_accum* acc, const pt_push_constants* pc);

/* Denoiser features (extended mode; no counterpart in the reference, which
 * shows the noisy running mean, src/shader.frag:81-99): traces the samples
 * [currentSamples - samplesPerFrame, currentSamples) of every pixel, as
 * pt_dispatch would, and writes 12 floats per pixel (PT_FEATURES): the means
 * over those samples of XYZ (the accumulator's texel units: exposure
 * apertureSize^2 * ISO applied), of the first hit's albedo (reflectance at
 * 450, 550 and 650 nm; 1 for glass and emitters; the normal-incidence
 * reflectance of a conductor), of its shading normal facing the camera and of
 * its distance (zeros for rays that miss); then the standard error of the
 * mean of Y (same units as XYZ) and the number of samples. Samples with a NaN
 * are dropped. Synchronous; `n_floats` must be width * height * 12. */
#define PT_FEATURES 12
pt_status pt_render_features(pt_ctx* ctx, const pt_scene* scene, const pt_push_constants* pc, float* host_feat,
                             size_t n_floats);

/* Counters summed over dispatches since creation or the last reset (synchronises). */
pt_status pt_counters_get(pt_ctx* ctx, pt_counters* out);
pt_status pt_counters_reset(pt_ctx* ctx);

/* Per-kernel timing with HIP events on the context's stream (the reference
 * has only ImGui frame times, src/pathtracer.cpp:2963-2964). While enabled,
 * every kernel launch is bracketed by two events; pt_profile_read
 * synchronises and returns, per kernel name, the launch count and the summed
 * GPU time in milliseconds, and with `reset` != 0 clears the totals. */
typedef struct pt_kernel_stat {
    char name[48];
    uint64_t launches;
    double total_ms;
} pt_kernel_stat;
pt_status pt_profile_enable(pt_ctx* ctx, int enable);
pt_status pt_profile_read(pt_ctx* ctx, pt_kernel_stat* out, int cap, int* n_out, int reset);

/* Display resolve of SaveRender / shader.frag (src/pathtracer.cpp:3491-3518,
 * src/shader.frag:81-99): XYZ -> Bradford E->D65 -> linear sRGB -> tonemap
 * (0 none, 1 Reinhard, 2 ACES, 3 DEUCES) -> sRGB OETF -> 8 bit. */
pt_status pt_resolve_srgb8(pt_ctx* ctx, pt_accum* acc, int tonemap, uint8_t* host_rgb, size_t n_bytes);

/* Diagnostics: evaluates the tracer's FP32 math functions (fn: 0 sin, 1 cos,
 * 2 exp, 3 log, 4 pow, 5 acos, 6 sqrt, 7 rsqrt) on the host build or on the
 * device build, so the two can be compared bit for bit. */
pt_status pt_math_eval_host(int fn, const float* x, const float* y, float* out, size_t n);
pt_status pt_math_eval_device(pt_ctx* ctx, int fn, const float* x, const float* y, float* out, size_t n);

/* Diagnostics with no counterpart in the reference: measured device peaks for
 * the rooflines. what = 0: a streaming copy of n floats (n a multiple of 4,
 * at most 2^30; 8 n bytes of HBM traffic per launch; kernel k_peak_copy).
 * what = 1: an FP32 fused multiply-add loop of n rounds (at most 2^20) over 8
 * independent chains per thread, in a grid of 16 blocks of 256 threads per CU
 * (kernel k_peak_fma). Launches one warm-up (k_peak_*_warmup) and then `reps`
 * timed launches on the context's stream, and returns when they are done; with
 * profiling on (pt_profile_enable) pt_profile_read gives their durations.
 * *units_out: the bytes (what 0) or flops (what 1) of one launch. */
pt_status pt_peak_probe(pt_ctx* ctx, int what, size_t n, int reps, double* units_out);

/* ---- Serving --------------------------------------------------------------
 * No counterpart in the reference, which renders one scene interactively until
 * the user stops (src/pathtracer.cpp:4005-4086). A server traces a stream of
 * render requests (extended mode) on one context: each request runs until
 * its relative RMSE estimate, sqrt(sum of per-pixel variance of the mean) /
 * sqrt(sum of mean^2) over pixels and XYZ channels, reaches its target or its
 * sample budget is spent. Scenes, BVHs and per-pixel moments of admitted
 * requests stay in device memory; admission stops at the server's HBM budget
 * and request count. One step traces one wave (continuous batching): every
 * active request that fits, in policy order, contributes a quantum of samples
 * per pixel, min_spp first, then its estimated remaining ones (from error ~
 * 1/sqrt(samples)); each request's samples are summed in sample order, so its
 * result is the same alone and in the mix. Times are seconds of the
 * server's monotonic clock (pt_server_now). A server is used by one host
 * thread at a time. */
typedef struct pt_server pt_server;
typedef struct pt_denoiser pt_denoiser; /* the denoiser section below */
#define PT_SERVE_SRPT 0 /* shortest estimated remaining GPU time first */
#define PT_SERVE_FIFO 1 /* arrival order, one request to completion at a time */
#define PT_SERVE_HRRN 2 /* highest response ratio first: (time since arrival + estimated remaining GPU
                           time) / estimated remaining GPU time, so a long request's priority grows
                           while it waits (shortest remaining first, with aging) */
#define PT_SERVE_IDLE 1 /* pt_server_step status: no request to trace */
#define PT_REQ_WAITING 0
#define PT_REQ_ACTIVE 1
#define PT_REQ_DONE 2
#define PT_REQ_FAILED 3

typedef struct pt_server_config {
    int64_t hbm_budget;    /* bytes of admitted scenes + moment buffers (0 = 64 GiB) */
    int32_t max_active;    /* requests admitted at once (0 = 16) */
    int32_t policy;        /* PT_SERVE_SRPT, PT_SERVE_FIFO or PT_SERVE_HRRN */
    int32_t min_spp;       /* samples per pixel before a request may be declared converged (0 = 8) */
    int32_t reserved;
} pt_server_config;

typedef struct pt_request_desc {
    const char* json;      /* extended-mode scene JSON; camera shot 1 is rendered */
    size_t len;
    int32_t width, height;
    int32_t spp_max;       /* sample budget per pixel */
    int32_t path_length;   /* maximum segments per path */
    float target_rel;      /* relative RMSE at which the request is done, e.g. 0.01 */
    int32_t denoise;       /* 1: may finish with the denoiser's image (pt_server_set_denoiser) */
} pt_request_desc;

typedef struct pt_request_status {
    int32_t id, state;     /* PT_REQ_* */
    int32_t spp;           /* samples per pixel traced */
    int32_t converged;     /* 1 when rel_error <= target_rel after at least min_spp samples */
    int32_t width, height;
    double rel_error;      /* last estimate for the image pt_server_read returns, -1 before two samples */
    double t_arrive, t_admit, t_done;
    double gpu_seconds;    /* wall time of its admission (scene parse, BVH build), of the steps that
                              traced it and of its denoiser runs */
    uint64_t segments;
    double rel_plain;      /* last estimate for the plain mean image */
    int32_t denoised;      /* 1 when pt_server_read returns the denoised image */
    int32_t reserved;
} pt_request_status;

pt_status pt_server_create(pt_ctx* ctx, const pt_server_config* cfg, pt_server** out);
void pt_server_destroy(pt_server* server);
const char* pt_server_last_error(const pt_server* server);
double pt_server_now(const pt_server* server);
/* Queues a request that arrived at t_arrive (server clock); *id receives its id. */
pt_status pt_server_submit(pt_server* server, const pt_request_desc* req, double t_arrive, int32_t* id);
/* Admits what fits, then traces one wave holding a quantum of samples of each
 * active request that fits (in policy order) and updates their estimates;
 * up to `cap` statuses of the traced requests go to stepped[], their number to
 * *n_stepped. Returns PT_OK, PT_SERVE_IDLE when nothing is admitted, or an error. */
pt_status pt_server_step(pt_server* server, pt_request_status* stepped, int32_t cap, int32_t* n_stepped);
pt_status pt_server_status(pt_server* server, int32_t id, pt_request_status* out);
/* The request's image as RGBA32F texels in the units of pt_accum: its mean XYZ,
 * or its denoised image when the status says `denoised`. */
pt_status pt_server_read(pt_server* server, int32_t id, float* host_rgba, size_t n_floats);
/* Attaches a denoiser (of the server's context; NULL detaches) for requests
 * with `denoise` set. Such a request also keeps the sums of its odd-index
 * samples, the first-hit features of a 4-sample camera-ray pass and the
 * denoised image (80 more bytes per pixel). After min_spp samples, while its
 * plain estimate misses the target, the error of its denoised image is
 * estimated from dual buffers: each half-sample image (even / odd sample
 * indices) is denoised and compared with the other, independent half, whose
 * variance is subtracted (the squared error of a denoised half, bias
 * included); half the variance of a denoised half, measured from the two
 * denoised halves, is taken off for the full sample count (phased in
 * linearly up to 64 samples per pixel, below which the network's output
 * variance falls more slowly than 1/samples). The
 * request finishes denoised when that estimate meets the target; at its
 * sample budget it returns the image with the lower estimate, so the plain
 * mean is the fallback. Resolutions the network does not take (width or
 * height above the denoiser's maximum or not a multiple of 4) are served
 * plain. */
pt_status pt_server_set_denoiser(pt_server* server, pt_denoiser* dn);
/* Frees a finished or waiting request's device memory and forgets it. */
pt_status pt_server_release(pt_server* server, int32_t id);
pt_status pt_server_counts(const pt_server* server, int32_t* waiting, int32_t* active, int64_t* bytes_in_use);

/* ---- Denoiser ------------------------------------------------------------
 * No counterpart in the reference. A 3-level U-Net (32/64/128 channels, 3x3
 * convolutions) that maps a low-sample image plus first-hit features to a
 * denoised XYZ image, trained and run on the GPU's matrix cores in BF16 with
 * FP32 accumulation and FP32 master weights (Adam).
 * Features: N images H x W (H, W multiples of 4), PT_FEATURES floats per
 * pixel as pt_render_features writes them: XYZ (3), albedo (3), normal (3),
 * distance (1), standard error of the mean of Y (1), samples (1). The network
 * sees log1p(XYZ), albedo, normal, 1 / (1 + distance), log1p(standard error)
 * and 1 / sqrt(samples), and predicts log1p(denoised XYZ) - log1p(input XYZ).
 * Radiance is best scaled to order one before the call. */
pt_status pt_dn_create(pt_ctx* ctx, int max_batch, int max_height, int max_width, uint64_t seed, pt_denoiser** out);
void pt_dn_destroy(pt_denoiser* dn);
const char* pt_dn_last_error(const pt_denoiser* dn);
int64_t pt_dn_param_count(const pt_denoiser* dn);
/* Training loss: the mean over pixels and XYZ of |log1p(denoised) -
 * log1p(target)| (the default) or of (denoised - target)^2, the squared error
 * whose sum the relative RMSE measures. */
#define PT_DN_LOSS_L1_LOG 0
#define PT_DN_LOSS_L2 1
pt_status pt_dn_set_loss(pt_denoiser* dn, int kind);
/* FP32 parameters: per layer, weights [9][Cout][Cin] then biases [Cout]. */
pt_status pt_dn_get_params(pt_denoiser* dn, float* host, size_t n);
pt_status pt_dn_set_params(pt_denoiser* dn, const float* host, size_t n);
/* features (N, H, W, PT_FEATURES) -> denoised XYZ (N, H, W, 3) */
pt_status pt_dn_denoise(pt_denoiser* dn, const float* features, int N, int H, int W, float* denoised);
/* Loss against target XYZ (N, H, W, 3) and the gradient of every parameter
 * (host_grads may be NULL). */
pt_status pt_dn_gradients(pt_denoiser* dn, const float* features, const float* target, int N, int H, int W,
                          float* loss, float* host_grads, size_t n_grads);
/* One Adam step (beta 0.9 / 0.999) on a batch. */
pt_status pt_dn_train_step(pt_denoiser* dn, const float* features, const float* target, int N, int H, int W, float lr,
                           float* loss);

#ifdef __cplusplus
}
#endif

#endif /* PATHTRACER_H */
