// ext_tracer.hip — PT_MODE_EXT: the extended, physically based spectral
// wavefront tracer for gfx950.
//
// Scene: spheres in a host-built W-wide BVH (W = 2, 4, 8) whose leaves index a
// leaf-ordered copy of the sphere records (one load per primitive test),
// traversed with a per-lane stack of 16-bit node indices in LDS;
// parallelograms (boxes become six) tested directly when there are at most 16
// of them, through their own BVH otherwise; infinite horizontal planes; SDF
// objects (mandelbulb, menger, blob, terrain) sphere-traced by one kernel per
// kind.
// Materials: Lambert with the reference's one-lobe spectrum or a sum of up to
// three Gaussians; smooth dielectric with Cauchy or Sellmeier dispersion;
// smooth conductor with a complex index n + ik. Emitters are the reference's
// normalised blackbodies; an optional uniform environment radiance.
// Wavelengths: hero wavelength uniform in [lambdaMin, lambdaMax) plus
// n_lambda - 1 rotated companions (Wilkie et al. 2014); a dispersive
// refraction keeps only the hero.
// Integrator: BSDF sampling + next-event estimation at Lambert vertices with
// the power heuristic, Russian roulette after `rr_depth` bounces.
// RNG: counter-based pcg_hash keyed on (seed, pixel, sample, bounce, dim), so a
// pixel's samples do not depend on how paths are batched into waves.
//
// Per bounce: intersect (+ classify into material queues) -> march<kind> per
// SDF kind [-> classify] -> shade_{miss, emit, lambert, dielectric, conductor}
// -> shadow (any-hit) -> next bounce. Queues are compacted with wave64 ballots;
// kernels run grid-stride over their queue.
#include <algorithm>
#include <map>
#include <cmath>
#include <cstring>
#include <vector>

#include "bvh.hpp"
#include "tracer.hpp"

namespace ptk {

using namespace ptd;

// in namespace ptk: the header holds device helpers only, no includes
#include "wave_queue.h"

// ---------------------------------------------------------------------------------
// Device scene.
enum : uint32_t { REF_SPHERE = 0u, REF_QUAD = 1u, REF_PLANE = 2u, REF_SDF = 3u, REF_LENS = 4u, REF_CYC = 5u };
constexpr uint32_t kRefShift = 29u, kRefMask = (1u << 29) - 1u, kMiss = 0xFFFFFFFFu;
enum : int { Q_MISS = 0, Q_EMIT = 1, Q_LAMBERT = 2, Q_DIELECTRIC = 3, Q_CONDUCTOR = 4, N_MATQ = 5 };
constexpr int kQuadBrute = 16;  // at most this many parallelograms are tested without a BVH
// Queue lengths: [0..1] extension queues, [2] shadow queue, [3..7] material
// queues, each counter on its own 256-byte line (device-scope atomics on one
// line serialise across the chip); kCntStride is in wave_queue.h.
// The shadow queue's length alternates between two counters by bounce parity:
// bounce b appends to CNT_SHADOW[b & 1], which the next bounce's trace reads
// while bounce b + 1 appends to the other one.
enum : int { CNT_SHADOW0 = 2, CNT_MAT = 3, CNT_SHADOW1 = 3 + N_MATQ, N_CNT = 4 + N_MATQ };
constexpr int shadow_cnt(int b) { return (b & 1) ? CNT_SHADOW1 : CNT_SHADOW0; }
// Persistent kernels' queue cursors: kSeg segments per queue, one line each
// (wave_queue.h).
// Cursor sets: persistent intersect, persistent shadow, and one per SDF kind for
// the persistent march kernels (CUR_MARCH + kind - 1).
enum : int { CUR_ISECT = 0, CUR_SHADOW = 1, CUR_MARCH = 2, N_CUR = 6 };

struct ExtMat {
    int type;        // pt::MaterialType
    int n;           // gaussian count | dispersion model
    int dispersive;  // dielectric index depends on wavelength
    int pad;
    float p[9];      // lobe (peak, sigma, invert) | gaussians (w, mean, sigma) x n | dispersion | (n, k)
};
struct ExtLight {
    int type;        // REF_SPHERE or REF_QUAD (sampled by NEE); REF_LENS or REF_CYC (found by BSDF sampling only)
    int prim;
    float T, lum;    // blackbody temperature, luminosity
    f3 p, e1, e2, n; // quad: corner, edges, unit normal; sphere: centre in p
    float area, radius;
};
struct ExtSdf {
    int kind, iters, mat, pad;
    float bailout;
    f3 lo, hi, c;
};
// The reference's lens (src/shader.comp:366-448) in its frame (object = v * M,
// x along the optical axis): two spheres of radius R with centres on the axis
// at +-(R - sp); cap 0 is the part of the first with x - centre <= -(R - h),
// cap 1 the part of the second with x - centre >= R - h (h the cap height).
// The solid is inside both spheres when converging, outside them otherwise.
struct ExtLens {
    f3 pos;
    m3 M;
    float R, h, sp, br2;  // sphere radius (2f), cap height, slice position, bounding radius^2
    int conv, mat, light, pad;
};
// The reference's Dupin cyclide (src/shader.comp:633-679): in its frame
// p = (v * M) / scale, the surface (|p|^2 + B^2 - D^2)^2 - 4 (A p.x - C D)^2
// - 4 B^2 p.z^2 = 0 inside a bounding sphere of radius^2 brad2.
struct ExtCyc {
    f3 pos;
    m3 M;
    f3 inv_scale;
    float A, B, C, D, brad2;
    int mat, light;
};
// A BVH over one primitive type with leaf-ordered primitive records.
struct BvhDev {
    const uint32_t* nodes;  // 8 * width words per node (bvh.hpp)
    const float4* prims;    // records in leaf order (1 float4 per sphere, 4 per quad)
    const int* ids;         // leaf slot -> primitive index
    int count;
};

struct ExtSceneDev {
    const float4* sph;  // (centre, radius) by sphere index
    const int2* sph_ml; // (material, light or -1)
    const float4* quad; // 4 float4 per quad: (Q, w.Q), (w, area), (e2 x w), (w x e1)
    const int2* quad_ml;
    const float* plane_y;
    const int2* plane_ml;
    const ExtSdf* sdf;
    const ExtMat* mat;
    const ExtLight* light;
    const ExtLens* lens;
    const ExtCyc* cyc;
    BvhDev sbvh, qbvh;
    int nsph, nquad, nplane, nsdf, nmat, nlight;
    int nlens, ncyc;
    int nlight_nee;     // lights [0, nlight_nee) are sampled by next-event estimation
    int quad_bvh;       // parallelograms go through qbvh
    int has_env;
    float env;
    int n_lambda;
    float lmin, lrange;
    int rr_depth;
    uint32_t seed;
    int camera;  // 0 lens, 1 pinhole
    int sdf_steps_max;
};

struct ExtFrame {
    int W, H, frame, spf, maxDepth;
    float camSize, apSize, apDist;
    f3 camPos, fwd;
    m3 camM;
    f3 lensPos;
    m3 lensM;
    float lensHalf, lensSlicePos, lensTwoF;
    int p0, pc, k0, m;
};

// A path's state that the shading kernels read and write, in one 64-byte
// record (one memory sector): they reach paths through the material queues,
// whose path indices grow sparse as paths end, so every field in an array of
// its own cost a sector of its own per path.
struct PathRec {
    float4 org;  // ray origin; w: hero wavelength
    float4 dir;  // ray direction; w unused
    float4 T;    // throughput per wavelength
    uint4 aux;   // x: RNG key, y: flags (bit 0 companions dropped, bit 1 last bounce specular), z: BSDF pdf bits,
                 // w: request slot
};

// A material-queue entry: the path and its hit, so that the shading kernels
// read the hit with the entry instead of gathering it.
struct MatEntry {
    int p;         // path index
    float t;       // hit distance
    uint32_t ref;  // primitive reference bits
    int pad;
};

// State of one extended wave.
struct ExtWave {
    PathRec* pr;   // by path index
    float4* acc;   // XYZ gathered so far (w unused)
    // (t, primitive reference bits) by entry of the current extension queue:
    // the traversal kernels write it in the order they take the entries and
    // the classification reads it in that order
    float2* hit;
    int* q[2];
    // the extension rays of queue c in queue order (entry e: the ray of path
    // q[c][e]), written where the shading kernels append, so the intersect
    // kernels read them coalesced; org / dir keep them by path for shading
    float4* rq_org[2];
    float4* rq_dir[2];
    MatEntry* mq[N_MATQ];
    float4* s_org; // shadow ray origin, w = tmax
    float4* s_dir; // direction, w = path index bits
    float4* s_xyz; // unoccluded contribution
    float4* aov0;  // first hit: albedo (xyz), distance (w)
    float4* aov1;  // first hit: shading normal (xyz)
    uint32_t* smask;  // SDF scenes, by entry like hit: bit k set when the ray enters the box of an SDF of kind k before its analytic hit
    float* sum;
    int* counts;   // N_CNT queue lengths, kCntStride apart
    int* cursor;   // N_CUR x kSeg persistent-kernel cursors, kCntStride apart
    unsigned long long* stats;
    int cap;
};

// Continuous batching: one wave holds the paths of several requests. A path's
// request slot is aux.w (a shadow record's is s_xyz.w); the slot indexes the
// batch's scene and frame tables. Kernels instantiated with MULTI = false read
// the one scene and frame of their arguments instead.
struct ExtBatch {
    const ExtSceneDev* scenes = nullptr;
    const ExtFrame* frames = nullptr;
};

// The scene / frame of request slot rid: the batch's tables with MULTI, the
// kernel's own arguments otherwise.
template <bool MULTI>
__device__ __forceinline__ const ExtSceneDev& scene_of(const ExtSceneDev& S0, const ExtBatch& B, uint32_t rid) {
    return MULTI ? B.scenes[rid] : S0;
}

// The ray of entry e of extension queue `cur`, from the queue-ordered records.
__device__ __forceinline__ void ray_of(const ExtWave& Wv, int cur, int e, f3& o, f3& d) {
    const float4 a = Wv.rq_org[cur][e], b = Wv.rq_dir[cur][e];
    o = mk3(a.x, a.y, a.z);
    d = mk3(b.x, b.y, b.z);
}
// The ray path p traces (w of the origin: the hero wavelength), from the
// path's record: raygen and the shading kernels write each ray both there and
// into its queue's records.
__device__ __forceinline__ void path_ray(const ExtWave& Wv, int p, float4& org, float4& dir) {
    org = Wv.pr[p].org;
    dir = Wv.pr[p].dir;
}

static __constant__ float c_cie[PT_CIE_SIZE] = {
#include "cie1931_2deg.inc"
};

constexpr int kBlk = 256;
constexpr float kPiF = 3.14159265358979f;
constexpr float kEps = 1e-6f;

// ---------------------------------------------------------------------------------
// Counter-based random numbers: pcg_hash of (path key, bounce, dimension).
__device__ __forceinline__ uint32_t path_key(uint32_t seed, uint32_t pixel, uint32_t sample) {
    return pcg(pcg(pcg(seed) + pixel) + sample);
}
__device__ __forceinline__ float rnd(uint32_t key, uint32_t bounce, uint32_t dim) {
    uint32_t h = pcg(key + pcg((bounce << 4) | dim));
    return (float)(h >> 8) * 5.9604644775390625e-8f;  // [0, 1) in steps of 2^-24
}
constexpr uint32_t kCamBounce = 0xFFFu;

__device__ __forceinline__ f3 cmf(float l) {
    float x = fminf(fmaxf(l - 360.0f, 0.0f), 439.999f);
    int i = (int)x;
    float t = x - (float)i;
    const float* a = &c_cie[3 * i];
    return mk3(a[0] + (a[3] - a[0]) * t, a[1] + (a[4] - a[1]) * t, a[2] + (a[5] - a[2]) * t);
}
__device__ __forceinline__ float planck_norm(float lnm, float T, float lum) {
    // The reference's emitter: Planck's law at lnm normalised by its Wien peak, times luminosity.
    float t = fmaxf(T, 0.0f);
    float l = lnm * 1e-9f;
    float l2 = l * l;
    float b = 1.1910429724e-16f / (l2 * l2 * l) / (__expf(0.014387768775f / (l * t)) - 1.0f);
    float t2 = t * t;
    float peak = 4.0956746759e-6f * (t2 * t2 * t);
    return (b / peak) * fmaxf(lum, 0.0f);
}
__device__ __forceinline__ float reflectance(const ExtMat& m, float l) {
    if (m.type == pt::MAT_LAMBERT_LOBE) {
        float x = (l - m.p[0]) / (2.0f * m.p[1] * m.p[1]);
        float r = __expf(-x * x);
        return m.p[2] != 0.0f ? 1.0f - r : r;
    }
    float s = 0.0f;
    for (int i = 0; i < m.n; ++i) {
        float d = (l - m.p[3 * i + 1]) / m.p[3 * i + 2];
        s += m.p[3 * i] * __expf(-0.5f * d * d);
    }
    return fminf(s, 1.0f);
}
__device__ __forceinline__ float ior(const ExtMat& m, float lnm) {
    float lu = lnm * 1e-3f;
    float l2 = lu * lu;
    if (m.n == pt::DISP_CAUCHY) return m.p[0] + m.p[1] / l2;
    float n2 = 1.0f + m.p[0] * l2 / (l2 - m.p[3]) + m.p[1] * l2 / (l2 - m.p[4]) + m.p[2] * l2 / (l2 - m.p[5]);
    return sqrtf(n2);
}
// Unpolarised Fresnel reflectance of a dielectric interface.
__device__ __forceinline__ float fresnel_dielectric(float cos_i, float cos_t, float eta_i, float eta_t) {
    float rs = (eta_i * cos_i - eta_t * cos_t) / (eta_i * cos_i + eta_t * cos_t);
    float rp = (eta_t * cos_i - eta_i * cos_t) / (eta_t * cos_i + eta_i * cos_t);
    return 0.5f * (rs * rs + rp * rp);
}
// Unpolarised Fresnel reflectance of a conductor with index n + ik seen from
// vacuum, with cos_t = sqrt(1 - sin^2 / eta^2) in complex arithmetic.
__device__ __forceinline__ float fresnel_conductor(float cos_i, float n, float k) {
    float s2 = fmaxf(1.0f - cos_i * cos_i, 0.0f);
    float er = n * n - k * k, ei = 2.0f * n * k;  // eta^2
    float den = er * er + ei * ei;
    float zr = 1.0f - s2 * er / den, zi = s2 * ei / den;
    float m = sqrtf(zr * zr + zi * zi);
    float cr = sqrtf(fmaxf(0.5f * (m + zr), 0.0f));
    float ci = copysignf(sqrtf(fmaxf(0.5f * (m - zr), 0.0f)), zi);
    float ect_r = n * cr - k * ci, ect_i = n * ci + k * cr;  // eta cos_t
    float a_r = cos_i - ect_r, a_i = -ect_i, b_r = cos_i + ect_r, b_i = ect_i;
    float rperp = (a_r * a_r + a_i * a_i) / (b_r * b_r + b_i * b_i);
    float eci_r = n * cos_i, eci_i = k * cos_i;  // eta cos_i
    float c_r = eci_r - cr, c_i = eci_i - ci, d_r = eci_r + cr, d_i = eci_i + ci;
    float rpar = (c_r * c_r + c_i * c_i) / (d_r * d_r + d_i * d_i);
    return 0.5f * (rperp + rpar);
}
// Orthonormal basis around n (Duff et al. 2017, "Building an orthonormal basis, revisited").
__device__ __forceinline__ void onb(f3 n, f3& t, f3& b) {
    float s = copysignf(1.0f, n.z);
    float a = -1.0f / (s + n.z);
    float c = n.x * n.y * a;
    t = mk3(1.0f + s * n.x * n.x * a, s * c, -s * n.x);
    b = mk3(c, s + n.y * n.y * a, -n.y);
}
__device__ __forceinline__ float offset_scale(f3 p) {
    return 5e-6f * fmaxf(1.0f, fmaxf(fabsf(p.x), fmaxf(fabsf(p.y), fabsf(p.z))));
}

// Hardware square root, reciprocal and reciprocal square root (about 1 ulp),
// for the extended mode's sphere tests and distance estimators. This file is
// built with correctly rounded division and square root, which the reference
// mode's bit parity with the CPU needs and which expand to 10-17 instructions
// each; the extended mode is checked against the FP64 oracle within tolerances.
__device__ __forceinline__ float hw_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ float hw_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float hw_len(f3 v) { return hw_sqrt(dot(v, v)); }
// Dot product with fused multiply-adds: two v_fma_f32 and one v_mul_f32
// instead of three multiplies and two adds, with fewer roundings. The file
// is built without floating-point contraction (for the reference mode), so
// the traversal's hot arithmetic writes its fused operations explicitly.
__device__ __forceinline__ float fdot(f3 a, f3 b) { return fmaf(a.x, b.x, fmaf(a.y, b.y, a.z * b.z)); }

// ---------------------------------------------------------------------------------
// Primitive intersection.
__device__ __forceinline__ bool isect_sphere(float4 s, f3 o, f3 d, float tmax, float& t) {
    f3 oc = o - mk3(s.x, s.y, s.z);
    float b = fdot(oc, d);
    // component of oc orthogonal to the ray (Haines et al., Ray Tracing Gems ch. 7)
    f3 f = mk3(fmaf(-b, d.x, oc.x), fmaf(-b, d.y, oc.y), fmaf(-b, d.z, oc.z));
    float disc = fmaf(s.w, s.w, -fdot(f, f));
    if (disc < 0.0f) return false;
    float q = hw_sqrt(disc);
    float t0 = -b - q;
    if (t0 > kEps && t0 < tmax) { t = t0; return true; }
    float t1 = -b + q;
    if (t1 > kEps && t1 < tmax) { t = t1; return true; }
    return false;
}
__device__ __forceinline__ bool isect_quad(const float4* q, f3 o, f3 d, float tmax, float& t) {
    const float4 q0 = q[0], q1 = q[1];
    f3 w = mk3(q1.x, q1.y, q1.z);
    float den = fdot(w, d);
    if (fabsf(den) < 1e-20f) return false;
    float tt = (q0.w - fdot(w, o)) * hw_rcp(den);
    if (!(tt > kEps && tt < tmax)) return false;
    const float4 q2 = q[2], q3 = q[3];
    f3 hp = mk3(fmaf(d.x, tt, o.x) - q0.x, fmaf(d.y, tt, o.y) - q0.y, fmaf(d.z, tt, o.z) - q0.z);
    float a = fdot(mk3(q2.x, q2.y, q2.z), hp);
    float b = fdot(mk3(q3.x, q3.y, q3.z), hp);
    if (a < 0.0f || a > 1.0f || b < 0.0f || b > 1.0f) return false;
    t = tt;
    return true;
}

// Stages the flat parallelogram list (quads, nq of them; nq = 0 for a list
// with its own BVH) that the calling wave's active lanes test: when all of
// them trace the same scene (always with one scene; in a batched wave, when
// its lanes hold one request's paths), the wave loads the list into its LDS
// slice, one 16-byte vector per lane (64 lanes: up to kQuadBrute
// parallelograms), and returns true (the same on every lane); otherwise each
// lane reads its own scene's list from global memory. Every lane of the wave
// calls it (the caller's loop is uniform).
__device__ __forceinline__ bool wave_quads(const float4* quads, int nq, bool active, float4* slice) {
    static_assert(4 * kQuadBrute <= 64, "one vector per lane");
    const uint64_t act = __ballot(active);
    if (act == 0ull) return false;
    const int first = __ffsll((unsigned long long)act) - 1;
    const uint64_t qa = (uint64_t)(uintptr_t)quads;
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)qa, first), hi = (uint32_t)__shfl((int)(uint32_t)(qa >> 32), first);
    const float4* q0 = reinterpret_cast<const float4*>((uintptr_t)(((uint64_t)hi << 32) | lo));
    const int n0 = __shfl(nq, first);
    if (__ballot(active && (quads != q0 || nq != n0)) != 0ull || n0 > kQuadBrute) return false;
    const int lane = (int)__lane_id();
    if (lane < 4 * n0) slice[lane] = q0[lane];
    // the lanes' LDS writes before any lane's reads (one wave: LDS operations in order)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    return true;
}

// The lens's nearest cap point in (kEps, tmax) (struct ExtLens). Each cap's
// sphere is solved like isect_sphere; a root counts when it lies on the cap.
__device__ __forceinline__ bool isect_lens(const ExtLens& L, f3 o, f3 d, float tmax, float& t) {
    const f3 oc0 = o - L.pos;
    {  // the bounding sphere
        const float b = fdot(oc0, d);
        const f3 f = mk3(fmaf(-b, d.x, oc0.x), fmaf(-b, d.y, oc0.y), fmaf(-b, d.z, oc0.z));
        if (fdot(f, f) > L.br2) return false;
    }
    const f3 lo = vmul(oc0, L.M), ld = vmul(d, L.M);
    const float R = L.R, rim = R - L.h;
    bool hit = false;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const f3 oc = mk3(lo.x - (k == 0 ? R - L.sp : L.sp - R), lo.y, lo.z);
        const float b = fdot(oc, ld);
        const f3 f = mk3(fmaf(-b, ld.x, oc.x), fmaf(-b, ld.y, oc.y), fmaf(-b, ld.z, oc.z));
        const float disc = fmaf(R, R, -fdot(f, f));
        if (disc < 0.0f) continue;
        const float q = hw_sqrt(disc);
        const float r0 = -b - q, r1 = -b + q;
        const float x0 = fmaf(ld.x, r0, oc.x), x1 = fmaf(ld.x, r1, oc.x);
        const bool c0 = k == 0 ? x0 <= -rim : x0 >= rim, c1 = k == 0 ? x1 <= -rim : x1 >= rim;
        if (c0 && r0 > kEps && r0 < tmax) {
            tmax = r0;
            hit = true;
        } else if (c1 && r1 > kEps && r1 < tmax) {
            tmax = r1;
            hit = true;
        }
    }
    if (hit) t = tmax;
    return hit;
}
// The lens solid's outward normal at a surface point p: from the centre of the
// cap sphere p lies on (the one whose distance is nearer R), outwards for a
// converging lens, inwards for a diverging one.
__device__ __forceinline__ f3 lens_normal(const ExtLens& L, f3 p) {
    const f3 lp = vmul(p - L.pos, L.M);
    const f3 a = mk3(lp.x - (L.R - L.sp), lp.y, lp.z), b = mk3(lp.x - (L.sp - L.R), lp.y, lp.z);
    const float da = fabsf(hw_sqrt(fdot(a, a)) - L.R), db = fabsf(hw_sqrt(fdot(b, b)) - L.R);
    f3 n = normalize(da <= db ? a : b);
    if (!L.conv) n = -n;
    return mvmul(L.M, n);
}

// Polynomial c[0] + c[1] x + ... + c[N] x^N (Horner, fused multiply-adds).
template <int N>
__device__ __forceinline__ float poly_eval(const float* c, float x) {
    float v = c[N];
#pragma unroll
    for (int k = N - 1; k >= 0; --k) v = fmaf(v, x, c[k]);
    return v;
}
// The sign changes of the degree-N polynomial c between consecutive entries of
// `ends` (ascending; at most NE of them), each found by bisection: `ends` are
// the interval ends and the real roots of the derivative between them, so the
// polynomial is monotonic on each interval and a sign change there is its only
// root. Writes the roots found, ascending, into `roots` (the first only with
// first_only) and returns their number.
template <int N, int NE>
__device__ __forceinline__ int poly_brackets(const float* c, const float* ends, int ne, float* roots, bool first_only) {
    int m = 0;
    float fa = poly_eval<N>(c, ends[0]);
    for (int i = 0; i + 1 < ne && i + 1 < NE; ++i) {
        float a = ends[i], b = ends[i + 1];
        const float fb = poly_eval<N>(c, b);
        if ((fa < 0.0f) != (fb < 0.0f)) {
            float fl = fa;
            for (int it = 0; it < 40; ++it) {
                const float mid = 0.5f * (a + b);
                if (!(mid > a && mid < b)) break;
                const float fm = poly_eval<N>(c, mid);
                if ((fm < 0.0f) == (fl < 0.0f)) {
                    a = mid;
                    fl = fm;
                } else {
                    b = mid;
                }
            }
            roots[m++] = 0.5f * (a + b);
            if (first_only) return m;
        }
        fa = fb;
    }
    return m;
}
// Smallest root of the quartic c[0..4] in (0, hi), by isolating the real roots
// of its derivatives in turn: its second derivative (a quadratic, in closed
// form), then its first derivative (a cubic) on the intervals between those,
// then the quartic itself on the intervals between the cubic's roots. A root
// where the quartic only touches zero (a tangency) is not found.
__device__ __forceinline__ bool quartic_first_root(const float c[5], float hi, float& root) {
    const float d1[4] = {c[1], 2.0f * c[2], 3.0f * c[3], 4.0f * c[4]};
    const float d2[3] = {d1[1], 2.0f * d1[2], 3.0f * d1[3]};
    float e2[4];
    int n2 = 0;
    e2[n2++] = 0.0f;
    {  // the roots of the quadratic d2 in (0, hi), ascending
        const float A = d2[2], B = d2[1], C = d2[0];
        if (A != 0.0f) {
            const float disc = fmaf(B, B, -4.0f * A * C);
            if (disc >= 0.0f) {
                const float q = -0.5f * (B + copysignf(hw_sqrt(disc), B));
                float r0 = q / A, r1 = q != 0.0f ? C / q : r0;
                if (r0 > r1) {
                    const float tmp = r0;
                    r0 = r1;
                    r1 = tmp;
                }
                if (r0 > 0.0f && r0 < hi) e2[n2++] = r0;
                if (r1 > 0.0f && r1 < hi && r1 != r0) e2[n2++] = r1;
            }
        } else if (B != 0.0f) {
            const float r0 = -C / B;
            if (r0 > 0.0f && r0 < hi) e2[n2++] = r0;
        }
    }
    e2[n2++] = hi;
    float e3[5], r3[3];
    int n3 = 0;
    e3[n3++] = 0.0f;
    const int m3 = poly_brackets<3, 4>(d1, e2, n2, r3, false);
    for (int i = 0; i < m3; ++i) e3[n3++] = r3[i];
    e3[n3++] = hi;
    float r4[1];
    if (poly_brackets<4, 5>(c, e3, n3, r4, true) == 0) return false;
    root = r4[0];
    return true;
}
// The cyclide's nearest hit in (kEps, tmax) (struct ExtCyc): the implicit
// function along the ray is a quartic in s = t - ta, ta the entry into the
// bounding sphere (expanding about ta keeps the coefficients of the short
// search interval well scaled in FP32).
__device__ __forceinline__ bool isect_cyclide(const ExtCyc& C, f3 o, f3 d, float tmax, float& t) {
    const f3 oc = o - C.pos;
    const float b = fdot(oc, d);
    const f3 f = mk3(fmaf(-b, d.x, oc.x), fmaf(-b, d.y, oc.y), fmaf(-b, d.z, oc.z));
    const float disc = C.brad2 - fdot(f, f);
    if (disc < 0.0f) return false;
    const float q = hw_sqrt(disc);
    const float ta = fmaxf(-b - q, kEps), tb = fminf(-b + q, tmax);
    if (!(ta < tb)) return false;
    const f3 lv = vmul(d, C.M), lo = vmul(oc, C.M);
    const f3 v = mk3(lv.x * C.inv_scale.x, lv.y * C.inv_scale.y, lv.z * C.inv_scale.z);
    const f3 p = mk3(fmaf(v.x, ta, lo.x * C.inv_scale.x), fmaf(v.y, ta, lo.y * C.inv_scale.y),
                     fmaf(v.z, ta, lo.z * C.inv_scale.z));
    // S(s) = s2 s^2 + s1 s + s0, L1(s) = l1 s + l0, L2(s) = m1 s + m0;
    // F = S^2 - 4 L1^2 - 4 B^2 L2^2
    const float B2 = C.B * C.B;
    const float s2 = fdot(v, v), s1 = 2.0f * fdot(p, v), s0 = fdot(p, p) + (B2 - C.D * C.D);
    const float l1 = C.A * v.x, l0 = fmaf(C.A, p.x, -C.C * C.D), m1 = v.z, m0 = p.z;
    const float c[5] = {
        fmaf(s0, s0, -4.0f * fmaf(l0, l0, B2 * m0 * m0)),
        2.0f * fmaf(s1, s0, -4.0f * fmaf(l1, l0, B2 * m1 * m0)),
        fmaf(s1, s1, fmaf(2.0f * s2, s0, -4.0f * fmaf(l1, l1, B2 * m1 * m1))),
        2.0f * s2 * s1,
        s2 * s2,
    };
    float s;
    if (!quartic_first_root(c, tb - ta, s)) return false;
    t = ta + s;
    return t > kEps && t < tmax;
}
// Outward normal (the world gradient of the implicit function) at p.
__device__ __forceinline__ f3 cyclide_normal(const ExtCyc& C, f3 p) {
    const f3 l = vmul(p - C.pos, C.M);
    const f3 x = mk3(l.x * C.inv_scale.x, l.y * C.inv_scale.y, l.z * C.inv_scale.z);
    const float S = fdot(x, x) + (C.B * C.B - C.D * C.D);
    const f3 g = mk3(4.0f * x.x * S - 8.0f * C.A * fmaf(C.A, x.x, -C.C * C.D), 4.0f * x.y * S,
                     4.0f * x.z * S - 8.0f * C.B * C.B * x.z);
    return normalize(mvmul(C.M, mk3(g.x * C.inv_scale.x, g.y * C.inv_scale.y, g.z * C.inv_scale.z)));
}
// Lenses and cyclides (a few per scene): the closest hit below t (ANY: the
// first). The k_ext_extras kernels apply it after the intersect kernels, so
// the traversal kernels of scenes without them carry none of its code.
template <bool ANY>
__device__ __forceinline__ bool trace_extras(const ExtSceneDev& S, f3 o, f3 d, float& t, uint32_t& ref) {
    bool hit = false;
    for (int i = 0; i < S.nlens; ++i) {  // uniform index
        float tl;
        if (isect_lens(S.lens[i], o, d, t, tl)) {
            t = tl;
            ref = (REF_LENS << kRefShift) | (uint32_t)i;
            hit = true;
            if (ANY) return true;
        }
    }
    for (int i = 0; i < S.ncyc; ++i) {
        float tc;
        if (isect_cyclide(S.cyc[i], o, d, t, tc)) {
            t = tc;
            ref = (REF_CYC << kRefShift) | (uint32_t)i;
            hit = true;
            if (ANY) return true;
        }
    }
    return hit;
}

// ---------------------------------------------------------------------------------
// Wide-BVH traversal. `stack` is this lane's column of the block's LDS stack
// (stride kBlk); STACK bounds the entries (checked on the host against the
// tree's exact worst case) and SE is the entry type (16-bit when the tree has
// fewer than 65536 nodes).
// Reciprocal direction for the slab tests (hardware v_rcp_f32, about 1 ulp:
// box bounds only decide which subtrees are visited, and a primitive's own
// test decides its hits).
__device__ __forceinline__ void ray_inv(f3 o, f3 d, f3& inv, f3& oi) {
    const float big = 1e30f;
    inv = mk3(fabsf(d.x) > 1e-30f ? hw_rcp(d.x) : copysignf(big, d.x),
              fabsf(d.y) > 1e-30f ? hw_rcp(d.y) : copysignf(big, d.y),
              fabsf(d.z) > 1e-30f ? hw_rcp(d.z) : copysignf(big, d.z));
    oi = o * inv;
}

// Visits `node`: tests its W children's boxes, the primitives of the leaf
// children (which can shrink tbest), and picks the next node: the nearest
// internal child the ray enters (kept in a register, the others pushed far to
// near), else the top of the stack, else -1 when the traversal is over. Only
// pushes and pops touch the LDS stack, so most visits' dependent chain is one
// node fetch. `hit` becomes the leaf slot of the closest hit so far. With ANY,
// returns true at the first hit, and children are taken in node order.
// Child count of a sphere-BVH leaf holding one sphere, stored in the parent's
// record: lo_x, lo_y, lo_z, hi_x are its centre and radius, child its leaf slot.
constexpr int kInlineSphere = -2;

template <int W, typename SE, bool ANY, bool QUAD>
__device__ __forceinline__ bool bvh_step(const BvhDev& B, f3 o, f3 d, f3 inv, f3 oi, float& tbest, int& hit,
                                         int& node, int& sp, SE* stack) {
    const float big = 1e30f;
    const float4* nd = reinterpret_cast<const float4*>(B.nodes + (size_t)node * (8 * W));
    float lo_x[W], lo_y[W], lo_z[W], hi_x[W], hi_y[W], hi_z[W];
    int kid[W], cnt[W];
    {
        const float* f = reinterpret_cast<const float*>(nd);
        const int* iv = reinterpret_cast<const int*>(nd);
#pragma unroll
        for (int j = 0; j < W; ++j) {
            lo_x[j] = f[0 * W + j]; lo_y[j] = f[1 * W + j]; lo_z[j] = f[2 * W + j];
            hi_x[j] = f[3 * W + j]; hi_y[j] = f[4 * W + j]; hi_z[j] = f[5 * W + j];
            kid[j] = iv[6 * W + j]; cnt[j] = iv[7 * W + j];
        }
    }
    float tn[W];
#pragma unroll
    for (int j = 0; j < W; ++j) {
        float tx0 = fmaf(lo_x[j], inv.x, -oi.x), tx1 = fmaf(hi_x[j], inv.x, -oi.x);
        float ty0 = fmaf(lo_y[j], inv.y, -oi.y), ty1 = fmaf(hi_y[j], inv.y, -oi.y);
        float tz0 = fmaf(lo_z[j], inv.z, -oi.z), tz1 = fmaf(hi_z[j], inv.z, -oi.z);
        float lo = fmaxf(fmaxf(fminf(tx0, tx1), fminf(ty0, ty1)), fmaxf(fminf(tz0, tz1), 0.0f));
        float hi = fminf(fminf(fmaxf(tx0, tx1), fmaxf(ty0, ty1)), fminf(fmaxf(tz0, tz1), tbest));
        tn[j] = (lo <= hi && cnt[j] >= 0) ? lo : big;
    }
    // Leaves first (they can shrink tbest), then internal children.
#pragma unroll
    for (int j = 0; j < W; ++j) {
        if (!QUAD && cnt[j] == kInlineSphere) {
            float t;
            if (isect_sphere(make_float4(lo_x[j], lo_y[j], lo_z[j], hi_x[j]), o, d, tbest, t)) {
                tbest = t;
                hit = kid[j];
                if (ANY) return true;
            }
        } else if (tn[j] < big && cnt[j] > 0) {
            for (int k = kid[j]; k < kid[j] + cnt[j]; ++k) {
                float t;
                bool h = QUAD ? isect_quad(B.prims + 4 * k, o, d, tbest, t) : isect_sphere(B.prims[k], o, d, tbest, t);
                if (h) {
                    tbest = t;
                    hit = k;
                    if (ANY) return true;
                }
            }
            tn[j] = big;
        }
    }
    int next = -1;
    if (ANY) {
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (tn[j] < tbest && cnt[j] == 0) {
                if (next < 0) next = kid[j];
                else stack[(sp++) * kBlk] = (SE)kid[j];
            }
        }
    } else {
        // entered internal children sorted far to near: ks[n - 1] is the nearest
        float ds[W];
        int ks[W];
        int n = 0;
#pragma unroll
        for (int j = 0; j < W; ++j) {
            if (tn[j] < tbest && cnt[j] == 0) {
                float t = tn[j];
                int c = kid[j];
                int p = n++;
                while (p > 0 && ds[p - 1] < t) {
                    ds[p] = ds[p - 1];
                    ks[p] = ks[p - 1];
                    --p;
                }
                ds[p] = t;
                ks[p] = c;
            }
        }
        if (n > 0) {
            for (int j = 0; j < n - 1; ++j) stack[(sp++) * kBlk] = (SE)ks[j];
            next = ks[n - 1];
        }
    }
    if (next < 0 && sp > 0) next = (int)stack[(--sp) * kBlk];
    node = next;
    return false;
}

// Whole traversal of one ray. Returns the leaf slot of the closest hit (or of
// any hit with ANY), -1 if none; tbest shrinks to the hit distance.
// `visits` counts the nodes visited.
template <int W, int STACK, typename SE, bool ANY, bool QUAD>
__device__ __forceinline__ int bvh_trace(const BvhDev& B, f3 o, f3 d, float& tbest, SE* stack, unsigned& visits) {
    f3 inv, oi;
    ray_inv(o, d, inv, oi);
    int hit = -1, sp = 0, node = 0;
    while (node >= 0) {
        ++visits;
        if (bvh_step<W, SE, ANY, QUAD>(B, o, d, inv, oi, tbest, hit, node, sp, stack)) return hit;
    }
    return hit;
}

// Closest (or any, with ANY) hit over parallelograms, spheres and planes.
template <int W, int STACK, typename SE, bool ANY>
// `staged`: the scene's flat parallelogram list is in the wave's LDS slice
// (wave_quads), else it is read from global memory.
__device__ __forceinline__ bool scene_trace(const ExtSceneDev& S, bool staged, const float4* slice, f3 o, f3 d,
                                            float& t, uint32_t& ref, SE* stack, unsigned& visits) {
    bool hit = false;
    if (S.quad_bvh) {
        int k = bvh_trace<W, STACK, SE, ANY, true>(S.qbvh, o, d, t, stack, visits);
        if (k >= 0) {
            ref = (REF_QUAD << kRefShift) | (uint32_t)S.qbvh.ids[k];
            hit = true;
            if (ANY) return true;
        }
    } else {
        // one loop per source (a branch uniform over the wave): a pointer that
        // may be either would make every load a generic one
        if (staged) {
            for (int i = 0; i < S.nquad; ++i) {
                float tq;
                if (isect_quad(slice + 4 * i, o, d, t, tq)) {
                    t = tq;
                    ref = (REF_QUAD << kRefShift) | (uint32_t)i;
                    hit = true;
                    if (ANY) return true;
                }
            }
        } else {
            for (int i = 0; i < S.nquad; ++i) {
                float tq;
                if (isect_quad(S.quad + 4 * i, o, d, t, tq)) {
                    t = tq;
                    ref = (REF_QUAD << kRefShift) | (uint32_t)i;
                    hit = true;
                    if (ANY) return true;
                }
            }
        }
    }
    if (S.nsph > 0) {
        int k = bvh_trace<W, STACK, SE, ANY, false>(S.sbvh, o, d, t, stack, visits);
        if (k >= 0) {
            ref = (REF_SPHERE << kRefShift) | (uint32_t)S.sbvh.ids[k];
            hit = true;
            if (ANY) return true;
        }
    }
    for (int i = 0; i < S.nplane; ++i) {
        float tp = (ld_ro(S.plane_y + i) - o.y) / d.y;
        if (tp > kEps && tp < t) {
            t = tp;
            ref = (REF_PLANE << kRefShift) | (uint32_t)i;
            hit = true;
            if (ANY) return true;
        }
    }
    return hit;
}

// ---------------------------------------------------------------------------------
// SDF distance estimators (extended mode: the shapes of sdfs/*.glsl in plain
// FP32 arithmetic).
// `it` counts the iterations executed (the loop leaves early past the bailout).
// One iteration of the mandelbulb's power-8 map, z <- z^8 + p in spherical form
// (the angles' multiples by Chebyshev polynomials of their cosines), with the
// running derivative dr <- 8 r^7 dr + 1; r2 = |z|^2 before and after. The
// march is bound by VALU issue, so the iteration is written for it: its
// multiply-adds are fused (this file is built without contraction), and it
// takes two reciprocal square roots, 1/r = rsq(r2) (then r = r2 * (1/r)) and
// 1/l = rsq(x^2 + z^2) with l = sqrt(x^2 + z^2) the polar radius, so that
// sin(phi) = l / r = (x^2 + z^2) * (1/l) * (1/r): a square root and two
// reciprocals would be four quarter-rate instructions.
__device__ __forceinline__ void mandelbulb_iter(f3& z, float& r2, float& dr, f3 p) {
    const float invr = hw_rsq(r2), r = r2 * invr;
    const float r4 = r2 * r2;
    dr = fmaf(8.0f * (r4 * r2 * r), dr, 1.0f);
    const float l2 = fmaf(z.x, z.x, z.z * z.z);
    const float invl = hw_rsq(l2);
    const float r8 = r4 * r4;
    const float cost = z.z * invl, cosp = z.y * invr, sint = z.x * invl, sinp = invr * (l2 * invl);
    const float ct2 = cost * cost, cp2 = cosp * cosp;
    const float cost8 = fmaf(fmaf(fmaf(fmaf(128.0f, ct2, -256.0f), ct2, 160.0f), ct2, -32.0f), ct2, 1.0f);
    const float cosp8 = fmaf(fmaf(fmaf(fmaf(128.0f, cp2, -256.0f), cp2, 160.0f), cp2, -32.0f), cp2, 1.0f);
    const float sint8 = fmaf(fmaf(fmaf(128.0f, ct2, -192.0f), ct2, 80.0f), ct2, -8.0f) * sint * cost;
    const float sinp8 = fmaf(fmaf(fmaf(128.0f, cp2, -192.0f), cp2, 80.0f), cp2, -8.0f) * sinp * cosp;
    z = mk3(fmaf(r8, sint8 * sinp8, p.x), fmaf(r8, cosp8, p.y), fmaf(r8, cost8 * sinp8, p.z));
    r2 = fdot(z, z);
}
// The distance estimate 0.5 log(r) r / dr from the map's last iterate (r^2 = r2).
__device__ __forceinline__ float mandelbulb_dist(float r2, float dr) {
    return 0.25f * __logf(r2) * hw_sqrt(r2) * hw_rcp(dr);
}
__device__ __forceinline__ float ext_mandelbulb(f3 p, int iters, float bailout, unsigned& it) {
    f3 z = p;
    float r2 = fdot(z, z);
    float dr = 1.0f;
    const float bail2 = bailout * bailout;
    for (int i = 0; i < iters; ++i) {
        ++it;
        mandelbulb_iter(z, r2, dr, p);
        if (r2 > bail2) break;
    }
    return mandelbulb_dist(r2, dr);
}
__device__ __forceinline__ float ext_menger(f3 p, int levels) {
    f3 q = abs3(p) - mk3(1.0f, 1.0f, 1.0f);
    float d = hw_len(max3(q, mk3(0.0f, 0.0f, 0.0f))) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
    float s = 1.0f;
    for (int m = 0; m < levels; ++m) {
        f3 ps = p * s;
        f3 a = mk3(fmaf(-2.0f, floorf(ps.x * 0.5f), ps.x) - 1.0f, fmaf(-2.0f, floorf(ps.y * 0.5f), ps.y) - 1.0f,
                   fmaf(-2.0f, floorf(ps.z * 0.5f), ps.z) - 1.0f);
        s *= 3.0f;
        f3 r = abs3(mk3(fmaf(-3.0f, fabsf(a.x), 1.0f), fmaf(-3.0f, fabsf(a.y), 1.0f), fmaf(-3.0f, fabsf(a.z), 1.0f)));
        float da = fmaxf(r.x, r.y), db = fmaxf(r.y, r.z), dc = fmaxf(r.z, r.x);
        d = fmaxf(d, (fminf(da, fminf(db, dc)) - 1.0f) * hw_rcp(s));
    }
    return d;
}
__device__ __forceinline__ float ext_smin(float x, float y) {
    const float k = 0.12f;
    float h = fmaxf(k - fabsf(x - y), 0.0f) * (1.0f / k);
    return fminf(x, y) - h * h * h * k * (1.0f / 6.0f);
}
__device__ __forceinline__ float ext_blob(f3 p) {
    float s1 = hw_len(p) - 1.0f;
    float s2 = hw_len(mk3(p.x, p.y - 1.0f, p.z)) - 0.7f;
    float s3 = hw_len(mk3(p.x - 0.25f, p.y - 1.25f, p.z - 0.5f)) - 0.25f;
    float s4 = hw_len(mk3(p.x + 0.25f, p.y - 1.25f, p.z - 0.5f)) - 0.25f;
    return ext_smin(ext_smin(ext_smin(s1, s2), s3), s4);
}
__device__ __forceinline__ float ext_terrain_h(float x) {
    return __sinf(0.0625f * x) + __sinf(0.125f * (x + 10.0f)) + 0.25f * __sinf(0.25f * x) +
           0.125f * __sinf(0.5f * x) + 0.25f * __sinf(x) + 0.25f * __sinf(2.0f * x + 1.0f) +
           0.125f * __sinf(4.0f * x + 2.0f) + 0.0625f * __sinf(8.0f * x + 1.0f) +
           0.03125f * __sinf(16.0f * x + 5.0f) + 0.015625f * __sinf(32.0f * x) + 0.0078125f * __sinf(64.0f * x);
}
// FP32 operations of the distance estimators as written above, counted one per
// add, subtract, multiply, divide, min, max, floor, sqrt, rsqrt, log and sin and
// two per fused multiply-add (abs and negation are source modifiers; compares
// and selects are not counted): mandelbulb 10 + 65 per iteration executed,
// menger 16 + 33 per level, blob 65, terrain 93. A march step adds 10: the
// point o + d t, its offset from the object's centre, and t += dist.
__device__ __forceinline__ unsigned de_step_flops(int kind, int iters) {
    switch (kind) {
        case PT_SDF_MANDELBULB: return 10u + 10u;
        case PT_SDF_MENGER: return 10u + 16u + 33u * (unsigned)iters;
        case PT_SDF_BLOB: return 10u + 65u;
        default: return 10u + 93u;
    }
}
__device__ __forceinline__ unsigned de_iter_flops(int kind) { return kind == PT_SDF_MANDELBULB ? 65u : 0u; }

// Distance to an SDF of `kind` centred at c with `iters` iterations (levels).
__device__ __forceinline__ float ext_de_at(int kind, f3 c, int iters, float bailout, f3 p, unsigned& it) {
    f3 q = p - c;
    switch (kind) {
        case PT_SDF_MANDELBULB: return ext_mandelbulb(q, iters, bailout, it);
        case PT_SDF_MENGER: return ext_menger(q, iters);
        case PT_SDF_BLOB: return ext_blob(q);
        default: return (q.y - (ext_terrain_h(q.x) + ext_terrain_h(q.z))) * 0.5f;  // gradient bound ~2
    }
}
template <int KIND>
__device__ __forceinline__ float ext_de(const ExtSdf& s, f3 p, unsigned& it) {
    return ext_de_at(KIND ? KIND : s.kind, s.c, s.iters, s.bailout, p, it);
}
// The ray's entry and exit distances of an SDF's box, clipped to (kEps, tmax);
// false when it does not enter before tmax. Every march and the intersect
// kernels' SDF box mask use this one test, so they agree bit for bit.
__device__ __forceinline__ bool sdf_box(const ExtSdf& s, f3 o, f3 d, float tmax, float& tin, float& tout) {
    f3 inv = mk3(hw_rcp(d.x), hw_rcp(d.y), hw_rcp(d.z));
    f3 t0 = (s.lo - o) * inv, t1 = (s.hi - o) * inv;
    f3 a = min3(t0, t1), b = max3(t0, t1);
    tin = fmaxf(fmaxf(fmaxf(a.x, a.y), a.z), kEps);
    tout = fminf(fminf(fminf(b.x, b.y), b.z), tmax);
    return tin < tout;
}
// Bit k: the ray enters the box of an SDF of kind k before tmax (its analytic
// hit). A march bounded by a nearer hit can only enter fewer boxes.
__device__ __forceinline__ uint32_t sdf_mask(const ExtSceneDev& S, f3 o, f3 d, float tmax) {
    uint32_t m = 0;
    for (int i = 0; i < S.nsdf; ++i) {
        float tin, tout;
        if (sdf_box(S.sdf[i], o, d, tmax, tin, tout)) m |= 1u << S.sdf[i].kind;
    }
    return m;
}

// The point at distance t along a ray (fused multiply-adds).
__device__ __forceinline__ f3 ray_at(f3 o, f3 d, float t) {
    return mk3(fmaf(d.x, t, o.x), fmaf(d.y, t, o.y), fmaf(d.z, t, o.z));
}

// Sphere tracing of one SDF object inside its box, up to tmax. Returns the hit
// distance or a negative value. Adds its steps to `steps` and their FP32
// operations (de_step_flops, de_iter_flops) to `flops`.
template <int KIND>
__device__ float ext_march(const ExtSdf& s, f3 o, f3 d, float tmax, int max_steps, unsigned& steps,
                           unsigned& exhausted, unsigned long long& flops) {
    float tin, tout;
    if (!sdf_box(s, o, d, tmax, tin, tout)) return -1.0f;
    float t = tin, res = -1.0f;
    unsigned n = 0, it = 0;
    bool open = true;  // neither a hit nor past the box's exit yet
    while (open && n < (unsigned)max_steps) {
        float dist = ext_de<KIND>(s, ray_at(o, d, t), it);
        ++n;
        if (dist < 1e-4f) {
            res = t;
            open = false;
        } else {
            t += dist;
            open = !(t > tout);
        }
    }
    if (open) ++exhausted;  // the step budget ran out before a hit or the box's exit: a miss
    steps += n;
    const int kind = KIND ? KIND : s.kind;
    flops += (unsigned long long)n * de_step_flops(kind, s.iters) + (unsigned long long)it * de_iter_flops(kind);
    return res;
}
template <int KIND>
__device__ f3 ext_sdf_normal(const ExtSdf& s, f3 p) {
    const float e = 1e-4f;
    unsigned it = 0;
    return normalize(mk3(ext_de<KIND>(s, p + mk3(e, 0, 0), it) - ext_de<KIND>(s, p - mk3(e, 0, 0), it),
                         ext_de<KIND>(s, p + mk3(0, e, 0), it) - ext_de<KIND>(s, p - mk3(0, e, 0), it),
                         ext_de<KIND>(s, p + mk3(0, 0, e), it) - ext_de<KIND>(s, p - mk3(0, 0, e), it)));
}

// ---------------------------------------------------------------------------------
// Hit attributes.
struct Surf {
    f3 p, n;   // position, geometric normal (unit, not face-forwarded)
    int mat, light;
};
__device__ __forceinline__ Surf surface(const ExtSceneDev& S, uint32_t ref, f3 o, f3 d, float t) {
    Surf s;
    s.p = o + d * t;
    uint32_t idx = ref & kRefMask;
    switch (ref >> kRefShift) {
        case REF_SPHERE: {
            float4 c = S.sph[idx];
            s.n = (s.p - mk3(c.x, c.y, c.z)) * (1.0f / c.w);
            int2 ml = S.sph_ml[idx];
            s.mat = ml.x;
            s.light = ml.y;
            break;
        }
        case REF_QUAD: {
            float4 w = S.quad[4 * idx + 1];
            s.n = normalize(mk3(w.x, w.y, w.z));
            int2 ml = S.quad_ml[idx];
            s.mat = ml.x;
            s.light = ml.y;
            break;
        }
        case REF_PLANE: {
            s.n = mk3(0.0f, 1.0f, 0.0f);
            int2 ml = S.plane_ml[idx];
            s.mat = ml.x;
            s.light = ml.y;
            break;
        }
        case REF_LENS: {
            const ExtLens& L = S.lens[idx];
            s.n = lens_normal(L, s.p);
            s.mat = L.mat;
            s.light = L.light;
            break;
        }
        case REF_CYC: {
            const ExtCyc& C = S.cyc[idx];
            s.n = cyclide_normal(C, s.p);
            s.mat = C.mat;
            s.light = C.light;
            break;
        }
        default: {
            const ExtSdf& q = S.sdf[idx];
            s.n = ext_sdf_normal<0>(q, s.p);
            s.mat = q.mat;
            s.light = -1;
            break;
        }
    }
    return s;
}
// True for a hit record that is neither a miss nor a primitive of the scene.
__device__ __forceinline__ bool ref_is_bad(const ExtSceneDev& S, uint32_t ref) {
    if (ref == kMiss) return false;
    const uint32_t kind = ref >> kRefShift;
    const uint32_t limit = kind == REF_SPHERE ? (uint32_t)S.nsph
                         : kind == REF_QUAD   ? (uint32_t)S.nquad
                         : kind == REF_PLANE  ? (uint32_t)S.nplane
                         : kind == REF_SDF    ? (uint32_t)S.nsdf
                         : kind == REF_LENS   ? (uint32_t)S.nlens
                         : kind == REF_CYC    ? (uint32_t)S.ncyc
                                              : 0u;
    return (ref & kRefMask) >= limit;
}
// Counts the lanes whose hit record is bad (pt_counters.bad_refs). Call from
// converged control flow.
__device__ __forceinline__ void count_bad_refs(unsigned long long* stats, bool bad) {
    const uint64_t m = __ballot(bad);
    if (m && __lane_id() == 0) atomicAdd(stat_slot(stats, kStatBadRefs), (unsigned long long)__popcll(m));
}
__device__ __forceinline__ int hit_class(const ExtSceneDev& S, uint32_t ref) {
    if (ref == kMiss) return Q_MISS;
    const uint32_t idx = ref & kRefMask;
    const uint32_t kind = ref >> kRefShift;
    // A reference to no primitive of the scene is classified as a miss, so no
    // later kernel indexes the scene tables with it; the classify passes count
    // such records in pt_counters.bad_refs.
    if (ref_is_bad(S, ref)) return Q_MISS;
    int mat, light;
    switch (kind) {
        case REF_SPHERE: { int2 ml = S.sph_ml[idx]; mat = ml.x; light = ml.y; break; }
        case REF_QUAD: { int2 ml = S.quad_ml[idx]; mat = ml.x; light = ml.y; break; }
        case REF_PLANE: { int2 ml = S.plane_ml[idx]; mat = ml.x; light = ml.y; break; }
        case REF_LENS: mat = S.lens[idx].mat; light = S.lens[idx].light; break;
        case REF_CYC: mat = S.cyc[idx].mat; light = S.cyc[idx].light; break;
        default: mat = S.sdf[idx].mat; light = -1; break;
    }
    if (light >= 0) return Q_EMIT;
    int type = S.mat[mat].type;
    if (type == pt::MAT_DIELECTRIC) return Q_DIELECTRIC;
    if (type == pt::MAT_CONDUCTOR) return Q_CONDUCTOR;
    return Q_LAMBERT;
}

__device__ __forceinline__ void wavelengths(const ExtSceneDev& S, float l0, float l[4]) {
    l[0] = l0;
    float step = S.lrange / (float)S.n_lambda;
    for (int j = 1; j < 4; ++j) {
        float x = l0 - S.lmin + step * (float)j;
        if (x >= S.lrange) x -= S.lrange;
        l[j] = S.lmin + x;
    }
}
// Radiance samples c_j at the path's wavelengths -> XYZ (uniform wavelength pdf).
__device__ __forceinline__ f3 to_xyz(const ExtSceneDev& S, const float l[4], f4 c) {
    float w = S.lrange / (float)S.n_lambda;
    f3 r = cmf(l[0]) * c.x;
    if (S.n_lambda > 1) r = r + cmf(l[1]) * c.y + cmf(l[2]) * c.z + cmf(l[3]) * c.w;
    return r * w;
}
__device__ __forceinline__ f3 xyz(float4 v) { return mk3(v.x, v.y, v.z); }
__device__ __forceinline__ f4 f4of(float4 v) { return f4{v.x, v.y, v.z, v.w}; }
__device__ __forceinline__ f4 ld4(const float4* a, int p) {
    float4 v = a[p];
    return f4{v.x, v.y, v.z, v.w};
}
__device__ __forceinline__ f3 ld3(const float4* a, int p) {
    float4 v = a[p];
    return mk3(v.x, v.y, v.z);
}
__device__ __forceinline__ void add_xyz(ExtWave& Wv, int p, f3 c) {
    float4 a = Wv.acc[p];
    Wv.acc[p] = make_float4(a.x + c.x, a.y + c.y, a.z + c.z, a.w);
}

// ---------------------------------------------------------------------------------
// Kernels.
__device__ __forceinline__ int* qlen(const ExtWave& W, int i) { return W.counts + i * kCntStride; }

// The per-wave sizes of a batched wave, in device memory so a captured graph
// of the wave's kernels replays for any batch.
struct ExtBatchHdr {
    int npaths, npix, nseg, pad;
};
__global__ void k_ext_counts_batch(int* counts, const ExtBatchHdr* hdr) {
    int i = threadIdx.x;
    if (i < N_CNT) counts[i * kCntStride] = (i == 0) ? hdr->npaths : 0;
}
__global__ void k_ext_counts(int* counts, int n0) {
    int i = threadIdx.x;
    if (i < N_CNT) counts[i * kCntStride] = (i == 0) ? n0 : 0;
}
// Start of bounce b: empty the next extension queue, the material queues, the
// shadow queue bounce b appends to (read by bounce b - 1 at the latest), and
// the persistent cursors.
__global__ void k_ext_reset(int* counts, int* cursor, int nxt, int shq) {
    int i = threadIdx.x;
    if (i == nxt || i == shq || (i >= CNT_MAT && i < CNT_MAT + N_MATQ)) counts[i * kCntStride] = 0;
    for (int j = i; j < N_CUR * kSeg; j += blockDim.x) cursor[j * kCntStride] = 0;
}

// BK7 Sellmeier index (the reference camera lens, src/shader.comp:1064-1073).
__device__ __forceinline__ float bk7_index(float lnm) {
    float l = lnm * 1e-3f;
    float l2 = l * l;
    float n2 = 1.0f + (1.03961212f * l2) / (l2 - 6.00069867e-3f) + (0.231792344f * l2) / (l2 - 2.00179144e-2f) +
               (1.01046945f * l2) / (l2 - 1.03560653e2f);
    return sqrtf(n2);
}

// Camera ray of sample k of pixel `pix` of frame F (scene S) into wave slot i,
// request slot rid.
__device__ __forceinline__ void gen_ray(const ExtSceneDev& S, const ExtFrame& F, int pix, int k, ExtWave& Wv, int i,
                                        uint32_t rid, PathRec& rec) {
    int gx = pix % F.W, gy = pix / F.W;
    uint32_t sample = (uint32_t)(F.frame - F.spf + k);
    uint32_t key = path_key(S.seed, (uint32_t)pix, sample);
    float ju = rnd(key, kCamBounce, 0), jv = rnd(key, kCamBounce, 1);
    float uvx = (2.0f * ((float)gx + ju) - (float)F.W) / (float)F.H;
    float uvy = (2.0f * ((float)(F.H - gy) - jv) - (float)F.H) / (float)F.H;
    float l0 = S.lmin + S.lrange * rnd(key, kCamBounce, 4);
    f3 o, d;
    if (S.camera == 1) {
        float h = 0.5f * F.camSize;
        o = F.camPos;
        d = normalize(vmul(mk3(uvx * h, uvy * h, F.apDist), F.camM));
    } else {
        // The reference's camera (sensor, aperture disk, BK7 bi-convex lens) at the hero wavelength.
        float sc = -F.camSize * 0.5f;
        o = F.camPos + vmul(mk3(uvx * sc, uvy * sc, 0.0f), F.camM);
        float ux = rnd(key, kCamBounce, 2), uy = rnd(key, kCamBounce, 3);
        float phi = 2.0f * kPiF * uy, dr = sqrtf(ux);
        float ap = 0.5f * F.apSize;
        f3 onAp = F.camPos + vmul(mk3(ap * dr * __cosf(phi), ap * dr * __sinf(phi), F.apDist), F.camM);
        d = normalize(onAp - o);
        float l = l0;
        for (int pass = 0; pass < 2; ++pass) {
            // the two spherical caps (radius 2f) of the lens
            float best = 1e6f;
            f3 nb = mk3(0.0f, 0.0f, 0.0f);
            int outside = 1;
            for (int sidx = 0; sidx < 2; ++sidx) {
                bool first = sidx == 0;
                float off = F.lensTwoF - F.lensHalf;
                f3 lo = vmul(o - F.lensPos, F.lensM);
                if (first) lo.x += F.lensSlicePos - F.lensHalf - off;
                else lo.x -= F.lensSlicePos - F.lensHalf - off;
                f3 dl = vmul(d, F.lensM);
                float b = 2.0f * dot(dl, lo), c = dot(lo, lo) - F.lensTwoF * F.lensTwoF;
                float disc = b * b - 4.0f * c;
                if (disc < 0.0f) continue;
                float sq = sqrtf(disc);
                float t1 = (-b - sq) * 0.5f, t2 = (-b + sq) * 0.5f;
                if (first) {
                    t1 = (fmaf(dl.x, t1, lo.x) > -off) ? 1e6f : t1;
                    t2 = (fmaf(dl.x, t2, lo.x) > -off) ? 1e6f : t2;
                } else {
                    t1 = (fmaf(dl.x, t1, lo.x) < off) ? 1e6f : t1;
                    t2 = (fmaf(dl.x, t2, lo.x) < off) ? 1e6f : t2;
                }
                float t = (t1 > 0.0f) ? t1 : 1e6f;
                int out = 1;
                if (t2 < t) { t = t2; out = -1; }
                if (t < 1e-4f || !(t < best)) continue;
                best = t;
                nb = mvmul(F.lensM, normalize(fma3s(dl, t, lo) * (float)out));
                outside = out;
            }
            float n1 = 1.0f, n2 = 1.0f;
            if (outside == 1) n2 = bk7_index(l);
            else n1 = bk7_index(l);
            float n12 = n1 / n2;
            l = l * n12;
            o = fma3s(d, best, o);
            d = refract(d, nb, n12);
        }
    }
    // the camera ray in queue 0's records for the traversal, and in the path's
    // record (stored by the caller) for the shading
    Wv.rq_org[0][i] = make_float4(o.x, o.y, o.z, l0);
    Wv.rq_dir[0][i] = make_float4(d.x, d.y, d.z, 0.0f);
    Wv.acc[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    rec.org = make_float4(o.x, o.y, o.z, l0);
    rec.dir = make_float4(d.x, d.y, d.z, 0.0f);
    rec.T = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    rec.aux = make_uint4(key, 2u, 0u, rid);  // the camera ray counts as specular for emitter MIS
}

// Stores the records of the calling wave's active lanes, which hold
// consecutive paths from the first lane's i on (a prefix of the wave), as
// contiguous 16-byte vectors through the wave's LDS slice (64 records, padded
// to 80 bytes a lane): stored lane by lane, a 64-byte record per lane is four
// stores 64 bytes apart, which raygen wrote at 4.4 TB/s against 5.1 TB/s
// for contiguous ones. Every lane of the wave calls it.
constexpr int kRecPad = 5;  // float4 per lane in the slice
__device__ __forceinline__ void store_recs(PathRec* pr, int i, bool active, const PathRec& r, float4* slice) {
    const int nact = (int)__popcll(__ballot(active));
    if (nact == 0) return;
    const int lane = (int)__lane_id();
    const int i0 = __shfl(i, 0);
    float4* mine = slice + lane * kRecPad;
    mine[0] = r.org;
    mine[1] = r.dir;
    mine[2] = r.T;
    mine[3] = make_float4(__uint_as_float(r.aux.x), __uint_as_float(r.aux.y), __uint_as_float(r.aux.z),
                          __uint_as_float(r.aux.w));
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    float4* dst = reinterpret_cast<float4*>(pr + i0);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int v = lane + 64 * k;  // vector v of the wave's records: record v / 4, field v % 4
        if (v < 4 * nact) dst[v] = slice[(v >> 2) * kRecPad + (v & 3)];
    }
    // every lane has read the slice before it is written again
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(kBlk) k_ext_raygen(ExtSceneDev S, ExtFrame F, ExtWave Wv) {
    __shared__ float4 s_rec[kBlk * kRecPad];
    const int n = F.pc * F.m;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[3], (unsigned long long)n);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {  // uniform rounds: every lane reaches store_recs
        const int i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        PathRec rec{};
        if (i < n) {
            gen_ray(S, F, F.p0 + i % F.pc, F.k0 + i / F.pc, Wv, i, 0u, rec);
            Wv.q[0][i] = i;
        }
        store_recs(Wv.pr, i, i < n, rec, s_rec + (threadIdx.x >> 6) * 64 * kRecPad);
    }
}

// One segment of a batched wave: request slot rid, its pixels [p0, p0 + pc) x
// samples [k0, k0 + m) of its frame, at wave slots [base, base + pc * m)
// (sample-major), and pix_base = the sum of pc over the earlier segments.
struct ExtSeg {
    int rid, base, p0, pc, k0, m, pix_base, pad;
};
// Index of the segment holding `i` of the ascending starts start(0..n-1).
template <typename Start>
__device__ __forceinline__ int seg_find(int n, int i, Start start) {
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (start(mid) <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ void __launch_bounds__(kBlk) k_ext_raygen_batch(ExtBatch B, const ExtSeg* segs, const ExtBatchHdr* hdr,
                                                            ExtWave Wv) {
    __shared__ float4 s_rec[kBlk * kRecPad];
    const int npaths = hdr->npaths, nseg = hdr->nseg;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[3], (unsigned long long)npaths);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (npaths + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {  // uniform rounds: every lane reaches store_recs
        const int i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        PathRec rec{};
        if (i < npaths) {
            const ExtSeg& g = segs[seg_find(nseg, i, [&](int k) { return segs[k].base; })];
            const int local = i - g.base;
            gen_ray(B.scenes[g.rid], B.frames[g.rid], g.p0 + local % g.pc, g.k0 + local / g.pc, Wv, i,
                    (uint32_t)g.rid, rec);
            Wv.q[0][i] = i;
        }
        store_recs(Wv.pr, i, i < npaths, rec, s_rec + (threadIdx.x >> 6) * 64 * kRecPad);
    }
}

// Appends the path to its material queue; every thread of the block calls it.
__device__ __forceinline__ void classify_append(const ExtSceneDev& S, ExtWave& Wv, bool active, int p, float t,
                                                uint32_t ref) {
    const int cls = active ? hit_class(S, ref) : -1;
    count_bad_refs(Wv.stats, active && ref_is_bad(S, ref));
    int* const ctr[N_MATQ] = {qlen(Wv, CNT_MAT + 0), qlen(Wv, CNT_MAT + 1), qlen(Wv, CNT_MAT + 2),
                              qlen(Wv, CNT_MAT + 3), qlen(Wv, CNT_MAT + 4)};
    int slot[N_MATQ];
    block_append<kBlk, N_MATQ>(ctr, cls >= 0 ? 1u << cls : 0u, slot);
    if (cls >= 0) Wv.mq[cls][slot[cls]] = MatEntry{p, t, ref, 0};
}

// Closest hit over parallelograms, spheres and planes; with CLASSIFY, the path
// is also appended to its material queue (scenes without SDFs).
template <int W, int STACK, typename SE, bool CLASSIFY, bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_intersect(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur) {
    __shared__ SE stack_mem[STACK * kBlk];
    __shared__ float4 s_wq[kBlk / 64][4 * kQuadBrute];
    SE* stack = stack_mem + threadIdx.x;
    const int n = *qlen(Wv, cur);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[0], (unsigned long long)n);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    unsigned visits = 0;
    for (int r = 0; r < rounds; ++r) {
        int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool active = idx < n;
        // the path index: its request slot, its material-queue entry
        int p = active && (MULTI || CLASSIFY) ? Wv.q[cur][idx] : 0;
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI && active ? Wv.pr[p].aux.w : 0u);
        const bool staged = wave_quads(S.quad, S.quad_bvh ? 0 : S.nquad, active, s_wq[threadIdx.x >> 6]);
        uint32_t ref = kMiss;
        float t = 1e30f;
        if (active) {
            f3 o, d;
            ray_of(Wv, cur, idx, o, d);
            scene_trace<W, STACK, SE, false>(S, staged, s_wq[threadIdx.x >> 6], o, d, t, ref, stack, visits);
            Wv.hit[idx] = make_float2(t, __uint_as_float(ref));
            if (S.nsdf > 0) Wv.smask[idx] = sdf_mask(S, o, d, t);
        }
        if (CLASSIFY) classify_append(S, Wv, active, p, t, ref);
    }
    wave_add_u64(stat_slot(Wv.stats, 11), visits);
}

// ---------------------------------------------------------------------------------
// Persistent traversal with lane refill. Each wave loops over the queue until it
// is drained: lanes whose ray has finished take the next queue entries (once at
// least kRefill lanes are idle, or none is busy), and every busy lane visits one
// BVH node per iteration. The wave's lanes stay busy whatever the spread of
// traversal lengths, which a one-ray-per-lane kernel loses to its longest ray.
// Per ray, the nodes visited and the primitive tests are those of bvh_trace, so
// hits are identical. The queue is split into min(kSeg, waves in the grid)
// segments, each with its own cursor on its own line; a wave takes `chunk` entries per atomic from its home
// segment (chunk = n / waves in the grid, clamped to [1, kChunkMax]) and hands
// them to its lanes from registers. Consecutive waves have different home
// segments, so the blocks of one XCD (blockIdx.x mod 8) share a few segments.
// Half the wave: on the 4096-sphere scene 32 measured 4.36 G segments/s, 16
// and 48 4.30 G, 8 4.15 G, 4 3.96 G (a refill's ray loads and parallelogram
// tests run at the width of the lanes it fills); the Cornell box and the SDF
// scene measured the same for 16, 32 and 48.
constexpr int kRefill = 32;

// The scene's parallelograms staged in LDS by the block (at most kQuadBrute:
// the host launches the persistent kernels only for scenes without a
// parallelogram BVH). Each refill tests them all; from LDS their loads are
// broadcasts, where from global memory each parallelogram was a chain of
// dependent vector loads on the refill's critical path.
__device__ __forceinline__ int stage_quads(const ExtSceneDev& S, float4* s_quad) {
    const int nq = S.nquad < kQuadBrute ? S.nquad : kQuadBrute;
    for (int i = (int)threadIdx.x; i < 4 * nq; i += kBlk) s_quad[i] = S.quad[i];
    __syncthreads();
    return nq;
}

// Persistent closest-hit kernel (parallelograms tested directly, spheres through
// their BVH, then planes); see the persistent traversal notes above. Writes the
// hits only: k_ext_classify appends the paths to the material queues (one
// atomic per wave and class for 64 paths, where appending as lanes finish
// would take one per finishing lane).
template <int W, int STACK, typename SE>
__global__ void __launch_bounds__(kBlk) k_ext_intersect_pt(ExtSceneDev S, ExtWave Wv, int cur) {
    __shared__ SE stack_mem[STACK * kBlk];
    __shared__ float4 s_quad[4 * kQuadBrute];
    SE* stack = stack_mem + threadIdx.x;
    const int nq = stage_quads(S, s_quad);
    const int n = *qlen(Wv, cur);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[0], (unsigned long long)n);
    int p = -1, sp = 0, node = -1, best = -1;  // p: the lane's queue entry (-1 idle)
    unsigned visits = 0;
    f3 o = mk3(0.0f, 0.0f, 0.0f), d = o, inv = o, oi = o;
    float t = 0.0f;
    uint32_t ref = kMiss;
    WaveChunk chunk = wave_chunk_init(Wv.cursor + CUR_ISECT * kSeg * kCntStride, n);
    while (true) {
        const uint64_t idle = __ballot(p < 0);
        if (!chunk.drained() && (__popcll(idle) >= kRefill || ~idle == 0ull)) {
            int e = wave_refill(idle, chunk);
            if (e >= 0) {
                p = e;
                ray_of(Wv, cur, e, o, d);
                t = 1e30f;
                ref = kMiss;
                best = -1;
                for (int i = 0; i < nq; ++i) {
                    float tq;
                    if (isect_quad(s_quad + 4 * i, o, d, t, tq)) {
                        t = tq;
                        ref = (REF_QUAD << kRefShift) | (uint32_t)i;
                    }
                }
                ray_inv(o, d, inv, oi);
                sp = 0;
                node = S.nsph > 0 ? 0 : -1;
            }
        }
        if (__ballot(p >= 0) == 0ull) {
            if (chunk.drained()) break;
            continue;
        }
        if (p >= 0 && node >= 0) {
            ++visits;
            bvh_step<W, SE, false, false>(S.sbvh, o, d, inv, oi, t, best, node, sp, stack);
        }
        if (p >= 0 && node < 0) {
            if (best >= 0) ref = (REF_SPHERE << kRefShift) | (uint32_t)S.sbvh.ids[best];
            for (int i = 0; i < S.nplane; ++i) {
                float tp = (ld_ro(S.plane_y + i) - o.y) / d.y;
                if (tp > kEps && tp < t) {
                    t = tp;
                    ref = (REF_PLANE << kRefShift) | (uint32_t)i;
                }
            }
            Wv.hit[p] = make_float2(t, __uint_as_float(ref));
            if (S.nsdf > 0) Wv.smask[p] = sdf_mask(S, o, d, t);
            p = -1;
        }
    }
    wave_add_u64(stat_slot(Wv.stats, 11), visits);
}

// Extension rays of queue `cur` (closest hit, then the material queues) and,
// in the same launch, the shadow rays of the previous bounce (queue length at
// counter `shq`; any hit, then the unoccluded contribution). Both only read the
// scene and write different path fields, and a path's shadow contribution is
// still added before the next bounce's shading, so the sums are unchanged. A
// bounce launch lasts as long as its slowest ray's chain of dependent node
// fetches; one launch instead of two pays that once per bounce. Scenes without
// SDF objects only (their shadow rays also march the SDFs).
template <int W, int STACK, typename SE, bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_trace(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur, int shq) {
    __shared__ SE stack_mem[STACK * kBlk];
    __shared__ float4 s_wq[kBlk / 64][4 * kQuadBrute];
    SE* stack = stack_mem + threadIdx.x;
    const int ne = *qlen(Wv, cur), ns = *qlen(Wv, shq);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        atomicAdd(&Wv.stats[0], (unsigned long long)ne);
        atomicAdd(&Wv.stats[1], (unsigned long long)ns);
    }
    const int n = ne + ns;
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    unsigned ev = 0, sv = 0;  // nodes visited by extension and shadow rays
    for (int r = 0; r < rounds; ++r) {
        const int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool ext = idx < ne, shadow = !ext && idx < n;
        int p = 0;
        uint32_t ref = kMiss, rid = 0u, srid = 0u;
        float t = 1e30f;
        float4 so = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sd = so, sx = so;
        if (ext) {
            p = Wv.q[cur][idx];
            if (MULTI) rid = Wv.pr[p].aux.w;
        } else if (shadow) {
            const int i = idx - ne;
            so = Wv.s_org[i];
            sd = Wv.s_dir[i];
            sx = Wv.s_xyz[i];
            if (MULTI) srid = __float_as_uint(sx.w);
        }
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, ext ? rid : srid);
        const bool staged = wave_quads(S.quad, S.quad_bvh ? 0 : S.nquad, ext || shadow, s_wq[threadIdx.x >> 6]);
        if (ext) {
            f3 o, d;
            ray_of(Wv, cur, idx, o, d);
            scene_trace<W, STACK, SE, false>(S, staged, s_wq[threadIdx.x >> 6], o, d, t, ref, stack, ev);
            Wv.hit[idx] = make_float2(t, __uint_as_float(ref));
        } else if (shadow) {
            f3 o = mk3(so.x, so.y, so.z), d = mk3(sd.x, sd.y, sd.z);
            float ts = so.w;
            uint32_t sref = kMiss;
            if (!scene_trace<W, STACK, SE, true>(S, staged, s_wq[threadIdx.x >> 6], o, d, ts, sref, stack, sv))
                add_xyz(Wv, __float_as_int(sd.w), mk3(sx.x, sx.y, sx.z));
        }
        classify_append(scene_of<MULTI>(S0, B, rid), Wv, ext, p, t, ref);
    }
    wave_add_u64(stat_slot(Wv.stats, 11), ev);
    wave_add_u64(stat_slot(Wv.stats, 12), sv);
}

// Sphere tracing of every SDF object of one kind against the current hit.
template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_march(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur) {
    const int n = *qlen(Wv, cur);
    unsigned steps = 0, exhausted = 0;
    unsigned long long flops = 0;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[Wv.q[cur][idx]].aux.w : 0u);
        f3 o, d;
        ray_of(Wv, cur, idx, o, d);
        float2 h = Wv.hit[idx];
        float t = h.x;
        uint32_t ref = __float_as_uint(h.y);
        bool upd = false;
        for (int s = 0; s < S.nsdf; ++s) {
            if (S.sdf[s].kind != KIND) continue;
            float th = ext_march<KIND>(S.sdf[s], o, d, t, S.sdf_steps_max, steps, exhausted, flops);
            if (th > 0.0f && th < t) {
                t = th;
                ref = (REF_SDF << kRefShift) | (uint32_t)s;
                upd = true;
            }
        }
        if (upd) Wv.hit[idx] = make_float2(t, __uint_as_float(ref));
    }
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
    wave_add_u64(stat_slot(Wv.stats, 13), exhausted);
    wave_add_u64(stat_slot(Wv.stats, kStatMarchFlops + KIND - 1), flops);
}

// Persistent sphere tracing with lane refill (the default march). A ray's march
// lasts from one step to the whole step budget, so a one-ray-per-lane wave runs
// at the pace of its longest march with most lanes idle. Here each wave loops
// over a segmented queue (WaveChunk, as the persistent traversal): every busy
// lane takes one distance-estimator step per iteration, and lanes whose march
// has ended take the next queue entries while at least kMarchRefill lanes are
// idle. Per ray, the SDFs of this kind are marched in scene order with the same
// box clip, steps and comparisons as ext_march, each bounded by the nearest hit
// so far, so hits, step counts and operation counts equal k_ext_march's.
constexpr int kMarchRefill = 8;

template <int KIND, bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_march_pt(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur) {
    const int n = *qlen(Wv, cur);
    unsigned steps = 0, exhausted = 0, it = 0;
    unsigned long long flops = 0;
    // Lane state: queue entry p (-1 idle), its ray, nearest hit (t, ref), index s of
    // the SDF being marched, and that march's position tm, exit tout, step
    // count and SDF parameters. `seek`: start the next SDF's march.
    int p = -1, s = -1, nsdf = 0, smax = 0, nst = 0;
    bool upd = false, seek = false;
    f3 o = mk3(0.0f, 0.0f, 0.0f), d = o, sc = o;
    float t = 0.0f, tm = 0.0f, tout = 0.0f, sbail = 0.0f;
    int siters = 0;
    uint32_t ref = kMiss, sflops = 0;
    const ExtSdf* sdfs = nullptr;
    WaveChunk chunk = wave_chunk_init(Wv.cursor + (CUR_MARCH + KIND - 1) * kSeg * kCntStride, n);
    while (true) {
        const uint64_t idle = __ballot(p < 0);
        if (!chunk.drained() && __popcll(idle) >= kMarchRefill) {
            const int e = wave_refill(idle, chunk);
            // rays that enter no box of this kind (the intersect kernel's mask)
            // go back to idle without loading their ray
            if (e >= 0 && ((Wv.smask[e] >> KIND) & 1u)) {
                p = e;
                const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[Wv.q[cur][e]].aux.w : 0u);
                sdfs = S.sdf;
                nsdf = S.nsdf;
                smax = S.sdf_steps_max;
                ray_of(Wv, cur, e, o, d);
                const float2 h = Wv.hit[p];
                t = h.x;
                ref = __float_as_uint(h.y);
                upd = false;
                s = -1;
                seek = true;
            }
        }
        // The next SDF of this kind whose box the ray enters before its nearest
        // hit; with none left, the hit is stored and the lane goes idle.
        if (seek) {
            seek = false;
            int ns = s + 1;
            bool found = false;
            float tin = 0.0f, tx = 0.0f;
            for (; ns < nsdf; ++ns) {
                const ExtSdf& e = sdfs[ns];
                if (e.kind != KIND) continue;
                if (sdf_box(e, o, d, t, tin, tx)) {
                    found = true;
                    break;
                }
            }
            s = ns;
            if (found) {
                const ExtSdf& e = sdfs[ns];
                tm = tin;
                tout = tx;
                nst = 0;
                sc = e.c;
                siters = e.iters;
                sbail = e.bailout;
                sflops = de_step_flops(KIND, e.iters);
            } else {
                if (upd) Wv.hit[p] = make_float2(t, __uint_as_float(ref));
                p = -1;
            }
        }
        const uint64_t busy = __ballot(p >= 0);
        if (busy == 0ull) {
            if (chunk.drained()) break;
            continue;
        }
        // fill the wave before stepping while the queue has entries
        if (!chunk.drained() && __popcll(~busy) >= kMarchRefill) continue;
        if (p >= 0) {
            if (nst < smax) {
                const float dist = ext_de_at(KIND, sc, siters, sbail, ray_at(o, d, tm), it);
                ++nst;
                ++steps;
                flops += sflops;
                if (dist < 1e-4f) {
                    if (tm > 0.0f && tm < t) {
                        t = tm;
                        ref = (REF_SDF << kRefShift) | (uint32_t)s;
                        upd = true;
                    }
                    seek = true;
                } else {
                    tm += dist;
                    seek = tm > tout;
                }
            } else {
                ++exhausted;  // the step budget ran out before a hit or the box's exit: a miss
                seek = true;
            }
        }
    }
    flops += (unsigned long long)it * de_iter_flops(KIND);
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
    wave_add_u64(stat_slot(Wv.stats, 13), exhausted);
    wave_add_u64(stat_slot(Wv.stats, kStatMarchFlops + KIND - 1), flops);
}

// Material queues from the hits of queue `cur`, in two passes over the block's
// entries: the first counts each class, so the block reserves its slots with one
// device atomic per class for all its rounds (atomics on one address serialise
// across the chip, and per-round reservations of a 2^25-path wave cost about
// 0.4 ms a launch); the second places the entries round by round from
// block-local running offsets.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_classify(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur) {
    constexpr int NW = kBlk / 64;
    __shared__ int s_base[N_MATQ];
    __shared__ int s_n[N_MATQ][NW];
    const int n = *qlen(Wv, cur);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    const int lane = (int)__lane_id(), wv = (int)threadIdx.x >> 6;
    if ((int)threadIdx.x < N_MATQ) s_base[threadIdx.x] = 0;
    __syncthreads();
    int cnt[N_MATQ] = {};
    int bad = 0;
    for (int r = 0; r < rounds; ++r) {
        const int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        if (idx < n) {
            const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[Wv.q[cur][idx]].aux.w : 0u);
            const uint32_t ref = __float_as_uint(Wv.hit[idx].y);
            const int cls = hit_class(S, ref);
            bad += ref_is_bad(S, ref) ? 1 : 0;
#pragma unroll
            for (int q = 0; q < N_MATQ; ++q) cnt[q] += cls == q ? 1 : 0;
        }
    }
    if (bad) atomicAdd(stat_slot(Wv.stats, kStatBadRefs), (unsigned long long)bad);
#pragma unroll
    for (int q = 0; q < N_MATQ; ++q) {
        int v = cnt[q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off);
        if (lane == 0 && v) atomicAdd(&s_base[q], v);
    }
    __syncthreads();
    if ((int)threadIdx.x < N_MATQ) {
        const int q = (int)threadIdx.x;
        const int tot = s_base[q];
        s_base[q] = tot ? atomicAdd(qlen(Wv, CNT_MAT + q), tot) : 0;
    }
    __syncthreads();
    for (int r = 0; r < rounds; ++r) {
        const int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        int p = 0, cls = -1;
        float2 h = make_float2(0.0f, 0.0f);
        if (idx < n) {
            p = Wv.q[cur][idx];
            const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[p].aux.w : 0u);
            h = Wv.hit[idx];
            cls = hit_class(S, __float_as_uint(h.y));
        }
        uint64_t m[N_MATQ];
#pragma unroll
        for (int q = 0; q < N_MATQ; ++q) {
            m[q] = __ballot(cls == q);
            if (lane == 0) s_n[q][wv] = (int)__popcll(m[q]);
        }
        __syncthreads();
        if (cls >= 0) {
            int off = s_base[cls];
            for (int w = 0; w < wv; ++w) off += s_n[cls][w];
            Wv.mq[cls][off + (int)lanes_below(m[cls])] = MatEntry{p, h.x, __float_as_uint(h.y), 0};
        }
        __syncthreads();  // every thread has read s_base before it advances
        if ((int)threadIdx.x < N_MATQ) {
            const int q = (int)threadIdx.x;
            int tot = 0;
#pragma unroll
            for (int w = 0; w < NW; ++w) tot += s_n[q][w];
            s_base[q] += tot;
        }
        __syncthreads();  // s_n and s_base are rewritten by the next round
    }
}

// First-hit features of the camera paths (queue 0 right after bounce 0's
// trace): an albedo triple (reflectance at 450 / 550 / 650 nm; 1 for glass and
// emitters; the normal-incidence reflectance of a conductor), the shading
// normal facing the camera and the hit distance; zeros for a miss.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_aov(ExtSceneDev S0, ExtBatch B, ExtWave Wv) {
    const int n = *qlen(Wv, 0);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const int p = Wv.q[0][idx];
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[p].aux.w : 0u);
        const float2 h = Wv.hit[idx];
        const uint32_t ref = __float_as_uint(h.y);
        float4 a0 = make_float4(0.0f, 0.0f, 0.0f, 0.0f), a1 = a0;
        if (hit_class(S, ref) != Q_MISS) {
            f3 o = xyz(Wv.rq_org[0][p]), d = xyz(Wv.rq_dir[0][p]);  // bounce 0: the camera ray
            Surf sf = surface(S, ref, o, d, h.x);
            f3 N = dot(sf.n, d) < 0.0f ? sf.n : -sf.n;
            f3 al = mk3(1.0f, 1.0f, 1.0f);
            const ExtMat& M = S.mat[sf.mat];
            if (sf.light < 0) {
                if (M.type == pt::MAT_CONDUCTOR) {
                    float f0 = fresnel_conductor(1.0f, M.p[0], M.p[1]);
                    al = mk3(f0, f0, f0);
                } else if (M.type != pt::MAT_DIELECTRIC) {
                    al = mk3(reflectance(M, 450.0f), reflectance(M, 550.0f), reflectance(M, 650.0f));
                }
            }
            a0 = make_float4(al.x, al.y, al.z, h.x);
            a1 = make_float4(N.x, N.y, N.z, 0.0f);
        }
        Wv.aov0[p] = a0;
        Wv.aov1[p] = a1;
    }
}

// Features sink: per pixel, in sample order, sums of the per-sample XYZ, the
// first-hit features and Y^2 into feat[pixel * 12 + (XYZ 0..2, albedo 3..5,
// normal 6..8, distance 9, Y^2 10)]; samples with a NaN are dropped and counted.
__global__ void __launch_bounds__(kBlk) k_ext_features(ExtFrame F, ExtWave Wv, float* feat, int* nsamp) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nans = 0;
    if (j < F.pc) {
        float* f = feat + (size_t)(F.p0 + j) * 12;
        float s[11];
        for (int c = 0; c < 11; ++c) s[c] = f[c];
        int n = nsamp[F.p0 + j];
        for (int k = 0; k < F.m; ++k) {
            const int p = k * F.pc + j;
            float4 a = Wv.acc[p];
            if (a.x != a.x || a.y != a.y || a.z != a.z) { ++nans; continue; }
            float4 b0 = Wv.aov0[p], b1 = Wv.aov1[p];
            s[0] += a.x; s[1] += a.y; s[2] += a.z;
            s[3] += b0.x; s[4] += b0.y; s[5] += b0.z;
            s[6] += b1.x; s[7] += b1.y; s[8] += b1.z;
            s[9] += b0.w;
            s[10] += a.y * a.y;
            ++n;
        }
        for (int c = 0; c < 11; ++c) f[c] = s[c];
        nsamp[F.p0 + j] = n;
    }
    wave_add_u64(stat_slot(Wv.stats, 4), nans);
}

// Russian roulette from depth rr_depth; returns false when the path ends.
__device__ __forceinline__ bool roulette(const ExtSceneDev& S, uint32_t key, int bounce, f4& T) {
    if (bounce + 1 < S.rr_depth) return true;
    float m = fmaxf(fmaxf(T.x, T.y), fmaxf(T.z, T.w));
    float q = fminf(m, 0.95f);
    if (!(rnd(key, (uint32_t)bounce, 6) < q)) return false;
    T = T * (1.0f / q);
    return true;
}

// Escaped paths (environment) and emitter hits (weighted by MIS against NEE).
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_shade_terminal(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int which) {
    const int n = *qlen(Wv, CNT_MAT + which);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[6 + which], (unsigned long long)n);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const MatEntry me = Wv.mq[which][idx];
        const int p = me.p;
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[p].aux.w : 0u);
        f4 T = f4of(Wv.pr[p].T);
        float4 org, dir;  // org.w: hero wavelength
        path_ray(Wv, p, org, dir);
        float l[4];
        wavelengths(S, org.w, l);
        if (which == Q_MISS) {
            if (S.has_env) add_xyz(Wv, p, to_xyz(S, l, T * S.env));
            continue;
        }
        f3 o = xyz(org), d = xyz(dir);
        Surf sf = surface(S, me.ref, o, d, me.t);
        const ExtLight& L = S.light[sf.light];
        uint4 aux = Wv.pr[p].aux;
        float w = 1.0f;
        // emitters that NEE does not sample (lenses, cyclides) count with weight 1
        if (!(aux.y & 2u) && sf.light < S.nlight_nee) {
            // pdf with which next-event estimation from the previous vertex picks this point
            float pl;
            if (L.type == REF_QUAD) {
                float cl = fabsf(dot(L.n, d));
                pl = (me.t * me.t) / fmaxf(cl * L.area, 1e-30f);
            } else {
                f3 dc = L.p - o;
                float sin2 = (L.radius * L.radius) / dot(dc, dc);
                float cmax = sqrtf(fmaxf(1.0f - sin2, 0.0f));
                pl = 1.0f / (2.0f * kPiF * fmaxf(1.0f - cmax, 1e-12f));
            }
            pl *= 1.0f / (float)S.nlight_nee;
            float pb = __uint_as_float(aux.z);
            w = pb * pb / (pb * pb + pl * pl);
        }
        f4 Le = f4{planck_norm(l[0], L.T, L.lum), 0.0f, 0.0f, 0.0f};
        if (S.n_lambda > 1)
            Le = f4{Le.x, planck_norm(l[1], L.T, L.lum), planck_norm(l[2], L.T, L.lum), planck_norm(l[3], L.T, L.lum)};
        add_xyz(Wv, p, to_xyz(S, l, T * Le * w));
    }
}

// Lambert vertices: next-event estimation (one shadow ray) + cosine sampling.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_shade_lambert(ExtSceneDev S0, ExtFrame F0, ExtBatch B, ExtWave Wv,
                                                            int cur, int bounce) {
    const int nxt = cur ^ 1;
    const int n = *qlen(Wv, CNT_MAT + Q_LAMBERT);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[6 + Q_LAMBERT], (unsigned long long)n);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool active = idx < n;
        int p = 0;
        bool cont = false, shadow = false;
        float4 ro = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rd = ro;  // the next ray, for its queue entry
        f3 sP = mk3(0, 0, 0), sD = sP, sC = sP;
        float smax = 0.0f;
        uint32_t rid = 0u;
        if (active) {
            const MatEntry me = Wv.mq[Q_LAMBERT][idx];
            p = me.p;
            uint4 aux = Wv.pr[p].aux;
            rid = aux.w;
            const ExtSceneDev& S = scene_of<MULTI>(S0, B, rid);
            const int maxDepth = MULTI ? B.frames[rid].maxDepth : F0.maxDepth;
            float4 org, dir;  // org.w: hero wavelength
            path_ray(Wv, p, org, dir);
            f3 o = xyz(org), d = xyz(dir);
            Surf sf = surface(S, me.ref, o, d, me.t);
            f3 N = dot(sf.n, d) < 0.0f ? sf.n : -sf.n;
            const ExtMat& M = S.mat[sf.mat];
            float l[4];
            wavelengths(S, org.w, l);
            const bool multi = S.n_lambda > 1;
            f4 rho = f4{reflectance(M, l[0]), 0.0f, 0.0f, 0.0f};
            if (multi) rho = f4{rho.x, reflectance(M, l[1]), reflectance(M, l[2]), reflectance(M, l[3])};
            f4 T = f4of(Wv.pr[p].T);
            uint32_t key = aux.x;
            f3 P = sf.p + N * offset_scale(sf.p);
            if (S.nlight_nee > 0) {
                float ul = rnd(key, (uint32_t)bounce, 3);
                int li = min((int)(ul * (float)S.nlight_nee), S.nlight_nee - 1);
                const ExtLight& L = S.light[li];
                float u4 = rnd(key, (uint32_t)bounce, 4), u5 = rnd(key, (uint32_t)bounce, 5);
                f3 wi;
                float dist, pl;
                bool ok = true;
                if (L.type == REF_QUAD) {
                    f3 y = L.p + L.e1 * u4 + L.e2 * u5;
                    f3 v = y - P;
                    float d2 = dot(v, v);
                    dist = sqrtf(d2);
                    wi = v * (1.0f / dist);
                    float cl = fabsf(dot(L.n, wi));
                    ok = cl > 1e-7f;
                    pl = d2 / fmaxf(cl * L.area, 1e-30f);
                } else {
                    f3 dc = L.p - P;
                    float dc2 = dot(dc, dc);
                    float sin2 = (L.radius * L.radius) / dc2;
                    ok = sin2 < 1.0f;
                    float cmax = sqrtf(fmaxf(1.0f - sin2, 0.0f));
                    float ct = 1.0f - u4 * (1.0f - cmax);
                    float st = sqrtf(fmaxf(1.0f - ct * ct, 0.0f));
                    float phi = 2.0f * kPiF * u5;
                    f3 ax = dc * rsqrtf(dc2), tb, bb;
                    onb(ax, tb, bb);
                    wi = tb * (st * __cosf(phi)) + bb * (st * __sinf(phi)) + ax * ct;
                    float th = 0.0f;
                    ok = ok && isect_sphere(make_float4(L.p.x, L.p.y, L.p.z, L.radius), P, wi, 1e30f, th);
                    dist = th;
                    pl = 1.0f / (2.0f * kPiF * fmaxf(1.0f - cmax, 1e-12f));
                }
                pl *= 1.0f / (float)S.nlight_nee;
                float cs = dot(N, wi);
                if (ok && cs > 0.0f) {
                    float pb = cs * (1.0f / kPiF);
                    float w = pl * pl / (pl * pl + pb * pb);
                    f4 Le = f4{planck_norm(l[0], L.T, L.lum), 0.0f, 0.0f, 0.0f};
                    if (multi)
                        Le = f4{Le.x, planck_norm(l[1], L.T, L.lum), planck_norm(l[2], L.T, L.lum),
                                planck_norm(l[3], L.T, L.lum)};
                    float g = cs * (1.0f / kPiF) / pl * w;
                    sC = to_xyz(S, l, T * rho * Le * g);
                    sP = P;
                    sD = wi;
                    smax = dist * (1.0f - 1e-4f);
                    shadow = true;
                }
            }
            // cosine-weighted hemisphere (Malley's method)
            float u0 = rnd(key, (uint32_t)bounce, 0), u1 = rnd(key, (uint32_t)bounce, 1);
            float rr = sqrtf(u0), phi = 2.0f * kPiF * u1;
            float z = sqrtf(fmaxf(1.0f - u0, 0.0f));
            f3 tb, bb;
            onb(N, tb, bb);
            f3 wo = tb * (rr * __cosf(phi)) + bb * (rr * __sinf(phi)) + N * z;
            T = T * rho;
            if (bounce + 1 < maxDepth && z > 0.0f && roulette(S, key, bounce, T)) {
                cont = true;
                ro = make_float4(P.x, P.y, P.z, org.w);
                rd = make_float4(wo.x, wo.y, wo.z, 0.0f);
                Wv.pr[p].org = ro;
                Wv.pr[p].dir = rd;
                Wv.pr[p].T = make_float4(T.x, T.y, T.z, T.w);
                Wv.pr[p].aux = make_uint4(key, aux.y & ~2u, __float_as_uint(z * (1.0f / kPiF)), aux.w);
            }
        }
        int* const ctr[2] = {qlen(Wv, shadow_cnt(bounce)), qlen(Wv, nxt)};
        int slot[2];
        block_append<kBlk, 2>(ctr, (shadow ? 1u : 0u) | (cont ? 2u : 0u), slot);
        if (shadow) {
            Wv.s_org[slot[0]] = make_float4(sP.x, sP.y, sP.z, smax);
            Wv.s_dir[slot[0]] = make_float4(sD.x, sD.y, sD.z, __int_as_float(p));
            Wv.s_xyz[slot[0]] = make_float4(sC.x, sC.y, sC.z, __uint_as_float(rid));
        }
        if (cont) {
            Wv.q[nxt][slot[1]] = p;
            Wv.rq_org[nxt][slot[1]] = ro;
            Wv.rq_dir[nxt][slot[1]] = rd;
        }
    }
}

// Smooth dielectric: Fresnel-weighted choice of reflection or refraction at the
// hero wavelength; a dispersive interface keeps only the hero.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_shade_dielectric(ExtSceneDev S0, ExtFrame F0, ExtBatch B, ExtWave Wv,
                                                               int cur, int bounce) {
    const int nxt = cur ^ 1;
    const int n = *qlen(Wv, CNT_MAT + Q_DIELECTRIC);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[6 + Q_DIELECTRIC], (unsigned long long)n);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool active = idx < n;
        int p = 0;
        bool cont = false;
        float4 ro = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rd = ro;  // the next ray, for its queue entry
        if (active) {
            const MatEntry me = Wv.mq[Q_DIELECTRIC][idx];
            p = me.p;
            uint4 aux = Wv.pr[p].aux;
            const ExtSceneDev& S = scene_of<MULTI>(S0, B, aux.w);
            const int maxDepth = MULTI ? B.frames[aux.w].maxDepth : F0.maxDepth;
            float4 org, dir;  // org.w: hero wavelength
            path_ray(Wv, p, org, dir);
            f3 o = xyz(org), d = xyz(dir);
            Surf sf = surface(S, me.ref, o, d, me.t);
            const ExtMat& M = S.mat[sf.mat];
            uint32_t key = aux.x;
            float l0 = org.w;
            f3 N = sf.n;
            float ci = -dot(d, N);
            bool entering = ci > 0.0f;
            if (!entering) { N = -N; ci = -ci; }
            float eta = ior(M, l0);
            float ei = entering ? 1.0f : eta, et = entering ? eta : 1.0f;
            float er = ei / et;
            float s2t = er * er * fmaxf(1.0f - ci * ci, 0.0f);
            float Fr = 1.0f, ct = 0.0f;
            if (s2t < 1.0f) {
                ct = sqrtf(1.0f - s2t);
                Fr = fresnel_dielectric(ci, ct, ei, et);
            }
            f3 wo, P;
            float os = offset_scale(sf.p);
            if (rnd(key, (uint32_t)bounce, 2) < Fr) {
                wo = d + N * (2.0f * ci);
                P = sf.p + N * os;
            } else {
                wo = normalize(d * er + N * (er * ci - ct));
                P = sf.p - N * os;
            }
            f4 T = f4of(Wv.pr[p].T);
            uint32_t fl = aux.y;
            if (M.dispersive && S.n_lambda > 1 && !(fl & 1u)) {
                T = f4{T.x * (float)S.n_lambda, 0.0f, 0.0f, 0.0f};
                fl |= 1u;
            }
            if (bounce + 1 < maxDepth && roulette(S, key, bounce, T)) {
                cont = true;
                ro = make_float4(P.x, P.y, P.z, l0);
                rd = make_float4(wo.x, wo.y, wo.z, 0.0f);
                Wv.pr[p].org = ro;
                Wv.pr[p].dir = rd;
                Wv.pr[p].T = make_float4(T.x, T.y, T.z, T.w);
                Wv.pr[p].aux = make_uint4(key, fl | 2u, aux.z, aux.w);
            }
        }
        int* const ctr[1] = {qlen(Wv, nxt)};
        int slot[1];
        block_append<kBlk, 1>(ctr, cont ? 1u : 0u, slot);
        if (cont) {
            Wv.q[nxt][slot[0]] = p;
            Wv.rq_org[nxt][slot[0]] = ro;
            Wv.rq_dir[nxt][slot[0]] = rd;
        }
    }
}

// Smooth conductor: mirror reflection weighted by the complex-index Fresnel term.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_shade_conductor(ExtSceneDev S0, ExtFrame F0, ExtBatch B, ExtWave Wv,
                                                              int cur, int bounce) {
    const int nxt = cur ^ 1;
    const int n = *qlen(Wv, CNT_MAT + Q_CONDUCTOR);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[6 + Q_CONDUCTOR], (unsigned long long)n);
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {
        int idx = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool active = idx < n;
        int p = 0;
        bool cont = false;
        float4 ro = make_float4(0.0f, 0.0f, 0.0f, 0.0f), rd = ro;  // the next ray, for its queue entry
        if (active) {
            const MatEntry me = Wv.mq[Q_CONDUCTOR][idx];
            p = me.p;
            uint4 aux = Wv.pr[p].aux;
            const ExtSceneDev& S = scene_of<MULTI>(S0, B, aux.w);
            const int maxDepth = MULTI ? B.frames[aux.w].maxDepth : F0.maxDepth;
            float4 org, dir;  // org.w: hero wavelength
            path_ray(Wv, p, org, dir);
            f3 o = xyz(org), d = xyz(dir);
            Surf sf = surface(S, me.ref, o, d, me.t);
            const ExtMat& M = S.mat[sf.mat];
            f3 N = dot(sf.n, d) < 0.0f ? sf.n : -sf.n;
            float ci = -dot(d, N);
            float Fr = fresnel_conductor(ci, M.p[0], M.p[1]);
            f4 T = f4of(Wv.pr[p].T) * Fr;
            f3 wo = d + N * (2.0f * ci);
            f3 P = sf.p + N * offset_scale(sf.p);
            if (bounce + 1 < maxDepth && roulette(S, aux.x, bounce, T)) {
                cont = true;
                ro = make_float4(P.x, P.y, P.z, org.w);
                rd = make_float4(wo.x, wo.y, wo.z, 0.0f);
                Wv.pr[p].org = ro;
                Wv.pr[p].dir = rd;
                Wv.pr[p].T = make_float4(T.x, T.y, T.z, T.w);
                Wv.pr[p].aux = make_uint4(aux.x, aux.y | 2u, aux.z, aux.w);
            }
        }
        int* const ctr[1] = {qlen(Wv, nxt)};
        int slot[1];
        block_append<kBlk, 1>(ctr, cont ? 1u : 0u, slot);
        if (cont) {
            Wv.q[nxt][slot[0]] = p;
            Wv.rq_org[nxt][slot[0]] = ro;
            Wv.rq_dir[nxt][slot[0]] = rd;
        }
    }
}

// Shadow rays: any hit before tmax over the scene and the SDFs; unoccluded
// contributions are added to the path.
template <int W, int STACK, typename SE, bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_shadow(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int shq) {
    __shared__ SE stack_mem[STACK * kBlk];
    __shared__ float4 s_wq[kBlk / 64][4 * kQuadBrute];
    SE* stack = stack_mem + threadIdx.x;
    const int n = *qlen(Wv, shq);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[1], (unsigned long long)n);
    unsigned steps = 0, visits = 0, exhausted = 0;
    unsigned long long flops = 0;
    const int stride = gridDim.x * blockDim.x;
    const int rounds = (n + stride - 1) / stride;
    for (int r = 0; r < rounds; ++r) {  // uniform rounds: every lane reaches wave_quads
        const int i = r * stride + blockIdx.x * blockDim.x + threadIdx.x;
        const bool active = i < n;
        float4 so = make_float4(0.0f, 0.0f, 0.0f, 0.0f), sd = so, sx = so;
        if (active) {
            so = Wv.s_org[i];
            sd = Wv.s_dir[i];
            sx = Wv.s_xyz[i];
        }
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI && active ? __float_as_uint(sx.w) : 0u);
        const bool staged = wave_quads(S.quad, S.quad_bvh ? 0 : S.nquad, active, s_wq[threadIdx.x >> 6]);
        if (!active) continue;
        f3 o = mk3(so.x, so.y, so.z), d = mk3(sd.x, sd.y, sd.z);
        float t = so.w;
        uint32_t ref = kMiss;
        bool blocked = scene_trace<W, STACK, SE, true>(S, staged, s_wq[threadIdx.x >> 6], o, d, t, ref, stack, visits);
        for (int s = 0; !blocked && s < S.nsdf; ++s)
            blocked = ext_march<0>(S.sdf[s], o, d, t, S.sdf_steps_max, steps, exhausted, flops) > 0.0f;
        if (!blocked) add_xyz(Wv, __float_as_int(sd.w), mk3(sx.x, sx.y, sx.z));
    }
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
    wave_add_u64(stat_slot(Wv.stats, 13), exhausted);
    wave_add_u64(stat_slot(Wv.stats, 12), visits);
    wave_add_u64(stat_slot(Wv.stats, kStatShadowMarchFlops), flops);
}

// Persistent any-hit kernel for the shadow queue; see the persistent traversal
// notes above. SDF objects are marched when the analytic scene leaves the ray
// unblocked.
template <int W, int STACK, typename SE>
__global__ void __launch_bounds__(kBlk) k_ext_shadow_pt(ExtSceneDev S, ExtWave Wv, int shq) {
    __shared__ SE stack_mem[STACK * kBlk];
    __shared__ float4 s_quad[4 * kQuadBrute];
    SE* stack = stack_mem + threadIdx.x;
    const int nq = stage_quads(S, s_quad);
    const int n = *qlen(Wv, shq);
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&Wv.stats[1], (unsigned long long)n);
    unsigned steps = 0, exhausted = 0;
    unsigned long long flops = 0;
    int i = -1, sp = 0, node = -1, best = -1, path = 0;
    unsigned visits = 0;
    bool blocked = false;
    f3 o = mk3(0.0f, 0.0f, 0.0f), d = o, inv = o, oi = o;
    float t = 0.0f;
    WaveChunk chunk = wave_chunk_init(Wv.cursor + CUR_SHADOW * kSeg * kCntStride, n);
    while (true) {
        const uint64_t idle = __ballot(i < 0);
        if (!chunk.drained() && (__popcll(idle) >= kRefill || ~idle == 0ull)) {
            int e = wave_refill(idle, chunk);
            if (e >= 0) {
                i = e;
                float4 so = Wv.s_org[i], sd = Wv.s_dir[i];
                o = mk3(so.x, so.y, so.z);
                d = mk3(sd.x, sd.y, sd.z);
                t = so.w;
                path = __float_as_int(sd.w);
                blocked = false;
                for (int k = 0; k < nq; ++k) {
                    float tq;
                    if (isect_quad(s_quad + 4 * k, o, d, t, tq)) {
                        blocked = true;
                        break;
                    }
                }
                ray_inv(o, d, inv, oi);
                sp = 0;
                node = !blocked && S.nsph > 0 ? 0 : -1;
            }
        }
        if (__ballot(i >= 0) == 0ull) {
            if (chunk.drained()) break;
            continue;
        }
        if (i >= 0 && node >= 0) {
            ++visits;
            if (bvh_step<W, SE, true, false>(S.sbvh, o, d, inv, oi, t, best, node, sp, stack)) {
                blocked = true;
                node = -1;
            }
        }
        if (i >= 0 && node < 0) {
            for (int k = 0; !blocked && k < S.nplane; ++k) {
                float tp = (ld_ro(S.plane_y + k) - o.y) / d.y;
                blocked = tp > kEps && tp < t;
            }
            for (int s = 0; !blocked && s < S.nsdf; ++s)
                blocked = ext_march<0>(S.sdf[s], o, d, t, S.sdf_steps_max, steps, exhausted, flops) > 0.0f;
            if (!blocked) add_xyz(Wv, path, ld3(Wv.s_xyz, i));
            i = -1;
        }
    }
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
    wave_add_u64(stat_slot(Wv.stats, 13), exhausted);
    wave_add_u64(stat_slot(Wv.stats, 12), visits);
    wave_add_u64(stat_slot(Wv.stats, kStatShadowMarchFlops), flops);
}

// Lenses and cyclides against the extension rays of queue `cur`, after the
// intersect kernel and before the SDF marches and the classification (scenes
// that have any): a closer hit replaces the record. The intersect kernels'
// SDF box mask was taken with the farther hit, so it can only name boxes a
// march then bounds by the nearer one.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_extras(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int cur) {
    const int n = *qlen(Wv, cur);
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < n; idx += gridDim.x * blockDim.x) {
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? Wv.pr[Wv.q[cur][idx]].aux.w : 0u);
        if (S.nlens + S.ncyc == 0) continue;
        f3 o, d;
        ray_of(Wv, cur, idx, o, d);
        const float2 h = Wv.hit[idx];
        float t = h.x;
        uint32_t ref = __float_as_uint(h.y);
        if (trace_extras<false>(S, o, d, t, ref)) Wv.hit[idx] = make_float2(t, __uint_as_float(ref));
    }
}
// Lenses and cyclides against the shadow rays of counter `shq`, before the
// shadow kernel: a blocked ray's contribution becomes zero.
template <bool MULTI>
__global__ void __launch_bounds__(kBlk) k_ext_extras_shadow(ExtSceneDev S0, ExtBatch B, ExtWave Wv, int shq) {
    const int n = *qlen(Wv, shq);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 sx = Wv.s_xyz[i];
        const ExtSceneDev& S = scene_of<MULTI>(S0, B, MULTI ? __float_as_uint(sx.w) : 0u);
        if (S.nlens + S.ncyc == 0) continue;
        const float4 so = Wv.s_org[i], sd = Wv.s_dir[i];
        float t = so.w;
        uint32_t ref = kMiss;
        if (trace_extras<true>(S, mk3(so.x, so.y, so.z), mk3(sd.x, sd.y, sd.z), t, ref))
            Wv.s_xyz[i] = make_float4(0.0f, 0.0f, 0.0f, sx.w);
    }
}

// Per pixel, in sample order: add each path's XYZ to the running sum.
__global__ void __launch_bounds__(kBlk) k_ext_accum(ExtFrame F, ExtWave Wv) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nans = 0;
    if (j < F.pc) {
        float ax = 0.0f, ay = 0.0f, az = 0.0f;
        if (F.k0 != 0) { ax = Wv.sum[j]; ay = Wv.sum[F.pc + j]; az = Wv.sum[2 * F.pc + j]; }
        for (int s = 0; s < F.m; ++s) {
            float4 a = Wv.acc[s * F.pc + j];
            if (a.x != a.x || a.y != a.y || a.z != a.z) { ++nans; continue; }
            ax += a.x; ay += a.y; az += a.z;
        }
        Wv.sum[j] = ax; Wv.sum[F.pc + j] = ay; Wv.sum[2 * F.pc + j] = az;
    }
    wave_add_u64(stat_slot(Wv.stats, 4), nans);
}

struct ExtFinish {
    int p0, pc, spf, currentSamples, frame, ISO;
    float apSize, FPS, persistence;
};
__global__ void __launch_bounds__(kBlk) k_ext_finish(ExtFinish F, const float* sum, float4* texels) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= F.pc) return;
    float e = F.apSize * F.apSize * (float)F.ISO / (float)F.spf;
    f3 out = mk3(sum[j], sum[F.pc + j], sum[2 * F.pc + j]) * e;
    float4 tx = texels[F.p0 + j];
    f3 in = mk3(tx.x, tx.y, tx.z);
    if (F.currentSamples == F.spf && F.frame > F.spf) {
        float wgt = ptm_pow(2.0f, -8.0f / (F.FPS * F.persistence));
        out = (1.0f - wgt) * out + wgt * in;
    } else {
        int u = F.currentSamples / F.spf;
        out = ((float)(u - 1) * in + out) / (float)u;
    }
    texels[F.p0 + j] = make_float4(out.x, out.y, out.z, 1.0f);
}

// Serving sink: per pixel, in sample order, the first and second moments of
// the per-sample XYZ over a request's samples (sum[c * P + pixel],
// sumsq[c * P + pixel]); samples with a NaN are dropped and counted.
__global__ void __launch_bounds__(kBlk) k_ext_moments(ExtFrame F, ExtWave Wv, int P, float* sum, float* sumsq,
                                                      int* nsamp) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nans = 0;
    if (j < F.pc) {
        const int px = F.p0 + j;
        float ax = sum[px], ay = sum[P + px], az = sum[2 * P + px];
        float qx = sumsq[px], qy = sumsq[P + px], qz = sumsq[2 * P + px];
        int n = nsamp[px];
        for (int s = 0; s < F.m; ++s) {
            float4 a = Wv.acc[s * F.pc + j];
            if (a.x != a.x || a.y != a.y || a.z != a.z) { ++nans; continue; }
            ax += a.x; ay += a.y; az += a.z;
            qx += a.x * a.x; qy += a.y * a.y; qz += a.z * a.z;
            ++n;
        }
        sum[px] = ax; sum[P + px] = ay; sum[2 * P + px] = az;
        sumsq[px] = qx; sumsq[P + px] = qy; sumsq[2 * P + px] = qz;
        nsamp[px] = n;
    }
    wave_add_u64(stat_slot(Wv.stats, 4), nans);
}

// A request's per-pixel moment buffers (k_ext_moments layout).
struct ExtMomentSink {
    float* sum;
    float* sumsq;
    int* nsamp;
    float* sum_odd;  // optional: sums and counts of the samples with an odd index
    int* n_odd;
    int P;
    int first;       // index of the quantum's first sample
};

// Batched serving sink: one thread per (segment, pixel) adds the pixel's
// samples of the segment, in sample order, to its request's moments.
__global__ void __launch_bounds__(kBlk) k_ext_moments_batch(const ExtSeg* segs, const ExtMomentSink* sinks,
                                                            const ExtBatchHdr* hdr, ExtWave Wv) {
    const int npix = hdr->npix, nseg = hdr->nseg;
    unsigned nans = 0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < npix; t += gridDim.x * blockDim.x) {
        const ExtSeg& g = segs[seg_find(nseg, t, [&](int k) { return segs[k].pix_base; })];
        const ExtMomentSink& K = sinks[g.rid];
        const int j = t - g.pix_base, px = g.p0 + j, P = K.P;
        float ax = K.sum[px], ay = K.sum[P + px], az = K.sum[2 * P + px];
        float qx = K.sumsq[px], qy = K.sumsq[P + px], qz = K.sumsq[2 * P + px];
        int n = K.nsamp[px];
        for (int s = 0; s < g.m; ++s) {
            float4 a = Wv.acc[g.base + s * g.pc + j];
            if (a.x != a.x || a.y != a.y || a.z != a.z) { ++nans; continue; }
            ax += a.x; ay += a.y; az += a.z;
            qx += a.x * a.x; qy += a.y * a.y; qz += a.z * a.z;
            ++n;
        }
        K.sum[px] = ax; K.sum[P + px] = ay; K.sum[2 * P + px] = az;
        K.sumsq[px] = qx; K.sumsq[P + px] = qy; K.sumsq[2 * P + px] = qz;
        K.nsamp[px] = n;
        if (K.sum_odd) {
            // the odd-index samples alone: two independent half images for
            // the denoiser's error estimate
            float ox = K.sum_odd[px], oy = K.sum_odd[P + px], oz = K.sum_odd[2 * P + px];
            int no = K.n_odd[px];
            for (int s = (K.first + 1) & 1; s < g.m; s += 2) {
                float4 a = Wv.acc[g.base + s * g.pc + j];
                if (a.x != a.x || a.y != a.y || a.z != a.z) continue;
                ox += a.x; oy += a.y; oz += a.z;
                ++no;
            }
            K.sum_odd[px] = ox; K.sum_odd[P + px] = oy; K.sum_odd[2 * P + px] = oz;
            K.n_odd[px] = no;
        }
    }
    wave_add_u64(stat_slot(Wv.stats, 4), nans);
}

// Relative RMSE estimate of a request's mean image: over pixels and XYZ
// channels, sum of (sample variance / n) and sum of mean^2. Per block, then
// one double atomic per block into line blockIdx.x % 64 of `out` (2 doubles
// per 256-byte line); the host adds the lines.
__global__ void __launch_bounds__(kBlk) k_ext_rel_error(int P, const float* sum, const float* sumsq, const int* nsamp,
                                                        double* out) {
    __shared__ double s_v[kBlk / 64], s_m[kBlk / 64];
    double v = 0.0, m = 0.0;
    for (int px = blockIdx.x * blockDim.x + threadIdx.x; px < P; px += gridDim.x * blockDim.x) {
        const int n = nsamp[px];
        if (n < 2) continue;
        for (int c = 0; c < 3; ++c) {
            const double s1 = sum[c * P + px], s2 = sumsq[c * P + px];
            const double mean = s1 / n;
            const double var = fmax(s2 - s1 * mean, 0.0) / (double)(n - 1);
            v += var / n;
            m += mean * mean;
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        v += __shfl_down(v, off);
        m += __shfl_down(m, off);
    }
    if (__lane_id() == 0) {
        s_v[threadIdx.x >> 6] = v;
        s_m[threadIdx.x >> 6] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tv = 0.0, tm = 0.0;
        for (int w = 0; w < kBlk / 64; ++w) { tv += s_v[w]; tm += s_m[w]; }
        double* line = out + (size_t)(blockIdx.x % 64) * 32;
        atomicAdd(&line[0], tv);
        atomicAdd(&line[1], tm);
    }
}

// ---------------------------------------------------------------------------------
// Host side.
struct ExtSceneHost {
    DevAlloc mem;
    ExtSceneDev dev{};
    int stack_need = 0;
    int width = 4;
    bool small_nodes = true;  // every BVH has < 65536 nodes: 16-bit stack entries
    std::vector<int> sdf_kinds;  // kinds present, one march launch per kind
};

struct ExtCtxState {
    DevAlloc mem;
    ExtWave wave{};
    int cu = 0;  // compute units of the context's device (grid sizing)
    DevAlloc batch_mem;  // batched waves: header + scene, frame, segment and sink tables of up to kMaxBatch requests
    // pinned host copy of those tables: one asynchronous upload per wave, and
    // an event marking when the staging copy may be rewritten
    void* batch_host = nullptr;
    hipEvent_t batch_done = nullptr;
    bool batch_pending = false;
    std::map<uint64_t, hipGraphExec_t> graphs;  // captured batched waves, by launch shape
};

static void drop_graphs(ExtCtxState* st) {
    for (auto& kv : st->graphs) (void)hipGraphExecDestroy(kv.second);
    st->graphs.clear();
}

static inline f3 v3(const float* a) { return mk3(a[0], a[1], a[2]); }

static int upload_ext(pt_ctx* ctx, pt_scene* s, ExtSceneHost* h) {
    const pt::SceneModel& m = s->model;
    std::vector<float4> sph, quad;
    std::vector<int2> sph_ml, quad_ml, plane_ml;
    std::vector<float> plane_y;
    std::vector<ExtLight> lights;
    auto light_of = [&](int lightID, int type, int prim) -> int {
        if (lightID <= 0) return -1;
        if ((size_t)lightID > m.lights.size()) return -2;
        ExtLight L{};
        L.type = type;
        L.prim = prim;
        L.T = m.lights[(size_t)lightID - 1].emission[0];
        L.lum = m.lights[(size_t)lightID - 1].emission[1];
        lights.push_back(L);
        return (int)lights.size() - 1;
    };
    for (const pt::RSphere& o : m.spheres) {
        int li = light_of(o.lightID, REF_SPHERE, (int)sph.size());
        if (li == -2) { ctx->err = "sphere lightID out of range"; return PT_ERR_INVALID_ARG; }
        if (li >= 0) { lights[(size_t)li].p = v3(o.pos); lights[(size_t)li].radius = o.radius; }
        sph.push_back(make_float4(o.pos[0], o.pos[1], o.pos[2], o.radius));
        sph_ml.push_back(make_int2(o.materialID - 1, li));
    }
    std::vector<pt::Aabb> qbox;
    auto add_quad = [&](f3 Q, f3 e1, f3 e2, int materialID, int lightID) -> int {
        f3 nn = cross(e1, e2);
        float nn2 = dot(nn, nn);
        if (!(nn2 > 0.0f)) return PT_OK;  // degenerate: skip
        f3 w = nn * (1.0f / nn2);
        f3 A = cross(e2, w), B = cross(w, e1);
        float area = sqrtf(nn2);
        int idx = (int)quad_ml.size();
        int li = light_of(lightID, REF_QUAD, idx);
        if (li == -2) { ctx->err = "quad lightID out of range"; return PT_ERR_INVALID_ARG; }
        if (li >= 0) {
            ExtLight& L = lights[(size_t)li];
            L.p = Q; L.e1 = e1; L.e2 = e2; L.n = nn * (1.0f / area); L.area = area;
        }
        quad.push_back(make_float4(Q.x, Q.y, Q.z, dot(w, Q)));
        quad.push_back(make_float4(w.x, w.y, w.z, area));
        quad.push_back(make_float4(A.x, A.y, A.z, 0.0f));
        quad.push_back(make_float4(B.x, B.y, B.z, 0.0f));
        quad_ml.push_back(make_int2(materialID - 1, li));
        pt::Aabb b;
        f3 c[4] = {Q, Q + e1, Q + e2, Q + e1 + e2};
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = 1e30f; b.hi[k] = -1e30f;
            for (int v = 0; v < 4; ++v) { b.lo[k] = std::min(b.lo[k], comp(c[v], k)); b.hi[k] = std::max(b.hi[k], comp(c[v], k)); }
            float pad = 1e-6f * std::max(1.0f, std::fabs(b.hi[k]));
            b.lo[k] -= pad; b.hi[k] += pad;
            if (!std::isfinite(b.lo[k]) || !std::isfinite(b.hi[k])) {
                ctx->err = "quad/box bounds overflow the float range";
                return PT_ERR_INVALID_ARG;
            }
        }
        qbox.push_back(b);
        return PT_OK;
    };
    for (const pt::Quad& q : m.quads) {
        int st = add_quad(v3(q.corner), v3(q.e1), v3(q.e2), q.materialID, q.lightID);
        if (st != PT_OK) return st;
    }
    for (const pt::RBox& bx : m.boxes) {
        // A rotated box as six outward-facing parallelograms (normal = e1 x e2);
        // the box frame is the reference's RotationMatrix (src/shader.comp:337-364).
        m3 M = rotation_deg(v3(bx.rotation));
        f3 hs = v3(bx.size) * 0.5f, c = v3(bx.pos);
        f3 ax = mvmul(M, mk3(1, 0, 0)), ay = mvmul(M, mk3(0, 1, 0)), az = mvmul(M, mk3(0, 0, 1));
        f3 X = ax * (2.0f * hs.x), Y = ay * (2.0f * hs.y), Z = az * (2.0f * hs.z);
        f3 o = c - ax * hs.x - ay * hs.y - az * hs.z;
        f3 o2 = o + X + Y + Z;
        f3 faces[6][3] = {{o, Z, Y}, {o, X, Z}, {o, Y, X}, {o2, -Y, -Z}, {o2, -Z, -X}, {o2, -X, -Y}};
        for (auto& f : faces) {
            int st = add_quad(f[0], f[1], f[2], bx.materialID, bx.lightID);
            if (st != PT_OK) return st;
        }
    }
    // Lights found by next-event estimation come first; the lenses' and
    // cyclides' (hit by BSDF sampling only) after them.
    const int nlight_nee = (int)lights.size();
    std::vector<ExtLens> lenses;
    std::vector<ExtCyc> cycs;
    for (const pt::RLens& o : m.lenses) {
        ExtLens L{};
        L.pos = v3(o.pos);
        L.M = rotation_deg(v3(o.rotation));
        const float r = o.radius, f = o.focalLength, th = o.thickness;
        L.R = 2.0f * f;
        if (!(r > 0.0f) || !(L.R >= r) || !std::isfinite(L.R) || !std::isfinite(th)) {
            ctx->err = "lens radius must be in (0, 2 focalLength]";
            return PT_ERR_INVALID_ARG;
        }
        L.h = L.R - std::sqrt(L.R * L.R - r * r);
        L.conv = o.isConverging ? 1 : 0;
        L.sp = o.isConverging ? 0.5f * th + L.h : -0.5f * th;
        // the caps lie within |x| <= |th| / 2 + h of the centre and within r of the axis
        const float ax = 0.5f * std::fabs(th) + L.h;
        L.br2 = ax * ax + r * r;
        L.mat = o.materialID - 1;
        L.light = light_of(o.lightID, REF_LENS, (int)lenses.size());
        if (L.light == -2) { ctx->err = "lens lightID out of range"; return PT_ERR_INVALID_ARG; }
        lenses.push_back(L);
    }
    for (const pt::RCyclide& o : m.cyclides) {
        ExtCyc C{};
        C.pos = v3(o.pos);
        C.M = rotation_deg(v3(o.rotation));
        if (!(o.scale[0] != 0.0f && o.scale[1] != 0.0f && o.scale[2] != 0.0f) || !(o.brad > 0.0f)) {
            ctx->err = "cyclide scale must be nonzero and its boundingRadius positive";
            return PT_ERR_INVALID_ARG;
        }
        C.inv_scale = mk3(1.0f / o.scale[0], 1.0f / o.scale[1], 1.0f / o.scale[2]);
        C.A = o.a; C.B = o.b; C.C = o.c; C.D = o.d;
        C.brad2 = o.brad;  // the reference's boundingRadius is compared with squared distances
        C.mat = o.materialID - 1;
        C.light = light_of(o.lightID, REF_CYC, (int)cycs.size());
        if (C.light == -2) { ctx->err = "cyclide lightID out of range"; return PT_ERR_INVALID_ARG; }
        cycs.push_back(C);
    }
    for (const pt::RPlane& pl : m.planes) {
        if (pl.lightID > 0) {
            ctx->err = "extended mode cannot sample an emissive infinite plane";
            return PT_ERR_STATE;
        }
        plane_y.push_back(pl.pos[1]);
        plane_ml.push_back(make_int2(pl.materialID - 1, -1));
    }
    std::vector<ExtSdf> sdfs;
    std::vector<int> kinds;
    for (const pt::RSdf& sd : m.sdfs) {
        ExtSdf e{};
        e.kind = sd.desc.kind;
        e.iters = sd.desc.iterations;
        e.bailout = sd.desc.bailout > 0.0f ? sd.desc.bailout : 16.0f;
        const float m0 = sd.desc.material[0];  // an id far outside the table is rejected below
        e.mat = (m0 > -1e9f && m0 < 1e9f) ? (int)std::floor(m0) : -1000000000;
        e.c = v3(sd.pos);
        e.lo = v3(sd.pos) - v3(sd.size) * 0.5f;
        e.hi = v3(sd.pos) + v3(sd.size) * 0.5f;
        sdfs.push_back(e);
        if (std::find(kinds.begin(), kinds.end(), e.kind) == kinds.end()) kinds.push_back(e.kind);
    }
    std::vector<ExtMat> mats;
    for (const pt::Material& mm : m.materials) {
        ExtMat e{};
        e.type = mm.type;
        switch (mm.type) {
            case pt::MAT_LAMBERT_LOBE:
                e.p[0] = mm.reflection[0]; e.p[1] = mm.reflection[1]; e.p[2] = mm.reflection[2];
                break;
            case pt::MAT_LAMBERT_GAUSS:
                e.n = mm.n_gauss;
                for (int k = 0; k < 9; ++k) e.p[k] = mm.gauss[k];
                break;
            case pt::MAT_DIELECTRIC:
                e.n = mm.dispersion;
                for (int k = 0; k < 6; ++k) e.p[k] = mm.disp[k];
                e.dispersive = mm.dispersion == pt::DISP_SELLMEIER || mm.disp[1] != 0.0f;
                break;
            default:
                e.p[0] = mm.cond_n; e.p[1] = mm.cond_k;
                break;
        }
        mats.push_back(e);
    }
    auto bad_mat = [&](int mi) { return mi < 0 || mi >= (int)mats.size(); };
    for (auto& ml : sph_ml) if (bad_mat(ml.x)) { ctx->err = "sphere materialID out of range"; return PT_ERR_INVALID_ARG; }
    for (auto& ml : quad_ml) if (bad_mat(ml.x)) { ctx->err = "quad/box materialID out of range"; return PT_ERR_INVALID_ARG; }
    for (auto& ml : plane_ml) if (bad_mat(ml.x)) { ctx->err = "plane materialID out of range"; return PT_ERR_INVALID_ARG; }
    for (auto& e : sdfs) if (bad_mat(e.mat)) { ctx->err = "sdf material out of range"; return PT_ERR_INVALID_ARG; }
    for (auto& e : lenses) if (bad_mat(e.mat)) { ctx->err = "lens materialID out of range"; return PT_ERR_INVALID_ARG; }
    for (auto& e : cycs) if (bad_mat(e.mat)) { ctx->err = "cyclide materialID out of range"; return PT_ERR_INVALID_ARG; }

    const int width = ctx->cfg.bvh_width ? ctx->cfg.bvh_width : 2;
    // Defaults measured on the 4096-sphere workload: binary nodes (64 B) with one
    // sphere per leaf beat 4- and 8-wide nodes and larger leaves.
    const int leaf = ctx->cfg.bvh_leaf ? ctx->cfg.bvh_leaf : 1;
    // Sphere BVH with leaf-ordered sphere records.
    std::vector<pt::Aabb> sbox(sph.size());
    std::vector<uint32_t> sref(sph.size());
    for (size_t i = 0; i < sph.size(); ++i) {
        for (int k = 0; k < 3; ++k) {
            float c = k == 0 ? sph[i].x : (k == 1 ? sph[i].y : sph[i].z);
            sbox[i].lo[k] = c - sph[i].w;
            sbox[i].hi[k] = c + sph[i].w;
            if (!std::isfinite(sbox[i].lo[k]) || !std::isfinite(sbox[i].hi[k])) {
                ctx->err = "sphere bounds overflow the float range";
                return PT_ERR_INVALID_ARG;
            }
        }
        sref[i] = (uint32_t)i;
    }
    pt::WideBvh sb = pt::build_wide_bvh(sbox, sref, width, leaf);
    std::vector<float4> sleaf(sb.prim_refs.size());
    std::vector<int> sids(sb.prim_refs.size());
    for (size_t k = 0; k < sb.prim_refs.size(); ++k) {
        sleaf[k] = sph[sb.prim_refs[k]];
        sids[k] = (int)sb.prim_refs[k];
    }
    // A leaf child of one sphere keeps that sphere in its parent's record
    // (lo_x, lo_y, lo_z, hi_x = centre, radius; count kInlineSphere), so the
    // traversal tests it without fetching the sphere record after the node.
    for (int n = 0; n < sb.n_nodes; ++n) {
        uint32_t* w = sb.words.data() + (size_t)n * 8 * width;
        for (int j = 0; j < width; ++j) {
            int32_t cnt, child;
            std::memcpy(&cnt, &w[7 * width + j], 4);
            std::memcpy(&child, &w[6 * width + j], 4);
            if (cnt != 1 || child < 0 || (size_t)child >= sleaf.size()) continue;
            const float4 c = sleaf[(size_t)child];
            std::memcpy(&w[0 * width + j], &c.x, 4);
            std::memcpy(&w[1 * width + j], &c.y, 4);
            std::memcpy(&w[2 * width + j], &c.z, 4);
            std::memcpy(&w[3 * width + j], &c.w, 4);
            const int32_t inl = kInlineSphere;
            std::memcpy(&w[7 * width + j], &inl, 4);
        }
    }
    const bool quad_bvh = quad_ml.size() > (size_t)kQuadBrute;
    pt::WideBvh qb;
    std::vector<float4> qleaf;
    std::vector<int> qids;
    if (quad_bvh) {
        std::vector<uint32_t> qref(quad_ml.size());
        for (size_t i = 0; i < qref.size(); ++i) qref[i] = (uint32_t)i;
        qb = pt::build_wide_bvh(qbox, qref, width, leaf);
        for (uint32_t r : qb.prim_refs) {
            for (int v = 0; v < 4; ++v) qleaf.push_back(quad[4 * (size_t)r + (size_t)v]);
            qids.push_back((int)r);
        }
    }
    int stack_need = std::max(sb.max_stack, quad_bvh ? qb.max_stack : 0);
    if (stack_need > 64) {
        ctx->err = "BVH too deep for the 64-entry traversal stack";
        return PT_ERR_CAPACITY;
    }
    h->stack_need = stack_need;
    h->width = width;
    h->small_nodes = sb.n_nodes < 65536 && (!quad_bvh || qb.n_nodes < 65536);
    h->sdf_kinds = kinds;

    size_t bytes = 0;
    auto sz = [&](size_t b) { bytes += ((b + 255) / 256) * 256; };
    sz(sph.size() * 16); sz(sph_ml.size() * 8); sz(quad.size() * 16); sz(quad_ml.size() * 8);
    sz(plane_y.size() * 4); sz(plane_ml.size() * 8); sz(sdfs.size() * sizeof(ExtSdf)); sz(mats.size() * sizeof(ExtMat));
    sz(lights.size() * sizeof(ExtLight)); sz(lenses.size() * sizeof(ExtLens)); sz(cycs.size() * sizeof(ExtCyc));
    sz(sb.words.size() * 4); sz(sleaf.size() * 16); sz(sids.size() * 4);
    sz(qb.words.size() * 4); sz(qleaf.size() * 16); sz(qids.size() * 4);
    int st = dev_alloc(ctx, h->mem, bytes);
    if (st != PT_OK) return st;
    char* p = static_cast<char*>(h->mem.p);
    bool fail = false;
    auto put = [&](const void* src, size_t b) -> void* {
        void* dst = p;
        if (b && hipMemcpy(dst, src, b, hipMemcpyHostToDevice) != hipSuccess) fail = true;
        p += ((b + 255) / 256) * 256;
        return dst;
    };
    ExtSceneDev& D = h->dev;
    D.sph = (const float4*)put(sph.data(), sph.size() * 16);
    D.sph_ml = (const int2*)put(sph_ml.data(), sph_ml.size() * 8);
    D.quad = (const float4*)put(quad.data(), quad.size() * 16);
    D.quad_ml = (const int2*)put(quad_ml.data(), quad_ml.size() * 8);
    D.plane_y = (const float*)put(plane_y.data(), plane_y.size() * 4);
    D.plane_ml = (const int2*)put(plane_ml.data(), plane_ml.size() * 8);
    D.sdf = (const ExtSdf*)put(sdfs.data(), sdfs.size() * sizeof(ExtSdf));
    D.mat = (const ExtMat*)put(mats.data(), mats.size() * sizeof(ExtMat));
    D.light = (const ExtLight*)put(lights.data(), lights.size() * sizeof(ExtLight));
    D.lens = (const ExtLens*)put(lenses.data(), lenses.size() * sizeof(ExtLens));
    D.cyc = (const ExtCyc*)put(cycs.data(), cycs.size() * sizeof(ExtCyc));
    D.sbvh.nodes = (const uint32_t*)put(sb.words.data(), sb.words.size() * 4);
    D.sbvh.prims = (const float4*)put(sleaf.data(), sleaf.size() * 16);
    D.sbvh.ids = (const int*)put(sids.data(), sids.size() * 4);
    D.sbvh.count = (int)sleaf.size();
    D.qbvh.nodes = (const uint32_t*)put(qb.words.data(), qb.words.size() * 4);
    D.qbvh.prims = (const float4*)put(qleaf.data(), qleaf.size() * 16);
    D.qbvh.ids = (const int*)put(qids.data(), qids.size() * 4);
    D.qbvh.count = (int)qids.size();
    if (fail) return hip_fail(ctx, hipGetLastError(), "extended scene upload");
    D.nsph = (int)sph.size();
    D.nquad = (int)quad_ml.size();
    D.quad_bvh = quad_bvh ? 1 : 0;
    D.nplane = (int)plane_y.size();
    D.nsdf = (int)sdfs.size();
    D.nmat = (int)mats.size();
    D.nlight = (int)lights.size();
    D.nlight_nee = nlight_nee;
    D.nlens = (int)lenses.size();
    D.ncyc = (int)cycs.size();
    D.has_env = m.env_enabled ? 1 : 0;
    D.env = m.env_radiance;
    D.n_lambda = m.render.n_lambda;
    D.lmin = m.render.lambda_min;
    D.lrange = m.render.lambda_max - m.render.lambda_min;
    D.rr_depth = m.render.rr_depth;
    D.seed = m.render.seed;
    D.camera = m.render.camera;
    D.sdf_steps_max = ctx->cfg.sdf_max_steps > 0 ? ctx->cfg.sdf_max_steps : 256;
    s->device_bytes += (int64_t)bytes;
    return PT_OK;
}

int ext_scene_build(pt_ctx* ctx, pt_scene* s) {
    if (s->ext) return PT_OK;
    ExtSceneHost* h = new ExtSceneHost();
    int st = upload_ext(ctx, s, h);
    if (st != PT_OK) {
        dev_free(ctx, h->mem);
        delete h;
        return st;
    }
    s->ext = h;
    return PT_OK;
}

void ext_scene_release(pt_ctx* ctx, pt_scene* s) {
    if (!s->ext) return;
    ExtSceneHost* h = static_cast<ExtSceneHost*>(s->ext);
    dev_free(ctx, h->mem);
    delete h;
    s->ext = nullptr;
}

template <typename T>
static T* carve_e(char*& p, size_t n) {
    T* r = reinterpret_cast<T*>(p);
    p += ((n * sizeof(T) + 255) / 256) * 256;
    return r;
}

static int ensure_ext_wave(pt_ctx* ctx, int cap) {
    ExtCtxState* st = static_cast<ExtCtxState*>(ctx->ext_state);
    if (st && st->wave.cap >= cap) return PT_OK;
    if (!st) {
        st = new ExtCtxState();
        ctx->ext_state = st;
    }
    drop_graphs(st);  // they hold the old wave's addresses
    dev_free(ctx, st->mem);
    size_t n = (size_t)cap;
    // per path: record (org, dir, T, aux) 64, acc 16, hit 8, extension queues 4 * 2,
    // material-queue entries 16 * N_MATQ, shadow org/dir/xyz 48, first-hit
    // features 32, SDF mask 4, sums 12, queue-ordered rays 64
    size_t per_path = 64 + 16 + 8 + 4 * 2 + 16 * N_MATQ + 48 + 32 + 4 + 12 + 64;
    // + 256-byte rounding of 26 arrays + the counters and cursors, one line each
    int rc = dev_alloc(ctx, st->mem, n * per_path + 28 * 256 + (size_t)(N_CNT + N_CUR * kSeg) * kCntStride * 4);
    if (rc != PT_OK) return rc;
    char* p = static_cast<char*>(st->mem.p);
    ExtWave& W = st->wave;
    W.pr = carve_e<PathRec>(p, n);
    W.acc = carve_e<float4>(p, n);
    W.hit = carve_e<float2>(p, n);
    W.q[0] = carve_e<int>(p, n); W.q[1] = carve_e<int>(p, n);
    for (int c = 0; c < 2; ++c) {
        W.rq_org[c] = carve_e<float4>(p, n);
        W.rq_dir[c] = carve_e<float4>(p, n);
    }
    for (int c = 0; c < N_MATQ; ++c) W.mq[c] = carve_e<MatEntry>(p, n);
    W.s_org = carve_e<float4>(p, n); W.s_dir = carve_e<float4>(p, n); W.s_xyz = carve_e<float4>(p, n);
    W.aov0 = carve_e<float4>(p, n); W.aov1 = carve_e<float4>(p, n);
    W.smask = carve_e<uint32_t>(p, n);
    W.sum = carve_e<float>(p, 3 * n);
    W.counts = carve_e<int>(p, (size_t)N_CNT * kCntStride);
    W.cursor = carve_e<int>(p, (size_t)N_CUR * kSeg * kCntStride);
    W.stats = ctx->d_stats;
    W.cap = cap;
    return PT_OK;
}

void ext_wave_release(pt_ctx* ctx) {
    ExtCtxState* st = static_cast<ExtCtxState*>(ctx->ext_state);
    if (!st) return;
    drop_graphs(st);
    if (st->batch_done) (void)hipEventSynchronize(st->batch_done);
    if (st->batch_done) (void)hipEventDestroy(st->batch_done);
    if (st->batch_host) (void)hipHostFree(st->batch_host);
    dev_free(ctx, st->batch_mem);
    dev_free(ctx, st->mem);
    delete st;
    ctx->ext_state = nullptr;
}

// What the bounce kernels of one wave need to know about its scene(s).
struct ExtPlan {
    ExtSceneDev S0{};  // the wave's scene (batched: unused by the MULTI kernels)
    ExtFrame F0{};
    ExtBatch B{};      // batched: per-request scene and frame tables in device memory
    bool multi = false;
    bool any_sdf = false;
    bool quad_bvh = false;  // a scene tests its parallelograms through a BVH
    bool extras = false;    // a scene has lenses or cyclides (the k_ext_extras kernels)
    std::vector<int> sdf_kinds;
    int width = 2, stack_need = 1;
    bool small_nodes = true;
    bool aov = false;  // record first-hit features after bounce 0's trace
    int cus = 256;     // compute units of the device
};

// Blocks per compute unit of the persistent march kernels (a separate setting
// from the queue kernels' grid, so the autotuner can pick it per scene). On the
// 1080p mandelbulb + menger scene with 2^25-path waves, 8 measured best
// (2 / 8 / 16: 2.56 / 3.52 / 3.52 G segments/s).
constexpr int kMarchBlocksPerCu = 8;

template <int KIND, bool MULTI>
static void launch_march(pt_ctx* ctx, const char* name, const ExtPlan& P, ExtWave& Wv, int cur, int grid) {
    if (ctx->cfg.march == 1) {
        launch(ctx, name, k_ext_march<KIND, MULTI>, dim3(grid), dim3(kBlk), P.S0, P.B, Wv, cur);
        return;
    }
    const int bpc = ctx->cfg.march_blocks_per_cu > 0 ? ctx->cfg.march_blocks_per_cu : kMarchBlocksPerCu;
    const int mgrid = std::max(1, std::min(grid, P.cus * bpc));
    launch(ctx, name, k_ext_march_pt<KIND, MULTI>, dim3(mgrid), dim3(kBlk), P.S0, P.B, Wv, cur);
}

// One bounce b (extension queue `cur`). Scenes without SDF objects trace the
// extension rays together with bounce b - 1's shadow rays (k_ext_trace); the
// last bounce's shadow rays are traced by launch_tail. Scenes with SDF objects,
// and persistent traversal, keep separate intersect / march / classify /
// shadow kernels.
template <int W, int STACK, typename SE, bool MULTI>
static void launch_bounce(pt_ctx* ctx, const ExtPlan& P, ExtWave& Wv, int cur, int b, int grid) {
    const ExtSceneDev& S = P.S0;
    // Persistent traversal unless one ray per lane is asked for (one scene only);
    // it tests parallelograms directly, so scenes with a parallelogram BVH take
    // the one-ray-per-lane kernels. With 2^25-path waves it measured 9-10 %
    // faster on the analytic scenes (4096 spheres, Cornell box) and even on SDFs.
    const bool persistent = !MULTI && ctx->cfg.traversal != 2 && !P.quad_bvh;
    // lenses and cyclides are tested by their own kernels between the traversal
    // and the classification, which the fused kernel does in one
    const bool fused = !persistent && !P.any_sdf && !P.extras;
    auto extras = [&]() {
        if (P.extras) launch(ctx, "k_ext_extras", k_ext_extras<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, cur);
    };
    launch(ctx, "k_ext_reset", k_ext_reset, dim3(1), dim3(64), Wv.counts, Wv.cursor, cur ^ 1, shadow_cnt(b));
    if (fused) {
        launch(ctx, "k_ext_trace", k_ext_trace<W, STACK, SE, MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, cur,
               shadow_cnt(b - 1));
    } else if (P.any_sdf) {
        if (persistent)
            launch(ctx, "k_ext_intersect", k_ext_intersect_pt<W, STACK, SE>, dim3(grid), dim3(kBlk), S, Wv, cur);
        else
            launch(ctx, "k_ext_intersect", k_ext_intersect<W, STACK, SE, false, MULTI>, dim3(grid), dim3(kBlk), S, P.B,
                   Wv, cur);
        extras();
        for (int kind : P.sdf_kinds) {
            switch (kind) {
                case PT_SDF_MANDELBULB: launch_march<PT_SDF_MANDELBULB, MULTI>(ctx, "k_ext_march_mandelbulb", P, Wv, cur, grid); break;
                case PT_SDF_MENGER: launch_march<PT_SDF_MENGER, MULTI>(ctx, "k_ext_march_menger", P, Wv, cur, grid); break;
                case PT_SDF_BLOB: launch_march<PT_SDF_BLOB, MULTI>(ctx, "k_ext_march_blob", P, Wv, cur, grid); break;
                default: launch_march<PT_SDF_TERRAIN, MULTI>(ctx, "k_ext_march_terrain", P, Wv, cur, grid); break;
            }
        }
        launch(ctx, "k_ext_classify", k_ext_classify<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, cur);
    } else {
        if (persistent)
            launch(ctx, "k_ext_intersect", k_ext_intersect_pt<W, STACK, SE>, dim3(grid), dim3(kBlk), S, Wv, cur);
        else
            launch(ctx, "k_ext_intersect", k_ext_intersect<W, STACK, SE, false, MULTI>, dim3(grid), dim3(kBlk), S, P.B,
                   Wv, cur);
        extras();
        launch(ctx, "k_ext_classify", k_ext_classify<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, cur);
    }
    if (b == 0 && P.aov) launch(ctx, "k_ext_aov", k_ext_aov<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv);
    launch(ctx, "k_ext_shade_miss", k_ext_shade_terminal<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, (int)Q_MISS);
    launch(ctx, "k_ext_shade_emit", k_ext_shade_terminal<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv, (int)Q_EMIT);
    launch(ctx, "k_ext_shade_lambert", k_ext_shade_lambert<MULTI>, dim3(grid), dim3(kBlk), S, P.F0, P.B, Wv, cur, b);
    launch(ctx, "k_ext_shade_dielectric", k_ext_shade_dielectric<MULTI>, dim3(grid), dim3(kBlk), S, P.F0, P.B, Wv, cur,
           b);
    launch(ctx, "k_ext_shade_conductor", k_ext_shade_conductor<MULTI>, dim3(grid), dim3(kBlk), S, P.F0, P.B, Wv, cur,
           b);
    // Shadow rays push children unordered: near-first ordering measured slower
    // on the 4096-sphere scene (most shadow rays are occluded early anyway).
    if (P.extras)
        launch(ctx, "k_ext_extras_shadow", k_ext_extras_shadow<MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv,
               shadow_cnt(b));
    if (persistent)
        launch(ctx, "k_ext_shadow", k_ext_shadow_pt<W, STACK, SE>, dim3(grid), dim3(kBlk), S, Wv, shadow_cnt(b));
    else if (!fused)
        launch(ctx, "k_ext_shadow", k_ext_shadow<W, STACK, SE, MULTI>, dim3(grid), dim3(kBlk), S, P.B, Wv,
               shadow_cnt(b));
}

// After the last bounce (`nb` bounces, extension queue nb & 1 empty): the
// shadow rays the last bounce emitted, when the bounces are fused.
template <int W, int STACK, typename SE, bool MULTI>
static void launch_tail(pt_ctx* ctx, const ExtPlan& P, ExtWave& Wv, int nb, int grid) {
    const bool persistent = !MULTI && ctx->cfg.traversal != 2 && !P.quad_bvh;
    if (persistent || P.any_sdf || P.extras || nb <= 0) return;
    launch(ctx, "k_ext_trace", k_ext_trace<W, STACK, SE, MULTI>, dim3(grid), dim3(kBlk), P.S0, P.B, Wv, nb & 1,
           shadow_cnt(nb - 1));
}

template <int W, int STACK, typename SE>
static void launch_step(pt_ctx* ctx, const ExtPlan& P, ExtWave& Wv, int cur, int b, int grid, bool tail) {
    if (P.multi) {
        if (tail) launch_tail<W, STACK, SE, true>(ctx, P, Wv, b, grid);
        else launch_bounce<W, STACK, SE, true>(ctx, P, Wv, cur, b, grid);
    } else {
        if (tail) launch_tail<W, STACK, SE, false>(ctx, P, Wv, b, grid);
        else launch_bounce<W, STACK, SE, false>(ctx, P, Wv, cur, b, grid);
    }
}

template <int W>
static void launch_bounce_w(pt_ctx* ctx, const ExtPlan& P, ExtWave& Wv, int cur, int b, int grid, bool tail) {
    if (P.small_nodes) {
        if (P.stack_need <= 32) launch_step<W, 32, uint16_t>(ctx, P, Wv, cur, b, grid, tail);
        else launch_step<W, 64, uint16_t>(ctx, P, Wv, cur, b, grid, tail);
    } else {
        if (P.stack_need <= 32) launch_step<W, 32, int>(ctx, P, Wv, cur, b, grid, tail);
        else launch_step<W, 64, int>(ctx, P, Wv, cur, b, grid, tail);
    }
}

// Bounces 0..depth-1, then the tail.
static void launch_bounces(pt_ctx* ctx, const ExtPlan& P, ExtWave& Wv, int depth, int grid) {
    for (int b = 0; b <= depth; ++b) {
        const int cur = b & 1;
        const bool tail = b == depth;  // b = the number of bounces traced
        switch (P.width) {
            case 2: launch_bounce_w<2>(ctx, P, Wv, cur, b, grid, tail); break;
            case 8: launch_bounce_w<8>(ctx, P, Wv, cur, b, grid, tail); break;
            default: launch_bounce_w<4>(ctx, P, Wv, cur, b, grid, tail); break;
        }
    }
}

// The frame of a dispatch: resolution, sample range, depth and camera of the
// push constants (lens frame as the reference builds it, src/shader.comp:1409-1465).
static ExtFrame frame_of(const pt_push_constants& pc) {
    ExtFrame F{};
    F.W = pc.resolution[0]; F.H = pc.resolution[1]; F.frame = pc.frame; F.spf = pc.samplesPerFrame;
    F.maxDepth = pc.pathLength;
    F.camSize = pc.cameraSize; F.apSize = pc.apertureSize; F.apDist = pc.apertureDist;
    F.camPos = mk3(pc.cameraPosX, pc.cameraPosY, pc.cameraPosZ);
    F.camM = rotation_deg(mk3(pc.cameraAngle[0], pc.cameraAngle[1], 0.0f));
    F.fwd = mk3(F.camM.c0.z, F.camM.c1.z, F.camM.c2.z);
    F.lensPos = F.camPos + F.fwd * pc.lensDistance;
    F.lensM = rotation_deg(mk3(0.0f, 90.0f - pc.cameraAngle[1], pc.cameraAngle[0]));
    float r = pc.lensRadius, f = pc.lensFocalLength;
    F.lensHalf = 2.0f * f - sqrtf(4.0f * f * f - r * r);
    F.lensSlicePos = 0.5f * pc.lensThickness + F.lensHalf;
    F.lensTwoF = 2.0f * f;
    return F;
}

// Compute units of the context's device (grid sizing), after ensure_ext_wave.
// Blocks of kBlk lanes per compute unit in the grid of the queue kernels.
static int blocks_per_cu(const pt_ctx* ctx) { return ctx->cfg.blocks_per_cu > 0 ? ctx->cfg.blocks_per_cu : 8; }

static int ctx_cus(pt_ctx* ctx) {
    ExtCtxState* cs = static_cast<ExtCtxState*>(ctx->ext_state);
    if (cs->cu == 0) {
        hipDeviceProp_t prop;
        cs->cu = (hipGetDeviceProperties(&prop, ctx->device) == hipSuccess && prop.multiProcessorCount > 0)
                     ? prop.multiProcessorCount
                     : 256;
    }
    return cs->cu;
}

// Where a frame's per-path XYZ goes: the reference's texel accumulate
// (running mean / persistence blend of the push constants), or a serving
// request's per-pixel moments.
struct ExtSink {
    pt_accum* acc = nullptr;
    const pt_push_constants* pc = nullptr;
    float* sum = nullptr;
    float* sumsq = nullptr;
    int* nsamp = nullptr;
    float* feat = nullptr;  // features sink: sums of XYZ, first-hit features and Y^2, 12 per pixel
};

// Default paths per wave of a single-scene dispatch: 2^27 paths, 28 GB of wave
// state. Each bounce launch ends with the latency chain of its slowest ray (a
// long BVH traversal or a 256-step march), which a wave of more paths amortises;
// measured on the 1080p benchmark scenes, 2^22 -> 2^25 paths per wave raised
// traced segments per second by 1.3x (spheres) and 2.4x (SDFs), 2^26 by a
// further 6 % (spheres, persistent traversal) and 2^27 by 1.4 % (spheres) and
// 8.7 % (SDFs).
constexpr int64_t kDispatchWave = (int64_t)1 << 27;

static int ext_run(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, const ExtSink& sink) {
    int st = ext_scene_build(ctx, const_cast<pt_scene*>(s));
    if (st != PT_OK) return st;
    const ExtSceneHost& H = *static_cast<const ExtSceneHost*>(s->ext);
    const int Wd = pc->resolution[0], Ht = pc->resolution[1], P = Wd * Ht;
    // Paths per wave: the configured capacity, or kDispatchWave; the wave is
    // allocated for what this dispatch needs, up to that capacity.
    int64_t cap = ctx->cfg.wave_paths > 0 ? ctx->cfg.wave_paths : kDispatchWave;
    cap = std::max<int64_t>(1024, std::min<int64_t>(cap, (int64_t)P * std::max(1, pc->samplesPerFrame)));
    st = ensure_ext_wave(ctx, (int)cap);
    if (st != PT_OK) return st;
    ExtWave& Wv = static_cast<ExtCtxState*>(ctx->ext_state)->wave;

    ExtFrame F = frame_of(*pc);
    const int grid_cap = ctx_cus(ctx) * blocks_per_cu(ctx);
    ExtPlan plan;
    plan.cus = ctx_cus(ctx);
    plan.S0 = H.dev;
    plan.any_sdf = H.dev.nsdf > 0;
    plan.quad_bvh = H.dev.quad_bvh != 0;
    plan.extras = H.dev.nlens + H.dev.ncyc > 0;
    plan.sdf_kinds = H.sdf_kinds;
    plan.width = H.width;
    plan.stack_need = H.stack_need;
    plan.small_nodes = H.small_nodes;
    plan.aov = sink.feat != nullptr;
    const int pcmax = (int)std::min<int64_t>(P, cap);
    for (int p0 = 0; p0 < P; p0 += pcmax) {
        const int pcur = std::min(pcmax, P - p0);
        const int m = (int)std::max<int64_t>(1, std::min<int64_t>(pc->samplesPerFrame, cap / pcur));
        for (int k0 = 0; k0 < pc->samplesPerFrame; k0 += m) {
            F.p0 = p0; F.pc = pcur; F.k0 = k0; F.m = std::min(m, pc->samplesPerFrame - k0);
            const int npaths = F.pc * F.m;
            const int grid = std::min((npaths + kBlk - 1) / kBlk, grid_cap);
            launch(ctx, "k_ext_counts", k_ext_counts, dim3(1), dim3(64), Wv.counts, npaths);
            launch(ctx, "k_ext_raygen", k_ext_raygen, dim3(grid), dim3(kBlk), H.dev, F, Wv);
            plan.F0 = F;
            launch_bounces(ctx, plan, Wv, pc->pathLength, grid);
            const dim3 pg((F.pc + kBlk - 1) / kBlk);
            if (sink.acc)
                launch(ctx, "k_ext_accum", k_ext_accum, pg, dim3(kBlk), F, Wv);
            else if (sink.feat)
                launch(ctx, "k_ext_features", k_ext_features, pg, dim3(kBlk), F, Wv, sink.feat, sink.nsamp);
            else
                launch(ctx, "k_ext_moments", k_ext_moments, pg, dim3(kBlk), F, Wv, P, sink.sum, sink.sumsq, sink.nsamp);
        }
        if (sink.acc) {
            ExtFinish Fin{};
            Fin.p0 = p0; Fin.pc = pcur; Fin.spf = pc->samplesPerFrame; Fin.currentSamples = pc->currentSamples;
            Fin.frame = pc->frame; Fin.ISO = pc->ISO; Fin.apSize = pc->apertureSize; Fin.FPS = pc->FPS;
            Fin.persistence = pc->persistence;
            launch(ctx, "k_ext_finish", k_ext_finish, dim3((pcur + kBlk - 1) / kBlk), dim3(kBlk), Fin,
                   (const float*)Wv.sum, reinterpret_cast<float4*>(sink.acc->d_texels));
        }
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "extended dispatch launch");
    return PT_OK;
}

int ext_dispatch(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc) {
    ExtSink sink;
    sink.acc = acc;
    sink.pc = pc;
    return ext_run(ctx, s, pc, sink);
}

int ext_dispatch_moments(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, float* sum, float* sumsq,
                         int* nsamp) {
    ExtSink sink;
    sink.pc = pc;
    sink.sum = sum;
    sink.sumsq = sumsq;
    sink.nsamp = nsamp;
    return ext_run(ctx, s, pc, sink);
}

int ext_dispatch_batch(pt_ctx* ctx, const ExtBatchItem* items, int n) {
    if (n <= 0) return PT_OK;
    if (n > kMaxBatch) {
        ctx->err = "batched wave: more than " + std::to_string(kMaxBatch) + " requests";
        return PT_ERR_CAPACITY;
    }
    ExtPlan plan;
    plan.multi = true;
    std::vector<ExtSceneDev> scenes((size_t)n);
    std::vector<ExtFrame> frames((size_t)n);
    std::vector<ExtSeg> segs((size_t)n);
    std::vector<ExtMomentSink> sinks((size_t)n);
    int64_t npaths = 0, npix = 0;
    int depth = 0;
    for (int i = 0; i < n; ++i) {
        const ExtBatchItem& it = items[i];
        int st = ext_scene_build(ctx, const_cast<pt_scene*>(it.scene));
        if (st != PT_OK) return st;
        const ExtSceneHost& H = *static_cast<const ExtSceneHost*>(it.scene->ext);
        if (i == 0) {
            plan.width = H.width;
        } else if (H.width != plan.width) {
            ctx->err = "batched wave: scenes with different BVH widths";
            return PT_ERR_STATE;
        }
        plan.stack_need = std::max(plan.stack_need, H.stack_need);
        plan.small_nodes = plan.small_nodes && H.small_nodes;
        plan.any_sdf = plan.any_sdf || H.dev.nsdf > 0;
        plan.quad_bvh = plan.quad_bvh || H.dev.quad_bvh;
        plan.extras = plan.extras || H.dev.nlens + H.dev.ncyc > 0;
        for (int k : H.sdf_kinds)
            if (std::find(plan.sdf_kinds.begin(), plan.sdf_kinds.end(), k) == plan.sdf_kinds.end())
                plan.sdf_kinds.push_back(k);
        const int P = it.pc.resolution[0] * it.pc.resolution[1];
        const int q = it.pc.samplesPerFrame;
        if (P <= 0 || q <= 0) {
            ctx->err = "batched wave: empty request";
            return PT_ERR_INVALID_ARG;
        }
        scenes[(size_t)i] = H.dev;
        frames[(size_t)i] = frame_of(it.pc);
        segs[(size_t)i] = ExtSeg{i, (int)npaths, 0, P, 0, q, (int)npix, 0};
        sinks[(size_t)i] = ExtMomentSink{it.sum, it.sumsq, it.nsamp, it.sum_odd, it.n_odd, P,
                                         it.pc.frame - it.pc.samplesPerFrame};
        npaths += (int64_t)P * q;
        npix += P;
        depth = std::max(depth, it.pc.pathLength);
    }
    int cap = ctx->cfg.wave_paths > 0 ? ctx->cfg.wave_paths : (1 << 22);
    cap = std::max(cap, 1024);
    if (npaths > cap) {
        ctx->err = "batched wave: " + std::to_string(npaths) + " paths exceed the wave capacity " + std::to_string(cap);
        return PT_ERR_CAPACITY;
    }
    int st = ensure_ext_wave(ctx, cap);
    if (st != PT_OK) return st;
    ExtCtxState* cs = static_cast<ExtCtxState*>(ctx->ext_state);
    ExtWave& Wv = cs->wave;
    // tables at fixed addresses: header, scenes, frames, segments, sinks
    const size_t o_sc = 256, o_fr = o_sc + (sizeof(ExtSceneDev) * kMaxBatch + 255) / 256 * 256,
                 o_sg = o_fr + (sizeof(ExtFrame) * kMaxBatch + 255) / 256 * 256,
                 o_sk = o_sg + (sizeof(ExtSeg) * kMaxBatch + 255) / 256 * 256,
                 tb = o_sk + (sizeof(ExtMomentSink) * kMaxBatch + 255) / 256 * 256;
    if (!cs->batch_mem.p) {
        st = dev_alloc(ctx, cs->batch_mem, tb);
        if (st != PT_OK) return st;
    }
    hipError_t e = hipSuccess;
    if (!cs->batch_host) {
        e = hipHostMalloc(&cs->batch_host, tb, hipHostMallocDefault);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&cs->batch_done, hipEventDisableTiming);
        if (e != hipSuccess) {
            if (cs->batch_host) (void)hipHostFree(cs->batch_host);
            cs->batch_host = nullptr;
            return hip_fail(ctx, e, "batched wave staging");
        }
    }
    // the previous wave's upload has left the staging copy before it is rewritten
    if (cs->batch_pending) e = hipEventSynchronize(cs->batch_done);
    if (e != hipSuccess) return hip_fail(ctx, e, "batched wave staging");
    cs->batch_pending = false;
    char* base = static_cast<char*>(cs->batch_mem.p);
    char* hb = static_cast<char*>(cs->batch_host);
    ExtBatchHdr hdr{(int)npaths, (int)npix, n, 0};
    std::memcpy(hb, &hdr, sizeof hdr);
    std::memcpy(hb + o_sc, scenes.data(), sizeof(ExtSceneDev) * (size_t)n);
    std::memcpy(hb + o_fr, frames.data(), sizeof(ExtFrame) * (size_t)n);
    std::memcpy(hb + o_sg, segs.data(), sizeof(ExtSeg) * (size_t)n);
    std::memcpy(hb + o_sk, sinks.data(), sizeof(ExtMomentSink) * (size_t)n);
    const ExtBatchHdr* d_hdr = reinterpret_cast<const ExtBatchHdr*>(base);
    const ExtSceneDev* d_scenes = reinterpret_cast<const ExtSceneDev*>(base + o_sc);
    const ExtFrame* d_frames = reinterpret_cast<const ExtFrame*>(base + o_fr);
    const ExtSeg* d_segs = reinterpret_cast<const ExtSeg*>(base + o_sg);
    const ExtMomentSink* d_sinks = reinterpret_cast<const ExtMomentSink*>(base + o_sk);
    // one upload of the used part of the tables (from pinned memory, so it is
    // asynchronous: no host wait before the wave is enqueued)
    e = hipMemcpyAsync(base, hb, o_sk + sizeof(ExtMomentSink) * (size_t)n, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(cs->batch_done, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "batched wave tables");
    cs->batch_pending = true;
    plan.B = ExtBatch{d_scenes, d_frames};
    plan.S0 = scenes[0];
    plan.F0 = frames[0];
    // a fixed grid (the wave's sizes come from the header), so one captured
    // graph per launch shape replays for every batch of that shape
    const int grid = (int)std::min<int64_t>(((int64_t)cap + kBlk - 1) / kBlk,
                                            (int64_t)ctx_cus(ctx) * blocks_per_cu(ctx));
    plan.cus = ctx_cus(ctx);
    auto launch_wave = [&]() {
        launch(ctx, "k_ext_counts", k_ext_counts_batch, dim3(1), dim3(64), Wv.counts, d_hdr);
        launch(ctx, "k_ext_raygen", k_ext_raygen_batch, dim3(grid), dim3(kBlk), plan.B, d_segs, d_hdr, Wv);
        launch_bounces(ctx, plan, Wv, depth, grid);
        launch(ctx, "k_ext_moments", k_ext_moments_batch, dim3(grid), dim3(kBlk), d_segs, d_sinks, d_hdr, Wv);
    };
    if (ctx->cfg.graph_mode != 0 || ctx->prof) {
        launch_wave();
    } else {
        uint32_t kinds = 0;
        for (int k : plan.sdf_kinds) kinds |= 1u << (k & 7);
        const uint64_t key = (uint64_t)plan.width | (uint64_t)(plan.stack_need <= 32) << 4 |
                             (uint64_t)plan.small_nodes << 5 | (uint64_t)plan.any_sdf << 6 | (uint64_t)kinds << 8 |
                             (uint64_t)depth << 16 | (uint64_t)grid << 32;
        auto it = cs->graphs.find(key);
        if (it == cs->graphs.end()) {
            hipGraph_t g = nullptr;
            hipGraphExec_t x = nullptr;
            e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
            launch_wave();
            e = hipStreamEndCapture(ctx->stream, &g);
            if (e == hipSuccess) e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) return hip_fail(ctx, e, "batched wave graph capture");
            it = cs->graphs.emplace(key, x).first;
        }
        e = hipGraphLaunch(it->second, ctx->stream);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipGraphLaunch");
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "batched dispatch launch");
    return PT_OK;
}

int ext_dispatch_features(pt_ctx* ctx, const pt_scene* s, const pt_push_constants* pc, float* feat, int* nsamp) {
    ExtSink sink;
    sink.pc = pc;
    sink.feat = feat;
    sink.nsamp = nsamp;
    return ext_run(ctx, s, pc, sink);
}

int ext_rel_error(pt_ctx* ctx, int P, const float* sum, const float* sumsq, const int* nsamp, double* d_out) {
    ExtCtxState* cs = static_cast<ExtCtxState*>(ctx->ext_state);
    const int cu = cs && cs->cu ? cs->cu : 256;
    const int grid = std::max(1, std::min((P + kBlk - 1) / kBlk, cu * 4));
    launch(ctx, "k_ext_rel_error", k_ext_rel_error, dim3(grid), dim3(kBlk), P, sum, sumsq, nsamp, d_out);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "relative error launch");
    return PT_OK;
}

}  // namespace ptk
