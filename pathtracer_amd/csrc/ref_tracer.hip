// ref_tracer.hip — PT_MODE_REF: the reference's spectral path tracer
// (src/shader.comp) as a wavefront of gfx950 kernels.
//
// The reference runs one invocation per pixel that loops over samples and
// bounces (src/shader.comp:1393-1533). Here one wave holds pixels
// [p0, p0 + pc) x samples [k0, k0 + m) as independent paths whose state lives
// in SoA arrays; every bounce runs
//     intersect (analytic closest hit) -> march<kind> (SDF sphere tracing)
//     -> shade (emission, BRDF sampling, NEE setup, Russian roulette)
//     -> shadow (closest-hit visibility with object-ID match) [-> march -> resolve]
// over a compacted queue of live paths, appended with wave64 ballots. The
// sequential PCG chain travels in the path state, so every path draws the
// same numbers in the same order as the shader, and each path's arithmetic is
// the shader's (same operations, same order, ptmath.h built-ins). Radiance is
// folded per pixel in sample order, so the texels match the reference method
// bit for bit wherever the operations are the same.
#include <algorithm>
#include <cstring>
#include <mutex>

#include "tracer.hpp"

namespace ptk {

using namespace ptd;

constexpr float kPi = 3.14159265358979f;  // PI, src/shader.comp:9 (binary32)
constexpr float kOneByThree = 0.3333333f; // ONEBYTHREE, src/shader.comp:10
constexpr int kBlock = 256;

__device__ __forceinline__ float ubo_at(const float* u, int k) { return (k >= 0 && k < kUboFloats) ? u[k] : 0.0f; }

// ---- spectral helpers (src/shader.comp:129-140, 960-974, 1030-1091) ----------------
__device__ __forceinline__ f4 sample_wavelengths(float lh) {
    float base = lh - 390.0f;
    return f4{390.0f + ptm_mod(base + (0.25f * 1.0f) * 330.0f, 330.0f),
              390.0f + ptm_mod(base + (0.25f * 2.0f) * 330.0f, 330.0f),
              390.0f + ptm_mod(base + (0.25f * 3.0f) * 330.0f, 330.0f),
              390.0f + ptm_mod(base + (0.25f * 4.0f) * 330.0f, 330.0f)};
}
__device__ __forceinline__ f3 wave_to_xyz(const float* u, float wave) {
    f3 r = mk3(0.0f, 0.0f, 0.0f);
    if (wave >= 360.0f && wave <= 800.0f) {
        int i3 = 3 * (int)(floorf(wave) - 360.0f);
        float t = wave - floorf(wave);
        int b = kUboCie + i3;
        r.x = ptm_mix(ubo_at(u, b), ubo_at(u, b + 3), t);
        r.y = ptm_mix(ubo_at(u, b + 1), ubo_at(u, b + 4), t);
        r.z = ptm_mix(ubo_at(u, b + 2), ubo_at(u, b + 5), t);
    }
    return r;
}
__device__ __forceinline__ float spd(float l, float peak, float sig, int invert) {
    float x = (l - peak) / (2.0f * sig * sig);
    float r = ptm_exp(-x * x);
    return ptm_mix(r, 1.0f - r, (float)invert);
}
__device__ __forceinline__ float planck(float l, float T) {
    return (1.1910429724e-16f * ptm_pow(l, -5.0f)) / (ptm_exp(0.014387768775f / (l * T)) - 1.0f);
}
__device__ __forceinline__ f4 emit(f4 l, float T, float lum) {
    float t = fmaxf(T, 0.0f);
    float peak = 4.0956746759e-6f * ptm_pow(t, 5.0f);
    float L = fmaxf(lum, 0.0f);
    return f4{(planck(l.x * 1e-9f, t) / peak) * L, (planck(l.y * 1e-9f, t) / peak) * L,
              (planck(l.z * 1e-9f, t) / peak) * L, (planck(l.w * 1e-9f, t) / peak) * L};
}
__host__ __device__ __forceinline__ float bk7(float l) {
    l *= 1e-3f;
    float l2 = l * l;
    float n2 = 1.0f;
    n2 += (1.03961212f * l2) / (l2 - 6.00069867e-3f);
    n2 += (0.231792344f * l2) / (l2 - 2.00179144e-2f);
    n2 += (1.01046945f * l2) / (l2 - 1.03560653e2f);
    return sqrtf(n2);
}
struct RMat { float peak, sigma, invert; };
__device__ __forceinline__ RMat material_mix(const float* u, float id) {
    int a = (int)floorf(id), b = (int)ceilf(id);
    float x = id - floorf(id);
    RMat r;
    r.peak = ptm_mix(ubo_at(u, kUboMat + 3 * a), ubo_at(u, kUboMat + 3 * b), x);
    r.sigma = ptm_mix(ubo_at(u, kUboMat + 3 * a + 1), ubo_at(u, kUboMat + 3 * b + 1), x);
    r.invert = ptm_mix(ubo_at(u, kUboMat + 3 * a + 2), ubo_at(u, kUboMat + 3 * b + 2), x);
    return r;
}
__device__ __forceinline__ void light_mix(const float* u, float id, float& T, float& lum) {
    int idx = (int)id;
    if (idx == -1) { T = 5500.0f; lum = 0.0f; return; }
    T = ubo_at(u, kUboLight + 2 * idx);
    lum = ubo_at(u, kUboLight + 2 * idx + 1);
}
__device__ __forceinline__ f4 brdf(f4 l, const RMat& m) {
    int inv = (int)m.invert;
    return f4{spd(l.x, m.peak, m.sigma, inv) / kPi, spd(l.y, m.peak, m.sigma, inv) / kPi,
              spd(l.z, m.peak, m.sigma, inv) / kPi, spd(l.w, m.peak, m.sigma, inv) / kPi};
}

// ---- analytic primitives (src/shader.comp:263-448, 450-541, 633-679) ----------------
struct Hit {
    float t;
    f3 n;
    float mat, light;
};

__device__ __forceinline__ bool bounding_sphere(f3 o, f3 d, f3 pos, float r2) {
    f3 lo = o - pos;
    float b = dot(d, lo);
    float c = dot(lo, lo) - r2;
    if (b * b < c) return false;
    if (b >= 0.0f && c >= 0.0f) return false;
    return true;
}

__device__ __forceinline__ bool hit_sphere(f3 o, f3 d, const RefSphere& s, Hit& h) {
    f3 lo = o - mk3(s.px, s.py, s.pz);
    float b = 2.0f * dot(d, lo);
    float c = dot(lo, lo) - s.r * s.r;
    float disc = b * b - 4.0f * c;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t1 = (-b - sq) * 0.5f;
    float t2 = (-b + sq) * 0.5f;
    float t = (t1 > 0.0f) ? t1 : t2;
    float outside = (t1 > 0.0f) ? 1.0f : -1.0f;
    if (t < 1e-4f) return false;
    if (t < h.t) {
        h.t = t;
        h.n = normalize(fma3s(d, t, lo) * outside);
        h.mat = s.mat;
        h.light = s.light;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool hit_plane(f3 o, f3 d, const RefPlane& p, Hit& h) {
    f3 lo = o - mk3(p.px, p.py, p.pz);
    float t = -lo.y / d.y;
    if (t < 1e-4f) return false;
    if (t < h.t) {
        h.t = t;
        float facing = 0.0f * d.x + 1.0f * d.y + 0.0f * d.z;  // faceforward(+Y, I, +Y)
        h.n = facing < 0.0f ? mk3(0.0f, 1.0f, 0.0f) : mk3(-0.0f, -1.0f, -0.0f);
        h.mat = p.mat;
        h.light = p.light;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool hit_box(f3 o, f3 d, const RefBox& b, Hit& h) {
    f3 lo = vmul(o - b.pos, b.M);
    f3 dl = vmul(d, b.M);
    f3 inv = mk3(1.0f / dl.x, 1.0f / dl.y, 1.0f / dl.z);
    f3 tmin = fma3(b.size, mk3(-0.5f, -0.5f, -0.5f), -lo) * inv;
    f3 tmax = fma3(b.size, mk3(0.5f, 0.5f, 0.5f), -lo) * inv;
    f3 lo3 = min3(tmin, tmax), hi3 = max3(tmin, tmax);
    float t1 = fmaxf(fmaxf(lo3.x, lo3.y), lo3.z);
    float t2 = fminf(fminf(hi3.x, hi3.y), hi3.z);
    float t = (t1 < 0.0f) ? t2 : t1;
    if (t1 > t2 || t < 1e-4f) return false;
    if (t < h.t) {
        h.t = t;
        f3 p = abs3((lo + dl * t) / b.size);
        float m = fmaxf(fmaxf(p.x, p.y), p.z);
        f3 st = mk3(ptm_step(m, p.x), ptm_step(m, p.y), ptm_step(m, p.z));
        f3 sg = -mk3(ptm_sign(dl.x), ptm_sign(dl.y), ptm_sign(dl.z));
        h.n = mvmul(b.M, st * sg);
        h.mat = b.mat;
        h.light = b.light;
        return true;
    }
    return false;
}

// One spherical cap of a lens (SphereSliceIntersection, src/shader.comp:366-417).
__device__ __forceinline__ bool hit_slice(f3 o, f3 d, f3 pos, const m3& M, float radius, float slice, bool first,
                                          float slicePos, bool sideInvert, float mat, float light, Hit& h,
                                          int& isOutside) {
    float off = radius - slice;
    f3 lo = vmul(o - pos, M);
    if (first) lo.x += slicePos - slice - off;
    else lo.x -= slicePos - slice - off;
    f3 dl = vmul(d, M);
    float b = 2.0f * dot(dl, lo);
    float c = dot(lo, lo) - radius * radius;
    float disc = b * b - 4.0f * c;
    float t = 1e6f;
    int out = 1;
    if (disc < 0.0f) return false;
    float sq = sqrtf(disc);
    float t1 = (-b - sq) * 0.5f;
    float t2 = (-b + sq) * 0.5f;
    if (first) {
        t1 = (fmaf(dl.x, t1, lo.x) > -off) ? 1e6f : t1;
        t2 = (fmaf(dl.x, t2, lo.x) > -off) ? 1e6f : t2;
    } else {
        t1 = (fmaf(dl.x, t1, lo.x) < off) ? 1e6f : t1;
        t2 = (fmaf(dl.x, t2, lo.x) < off) ? 1e6f : t2;
    }
    t = (t1 > 0.0f) ? t1 : t;
    if (t2 < t) { t = t2; out = -1; }
    if (t < 1e-4f) return false;
    if (t < h.t) {
        h.t = t;
        h.n = mvmul(M, normalize(fma3s(dl, t, lo) * (float)out));
        isOutside = (!sideInvert) ? out : -out;
        h.mat = mat;
        h.light = light;
        return true;
    }
    return false;
}

__device__ __forceinline__ bool hit_lens(f3 o, f3 d, const RefLens& L, Hit& h, int& isOutside) {
    bool any = false;
    for (int i = 0; i < 2; ++i)
        if (hit_slice(o, d, L.pos, L.M, L.twoF, L.half, i == 0, L.slicePos, !L.conv, L.mat, L.light, h, isOutside))
            any = true;
    return any;
}

__device__ __forceinline__ float eval_cubic(float a, float b, float c, float d, float x) {
    return x * (x * (x * a + b) + c) + d;
}
__device__ __forceinline__ float eval_quad(float a, float b, float c, float x) { return x * (x * a + b) + c; }
__device__ __forceinline__ float eval_quartic(float a, float b, float c, float d, float e, float x) {
    return x * (x * (x * (x * a + b) + c) + d) + e;
}
// SolveCubic (src/shader.comp:474-503); only the first root is used by the caller.
__device__ float cubic_first_root(float b, float c, float d) {
    float bd3 = b * kOneByThree;
    float Q = c * kOneByThree - bd3 * bd3;
    float R = 0.5f * bd3 * c - bd3 * bd3 * bd3 - 0.5f * d;
    float D = Q * Q * Q + R * R;
    float x;
    if (D > 0.0f) {
        float u = R + sqrtf(D);
        float v = R - sqrtf(D);
        float S = ptm_sign(u) * ptm_pow(fabsf(u), kOneByThree);
        float T = ptm_sign(v) * ptm_pow(fabsf(v), kOneByThree);
        x = S + T - bd3;
    } else {
        float sq = sqrtf(-Q);
        float th = ptm_acos(R / (sq * sq * sq)) * kOneByThree;
        x = (2.0f * sq) * ptm_cos(th) - bd3;
    }
    for (int i = 0; i < 2; ++i) x -= eval_cubic(1.0f, b, c, d, x) / eval_quad(3.0f, 2.0f * b, c, x);
    return x;
}
// SolveQuartic (src/shader.comp:505-541); returns a bit mask of real roots.
__device__ int solve_quartic(float a, float b, float c, float d, float e, float r[4]) {
    float inva = 1.0f / a;
    float inva2 = inva * 0.5f;
    float inva2a2 = inva2 * inva2;
    float bb = b * b;
    float p = -1.5f * bb * inva2a2 + c * inva;
    float q = bb * b * inva2a2 * inva2 - b * c * inva * inva2 + d * inva;
    float rr = -0.1875f * bb * bb * inva2a2 * inva2a2 + 0.5f * c * bb * inva2a2 * inva2 - b * d * inva2a2 + e * inva;
    float s = cubic_first_root(0.5f * -p, -rr, 0.5f * p * rr - 0.125f * q * q);
    float s2p = 2.0f * s - p;
    if (s2p < 0.0f) return 0;
    float is2p = -2.0f * s - p;
    float sq = sqrtf(s2p);
    float q2 = 2.0f * q / sq;
    float add = is2p + q2, sub = is2p - q2;
    float bd4a = 0.25f * inva * b;
    int mask = 0;
    if (add >= 0.0f) {
        float sa = sqrtf(add);
        r[0] = 0.5f * (-sq + 1.0f * sa) - bd4a;
        r[1] = 0.5f * (-sq + -1.0f * sa) - bd4a;
        for (int k = 0; k < 2; ++k)
            r[k] -= eval_quartic(a, b, c, d, e, r[k]) / eval_cubic(4.0f * a, 3.0f * b, 2.0f * c, d, r[k]);
        mask |= 3;
    }
    if (sub >= 0.0f) {
        float ss = sqrtf(sub);
        r[2] = 0.5f * (sq + 1.0f * ss) - bd4a;
        r[3] = 0.5f * (sq + -1.0f * ss) - bd4a;
        for (int k = 2; k < 4; ++k)
            r[k] -= eval_quartic(a, b, c, d, e, r[k]) / eval_cubic(4.0f * a, 3.0f * b, 2.0f * c, d, r[k]);
        mask |= 12;
    }
    return mask;
}

__device__ bool hit_cyclide(f3 ro, f3 rd, const RefCyclide& C, Hit& h) {
    f3 ol = vmul(ro - C.pos, C.M) / C.scale;
    f3 dl = vmul(rd, C.M) / C.scale;
    f3 o = mk3(ol.x, ol.z, ol.y);
    f3 d = mk3(dl.x, dl.z, dl.y);
    const float A = C.a, B = C.b, Cc = C.c, D = C.d;
    f3 dd = d * d, oo = o * o, od = o * d;
    f3 dyzx = mk3(d.y, d.z, d.x), dzxy = mk3(d.z, d.x, d.y);
    f3 oyzx = mk3(o.y, o.z, o.x), ozxy = mk3(o.z, o.x, o.y);
    float a4 = dot(dd, dd) + 2.0f * dot(dd, dyzx * dyzx);
    float a3 = 4.0f * (dot(o, d * d * d) + dot(od, dyzx * dyzx) + dot(od, dzxy * dzxy));
    float a2 = 6.0f * dot(oo, dd) + 8.0f * dot(od, oyzx * dyzx) + 2.0f * (dot(oo, dyzx * dyzx) + dot(oo, dzxy * dzxy)) +
               2.0f * (B * B - D * D) * dot(d, d) - 4.0f * (A * A * d.x * d.x + B * B * d.y * d.y);
    float a1 = 4.0f * (dot(o * o * o, d) + dot(oo, oyzx * dyzx) + dot(oo, ozxy * dzxy) + 2.0f * A * Cc * D * d.x +
                       (B * B - D * D) * dot(o, d) - 2.0f * (A * A * o.x * d.x + B * B * o.y * d.y));
    float a0 = dot(oo, oo) + 2.0f * dot(oo, oyzx * oyzx) + B * B * B * B + D * D * D * D - 2.0f * B * B * D * D -
               4.0f * Cc * Cc * D * D + 8.0f * A * Cc * D * o.x + 2.0f * (B * B - D * D) * dot(o, o) -
               4.0f * (A * A * o.x * o.x + B * B * o.y * o.y);
    float r[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int real = solve_quartic(a4, a3, a2, a1, a0, r);
    float t = 1e6f;
    for (int i = 0; i < 4; ++i)
        if (((real >> i) & 1) && r[i] < t && r[i] > 0.0f) t = r[i];
    if (t < h.t) {
        h.t = t;
        float x = o.x + d.x * t, y = o.y + d.y * t, z = o.z + d.z * t;
        float term1 = x * x + y * y + z * z + B * B - D * D;
        f3 n = mk3(4.0f * (x * term1 - 2.0f * A * (A * x - Cc * D)), 4.0f * z * term1,
                   4.0f * y * (term1 - 2.0f * B * B));
        h.n = normalize(n);
        h.mat = C.mat;
        h.light = C.light;
        return true;
    }
    return false;
}

// Closest analytic hit over spheres, planes, boxes, lenses, cyclides in the
// reference's order with strict '<' (src/shader.comp:862-922, 1131-1201).
// Returns the object index in packing order, or -1.
__device__ int closest_analytic(const RefSceneDev& S, f3 o, f3 d, Hit& h) {
    int id = -1, base = 0;
    for (int i = 0; i < S.ns; ++i)
        if (hit_sphere(o, d, S.sph[i], h)) id = i;
    base += S.ns;
    for (int i = 0; i < S.np; ++i)
        if (hit_plane(o, d, S.pln[i], h)) id = base + i;
    base += S.np;
    for (int i = 0; i < S.nb; ++i) {
        const RefBox& b = S.box[i];
        if (!bounding_sphere(o, d, b.pos, b.bsr2)) continue;
        if (hit_box(o, d, b, h)) id = base + i;
    }
    base += S.nb;
    for (int i = 0; i < S.nl; ++i) {
        const RefLens& L = S.lens[i];
        if (!bounding_sphere(o, d, L.pos, L.br2)) continue;
        int outside = 1;
        if (hit_lens(o, d, L, h, outside)) id = base + i;
    }
    base += S.nl;
    for (int i = 0; i < S.nc; ++i) {
        const RefCyclide& C = S.cyc[i];
        if (!bounding_sphere(o, d, C.pos, C.brad2)) continue;
        if (hit_cyclide(o, d, C, h)) id = base + i;
    }
    return id;
}

// ---- SDFs (sdfs/{mandelbulb,menger,blob,terrain}.glsl) -----------------------------
__device__ __forceinline__ float de_mandelbulb(f3 p, int iters, float bailout) {
    f3 z = p;
    float r = length(z);
    float dr = 1.0f;
    for (int i = 0; i < iters; ++i) {
        dr = 8.0f * ptm_pow(r, 7.0f) * dr + 1.0f;
        float invr = 1.0f / r;
        float invl = ptm_rsqrt(z.x * z.x + z.z * z.z);
        r *= r;
        r *= r;
        r *= r;
        float cost = z.z * invl, cosp = z.y * invr, sint = z.x * invl, sinp = invr / invl;
        float ct2 = cost * cost, cp2 = cosp * cosp;
        float cost8 = fmaf(fmaf(fmaf(fmaf(128.0f, ct2, -256.0f), ct2, 160.0f), ct2, -32.0f), ct2, 1.0f);
        float cosp8 = fmaf(fmaf(fmaf(fmaf(128.0f, cp2, -256.0f), cp2, 160.0f), cp2, -32.0f), cp2, 1.0f);
        float sint8 = fmaf(fmaf(fmaf(128.0f, ct2, -192.0f), ct2, 80.0f), ct2, -8.0f) * sint * cost;
        float sinp8 = fmaf(fmaf(fmaf(128.0f, cp2, -192.0f), cp2, 80.0f), cp2, -8.0f) * sinp * cosp;
        z = r * mk3(sint8 * sinp8, cosp8, cost8 * sinp8);
        z = z + p;
        r = length(z);
        if (r > bailout) break;
    }
    return 0.5f * ptm_log(r) * r / dr;
}
__device__ __forceinline__ float de_menger(f3 p, int levels) {
    f3 q = abs3(p) - mk3(1.0f, 1.0f, 1.0f);
    float d = length(max3(q, mk3(0.0f, 0.0f, 0.0f))) + fminf(fmaxf(q.x, fmaxf(q.y, q.z)), 0.0f);
    float s = 1.0f;
    for (int m = 0; m < levels; ++m) {
        f3 ps = p * s;
        f3 a = mk3(ptm_mod(ps.x, 2.0f) - 1.0f, ptm_mod(ps.y, 2.0f) - 1.0f, ptm_mod(ps.z, 2.0f) - 1.0f);
        s *= 3.0f;
        f3 aa = abs3(a);
        f3 r = abs3(mk3(1.0f - 3.0f * aa.x, 1.0f - 3.0f * aa.y, 1.0f - 3.0f * aa.z));
        float da = fmaxf(r.x, r.y), db = fmaxf(r.y, r.z), dc = fmaxf(r.z, r.x);
        float c = (fminf(da, fminf(db, dc)) - 1.0f) / s;
        d = fmaxf(d, c);
    }
    return d;
}
__device__ __forceinline__ float blob_sphere(f3 p, int k) {
    if (k == 0) return length(p) - 1.0f;
    if (k == 1) return length(mk3(p.x, p.y - 1.0f, p.z)) - 0.7f;
    if (k == 2) return length(mk3(p.x - 0.25f, p.y - 1.25f, p.z - 0.5f)) - 0.25f;
    return length(mk3(p.x + 0.25f, p.y - 1.25f, p.z - 0.5f)) - 0.25f;
}
// smin (src/shader.comp:685-702), after iquilezles.org/articles/smin
__device__ __forceinline__ float smin1(float x, float y) {
    float k = 0.02f;
    k *= 6.0f;
    float h = fmaxf(k - fabsf(x - y), 0.0f) / k;
    float m = h * h * h * 0.5f;
    float s = m * k * kOneByThree;
    return fminf(x, y) - s;
}
__device__ __forceinline__ void smin2(float& xd, float& xm, float yd, float ym) {
    float k = 0.02f;
    k *= 6.0f;
    float h = fmaxf(k - fabsf(xd - yd), 0.0f) / k;
    float m = h * h * h * 0.5f;
    float s = m * k * kOneByThree;
    if (xd < yd) { xd = xd - s; xm = xm + (ym - xm) * m; }
    else { xd = yd - s; xm = xm + (ym - xm) * (1.0f - m); }
}
__device__ __forceinline__ float de_blob(f3 p) {
    return smin1(smin1(smin1(blob_sphere(p, 0), blob_sphere(p, 1)), blob_sphere(p, 2)), blob_sphere(p, 3));
}
__device__ __forceinline__ float terrain_height(float x) {
    return ptm_sin(0.0625f * x) + ptm_sin(0.125f * (x + 10.0f)) + 0.25f * ptm_sin(0.25f * x) +
           0.125f * ptm_sin(0.5f * x) + 0.25f * ptm_sin(x) + 0.25f * ptm_sin(2.0f * x + 1.0f) +
           0.125f * ptm_sin(4.0f * x + 2.0f) + 0.0625f * ptm_sin(8.0f * x + 1.0f) +
           0.03125f * ptm_sin(16.0f * x + 5.0f) + 0.015625f * ptm_sin(32.0f * x) + 0.0078125f * ptm_sin(64.0f * x);
}
__device__ __forceinline__ float de_terrain(f3 p) { return p.y - (terrain_height(p.x) + terrain_height(p.z)); }

template <int KIND>
__device__ __forceinline__ float de_eval(const pt_sdf_desc& d, f3 p) {
    const int kind = KIND ? KIND : d.kind;
    switch (kind) {
        case PT_SDF_MANDELBULB: return de_mandelbulb(p, d.iterations, d.bailout);
        case PT_SDF_MENGER: return de_menger(p, d.iterations);
        case PT_SDF_BLOB: return de_blob(p);
        case PT_SDF_TERRAIN: return de_terrain(p);
        default: return kMaxDist;
    }
}
__device__ __forceinline__ float sdf_material(const pt_sdf_desc& d, f3 p) {
    if (d.material_kind == PT_SDFMAT_RADIAL) {
        float f = dot(p, p);
        return ptm_mix(d.material[0], d.material[1], f / (d.material[2] + f));
    }
    if (d.material_kind == PT_SDFMAT_BLOB) {
        float xd = blob_sphere(p, 0), xm = d.material[0];
        smin2(xd, xm, blob_sphere(p, 1), d.material[1]);
        smin2(xd, xm, blob_sphere(p, 2), d.material[2]);
        smin2(xd, xm, blob_sphere(p, 3), d.material[3]);
        return xm;
    }
    return d.material[0];
}

// The spliced SDF(): min over the SDFs of the set (src/pathtracer.cpp:2004-2054).
template <int KIND>
__device__ __forceinline__ float sdf_set(const RefSceneDev& S, f3 p, uint32_t set, unsigned& steps) {
    float s = kMaxDist;
    for (int i = 0; i < S.nsdf; ++i) {
        uint32_t code = 1u << (uint32_t)i;
        if ((set & code) == code) s = fminf(s, de_eval<KIND>(S.sdf[i].d, p - S.sdf[i].pos));
    }
    ++steps;
    return s;
}
// The spliced SDFMATERIAL(): material of the nearest SDF, ties to the later one.
template <int KIND>
__device__ float sdf_set_material(const RefSceneDev& S, f3 p, uint32_t set) {
    float s = kMaxDist, m = 0.0f;
    for (int i = 0; i < S.nsdf; ++i) {
        uint32_t code = 1u << (uint32_t)i;
        if ((set & code) == code) {
            f3 q = p - S.sdf[i].pos;
            float di = de_eval<KIND>(S.sdf[i].d, q);
            m = (s < di) ? m : sdf_material(S.sdf[i].d, q);
            s = fminf(s, di);
        }
    }
    return m;
}
// SearchSDF (src/shader.comp:732-777).
__device__ bool search_sdf(const RefSceneDev& S, f3 p, f3 invdir, float& tmn, float& tmx, uint32_t& set) {
    bool found = false;
    set = 0;
    for (int i = 0; i < S.nsdf; ++i) {
        f3 lo = p - S.sdf[i].pos;
        f3 t0 = fma3(S.sdf[i].size, mk3(-0.5f, -0.5f, -0.5f), -lo) * invdir;
        f3 t1 = fma3(S.sdf[i].size, mk3(0.5f, 0.5f, 0.5f), -lo) * invdir;
        f3 a = min3(t0, t1), b = max3(t0, t1);
        float bx = fmaxf(fmaxf(a.x, a.y), a.z);
        float by = fminf(fminf(b.x, b.y), b.z);
        if (bx > by || by < 0.0f) continue;
        uint32_t bit = 1u << (uint32_t)i;
        if (bx < tmn) {
            found = true;
            if (by < tmn) { tmn = bx; tmx = by; set = bit; }
            else if (by < tmx) { tmn = bx; set += bit; }
            else { tmn = bx; tmx = by; set += bit; }
        } else if (bx < tmx) {
            found = true;
            if (by > tmx) tmx = by;
            set += bit;
        }
    }
    return found;
}
// SphereTracing (src/shader.comp:779-860): over-relaxed sphere tracing after
// Keinert et al., "Enhanced Sphere Tracing" (2014).
template <int KIND>
__device__ bool sphere_trace(const RefSceneDev& S, f3 o, f3 d, Hit& h, unsigned& steps) {
    float t = 1e-3f, insT = 0.0f;
    const float omegaMax = 1.70f, omegaSpeed = 0.20f;
    float omega = omegaMax, prev = 0.0f;
    f3 p = o;
    f3 invdir = mk3(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
    int points = 0;
    float tmn = kMaxDist, tmx = kMaxDist;
    uint32_t set = 0;
    if (!search_sdf(S, p, invdir, tmn, tmx, set)) return false;
    t = fmaxf(tmn, t);
    p = fma3s(d, t, o);
    float k = ptm_sign(sdf_set<KIND>(S, o, set, steps));
    for (int i = 0; i < 512; ++i) {
        float rad = sdf_set<KIND>(S, p, set, steps);
        if (insT > (fabsf(prev) + fabsf(rad))) {
            t -= insT;
            omega = 1.0f;
            insT = prev * omega * k;
            t += insT;
            p = fma3s(d, t, o);
            continue;
        }
        if (fabsf(rad) < 1e-4f) break;
        if (t > tmx) points += 1;
        else points = 0;
        if (points >= 2) {
            t = tmx + 1e-3f;
            tmn = kMaxDist;
            tmx = kMaxDist;
            if (search_sdf(S, fma3s(d, t, o), invdir, tmn, tmx, set)) {
                tmn += t;
                tmx += t;
                t = fmaxf(tmn, t);
                p = fma3s(d, t, o);
                continue;
            }
            return false;
        }
        insT = rad * omega * k;
        t += insT;
        p = fma3s(d, t, o);
        float f = fminf(rad / prev, 0.99f);
        omega += omegaSpeed * (fminf(1.0f / (1.0f - f), omegaMax) - omega);
        prev = rad;
    }
    if (t < h.t) {
        h.t = t - 1e-3f;
        p = fma3s(d, t, o);
        const float e = 1e-4f;
        f3 n = mk3(sdf_set<KIND>(S, p + mk3(e, 0.0f, 0.0f), set, steps) - sdf_set<KIND>(S, p - mk3(e, 0.0f, 0.0f), set, steps),
                   sdf_set<KIND>(S, p + mk3(0.0f, e, 0.0f), set, steps) - sdf_set<KIND>(S, p - mk3(0.0f, e, 0.0f), set, steps),
                   sdf_set<KIND>(S, p + mk3(0.0f, 0.0f, e), set, steps) - sdf_set<KIND>(S, p - mk3(0.0f, 0.0f, e), set, steps));
        h.n = normalize(n);
        h.mat = sdf_set_material<KIND>(S, p, set);
        h.light = -1.0f;
        return true;
    }
    return false;
}

// ---- sampling (src/shader.comp:976-1028, 1093-1119, 1225-1296) ----------------------
__device__ __forceinline__ f3 sample_unit_sphere(uint32_t& seed) {
    float rx = rand_chain(seed);
    float ry = rand_chain(seed);
    float phi = 2.0f * kPi * ry;
    float sinT = 2.0f * rx - 1.0f;
    float cosT = sqrtf(fmaf(-sinT, sinT, 1.0f));
    return mk3(ptm_cos(phi) * cosT, ptm_sin(phi) * cosT, sinT);
}
__device__ __forceinline__ f3 sample_cone(uint32_t& seed, float cosMax) {
    float rx = rand_chain(seed);
    float ry = rand_chain(seed);
    float cosA = 2.0f * cosMax * cosMax - 1.0f;
    float phi = 2.0f * kPi * ry;
    float ct = (1.0f - cosA) * rx + cosA;
    float st = sqrtf(fmaf(-ct, ct, 1.0f));
    return normalize(mk3(ptm_cos(phi) * st, ptm_sin(phi) * st, ct + 1.0f));
}
__device__ __forceinline__ f3 to_world(f3 v, f3 n) {
    f3 b1 = mk3(0.0f, -1.0f, 0.0f), b2 = mk3(-1.0f, 0.0f, 0.0f);
    if (n.z >= -0.9999999f) {
        float a = 1.0f / (1.0f + n.z);
        float b = -n.x * n.y * a;
        b1 = mk3(1.0f - (n.x * n.x * a), b, -n.x);
        b2 = mk3(b, 1.0f - (n.y * n.y * a), -n.y);
    }
    return b1 * v.x + b2 * v.y + n * v.z;
}
__device__ __forceinline__ float mis2(float p1, float p2) { return p1 * p1 / (p1 * p1 + p2 * p2); }

// SampleRandomLightSource (src/shader.comp:1225-1285), reading the packed
// objects like the shader, including its lens/cyclide count comparisons.
__device__ int sample_light(const RefSceneDev& S, uint32_t& seed, float& brad, f3& pos, float& lid) {
    const float* u = S.ubo;
    int rl = (int)floorf(rand_chain(seed) * S.numLights);
    int rid = (int)ubo_at(u, kUboLightId + rl);
    int off = kUboObj;
    auto P3 = [&](int k) { return mk3(ubo_at(u, k), ubo_at(u, k + 1), ubo_at(u, k + 2)); };
    auto ID = [&](int k) { return (float)((int)ubo_at(u, k) - 1); };
    if (rid < S.ns) {
        int k = off + 6 * rid;
        brad = ubo_at(u, k + 3); pos = P3(k); lid = ID(k + 5);
        return (int)ubo_at(u, kUboLightId + rl);
    }
    rid -= S.ns; off += 6 * S.ns;
    if (rid < S.np) {
        int k = off + 5 * rid;
        brad = 1e5f; pos = P3(k); lid = ID(k + 4);
        return (int)ubo_at(u, kUboLightId + rl);
    }
    rid -= S.np; off += 5 * S.np;
    if (rid < S.nb) {
        int k = off + 11 * rid;
        brad = 0.5f * length(P3(k + 6)); pos = P3(k); lid = ID(k + 10);
        return (int)ubo_at(u, kUboLightId + rl);
    }
    rid -= S.nb; off += 11 * S.nb;
    if (rid < S.nb) {
        int k = off + 12 * rid;
        float rad = ubo_at(u, k + 6), th = ubo_at(u, k + 8);
        brad = sqrtf((rad * rad) + (0.25f * th * th)); pos = P3(k); lid = ID(k + 11);
        return (int)ubo_at(u, kUboLightId + rl);
    }
    rid -= S.nl; off += 12 * S.nl;
    if (rid < S.nl) {
        int k = off + 16 * rid;
        brad = sqrtf(ubo_at(u, k + 13)); pos = P3(k); lid = ID(k + 15);
        return (int)ubo_at(u, kUboLightId + rl);
    }
    return 0;
}

// ---- kernels ---------------------------------------------------------------------------
__global__ void k_ref_counts(int* counts, int c0, int c1, int c2) {
    counts[0] = c0;
    counts[1] = c1;
    counts[2] = c2;
}
__global__ void k_ref_reset(int* counts, int which) {
    counts[which] = 0;
    counts[2] = 0;
}

// Camera rays: Rendering + Scene up to TracePath (src/shader.comp:1446-1476, 1509-1516).
__global__ void __launch_bounds__(kBlock) k_ref_raygen(RefFrame F, RefWave Wv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = F.pc * F.m;
    if (i == 0) atomicAdd(&Wv.stats[3], (unsigned long long)n);
    if (i >= n) return;
    int pix = F.p0 + i % F.pc;
    int k = F.k0 + i / F.pc;
    int gx = pix % F.W, gy = pix / F.W;
    uint32_t x = (uint32_t)gx, y = (uint32_t)(F.H - gy);
    float uvx = (2.0f * (float)x - (float)F.W) / (float)F.H;
    float uvy = (2.0f * (float)y - (float)F.H) / (float)F.H;
    uint32_t seed = pcg((uint32_t)(F.frame - F.spf + k));
    seed += x + (uint32_t)F.W * y;
    float jx = 2.0f * rand_chain(seed) - 0.5f;
    float jy = 2.0f * rand_chain(seed) - 0.5f;
    uvx += jx / (float)F.W;
    uvy += jy / (float)F.H;
    float sc = -F.camSize * 0.5f;
    uvx *= sc;
    uvy *= sc;
    f3 o = F.camPos + vmul(mk3(uvx, uvy, 0.0f), F.camM);
    float ux = rand_chain(seed), uy = rand_chain(seed);
    float phi = 2.0f * kPi * uy;
    float dr = sqrtf(ux);
    float dsx = dr * ptm_cos(phi), dsy = dr * ptm_sin(phi);
    float ap = 0.5f * F.apSize;
    f3 onAp = F.camPos + vmul(mk3(ap * dsx, ap * dsy, F.apDist), F.camM);
    f3 d = normalize(onAp - o);
    float lh = ptm_mix(360.0f, 800.0f, rand_chain(seed));
    // TracePathLens: two refractions through the BK7 bi-convex lens at the hero wavelength.
    float l = lh;
    for (int pass = 0; pass < 2; ++pass) {
        Hit h;
        h.t = 1e6f;
        h.n = mk3(0.0f, 0.0f, 0.0f);
        int outside = 1;
        for (int s = 0; s < 2; ++s)
            hit_slice(o, d, F.lensPos, F.lensM, F.lensTwoF, F.lensHalf, s == 0, F.lensSlicePos, false, 0.0f, -1.0f, h,
                      outside);
        float n1, n2;
        if (outside == 1) { n1 = 1.0f; n2 = bk7(l); }
        else { n1 = bk7(l); n2 = 1.0f; }
        float n12 = n1 / n2;
        l = l * n12;
        o = fma3s(d, h.t, o);
        d = refract(d, h.n, n12);
    }
    Wv.ox[i] = o.x; Wv.oy[i] = o.y; Wv.oz[i] = o.z;
    Wv.dx[i] = d.x; Wv.dy[i] = d.y; Wv.dz[i] = d.z;
    Wv.seed[i] = seed;
    Wv.lh[i] = lh;
    Wv.T[i] = make_float4(1.0f, 1.0f, 1.0f, 1.0f);
    Wv.L[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    Wv.w[i] = 1.0f;
    Wv.q[0][i] = i;
}

// Closest analytic hit of every live path (Intersection without SphereTracing).
__global__ void __launch_bounds__(kBlock) k_ref_intersect(RefSceneDev S, RefWave Wv, int cur) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int n = Wv.counts[cur];
    if (idx == 0) atomicAdd(&Wv.stats[0], (unsigned long long)n);
    if (idx >= n) return;
    int p = Wv.q[cur][idx];
    f3 o = mk3(Wv.ox[p], Wv.oy[p], Wv.oz[p]);
    f3 d = mk3(Wv.dx[p], Wv.dy[p], Wv.dz[p]);
    Hit h;
    h.t = kMaxDist;
    h.n = mk3(0.0f, 0.0f, 0.0f);
    h.mat = 0.0f;
    h.light = -1.0f;
    closest_analytic(S, o, d, h);
    Wv.ht[p] = h.t;
    Wv.hnx[p] = h.n.x; Wv.hny[p] = h.n.y; Wv.hnz[p] = h.n.z;
    Wv.hmat[p] = h.mat;
    Wv.hlight[p] = h.light;
}

// SphereTracing of the extension rays against the current closest hit.
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_ref_march(RefSceneDev S, RefWave Wv, int cur) {
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int n = Wv.counts[cur];
    unsigned steps = 0;
    if (idx < n) {
        int p = Wv.q[cur][idx];
        f3 o = mk3(Wv.ox[p], Wv.oy[p], Wv.oz[p]);
        f3 d = mk3(Wv.dx[p], Wv.dy[p], Wv.dz[p]);
        Hit h;
        h.t = Wv.ht[p];
        if (sphere_trace<KIND>(S, o, d, h, steps)) {
            Wv.ht[p] = h.t;
            Wv.hnx[p] = h.n.x; Wv.hny[p] = h.n.y; Wv.hnz[p] = h.n.z;
            Wv.hmat[p] = h.mat;
            Wv.hlight[p] = h.light;
        }
    }
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
}

// TraceRay after the hit (src/shader.comp:1352-1391) with SampleLightSource
// (src/shader.comp:1298-1343) up to the visibility test.
__global__ void __launch_bounds__(kBlock) k_ref_shade(RefSceneDev S, RefFrame F, RefWave Wv, int cur, int bounce) {
    const int nxt = cur ^ 1;
    int idx = blockIdx.x * blockDim.x + threadIdx.x;
    int n = Wv.counts[cur];
    bool active = idx < n;
    bool shadow = false, cont = false;
    int p = 0, lobj = 0;
    f3 no = mk3(0.0f, 0.0f, 0.0f), sd = no;
    f4 contrib = mk4(0.0f);
    if (active) {
        p = Wv.q[cur][idx];
        float t = Wv.ht[p];
        float4 Tv = Wv.T[p];
        f4 T = f4{Tv.x, Tv.y, Tv.z, Tv.w};
        float w = Wv.w[p];
        uint32_t seed = Wv.seed[p];
        f4 l = sample_wavelengths(Wv.lh[p]);
        RMat mat = material_mix(S.ubo, Wv.hmat[p]);
        float LT, LL;
        light_mix(S.ubo, Wv.hlight[p], LT, LL);
        if (t < kMaxDist) {
            if (LL > 0.0f) {
                f4 e = emit(l, LT, LL) * T * w;
                float4 Lp = Wv.L[p];
                Wv.L[p] = make_float4(Lp.x + e.x, Lp.y + e.y, Lp.z + e.z, Lp.w + e.w);
            } else {
                f3 o = mk3(Wv.ox[p], Wv.oy[p], Wv.oz[p]);
                f3 d = mk3(Wv.dx[p], Wv.dy[p], Wv.dz[p]);
                f3 nrm = mk3(Wv.hnx[p], Wv.hny[p], Wv.hnz[p]);
                no = fma3s(d, t, o);
                f3 nd = normalize(nrm + sample_unit_sphere(seed));
                float pdf = dot(nd, nrm) / kPi;
                if (S.numLights > 0.0f) {
                    float brad = 0.0f, lid = -1.0f;
                    f3 lpos = mk3(0.0f, 0.0f, 0.0f);
                    lobj = sample_light(S, seed, brad, lpos, lid);
                    float invd = 1.0f / length(lpos - no);
                    f3 ldir = (lpos - no) * invd;
                    float sinMax = fminf(brad * invd, 1.0f);
                    float cosMax = sqrtf(1.0f - sinMax * sinMax);
                    sd = to_world(sample_cone(seed, cosMax), ldir);
                    float lpdf = 1.0f / S.numLights;
                    lpdf *= dot(sd, ldir) / (kPi * (1.0f - cosMax * cosMax));
                    w = mis2(pdf, lpdf);
                    float ct = dot(sd, nrm);
                    float death = 1.25f * fmaxf(w - 0.2f, 0.0f);
                    if (ct >= 0.0f) {
                        if (rand_chain(seed) > death) {
                            float T2, L2;
                            light_mix(S.ubo, lid, T2, L2);
                            f4 f = brdf(l, mat);
                            f4 g = f4{f.x * ct / lpdf, f.y * ct / lpdf, f.z * ct / lpdf, f.w * ct / lpdf};
                            contrib = emit(l, T2, L2) * (T * g) * (1.0f - w);
                            shadow = true;
                        } else {
                            w = 1.0f;
                        }
                    }
                } else {
                    w = mis2(pdf, 0.0f);
                }
                float ct = dot(nd, nrm);
                f4 f = brdf(l, mat);
                T = T * f4{f.x * ct / pdf, f.y * ct / pdf, f.z * ct / pdf, f.w * ct / pdf};
                float pr = ptm_clamp(fmaxf(T.x, fmaxf(T.y, fmaxf(T.z, T.w))), 0.0f, 0.99f);
                if (!(rand_chain(seed) > pr)) {
                    T = T * (1.0f / pr);
                    Wv.ox[p] = no.x; Wv.oy[p] = no.y; Wv.oz[p] = no.z;
                    Wv.dx[p] = nd.x; Wv.dy[p] = nd.y; Wv.dz[p] = nd.z;
                    cont = bounce + 1 < F.pathLength;
                }
                Wv.T[p] = make_float4(T.x, T.y, T.z, T.w);
                Wv.w[p] = w;
                Wv.seed[p] = seed;
            }
        }
    }
    int slot = wave_append(&Wv.counts[2], shadow);
    if (shadow) {
        Wv.s_path[slot] = p;
        Wv.sox[slot] = no.x; Wv.soy[slot] = no.y; Wv.soz[slot] = no.z;
        Wv.sdx[slot] = sd.x; Wv.sdy[slot] = sd.y; Wv.sdz[slot] = sd.z;
        Wv.s_lobj[slot] = lobj;
        Wv.s_c[slot] = make_float4(contrib.x, contrib.y, contrib.z, contrib.w);
    }
    int nslot = wave_append(&Wv.counts[nxt], cont);
    if (cont) Wv.q[nxt][nslot] = p;
}

__device__ __forceinline__ void add_contrib(RefWave& Wv, int i) {
    int p = Wv.s_path[i];
    float4 c = Wv.s_c[i], Lp = Wv.L[p];
    Wv.L[p] = make_float4(Lp.x + c.x, Lp.y + c.y, Lp.z + c.z, Lp.w + c.w);
}

// LightSourceVisibilityCheck, analytic part (src/shader.comp:1121-1201). With
// no SDFs in the scene the visible contributions are added here directly.
__global__ void __launch_bounds__(kBlock) k_ref_shadow(RefSceneDev S, RefWave Wv, int resolve) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = Wv.counts[2];
    if (i == 0) atomicAdd(&Wv.stats[1], (unsigned long long)n);
    if (i >= n) return;
    f3 o = mk3(Wv.sox[i], Wv.soy[i], Wv.soz[i]);
    f3 d = mk3(Wv.sdx[i], Wv.sdy[i], Wv.sdz[i]);
    Hit h;
    h.t = kMaxDist;
    int id = closest_analytic(S, o, d, h);
    if (resolve) {
        if (id == Wv.s_lobj[i]) add_contrib(Wv, i);
    } else {
        Wv.s_t[i] = h.t;
        Wv.s_id[i] = id;
    }
}

// SphereTracing of the shadow rays: an SDF closer than the analytic hit sets
// the object index to -1 (src/shader.comp:1203-1205), then visibility resolves.
template <int KIND>
__global__ void __launch_bounds__(kBlock) k_ref_shadow_march(RefSceneDev S, RefWave Wv) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int n = Wv.counts[2];
    unsigned steps = 0;
    if (i < n) {
        f3 o = mk3(Wv.sox[i], Wv.soy[i], Wv.soz[i]);
        f3 d = mk3(Wv.sdx[i], Wv.sdy[i], Wv.sdz[i]);
        Hit h;
        h.t = Wv.s_t[i];
        int id = Wv.s_id[i];
        if (sphere_trace<KIND>(S, o, d, h, steps)) id = -1;
        if (id == Wv.s_lobj[i]) add_contrib(Wv, i);
    }
    wave_add_u64(stat_slot(Wv.stats, 2), steps);
}

// Per pixel, in sample order: XYZ of each path (src/shader.comp:1477-1487)
// added to the running sum (src/shader.comp:1513-1516).
__global__ void __launch_bounds__(kBlock) k_ref_accum(RefSceneDev S, RefFrame F, RefWave Wv) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned nans = 0;
    if (j < F.pc) {
        f3 acc = mk3(0.0f, 0.0f, 0.0f);
        if (F.k0 != 0) acc = mk3(Wv.sum[j], Wv.sum[F.pc + j], Wv.sum[2 * F.pc + j]);
        for (int s = 0; s < F.m; ++s) {
            int path = s * F.pc + j;
            float4 R = Wv.L[path];
            f4 l = sample_wavelengths(Wv.lh[path]);
            f3 X0 = wave_to_xyz(S.ubo, l.x), X1 = wave_to_xyz(S.ubo, l.y);
            f3 X2 = wave_to_xyz(S.ubo, l.z), X3 = wave_to_xyz(S.ubo, l.w);
            f3 col = R.x * X0 + R.y * X1 + R.z * X2 + R.w * X3;
            col = mk3(0.0f, 0.0f, 0.0f) + (col * (720.0f - 390.0f)) * 0.25f;
            if (col.x != col.x || col.y != col.y || col.z != col.z) {
                col = mk3(0.0f, 0.0f, 0.0f);
                ++nans;
            }
            acc = acc + col;
        }
        Wv.sum[j] = acc.x;
        Wv.sum[F.pc + j] = acc.y;
        Wv.sum[2 * F.pc + j] = acc.z;
    }
    wave_add_u64(stat_slot(Wv.stats, 4), nans);
}

struct RefFinish {
    int p0, pc, spf, currentSamples, frame, ISO;
    float apSize, FPS, persistence;
};
// Exposure and Accumulate into the texel buffer (src/shader.comp:1492-1532).
__global__ void __launch_bounds__(kBlock) k_ref_finish(RefFinish F, const float* sum, float4* texels) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= F.pc) return;
    f3 out = mk3(sum[j], sum[F.pc + j], sum[2 * F.pc + j]);
    out = out / (float)F.spf;
    out = out * (F.apSize * F.apSize * (float)F.ISO);
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

// ---- host side ---------------------------------------------------------------------------
template <typename T>
static T* carve(char*& p, size_t n) {
    T* r = reinterpret_cast<T*>(p);
    p += ((n * sizeof(T) + 255) / 256) * 256;
    return r;
}

constexpr size_t round256(size_t bytes) { return ((bytes + 255) / 256) * 256; }
// Bytes of a wave of `cap` paths as carve_ref_wave lays it out.
constexpr size_t ref_wave_bytes(int cap) {
    const size_t f = (size_t)cap;
    return 6 * round256(f * 4)                                            // ray
           + 2 * round256(f * 4)                                          // seed, lh
           + 2 * round256(f * 16) + round256(f * 4)                       // T, L, w
           + 6 * round256(f * 4)                                          // hit
           + 2 * round256(f * 4)                                          // queues
           + round256(f * 4)                                              // s_path
           + 6 * round256(f * 4)                                          // shadow ray
           + round256(f * 4) + round256(f * 16) + 2 * round256(f * 4)     // s_lobj, s_c, s_t, s_id
           + round256(f * 12)                                             // sums
           + round256(64);                                                // counts
}

// Lays a wave of `cap` paths out over `base` (ref_wave_bytes(cap) bytes).
static void carve_ref_wave(char* p, int cap, unsigned long long* stats, RefWave& W) {
    size_t n = (size_t)cap;
    W.ox = carve<float>(p, n); W.oy = carve<float>(p, n); W.oz = carve<float>(p, n);
    W.dx = carve<float>(p, n); W.dy = carve<float>(p, n); W.dz = carve<float>(p, n);
    W.seed = carve<uint32_t>(p, n); W.lh = carve<float>(p, n);
    W.T = carve<float4>(p, n); W.L = carve<float4>(p, n); W.w = carve<float>(p, n);
    W.ht = carve<float>(p, n); W.hnx = carve<float>(p, n); W.hny = carve<float>(p, n); W.hnz = carve<float>(p, n);
    W.hmat = carve<float>(p, n); W.hlight = carve<float>(p, n);
    W.q[0] = carve<int>(p, n); W.q[1] = carve<int>(p, n);
    W.s_path = carve<int>(p, n);
    W.sox = carve<float>(p, n); W.soy = carve<float>(p, n); W.soz = carve<float>(p, n);
    W.sdx = carve<float>(p, n); W.sdy = carve<float>(p, n); W.sdz = carve<float>(p, n);
    W.s_lobj = carve<int>(p, n); W.s_c = carve<float4>(p, n); W.s_t = carve<float>(p, n); W.s_id = carve<int>(p, n);
    W.sum = carve<float>(p, 3 * n);
    W.counts = carve<int>(p, 16);
    W.stats = stats;
    W.cap = cap;
}

static int ensure_ref_wave(pt_ctx* ctx, int cap) {
    if (ctx->ref_wave.cap >= cap && ctx->ref_mem.p) return PT_OK;
    dev_free(ctx, ctx->ref_mem);
    int st = dev_alloc(ctx, ctx->ref_mem, ref_wave_bytes(cap));
    if (st != PT_OK) return st;
    carve_ref_wave(static_cast<char*>(ctx->ref_mem.p), cap, ctx->d_stats, ctx->ref_wave);
    return PT_OK;
}

// ---- the allocation-free dispatch's wave --------------------------------------------
// A wave of kStaticWavePaths paths in a static device array: pt_dispatch_noalloc
// traces through it instead of allocating. Each device holds its own copy for
// the life of the process. Every launch writes the records it reads (queue
// counts by k_ref_counts, sums on the first sample chunk), so nothing needs
// clearing between calls; calls that overlap on one device race on it.
constexpr int kStaticWavePaths = 1 << 16;
constexpr size_t kStaticWaveBytes = ref_wave_bytes(kStaticWavePaths);  // 11,010,304 B
static __device__ __align__(256) char scratch[kStaticWaveBytes];
constexpr int kMaxStaticDevices = 64;
static std::mutex g_scratch_mu;
static char* g_scratch[kMaxStaticDevices] = {};

int ref_static_setup(pt_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxStaticDevices) {
        ctx->err = "static dispatch wave: device ordinal beyond 63";
        return PT_ERR_CAPACITY;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[ctx->device]) return PT_OK;
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) return hip_fail(ctx, e != hipSuccess ? e : hipErrorInvalidSymbol, "static dispatch wave");
    g_scratch[ctx->device] = static_cast<char*>(ptr);
    return PT_OK;
}

void ref_wave_release(pt_ctx* ctx) {
    dev_free(ctx, ctx->ref_mem);
    ctx->ref_wave = RefWave{};
}

int ref_scene_upload(pt_ctx* ctx, pt_scene* s) {
    const float* u = reinterpret_cast<const float*>(&s->ubo);
    auto U = [&](int k) { return (k >= 0 && k < kUboFloats) ? u[k] : 0.0f; };
    auto ID = [&](int k) { return (float)((int)U(k) - 1); };
    int ns = (int)U(0), np = (int)U(1), nb = (int)U(2), nl = (int)U(3), nc = (int)U(4), nsdf = (int)U(5);
    if (ns < 0 || np < 0 || nb < 0 || nl < 0 || nc < 0 || nsdf < 0 || nsdf > PT_MAX_SDF_OBJECTS ||
        6 * ns + 5 * np + 11 * nb + 12 * nl + 16 * nc > PT_MAX_OBJECTS_SIZE || (size_t)nsdf != s->sdfs.size()) {
        ctx->err = "packed scene object counts are inconsistent with its arrays";
        return PT_ERR_INVALID_ARG;
    }
    std::vector<RefSphere> sph((size_t)ns);
    std::vector<RefPlane> pln((size_t)np);
    std::vector<RefBox> box((size_t)nb);
    std::vector<RefLens> lens((size_t)nl);
    std::vector<RefCyclide> cyc((size_t)nc);
    std::vector<RefSdf> sdf((size_t)nsdf);
    int off = kUboObj;
    for (int i = 0; i < ns; ++i, off += 6)
        sph[(size_t)i] = RefSphere{U(off), U(off + 1), U(off + 2), U(off + 3), ID(off + 4), ID(off + 5)};
    for (int i = 0; i < np; ++i, off += 5)
        pln[(size_t)i] = RefPlane{U(off), U(off + 1), U(off + 2), ID(off + 3), ID(off + 4)};
    for (int i = 0; i < nb; ++i, off += 11) {
        RefBox& b = box[(size_t)i];
        b.pos = mk3(U(off), U(off + 1), U(off + 2));
        b.M = rotation_deg(mk3(U(off + 3), U(off + 4), U(off + 5)));
        b.size = mk3(U(off + 6), U(off + 7), U(off + 8));
        b.bsr2 = 0.25f * dot(b.size, b.size);
        b.mat = ID(off + 9);
        b.light = ID(off + 10);
    }
    for (int i = 0; i < nl; ++i, off += 12) {
        RefLens& L = lens[(size_t)i];
        L.pos = mk3(U(off), U(off + 1), U(off + 2));
        L.M = rotation_deg(mk3(U(off + 3), U(off + 4), U(off + 5)));
        float r = U(off + 6), f = U(off + 7), th = U(off + 8);
        L.conv = U(off + 9) != 0.0f;
        L.mat = ID(off + 10);
        L.light = ID(off + 11);
        L.radius = r;
        L.half = 2.0f * f - sqrtf(4.0f * f * f - r * r);
        L.slicePos = 0.5f * (L.conv ? th : -th);
        if (L.conv) L.slicePos += L.half;
        L.twoF = 2.0f * f;
        if (L.conv) {
            L.br2 = (r * r) + (0.25f * th * th);
        } else {
            float br = 0.5f * th + 2.0f * f - sqrtf(4.0f * f * f - r * r);
            L.br2 = br * br + r * r;
        }
    }
    for (int i = 0; i < nc; ++i, off += 16) {
        RefCyclide& C = cyc[(size_t)i];
        C.pos = mk3(U(off), U(off + 1), U(off + 2));
        C.M = rotation_deg(mk3(U(off + 3), U(off + 4), U(off + 5)));
        C.scale = mk3(U(off + 6), U(off + 7), U(off + 8));
        C.a = U(off + 9); C.b = U(off + 10); C.c = U(off + 11); C.d = U(off + 12);
        C.brad2 = U(off + 13);
        C.mat = ID(off + 14);
        C.light = ID(off + 15);
    }
    int kind = -1;
    for (int i = 0; i < nsdf; ++i) {
        RefSdf& q = sdf[(size_t)i];
        int k = kUboSdf + 6 * i;
        q.pos = mk3(U(k), U(k + 1), U(k + 2));
        q.size = mk3(U(k + 3), U(k + 4), U(k + 5));
        q.d = s->sdfs[(size_t)i];
        if (q.d.kind < PT_SDF_MANDELBULB || q.d.kind > PT_SDF_TERRAIN || q.d.iterations < 0 || q.d.iterations > 64) {
            ctx->err = "sdf descriptor " + std::to_string(i) + " has an unknown kind or iteration count";
            return PT_ERR_UNSUPPORTED_SDF;
        }
        kind = (kind == -1 || kind == q.d.kind) ? q.d.kind : 0;
    }
    size_t bytes = 0;
    auto sz = [&](size_t b) { bytes += ((b + 255) / 256) * 256; };
    sz(sizeof(pt_packed_scene)); sz(sph.size() * sizeof(RefSphere)); sz(pln.size() * sizeof(RefPlane));
    sz(box.size() * sizeof(RefBox)); sz(lens.size() * sizeof(RefLens)); sz(cyc.size() * sizeof(RefCyclide));
    sz(sdf.size() * sizeof(RefSdf));
    dev_free(ctx, s->ref_mem);
    int st = dev_alloc(ctx, s->ref_mem, bytes);
    if (st != PT_OK) return st;
    char* base = static_cast<char*>(s->ref_mem.p);
    char* p = base;
    auto put = [&](const void* src, size_t b) -> void* {
        void* dst = p;
        if (b) {
            hipError_t e = hipMemcpy(dst, src, b, hipMemcpyHostToDevice);
            if (e != hipSuccess) return nullptr;
        }
        p += ((b + 255) / 256) * 256;
        return dst;
    };
    RefSceneDev& D = s->ref_dev;
    D.ubo = static_cast<const float*>(put(&s->ubo, sizeof(pt_packed_scene)));
    D.sph = static_cast<const RefSphere*>(put(sph.data(), sph.size() * sizeof(RefSphere)));
    D.pln = static_cast<const RefPlane*>(put(pln.data(), pln.size() * sizeof(RefPlane)));
    D.box = static_cast<const RefBox*>(put(box.data(), box.size() * sizeof(RefBox)));
    D.lens = static_cast<const RefLens*>(put(lens.data(), lens.size() * sizeof(RefLens)));
    D.cyc = static_cast<const RefCyclide*>(put(cyc.data(), cyc.size() * sizeof(RefCyclide)));
    D.sdf = static_cast<const RefSdf*>(put(sdf.data(), sdf.size() * sizeof(RefSdf)));
    if (!D.ubo) return hip_fail(ctx, hipGetLastError(), "scene upload");
    D.ns = ns; D.np = np; D.nb = nb; D.nl = nl; D.nc = nc; D.nsdf = nsdf;
    D.numLights = U(6);
    D.sdf_kind = kind < 0 ? 0 : kind;
    s->device_bytes += (int64_t)bytes;
    return PT_OK;
}

static RefFrame make_frame(const pt_push_constants* pc) {
    RefFrame F{};
    F.W = pc->resolution[0];
    F.H = pc->resolution[1];
    F.frame = pc->frame;
    F.spf = pc->samplesPerFrame;
    F.pathLength = pc->pathLength;
    F.camSize = pc->cameraSize;
    F.apSize = pc->apertureSize;
    F.apDist = pc->apertureDist;
    F.camPos = mk3(pc->cameraPosX, pc->cameraPosY, pc->cameraPosZ);
    F.camM = rotation_deg(mk3(pc->cameraAngle[0], pc->cameraAngle[1], 0.0f));
    F.fwd = mk3(F.camM.c0.z, F.camM.c1.z, F.camM.c2.z);
    // The camera lens object of TracePathLens (src/shader.comp:1411-1418).
    F.lensPos = F.camPos + F.fwd * pc->lensDistance;
    F.lensM = rotation_deg(mk3(0.0f, 90.0f - pc->cameraAngle[1], pc->cameraAngle[0]));
    float r = pc->lensRadius, f = pc->lensFocalLength;
    F.lensRadius = r;
    F.lensHalf = 2.0f * f - sqrtf(4.0f * f * f - r * r);
    F.lensSlicePos = 0.5f * pc->lensThickness;
    F.lensSlicePos += F.lensHalf;
    F.lensTwoF = 2.0f * f;
    return F;
}

template <int KIND>
static void launch_march(pt_ctx* ctx, const RefSceneDev& S, RefWave& W, int cur, int grid) {
    launch(ctx, "k_ref_march", k_ref_march<KIND>, dim3(grid), dim3(kBlock), S, W, cur);
}
template <int KIND>
static void launch_shadow_march(pt_ctx* ctx, const RefSceneDev& S, RefWave& W, int grid) {
    launch(ctx, "k_ref_shadow_march", k_ref_shadow_march<KIND>, dim3(grid), dim3(kBlock), S, W);
}

// One dispatch through the wave Wv (capacity Wv.cap paths): pixels in chunks of
// at most cap, each chunk's samples in waves that fill it.
static int ref_dispatch_wave(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc, RefWave& Wv) {
    const RefSceneDev& S = s->ref_dev;
    const int W = pc->resolution[0], H = pc->resolution[1];
    const int P = W * H;
    const int cap = Wv.cap;
    RefFrame F = make_frame(pc);
    const int pcmax = std::min(P, cap);
    for (int p0 = 0; p0 < P; p0 += pcmax) {
        const int pcur = std::min(pcmax, P - p0);
        const int m = std::max(1, std::min(pc->samplesPerFrame, cap / pcur));
        for (int k0 = 0; k0 < pc->samplesPerFrame; k0 += m) {
            F.p0 = p0;
            F.pc = pcur;
            F.k0 = k0;
            F.m = std::min(m, pc->samplesPerFrame - k0);
            const int npaths = F.pc * F.m;
            const int grid = (npaths + kBlock - 1) / kBlock;
            launch(ctx, "k_ref_counts", k_ref_counts, dim3(1), dim3(1), Wv.counts, npaths, 0, 0);
            launch(ctx, "k_ref_raygen", k_ref_raygen, dim3(grid), dim3(kBlock), F, Wv);
            for (int b = 0; b < pc->pathLength; ++b) {
                const int cur = b & 1;
                launch(ctx, "k_ref_reset", k_ref_reset, dim3(1), dim3(1), Wv.counts, cur ^ 1);
                launch(ctx, "k_ref_intersect", k_ref_intersect, dim3(grid), dim3(kBlock), S, Wv, cur);
                if (S.nsdf > 0) {
                    switch (S.sdf_kind) {
                        case PT_SDF_MANDELBULB: launch_march<PT_SDF_MANDELBULB>(ctx, S, Wv, cur, grid); break;
                        case PT_SDF_MENGER: launch_march<PT_SDF_MENGER>(ctx, S, Wv, cur, grid); break;
                        case PT_SDF_BLOB: launch_march<PT_SDF_BLOB>(ctx, S, Wv, cur, grid); break;
                        case PT_SDF_TERRAIN: launch_march<PT_SDF_TERRAIN>(ctx, S, Wv, cur, grid); break;
                        default: launch_march<0>(ctx, S, Wv, cur, grid); break;
                    }
                }
                launch(ctx, "k_ref_shade", k_ref_shade, dim3(grid), dim3(kBlock), S, F, Wv, cur, b);
                launch(ctx, "k_ref_shadow", k_ref_shadow, dim3(grid), dim3(kBlock), S, Wv, S.nsdf > 0 ? 0 : 1);
                if (S.nsdf > 0) {
                    switch (S.sdf_kind) {
                        case PT_SDF_MANDELBULB: launch_shadow_march<PT_SDF_MANDELBULB>(ctx, S, Wv, grid); break;
                        case PT_SDF_MENGER: launch_shadow_march<PT_SDF_MENGER>(ctx, S, Wv, grid); break;
                        case PT_SDF_BLOB: launch_shadow_march<PT_SDF_BLOB>(ctx, S, Wv, grid); break;
                        case PT_SDF_TERRAIN: launch_shadow_march<PT_SDF_TERRAIN>(ctx, S, Wv, grid); break;
                        default: launch_shadow_march<0>(ctx, S, Wv, grid); break;
                    }
                }
            }
            launch(ctx, "k_ref_accum", k_ref_accum, dim3((F.pc + kBlock - 1) / kBlock), dim3(kBlock), S, F, Wv);
        }
        RefFinish Fin{};
        Fin.p0 = p0;
        Fin.pc = pcur;
        Fin.spf = pc->samplesPerFrame;
        Fin.currentSamples = pc->currentSamples;
        Fin.frame = pc->frame;
        Fin.ISO = pc->ISO;
        Fin.apSize = pc->apertureSize;
        Fin.FPS = pc->FPS;
        Fin.persistence = pc->persistence;
        launch(ctx, "k_ref_finish", k_ref_finish, dim3((pcur + kBlock - 1) / kBlock), dim3(kBlock), Fin,
               (const float*)Wv.sum, reinterpret_cast<float4*>(acc->d_texels));
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(ctx, e, "reference dispatch launch");
    return PT_OK;
}

int ref_dispatch(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc) {
    int cap = ctx->cfg.wave_paths > 0 ? ctx->cfg.wave_paths : (1 << 21);
    cap = std::max(cap, 1024);
    int st = ensure_ref_wave(ctx, cap);
    if (st != PT_OK) return st;
    return ref_dispatch_wave(ctx, s, acc, pc, ctx->ref_wave);
}

int ref_dispatch_static(pt_ctx* ctx, const pt_scene* s, pt_accum* acc, const pt_push_constants* pc) {
    char* base = nullptr;
    if (ctx->device >= 0 && ctx->device < kMaxStaticDevices) {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        base = g_scratch[ctx->device];
    }
    if (!base) {
        ctx->err = "pt_dispatch_noalloc: the static wave of this device is not set up";
        return PT_ERR_STATE;
    }
    RefWave Wv{};
    carve_ref_wave(base, kStaticWavePaths, ctx->d_stats, Wv);
    return ref_dispatch_wave(ctx, s, acc, pc, Wv);
}

}  // namespace ptk
