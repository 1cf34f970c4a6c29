/*
 * ptext.c — ORACLE (test infrastructure only; never linked into the product).
 *
 * FP64 CPU tracer of the extended mode, written independently of the GPU
 * kernels from textbook radiometry, for convergence tests (the GPU image must
 * agree with this one within the Monte Carlo bound) and as the CPU baseline of
 * bench.py. It shares with the GPU only the scene definition and the
 * random-number specification (pcg_hash of (seed, pixel, sample, bounce, dim)),
 * so with the same seed both trace nearly the same paths and with another
 * seed they are independent estimators of the same image.
 *
 * Physics (the extended mode's definition):
 *  - spheres, parallelograms (boxes arrive as six), horizontal planes, SDFs
 *    (mandelbulb, menger, blob, terrain) by sphere tracing, the reference's
 *    lenses (two spherical caps, src/shader.comp:366-448) and Dupin cyclides
 *    (the quartic of src/shader.comp:633-679, its real roots isolated by the
 *    roots of its derivatives and bisection, not by a closed-form solver);
 *  - Lambert with the reference's one-lobe spectrum (src/shader.comp:1030-1038)
 *    or a clamped sum of Gaussians; smooth dielectric (Cauchy A + B / lambda^2,
 *    lambda in um, or Sellmeier) with exact unpolarised Fresnel; smooth
 *    conductor with the complex-index Fresnel reflectance;
 *  - emitters: Planck's law normalised at its Wien peak times luminosity
 *    (src/shader.comp:1040-1055), two-sided; uniform environment radiance;
 *    emissive lenses and cyclides are found by BSDF sampling only (weight 1);
 *  - hero wavelength uniform in [lmin, lmin + range) with n - 1 companions
 *    rotated by range / n; a dispersive refraction keeps the hero only (its
 *    throughput times n);
 *  - next-event estimation at Lambert vertices (uniform light choice, area
 *    sampling of parallelograms, solid-angle cone sampling of spheres), power
 *    heuristic against cosine BSDF sampling; Russian roulette from depth
 *    rr_depth with q = min(max throughput, 0.95). With nee = 0, BSDF sampling
 *    alone (emitters hit with weight 1): an independent estimator of the same
 *    image, whose agreement pins the light pdfs;
 *  - camera: a pinhole, or the reference's camera: a point on the sensor, a
 *    point on the aperture disk, then refraction at the two spherical faces of
 *    its biconvex BK7 lens (src/shader.comp:1409-1465).
 * Parallel over rows with OpenMP.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

typedef struct {
    double x, y, z;
} d3;
static d3 D3(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static d3 add(d3 a, d3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
static d3 sub(d3 a, d3 b) { return D3(a.x - b.x, a.y - b.y, a.z - b.z); }
static d3 mul(d3 a, double s) { return D3(a.x * s, a.y * s, a.z * s); }
static double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static d3 cross(d3 a, d3 b) { return D3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
static d3 unit(d3 a) { return mul(a, 1.0 / sqrt(dot(a, a))); }
static double comp(d3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

/* Scene passed from Python (oracle/ptext.py). */
typedef struct {
    int nsph, nquad, nplane, nsdf, nmat, nlight;
    const double* sph;     /* cx cy cz r */
    const int* sph_ml;     /* material (0-based), light (-1 none) */
    const double* quad;    /* Q(3) e1(3) e2(3) */
    const int* quad_ml;
    const double* plane_y;
    const int* plane_ml;
    const double* sdf;     /* kind iters bailout mat cx cy cz sx sy sz */
    const double* mat;     /* type n dispersive p[9] */
    const double* light;   /* type(0 sphere, 1 quad) prim T lum */
    const float* cie;      /* 441 x 3 */
    int has_env;
    double env;
    int n_lambda;
    double lmin, lrange;
    int rr_depth;
    uint32_t seed;
    int W, H, max_depth, sdf_steps;
    double cam_pos[3], cam_angle[2], cam_size, ap_dist;
    /* camera model (0: the reference's sensor + aperture + BK7 lens, 1: pinhole)
     * and the lens parameters of the scene's camera */
    int camera;
    double ap_size, lens_dist, lens_radius, lens_focal, lens_thick;
    int nee; /* 1: NEE with MIS at diffuse vertices; 0: BSDF sampling only */
    uint64_t lens_missed; /* camera rays that missed a lens face (written) */
    int nlens, ncyc;
    /* lens: pos(3) rotation(3, degrees) radius focalLength thickness isConverging mat light */
    const double* lens;
    /* cyclide: pos(3) rotation(3) scale(3) a b c d boundingRadius (squared) mat light */
    const double* cyc;
    int nlight_nee; /* lights [0, nlight_nee) are sampled by NEE (spheres, parallelograms) */
} Scene;

enum { T_SPHERE = 0, T_QUAD = 1, T_PLANE = 2, T_SDF = 3, T_LENS = 4, T_CYC = 5 };
#define LENS_REC 12
#define CYC_REC 16
enum { M_LOBE = 0, M_GAUSS = 1, M_DIEL = 2, M_COND = 3 };

/* ---- random numbers: pcg_hash (Jarzynski & Olano 2020) keyed per path ---------- */
static uint32_t pcg(uint32_t v) {
    uint32_t s = v * 747796405u + 2891336453u;
    uint32_t w = ((s >> ((s >> 28u) + 4u)) ^ s) * 277803737u;
    return (w >> 22u) ^ w;
}
static uint32_t key_of(uint32_t seed, uint32_t pixel, uint32_t sample) { return pcg(pcg(pcg(seed) + pixel) + sample); }
static double rnd(uint32_t key, uint32_t bounce, uint32_t dim) {
    return (double)(pcg(key + pcg((bounce << 4) | dim)) >> 8) * (1.0 / 16777216.0);
}
#define CAM 0xFFFu

/* ---- BVH (binary, median split on the longest centroid axis) --------------------- */
typedef struct { double lo[3], hi[3]; } box;
typedef struct { box b; int left, right, first, count; } bnode;
typedef struct { int type, idx; } pref;

typedef struct {
    const Scene* s;
    bnode* nodes;
    int nnodes;
    pref* prims;
    int nprims;
} Accel;

static box prim_box(const Scene* s, pref p) {
    box b;
    if (p.type == T_SPHERE) {
        const double* q = s->sph + 4 * p.idx;
        for (int k = 0; k < 3; ++k) { b.lo[k] = q[k] - q[3]; b.hi[k] = q[k] + q[3]; }
    } else {
        const double* q = s->quad + 9 * p.idx;
        d3 Q = D3(q[0], q[1], q[2]), e1 = D3(q[3], q[4], q[5]), e2 = D3(q[6], q[7], q[8]);
        d3 c[4] = {Q, add(Q, e1), add(Q, e2), add(add(Q, e1), e2)};
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = 1e300; b.hi[k] = -1e300;
            for (int v = 0; v < 4; ++v) {
                b.lo[k] = fmin(b.lo[k], comp(c[v], k)) - 1e-9;
                b.hi[k] = fmax(b.hi[k], comp(c[v], k)) + 1e-9;
            }
        }
    }
    return b;
}
static int sort_axis;
static const Scene* sort_scene;
static double centroid(const Scene* s, pref p, int k) {
    box b = prim_box(s, p);
    return 0.5 * (b.lo[k] + b.hi[k]);
}
static int cmp_pref(const void* a, const void* b) {
    double ca = centroid(sort_scene, *(const pref*)a, sort_axis), cb = centroid(sort_scene, *(const pref*)b, sort_axis);
    return ca < cb ? -1 : (ca > cb ? 1 : 0);
}
static int build(Accel* A, int first, int count) {
    int id = A->nnodes++;
    bnode* n = &A->nodes[id];
    box b = {{1e300, 1e300, 1e300}, {-1e300, -1e300, -1e300}};
    double clo[3] = {1e300, 1e300, 1e300}, chi[3] = {-1e300, -1e300, -1e300};
    for (int i = first; i < first + count; ++i) {
        box pb = prim_box(A->s, A->prims[i]);
        for (int k = 0; k < 3; ++k) {
            b.lo[k] = fmin(b.lo[k], pb.lo[k]);
            b.hi[k] = fmax(b.hi[k], pb.hi[k]);
            double c = 0.5 * (pb.lo[k] + pb.hi[k]);
            clo[k] = fmin(clo[k], c);
            chi[k] = fmax(chi[k], c);
        }
    }
    n->b = b;
    n->first = first;
    n->count = count;
    n->left = n->right = -1;
    if (count <= 2) return id;
    int axis = 0;
    for (int k = 1; k < 3; ++k)
        if (chi[k] - clo[k] > chi[axis] - clo[axis]) axis = k;
    sort_axis = axis;
    sort_scene = A->s;
    qsort(A->prims + first, (size_t)count, sizeof(pref), cmp_pref);
    int half = count / 2;
    int l = build(A, first, half);
    int r = build(A, first + half, count - half);
    A->nodes[id].left = l;
    A->nodes[id].right = r;
    return id;
}
static void accel_build(Accel* A, const Scene* s) {
    A->s = s;
    A->nprims = s->nsph + s->nquad;
    A->prims = (pref*)malloc(sizeof(pref) * (size_t)(A->nprims > 0 ? A->nprims : 1));
    for (int i = 0; i < s->nsph; ++i) { A->prims[i].type = T_SPHERE; A->prims[i].idx = i; }
    for (int i = 0; i < s->nquad; ++i) { A->prims[s->nsph + i].type = T_QUAD; A->prims[s->nsph + i].idx = i; }
    A->nodes = (bnode*)malloc(sizeof(bnode) * (size_t)(2 * A->nprims + 2));
    A->nnodes = 0;
    if (A->nprims > 0) build(A, 0, A->nprims);
}
static void accel_free(Accel* A) {
    free(A->nodes);
    free(A->prims);
}

/* ---- intersection ----------------------------------------------------------------- */
#define EPS 1e-9
static int hit_sphere(const double* q, d3 o, d3 d, double tmax, double* t) {
    d3 oc = sub(o, D3(q[0], q[1], q[2]));
    double b = dot(oc, d);
    d3 f = sub(oc, mul(d, b));
    double disc = q[3] * q[3] - dot(f, f);
    if (disc < 0) return 0;
    double s = sqrt(disc);
    double t0 = -b - s, t1 = -b + s;
    if (t0 > EPS && t0 < tmax) { *t = t0; return 1; }
    if (t1 > EPS && t1 < tmax) { *t = t1; return 1; }
    return 0;
}
static int hit_quad(const double* q, d3 o, d3 d, double tmax, double* t) {
    d3 Q = D3(q[0], q[1], q[2]), e1 = D3(q[3], q[4], q[5]), e2 = D3(q[6], q[7], q[8]);
    d3 n = cross(e1, e2);
    double den = dot(n, d);
    if (fabs(den) < 1e-300) return 0;
    double tt = dot(n, sub(Q, o)) / den;
    if (!(tt > EPS && tt < tmax)) return 0;
    d3 hp = sub(add(o, mul(d, tt)), Q);
    double nn = dot(n, n);
    double a = dot(cross(hp, e2), n) / nn, b = dot(cross(e1, hp), n) / nn;
    if (a < 0 || a > 1 || b < 0 || b > 1) return 0;
    *t = tt;
    return 1;
}
static d3 rot_row(d3 a, d3 v);
static d3 rot_col(d3 a, d3 v);

/* A lens (src/shader.comp:366-448): in its frame (x on the optical axis, local =
 * v * RotationMatrix), two spheres of radius R = 2f with centres on the axis at
 * +-(R - slicePos) and cap height h = R - sqrt(R^2 - r^2); cap 0 is the part of
 * the first sphere with x - centre <= -(R - h), cap 1 the part of the second
 * with x - centre >= R - h. slicePos = thickness / 2 + h for a converging lens
 * (the solid inside both spheres), -thickness / 2 for a diverging one (outside
 * them). The nearest cap point beyond EPS; *n is the solid's outward normal. */
static int hit_lens(const double* q, d3 o, d3 d, double tmax, double* t, d3* n) {
    d3 pos = D3(q[0], q[1], q[2]), ang = D3(q[3], q[4], q[5]);
    double r = q[6], f = q[7], th = q[8];
    int conv = q[9] != 0.0;
    double R = 2.0 * f, h = R - sqrt(R * R - r * r), sp = conv ? 0.5 * th + h : -0.5 * th;
    d3 lo = rot_row(ang, sub(o, pos)), ld = rot_row(ang, d);
    double best = tmax;
    int hit = 0;
    for (int k = 0; k < 2; ++k) {
        double cx = k == 0 ? R - sp : sp - R;
        d3 oc = sub(lo, D3(cx, 0.0, 0.0));
        double b = dot(oc, ld), c = dot(oc, oc) - R * R, disc = b * b - c;
        if (disc < 0.0) continue;
        double sq = sqrt(disc), roots[2] = {-b - sq, -b + sq};
        for (int i = 0; i < 2; ++i) {
            double tt = roots[i];
            if (!(tt > EPS && tt < best)) continue;
            double xr = oc.x + ld.x * tt;
            if (k == 0 ? !(xr <= -(R - h)) : !(xr >= R - h)) continue;
            best = tt;
            hit = 1;
            d3 nl = unit(add(oc, mul(ld, tt)));
            *n = rot_col(ang, conv ? nl : mul(nl, -1.0));
            break;
        }
    }
    if (hit) *t = best;
    return hit;
}

/* p(t) = c[0] + c[1] t + ... + c[n] t^n (Horner). */
static double poly_at(const double* c, int n, double t) {
    double v = c[n];
    for (int k = n - 1; k >= 0; --k) v = v * t + c[k];
    return v;
}
/* The real roots of p in (lo, hi), ascending, at most n: the roots of p' split
 * (lo, hi) into intervals on which p is monotonic, and an interval whose ends
 * differ in sign holds one root, found by bisection to the last bit. A double
 * root (a tangency, where p does not change sign) is not found. */
static int poly_roots(const double* c, int n, double lo, double hi, double* out) {
    if (n <= 1) {
        if (n == 1 && c[1] != 0.0) {
            double t = -c[0] / c[1];
            if (t > lo && t < hi) { out[0] = t; return 1; }
        }
        return 0;
    }
    double dc[4], crit[4], ends[6];
    for (int k = 0; k < n; ++k) dc[k] = (k + 1) * c[k + 1];
    int nc = poly_roots(dc, n - 1, lo, hi, crit), ne = 0, m = 0;
    ends[ne++] = lo;
    for (int i = 0; i < nc; ++i) ends[ne++] = crit[i];
    ends[ne++] = hi;
    for (int i = 0; i + 1 < ne; ++i) {
        double a = ends[i], b = ends[i + 1], fa = poly_at(c, n, a), fb = poly_at(c, n, b);
        if ((fa < 0.0) == (fb < 0.0)) continue;
        for (int it = 0; it < 200; ++it) {
            double mid = 0.5 * (a + b);
            if (mid <= a || mid >= b) break;
            double fm = poly_at(c, n, mid);
            if ((fm < 0.0) == (fa < 0.0)) { a = mid; fa = fm; } else { b = mid; }
        }
        out[m++] = 0.5 * (a + b);
    }
    return m;
}
/* A Dupin cyclide (src/shader.comp:633-679): in its frame, p = (v * RotationMatrix
 * - ...) / scale, the surface F(p) = (|p|^2 + b^2 - d^2)^2 - 4 (a p.x - c d)^2
 * - 4 b^2 p.z^2 = 0, searched inside its bounding sphere (radius^2 = brad).
 * F along the ray is a quartic in t (the same t as the world ray's); its
 * smallest root beyond EPS is the hit, and *n the world gradient of F
 * (outward: F > 0 far away). */
static int hit_cyclide(const double* q, d3 o, d3 d, double tmax, double* t, d3* n) {
    d3 pos = D3(q[0], q[1], q[2]), ang = D3(q[3], q[4], q[5]), sc = D3(q[6], q[7], q[8]);
    double A = q[9], B = q[10], Cc = q[11], D = q[12], brad = q[13];
    d3 oc = sub(o, pos);
    double bb = dot(oc, d), cc = dot(oc, oc) - brad, disc = bb * bb - cc;
    if (disc < 0.0) return 0;
    double sq = sqrt(disc), ta = fmax(-bb - sq, EPS), tb = fmin(-bb + sq, tmax);
    if (!(ta < tb)) return 0;
    d3 lp = rot_row(ang, oc), lv = rot_row(ang, d);
    d3 p = D3(lp.x / sc.x, lp.y / sc.y, lp.z / sc.z), v = D3(lv.x / sc.x, lv.y / sc.y, lv.z / sc.z);
    /* S(t) = s2 t^2 + s1 t + s0, L1(t) = l1 t + l0, L2(t) = m1 t + m0 */
    double s2 = dot(v, v), s1 = 2.0 * dot(p, v), s0 = dot(p, p) + B * B - D * D;
    double l1 = A * v.x, l0 = A * p.x - Cc * D, m1 = v.z, m0 = p.z;
    double c[5];
    c[4] = s2 * s2;
    c[3] = 2.0 * s2 * s1;
    c[2] = s1 * s1 + 2.0 * s2 * s0 - 4.0 * l1 * l1 - 4.0 * B * B * m1 * m1;
    c[1] = 2.0 * s1 * s0 - 8.0 * l1 * l0 - 8.0 * B * B * m1 * m0;
    c[0] = s0 * s0 - 4.0 * l0 * l0 - 4.0 * B * B * m0 * m0;
    double roots[4];
    if (poly_roots(c, 4, ta, tb, roots) == 0) return 0;
    double tt = roots[0];
    d3 x = add(p, mul(v, tt));
    double S = dot(x, x) + B * B - D * D;
    d3 g = D3(4.0 * x.x * S - 8.0 * A * (A * x.x - Cc * D), 4.0 * x.y * S, 4.0 * x.z * S - 8.0 * B * B * x.z);
    *n = unit(rot_col(ang, D3(g.x / sc.x, g.y / sc.y, g.z / sc.z)));
    *t = tt;
    return 1;
}

static int slab(const box* b, d3 o, d3 inv, double tmax) {
    double t0 = 0.0, t1 = tmax;
    for (int k = 0; k < 3; ++k) {
        double a = (b->lo[k] - comp(o, k)) * comp(inv, k), c = (b->hi[k] - comp(o, k)) * comp(inv, k);
        if (a > c) { double tmp = a; a = c; c = tmp; }
        t0 = fmax(t0, a);
        t1 = fmin(t1, c);
    }
    return t0 <= t1;
}

/* SDF distance estimators (the shapes of the reference's sdfs/ GLSL files, in binary64). */
static double de_mandelbulb(d3 p, int iters, double bail) {
    d3 z = p;
    double r = sqrt(dot(z, z)), dr = 1.0;
    for (int i = 0; i < iters; ++i) {
        dr = 8.0 * pow(r, 7.0) * dr + 1.0;
        double invr = 1.0 / r, invl = 1.0 / sqrt(z.x * z.x + z.z * z.z);
        double r8 = pow(r, 8.0);
        double ct = z.z * invl, cp = z.y * invr, st = z.x * invl, sp = invr / invl;
        /* cos 8t, sin 8t from cos t, sin t (Chebyshev), as the shader's power-8 formula */
        double c2 = ct * ct, p2 = cp * cp;
        double ct8 = (((128 * c2 - 256) * c2 + 160) * c2 - 32) * c2 + 1;
        double cp8 = (((128 * p2 - 256) * p2 + 160) * p2 - 32) * p2 + 1;
        double st8 = (((128 * c2 - 192) * c2 + 80) * c2 - 8) * st * ct;
        double sp8 = (((128 * p2 - 192) * p2 + 80) * p2 - 8) * sp * cp;
        z = add(mul(D3(st8 * sp8, cp8, ct8 * sp8), r8), p);
        r = sqrt(dot(z, z));
        if (r > bail) break;
    }
    return 0.5 * log(r) * r / dr;
}
static double de_menger(d3 p, int levels) {
    d3 q = D3(fabs(p.x) - 1, fabs(p.y) - 1, fabs(p.z) - 1);
    d3 qm = D3(fmax(q.x, 0), fmax(q.y, 0), fmax(q.z, 0));
    double d = sqrt(dot(qm, qm)) + fmin(fmax(q.x, fmax(q.y, q.z)), 0.0);
    double s = 1.0;
    for (int m = 0; m < levels; ++m) {
        double a[3] = {p.x * s, p.y * s, p.z * s};
        for (int k = 0; k < 3; ++k) a[k] = a[k] - 2.0 * floor(a[k] / 2.0) - 1.0;
        s *= 3.0;
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = fabs(1.0 - 3.0 * fabs(a[k]));
        double da = fmax(r[0], r[1]), db = fmax(r[1], r[2]), dc = fmax(r[2], r[0]);
        d = fmax(d, (fmin(da, fmin(db, dc)) - 1.0) / s);
    }
    return d;
}
static double smin_d(double x, double y) {
    const double k = 0.12;
    double h = fmax(k - fabs(x - y), 0.0) / k;
    return fmin(x, y) - h * h * h * k / 6.0;
}
static double de_blob(d3 p) {
    double s1 = sqrt(dot(p, p)) - 1.0;
    d3 a = D3(p.x, p.y - 1.0, p.z), b = D3(p.x - 0.25, p.y - 1.25, p.z - 0.5), c = D3(p.x + 0.25, p.y - 1.25, p.z - 0.5);
    return smin_d(smin_d(smin_d(s1, sqrt(dot(a, a)) - 0.7), sqrt(dot(b, b)) - 0.25), sqrt(dot(c, c)) - 0.25);
}
static double terrain_h(double x) {
    return sin(0.0625 * x) + sin(0.125 * (x + 10.0)) + 0.25 * sin(0.25 * x) + 0.125 * sin(0.5 * x) + 0.25 * sin(x) +
           0.25 * sin(2.0 * x + 1.0) + 0.125 * sin(4.0 * x + 2.0) + 0.0625 * sin(8.0 * x + 1.0) +
           0.03125 * sin(16.0 * x + 5.0) + 0.015625 * sin(32.0 * x) + 0.0078125 * sin(64.0 * x);
}
static double de(const double* q, d3 p) {
    d3 c = sub(p, D3(q[4], q[5], q[6]));
    switch ((int)q[0]) {
        case 1: return de_mandelbulb(c, (int)q[1], q[2]);
        case 2: return de_menger(c, (int)q[1]);
        case 3: return de_blob(c);
        default: return 0.5 * (c.y - (terrain_h(c.x) + terrain_h(c.z)));
    }
}
static double march(const double* q, d3 o, d3 d, double tmax, int max_steps, uint64_t* steps) {
    double tin = EPS, tout = tmax;
    for (int k = 0; k < 3; ++k) {
        double lo = q[4 + k] - 0.5 * q[7 + k], hi = q[4 + k] + 0.5 * q[7 + k];
        double a = (lo - comp(o, k)) / comp(d, k), b = (hi - comp(o, k)) / comp(d, k);
        if (a > b) { double t = a; a = b; b = t; }
        tin = fmax(tin, a);
        tout = fmin(tout, b);
    }
    if (!(tin < tout)) return -1.0;
    double t = tin;
    for (int i = 0; i < max_steps; ++i) {
        double dist = de(q, add(o, mul(d, t)));
        (*steps)++;
        if (dist < 1e-4) return t;
        t += dist;
        if (t > tout) return -1.0;
    }
    return -1.0;
}

typedef struct { double t; int type, idx; d3 n; /* lens, cyclide: the normal found with the hit */ } hitrec;

static int trace(const Accel* A, d3 o, d3 d, double tmax, int any, hitrec* h, uint64_t* steps) {
    const Scene* s = A->s;
    h->t = tmax;
    h->type = -1;
    d3 inv = D3(1.0 / d.x, 1.0 / d.y, 1.0 / d.z);
    int stack[128], sp = 0;
    if (A->nprims > 0) stack[sp++] = 0;
    while (sp > 0) {
        const bnode* n = &A->nodes[stack[--sp]];
        if (!slab(&n->b, o, inv, h->t)) continue;
        if (n->left < 0) {
            for (int i = n->first; i < n->first + n->count; ++i) {
                pref p = A->prims[i];
                double t;
                int ok = p.type == T_SPHERE ? hit_sphere(s->sph + 4 * p.idx, o, d, h->t, &t)
                                            : hit_quad(s->quad + 9 * p.idx, o, d, h->t, &t);
                if (ok) {
                    h->t = t;
                    h->type = p.type;
                    h->idx = p.idx;
                    if (any) return 1;
                }
            }
        } else {
            stack[sp++] = n->left;
            stack[sp++] = n->right;
        }
    }
    for (int i = 0; i < s->nplane; ++i) {
        double t = (s->plane_y[i] - o.y) / d.y;
        if (t > EPS && t < h->t) {
            h->t = t;
            h->type = T_PLANE;
            h->idx = i;
            if (any) return 1;
        }
    }
    for (int i = 0; i < s->nlens; ++i) {
        double t;
        d3 nn;
        if (hit_lens(s->lens + LENS_REC * i, o, d, h->t, &t, &nn)) {
            h->t = t;
            h->type = T_LENS;
            h->idx = i;
            h->n = nn;
            if (any) return 1;
        }
    }
    for (int i = 0; i < s->ncyc; ++i) {
        double t;
        d3 nn;
        if (hit_cyclide(s->cyc + CYC_REC * i, o, d, h->t, &t, &nn)) {
            h->t = t;
            h->type = T_CYC;
            h->idx = i;
            h->n = nn;
            if (any) return 1;
        }
    }
    for (int i = 0; i < s->nsdf; ++i) {
        double t = march(s->sdf + 10 * i, o, d, h->t, s->sdf_steps, steps);
        if (t > 0 && t < h->t) {
            h->t = t;
            h->type = T_SDF;
            h->idx = i;
            if (any) return 1;
        }
    }
    return h->type >= 0;
}

/* ---- spectra and BSDFs ------------------------------------------------------------- */
static void cmf(const Scene* s, double l, double out[3]) {
    double x = fmin(fmax(l - 360.0, 0.0), 439.999);
    int i = (int)x;
    double t = x - i;
    for (int k = 0; k < 3; ++k) out[k] = s->cie[3 * i + k] + ((double)s->cie[3 * i + 3 + k] - s->cie[3 * i + k]) * t;
}
static double planck_norm(double lnm, double T, double lum) {
    T = fmax(T, 0.0);
    double l = lnm * 1e-9;
    double b = 1.1910429724e-16 / pow(l, 5.0) / (exp(0.014387768775 / (l * T)) - 1.0);
    double peak = 4.0956746759e-6 * pow(T, 5.0);
    return b / peak * fmax(lum, 0.0);
}
static double reflect_spec(const double* m, double l) {
    if ((int)m[0] == M_LOBE) {
        double x = (l - m[3]) / (2.0 * m[4] * m[4]);
        double r = exp(-x * x);
        return m[5] != 0.0 ? 1.0 - r : r;
    }
    double s = 0.0;
    for (int i = 0; i < (int)m[1]; ++i) {
        double d = (l - m[3 + 3 * i + 1]) / m[3 + 3 * i + 2];
        s += m[3 + 3 * i] * exp(-0.5 * d * d);
    }
    return fmin(s, 1.0);
}
static double ior(const double* m, double lnm) {
    double l2 = (lnm * 1e-3) * (lnm * 1e-3);
    if ((int)m[1] == 0) return m[3] + m[4] / l2;
    return sqrt(1.0 + m[3] * l2 / (l2 - m[6]) + m[4] * l2 / (l2 - m[7]) + m[5] * l2 / (l2 - m[8]));
}
static double fresnel_diel(double ci, double eta_i, double eta_t) {
    double st2 = (eta_i / eta_t) * (eta_i / eta_t) * fmax(0.0, 1.0 - ci * ci);
    if (st2 >= 1.0) return 1.0;
    double ct = sqrt(1.0 - st2);
    double rs = (eta_i * ci - eta_t * ct) / (eta_i * ci + eta_t * ct);
    double rp = (eta_t * ci - eta_i * ct) / (eta_t * ci + eta_i * ct);
    return 0.5 * (rs * rs + rp * rp);
}
/* Conductor: unpolarised reflectance from the s and p amplitude ratios with
 * a complex index (Born & Wolf), written with C99 complex arithmetic. */
static double fresnel_cond(double ci, double n, double k) {
    double complex eta = n + k * I;
    double complex ct = csqrt(1.0 - (1.0 - ci * ci) / (eta * eta));
    double complex rs = (ci - eta * ct) / (ci + eta * ct);
    double complex rp = (eta * ci - ct) / (eta * ci + ct);
    return 0.5 * (creal(rs * conj(rs)) + creal(rp * conj(rp)));
}
/* Tangent frame of Duff et al. 2017 ("Building an orthonormal basis,
 * revisited"), the frame the extended mode's sampling is specified in, so
 * that equal random numbers give equal directions. */
static void basis(d3 n, d3* t, d3* b) {
    double s = copysign(1.0, n.z);
    double a = -1.0 / (s + n.z);
    double c = n.x * n.y * a;
    *t = D3(1.0 + s * n.x * n.x * a, s * c, -s * n.x);
    *b = D3(c, s + n.y * n.y * a, -n.y);
}

static void surface(const Scene* s, const hitrec* h, d3 p, d3* n, int* mat, int* light) {
    switch (h->type) {
        case T_SPHERE: {
            const double* q = s->sph + 4 * h->idx;
            *n = mul(sub(p, D3(q[0], q[1], q[2])), 1.0 / q[3]);
            *mat = s->sph_ml[2 * h->idx];
            *light = s->sph_ml[2 * h->idx + 1];
            break;
        }
        case T_QUAD: {
            const double* q = s->quad + 9 * h->idx;
            *n = unit(cross(D3(q[3], q[4], q[5]), D3(q[6], q[7], q[8])));
            *mat = s->quad_ml[2 * h->idx];
            *light = s->quad_ml[2 * h->idx + 1];
            break;
        }
        case T_PLANE:
            *n = D3(0, 1, 0);
            *mat = s->plane_ml[2 * h->idx];
            *light = s->plane_ml[2 * h->idx + 1];
            break;
        case T_LENS: {
            const double* q = s->lens + LENS_REC * h->idx;
            *n = h->n;
            *mat = (int)q[10];
            *light = (int)q[11];
            break;
        }
        case T_CYC: {
            const double* q = s->cyc + CYC_REC * h->idx;
            *n = h->n;
            *mat = (int)q[14];
            *light = (int)q[15];
            break;
        }
        default: {
            const double* q = s->sdf + 10 * h->idx;
            double e = 1e-5;
            *n = unit(D3(de(q, add(p, D3(e, 0, 0))) - de(q, sub(p, D3(e, 0, 0))),
                         de(q, add(p, D3(0, e, 0))) - de(q, sub(p, D3(0, e, 0))),
                         de(q, add(p, D3(0, 0, e))) - de(q, sub(p, D3(0, 0, e)))));
            *mat = (int)q[3];
            *light = -1;
        }
    }
}

/* Solid-angle pdf with which NEE from x picks point y on light li. */
static double light_pdf(const Scene* s, int li, d3 x, d3 dir, double dist) {
    const double* L = s->light + 4 * li;
    double pdf;
    if ((int)L[0] == 1) {
        const double* q = s->quad + 9 * (int)L[1];
        d3 nn = cross(D3(q[3], q[4], q[5]), D3(q[6], q[7], q[8]));
        double area = sqrt(dot(nn, nn));
        double cl = fabs(dot(nn, dir)) / area;
        pdf = dist * dist / (cl * area);
    } else {
        const double* q = s->sph + 4 * (int)L[1];
        d3 dc = sub(D3(q[0], q[1], q[2]), x);
        double cmax = sqrt(fmax(0.0, 1.0 - q[3] * q[3] / dot(dc, dc)));
        pdf = 1.0 / (2.0 * M_PI * (1.0 - cmax));
    }
    return pdf / s->nlight_nee;
}

/* One path: returns its XYZ estimate. */
static void trace_path(const Scene* s, const Accel* A, uint32_t key, d3 o, d3 d, double l0, double out[3],
                       uint64_t* cnt) {
    int nl = s->n_lambda;
    double lam[4], T[4] = {1, 1, 1, 1};
    lam[0] = l0;
    for (int j = 1; j < 4; ++j) {
        double x = l0 - s->lmin + s->lrange / nl * j;
        if (x >= s->lrange) x -= s->lrange;
        lam[j] = s->lmin + x;
    }
    double acc[4] = {0, 0, 0, 0};
    int specular = 1, collapsed = 0;
    double pdf_prev = 0.0;
    for (int b = 0; b < s->max_depth; ++b) {
        hitrec h;
        cnt[0]++;
        if (!trace(A, o, d, 1e300, 0, &h, &cnt[2])) {
            if (s->has_env)
                for (int j = 0; j < nl; ++j) acc[j] += T[j] * s->env;
            break;
        }
        d3 p = add(o, mul(d, h.t)), n;
        int mi, li;
        surface(s, &h, p, &n, &mi, &li);
        if (li >= 0) {
            double w = 1.0;
            if (!specular && s->nee && li < s->nlight_nee) {
                double pl = light_pdf(s, li, o, d, h.t);
                w = pdf_prev * pdf_prev / (pdf_prev * pdf_prev + pl * pl);
            }
            const double* L = s->light + 4 * li;
            for (int j = 0; j < nl; ++j) acc[j] += T[j] * planck_norm(lam[j], L[2], L[3]) * w;
            break;
        }
        const double* m = s->mat + 12 * mi;
        int type = (int)m[0];
        double scale = 1e-4 * fmax(1.0, fmax(fabs(p.x), fmax(fabs(p.y), fabs(p.z)))) * 0.05;
        if (type == M_LOBE || type == M_GAUSS) {
            d3 N = dot(n, d) < 0 ? n : mul(n, -1.0);
            double rho[4];
            for (int j = 0; j < 4; ++j) rho[j] = reflect_spec(m, lam[j]);
            d3 P = add(p, mul(N, scale));
            if (s->nlight_nee > 0 && s->nee) {
                int k = (int)(rnd(key, b, 3) * s->nlight_nee);
                if (k >= s->nlight_nee) k = s->nlight_nee - 1;
                const double* L = s->light + 4 * k;
                double u4 = rnd(key, b, 4), u5 = rnd(key, b, 5);
                d3 wi;
                double dist = 0.0;
                int ok = 1;
                if ((int)L[0] == 1) {
                    const double* q = s->quad + 9 * (int)L[1];
                    d3 y = add(add(D3(q[0], q[1], q[2]), mul(D3(q[3], q[4], q[5]), u4)), mul(D3(q[6], q[7], q[8]), u5));
                    d3 v = sub(y, P);
                    dist = sqrt(dot(v, v));
                    wi = mul(v, 1.0 / dist);
                } else {
                    const double* q = s->sph + 4 * (int)L[1];
                    d3 dc = sub(D3(q[0], q[1], q[2]), P);
                    double dc2 = dot(dc, dc);
                    ok = q[3] * q[3] < dc2;
                    double cmax = sqrt(fmax(0.0, 1.0 - q[3] * q[3] / dc2));
                    double ct = 1.0 - u4 * (1.0 - cmax), st = sqrt(fmax(0.0, 1.0 - ct * ct)), ph = 2.0 * M_PI * u5;
                    d3 ax = mul(dc, 1.0 / sqrt(dc2)), t1, t2;
                    basis(ax, &t1, &t2);
                    wi = add(add(mul(t1, st * cos(ph)), mul(t2, st * sin(ph))), mul(ax, ct));
                    double th;
                    ok = ok && hit_sphere(q, P, wi, 1e300, &th);
                    dist = ok ? th : 0.0;
                }
                double cs = dot(N, wi);
                if (ok && cs > 0.0) {
                    hitrec sh;
                    cnt[1]++;
                    if (!trace(A, P, wi, dist * (1.0 - 1e-4), 1, &sh, &cnt[2])) {
                        double pl = light_pdf(s, k, P, wi, dist);
                        double pb = cs / M_PI;
                        double w = pl * pl / (pl * pl + pb * pb);
                        for (int j = 0; j < nl; ++j)
                            acc[j] += T[j] * rho[j] / M_PI * cs * planck_norm(lam[j], L[2], L[3]) / pl * w;
                    }
                }
            }
            double u0 = rnd(key, b, 0), u1 = rnd(key, b, 1);
            double r = sqrt(u0), ph = 2.0 * M_PI * u1, z = sqrt(fmax(0.0, 1.0 - u0));
            d3 t1, t2;
            basis(N, &t1, &t2);
            d = add(add(mul(t1, r * cos(ph)), mul(t2, r * sin(ph))), mul(N, z));
            o = P;
            for (int j = 0; j < 4; ++j) T[j] *= rho[j];
            pdf_prev = z / M_PI;
            specular = 0;
            if (!(z > 0.0)) break;
        } else if (type == M_DIEL) {
            d3 N = n;
            double ci = -dot(d, N);
            int entering = ci > 0;
            if (!entering) { N = mul(N, -1.0); ci = -ci; }
            double eta = ior(m, lam[0]);
            double ei = entering ? 1.0 : eta, et = entering ? eta : 1.0;
            double F = fresnel_diel(ci, ei, et);
            if (rnd(key, b, 2) < F) {
                d = add(d, mul(N, 2.0 * ci));
                o = add(p, mul(N, scale));
            } else {
                double er = ei / et;
                double ct = sqrt(1.0 - er * er * (1.0 - ci * ci));
                d = unit(add(mul(d, er), mul(N, er * ci - ct)));
                o = sub(p, mul(N, scale));
            }
            if ((int)m[2] && nl > 1 && !collapsed) {
                T[0] *= nl;
                T[1] = T[2] = T[3] = 0.0;
                collapsed = 1;
            }
            specular = 1;
        } else {
            d3 N = dot(n, d) < 0 ? n : mul(n, -1.0);
            double ci = -dot(d, N);
            double F = fresnel_cond(ci, m[3], m[4]);
            for (int j = 0; j < 4; ++j) T[j] *= F;
            d = add(d, mul(N, 2.0 * ci));
            o = add(p, mul(N, scale));
            specular = 1;
        }
        if (b + 1 >= s->rr_depth) {
            double mx = fmax(fmax(T[0], T[1]), fmax(T[2], T[3]));
            double q = fmin(mx, 0.95);
            if (!(rnd(key, b, 6) < q)) break;
            for (int j = 0; j < 4; ++j) T[j] /= q;
        }
    }
    double w = s->lrange / nl;
    out[0] = out[1] = out[2] = 0.0;
    for (int j = 0; j < nl; ++j) {
        double c[3];
        cmf(s, lam[j], c);
        for (int k = 0; k < 3; ++k) out[k] += acc[j] * c[k] * w;
    }
}

/* v * RotationMatrix(a) for GLSL's RotationMatrix (src/shader.comp:142-152),
 * M = mX mY mZ: rotate about x, then y, then z (angles in degrees). */
static d3 rot_row(d3 a, d3 v) {
    double x = a.x * M_PI / 180.0, y = a.y * M_PI / 180.0, z = a.z * M_PI / 180.0;
    d3 r1 = D3(v.x, v.y * cos(x) - v.z * sin(x), v.y * sin(x) + v.z * cos(x));
    d3 r2 = D3(r1.x * cos(y) + r1.z * sin(y), r1.y, -r1.x * sin(y) + r1.z * cos(y));
    return D3(r2.x * cos(z) - r2.y * sin(z), r2.x * sin(z) + r2.y * cos(z), r2.z);
}
/* Inverse of rot_row (the matrix is orthogonal): M v. */
static d3 rot_col(d3 a, d3 v) {
    double x = a.x * M_PI / 180.0, y = a.y * M_PI / 180.0, z = a.z * M_PI / 180.0;
    d3 r2 = D3(v.x * cos(z) + v.y * sin(z), -v.x * sin(z) + v.y * cos(z), v.z);
    d3 r1 = D3(r2.x * cos(y) - r2.z * sin(y), r2.y, r2.x * sin(y) + r2.z * cos(y));
    return D3(r1.x, r1.y * cos(x) + r1.z * sin(x), -r1.y * sin(x) + r1.z * cos(x));
}
/* Schott N-BK7, Sellmeier (wavelength in nm). */
static double bk7(double lnm) {
    double l2 = (lnm * 1e-3) * (lnm * 1e-3);
    return sqrt(1.0 + 1.03961212 * l2 / (l2 - 6.00069867e-3) + 0.231792344 * l2 / (l2 - 2.00179144e-2) +
                1.01046945 * l2 / (l2 - 103.560653));
}
/* Snell's law for unit d and a unit normal n facing against d, eta = n1 / n2;
 * the zero vector on total internal reflection (GLSL refract). */
static d3 refract_d(d3 d, d3 n, double eta) {
    double c = dot(n, d);
    double k = 1.0 - eta * eta * (1.0 - c * c);
    if (k < 0.0) return D3(0, 0, 0);
    return sub(mul(d, eta), mul(n, eta * c + sqrt(k)));
}
/* The reference's camera ray at hero wavelength l0: sensor point, aperture
 * point, then two refractions at the faces of the biconvex lens (radius of
 * curvature 2f, sagitta h = 2f - sqrt(4f^2 - r^2), faces at object x <= -t/2
 * and x >= t/2 with their spheres centred at x = +-(2f - t/2 - h); object
 * axes RotationMatrix(0, 90 - yaw, pitch) about a point lensDistance ahead).
 * Entering the glass the index ratio is 1 / n(l) and the wavelength is
 * rescaled by it before the second face, as the reference does
 * (src/shader.comp:1409-1444). Returns 0 when a face is missed. */
static int lens_ray(const Scene* s, uint32_t key, double uvx, double uvy, double l0, d3* po, d3* pd) {
    d3 ca = D3(s->cam_angle[0], s->cam_angle[1], 0.0);
    d3 cam = D3(s->cam_pos[0], s->cam_pos[1], s->cam_pos[2]);
    double sc = -0.5 * s->cam_size;
    d3 o = add(cam, rot_row(ca, D3(uvx * sc, uvy * sc, 0.0)));
    double ux = rnd(key, CAM, 2), uy = rnd(key, CAM, 3);
    double r = 0.5 * s->ap_size * sqrt(ux), ph = 2.0 * M_PI * uy;
    d3 ap = add(cam, rot_row(ca, D3(r * cos(ph), r * sin(ph), s->ap_dist)));
    d3 d = unit(sub(ap, o));
    d3 lp = add(cam, mul(rot_row(ca, D3(0, 0, 1)), s->lens_dist));
    d3 la = D3(0.0, 90.0 - s->cam_angle[1], s->cam_angle[0]);
    double R = 2.0 * s->lens_focal;
    double h = R - sqrt(R * R - s->lens_radius * s->lens_radius);
    double half_t = 0.5 * s->lens_thick;
    double cx[2] = {R - half_t - h, -(R - half_t - h)};
    double l = l0;
    for (int pass = 0; pass < 2; ++pass) {
        d3 lo = rot_row(la, sub(o, lp)), ld = rot_row(la, d);
        double best = INFINITY;
        d3 bn = D3(0, 0, 0);
        int entering = 0;
        for (int k = 0; k < 2; ++k) {
            d3 oc = sub(lo, D3(cx[k], 0, 0));
            double b = dot(oc, ld), c = dot(oc, oc) - R * R, disc = b * b - c;
            if (disc < 0) continue;
            double sq = sqrt(disc);
            double ts[2] = {-b - sq, -b + sq};
            for (int j = 0; j < 2; ++j) {
                double t = ts[j];
                if (!(t > 1e-4) || t >= best) continue;
                double x = lo.x + t * ld.x;
                if (k == 0 ? x > -half_t : x < half_t) continue;
                best = t;
                d3 rad = mul(sub(add(lo, mul(ld, t)), D3(cx[k], 0, 0)), 1.0 / R);
                entering = dot(ld, rad) < 0;
                bn = entering ? rad : mul(rad, -1.0);
            }
        }
        if (!(best < INFINITY)) return 0;
        double eta;
        if (entering) eta = 1.0 / bk7(l);
        else eta = bk7(l);
        l *= eta;
        o = add(o, mul(d, best));
        d = refract_d(d, rot_col(la, bn), eta);
    }
    *po = o;
    *pd = d;
    return 1;
}

/* GLSL RotationMatrix(angle) (src/shader.comp:142-152) applied as v * M. */
static d3 cam_dir(const Scene* s, d3 v) {
    double ax = s->cam_angle[0] * M_PI / 180.0, ay = s->cam_angle[1] * M_PI / 180.0;
    /* rows of M^T = columns of M: M = Rx^T Ry^T (Rz = identity for the camera) */
    double cx = cos(ax), sx = sin(ax), cy = cos(ay), sy = sin(ay);
    /* v * M with M = mX * mY: (v*mX)*mY */
    d3 a = D3(v.x, v.y * cx - v.z * sx, v.y * sx + v.z * cx);
    d3 r = D3(a.x * cy + a.z * sy, a.y, -a.x * sy + a.z * cy);
    return r;
}

/* Renders rows [y0, y1) of the image, `spp` samples per pixel with global
 * sample indices first_sample .. first_sample + spp - 1. sum[3 * (W * (y - y0) + x)]
 * receives the XYZ sum over samples, sq the sum of squares; cnt[0..2] add
 * segments, shadow rays and SDF steps. */
int ptext_render(Scene* s, int first_sample, int spp, int y0, int y1, double* sum, double* sq, uint64_t* cnt,
                 int threads) {
    if (!s || !sum || y0 < 0 || y1 > s->H || y0 >= y1 || spp <= 0) return -1;
    Accel A;
    accel_build(&A, s);
    uint64_t c0 = 0, c1 = 0, c2 = 0, missed = 0;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : c0, c1, c2, missed)
    for (int gy = y0; gy < y1; ++gy) {
        uint64_t lc[3] = {0, 0, 0};
        for (int gx = 0; gx < s->W; ++gx) {
            uint32_t pix = (uint32_t)(gx + s->W * gy);
            double acc[3] = {0, 0, 0}, acc2[3] = {0, 0, 0};
            for (int k = 0; k < spp; ++k) {
                uint32_t key = key_of(s->seed, pix, (uint32_t)(first_sample + k));
                double ju = rnd(key, CAM, 0), jv = rnd(key, CAM, 1);
                double uvx = (2.0 * (gx + ju) - s->W) / s->H;
                double uvy = (2.0 * ((s->H - gy) - jv) - s->H) / s->H;
                double l0 = s->lmin + s->lrange * rnd(key, CAM, 4);
                d3 o, d;
                double x[3] = {0, 0, 0};
                if (s->camera == 1) {
                    double h = 0.5 * s->cam_size;
                    d = unit(cam_dir(s, D3(uvx * h, uvy * h, s->ap_dist)));
                    o = D3(s->cam_pos[0], s->cam_pos[1], s->cam_pos[2]);
                    trace_path(s, &A, key, o, d, l0, x, lc);
                } else if (lens_ray(s, key, uvx, uvy, l0, &o, &d)) {
                    trace_path(s, &A, key, o, d, l0, x, lc);
                } else {
                    ++missed;
                }
                for (int c = 0; c < 3; ++c) {
                    acc[c] += x[c];
                    acc2[c] += x[c] * x[c];
                }
            }
            size_t base = 3 * ((size_t)s->W * (size_t)(gy - y0) + (size_t)gx);
            for (int c = 0; c < 3; ++c) {
                sum[base + c] += acc[c];
                if (sq) sq[base + c] += acc2[c];
            }
        }
        c0 += lc[0]; c1 += lc[1]; c2 += lc[2];
    }
    if (cnt) { cnt[0] += c0; cnt[1] += c1; cnt[2] += c2; }
    s->lens_missed = missed;
    accel_free(&A);
    return 0;
}

/* Closed-form helpers for tests. */
double ptext_fresnel_dielectric(double ci, double eta_i, double eta_t) { return fresnel_diel(ci, eta_i, eta_t); }
double ptext_fresnel_conductor(double ci, double n, double k) { return fresnel_cond(ci, n, k); }
double ptext_planck(double lnm, double T, double lum) { return planck_norm(lnm, T, lum); }

/* Probes of single functions, for tests/test_oracle_ext_pins.py. */
int ptext_probe_quad(const double q[9], const double o[3], const double d[3], double* t) {
    return hit_quad(q, D3(o[0], o[1], o[2]), D3(d[0], d[1], d[2]), 1e300, t);
}
int ptext_probe_sphere(const double q[4], const double o[3], const double d[3], double* t) {
    return hit_sphere(q, D3(o[0], o[1], o[2]), D3(d[0], d[1], d[2]), 1e300, t);
}
/* m = one material row (type n dispersive p[9]) */
int ptext_probe_lens(const double q[12], const double o[3], const double d[3], double* t, double n[3]) {
    d3 nn = D3(0, 0, 0);
    int hit = hit_lens(q, D3(o[0], o[1], o[2]), D3(d[0], d[1], d[2]), 1e300, t, &nn);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    return hit;
}
int ptext_probe_cyclide(const double q[16], const double o[3], const double d[3], double* t, double n[3]) {
    d3 nn = D3(0, 0, 0);
    int hit = hit_cyclide(q, D3(o[0], o[1], o[2]), D3(d[0], d[1], d[2]), 1e300, t, &nn);
    n[0] = nn.x; n[1] = nn.y; n[2] = nn.z;
    return hit;
}
int ptext_probe_poly_roots(const double* c, int n, double lo, double hi, double* out) {
    return poly_roots(c, n, lo, hi, out);
}
double ptext_probe_ior(const double m[12], double lnm) { return ior(m, lnm); }
double ptext_probe_reflectance(const double m[12], double lnm) { return reflect_spec(m, lnm); }
/* Sphere tracing of one SDF record q (kind iters bailout mat cx cy cz sx sy sz)
 * along o + t d; returns t (-1 for no hit) and the estimator evaluations. */
double ptext_probe_march(const double q[10], const double o[3], const double d[3], int max_steps, uint64_t* steps) {
    uint64_t n = 0;
    double t = march(q, D3(o[0], o[1], o[2]), D3(d[0], d[1], d[2]), 1e300, max_steps, &n);
    if (steps) *steps = n;
    return t;
}
double ptext_probe_de(const double q[10], const double p[3]) { return de(q, D3(p[0], p[1], p[2])); }
/* The camera ray of (key, uv, hero wavelength); returns 0 if the lens is missed. */
int ptext_probe_camera(const Scene* s, uint32_t key, double uvx, double uvy, double l0, double o[3], double d[3]) {
    d3 po, pd;
    int ok = lens_ray(s, key, uvx, uvy, l0, &po, &pd);
    o[0] = po.x; o[1] = po.y; o[2] = po.z;
    d[0] = pd.x; d[1] = pd.y; d[2] = pd.z;
    return ok;
}
